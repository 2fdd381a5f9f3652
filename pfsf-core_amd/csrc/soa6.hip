// SoA-6 operator of the drop-in tick (the six face conductivities per voxel
// that pfsf_register_island_buffers delivers, SoA order cond[d * n + i],
// d = -x, +x, -y, +y, -z, +z; PFSFDataBuilder.java:124-132,
// PFSFConductivity.java:37-91).
//
// The reference's GPU form (SH/pcg_matvec.comp.glsl:42-128,
// SH/rbgs_smooth.comp.glsl:116-163) builds every weight of voxel i from i's
// own faces (edge EDGE_P sqrt(c_a c_b), corner CORNER_P cbrt(c_x c_y c_z)), so
// a_ij != a_ji.  Here each pair takes the mean of the two endpoints' half
// weights, which is symmetric and equals the shader's weight whenever both
// endpoints agree:
//   face   a = 0.5 (c_i[d] + c_j[d'])
//   edge   a = EDGE_P   0.5 (sqrt(c_i[da] c_i[db]) + sqrt(c_j[da'] c_j[db']))
//   corner a = CORNER_P 0.5 (cbrt(c_i[dx] c_i[dy] c_i[dz]) + cbrt(...j...))
// (' = the opposite direction), c = cond when 0 < cond <= FLT_MAX, else 0,
// FP64 with products formed left to right.  Pairs with an air or out-of-grid
// endpoint have no coupling.  The cube root uses correctly rounded basic
// operations only (a chord guess and six Newton steps), so the couplings are
// the same bits on any conforming FP64 implementation.
//
// The 13 couplings of a voxel to its forward neighbours (the even offsets of
// the stencil order PFSFStencil.java:70-84) are tabulated once per upload in
// cw[(k / 2) * n + i]; an odd offset reads its opposite's entry at the
// neighbour.  Every kernel here sums in stencil order with no contraction
// (-ffp-contract=off semantics via the pragma), so the failure scan, the
// history update and RBGS are bit-identical to the CPU restatement.
#pragma clang fp contract(off)

#include <math.h>

#include <algorithm>

#include "batch_solver.h"
#include "soa6.h"

namespace pfsf {
namespace {

constexpr int BS = 256;
constexpr int kOpp[26] = {1,  0,  3,  2,  5,  4,  9,  8,  7,  6,  13, 12, 11,
                          10, 17, 16, 15, 14, 25, 24, 23, 22, 21, 20, 19, 18};

__device__ __forceinline__ bool solid(uint8_t t) { return t != 0 && t != 2; }

__device__ __forceinline__ int64_t flat3(int x, int y, int z, pfsf_dims d) {
  return (int64_t)x + (int64_t)d.lx * ((int64_t)y + (int64_t)d.ly * z);
}

struct P3 {
  int x, y, z;
};

__device__ __forceinline__ P3 unflat3(int64_t i, pfsf_dims d) {
  P3 p;
  p.x = (int)(i % d.lx);
  const int64_t r = i / d.lx;
  p.y = (int)(r % d.ly);
  p.z = (int)(r / d.ly);
  return p;
}

__device__ __forceinline__ bool inside(int x, int y, int z, pfsf_dims d) {
  return x >= 0 && x < d.lx && y >= 0 && y < d.ly && z >= 0 && z < d.lz;
}

// Cube root of v >= 0 from frexp / ldexp (exact) and correctly rounded
// +, *, /: v = m 2^(3q), m in [0.5, 4); chord guess, six Newton steps.
__device__ double cbrt_exact(double v) {
  if (!(v > 0.0)) return 0.0;
  if (v > 1.7976931348623157e308) return v;
  int e;
  double m = frexp(v, &e);
  int r = e % 3;
  if (r < 0) r += 3;
  const int q = (e - r) / 3;
  m = ldexp(m, r);
  double y = 0.79370052598409979 + 0.22677165 * (m - 0.5);
#pragma unroll 1
  for (int it = 0; it < 6; it++) y = (2.0 * y + m / (y * y)) / 3.0;
  return ldexp(y, q);
}

__device__ __forceinline__ double cval(const float* __restrict__ cond, int64_t n, int dir,
                                       int64_t i) {
  const float c = cond[dir * n + i];
  return (c > 0.0f && c <= 3.40282347e38f) ? (double)c : 0.0;
}

// Half weight of voxel i toward offset o (flip = -1: toward -o).
__device__ __forceinline__ double half_w(const float* __restrict__ cond, int64_t n, int64_t i,
                                         const int (&o)[3], int flip) {
  int dirs[3], nd = 0;
#pragma unroll
  for (int a = 0; a < 3; a++)
    if (o[a] != 0) dirs[nd++] = 2 * a + (o[a] * flip > 0 ? 1 : 0);
  if (nd == 1) return cval(cond, n, dirs[0], i);
  if (nd == 2) return sqrt(cval(cond, n, dirs[0], i) * cval(cond, n, dirs[1], i));
  return cbrt_exact(cval(cond, n, dirs[0], i) * cval(cond, n, dirs[1], i) *
                    cval(cond, n, dirs[2], i));
}

__global__ void __launch_bounds__(BS)
    k_s6_couple(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ cond,
                double* __restrict__ cw) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const P3 p = unflat3(i, d);
  const bool live = type[i] != 0;
#pragma unroll 1
  for (int k = 0; k < 26; k += 2) {
    const int o[3] = {O[k][0], O[k][1], O[k][2]};
    const int nx = p.x + o[0], ny = p.y + o[1], nz = p.z + o[2];
    double a = 0.0;
    if (live && inside(nx, ny, nz, d)) {
      const int64_t j = flat3(nx, ny, nz, d);
      if (type[j] != 0) {
        const double h = 0.5 * (half_w(cond, n, i, o, 1) + half_w(cond, n, j, o, -1));
        const int nd = (o[0] != 0) + (o[1] != 0) + (o[2] != 0);
        a = nd == 1 ? h : (nd == 2 ? (double)kEdgeP * h : (double)kCornerP * h);
      }
    }
    cw[(k / 2) * n + i] = a;
  }
}

__device__ __forceinline__ double wgt(const double* __restrict__ cw, int64_t n, int k, int64_t i,
                                      int64_t j) {
  return (k & 1) ? cw[(kOpp[k] / 2) * n + j] : cw[(k / 2) * n + i];
}

__global__ void __launch_bounds__(BS)
    k_s6_apply_f64(pfsf_dims d, const uint8_t* __restrict__ type, const double* __restrict__ cw,
                   const double* __restrict__ x, double* __restrict__ y) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (!solid(type[i])) {
    y[i] = 0.0;
    return;
  }
  const P3 p = unflat3(i, d);
  const double xi = x[i];
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 26; k++) {
    const int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
    if (!inside(nx, ny, nz, d)) continue;
    const int64_t j = flat3(nx, ny, nz, d);
    const uint8_t tj = type[j];
    if (tj == 0) continue;
    const double xj = solid(tj) ? x[j] : 0.0;
    acc += wgt(cw, n, k, i, j) * (xi - xj);
  }
  y[i] = acc;
}

__global__ void __launch_bounds__(BS)
    k_s6_failure_scan(pfsf_dims d, const uint8_t* __restrict__ type,
                      const double* __restrict__ cw, const float* __restrict__ phi,
                      const float* __restrict__ rcomp, const float* __restrict__ rtens,
                      float phi_orphan, uint8_t* __restrict__ code) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (!solid(type[i])) {
    code[i] = 0;
    return;
  }
  const P3 p = unflat3(i, d);
  const float phii = phi[i];
  const double pi = (double)phii;
  double diag = 0.0, fin = 0.0, fout = 0.0;
#pragma unroll
  for (int k = 0; k < 26; k++) {
    const int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
    if (!inside(nx, ny, nz, d)) continue;
    const int64_t j = flat3(nx, ny, nz, d);
    if (type[j] == 0) continue;
    const double a = wgt(cw, n, k, i, j);
    diag += a;
    const double dp = (double)phi[j] - pi;
    if (dp > 0.0)
      fin += a * dp;
    else if (dp < 0.0)
      fout += a * (-dp);
  }
  uint8_t c = 0;
  if (diag == 0.0 || phii > phi_orphan)
    c = 3;
  else if (rcomp[i] > 0.0f && fin > (double)rcomp[i])
    c = 2;
  else if (rtens[i] > 0.0f && fout > (double)rtens[i])
    c = 4;
  code[i] = c;
}

__device__ __forceinline__ float grad1(const uint8_t* __restrict__ type,
                                       const float* __restrict__ phi, int64_t i, bool hm, bool hp,
                                       int64_t jm, int64_t jp) {
  const bool om = hm && type[jm] != 0;
  const bool op = hp && type[jp] != 0;
  if (om && op) return (phi[jp] - phi[jm]) * 0.5f;
  if (op) return phi[jp] - phi[i];
  if (om) return phi[i] - phi[jm];
  return 0.0f;
}

// psi = 0.5 (kx gx^2 + ky gy^2 + kz gz^2), k_a = 0.5 (c[-a] + c[+a]); FP32.
__global__ void __launch_bounds__(BS)
    k_s6_history(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ cond,
                 const float* __restrict__ phi, float* __restrict__ H,
                 float* __restrict__ psi_out) {
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (!solid(type[i])) {
    if (psi_out) psi_out[i] = 0.0f;
    return;
  }
  const P3 p = unflat3(i, d);
  const int64_t sy = d.lx, sz = (int64_t)d.lx * d.ly;
  const float gx = grad1(type, phi, i, p.x > 0, p.x + 1 < d.lx, i - 1, i + 1);
  const float gy = grad1(type, phi, i, p.y > 0, p.y + 1 < d.ly, i - sy, i + sy);
  const float gz = grad1(type, phi, i, p.z > 0, p.z + 1 < d.lz, i - sz, i + sz);
  float k[3];
#pragma unroll
  for (int a = 0; a < 3; a++)
    k[a] = 0.5f * ((float)cval(cond, n, 2 * a, i) + (float)cval(cond, n, 2 * a + 1, i));
  float e = k[0] * (gx * gx);
  e = e + k[1] * (gy * gy);
  e = e + k[2] * (gz * gz);
  const float psi = 0.5f * e;
  if (psi_out) psi_out[i] = psi;
  if (H && psi > H[i]) H[i] = psi;
}

// One colour pass of the 8-colour Gauss-Seidel (colour (x&1)|(y&1)<<1|(z&1)<<2,
// SH/rbgs_smooth.comp.glsl:86-87) with the SoA-6 couplings.
__global__ void __launch_bounds__(BS)
    k_s6_rbgs_color(pfsf_dims d, const uint8_t* __restrict__ type, const double* __restrict__ cw,
                    const float* __restrict__ src, float* phi, int c) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  const int hx = (d.lx - cx + 1) >> 1, hy = (d.ly - cy + 1) >> 1, hz = (d.lz - cz + 1) >> 1;
  const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (hx <= 0 || hy <= 0 || hz <= 0 || t >= (int64_t)hx * hy * hz) return;
  P3 p;
  p.x = 2 * (int)(t % hx) + cx;
  const int64_t r = t / hx;
  p.y = 2 * (int)(r % hy) + cy;
  p.z = 2 * (int)(r / hy) + cz;
  const int64_t n = nvox(d);
  const int64_t i = flat3(p.x, p.y, p.z, d);
  if (!solid(type[i])) {
    phi[i] = 0.0f;
    return;
  }
  double diag = 0.0, off = 0.0;
#pragma unroll
  for (int k = 0; k < 26; k++) {
    const int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
    if (!inside(nx, ny, nz, d)) continue;
    const int64_t j = flat3(nx, ny, nz, d);
    if (type[j] == 0) continue;
    const double a = wgt(cw, n, k, i, j);
    diag += a;
    off += a * (double)phi[j];
  }
  const double pj = (diag > 1e-15) ? ((double)src[i] + off) / diag : (double)phi[i];
  phi[i] = (float)pj;
}

// ── PCG of the SoA-6 operator ──────────────────────────────────────────
__device__ __forceinline__ double block_reduce(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BS / 64; w++) t += sh[w];
  return t;
}

__device__ __forceinline__ void put_partial(double v, double* part) {
  __shared__ double sh[BS / 64];
  const double t = block_reduce(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Diagonal and the unknowns: solid, anchored (reach) and D > 0.
__global__ void __launch_bounds__(BS)
    k_s6_setup(pfsf_dims d, const uint8_t* __restrict__ type, const double* __restrict__ cw,
               const uint8_t* __restrict__ reach, double* __restrict__ dg,
               uint8_t* __restrict__ fr) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  if (solid(type[i])) {
    const P3 p = unflat3(i, d);
#pragma unroll
    for (int k = 0; k < 26; k++) {
      const int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
      if (!inside(nx, ny, nz, d)) continue;
      const int64_t j = flat3(nx, ny, nz, d);
      if (type[j] == 0) continue;
      acc += wgt(cw, n, k, i, j);
    }
  }
  dg[i] = acc;
  fr[i] = (solid(type[i]) && reach[i] && acc > 0.0) ? 1 : 0;
}

// x = free ? guess : 0 (guess NULL: 0); partial ||b||^2 over the free set.
__global__ void __launch_bounds__(BS)
    k_s6_begin(int64_t n, const uint8_t* __restrict__ fr, const double* __restrict__ guess,
               double* __restrict__ x, const float* __restrict__ b, double* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  double bb = 0.0;
  if (i < n) {
    const bool f = fr[i] != 0;
    x[i] = (f && guess) ? guess[i] : 0.0;
    const double bi = f ? (double)b[i] : 0.0;
    bb = bi * bi;
  }
  put_partial(bb, part);
}

// y = A v over the free rows (free neighbours at v, anchors 0); with b:
// y = b - A v on the free rows (the residual).  Partial: v.y (b == NULL) or
// y.y (residual).
__global__ void __launch_bounds__(BS)
    k_s6_apply(pfsf_dims d, const uint8_t* __restrict__ fr, const double* __restrict__ cw,
               const double* __restrict__ v, const float* __restrict__ b, double* __restrict__ y,
               double* __restrict__ part, const S6Scalars* __restrict__ sc) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  if (sc && sc->done) return;
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  double pv = 0.0;
  if (i < n) {
    double out = 0.0;
    if (fr[i]) {
      const P3 p = unflat3(i, d);
      const double vi = v[i];
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 26; k++) {
        const int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
        if (!inside(nx, ny, nz, d)) continue;
        const int64_t j = flat3(nx, ny, nz, d);
        const double w = wgt(cw, n, k, i, j);
        const double vj = fr[j] ? v[j] : 0.0;
        acc += w * (vi - vj);
      }
      out = b ? (double)b[i] - acc : acc;
      pv = b ? out * out : vi * out;
    }
    y[i] = out;
  }
  put_partial(pv, part);
}

__device__ __forceinline__ double sum_partials(const double* __restrict__ part, unsigned nb) {
  __shared__ double sh[BS / 64];
  double a = 0.0;
  for (unsigned k = threadIdx.x; k < nb; k += BS) a += part[k];
  return block_reduce(a, sh);  // valid on thread 0
}

enum S6Fin { S6_BB = 0, S6_RR0, S6_PQ, S6_RR, S6_RZ };

__global__ void __launch_bounds__(BS)
    k_s6_fin(int mode, const double* __restrict__ part, unsigned nb, S6Scalars* sc, double tol) {
  if (mode >= S6_PQ && sc->done) return;
  const double v = sum_partials(part, nb);
  if (threadIdx.x != 0) return;
  switch (mode) {
    case S6_BB:
      sc->bb = v;
      break;
    case S6_RR0:  // a pass starts: r = b - A x
      sc->rr = v;
      sc->stop2 = tol * tol * (sc->bb > 0.0 ? sc->bb : 1.0);
      sc->done = (v <= sc->stop2) ? 1 : 0;
      sc->first = 1;
      break;
    case S6_PQ:
      sc->pq = v;
      if (v > 0.0) {
        sc->alpha = sc->rz / v;
      } else {
        sc->alpha = 0.0;
        sc->done = 2;
      }
      break;
    case S6_RR:
      sc->rr = v;
      sc->iters += 1;
      if (v <= sc->stop2 || sc->iters >= sc->max_iters) sc->done = 1;
      break;
    case S6_RZ:
      sc->beta = sc->first ? 0.0 : v / sc->rz;
      sc->first = 0;
      sc->rz = v;
      break;
  }
}

// x += alpha p, r -= alpha q on the free set; the V-cycle's FP32 input
// (zero outside both free sets); partial r.r.
__global__ void __launch_bounds__(BS)
    k_s6_update(int64_t n, const uint8_t* __restrict__ fr, const uint8_t* __restrict__ mgfr,
                double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                const double* __restrict__ q, float* __restrict__ rmg, double* __restrict__ part,
                const S6Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  double rr = 0.0;
  if (i < n) {
    float rm = 0.0f;
    if (fr[i]) {
      const double a = sc->alpha;
      x[i] += a * p[i];
      const double ri = r[i] - a * q[i];
      r[i] = ri;
      rr = ri * ri;
      rm = mgfr[i] ? (float)ri : 0.0f;
    }
    rmg[i] = rm;
  }
  put_partial(rr, part);
}

// The V-cycle's FP32 input from r (the start of a pass).
__global__ void __launch_bounds__(BS)
    k_s6_to_mg(int64_t n, const uint8_t* __restrict__ fr, const uint8_t* __restrict__ mgfr,
               const double* __restrict__ r, float* __restrict__ rmg) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) rmg[i] = (fr[i] && mgfr[i]) ? (float)r[i] : 0.0f;
}

// z = V-cycle output on its free set, a Jacobi step r / D on unknowns outside
// it, 0 elsewhere; partial r.z.
__global__ void __launch_bounds__(BS)
    k_s6_zfix(int64_t n, const uint8_t* __restrict__ fr, const uint8_t* __restrict__ mgfr,
              const float* __restrict__ zmg, const double* __restrict__ r,
              const double* __restrict__ dg, double* __restrict__ z, double* __restrict__ part,
              const S6Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  double rz = 0.0;
  if (i < n) {
    double zi = 0.0;
    if (fr[i]) {
      zi = mgfr[i] ? (double)zmg[i] : r[i] / dg[i];
      rz = r[i] * zi;
    }
    z[i] = zi;
  }
  put_partial(rz, part);
}

// p = z + beta p.
__global__ void __launch_bounds__(BS)
    k_s6_dir(int64_t n, const double* __restrict__ z, double* __restrict__ p,
             const S6Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) p[i] = sc->beta == 0.0 ? z[i] : z[i] + sc->beta * p[i];
}

__global__ void __launch_bounds__(BS)
    k_s6_end(int64_t n, const double* __restrict__ x, double* __restrict__ x64,
             float* __restrict__ phi32) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (x64 && x64 != x) x64[i] = x[i];
  if (phi32) phi32[i] = (float)x[i];
}

__global__ void k_s6_start(S6Scalars* sc, int max_iters) {
  sc->done = 0;
  sc->iters = 0;
  sc->max_iters = max_iters;
  sc->first = 1;
}

// The next pass may run only when the last one left room and made progress.
__global__ void k_s6_pass(S6Scalars* sc) {
  if (sc->iters < sc->max_iters) sc->done = 0;
}

// ── batched SoA-6 PCG (Soa6Batch) ──────────────────────────────────────
// One 256-cell chunk of one island per work-group; per-island scalars; an
// island whose `done` says so takes no part (its work-groups return at
// once).  The per-cell work is that of the single-island kernels above.
constexpr int kS6Finished = 3;

struct ChunkCell {
  int k;      // island
  int64_t i;  // cell (valid when < I.n)
};

__device__ __forceinline__ ChunkCell chunk_cell(const S6Chunk* __restrict__ ch) {
  const S6Chunk c = ch[blockIdx.x];
  return ChunkCell{c.island, (int64_t)c.begin + threadIdx.x};
}

__global__ void __launch_bounds__(BS)
    ks6_begin(const S6Chunk* __restrict__ ch, const S6Isl* __restrict__ isl,
              double* __restrict__ part) {
  const ChunkCell c = chunk_cell(ch);
  const S6Isl& I = isl[c.k];
  double bb = 0.0;
  if (c.i < I.n) {
    const bool f = I.fr[c.i] != 0;
    I.x[c.i] = (f && I.guess) ? I.guess[c.i] : 0.0;
    const double bi = f ? (double)I.b[c.i] : 0.0;
    bb = bi * bi;
  }
  put_partial(bb, part);
}

__global__ void ks6_start(S6Scalars* sc, int n, int max_iters) {
  const int k = blockIdx.x * BS + threadIdx.x;
  if (k >= n) return;
  sc[k].done = 0;
  sc[k].iters = 0;
  sc[k].max_iters = max_iters;
  sc[k].first = 1;
  sc[k].pass_end = 0;
}

// resid: r = b - A x and r.r partials on every unfinished island; else
// q = A p and p.q partials on every iterating island (as k_s6_apply).
__global__ void __launch_bounds__(BS)
    ks6_apply(const S6Chunk* __restrict__ ch, const S6Isl* __restrict__ isl,
              const S6Scalars* __restrict__ sc, int resid, double* __restrict__ part) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  const ChunkCell c = chunk_cell(ch);
  const int dn = sc[c.k].done;
  if (resid ? dn == kS6Finished : dn != 0) return;
  const S6Isl& I = isl[c.k];
  const pfsf_dims d = I.d;
  const int64_t n = I.n, i = c.i;
  const double* __restrict__ v = resid ? I.x : I.p;
  double* __restrict__ y = resid ? I.r : I.q;
  double pv = 0.0;
  if (i < n) {
    double out = 0.0;
    if (I.fr[i]) {
      const P3 q3 = unflat3(i, d);
      const double vi = v[i];
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 26; k++) {
        const int nx = q3.x + O[k][0], ny = q3.y + O[k][1], nz = q3.z + O[k][2];
        if (!inside(nx, ny, nz, d)) continue;
        const int64_t j = flat3(nx, ny, nz, d);
        const double w = wgt(I.cw, n, k, i, j);
        const double vj = I.fr[j] ? v[j] : 0.0;
        acc += w * (vi - vj);
      }
      out = resid ? (double)I.b[i] - acc : acc;
      pv = resid ? out * out : vi * out;
    }
    y[i] = out;
  }
  put_partial(pv, part);
}

// Per island (one work-group each): the scalar step of k_s6_fin from the
// island's chunk partials.  With vsc, the V-cycle of an island that stops
// is switched off (vsc[k].done).
__global__ void __launch_bounds__(BS)
    ks6_fin(int mode, const S6Isl* __restrict__ isl, const double* __restrict__ part,
            S6Scalars* sc, double tol, PcgScalars* vsc) {
  const int k = blockIdx.x;
  S6Scalars* s = sc + k;
  const int dn = s->done;
  if (mode == S6_RR0 ? dn == kS6Finished : (mode >= S6_PQ && dn != 0)) return;
  const S6Isl& I = isl[k];
  const double v = sum_partials(part + I.cb, (unsigned)I.cn);
  if (threadIdx.x != 0) return;
  switch (mode) {
    case S6_BB:
      s->bb = v;
      break;
    case S6_RR0:
      s->pass_end = dn;
      s->rr = v;
      s->stop2 = tol * tol * (s->bb > 0.0 ? s->bb : 1.0);
      s->done = (v <= s->stop2) ? 1 : 0;
      s->first = 1;
      break;
    case S6_PQ:
      s->pq = v;
      if (v > 0.0) {
        s->alpha = s->rz / v;
      } else {
        s->alpha = 0.0;
        s->done = 2;
      }
      break;
    case S6_RR:
      s->rr = v;
      s->iters += 1;
      if (v <= s->stop2 || s->iters >= s->max_iters) s->done = 1;
      break;
    case S6_RZ:
      s->beta = s->first ? 0.0 : v / s->rz;
      s->first = 0;
      s->rz = v;
      break;
  }
  if (vsc) vsc[k].done = s->done != 0 ? 1 : 0;
}

// A pass starts for the islands with flags[k] (those left with room
// continue, as k_s6_pass); the others are finished for this solve.
__global__ void ks6_pass(S6Scalars* sc, int n, const int* __restrict__ flags, PcgScalars* vsc) {
  const int k = blockIdx.x * BS + threadIdx.x;
  if (k >= n) return;
  if (!flags[k]) {
    sc[k].done = kS6Finished;
  } else if (sc[k].iters < sc[k].max_iters) {
    sc[k].done = 0;
  }
  vsc[k].done = sc[k].done != 0 ? 1 : 0;
}

__global__ void __launch_bounds__(BS)
    ks6_to_mg(const S6Chunk* __restrict__ ch, const S6Isl* __restrict__ isl,
              const S6Scalars* __restrict__ sc) {
  const ChunkCell c = chunk_cell(ch);
  if (sc[c.k].done) return;
  const S6Isl& I = isl[c.k];
  if (c.i < I.n) I.rmg[c.i] = (I.fr[c.i] && I.mgfr[c.i]) ? (float)I.r[c.i] : 0.0f;
}

__global__ void __launch_bounds__(BS)
    ks6_update(const S6Chunk* __restrict__ ch, const S6Isl* __restrict__ isl,
               const S6Scalars* __restrict__ sc, double* __restrict__ part) {
  const ChunkCell c = chunk_cell(ch);
  const S6Scalars& S = sc[c.k];
  if (S.done) return;
  const S6Isl& I = isl[c.k];
  double rr = 0.0;
  if (c.i < I.n) {
    const int64_t i = c.i;
    float rm = 0.0f;
    if (I.fr[i]) {
      const double a = S.alpha;
      I.x[i] += a * I.p[i];
      const double ri = I.r[i] - a * I.q[i];
      I.r[i] = ri;
      rr = ri * ri;
      rm = I.mgfr[i] ? (float)ri : 0.0f;
    }
    I.rmg[i] = rm;
  }
  put_partial(rr, part);
}

__global__ void __launch_bounds__(BS)
    ks6_zfix(const S6Chunk* __restrict__ ch, const S6Isl* __restrict__ isl,
             const S6Scalars* __restrict__ sc, double* __restrict__ part) {
  const ChunkCell c = chunk_cell(ch);
  if (sc[c.k].done) return;
  const S6Isl& I = isl[c.k];
  double rz = 0.0;
  if (c.i < I.n) {
    const int64_t i = c.i;
    double zi = 0.0;
    if (I.fr[i]) {
      zi = I.mgfr[i] ? (double)I.zmg[i] : I.r[i] / I.dg[i];
      rz = I.r[i] * zi;
    }
    I.z[i] = zi;
  }
  put_partial(rz, part);
}

__global__ void __launch_bounds__(BS)
    ks6_dir(const S6Chunk* __restrict__ ch, const S6Isl* __restrict__ isl,
            const S6Scalars* __restrict__ sc) {
  const ChunkCell c = chunk_cell(ch);
  const S6Scalars& S = sc[c.k];
  if (S.done) return;
  const S6Isl& I = isl[c.k];
  if (c.i < I.n) I.p[c.i] = S.beta == 0.0 ? I.z[c.i] : I.z[c.i] + S.beta * I.p[c.i];
}

__global__ void __launch_bounds__(BS)
    ks6_end(const S6Chunk* __restrict__ ch, const S6Isl* __restrict__ isl) {
  const ChunkCell c = chunk_cell(ch);
  const S6Isl& I = isl[c.k];
  if (c.i >= I.n) return;
  const double v = I.x[c.i];
  if (I.xio && I.xio != I.x) I.xio[c.i] = v;
  if (I.phio) I.phio[c.i] = (float)v;
}

// Islands still iterating (done == 0), into live[0].
__global__ void __launch_bounds__(1024) ks6_live(const S6Scalars* __restrict__ sc, int n, int* live) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int c = 0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) c += sc[k].done == 0;
  if (c) atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) live[0] = cnt;
}

}  // namespace

// ── launch wrappers ────────────────────────────────────────────────────
hipError_t launch_soa6_couplings(pfsf_dims d, const uint8_t* type, const float* cond,
                                 double* cw, hipStream_t s) {
  if (nvox(d) == 0) return hipSuccess;
  k_s6_couple<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, cond, cw);
  return hipGetLastError();
}

hipError_t launch_soa6_apply_f64(pfsf_dims d, const uint8_t* type, const double* cw,
                                 const double* x, double* y, hipStream_t s) {
  if (nvox(d) == 0) return hipSuccess;
  k_s6_apply_f64<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, cw, x, y);
  return hipGetLastError();
}

hipError_t launch_soa6_failure_scan(pfsf_dims d, const uint8_t* type, const double* cw,
                                    const float* phi, const float* rcomp, const float* rtens,
                                    float phi_orphan, uint8_t* code, hipStream_t s) {
  if (nvox(d) == 0) return hipSuccess;
  k_s6_failure_scan<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, cw, phi, rcomp, rtens,
                                                           phi_orphan, code);
  return hipGetLastError();
}

hipError_t launch_soa6_history(pfsf_dims d, const uint8_t* type, const float* cond,
                               const float* phi, float* H, float* psi, hipStream_t s) {
  if (nvox(d) == 0) return hipSuccess;
  k_s6_history<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, cond, phi, H, psi);
  return hipGetLastError();
}

hipError_t launch_soa6_rbgs(pfsf_dims d, const uint8_t* type, const double* cw,
                            const float* src, float* phi, int sweeps, hipStream_t s) {
  for (int sw = 0; sw < sweeps; sw++)
    for (int c = 0; c < 8; c++) {
      const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
      const int64_t hn = (int64_t)((d.lx - cx + 1) >> 1) * ((d.ly - cy + 1) >> 1) *
                         ((d.lz - cz + 1) >> 1);
      if (hn <= 0) continue;
      k_s6_rbgs_color<<<blocks_for(hn, BS), BS, 0, s>>>(d, type, cw, src, phi, c);
    }
  return hipGetLastError();
}

// ── Soa6Solver ─────────────────────────────────────────────────────────
static inline bool ck(hipError_t e, hipError_t& keep) {
  if (e != hipSuccess && keep == hipSuccess) keep = e;
  return e != hipSuccess;
}

Soa6Solver::~Soa6Solver() { release(); }

void Soa6Solver::release() {
  if (mg_) (void)hipStreamSynchronize(mg_->stream());
  for (void* p : owned_) (void)hipFree(p);
  owned_.clear();
  if (sc_host_) (void)hipHostFree(sc_host_);
  sc_host_ = nullptr;
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  ev0_ = ev1_ = nullptr;
  mg_.reset();
}

int64_t Soa6Solver::device_bytes() const {
  // D, r, p, q, z (FP64), own iterate, free and reach masks, partials
  return n_ * (5 * 8 + 8 + 2) + 8 * (int64_t)nblk_ + 256;
}

hipError_t Soa6Solver::create(pfsf_dims d, const uint8_t* type, const float* cond,
                              const double* cw, const float* proxy_kappa, bool multigrid,
                              double tol, int max_iters, hipStream_t caller, bool batchable) {
  hipError_t err = hipSuccess;
  d_ = d;
  n_ = nvox(d);
  if (n_ > kMaxSolverCells) return hipErrorInvalidValue;
  type_ = type;
  cond_ = cond;
  cw_ = cw;
  kappa_ = proxy_kappa;
  tol_ = tol > 0 ? tol : 1e-6;
  max_iters_ = max_iters > 0 ? max_iters : 500;
  caller_ = caller;
  pfsf_solver_desc sd{};
  sd.dims = d;
  sd.type = type;
  sd.kappa = proxy_kappa;
  sd.mg_levels = multigrid ? 0 : 1;
  sd.precision = PFSF_PREC_FP32;
  sd.tol = tol_;
  sd.max_iters = max_iters_;
  sd.stream = caller;
  mg_ = std::make_unique<MgSolver>();
  if (ck(mg_->create(sd, batchable && multigrid), err)) return err;
  nblk_ = blocks_for(n_, BS);
  auto take = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (ck(hipMalloc(&p, std::max<size_t>(bytes, 8)), err)) return nullptr;
    owned_.push_back(p);
    return p;
  };
  dg_ = (double*)take(8 * n_);
  r_ = (double*)take(8 * n_);
  p_ = (double*)take(8 * n_);
  q_ = (double*)take(8 * n_);
  z_ = (double*)take(8 * n_);
  xo_ = (double*)take(8 * n_);
  free_ = (uint8_t*)take(n_);
  part_ = (double*)take(8 * (size_t)nblk_);
  sc_ = (S6Scalars*)take(sizeof(S6Scalars));
  if (err != hipSuccess) return err;
  if (ck(hipHostMalloc((void**)&sc_host_, sizeof(S6Scalars)), err)) return err;
  ck(hipEventCreate(&ev0_), err);
  ck(hipEventCreate(&ev1_), err);
  if (err != hipSuccess) return err;
  hipStream_t s = mg_->stream();
  ck(hipMemsetAsync(p_, 0, 8 * n_, s), err);
  ck(hipMemsetAsync(z_, 0, 8 * n_, s), err);
  if (err != hipSuccess) return err;
  return refresh();  // the proxy was just built from the same type and kappa
}

hipError_t Soa6Solver::rebuild() {
  hipError_t err = hipSuccess;
  hipStream_t s = mg_->stream();
  // the caller's stream holds the uploads of type and kappa
  ck(hipEventRecord(ev0_, caller_), err);
  ck(hipStreamWaitEvent(s, ev0_, 0), err);
  // the proxy's build also forms the orphan mask of type (MgSolver::reach)
  ck(mg_->update(type_, kappa_), err);
  return err;
}

hipError_t Soa6Solver::refresh() {
  hipError_t err = hipSuccess;
  hipStream_t s = mg_->stream();
  ck(hipEventRecord(ev0_, caller_), err);
  ck(hipStreamWaitEvent(s, ev0_, 0), err);
  k_s6_setup<<<nblk_, BS, 0, s>>>(d_, type_, cw_, mg_->reach(), dg_, free_);
  ck(hipGetLastError(), err);
  return err;
}

// r = b - A x on the free set; *rel = ||r|| / ||b|| (synchronises).
hipError_t Soa6Solver::residual(const float* b, double* x, double* rel) {
  hipError_t err = hipSuccess;
  hipStream_t s = mg_->stream();
  k_s6_apply<<<nblk_, BS, 0, s>>>(d_, free_, cw_, x, b, r_, part_, nullptr);
  k_s6_fin<<<1, BS, 0, s>>>(S6_RR0, part_, nblk_, sc_, tol_);
  ck(hipGetLastError(), err);
  ck(hipMemcpyAsync(sc_host_, sc_, sizeof(S6Scalars), hipMemcpyDeviceToHost, s), err);
  ck(hipStreamSynchronize(s), err);
  const double bb = sc_host_->bb > 0.0 ? sc_host_->bb : 1.0;
  *rel = sqrt(sc_host_->rr / bb);
  return err;
}

hipError_t Soa6Solver::queue_iteration(double* x) {
  hipStream_t s = mg_->stream();
  k_s6_apply<<<nblk_, BS, 0, s>>>(d_, free_, cw_, p_, nullptr, q_, part_, sc_);
  k_s6_fin<<<1, BS, 0, s>>>(S6_PQ, part_, nblk_, sc_, tol_);
  k_s6_update<<<nblk_, BS, 0, s>>>(n_, free_, mg_->free_mask(), x, r_, p_, q_,
                                   mg_->precond_in(), part_, sc_);
  k_s6_fin<<<1, BS, 0, s>>>(S6_RR, part_, nblk_, sc_, tol_);
  hipError_t e = mg_->precond_apply();
  if (e != hipSuccess) return e;
  k_s6_zfix<<<nblk_, BS, 0, s>>>(n_, free_, mg_->free_mask(), mg_->precond_out(), r_, dg_, z_,
                                 part_, sc_);
  k_s6_fin<<<1, BS, 0, s>>>(S6_RZ, part_, nblk_, sc_, tol_);
  k_s6_dir<<<nblk_, BS, 0, s>>>(n_, z_, p_, sc_);
  return hipGetLastError();
}

// Passes of FP64 PCG from the true residual; each pass stops when its
// recursive residual reaches the tolerance, then the true residual decides.
hipError_t Soa6Solver::solve(const float* b, double* x64, float* phi32, pfsf_solve_info* info) {
  hipError_t err = hipSuccess;
  hipStream_t s = mg_->stream();
  double* x = x64 ? x64 : xo_;
  ck(hipEventRecord(ev0_, caller_), err);
  ck(hipStreamWaitEvent(s, ev0_, 0), err);
  ck(hipEventRecord(ev0_, s), err);
  k_s6_begin<<<nblk_, BS, 0, s>>>(n_, free_, x64, x, b, part_);
  k_s6_fin<<<1, BS, 0, s>>>(S6_BB, part_, nblk_, sc_, tol_);
  k_s6_start<<<1, 1, 0, s>>>(sc_, max_iters_);
  ck(mg_->precond_begin(), err);
  ck(hipGetLastError(), err);
  double rel = 0.0;
  ck(residual(b, x, &rel), err);
  int total = 0, passes = 0;
  while (err == hipSuccess && rel > tol_ && total < max_iters_ && passes < 8) {
    // pass start: z = M r, p = z
    k_s6_pass<<<1, 1, 0, s>>>(sc_);
    k_s6_to_mg<<<nblk_, BS, 0, s>>>(n_, free_, mg_->free_mask(), r_, mg_->precond_in());
    ck(mg_->precond_apply(), err);
    k_s6_zfix<<<nblk_, BS, 0, s>>>(n_, free_, mg_->free_mask(), mg_->precond_out(), r_, dg_, z_,
                                   part_, sc_);
    k_s6_fin<<<1, BS, 0, s>>>(S6_RZ, part_, nblk_, sc_, tol_);
    k_s6_dir<<<nblk_, BS, 0, s>>>(n_, z_, p_, sc_);
    ck(hipGetLastError(), err);
    int chunk = 4, before = sc_host_->iters;
    for (;;) {
      for (int k = 0; k < chunk && err == hipSuccess; k++) ck(queue_iteration(x), err);
      ck(hipMemcpyAsync(sc_host_, sc_, sizeof(S6Scalars), hipMemcpyDeviceToHost, s), err);
      ck(hipStreamSynchronize(s), err);
      if (err != hipSuccess || sc_host_->done) break;
      chunk = std::min(chunk * 2, 32);
    }
    const int it = sc_host_->iters - before;
    total = sc_host_->iters;
    passes++;
    const bool breakdown = sc_host_->done == 2;
    double nrel = rel;
    ck(residual(b, x, &nrel), err);
    const bool stalled = it == 0 || (breakdown && nrel >= rel);
    rel = nrel;
    if (stalled) break;
  }
  k_s6_end<<<nblk_, BS, 0, s>>>(n_, x, x64, phi32);
  ck(hipGetLastError(), err);
  ck(hipEventRecord(ev1_, s), err);
  ck(hipEventSynchronize(ev1_), err);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  ck(hipStreamWaitEvent(caller_, ev1_, 0), err);
  if (info) {
    info->iterations = total;
    info->refinements = passes;
    info->converged = rel <= tol_ ? 1 : 0;
    info->levels = mg_->levels();
    info->rel_residual = rel;
    info->solve_ms = ms;
  }
  return err;
}

// ── Soa6Batch ──────────────────────────────────────────────────────────
Soa6Batch::~Soa6Batch() { release(); }

void Soa6Batch::release() {
  if (vc_ && vc_->stream()) (void)hipStreamSynchronize(vc_->stream());
  for (void* q : owned_) (void)hipFree(q);
  owned_.clear();
  if (sc_host_) (void)hipHostFree(sc_host_);
  if (flags_host_) (void)hipHostFree(flags_host_);
  if (live_host_) (void)hipHostFree(live_host_);
  sc_host_ = nullptr;
  flags_host_ = live_host_ = nullptr;
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  ev0_ = ev1_ = nullptr;
  vc_.reset();
}

hipError_t Soa6Batch::attach(const std::vector<Soa6Solver*>& solvers, double tol, int max_iters,
                             hipStream_t caller) {
  hipError_t err = hipSuccess;
  if (solvers.empty()) return hipErrorInvalidValue;
  sol_ = solvers;
  caller_ = caller;
  tol_ = tol > 0 ? tol : 1e-6;
  max_iters_ = max_iters > 0 ? max_iters : 500;
  std::vector<MgSolver*> proxies;
  for (Soa6Solver* v : sol_) proxies.push_back(v->mg_.get());
  vc_ = std::make_unique<BatchSolver>();
  if (ck(vc_->attach(proxies, tol_, max_iters_, caller_), err)) return err;
  const int n = (int)sol_.size();
  host_.assign(n, S6Isl{});
  std::vector<S6Chunk> ch;
  for (int k = 0; k < n; k++) {
    Soa6Solver& v = *sol_[k];
    S6Isl& I = host_[k];
    I.d = v.d_;
    I.n = v.n_;
    I.fr = v.free_;
    I.mgfr = v.mg_->free_mask();
    I.cw = v.cw_;
    I.dg = v.dg_;
    I.r = v.r_;
    I.p = v.p_;
    I.q = v.q_;
    I.z = v.z_;
    I.rmg = v.mg_->precond_in();
    I.zmg = v.mg_->precond_out();
    I.cb = (int)ch.size();
    for (int64_t c0 = 0; c0 < v.n_; c0 += BS) ch.push_back(S6Chunk{k, (int)c0});
    I.cn = (int)ch.size() - I.cb;
  }
  nch_ = (int)ch.size();
  auto take = [&](size_t bytes) -> void* {
    void* q = nullptr;
    if (ck(hipMalloc(&q, std::max<size_t>(bytes, 8)), err)) return nullptr;
    owned_.push_back(q);
    return q;
  };
  isl_ = (S6Isl*)take(sizeof(S6Isl) * n);
  ch_ = (S6Chunk*)take(sizeof(S6Chunk) * std::max(nch_, 1));
  part_ = (double*)take(sizeof(double) * std::max(nch_, 1));
  sc_ = (S6Scalars*)take(sizeof(S6Scalars) * n);
  flags_ = (int*)take(sizeof(int) * n);
  live_ = (int*)take(sizeof(int));
  if (err != hipSuccess) return err;
  ck(hipHostMalloc((void**)&sc_host_, sizeof(S6Scalars) * n), err);
  ck(hipHostMalloc((void**)&flags_host_, sizeof(int) * n), err);
  ck(hipHostMalloc((void**)&live_host_, sizeof(int)), err);
  ck(hipEventCreate(&ev0_), err);
  ck(hipEventCreate(&ev1_), err);
  if (err != hipSuccess) return err;
  if (nch_ > 0)
    ck(hipMemcpyAsync(ch_, ch.data(), sizeof(S6Chunk) * nch_, hipMemcpyHostToDevice,
                      vc_->stream()),
       err);
  ck(hipStreamSynchronize(vc_->stream()), err);
  return err;
}

// One PCG iteration of every iterating island (Soa6Solver::queue_iteration).
hipError_t Soa6Batch::iteration() {
  hipStream_t s = vc_->stream();
  const int n = (int)sol_.size();
  PcgScalars* vsc = vc_->island_scalars();
  ks6_apply<<<nch_, BS, 0, s>>>(ch_, isl_, sc_, 0, part_);
  ks6_fin<<<n, BS, 0, s>>>(S6_PQ, isl_, part_, sc_, tol_, vsc);
  ks6_update<<<nch_, BS, 0, s>>>(ch_, isl_, sc_, part_);
  ks6_fin<<<n, BS, 0, s>>>(S6_RR, isl_, part_, sc_, tol_, vsc);
  hipError_t e = vc_->apply_preconditioner();
  if (e != hipSuccess) return e;
  ks6_zfix<<<nch_, BS, 0, s>>>(ch_, isl_, sc_, part_);
  ks6_fin<<<n, BS, 0, s>>>(S6_RZ, isl_, part_, sc_, tol_, nullptr);
  ks6_dir<<<nch_, BS, 0, s>>>(ch_, isl_, sc_);
  return hipGetLastError();
}

// Soa6Solver::solve for every island at once: per island the same passes
// (from the FP64 true residual, stop at the tolerance, a stalled pass ends
// the island's solve), the host following the count of islands still
// iterating after every few iterations.
hipError_t Soa6Batch::solve(const float* const* b, double* const* x64, float* const* phi32,
                            pfsf_solve_info* per_island) {
  hipError_t err = hipSuccess;
  const int n = (int)sol_.size();
  if (n == 0 || nch_ == 0) return hipSuccess;
  hipStream_t s = vc_->stream();
  PcgScalars* vsc = vc_->island_scalars();
  for (int k = 0; k < n; k++) {
    S6Isl& I = host_[k];
    I.b = b[k];
    I.guess = x64 ? x64[k] : nullptr;
    I.x = (x64 && x64[k]) ? x64[k] : sol_[k]->xo_;
    I.xio = x64 ? x64[k] : nullptr;
    I.phio = phi32 ? phi32[k] : nullptr;
  }
  ck(hipEventRecord(ev0_, caller_), err);  // the islands' arrays were uploaded there
  ck(hipStreamWaitEvent(s, ev0_, 0), err);
  // each solver's set-up (diagonal, free set: Soa6Solver::refresh) runs on
  // its proxy's stream
  for (Soa6Solver* v : sol_) {
    ck(hipEventRecord(ev0_, v->mg_->stream()), err);
    ck(hipStreamWaitEvent(s, ev0_, 0), err);
  }
  ck(hipEventRecord(ev0_, s), err);
  ck(hipMemcpyAsync(isl_, host_.data(), sizeof(S6Isl) * n, hipMemcpyHostToDevice, s), err);
  const unsigned nbi = (unsigned)((n + BS - 1) / BS);
  ks6_begin<<<nch_, BS, 0, s>>>(ch_, isl_, part_);
  ks6_fin<<<n, BS, 0, s>>>(S6_BB, isl_, part_, sc_, tol_, nullptr);
  ks6_start<<<nbi, BS, 0, s>>>(sc_, n, max_iters_);
  ks6_apply<<<nch_, BS, 0, s>>>(ch_, isl_, sc_, 1, part_);
  ks6_fin<<<n, BS, 0, s>>>(S6_RR0, isl_, part_, sc_, tol_, nullptr);
  ck(hipGetLastError(), err);
  ck(hipMemcpyAsync(sc_host_, sc_, sizeof(S6Scalars) * n, hipMemcpyDeviceToHost, s), err);
  ck(hipStreamSynchronize(s), err);
  std::vector<double> rel(n, 0.0);
  std::vector<int> passes(n, 0), fin(n, 0);
  auto rel_of = [&](int k) {
    const double bb = sc_host_[k].bb > 0.0 ? sc_host_[k].bb : 1.0;
    return sqrt(sc_host_[k].rr / bb);
  };
  for (int k = 0; k < n; k++) {
    rel[k] = rel_of(k);
    fin[k] = !(rel[k] > tol_) || sc_host_[k].iters >= max_iters_;
  }
  std::vector<int> before(n, 0);
  for (int pass = 0; pass < 8 && err == hipSuccess; pass++) {
    int active = 0;
    for (int k = 0; k < n; k++) {
      flags_host_[k] = fin[k] ? 0 : 1;
      active += flags_host_[k];
      before[k] = sc_host_[k].iters;
    }
    if (active == 0) break;
    ck(hipMemcpyAsync(flags_, flags_host_, sizeof(int) * n, hipMemcpyHostToDevice, s), err);
    ks6_pass<<<nbi, BS, 0, s>>>(sc_, n, flags_, vsc);
    // pass start: z = M r, p = z
    ks6_to_mg<<<nch_, BS, 0, s>>>(ch_, isl_, sc_);
    ck(vc_->apply_preconditioner(), err);
    ks6_zfix<<<nch_, BS, 0, s>>>(ch_, isl_, sc_, part_);
    ks6_fin<<<n, BS, 0, s>>>(S6_RZ, isl_, part_, sc_, tol_, nullptr);
    ks6_dir<<<nch_, BS, 0, s>>>(ch_, isl_, sc_);
    ck(hipGetLastError(), err);
    int chunk = 4;
    for (int queued = 0; err == hipSuccess && queued < max_iters_;) {
      for (int k = 0; k < chunk && err == hipSuccess; k++) ck(iteration(), err);
      queued += chunk;
      ks6_live<<<1, 1024, 0, s>>>(sc_, n, live_);
      ck(hipMemcpyAsync(live_host_, live_, sizeof(int), hipMemcpyDeviceToHost, s), err);
      ck(hipStreamSynchronize(s), err);
      if (err != hipSuccess || *live_host_ == 0) break;
      chunk = std::min(chunk * 2, 32);
    }
    // the true residual of every island that took part
    ks6_apply<<<nch_, BS, 0, s>>>(ch_, isl_, sc_, 1, part_);
    ks6_fin<<<n, BS, 0, s>>>(S6_RR0, isl_, part_, sc_, tol_, nullptr);
    ck(hipGetLastError(), err);
    ck(hipMemcpyAsync(sc_host_, sc_, sizeof(S6Scalars) * n, hipMemcpyDeviceToHost, s), err);
    ck(hipStreamSynchronize(s), err);
    for (int k = 0; k < n && err == hipSuccess; k++) {
      if (fin[k]) continue;
      const int it = sc_host_[k].iters - before[k];
      passes[k]++;
      const bool breakdown = sc_host_[k].pass_end == 2;
      const double nrel = rel_of(k);
      const bool stalled = it == 0 || (breakdown && nrel >= rel[k]);
      rel[k] = nrel;
      if (stalled || !(rel[k] > tol_) || sc_host_[k].iters >= max_iters_) fin[k] = 1;
    }
  }
  ks6_end<<<nch_, BS, 0, s>>>(ch_, isl_);
  ck(hipGetLastError(), err);
  ck(hipEventRecord(ev1_, s), err);
  ck(hipEventSynchronize(ev1_), err);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  ck(hipStreamWaitEvent(caller_, ev1_, 0), err);
  if (per_island)
    for (int k = 0; k < n; k++) {
      pfsf_solve_info& f = per_island[k];
      f.iterations = sc_host_[k].iters;
      f.refinements = passes[k];
      f.converged = rel[k] <= tol_ ? 1 : 0;
      f.levels = sol_[k]->mg_->levels();
      f.rel_residual = rel[k];
      f.solve_ms = ms;
    }
  return err;
}

}  // namespace pfsf
