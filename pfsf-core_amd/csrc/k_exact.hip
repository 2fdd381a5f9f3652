// Bit-exact kernels of the structural solve: every product and sum rounds
// separately (this file is built with -ffp-contract=off), coefficients are
// formed exactly as the Java CPU reference forms them, and neighbour sums run
// in PFSFStencil offset order, so results match the CPU restatement bit for
// bit on identical inputs.
//
//   Jacobi step / RBGS    VoxelPhysicsCpuReference.java:66-118,
//                         SH/rbgs_smooth.comp.glsl:86-87 (colour order)
//   failure scan          SH/failure_scan.comp.glsl:36-176 (code precedence)
//   history update        phase-field H = max(H, psi_e), psi_e = 0.5 k |grad phi|^2
//   compaction            SH/failure_compact.comp.glsl:27-45,
//                         NP/src/core/island_buffer.cpp:930-962 (tuple format)
//   downsample            NP/src/compute/downsample.cpp:39-98
//   normalisation         NP/src/compute/normalize.cpp:29-63
#pragma clang fp contract(off)

#include <algorithm>
#include <mutex>
#include <vector>

#include "pfsf/pfsf_diagnostics.h"
#include "pfsf_internal.h"

namespace pfsf {
namespace {

constexpr int BS = 256;

__device__ __forceinline__ float kapf(float k) { return k > 0.0f ? k : 0.0f; }

// Coupling of offset class k (0..5 face, 6..17 edge, 18..25 corner),
// Java expression order: float sum / product, double scale, IEEE sqrt.
__device__ __forceinline__ double couple_exact(int k, float ki, float kj) {
  if (k < 6) {
    float s = ki + kj;
    return 0.5 * (double)s;
  }
  float p = ki * kj;
  double r = sqrt((double)p);
  return r * (k < 18 ? (double)kEdgeP : (double)kCornerP);
}

__device__ __forceinline__ bool is_solid(uint8_t t) { return t != 0 && t != 2; }

struct XYZ {
  int x, y, z;
};

__device__ __forceinline__ XYZ unflat(int64_t i, pfsf_dims d) {
  XYZ p;
  p.x = (int)(i % d.lx);
  int64_t r = i / d.lx;
  p.y = (int)(r % d.ly);
  p.z = (int)(r / d.ly);
  return p;
}

__device__ __forceinline__ int64_t flat(int x, int y, int z, pfsf_dims d) {
  return (int64_t)x + (int64_t)d.lx * ((int64_t)y + (int64_t)d.ly * z);
}

// The jacobiStep per-voxel update on a neighbourhood read from `phi`.  All 26
// neighbours are loaded first (a neighbour outside the grid reads the voxel
// itself and is dropped), so a thread waits for memory once; a dropped or air
// neighbour leaves both sums as they are (a select, the same bits as skipping
// it), and the FP64 operations run in offset order as in the restatement.
__device__ __forceinline__ float jacobi_voxel(int64_t i, XYZ p, pfsf_dims d,
                                              const uint8_t* __restrict__ type,
                                              const float* __restrict__ kappa,
                                              const float* __restrict__ src, const float* phi,
                                              float omega) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  float kn[26], pn[26];
  bool in[26];
  uint8_t tn[26];
#pragma unroll
  for (int k = 0; k < 26; k++) {
    int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
    in[k] = nx >= 0 && nx < d.lx && ny >= 0 && ny < d.ly && nz >= 0 && nz < d.lz;
    int64_t j = in[k] ? flat(nx, ny, nz, d) : i;
    tn[k] = type[j];
    kn[k] = kappa[j];
    pn[k] = phi[j];
  }
  double diag = 0.0, off = 0.0;
  float ki = kapf(kappa[i]);
#pragma unroll
  for (int k = 0; k < 26; k++) {
    const bool use = in[k] && tn[k] != 0;
    double a = couple_exact(k, ki, kapf(kn[k]));
    diag = use ? diag + a : diag;
    off = use ? off + a * (double)pn[k] : off;
  }
  double prev = (double)phi[i];
  double b = (double)src[i];
  double pj = (diag > 1e-15) ? (b + off) / diag : prev;
  return (float)((double)omega * (pj - prev) + prev);
}

__global__ void __launch_bounds__(BS)
    k_jacobi_exact(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                   const float* __restrict__ src, const float* __restrict__ prev,
                   float* __restrict__ out, float omega) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= nvox(d)) return;
  uint8_t t = type[i];
  if (!is_solid(t)) {
    out[i] = 0.0f;
    return;
  }
  out[i] = jacobi_voxel(i, unflat(i, d), d, type, kappa, src, prev, omega);
}

// One colour pass of the in-place 8-colour Gauss-Seidel.  Threads cover only
// the voxels of colour c: (x&1, y&1, z&1) == (c&1, c>>1&1, c>>2&1).  No voxel
// of colour c neighbours another of colour c, so the pass is race-free and its
// result does not depend on the order within the colour.
__global__ void __launch_bounds__(BS)
    k_rbgs_color(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                 const float* __restrict__ src, float* phi, int c) {
  int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  int hx = (d.lx - cx + 1) >> 1, hy = (d.ly - cy + 1) >> 1, hz = (d.lz - cz + 1) >> 1;
  int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (hx <= 0 || hy <= 0 || hz <= 0 || t >= (int64_t)hx * hy * hz) return;
  XYZ p;
  p.x = 2 * (int)(t % hx) + cx;
  int64_t r = t / hx;
  p.y = 2 * (int)(r % hy) + cy;
  p.z = 2 * (int)(r / hy) + cz;
  int64_t i = flat(p.x, p.y, p.z, d);
  if (!is_solid(type[i])) {
    phi[i] = 0.0f;
    return;
  }
  phi[i] = jacobi_voxel(i, p, d, type, kappa, src, phi, 1.0f);
}

// ── RBGS in one launch ──────────────────────────────────────────────────
// Both single-launch forms let a work-group own whole z-planes for the run.
// Colours 0-3 lie on even planes and read only odd planes across z (4-7 the
// reverse), so a sweep is an even-plane phase then an odd-plane phase, and a
// work-group only ever waits for its two slab neighbours.  Every wait is
// bounded: one that runs out raises *abort, every work-group then stops and
// the host reports a time-out.
constexpr unsigned kRbgsSpinLimit = 1u << 25;

// One more round of a bounded wait: false once *abort is raised (by this
// wait running out or by another work-group).
__device__ __forceinline__ bool rbgs_spin(unsigned* spins, unsigned* abort) {
  if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
  if (++*spins > kRbgsSpinLimit) {
    __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
  }
  __builtin_amdgcn_s_sleep(1);
  return true;
}

// Plane form, for planes of at most 32 x 32 voxels: work-group z owns plane
// z.  Its LDS holds planes z-1, z, z+1 with a one-voxel border; the border,
// missing planes and air read type 0, kappa 0 and phi 0.  The couplings are
// formed once (couple_exact, so the same bits as forming them per sweep):
// faces as their float sum kappa_i + kappa_j, edges in FP64 in LDS, corners
// and the ordered diagonal in registers.  A coupling to a dropped neighbour
// (outside the grid or air) is stored as 0 and that neighbour's phi reads 0,
// so the sweep adds 0 * 0 = +0 where the restatement skips the term, which
// leaves the sum unchanged (it starts at +0 and a round-to-nearest sum is
// never -0 unless both addends are).  A sweep then reads only the 26
// neighbour phi of each voxel.
//
// Planes pass between work-groups as 8-byte words {phi bits, tag} stored and
// loaded with agent-scope atomics, so a word's data and its readiness arrive
// together and no fence is needed.  Tag 1 is the starting plane and sweep s
// publishes tag s + 2; at sweep s an even plane waits for its neighbours'
// tag s + 1 and an odd plane for tag s + 2.  A writer never overwrites a word
// still to be read: it publishes sweep s + 1 only after reading what its
// reader published once it had read sweep s.
constexpr int kPlaneSide = 32;
constexpr int kPlaneVox = kPlaneSide * kPlaneSide;
constexpr int kPlanePad = (kPlaneSide + 2) * (kPlaneSide + 2);
constexpr int kPlaneWords = 2 * kPlaneVox / BS;  // neighbour words per thread
static_assert(kPlaneVox / 4 <= BS, "one voxel of each colour per thread");

__global__ void __launch_bounds__(BS)
    k_rbgs_planes(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                  const float* __restrict__ src, float* __restrict__ phi, int sweeps,
                  uint64_t* words, unsigned* abort) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  __shared__ double s_edge[12 * kPlaneVox];
  __shared__ float s_face[6 * kPlaneVox];
  __shared__ float s_kap[3 * kPlanePad];
  __shared__ float s_phi[3 * kPlanePad];
  __shared__ float s_src[kPlaneVox];
  __shared__ uint8_t s_typ[3 * kPlanePad];
  __shared__ int go;
  const int z = (int)blockIdx.x;
  const int P = d.lx * d.ly, W = d.lx + 2, PP = W * (d.ly + 2);
  for (int t = threadIdx.x; t < 3 * PP; t += BS) {
    const int pl = t / PP, o = t - pl * PP, zz = z + pl - 1;
    const int x = o % W - 1, y = o / W - 1;
    const bool in = zz >= 0 && zz < d.lz && x >= 0 && x < d.lx && y >= 0 && y < d.ly;
    const int64_t gi = in ? (int64_t)x + (int64_t)d.lx * y + (int64_t)P * zz : 0;
    const uint8_t ty = in ? type[gi] : (uint8_t)0;
    s_typ[pl * kPlanePad + o] = ty;
    s_kap[pl * kPlanePad + o] = ty != 0 ? kapf(kappa[gi]) : 0.0f;
    s_phi[pl * kPlanePad + o] = ty != 0 ? phi[gi] : 0.0f;
    if (pl == 1 && in) s_src[x + d.lx * y] = src[gi];
  }
  if (threadIdx.x == 0) go = 1;
  __syncthreads();
  // thread t owns in-plane voxel t of each of the plane's four colours
  double cor[4][8], dgv[4], bv[4];
  bool has[4], solid[4];
  int own[4], lds[4];
#pragma unroll
  for (int cc = 0; cc < 4; cc++) {
    const int cx = cc & 1, cy = (cc >> 1) & 1;
    const int hx = (d.lx - cx + 1) >> 1, hy = (d.ly - cy + 1) >> 1;
    const int r = threadIdx.x;
    has[cc] = r < hx * hy;
    const int x = has[cc] ? 2 * (r % hx) + cx : 0, y = has[cc] ? 2 * (r / hx) + cy : 0;
    const int o = x + d.lx * y, lo = kPlanePad + (x + 1) + W * (y + 1);
    own[cc] = o;
    lds[cc] = lo;
    const float ki = s_kap[lo];
    double dg = 0.0;
#pragma unroll
    for (int k = 0; k < 26; k++) {
      const int li = lo + O[k][2] * kPlanePad + O[k][0] + W * O[k][1];
      const bool use = s_typ[li] != 0;
      const float kj = s_kap[li];
      const double a = couple_exact(k, ki, kj);
      dg = use ? dg + a : dg;
      if (k < 6) {
        if (has[cc]) s_face[k * kPlaneVox + o] = use ? ki + kj : 0.0f;
      } else if (k < 18) {
        if (has[cc]) s_edge[(k - 6) * kPlaneVox + o] = use ? a : 0.0;
      } else {
        cor[cc][k - 18] = use ? a : 0.0;
      }
    }
    dgv[cc] = dg;
    bv[cc] = has[cc] ? (double)s_src[o] : 0.0;
    solid[cc] = is_solid(s_typ[lo]);
  }
  uint64_t* mine = words + (int64_t)P * z;
  for (int o = threadIdx.x; o < P; o += BS) {
    const float v = s_phi[kPlanePad + (o % d.lx + 1) + W * (o / d.lx + 1)];
    __hip_atomic_store(mine + o, (uint64_t)__float_as_uint(v) | (uint64_t(1) << 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  for (int s = 0; s < sweeps; s++) {
    const uint32_t want = (uint32_t)(s + 1 + (z & 1));
    uint64_t w[kPlaneWords];
    uint32_t pending = 0;
#pragma unroll
    for (int q = 0; q < kPlaneWords; q++) {
      const int t = threadIdx.x + q * BS, h = t >= P, o = t - h * P, zz = z + (h ? 1 : -1);
      const bool need = t < 2 * P && zz >= 0 && zz < d.lz;
      w[q] = need ? __hip_atomic_load(words + (int64_t)P * zz + o, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT)
                  : (uint64_t)want << 32;
      pending |= (uint32_t)(w[q] >> 32) != want ? 1u << q : 0u;
    }
    unsigned spins = 0;
    while (pending) {
      if (!rbgs_spin(&spins, abort)) {
        go = 0;
        break;
      }
#pragma unroll
      for (int q = 0; q < kPlaneWords; q++) {
        if (!((pending >> q) & 1u)) continue;
        const int t = threadIdx.x + q * BS, h = t >= P, o = t - h * P, zz = z + (h ? 1 : -1);
        w[q] = __hip_atomic_load(words + (int64_t)P * zz + o, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
        if ((uint32_t)(w[q] >> 32) == want) pending &= ~(1u << q);
      }
    }
    __syncthreads();
    if (!go) return;
#pragma unroll
    for (int q = 0; q < kPlaneWords; q++) {
      const int t = threadIdx.x + q * BS, h = t >= P, o = t - h * P, zz = z + (h ? 1 : -1);
      if (t < 2 * P && zz >= 0 && zz < d.lz) {
        const int li = (h ? 2 : 0) * kPlanePad + (o % d.lx + 1) + W * (o / d.lx + 1);
        if (s_typ[li] != 0) s_phi[li] = __uint_as_float((uint32_t)w[q]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
      if (has[cc]) {
        const int o = own[cc], lo = lds[cc];
        double off = 0.0;
#pragma unroll
        for (int k = 0; k < 26; k++) {
          const float pn = s_phi[lo + O[k][2] * kPlanePad + O[k][0] + W * O[k][1]];
          double a;
          if (k < 6)
            a = 0.5 * (double)s_face[k * kPlaneVox + o];
          else if (k < 18)
            a = s_edge[(k - 6) * kPlaneVox + o];
          else
            a = cor[cc][k - 18];
          off = off + a * (double)pn;
        }
        const double prev = (double)s_phi[lo];
        const double pj = (dgv[cc] > 1e-15) ? (bv[cc] + off) / dgv[cc] : prev;
        const float v = (float)(1.0 * (pj - prev) + prev);
        s_phi[lo] = solid[cc] ? v : 0.0f;
      }
      __syncthreads();
    }
    const uint64_t tag = (uint64_t)(uint32_t)(s + 2) << 32;
    for (int o = threadIdx.x; o < P; o += BS) {
      const float v = s_phi[kPlanePad + (o % d.lx + 1) + W * (o / d.lx + 1)];
      __hip_atomic_store(mine + o, (uint64_t)__float_as_uint(v) | tag, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      if (s + 1 == sweeps) phi[(int64_t)o + (int64_t)P * z] = v;
    }
  }
}

// Table form, for larger planes: the couplings live in a device table.
// Voxel (x, y, z) of colour c = (x&1) | (y&1)<<1 | (z&1)<<2 sits at
// colour_base(c) + (x>>1) + hx_c * ((y>>1) + hy_c * (z>>1)): a colour pass
// over a plane reads its table rows contiguously.
__host__ __device__ inline int64_t colour_base(pfsf_dims d, int c) {
  int64_t b = 0;
  for (int k = 0; k < c; k++) {
    const int cx = k & 1, cy = (k >> 1) & 1, cz = (k >> 2) & 1;
    b += (int64_t)((d.lx - cx + 1) >> 1) * ((d.ly - cy + 1) >> 1) * ((d.lz - cz + 1) >> 1);
  }
  return b;
}

// Per voxel, at its colour-major position: the 26 couplings (couple_exact in
// stencil order, 0 where the neighbour is outside the grid or air), the
// ordered diagonal, a mask (bit k: neighbour k used; bit 26: solid) and the
// source.
__global__ void __launch_bounds__(BS)
    k_rbgs_table(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                 const float* __restrict__ src, double* __restrict__ coef,
                 double* __restrict__ diag, uint32_t* __restrict__ mask,
                 float* __restrict__ srcc) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const XYZ p = unflat(i, d);
  const int c = (p.x & 1) | ((p.y & 1) << 1) | ((p.z & 1) << 2);
  const int hx = (d.lx - (p.x & 1) + 1) >> 1, hy = (d.ly - (p.y & 1) + 1) >> 1;
  const int64_t o = colour_base(d, c) + (p.x >> 1) + (int64_t)hx * ((p.y >> 1) + (int64_t)hy * (p.z >> 1));
  const float ki = kapf(kappa[i]);
  double dg = 0.0;
  uint32_t m = is_solid(type[i]) ? 1u << 26 : 0u;
#pragma unroll
  for (int k = 0; k < 26; k++) {
    const int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
    const bool in = nx >= 0 && nx < d.lx && ny >= 0 && ny < d.ly && nz >= 0 && nz < d.lz;
    const int64_t j = in ? flat(nx, ny, nz, d) : i;
    const bool use = in && type[j] != 0;
    const double a = couple_exact(k, ki, kapf(kappa[j]));
    dg = use ? dg + a : dg;
    m |= use ? 1u << k : 0u;
    coef[k * n + o] = use ? a : 0.0;
  }
  diag[o] = dg;
  mask[o] = m;
  srcc[o] = src[i];
}

// Table form: work-group g owns planes [g * ppw, (g + 1) * ppw).  Before
// phase p (p = 2 * sweep + parity) it waits until both slab neighbours have
// published p (so it reads their newest planes and they no longer read the
// planes it overwrites); after it, it publishes p + 1.  The neighbour phi
// addresses come from the coordinates alone, so all loads of a voxel issue
// at once.
__global__ void __launch_bounds__(BS)
    k_rbgs_slabs(pfsf_dims d, const double* __restrict__ coef, const double* __restrict__ diag,
                 const uint32_t* __restrict__ mask, const float* __restrict__ srcc, float* phi,
                 int sweeps, int ppw, unsigned* flags, unsigned* abort) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  __shared__ int go;
  const int g = (int)blockIdx.x, G = (int)gridDim.x;
  const int z0 = g * ppw, z1 = min(d.lz, z0 + ppw);
  const int64_t n = nvox(d);
  for (int s = 0; s < sweeps; s++) {
    for (int half = 0; half < 2; half++) {
      const unsigned p = 2u * (unsigned)s + (unsigned)half;
      if (threadIdx.x == 0) {
        bool ok = true;
        unsigned spins = 0;
        while (ok && g > 0 &&
               __hip_atomic_load(flags + g - 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < p)
          ok = rbgs_spin(&spins, abort);
        while (ok && g + 1 < G &&
               __hip_atomic_load(flags + g + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < p)
          ok = rbgs_spin(&spins, abort);
        go = ok;
      }
      __syncthreads();
      if (!go) return;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      const int zs = z0 + ((z0 & 1) != half ? 1 : 0);  // first plane of this parity
      const int nzp = zs < z1 ? (z1 - zs + 1) / 2 : 0;
      for (int cc = 0; cc < 4; cc++) {
        const int cx = cc & 1, cy = (cc >> 1) & 1;
        const int hx = (d.lx - cx + 1) >> 1, hy = (d.ly - cy + 1) >> 1;
        const int per = hx * hy;
        const int64_t cb = colour_base(d, cc + 4 * half);
        for (int t = threadIdx.x; t < per * nzp; t += BS) {
          const int pl = t / per, r = t - pl * per;
          const int x = 2 * (r % hx) + cx, y = 2 * (r / hx) + cy, z = zs + 2 * pl;
          const int64_t i = flat(x, y, z, d);
          const int64_t o = cb + r + (int64_t)per * (z >> 1);
          const uint32_t m = mask[o];
          double a[26];
          float pk[26];
#pragma unroll
          for (int k = 0; k < 26; k++) {
            const int nx = x + O[k][0], ny = y + O[k][1], nz = z + O[k][2];
            const bool in = nx >= 0 && nx < d.lx && ny >= 0 && ny < d.ly && nz >= 0 && nz < d.lz;
            a[k] = coef[k * n + o];
            pk[k] = phi[in ? flat(nx, ny, nz, d) : i];
          }
          const double dg = diag[o];
          const double prev = (double)phi[i];
          const double b = (double)srcc[o];
          double off = 0.0;
#pragma unroll
          for (int k = 0; k < 26; k++) off = (m >> k) & 1u ? off + a[k] * (double)pk[k] : off;
          const double pj = (dg > 1e-15) ? (b + off) / dg : prev;
          const float v = (float)(1.0 * (pj - prev) + prev);
          phi[i] = (m >> 26) & 1u ? v : 0.0f;
        }
        __syncthreads();
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_store(flags + g, p + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Macro-block residual map (SH/failure_scan.comp.glsl:183-195, 8^3 blocks):
// the FP64 residual b - A phi of k_apply_f64's operator at each solid voxel
// (phi widened from FP32), |r| rounded to FP32 and maximised per block as
// IEEE bits (non-negative floats order as unsigned integers).
__global__ void __launch_bounds__(BS)
    k_macro_residual(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                     const float* __restrict__ phi, const float* __restrict__ b,
                     unsigned* __restrict__ blk) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= nvox(d) || !is_solid(type[i])) return;
  const XYZ p = unflat(i, d);
  const float ki = kapf(kappa[i]);
  const double xi = (double)phi[i];
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 26; k++) {
    const int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
    if (nx < 0 || nx >= d.lx || ny < 0 || ny >= d.ly || nz < 0 || nz >= d.lz) continue;
    const int64_t j = flat(nx, ny, nz, d);
    const uint8_t tj = type[j];
    if (tj == 0) continue;
    const double xj = is_solid(tj) ? (double)phi[j] : 0.0;
    acc += couple_exact(k, ki, kapf(kappa[j])) * (xi - xj);
  }
  const float r = (float)fabs((double)b[i] - acc);
  const int bx = (d.lx + 7) / 8, by = (d.ly + 7) / 8;
  atomicMax(blk + (p.x / 8) + (int64_t)bx * ((p.y / 8) + (int64_t)by * (p.z / 8)),
            __float_as_uint(r));
}

__global__ void k_zero_u32(unsigned* a, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) a[i] = 0u;
}

// Failure scan over 32 x 8 x 4 tiles of the whole grid, one 256-thread
// work-group each.  The tile and its one-voxel halo are staged in LDS (the
// coupling kappa, -1 for air and cells outside the grid, and phi), then each
// thread classifies its column of four voxels with the neighbour loop of the
// per-voxel form: offsets in PFSFStencil order, the same FP64 operations in
// the same order, so the codes are bit-identical to the restatement; only the
// 26 neighbour loads come from LDS instead of global memory.
constexpr int kFsTX = 32, kFsTY = 8, kFsTZ = 4;
constexpr int kFsHX = kFsTX + 2, kFsHY = kFsTY + 2, kFsHZ = kFsTZ + 2;
constexpr int kFsHN = kFsHX * kFsHY * kFsHZ;
static_assert(kFsTX * kFsTY == BS, "one thread per column of a tile");

__global__ void __launch_bounds__(BS)
    k_failure_scan(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                   const float* __restrict__ phi, const float* __restrict__ rcomp,
                   const float* __restrict__ rtens, float phi_orphan, uint8_t* __restrict__ code) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  __shared__ float K[kFsHN], P[kFsHN];
  const int ntx = (d.lx + kFsTX - 1) / kFsTX, nty = (d.ly + kFsTY - 1) / kFsTY;
  const int bt = (int)blockIdx.x;
  const int x0 = (bt % ntx) * kFsTX, y0 = ((bt / ntx) % nty) * kFsTY,
            z0 = (bt / (ntx * nty)) * kFsTZ;
  for (int e = threadIdx.x; e < kFsHN; e += BS) {
    const int ex = e % kFsHX, ey = (e / kFsHX) % kFsHY, ez = e / (kFsHX * kFsHY);
    const int x = x0 + ex - 1, y = y0 + ey - 1, z = z0 + ez - 1;
    const bool in = x >= 0 && x < d.lx && y >= 0 && y < d.ly && z >= 0 && z < d.lz;
    const int64_t j = in ? flat(x, y, z, d) : 0;
    const uint8_t tj = type[j];
    const float kj = kappa[j];
    const float pj = phi[j];
    K[e] = (in && tj != 0) ? kapf(kj) : -1.0f;
    P[e] = in ? pj : 0.0f;
  }
  __syncthreads();
  const int tx = threadIdx.x % kFsTX, ty = threadIdx.x / kFsTX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= d.lx || y >= d.ly) return;
  for (int tz = 0; tz < kFsTZ; tz++) {
    const int z = z0 + tz;
    if (z >= d.lz) break;
    const int64_t i = flat(x, y, z, d);
    if (!is_solid(type[i])) {
      code[i] = 0;
      continue;
    }
    const int c = (tx + 1) + kFsHX * ((ty + 1) + kFsHY * (tz + 1));
    const float ki = K[c];
    const float phii = P[c];
    const double pi = (double)phii;
    double diag = 0.0, fin = 0.0, fout = 0.0;
#pragma unroll
    for (int k = 0; k < 26; k++) {
      const int o = O[k][0] + kFsHX * (O[k][1] + kFsHY * O[k][2]);
      const float kj = K[c + o];
      if (kj < 0.0f) continue;
      double a = couple_exact(k, ki, kj);
      diag += a;
      double dp = (double)P[c + o] - pi;
      if (dp > 0.0)
        fin += a * dp;
      else if (dp < 0.0)
        fout += a * (-dp);
    }
    uint8_t cd = 0;
    if (diag == 0.0 || phii > phi_orphan)
      cd = 3;
    else if (rcomp[i] > 0.0f && fin > (double)rcomp[i])
      cd = 2;
    else if (rtens[i] > 0.0f && fout > (double)rtens[i])
      cd = 4;
    code[i] = cd;
  }
}

__device__ __forceinline__ float axis_grad(const uint8_t* __restrict__ type,
                                           const float* __restrict__ phi, int64_t i, bool hm,
                                           bool hp, int64_t jm, int64_t jp) {
  bool om = hm && type[jm] != 0;
  bool op = hp && type[jp] != 0;
  if (om && op) return (phi[jp] - phi[jm]) * 0.5f;
  if (op) return phi[jp] - phi[i];
  if (om) return phi[i] - phi[jm];
  return 0.0f;
}

__global__ void __launch_bounds__(BS)
    k_history(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
              const float* __restrict__ phi, float* __restrict__ H, float* __restrict__ psi_out) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= nvox(d)) return;
  if (!is_solid(type[i])) {
    if (psi_out) psi_out[i] = 0.0f;
    return;
  }
  XYZ p = unflat(i, d);
  int64_t sy = d.lx, sz = (int64_t)d.lx * d.ly;
  float gx = axis_grad(type, phi, i, p.x > 0, p.x + 1 < d.lx, i - 1, i + 1);
  float gy = axis_grad(type, phi, i, p.y > 0, p.y + 1 < d.ly, i - sy, i + sy);
  float gz = axis_grad(type, phi, i, p.z > 0, p.z + 1 < d.lz, i - sz, i + sz);
  float g2 = gx * gx;
  g2 = g2 + gy * gy;
  g2 = g2 + gz * gz;
  float h = 0.5f * kapf(kappa[i]);
  float psi = h * g2;
  if (psi_out) psi_out[i] = psi;
  if (H && psi > H[i]) H[i] = psi;
}

__global__ void __launch_bounds__(BS)
    k_apply_f64(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                const double* __restrict__ x, double* __restrict__ y) {
  constexpr int O[26][3] = PFSF_OFFSETS;
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= nvox(d)) return;
  if (!is_solid(type[i])) {
    y[i] = 0.0;
    return;
  }
  XYZ p = unflat(i, d);
  float ki = kapf(kappa[i]);
  double xi = x[i], acc = 0.0;
#pragma unroll
  for (int k = 0; k < 26; k++) {
    int nx = p.x + O[k][0], ny = p.y + O[k][1], nz = p.z + O[k][2];
    if (nx < 0 || nx >= d.lx || ny < 0 || ny >= d.ly || nz < 0 || nz >= d.lz) continue;
    int64_t j = flat(nx, ny, nz, d);
    uint8_t tj = type[j];
    if (tj == 0) continue;
    double xj = is_solid(tj) ? x[j] : 0.0;
    acc += couple_exact(k, ki, kapf(kappa[j])) * (xi - xj);
  }
  y[i] = acc;
}

// ── Ordered compaction of failure codes into the tuple wire format ───────
constexpr int CB = 256, CPT = 4, CSEG = CB * CPT;  // voxels per segment

__device__ __forceinline__ int block_excl_scan(int v, int* sh, int* total) {
  // Inclusive Hillis-Steele scan over CB threads in LDS.
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < CB; o <<= 1) {
    int a = (t >= o) ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  int incl = sh[t];
  *total = sh[CB - 1];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(CB)
    k_compact_count(int64_t n, const uint8_t* __restrict__ code, int* __restrict__ counts) {
  __shared__ int sh[CB];
  int64_t base = (int64_t)blockIdx.x * CSEG + (int64_t)threadIdx.x * CPT;
  int c = 0;
  for (int e = 0; e < CPT; e++)
    if (base + e < n && code[base + e] != 0) c++;
  int tot;
  block_excl_scan(c, sh, &tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(CB)
    k_compact_scan(int nb, int* __restrict__ counts, const int32_t* __restrict__ append_hdr) {
  // Exclusive scan of nb block counts in place; counts[nb] = total,
  // counts[nb + 1] = the tuples already in the output (append mode) or 0.
  __shared__ int sh[CB];
  int per = (nb + CB - 1) / CB;
  int b0 = threadIdx.x * per;
  int s = 0;
  for (int e = 0; e < per; e++)
    if (b0 + e < nb) s += counts[b0 + e];
  int tot;
  int ex = block_excl_scan(s, sh, &tot);
  for (int e = 0; e < per; e++)
    if (b0 + e < nb) {
      int v = counts[b0 + e];
      counts[b0 + e] = ex;
      ex += v;
    }
  if (threadIdx.x == 0) {
    counts[nb] = tot;
    counts[nb + 1] = append_hdr ? append_hdr[0] : 0;
  }
}

__global__ void __launch_bounds__(CB)
    k_compact_write(pfsf_dims d, const uint8_t* __restrict__ code, const int* __restrict__ offs,
                    int nb, pfsf_pos origin, int32_t* __restrict__ out, int32_t capacity) {
  __shared__ int sh[CB];
  int64_t n = nvox(d);
  int64_t base = (int64_t)blockIdx.x * CSEG + (int64_t)threadIdx.x * CPT;
  int c = 0;
  for (int e = 0; e < CPT; e++)
    if (base + e < n && code[base + e] != 0) c++;
  int tot;
  const int prior = offs[nb + 1];
  int rank = prior + offs[blockIdx.x] + block_excl_scan(c, sh, &tot);
  for (int e = 0; e < CPT; e++) {
    int64_t i = base + e;
    if (i < n && code[i] != 0) {
      if (rank < capacity) {
        XYZ p = unflat(i, d);
        int32_t* ev = out + 1 + 4 * (int64_t)rank;
        ev[0] = origin.x + p.x;
        ev[1] = origin.y + p.y;
        ev[2] = origin.z + p.z;
        ev[3] = (int32_t)code[i];
      }
      rank++;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    int total = prior + offs[nb];
    out[0] = total < capacity ? total : capacity;
  }
}

__global__ void __launch_bounds__(BS)
    k_downsample(pfsf_dims f, const float* __restrict__ fine, const uint8_t* __restrict__ ft,
                 float* __restrict__ coarse, uint8_t* __restrict__ ct) {
  pfsf_dims c{(f.lx + 1) / 2, (f.ly + 1) / 2, (f.lz + 1) / 2};
  int64_t ci = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (ci >= nvox(c)) return;
  XYZ p = unflat(ci, c);
  float sum = 0.0f;
  int cnt = 0, na = 0, ns = 0, nk = 0;
  for (int dz = 0; dz < 2; dz++) {
    int zf = 2 * p.z + dz;
    if (zf >= f.lz) continue;
    for (int dy = 0; dy < 2; dy++) {
      int yf = 2 * p.y + dy;
      if (yf >= f.ly) continue;
      for (int dx = 0; dx < 2; dx++) {
        int xf = 2 * p.x + dx;
        if (xf >= f.lx) continue;
        int64_t fi = flat(xf, yf, zf, f);
        sum += fine[fi];
        cnt++;
        uint8_t t = ft[fi];
        if (t == 2) nk++;
        else if (t == 1) ns++;
        else na++;
      }
    }
  }
  coarse[ci] = cnt > 0 ? sum / (float)cnt : 0.0f;
  if (ct) ct[ci] = (nk >= ns && nk >= na) ? 2 : (ns >= na ? 1 : 0);
}

// ── island data prep (PFSFDataBuilder.java:77-133) ─────────────────────
// Java's Math.min / Math.max on doubles: NaN if either operand is NaN.
__device__ __forceinline__ double jmin(double a, double b) {
  return (a != a || b != b) ? (double)NAN : (a < b ? a : b);
}
__device__ __forceinline__ double jmax(double a, double b) {
  return (a != a || b != b) ? (double)NAN : (a > b ? a : b);
}
__device__ __forceinline__ float jminf(float a, float b) {
  return (a != a || b != b) ? NAN : (a < b ? a : b);
}
__device__ __forceinline__ float jmaxf(float a, float b) {
  return (a != a || b != b) ? NAN : (a > b ? a : b);
}

// Face conductivity between the materials of two member voxels
// (PFSFConductivity.sigma, PFSFConductivity.java:36-97, force arm 0, no wind
// bias): the weaker compressive strength for a vertical (y) face; for a
// horizontal face that base times min(1, mean tensile strength /
// max(base, 1)), 0 when not finite.  An indestructible material counts 1e6
// for both strengths.  Strengths are compared in FP64, the base and the ratio
// rounded to FP32 and multiplied in FP32, as in the Java.
__device__ __forceinline__ float face_sigma(const pfsf_material& mi, bool indi,
                                            const pfsf_material& mj, bool indj, bool vertical) {
  const double rci = indi ? 1e6 : (double)mi.rcomp;
  const double rcj = indj ? 1e6 : (double)mj.rcomp;
  const float base = (float)jmin(rci, rcj);
  if (base <= 0.0f) return 0.0f;
  if (vertical) return base;
  const double rti = indi ? 1e6 : (double)mi.rtens;
  const double rtj = indj ? 1e6 : (double)mj.rtens;
  const double avg = (rti + rtj) / 2.0;
  const float ratio = (float)jmin(1.0, avg / jmax((double)base, 1.0));
  const float r = base * ratio;
  return (isnan(r) || isinf(r)) ? 0.0f : r;
}

// One voxel of the island arrays from a material index grid (index outside
// [0, nmat) = air): type (anchor from the anchor flags), self-weight source
// density * fill * g * 1 m^3 (FP64, rounded to FP32), rcomp, rtens, max_phi
// (+inf) and the six face conductivities sigma * sqrt(curing clamped to
// [0, 1]) in SoA order dir * n + i, directions -x, +x, -y, +y, -z, +z
// (PFSFConductivity.dirToIndex, y vertical).  Air voxels get zeros.
__global__ void __launch_bounds__(BS)
    k_build_island(pfsf_dims d, const int32_t* __restrict__ mat_idx,
                   const pfsf_material* __restrict__ mats, int32_t nmat,
                   const uint8_t* __restrict__ indestructible, const uint8_t* __restrict__ anchor,
                   const float* __restrict__ curing, const float* __restrict__ fill,
                   uint8_t* __restrict__ type, float* __restrict__ source,
                   float* __restrict__ cond, float* __restrict__ rcomp,
                   float* __restrict__ rtens, float* __restrict__ max_phi) {
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const int32_t m = mat_idx[i];
  if (m < 0 || m >= nmat) {
    type[i] = 0;
    source[i] = 0.0f;
    rcomp[i] = 0.0f;
    rtens[i] = 0.0f;
    max_phi[i] = 0.0f;
    for (int dir = 0; dir < 6; dir++) cond[dir * n + i] = 0.0f;
    return;
  }
  const pfsf_material mi = mats[m];
  const bool indi = indestructible && indestructible[m];
  const float hdeg = curing ? jmaxf(0.0f, jminf(1.0f, curing[i])) : 1.0f;
  const float scale = (float)sqrt((double)hdeg);
  const float fr = fill ? fill[i] : 1.0f;
  source[i] = (float)((double)mi.density * (double)fr * 9.81 * 1.0);
  type[i] = (anchor && anchor[i]) ? 2 : 1;
  max_phi[i] = INFINITY;
  rcomp[i] = mi.rcomp;
  rtens[i] = mi.rtens;
  const XYZ p = unflat(i, d);
  constexpr int DV[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
#pragma unroll
  for (int dir = 0; dir < 6; dir++) {
    const int nx = p.x + DV[dir][0], ny = p.y + DV[dir][1], nz = p.z + DV[dir][2];
    float sg = 0.0f;
    if (nx >= 0 && nx < d.lx && ny >= 0 && ny < d.ly && nz >= 0 && nz < d.lz) {
      const int32_t mj = mat_idx[flat(nx, ny, nz, d)];
      if (mj >= 0 && mj < nmat)
        sg = face_sigma(mi, indi, mats[mj], indestructible && indestructible[mj], dir == 2 || dir == 3);
    }
    cond[dir * n + i] = sg * scale;
  }
}

__global__ void __launch_bounds__(BS)
    k_kappa_soa6(const float* __restrict__ cond, int64_t n, float* __restrict__ kappa) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  // non-finite faces count as 0, so the maximum is always finite and >= 0
  float m = 0.0f;
  for (int dd = 0; dd < 6; dd++) {
    const float v = cond[dd * n + i];
    m = (v > m && v <= 3.40282347e38f) ? v : m;
  }
  kappa[i] = m;
}

__global__ void __launch_bounds__(BS)
    k_max_bits(const float* __restrict__ a, int64_t n, unsigned* __restrict__ mx) {
  // max over positive finite values via their (monotone) IEEE bit patterns
  __shared__ unsigned sh[BS];
  unsigned m = 0;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    float v = a[i];
    if (v > 0.0f) {
      unsigned b = __float_as_uint(v);
      m = b > m ? b : m;
    }
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = BS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(mx, sh[0]);
}

__global__ void __launch_bounds__(BS)
    k_scale3(float* a, float* b, float* c, float* cond, int64_t n, const unsigned* mxb) {
  float smax = __uint_as_float(*mxb);
  if (!(smax > 1.0f)) return;
  float inv = 1.0f / smax;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < 6 * n;
       i += (int64_t)gridDim.x * BS) {
    if (i < n) {
      a[i] *= inv;
      b[i] *= inv;
      c[i] *= inv;
    }
    cond[i] *= inv;
  }
}

// Sparse voxel re-upload: 48-byte records {flat, source, type, max_phi, rcomp,
// rtens, cond[6]} scattered into the island's device arrays
// (SH/sparse_scatter.comp.glsl:21-67).  Out-of-range indices are skipped.
__global__ void __launch_bounds__(64)
    k_sparse_scatter(const uint32_t* __restrict__ rec, int count, int64_t n,
                     float* __restrict__ source, float* __restrict__ cond,
                     uint8_t* __restrict__ type, float* __restrict__ max_phi,
                     float* __restrict__ rcomp, float* __restrict__ rtens) {
  int u = blockIdx.x * 64 + threadIdx.x;
  if (u >= count) return;
  const uint32_t* r = rec + 12 * (int64_t)u;
  int64_t i = (int64_t)(int32_t)r[0];
  if (i < 0 || i >= n) return;
  source[i] = __uint_as_float(r[1]);
  type[i] = (uint8_t)r[2];
  if (max_phi) max_phi[i] = __uint_as_float(r[3]);
  rcomp[i] = __uint_as_float(r[4]);
  rtens[i] = __uint_as_float(r[5]);
  for (int dd = 0; dd < 6; dd++) cond[dd * n + i] = __uint_as_float(r[6 + dd]);
}

// Phase-field damage step (SH/phase_field_evolve.comp.glsl:50-168, hybrid
// AT2 form of Ambati et al. 2015): the FP32 expression order of the CPU
// restatement, reading d_in and writing d_out (a Jacobi step; the shader's
// in-place update makes neighbours race).  Voxels with a failure code from
// the previous scan are frozen; code becomes 1 (CANTILEVER) where d > 0.95.
__global__ void __launch_bounds__(BS)
    k_phase_field(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ cond,
                  const float* __restrict__ phi, const float* __restrict__ H,
                  const float* __restrict__ hydration, uint8_t* __restrict__ code,
                  const float* __restrict__ d_in, float* __restrict__ d_out, float l0, float gc,
                  float relax, int split) {
  constexpr int DV[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  const int64_t n = nvox(d);
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const float di = d_in[i];
  float dn = di;
  const uint8_t c0 = code[i];
  if (is_solid(type[i]) && c0 == 0) {
    const XYZ p = unflat(i, d);
    const float pi = phi[i];
    float ssum = 0.0f, lp = 0.0f, ld = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const int nx = p.x + DV[k][0], ny = p.y + DV[k][1], nz = p.z + DV[k][2];
      if (nx < 0 || nx >= d.lx || ny < 0 || ny >= d.ly || nz < 0 || nz >= d.lz) continue;
      const int64_t j = flat(nx, ny, nz, d);
      if (type[j] == 0) continue;
      const float pj = phi[j], dj = d_in[j];
      if (isfinite(pj)) {
        lp += pj - pi;
        ssum += cond[k * n + i];
      }
      if (isfinite(dj)) ld += dj - di;
    }
    const float hv = H[i];
    if (hv > 0.0f) {
      const float hy = hydration ? hydration[i] : 1.0f;
      const float hd = hy > 0.01f ? (hy < 1.0f ? hy : 1.0f) : 0.01f;
      const float gce = gc * (hd * (float)sqrt((double)hd));
      float he = hv;
      if (split) {
        const float sd = ssum > 1e-12f ? ssum : 1e-12f;
        const float fo = lp > 0.0f ? lp : 0.0f;
        const float pp = fo / sd;
        he = pp > hv ? pp : hv;
      }
      const float num = 2.0f * he + gce * l0 * ld;
      float den = 2.0f * he + gce / l0;
      den = den > 1e-8f ? den : 1e-8f;
      float q = num / den;
      q = q > di ? q : di;
      q = q < 1.0f ? q : 1.0f;
      float v = di * (1.0f - relax) + q * relax;
      v = v > 0.0f ? v : 0.0f;
      v = v < 1.0f ? v : 1.0f;
      v = v > di ? v : di;  // the mix rounds; damage never decreases
      dn = isfinite(v) ? v : di;
      if (dn > 0.95f) code[i] = 1;
    }
  }
  d_out[i] = dn;
}

// After the failure scan: a voxel the scan leaves at 0 whose damage exceeds
// the fracture threshold keeps CANTILEVER (the scan's codes take precedence).
__global__ void __launch_bounds__(BS)
    k_mark_fractured(int64_t n, const uint8_t* __restrict__ type, const float* __restrict__ dmg,
                     uint8_t* __restrict__ code) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n && is_solid(type[i]) && code[i] == 0 && dmg[i] > 0.95f) code[i] = 1;
}

}  // namespace

hipError_t launch_phase_field(pfsf_dims d, const uint8_t* type, const float* cond,
                              const float* phi, const float* H, const float* hydration,
                              uint8_t* code, const float* d_in, float* d_out, float l0, float gc,
                              float relax, int split, hipStream_t s) {
  if (nvox(d) == 0) return hipSuccess;
  k_phase_field<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, cond, phi, H, hydration, code,
                                                       d_in, d_out, l0, gc, relax, split);
  return hipGetLastError();
}

hipError_t launch_mark_fractured(int64_t n, const uint8_t* type, const float* dmg, uint8_t* code,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_mark_fractured<<<blocks_for(n, BS), BS, 0, s>>>(n, type, dmg, code);
  return hipGetLastError();
}

hipError_t launch_sparse_scatter(const uint32_t* rec, int count, int64_t n, float* source,
                                 float* cond, uint8_t* type, float* max_phi, float* rcomp,
                                 float* rtens, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  k_sparse_scatter<<<(count + 63) / 64, 64, 0, s>>>(rec, count, n, source, cond, type, max_phi,
                                                     rcomp, rtens);
  return hipGetLastError();
}

hipError_t launch_jacobi_exact(pfsf_dims d, const uint8_t* type, const float* kappa,
                               const float* src, const float* prev, float* out, float omega,
                               hipStream_t s) {
  k_jacobi_exact<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, kappa, src, prev, out, omega);
  return hipGetLastError();
}

// Chebyshev-accelerated Jacobi (SH/jacobi_smooth.comp.glsl:215-226 with the
// schedule of NP/src/compute/chebyshev.cpp:30-57): step k is the jacobiStep
// with omega_k = pfsf_chebyshev_omega(k, rho), alternating between phi and
// tmp; the result ends in phi.
hipError_t launch_chebyshev_jacobi(pfsf_dims d, const uint8_t* type, const float* kappa,
                                   const float* src, float* phi, float* tmp, int steps,
                                   float rho, hipStream_t s) {
  float* cur = phi;
  float* nxt = tmp;
  for (int k = 0; k < steps; k++) {
    const float w = pfsf_chebyshev_omega(k, rho);
    k_jacobi_exact<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, kappa, src, cur, nxt, w);
    float* t = cur;
    cur = nxt;
    nxt = t;
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && cur != phi)
    e = hipMemcpyAsync(phi, cur, sizeof(float) * (size_t)nvox(d), hipMemcpyDeviceToDevice, s);
  return e;
}

// One sweep: the eight colour passes in order.
static void rbgs_sweep(pfsf_dims d, const uint8_t* type, const float* kappa, const float* src,
                       float* phi, hipStream_t s) {
  for (int c = 0; c < 8; c++) {
    int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    int64_t hn = (int64_t)((d.lx - cx + 1) >> 1) * ((d.ly - cy + 1) >> 1) * ((d.lz - cz + 1) >> 1);
    if (hn <= 0) continue;
    k_rbgs_color<<<blocks_for(hn, BS), BS, 0, s>>>(d, type, kappa, src, phi, c);
  }
}

// Grids beyond both single-launch forms replay runs of kRbgsGraphSweeps
// sweeps from a HIP graph holding exactly those colour launches, captured
// once per device and argument set on a private stream: one graph launch
// replaces 64 host launches.  The kernels, their arguments and their
// order are those of the direct launches, so the iterates are the same bits.
// The last kRbgsGraphCache argument sets stay instantiated.
constexpr int kRbgsGraphSweeps = 8;
constexpr size_t kRbgsGraphCache = 8;

struct RbgsGraph {
  int dev;
  pfsf_dims d;
  const uint8_t* type;
  const float *kappa, *src;
  float* phi;
  hipGraphExec_t exec;
};

static hipError_t rbgs_graph(pfsf_dims d, const uint8_t* type, const float* kappa,
                             const float* src, float* phi, hipGraphExec_t* out) {
  static std::mutex mu;
  static std::vector<RbgsGraph> cache;  // least recently used first
  std::lock_guard<std::mutex> lock(mu);
  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return err;
  for (size_t k = 0; k < cache.size(); k++) {
    const RbgsGraph& e = cache[k];
    if (e.dev == dev && e.d.lx == d.lx && e.d.ly == d.ly && e.d.lz == d.lz && e.type == type &&
        e.kappa == kappa && e.src == src && e.phi == phi) {
      RbgsGraph hit = e;
      cache.erase(cache.begin() + k);
      cache.push_back(hit);
      *out = hit.exec;
      return hipSuccess;
    }
  }
  hipStream_t cs = nullptr;
  err = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (err != hipSuccess) return err;
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  err = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
  if (err == hipSuccess) {
    for (int k = 0; k < kRbgsGraphSweeps; k++) rbgs_sweep(d, type, kappa, src, phi, cs);
    hipError_t lerr = hipGetLastError();
    err = hipStreamEndCapture(cs, &g);
    if (err == hipSuccess) err = lerr;
  }
  if (err == hipSuccess) err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cs);
  if (err != hipSuccess) return err;
  if (cache.size() >= kRbgsGraphCache) {
    // the evicted graph may still be queued on some stream
    (void)hipDeviceSynchronize();
    (void)hipGraphExecDestroy(cache.front().exec);
    cache.erase(cache.begin());
  }
  cache.push_back(RbgsGraph{dev, d, type, kappa, src, phi, exec});
  *out = exec;
  return hipSuccess;
}

// Single-launch forms (stream not under capture).  The plane form takes
// planes of at most kPlaneSide x kPlaneSide and no more planes than it can
// keep resident (its LDS allows one work-group per CU), with an 8 B word per
// voxel as workspace; the table form takes up to kRbgsTableMax voxels (its
// table holds 224 B per voxel) in work-groups of one or more whole planes,
// at most one per CU.  Both launch cooperatively, so the launch fails rather
// than start work-groups that could not all be resident, and both
// synchronise the stream to read the abort word.
constexpr int64_t kRbgsTableMax = int64_t(1) << 21;

static hipError_t device_cus(int* cus) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
  return e;
}

static hipError_t finish_single_launch(hipError_t e, char* ws, const unsigned* abort,
                                       hipStream_t s) {
  unsigned ab_host = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&ab_host, abort, sizeof(unsigned), hipMemcpyDeviceToHost, s);
  (void)hipFreeAsync(ws, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess && ab_host) e = hipErrorLaunchTimeOut;
  return e;
}

static bool rbgs_planes_fit(pfsf_dims d, int cus) {
  if (d.lx > kPlaneSide || d.ly > kPlaneSide) return false;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rbgs_planes, BS, 0) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return (int64_t)d.lz <= (int64_t)per_cu * cus;
}

static hipError_t rbgs_planes(pfsf_dims d, const uint8_t* type, const float* kappa,
                              const float* src, float* phi, int sweeps, hipStream_t s) {
  const int64_t n = nvox(d);
  char* ws = nullptr;
  hipError_t e = hipMallocAsync((void**)&ws, sizeof(uint64_t) * (size_t)n + 256, s);
  if (e != hipSuccess) return e;
  uint64_t* words = (uint64_t*)ws;
  unsigned* abort = (unsigned*)(words + n);
  k_zero_u32<<<blocks_for(2 * n + 1, 256), 256, 0, s>>>((unsigned*)ws, (int)(2 * n + 1));
  e = hipGetLastError();
  if (e == hipSuccess) {
    const uint8_t* t = type;
    const float *k = kappa, *b = src;
    float* ph = phi;
    int sw = sweeps;
    uint64_t* wd = words;
    unsigned* ab = abort;
    void* args[] = {&d, &t, &k, &b, &ph, &sw, &wd, &ab};
    e = hipLaunchCooperativeKernel((const void*)k_rbgs_planes, dim3(d.lz), dim3(BS), args, 0, s);
  }
  return finish_single_launch(e, ws, abort, s);
}

static hipError_t rbgs_table(pfsf_dims d, const uint8_t* type, const float* kappa,
                             const float* src, float* phi, int sweeps, int cus, hipStream_t s) {
  const int ppw = (d.lz + std::max(cus, 1) - 1) / std::max(cus, 1);
  const int G = (d.lz + ppw - 1) / ppw;
  const int64_t n = nvox(d);
  const size_t bytes = sizeof(double) * 27 * (size_t)n + sizeof(uint32_t) * (size_t)n +
                       sizeof(float) * (size_t)n + sizeof(unsigned) * (size_t)(G + 1);
  char* ws = nullptr;
  hipError_t e = hipMallocAsync((void**)&ws, bytes, s);
  if (e != hipSuccess) return e;
  double* coef = (double*)ws;
  double* diag = coef + 26 * n;
  uint32_t* mask = (uint32_t*)(diag + n);
  float* srcc = (float*)(mask + n);
  unsigned* flags = (unsigned*)(srcc + n);
  unsigned* abort = flags + G;
  k_rbgs_table<<<blocks_for(n, BS), BS, 0, s>>>(d, type, kappa, src, coef, diag, mask, srcc);
  k_zero_u32<<<(G + 1 + 255) / 256, 256, 0, s>>>(flags, G + 1);
  e = hipGetLastError();
  if (e == hipSuccess) {
    const double *cf = coef, *dg = diag;
    const uint32_t* mk = mask;
    const float* sr = srcc;
    float* ph = phi;
    int sw = sweeps, pw = ppw;
    unsigned *fl = flags, *ab = abort;
    void* args[] = {&d, &cf, &dg, &mk, &sr, &ph, &sw, &pw, &fl, &ab};
    e = hipLaunchCooperativeKernel((const void*)k_rbgs_slabs, dim3(G), dim3(BS), args, 0, s);
  }
  return finish_single_launch(e, ws, abort, s);
}

hipError_t launch_rbgs_exact(pfsf_dims d, const uint8_t* type, const float* kappa,
                             const float* src, float* phi, int sweeps, hipStream_t s) {
  if (sweeps <= 0) return hipSuccess;
  bool capturing = true;  // a stream under capture takes the direct launches
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) == hipSuccess)
    capturing = cap != hipStreamCaptureStatusNone;
  else
    (void)hipGetLastError();
  int sw = 0;
  if (!capturing) {
    int cus = 0;
    hipError_t err = device_cus(&cus);
    if (err != hipSuccess) return err;
    if (rbgs_planes_fit(d, cus)) return rbgs_planes(d, type, kappa, src, phi, sweeps, s);
    if (nvox(d) <= kRbgsTableMax) return rbgs_table(d, type, kappa, src, phi, sweeps, cus, s);
    if (sweeps >= kRbgsGraphSweeps) {
      hipGraphExec_t exec = nullptr;
      err = rbgs_graph(d, type, kappa, src, phi, &exec);
      if (err != hipSuccess) return err;
      for (; sw + kRbgsGraphSweeps <= sweeps; sw += kRbgsGraphSweeps) {
        err = hipGraphLaunch(exec, s);
        if (err != hipSuccess) return err;
      }
    }
  }
  for (; sw < sweeps; sw++) rbgs_sweep(d, type, kappa, src, phi, s);
  return hipGetLastError();
}

hipError_t launch_macro_residual(pfsf_dims d, const uint8_t* type, const float* kappa,
                                const float* phi, const float* b, float* blk, hipStream_t s) {
  const int64_t nb = (int64_t)((d.lx + 7) / 8) * ((d.ly + 7) / 8) * ((d.lz + 7) / 8);
  if (nb == 0) return hipSuccess;
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  k_zero_u32<<<(unsigned)((nb + 255) / 256), 256, 0, s>>>((unsigned*)blk, (int)nb);
  k_macro_residual<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, kappa, phi, b,
                                                          (unsigned*)blk);
  return hipGetLastError();
}

hipError_t launch_failure_scan(pfsf_dims d, const uint8_t* type, const float* kappa,
                               const float* phi, const float* rcomp, const float* rtens,
                               float phi_orphan, uint8_t* code, hipStream_t s) {
  const int64_t tiles = (int64_t)((d.lx + kFsTX - 1) / kFsTX) * ((d.ly + kFsTY - 1) / kFsTY) *
                        ((d.lz + kFsTZ - 1) / kFsTZ);
  if (tiles == 0) return hipSuccess;
  if (tiles > 0x7fffffff) return hipErrorInvalidValue;
  k_failure_scan<<<(unsigned)tiles, BS, 0, s>>>(d, type, kappa, phi, rcomp, rtens, phi_orphan,
                                                code);
  return hipGetLastError();
}

hipError_t launch_history(pfsf_dims d, const uint8_t* type, const float* kappa,
                          const float* phi, float* H, float* psi, hipStream_t s) {
  k_history<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, kappa, phi, H, psi);
  return hipGetLastError();
}

hipError_t launch_apply_f64(pfsf_dims d, const uint8_t* type, const float* kappa,
                            const double* x, double* y, hipStream_t s) {
  k_apply_f64<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, type, kappa, x, y);
  return hipGetLastError();
}

int64_t compact_scratch_elems(pfsf_dims d) { return (int64_t)blocks_for(nvox(d), CSEG) + 2; }

hipError_t launch_compact(pfsf_dims d, const uint8_t* code, pfsf_pos origin, int32_t* out,
                          int32_t capacity, int32_t* scratch, int64_t scratch_elems,
                          hipStream_t s, bool append) {
  int nb = (int)blocks_for(nvox(d), CSEG);
  if (scratch_elems < nb + 2) return hipErrorInvalidValue;
  k_compact_count<<<nb, CB, 0, s>>>(nvox(d), code, scratch);
  k_compact_scan<<<1, CB, 0, s>>>(nb, scratch, append ? out : nullptr);
  k_compact_write<<<nb, CB, 0, s>>>(d, code, scratch, nb, origin, out, capacity);
  return hipGetLastError();
}

// ── Compaction without allocation ───────────────────────────────────────
// The count/scan scratch of pfsf_hip_compact_failures_noalloc: a library-owned
// static device array (one copy per GPU, held until the process ends).  Host
// code never passes `scratch` itself (a host-side stand-in): it passes the
// device address hipGetSymbolAddress returns, looked up once per device on the
// first call with that device current.
constexpr int64_t kNoAllocScratchInts = int64_t(1) << 18;  // 1 MiB
static __device__ __align__(256) int32_t scratch[kNoAllocScratchInts];

static hipError_t noalloc_scratch(int32_t** out) {
  static std::mutex mu;
  static void* per_dev[64] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!per_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    per_dev[dev] = p;
  }
  *out = (int32_t*)per_dev[dev];
  return hipSuccess;
}

hipError_t launch_downsample(pfsf_dims fine, const float* field, const uint8_t* type,
                             float* coarse, uint8_t* ctype, hipStream_t s) {
  pfsf_dims c{(fine.lx + 1) / 2, (fine.ly + 1) / 2, (fine.lz + 1) / 2};
  k_downsample<<<blocks_for(nvox(c), BS), BS, 0, s>>>(fine, field, type, coarse, ctype);
  return hipGetLastError();
}

hipError_t launch_build_island_data(pfsf_dims d, const int32_t* mat_idx,
                                   const pfsf_material* mats, int32_t nmat,
                                   const uint8_t* indestructible, const uint8_t* anchor,
                                   const float* curing, const float* fill, uint8_t* type,
                                   float* source, float* cond, float* rcomp, float* rtens,
                                   float* max_phi, hipStream_t s) {
  k_build_island<<<blocks_for(nvox(d), BS), BS, 0, s>>>(d, mat_idx, mats, nmat, indestructible,
                                                        anchor, curing, fill, type, source, cond,
                                                        rcomp, rtens, max_phi);
  return hipGetLastError();
}

hipError_t launch_kappa_from_soa6(const float* cond, int64_t n, float* kappa, hipStream_t s) {
  k_kappa_soa6<<<blocks_for(n, BS), BS, 0, s>>>(cond, n, kappa);
  return hipGetLastError();
}

hipError_t launch_max_soa6(const float* cond, int64_t n6, unsigned* max_bits, hipStream_t s) {
  unsigned nb = blocks_for(n6, BS);
  if (nb > 1024) nb = 1024;
  k_max_bits<<<nb, BS, 0, s>>>(cond, n6, max_bits);
  return hipGetLastError();
}

hipError_t launch_scale3(float* a, float* b, float* c, float* cond, int64_t n,
                         const unsigned* max_bits, hipStream_t s) {
  unsigned nb = blocks_for(6 * n, BS);
  if (nb > 4096) nb = 4096;
  k_scale3<<<nb, BS, 0, s>>>(a, b, c, cond, n, max_bits);
  return hipGetLastError();
}

}  // namespace pfsf

extern "C" pfsf_result pfsf_hip_compact_failures_noalloc(pfsf_dims d, const uint8_t* code,
                                                        pfsf_pos origin, int32_t* out,
                                                        int32_t capacity, void* stream) {
  using namespace pfsf;
  if (d.lx <= 0 || d.ly <= 0 || d.lz <= 0 || !code || !out || capacity < 0)
    return PFSF_ERROR_INVALID_ARG;
  const int64_t ne = compact_scratch_elems(d);
  if (ne > kNoAllocScratchInts) return PFSF_ERROR_INVALID_ARG;  // grid beyond the scratch size
  int32_t* ws = nullptr;
  hipError_t e = noalloc_scratch(&ws);
  if (e == hipSuccess) e = launch_compact(d, code, origin, out, capacity, ws, ne, (hipStream_t)stream);
  if (e == hipSuccess) return PFSF_OK;
  (void)hipGetLastError();
  return PFSF_ERROR_VULKAN;
}
