// FP32 multigrid-preconditioned PCG with an FP64 iterate and FP64 restarts.
//
// Operator (every level), in difference form:
//   y_i = s_l * sum_{j free} w(k_i, k_j) (x_i - x_j) + m_i x_i
// over the 26 neighbours, w = (k_i+k_j)/2 on faces, 0.5*sqrt(k_i)sqrt(k_j) on
// edges, 0.1666667*sqrt(k_i)sqrt(k_j) on corners (the coupling rule of
// VoxelPhysicsCpuReference.java:97-105); m_i ("anchor mass") is the sum of the
// couplings to Dirichlet neighbours.  Forming x_i - x_j before the product
// keeps the FP32 operator accurate on the smooth vectors PCG produces (the
// expanded form D x_i - sum w x_j cancels catastrophically there).  The free
// set (solid voxels connected to an anchor, with a non-zero diagonal) is
// encoded in the kappa array itself: kappa < 0 marks a cell outside it.
//
// Coarse levels are rediscretised: a coarse cell is free when any child is,
// its kappa is the sum of its free children's kappa / 8, couplings are scaled
// by 2 per level (the h-scaling of a cell-centred 3D Laplacian), and the
// children's anchor masses add up.  Prolongation is piecewise constant and
// restriction its transpose; damped-Jacobi pre/post smoothing with the same
// weight and a fixed-degree Chebyshev coarsest solve keep the V-cycle a
// symmetric positive-definite preconditioner.
//
// Hot kernels work on 32 x 8 x TZ voxel tiles staged in LDS with a one-voxel
// halo (value, kappa, sqrt kappa), so every output reads its 26 neighbours
// from LDS.  TZ is chosen per level so that small levels still launch enough
// work-groups.  Work-groups cover only tiles holding free voxels; the tile
// list is ordered so that consecutive work-groups on one XCD (work-groups are
// dealt to the 8 XCDs round-robin) take neighbouring tiles and share halos in
// that XCD's L2.  Dot products are FP64: each work-group writes a partial and
// a one-block finalize kernel sums them in a fixed order (deterministic, and
// no device-scope fence per work-group).
//
// Reference cycle for contrast: SH/mg_restrict.comp.glsl:118-195,
// SH/mg_prolong.comp.glsl:39-80, NP/src/solver/dispatcher_vcycle.cpp:342-392
// (fixed 16-step schedule, no tolerance); PCG kernels SH/pcg_*.comp.glsl,
// NP/src/solver/dispatcher_pcg.cpp:254-297 (FP32 dots, N <= 512^2).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include <hip/hip_ext.h>

#include "mg_solver.h"

namespace pfsf {
namespace {

constexpr int BS = 256;
constexpr int kMaxWaves = 16;  // waves of the largest work-group (1024 threads)
// Waves per SIMD the tiled stencil kernels are register-budgeted for (LDS
// admits six 24 KB tiles per CU).
#ifndef PFSF_WPE
#define PFSF_WPE 4
#endif

constexpr int TX = kTX, TY = kTY;
constexpr int HX = TX + 2, HY = TY + 2, HXY = HX * HY;
static_assert(TX * TY == BS, "one thread per (x, y) column of a tile");

// ── storage helpers (FP32 or BF16 kappa) ───────────────────────────────
__device__ __forceinline__ float ldk(const float* p, int64_t i) { return p[i]; }
__device__ __forceinline__ float ldk(const uint16_t* p, int64_t i) {
  return __uint_as_float(((unsigned)p[i]) << 16);
}
__device__ __forceinline__ void stk(float* p, int64_t i, float v) { p[i] = v; }
__device__ __forceinline__ void stk(uint16_t* p, int64_t i, float v) {
  unsigned u = __float_as_uint(v);
  unsigned r = u + 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  p[i] = (uint16_t)(r >> 16);
}

// Loads and stores at a 32-bit element index.  The byte offset is formed in
// 32 bits (the solver's arrays stay below 4 GiB, checked at creation), so the
// access is the scalar kernel-argument base plus one 32-bit vector offset.
template <typename T>
__device__ __forceinline__ T ld(const T* p, int i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) +
                                     (unsigned)i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(T* p, int i, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + (unsigned)i * (unsigned)sizeof(T)) = v;
}
__device__ __forceinline__ float ldk32(const float* p, int i) { return ld(p, i); }
__device__ __forceinline__ float ldk32(const uint16_t* p, int i) {
  return __uint_as_float(((unsigned)ld(p, i)) << 16);
}

__device__ __forceinline__ int64_t flat(int x, int y, int z, pfsf_dims d) {
  return (int64_t)x + (int64_t)d.lx * ((int64_t)y + (int64_t)d.ly * z);
}

__device__ __forceinline__ void unflat(int64_t i, pfsf_dims d, int& x, int& y, int& z) {
  x = (int)(i % d.lx);
  int64_t r = i / d.lx;
  y = (int)(r % d.ly);
  z = (int)(r / d.ly);
}

__device__ __forceinline__ float kapf(float k) { return k > 0.0f ? k : 0.0f; }

// Hardware reciprocal (1 ulp) for the smoothers' D^-1: every use of a given
// diagonal entry gets the same value, so pre- and post-smoothing stay
// transposes of each other.
__device__ __forceinline__ float rcp(float v) { return __builtin_amdgcn_rcpf(v); }

template <typename KT>
struct Lv {
  pfsf_dims d;
  const KT* kap;     // kappa on the free set, -1 elsewhere
  const float* dg;   // diagonal (0 outside the free set)
  const float* am;   // anchor mass
  float scale;
  float omega;
};

struct TL {
  int ntx, nty;
  const int* ids;
};

// ── gather form of the operator (setup) ────────────────────────────────
// sum over free neighbours of w_ij * v(j).
template <typename KT, typename F>
__device__ __forceinline__ float offsum(const Lv<KT>& L, int x, int y, int z, float ki, F v) {
  const pfsf_dims d = L.d;
  const float si = sqrtf(ki);
  float af = 0.f, ae = 0.f, ac = 0.f;
  for (int dz = -1; dz <= 1; dz++) {
    int nz = z + dz;
    if (nz < 0 || nz >= d.lz) continue;
    for (int dy = -1; dy <= 1; dy++) {
      int ny = y + dy;
      if (ny < 0 || ny >= d.ly) continue;
      for (int dx = -1; dx <= 1; dx++) {
        int nx = x + dx;
        if (nx < 0 || nx >= d.lx) continue;
        int cls = (dx != 0) + (dy != 0) + (dz != 0);
        if (cls == 0) continue;
        int64_t j = flat(nx, ny, nz, d);
        float kj = ldk(L.kap, j);
        if (kj < 0.f) continue;
        float vj = v(j);
        if (cls == 1)
          af += (ki + kj) * vj;
        else if (cls == 2)
          ae += sqrtf(kj) * vj;
        else
          ac += sqrtf(kj) * vj;
      }
    }
  }
  return L.scale * (0.5f * af + si * (kEdgeP * ae + kCornerP * ac));
}

// ── LDS tile machinery ─────────────────────────────────────────────────
// Grids hold fewer than 2^31 cells (checked at creation), so the hot kernels
// index with 32-bit integers.
__device__ __forceinline__ int flat32(int x, int y, int z, pfsf_dims d) {
  return x + d.lx * (y + d.ly * z);
}

template <int TZ>
struct Tile {
  static constexpr int HZ = TZ + 2, NROWS = HY * HZ, HN = HXY * HZ;
  float v[HN];  // value (0 outside the free set)
  float k[HN];  // kappa (< 0 outside the free set)
  float s[HN];  // sqrt kappa (0 outside the free set); hardware square root,
                // used for both ends of every coupling so A stays symmetric
};

template <int TZ>
__device__ __forceinline__ bool tile_origin(const TL& M, int& x0, int& y0, int& z0) {
  int t = M.ids[blockIdx.x];
  if (t < 0) return false;
  int tx = t % M.ntx;
  int r = t / M.ntx;
  x0 = tx * TX;
  y0 = (r % M.nty) * TY;
  z0 = (r / M.nty) * TZ;
  return true;
}

// Halo staging in two phases.  stage_load issues every global load of the
// tile's halo box (kappa plus whatever raw operands ld(j, x, y, z) reads) and
// returns them in registers; stage_put combines them with fin(raw) and writes
// LDS, so a wave waits for memory once, after all loads are in flight.
//
// Rows of the box: the 32 lanes of a half-wave take the 32 interior columns of
// one (y, z) row, eight rows per pass; one more pass covers the two halo
// columns of every row.  A cell outside the grid is redirected to cell 0
// (always in bounds) and discarded, so no load is behind a branch; fin may
// see garbage or divide by zero outside the free set, and the select drops
// it.  Cells outside the grid or the free set hold v = 0, k = -1, s = 0.
template <int TZ>
struct Stage {
  static constexpr int NROWS = Tile<TZ>::NROWS;
  static constexpr int RPP = BS / TX;
  static constexpr int NP = (NROWS + RPP - 1) / RPP + 1;  // passes incl. halo columns
  static_assert(2 * NROWS <= BS, "halo columns fit one pass");
};

template <int TZ, typename KT, typename Raw, typename LD>
__device__ __forceinline__ void stage_load(const Lv<KT>& L, int x0, int y0, int z0, LD ld,
                                           float (&kk)[Stage<TZ>::NP],
                                           Raw (&raw)[Stage<TZ>::NP]) {
  constexpr int NROWS = Stage<TZ>::NROWS, RPP = Stage<TZ>::RPP, NP = Stage<TZ>::NP;
  const pfsf_dims d = L.d;
  const int t = threadIdx.x;
  auto cell = [&](int p, int gx, int gy, int gz, bool ok) {
    const int cx = ok ? gx : 0, cy = ok ? gy : 0, cz = ok ? gz : 0;
    const int j = flat32(cx, cy, cz, d);
    const float kj = ldk32(L.kap, j);
    raw[p] = ld(j, cx, cy, cz);
    kk[p] = ok ? kj : -1.f;
  };
  const int tx = t % TX, tr = t / TX;
  const int gx = x0 + tx;
  const bool inx = gx < d.lx;
#pragma unroll
  for (int p = 0; p < NP - 1; p++) {
    const int row = tr + p * RPP;
    const int ey = row % HY, ez = row / HY;
    const int gy = y0 + ey - 1, gz = z0 + ez - 1;
    cell(p, gx, gy, gz, row < NROWS && inx && gy >= 0 && gy < d.ly && gz >= 0 && gz < d.lz);
  }
  const int hrow = t >> 1, ex = (t & 1) ? HX - 1 : 0;
  const int ey = hrow % HY, ez = hrow / HY;
  const int hx = x0 + ex - 1, gy = y0 + ey - 1, gz = z0 + ez - 1;
  cell(NP - 1, hx, gy, gz,
       hrow < NROWS && hx >= 0 && hx < d.lx && gy >= 0 && gy < d.ly && gz >= 0 && gz < d.lz);
}

template <int TZ>
__device__ __forceinline__ void put_cell(Tile<TZ>& T, int e, float k, float v) {
  T.v[e] = k >= 0.f ? v : 0.f;
  T.k[e] = k;
  T.s[e] = k > 0.f ? __builtin_amdgcn_sqrtf(k) : 0.f;
}

template <int TZ, typename Raw, typename FIN>
__device__ __forceinline__ void stage_put(Tile<TZ>& T, const float (&kk)[Stage<TZ>::NP],
                                          const Raw (&raw)[Stage<TZ>::NP], FIN fin) {
  constexpr int NROWS = Stage<TZ>::NROWS, RPP = Stage<TZ>::RPP, NP = Stage<TZ>::NP;
  const int t = threadIdx.x;
  const int tx = t % TX, tr = t / TX;
#pragma unroll
  for (int p = 0; p < NP - 1; p++) {
    const int row = tr + p * RPP;
    if (p < NP - 2 || row < NROWS) put_cell<TZ>(T, row * HX + tx + 1, kk[p], fin(raw[p]));
  }
  const int hrow = t >> 1, ex = (t & 1) ? HX - 1 : 0;
  if (hrow < NROWS) put_cell<TZ>(T, hrow * HX + ex, kk[NP - 1], fin(raw[NP - 1]));
}

// One-phase staging: val(j, x, y, z) is the staged value itself.
template <int TZ, typename KT, typename F>
__device__ __forceinline__ void stage(Tile<TZ>& T, const Lv<KT>& L, int x0, int y0, int z0,
                                      F val) {
  float kk[Stage<TZ>::NP], vv[Stage<TZ>::NP];
  stage_load<TZ>(L, x0, y0, z0, val, kk, vv);
  stage_put<TZ>(T, kk, vv, [](float v) { return v; });
}

// (A v)_i at LDS cell c of a box with row stride sy and plane stride sz:
// difference form over free neighbours + anchor mass.  With constant strides
// every read is one LDS instruction with an immediate offset from one base.
__device__ __forceinline__ float stencil_apply(const float* __restrict__ v,
                                               const float* __restrict__ k,
                                               const float* __restrict__ s, int c, int sy,
                                               int sz, float ki, float mi, float scale) {
  const float vc = v[c];
  float af = 0.f, ae = 0.f, ac = 0.f;
  auto face = [&](int o) {
    float kj = k[c + o];
    float w = kj >= 0.f ? ki + kj : 0.f;
    af += w * (vc - v[c + o]);
  };
  auto edge = [&](int o) { ae += s[c + o] * (vc - v[c + o]); };
  auto corner = [&](int o) { ac += s[c + o] * (vc - v[c + o]); };
  face(-1);
  face(1);
  face(-sy);
  face(sy);
  face(-sz);
  face(sz);
  edge(-1 - sy);
  edge(1 - sy);
  edge(-1 + sy);
  edge(1 + sy);
  edge(-1 - sz);
  edge(1 - sz);
  edge(-1 + sz);
  edge(1 + sz);
  edge(-sy - sz);
  edge(sy - sz);
  edge(-sy + sz);
  edge(sy + sz);
  corner(-1 - sy - sz);
  corner(1 - sy - sz);
  corner(-1 + sy - sz);
  corner(1 + sy - sz);
  corner(-1 - sy + sz);
  corner(1 - sy + sz);
  corner(-1 + sy + sz);
  corner(1 + sy + sz);
  return scale * (0.5f * af + s[c] * (kEdgeP * ae + kCornerP * ac)) + mi * vc;
}

template <int TZ>
__device__ __forceinline__ float tile_apply(const Tile<TZ>& T, int c, float ki, float mi,
                                            float scale) {
  return stencil_apply(T.v, T.k, T.s, c, HX, HXY, ki, mi, scale);
}

__device__ __forceinline__ int tcell(int tx, int ty, int tz) {
  return (tz + 1) * HXY + (ty + 1) * HX + (tx + 1);
}

// ── FP64 reductions ────────────────────────────────────────────────────
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double (*sh)[BS / 64]) {
#pragma unroll
  for (int k = 0; k < N; k++) {
    double a = v[k];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      double a = 0.0;
      for (int w = 0; w < BS / 64; w++) a += sh[k][w];
      v[k] = a;
    }
  }
}

// Block partials go to partials[k * gridDim.x + blockIdx.x]; a one-block
// finalize kernel sums them in a fixed order, so results are deterministic
// and the kernel boundary orders the writes.
template <int N>
__device__ __forceinline__ void write_partials(double (&v)[N], double* partials,
                                               double (*sh)[BS / 64]) {
  block_sum<N>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; k++) partials[(int64_t)k * gridDim.x + blockIdx.x] = v[k];
  }
}

// Sum of n FP64 partials, returned to every thread of the work-group.  Every
// work-group adds in the same order, so each one obtains the same value: a
// consumer kernel can finish a producer's reduction itself instead of
// waiting for a separate one-block kernel.
__device__ __forceinline__ double block_total(const double* __restrict__ part, unsigned n) {
  __shared__ double tot_sh[kMaxWaves];
  double a = 0.0;
  for (unsigned b = threadIdx.x; b < n; b += blockDim.x) a += part[b];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) tot_sh[threadIdx.x >> 6] = a;
  __syncthreads();
  double t = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int w = 0; w < nw; w++) t += tot_sh[w];
  return t;
}

// One-work-group reductions that remain separate kernels: the Jacobi
// (single-level) iteration, a coarsest level too large for LDS, and the
// FP64 norms of the solve set-up and the refinement passes.
enum FinMode {
  FIN_ALPHA = 0,  // p.q -> alpha = rz / pq
  FIN_UPDATE,     // r.r -> rr, iteration count, convergence
  FIN_BNORM,      // b.b -> bb
  FIN_IR,         // r.r -> rr, rel
};

__global__ void __launch_bounds__(BS)
    k_finalize(int mode, const double* __restrict__ partials, unsigned nblk, PcgScalars* sc,
               IrScalars* ir) {
  if (mode <= FIN_UPDATE && sc->done) return;
  const double v = block_total(partials, nblk);
  if (threadIdx.x != 0) return;
  switch (mode) {
    case FIN_ALPHA:
      sc->pq = v;
      if (v > 0.0) {
        sc->alpha = sc->rz / v;
        sc->have_dir = 1;
      } else {
        sc->alpha = 0.0;
        sc->done = 2;  // breakdown: the direction carries no energy
      }
      break;
    case FIN_UPDATE:
      sc->rr = v;
      sc->iters += 1;
      if (v <= sc->stop2 || sc->iters >= sc->max_iters) sc->done = 1;
      break;
    case FIN_BNORM:
      ir->bb = v;
      break;
    case FIN_IR:
      ir->rr = v;
      ir->rel = ir->bb > 0.0 ? sqrt(v / ir->bb) : sqrt(v);
      break;
  }
}

// ── PCG + V-cycle kernels ──────────────────────────────────────────────
// The output loops below evaluate the stencil for every cell of the column
// and select the result (a cell outside the free set yields 0), so no branch
// depends on the free mask; only the stores are guarded by the grid bounds.

// p_out = z + beta p_in (z alone for the first direction); q = A p_out; p.q.
// beta = (r.z) / (r.z)_previous is formed here from the r.z partials of the
// preconditioner (rz_n work-groups); work-group 0 publishes r.z in slot
// `slot` (the iteration parity) and beta.
template <int TZ, typename KT>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    k_spmv_dir(Lv<KT> L, TL M, const float* __restrict__ z, const float* __restrict__ pin,
               float* __restrict__ pout, float* __restrict__ q, PcgScalars* sc,
               const double* __restrict__ rz_part, unsigned rz_n, int slot, double* partials) {
  if (sc->done) return;
  __shared__ Tile<TZ> T;
  __shared__ double sh[1][BS / 64];
  double acc[1] = {0.0};
  const bool first = !sc->have_dir;
  const double rz_old = sc->rzs[slot ^ 1];
  const double rz = block_total(rz_part, rz_n);
  int x0 = 0, y0 = 0, z0 = 0;
  const bool have = tile_origin<TZ>(M, x0, y0, z0);
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const pfsf_dims d = L.d;
  const int gx = x0 + tx, gy = y0 + ty, sz = d.lx * d.ly;
  const bool col = gx < d.lx && gy < d.ly;
  const int i0 = flat32(gx, gy, z0, d);
  bool inb[TZ];
  float amr[TZ];  // own-voxel operands (cell 0 stands in for a cell outside the grid)
  float kk[Stage<TZ>::NP];
  float2 raw[Stage<TZ>::NP];
  if (have) {
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      inb[tz] = col && z0 + tz < d.lz;
      amr[tz] = ld(L.am, inb[tz] ? i0 + tz * sz : 0);
    }
    stage_load<TZ>(L, x0, y0, z0,
                   [&](int j, int, int, int) { return make_float2(ld(z, j), ld(pin, j)); }, kk,
                   raw);
  }
  const double beta64 = (!first && rz_old > 0.0) ? rz / rz_old : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rzs[slot] = rz;
    sc->rz = rz;
    sc->beta = beta64;
  }
  if (have) {
    // p_in is finite everywhere (zero-initialised), so beta = 0 drops it
    const float beta = (float)beta64;
    stage_put<TZ>(T, kk, raw, [&](float2 v) { return v.x + beta * v.y; });
    __syncthreads();
    const int c0 = tcell(tx, ty, 0);
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      const int c = c0 + tz * HXY;
      const float ki = T.k[c], pv = T.v[c];
      float qi = tile_apply<TZ>(T, c, ki, amr[tz], L.scale);
      qi = ki >= 0.f ? qi : 0.f;
      acc[0] += (double)pv * qi;
      if (inb[tz]) {
        st(pout, i0 + tz * sz, pv);
        st(q, i0 + tz * sz, qi);
      }
    }
  }
  write_partials<1>(acc, partials, sh);
}

// Sum over each 2x2x2 group of children held by one tile, from per-thread
// values res[tz].  Lanes tx^1 and ty^1 (= lane^32) hold the x and y
// siblings.  Writes rc for every in-bounds coarse cell.
template <int TZ>
__device__ __forceinline__ void restrict_tile(const float (&res)[TZ], int x0, int y0, int z0,
                                              int tx, int ty, pfsf_dims c, float* rc) {
  const int X = (x0 + tx) >> 1, Y = (y0 + ty) >> 1;
  const bool lead = !(tx & 1) && !(ty & 1) && X < c.lx && Y < c.ly;
#pragma unroll
  for (int k = 0; k < TZ / 2; k++) {
    float a = res[2 * k] + res[2 * k + 1];
    a = a + __shfl_xor(a, 1, 64);
    a = a + __shfl_xor(a, 32, 64);
    const int Z = (z0 >> 1) + k;
    if (lead && Z < c.lz) st(rc, flat32(X, Y, Z, c), a);
  }
}

// Down-sweep from level l: the pre-smoothed iterate x = omega D^-1 r is formed
// on the fly, and the coarse right-hand side is rc = P^T (r - A x).
// With UPDATE (finest level inside an iteration): first x64 += alpha p and
// r_out = r_in - alpha q, then the same on r_out, plus r.r for convergence.
// alpha = (r.z) / (p.q) is formed here from the p.q partials of the SpMV
// (pq_n work-groups); work-group 0 publishes it (or flags a breakdown when
// p.q <= 0).
template <int TZ, typename KT, bool UPDATE>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    k_down(Lv<KT> L, TL M, pfsf_dims c, const float* __restrict__ rin, float* __restrict__ rout,
           const float* __restrict__ q, const float* __restrict__ p, double* __restrict__ x,
           float* __restrict__ rc, PcgScalars* sc, const double* __restrict__ pq_part,
           unsigned pq_n, double* partials) {
  if (sc->done) return;
  __shared__ Tile<TZ> T;
  __shared__ double sh[1][BS / 64];
  double acc[1] = {0.0};
  const double rz = UPDATE ? sc->rz : 0.0;
  const double pq0 = UPDATE ? block_total(pq_part, pq_n) : 0.0;
  const pfsf_dims d = L.d;
  const float om = L.omega;
  int x0 = 0, y0 = 0, z0 = 0;
  const bool have = tile_origin<TZ>(M, x0, y0, z0);
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int gx = x0 + tx, gy = y0 + ty, sz = d.lx * d.ly;
  const bool col = gx < d.lx && gy < d.ly;
  const int i0 = flat32(gx, gy, z0, d);
  // own-voxel operands (cell 0 stands in for a cell outside the grid; its
  // values are never used)
  bool inb[TZ];
  float ro[TZ], qo[TZ], amr[TZ], pr[TZ];
  double xr[TZ];
  float kk[Stage<TZ>::NP];
  float3 raw[Stage<TZ>::NP];  // r_in, q, diagonal
  if (have) {
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      inb[tz] = col && z0 + tz < d.lz;
      const int i = inb[tz] ? i0 + tz * sz : 0;
      ro[tz] = ld(rin, i);
      qo[tz] = UPDATE ? ld(q, i) : 0.f;
      amr[tz] = ld(L.am, i);
      pr[tz] = UPDATE ? ld(p, i) : 0.f;
      xr[tz] = UPDATE ? ld(x, i) : 0.0;
    }
    stage_load<TZ>(L, x0, y0, z0,
                   [&](int j, int, int, int) {
                     return make_float3(ld(rin, j), UPDATE ? ld(q, j) : 0.f, ld(L.dg, j));
                   },
                   kk, raw);
  }
  double alpha64 = 0.0;
  if (UPDATE) {
    const double pq = pq0;
    const bool ok = pq > 0.0;
    alpha64 = ok ? rz / pq : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->pq = pq;
      sc->alpha = alpha64;
      sc->have_dir = 1;
      if (!ok) sc->done = 2;  // breakdown: the direction carries no energy
    }
    if (!ok) return;
  }
  if (have) {
    const float alpha = (float)alpha64;
    stage_put<TZ>(T, kk, raw, [&](float3 v) {
      const float rj = UPDATE ? v.x - alpha * v.y : v.x;
      return om * rj * rcp(v.z);
    });
    float rr_[TZ];
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) rr_[tz] = UPDATE ? ro[tz] - alpha * qo[tz] : ro[tz];
    if (UPDATE) {
      // p and q are zero outside the free set, so the update leaves x there
      // unchanged and carries r there over
#pragma unroll
      for (int tz = 0; tz < TZ; tz++) {
        if (inb[tz]) {
          st(x, i0 + tz * sz, xr[tz] + alpha64 * (double)pr[tz]);
          st(rout, i0 + tz * sz, rr_[tz]);
        }
      }
    }
    __syncthreads();
    const int c0 = tcell(tx, ty, 0);
    float res[TZ];
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      const int cc = c0 + tz * HXY;
      const float ki = T.k[cc], ri = rr_[tz];
      const bool fr = ki >= 0.f;
      const float ax = tile_apply<TZ>(T, cc, ki, amr[tz], L.scale);
      res[tz] = fr ? ri - ax : 0.f;
      if (UPDATE) acc[0] += fr ? (double)ri * ri : 0.0;
    }
    restrict_tile<TZ>(res, x0, y0, z0, tx, ty, c, rc);
  }
  if (UPDATE) write_partials<1>(acc, partials, sh);
}

// Up-sweep at level l: x~ = omega D^-1 r + P e_coarse on the free set, then
// one damped-Jacobi post-smoothing step out = x~ + omega D^-1 (r - A x~).
// At the finest level it also accumulates r.z.
template <int TZ, typename KT, bool FINEST>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    k_up(Lv<KT> L, TL M, pfsf_dims c, const float* __restrict__ r, const float* __restrict__ ec,
         float* __restrict__ out, const PcgScalars* sc, double* partials) {
  if (sc->done) return;
  __shared__ Tile<TZ> T;
  __shared__ double sh[1][BS / 64];
  double acc[1] = {0.0};
  int x0, y0, z0;
  const pfsf_dims d = L.d;
  const float om = L.omega;
  if (tile_origin<TZ>(M, x0, y0, z0)) {
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int gx = x0 + tx, gy = y0 + ty, sz = d.lx * d.ly;
    const bool col = gx < d.lx && gy < d.ly;
    const int i0 = flat32(gx, gy, z0, d);
    // own-voxel operands, loaded before the barrier (cell 0 stands in for a
    // cell outside the grid; its values are never used)
    bool inb[TZ];
    float rr_[TZ], amr[TZ], dgr[TZ];
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      inb[tz] = col && z0 + tz < d.lz;
      const int i = inb[tz] ? i0 + tz * sz : 0;
      rr_[tz] = ld(r, i);
      amr[tz] = ld(L.am, i);
      dgr[tz] = ld(L.dg, i);
    }
    stage<TZ>(T, L, x0, y0, z0, [&](int j, int gx_, int gy_, int gz_) {
      return om * ld(r, j) * rcp(ld(L.dg, j)) + ld(ec, flat32(gx_ >> 1, gy_ >> 1, gz_ >> 1, c));
    });
    __syncthreads();
    const int c0 = tcell(tx, ty, 0);
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      const int cc = c0 + tz * HXY;
      const float ki = T.k[cc], xi = T.v[cc], ri = rr_[tz];
      const float ax = tile_apply<TZ>(T, cc, ki, amr[tz], L.scale);
      float zi = xi + om * (ri - ax) * rcp(dgr[tz]);
      zi = ki >= 0.f ? zi : 0.f;
      if (FINEST) acc[0] += ki >= 0.f ? (double)ri * zi : 0.0;
      if (inb[tz]) st(out, i0 + tz * sz, zi);
    }
  }
  if (FINEST) write_partials<1>(acc, partials, sh);
}

// Jacobi-preconditioned variant of the update (single-level solve):
// x64 += alpha p; r_out = r_in - alpha q; z = D^-1 r_out; r.r and r.z.
__global__ void __launch_bounds__(BS)
    k_update_jacobi(int64_t n, const float* __restrict__ dg, double* __restrict__ x,
                    const float* __restrict__ rin, float* __restrict__ rout,
                    const float* __restrict__ p, const float* __restrict__ q, float* __restrict__ z,
                    const PcgScalars* sc, double* partials) {
  if (sc->done) return;
  __shared__ double sh[2][BS / 64];
  double acc[2] = {0.0, 0.0};
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  float alpha = (float)sc->alpha;
  if (i < n) {
    float D = dg[i];
    float ri = rin[i] - alpha * q[i];
    x[i] = x[i] + sc->alpha * (double)p[i];
    rout[i] = ri;
    float zi = D > 0.f ? ri / D : 0.f;
    z[i] = zi;
    acc[0] = (double)ri * ri;
    acc[1] = (double)ri * zi;
  }
  write_partials<2>(acc, partials, sh);
}

// z = D^-1 r and r.z (Jacobi preconditioner at the start of a solve pass).
__global__ void __launch_bounds__(BS)
    k_jacobi_apply(int64_t n, const float* __restrict__ r, const float* __restrict__ dg,
                   float* __restrict__ z, double* partials) {
  __shared__ double sh[1][BS / 64];
  double acc[1] = {0.0};
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) {
    float D = dg[i];
    float zi = D > 0.f ? r[i] / D : 0.f;
    z[i] = zi;
    acc[0] = (double)r[i] * zi;
  }
  write_partials<1>(acc, partials, sh);
}

// One Chebyshev step on the coarsest level (from x = 0), tiled:
//   first:  d = c2 D^-1 r, x = d
//   else :  d = c1 d + c2 D^-1 (r - A x_in), x_out = x_in + d
template <int TZ, typename KT>
__global__ void __launch_bounds__(BS)
    k_cheb_tiled(Lv<KT> L, TL M, const float* __restrict__ r, const float* __restrict__ xin,
                 float* __restrict__ xout, float* __restrict__ dv, float c1, float c2, int first,
                 const PcgScalars* sc) {
  if (sc->done) return;
  __shared__ Tile<TZ> T;
  int x0, y0, z0;
  if (!tile_origin<TZ>(M, x0, y0, z0)) return;
  const pfsf_dims d = L.d;
  stage<TZ>(T, L, x0, y0, z0, [&](int j, int, int, int) { return first ? 0.f : xin[j]; });
  __syncthreads();
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int gx = x0 + tx, gy = y0 + ty;
  if (gx >= d.lx || gy >= d.ly) return;
  for (int tz = 0; tz < TZ; tz++) {
    int gz = z0 + tz;
    if (gz >= d.lz) break;
    const int i = flat32(gx, gy, gz, d);
    int cc = tcell(tx, ty, tz);
    float ki = T.k[cc];
    if (ki < 0.f) {
      xout[i] = 0.f;
      if (first) dv[i] = 0.f;
      continue;
    }
    float D = L.dg[i];
    if (first) {
      float di = c2 * r[i] * rcp(D);
      dv[i] = di;
      xout[i] = di;
    } else {
      float xi = T.v[cc];
      float ax = tile_apply<TZ>(T, cc, ki, L.am[i], L.scale);
      float di = c1 * dv[i] + c2 * (r[i] - ax) * rcp(D);
      dv[i] = di;
      xout[i] = xi + di;
    }
  }
}

// Whole Chebyshev solve of the coarsest level held in one work-group's LDS
// (n <= kMaxLdsCoarse cells, one cell per thread).  Kappa, sqrt kappa and the
// iterate sit in LDS boxes with a one-cell border (kappa -1, value 0), so the
// 26 neighbour reads need no bounds tests; the diagonal, anchor mass, rhs and
// direction of a thread's own cell stay in registers.
constexpr int kMaxLdsCoarse = 1024;
constexpr int kMaxLdsPadded = 2048;

__host__ __device__ inline int padded_cells(pfsf_dims d) {
  return (d.lx + 2) * (d.ly + 2) * (d.lz + 2);
}

//
// With rr_n > 0 it first finishes the PCG iteration's convergence test from
// the r.r partials of the finest update (rr_n work-groups): iteration count,
// r.r, and the stop flag, which also ends this V-cycle.
template <typename KT>
__global__ void __launch_bounds__(kMaxLdsCoarse)
    k_cheb_lds(Lv<KT> L, const float* __restrict__ r, float* __restrict__ out,
               const float* __restrict__ coef, int steps, PcgScalars* sc,
               const double* __restrict__ rr_part, unsigned rr_n) {
  if (sc->done) return;
  if (rr_n > 0) {
    const double rr = block_total(rr_part, rr_n);
    const int it = sc->iters + 1;
    const bool stop = rr <= sc->stop2 || it >= sc->max_iters;
    __syncthreads();  // every thread has read the scalars
    if (threadIdx.x == 0) {
      sc->rr = rr;
      sc->iters = it;
      if (stop) sc->done = 1;
    }
    if (stop) return;
  }
  __shared__ float sk[kMaxLdsPadded], ss[kMaxLdsPadded], sx[2][kMaxLdsPadded];
  const pfsf_dims d = L.d;
  const int sy = d.lx + 2, sz = sy * (d.ly + 2), np = sz * (d.lz + 2);
  for (int e = threadIdx.x; e < np; e += blockDim.x) {
    sk[e] = -1.f;
    ss[e] = 0.f;
    sx[0][e] = 0.f;
    sx[1][e] = 0.f;
  }
  __syncthreads();
  const int n = d.lx * d.ly * d.lz;
  const int i = threadIdx.x;
  const bool act = i < n;
  int e = 0;
  float ki = -1.f, D = 1.f, m = 0.f, ri = 0.f, xi = 0.f, di = 0.f;
  if (act) {
    const int x = i % d.lx, y = (i / d.lx) % d.ly, z = i / (d.lx * d.ly);
    e = (x + 1) + (y + 1) * sy + (z + 1) * sz;
    ki = ldk(L.kap, i);
    if (ki >= 0.f) {
      D = L.dg[i];
      m = L.am[i];
      ri = r[i];
      di = coef[1] * ri * rcp(D);
      xi = di;
    }
    sk[e] = ki;
    ss[e] = ki > 0.f ? __builtin_amdgcn_sqrtf(ki) : 0.f;
    sx[0][e] = xi;
  }
  const bool live = act && ki >= 0.f;
  __syncthreads();
  int cur = 0;
  for (int st = 1; st < steps; st++) {
    const float c1 = coef[2 * st], c2 = coef[2 * st + 1];
    if (live) {
      const float ax = stencil_apply(sx[cur], sk, ss, e, sy, sz, ki, m, L.scale);
      di = c1 * di + c2 * (ri - ax) * rcp(D);
      xi = xi + di;
      sx[cur ^ 1][e] = xi;
    }
    __syncthreads();
    cur ^= 1;
  }
  if (act) out[i] = live ? xi : 0.f;
}

// ── setup kernels ──────────────────────────────────────────────────────
// Fine level: free mask, kappa (-1 outside the free set) and anchor mass
// (FP32 coupling rule summed in FP64).
__global__ void __launch_bounds__(BS)
    k_fine_setup(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                 const uint8_t* __restrict__ reach, float* __restrict__ kap0,
                 float* __restrict__ am0, uint8_t* __restrict__ free0) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= (int64_t)d.lx * d.ly * d.lz) return;
  bool fr = reach[i] != 0;  // solid and connected to an anchor
  float k = fr ? kapf(kappa[i]) : -1.0f;
  kap0[i] = k;
  free0[i] = fr ? 1 : 0;
  if (!fr) {
    am0[i] = 0.0f;
    return;
  }
  int x, y, z;
  unflat(i, d, x, y, z);
  float si = sqrtf(k);
  double m = 0.0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int cls = (dx != 0) + (dy != 0) + (dz != 0);
        if (cls == 0) continue;
        int nx = x + dx, ny = y + dy, nz = z + dz;
        if (nx < 0 || nx >= d.lx || ny < 0 || ny >= d.ly || nz < 0 || nz >= d.lz) continue;
        int64_t j = flat(nx, ny, nz, d);
        if (type[j] != 2) continue;  // anchors only
        float kj = kapf(kappa[j]);
        float w = cls == 1 ? 0.5f * (k + kj) : (cls == 2 ? kEdgeP : kCornerP) * si * sqrtf(kj);
        m += (double)w;
      }
  am0[i] = (float)m;
}

// Coarse cell: free if any child is free; kappa = sum(free children kappa)/8
// (-1 when none is free); anchor mass = sum of children's anchor mass.
template <typename KF, typename KC>
__global__ void __launch_bounds__(BS)
    k_coarsen(pfsf_dims f, const KF* __restrict__ kf, const float* __restrict__ amf,
              const uint8_t* __restrict__ freef, pfsf_dims c, KC* __restrict__ kc,
              float* __restrict__ amc, uint8_t* __restrict__ freec) {
  int64_t ci = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (ci >= (int64_t)c.lx * c.ly * c.lz) return;
  int X, Y, Z;
  unflat(ci, c, X, Y, Z);
  float ks = 0.f, ms = 0.f;
  int nf = 0;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        int x = 2 * X + dx, y = 2 * Y + dy, z = 2 * Z + dz;
        if (x >= f.lx || y >= f.ly || z >= f.lz) continue;
        int64_t i = flat(x, y, z, f);
        if (!freef[i]) continue;
        nf++;
        ks += ldk(kf, i);
        ms += amf[i];
      }
  stk(kc, ci, nf > 0 ? ks * 0.125f : -1.0f);
  amc[ci] = ms;
  freec[ci] = nf > 0 ? 1 : 0;
}

// Diagonal of a level: scale * sum over free neighbours of w + anchor mass.
template <typename KT>
__global__ void __launch_bounds__(BS)
    k_level_diag(Lv<KT> L, const uint8_t* __restrict__ fr, float* __restrict__ dg) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  pfsf_dims d = L.d;
  if (i >= (int64_t)d.lx * d.ly * d.lz) return;
  if (!fr[i]) {
    dg[i] = 0.f;
    return;
  }
  int x, y, z;
  unflat(i, d, x, y, z);
  float s = offsum(L, x, y, z, ldk(L.kap, i), [&](int64_t) { return 1.0f; });
  float D = s + L.am[i];
  dg[i] = D > 0.f ? D : 0.f;
}

// The free set follows the diagonal: cells whose couplings all vanish leave
// the solve (mask cleared, kappa flagged negative).
template <typename KT>
__global__ void __launch_bounds__(BS)
    k_sync_free(int64_t n, const float* __restrict__ dg, uint8_t* __restrict__ fr, KT* kap) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  bool f = dg[i] > 0.f;
  fr[i] = f ? 1 : 0;
  if (!f) stk(kap, i, -1.0f);
}

// Tiles that hold at least one free voxel.
__global__ void __launch_bounds__(BS)
    k_tile_flags(pfsf_dims d, const uint8_t* __restrict__ fr, int ntx, int nty, int tz,
                 int* __restrict__ flags) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= (int64_t)d.lx * d.ly * d.lz) return;
  if (!fr[i]) return;
  int x, y, z;
  unflat(i, d, x, y, z);
  int t = x / TX + ntx * (y / TY + nty * (z / tz));
  flags[t] = 1;  // benign race: every writer stores 1
}

__global__ void __launch_bounds__(BS) k_fill(int64_t n, float* __restrict__ v, float a) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) v[i] = a;
}

// ── FP64 residual of the iterate ───────────────────────────────────────
__device__ __forceinline__ double couple_f64(int cls, float ki, float kj) {
  if (cls == 1) {
    float s = ki + kj;
    return 0.5 * (double)s;
  }
  float p = ki * kj;
  double r = sqrt((double)p);
  return r * (cls == 2 ? (double)kEdgeP : (double)kCornerP);
}

// True residual of the FP64 iterate: r = b - A x with the exact coupling rule
// (FP64), written as r32 = (float) r, plus partial ||r||^2.
__global__ void __launch_bounds__(BS)
    k_true_residual(pfsf_dims d, const uint8_t* __restrict__ type,
                    const float* __restrict__ kappa, const uint8_t* __restrict__ fr,
                    const float* __restrict__ b, const double* __restrict__ x,
                    float* __restrict__ r32, double* partials) {
  __shared__ double sh[1][BS / 64];
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  double acc[1] = {0.0};
  if (i < (int64_t)d.lx * d.ly * d.lz) {
    float ri32 = 0.f;
    if (fr[i]) {
      double xi = x[i];
      int px, py, pz;
      unflat(i, d, px, py, pz);
      float ki = kapf(kappa[i]);
      double ax = 0.0;
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int cls = (dx != 0) + (dy != 0) + (dz != 0);
            if (cls == 0) continue;
            int nx = px + dx, ny = py + dy, nz = pz + dz;
            if (nx < 0 || nx >= d.lx || ny < 0 || ny >= d.ly || nz < 0 || nz >= d.lz) continue;
            int64_t j = flat(nx, ny, nz, d);
            if (type[j] == 0) continue;
            double xj = fr[j] ? x[j] : 0.0;
            ax += couple_f64(cls, ki, kapf(kappa[j])) * (xi - xj);
          }
      double ri = (double)b[i] - ax;
      ri32 = (float)ri;
      acc[0] = ri * ri;
    }
    r32[i] = ri32;
  }
  write_partials<1>(acc, partials, sh);
}

// Iterate outside the free set is zero (the caller's guess may not be).
__global__ void __launch_bounds__(BS)
    k_mask64(int64_t n, const uint8_t* __restrict__ fr, double* __restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n && !fr[i]) x[i] = 0.0;
}

// ||b||^2 over the free set.
__global__ void __launch_bounds__(BS)
    k_bnorm(int64_t n, const uint8_t* __restrict__ fr, const float* __restrict__ b,
            double* partials) {
  __shared__ double sh[1][BS / 64];
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n && fr[i]) acc[0] = (double)b[i] * (double)b[i];
  write_partials<1>(acc, partials, sh);
}

__global__ void __launch_bounds__(BS)
    k_to_f32(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

__global__ void __launch_bounds__(BS) k_zero64(int64_t n, double* __restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) v[i] = 0.0;
}

__global__ void __launch_bounds__(BS)
    k_copy64(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// (Re)start the PCG recursion from the current true residual: no search
// direction yet; the recursion stops once its residual norm falls below
// stop_rel * ||r_true|| (the host then checks the true residual).
__global__ void k_pcg_begin(PcgScalars* sc, const IrScalars* ir, double stop_rel, int max_iters,
                            int reset_count) {
  sc->done = 0;
  if (reset_count) sc->iters = 0;
  sc->have_dir = 0;
  sc->rz = 0.0;
  sc->rzs[0] = sc->rzs[1] = 0.0;
  sc->beta = 0.0;
  sc->alpha = 0.0;
  sc->rr = ir->rr;
  sc->stop2 = ir->rr * stop_rel * stop_rel;
  sc->max_iters = max_iters;
}

// Profiling: keep every kernel live (no convergence stop) for timed replays.
__global__ void k_profile_begin(PcgScalars* sc) {
  sc->done = 0;
  sc->iters = 0;
  sc->stop2 = -1.0;
  sc->max_iters = 0x7fffffff;
}

}  // namespace

// ══════════════════════════════ host side ══════════════════════════════

// Records the first failure in `keep`; returns true when e failed.
static inline bool chk(hipError_t e, hipError_t& keep) {
  if (e == hipSuccess) return false;
  if (keep == hipSuccess) keep = e;
  return true;
}

template <typename T>
static T* dalloc(int64_t n, std::vector<void*>& owned, hipError_t& err) {
  void* p = nullptr;
  if (chk(hipMalloc(&p, (size_t)std::max<int64_t>(n, 1) * sizeof(T)), err)) return nullptr;
  owned.push_back(p);
  return (T*)p;
}

MgSolver::~MgSolver() { release(); }

void MgSolver::release() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (void* p : owned_) (void)hipFree(p);
  owned_.clear();
  if (sc_host_) (void)hipHostFree(sc_host_);
  sc_host_ = nullptr;
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  ev0_ = ev1_ = nullptr;
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  levels_.clear();
}

template <typename KT>
static Lv<KT> view(const MgLevel& l) {
  Lv<KT> v;
  v.d = l.d;
  v.kap = (const KT*)l.kap;
  v.dg = l.dg;
  v.am = l.am;
  v.scale = l.scale;
  v.omega = l.omega;
  return v;
}

static TL tl(const MgLevel& l) { return TL{l.tiles.ntx, l.tiles.nty, l.tiles.ids}; }

static void fill0(float* p, int64_t n, hipStream_t s) {
  k_fill<<<blocks_for(n, BS), BS, 0, s>>>(n, p, 0.f);
}

// Tile depth per level: the deepest of {8, 4} that still gives about eight
// tiles per CU, else 2 (small levels).  PFSF_TILE_Z=8|4|2 forces the depth of
// the finest level (tuning runs).
static int pick_tz(pfsf_dims d, bool finest) {
  if (finest) {
    const char* env = getenv("PFSF_TILE_Z");
    int v = env ? atoi(env) : 0;
    if (v == 8 || v == 4 || v == 2) return v;
  }
  for (int tz : {8, 4}) {
    int64_t nt = (int64_t)((d.lx + TX - 1) / TX) * ((d.ly + TY - 1) / TY) * ((d.lz + tz - 1) / tz);
    if (nt >= 2048) return tz;
  }
  return 2;
}

hipError_t MgSolver::create(const pfsf_solver_desc& desc) {
  hipError_t err = hipSuccess;
  desc_ = desc;
  d0_ = desc.dims;
  n0_ = nvox(d0_);
  type_ = desc.type;
  kappa_ = desc.kappa;
  bf16_ = desc.precision == PFSF_PREC_BF16_COARSE;
  coarse_sweeps_ = desc.coarse_sweeps > 0 ? desc.coarse_sweeps : 12;
  // The solver runs on a stream of its own (default priority) so that its
  // iterations can be captured into a HIP graph; solve() orders it after the
  // caller's stream and waits for it before returning.
  if (chk(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), err)) return err;

  int want = desc.mg_levels;
  if (want <= 0) {
    want = 1;
    pfsf_dims d = d0_;
    while (std::max(d.lx, std::max(d.ly, d.lz)) > 8 && want < 8) {
      d = {(d.lx + 1) / 2, (d.ly + 1) / 2, (d.lz + 1) / 2};
      want++;
    }
  }
  nlev_ = std::max(1, std::min(want, 10));

  sc_ = dalloc<PcgScalars>(1, owned_, err);
  ir_ = dalloc<IrScalars>(1, owned_, err);
  reach_ = dalloc<uint8_t>(n0_, owned_, err);
  flag_ = dalloc<int>(1, owned_, err);
  x_ = dalloc<double>(n0_, owned_, err);
  for (int k = 0; k < 2; k++) {
    r_[k] = dalloc<float>(n0_, owned_, err);
    p_[k] = dalloc<float>(n0_, owned_, err);
  }
  q_ = dalloc<float>(n0_, owned_, err);
  z_ = dalloc<float>(n0_, owned_, err);
  if (err != hipSuccess) return err;
  if (chk(hipHostMalloc((void**)&sc_host_, sizeof(PcgScalars) + sizeof(IrScalars)), err))
    return err;
  chk(hipEventCreate(&ev0_), err);
  chk(hipEventCreate(&ev1_), err);

  levels_.resize(nlev_);
  pfsf_dims d = d0_;
  int64_t maxblk = blocks_for(n0_, BS);
  for (int l = 0; l < nlev_; l++) {
    MgLevel& L = levels_[l];
    L.d = d;
    L.n = nvox(d);
    L.bf16 = (l > 0) && bf16_;
    size_t ks = L.bf16 ? sizeof(uint16_t) : sizeof(float);
    L.kap = dalloc<uint8_t>(L.n * (int64_t)ks, owned_, err);
    L.dg = dalloc<float>(L.n, owned_, err);
    L.am = dalloc<float>(L.n, owned_, err);
    L.fr = dalloc<uint8_t>(L.n, owned_, err);
    L.scale = (float)(1 << l);
    if (l > 0) {
      L.r = dalloc<float>(L.n, owned_, err);
      L.e = dalloc<float>(L.n, owned_, err);
    }
    if (l == nlev_ - 1 && nlev_ > 1) {
      L.xb = dalloc<float>(L.n, owned_, err);
      L.dv = dalloc<float>(L.n, owned_, err);
    }
    L.tiles.tz = pick_tz(d, l == 0);
    L.tiles.ntx = (d.lx + TX - 1) / TX;
    L.tiles.nty = (d.ly + TY - 1) / TY;
    L.tiles.ntz = (d.lz + L.tiles.tz - 1) / L.tiles.tz;
    int64_t nt = (int64_t)L.tiles.ntx * L.tiles.nty * L.tiles.ntz;
    L.tiles.ids = dalloc<int>(nt + 8, owned_, err);
    maxblk = std::max(maxblk, nt + 8);
    d = {(d.lx + 1) / 2, (d.ly + 1) / 2, (d.lz + 1) / 2};
  }
  npartials_ = 4 * maxblk + 64;
  partials_ = dalloc<double>(npartials_, owned_, err);
  ppq_ = partials_;
  prr_ = partials_ + maxblk;
  prz_ = partials_ + 3 * maxblk;
  coef_ = dalloc<float>(2 * 64, owned_, err);
  if (err != hipSuccess) return err;
  // every work vector starts at zero: tiles without free voxels are never
  // written and must read as zero from their neighbours
  for (int k = 0; k < 2; k++) {
    fill0(r_[k], n0_, stream_);
    fill0(p_[k], n0_, stream_);
  }
  fill0(q_, n0_, stream_);
  fill0(z_, n0_, stream_);
  k_zero64<<<blocks_for(n0_, BS), BS, 0, stream_>>>(n0_, x_);
  for (int l = 1; l < nlev_; l++) {
    fill0(levels_[l].r, levels_[l].n, stream_);
    fill0(levels_[l].e, levels_[l].n, stream_);
  }
  chk(hipGetLastError(), err);
  if (err != hipSuccess) return err;
  return build();
}

// Active-tile list in XCD-grouped launch order: sorted tile ids are split
// into 8 contiguous chunks and work-group g takes chunk g % 8, entry g / 8.
hipError_t MgSolver::build_tiles(MgLevel& L) {
  hipError_t err = hipSuccess;
  int64_t nt = (int64_t)L.tiles.ntx * L.tiles.nty * L.tiles.ntz;
  std::vector<int> flags(nt, 0);
  chk(hipMemsetAsync(L.tiles.ids, 0, sizeof(int) * nt, stream_), err);
  k_tile_flags<<<blocks_for(L.n, BS), BS, 0, stream_>>>(L.d, L.fr, L.tiles.ntx, L.tiles.nty,
                                                       L.tiles.tz, L.tiles.ids);
  chk(hipGetLastError(), err);
  chk(hipMemcpyAsync(flags.data(), L.tiles.ids, sizeof(int) * nt, hipMemcpyDeviceToHost,
                     stream_),
      err);
  chk(hipStreamSynchronize(stream_), err);
  if (err != hipSuccess) return err;
  std::vector<int> act;
  for (int64_t t = 0; t < nt; t++)
    if (flags[t]) act.push_back((int)t);
  if (act.empty()) act.push_back(0);  // keep launches non-empty
  int T = (int)act.size();
  int chunk = (T + 7) / 8;
  std::vector<int> ids(8 * chunk, -1);
  for (int g = 0; g < 8 * chunk; g++) {
    int src = (g % 8) * chunk + g / 8;
    ids[g] = src < T ? act[src] : -1;
  }
  L.tiles.nactive = T;
  L.tiles.nlaunch = 8 * chunk;
  chk(hipMemcpyAsync(L.tiles.ids, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice,
                     stream_),
      err);
  chk(hipStreamSynchronize(stream_), err);
  return err;
}

hipError_t MgSolver::build() {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  chk(anchored_mask(d0_, type_, reach_, flag_, s), err);
  MgLevel& F = levels_[0];
  k_fine_setup<<<blocks_for(n0_, BS), BS, 0, s>>>(d0_, type_, kappa_, reach_, (float*)F.kap,
                                                  F.am, F.fr);
  for (int l = 1; l < nlev_; l++) {
    MgLevel& A = levels_[l - 1];
    MgLevel& B = levels_[l];
    unsigned nb = blocks_for(B.n, BS);
    if (!A.bf16 && !B.bf16)
      k_coarsen<float, float><<<nb, BS, 0, s>>>(A.d, (const float*)A.kap, A.am, A.fr, B.d,
                                                (float*)B.kap, B.am, B.fr);
    else if (!A.bf16 && B.bf16)
      k_coarsen<float, uint16_t><<<nb, BS, 0, s>>>(A.d, (const float*)A.kap, A.am, A.fr, B.d,
                                                   (uint16_t*)B.kap, B.am, B.fr);
    else
      k_coarsen<uint16_t, uint16_t><<<nb, BS, 0, s>>>(A.d, (const uint16_t*)A.kap, A.am, A.fr,
                                                      B.d, (uint16_t*)B.kap, B.am, B.fr);
  }
  // diagonals need every level's free flags: computed after the coarsening
  for (int l = 0; l < nlev_; l++) {
    MgLevel& L = levels_[l];
    unsigned nb = blocks_for(L.n, BS);
    if (L.bf16) {
      k_level_diag<uint16_t><<<nb, BS, 0, s>>>(view<uint16_t>(L), L.fr, L.dg);
      k_sync_free<uint16_t><<<nb, BS, 0, s>>>(L.n, L.dg, L.fr, (uint16_t*)L.kap);
    } else {
      k_level_diag<float><<<nb, BS, 0, s>>>(view<float>(L), L.fr, L.dg);
      k_sync_free<float><<<nb, BS, 0, s>>>(L.n, L.dg, L.fr, (float*)L.kap);
    }
  }
  chk(hipGetLastError(), err);
  for (int l = 0; l < nlev_ && err == hipSuccess; l++) chk(build_tiles(levels_[l]), err);
  // Smoother weights and the Chebyshev interval use the Gershgorin bound
  // lambda_max(D^-1 A) <= 2 (each diagonal is at least the sum of its row's
  // off-diagonal magnitudes).  A bound, unlike an eigenvalue estimate, keeps
  // every smoother contracting and the V-cycle positive definite.
  for (int l = 0; l < nlev_; l++) {
    MgLevel& L = levels_[l];
    L.lmax = 2.0f;
    // damped Jacobi = degree-1 Chebyshev on [0.1 lmax, lmax]: omega = 1/theta
    L.omega = 1.0f / (0.55f * L.lmax);
  }
  {
    MgLevel& C = levels_[nlev_ - 1];
    double lmax = C.lmax, lmin = C.lmax / 30.0;
    double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), sg = th / de;
    std::vector<float> cf(2 * 64, 0.f);
    int st = std::min(coarse_sweeps_, 63);
    double rho = 1.0 / sg;
    cf[1] = (float)(1.0 / th);
    for (int k = 1; k < st; k++) {
      double rn = 1.0 / (2.0 * sg - rho);
      cf[2 * k] = (float)(rn * rho);
      cf[2 * k + 1] = (float)(2.0 * rn / de);
      rho = rn;
    }
    coef_host_ = cf;
    chk(hipMemcpyAsync(coef_, cf.data(), sizeof(float) * cf.size(), hipMemcpyHostToDevice, s),
        err);
  }
  chk(hipStreamSynchronize(s), err);
  return err;
}

hipError_t MgSolver::update(const uint8_t* type, const float* kappa) {
  type_ = type;
  kappa_ = kappa;
  return build();
}

// ── launch helpers: dispatch on tile depth and kappa storage ───────────
#define PFSF_TZ_DISPATCH(L, CALL) \
  switch ((L).tiles.tz) {         \
    case 8: {                     \
      constexpr int TZV = 8;      \
      CALL;                       \
    } break;                      \
    case 4: {                     \
      constexpr int TZV = 4;      \
      CALL;                       \
    } break;                      \
    default: {                    \
      constexpr int TZV = 2;      \
      CALL;                       \
    } break;                      \
  }

template <typename KT>
static void launch_down_t(const MgLevel& A, const MgLevel& B, const float* r, PcgScalars* sc,
                          hipStream_t s) {
  PFSF_TZ_DISPATCH(A, (k_down<TZV, KT, false><<<A.tiles.nlaunch, BS, 0, s>>>(
                           view<KT>(A), tl(A), B.d, r, nullptr, nullptr, nullptr, nullptr, B.r,
                           sc, nullptr, 0u, nullptr)));
}

static void launch_down(const MgLevel& A, const MgLevel& B, const float* r, PcgScalars* sc,
                        hipStream_t s) {
  if (A.bf16)
    launch_down_t<uint16_t>(A, B, r, sc, s);
  else
    launch_down_t<float>(A, B, r, sc, s);
}

template <typename KT>
static void launch_up_t(const MgLevel& A, const MgLevel& B, const float* r, float* out,
                        PcgScalars* sc, hipStream_t s) {
  PFSF_TZ_DISPATCH(A, (k_up<TZV, KT, false><<<A.tiles.nlaunch, BS, 0, s>>>(
                           view<KT>(A), tl(A), B.d, r, B.e, out, sc, nullptr)));
}

static void launch_up(const MgLevel& A, const MgLevel& B, const float* r, float* out,
                      PcgScalars* sc, hipStream_t s) {
  if (A.bf16)
    launch_up_t<uint16_t>(A, B, r, out, sc, s);
  else
    launch_up_t<float>(A, B, r, out, sc, s);
}

// Coarse part of the V-cycle: down-sweeps from level `from` (whose residual
// is already in levels_[from].r), the coarsest solve, and up-sweeps back to
// level `from`.  On return levels_[from].e holds that level's correction.
// With `check`, the iteration's convergence test (r.r partials of the finest
// update in prr_) runs first: inside the one-work-group coarsest solve when
// it fits LDS, else as a finalize kernel.
hipError_t MgSolver::vcycle_below(int from, bool check) {
  hipStream_t s = stream_;
  const unsigned G = levels_[0].tiles.nlaunch;
  MgLevel& C = levels_[nlev_ - 1];
  const bool lds = C.n <= kMaxLdsCoarse && padded_cells(C.d) <= kMaxLdsPadded;
  if (check && !lds) k_finalize<<<1, BS, 0, s>>>(FIN_UPDATE, prr_, G, sc_, ir_);
  for (int l = from; l < nlev_ - 1; l++)
    launch_down(levels_[l], levels_[l + 1], levels_[l].r, sc_, s);
  int st = std::min(coarse_sweeps_, 63);
  const unsigned rr_n = (check && lds) ? G : 0u;
  if (lds) {
    if (C.bf16)
      k_cheb_lds<uint16_t><<<1, kMaxLdsCoarse, 0, s>>>(view<uint16_t>(C), C.r, C.e, coef_, st,
                                                        sc_, prr_, rr_n);
    else
      k_cheb_lds<float><<<1, kMaxLdsCoarse, 0, s>>>(view<float>(C), C.r, C.e, coef_, st, sc_,
                                                     prr_, rr_n);
  } else {
    float* xin = C.xb;
    float* xout = C.e;
    if (st % 2 == 0) std::swap(xin, xout);  // the last step writes C.e
    for (int k = 0; k < st; k++) {
      float c1 = coef_host_[2 * k], c2 = coef_host_[2 * k + 1];
      if (C.bf16) {
        PFSF_TZ_DISPATCH(C, (k_cheb_tiled<TZV, uint16_t><<<C.tiles.nlaunch, BS, 0, s>>>(
                                 view<uint16_t>(C), tl(C), C.r, xin, xout, C.dv, c1, c2, k == 0,
                                 sc_)));
      } else {
        PFSF_TZ_DISPATCH(C, (k_cheb_tiled<TZV, float><<<C.tiles.nlaunch, BS, 0, s>>>(
                                 view<float>(C), tl(C), C.r, xin, xout, C.dv, c1, c2, k == 0,
                                 sc_)));
      }
      std::swap(xin, xout);
    }
  }
  for (int l = nlev_ - 2; l >= from; l--)
    launch_up(levels_[l], levels_[l + 1], levels_[l].r, levels_[l].e, sc_, s);
  return hipGetLastError();
}

// z = M r_[0] and r.z at the start of a solve pass.
hipError_t MgSolver::precondition() {
  hipStream_t s = stream_;
  MgLevel& F = levels_[0];
  if (nlev_ == 1) {
    // r.z goes where the Jacobi update leaves it, behind its r.r partials
    unsigned nb = blocks_for(n0_, BS);
    k_jacobi_apply<<<nb, BS, 0, s>>>(n0_, r_[0], F.dg, z_, prr_ + nb);
    return hipGetLastError();
  }
  MgLevel& B = levels_[1];
  launch_down(F, B, r_[0], sc_, s);
  (void)vcycle_below(1, false);
  PFSF_TZ_DISPATCH(F, (k_up<TZV, float, true><<<F.tiles.nlaunch, BS, 0, s>>>(
                          view<float>(F), tl(F), B.d, r_[0], B.e, z_, sc_, prz_)));
  return hipGetLastError();
}

// One PCG iteration reading r_[c], p_[c] and writing r_[1-c], p_[1-c]; the
// FP64 iterate is updated in place.
hipError_t MgSolver::iteration(int c) {
  hipStream_t s = stream_;
  MgLevel& F = levels_[0];
  const unsigned G = F.tiles.nlaunch;
  if (nlev_ == 1) {
    const unsigned nb = blocks_for(n0_, BS);
    PFSF_TZ_DISPATCH(F, (k_spmv_dir<TZV, float><<<G, BS, 0, s>>>(
                            view<float>(F), tl(F), z_, p_[c], p_[1 - c], q_, sc_, prr_ + nb, nb,
                            c, ppq_)));
    k_finalize<<<1, BS, 0, s>>>(FIN_ALPHA, ppq_, G, sc_, ir_);
    k_update_jacobi<<<nb, BS, 0, s>>>(n0_, F.dg, x_, r_[c], r_[1 - c], p_[1 - c], q_, z_, sc_,
                                      prr_);
    k_finalize<<<1, BS, 0, s>>>(FIN_UPDATE, prr_, nb, sc_, ir_);
    return hipGetLastError();
  }
  MgLevel& B = levels_[1];
  PFSF_TZ_DISPATCH(F, (k_spmv_dir<TZV, float><<<G, BS, 0, s>>>(
                          view<float>(F), tl(F), z_, p_[c], p_[1 - c], q_, sc_, prz_, G, c,
                          ppq_)));
  PFSF_TZ_DISPATCH(F, (k_down<TZV, float, true><<<G, BS, 0, s>>>(
                          view<float>(F), tl(F), B.d, r_[c], r_[1 - c], q_, p_[1 - c], x_, B.r,
                          sc_, ppq_, G, prr_)));
  (void)vcycle_below(1, true);
  PFSF_TZ_DISPATCH(F, (k_up<TZV, float, true><<<G, BS, 0, s>>>(
                          view<float>(F), tl(F), B.d, r_[1 - c], B.e, z_, sc_, prz_)));
  return hipGetLastError();
}

// 2*pairs PCG iterations.  The pair (ping, pong) is captured once into a HIP
// graph and replayed; the capture holds kernel launches only.
hipError_t MgSolver::chunk(int pairs) {
  hipError_t err = hipSuccess;
  if (use_graph_ && !graph_exec_) {
    if (!chk(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), err)) {
      (void)iteration(0);
      (void)iteration(1);
      hipGraph_t g = nullptr;
      chk(hipStreamEndCapture(stream_, &g), err);
      if (err == hipSuccess && !chk(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0), err))
        graph_ = g;
      else if (g)
        (void)hipGraphDestroy(g);
    }
    if (err != hipSuccess) {  // fall back to direct launches
      (void)hipGetLastError();
      use_graph_ = false;
      graph_exec_ = nullptr;
      err = hipSuccess;
    }
  }
  for (int k = 0; k < pairs && err == hipSuccess; k++) {
    if (use_graph_) {
      chk(hipGraphLaunch(graph_exec_, stream_), err);
    } else {
      chk(iteration(0), err);
      chk(iteration(1), err);
    }
  }
  return err;
}

// True FP64 residual of x_ into r_[0]; returns ||r|| / ||b|| (synchronises).
hipError_t MgSolver::true_residual(const float* b, double* rel) {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  unsigned nb0 = blocks_for(n0_, BS);
  k_true_residual<<<nb0, BS, 0, s>>>(d0_, type_, kappa_, levels_[0].fr, b, x_, r_[0], partials_);
  k_finalize<<<1, BS, 0, s>>>(FIN_IR, partials_, nb0, sc_, ir_);
  IrScalars* irh = (IrScalars*)(sc_host_ + 1);
  chk(hipMemcpyAsync(irh, ir_, sizeof(IrScalars), hipMemcpyDeviceToHost, s), err);
  chk(hipStreamSynchronize(s), err);
  *rel = irh->rel;
  return err;
}

// Solve passes: PCG on FP32 vectors with the iterate accumulated in FP64,
// started from the true FP64 residual; a pass ends when its recursive
// residual falls below half the target (or the FP32 resolution limit), then
// the true residual decides whether a further pass (a restart) is needed.
hipError_t MgSolver::solve(const float* b, double* x64, float* phi32, pfsf_solve_info* info) {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  hipStream_t caller = (hipStream_t)desc_.stream;
  const double tol = desc_.tol > 0 ? desc_.tol : 1e-6;
  const int max_it = desc_.max_iters > 0 ? desc_.max_iters : 1000;
  unsigned nb0 = blocks_for(n0_, BS);
  MgLevel& F = levels_[0];
  chk(hipEventRecord(ev0_, caller), err);  // order after the caller's stream
  chk(hipStreamWaitEvent(s, ev0_, 0), err);
  chk(hipEventRecord(ev0_, s), err);
  if (x64) {
    k_copy64<<<nb0, BS, 0, s>>>(n0_, x64, x_);
    k_mask64<<<nb0, BS, 0, s>>>(n0_, F.fr, x_);
  } else {
    k_zero64<<<nb0, BS, 0, s>>>(n0_, x_);
  }
  k_bnorm<<<nb0, BS, 0, s>>>(n0_, F.fr, b, partials_);
  k_finalize<<<1, BS, 0, s>>>(FIN_BNORM, partials_, nb0, sc_, ir_);
  double rel = 0.0;
  chk(true_residual(b, &rel), err);
  int passes = 0, total = 0;
  const int pairs_per_check = 4;  // 8 iterations between host checks
  for (; rel > tol && total < max_it && err == hipSuccess && passes < 16;) {
    double stop_rel = std::max(0.5 * tol / rel, 2e-6);
    k_pcg_begin<<<1, 1, 0, s>>>(sc_, ir_, stop_rel, max_it - total, 1);
    chk(precondition(), err);
    int done = 0;
    // the device caps the pass at max_it - total iterations; the number of
    // host checks is capped independently
    const int max_checks = (max_it - total) / (2 * pairs_per_check) + 2;
    for (int c = 0; !done && err == hipSuccess && c < max_checks; c++) {
      chk(chunk(pairs_per_check), err);
      chk(hipMemcpyAsync(sc_host_, sc_, sizeof(PcgScalars), hipMemcpyDeviceToHost, s), err);
      chk(hipStreamSynchronize(s), err);
      done = sc_host_->done;
    }
    const int it = sc_host_->iters;
    total += it;
    passes++;
    double nrel = rel;
    chk(true_residual(b, &nrel), err);
    bool stalled = (it == 0) || (sc_host_->done == 2 && nrel >= rel);
    rel = nrel;
    if (stalled) break;
  }
  if (x64) k_copy64<<<nb0, BS, 0, s>>>(n0_, x_, x64);
  if (phi32) k_to_f32<<<nb0, BS, 0, s>>>(n0_, x_, phi32);
  chk(hipEventRecord(ev1_, s), err);
  chk(hipEventSynchronize(ev1_), err);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  chk(hipStreamWaitEvent(caller, ev1_, 0), err);
  if (info) {
    info->iterations = total;
    info->refinements = passes;
    info->converged = rel <= tol ? 1 : 0;
    info->levels = nlev_;
    info->rel_residual = rel;
    info->solve_ms = ms;
  }
  return err;
}

// Average device time of one launch of a solver kernel over `reps` launches
// on the solver stream.  which: 0 finest SpMV + direction + p.q, 1 finest
// update + down-sweep (or the Jacobi update), 2 second-level down-sweep,
// 3 finest up-sweep + r.z: each launch is bracketed by the kernel's own
// start/stop timestamps (hipExtLaunchKernel events), so the figure is the
// kernel duration alone.  4: one whole PCG iteration, timed between events
// around `reps` back-to-back iterations (launch gaps included).  Call after a
// solve so the vectors hold finite data.
hipError_t MgSolver::profile(int which, int reps, float* avg_ms) {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  MgLevel& F = levels_[0];
  const unsigned G = F.tiles.nlaunch;
  if (which < 0 || which > 4 || reps <= 0) return hipErrorInvalidValue;
  if ((which == 2 || which == 3) && nlev_ < 2) return hipErrorInvalidValue;
  if (which == 2 && nlev_ < 3) return hipErrorInvalidValue;
  k_profile_begin<<<1, 1, 0, s>>>(sc_);
  if (which == 4) {
    (void)iteration(0);  // warm
    chk(hipEventRecord(ev0_, s), err);
    for (int k = 0; k < reps; k++) (void)iteration(0);
    chk(hipEventRecord(ev1_, s), err);
    chk(hipEventSynchronize(ev1_), err);
    float ms = 0.f;
    chk(hipEventElapsedTime(&ms, ev0_, ev1_), err);
    *avg_ms = ms / (float)reps;
    return err;
  }
  auto ext = [&](auto kern, unsigned grid, hipEvent_t a, hipEvent_t b, auto... args) {
    hipExtLaunchKernelGGL(kern, dim3(grid), dim3(BS), 0, s, a, b, 0, args...);
  };
  auto once = [&](hipEvent_t a, hipEvent_t b) {
    switch (which) {
      case 0:
        PFSF_TZ_DISPATCH(F, ext(k_spmv_dir<TZV, float>, G, a, b, view<float>(F), tl(F),
                                (const float*)z_, (const float*)p_[0], p_[1], q_, sc_,
                                (const double*)prz_, G, 0, ppq_));
        break;
      case 1:
        if (nlev_ == 1) {
          const unsigned nb = blocks_for(n0_, BS);
          ext(k_update_jacobi, nb, a, b, n0_, (const float*)F.dg, x_, (const float*)r_[0], r_[1],
              (const float*)p_[1], (const float*)q_, z_, (const PcgScalars*)sc_, prr_);
        } else {
          PFSF_TZ_DISPATCH(F, ext(k_down<TZV, float, true>, G, a, b, view<float>(F), tl(F),
                                  levels_[1].d, (const float*)r_[0], r_[1], (const float*)q_,
                                  (const float*)p_[1], x_, levels_[1].r, sc_,
                                  (const double*)ppq_, G, prr_));
        }
        break;
      case 2: {
        const MgLevel& A = levels_[1];
        const MgLevel& B = levels_[2];
        if (A.bf16) {
          PFSF_TZ_DISPATCH(A, ext(k_down<TZV, uint16_t, false>, A.tiles.nlaunch, a, b,
                                  view<uint16_t>(A), tl(A), B.d, (const float*)A.r,
                                  (float*)nullptr, (const float*)nullptr, (const float*)nullptr,
                                  (double*)nullptr, B.r, sc_, (const double*)nullptr, 0u,
                                  (double*)nullptr));
        } else {
          PFSF_TZ_DISPATCH(A, ext(k_down<TZV, float, false>, A.tiles.nlaunch, a, b,
                                  view<float>(A), tl(A), B.d, (const float*)A.r,
                                  (float*)nullptr, (const float*)nullptr, (const float*)nullptr,
                                  (double*)nullptr, B.r, sc_, (const double*)nullptr, 0u,
                                  (double*)nullptr));
        }
        break;
      }
      default:
        PFSF_TZ_DISPATCH(F, ext(k_up<TZV, float, true>, G, a, b, view<float>(F), tl(F),
                                levels_[1].d, (const float*)r_[1], (const float*)levels_[1].e,
                                z_, (const PcgScalars*)sc_, prz_));
    }
  };
  once(nullptr, nullptr);  // warm
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  for (auto& e : ev) chk(hipEventCreate(&e), err);
  if (err == hipSuccess) {
    for (int k = 0; k < reps; k++) once(ev[2 * k], ev[2 * k + 1]);
    chk(hipGetLastError(), err);
    chk(hipEventSynchronize(ev.back()), err);
  }
  double total = 0.0;
  for (int k = 0; k < reps && err == hipSuccess; k++) {
    float ms = 0.f;
    chk(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]), err);
    total += ms;
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  *avg_ms = (float)(total / reps);
  return err;
}

}  // namespace pfsf
