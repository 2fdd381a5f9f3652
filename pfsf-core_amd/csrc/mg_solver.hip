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
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include <hip/hip_ext.h>

#include "mg_kernels.h"
#include "mg_solver.h"
#include "mg_stream.h"

namespace pfsf {
namespace {

// ── gather form of the operator (setup) ────────────────────────────────
// sum over free neighbours (within the coupling mask mk) of w_ij * v(j).
template <typename KT, typename F>
__device__ __forceinline__ float offsum(const Lv<KT>& L, int x, int y, int z, float ki,
                                        uint32_t mk, F v) {
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
        if (cls == 0 || !((mk >> nbit(dx, dy, dz)) & 1u)) continue;
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

// One-work-group reductions that remain separate kernels: the Jacobi
// (single-level) iteration, a coarsest level too large for LDS, and the
// FP64 residual norm of the refinement passes.
enum FinMode {
  FIN_ALPHA = 0,  // p.q -> alpha = rz / pq
  FIN_UPDATE,     // r.r -> rr, iteration count, convergence
  FIN_IR,         // r.r -> rr, rel
};

__global__ void __launch_bounds__(BS)
    k_finalize(int mode, const double* __restrict__ partials, unsigned nblk, PcgScalars* sc,
               IrScalars* ir, unsigned long long* mirror) {
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
      if (mirror) publish_progress(mirror, sc->iters, sc->done, v);
      break;
    case FIN_IR:
      ir->rr = v;
      ir->rel = ir->bb > 0.0 ? sqrt(v / ir->bb) : sqrt(v);
      break;
  }
}

// ── single-system launch wrappers ──────────────────────────────────────
// Work-group g takes tile M.ids[g]; work-group 0 publishes the PCG scalars
// and partials go to partials[g].
template <int TZ, typename KT>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    k_spmv_dir(Lv<KT> L, TL M, const float* __restrict__ z, const float* __restrict__ pin,
               float* __restrict__ pout, float* __restrict__ q, PcgScalars* sc,
               const double* __restrict__ rz_part, unsigned rz_n, int slot, double* partials) {
  int x0 = 0, y0 = 0, z0 = 0;
  const bool have = tile_origin<TZ>(M, x0, y0, z0);
  spmv_body<TZ, KT>(L, have, x0, y0, z0, z, pin, pout, q, sc, rz_part, rz_n, slot,
                    blockIdx.x == 0, partials, blockIdx.x);
}

// MASK: the level carries a coupling mask (every level below the finest).
// With rr_n > 0 the launch has one work-group more than tiles, and that last
// work-group runs the PCG iteration's convergence test from the r.r
// partials of the finest update (mirrored to the host words) while the
// others sweep their tiles: the test leaves the critical path of the
// coarsest solve, and every later launch of the iteration reads the stop.
template <int TZ, typename KT, bool UPDATE, bool MASK = false>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    k_down(Lv<KT> L, TL M, pfsf_dims c, const float* __restrict__ rin, float* __restrict__ rout,
           const float* __restrict__ q, const float* __restrict__ p, double* __restrict__ x,
           float* __restrict__ rc, PcgScalars* sc, const double* __restrict__ pq_part,
           unsigned pq_n, double* partials, const double* __restrict__ rr_part = nullptr,
           unsigned rr_n = 0u, unsigned long long* mirror = nullptr) {
  pin_args(L, M, c, rin, rout, q, p, x, rc, sc, pq_part, pq_n, partials, rr_part, rr_n, mirror);
  if (rr_n > 0 && blockIdx.x == gridDim.x - 1) {
    if (!sc->done) (void)convergence_test(sc, rr_part, rr_n, mirror);
    return;
  }
  int x0 = 0, y0 = 0, z0 = 0;
  const bool have = tile_origin<TZ>(M, x0, y0, z0);
  down_body<TZ, KT, UPDATE, MASK>(L, have, x0, y0, z0, c, rin, rout, q, p, x, rc, sc, pq_part,
                                  pq_n, blockIdx.x == 0, partials, blockIdx.x);
}

template <int TZ, typename KT, bool FINEST, bool MASK = false>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    k_up(Lv<KT> L, TL M, pfsf_dims c, const float* __restrict__ r, const float* __restrict__ ec,
         float* __restrict__ out, const PcgScalars* sc, double* partials) {
  pin_args(L, M, c, r, ec, out, sc, partials);
  int x0 = 0, y0 = 0, z0 = 0;
  const bool have = tile_origin<TZ>(M, x0, y0, z0);
  up_body<TZ, KT, FINEST, MASK>(L, have, x0, y0, z0, c, r, ec, out, sc, partials, blockIdx.x);
}

template <typename KT>
__global__ void __launch_bounds__(kMaxLdsCoarse)
    k_cheb_lds(Lv<KT> L, const float* __restrict__ r, float* __restrict__ out,
               const float* __restrict__ coef, int steps, PcgScalars* sc,
               const double* __restrict__ rr_part, unsigned rr_n, unsigned long long* mirror) {
  pin_args(L, r, out, coef, steps, sc, rr_part, rr_n, mirror);
  cheb_lds_body<KT>(L, r, out, coef, steps, sc, rr_part, rr_n, mirror);
}

// ── coarsest level as a dense matrix ──────────────────────────────────
// The coarsest solve is a fixed polynomial in the coarsest operator: `steps`
// Chebyshev steps from a zero guess give x = P r with P = p(D^-1 A) D^-1,
// symmetric in exact arithmetic.  P is formed once per operator (column j =
// the steps applied to the unit vector of cell j), symmetrised, and applied
// each V-cycle as a matrix-vector product spread over many work-groups, in
// place of a one-work-group chain of `steps` dependent stencil passes.  Rows
// are padded to ld (a multiple of 256) with zeros; cells outside the free
// set have zero rows and columns.

// Column j of P, stored as row j (blockIdx.x = j; kMaxLdsCoarse threads).
template <typename KT>
__global__ void __launch_bounds__(kMaxLdsCoarse)
    k_cheb_columns(Lv<KT> L, const float* __restrict__ coef, int steps, float* __restrict__ pm,
                   int ld) {
  const pfsf_dims d = L.d;
  const int n = d.lx * d.ly * d.lz;
  const int j = blockIdx.x, i = threadIdx.x;
  if (ldk(L.kap, j) < 0.f) return;  // outside the free set: a zero column (pre-filled)
  const bool act = i < n;
  const int ic = act ? i : 0;
  const float k_ld = ldk(L.kap, ic);
  const float D_ld = L.dg[ic], m_ld = L.am[ic];
  const uint32_t mk_ld = L.msk ? L.msk[ic] : kAllCouplings;
  const float c_ld = i < 2 * steps ? coef[i] : 0.f;
  const float xi = cheb_lds_run(d, L.scale, act, k_ld, D_ld, m_ld, mk_ld, i == j ? 1.f : 0.f,
                                c_ld, steps);
  if (act) pm[(int64_t)j * ld + i] = xi;
}

// P <- (P + P^T) / 2: thread (i, j) with i < j owns both entries.
__global__ void __launch_bounds__(BS) k_sym_dense(float* __restrict__ pm, int n, int ld) {
  const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  if (i >= j) return;
  const float a = pm[(int64_t)i * ld + j], b = pm[(int64_t)j * ld + i];
  const float v = 0.5f * (a + b);
  pm[(int64_t)i * ld + j] = v;
  pm[(int64_t)j * ld + i] = v;
}

// out = P r: one row per wave, NP passes of 64 lanes x 4 columns (ld = 256 NP).
// With rr_n > 0, work-group 0 first runs the PCG iteration's convergence test
// (as the one-work-group coarsest solve does); the other work-groups form
// their rows regardless, and the up-sweeps behind this launch read the stop.
template <int NP>
__global__ void __launch_bounds__(BS)
    k_cheb_dense(const float* __restrict__ pm, int n, const float* __restrict__ r,
                 float* __restrict__ out, PcgScalars* sc, const double* __restrict__ rr_part,
                 unsigned rr_n, unsigned long long* mirror) {
  pin_args(pm, n, r, out, sc, rr_part, rr_n, mirror);
  constexpr int ld = 256 * NP;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  const int rw = row < n ? row : 0;  // row 0 stands in for a row past the end
  float4 a[NP];
  float rv[NP][4];
#pragma unroll
  for (int p = 0; p < NP; p++) {
    const int k = p * 256 + lane * 4;
    a[p] = *reinterpret_cast<const float4*>(pm + (int64_t)rw * ld + k);
#pragma unroll
    for (int c = 0; c < 4; c++) rv[p][c] = r[min(k + c, n - 1)];  // zero-padded columns
  }
  if (sc->done) return;
  if (rr_n > 0 && blockIdx.x == 0 && convergence_test(sc, rr_part, rr_n, mirror)) return;
  float acc = 0.f;
#pragma unroll
  for (int p = 0; p < NP; p++) {
    const int k = p * 256 + lane * 4;
    acc += a[p].x * (k < n ? rv[p][0] : 0.f);
    acc += a[p].y * (k + 1 < n ? rv[p][1] : 0.f);
    acc += a[p].z * (k + 2 < n ? rv[p][2] : 0.f);
    acc += a[p].w * (k + 3 < n ? rv[p][3] : 0.f);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0 && row < n) out[row] = acc;
}

// ── down-sweep folded into the dense coarsest solve ────────────────────
// The coarsest rhs is r_c = R r_b with R = P^T (I - A S) the down-sweep of
// the level b above (S = omega D^-1: pre-smoothing from a zero guess, then
// the restricted residual), so the coarsest correction is e = P R r_b =
// G r_b.  G is formed once per operator; each V-cycle then reads r_b and
// skips the level-b down-sweep launch.  Column j of R, for a free cell j,
// holds 1 - omega at parent(j) and omega w_kj / D_j at parent(k) for every
// coupled free neighbour k (w as in stencil_apply), summed into the at most
// 8 parents of j's neighbourhood (slot = per-axis offset from
// ((x - 1) >> 1, (y - 1) >> 1, (z - 1) >> 1)).
template <typename KT>
__global__ void __launch_bounds__(BS)
    k_fold_columns(Lv<KT> L, const uint8_t* __restrict__ fr, pfsf_dims c, int* __restrict__ par,
                   float* __restrict__ val) {
  const pfsf_dims d = L.d;
  const int64_t j = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (j >= (int64_t)d.lx * d.ly * d.lz) return;
  int x, y, z;
  unflat(j, d, x, y, z);
  const int bx = (x - 1) >> 1, by = (y - 1) >> 1, bz = (z - 1) >> 1;
  auto slot = [&](int px, int py, int pz) { return (px - bx) + 2 * (py - by) + 4 * (pz - bz); };
  float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (fr[j]) {
    const float cj = L.omega / L.dg[j];
    v[slot(x >> 1, y >> 1, z >> 1)] += 1.f - L.omega;
    const float kj = ldk(L.kap, j), sj = sqrtf(kj);
    const uint32_t mk = L.msk ? L.msk[j] : kAllCouplings;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int cls = (dx != 0) + (dy != 0) + (dz != 0);
          const int nx = x + dx, ny = y + dy, nz = z + dz;
          if (cls == 0 || nx < 0 || ny < 0 || nz < 0 || nx >= d.lx || ny >= d.ly || nz >= d.lz)
            continue;
          if (!((mk >> nbit(dx, dy, dz)) & 1u)) continue;
          const int64_t k = flat(nx, ny, nz, d);
          if (!fr[k]) continue;
          const float kk = ldk(L.kap, k);
          const float w = cls == 1 ? L.scale * 0.5f * (kj + kk)
                                   : L.scale * sj * (cls == 2 ? kEdgeP : kCornerP) * sqrtf(kk);
          v[slot(nx >> 1, ny >> 1, nz >> 1)] += w * cj;
        }
  }
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int px = bx + (q & 1), py = by + ((q >> 1) & 1), pz = bz + (q >> 2);
    const bool in = px >= 0 && py >= 0 && pz >= 0 && px < c.lx && py < c.ly && pz < c.lz;
    par[j * 8 + q] = in && v[q] != 0.f ? (int)flat(px, py, pz, c) : -1;
    val[j * 8 + q] = v[q];
  }
}

// G[i][j] = sum over j's parent slots of P[p][i] val (P symmetric).
__global__ void __launch_bounds__(BS)
    k_fold_matrix(const float* __restrict__ pm, int ldp, int nc, const int* __restrict__ par,
                  const float* __restrict__ val, int nb, float* __restrict__ g, int ldg) {
  const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (t >= (int64_t)nc * nb) return;
  const int i = (int)(t / nb), j = (int)(t % nb);
  float a = 0.f;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int p = par[j * 8 + q];
    if (p >= 0) a += pm[(int64_t)p * ldp + i] * val[j * 8 + q];
  }
  g[(int64_t)i * ldg + j] = a;
}

// out = G r_b: r_b staged in LDS, one row of G per wave, NP passes of 64
// lanes x 4 columns (ld = 256 NP, padded with zeros).  With rr_n > 0,
// work-group 0 first runs the PCG iteration's convergence test.
template <int NP>
__global__ void __launch_bounds__(BS)
    k_fold_dense(const float* __restrict__ g, int nc, int nb, const float* __restrict__ rb,
                 float* __restrict__ out, PcgScalars* sc, const double* __restrict__ rr_part,
                 unsigned rr_n, unsigned long long* mirror) {
  pin_args(g, nc, nb, rb, out, sc, rr_part, rr_n, mirror);
  constexpr int ld = 256 * NP;
  __shared__ float rs[ld];
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  const int rw = row < nc ? row : 0;  // row 0 stands in for a row past the end
  float rv[ld / BS];
#pragma unroll
  for (int k = 0; k < ld / BS; k++) {
    const int e = k * BS + (int)threadIdx.x;
    rv[k] = rb[min(e, nb - 1)];
  }
  float4 a[NP];
#pragma unroll
  for (int p = 0; p < NP; p++)
    a[p] = *reinterpret_cast<const float4*>(g + (int64_t)rw * ld + p * 256 + lane * 4);
  if (sc->done) return;
  if (rr_n > 0 && blockIdx.x == 0 && convergence_test(sc, rr_part, rr_n, mirror)) return;
#pragma unroll
  for (int k = 0; k < ld / BS; k++) {
    const int e = k * BS + (int)threadIdx.x;
    rs[e] = e < nb ? rv[k] : 0.f;
  }
  __syncthreads();
  float acc = 0.f;
#pragma unroll
  for (int p = 0; p < NP; p++) {
    const float4 r4 = *reinterpret_cast<const float4*>(rs + p * 256 + lane * 4);
    acc += a[p].x * r4.x;
    acc += a[p].y * r4.y;
    acc += a[p].z * r4.z;
    acc += a[p].w * r4.w;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0 && row < nc) out[row] = acc;
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
    if (!is_free_s(T.s[cc])) {
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
      float ax = tile_apply<TZ, true>(T, cc, L.am[i], L.scale, L.msk ? L.msk[i] : kAllCouplings);
      float di = c1 * dv[i] + c2 * (r[i] - ax) * rcp(D);
      dv[i] = di;
      xout[i] = xi + di;
    }
  }
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

// Coupling mask of a coarse level: bit nbit(D) of coarse cell I is set when
// some free child i of I is coupled at the finer level (free neighbour j,
// within the finer level's own mask) to a child j of the coarse neighbour
// I + D.  Rediscretised couplings are kept only where the finer level has a
// coupling, so structures separated by air at the finer level stay separate
// (the relation is symmetric, so the operator stays symmetric).
__global__ void __launch_bounds__(BS)
    k_coarse_mask(pfsf_dims f, const uint8_t* __restrict__ freef, const uint32_t* __restrict__ mf,
                  pfsf_dims c, const uint8_t* __restrict__ freec, uint32_t* __restrict__ mc) {
  int64_t ci = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (ci >= nvox(c)) return;
  uint32_t m = 0;
  if (freec[ci]) {
    int X, Y, Z;
    unflat(ci, c, X, Y, Z);
    for (int cz = 0; cz < 2; cz++)
      for (int cy = 0; cy < 2; cy++)
        for (int cx = 0; cx < 2; cx++) {
          const int x = 2 * X + cx, y = 2 * Y + cy, z = 2 * Z + cz;
          if (x >= f.lx || y >= f.ly || z >= f.lz) continue;
          const int64_t i = flat(x, y, z, f);
          if (!freef[i]) continue;
          const uint32_t mi = mf ? mf[i] : kAllCouplings;
          for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
              for (int dx = -1; dx <= 1; dx++) {
                const int nx = x + dx, ny = y + dy, nz = z + dz;
                if ((dx | dy | dz) == 0 || !((mi >> nbit(dx, dy, dz)) & 1u)) continue;
                if (nx < 0 || nx >= f.lx || ny < 0 || ny >= f.ly || nz < 0 || nz >= f.lz)
                  continue;
                if (!freef[flat(nx, ny, nz, f)]) continue;
                const int DX = (nx >> 1) - X, DY = (ny >> 1) - Y, DZ = (nz >> 1) - Z;
                if ((DX | DY | DZ) != 0) m |= 1u << nbit(DX, DY, DZ);
              }
        }
  }
  mc[ci] = m;
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
  const uint32_t mk = L.msk ? L.msk[i] : kAllCouplings;
  float s = offsum(L, x, y, z, ldk(L.kap, i), mk, [&](int64_t) { return 1.0f; });
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

// flag = 1 where a level's free set differs from its copy before.
__global__ void __launch_bounds__(BS)
    k_fr_moved(int64_t n, const uint8_t* __restrict__ before, const uint8_t* __restrict__ fr,
               int* __restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n && before[i] != fr[i]) *flag = 1;
}

// Anchor-coded kappa of the finest level (MgLevel::ks) from the final free
// set: free kappa as is, anchors -max(kappa, FLT_MIN) (an anchor of kappa 0
// still couples through its face weight), everything else -inf.
__global__ void __launch_bounds__(BS)
    k_stream_kappa(int64_t n, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                   const uint8_t* __restrict__ fr, const float* __restrict__ kap0,
                   float* __restrict__ ks) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  float v = -INFINITY;
  if (fr[i])
    v = kap0[i];
  else if (type[i] == 2)
    v = -fmaxf(kapf(kappa[i]), FLT_MIN);
  ks[i] = v;
}

// Cell bits of every listed column (see kCellWords): one work-group per list
// slot, one thread per word.  A cell is free where fr is set and couples
// where it is free or an anchor (the finest level's anchor-coded kappa);
// cells outside the grid have no bits.
__global__ void __launch_bounds__(BS)
    k_cell_bits(pfsf_dims d, const uint8_t* __restrict__ type, const uint8_t* __restrict__ fr,
                const int* __restrict__ ids, int ntx, int nty, int tz,
                uint32_t* __restrict__ cells) {
  const int g = blockIdx.x;
  const int t = ids[g];
  for (int w = threadIdx.x; w < kCellWords; w += BS) {
    uint32_t bits = 0;
    if (t >= 0) {
      const int r = t / ntx;
      const int x = (t % ntx) * TX - 1 + w % HX, y = (r % nty) * TY - 1 + w / HX;
      const int z0 = (r / nty) * tz;
      if (x >= 0 && x < d.lx && y >= 0 && y < d.ly)
        for (int pl = 0; pl < tz + 2; pl++) {
          const int z = z0 - 1 + pl;
          if (z < 0 || z >= d.lz) continue;
          const int i = flat32(x, y, z, d);
          const bool f = fr[i] != 0, a = f || type[i] == 2;
          bits |= (f ? 1u << pl : 0u) | (a ? 1u << (16 + pl) : 0u);
        }
    }
    cells[(int64_t)g * kCellWords + w] = bits;
  }
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

// True residual of the FP64 iterate: r = b - A x with the exact coupling rule
// (FP64), written as r32 = (float) r, plus partial ||r||^2 per work-group
// (residual_stream_body): work-group g streams the column of list slot g.
template <int TZ>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    k_true_residual(pfsf_dims d, TL M, const uint8_t* __restrict__ type,
                    const float* __restrict__ kappa, const uint8_t* __restrict__ fr,
                    const float* __restrict__ b, const double* __restrict__ x,
                    float* __restrict__ r32, double* partials, const int* nz, int seq) {
  pin_args(d, M, type, kappa, fr, b, x, r32, partials, nz, seq);
  // nz holds the number of the last solve whose initial guess had a nonzero
  const bool xzero = nz && *nz != seq;
  residual_stream_body<TZ>(d, SlotColumn{M}, type, kappa, fr, b, x, r32, partials, blockIdx.x,
                           xzero);
}

// Solve set-up over all cells, one pass (grid-stride, at most kBeginBlocks
// work-groups): the FP64 iterate takes the caller's guess on the free set (0
// elsewhere, or everywhere without a guess), each work-group writes a ||b||^2
// partial over the free set and a flag telling whether it saw a nonzero
// guess (combined by k_begin_fin).
constexpr unsigned kBeginBlocks = 1024;

__global__ void __launch_bounds__(BS)
    k_begin(int64_t n, const double* __restrict__ xin, const uint8_t* __restrict__ fr,
            const float* __restrict__ b, double* __restrict__ x, int* __restrict__ gflag,
            double* partials) {
  __shared__ double sh[1][BS / 64];
  bool nonzero = false;
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    const bool f = fr[i] != 0;
    const double v = (xin && f) ? xin[i] : 0.0;
    x[i] = v;
    nonzero |= v != 0.0;
    if (f) acc[0] += (double)b[i] * (double)b[i];
  }
  const int any = __syncthreads_or(nonzero);
  if (threadIdx.x == 0) gflag[blockIdx.x] = any;
  write_partials<1>(acc, partials, sh);
}

// ||b||^2 from k_begin's partials, and *nz = seq (the number of this solve)
// when any of its blocks saw a nonzero initial guess.
__global__ void __launch_bounds__(BS)
    k_begin_fin(const double* __restrict__ partials, const int* __restrict__ gflag,
                unsigned nblk, IrScalars* ir, int* nz, int seq) {
  const double bb = block_total(partials, nblk);
  int any = 0;
  for (unsigned k = threadIdx.x; k < nblk; k += BS) any |= gflag[k];
  if (__syncthreads_or(any) && threadIdx.x == 0) *nz = seq;
  if (threadIdx.x == 0) ir->bb = bb;
}

// Solution copies: the FP64 iterate to the caller's FP64 and/or FP32 arrays.
__global__ void __launch_bounds__(BS)
    k_end(int64_t n, const double* __restrict__ x, double* __restrict__ x64,
          float* __restrict__ phi32) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  if (x64) x64[i] = v;
  if (phi32) phi32[i] = (float)v;
}

__global__ void __launch_bounds__(BS) k_zero64(int64_t n, double* __restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) v[i] = 0.0;
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
  sc->xa = 0.0;
  sc->rr = ir->rr;
  sc->stop2 = ir->rr * stop_rel * stop_rel;
  sc->max_iters = max_iters;
}

// The iterate update a pass leaves pending: when its last iteration (or the
// last one before a breakdown) was an even one, x += xa p (p = p_[1], that
// iteration's direction, zero off the free set), the FMA of a single update.
// A pass has run an odd number of convergence tests exactly then.
__global__ void __launch_bounds__(BS)
    k_x_pending(int64_t n, const PcgScalars* __restrict__ sc, const float* __restrict__ p,
                double* __restrict__ x) {
  if ((sc->iters & 1) == 0) return;
  const double a = sc->xa;
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) x[i] = fma(a, (double)p[i], x[i]);
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

void MgSolver::drop_graphs() {
  for (int k = 0; k < 6; k++) {
    if (graph_exec_[k]) (void)hipGraphExecDestroy(graph_exec_[k]);
    if (graph_[k]) (void)hipGraphDestroy(graph_[k]);
    graph_exec_[k] = nullptr;
    graph_[k] = nullptr;
  }
}

void MgSolver::release() {
  drop_graphs();
  if (stream_ && owns_stream_) (void)hipStreamSynchronize(stream_);
  for (void* p : owned_) (void)hipFree(p);
  owned_.clear();
  if (sc_host_) (void)hipHostFree(sc_host_);
  sc_host_ = nullptr;
  if (mirror_) (void)hipHostFree(mirror_);
  mirror_ = nullptr;
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  ev0_ = ev1_ = nullptr;
  if (stream_ && owns_stream_) (void)hipStreamDestroy(stream_);
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
  v.msk = l.msk;
  v.scale = l.scale;
  v.omega = l.omega;
  return v;
}

// The finest level as the plane-streaming kernels read it: the anchor-coded
// kappa, no anchor mass.
static Lv<float> sview(const MgLevel& l) {
  Lv<float> v = view<float>(l);
  v.kap = l.ks;
  v.am = nullptr;
  return v;
}

static TL tl(const MgLevel& l) {
  if (l.tiles.dense)
    return TL{l.tiles.ntx, l.tiles.nty, nullptr, l.tiles.ntx * l.tiles.nty * l.tiles.ntz,
              nullptr};
  return TL{l.tiles.ntx, l.tiles.nty, l.tiles.ids, l.tiles.nlaunch, l.tiles.cells};
}

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
    if (v == 16 || v == 8 || v == 4 || v == 2) return v;
  }
  // the finest level streams its columns plane by plane: deeper columns
  // load fewer halo planes, and about two columns per CU keep it busy
  const int64_t want = finest ? 512 : 2048;
  for (int tz : {8, 4}) {
    int64_t nt = (int64_t)((d.lx + TX - 1) / TX) * ((d.ly + TY - 1) / TY) * ((d.lz + tz - 1) / tz);
    if (nt >= want) return tz;
  }
  return 2;
}

// Level count and per-level shapes (host only).  Levels are added until the
// largest dimension is at most 8, never fewer than min_levels; tz_fine /
// tz_coarse fix the tile depth (0: chosen from the level's size).
void MgSolver::configure(const pfsf_solver_desc& desc, int min_levels, int tz_fine,
                         int tz_coarse) {
  desc_ = desc;
  d0_ = desc.dims;
  n0_ = nvox(d0_);
  type_ = desc.type;
  kappa_ = desc.kappa;
  bf16_ = desc.precision == PFSF_PREC_BF16_COARSE;
  coarse_sweeps_ = desc.coarse_sweeps > 0 ? desc.coarse_sweeps : kCoarseSweeps;
  int want = desc.mg_levels;
  if (want <= 0) {
    want = 1;
    pfsf_dims d = d0_;
    while (std::max(d.lx, std::max(d.ly, d.lz)) > 8 && want < 8) {
      d = {(d.lx + 1) / 2, (d.ly + 1) / 2, (d.lz + 1) / 2};
      want++;
    }
  }
  nlev_ = std::max(min_levels, std::min(want, 10));
  levels_.assign(nlev_, MgLevel{});
  pfsf_dims d = d0_;
  for (int l = 0; l < nlev_; l++) {
    MgLevel& L = levels_[l];
    L.d = d;
    L.n = nvox(d);
    L.bf16 = (l > 0) && bf16_;
    L.scale = (float)(1 << l);
    const int tzf = l == 0 ? tz_fine : tz_coarse;
    // masked coarse stencils need more registers: coarse tiles stay 2 deep
    L.tiles.tz = tzf > 0 ? tzf : (l == 0 ? pick_tz(d, true) : 2);
    L.tiles.ntx = (d.lx + TX - 1) / TX;
    L.tiles.nty = (d.ly + TY - 1) / TY;
    L.tiles.ntz = (d.lz + L.tiles.tz - 1) / L.tiles.tz;
    d = {(d.lx + 1) / 2, (d.ly + 1) / 2, (d.lz + 1) / 2};
  }
}

void* DevPool::take(size_t bytes) {
  const size_t off = (used + 255) & ~(size_t)255;
  used = off + std::max<size_t>(bytes, 1);
  return base ? base + off : nullptr;
}

template <typename T>
T* MgSolver::alloc(int64_t n, hipError_t& err) {
  if (pool_) return (T*)pool_->take((size_t)std::max<int64_t>(n, 1) * sizeof(T));
  return dalloc<T>(n, owned_, err);
}

// Device arrays.  `standalone` adds what only solve()/profile() use: the
// PCG scalars, partial sums and Chebyshev coefficients.
hipError_t MgSolver::allocate(bool standalone) {
  hipError_t err = hipSuccess;
  reach_ = alloc<uint8_t>(n0_, err);
  flag_ = alloc<int>(1, err);
  x_ = alloc<double>(n0_, err);
  for (int k = 0; k < 2; k++) {
    r_[k] = alloc<float>(n0_, err);
    p_[k] = alloc<float>(n0_, err);
  }
  q_ = alloc<float>(n0_, err);
  z_ = alloc<float>(n0_, err);
  int64_t maxblk = blocks_for(n0_, BS);
  for (int l = 0; l < nlev_; l++) {
    MgLevel& L = levels_[l];
    size_t ks = L.bf16 ? sizeof(uint16_t) : sizeof(float);
    L.kap = alloc<uint8_t>(L.n * (int64_t)ks, err);
    L.dg = alloc<float>(L.n, err);
    L.am = alloc<float>(L.n, err);
    if (l == 0 && nlev_ > 1) L.ks = alloc<float>(L.n, err);
    L.fr = alloc<uint8_t>(L.n, err);
    if (l > 0) {
      L.r = alloc<float>(L.n, err);
      L.e = alloc<float>(L.n, err);
      L.msk = alloc<uint32_t>(L.n, err);
    }
    if (l == nlev_ - 1 && nlev_ > 1) {
      L.xb = alloc<float>(L.n, err);
      L.dv = alloc<float>(L.n, err);
      if (dense_ok_ && L.n <= kMaxLdsCoarse && padded_cells(L.d) <= kMaxLdsPadded) {
        L.pm_ld = (int)((L.n + 255) / 256 * 256);
        L.pm = alloc<float>(L.n * L.pm_ld, err);
        const MgLevel& B = levels_[l - 1];
        if (nlev_ >= 4 && B.n <= kMaxFoldCells) {
          L.fg_ld = fold_ld(B.n);
          L.fg = alloc<float>(L.n * L.fg_ld, err);
          L.fpar = alloc<int>(8 * B.n, err);
          L.fval = alloc<float>(8 * B.n, err);
        }
      }
    }
    int64_t nt = (int64_t)L.tiles.ntx * L.tiles.nty * L.tiles.ntz;
    L.tiles.ids = alloc<int>(nt + 8, err);
    // cell bits of the streaming columns (standalone solvers only: a batch
    // streams its columns through its own tile records)
    if (l == 0 && nlev_ > 1 && standalone && L.tiles.tz + 2 <= kCellMaxPlanes)
      L.tiles.cells = alloc<uint32_t>((nt + 8) * kCellWords, err);
    maxblk = std::max(maxblk, nt + 8);
  }
  if (standalone) {
    sc_ = alloc<PcgScalars>(1, err);
    ir_ = alloc<IrScalars>(1, err);
    npartials_ = 4 * maxblk + 64;
    partials_ = alloc<double>(npartials_, err);
    ppq_ = partials_;
    prr_ = partials_ + maxblk;
    prz_ = partials_ + 3 * maxblk;
    coef_ = alloc<float>(2 * 64, err);
    nz_flag_ = alloc<int>(1, err);
    guess_flags_ = alloc<int>(blocks_for(n0_, BS), err);
    if (nz_flag_) chk(hipMemsetAsync(nz_flag_, 0, sizeof(int), stream_), err);
  }
  return err;
}

// Every work vector starts at zero: tiles without free voxels are never
// written and must read as zero from their neighbours.
void MgSolver::zero_vectors() {
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
}

hipError_t MgSolver::create(const pfsf_solver_desc& desc, bool batchable) {
  hipError_t err = hipSuccess;
  if (nvox(desc.dims) > kMaxSolverCells) return hipErrorInvalidValue;
  if (batchable) {
    configure(desc, 2, kBatchTZFine, kBatchTZCoarse);
  } else {
    configure(desc, 1, 0, 0);
    dense_ok_ = true;
    const MgLevel& C = levels_.back();
    if (nlev_ > 1 && desc.coarse_sweeps <= 0 && C.n <= kMaxLdsCoarse &&
        padded_cells(C.d) <= kMaxLdsPadded)
      coarse_sweeps_ = kCoarseSweepsDense;
  }
  // The solver runs on a stream of its own (default priority) so that its
  // iterations can be captured into a HIP graph; solve() orders it after the
  // caller's stream and waits for it before returning.
  if (chk(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), err)) return err;
  owns_stream_ = true;
  if (chk(allocate(true), err)) return err;
  if (chk(hipHostMalloc((void**)&sc_host_, sizeof(PcgScalars) + sizeof(IrScalars)), err))
    return err;
  // device stores must reach the host while a graph is still running
  if (chk(hipHostMalloc((void**)&mirror_, 2 * sizeof(unsigned long long),
                        hipHostMallocMapped | hipHostMallocCoherent),
          err))
    return err;
  mirror_[0] = mirror_[1] = 0;
  chk(hipEventCreate(&ev0_), err);
  chk(hipEventCreate(&ev1_), err);
  if (err != hipSuccess) return err;
  zero_vectors();
  chk(hipGetLastError(), err);
  if (err != hipSuccess) return err;
  return build();
}

void MgSolver::configure_pooled(const pfsf_solver_desc& desc, int tz_fine, int tz_coarse) {
  configure(desc, 2, tz_fine, tz_coarse);
}

void MgSolver::allocate_pooled(DevPool& pool) {
  hipError_t err = hipSuccess;
  pool_ = &pool;
  (void)allocate(false);
  (void)err;
  pool_ = nullptr;
}

hipError_t MgSolver::build_pooled(hipStream_t s) {
  stream_ = s;
  owns_stream_ = false;
  zero_vectors();
  hipError_t err = hipGetLastError();
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
  L.tiles.host_active = act;
  // a coarse level whose tiles are nearly all active launches over all of
  // them (empty tiles find no free voxel and write zeros)
  L.tiles.dense = dense_ok_ && &L != &levels_[0] && (int64_t)T * 8 >= nt * 7;
  if (L.tiles.dense) L.tiles.nlaunch = (int)((nt + 7) / 8 * 8);
  chk(hipMemcpyAsync(L.tiles.ids, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice,
                     stream_),
      err);
  if (L.tiles.cells)
    k_cell_bits<<<L.tiles.nlaunch, BS, 0, stream_>>>(L.d, type_, L.fr, L.tiles.ids, L.tiles.ntx,
                                                     L.tiles.nty, L.tiles.tz, L.tiles.cells);
  chk(hipGetLastError(), err);
  chk(hipStreamSynchronize(stream_), err);
  return err;
}

// type and kappa may still be in flight on the caller's stream (an upload
// queued just before): set-up work waits for it.
hipError_t MgSolver::wait_caller() {
  hipError_t err = hipSuccess;
  hipStream_t caller = (hipStream_t)desc_.stream;
  if (caller == stream_) return err;
  hipEvent_t ev = nullptr;
  chk(hipEventCreateWithFlags(&ev, hipEventDisableTiming), err);
  if (ev) {
    chk(hipEventRecord(ev, caller), err);
    chk(hipStreamWaitEvent(stream_, ev, 0), err);
    (void)hipEventDestroy(ev);
  }
  return err;
}

hipError_t MgSolver::build() {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  drop_graphs();
  if (chk(wait_caller(), err)) return err;
  // q_ is free until the first solve: it carries the component parents and
  // is cleared again afterwards (every work vector starts at zero)
  chk(anchored_mask(d0_, type_, reach_, (uint32_t*)q_, flag_, s), err);
  fill0(q_, n0_, s);
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
    // Chebyshev interval of the coarsest solve: [lmax / 100, lmax].  Measured
    // against [lmax / 30, lmax] at 6 steps: 43 -> 40 PCG iterations at 128^3,
    // 82 -> 79 at 256^3, 30 -> 25 at 96^3, at the same cost per iteration.
    MgLevel& C = levels_[nlev_ - 1];
    double lmax = C.lmax, lmin = C.lmax / 100.0;
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
    if (coef_)
      chk(hipMemcpyAsync(coef_, coef_host_.data(), sizeof(float) * coef_host_.size(),
                         hipMemcpyHostToDevice, s),
          err);
  }
  setup_operator(s);
  chk(hipGetLastError(), err);
  for (int l = 0; l < nlev_ && err == hipSuccess; l++) chk(build_tiles(levels_[l]), err);
  chk(hipStreamSynchronize(s), err);
  return err;
}

// The kappa-dependent set-up of every level: fine kappa and anchor weights,
// coarsened operators, diagonals (the free set follows them) and the finest
// level's anchor-coded kappa.
void MgSolver::setup_operator(hipStream_t s) {
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
  for (int l = 1; l < nlev_; l++) {
    const MgLevel& A = levels_[l - 1];
    MgLevel& B = levels_[l];
    k_coarse_mask<<<blocks_for(B.n, BS), BS, 0, s>>>(A.d, A.fr, A.msk, B.d, B.fr, B.msk);
  }
  // diagonals need every level's free flags and masks: computed after both
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
  if (F.ks)
    k_stream_kappa<<<blocks_for(n0_, BS), BS, 0, s>>>(n0_, type_, kappa_, F.fr,
                                                      (const float*)F.kap, F.ks);
  MgLevel& C = levels_[nlev_ - 1];
  if (C.pm) {  // the coarsest polynomial (coefficients already on the stream)
    const int st = std::min(coarse_sweeps_, 63);
    const int64_t nn = C.n * (int64_t)C.pm_ld;
    fill0(C.pm, nn, s);
    if (C.bf16)
      k_cheb_columns<uint16_t><<<(unsigned)C.n, kMaxLdsCoarse, 0, s>>>(view<uint16_t>(C), coef_,
                                                                        st, C.pm, C.pm_ld);
    else
      k_cheb_columns<float><<<(unsigned)C.n, kMaxLdsCoarse, 0, s>>>(view<float>(C), coef_, st,
                                                                     C.pm, C.pm_ld);
    k_sym_dense<<<blocks_for(C.n * C.n, BS), BS, 0, s>>>(C.pm, (int)C.n, C.pm_ld);
    if (C.fg) {  // the level above's down-sweep folded in: G = P R
      const MgLevel& B = levels_[nlev_ - 2];
      if (B.bf16)
        k_fold_columns<uint16_t><<<blocks_for(B.n, BS), BS, 0, s>>>(view<uint16_t>(B), B.fr, C.d,
                                                                     C.fpar, C.fval);
      else
        k_fold_columns<float><<<blocks_for(B.n, BS), BS, 0, s>>>(view<float>(B), B.fr, C.d,
                                                                  C.fpar, C.fval);
      fill0(C.fg, C.n * (int64_t)C.fg_ld, s);
      k_fold_matrix<<<blocks_for(C.n * B.n, BS), BS, 0, s>>>(C.pm, C.pm_ld, (int)C.n, C.fpar,
                                                              C.fval, (int)B.n, C.fg, C.fg_ld);
    }
  }
}

hipError_t MgSolver::update(const uint8_t* type, const float* kappa) {
  type_ = type;
  kappa_ = kappa;
  return build();
}

hipError_t MgSolver::refresh_kappa(const float* kappa) {
  hipError_t err = hipSuccess;
  if (pool_ || !sc_) return hipErrorNotSupported;
  hipStream_t s = stream_;
  int64_t total = 0;
  for (const MgLevel& L : levels_) total += L.n;
  if (!fr_snap_) {
    fr_snap_ = alloc<uint8_t>(total, err);
    if (err != hipSuccess) return err;
  }
  kappa_ = kappa;
  if (chk(wait_caller(), err)) return err;
  int64_t off = 0;
  for (const MgLevel& L : levels_) {
    chk(hipMemcpyAsync(fr_snap_ + off, L.fr, L.n, hipMemcpyDeviceToDevice, s), err);
    off += L.n;
  }
  chk(hipMemsetAsync(flag_, 0, sizeof(int), s), err);
  setup_operator(s);
  off = 0;
  for (const MgLevel& L : levels_) {
    k_fr_moved<<<blocks_for(L.n, BS), BS, 0, s>>>(L.n, fr_snap_ + off, L.fr, flag_);
    off += L.n;
  }
  chk(hipGetLastError(), err);
  int moved = 0;
  chk(hipMemcpyAsync(&moved, flag_, sizeof(int), hipMemcpyDeviceToHost, s), err);
  chk(hipStreamSynchronize(s), err);
  if (err != hipSuccess) return err;
  return moved ? build() : hipSuccess;
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

// The finest level also streams 16-plane columns.
#define PFSF_TZF_DISPATCH(L, CALL) \
  if ((L).tiles.tz == 16) {        \
    constexpr int TZV = 16;        \
    CALL;                          \
  } else                           \
    PFSF_TZ_DISPATCH(L, CALL)

// The finest level's streaming kernels with their column walk chosen at run
// time (LOOPV: a long column list walked in strides, else one column per
// work-group).
#define PFSF_STREAM_DISPATCH(L, LOOP, CALL) \
  if (LOOP) {                               \
    constexpr bool LOOPV = true;            \
    PFSF_TZF_DISPATCH(L, CALL);             \
  } else {                                  \
    constexpr bool LOOPV = false;           \
    PFSF_TZF_DISPATCH(L, CALL);             \
  }

// With rr_n > 0 the launch carries the convergence test (see k_down).
template <typename KT>
static void launch_down_t(const MgLevel& A, const MgLevel& B, const float* r, PcgScalars* sc,
                          hipStream_t s, const double* rr_part, unsigned rr_n,
                          unsigned long long* mirror) {
  const unsigned g = A.tiles.nlaunch + (rr_n > 0 ? 1u : 0u);
  if (A.msk) {
    PFSF_TZ_DISPATCH(A, (k_down<TZV, KT, false, true><<<g, BS, 0, s>>>(
                             view<KT>(A), tl(A), B.d, r, nullptr, nullptr, nullptr, nullptr, B.r,
                             sc, nullptr, 0u, nullptr, rr_part, rr_n, mirror)));
  } else {
    PFSF_TZ_DISPATCH(A, (k_down<TZV, KT, false><<<g, BS, 0, s>>>(
                             view<KT>(A), tl(A), B.d, r, nullptr, nullptr, nullptr, nullptr, B.r,
                             sc, nullptr, 0u, nullptr, rr_part, rr_n, mirror)));
  }
}

static void launch_down(const MgLevel& A, const MgLevel& B, const float* r, PcgScalars* sc,
                        hipStream_t s, const double* rr_part = nullptr, unsigned rr_n = 0u,
                        unsigned long long* mirror = nullptr) {
  if (A.bf16)
    launch_down_t<uint16_t>(A, B, r, sc, s, rr_part, rr_n, mirror);
  else
    launch_down_t<float>(A, B, r, sc, s, rr_part, rr_n, mirror);
}

// A coarse level launched over all its tiles runs its up-sweep on one-plane
// tiles: twice the work-groups of the 2-plane tiles, each thread one cell.
// The up-sweep restricts nothing, so it needs no z-pair in a thread, and
// these small launches are bound by one wave's chain of loads, LDS and
// stencil work, which halves (128^3: 3 up-sweeps, iteration 68.4 -> 65.6 us;
// the same per-cell arithmetic, so the same bits).
template <typename KT>
static void launch_up_t(const MgLevel& A, const MgLevel& B, const float* r, float* out,
                        PcgScalars* sc, hipStream_t s) {
  if (A.msk && A.tiles.dense) {
    const int nt = A.tiles.ntx * A.tiles.nty * A.d.lz;
    k_up<1, KT, false, true><<<(unsigned)((nt + 7) / 8 * 8), BS, 0, s>>>(
        view<KT>(A), TL{A.tiles.ntx, A.tiles.nty, nullptr, nt, nullptr}, B.d, r, B.e, out, sc,
        nullptr);
    return;
  }
  if (A.msk) {
    PFSF_TZ_DISPATCH(A, (k_up<TZV, KT, false, true><<<A.tiles.nlaunch, BS, 0, s>>>(
                             view<KT>(A), tl(A), B.d, r, B.e, out, sc, nullptr)));
  } else {
    PFSF_TZ_DISPATCH(A, (k_up<TZV, KT, false><<<A.tiles.nlaunch, BS, 0, s>>>(
                             view<KT>(A), tl(A), B.d, r, B.e, out, sc, nullptr)));
  }
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
  const unsigned G = fine_groups();  // r.r partials of the finest update
  MgLevel& C = levels_[nlev_ - 1];
  const bool lds = C.n <= kMaxLdsCoarse && padded_cells(C.d) <= kMaxLdsPadded;
  int st = std::min(coarse_sweeps_, 63);
  // the convergence test runs beside the first coarse down-sweep when there
  // is one, else inside the one-work-group coarsest solve
  const bool test_in_down = check && from < nlev_ - 1;
  const unsigned rr_n = (check && lds && !test_in_down) ? G : 0u;
  if (check && !lds && !test_in_down)
    k_finalize<<<1, BS, 0, s>>>(FIN_UPDATE, prr_, G, sc_, ir_, mirror_);
  // with the fold, the level above the coarsest is restricted inside the
  // coarsest solve (k_fold_dense reads its residual)
  const bool fold = C.fg && from < nlev_ - 2;
  for (int l = from; l < nlev_ - (fold ? 2 : 1); l++) {
    if (l == from && test_in_down)
      launch_down(levels_[l], levels_[l + 1], levels_[l].r, sc_, s, prr_, G, mirror_);
    else
      launch_down(levels_[l], levels_[l + 1], levels_[l].r, sc_, s);
  }
  if (fold) {
    const MgLevel& B = levels_[nlev_ - 2];
    const unsigned g = (unsigned)((C.n + BS / 64 - 1) / (BS / 64));
    switch (C.fg_ld / 256) {
      case 4:
        k_fold_dense<4><<<g, BS, 0, s>>>(C.fg, (int)C.n, (int)B.n, B.r, C.e, sc_, prr_, rr_n,
                                         mirror_);
        break;
      case 8:
        k_fold_dense<8><<<g, BS, 0, s>>>(C.fg, (int)C.n, (int)B.n, B.r, C.e, sc_, prr_, rr_n,
                                         mirror_);
        break;
      case 16:
        k_fold_dense<16><<<g, BS, 0, s>>>(C.fg, (int)C.n, (int)B.n, B.r, C.e, sc_, prr_, rr_n,
                                          mirror_);
        break;
      default:
        k_fold_dense<32><<<g, BS, 0, s>>>(C.fg, (int)C.n, (int)B.n, B.r, C.e, sc_, prr_, rr_n,
                                          mirror_);
        break;
    }
  } else if (C.pm) {
    const unsigned g = (unsigned)((C.n + BS / 64 - 1) / (BS / 64));
    switch (C.pm_ld / 256) {
      case 1:
        k_cheb_dense<1><<<g, BS, 0, s>>>(C.pm, (int)C.n, C.r, C.e, sc_, prr_, rr_n, mirror_);
        break;
      case 2:
        k_cheb_dense<2><<<g, BS, 0, s>>>(C.pm, (int)C.n, C.r, C.e, sc_, prr_, rr_n, mirror_);
        break;
      case 3:
        k_cheb_dense<3><<<g, BS, 0, s>>>(C.pm, (int)C.n, C.r, C.e, sc_, prr_, rr_n, mirror_);
        break;
      default:
        k_cheb_dense<4><<<g, BS, 0, s>>>(C.pm, (int)C.n, C.r, C.e, sc_, prr_, rr_n, mirror_);
        break;
    }
  } else if (lds) {
    if (C.bf16)
      k_cheb_lds<uint16_t><<<1, kMaxLdsCoarse, 0, s>>>(view<uint16_t>(C), C.r, C.e, coef_, st,
                                                        sc_, prr_, rr_n, mirror_);
    else
      k_cheb_lds<float><<<1, kMaxLdsCoarse, 0, s>>>(view<float>(C), C.r, C.e, coef_, st, sc_,
                                                     prr_, rr_n, mirror_);
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

// z = M r_[c] and r.z at the start of a solve segment.
hipError_t MgSolver::precondition(int c) {
  hipStream_t s = stream_;
  MgLevel& F = levels_[0];
  if (nlev_ == 1) {
    // r.z goes where the Jacobi update leaves it, behind its r.r partials
    unsigned nb = blocks_for(n0_, BS);
    k_jacobi_apply<<<nb, BS, 0, s>>>(n0_, r_[c], F.dg, z_, prr_ + nb);
    return hipGetLastError();
  }
  MgLevel& B = levels_[1];
  const unsigned G = fine_groups();
  const bool loop = fine_loop();
  PFSF_STREAM_DISPATCH(F, loop, (ks_down<TZV, false, LOOPV><<<G, BS, 0, s>>>(
                                    sview(F), tl(F), B.d, r_[c], nullptr, nullptr, nullptr,
                                    nullptr, nullptr, B.r, sc_, nullptr, 0u, nullptr)));
  (void)vcycle_below(1, false);
  PFSF_STREAM_DISPATCH(F, loop, (ks_up<TZV, true, LOOPV><<<G, BS, 0, s>>>(
                                    sview(F), tl(F), B.d, r_[c], B.e, z_, sc_, prz_)));
  return hipGetLastError();
}

hipError_t MgSolver::precond_begin() {
  k_profile_begin<<<1, 1, 0, stream_>>>(sc_);
  return hipGetLastError();
}

// One PCG iteration reading r_[c], p_[c] and writing r_[1-c], p_[1-c]
// (iterations of a pass alternate c = 0, 1, ...).  The FP64 iterate is
// updated in place by the odd iterations, with the even one's direction
// p_[1] and their own p_[0] (see ks_down).
hipError_t MgSolver::iteration(int c) {
  hipStream_t s = stream_;
  MgLevel& F = levels_[0];
  const unsigned G = F.tiles.nlaunch;
  if (nlev_ == 1) {
    const unsigned nb = blocks_for(n0_, BS);
    PFSF_TZ_DISPATCH(F, (k_spmv_dir<TZV, float><<<G, BS, 0, s>>>(
                            view<float>(F), tl(F), z_, p_[c], p_[1 - c], q_, sc_, prr_ + nb, nb,
                            c, ppq_)));
    k_finalize<<<1, BS, 0, s>>>(FIN_ALPHA, ppq_, G, sc_, ir_, nullptr);
    k_update_jacobi<<<nb, BS, 0, s>>>(n0_, F.dg, x_, r_[c], r_[1 - c], p_[1 - c], q_, z_, sc_,
                                      prr_);
    k_finalize<<<1, BS, 0, s>>>(FIN_UPDATE, prr_, nb, sc_, ir_, mirror_);
    return hipGetLastError();
  }
  MgLevel& B = levels_[1];
  const unsigned GS = fine_groups();
  const bool loop = fine_loop();
  PFSF_STREAM_DISPATCH(F, loop, (ks_spmv<TZV, LOOPV><<<GS, BS, 0, s>>>(
                                    sview(F), tl(F), z_, p_[c], p_[1 - c], q_, sc_, prz_, GS, c,
                                    ppq_)));
  if (c) {
    PFSF_STREAM_DISPATCH(F, loop, (ks_down<TZV, true, LOOPV, kXPair><<<GS, BS, 0, s>>>(
                                      sview(F), tl(F), B.d, r_[c], r_[1 - c], q_, p_[1 - c],
                                      p_[c], x_, B.r, sc_, ppq_, GS, prr_)));
  } else {
    PFSF_STREAM_DISPATCH(F, loop, (ks_down<TZV, true, LOOPV, kXDefer><<<GS, BS, 0, s>>>(
                                      sview(F), tl(F), B.d, r_[c], r_[1 - c], q_, p_[1 - c],
                                      nullptr, x_, B.r, sc_, ppq_, GS, prr_)));
  }
  (void)vcycle_below(1, true);
  PFSF_STREAM_DISPATCH(F, loop, (ks_up<TZV, true, LOOPV><<<GS, BS, 0, s>>>(
                                    sview(F), tl(F), B.d, r_[1 - c], B.e, z_, sc_, prz_)));
  return hipGetLastError();
}

// n = 1, 2 or 4 PCG iterations, the first of parity c.  Each sequence is
// captured once into a HIP graph and replayed; the capture holds kernel
// launches only.  Without graph support the kernels are launched directly.
hipError_t MgSolver::launch_iters(int n, int c) {
  hipError_t err = hipSuccess;
  const int k = 2 * (n == 4 ? 2 : n == 2 ? 1 : 0) + c;
  if (use_graph_ && !graph_exec_[k]) {
    if (!chk(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), err)) {
      for (int i = 0; i < n; i++) (void)iteration((c + i) & 1);
      hipGraph_t g = nullptr;
      chk(hipStreamEndCapture(stream_, &g), err);
      if (err == hipSuccess &&
          !chk(hipGraphInstantiate(&graph_exec_[k], g, nullptr, nullptr, 0), err))
        graph_[k] = g;
      else if (g)
        (void)hipGraphDestroy(g);
    }
    if (err != hipSuccess) {  // fall back to direct launches
      (void)hipGetLastError();
      drop_graphs();
      use_graph_ = false;
      err = hipSuccess;
    }
  }
  if (use_graph_) return hipGraphLaunch(graph_exec_[k], stream_);
  for (int i = 0; i < n && err == hipSuccess; i++) chk(iteration((c + i) & 1), err);
  return err;
}

// Waits until the mirrored progress word shows at least `target` convergence
// tests of the running pass, or a stop.  The word is stored by the test
// itself, so the host follows the iteration without a copy on the stream.
// A breakdown (p.q <= 0) stops the pass without a convergence test: once the
// stream has drained, the device scalars are copied back and decide.
hipError_t MgSolver::wait_progress(int target, int* iters, int* done, double* rr) {
  // The stream is queried only after 2 ms without a new word: a query can
  // place a marker on the queue, which would cost every iteration a packet.
  using clk = std::chrono::steady_clock;
  unsigned long long last = ~0ull;
  clk::time_point since = clk::now();
  for (unsigned spin = 0;; spin++) {
    const unsigned long long w = __atomic_load_n(mirror_, __ATOMIC_ACQUIRE);
    *iters = (int)(w >> 8);
    *done = (int)(w & 0xffu);
    if (*done || *iters >= target) {
      const unsigned long long rb = __atomic_load_n(mirror_ + 1, __ATOMIC_ACQUIRE);
      memcpy(rr, &rb, sizeof(double));
      return hipSuccess;
    }
    if (w != last) {
      last = w;
      since = clk::now();
      continue;
    }
    if ((spin & 255u) != 255u || clk::now() - since < std::chrono::milliseconds(2)) continue;
    const hipError_t q = hipStreamQuery(stream_);
    if (q == hipErrorNotReady) {
      since = clk::now();
      continue;
    }
    if (q != hipSuccess) return q;
    hipError_t err = hipSuccess;
    chk(hipMemcpyAsync(sc_host_, sc_, sizeof(PcgScalars), hipMemcpyDeviceToHost, stream_), err);
    chk(hipStreamSynchronize(stream_), err);
    *iters = sc_host_->iters;
    *done = sc_host_->done;
    *rr = sc_host_->rr;
    if (!*done && *iters < target) *done = 3;  // drained short of the target: stop
    return err;
  }
}

// True FP64 residual of x_ into r_[c]; returns ||r|| / ||b|| (synchronises).
// With seq > 0 (the solve's first residual) an all-zero initial guess, as
// flagged by k_load_guess, takes the residual b without the stencil.
hipError_t MgSolver::true_residual(const float* b, double* rel, int c, int seq) {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  const MgLevel& F = levels_[0];
  const int* nz = seq > 0 ? nz_flag_ : nullptr;
  // one work-group per column: nlaunch partials, within the partial buffer
  // (4 x the largest launch), not live between passes
  const unsigned G = (unsigned)F.tiles.nlaunch;
  if ((int64_t)G > npartials_) return hipErrorInvalidValue;
  PFSF_TZF_DISPATCH(F, (k_true_residual<TZV><<<G, BS, 0, s>>>(d0_, tl(F), type_, kappa_, F.fr, b,
                                                              x_, r_[c], partials_, nz, seq)));
  k_finalize<<<1, BS, 0, s>>>(FIN_IR, partials_, G, sc_, ir_, nullptr);
  IrScalars* irh = (IrScalars*)(sc_host_ + 1);
  chk(hipMemcpyAsync(irh, ir_, sizeof(IrScalars), hipMemcpyDeviceToHost, s), err);
  chk(hipStreamSynchronize(s), err);
  *rel = irh->rel;
  return err;
}

// Solve passes: PCG on FP32 vectors with the iterate accumulated in FP64,
// started from the true FP64 residual; a pass ends when its recursive
// residual falls below 0.85 of the target (or the FP32 resolution limit),
// then the true residual decides whether a further pass (a restart) is
// needed.  A pass that starts near the target tracks the true residual
// closely (its FP32 recursion error scales with its own small start), so the
// margin only has to cover that: 0.85 against 0.5 saved one iteration of
// the 128^3 solve's second pass (35 -> 33 iterations in all, final residual
// 6.3e-7) and the 256^3 one (71 -> 70).
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
  solve_seq_ = solve_seq_ % 0x7ffffffe + 1;  // never 0, the flag's initial value
  const int seq = solve_seq_;
  const unsigned nbg = std::min(nb0, kBeginBlocks);
  k_begin<<<nbg, BS, 0, s>>>(n0_, x64, F.fr, b, x_, guess_flags_, partials_);
  k_begin_fin<<<1, BS, 0, s>>>(partials_, guess_flags_, nbg, ir_, nz_flag_, seq);
  double rel = 0.0;
  chk(true_residual(b, &rel, 0, seq), err);
  int passes = 0, total = 0;
  const IrScalars* irh = (const IrScalars*)(sc_host_ + 1);
  for (; rel > tol && total < max_it && err == hipSuccess && passes < 16;) {
    const double floor = passes == 0 ? floor0_ : (double)PFSF_PASS_FLOOR;
    const double stop_rel = std::max(0.85 * tol / rel, floor);
    const double rel_start = rel;
    mirror_[0] = mirror_[1] = 0;  // the stream is idle here (true_residual synchronised)
    k_pcg_begin<<<1, 1, 0, s>>>(sc_, ir_, stop_rel, max_it - total, 1);
    chk(precondition(0), err);
    // The device stops the pass at its tolerance or after `cap` iterations;
    // the host only decides how much work to queue.  It follows every
    // convergence test through the mirror and tops the queue up while at
    // most one untested iteration is left, by the iterations that the
    // contraction rate still predicts (before this pass has a rate of its
    // own, the rate of the same pass of the previous solve).  Queued iterations past the stop still cost their
    // launches (their kernels return at once), so the queue holds no more
    // than the prediction asks for.
    const int cap = max_it - total;
    const double rr0 = irh->rr, stop2 = rr0 * stop_rel * stop_rel;
    int launched = 0, it = 0, done = 0;
    double rr = rr0;
    std::vector<double> hist;  // r.r after each test
    auto predict = [&](double rho) {  // iterations from r.r = rr to the stop
      if (!(rho > 0.0 && rho < 1.0) || rr <= stop2) return 8;
      return std::max(1, (int)ceil(log(stop2 / rr) / log(rho)));
    };
    auto launch = [&](int want) {  // queue `want` iterations (1, 2 or 4 per graph)
      want = std::max(1, std::min(want, cap + 1 - launched));
      while (want > 0 && err == hipSuccess) {
        const int n = want >= 4 ? 4 : want >= 2 ? 2 : 1;
        chk(launch_iters(n, launched & 1), err);
        launched += n;
        want -= n;
      }
    };
    const double rho_last = rho_pass_[std::min(passes, 3)];
    launch(std::min(4, predict(rho_last)));
    while (!done && err == hipSuccess) {
      chk(wait_progress(it + 1, &it, &done, &rr), err);
      if (done || err != hipSuccess) break;
      hist.resize(it, rr);  // tests the host did not see take the latest r.r
      hist[it - 1] = rr;
      const int pending = launched - it;  // iterations not yet tested
      if (pending > 1 || launched > cap) continue;
      // rate over the last three tests, else since the pass started
      double rho = it >= 4 && hist[it - 4] > rr ? pow(rr / hist[it - 4], 1.0 / 3.0)
                                                : pow(rr / rr0, 1.0 / it);
      if (it < 4 && rho_last > 0.0) rho = std::min(rho, rho_last);
      const int more = predict(rho);
      if (more > pending) launch(std::min(more - pending, 8));
    }
    if (it > 0 && rr > 0.0 && rr < rr0) rho_pass_[std::min(passes, 3)] = pow(rr / rr0, 1.0 / it);
    // the true residual is queued behind any iterations still in flight
    sc_host_->done = done;
    sc_host_->iters = it;
    total += it;
    passes++;
    // a restart from the true residual: continuing the recursion with the
    // replaced residual and the old direction (residual replacement) was
    // measured slower here, 48 iterations against 43 at 128^3
    if (nlev_ > 1) k_x_pending<<<nb0, BS, 0, s>>>(n0_, sc_, p_[1], x_);
    double nrel = rel;
    chk(true_residual(b, &nrel, 0), err);
    bool stalled = (it == 0) || (sc_host_->done == 2 && nrel >= rel);
    // The FP32 recursion drifts from the true residual: a first pass whose
    // true residual ended well above its recursive one ran iterations that
    // did not lower the true residual.  The next solve's first pass then
    // stops at half the true drop this one reached (a restart from the true
    // residual converges from there); a pass with little drift lets the
    // floor decay back.  128^3 (drift 1.8x) keeps 2e-6; the 256^3 island at
    // sigma 1.5 (true 8e-5 against 2e-6) stops its first pass near 4e-5:
    // 69 -> 65 iterations.
    if (passes == 1 && done == 1 && it > 0 && rr > 0.0 && rr0 > 0.0 && rel_start > 0.0) {
      const double rec = rel_start * sqrt(rr / rr0);  // recursive relative residual
      const double drift = nrel / rec;
      if (drift > 3.0)
        floor0_ = std::min(std::max(0.5 * nrel / rel_start, (double)PFSF_PASS_FLOOR), 1e-2);
      else if (drift < 1.5)
        floor0_ = std::max(0.5 * floor0_, (double)PFSF_PASS_FLOOR);
    }
    rel = nrel;
    if (stalled) break;
  }
  if (x64 || phi32) k_end<<<nb0, BS, 0, s>>>(n0_, x_, x64, phi32);
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
// update + down-sweep of an odd iteration (r and the FP64 iterate with two
// directions; or the Jacobi update), 5 the same of an even iteration (r
// only, the iterate update left pending), 2 second-level down-sweep, 3
// finest up-sweep + r.z: each launch is bracketed by the kernel's own
// start/stop timestamps (hipExtLaunchKernel events), so the figure is the
// kernel duration alone.  4: one whole PCG iteration, the average over
// `reps` (rounded up to a multiple of 4) back-to-back iterations launched as
// the solve launches them (graphs of four iterations of alternating parity),
// timed between events, launch gaps included.  Call after a solve so the
// vectors hold finite data.
hipError_t MgSolver::profile(int which, int reps, float* avg_ms) {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  MgLevel& F = levels_[0];
  const unsigned G = F.tiles.nlaunch;  // tiled launches
  const unsigned GS = nlev_ > 1 ? fine_groups() : G;  // streaming launches
  const bool loop = nlev_ > 1 && fine_loop();
  if (which < 0 || which > 5 || reps <= 0) return hipErrorInvalidValue;
  if ((which == 2 || which == 3 || which == 5) && nlev_ < 2) return hipErrorInvalidValue;
  if (which == 2 && nlev_ < 3) return hipErrorInvalidValue;
  k_profile_begin<<<1, 1, 0, s>>>(sc_);
  if (which == 4) {
    // the solve's own launches: captured graphs of four iterations (parities
    // 0, 1, 0, 1) where graphs are in use, else direct launches
    const int n4 = (reps + 3) / 4;
    chk(launch_iters(4, 0), err);  // warm (and capture)
    chk(hipEventRecord(ev0_, s), err);
    for (int k = 0; k < n4 && err == hipSuccess; k++) chk(launch_iters(4, 0), err);
    chk(hipEventRecord(ev1_, s), err);
    chk(hipEventSynchronize(ev1_), err);
    float ms = 0.f;
    chk(hipEventElapsedTime(&ms, ev0_, ev1_), err);
    *avg_ms = ms / (float)(4 * n4);
    return err;
  }
  auto ext = [&](auto kern, unsigned grid, hipEvent_t a, hipEvent_t b, auto... args) {
    hipExtLaunchKernelGGL(kern, dim3(grid), dim3(BS), 0, s, a, b, 0, args...);
  };
  auto once = [&](hipEvent_t a, hipEvent_t b) {
    switch (which) {
      case 0:
        if (nlev_ == 1) {  // the tiled SpMV of the Jacobi-PCG iteration
          const unsigned nb = blocks_for(n0_, BS);
          PFSF_TZ_DISPATCH(F, ext(k_spmv_dir<TZV, float>, G, a, b, view<float>(F), tl(F),
                                  (const float*)z_, (const float*)p_[0], p_[1], q_, sc_,
                                  (const double*)(prr_ + nb), nb, 0, ppq_));
        } else {
          PFSF_STREAM_DISPATCH(F, loop, ext(ks_spmv<TZV, LOOPV>, GS, a, b, sview(F), tl(F),
                                            (const float*)z_, (const float*)p_[0], p_[1], q_,
                                            sc_, (const double*)prz_, GS, 0, ppq_));
        }
        break;
      case 1:
        if (nlev_ == 1) {
          const unsigned nb = blocks_for(n0_, BS);
          ext(k_update_jacobi, nb, a, b, n0_, (const float*)F.dg, x_, (const float*)r_[0], r_[1],
              (const float*)p_[1], (const float*)q_, z_, (const PcgScalars*)sc_, prr_);
        } else {
          PFSF_STREAM_DISPATCH(F, loop, ext(ks_down<TZV, true, LOOPV, kXPair>, GS, a, b,
                                            sview(F), tl(F), levels_[1].d, (const float*)r_[0],
                                            r_[1], (const float*)q_, (const float*)p_[1],
                                            (const float*)p_[0], x_, levels_[1].r, sc_,
                                            (const double*)ppq_, GS, prr_));
        }
        break;
      case 5:
        PFSF_STREAM_DISPATCH(F, loop, ext(ks_down<TZV, true, LOOPV, kXDefer>, GS, a, b,
                                          sview(F), tl(F), levels_[1].d, (const float*)r_[0],
                                          r_[1], (const float*)q_, (const float*)p_[1],
                                          (const float*)nullptr, x_, levels_[1].r, sc_,
                                          (const double*)ppq_, GS, prr_));
        break;
      case 2: {
        const MgLevel& A = levels_[1];
        const MgLevel& B = levels_[2];
        if (A.bf16) {
          PFSF_TZ_DISPATCH(A, ext(k_down<TZV, uint16_t, false, true>, A.tiles.nlaunch, a, b,
                                  view<uint16_t>(A), tl(A), B.d, (const float*)A.r,
                                  (float*)nullptr, (const float*)nullptr, (const float*)nullptr,
                                  (double*)nullptr, B.r, sc_, (const double*)nullptr, 0u,
                                  (double*)nullptr, (const double*)nullptr, 0u,
                                  (unsigned long long*)nullptr));
        } else {
          PFSF_TZ_DISPATCH(A, ext(k_down<TZV, float, false, true>, A.tiles.nlaunch, a, b,
                                  view<float>(A), tl(A), B.d, (const float*)A.r,
                                  (float*)nullptr, (const float*)nullptr, (const float*)nullptr,
                                  (double*)nullptr, B.r, sc_, (const double*)nullptr, 0u,
                                  (double*)nullptr, (const double*)nullptr, 0u,
                                  (unsigned long long*)nullptr));
        }
        break;
      }
      default:
        PFSF_STREAM_DISPATCH(F, loop, ext(ks_up<TZV, true, LOOPV>, GS, a, b, sview(F), tl(F),
                                          levels_[1].d, (const float*)r_[1],
                                          (const float*)levels_[1].e, z_,
                                          (const PcgScalars*)sc_, prz_));
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
