// Device building blocks of the tiled multigrid-PCG stencil kernels, shared
// by the single-system solver (mg_solver.hip) and the multi-island batch
// solver (batch_solver.hip).  The kernel bodies take the level view, the tile
// origin and every vector explicitly, so a launch wrapper can resolve them
// either from its arguments (one system) or from per-tile descriptor tables
// (many systems in one launch).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

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
// Raw buffer access (MUBUF) to an FP32 array: the base sits in a 128-bit
// resource in SGPRs and each access adds one 32-bit vector byte offset, so a
// load or store costs no 64-bit address arithmetic (the compiler does not
// always fold ld()/st() into the SGPR-base form).  Accesses at or past
// `bytes` read 0 and write nothing.  0x00020000: the resource word of raw
// 32-bit buffer access on gfx9-family targets.
struct Buf {
  __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ Buf buf_of(const void* p, int64_t bytes) {
  return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0,
                                               (int)(uint32_t)bytes, 0x00020000)};
}
__device__ __forceinline__ float bld(const Buf& b, int i) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(b.r, (int)((unsigned)i * 4u), 0, 0));
}
__device__ __forceinline__ void bst(const Buf& b, int i, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), b.r, (int)((unsigned)i * 4u), 0, 0);
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

// ── kernel-argument batching ───────────────────────────────────────────
// Kernel arguments arrive by scalar loads from the kernarg segment, issued
// where a value is first used.  A use behind a branch, or after a wait for
// another argument, puts a separate memory round trip into the prologue; the
// iteration kernels showed five or six dependent round trips before their
// first vector load.  pin_args(...) at kernel entry makes every argument live
// in SGPRs at that point, so all of them load as one batch.
__device__ __forceinline__ void pin1(const void* v) { asm volatile("" ::"s"(v)); }
__device__ __forceinline__ void pin1(int v) { asm volatile("" ::"s"(v)); }
__device__ __forceinline__ void pin1(unsigned v) { asm volatile("" ::"s"(v)); }
__device__ __forceinline__ void pin1(float v) { asm volatile("" ::"s"(v)); }
__device__ __forceinline__ void pin1(pfsf_dims d) {
  pin1(d.lx);
  pin1(d.ly);
  pin1(d.lz);
}
template <typename... T>
__device__ __forceinline__ void pin_args(const T&... a) {
  (pin1(a), ...);
}

template <typename KT>
struct Lv {
  pfsf_dims d;
  const KT* kap;        // kappa on the free set, -1 elsewhere
  const float* dg;      // diagonal (0 outside the free set)
  const float* am;      // anchor mass
  const uint32_t* msk;  // coarse levels: coupling mask (nbit order); null = all
  float scale;
  float omega;
};

// Bit of neighbour offset (dx, dy, dz) in a coupling mask; the opposite
// offset has bit 26 - nbit.
__host__ __device__ constexpr int nbit(int dx, int dy, int dz) {
  return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
}
constexpr uint32_t kAllCouplings = 0x07FFFFFFu;

struct TL {
  int ntx, nty;
  const int* ids;  // null: every tile of the level, in XCD-chunked order (tile_origin)
  int n;  // entries in ids (the launch list, a multiple of 8); without ids, the tile count
  const uint32_t* cells;  // per-slot cell bits of the finest streaming level, or null
};

template <typename KT>
__device__ __forceinline__ void pin1(const Lv<KT>& L) {
  pin1(L.d);
  pin1(L.kap);
  pin1(L.dg);
  pin1(L.am);
  pin1(L.msk);
  pin1(L.scale);
  pin1(L.omega);
}
__device__ __forceinline__ void pin1(const TL& M) {
  pin1(M.ntx);
  pin1(M.nty);
  pin1(M.ids);
  pin1(M.n);
  pin1(M.cells);
}

// Cell bits of one column of the finest level (list slot g, words
// cells[g * kCellWords ...]): word ey * HX + ex describes the cell
// (x0 - 1 + ex, y0 - 1 + ey) of the column's plane footprint with its halo,
// in the layout of an LDS plane; bit pl (plane z0 - 1 + pl) is set when the
// cell is free, bit 16 + pl when it is free or an anchor (a cell whose kappa
// coding couples).  Columns of up to 14 planes.
constexpr int kCellWords = HXY;
constexpr int kCellMaxPlanes = 16;

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
  float s[HN];  // sqrt kappa on the free set, -0.0 elsewhere (see free_s)
};

// LDS tiles hold sqrt(kappa) only: kappa = s^2 for the face weights, and a
// cell outside the free set holds s = -0.0, so its edge and corner products
// vanish and its sign bit (set by no free cell, not even one with kappa = 0)
// marks it for the face test.  The hardware square root is used for both
// ends of every coupling, so the operator stays symmetric.
__device__ __forceinline__ float free_s(float k) {
  return k >= 0.f ? __builtin_amdgcn_sqrtf(k) : -0.0f;
}
__device__ __forceinline__ bool is_free_s(float s) { return __float_as_int(s) >= 0; }

// Work-group g takes list entry g, or, without a list, tile (g % 8) * chunk
// + g / 8 of the whole level (chunk = ceil(tiles / 8), the launch being
// 8 chunk work-groups): the same XCD-chunked order without the dependent load
// of a list entry before the operand loads.
template <int TZ>
__device__ __forceinline__ bool tile_origin(const TL& M, int& x0, int& y0, int& z0) {
  int t;
  if (M.ids) {
    t = M.ids[blockIdx.x];
    if (t < 0) return false;
  } else {
    t = (int)(blockIdx.x & 7u) * ((M.n + 7) >> 3) + (int)(blockIdx.x >> 3);
    if (t >= M.n) return false;
  }
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
  T.s[e] = free_s(k);
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
// difference form over free neighbours + anchor mass.  v holds values (0 off
// the free set), s the free_s coding of kappa.  With constant strides every
// read is one LDS instruction with an immediate offset from one base.
template <bool MASK = false>
__device__ __forceinline__ float stencil_apply(const float* __restrict__ v,
                                               const float* __restrict__ s, int c, int sy,
                                               int sz, float mi, float scale,
                                               uint32_t mk = kAllCouplings) {
  const float vc = v[c], si = s[c];
  const float ki = si * si;
  float af = 0.f, ae = 0.f, ac = 0.f;
  // A masked-off neighbour contributes exactly like one outside the free set.
  // The mask is applied as a bitwise AND on values that are always loaded, so
  // no neighbour access becomes a branch.
  auto keep = [&](float x, int b) {
    if (!MASK) return x;
    return __uint_as_float(__float_as_uint(x) & (0u - ((mk >> b) & 1u)));
  };
  auto face = [&](int o, int b) {
    const float sj = s[c + o];
    const float w = is_free_s(sj) ? fmaf(sj, sj, ki) : 0.f;
    af += keep(w, b) * (vc - v[c + o]);
  };
  auto edge = [&](int o, int b) { ae += keep(s[c + o], b) * (vc - v[c + o]); };
  auto corner = [&](int o, int b) { ac += keep(s[c + o], b) * (vc - v[c + o]); };
  face(-1, nbit(-1, 0, 0));
  face(1, nbit(1, 0, 0));
  face(-sy, nbit(0, -1, 0));
  face(sy, nbit(0, 1, 0));
  face(-sz, nbit(0, 0, -1));
  face(sz, nbit(0, 0, 1));
  edge(-1 - sy, nbit(-1, -1, 0));
  edge(1 - sy, nbit(1, -1, 0));
  edge(-1 + sy, nbit(-1, 1, 0));
  edge(1 + sy, nbit(1, 1, 0));
  edge(-1 - sz, nbit(-1, 0, -1));
  edge(1 - sz, nbit(1, 0, -1));
  edge(-1 + sz, nbit(-1, 0, 1));
  edge(1 + sz, nbit(1, 0, 1));
  edge(-sy - sz, nbit(0, -1, -1));
  edge(sy - sz, nbit(0, 1, -1));
  edge(-sy + sz, nbit(0, -1, 1));
  edge(sy + sz, nbit(0, 1, 1));
  corner(-1 - sy - sz, nbit(-1, -1, -1));
  corner(1 - sy - sz, nbit(1, -1, -1));
  corner(-1 + sy - sz, nbit(-1, 1, -1));
  corner(1 + sy - sz, nbit(1, 1, -1));
  corner(-1 - sy + sz, nbit(-1, -1, 1));
  corner(1 - sy + sz, nbit(1, -1, 1));
  corner(-1 + sy + sz, nbit(-1, 1, 1));
  corner(1 + sy + sz, nbit(1, 1, 1));
  return scale * (0.5f * af + si * (kEdgeP * ae + kCornerP * ac)) + mi * vc;
}

// stencil_apply's couplings at one cell, formed once for a cell whose
// kappa coding and mask do not change (the coarsest solve's steps): the
// masked face weights and the edge / corner neighbour codings, in
// stencil_apply's order.
struct CellCouplings {
  float f[6], e[12], c[8];
  float si;
};

template <bool MASK>
__device__ __forceinline__ void cell_couplings(const float* __restrict__ s, int c, int sy, int sz,
                                               uint32_t mk, CellCouplings& W) {
  const float si = s[c], ki = si * si;
  W.si = si;
  auto keep = [&](float x, int b) {
    if (!MASK) return x;
    return __uint_as_float(__float_as_uint(x) & (0u - ((mk >> b) & 1u)));
  };
  auto face = [&](int k, int o, int b) {
    const float sj = s[c + o];
    W.f[k] = keep(is_free_s(sj) ? fmaf(sj, sj, ki) : 0.f, b);
  };
  face(0, -1, nbit(-1, 0, 0));
  face(1, 1, nbit(1, 0, 0));
  face(2, -sy, nbit(0, -1, 0));
  face(3, sy, nbit(0, 1, 0));
  face(4, -sz, nbit(0, 0, -1));
  face(5, sz, nbit(0, 0, 1));
  const int eo[12] = {-1 - sy, 1 - sy, -1 + sy, 1 + sy, -1 - sz, 1 - sz,
                      -1 + sz, 1 + sz, -sy - sz, sy - sz, -sy + sz, sy + sz};
  const int eb[12] = {nbit(-1, -1, 0), nbit(1, -1, 0), nbit(-1, 1, 0), nbit(1, 1, 0),
                      nbit(-1, 0, -1), nbit(1, 0, -1), nbit(-1, 0, 1), nbit(1, 0, 1),
                      nbit(0, -1, -1), nbit(0, 1, -1), nbit(0, -1, 1), nbit(0, 1, 1)};
#pragma unroll
  for (int k = 0; k < 12; k++) W.e[k] = keep(s[c + eo[k]], eb[k]);
  const int co[8] = {-1 - sy - sz, 1 - sy - sz, -1 + sy - sz, 1 + sy - sz,
                     -1 - sy + sz, 1 - sy + sz, -1 + sy + sz, 1 + sy + sz};
  const int cb[8] = {nbit(-1, -1, -1), nbit(1, -1, -1), nbit(-1, 1, -1), nbit(1, 1, -1),
                     nbit(-1, -1, 1), nbit(1, -1, 1), nbit(-1, 1, 1), nbit(1, 1, 1)};
#pragma unroll
  for (int k = 0; k < 8; k++) W.c[k] = keep(s[c + co[k]], cb[k]);
}

// stencil_apply with the couplings of cell_couplings: the same products and
// sums in the same order, so the same bits.
__device__ __forceinline__ float stencil_apply_w(const float* __restrict__ v, int c, int sy,
                                                 int sz, const CellCouplings& W, float mi,
                                                 float scale) {
  const float vc = v[c];
  float af = 0.f, ae = 0.f, ac = 0.f;
  const int fo[6] = {-1, 1, -sy, sy, -sz, sz};
#pragma unroll
  for (int k = 0; k < 6; k++) af += W.f[k] * (vc - v[c + fo[k]]);
  const int eo[12] = {-1 - sy, 1 - sy, -1 + sy, 1 + sy, -1 - sz, 1 - sz,
                      -1 + sz, 1 + sz, -sy - sz, sy - sz, -sy + sz, sy + sz};
#pragma unroll
  for (int k = 0; k < 12; k++) ae += W.e[k] * (vc - v[c + eo[k]]);
  const int co[8] = {-1 - sy - sz, 1 - sy - sz, -1 + sy - sz, 1 + sy - sz,
                     -1 - sy + sz, 1 - sy + sz, -1 + sy + sz, 1 + sy + sz};
#pragma unroll
  for (int k = 0; k < 8; k++) ac += W.c[k] * (vc - v[c + co[k]]);
  return scale * (0.5f * af + W.si * (kEdgeP * ae + kCornerP * ac)) + mi * vc;
}

template <int TZ, bool MASK = false>
__device__ __forceinline__ float tile_apply(const Tile<TZ>& T, int c, float mi, float scale,
                                            uint32_t mk = kAllCouplings) {
  return stencil_apply<MASK>(T.v, T.s, c, HX, HXY, mi, scale, mk);
}

// The off-diagonal sum of stencil_apply, scale sum_j w_ij v_j over the
// (masked) couplings, so that (A v)_i = D_i v_i - offsum_apply with D_i the
// level's diagonal (k_level_diag: the same weights plus the anchor mass).
// Every cell off the free set holds v = 0, so its term vanishes without
// stencil_apply's coupling test; a masked-off free neighbour is cut by the
// mask.  The smoothers use it where the centre term cancels in closed form.
template <bool MASK = false>
__device__ __forceinline__ float offsum_apply(const float* __restrict__ v,
                                              const float* __restrict__ s, int c, int sy, int sz,
                                              float scale, uint32_t mk = kAllCouplings) {
  const float si = s[c];
  const float ki = si * si;
  float af = 0.f, ae = 0.f, ac = 0.f;
  auto keep = [&](float x, int b) {
    if (!MASK) return x;
    return __uint_as_float(__float_as_uint(x) & (0u - ((mk >> b) & 1u)));
  };
  auto face = [&](int o, int b) {
    const float sj = s[c + o];
    af = fmaf(keep(fmaf(sj, sj, ki), b), v[c + o], af);
  };
  auto edge = [&](int o, int b) { ae = fmaf(keep(s[c + o], b), v[c + o], ae); };
  auto corner = [&](int o, int b) { ac = fmaf(keep(s[c + o], b), v[c + o], ac); };
  face(-1, nbit(-1, 0, 0));
  face(1, nbit(1, 0, 0));
  face(-sy, nbit(0, -1, 0));
  face(sy, nbit(0, 1, 0));
  face(-sz, nbit(0, 0, -1));
  face(sz, nbit(0, 0, 1));
  edge(-1 - sy, nbit(-1, -1, 0));
  edge(1 - sy, nbit(1, -1, 0));
  edge(-1 + sy, nbit(-1, 1, 0));
  edge(1 + sy, nbit(1, 1, 0));
  edge(-1 - sz, nbit(-1, 0, -1));
  edge(1 - sz, nbit(1, 0, -1));
  edge(-1 + sz, nbit(-1, 0, 1));
  edge(1 + sz, nbit(1, 0, 1));
  edge(-sy - sz, nbit(0, -1, -1));
  edge(sy - sz, nbit(0, 1, -1));
  edge(-sy + sz, nbit(0, -1, 1));
  edge(sy + sz, nbit(0, 1, 1));
  corner(-1 - sy - sz, nbit(-1, -1, -1));
  corner(1 - sy - sz, nbit(1, -1, -1));
  corner(-1 + sy - sz, nbit(-1, 1, -1));
  corner(1 + sy - sz, nbit(1, 1, -1));
  corner(-1 - sy + sz, nbit(-1, -1, 1));
  corner(1 - sy + sz, nbit(1, -1, 1));
  corner(-1 + sy + sz, nbit(-1, 1, 1));
  corner(1 + sy + sz, nbit(1, 1, 1));
  return scale * (0.5f * af + si * (kEdgeP * ae + kCornerP * ac));
}

template <int TZ, bool MASK = false>
__device__ __forceinline__ float tile_offsum(const Tile<TZ>& T, int c, float scale,
                                             uint32_t mk = kAllCouplings) {
  return offsum_apply<MASK>(T.v, T.s, c, HX, HXY, scale, mk);
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

// Block partials go to partials[k * gridDim.x + blockIdx.x]; whoever needs
// the total sums them in a fixed order, so results are deterministic and the
// kernel boundary orders the writes.
template <int N>
__device__ __forceinline__ void write_partials(double (&v)[N], double* partials,
                                               double (*sh)[BS / 64]) {
  block_sum<N>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; k++) partials[(int64_t)k * gridDim.x + blockIdx.x] = v[k];
  }
}

// One FP64 partial per work-group at partials[slot].
__device__ __forceinline__ void write_partial_at(double v, double* partials,
                                                 double (*sh)[BS / 64], unsigned slot) {
  double a[1] = {v};
  block_sum<1>(a, sh);
  if (threadIdx.x == 0) partials[slot] = a[0];
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

// ── PCG + V-cycle kernels ──────────────────────────────────────────────
// The output loops below evaluate the stencil for every cell of the column
// and select the result (a cell outside the free set yields 0), so no branch
// depends on the free mask; only the stores are guarded by the grid bounds.

// p_out = z + beta p_in (z alone for the first direction); q = A p_out; p.q.
// beta = (r.z) / (r.z)_previous is formed here from the r.z partials of the
// preconditioner (rz_n work-groups); work-group 0 publishes r.z in slot
// `slot` (the iteration parity) and beta when `publish` (one work-group per
// system).  The p.q partial goes to partials[pslot].
template <int TZ, typename KT>
__device__ __forceinline__ void spmv_body(const Lv<KT>& L, bool have, int x0, int y0, int z0,
                                          const float* __restrict__ z,
                                          const float* __restrict__ pin, float* __restrict__ pout,
                                          float* __restrict__ q, PcgScalars* sc,
                                          const double* __restrict__ rz_part, unsigned rz_n,
                                          int slot, bool publish, double* partials,
                                          unsigned pslot) {
  if (sc->done) return;
  __shared__ Tile<TZ> T;
  __shared__ double sh[1][BS / 64];
  double acc[1] = {0.0};
  const bool first = !sc->have_dir;
  const double rz_old = sc->rzs[slot ^ 1];
  const double rz = block_total(rz_part, rz_n);
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
  if (publish && threadIdx.x == 0) {
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
      const float pv = T.v[c];
      float qi = tile_apply<TZ>(T, c, amr[tz], L.scale);
      qi = is_free_s(T.s[c]) ? qi : 0.f;
      acc[0] += (double)pv * qi;
      if (inb[tz]) {
        st(pout, i0 + tz * sz, pv);
        st(q, i0 + tz * sz, qi);
      }
    }
  }
  write_partial_at(acc[0], partials, sh, pslot);
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
// (pq_n work-groups); with `publish` the work-group stores it (or flags a
// breakdown when p.q <= 0).  The r.r partial goes to partials[pslot].
template <int TZ, typename KT, bool UPDATE, bool MASK = false>
__device__ __forceinline__ void down_body(const Lv<KT>& L, bool have, int x0, int y0, int z0,
                                          pfsf_dims c, const float* __restrict__ rin,
                                          float* __restrict__ rout, const float* __restrict__ q,
                                          const float* __restrict__ p, double* __restrict__ x,
                                          float* __restrict__ rc, PcgScalars* sc,
                                          const double* __restrict__ pq_part, unsigned pq_n,
                                          bool publish, double* partials, unsigned pslot) {
  // the done flag is tested once the operand loads are in flight
  const int done = sc->done;
  __shared__ Tile<TZ> T;
  __shared__ double sh[1][BS / 64];
  double acc[1] = {0.0};
  const double rz = UPDATE ? sc->rz : 0.0;
  const double pq0 = UPDATE ? block_total(pq_part, pq_n) : 0.0;
  const pfsf_dims d = L.d;
  const float om = L.omega;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int gx = x0 + tx, gy = y0 + ty, sz = d.lx * d.ly;
  const bool col = gx < d.lx && gy < d.ly;
  const int i0 = flat32(gx, gy, z0, d);
  // own-voxel operands (cell 0 stands in for a cell outside the grid; its
  // values are never used)
  bool inb[TZ];
  float ro[TZ], qo[TZ], pr[TZ];
  double xr[TZ];
  uint32_t mkr[TZ];
  float kk[Stage<TZ>::NP];
  float3 raw[Stage<TZ>::NP];  // r_in, q, diagonal
  if (have) {
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      inb[tz] = col && z0 + tz < d.lz;
      const int i = inb[tz] ? i0 + tz * sz : 0;
      ro[tz] = ld(rin, i);
      qo[tz] = UPDATE ? ld(q, i) : 0.f;
      pr[tz] = UPDATE ? ld(p, i) : 0.f;
      xr[tz] = UPDATE ? ld(x, i) : 0.0;
      mkr[tz] = MASK ? L.msk[i] : kAllCouplings;
    }
    stage_load<TZ>(L, x0, y0, z0,
                   [&](int j, int, int, int) {
                     return make_float3(ld(rin, j), UPDATE ? ld(q, j) : 0.f, ld(L.dg, j));
                   },
                   kk, raw);
  }
  if (done) return;
  double alpha64 = 0.0;
  if (UPDATE) {
    const double pq = pq0;
    const bool ok = pq > 0.0;
    alpha64 = ok ? rz / pq : 0.0;
    if (publish && threadIdx.x == 0) {
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
      const float ri = rr_[tz];
      const bool fr = is_free_s(T.s[cc]);
      // r - A x with x = omega D^-1 r: (1 - omega) r + sum_j w_ij x_j
      const float os = tile_offsum<TZ, MASK>(T, cc, L.scale, mkr[tz]);
      res[tz] = fr ? fmaf(1.f - om, ri, os) : 0.f;
      if (UPDATE) acc[0] += fr ? (double)ri * ri : 0.0;
    }
    restrict_tile<TZ>(res, x0, y0, z0, tx, ty, c, rc);
  }
  if (UPDATE) write_partial_at(acc[0], partials, sh, pslot);
}

// Up-sweep at level l: x~ = omega D^-1 r + P e_coarse on the free set, then
// one damped-Jacobi post-smoothing step out = x~ + omega D^-1 (r - A x~).
// At the finest level it also accumulates r.z.
template <int TZ, typename KT, bool FINEST, bool MASK = false>
__device__ __forceinline__ void up_body(const Lv<KT>& L, bool have, int x0, int y0, int z0,
                                        pfsf_dims c, const float* __restrict__ r,
                                        const float* __restrict__ ec, float* __restrict__ out,
                                        const PcgScalars* sc, double* partials, unsigned pslot) {
  // the done flag is tested once the own-cell loads are in flight
  const int done = sc->done;
  __shared__ Tile<TZ> T;
  __shared__ double sh[1][BS / 64];
  double acc[1] = {0.0};
  const pfsf_dims d = L.d;
  const float om = L.omega;
  if (have) {
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int gx = x0 + tx, gy = y0 + ty, sz = d.lx * d.ly;
    const bool col = gx < d.lx && gy < d.ly;
    const int i0 = flat32(gx, gy, z0, d);
    // own-voxel operands, loaded before the barrier (cell 0 stands in for a
    // cell outside the grid; its values are never used)
    bool inb[TZ];
    float rr_[TZ], dgr[TZ];
    uint32_t mkr[TZ];
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      inb[tz] = col && z0 + tz < d.lz;
      const int i = inb[tz] ? i0 + tz * sz : 0;
      rr_[tz] = ld(r, i);
      dgr[tz] = ld(L.dg, i);
      mkr[tz] = MASK ? L.msk[i] : kAllCouplings;
    }
    // the halo loads are in flight before the done flag is tested
    float kk[Stage<TZ>::NP];
    float3 raw[Stage<TZ>::NP];  // r, diagonal, coarse correction
    stage_load<TZ>(L, x0, y0, z0,
                   [&](int j, int gx_, int gy_, int gz_) {
                     return make_float3(ld(r, j), ld(L.dg, j),
                                        ld(ec, flat32(gx_ >> 1, gy_ >> 1, gz_ >> 1, c)));
                   },
                   kk, raw);
    if (done) return;
    stage_put<TZ>(T, kk, raw, [&](float3 v) { return om * v.x * rcp(v.y) + v.z; });
    __syncthreads();
    const int c0 = tcell(tx, ty, 0);
#pragma unroll
    for (int tz = 0; tz < TZ; tz++) {
      const int cc = c0 + tz * HXY;
      const float xi = T.v[cc], ri = rr_[tz];
      const bool fr = is_free_s(T.s[cc]);
      // x + omega D^-1 (r - A x) = (1 - omega) x + omega D^-1 (r + sum_j w_ij x_j)
      const float os = tile_offsum<TZ, MASK>(T, cc, L.scale, mkr[tz]);
      float zi = fmaf(1.f - om, xi, om * (ri + os) * rcp(dgr[tz]));
      zi = fr ? zi : 0.f;
      if (FINEST) acc[0] += fr ? (double)ri * zi : 0.0;
      if (inb[tz]) st(out, i0 + tz * sz, zi);
    }
  } else if (done) {
    return;
  }
  if (FINEST) write_partial_at(acc[0], partials, sh, pslot);
}

// Whole Chebyshev solve of the coarsest level held in one work-group's LDS
// (n <= kMaxLdsCoarse cells, one cell per thread).  The free_s coding of kappa
// and the iterate sit in LDS boxes with a one-cell border (outside the free
// set, value 0), so the
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
// r.r, and the stop flag, which also ends this V-cycle.  With `mirror` the
// test also stores {(iterations << 8) | stop, r.r} into those host-visible
// words, so a host can follow the iteration without copying scalars back.
//
// Every global operand (the cell's kappa, diagonal, anchor mass, mask, rhs
// and the step coefficients) is requested before the partials are summed, so
// the two memory round trips overlap.
__device__ __forceinline__ void publish_progress(unsigned long long* mirror, int iters, int done,
                                                 double rr) {
  // system scope, no release: each word carries its information by itself
  __hip_atomic_store(mirror + 1, (unsigned long long)__double_as_longlong(rr), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(mirror, ((unsigned long long)(unsigned)iters << 8) | (unsigned)done,
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The PCG iteration's convergence test from the r.r partials of the finest
// update (rr_n work-groups): iteration count, r.r and the stop flag (also
// mirrored to the host words when given).  Returns the stop decision, the
// same in every thread of the work-group.
__device__ __forceinline__ bool convergence_test(PcgScalars* sc, const double* __restrict__ rr_part,
                                                 unsigned rr_n, unsigned long long* mirror) {
  const double rr = block_total(rr_part, rr_n);
  const int it = sc->iters + 1;
  const bool stop = rr <= sc->stop2 || it >= sc->max_iters;
  __syncthreads();  // every thread has read the scalars
  if (threadIdx.x == 0) {
    sc->rr = rr;
    sc->iters = it;
    if (stop) sc->done = 1;
    if (mirror) publish_progress(mirror, it, stop ? 1 : 0, rr);
  }
  return stop;
}

// The Chebyshev steps of the coarsest level held in LDS: thread i owns cell
// i (act: i < n) and brings its own operands (kappa, diagonal, anchor mass,
// coupling mask, rhs) and step coefficient c_ld = coef[i] (i < 2 * steps);
// returns the cell's iterate, 0 outside the free set.  Every thread of the
// work-group takes part (barriers).
__device__ __forceinline__ float cheb_lds_run(pfsf_dims d, float scale, bool act, float k_ld,
                                              float D_ld, float m_ld, uint32_t mk_ld, float r_ld,
                                              float c_ld, int steps) {
  __shared__ float ss[kMaxLdsPadded], sx[2][kMaxLdsPadded];
  __shared__ float cf[128];
  const int i = threadIdx.x;
  const int sy = d.lx + 2, sz = sy * (d.ly + 2), np = sz * (d.lz + 2);
  for (int e = threadIdx.x; e < np; e += blockDim.x) {
    ss[e] = -0.0f;
    sx[0][e] = 0.f;
    sx[1][e] = 0.f;
  }
  if (i < 2 * steps) cf[i] = c_ld;
  __syncthreads();
  int e = 0;
  float ki = -1.f, D = 1.f, m = 0.f, ri = 0.f, xi = 0.f, di = 0.f;
  uint32_t mk = kAllCouplings;
  if (act) {
    const int x = i % d.lx, y = (i / d.lx) % d.ly, z = i / (d.lx * d.ly);
    e = (x + 1) + (y + 1) * sy + (z + 1) * sz;
    ki = k_ld;
    if (ki >= 0.f) {
      D = D_ld;
      m = m_ld;
      mk = mk_ld;
      ri = r_ld;
      di = cf[1] * ri * rcp(D);
      xi = di;
    }
    ss[e] = free_s(ki);
    sx[0][e] = xi;
  }
  const bool live = act && ki >= 0.f;
  __syncthreads();
  // the cell's couplings do not change across the steps (an idle thread's
  // e is not a cell of the box: it forms none)
  CellCouplings W;
  if (live) cell_couplings<true>(ss, e, sy, sz, mk, W);
  int cur = 0;
  for (int st = 1; st < steps; st++) {
    const float c1 = cf[2 * st], c2 = cf[2 * st + 1];
    if (live) {
      const float ax = stencil_apply_w(sx[cur], e, sy, sz, W, m, scale);
      di = c1 * di + c2 * (ri - ax) * rcp(D);
      xi = xi + di;
      sx[cur ^ 1][e] = xi;
    }
    __syncthreads();
    cur ^= 1;
  }
  return live ? xi : 0.f;
}

template <typename KT>
__device__ __forceinline__ void cheb_lds_body(const Lv<KT>& L, const float* __restrict__ r,
                                              float* __restrict__ out,
                                              const float* __restrict__ coef, int steps,
                                              PcgScalars* sc, const double* __restrict__ rr_part,
                                              unsigned rr_n,
                                              unsigned long long* mirror = nullptr) {
  const pfsf_dims d = L.d;
  const int n = d.lx * d.ly * d.lz;
  const int i = threadIdx.x;
  const bool act = i < n;
  const int ic = act ? i : 0;  // cell 0 stands in for an idle thread
  const float k_ld = ldk(L.kap, ic);
  const float D_ld = L.dg[ic], m_ld = L.am[ic], r_ld = r[ic];
  const uint32_t mk_ld = L.msk ? L.msk[ic] : kAllCouplings;
  const float c_ld = i < 2 * steps ? coef[i] : 0.f;
  if (sc->done) return;
  if (rr_n > 0 && convergence_test(sc, rr_part, rr_n, mirror)) return;
  const float xi = cheb_lds_run(d, L.scale, act, k_ld, D_ld, m_ld, mk_ld, r_ld, c_ld, steps);
  if (act) out[i] = xi;
}

// ── FP64 coupling of the exact operator ────────────────────────────────
// Face 1/2 fl(ki + kj), edge / corner kE / kC sqrt(fl(ki kj)) in FP64 (the
// reference's float product, root in double).
__device__ __forceinline__ double couple_f64(int cls, float ki, float kj) {
  if (cls == 1) {
    float s = ki + kj;
    return 0.5 * (double)s;
  }
  float p = ki * kj;
  double r = sqrt((double)p);
  return r * (cls == 2 ? (double)kEdgeP : (double)kCornerP);
}

}  // namespace
}  // namespace pfsf
