// Plane-streaming kernels of the finest level (single-system and batched solver).
//
// A work-group owns a 32 x 8 column of the finest grid over TZ planes and
// walks it plane by plane.  The staged value and the coupling coding of kappa
// of each plane (with its one-cell halo ring, 34 x 10 cells) go into one of
// two LDS plane slots; every thread then copies its 3 x 3 neighbourhood of
// that plane into a register window holding the last three planes, so each
// output reads 18 LDS words instead of 52.  The loads of the next PD planes
// are in flight while the current plane's stencil runs, so memory traffic,
// LDS work and stores overlap inside one work-group instead of running as
// one load phase, one compute phase and one store phase over the whole grid.
// Each z-halo plane is loaded once per column rather than once per
// 32 x 8 x TZ tile.
//
// kappa comes in the anchor-coded form of the finest level (free kappa as
// is, -kappa on anchors, -inf elsewhere): an anchor is staged as a coupled
// cell of value 0, so the stencil forms the couplings to anchors itself and
// no anchor-mass array is read.  The operator is that of spmv_body /
// down_body / up_body (same couplings, same operand order, the anchor terms
// summed inside the stencil instead of added as a precomputed mass).
#pragma once

#include <type_traits>
#include <utility>

#include "mg_kernels.h"

namespace pfsf {
namespace {

// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>): the plane
// loop is unrolled by construction, so every ring slot, register buffer and
// parity test is a compile-time constant (no indexed private arrays).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr int kRing = 2;   // LDS plane slots: one read into registers, one written
// planes whose loads are in flight ahead of the staged one
#ifndef PFSF_PD
#define PFSF_PD 2
#endif
constexpr int kPD = PFSF_PD;
constexpr int kHaloCells = 2 * HX + 2 * TY;  // halo ring of one plane (84 cells)
static_assert(kHaloCells <= BS, "one halo cell per thread at most");

struct PlaneRing {
  float v[kRing][HXY];  // staged value (0 off the free set)
  float s[kRing][HXY];  // coupling coding of kappa (see ring_put)
};

// Thread geometry of a plane: the interior cell (tx, ty) and, for threads
// t < 84, one cell of the halo ring (top row, bottom row, left, right).
// mo / mh: the cell bits (TL::cells) of the own cell and of the halo cell
// (none for a thread without one), all set where a column carries none.
struct PlaneGeom {
  int gx, gy, hx, hy;
  bool inxy, hown, hxy;
  int oxy, hxy_off;  // flat offsets of the two cells inside a plane (0 when outside)
  int cxy, chxy;     // their coarse parents' offsets inside a coarse plane (plane_geom_coarse)
  int li, lh;  // LDS offsets inside a plane (and cell-bit words of a column)
  uint32_t mo, mh;
};

__device__ __forceinline__ PlaneGeom plane_geom(int x0, int y0, pfsf_dims d) {
  PlaneGeom g;
  const int t = threadIdx.x, tx = t % TX, ty = t / TX;
  g.gx = x0 + tx;
  g.gy = y0 + ty;
  g.inxy = g.gx < d.lx && g.gy < d.ly;
  g.li = (ty + 1) * HX + tx + 1;
  int ex, ey;
  if (t < HX) {
    ex = t;
    ey = 0;
  } else if (t < 2 * HX) {
    ex = t - HX;
    ey = HY - 1;
  } else if (t < 2 * HX + TY) {
    ex = 0;
    ey = t - 2 * HX + 1;
  } else {
    ex = HX - 1;
    ey = t - 2 * HX - TY + 1;
  }
  g.hown = t < kHaloCells;
  g.hx = x0 + ex - 1;
  g.hy = y0 + ey - 1;
  g.hxy = g.hown && g.hx >= 0 && g.hx < d.lx && g.hy >= 0 && g.hy < d.ly;
  // a plane's flat offsets once per column, so a plane adds only z * lx * ly
  g.oxy = g.inxy ? g.gx + d.lx * g.gy : 0;
  g.hxy_off = g.hxy ? g.hx + d.lx * g.hy : 0;
  g.cxy = g.chxy = 0;
  g.lh = ey * HX + ex;
  g.mo = g.mh = ~0u;
  return g;
}

// The coarse parents' offsets (up-sweep: x >> 1, y >> 1 inside a plane of
// the coarse grid c).
__device__ __forceinline__ void plane_geom_coarse(PlaneGeom& g, pfsf_dims c) {
  g.cxy = g.inxy ? (g.gx >> 1) + c.lx * (g.gy >> 1) : 0;
  g.chxy = g.hxy ? (g.hx >> 1) + c.lx * (g.hy >> 1) : 0;
}

// The cell bits of list slot `slot` into g (none: every cell is read).
__device__ __forceinline__ void cell_bits(const TL& M, int slot, PlaneGeom& g) {
  if (!M.cells) return;
  const uint32_t* w = M.cells + (int64_t)slot * kCellWords;
  g.mo = w[g.li];
  g.mh = g.hown ? w[g.lh] : 0u;
}

// Plane pl of a column: whether a cell's value (a free voxel) and its kappa
// coding (a free voxel or an anchor) must be read.  Columns deeper than the
// cell bits describe read every cell.
template <int TZ>
__device__ __forceinline__ bool cell_free(uint32_t m, int pl) {
  if constexpr (TZ + 2 > kCellMaxPlanes) return true;
  return (m >> pl) & 1u;
}
template <int TZ>
__device__ __forceinline__ bool cell_coupled(uint32_t m, int pl) {
  if constexpr (TZ + 2 > kCellMaxPlanes) return true;
  return (m >> (16 + pl)) & 1u;
}

// A thread's 3 x 3 neighbourhood of one plane, index (dy + 1) * 3 + (dx + 1).
struct Win {
  float v[9], s[9];
};

__device__ __forceinline__ void read_win(const PlaneRing& R, int slot, int c, Win& w) {
#pragma unroll
  for (int dy = -1; dy <= 1; dy++)
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
      w.v[(dy + 1) * 3 + dx + 1] = R.v[slot][c + dy * HX + dx];
      w.s[(dy + 1) * 3 + dx + 1] = R.s[slot][c + dy * HX + dx];
    }
}

// (A v) at the centre of plane window W0 with neighbour planes Wm (z-1) and
// Wp (z+1), in the operand order of stencil_apply; anchors are coupled cells
// of value 0 in the window, so their couplings are part of the sums.
__device__ __forceinline__ float stencil_win(const Win& Wm, const Win& W0, const Win& Wp,
                                             float scale) {
  const float vc = W0.v[4], si = W0.s[4];
  const float ki = si * si;
  float af = 0.f, ae = 0.f, ac = 0.f;
  auto face = [&](const Win& W, int k) {
    const float sj = W.s[k];
    const float w = is_free_s(sj) ? fmaf(sj, sj, ki) : 0.f;
    af += w * (vc - W.v[k]);
  };
  auto edge = [&](const Win& W, int k) { ae += W.s[k] * (vc - W.v[k]); };
  auto corner = [&](const Win& W, int k) { ac += W.s[k] * (vc - W.v[k]); };
  face(W0, 3);  // -x
  face(W0, 5);  // +x
  face(W0, 1);  // -y
  face(W0, 7);  // +y
  face(Wm, 4);  // -z
  face(Wp, 4);  // +z
  edge(W0, 0);
  edge(W0, 2);
  edge(W0, 6);
  edge(W0, 8);
  edge(Wm, 3);
  edge(Wm, 5);
  edge(Wp, 3);
  edge(Wp, 5);
  edge(Wm, 1);
  edge(Wm, 7);
  edge(Wp, 1);
  edge(Wp, 7);
  corner(Wm, 0);
  corner(Wm, 2);
  corner(Wm, 6);
  corner(Wm, 8);
  corner(Wp, 0);
  corner(Wp, 2);
  corner(Wp, 6);
  corner(Wp, 8);
  return scale * (0.5f * af + si * (kEdgeP * ae + kCornerP * ac));
}

// The off-diagonal sum of stencil_win, sum_j w_ij v_j, so that (A v)_i =
// D_i v_i - offsum_win with D_i the level's diagonal (k_level_diag sums the
// same weights over free neighbours and anchors).  Every cell off the free
// set is staged at 0, so its term vanishes without the coupling test of
// stencil_win: one FMA per neighbour (two per face).  The V-cycle kernels use
// it where the centre term cancels in closed form; q = A p keeps stencil_win,
// whose differences are exact for neighbouring values within a factor 2.
__device__ __forceinline__ float offsum_win(const Win& Wm, const Win& W0, const Win& Wp,
                                            float scale) {
  const float si = W0.s[4];
  const float ki = si * si;
  float af = 0.f, ae = 0.f, ac = 0.f;
  auto face = [&](const Win& W, int k) { af = fmaf(fmaf(W.s[k], W.s[k], ki), W.v[k], af); };
  auto edge = [&](const Win& W, int k) { ae = fmaf(W.s[k], W.v[k], ae); };
  auto corner = [&](const Win& W, int k) { ac = fmaf(W.s[k], W.v[k], ac); };
  face(W0, 3);
  face(W0, 5);
  face(W0, 1);
  face(W0, 7);
  face(Wm, 4);
  face(Wp, 4);
  edge(W0, 0);
  edge(W0, 2);
  edge(W0, 6);
  edge(W0, 8);
  edge(Wm, 3);
  edge(Wm, 5);
  edge(Wp, 3);
  edge(Wp, 5);
  edge(Wm, 1);
  edge(Wm, 7);
  edge(Wp, 1);
  edge(Wp, 7);
  corner(Wm, 0);
  corner(Wm, 2);
  corner(Wm, 6);
  corner(Wm, 8);
  corner(Wp, 0);
  corner(Wp, 2);
  corner(Wp, 6);
  corner(Wp, 8);
  return scale * (0.5f * af + si * (kEdgeP * ae + kCornerP * ac));
}

// Flat indices of a plane's two cells (cell 0 stands in for a cell outside
// the grid; its values are never used).
struct PlaneIdx {
  int i, h;
  bool iv, hv;
};
__device__ __forceinline__ PlaneIdx plane_idx(const PlaneGeom& g, int z, pfsf_dims d) {
  PlaneIdx p;
  const bool zin = z >= 0 && z < d.lz;
  p.iv = g.inxy && zin;
  p.hv = g.hxy && zin;
  const int zo = z * (d.lx * d.ly);
  p.i = p.iv ? g.oxy + zo : 0;
  p.h = p.hv ? g.hxy_off + zo : 0;
  return p;
}

// Stages one cell from its anchor-coded kappa k: free cells (k >= 0) keep
// their value, anchors (finite k < 0) and the rest hold 0; the coupling
// coding is sqrt|k| with the sign bit clear for free cells and anchors and
// -0.0 for cells that couple to nothing (k = -inf), so is_free_s() reads
// "coupled" here.
__device__ __forceinline__ void ring_put(PlaneRing& R, int slot, int off, float k, float v) {
  R.v[slot][off] = k >= 0.f ? v : 0.f;
  R.s[slot][off] = k != -INFINITY ? __builtin_amdgcn_sqrtf(fabsf(k)) : -0.0f;
}

// Restriction over the z-pair of planes (zo even, zo + 1) held by one thread
// in `a`, then over the x and y siblings (lanes tx^1 and lane^32).
__device__ __forceinline__ void restrict_pair(float a, int tx, int ty, int gx, int gy, int zo,
                                              pfsf_dims c, float* rc) {
  a = a + __shfl_xor(a, 1, 64);
  a = a + __shfl_xor(a, 32, 64);
  const int X = gx >> 1, Y = gy >> 1, Z = zo >> 1;
  if (!(tx & 1) && !(ty & 1) && X < c.lx && Y < c.ly && Z < c.lz) st(rc, flat32(X, Y, Z, c), a);
}

// Columns of one work-group.  get<TZ>(k, x0, y0, z0) gives its k-th column
// (k = 0, 1, ...): 1 = a column at (x0, y0, z0), 0 = an empty slot of the
// list (nothing to stream), -1 = no columns left.  A work-group carries its
// PCG scalars, LDS ring and partial sum across its columns.
//
// StrideColumns: list slots blockIdx.x + k * gridDim.x of a single system's
// tile list, so a grid of at most a few work-groups per CU walks a large
// grid's columns and the consumers of its partial sums read that many
// partials, not one per column.  gridDim.x is a multiple of 8 and slot p of
// the list belongs to XCD chunk p mod 8, so all columns of a work-group come
// from the chunk of the XCD it runs on.
struct StrideColumns {
  TL M;
  __device__ __forceinline__ void bits(int k, PlaneGeom& g) const {
    cell_bits(M, (int)blockIdx.x + k * (int)gridDim.x, g);
  }
  template <int TZ>
  __device__ __forceinline__ int get(int k, int& x0, int& y0, int& z0) const {
    const int i = (int)blockIdx.x + k * (int)gridDim.x;
    if (i >= M.n) return -1;
    const int t = M.ids[i];
    if (t < 0) return 0;
    const int r = t / M.ntx;
    x0 = (t % M.ntx) * TX;
    y0 = (r % M.nty) * TY;
    z0 = (r / M.nty) * TZ;
    return 1;
  }
};

// SlotColumn: the column of list slot blockIdx.x, one per work-group.
struct SlotColumn {
  TL M;
  __device__ __forceinline__ void bits(int, PlaneGeom& g) const {
    cell_bits(M, (int)blockIdx.x, g);
  }
  template <int TZ>
  __device__ __forceinline__ int get(int k, int& x0, int& y0, int& z0) const {
    if (k > 0) return -1;
    return tile_origin<TZ>(M, x0, y0, z0) ? 1 : 0;
  }
};

// OneColumn: the column of a batch tile record.
struct OneColumn {
  int x0, y0, z0;
  __device__ __forceinline__ void bits(int, PlaneGeom&) const {}
  template <int TZ>
  __device__ __forceinline__ int get(int k, int& a, int& b, int& c) const {
    if (k > 0) return -1;
    a = x0;
    b = y0;
    c = z0;
    return 1;
  }
};

// ── p_out = z + beta p_in; q = A p_out; p.q (see spmv_body) ─────────────
struct SpRaw {
  PlaneIdx ix;
  float k, z, p, hk, hz, hp;
};

// The arrays of one finest-level system as raw buffers (loads through
// 32-bit offsets; see Buf).
struct SpBufs {
  Buf k, z, p, po, q;
};

// Every cell of a plane is read: gating these loads by the column's cell
// bits (as dn_load does) made this kernel and ks_up slower at 128^3 (12.5 ->
// 13.1 us, 11.2 -> 12.4 us), though they read fewer bytes.  Their stores
// cover every cell of a column too: storing only the free cells (whose
// values are the only ones read) left 128^3 unchanged and made the 256^3
// solve 1.4 % slower.
template <int TZ>
__device__ __forceinline__ void sp_load(SpRaw& R, const PlaneGeom& g, int zpl, pfsf_dims d,
                                        const SpBufs& B) {
  R.ix = plane_idx(g, zpl, d);
  R.k = bld(B.k, R.ix.i);
  R.z = bld(B.z, R.ix.i);
  R.p = bld(B.p, R.ix.i);
  R.hk = bld(B.k, R.ix.h);  // halo cell (cell 0 for threads without one: unused)
  R.hz = bld(B.z, R.ix.h);
  R.hp = bld(B.p, R.ix.h);
}

// The columns of `cols` of one system; `publish` (one work-group per system)
// stores the PCG scalars, the p.q partial over the work-group's columns goes
// to partials[pslot].
template <int TZ, typename Cols>
__device__ __forceinline__ void spmv_stream_body(
    const Lv<float>& L, const Cols& cols, const float* __restrict__ z,
    const float* __restrict__ pin, float* __restrict__ pout, float* __restrict__ q, PcgScalars* sc,
    const double* __restrict__ rz_part, unsigned rz_n, int slot, bool publish, double* partials,
    unsigned pslot) {
  __shared__ PlaneRing R;
  __shared__ double sh[1][BS / 64];
  static_assert((TZ + 2) % kRing == 0, "ring slots continue across columns");
  // the first planes' loads do not depend on beta: they are in flight while
  // the done flag and the r.z partials are read
  const int64_t nb = (int64_t)L.d.lx * L.d.ly * L.d.lz * 4;
  const SpBufs B{buf_of(L.kap, nb), buf_of(z, nb), buf_of(pin, nb), buf_of(pout, nb),
                 buf_of(q, nb)};
  int x0 = 0, y0 = 0, z0 = 0;
  int col = cols.template get<TZ>(0, x0, y0, z0);
  PlaneGeom g = plane_geom(x0, y0, L.d);
  SpRaw buf[kPD];
  if (col > 0)
    static_for<kPD>([&](auto S_) {
      constexpr int s = decltype(S_)::value;
      sp_load<TZ>(buf[s], g, z0 - 1 + s, L.d, B);
    });
  if (sc->done) return;
  const bool first = !sc->have_dir;
  const double rz_old = sc->rzs[slot ^ 1];
  const double rz = block_total(rz_part, rz_n);
  const double beta64 = (!first && rz_old > 0.0) ? rz / rz_old : 0.0;
  if (publish && threadIdx.x == 0) {
    sc->rzs[slot] = rz;
    sc->rz = rz;
    sc->beta = beta64;
  }
  double acc = 0.0;
  const float beta = (float)beta64;
  for (int k = 1; col >= 0; k++) {
    if (col > 0) {
      float pv_own[2] = {0.f, 0.f};
      bool fr_own[2] = {false, false};
      PlaneIdx ix_own[2];
      Win W[3];
      static_for<TZ + 2>([&](auto S_) {
        constexpr int s = decltype(S_)::value;
        const SpRaw cur = buf[s % kPD];
        if constexpr (s + kPD < TZ + 2) sp_load<TZ>(buf[s % kPD], g, z0 - 1 + s + kPD, L.d, B);
        const int sl = s % kRing;
        const float ki = cur.ix.iv ? cur.k : -INFINITY;
        const float pv = ki >= 0.f ? cur.z + beta * cur.p : 0.f;
        ring_put(R, sl, g.li, ki, pv);
        if (g.hown) {
          const float kh = cur.ix.hv ? cur.hk : -INFINITY;
          ring_put(R, sl, g.lh, kh, cur.hz + beta * cur.hp);
        }
        if constexpr (s >= 1 && s <= TZ) {
          if (cur.ix.iv) bst(B.po, cur.ix.i, pv);
          pv_own[s & 1] = pv;
          fr_own[s & 1] = ki >= 0.f;
          ix_own[s & 1] = cur.ix;
        }
        __syncthreads();
        read_win(R, sl, g.li, W[s % 3]);
        if constexpr (s >= 2) {
          const int o = (s - 1) & 1;
          float qi = stencil_win(W[(s - 2) % 3], W[(s - 1) % 3], W[s % 3], L.scale);
          qi = fr_own[o] ? qi : 0.f;
          acc += (double)pv_own[o] * qi;
          if (ix_own[o].iv) bst(B.q, ix_own[o].i, qi);
        }
      });
    }
    col = cols.template get<TZ>(k, x0, y0, z0);
    if (col > 0) {
      g = plane_geom(x0, y0, L.d);
      static_for<kPD>([&](auto S_) {
        constexpr int s = decltype(S_)::value;
        sp_load<TZ>(buf[s], g, z0 - 1 + s, L.d, B);
      });
    }
  }
  write_partial_at(acc, partials, sh, pslot);
}

// LOOP: the work-groups walk the list in strides (StrideColumns); else one
// column per work-group (SlotColumn).
template <int TZ, bool LOOP>
__global__ void __launch_bounds__(BS)
    ks_spmv(Lv<float> L, TL M, const float* __restrict__ z, const float* __restrict__ pin,
            float* __restrict__ pout, float* __restrict__ q, PcgScalars* sc,
            const double* __restrict__ rz_part, unsigned rz_n, int slot, double* partials) {
  pin_args(L, M, z, pin, pout, q, sc, rz_part, rz_n, slot, partials);
  if constexpr (LOOP)
    spmv_stream_body<TZ>(L, StrideColumns{M}, z, pin, pout, q, sc, rz_part, rz_n, slot,
                         blockIdx.x == 0, partials, blockIdx.x);
  else
    spmv_stream_body<TZ>(L, SlotColumn{M}, z, pin, pout, q, sc, rz_part, rz_n, slot,
                         blockIdx.x == 0, partials, blockIdx.x);
}

// ── down-sweep of the finest level (see down_body) ─────────────────────
// With UPDATE: r_out = r_in - alpha q (stored as each owned plane is
// staged), r.r partials; then x = omega D^-1 r_out and the coarse rhs
// P^T (r_out - A x).  The FP64 iterate x64 is updated by mode:
//   kXNow   x64 += alpha p;
//   kXDefer (even iterations of a pass) alpha p stays pending (PcgScalars::xa);
//   kXPair  (odd iterations) x64 + xa pprev + alpha p, pprev being the even
//           iteration's direction, still in the other direction buffer,
// with the FMAs of single updates, so pairs give the same iterate while
// x64 and the directions are streamed every other iteration only.
enum XMode { kXNow = 0, kXDefer = 1, kXPair = 2 };

struct DnRaw {
  PlaneIdx ix;
  float k, r, q, dg, p, pp, hk, hr, hq, hdg;
  double x;
  bool fo;  // the own cell is a free voxel (its updates are stored)
};

// Only the cells a value or a coupling comes from are read (the column's
// cell bits): a cell that is not free takes the values staged for such cells
// (0, kappa -inf where nothing couples, diagonal 1), and its updates are not
// stored (x and r stay as they are there).  A lane's loads sit behind its
// own bits, so a wave issues only the loads of its free cells and the line
// segments that hold none are not fetched.
template <int TZ, bool UPDATE, int XM>
__device__ __forceinline__ void dn_load(DnRaw& R, const PlaneGeom& g, int zpl, int pl,
                                        const Lv<float>& L, const float* rin, const float* q,
                                        const float* p, const float* pprev, const double* x) {
  R.ix = plane_idx(g, zpl, L.d);
  const bool fo = cell_free<TZ>(g.mo, pl), co = cell_coupled<TZ>(g.mo, pl);
  const bool fh = cell_free<TZ>(g.mh, pl), ch = cell_coupled<TZ>(g.mh, pl);
  R.fo = fo;
  R.k = R.hk = -INFINITY;
  R.r = R.q = R.p = R.pp = R.hr = R.hq = 0.f;
  R.dg = R.hdg = 1.f;
  R.x = 0.0;
  if (co) R.k = ld(L.kap, R.ix.i);
  if (fo) {
    R.r = ld(rin, R.ix.i);
    if (UPDATE) R.q = ld(q, R.ix.i);
    R.dg = ld(L.dg, R.ix.i);
    if (UPDATE && XM != kXDefer) {
      R.p = ld(p, R.ix.i);
      if (XM == kXPair) R.pp = ld(pprev, R.ix.i);
      R.x = ld(x, R.ix.i);
    }
  }
  if (ch) R.hk = ld(L.kap, R.ix.h);  // halo cell (cell 0 for threads without one: unused)
  if (fh) {
    R.hr = ld(rin, R.ix.h);
    if (UPDATE) R.hq = ld(q, R.ix.h);
    R.hdg = ld(L.dg, R.ix.h);
  }
}

template <int TZ, bool UPDATE, int XM, typename Cols>
__device__ __forceinline__ void down_stream_body(
    const Lv<float>& L, const Cols& cols, pfsf_dims c,
    const float* __restrict__ rin, float* __restrict__ rout, const float* __restrict__ q,
    const float* __restrict__ p, const float* __restrict__ pprev, double* __restrict__ x,
    float* __restrict__ rc, PcgScalars* sc,
    const double* __restrict__ pq_part, unsigned pq_n, bool publish, double* partials,
    unsigned pslot) {
  __shared__ PlaneRing R;
  __shared__ double sh[1][BS / 64];
  static_assert((TZ + 2) % kRing == 0, "ring slots continue across columns");
  // the first planes' loads do not depend on alpha: they are in flight while
  // the done flag and the p.q partials are read
  int x0 = 0, y0 = 0, z0 = 0;
  int col = cols.template get<TZ>(0, x0, y0, z0);
  PlaneGeom g = plane_geom(x0, y0, L.d);
  if (col > 0) cols.bits(0, g);
  DnRaw buf[kPD];
  if (col > 0)
    static_for<kPD>([&](auto S_) {
      constexpr int s = decltype(S_)::value;
      dn_load<TZ, UPDATE, XM>(buf[s], g, z0 - 1 + s, s, L, rin, q, p, pprev, x);
    });
  if (sc->done) return;
  double alpha64 = 0.0, xa = 0.0;
  if (UPDATE) {
    const double rz = sc->rz;
    const double pq = block_total(pq_part, pq_n);
    const bool ok = pq > 0.0;
    alpha64 = ok ? rz / pq : 0.0;
    // written by the even iteration before (kXPair) or by this one (kXDefer)
    if constexpr (XM == kXPair) xa = sc->xa;
    if (publish && threadIdx.x == 0) {
      sc->pq = pq;
      sc->alpha = alpha64;
      if constexpr (XM == kXDefer) sc->xa = alpha64;
      sc->have_dir = 1;
      if (!ok) sc->done = 2;  // breakdown: the direction carries no energy
    }
    if (!ok) return;  // an update left pending stays so (k_x_pending)
  }
  double acc = 0.0;
  const float alpha = (float)alpha64, om = L.omega, om1 = 1.f - om;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  for (int k = 1; col >= 0; k++) {
    if (col > 0) {
      float r_own[2] = {0.f, 0.f};
      bool fr_own[2] = {false, false};
      float pair = 0.f;
      Win W[3];
      static_for<TZ + 2>([&](auto S_) {
        constexpr int s = decltype(S_)::value;
        const DnRaw cur = buf[s % kPD];
        if constexpr (s + kPD < TZ + 2)
          dn_load<TZ, UPDATE, XM>(buf[s % kPD], g, z0 - 1 + s + kPD, s + kPD, L, rin, q, p,
                                  pprev, x);
        const int sl = s % kRing;
        const float ki = cur.ix.iv ? cur.k : -INFINITY;
        const float ri = UPDATE ? cur.r - alpha * cur.q : cur.r;
        ring_put(R, sl, g.li, ki, om * ri * rcp(cur.dg));
        if (g.hown) {
          const float kh = cur.ix.hv ? cur.hk : -INFINITY;
          const float rh = UPDATE ? cur.hr - alpha * cur.hq : cur.hr;
          ring_put(R, sl, g.lh, kh, om * rh * rcp(cur.hdg));
        }
        if constexpr (s >= 1 && s <= TZ) {
          const bool fr = ki >= 0.f;
          if (UPDATE) {
            // r and x change on the free set only
            if (cur.ix.iv && cur.fo) {
              if constexpr (XM == kXPair)
                st(x, cur.ix.i, fma(alpha64, (double)cur.p, fma(xa, (double)cur.pp, cur.x)));
              else if constexpr (XM == kXNow)
                st(x, cur.ix.i, fma(alpha64, (double)cur.p, cur.x));
              st(rout, cur.ix.i, ri);
            }
            acc += fr ? (double)ri * ri : 0.0;
          }
          r_own[s & 1] = ri;
          fr_own[s & 1] = fr;
        }
        __syncthreads();
        read_win(R, sl, g.li, W[s % 3]);
        if constexpr (s >= 2) {
          const int o = (s - 1) & 1;
          // r - A x with x = omega D^-1 r: (1 - omega) r + sum_j w_ij x_j
          const float os = offsum_win(W[(s - 2) % 3], W[(s - 1) % 3], W[s % 3], L.scale);
          const float res = fr_own[o] ? fmaf(om1, r_own[o], os) : 0.f;
          const int zo = z0 + s - 2;  // z0 is even: zo has the parity of s
          if constexpr ((s & 1) == 0) {
            pair = res;
          } else {
            restrict_pair(pair + res, tx, ty, g.gx, g.gy, zo - 1, c, rc);
          }
        }
      });
    }
    col = cols.template get<TZ>(k, x0, y0, z0);
    if (col > 0) {
      g = plane_geom(x0, y0, L.d);
      cols.bits(k, g);
      static_for<kPD>([&](auto S_) {
        constexpr int s = decltype(S_)::value;
        dn_load<TZ, UPDATE, XM>(buf[s], g, z0 - 1 + s, s, L, rin, q, p, pprev, x);
      });
    }
  }
  if (UPDATE) write_partial_at(acc, partials, sh, pslot);
}

// One column per work-group: every column of a grid up to 1,024 columns is
// resident at four waves per SIMD, so registers stay within 128 (the pair
// update's two directions would take 129).  The column-loop form runs three
// work-groups per CU by design.
template <int TZ, bool UPDATE, bool LOOP, int XM = kXNow>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(LOOP ? 1 : PFSF_WPE)))
    ks_down(Lv<float> L, TL M, pfsf_dims c, const float* __restrict__ rin,
            float* __restrict__ rout, const float* __restrict__ q, const float* __restrict__ p,
            const float* __restrict__ pprev, double* __restrict__ x, float* __restrict__ rc,
            PcgScalars* sc, const double* __restrict__ pq_part, unsigned pq_n,
            double* partials) {
  pin_args(L, M, c, rin, rout, q, p, pprev, x, rc, sc, pq_part, pq_n, partials);
  if constexpr (LOOP)
    down_stream_body<TZ, UPDATE, XM>(L, StrideColumns{M}, c, rin, rout, q, p, pprev, x, rc, sc,
                                     pq_part, pq_n, blockIdx.x == 0, partials, blockIdx.x);
  else
    down_stream_body<TZ, UPDATE, XM>(L, SlotColumn{M}, c, rin, rout, q, p, pprev, x, rc, sc,
                                     pq_part, pq_n, blockIdx.x == 0, partials, blockIdx.x);
}

// ── up-sweep of the finest level (see up_body) ─────────────────────────
// x~ = omega D^-1 r + P e_coarse on the free set; out = x~ + omega D^-1
// (r - A x~); with FINEST the r.z partials.
struct UpRaw {
  PlaneIdx ix;
  float k, r, dg, ec, hk, hr, hdg, hec;
};

struct UpBufs {
  Buf k, r, dg, ec, out;
};

template <int TZ>
__device__ __forceinline__ void up_load(UpRaw& R, const PlaneGeom& g, int zpl, pfsf_dims d,
                                        pfsf_dims c, const UpBufs& B) {
  R.ix = plane_idx(g, zpl, d);
  const int zco = (zpl >> 1) * (c.lx * c.ly);  // used only for planes inside the grid
  R.k = bld(B.k, R.ix.i);
  R.r = bld(B.r, R.ix.i);
  R.dg = bld(B.dg, R.ix.i);
  R.ec = bld(B.ec, R.ix.iv ? g.cxy + zco : 0);
  R.hk = bld(B.k, R.ix.h);  // halo cell (cell 0 for threads without one: unused)
  R.hr = bld(B.r, R.ix.h);
  R.hdg = bld(B.dg, R.ix.h);
  R.hec = bld(B.ec, R.ix.hv ? g.chxy + zco : 0);
}

template <int TZ, bool FINEST, typename Cols>
__device__ __forceinline__ void up_stream_body(const Lv<float>& L, const Cols& cols,
                                               pfsf_dims c, const float* __restrict__ r,
                                               const float* __restrict__ ec,
                                               float* __restrict__ out, const PcgScalars* sc,
                                               double* partials, unsigned pslot) {
  __shared__ PlaneRing R;
  __shared__ double sh[1][BS / 64];
  static_assert((TZ + 2) % kRing == 0, "ring slots continue across columns");
  const int64_t nb = (int64_t)L.d.lx * L.d.ly * L.d.lz * 4;
  const UpBufs B{buf_of(L.kap, nb), buf_of(r, nb), buf_of(L.dg, nb),
                 buf_of(ec, (int64_t)c.lx * c.ly * c.lz * 4), buf_of(out, nb)};
  int x0 = 0, y0 = 0, z0 = 0;
  int col = cols.template get<TZ>(0, x0, y0, z0);
  PlaneGeom g = plane_geom(x0, y0, L.d);
  plane_geom_coarse(g, c);
  UpRaw buf[kPD];
  if (col > 0)  // in flight while the done flag is read
    static_for<kPD>([&](auto S_) {
      constexpr int s = decltype(S_)::value;
      up_load<TZ>(buf[s], g, z0 - 1 + s, L.d, c, B);
    });
  if (sc->done) return;
  double acc = 0.0;
  const float om = L.omega, om1 = 1.f - om;
  for (int k = 1; col >= 0; k++) {
    if (col > 0) {
      float r_own[2] = {0.f, 0.f}, dg_own[2] = {1.f, 1.f};
      bool fr_own[2] = {false, false};
      PlaneIdx ix_own[2];
      Win W[3];
      static_for<TZ + 2>([&](auto S_) {
        constexpr int s = decltype(S_)::value;
        const UpRaw cur = buf[s % kPD];
        if constexpr (s + kPD < TZ + 2) up_load<TZ>(buf[s % kPD], g, z0 - 1 + s + kPD, L.d, c, B);
        const int sl = s % kRing;
        const float ki = cur.ix.iv ? cur.k : -INFINITY;
        ring_put(R, sl, g.li, ki, om * cur.r * rcp(cur.dg) + cur.ec);
        if (g.hown) {
          const float kh = cur.ix.hv ? cur.hk : -INFINITY;
          ring_put(R, sl, g.lh, kh, om * cur.hr * rcp(cur.hdg) + cur.hec);
        }
        if constexpr (s >= 1 && s <= TZ) {
          r_own[s & 1] = cur.r;
          dg_own[s & 1] = cur.dg;
          fr_own[s & 1] = ki >= 0.f;
          ix_own[s & 1] = cur.ix;
        }
        __syncthreads();
        read_win(R, sl, g.li, W[s % 3]);
        if constexpr (s >= 2) {
          const int o = (s - 1) & 1;
          // x + omega D^-1 (r - A x) = (1 - omega) x + omega D^-1 (r + sum_j w_ij x_j)
          const float os = offsum_win(W[(s - 2) % 3], W[(s - 1) % 3], W[s % 3], L.scale);
          const float xi = W[(s - 1) % 3].v[4];
          float zi = fmaf(om1, xi, om * (r_own[o] + os) * rcp(dg_own[o]));
          zi = fr_own[o] ? zi : 0.f;
          if (FINEST) acc += fr_own[o] ? (double)r_own[o] * zi : 0.0;
          if (ix_own[o].iv) bst(B.out, ix_own[o].i, zi);
        }
      });
    }
    col = cols.template get<TZ>(k, x0, y0, z0);
    if (col > 0) {
      g = plane_geom(x0, y0, L.d);
      plane_geom_coarse(g, c);
      static_for<kPD>([&](auto S_) {
        constexpr int s = decltype(S_)::value;
        up_load<TZ>(buf[s], g, z0 - 1 + s, L.d, c, B);
      });
    }
  }
  if (FINEST) write_partial_at(acc, partials, sh, pslot);
}

template <int TZ, bool FINEST, bool LOOP>
__global__ void __launch_bounds__(BS)
    ks_up(Lv<float> L, TL M, pfsf_dims c, const float* __restrict__ r,
          const float* __restrict__ ec, float* __restrict__ out, const PcgScalars* sc,
          double* partials) {
  pin_args(L, M, c, r, ec, out, sc, partials);
  if constexpr (LOOP)
    up_stream_body<TZ, FINEST>(L, StrideColumns{M}, c, r, ec, out, sc, partials, blockIdx.x);
  else
    up_stream_body<TZ, FINEST>(L, SlotColumn{M}, c, r, ec, out, sc, partials, blockIdx.x);
}

// ── FP64 true residual of the iterate, plane streaming ─────────────────
// r = b - A x over the columns of `cols` with the exact coupling rule
// (couple_f64: face 1/2 fl(ki + kj), edge / corner kE / kC sqrt(fl(ki kj))
// in FP64, every non-air neighbour coupled, x 0 off the free set), written
// as r32 = (float) r on free cells, plus the work-group's ||r||^2 partial;
// with xzero (the iterate is 0 everywhere) r = b without the stencil.
//
// A plane's cells are staged once with their coupling kappa (-1 on air),
// 1 / kappa, FP64 sqrt(kappa) and x, so a coupling takes no square root:
// with p = fl(ki kj) and e = ki kj - p, the exact error of that FP32 product
// (one FMA), sqrt(p) = sqrt(ki) sqrt(kj) (1 - t / 2 - t^2 / 8 - ...) for
// t = e / (ki kj), |t| < 2^-23.  The first-order form is within 1e-14 of
// the correctly rounded root (t^2 / 8 and the 1-ulp FP32 reciprocals);
// the convergence test needs the residual to far better than its 1e-6
// target, not to the last bit.  Air stages 1 / kappa = sqrt = x = 0, so an
// edge or corner term vanishes without a test; a face keeps one (its weight
// 1/2 (ki + kj) does not vanish by itself).  Four LDS plane slots take one
// barrier per plane: the slot written at plane s was last read two planes
// earlier.
struct ResRing {
  float2 kr[4][HXY];   // kappa (-1 on air), 1 / kappa (0 on air or kappa 0)
  double2 xs[4][HXY];  // x (0 off the free set), sqrt(kappa) (0 on air)
};

struct ResRaw {
  PlaneIdx ix;
  float k, b, hk;
  double x, hx;
  bool fo, ho, ft, ht;  // free (own, halo); non-air (own, halo)
};

__device__ __forceinline__ void res_load(ResRaw& R, const PlaneGeom& g, int zpl, pfsf_dims d,
                                         const uint8_t* __restrict__ type,
                                         const float* __restrict__ kappa,
                                         const uint8_t* __restrict__ fr,
                                         const float* __restrict__ b,
                                         const double* __restrict__ x) {
  R.ix = plane_idx(g, zpl, d);
  R.k = ld(kappa, R.ix.i);
  R.b = ld(b, R.ix.i);
  R.x = ld(x, R.ix.i);
  R.ft = ld(type, R.ix.i) != 0;
  R.fo = ld(fr, R.ix.i) != 0;
  R.hk = ld(kappa, R.ix.h);  // halo cell (cell 0 for threads without one: unused)
  R.hx = ld(x, R.ix.h);
  R.ht = ld(type, R.ix.h) != 0;
  R.ho = ld(fr, R.ix.h) != 0;
}

__device__ __forceinline__ void res_put(ResRing& T, int sl, int off, bool in, bool nonair,
                                        bool fre, float kv, double xv) {
  const float k = (in && nonair) ? kapf(kv) : -1.f;
  T.kr[sl][off] = make_float2(k, k > 0.f ? 1.f / k : 0.f);
  T.xs[sl][off] = make_double2((in && fre) ? xv : 0.0, k >= 0.f ? sqrt((double)k) : 0.0);
}

template <int TZ, typename Cols>
__device__ __forceinline__ void residual_stream_body(
    pfsf_dims d, const Cols& cols, const uint8_t* __restrict__ type,
    const float* __restrict__ kappa, const uint8_t* __restrict__ fr, const float* __restrict__ b,
    const double* __restrict__ x, float* __restrict__ r32, double* partials, unsigned pslot,
    bool xzero) {
  __shared__ ResRing T;
  __shared__ double sh[1][BS / 64];
  double acc = 0.0;
  int x0 = 0, y0 = 0, z0 = 0;
  int col = cols.template get<TZ>(0, x0, y0, z0);
  for (int k = 1; col >= 0; k++) {
    if (col > 0) {
      const PlaneGeom g = plane_geom(x0, y0, d);
      if (xzero) {
        for (int tz = 0; tz < TZ; tz++) {
          const PlaneIdx ix = plane_idx(g, z0 + tz, d);
          if (ix.iv && ld(fr, ix.i)) {
            const double ri = (double)ld(b, ix.i);
            st(r32, ix.i, (float)ri);
            acc += ri * ri;
          }
        }
      } else {
        // a run-time plane loop: unrolled, the compiler interleaves planes
        // and needs ~148 VGPRs at TZ = 8 (spills below 128)
        ResRaw cur, nx1, nx2;
        res_load(cur, g, z0 - 1, d, type, kappa, fr, b, x);
        res_load(nx1, g, z0, d, type, kappa, fr, b, x);
        bool f_prev = false;
        float b_prev = 0.f;
        int i_prev = 0;
#pragma unroll 1
        for (int s = 0; s < TZ + 2; s++) {
          if (s + 2 < TZ + 2) res_load(nx2, g, z0 + 1 + s, d, type, kappa, fr, b, x);
          const int sl = s & 3;
          res_put(T, sl, g.li, cur.ix.iv, cur.ft, cur.fo, cur.k, cur.x);
          if (g.hown) res_put(T, sl, g.lh, cur.ix.hv, cur.ht, cur.ho, cur.hk, cur.hx);
          __syncthreads();
          if (s >= 2 && f_prev) {
            // plane s - 1: slots (s - 2) & 3, (s - 1) & 3, s & 3 for dz = -1, 0, 1
            const int c = g.li;
            const int sc = (s - 1) & 3;
            const float2 kri = T.kr[sc][c];
            const double2 xsi = T.xs[sc][c];
            const float ki = kri.x, hi = -0.5f * kri.y;
            const double xi = xsi.x, si = xsi.y;
            double af = 0.0, ae = 0.0, ac = 0.0;
            static_for<27>([&](auto N_) {
              constexpr int n = decltype(N_)::value;
              constexpr int dx = n % 3 - 1, dy = (n / 3) % 3 - 1, dz = n / 9 - 1;
              constexpr int cls = (dx != 0) + (dy != 0) + (dz != 0);
              if constexpr (cls != 0) {
                const int slot = (s - 1 + dz) & 3;
                const int cj = c + dy * HX + dx;
                const float2 krj = T.kr[slot][cj];
                const double2 xsj = T.xs[slot][cj];
                const double dxj = xi - xsj.x;
                if constexpr (cls == 1) {
                  const float w = krj.x >= 0.f ? ki + krj.x : 0.f;
                  af = fma((double)w, dxj, af);
                } else {
                  const float p = ki * krj.x;
                  const float e = fmaf(ki, krj.x, -p);
                  const float t = (e * hi) * krj.y;
                  const double sq = si * xsj.y;
                  const double w = fma(sq, (double)t, sq);
                  if constexpr (cls == 2)
                    ae = fma(w, dxj, ae);
                  else
                    ac = fma(w, dxj, ac);
                }
              }
            });
            const double ax = 0.5 * af + (double)kEdgeP * ae + (double)kCornerP * ac;
            const double ri = (double)b_prev - ax;
            st(r32, i_prev, (float)ri);
            acc += ri * ri;
          }
          f_prev = s >= 1 && s <= TZ && cur.ix.iv && cur.fo;
          b_prev = cur.b;
          i_prev = cur.ix.i;
          cur = nx1;
          nx1 = nx2;
        }
        __syncthreads();  // the next column's first planes reuse the slots read last
      }
    }
    col = cols.template get<TZ>(k, x0, y0, z0);
  }
  write_partial_at(acc, partials, sh, pslot);
}

}  // namespace
}  // namespace pfsf
