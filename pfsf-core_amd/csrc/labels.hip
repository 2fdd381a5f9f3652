// Connected components of a voxel grid by concurrent union-find, the
// component records built on them, and the anchored-voxel mask.
//
// Reference behaviour.  LabelPropagation.shiloachVishkin
// (LabelPropagation.java:199-270) iterates hook-to-min and pointer jumping
// to its fixed point, at which every live voxel (solid or anchor) carries
// 1 + the smallest flat index of its component and air carries 0xFFFFFFFF.
// The GPU pipeline label_prop_{init,iterate,summarise_alloc,
// summarise_aggregate}.comp.glsl (recorded by PFSFLabelPropRecorder.java:
// 75-190, replayed by PFSFLabelPropCpuSimulator.java:82-238) then gives each
// root a 12-word record {root label, count, anchored, 0, min xyz, 0, max xyz,
// 0}, at most max_components of them, and flags the overflow.
//
// Here the fixed point is reached without iterating to convergence:
//   1. k_cc_tile    each 32x8x4 tile unions its live voxels in LDS (CAS on
//                   LDS words) and writes, for every voxel, the global index
//                   of its tile-local root;
//   2. k_cc_cross   voxels on a tile face union with their live neighbours in
//                   other tiles on the global parent array (agent-scope CAS);
//   3. k_cc_flatten every voxel's parent becomes its root.
// A link always hangs the larger root under the smaller one, so parent[i] <=
// i throughout and the surviving root of a component is its smallest index,
// whatever order the work-groups run in: labels are deterministic and equal
// the Shiloach-Vishkin fixed point.  Roots take record slots in ascending
// flat order (the order of the CPU replay; the shader's atomic counter gives
// no fixed order), and the per-record sums, ors, mins and maxes are
// order-independent, so the records are deterministic too.
//
// Every union loop ends: a failed CAS means the target root was linked by
// another thread, which happens at most once per root.  The loops still
// carry an iteration guard that raises an error flag instead of spinning.
#include <stdint.h>

#include "pfsf_internal.h"

namespace pfsf {
namespace {

constexpr int LTX = 32, LTY = 8, LTZ = 4, LT = LTX * LTY * LTZ;
constexpr int LBS = 256;                  // threads per tile work-group
constexpr int SEG = 1024;                 // flat cells per summarise work-group
constexpr int AGG_ROUNDS = 8;             // SEG-sized rounds per aggregate work-group
constexpr uint32_t kNone = 0xFFFFFFFFu;   // air / unset
constexpr uint32_t kFlag = 0x80000000u;   // anchored mark on a root's parent word
constexpr uint32_t kIdx = 0x7FFFFFFFu;
constexpr uint32_t kGlobalGuard = 1u << 16;

// Backward neighbours (smaller flat index): every adjacent pair is united
// once, from its larger member.  Faces first, so conn = 6 takes 3 entries.
__constant__ int8_t kBack[13][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1},  {-1, -1, 0}, {1, -1, 0},
                                    {-1, 0, -1}, {1, 0, -1}, {0, -1, -1}, {0, 1, -1},
                                    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}};

// live rule 0: solid or anchor (label propagation); 1: any non-air type
// (reachability over non-air voxels, as the solver's free set uses).
__device__ __forceinline__ bool live_of(uint8_t t, int rule) {
  return rule ? t != 0 : (t == 1 || t == 2);
}

struct LdsOps {
  uint32_t* a;
  __device__ uint32_t ld(uint32_t i) const {
    return __hip_atomic_load(a + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ void st(uint32_t i, uint32_t v) const {
    __hip_atomic_store(a + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ bool cas(uint32_t i, uint32_t expect, uint32_t v) const {
    return __hip_atomic_compare_exchange_strong(a + i, &expect, v, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
};

struct GlobalOps {
  uint32_t* a;
  __device__ uint32_t ld(uint32_t i) const {
    return __hip_atomic_load(a + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ void st(uint32_t i, uint32_t v) const {
    __hip_atomic_store(a + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ bool cas(uint32_t i, uint32_t expect, uint32_t v) const {
    return __hip_atomic_compare_exchange_strong(a + i, &expect, v, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
};

// Root of v with path halving.  Parents only ever move to smaller ancestors,
// so every step strictly decreases the index and the walk ends.
template <typename Ops>
__device__ __forceinline__ uint32_t find_root(const Ops& o, uint32_t v) {
  uint32_t p = o.ld(v);
  while (p != v) {
    const uint32_t g = o.ld(p);
    if (g == p) return p;
    o.st(v, g);
    v = g;
    p = o.ld(v);
  }
  return v;
}

// Union of the sets of a and b: the larger root is hung under the smaller.
// Returns false only if the guard ran out (never expected).
template <typename Ops>
__device__ __forceinline__ bool unite(const Ops& o, uint32_t a, uint32_t b, uint32_t guard) {
  for (uint32_t it = 0; it < guard; it++) {
    a = find_root(o, a);
    b = find_root(o, b);
    if (a == b) return true;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    if (o.cas(a, a, b)) return true;
  }
  return false;
}

__device__ __forceinline__ void tile_of(int b, int ntx, int nty, int& x0, int& y0, int& z0) {
  x0 = (b % ntx) * LTX;
  y0 = ((b / ntx) % nty) * LTY;
  z0 = (b / (ntx * nty)) * LTZ;
}

// Pass 1: unions inside one tile in LDS; parent[i] = tile-local root (global
// flat index) for live voxels, kNone otherwise.
__global__ void __launch_bounds__(LBS)
    k_cc_tile(pfsf_dims d, const uint8_t* __restrict__ type, int rule, int conn, int ntx, int nty,
              uint32_t* __restrict__ par, int* __restrict__ err) {
  __shared__ uint32_t lp[LT];
  int x0, y0, z0;
  tile_of(blockIdx.x, ntx, nty, x0, y0, z0);
  const int tx = threadIdx.x % LTX, ty = threadIdx.x / LTX;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool col = gx < d.lx && gy < d.ly;
  const uint32_t sz = (uint32_t)d.lx * (uint32_t)d.ly;
  const uint32_t g0 = (uint32_t)gx + (uint32_t)d.lx * (uint32_t)gy + sz * (uint32_t)z0;
  bool in[LTZ], lv[LTZ];
#pragma unroll
  for (int k = 0; k < LTZ; k++) {
    in[k] = col && z0 + k < d.lz;
    lv[k] = in[k] && live_of(type[g0 + k * sz], rule);
    lp[tx + LTX * (ty + LTY * k)] = lv[k] ? (uint32_t)(tx + LTX * (ty + LTY * k)) : kNone;
  }
  __syncthreads();
  const LdsOps o{lp};
  const int nb = conn == 6 ? 3 : 13;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < LTZ; k++) {
    if (!lv[k]) continue;
    const uint32_t c = tx + LTX * (ty + LTY * k);
    for (int e = 0; e < nb; e++) {
      const int nx = tx + kBack[e][0], ny = ty + kBack[e][1], nz = k + kBack[e][2];
      if (nx < 0 || nx >= LTX || ny < 0 || ny >= LTY || nz < 0) continue;
      const uint32_t nc = nx + LTX * (ny + LTY * nz);
      if (o.ld(nc) == kNone) continue;  // air, or outside the grid
      ok &= unite(o, c, nc, LT);
    }
  }
  if (!ok) *err = 1;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < LTZ; k++) {
    if (!in[k]) continue;
    uint32_t v = kNone;
    if (lv[k]) {
      const uint32_t r = find_root(o, tx + LTX * (ty + LTY * k));
      const uint32_t rx = r % LTX, ry = (r / LTX) % LTY, rz = r / (LTX * LTY);
      v = (uint32_t)(x0 + rx) + (uint32_t)d.lx * (uint32_t)(y0 + ry) + sz * (uint32_t)(z0 + rz);
    }
    par[g0 + k * sz] = v;
  }
}

// Pass 2: unions across tile faces on the global parent array.
__global__ void __launch_bounds__(LBS)
    k_cc_cross(pfsf_dims d, const uint8_t* __restrict__ type, int rule, int conn, int ntx,
               int nty, uint32_t* par, int* __restrict__ err) {
  int x0, y0, z0;
  tile_of(blockIdx.x, ntx, nty, x0, y0, z0);
  const int tx = threadIdx.x % LTX, ty = threadIdx.x / LTX;
  const int gx = x0 + tx, gy = y0 + ty;
  if (gx >= d.lx || gy >= d.ly) return;
  // only the tile's low faces, and its high x/y faces (offsets +1 in x or
  // y with a negative y or z component), have backward neighbours outside
  const bool edge_xy = tx == 0 || ty == 0 || tx == LTX - 1 || ty == LTY - 1;
  const uint32_t sz = (uint32_t)d.lx * (uint32_t)d.ly;
  const GlobalOps o{par};
  const int nb = conn == 6 ? 3 : 13;
  bool ok = true;
  for (int k = 0; k < LTZ; k++) {
    const int gz = z0 + k;
    if (gz >= d.lz) break;
    if (!edge_xy && k != 0) continue;
    const uint32_t i = (uint32_t)gx + (uint32_t)d.lx * (uint32_t)gy + sz * (uint32_t)gz;
    if (!live_of(type[i], rule)) continue;
    for (int e = 0; e < nb; e++) {
      const int lx_ = tx + kBack[e][0], ly_ = ty + kBack[e][1], lz_ = k + kBack[e][2];
      if (lx_ >= 0 && lx_ < LTX && ly_ >= 0 && ly_ < LTY && lz_ >= 0) continue;  // same tile
      const int nx = gx + kBack[e][0], ny = gy + kBack[e][1], nz = gz + kBack[e][2];
      if (nx < 0 || nx >= d.lx || ny < 0 || ny >= d.ly || nz < 0) continue;
      const uint32_t j = (uint32_t)nx + (uint32_t)d.lx * (uint32_t)ny + sz * (uint32_t)nz;
      if (!live_of(type[j], rule)) continue;
      ok &= unite(o, i, j, kGlobalGuard);
    }
  }
  if (!ok) *err = 1;
}

// Pass 3: parent = root.  The walk does not halve paths: a thread writes
// only its own entry, so no stale ancestor can overwrite a finished one.
__global__ void __launch_bounds__(LBS) k_cc_flatten(uint32_t n, uint32_t* par) {
  const uint32_t i = blockIdx.x * LBS + threadIdx.x;
  if (i >= n) return;
  const GlobalOps o{par};
  const uint32_t v = o.ld(i);
  if (v == kNone || v == i) return;
  uint32_t r = v, p = o.ld(r);
  while (p != r) {
    r = p;
    p = o.ld(r);
  }
  if (r != v) o.st(i, r);
}

// ── anchored mask ──────────────────────────────────────────────────────
// Anchor voxels mark their root's parent word; solid voxels read the mark.
__global__ void __launch_bounds__(LBS)
    k_cc_mark_anchored(uint32_t n, const uint8_t* __restrict__ type, uint32_t* par) {
  const uint32_t i = blockIdx.x * LBS + threadIdx.x;
  if (i >= n || type[i] != 2) return;
  const GlobalOps o{par};
  const uint32_t r = o.ld(i) & kIdx;
  if (!(o.ld(r) & kFlag)) __hip_atomic_fetch_or(par + r, kFlag, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(LBS)
    k_cc_reach(uint32_t n, const uint8_t* __restrict__ type, const uint32_t* __restrict__ par,
               uint8_t* __restrict__ reach) {
  const uint32_t i = blockIdx.x * LBS + threadIdx.x;
  if (i >= n) return;
  const uint8_t t = type[i];
  uint8_t v = 0;
  if (t != 0 && t != 2) v = (par[par[i] & kIdx] & kFlag) ? 1 : 0;
  reach[i] = v;
}

// ── component records ──────────────────────────────────────────────────
__device__ __forceinline__ int seg_excl_scan(int v, int* sh, int* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < LBS; o <<= 1) {
    const int a = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  const int incl = sh[t];
  *total = sh[LBS - 1];
  __syncthreads();
  return incl - v;
}

// Roots per SEG-cell segment (4 consecutive cells per thread).
__global__ void __launch_bounds__(LBS)
    k_lab_count(uint32_t n, const uint32_t* __restrict__ par, int* __restrict__ counts) {
  __shared__ int sh[LBS];
  const uint32_t base = blockIdx.x * SEG + threadIdx.x * 4;
  int c = 0;
  for (int e = 0; e < 4; e++) c += (base + e < n && par[base + e] == base + e) ? 1 : 0;
  int tot;
  seg_excl_scan(c, sh, &tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// Exclusive scan of the segment counts in place; hdr = {allocated slots,
// overflow, total roots}.
__global__ void __launch_bounds__(LBS)
    k_lab_scan(int nseg, int* __restrict__ counts, int maxc, uint32_t* __restrict__ hdr) {
  __shared__ int sh[LBS];
  const int per = (nseg + LBS - 1) / LBS;
  const int s0 = threadIdx.x * per;
  int s = 0;
  for (int e = 0; e < per; e++)
    if (s0 + e < nseg) s += counts[s0 + e];
  int tot;
  int ex = seg_excl_scan(s, sh, &tot);
  for (int e = 0; e < per; e++)
    if (s0 + e < nseg) {
      const int v = counts[s0 + e];
      counts[s0 + e] = ex;
      ex += v;
    }
  if (threadIdx.x == 0) {
    hdr[0] = (uint32_t)(tot < maxc ? tot : maxc);
    hdr[1] = tot > maxc ? 1u : 0u;
    hdr[2] = (uint32_t)tot;
  }
}

// Slots for roots in ascending flat order; records of the first maxc roots
// start as {label, 0, 0, 0, ~0, ~0, ~0, 0, 0, 0, 0, 0}.
__global__ void __launch_bounds__(LBS)
    k_lab_alloc(uint32_t n, const uint32_t* __restrict__ par, const int* __restrict__ offs,
                int maxc, uint32_t* __restrict__ slot, uint32_t* __restrict__ rec) {
  __shared__ int sh[LBS];
  const uint32_t base = blockIdx.x * SEG + threadIdx.x * 4;
  int c = 0;
  for (int e = 0; e < 4; e++) c += (base + e < n && par[base + e] == base + e) ? 1 : 0;
  int tot;
  int rank = offs[blockIdx.x] + seg_excl_scan(c, sh, &tot);
  for (int e = 0; e < 4; e++) {
    const uint32_t i = base + e;
    if (i >= n || par[i] != i) continue;
    if (rank < maxc) {
      slot[i] = (uint32_t)rank;
      uint32_t* r = rec + 12 * (size_t)rank;
      r[0] = i + 1;
      r[4] = r[5] = r[6] = kNone;
    } else {
      slot[i] = kNone;
    }
    rank++;
  }
}

__global__ void __launch_bounds__(LBS) k_lab_zero(uint32_t n, uint32_t* __restrict__ a) {
  const uint32_t i = blockIdx.x * LBS + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Per-record count / anchored / bounding box, and parent -> label in place.
// A work-group folds AGG_ROUNDS*SEG cells into an LDS table of maxc records
// (each thread first merges runs of its own cells with the same slot), then
// merges the touched records into global memory with atomics.
struct AggAcc {
  uint32_t s, cnt, anc, mn[3], mx[3];
};

__device__ __forceinline__ void agg_flush(const AggAcc& a, uint32_t* tab, int maxc) {
  if (a.s == kNone || a.cnt == 0) return;
  atomicAdd(tab + a.s, a.cnt);
  if (a.anc) atomicOr(tab + maxc + a.s, 1u);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    atomicMin(tab + (2 + c) * maxc + a.s, a.mn[c]);
    atomicMax(tab + (5 + c) * maxc + a.s, a.mx[c]);
  }
}

__global__ void __launch_bounds__(LBS)
    k_lab_aggregate(pfsf_dims d, const uint8_t* __restrict__ type, uint32_t* __restrict__ par,
                    const uint32_t* __restrict__ slot, int maxc, uint32_t* __restrict__ rec) {
  extern __shared__ uint32_t tab[];  // 8 x maxc: count, anchored, min xyz, max xyz
  for (int s = threadIdx.x; s < maxc; s += LBS) {
    tab[s] = 0;
    tab[maxc + s] = 0;
    for (int c = 0; c < 3; c++) {
      tab[(2 + c) * maxc + s] = kNone;
      tab[(5 + c) * maxc + s] = 0;
    }
  }
  __syncthreads();
  const uint32_t n = (uint32_t)d.lx * (uint32_t)d.ly * (uint32_t)d.lz;
  AggAcc a{kNone, 0, 0, {kNone, kNone, kNone}, {0, 0, 0}};
  for (int rd = 0; rd < AGG_ROUNDS; rd++) {
    const uint32_t base = (blockIdx.x * AGG_ROUNDS + rd) * SEG + threadIdx.x * 4;
    for (int e = 0; e < 4; e++) {
      const uint32_t i = base + e;
      if (i >= n) break;
      const uint32_t r = par[i];
      if (r == kNone) continue;
      par[i] = r + 1;  // label = root + 1
      const uint32_t s = slot[r];
      if (s == kNone) continue;
      if (s != a.s) {
        agg_flush(a, tab, maxc);
        a = AggAcc{s, 0, 0, {kNone, kNone, kNone}, {0, 0, 0}};
      }
      const uint32_t x = i % (uint32_t)d.lx, q = i / (uint32_t)d.lx;
      const uint32_t y = q % (uint32_t)d.ly, z = q / (uint32_t)d.ly;
      const uint32_t xyz[3] = {x, y, z};
      a.cnt++;
      a.anc |= type[i] == 2 ? 1u : 0u;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        a.mn[c] = min(a.mn[c], xyz[c]);
        a.mx[c] = max(a.mx[c], xyz[c]);
      }
    }
  }
  agg_flush(a, tab, maxc);
  __syncthreads();
  for (int s = threadIdx.x; s < maxc; s += LBS) {
    if (tab[s] == 0) continue;
    uint32_t* r = rec + 12 * (size_t)s;
    atomicAdd(r + 1, tab[s]);
    if (tab[maxc + s]) atomicOr(r + 2, 1u);
    for (int c = 0; c < 3; c++) {
      atomicMin(r + 4 + c, tab[(2 + c) * maxc + s]);
      atomicMax(r + 8 + c, tab[(5 + c) * maxc + s]);
    }
  }
}

hipError_t union_find(pfsf_dims d, const uint8_t* type, int rule, int conn, uint32_t* par,
                      int* err, hipStream_t s) {
  const int ntx = (d.lx + LTX - 1) / LTX, nty = (d.ly + LTY - 1) / LTY,
            ntz = (d.lz + LTZ - 1) / LTZ;
  const unsigned nt = (unsigned)ntx * nty * ntz;
  const uint32_t n = (uint32_t)nvox(d);
  k_cc_tile<<<nt, LBS, 0, s>>>(d, type, rule, conn, ntx, nty, par, err);
  k_cc_cross<<<nt, LBS, 0, s>>>(d, type, rule, conn, ntx, nty, par, err);
  k_cc_flatten<<<blocks_for(n, LBS), LBS, 0, s>>>(n, par);
  return hipGetLastError();
}

}  // namespace

int64_t label_scratch_words(pfsf_dims d) {
  const int64_t n = nvox(d);
  return n + (n + SEG - 1) / SEG + 8;
}

hipError_t label_components(pfsf_dims d, const uint8_t* type, int conn, uint32_t* id,
                            uint32_t* rec, int maxc, uint32_t* scratch, uint32_t* hdr_host,
                            hipStream_t s) {
  const int64_t n64 = nvox(d);
  if (n64 <= 0 || n64 > kMaxLabelCells || maxc < 1 || maxc > kMaxLabelComponents)
    return hipErrorInvalidValue;
  const uint32_t n = (uint32_t)n64;
  const int nseg = (int)((n64 + SEG - 1) / SEG);
  uint32_t* slot = scratch;
  int* counts = (int*)(scratch + n);
  uint32_t* hdr = scratch + n + nseg;  // {slots, overflow, roots, error}
  const unsigned nrec = 12u * (unsigned)maxc;
  k_lab_zero<<<blocks_for(nrec, LBS), LBS, 0, s>>>(nrec, rec);
  k_lab_zero<<<1, LBS, 0, s>>>(4, hdr);
  hipError_t e = union_find(d, type, 0, conn, id, (int*)(hdr + 3), s);
  if (e != hipSuccess) return e;
  k_lab_count<<<nseg, LBS, 0, s>>>(n, id, counts);
  k_lab_scan<<<1, LBS, 0, s>>>(nseg, counts, maxc, hdr);
  k_lab_alloc<<<nseg, LBS, 0, s>>>(n, id, counts, maxc, slot, rec);
  const unsigned nagg = (unsigned)((n64 + (int64_t)SEG * AGG_ROUNDS - 1) / ((int64_t)SEG * AGG_ROUNDS));
  k_lab_aggregate<<<nagg, LBS, 8 * sizeof(uint32_t) * (size_t)maxc, s>>>(d, type, id, slot, maxc,
                                                                        rec);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(hdr_host, hdr, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  return hdr_host[3] ? hipErrorUnknown : hipSuccess;
}

hipError_t anchored_mask(pfsf_dims d, const uint8_t* type, uint8_t* reach, uint32_t* scratch,
                         int* dflag, hipStream_t s) {
  const int64_t n64 = nvox(d);
  if (n64 <= 0 || n64 > kMaxLabelCells) return hipErrorInvalidValue;
  const uint32_t n = (uint32_t)n64;
  k_lab_zero<<<1, LBS, 0, s>>>(1, (uint32_t*)dflag);
  hipError_t e = union_find(d, type, 1, 26, scratch, dflag, s);
  if (e != hipSuccess) return e;
  k_cc_mark_anchored<<<blocks_for(n, LBS), LBS, 0, s>>>(n, type, scratch);
  k_cc_reach<<<blocks_for(n, LBS), LBS, 0, s>>>(n, type, scratch, reach);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  int hflag = 0;
  e = hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  return hflag ? hipErrorUnknown : hipSuccess;
}

}  // namespace pfsf
