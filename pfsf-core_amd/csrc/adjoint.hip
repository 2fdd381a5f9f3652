// Load-path loss and its adjoint gradient dL/dkappa.
//
// A load path applies b_t = (t / T) b, t = 1..T, to one island; each state
// solves A(kappa) phi_t = b_t with the multigrid PCG solver (warm-started
// from phi_{t-1} * t / (t-1), exact for a linear path).  The loss is
//   L = sum_t sum_i softplus(f_i(phi_t) - rcomp_i),
//   f_i = sum_j a_ij (phi_j - phi_i)^+        (the crushing flux of the
// failure scan, SH/failure_scan.comp.glsl:116-158, over non-air neighbours
// of a free voxel i; phi = 0 off the free set).
//
// Gradient: dL/dkappa_m = sum_t [ dL_t/dkappa_m |phi - lambda_t^T dA/dkappa_m phi_t ]
// with A lambda_t = dL_t/dphi_t.  A is symmetric, so the adjoint solve reuses
// the same solver object (hierarchy, kernels, tolerance).  The forward pass
// keeps phi every `checkpoint_every` steps; the backward pass walks the
// segments in reverse, re-solving each segment's states from its checkpoint
// and then running that segment's adjoint solves newest first.
//
// The forward loss is FP32 (couplings, flux, softplus) with FP64 sums, as is
// the sigmoid weight the backward pass uses; the FP64 variant of the loss
// exists for finite-difference checks.  Gradient sums are FP64.
#include <math.h>

#include <algorithm>
#include <vector>

#include "mg_kernels.h"
#include "mg_solver.h"

namespace pfsf {
namespace {

template <typename T>
__device__ __forceinline__ T couple_t(int cls, float ki, float kj);
template <>
__device__ __forceinline__ float couple_t<float>(int cls, float ki, float kj) {
  if (cls == 1) return 0.5f * (ki + kj);
  return (cls == 2 ? kEdgeP : kCornerP) * sqrtf(ki * kj);
}
template <>
__device__ __forceinline__ double couple_t<double>(int cls, float ki, float kj) {
  return couple_f64(cls, ki, kj);
}

// d a_ij / d kappa_i of the real-valued coupling formulas.
__device__ __forceinline__ double dcouple(int cls, double ki, double kj) {
  if (cls == 1) return 0.5;
  const double c = cls == 2 ? (double)kEdgeP : (double)kCornerP;
  return ki > 0.0 ? c * 0.5 * sqrt(kj / ki) : 0.0;
}

// softplus(u) = log(1 + e^u) and its derivative, the logistic sigmoid.
__device__ __forceinline__ float softplus_t(float u) {
  return (u > 0.f ? u : 0.f) + log1pf(expf(-fabsf(u)));
}
__device__ __forceinline__ double softplus_t(double u) {
  return (u > 0.0 ? u : 0.0) + log1p(exp(-fabs(u)));
}
__device__ __forceinline__ float sigmoid_t(float u) { return 1.f / (1.f + expf(-u)); }
__device__ __forceinline__ double sigmoid_t(double u) { return 1.0 / (1.0 + exp(-u)); }

// Visits the in-bounds non-air neighbours of cell i: f(cls, j).
template <typename F>
__device__ __forceinline__ void for_neighbours(pfsf_dims d, const uint8_t* __restrict__ type,
                                               int i, F f) {
  const int px = i % d.lx, r = i / d.lx;
  const int py = r % d.ly, pz = r / d.ly;
  const int sz = d.lx * d.ly;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int cls = (dx != 0) + (dy != 0) + (dz != 0);
        if (cls == 0) continue;
        const int nx = px + dx, ny = py + dy, nz = pz + dz;
        if (nx < 0 || nx >= d.lx || ny < 0 || ny >= d.ly || nz < 0 || nz >= d.lz) continue;
        const int j = i + dx + dy * d.lx + dz * sz;
        if (type[j] == 0) continue;
        f(cls, j);
      }
}

// Loss terms of one state (partials of sum softplus) and, optionally, the
// sigmoid weights w_i = softplus'(f_i - rcomp_i).
template <typename T>
__global__ void __launch_bounds__(BS)
    k_flux_loss(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                const uint8_t* __restrict__ fr, const double* __restrict__ x,
                const float* __restrict__ rcomp, float* __restrict__ w, double* partials) {
  __shared__ double sh[1][BS / 64];
  const int64_t n = nvox(d);
  const int i = (int)(blockIdx.x * BS + threadIdx.x);
  double acc[1] = {0.0};
  if (i < n) {
    float wi = 0.f;
    if (fr[i]) {
      const float ki = kapf(kappa[i]);
      const T pi = (T)x[i];
      T f = T(0);
      for_neighbours(d, type, i, [&](int cls, int j) {
        const T pj = fr[j] ? (T)x[j] : T(0);
        if (pj > pi) f += couple_t<T>(cls, ki, kapf(kappa[j])) * (pj - pi);
      });
      const T u = f - (T)rcomp[i];
      acc[0] = (double)softplus_t(u);
      wi = (float)sigmoid_t(u);
    }
    if (w) w[i] = wi;
  }
  write_partials<1>(acc, partials, sh);
}

// Adjoint right-hand side g = dL_t/dphi (free voxels; FP32, as the forward):
//   g_k = -w_k sum_j a_kj [phi_j > phi_k] + sum_j w_j a_kj [phi_k > phi_j].
__global__ void __launch_bounds__(BS)
    k_flux_grad_phi(pfsf_dims d, const uint8_t* __restrict__ type,
                    const float* __restrict__ kappa, const uint8_t* __restrict__ fr,
                    const double* __restrict__ x, const float* __restrict__ w,
                    float* __restrict__ g) {
  const int64_t n = nvox(d);
  const int i = (int)(blockIdx.x * BS + threadIdx.x);
  if (i >= n) return;
  float acc = 0.f;
  if (fr[i]) {
    const float ki = kapf(kappa[i]);
    const float pi = (float)x[i], wi = w[i];
    for_neighbours(d, type, i, [&](int cls, int j) {
      const float a = couple_t<float>(cls, ki, kapf(kappa[j]));
      const float pj = fr[j] ? (float)x[j] : 0.f;
      if (pj > pi) acc -= wi * a;
      if (pi > pj && fr[j]) acc += w[j] * a;
    });
  }
  g[i] = acc;
}

// grad_m += sum_j da_mj/dkappa_m [ w_m (phi_j - phi_m)^+ + w_j (phi_m - phi_j)^+
//                                 - (lam_m - lam_j)(phi_m - phi_j) ]
// over non-air neighbours; phi, lam, w are 0 off the free set.
__global__ void __launch_bounds__(BS)
    k_grad_kappa(pfsf_dims d, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                 const uint8_t* __restrict__ fr, const double* __restrict__ x,
                 const double* __restrict__ lam, const float* __restrict__ w,
                 double* __restrict__ grad) {
  const int64_t n = nvox(d);
  const int i = (int)(blockIdx.x * BS + threadIdx.x);
  if (i >= n || type[i] == 0) return;
  const double ki = (double)kapf(kappa[i]);
  const bool fi = fr[i] != 0;
  const double pi = fi ? x[i] : 0.0, li = fi ? lam[i] : 0.0, wi = fi ? (double)w[i] : 0.0;
  double acc = 0.0;
  for_neighbours(d, type, i, [&](int cls, int j) {
    const bool fj = fr[j] != 0;
    const double pj = fj ? x[j] : 0.0, lj = fj ? lam[j] : 0.0, wj = fj ? (double)w[j] : 0.0;
    const double up = pj > pi ? pj - pi : 0.0, dn = pi > pj ? pi - pj : 0.0;
    acc += dcouple(cls, ki, (double)kapf(kappa[j])) * (wi * up + wj * dn - (li - lj) * (pi - pj));
  });
  grad[i] += acc;
}

__global__ void __launch_bounds__(BS)
    k_scaled(int64_t n, const float* __restrict__ a, float s, float* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = s * a[i];
}

__global__ void __launch_bounds__(BS) k_scale64(int64_t n, double s, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] *= s;
}

__global__ void __launch_bounds__(BS)
    k_copy64a(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void __launch_bounds__(BS) k_zero64a(int64_t n, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = 0.0;
}

// acc[slot] += sum of n partials (one work-group).
__global__ void __launch_bounds__(BS)
    k_sum_into(const double* __restrict__ part, unsigned n, double* acc, int slot) {
  const double t = block_total(part, n);
  if (threadIdx.x == 0) acc[slot] += t;
}

// ── damage-coupled load path (desc.damage) ──────────────────────────────
// State per voxel in FP64: damage d and history H.  Between states:
//   kappa_t = kappa * g(d),  g(d) = (1 - kDmgRes) (1 - d)^2 + kDmgRes
//   psi = 0.5 kappa_t (gx^2 + gy^2 + gz^2)  (per-axis differences over
//         non-air face neighbours, one-sided at a missing one, as the history
//         update of a tick), H = max(H, psi) on free voxels
//   d' = one hybrid AT2 damage step of the tick's phase field (split on,
//        hydration 1, conductivity kappa_t on every face) on free voxels with
//        H > 0.
// The reverse passes below are the exact derivatives of these expressions
// on the branches the forward state takes.
constexpr double kDmgRes = 1e-3;

__device__ __forceinline__ double dmg_g(double d) {
  return (1.0 - kDmgRes) * ((1.0 - d) * (1.0 - d)) + kDmgRes;
}
__device__ __forceinline__ double dmg_dg(double d) { return -2.0 * (1.0 - kDmgRes) * (1.0 - d); }

// kappa_t of every non-air voxel (FP32, the solver's operator).
__global__ void __launch_bounds__(BS)
    k_kdeg(int64_t n, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
           const double* __restrict__ d, float* __restrict__ kt) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) kt[i] = type[i] != 0 ? (float)((double)kapf(kappa[i]) * dmg_g(d[i])) : 0.f;
}

// Face neighbour k of cell i in order -x, +x, -y, +y, -z, +z: index, or -1
// outside the grid.
__device__ __forceinline__ int face_nbr(pfsf_dims d, int i, int k) {
  const int px = i % d.lx, r = i / d.lx;
  const int py = r % d.ly, pz = r / d.ly;
  const int s = (k & 1) ? 1 : -1;
  const int ax = k >> 1;
  const int c = ax == 0 ? px : ax == 1 ? py : pz;
  const int lim = ax == 0 ? d.lx : ax == 1 ? d.ly : d.lz;
  if (c + s < 0 || c + s >= lim) return -1;
  const int step = ax == 0 ? 1 : ax == 1 ? d.lx : d.lx * d.ly;
  return i + s * step;
}

// The state value of cell j (phi or d): 0 off the free set.
__device__ __forceinline__ double fval(const uint8_t* fr, const double* v, int j) {
  return fr[j] ? v[j] : 0.0;
}

// Per-axis difference of the history update at free cell i: g, and whether
// the - / + neighbours are non-air.
__device__ __forceinline__ double axis_diff(pfsf_dims d, const uint8_t* type, const uint8_t* fr,
                                            const double* x, int i, int ax, bool& om, bool& op) {
  const int jm = face_nbr(d, i, 2 * ax), jp = face_nbr(d, i, 2 * ax + 1);
  om = jm >= 0 && type[jm] != 0;
  op = jp >= 0 && type[jp] != 0;
  const double pi = fval(fr, x, i);
  const double pm = om ? fval(fr, x, jm) : 0.0, pp = op ? fval(fr, x, jp) : 0.0;
  if (om && op) return (pp - pm) * 0.5;
  if (op) return pp - pi;
  if (om) return pi - pm;
  return 0.0;
}

// psi of the state x and H = max(H, psi) on free cells (psi kept).
__global__ void __launch_bounds__(BS)
    k_psi_hist(pfsf_dims d, const uint8_t* __restrict__ type, const uint8_t* __restrict__ fr,
               const double* __restrict__ x, const float* __restrict__ kt,
               double* __restrict__ psi, double* __restrict__ H) {
  const int i = (int)(blockIdx.x * BS + threadIdx.x);
  if (i >= nvox(d)) return;
  double p = 0.0;
  if (fr[i]) {
    double g2 = 0.0;
    for (int ax = 0; ax < 3; ax++) {
      bool om, op;
      const double g = axis_diff(d, type, fr, x, i, ax, om, op);
      g2 += g * g;
    }
    p = 0.5 * (double)kt[i] * g2;
    H[i] = fmax(H[i], p);
  }
  psi[i] = p;
}

// The damage step's intermediates at free cell i (dm: damage before, Hn:
// history after this state's update, x: the state, kt: conductivity).
struct PfState {
  double nf, lp, ld, ssum, ss, pp, he, num, den0, den, q0, q1, v0, v1, v2, d;
};

__device__ __forceinline__ PfState pf_state(pfsf_dims d, const uint8_t* type, const uint8_t* fr,
                                            const double* x, const double* dm, double Hn,
                                            double kti, int i) {
  PfState s;
  s.nf = s.lp = s.ld = 0.0;
  const double pi = fval(fr, x, i), di = fval(fr, dm, i);
  for (int k = 0; k < 6; k++) {
    const int j = face_nbr(d, i, k);
    if (j < 0 || type[j] == 0) continue;
    s.nf += 1.0;
    s.lp += fval(fr, x, j) - pi;
    s.ld += fval(fr, dm, j) - di;
  }
  const double c1 = (double)kPfGc * (double)kPfL0, c2 = (double)kPfGc / (double)kPfL0;
  const double r = (double)kPfRelax;
  s.ssum = s.nf * kti;
  s.ss = fmax(s.ssum, 1e-12);
  s.pp = fmax(s.lp, 0.0) / s.ss;
  s.he = fmax(s.pp, Hn);
  s.num = 2.0 * s.he + c1 * s.ld;
  s.den0 = 2.0 * s.he + c2;
  s.den = fmax(s.den0, 1e-8);
  s.q0 = s.num / s.den;
  s.q1 = fmax(s.q0, di);
  const double q = fmin(s.q1, 1.0);
  s.v0 = di * (1.0 - r) + q * r;
  s.v1 = fmin(fmax(s.v0, 0.0), 1.0);
  s.v2 = fmax(s.v1, di);
  s.d = di;
  return s;
}

__global__ void __launch_bounds__(BS)
    k_pf64(pfsf_dims d, const uint8_t* __restrict__ type, const uint8_t* __restrict__ fr,
           const double* __restrict__ x, const double* __restrict__ dm,
           const double* __restrict__ H, const float* __restrict__ kt, double* __restrict__ dn) {
  const int i = (int)(blockIdx.x * BS + threadIdx.x);
  if (i >= nvox(d)) return;
  const bool active = fr[i] && H[i] > 0.0;
  dn[i] = active ? pf_state(d, type, fr, x, dm, H[i], (double)kt[i], i).v2 : dm[i];
}

// Reverse pass of the damage step, own-cell part: Dprev = the adjoint of the
// damage before (own terms), Hb += adjoint of H, gphi / kbar += own terms;
// dld, dlp: the adjoints of the cell's damage and phi Laplacians, which the
// gather adds to its face neighbours.
__global__ void __launch_bounds__(BS)
    k_pf_vjp_own(pfsf_dims d, const uint8_t* __restrict__ type, const uint8_t* __restrict__ fr,
                 const double* __restrict__ x, const double* __restrict__ dm,
                 const double* __restrict__ H, const float* __restrict__ kt,
                 const double* __restrict__ Db, double* __restrict__ Dprev,
                 double* __restrict__ Hb, double* __restrict__ gphi, double* __restrict__ kbar,
                 double* __restrict__ dld, double* __restrict__ dlp) {
  const int i = (int)(blockIdx.x * BS + threadIdx.x);
  if (i >= nvox(d)) return;
  if (!(fr[i] && H[i] > 0.0)) {  // inactive: the damage passes through
    Dprev[i] = Db[i];
    dld[i] = dlp[i] = 0.0;
    return;
  }
  const PfState s = pf_state(d, type, fr, x, dm, H[i], (double)kt[i], i);
  const double r = (double)kPfRelax, c1 = (double)kPfGc * (double)kPfL0;
  const double db = Db[i];
  const bool m2 = s.v1 > s.d;
  const double dv1 = m2 ? db : 0.0;
  double dd = m2 ? 0.0 : db;
  const double dv0 = (s.v0 > 0.0 && s.v0 < 1.0) ? dv1 : 0.0;
  dd += dv0 * (1.0 - r);
  const double dq = dv0 * r;
  const double dq1 = s.q1 < 1.0 ? dq : 0.0;
  const bool m0 = s.q0 > s.d;
  const double dq0 = m0 ? dq1 : 0.0;
  if (!m0) dd += dq1;
  const double dnum = dq0 / s.den;
  const double dden = s.den0 > 1e-8 ? -dq0 * s.q0 / s.den : 0.0;
  const double dhe = 2.0 * dnum + 2.0 * dden;
  const double dl = c1 * dnum;
  const bool mp = s.pp > H[i];
  const double dpp = mp ? dhe : 0.0;
  const double dh = mp ? 0.0 : dhe;
  const double dp = s.lp > 0.0 ? dpp / s.ss : 0.0;
  const double dss = (s.lp > 0.0 && s.ssum > 1e-12) ? -dpp * s.lp / (s.ss * s.ss) : 0.0;
  Dprev[i] = dd - s.nf * dl;
  Hb[i] += dh;
  gphi[i] += -s.nf * dp;
  kbar[i] += s.nf * dss;
  dld[i] = dl;
  dlp[i] = dp;
}

// Gather part: every non-air cell receives the Laplacian adjoints of its face
// neighbours (a cell's Laplacians hold +1 of each non-air face neighbour).
__global__ void __launch_bounds__(BS)
    k_pf_vjp_gather(pfsf_dims d, const uint8_t* __restrict__ type,
                    const double* __restrict__ dld, const double* __restrict__ dlp,
                    double* __restrict__ Dprev, double* __restrict__ gphi) {
  const int j = (int)(blockIdx.x * BS + threadIdx.x);
  if (j >= nvox(d) || type[j] == 0) return;
  double a = 0.0, b = 0.0;
  for (int k = 0; k < 6; k++) {
    const int i = face_nbr(d, j, k);
    if (i < 0) continue;
    a += dld[i];
    b += dlp[i];
  }
  Dprev[j] += a;
  gphi[j] += b;
}

// H_t = max(H_{t-1}, psi): the adjoint goes to psi where psi won, else to
// H_{t-1}.
__global__ void __launch_bounds__(BS)
    k_hist_vjp(int64_t n, const uint8_t* __restrict__ fr, const double* __restrict__ psi,
               const double* __restrict__ Hprev, double* __restrict__ Hb,
               double* __restrict__ psib) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const bool m = fr[i] && psi[i] > Hprev[i];
  psib[i] = m ? Hb[i] : 0.0;
  if (m) Hb[i] = 0.0;
}

// Reverse pass of psi, own-cell part: kbar += d psi / d kappa_t, gphi += the
// own terms of the one-sided differences; cp / cm (per axis): what the + and
// - neighbours receive (the gather adds them).
__global__ void __launch_bounds__(BS)
    k_psi_vjp_own(pfsf_dims d, const uint8_t* __restrict__ type, const uint8_t* __restrict__ fr,
                  const double* __restrict__ x, const float* __restrict__ kt,
                  const double* __restrict__ psib, double* __restrict__ gphi,
                  double* __restrict__ kbar, double* __restrict__ cp, double* __restrict__ cm) {
  const int i = (int)(blockIdx.x * BS + threadIdx.x);
  const int64_t n = nvox(d);
  if (i >= n) return;
  if (!fr[i] || psib[i] == 0.0) {
    for (int ax = 0; ax < 3; ax++) cp[ax * n + i] = cm[ax * n + i] = 0.0;
    return;
  }
  const double pb = psib[i], k = (double)kt[i];
  double dk = 0.0, own = 0.0;
  for (int ax = 0; ax < 3; ax++) {
    bool om, op;
    const double g = axis_diff(d, type, fr, x, i, ax, om, op);
    dk += 0.5 * g * g * pb;
    const double dg = k * g * pb;
    const bool both = om && op;
    cp[ax * n + i] = both ? 0.5 * dg : (op ? dg : 0.0);
    cm[ax * n + i] = both ? -0.5 * dg : (om ? -dg : 0.0);
    if (op && !om) own -= dg;
    if (om && !op) own += dg;
  }
  kbar[i] += dk;
  gphi[i] += own;
}

__global__ void __launch_bounds__(BS)
    k_psi_vjp_gather(pfsf_dims d, const double* __restrict__ cp, const double* __restrict__ cm,
                     double* __restrict__ gphi) {
  const int j = (int)(blockIdx.x * BS + threadIdx.x);
  const int64_t n = nvox(d);
  if (j >= n) return;
  double a = 0.0;
  for (int ax = 0; ax < 3; ax++) {
    const int im = face_nbr(d, j, 2 * ax), ip = face_nbr(d, j, 2 * ax + 1);
    if (im >= 0) a += cp[ax * n + im];  // j is the + neighbour of im
    if (ip >= 0) a += cm[ax * n + ip];  // j is the - neighbour of ip
  }
  gphi[j] += a;
}

// The adjoint right-hand side of a state: the flux term g plus the damage
// and history terms, on the free set (FP32, the solver's right-hand side).
__global__ void __launch_bounds__(BS)
    k_adj_rhs(int64_t n, const uint8_t* __restrict__ fr, const float* __restrict__ g,
              const double* __restrict__ gphi, float* __restrict__ rhs) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) rhs[i] = fr[i] ? (float)((double)g[i] + gphi[i]) : 0.f;
}

// kappa_t = kappa g(d_{t-1}): grad += kbar g(d), and the adjoint of the
// damage before this state, Db = Dprev + kbar kappa g'(d).
__global__ void __launch_bounds__(BS)
    k_dmg_chain(int64_t n, const uint8_t* __restrict__ type, const float* __restrict__ kappa,
                const double* __restrict__ dm, const double* __restrict__ kbar,
                const double* __restrict__ Dprev, double* __restrict__ grad,
                double* __restrict__ Db) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double kb = type[i] != 0 ? kbar[i] : 0.0;
  grad[i] += kb * dmg_g(dm[i]);
  Db[i] = Dprev[i] + kb * (double)kapf(kappa[i]) * dmg_dg(dm[i]);
}

// Voxels whose kappa does not enter the operator get 0.
__global__ void __launch_bounds__(BS)
    k_grad_mask(int64_t n, const float* __restrict__ kappa, double* __restrict__ grad) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n && !(kappa[i] > 0.f)) grad[i] = 0.0;
}

}  // namespace

static inline bool chk(hipError_t e, hipError_t& keep) {
  if (e == hipSuccess) return false;
  if (keep == hipSuccess) keep = e;
  return true;
}

// One load-path evaluation: loss (FP32 or FP64 flux) and, when grad != null,
// the adjoint gradient into grad (device, N doubles, overwritten).
hipError_t load_path_run(const pfsf_load_path_desc& desc, bool fp64, double* grad,
                         pfsf_load_path_info* info) {
  hipError_t err = hipSuccess;
  const pfsf_dims d = desc.dims;
  const int64_t n = nvox(d);
  const int T = desc.steps > 0 ? desc.steps : 50;
  const int K = desc.checkpoint_every > 0 ? std::min(desc.checkpoint_every, T) : 10;
  const int nseg = (T + K - 1) / K;
  const bool dmg = desc.damage != 0;
  pfsf_solver_desc sd{};
  sd.dims = d;
  sd.type = desc.type;
  sd.kappa = desc.kappa;
  sd.precision = PFSF_PREC_FP32;
  sd.tol = desc.tol > 0 ? desc.tol : 1e-6;
  sd.max_iters = desc.max_iters > 0 ? desc.max_iters : 2000;
  sd.stream = desc.stream;
  MgSolver S;
  if (chk(S.create(sd), err)) return err;
  hipStream_t s = (hipStream_t)desc.stream;
  const unsigned nb = blocks_for(n, BS);
  std::vector<void*> mem;
  auto alloc = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (chk(hipMalloc(&p, std::max<size_t>(bytes, 1)), err)) return nullptr;
    mem.push_back(p);
    return p;
  };
  auto d64 = [&](size_t count) { return (double*)alloc(count * sizeof(double)); };
  float* bt = (float*)alloc(n * sizeof(float));
  double* x = d64(n);
  double* ckpt = d64((size_t)nseg * n);
  double* seg = grad ? d64((size_t)K * n) : nullptr;
  float* w = grad ? (float*)alloc(n * sizeof(float)) : nullptr;
  float* g = grad ? (float*)alloc(n * sizeof(float)) : nullptr;
  double* lam = grad ? d64(n) : nullptr;
  double* part = d64(nb);
  double* acc = d64(2);
  // damage state: d and H (and their checkpoints and per-step copies), the
  // operator kappa_t, and the reverse pass's work arrays
  double *dm = nullptr, *dn = nullptr, *H = nullptr, *psi = nullptr;
  double *ck_d = nullptr, *ck_h = nullptr, *seg_d = nullptr, *seg_h = nullptr;
  double *Db = nullptr, *Hb = nullptr, *Dprev = nullptr, *gphi = nullptr, *kbar = nullptr;
  double *dld = nullptr, *dlp = nullptr, *psib = nullptr, *cpm = nullptr;
  float *kt = nullptr, *rhs = nullptr;
  if (dmg) {
    dm = d64(n);
    dn = d64(n);
    H = d64(n);
    psi = d64(n);
    ck_d = d64((size_t)nseg * n);
    ck_h = d64((size_t)nseg * n);
    kt = (float*)alloc(n * sizeof(float));
    if (grad) {
      seg_d = d64((size_t)K * n);
      seg_h = d64((size_t)K * n);
      Db = d64(n);
      Hb = d64(n);
      Dprev = d64(n);
      gphi = d64(n);
      kbar = d64(n);
      dld = d64(n);
      dlp = d64(n);
      psib = d64(n);
      cpm = d64(6 * (size_t)n);
      rhs = (float*)alloc(n * sizeof(float));
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  chk(hipEventCreate(&e0), err);
  chk(hipEventCreate(&e1), err);
  pfsf_load_path_info out{};
  if (err == hipSuccess) {
    chk(hipEventRecord(e0, s), err);
    chk(hipMemsetAsync(acc, 0, 2 * sizeof(double), s), err);
    k_zero64a<<<nb, BS, 0, s>>>(n, x);
    if (dmg) {
      k_zero64a<<<nb, BS, 0, s>>>(n, dm);
      k_zero64a<<<nb, BS, 0, s>>>(n, H);
    }
    const uint8_t* fr = S.free_mask();
    const float* kap = desc.kappa;  // the operator of the current state
    auto step = [&](int t) {  // x: state t-1 -> state t
      if (dmg) {
        // kappa_t from the damage before this state; the hierarchy follows
        k_kdeg<<<nb, BS, 0, s>>>(n, desc.type, desc.kappa, dm, kt);
        chk(S.refresh_kappa(kt), err);
        kap = kt;
      }
      if (t > 1) k_scale64<<<nb, BS, 0, s>>>(n, (double)t / (double)(t - 1), x);
      k_scaled<<<nb, BS, 0, s>>>(n, desc.source, (float)((double)t / T), bt);
      pfsf_solve_info si{};
      chk(S.solve(bt, x, nullptr, &si), err);
      out.pcg_iterations += si.iterations;
    };
    auto advance = [&]() {  // after state t: H and d of state t
      k_psi_hist<<<nb, BS, 0, s>>>(d, desc.type, fr, x, kt, psi, H);
      k_pf64<<<nb, BS, 0, s>>>(d, desc.type, fr, x, dm, H, kt, dn);
      std::swap(dm, dn);
    };
    auto save = [&](double* dst_x, double* dst_d, double* dst_h) {
      k_copy64a<<<nb, BS, 0, s>>>(n, x, dst_x);
      if (dmg) {
        k_copy64a<<<nb, BS, 0, s>>>(n, dm, dst_d);
        k_copy64a<<<nb, BS, 0, s>>>(n, H, dst_h);
      }
    };
    // forward: loss and checkpoints (state at t = 0, K, 2K, ...)
    for (int t = 1; t <= T && err == hipSuccess; t++) {
      if ((t - 1) % K == 0) {
        const size_t c = (size_t)((t - 1) / K) * n;
        save(ckpt + c, dmg ? ck_d + c : nullptr, dmg ? ck_h + c : nullptr);
      }
      step(t);
      out.forward_solves++;
      if (fp64)
        k_flux_loss<double><<<nb, BS, 0, s>>>(d, desc.type, kap, fr, x, desc.rcomp, nullptr,
                                              part);
      else
        k_flux_loss<float><<<nb, BS, 0, s>>>(d, desc.type, kap, fr, x, desc.rcomp, nullptr,
                                             part);
      k_sum_into<<<1, BS, 0, s>>>(part, nb, acc, 0);
      if (dmg) advance();
    }
    out.checkpoints = nseg;
    // backward: segments newest first
    if (grad && err == hipSuccess) {
      k_zero64a<<<nb, BS, 0, s>>>(n, grad);
      k_zero64a<<<nb, BS, 0, s>>>(n, lam);
      if (dmg) {
        k_zero64a<<<nb, BS, 0, s>>>(n, Db);
        k_zero64a<<<nb, BS, 0, s>>>(n, Hb);
      }
      for (int sg = nseg - 1; sg >= 0 && err == hipSuccess; sg--) {
        const int t0 = sg * K, t1 = std::min(T, t0 + K);
        k_copy64a<<<nb, BS, 0, s>>>(n, ckpt + (size_t)sg * n, x);
        if (dmg) {
          k_copy64a<<<nb, BS, 0, s>>>(n, ck_d + (size_t)sg * n, dm);
          k_copy64a<<<nb, BS, 0, s>>>(n, ck_h + (size_t)sg * n, H);
        }
        for (int t = t0 + 1; t <= t1 && err == hipSuccess; t++) {
          const size_t o = (size_t)(t - t0 - 1) * n;
          if (dmg) {  // the state before step t
            k_copy64a<<<nb, BS, 0, s>>>(n, dm, seg_d + o);
            k_copy64a<<<nb, BS, 0, s>>>(n, H, seg_h + o);
          }
          step(t);
          out.recompute_solves++;
          k_copy64a<<<nb, BS, 0, s>>>(n, x, seg + o);
          if (dmg && t < t1) advance();
        }
        for (int t = t1; t > t0 && err == hipSuccess; t--) {
          const size_t o = (size_t)(t - t0 - 1) * n;
          const double* ph = seg + o;
          if (dmg) {
            const double* dprev = seg_d + o;
            const double* hprev = seg_h + o;
            // the state of step t again: kappa_t, psi_t, H_t (into H)
            k_kdeg<<<nb, BS, 0, s>>>(n, desc.type, desc.kappa, dprev, kt);
            k_copy64a<<<nb, BS, 0, s>>>(n, hprev, H);
            k_psi_hist<<<nb, BS, 0, s>>>(d, desc.type, fr, ph, kt, psi, H);
            k_zero64a<<<nb, BS, 0, s>>>(n, gphi);
            k_zero64a<<<nb, BS, 0, s>>>(n, kbar);
            // damage step, history, psi (reverse order of the forward)
            k_pf_vjp_own<<<nb, BS, 0, s>>>(d, desc.type, fr, ph, dprev, H, kt, Db, Dprev, Hb,
                                            gphi, kbar, dld, dlp);
            k_pf_vjp_gather<<<nb, BS, 0, s>>>(d, desc.type, dld, dlp, Dprev, gphi);
            k_hist_vjp<<<nb, BS, 0, s>>>(n, fr, psi, hprev, Hb, psib);
            k_psi_vjp_own<<<nb, BS, 0, s>>>(d, desc.type, fr, ph, kt, psib, gphi, kbar, cpm,
                                             cpm + 3 * n);
            k_psi_vjp_gather<<<nb, BS, 0, s>>>(d, cpm, cpm + 3 * n, gphi);
            chk(S.refresh_kappa(kt), err);
            kap = kt;
          }
          k_flux_loss<float><<<nb, BS, 0, s>>>(d, desc.type, kap, fr, ph, desc.rcomp, w, part);
          k_flux_grad_phi<<<nb, BS, 0, s>>>(d, desc.type, kap, fr, ph, w, g);
          if (dmg) k_adj_rhs<<<nb, BS, 0, s>>>(n, fr, g, gphi, rhs);
          pfsf_solve_info si{};
          chk(S.solve(dmg ? rhs : g, lam, nullptr, &si), err);  // warm start: lambda of t+1
          out.pcg_iterations += si.iterations;
          out.adjoint_solves++;
          if (dmg) {
            k_grad_kappa<<<nb, BS, 0, s>>>(d, desc.type, kap, fr, ph, lam, w, kbar);
            k_dmg_chain<<<nb, BS, 0, s>>>(n, desc.type, desc.kappa, seg_d + o, kbar, Dprev,
                                          grad, Db);
          } else {
            k_grad_kappa<<<nb, BS, 0, s>>>(d, desc.type, kap, fr, ph, lam, w, grad);
          }
        }
      }
      if (dmg) k_grad_mask<<<nb, BS, 0, s>>>(n, desc.kappa, grad);
    }
    chk(hipGetLastError(), err);
    chk(hipEventRecord(e1, s), err);
    chk(hipEventSynchronize(e1), err);
    double h_acc[2] = {0.0, 0.0};
    chk(hipMemcpy(h_acc, acc, sizeof(h_acc), hipMemcpyDeviceToHost), err);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    out.loss = h_acc[0];
    out.ms = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* p : mem) (void)hipFree(p);
  if (info) *info = out;
  return err;
}

}  // namespace pfsf
