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
  float* bt = (float*)alloc(n * sizeof(float));
  double* x = (double*)alloc(n * sizeof(double));
  double* ckpt = (double*)alloc((size_t)nseg * n * sizeof(double));
  double* seg = grad ? (double*)alloc((size_t)K * n * sizeof(double)) : nullptr;
  float* w = grad ? (float*)alloc(n * sizeof(float)) : nullptr;
  float* g = grad ? (float*)alloc(n * sizeof(float)) : nullptr;
  double* lam = grad ? (double*)alloc(n * sizeof(double)) : nullptr;
  double* part = (double*)alloc(nb * sizeof(double));
  double* acc = (double*)alloc(2 * sizeof(double));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  chk(hipEventCreate(&e0), err);
  chk(hipEventCreate(&e1), err);
  pfsf_load_path_info out{};
  if (err == hipSuccess) {
    chk(hipEventRecord(e0, s), err);
    chk(hipMemsetAsync(acc, 0, 2 * sizeof(double), s), err);
    k_zero64a<<<nb, BS, 0, s>>>(n, x);
    const uint8_t* fr = S.free_mask();
    auto step = [&](int t) {  // x: state t-1 -> state t
      if (t > 1) k_scale64<<<nb, BS, 0, s>>>(n, (double)t / (double)(t - 1), x);
      k_scaled<<<nb, BS, 0, s>>>(n, desc.source, (float)((double)t / T), bt);
      pfsf_solve_info si{};
      chk(S.solve(bt, x, nullptr, &si), err);
      out.pcg_iterations += si.iterations;
    };
    // forward: loss and checkpoints (state at t = 0, K, 2K, ...)
    for (int t = 1; t <= T && err == hipSuccess; t++) {
      if ((t - 1) % K == 0)
        k_copy64a<<<nb, BS, 0, s>>>(n, x, ckpt + (size_t)((t - 1) / K) * n);
      step(t);
      out.forward_solves++;
      if (fp64)
        k_flux_loss<double><<<nb, BS, 0, s>>>(d, desc.type, desc.kappa, fr, x, desc.rcomp,
                                              nullptr, part);
      else
        k_flux_loss<float><<<nb, BS, 0, s>>>(d, desc.type, desc.kappa, fr, x, desc.rcomp,
                                             nullptr, part);
      k_sum_into<<<1, BS, 0, s>>>(part, nb, acc, 0);
    }
    out.checkpoints = nseg;
    // backward: segments newest first
    if (grad && err == hipSuccess) {
      k_zero64a<<<nb, BS, 0, s>>>(n, grad);
      k_zero64a<<<nb, BS, 0, s>>>(n, lam);
      for (int sg = nseg - 1; sg >= 0 && err == hipSuccess; sg--) {
        const int t0 = sg * K, t1 = std::min(T, t0 + K);
        k_copy64a<<<nb, BS, 0, s>>>(n, ckpt + (size_t)sg * n, x);
        for (int t = t0 + 1; t <= t1 && err == hipSuccess; t++) {
          step(t);
          out.recompute_solves++;
          k_copy64a<<<nb, BS, 0, s>>>(n, x, seg + (size_t)(t - t0 - 1) * n);
        }
        for (int t = t1; t > t0 && err == hipSuccess; t--) {
          const double* ph = seg + (size_t)(t - t0 - 1) * n;
          k_flux_loss<float><<<nb, BS, 0, s>>>(d, desc.type, desc.kappa, fr, ph, desc.rcomp, w,
                                               part);
          k_flux_grad_phi<<<nb, BS, 0, s>>>(d, desc.type, desc.kappa, fr, ph, w, g);
          pfsf_solve_info si{};
          chk(S.solve(g, lam, nullptr, &si), err);  // warm start: lambda of step t+1
          out.pcg_iterations += si.iterations;
          out.adjoint_solves++;
          k_grad_kappa<<<nb, BS, 0, s>>>(d, desc.type, desc.kappa, fr, ph, lam, w, grad);
        }
      }
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
