// Host primitives of the reference's pfsf_compute.h / pfsf_diagnostics.h
// surface (stateless, host buffers).  The arithmetic follows the reference's
// definitions step for step, so results match its own build bit for bit:
// normalisation NP/src/compute/normalize.cpp:29-63 (PFSFDataBuilder.java:
// 222-278), downsample NP/src/compute/downsample.cpp:39-98
// (PFSFDataBuilder.java:430-471), Chebyshev schedule
// NP/src/compute/chebyshev.cpp:30-101 (PFSFScheduler.java:45-123).
#include <math.h>
#include <stdint.h>

#include "pfsf/pfsf_compute.h"
#include "pfsf/pfsf_diagnostics.h"

namespace {

uint8_t vote(int anchors, int solids, int airs) {
  if (anchors >= solids && anchors >= airs) return PFSF_VOXEL_ANCHOR;
  return solids >= airs ? PFSF_VOXEL_SOLID : PFSF_VOXEL_AIR;
}

// One step of the schedule's recurrence from w (FP32).
float omega_next(float rho2, float w) { return 4.0f / (4.0f - rho2 * w); }

}  // namespace

extern "C" {

void pfsf_normalize_soa6(float* source, float* rcomp, float* rtens, float* conductivity,
                         const float* hydration, int32_t n, float* out_sigma_max) {
  (void)hydration;
  float smax = 1.0f;
  const bool ok = n > 0 && source && rcomp && rtens && conductivity;
  if (ok) {
    const int64_t n6 = 6 * (int64_t)n;
    for (int64_t j = 0; j < n6; j++) smax = conductivity[j] > smax ? conductivity[j] : smax;
    if (smax > 1.0f) {
      const float inv = 1.0f / smax;
      for (int64_t j = 0; j < n; j++) {
        source[j] *= inv;
        rcomp[j] *= inv;
        rtens[j] *= inv;
      }
      for (int64_t j = 0; j < n6; j++) conductivity[j] *= inv;
    }
  }
  if (out_sigma_max) *out_sigma_max = smax;
}

void pfsf_downsample_2to1(const float* fine, const uint8_t* fine_type, int32_t lxf, int32_t lyf,
                          int32_t lzf, float* coarse, uint8_t* coarse_type) {
  if (!fine || !coarse || lxf <= 0 || lyf <= 0 || lzf <= 0) return;
  const int64_t cx = (lxf + 1) / 2, cy = (lyf + 1) / 2, cz = (lzf + 1) / 2;
  const int64_t nc = cx * cy * cz;
  for (int64_t c = 0; c < nc; c++) {
    const int64_t x0 = 2 * (c % cx), y0 = 2 * ((c / cx) % cy), z0 = 2 * (c / (cx * cy));
    float sum = 0.0f;
    int cnt = 0, na = 0, ns = 0, nk = 0;
    for (int64_t z = z0; z < z0 + 2 && z < lzf; z++)
      for (int64_t y = y0; y < y0 + 2 && y < lyf; y++)
        for (int64_t x = x0; x < x0 + 2 && x < lxf; x++) {
          const int64_t f = x + lxf * (y + lyf * z);
          sum += fine[f];
          cnt++;
          if (!fine_type) continue;
          const uint8_t t = fine_type[f];
          nk += t == PFSF_VOXEL_ANCHOR;
          ns += t == PFSF_VOXEL_SOLID;
          na += t != PFSF_VOXEL_ANCHOR && t != PFSF_VOXEL_SOLID;
        }
    coarse[c] = cnt > 0 ? sum / (float)cnt : 0.0f;
    if (coarse_type) coarse_type[c] = vote(nk, ns, na);
  }
}

float pfsf_chebyshev_omega(int32_t iter, float rho_spec) {
  const float rho2 = rho_spec * rho_spec;
  if (iter <= 0 || rho2 >= 1.0f) return 1.0f;
  float prev = 1.0f, w = 2.0f / (2.0f - rho2);
  if (iter == 1) return w;  // the first weight is returned uncapped
  for (int32_t k = 2; k <= iter; k++) {
    if (4.0f - rho2 * w < PFSF_OMEGA_DENOM_EPSILON) {
      w = prev;
      break;
    }
    const float next = omega_next(rho2, w);
    if (isnan(next) || next > 2.0f) {
      w = prev;
      break;
    }
    prev = w;
    w = next;
  }
  return w < PFSF_MAX_OMEGA ? w : PFSF_MAX_OMEGA;
}

int32_t pfsf_precompute_omega_table(float rho_spec, float* out, int32_t capacity) {
  if (!out || capacity <= 0) return PFSF_ERROR_INVALID_ARG;
  const float rho2 = rho_spec * rho_spec;
  for (int32_t k = 0; k < capacity; k++) {
    if (k == 0) {
      out[0] = 1.0f;
    } else if (k == 1) {
      out[1] = rho2 < 1.0f ? 2.0f / (2.0f - rho2) : 1.0f;
    } else if (4.0f - rho2 * out[k - 1] < PFSF_OMEGA_DENOM_EPSILON) {
      out[k] = out[k - 1];
    } else {
      float w = omega_next(rho2, out[k - 1]);
      w = w > PFSF_MAX_OMEGA ? PFSF_MAX_OMEGA : w;
      out[k] = isnan(w) ? 1.0f : w;
    }
  }
  return capacity;
}

float pfsf_estimate_spectral_radius(int32_t l_max, float safety_margin) {
  if (l_max <= 1) return 0.5f;
  return (float)(cos(3.14159265358979323846 / (double)l_max) * (double)safety_margin);
}

}  // extern "C"
