/*
 * pfsf_oracle.c — CPU restatement of PFSF-CORE's per-tick structural solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against; it is built into oracle/_build/liboracle.so and may be
 * loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg.  Nothing in the product library links or calls it.
 *
 * Canonical operator (26-point, per-voxel conductivity kappa, symmetric):
 *   Block Reality/api/src/test/java/com/blockreality/api/physics/validation/
 *     VoxelPhysicsCpuReference.java:66-118   (jacobiStep: loop order z,y,x;
 *       offsets in PFSFStencil order; air neighbours skipped; anchors and air
 *       rows forced to 0; double accumulation, float store)
 *   .../physics/pfsf/PFSFStencil.java:42-43,70-84  (EDGE_P=0.5f,
 *       CORNER_P=0.1666667f, NEIGHBOR_OFFSETS order)
 *   .../physics/effective/EnergyEvaluatorCPU.java:166-185 (same weights)
 *
 * Coupling between voxel i and an in-bounds non-air neighbour j, evaluated in
 * exactly the Java expression order (float products/sums, double scaling):
 *   face   : 0.5 * (double)(ki + kj)             ki,kj = max(0, kappa) as float
 *   edge   : sqrt((double)(ki * kj)) * (double)EDGE_P
 *   corner : sqrt((double)(ki * kj)) * (double)CORNER_P
 *
 * Build with -ffp-contract=off: every product and sum rounds separately, as
 * in Java.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define T_AIR 0
#define T_SOLID 1
#define T_ANCHOR 2

static const float EDGE_P = 0.5f;
static const float CORNER_P = 0.1666667f;

/* PFSFStencil.java:70-84 — order matters for the double accumulation. */
static const int OFFS[26][3] = {
    {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0},
    {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},  {-1, 0, -1},
    {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1},
    {1, 1, 1},   {1, 1, -1},  {1, -1, 1},  {1, -1, -1},
    {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1},
};

static inline int64_t IDX(int x, int y, int z, int lx, int ly) {
  return (int64_t)x + (int64_t)lx * ((int64_t)y + (int64_t)ly * (int64_t)z);
}

static inline int is_solid(uint8_t t) { return t != T_AIR && t != T_ANCHOR; }

/* Coupling a_ij for offset class k (0..5 face, 6..17 edge, 18..25 corner). */
static inline double coupling(int k, float ki, float kj) {
  if (k < 6) {
    float s = ki + kj;
    return 0.5 * (double)s;
  }
  float p = ki * kj;
  double r = sqrt((double)p);
  return (k < 18) ? r * (double)EDGE_P : r * (double)CORNER_P;
}

static inline float kap(const float* kappa, int64_t i) {
  float k = kappa[i];
  return k > 0.0f ? k : 0.0f; /* Math.max(0f, sigma) */
}

/* One Jacobi step, VoxelPhysicsCpuReference.jacobiStep (java:66-118). */
void or_jacobi_step(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                    const float* source, const float* phi_prev, float* phi_out, float omega) {
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        uint8_t t = type[i];
        if (t == T_ANCHOR || t == T_AIR) {
          phi_out[i] = 0.0f;
          continue;
        }
        double diag = 0.0, off = 0.0;
        float ki = kap(kappa, i);
        for (int k = 0; k < 26; k++) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          if (type[j] == T_AIR) continue;
          double a = coupling(k, ki, kap(kappa, j));
          diag += a;
          off += a * (double)phi_prev[j];
        }
        double b = (double)source[i];
        double pj = (diag > 1e-15) ? (b + off) / diag : (double)phi_prev[i];
        phi_out[i] = (float)((double)omega * (pj - (double)phi_prev[i]) + (double)phi_prev[i]);
      }
}

/* Threads of the parallel per-voxel loops (RBGS colour passes, diagonal,
 * operator, failure scan).  Every reduction stays serial and each voxel's
 * value is formed by one thread in the same order, so results are the same
 * bits for any thread count.  Returns the count now in effect. */
int or_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

/* Island data prep, PFSFDataBuilder.buildFromLookups (PFSFDataBuilder.java:
 * 77-133) over a material index grid (index outside [0, nmat) = air), with
 * the conductivity rule PFSFConductivity.sigma (PFSFConductivity.java:36-97)
 * at force arm 0 and no wind bias.  mat holds nmat rows {density, rcomp,
 * rtens, indestructible (0/1)}; anchor, curing and fill may be NULL.  Java's
 * double/float steps are kept: Math.min/max on doubles (NaN if either is
 * NaN), the base and the tension ratio rounded to float, sigma * sqrt(h)
 * in float, the source in double rounded to float. */
static double j_min(double a, double b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }
static double j_max(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }
static float j_minf(float a, float b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }
static float j_maxf(float a, float b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }

static float or_sigma(const float* mi, const float* mj, int vertical) {
  double rci = mi[3] != 0.0f ? 1e6 : (double)mi[1];
  double rcj = mj[3] != 0.0f ? 1e6 : (double)mj[1];
  float base = (float)j_min(rci, rcj);
  if (base <= 0) return 0.0f;
  if (vertical) return base;
  double rti = mi[3] != 0.0f ? 1e6 : (double)mi[2];
  double rtj = mj[3] != 0.0f ? 1e6 : (double)mj[2];
  double avg = (rti + rtj) / 2.0;
  float ratio = (float)j_min(1.0, avg / j_max(base, 1.0));
  float sigmaH = base * ratio;
  float decay = 1.0f;
  float result = sigmaH * decay;
  if (isnan(result) || isinf(result)) return 0.0f;
  return result;
}

void or_build_island_data(int lx, int ly, int lz, const int32_t* mat_idx, const float* mat,
                          int nmat, const uint8_t* anchor, const float* curing, const float* fill,
                          uint8_t* type, float* source, float* cond, float* rcomp, float* rtens,
                          float* max_phi) {
  static const int DV[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  int64_t n = (int64_t)lx * ly * lz;
  memset(type, 0, (size_t)n);
  memset(source, 0, sizeof(float) * (size_t)n);
  memset(cond, 0, sizeof(float) * (size_t)(6 * n));
  memset(rcomp, 0, sizeof(float) * (size_t)n);
  memset(rtens, 0, sizeof(float) * (size_t)n);
  memset(max_phi, 0, sizeof(float) * (size_t)n);
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        int m = mat_idx[i];
        if (m < 0 || m >= nmat) continue;
        const float* mi = mat + 4 * m;
        float fillRatio = fill ? fill[i] : 1.0f;
        float hDeg = curing ? j_maxf(0.0f, j_minf(1.0f, curing[i])) : 1.0f;
        float sigmaScale = (float)sqrt((double)hDeg);
        source[i] = (float)((double)mi[0] * fillRatio * 9.81 * 1.0);
        type[i] = (anchor && anchor[i]) ? T_ANCHOR : T_SOLID;
        max_phi[i] = INFINITY;
        rcomp[i] = mi[1];
        rtens[i] = mi[2];
        for (int dir = 0; dir < 6; dir++) {
          int nx = x + DV[dir][0], ny = y + DV[dir][1], nz = z + DV[dir][2];
          float sigma = 0.0f;
          if (nx >= 0 && nx < lx && ny >= 0 && ny < ly && nz >= 0 && nz < lz) {
            int mj = mat_idx[IDX(nx, ny, nz, lx, ly)];
            if (mj >= 0 && mj < nmat) sigma = or_sigma(mi, mat + 4 * mj, dir == 2 || dir == 3);
          }
          cond[dir * n + i] = sigma * sigmaScale;
        }
      }
}

/* In-place 8-colour Gauss-Seidel with the jacobiStep per-voxel update.
 * Colour = (x&1) | (y&1)<<1 | (z&1)<<2, colours 0..7 ascending
 * (SH/rbgs_smooth.comp.glsl:86-87, NP/src/solver/jacobi_solver.cpp:163-175).
 * No damping and no macro-block skip. */
void or_rbgs(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
             const float* source, float* phi, int sweeps) {
  for (int s = 0; s < sweeps; s++)
    for (int c = 0; c < 8; c++) {
#pragma omp parallel for schedule(static)
      for (int z = (c >> 2) & 1; z < lz; z += 2)
        for (int y = (c >> 1) & 1; y < ly; y += 2)
          for (int x = c & 1; x < lx; x += 2) {
            int64_t i = IDX(x, y, z, lx, ly);
            uint8_t t = type[i];
            if (t == T_ANCHOR || t == T_AIR) {
              phi[i] = 0.0f;
              continue;
            }
            double diag = 0.0, off = 0.0;
            float ki = kap(kappa, i);
            for (int k = 0; k < 26; k++) {
              int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
              if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
              int64_t j = IDX(nx, ny, nz, lx, ly);
              if (type[j] == T_AIR) continue;
              double a = coupling(k, ki, kap(kappa, j));
              diag += a;
              off += a * (double)phi[j];
            }
            double b = (double)source[i];
            double pj = (diag > 1e-15) ? (b + off) / diag : (double)phi[i];
            phi[i] = (float)((pj - (double)phi[i]) + (double)phi[i]);
          }
    }
}

/* Diagonal D_i = sum_j a_ij over non-air neighbours (anchors included); 0 on
 * non-solid rows.  Same accumulation order as jacobiStep's diagA. */
void or_diag(int lx, int ly, int lz, const uint8_t* type, const float* kappa, double* d) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        if (!is_solid(type[i])) {
          d[i] = 0.0;
          continue;
        }
        double diag = 0.0;
        float ki = kap(kappa, i);
        for (int k = 0; k < 26; k++) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          if (type[j] == T_AIR) continue;
          diag += coupling(k, ki, kap(kappa, j));
        }
        d[i] = diag;
      }
}

/* y = A x over solid rows: y_i = sum_j a_ij (x_i - x_j), x_j := 0 for anchor
 * neighbours (Dirichlet), air skipped.  y = 0 on non-solid rows. */
void or_apply(int lx, int ly, int lz, const uint8_t* type, const float* kappa, const double* x,
              double* y) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < lz; z++)
    for (int yy = 0; yy < ly; yy++)
      for (int xx = 0; xx < lx; xx++) {
        int64_t i = IDX(xx, yy, z, lx, ly);
        if (!is_solid(type[i])) {
          y[i] = 0.0;
          continue;
        }
        double acc = 0.0;
        float ki = kap(kappa, i);
        double xi = x[i];
        for (int k = 0; k < 26; k++) {
          int nx = xx + OFFS[k][0], ny = yy + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          uint8_t tj = type[j];
          if (tj == T_AIR) continue;
          double xj = is_solid(tj) ? x[j] : 0.0;
          acc += coupling(k, ki, kap(kappa, j)) * (xi - xj);
        }
        y[i] = acc;
      }
}

/* ||b - A x||_2 / ||b||_2 over solid voxels, FP64.  Writes r if non-NULL. */
double or_residual(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                   const float* b, const double* x, double* r) {
  int64_t n = (int64_t)lx * ly * lz;
  double* ax = (double*)malloc(sizeof(double) * (size_t)n);
  or_apply(lx, ly, lz, type, kappa, x, ax);
  double rr = 0.0, bb = 0.0;
  for (int64_t i = 0; i < n; i++) {
    double ri = 0.0;
    if (is_solid(type[i])) {
      ri = (double)b[i] - ax[i];
      bb += (double)b[i] * (double)b[i];
    }
    rr += ri * ri;
    if (r) r[i] = ri;
  }
  free(ax);
  return bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
}

/* FP64 Jacobi-preconditioned CG on the solid rows to ||r||/||b|| <= tol.
 * x holds the initial guess on entry and the solution on exit.  Returns the
 * iteration count; *relres gets the final true relative residual. */
int or_pcg(int lx, int ly, int lz, const uint8_t* type, const float* kappa, const float* b,
           double* x, double tol, int maxit, double* relres) {
  int64_t n = (int64_t)lx * ly * lz;
  double* d = (double*)malloc(sizeof(double) * (size_t)n);
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  or_diag(lx, ly, lz, type, kappa, d);
  for (int64_t i = 0; i < n; i++)
    if (!is_solid(type[i]) || d[i] <= 0.0) x[i] = 0.0;
  or_residual(lx, ly, lz, type, kappa, b, x, r);
  double bb = 0.0;
  for (int64_t i = 0; i < n; i++)
    if (is_solid(type[i])) bb += (double)b[i] * (double)b[i];
  double bn = sqrt(bb > 0.0 ? bb : 1.0);
  double rz = 0.0;
  for (int64_t i = 0; i < n; i++) {
    z[i] = (d[i] > 0.0) ? r[i] / d[i] : 0.0;
    p[i] = z[i];
    rz += r[i] * z[i];
  }
  int it = 0;
  double rn = 0.0;
  for (int64_t i = 0; i < n; i++) rn += r[i] * r[i];
  rn = sqrt(rn);
  while (rn / bn > tol && it < maxit) {
    or_apply(lx, ly, lz, type, kappa, p, q);
    double pq = 0.0;
    for (int64_t i = 0; i < n; i++) pq += p[i] * q[i];
    if (pq <= 0.0) break;
    double alpha = rz / pq;
    double rz2 = 0.0;
    rn = 0.0;
    for (int64_t i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = (d[i] > 0.0) ? r[i] / d[i] : 0.0;
      rz2 += r[i] * z[i];
      rn += r[i] * r[i];
    }
    rn = sqrt(rn);
    double beta = rz2 / rz;
    rz = rz2;
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    it++;
  }
  if (relres) *relres = or_residual(lx, ly, lz, type, kappa, b, x, NULL);
  free(d); free(r); free(z); free(p); free(q);
  return it;
}

/* Failure scan on the canonical operator.  Code precedence follows
 * SH/failure_scan.comp.glsl:41-176:
 *   non-solid -> 0; no coupling (D_i == 0) or phi > 1e6 -> NO_SUPPORT(3);
 *   rcomp > 0 && flux_in > rcomp -> CRUSHING(2);
 *   rtens > 0 && flux_out > rtens -> TENSION(4); else 0.
 * flux_in  = sum a_ij (phi_j - phi_i) over phi_j > phi_i,
 * flux_out = sum a_ij (phi_i - phi_j) over phi_j < phi_i, all non-air
 * neighbours (anchors at their stored phi), FP64 in stencil order. */
void or_failure_scan(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                     const float* phi, const float* rcomp, const float* rtens, float phi_orphan,
                     uint8_t* code) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        code[i] = 0;
        if (!is_solid(type[i])) continue;
        float ki = kap(kappa, i);
        double pi = (double)phi[i];
        double diag = 0.0, fin = 0.0, fout = 0.0;
        for (int k = 0; k < 26; k++) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          if (type[j] == T_AIR) continue;
          double a = coupling(k, ki, kap(kappa, j));
          diag += a;
          double dp = (double)phi[j] - pi;
          if (dp > 0.0)
            fin += a * dp;
          else if (dp < 0.0)
            fout += a * (-dp);
        }
        if (diag == 0.0 || phi[i] > phi_orphan) {
          code[i] = 3;
          continue;
        }
        if (rcomp[i] > 0.0f && fin > (double)rcomp[i]) {
          code[i] = 2;
          continue;
        }
        if (rtens[i] > 0.0f && fout > (double)rtens[i]) {
          code[i] = 4;
          continue;
        }
      }
}

/* Elastic energy density psi_e = 0.5 * kappa * |grad phi|^2 (float).
 * Per axis: central difference when both face neighbours are non-air and in
 * bounds, one-sided when only one is, 0 when neither.  Only solid voxels get a
 * value; others 0.  Expression order is fixed so the GPU can match bitwise. */
static inline float axis_grad(const uint8_t* type, const float* phi, int64_t i, int has_m,
                              int has_p, int64_t jm, int64_t jp) {
  int om = has_m && type[jm] != T_AIR;
  int op = has_p && type[jp] != T_AIR;
  if (om && op) return (phi[jp] - phi[jm]) * 0.5f;
  if (op) return phi[jp] - phi[i];
  if (om) return phi[i] - phi[jm];
  return 0.0f;
}

void or_psi(int lx, int ly, int lz, const uint8_t* type, const float* kappa, const float* phi,
            float* psi) {
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        if (!is_solid(type[i])) {
          psi[i] = 0.0f;
          continue;
        }
        int64_t sx = 1, sy = lx, sz = (int64_t)lx * ly;
        float gx = axis_grad(type, phi, i, x > 0, x + 1 < lx, i - sx, i + sx);
        float gy = axis_grad(type, phi, i, y > 0, y + 1 < ly, i - sy, i + sy);
        float gz = axis_grad(type, phi, i, z > 0, z + 1 < lz, i - sz, i + sz);
        float g2 = gx * gx;
        g2 = g2 + gy * gy;
        g2 = g2 + gz * gz;
        float h = 0.5f * kap(kappa, i);
        psi[i] = h * g2;
      }
}

/* H = max(H, psi_e) on solid voxels (monotone history field). */
void or_history_update(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                       const float* phi, float* H) {
  int64_t n = (int64_t)lx * ly * lz;
  float* psi = (float*)malloc(sizeof(float) * (size_t)n);
  or_psi(lx, ly, lz, type, kappa, phi, psi);
  for (int64_t i = 0; i < n; i++)
    if (is_solid(type[i]) && psi[i] > H[i]) H[i] = psi[i];
  free(psi);
}

/* 26-connected reachability from anchors over non-air voxels
 * (VoxelPhysicsCpuReference.findOrphans, java:138-166).  reach[i]=1 for
 * reached voxels.  Returns the number of orphan solid voxels. */
int64_t or_anchored(int lx, int ly, int lz, const uint8_t* type, uint8_t* reach) {
  int64_t n = (int64_t)lx * ly * lz;
  int64_t* q = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t head = 0, tail = 0;
  for (int64_t i = 0; i < n; i++) {
    reach[i] = 0;
    if (type[i] == T_ANCHOR) {
      reach[i] = 1;
      q[tail++] = i;
    }
  }
  while (head < tail) {
    int64_t i = q[head++];
    int x = (int)(i % lx), y = (int)((i / lx) % ly), z = (int)(i / ((int64_t)lx * ly));
    for (int k = 0; k < 26; k++) {
      int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
      if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
      int64_t j = IDX(nx, ny, nz, lx, ly);
      if (type[j] == T_AIR || reach[j]) continue;
      reach[j] = 1;
      q[tail++] = j;
    }
  }
  int64_t orphans = 0;
  for (int64_t i = 0; i < n; i++)
    if (is_solid(type[i]) && !reach[i]) orphans++;
  free(q);
  return orphans;
}

/* sigma-max normalisation, PFSFDataBuilder.normalizeSoA6JavaRef
 * (PFSFDataBuilder.java:257-278) / NP/src/compute/normalize.cpp:29-63:
 * s = max(1, max cond); if s > 1 divide source, rcomp, rtens, cond by s
 * (multiplication by the float reciprocal, as the native path does). */
float or_normalize_soa6(float* source, float* rcomp, float* rtens, float* cond, int32_t n) {
  float smax = 1.0f;
  for (int64_t j = 0; j < 6 * (int64_t)n; j++)
    if (cond[j] > smax) smax = cond[j];
  if (smax > 1.0f) {
    float inv = 1.0f / smax;
    for (int64_t j = 0; j < n; j++) {
      source[j] *= inv;
      rcomp[j] *= inv;
      rtens[j] *= inv;
    }
    for (int64_t j = 0; j < 6 * (int64_t)n; j++) cond[j] *= inv;
  }
  return smax;
}

/* Chebyshev omega schedule, PFSFScheduler.computeOmega /
 * NP/src/compute/chebyshev.cpp:30-57: w0=1, w1=2/(2-rho^2),
 * wk=4/(4-rho^2 w_{k-1}); stop at the previous value when the denominator
 * drops below 0.01 or the result leaves (0,2]; clamp at 1.98. */
float or_chebyshev_omega(int32_t iter, float rho) {
  if (iter <= 0) return 1.0f;
  float r2 = rho * rho;
  if (r2 >= 1.0f) return 1.0f;
  if (iter == 1) return 2.0f / (2.0f - r2);
  float wp = 1.0f, w = 2.0f / (2.0f - r2);
  for (int32_t k = 2; k <= iter; k++) {
    float den = 4.0f - r2 * w;
    if (den < 0.01f) {
      w = wp;
      break;
    }
    float wn = 4.0f / den;
    if (wn > 2.0f || wn != wn) {
      w = wp;
      break;
    }
    wp = w;
    w = wn;
  }
  return w < 1.98f ? w : 1.98f;
}

/* 2:1 coarse downsample of a per-voxel float field and voxel types
 * (NP/src/compute/downsample.cpp:39-98): float mean over in-bounds children,
 * type = anchor if anchors >= solids and >= airs, else solid if solids >= airs,
 * else air. */
void or_downsample_2to1(const float* fine, const uint8_t* ftype, int lxf, int lyf, int lzf,
                        float* coarse, uint8_t* ctype) {
  int lxc = (lxf + 1) / 2, lyc = (lyf + 1) / 2, lzc = (lzf + 1) / 2;
  for (int zc = 0; zc < lzc; zc++)
    for (int yc = 0; yc < lyc; yc++)
      for (int xc = 0; xc < lxc; xc++) {
        float sum = 0.0f;
        int cnt = 0, na = 0, ns = 0, nk = 0;
        for (int dz = 0; dz < 2; dz++) {
          int zf = 2 * zc + dz;
          if (zf >= lzf) continue;
          for (int dy = 0; dy < 2; dy++) {
            int yf = 2 * yc + dy;
            if (yf >= lyf) continue;
            for (int dx = 0; dx < 2; dx++) {
              int xf = 2 * xc + dx;
              if (xf >= lxf) continue;
              int64_t f = IDX(xf, yf, zf, lxf, lyf);
              sum += fine[f];
              cnt++;
              uint8_t t = ftype[f];
              if (t == T_ANCHOR) nk++;
              else if (t == T_SOLID) ns++;
              else na++;
            }
          }
        }
        int64_t c = IDX(xc, yc, zc, lxc, lyc);
        coarse[c] = cnt > 0 ? sum / (float)cnt : 0.0f;
        if (nk >= ns && nk >= na) ctype[c] = T_ANCHOR;
        else if (ns >= na) ctype[c] = T_SOLID;
        else ctype[c] = T_AIR;
      }
}

/* Adjoint-ready helpers.  d(a_ij)/d(kappa_i) for the three coupling classes,
 * on the real-valued formulas a = (ki+kj)/2 and c*sqrt(ki kj). */
void or_energy_grad_kappa(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                          const double* u, const double* v, double* g) {
  /* g_i = sum over couplings touching i of d a_ij/d kappa_i * (u_i-u_j)(v_i-v_j),
   * with u,v := 0 at anchors.  This is dL/dkappa for L = -lambda^T A phi pairs. */
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        g[i] = 0.0;
        if (type[i] == T_AIR) continue;
        double ki = (double)kap(kappa, i);
        double ui = is_solid(type[i]) ? u[i] : 0.0;
        double vi = is_solid(type[i]) ? v[i] : 0.0;
        double acc = 0.0;
        for (int k = 0; k < 26; k++) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          uint8_t tj = type[j];
          if (tj == T_AIR) continue;
          if (!is_solid(type[i]) && !is_solid(tj)) continue;
          double kj = (double)kap(kappa, j);
          double da;
          if (k < 6)
            da = 0.5;
          else {
            double c = (k < 18) ? (double)EDGE_P : (double)CORNER_P;
            da = (ki > 0.0) ? c * 0.5 * sqrt(kj / ki) : 0.0;
          }
          double uj = is_solid(tj) ? u[j] : 0.0;
          double vj = is_solid(tj) ? v[j] : 0.0;
          acc += da * (ui - uj) * (vi - vj);
        }
        g[i] = acc;
      }
}

/* ── Load-path loss (adjoint workload) ─────────────────────────────────────
 * Not a reference function: the reference has no adjoint.  The loss is
 * defined on the reference's crushing flux (the failure scan's flux_in,
 * SH/failure_scan.comp.glsl:116-158 restated above):
 *   f_i = sum_j a_ij (phi_j - phi_i)^+   over non-air neighbours j of a free i
 *   L   = sum_i softplus(f_i - rcomp_i)  over free voxels i
 * phi is 0 outside the free set (mask `fr`).  FP64 throughout; the pieces
 * below are the exact partial derivatives of L, checked against finite
 * differences in the tests. */
static inline double softplus(double u) { return (u > 0 ? u : 0.0) + log1p(exp(-fabs(u))); }

double or_flux_loss(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                    const uint8_t* fr, const double* phi, const float* rcomp, double* w) {
  double L = 0.0;
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        if (w) w[i] = 0.0;
        if (!fr[i]) continue;
        float ki = kap(kappa, i);
        double pi = phi[i], f = 0.0;
        for (int k = 0; k < 26; k++) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          if (type[j] == T_AIR) continue;
          double pj = fr[j] ? phi[j] : 0.0;
          if (pj > pi) f += coupling(k, ki, kap(kappa, j)) * (pj - pi);
        }
        double u = f - (double)rcomp[i];
        L += softplus(u);
        if (w) w[i] = 1.0 / (1.0 + exp(-u));
      }
  return L;
}

/* g_k = dL/dphi_k = -w_k sum_j a_kj [phi_j > phi_k] + sum_j w_j a_kj [phi_k > phi_j]
 * (free k; 0 elsewhere; w = 0 off the free set). */
void or_flux_grad_phi(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                      const uint8_t* fr, const double* phi, const double* w, double* g) {
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        g[i] = 0.0;
        if (!fr[i]) continue;
        float ki = kap(kappa, i);
        double pi = phi[i], acc = 0.0;
        for (int k = 0; k < 26; k++) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          if (type[j] == T_AIR) continue;
          double a = coupling(k, ki, kap(kappa, j));
          double pj = fr[j] ? phi[j] : 0.0;
          if (pj > pi) acc -= w[i] * a;
          if (pi > pj && fr[j]) acc += w[j] * a;
        }
        g[i] = acc;
      }
}

/* Explicit part of dL/dkappa_m at fixed phi:
 *   sum_j da_mj/dkappa_m [w_m (phi_j - phi_m)^+ + w_j (phi_m - phi_j)^+]
 * over non-air neighbours (da as in or_energy_grad_kappa). */
void or_flux_grad_kappa(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                        const uint8_t* fr, const double* phi, const double* w, double* g) {
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        g[i] = 0.0;
        if (type[i] == T_AIR) continue;
        double ki = (double)kap(kappa, i);
        double pi = fr[i] ? phi[i] : 0.0, wi = fr[i] ? w[i] : 0.0, acc = 0.0;
        for (int k = 0; k < 26; k++) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          if (type[j] == T_AIR) continue;
          double kj = (double)kap(kappa, j);
          double da;
          if (k < 6)
            da = 0.5;
          else {
            double c = (k < 18) ? (double)EDGE_P : (double)CORNER_P;
            da = (ki > 0.0) ? c * 0.5 * sqrt(kj / ki) : 0.0;
          }
          double pj = fr[j] ? phi[j] : 0.0, wj = fr[j] ? w[j] : 0.0;
          double up = pj > pi ? pj - pi : 0.0, dn = pi > pj ? pi - pj : 0.0;
          acc += da * (wi * up + wj * dn);
        }
        g[i] = acc;
      }
}

/* ── Connected components (label propagation) ──────────────────────────
 * Live voxels are solid (1) or anchor (2); other types are not labelled.
 * conn = 26 uses the 26 offsets of PFSFStencil order, conn = 6 the faces
 * only (LabelPropagation.java:277-288). */
#define LBL_NONE 0xFFFFFFFFu

static inline int lbl_live(uint8_t t) { return t == 1 || t == T_ANCHOR; }

/* LabelPropagation.shiloachVishkin (LabelPropagation.java:199-270): labels
 * start at i+1 (air LBL_NONE); each round is a hook-to-min sweep in z,y,x
 * order that reads the array it is updating, then one pointer jump that
 * reads the pre-jump labels; rounds repeat until neither changes a label
 * (cap max(64, 2n)).  Returns the number of rounds. */
int or_label_sv(int lx, int ly, int lz, const uint8_t* type, int conn, uint32_t* id) {
  int64_t n = (int64_t)lx * ly * lz;
  uint32_t* prev = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) id[i] = lbl_live(type[i]) ? (uint32_t)(i + 1) : LBL_NONE;
  int64_t cap = 2 * n > 64 ? 2 * n : 64;
  int rounds = 0, changed = 1;
  while (changed && rounds < cap) {
    rounds++;
    changed = 0;
    for (int z = 0; z < lz; z++)
      for (int y = 0; y < ly; y++)
        for (int x = 0; x < lx; x++) {
          int64_t i = IDX(x, y, z, lx, ly);
          uint32_t me = id[i];
          if (me == LBL_NONE) continue;
          uint32_t mn = me;
          for (int k = 0; k < (conn == 6 ? 6 : 26); k++) {
            int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
            if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
            uint32_t nl = id[IDX(nx, ny, nz, lx, ly)];
            if (nl != LBL_NONE && nl < mn) mn = nl;
          }
          if (mn < me) {
            id[i] = mn;
            changed = 1;
          }
        }
    memcpy(prev, id, sizeof(uint32_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
      uint32_t l = prev[i];
      if (l == LBL_NONE) continue;
      int64_t p = (int64_t)l - 1;
      if (p < 0 || p >= n) continue;
      uint32_t pl = prev[p];
      if (pl != LBL_NONE && pl < l) {
        id[i] = pl;
        changed = 1;
      }
    }
  }
  free(prev);
  return rounds;
}

/* Breadth-first partition (LabelPropagation.bfsComponents,
 * LabelPropagation.java:121-168, on the flat grid): every live voxel gets
 * 1 + the smallest flat index of its component.  Returns the number of
 * components; anchored[c] is not produced here (see or_label_summarise). */
int64_t or_label_bfs(int lx, int ly, int lz, const uint8_t* type, int conn, uint32_t* id) {
  int64_t n = (int64_t)lx * ly * lz, comps = 0;
  int64_t* q = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) id[i] = LBL_NONE;
  for (int64_t s = 0; s < n; s++) {
    if (!lbl_live(type[s]) || id[s] != LBL_NONE) continue;
    /* s is the first unvisited voxel of its component in flat order, so it
     * is the component's smallest index */
    uint32_t lab = (uint32_t)(s + 1);
    int64_t head = 0, tail = 0;
    q[tail++] = s;
    id[s] = lab;
    comps++;
    while (head < tail) {
      int64_t i = q[head++];
      int x = (int)(i % lx), y = (int)((i / lx) % ly), z = (int)(i / ((int64_t)lx * ly));
      for (int k = 0; k < (conn == 6 ? 6 : 26); k++) {
        int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
        if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
        int64_t j = IDX(nx, ny, nz, lx, ly);
        if (!lbl_live(type[j]) || id[j] != LBL_NONE) continue;
        id[j] = lab;
        q[tail++] = j;
      }
    }
  }
  free(q);
  return comps;
}

/* Component records of converged labels: label_prop_summarise_alloc +
 * _aggregate as replayed by PFSFLabelPropCpuSimulator
 * (PFSFLabelPropCpuSimulator.java:164-238).  Roots (id[i] == i+1) take
 * slots in ascending flat order; the first maxc get a 12-word record
 * {root label, count, anchored, 0, min x, min y, min z, 0, max x, max y,
 * max z, 0}; later roots set *overflow.  rec (maxc*12 words) is zeroed first,
 * as the recorder clears it.  Returns the number of allocated slots. */
int or_label_summarise(int lx, int ly, int lz, const uint8_t* type, const uint32_t* id,
                       int maxc, uint32_t* rec, int* overflow) {
  int64_t n = (int64_t)lx * ly * lz;
  uint32_t* slot = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  memset(rec, 0, sizeof(uint32_t) * 12 * (size_t)maxc);
  int num = 0;
  *overflow = 0;
  for (int64_t i = 0; i < n; i++) {
    slot[i] = LBL_NONE;
    if (id[i] == LBL_NONE || id[i] != (uint32_t)(i + 1)) continue;
    if (num >= maxc) {
      *overflow = 1;
      continue;
    }
    slot[i] = (uint32_t)num;
    uint32_t* r = rec + 12 * (size_t)num;
    r[0] = id[i];
    r[4] = r[5] = r[6] = 0xFFFFFFFFu;
    num++;
  }
  for (int64_t i = 0; i < n; i++) {
    uint32_t l = id[i];
    if (l == LBL_NONE) continue;
    int64_t root = (int64_t)l - 1;
    if (root < 0 || root >= n || slot[root] == LBL_NONE) continue;
    uint32_t* r = rec + 12 * (size_t)slot[root];
    uint32_t x = (uint32_t)(i % lx), y = (uint32_t)((i / lx) % ly),
             z = (uint32_t)(i / ((int64_t)lx * ly));
    r[1]++;
    if (type[i] == T_ANCHOR) r[2] |= 1u;
    if (x < r[4]) r[4] = x;
    if (y < r[5]) r[5] = y;
    if (z < r[6]) r[6] = z;
    if (x > r[8]) r[8] = x;
    if (y > r[9]) r[9] = y;
    if (z > r[10]) r[10] = z;
  }
  free(slot);
  return num;
}

/* ── SoA-6 operator (face conductivities per voxel) ─────────────────────
 * The drop-in tick receives six face conductivities per voxel, SoA order
 * cond[d * n + i], d = -x, +x, -y, +y, -z, +z (PFSFDataBuilder.java:124-132,
 * PFSFConductivity.java:37-91).  The reference GPU operator
 * (SH/pcg_matvec.comp.glsl:42-128, SH/rbgs_smooth.comp.glsl:116-163) builds
 * edge and corner weights from voxel i's own faces only, so it is not
 * symmetric.  This operator takes the mean of the two endpoints' half
 * weights, which makes it symmetric and equal to the shader's weight when
 * both endpoints agree:
 *   face   a = 0.5 * (c_i[d] + c_j[opp d])
 *   edge   a = EDGE_P   * 0.5 * (sqrt(c_i[da] c_i[db]) + sqrt(c_j[da'] c_j[db']))
 *   corner a = CORNER_P * 0.5 * (cbrt(c_i[dx] c_i[dy] c_i[dz]) + cbrt(c_j[..']))
 * with d the face directions of the offset's components (' = opposite sign),
 * c = cond where 0 < cond <= FLT_MAX and 0 otherwise, all in FP64 with the
 * products formed left to right.  Pairs with an air or out-of-grid endpoint
 * have no coupling.  The 13 couplings of a voxel to its forward neighbours
 * (the even offsets of the stencil order: dz > 0, or dz = 0 and dy > 0, or
 * dz = dy = 0 and dx > 0) are tabulated as cw[(k / 2) * n + i]; an odd
 * (backward) offset k of voxel i reads the entry of its opposite offset at
 * the neighbour. */
static const int OPP[26] = {1,  0,  3,  2,  5,  4,  9,  8,  7,  6,  13, 12, 11,
                            10, 17, 16, 15, 14, 25, 24, 23, 22, 21, 20, 19, 18};

/* Cube root from basic IEEE operations only (each correctly rounded), so
 * that any conforming FP64 implementation gives the same bits: v = m 2^(3q)
 * with m in [0.5, 4), a chord of cbrt over that interval as the first guess
 * (error below 11 %), then six Newton steps y <- (2y + m / y^2) / 3, which
 * converge quadratically to within an ulp of cbrt(m). */
double or_cbrt(double v) {
  if (!(v > 0.0)) return 0.0;
  if (v > 1.7976931348623157e308) return v;
  int e;
  double m = frexp(v, &e);
  int r = e % 3;
  if (r < 0) r += 3;
  int q = (e - r) / 3;
  m = ldexp(m, r);
  double y = 0.79370052598409979 + 0.22677165 * (m - 0.5);
  for (int it = 0; it < 6; it++) y = (2.0 * y + m / (y * y)) / 3.0;
  return ldexp(y, q);
}

static inline double soa6_c(const float* cond, int64_t n, int d, int64_t i) {
  float c = cond[d * n + i];
  return (c > 0.0f && c <= 3.40282347e38f) ? (double)c : 0.0;
}

/* Face direction index of one offset component: axis a (0 x, 1 y, 2 z),
 * sign s != 0. */
static inline int soa6_dir(int a, int s) { return 2 * a + (s > 0 ? 1 : 0); }

/* Half weight of voxel i toward offset (dx,dy,dz) (sign flipped with
 * flip = -1: the neighbour's half toward i). */
static double soa6_half(const float* cond, int64_t n, int64_t i, const int* o, int flip) {
  int dirs[3], nd = 0;
  for (int a = 0; a < 3; a++)
    if (o[a] != 0) dirs[nd++] = soa6_dir(a, o[a] * flip);
  if (nd == 1) return soa6_c(cond, n, dirs[0], i);
  if (nd == 2) return sqrt(soa6_c(cond, n, dirs[0], i) * soa6_c(cond, n, dirs[1], i));
  return or_cbrt(soa6_c(cond, n, dirs[0], i) * soa6_c(cond, n, dirs[1], i) *
                 soa6_c(cond, n, dirs[2], i));
}

void or_soa6_couplings(int lx, int ly, int lz, const uint8_t* type, const float* cond,
                       double* cw) {
  int64_t n = (int64_t)lx * ly * lz;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        for (int k = 0; k < 26; k += 2) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          double a = 0.0;
          if (type[i] != T_AIR && nx >= 0 && nx < lx && ny >= 0 && ny < ly && nz >= 0 &&
              nz < lz) {
            int64_t j = IDX(nx, ny, nz, lx, ly);
            if (type[j] != T_AIR) {
              double h = 0.5 * (soa6_half(cond, n, i, OFFS[k], 1) +
                                soa6_half(cond, n, j, OFFS[k], -1));
              int nd = (OFFS[k][0] != 0) + (OFFS[k][1] != 0) + (OFFS[k][2] != 0);
              a = nd == 1 ? h : (nd == 2 ? (double)EDGE_P * h : (double)CORNER_P * h);
            }
          }
          cw[(k / 2) * n + i] = a;
        }
      }
}

/* Coupling of voxel i to its in-grid neighbour j at offset k. */
static inline double soa6_at(const double* cw, int64_t n, int k, int64_t i, int64_t j) {
  return (k & 1) ? cw[(OPP[k] / 2) * n + j] : cw[(k / 2) * n + i];
}

void or_soa6_apply(int lx, int ly, int lz, const uint8_t* type, const double* cw,
                   const double* x, double* y) {
  int64_t n = (int64_t)lx * ly * lz;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < lz; z++)
    for (int yy = 0; yy < ly; yy++)
      for (int xx = 0; xx < lx; xx++) {
        int64_t i = IDX(xx, yy, z, lx, ly);
        if (!is_solid(type[i])) {
          y[i] = 0.0;
          continue;
        }
        double acc = 0.0, xi = x[i];
        for (int k = 0; k < 26; k++) {
          int nx = xx + OFFS[k][0], ny = yy + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          uint8_t tj = type[j];
          if (tj == T_AIR) continue;
          double xj = is_solid(tj) ? x[j] : 0.0;
          acc += soa6_at(cw, n, k, i, j) * (xi - xj);
        }
        y[i] = acc;
      }
}

void or_soa6_diag(int lx, int ly, int lz, const uint8_t* type, const double* cw, double* d) {
  int64_t n = (int64_t)lx * ly * lz;
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        double acc = 0.0;
        if (is_solid(type[i]))
          for (int k = 0; k < 26; k++) {
            int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
            if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
            int64_t j = IDX(nx, ny, nz, lx, ly);
            if (type[j] == T_AIR) continue;
            acc += soa6_at(cw, n, k, i, j);
          }
        d[i] = acc;
      }
}

/* Failure scan of or_failure_scan with the SoA-6 couplings (same precedence,
 * FP64 fluxes in stencil order). */
void or_soa6_failure_scan(int lx, int ly, int lz, const uint8_t* type, const double* cw,
                          const float* phi, const float* rcomp, const float* rtens,
                          float phi_orphan, uint8_t* code) {
  int64_t n = (int64_t)lx * ly * lz;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        code[i] = 0;
        if (!is_solid(type[i])) continue;
        double pi = (double)phi[i];
        double diag = 0.0, fin = 0.0, fout = 0.0;
        for (int k = 0; k < 26; k++) {
          int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          if (type[j] == T_AIR) continue;
          double a = soa6_at(cw, n, k, i, j);
          diag += a;
          double dp = (double)phi[j] - pi;
          if (dp > 0.0)
            fin += a * dp;
          else if (dp < 0.0)
            fout += a * (-dp);
        }
        if (diag == 0.0 || phi[i] > phi_orphan) {
          code[i] = 3;
          continue;
        }
        if (rcomp[i] > 0.0f && fin > (double)rcomp[i]) {
          code[i] = 2;
          continue;
        }
        if (rtens[i] > 0.0f && fout > (double)rtens[i]) code[i] = 4;
      }
}

/* Elastic energy density with a per-axis conductivity, the anisotropic form
 * of 0.5 * kappa * |grad phi|^2:
 *   psi = 0.5 * (kx gx^2 + ky gy^2 + kz gz^2),  k_a = 0.5 * (c[-a] + c[+a]),
 * gradients as in or_psi; FP32 in this order. */
void or_soa6_psi(int lx, int ly, int lz, const uint8_t* type, const float* cond,
                 const float* phi, float* psi) {
  int64_t n = (int64_t)lx * ly * lz;
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        if (!is_solid(type[i])) {
          psi[i] = 0.0f;
          continue;
        }
        int64_t sy = lx, sz = (int64_t)lx * ly;
        float gx = axis_grad(type, phi, i, x > 0, x + 1 < lx, i - 1, i + 1);
        float gy = axis_grad(type, phi, i, y > 0, y + 1 < ly, i - sy, i + sy);
        float gz = axis_grad(type, phi, i, z > 0, z + 1 < lz, i - sz, i + sz);
        float k[3];
        for (int a = 0; a < 3; a++)
          k[a] = 0.5f * ((float)soa6_c(cond, n, 2 * a, i) + (float)soa6_c(cond, n, 2 * a + 1, i));
        float e = k[0] * (gx * gx);
        e = e + k[1] * (gy * gy);
        e = e + k[2] * (gz * gz);
        psi[i] = 0.5f * e;
      }
}

void or_soa6_history_update(int lx, int ly, int lz, const uint8_t* type, const float* cond,
                            const float* phi, float* H) {
  int64_t n = (int64_t)lx * ly * lz;
  float* psi = (float*)malloc(sizeof(float) * (size_t)n);
  or_soa6_psi(lx, ly, lz, type, cond, phi, psi);
  for (int64_t i = 0; i < n; i++)
    if (is_solid(type[i]) && psi[i] > H[i]) H[i] = psi[i];
  free(psi);
}

/* 8-colour Gauss-Seidel of or_rbgs with the SoA-6 couplings. */
void or_soa6_rbgs(int lx, int ly, int lz, const uint8_t* type, const double* cw,
                  const float* source, float* phi, int sweeps) {
  int64_t n = (int64_t)lx * ly * lz;
  for (int s = 0; s < sweeps; s++)
    for (int c = 0; c < 8; c++) {
#pragma omp parallel for schedule(static)
      for (int z = (c >> 2) & 1; z < lz; z += 2)
        for (int y = (c >> 1) & 1; y < ly; y += 2)
          for (int x = c & 1; x < lx; x += 2) {
            int64_t i = IDX(x, y, z, lx, ly);
            if (!is_solid(type[i])) {
              phi[i] = 0.0f;
              continue;
            }
            double diag = 0.0, off = 0.0;
            for (int k = 0; k < 26; k++) {
              int nx = x + OFFS[k][0], ny = y + OFFS[k][1], nz = z + OFFS[k][2];
              if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
              int64_t j = IDX(nx, ny, nz, lx, ly);
              if (type[j] == T_AIR) continue;
              double a = soa6_at(cw, n, k, i, j);
              diag += a;
              off += a * (double)phi[j];
            }
            double pj = (diag > 1e-15) ? ((double)source[i] + off) / diag : (double)phi[i];
            phi[i] = (float)pj;
          }
    }
}

/* ── Phase-field damage evolution ──────────────────────────────────────
 * SH/phase_field_evolve.comp.glsl:50-168 (Ambati 2015 hybrid AT2 form), with
 * the engine's parameters (NP/src/solver/dispatcher.cpp:240-248: l0 =
 * max(PHASE_FIELD_L0, 2), G_c = G_C_CONCRETE, relax = PHASE_FIELD_RELAX,
 * spectral split on).  Per solid voxel whose previous failure code is 0, over
 * the in-grid non-air face neighbours (order -x, +x, -y, +y, -z, +z):
 *   lap_phi += phi_j - phi_i and sumSigma += cond[d][i] for finite phi_j,
 *   lap_d   += d_j - d_i for finite d_j;
 * H <= 0 (or NaN) leaves d; else
 *   h = hydration clamped to [0.01, 1] (1 without a field), Gc = gc h sqrt(h)
 *   H_eff = split ? max(H, max(lap_phi, 0) / max(sumSigma, 1e-12)) : H
 *   d' = clamp((2 H_eff + Gc l0 lap_d) / max(2 H_eff + Gc / l0, 1e-8), d_i, 1)
 *   d' = max(clamp(d_i (1 - relax) + d' relax, 0, 1), d_i), d_i if not finite
 *   (the shader's mix can round below d_i by an ulp; damage is irreversible),
 * FP32 in this order.  The shader updates d in place while neighbours read
 * it; here every voxel reads d_in and writes d_out (a Jacobi step), so the
 * result does not depend on thread order.  frac[i] = 1 where d' > 0.95
 * (FAIL_CANTILEVER), else 0. */
void or_phase_field(int lx, int ly, int lz, const uint8_t* type, const float* cond,
                    const float* phi, const float* H, const float* hydration, const uint8_t* code,
                    const float* d_in, float* d_out, uint8_t* frac, float l0, float gc,
                    float relax, int split) {
  static const int DV[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  int64_t n = (int64_t)lx * ly * lz;
  for (int z = 0; z < lz; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int64_t i = IDX(x, y, z, lx, ly);
        float di = d_in[i];
        d_out[i] = di;
        frac[i] = 0;
        if (!is_solid(type[i]) || code[i] != 0) continue;
        float pi = phi[i], ssum = 0.0f, lp = 0.0f, ld = 0.0f;
        for (int d = 0; d < 6; d++) {
          int nx = x + DV[d][0], ny = y + DV[d][1], nz = z + DV[d][2];
          if (nx < 0 || nx >= lx || ny < 0 || ny >= ly || nz < 0 || nz >= lz) continue;
          int64_t j = IDX(nx, ny, nz, lx, ly);
          if (type[j] == T_AIR) continue;
          float pj = phi[j], dj = d_in[j];
          if (isfinite(pj)) {
            lp += pj - pi;
            ssum += cond[d * n + i];
          }
          if (isfinite(dj)) ld += dj - di;
        }
        float hv = H[i];
        if (!(hv > 0.0f)) continue;
        float hy = hydration ? hydration[i] : 1.0f;
        float hd = hy > 0.01f ? (hy < 1.0f ? hy : 1.0f) : 0.01f;
        float gce = gc * (hd * (float)sqrt((double)hd));
        float he = hv;
        if (split) {
          float sd = ssum > 1e-12f ? ssum : 1e-12f;
          float fo = lp > 0.0f ? lp : 0.0f;
          float pp = fo / sd;
          he = pp > hv ? pp : hv;
        }
        float num = 2.0f * he + gce * l0 * ld;
        float den = 2.0f * he + gce / l0;
        den = den > 1e-8f ? den : 1e-8f;
        float q = num / den;
        q = q > di ? q : di;
        q = q < 1.0f ? q : 1.0f;
        float dn = di * (1.0f - relax) + q * relax;
        dn = dn > 0.0f ? dn : 0.0f;
        dn = dn < 1.0f ? dn : 1.0f;
        dn = dn > di ? dn : di; /* the mix rounds; damage never decreases */
        if (!isfinite(dn)) dn = di;
        d_out[i] = dn;
        frac[i] = dn > 0.95f ? 1 : 0;
      }
}

/* Macro-block residual map (SH/failure_scan.comp.glsl:183-195, blocks of
 * MACRO_BLOCK_SIZE = 8): for each 8^3 block the largest |b_i - (A phi)_i|
 * over its solid voxels, the residual in FP64 as or_apply forms it (phi
 * widened from FP32), rounded to FP32; 0 for blocks without a solid voxel.
 * The shader's |flux_in - flux_out + rho| is the same quantity. */
void or_macro_residual(int lx, int ly, int lz, const uint8_t* type, const float* kappa,
                       const float* phi, const float* b, float* blk) {
  int64_t n = (int64_t)lx * ly * lz;
  int bx = (lx + 7) / 8, by = (ly + 7) / 8, bz = (lz + 7) / 8;
  double* x = (double*)malloc(sizeof(double) * (size_t)n);
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; i++) x[i] = (double)phi[i];
  or_apply(lx, ly, lz, type, kappa, x, y);
  memset(blk, 0, sizeof(float) * (size_t)bx * by * bz);
  for (int z = 0; z < lz; z++)
    for (int yy = 0; yy < ly; yy++)
      for (int xx = 0; xx < lx; xx++) {
        int64_t i = IDX(xx, yy, z, lx, ly);
        if (!is_solid(type[i])) continue;
        float r = (float)fabs((double)b[i] - y[i]);
        int64_t k = (xx / 8) + (int64_t)bx * ((yy / 8) + (int64_t)by * (z / 8));
        if (r > blk[k]) blk[k] = r;
      }
  free(x);
  free(y);
}
