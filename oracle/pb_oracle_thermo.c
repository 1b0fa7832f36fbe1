/*
 * pb_oracle_thermo.c -- CPU restatement of the MD thermostat helpers:
 * GaussianThermo::TimeEvolution (Newton-Raphson form), adjustVelocities and
 * constTemp.  TEST INFRASTRUCTURE ONLY (see pb_oracle.h).  FP64 throughout;
 * xyz-interleaved arrays of n particles.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include "pb_oracle.h"

/* GaussianThermo::TimeEvolution, GaussianThermo.cu:29-83, with the kernels
 * calcGaussianThermoA1_F4 (kernelfuncs.cu:372-405),
 * calcGaussianThermoFoverF_F4 (:407-454) and
 * propagateVelocityVerletGaussianThermo_F4 (:456-477).
 *
 * 1) v' = v + (F + Fold) dt/2m ;  A = v' - v xi dt/2 ;  v <- v'
 * 2) repeat: per component f = A - (dt/2 xi + 1) v,
 *            f' = -dt/2 (xi + (F - 2 xi m v) v / sum m v^2) - 1,
 *            v -= f/f' ;  v0 = sqrt(sum (f/f')^2 / 3n) ;
 *            xi = sum F.v / sum m v^2   (with the v just updated)
 *    until v0 <= tol (at most 50 passes, as the reference's loop)
 * 3) r += v (dt - xi dt^2/2) + F dt^2/2m ;  Fold = F
 *
 * Deviation: innerProduct_F4 (kernelfuncs.cu:479-504) assigns each
 * thread's last product instead of adding it, so the reference's sum F.v
 * covers one particle per thread once n exceeds the thread count; the full
 * sum is used here.  Returns the number of Newton passes, or -1 if the
 * loop did not converge (the reference prints and exits).
 * pinned by: the closed-form fixed point v = A (1 - dt/2 F.A / m A.A) and
 * the isokinetic property (tests/test_oracle_thermo.py). */
int pbo_gt_step(double *r, double *v, const double *F, double *Fold, const double *minv,
                uint32_t n, double dt, double *xi_io, double tol) {
  const double dt2 = 0.5 * dt;
  double xi = *xi_io;
  double *A = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  if (!A) return -1;
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      const size_t k = 3 * (size_t)i + c;
      const double vp = v[k] + (F[k] + Fold[k]) * dt2 * minv[i];
      A[k] = vp - v[k] * xi * dt2;
      v[k] = vp;
    }
  double mv2 = 0.0;
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) mv2 += v[3 * (size_t)i + c] * v[3 * (size_t)i + c] / minv[i];
  int it = 0;
  for (;;) {
    if (it++ > 50) {
      free(A);
      return -1;
    }
    const double mv2inv = 1.0 / mv2;
    double s = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      const double m = 1.0 / minv[i];
      for (int c = 0; c < 3; ++c) {
        const size_t k = 3 * (size_t)i + c;
        const double f1 = A[k] - (dt2 * xi + 1.0) * v[k];
        const double f2 = -dt2 * (xi + (F[k] - 2.0 * xi * m * v[k]) * v[k] * mv2inv) - 1.0;
        const double q = f1 / f2;
        v[k] -= q;
        s += q * q;
      }
    }
    const double v0 = sqrt(s / (3.0 * n));
    double fv = 0.0;
    mv2 = 0.0;
    for (uint32_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        const size_t k = 3 * (size_t)i + c;
        fv += F[k] * v[k];
        mv2 += v[k] * v[k] / minv[i];
      }
    xi = fv / mv2;
    if (!(v0 > tol)) break;
  }
  free(A);
  const double dt22 = 0.5 * dt * dt;
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      const size_t k = 3 * (size_t)i + c;
      r[k] += v[k] * (dt - xi * dt22) + F[k] * dt22 * minv[i];
      Fold[k] = F[k];
    }
  *xi_io = xi;
  return it;
}

/* adjustVelocity_F4 / adjustVelocity_VD (kernelfuncs.cu:546-613, 624-699):
 * per axis, mean and variance E[v^2] - E[v]^2 over all n particles, then
 * v = (v - mean) * v1 / sqrt(var).  stats (may be NULL) receives the means
 * and variances before the adjustment.
 * pinned by: zero mean and variance v1^2 afterwards, on data with a known
 * mean and spread (tests/test_oracle_thermo.py). */
void pbo_adjust_velocities(double *v, uint32_t n, double v1, double *stats) {
  double mean[3] = {0, 0, 0}, sq[3] = {0, 0, 0}, sc[3];
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      const double x = v[3 * (size_t)i + c];
      mean[c] += x;
      sq[c] += x * x;
    }
  for (int c = 0; c < 3; ++c) {
    mean[c] /= n;
    const double var = sq[c] / n - mean[c] * mean[c];
    sc[c] = v1 / sqrt(var);
    if (stats) {
      stats[c] = mean[c];
      stats[3 + c] = var;
    }
  }
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      const size_t k = 3 * (size_t)i + c;
      v[k] = (v[k] - mean[c]) * sc[c];
    }
}

/* cudaParticleMD::constTemp, cudaParticleMD.cu:140-155 with
 * correctConstTemp_F4 (kernelfuncs.cu:533-544): lambda = sum F.v / sum m v^2,
 * F -= lambda m v (skipped when lambda == 0).  Returns sum m v^2; *lambda_out
 * receives lambda.
 * pinned by: sum F.v = 0 after the correction, and lambda on a closed-form
 * two-particle case (tests/test_oracle_thermo.py). */
double pbo_const_temp(const double *v, double *F, const double *minv, uint32_t n,
                      double *lambda_out) {
  double fv = 0.0, mv2 = 0.0;
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      const size_t k = 3 * (size_t)i + c;
      fv += F[k] * v[k];
      mv2 += v[k] * v[k] / minv[i];
    }
  const double lam = fv / mv2;
  if (lam != 0.0)
    for (uint32_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        const size_t k = 3 * (size_t)i + c;
        F[k] -= lam * v[k] / minv[i];
      }
  if (lambda_out) *lambda_out = lam;
  return mv2;
}
