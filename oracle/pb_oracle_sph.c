/*
 * pb_oracle_sph.c -- CPU restatement of cudaParticleSPH_NS (reference
 * physics: Lucy kernel, p = c^2 (rho - rho0), physical viscosity).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker; the product never
 * links it.  FP64 throughout; O(N^2) pair search (no cells), so the pair set
 * is fixed by the mathematics (every j with |r_ij| < h, no periodic image)
 * and not by any cell bookkeeping.
 */
#include "pb_oracle.h"
#include <math.h>

/* w0 = 105 / (16 pi h^7): cudaParticleSPHBase::setSPHProperties,
 * cudaParticleSPHBase.cu:119-124 */
double pbo_sph_w0(double h) {
  return 105.0 / (16.0 * M_PI * h * h * h * h * h * h * h);
}

/* SPHKernelLucy, kernelinline.hh:173-189: W = w0 (h + 3r)(h - r)^3, 0 beyond h */
double pbo_sph_lucy_w(double r, double h, double w0) {
  double t = h - r;
  if (0 > t) return 0.0;
  t = t * t * t;
  return w0 * (h + 3.0 * r) * t;
}

/* SPHKernelLucyDW, kernelinline.hh:200-215: (dW/dr)/r = w1 (h - r)^2,
 * w1 = -12 w0 */
double pbo_sph_lucy_dw(double r, double h, double w1) {
  const double t = h - r;
  if (0 > t) return 0.0;
  return w1 * t * t;
}

/* SPHcalcDensity<SPHKernelLucy> (kerneltemplate.hh:636-656) driven over
 * every j including i itself (the reference functor has no i != j test),
 * with the non-periodic distance2 (kernelinline.hh:19-29), followed by
 * calcDensityPost (cudaParticleSPH_NS.cu:184-207):
 *   num = sum W, rho = sum m_j W, mu = (sum mu0_j m_j W) / rho,
 *   rhoinv = 1 / rho.
 * r xyz-interleaved; num_out 3 per particle (num, rho, mu). */
void pbo_sph_density(const double *r, const double *m, const double *mu0, uint32_t n,
                     double h, double *num_out, double *rhoinv) {
  const double w0 = pbo_sph_w0(h);
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
      const double dx = r[3 * i] - r[3 * j], dy = r[3 * i + 1] - r[3 * j + 1],
                   dz = r[3 * i + 2] - r[3 * j + 2];
      const double R = sqrt(dx * dx + dy * dy + dz * dz);
      if (!(R < h)) continue;
      const double W = pbo_sph_lucy_w(R, h, w0);
      s0 += W;
      s1 += m[j] * W;
      s2 += mu0[j] * m[j] * W;
    }
    num_out[3 * i] = s0;
    num_out[3 * i + 1] = s1;
    rhoinv[i] = 1.0 / s1;
    num_out[3 * i + 2] = s2 * rhoinv[i];
  }
}

/* SPHNavierStokes<SPHKernelLucyDW> (kerneltemplate.hh:666-709) over every
 * j != i, with rho0 = 1.0 as cudaParticleSPH_NS::calcAcceleration sets it
 * (cudaParticleSPH_NS.cu:91):
 *   mj  = m_j w1 (h - R)^2
 *   p1  = c2_i rhoinv_i (1 - rho0 rhoinv_i) + c2_j rhoinv_j (1 - rho0 rhoinv_j)
 *   v1  = (mu_i + mu_j) rhoinv_i rhoinv_j
 *   a_i = sum mj ((r_j - r_i) p1 - (v_j - v_i) v1)
 * mu is the field value (num.z after calcDensityPost).  a xyz-interleaved,
 * overwritten (the reference clears a first). */
void pbo_sph_accel(const double *r, const double *v, const double *m, const double *c2,
                   const double *rhoinv, const double *mu, uint32_t n, double h,
                   double rho0, double *a) {
  const double w1 = -12.0 * pbo_sph_w0(h);
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    double ax = 0.0, ay = 0.0, az = 0.0;
    const double pi_ = c2[i] * rhoinv[i] * (1.0 - rho0 * rhoinv[i]);
    for (uint32_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = r[3 * j] - r[3 * i], dy = r[3 * j + 1] - r[3 * i + 1],
                   dz = r[3 * j + 2] - r[3 * i + 2];
      const double R = sqrt(dx * dx + dy * dy + dz * dz);
      if (!(R < h)) continue;
      const double mj = m[j] * pbo_sph_lucy_dw(R, h, w1);
      const double p1 = pi_ + c2[j] * rhoinv[j] * (1.0 - rho0 * rhoinv[j]);
      const double v1 = (mu[i] + mu[j]) * rhoinv[i] * rhoinv[j];
      ax += mj * (dx * p1 - (v[3 * j] - v[3 * i]) * v1);
      ay += mj * (dy * p1 - (v[3 * j + 1] - v[3 * i + 1]) * v1);
      az += mj * (dz * p1 - (v[3 * j + 2] - v[3 * i + 2]) * v1);
    }
    a[3 * i] = ax;
    a[3 * i + 1] = ay;
    a[3 * i + 2] = az;
  }
}

/* propagateLeapFrog_F4 (kernelfuncs.cu:270-296), FP64:
 * v += a dt move ; r += v dt move. */
void pbo_lf_step(double *r, double *v, const double *a, const uint8_t *move, uint32_t n,
                 double dt) {
  for (uint32_t i = 0; i < n; ++i) {
    const double d = dt * (double)move[i];
    for (int c = 0; c < 3; ++c) {
      v[3 * i + c] += a[3 * i + c] * d;
      r[3 * i + c] += v[3 * i + c] * d;
    }
  }
}

/* rollbackLeapFrog_F4 (kernelfuncs.cu:298-327), FP64: with v = v(t+dt/2)
 * and a = a(t), r(t+dt) -> r(t) = r - v dt, v -> v(t-dt/2) = v - a dt, then
 * r(t) -> r(t-dt) = r(t) - v(t-dt/2) dt; fixed particles stay. */
void pbo_lf_rollback(double *r, double *v, const double *a, const uint8_t *move, uint32_t n,
                     double dt) {
  for (uint32_t i = 0; i < n; ++i) {
    const double d = dt * (double)move[i];
    for (int c = 0; c < 3; ++c) {
      const double rt = r[3 * i + c] - v[3 * i + c] * d;
      v[3 * i + c] -= a[3 * i + c] * d;
      r[3 * i + c] = rt - v[3 * i + c] * d;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Config physics of the dam-break workload (weakly compressible SPH).  The
 * reference's kernels hold no such form; this follows the published method:
 *   kernel   M4 cubic spline (Monaghan & Lattanzio 1985), 3-D, support 2h:
 *            W = 1/(pi h^3) * { 1 - 3/2 q^2 + 3/4 q^3      0 <= q < 1
 *                             { 1/4 (2 - q)^3              1 <= q < 2
 *                             { 0                          q >= 2,  q = r/h
 *   density  rho_i = sum_j m_j W_ij, j = i included (summation density)
 *   EOS      Tait (Monaghan 1994): p = B ((rho/rho0)^gamma - 1),
 *            B = rho0 c0^2 / gamma, then max(p, pmin)
 *   momentum a_i = -sum_{j != i} m_j (p_i/rho_i^2 + p_j/rho_j^2 + Pi_ij)
 *                  grad_i W_ij
 *   viscosity Pi_ij = -alpha c0 mu_ij / rho_bar_ij when v_ij . r_ij < 0,
 *            else 0; mu_ij = h v_ij . r_ij / (r_ij^2 + 0.01 h^2),
 *            rho_bar_ij = (rho_i + rho_j)/2 (Monaghan 1992, beta = 0, the
 *            mean sound speed taken as c0).
 * Non-periodic distances, O(N^2) pair search.
 * ------------------------------------------------------------------------- */

/* W(r) of the M4 spline, written piecewise */
double pbo_wc_w(double r, double h) {
  const double q = r / h, s = 1.0 / (M_PI * h * h * h);
  if (q < 1.0) return s * (1.0 - 1.5 * q * q + 0.75 * q * q * q);
  if (q < 2.0) {
    const double t = 2.0 - q;
    return s * 0.25 * t * t * t;
  }
  return 0.0;
}

/* (dW/dr)/r of the M4 spline: 1/(pi h^5) * (-3 + 9/4 q) for q < 1,
 * 1/(pi h^5) * (-3/4 (2 - q)^2 / q) for 1 <= q < 2, 0 beyond */
double pbo_wc_dw(double r, double h) {
  const double q = r / h, s = 1.0 / (M_PI * h * h * h * h * h);
  if (q < 1.0) return s * (-3.0 + 2.25 * q);
  if (q < 2.0) {
    const double t = 2.0 - q;
    return s * (-0.75 * t * t / q);
  }
  return 0.0;
}

/* summation density and Tait pressure; num_out 3 per particle
 * (sum W, rho, p), rhoinv = 1/rho */
void pbo_wc_density(const double *r, const double *m, uint32_t n, double h, double rho0,
                    double c0, double gamma, double pmin, double *num_out, double *rhoinv) {
  const double B = rho0 * c0 * c0 / gamma, sup = 2.0 * h;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
      const double dx = r[3 * i] - r[3 * j], dy = r[3 * i + 1] - r[3 * j + 1],
                   dz = r[3 * i + 2] - r[3 * j + 2];
      const double R = sqrt(dx * dx + dy * dy + dz * dz);
      if (!(R < sup)) continue;
      const double W = pbo_wc_w(R, h);
      s0 += W;
      s1 += m[j] * W;
    }
    double p = B * (pow(s1 / rho0, gamma) - 1.0);
    if (p < pmin) p = pmin;
    num_out[3 * i] = s0;
    num_out[3 * i + 1] = s1;
    num_out[3 * i + 2] = p;
    rhoinv[i] = 1.0 / s1;
  }
}

/* pressure + artificial-viscosity acceleration (gravity excluded);
 * rho and p from pbo_wc_density; a overwritten */
void pbo_wc_accel(const double *r, const double *v, const double *m, const double *rho,
                  const double *p, uint32_t n, double h, double c0, double alpha, double *a) {
  const double sup = 2.0 * h, eh2 = 0.01 * h * h;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    double ax = 0.0, ay = 0.0, az = 0.0;
    const double Pi = p[i] / (rho[i] * rho[i]);
    for (uint32_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = r[3 * i] - r[3 * j], dy = r[3 * i + 1] - r[3 * j + 1],
                   dz = r[3 * i + 2] - r[3 * j + 2];
      const double R2 = dx * dx + dy * dy + dz * dz, R = sqrt(R2);
      if (!(R < sup)) continue;
      const double vr = (v[3 * i] - v[3 * j]) * dx + (v[3 * i + 1] - v[3 * j + 1]) * dy +
                        (v[3 * i + 2] - v[3 * j + 2]) * dz;
      double visc = 0.0;
      if (vr < 0.0) {
        const double mu = h * vr / (R2 + eh2);
        visc = -alpha * c0 * mu / (0.5 * (rho[i] + rho[j]));
      }
      const double f = -m[j] * (Pi + p[j] / (rho[j] * rho[j]) + visc) * pbo_wc_dw(R, h);
      ax += f * dx;
      ay += f * dy;
      az += f * dz;
    }
    a[3 * i] = ax;
    a[3 * i + 1] = ay;
    a[3 * i + 2] = az;
  }
}
