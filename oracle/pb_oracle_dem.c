/*
 * pb_oracle_dem.c -- CPU restatement of cudaParticleDEM (+ Rotation).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker; the product never
 * links it.  FP64 throughout; the contact search is O(N^2) brute force, so
 * the contact set is fixed by the geometry alone.
 *
 * Contact semantics (documented deviation): the reference counts contacts
 * with |r_ij| < r_i + r_j (calcCoord_F4, kernelinline.hh:371-382) but lists
 * them with <= (makeJlist_WithBlock*_F4, kernelfuncs.cu:1158, 1206), which
 * overruns a row on an exact touch; the force functor only acts for <.
 * Here both use the strict <, and rows are sorted by j (sortColIdx,
 * kernelfuncs.cu:1262-1290, exits early on short rows and can leave later
 * rows unsorted; sorted rows are the canonical form).
 */
#include "pb_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Contact CSR: rows for movable i only (fixed particles get empty rows, as
 * calcF_IJpairWithBlock2/4_F4 skip move[i] == 0, kerneltemplate.hh:331,
 * 388), j != i with |r_ij|^2 < (rad_i + rad_j)^2, j ascending.
 * rowptr has n+1 entries; colidx may be NULL to count only.  Returns nnz. */
uint32_t pbo_dem_contacts(const double *r, const double *rad, const uint8_t *move, uint32_t n,
                          uint32_t *rowptr, uint32_t *colidx) {
  uint32_t nnz = 0;
  rowptr[0] = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (move[i]) {
      for (uint32_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const double dx = r[3 * i] - r[3 * j], dy = r[3 * i + 1] - r[3 * j + 1],
                     dz = r[3 * i + 2] - r[3 * j + 2];
        const double R2 = dx * dx + dy * dy + dz * dz;
        const double rc = rad[i] + rad[j];
        if (R2 < rc * rc) {
          if (colidx) colidx[nnz] = j;
          ++nnz;
        }
      }
    }
    rowptr[i + 1] = nnz;
  }
  return nnz;
}

/* succeedPrevState, kernelfuncs.cu:1292-1309: a contact (i, j) present in
 * both the previous and the current CSR keeps its tangential history;
 * new contacts start from 0 (the reference clears val first, :161). */
void pbo_dem_carry(const uint32_t *prow, const uint32_t *pcol, const double *pval,
                   const uint32_t *crow, const uint32_t *ccol, double *cval, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t q = crow[i]; q < crow[i + 1]; ++q) cval[q] = 0.0;
    for (uint32_t p = prow[i]; p < prow[i + 1]; ++p)
      for (uint32_t q = crow[i]; q < crow[i + 1]; ++q)
        if (ccol[q] == pcol[p]) cval[q] = pval[p];
  }
}

/* DEMContact_F4 (kernelinline.hh:235-363) over the CSR rows
 * (calcF_IJpairWithList_F4, kerneltemplate.hh:135-168), FP64:
 *   n    = (r_i - r_j)/|r_ij|, Rij = -rad_i/(rad_i+rad_j) (r_i - r_j)
 *   xi   = rad_i + rad_j - |r_ij|, C3 = v_ij . n
 *   Kn   = 2/3 E sqrt(Rbar xi), Fn = n (Kn xi - gamma2 sqrt(mbar Kn) C3)
 *   vt   = v_ij - C3 n + (w_i + w_j) x Rij
 *   if all three components of vt are non-zero:
 *     Fs = -mu F0 min(xi_s/xi_s_max, 1) vt/|vt|, xi_s_max = mu s2 xi,
 *     xi_s += |vt| dt (history val of the contact, updated in place)
 *   Ts   = Rij x Fs
 *   if all three components of w_i are non-zero:
 *     Tr = -mu_r rad_i F0 w_i/|w_i|
 *   F_i += (Fn + Fs) Q, T_i += (Ts + Tr) Q, Q = 2 - move_j.
 * E is the effective modulus as setDEMProperties stores it
 * (E = Y (1 - sigma^2), cudaParticleDEM.cu:225), s2 = (2-sigma)/2/(1-sigma),
 * gamma2 as passed.  F and T (xyz-interleaved) are overwritten; rows of
 * fixed particles are empty so their F and T are zero.
 * Returns the number of contacts violating the Cundall-Strack bound
 * dt < 2 sqrt(mbar/Kn) (the reference asserts it unless autotuning). */
uint32_t pbo_dem_force(const double *r, const double *v, const double *w, const double *rad,
                       const double *minv, const uint8_t *move, uint32_t n,
                       const uint32_t *rowptr, const uint32_t *colidx, double *val, double E,
                       double mu, double s2, double gamma2, double mu_r, double dt, double *F,
                       double *T) {
  uint32_t bad = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    for (uint32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const uint32_t j = colidx[q];
      const double d[3] = {r[3 * i] - r[3 * j], r[3 * i + 1] - r[3 * j + 1],
                           r[3 * i + 2] - r[3 * j + 2]};
      const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double rs = rad[i] + rad[j];
      if (!(R2 < rs * rs) || j == i) continue;
      const double R = sqrt(R2);
      const double nv[3] = {d[0] / R, d[1] / R, d[2] / R};
      const double c = -rad[i] / rs;
      const double Rij[3] = {c * d[0], c * d[1], c * d[2]};
      const double vij[3] = {v[3 * i] - v[3 * j], v[3 * i + 1] - v[3 * j + 1],
                             v[3 * i + 2] - v[3 * j + 2]};
      const double xi = rs - R;
      const double C3 = vij[0] * nv[0] + vij[1] * nv[1] + vij[2] * nv[2];
      const double Rbar = rad[i] * rad[j] / rs;
      const double mbar = 1.0 / (minv[i] + minv[j]);
      const double Kn = 2.0 / 3.0 * E * sqrt(Rbar * xi);
      const double C4 = Kn * xi - gamma2 * sqrt(mbar * Kn) * C3;
      const double Fn[3] = {nv[0] * C4, nv[1] * C4, nv[2] * C4};
      if (!(dt < 2.0 * sqrt(mbar / Kn))) ++bad;
      const double F0 = sqrt(Fn[0] * Fn[0] + Fn[1] * Fn[1] + Fn[2] * Fn[2]);
      const double ws[3] = {w[3 * i] + w[3 * j], w[3 * i + 1] + w[3 * j + 1],
                            w[3 * i + 2] + w[3 * j + 2]};
      double vt[3] = {vij[0] - C3 * nv[0], vij[1] - C3 * nv[1], vij[2] - C3 * nv[2]};
      vt[0] += ws[1] * Rij[2] - ws[2] * Rij[1];
      vt[1] += ws[2] * Rij[0] - ws[0] * Rij[2];
      vt[2] += ws[0] * Rij[1] - ws[1] * Rij[0];
      double Fs[3] = {0, 0, 0};
      if (vt[0] != 0 && vt[1] != 0 && vt[2] != 0) {
        const double xsm = mu * s2 * xi;
        const double vn = sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
        const double xs = val[q];
        const double v1 = -mu * F0 * ((xs < xsm) ? (xs / xsm) : 1.0) / vn;
        Fs[0] = v1 * vt[0];
        Fs[1] = v1 * vt[1];
        Fs[2] = v1 * vt[2];
        val[q] += vn * dt;
      }
      const double Ts[3] = {Rij[1] * Fs[2] - Rij[2] * Fs[1], Rij[2] * Fs[0] - Rij[0] * Fs[2],
                            Rij[0] * Fs[1] - Rij[1] * Fs[0]};
      double Tr[3] = {0, 0, 0};
      const double *wi = &w[3 * i];
      if (wi[0] != 0 && wi[1] != 0 && wi[2] != 0) {
        const double wn = sqrt(wi[0] * wi[0] + wi[1] * wi[1] + wi[2] * wi[2]);
        for (int k = 0; k < 3; ++k) Tr[k] = -mu_r * rad[i] * F0 * wi[k] / wn;
      }
      const double Q = 2.0 - (double)move[j];
      for (int k = 0; k < 3; ++k) {
        f[k] += (Fn[k] + Fs[k]) * Q;
        t[k] += (Ts[k] + Tr[k]) * Q;
      }
    }
    for (int k = 0; k < 3; ++k) {
      F[3 * i + k] = f[k];
      T[3 * i + k] = t[k];
    }
  }
  return bad;
}

/* propagateEulerianRotation_F4, kernelfuncs.cu:966-986, FP64:
 * L += T dt move ; w = L m0inv move, with m0inv = minv / (0.4 rad^2)
 * (cudaParticleRotation::setInertia, cudaParticleRotation.cu:35-61). */
void pbo_rot_step(double *w, double *L, const double *T, const double *m0inv,
                  const uint8_t *move, uint32_t n, double dt) {
  for (uint32_t i = 0; i < n; ++i) {
    const double d = dt * (double)move[i], m0 = m0inv[i] * (double)move[i];
    for (int k = 0; k < 3; ++k) {
      L[3 * i + k] += T[3 * i + k] * d;
      w[3 * i + k] = L[3 * i + k] * m0;
    }
  }
}

/* Linear spring-dashpot contact with a tangential spring and Coulomb
 * friction (the granular settle workload's law; the reference has only the
 * Hertz law above, so this restates the published method: Cundall & Strack,
 * Geotechnique 29 (1979) 47-65, in the "Hookean with shear history" form of
 * granular MD codes).  Rows, contacts and the movable/fixed rules are those
 * of pbo_dem_force; FP64 throughout.  Per contact (i, j):
 *   d = r_i - r_j, R = |d|, n = d/R, delta = rad_i + rad_j - R > 0
 *   1/meff = minv_i + (move_j ? minv_j : 0)   (a fixed sphere is a wall)
 *   gamma_n = 2 zeta sqrt(meff kn),  zeta = -ln e / sqrt(pi^2 + ln^2 e)
 *   v_ij = v_i - v_j, vn = v_ij . n
 *   Fn = (kn delta - gamma_n vn) n          (not clamped at zero)
 *   W = rad_i w_i + rad_j w_j, v_c = v_ij - W x n, vt = v_c - vn n
 *   s  = hist - (hist . n) n + vt dt        (spring kept in the tangent plane)
 *   Ft = -kt s; if |Ft| > mu |Fn|: Ft *= mu |Fn| / |Ft|, s = -Ft / kt
 *   hist = s (3 doubles per contact, updated in place)
 *   F_i += Fn + Ft, T_i += (-rad_i n) x Ft
 * zeta is passed (the caller derives it from the restitution coefficient).
 * Returns the contacts with dt >= 2 sqrt(meff / kn) (the stability bound of
 * the explicit update, the counterpart of the Hertz law's check). */
uint32_t pbo_dem_force_lsd(const double *r, const double *v, const double *w, const double *rad,
                           const double *minv, const uint8_t *move, uint32_t n,
                           const uint32_t *rowptr, const uint32_t *colidx, double *hist,
                           double kn, double zeta, double kt, double mu, double dt, double *F,
                           double *T) {
  uint32_t bad = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    for (uint32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const uint32_t j = colidx[q];
      const double d[3] = {r[3 * i] - r[3 * j], r[3 * i + 1] - r[3 * j + 1],
                           r[3 * i + 2] - r[3 * j + 2]};
      const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double rs = rad[i] + rad[j];
      if (!(R2 < rs * rs) || j == i) continue;
      const double R = sqrt(R2);
      const double nv[3] = {d[0] / R, d[1] / R, d[2] / R};
      const double delta = rs - R;
      const double meff = 1.0 / (minv[i] + (move[j] ? minv[j] : 0.0));
      const double gn = 2.0 * zeta * sqrt(meff * kn);
      if (!(dt < 2.0 * sqrt(meff / kn))) ++bad;
      const double vij[3] = {v[3 * i] - v[3 * j], v[3 * i + 1] - v[3 * j + 1],
                             v[3 * i + 2] - v[3 * j + 2]};
      const double vn = vij[0] * nv[0] + vij[1] * nv[1] + vij[2] * nv[2];
      const double fn = kn * delta - gn * vn;
      const double W[3] = {rad[i] * w[3 * i] + rad[j] * w[3 * j],
                           rad[i] * w[3 * i + 1] + rad[j] * w[3 * j + 1],
                           rad[i] * w[3 * i + 2] + rad[j] * w[3 * j + 2]};
      const double Wxn[3] = {W[1] * nv[2] - W[2] * nv[1], W[2] * nv[0] - W[0] * nv[2],
                             W[0] * nv[1] - W[1] * nv[0]};
      double vt[3];
      for (int k = 0; k < 3; ++k) vt[k] = vij[k] - Wxn[k] - vn * nv[k];
      double *h = &hist[3 * (size_t)q];
      const double hn = h[0] * nv[0] + h[1] * nv[1] + h[2] * nv[2];
      double s[3], ft[3];
      for (int k = 0; k < 3; ++k) {
        s[k] = h[k] - hn * nv[k] + vt[k] * dt;
        ft[k] = -kt * s[k];
      }
      const double ftm = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
      const double cap = mu * fabs(fn);
      if (ftm > cap) {
        const double c = cap / ftm;
        for (int k = 0; k < 3; ++k) {
          ft[k] *= c;
          s[k] = -ft[k] / kt;
        }
      }
      for (int k = 0; k < 3; ++k) h[k] = s[k];
      const double nxf[3] = {nv[1] * ft[2] - nv[2] * ft[1], nv[2] * ft[0] - nv[0] * ft[2],
                             nv[0] * ft[1] - nv[1] * ft[0]};
      for (int k = 0; k < 3; ++k) {
        f[k] += fn * nv[k] + ft[k];
        t[k] += -rad[i] * nxf[k];
      }
    }
    for (int k = 0; k < 3; ++k) {
      F[3 * i + k] = f[k];
      T[3 * i + k] = t[k];
    }
  }
  return bad;
}

/* pbo_dem_carry for histories of width hw doubles per contact */
void pbo_dem_carry_w(const uint32_t *prow, const uint32_t *pcol, const double *pval,
                     const uint32_t *crow, const uint32_t *ccol, double *cval, uint32_t n,
                     uint32_t hw) {
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t q = crow[i]; q < crow[i + 1]; ++q)
      for (uint32_t k = 0; k < hw; ++k) cval[(size_t)q * hw + k] = 0.0;
    for (uint32_t p = prow[i]; p < prow[i + 1]; ++p)
      for (uint32_t q = crow[i]; q < crow[i + 1]; ++q)
        if (ccol[q] == pcol[p])
          for (uint32_t k = 0; k < hw; ++k) cval[(size_t)q * hw + k] = pval[(size_t)p * hw + k];
  }
}

static int cmp_u32(const void *a, const void *b) {
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

/* pbo_dem_contacts over a cell grid (cells of edge >= cellsize, at least
 * twice the largest radius, covering cell[6]; a centre outside the box is
 * clamped into the edge cell), for large N: the same contact test and the
 * same CSR (rows of movable i, columns ascending).  colidx may be NULL to
 * count only; returns nnz, or UINT32_MAX on a bad grid. */
uint32_t pbo_dem_contacts_cells(const double *r, const double *rad, const uint8_t *move,
                                uint32_t n, const double cell[6], double cellsize,
                                uint32_t *rowptr, uint32_t *colidx) {
  uint32_t nb[3];
  for (int a = 0; a < 3; ++a) {
    const double L = cell[2 * a + 1] - cell[2 * a];
    if (!(L > 0) || !(cellsize > 0)) return UINT32_MAX;
    nb[a] = (uint32_t)floor(L / cellsize);
    if (nb[a] < 1) nb[a] = 1;
  }
  const size_t nc = (size_t)nb[0] * nb[1] * nb[2];
  uint32_t *cid = malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *start = calloc(nc + 1, sizeof(uint32_t));
  uint32_t *list = malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *fillp = malloc(sizeof(uint32_t) * (nc + 1));
  uint32_t *row = malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) {
      const double L = cell[2 * a + 1] - cell[2 * a];
      double x = (r[3 * i + a] - cell[2 * a]) / L * nb[a];
      long k = (long)floor(x);
      if (k < 0) k = 0;
      if (k >= (long)nb[a]) k = nb[a] - 1;
      c[a] = (uint32_t)k;
    }
    cid[i] = (c[0] * nb[1] + c[1]) * nb[2] + c[2];
    ++start[cid[i] + 1];
  }
  for (size_t c = 0; c < nc; ++c) start[c + 1] += start[c];
  memcpy(fillp, start, sizeof(uint32_t) * (nc + 1));
  for (uint32_t i = 0; i < n; ++i) list[fillp[cid[i]]++] = i;
  uint32_t nnz = 0;
  rowptr[0] = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t cnt = 0;
    if (move[i]) {
      const uint32_t c = cid[i];
      const int c0 = (int)(c / (nb[1] * nb[2])), c1 = (int)((c / nb[2]) % nb[1]),
                c2 = (int)(c % nb[2]);
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          for (int d = -1; d <= 1; ++d) {
            const int x = c0 + a, y = c1 + b, z = c2 + d;
            if (x < 0 || y < 0 || z < 0 || x >= (int)nb[0] || y >= (int)nb[1] || z >= (int)nb[2])
              continue;
            const uint32_t J = ((uint32_t)x * nb[1] + (uint32_t)y) * nb[2] + (uint32_t)z;
            for (uint32_t q = start[J]; q < start[J + 1]; ++q) {
              const uint32_t j = list[q];
              if (j == i) continue;
              const double dx = r[3 * i] - r[3 * j], dy = r[3 * i + 1] - r[3 * j + 1],
                           dz = r[3 * i + 2] - r[3 * j + 2];
              const double R2 = dx * dx + dy * dy + dz * dz;
              const double rc = rad[i] + rad[j];
              if (R2 < rc * rc) row[cnt++] = j;
            }
          }
      qsort(row, cnt, sizeof(uint32_t), cmp_u32);
      if (colidx) memcpy(colidx + nnz, row, sizeof(uint32_t) * cnt);
    }
    nnz += cnt;
    rowptr[i + 1] = nnz;
  }
  free(cid);
  free(start);
  free(list);
  free(fillp);
  free(row);
  return nnz;
}
