/*
 * pb_oracle.c -- CPU restatement of the cudaParticleBase LJ / binning /
 * neighbour-count / g(r) paths.  TEST INFRASTRUCTURE ONLY (see
 * pb_oracle.h): the product never links or loads it.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  The
 * f32-exact functions rely on IEEE single precision with no contraction and
 * correctly rounded division and sqrt (SSE on x86-64).
 */
#include "pb_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#pragma STDC FP_CONTRACT OFF

/* CUDA's float->uint conversion saturates (negative/NaN -> 0, >= 2^32 ->
 * UINT_MAX); calcBID_F4 relies on that for particles below the box. */
static inline uint32_t f2u_sat(float q) {
  if (!(q > 0.0f)) return 0u;
  if (q >= 4294967296.0f) return UINT32_MAX;
  return (uint32_t)q;
}

int pbo_setup_cells(const float cell[6], float rmax, int periodic,
                    uint32_t blocknum[3], float blocklen[6]) {
  /* cudaCutoffBlock.cu:97-113 */
  int ok = 0;
  for (int i = 0; i < 3; ++i) {
    const float c = cell[i * 2 + 1] - cell[i * 2];
    if (periodic) {
      blocknum[i] = f2u_sat(c / rmax);
      blocklen[i] = c / (float)blocknum[i];
    } else {
      blocklen[i] = rmax;
      blocknum[i] = f2u_sat(c / rmax) + 1u;
    }
    if (blocknum[i] <= 2) ok = -1;   /* assert(blocknum[i]>2), :109 */
    blocklen[i + 3] = cell[i * 2];
  }
  return ok;
}

void pbo_neighbor_table(const uint32_t blocknum[3], int w, int periodic,
                        uint32_t *table) {
  /* calc_BlockNeighbor, cudaCutoffBlock.cu:17-65 */
  const long bx = blocknum[0], by = blocknum[1], bz = blocknum[2];
  const int J27 = (2 * w + 1) * (2 * w + 1) * (2 * w + 1);
  for (long i = 0; i < bx; ++i)
    for (long j = 0; j < by; ++j)
      for (long k = 0; k < bz; ++k) {
        const long I = k + j * bz + i * bz * by;
        int L = 0;
        for (int i1 = -w; i1 < w + 1; ++i1)
          for (int j1 = -w; j1 < w + 1; ++j1)
            for (int k1 = -w; k1 < w + 1; ++k1) {
              long i2 = i + i1, j2 = j + j1, k2 = k + k1;
              if (i2 == bx) i2 = 0;
              if (j2 == by) j2 = 0;
              if (k2 == bz) k2 = 0;
              if (i2 == -1) i2 = bx - 1;
              if (j2 == -1) j2 = by - 1;
              if (k2 == -1) k2 = bz - 1;
              if (i2 == bx + 1) i2 = 1;
              if (j2 == by + 1) j2 = 1;
              if (k2 == bz + 1) k2 = 1;
              if (i2 == -2) i2 = bx - 2;
              if (j2 == -2) j2 = by - 2;
              if (k2 == -2) k2 = bz - 2;
              uint32_t J = UINT32_MAX;
              if (periodic || ((i + i1) >= 0 && (i + i1) < bx && (j + j1) >= 0 &&
                               (j + j1) < by && (k + k1) >= 0 && (k + k1) < bz))
                J = (uint32_t)(k2 + j2 * bz + i2 * bz * by);
              table[I * J27 + L] = J;
              ++L;
            }
      }
}

static inline uint32_t clamp_axis(uint32_t d, uint32_t c) {
  /* kernelfuncs.cu:821-835 */
  if (d == c) return c - 1;
  if (d > c) return 0;
  return d;
}

void pbo_calc_bid(const float *r, uint32_t n, const float blocklen[6],
                  const uint32_t blocknum[3], uint32_t *bid) {
  /* calcBID_F4, kernelfuncs.cu:807-841 (pid = identity) */
  const uint32_t CX = blocknum[2];
  const uint32_t CY = blocknum[1] * blocknum[2];
  for (uint32_t i = 0; i < n; ++i) {
    const float qx = (r[3 * i + 0] - blocklen[3]) / blocklen[0];
    const float qy = (r[3 * i + 1] - blocklen[4]) / blocklen[1];
    const float qz = (r[3 * i + 2] - blocklen[5]) / blocklen[2];
    const uint32_t d0 = clamp_axis(f2u_sat(qx), blocknum[0]);
    const uint32_t d1 = clamp_axis(f2u_sat(qy), blocknum[1]);
    const uint32_t d2 = clamp_axis(f2u_sat(qz), blocknum[2]);
    bid[i] = d2 + d1 * CX + d0 * CY;
  }
}

void pbo_sort_bins(const uint32_t *bid, uint32_t n, uint32_t ncell,
                   uint32_t *pid, uint32_t *bid_sorted, uint32_t *bindex) {
  uint32_t *cnt = (uint32_t *)calloc((size_t)ncell + 1, sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) cnt[bid[i]]++;
  uint32_t acc = 0;
  for (uint32_t c = 0; c < ncell; ++c) {
    const uint32_t k = cnt[c];
    bindex[c] = k ? acc : UINT32_MAX;
    cnt[c] = acc;
    acc += k;
  }
  bindex[ncell] = n;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t s = cnt[bid[i]]++;
    pid[s] = i;
    bid_sorted[s] = bid[i];
  }
  free(cnt);
}

/* ------------------------------------------------------------------ LJ */

void pbo_lj_make_params(const double *sigma_eps, int ntypes, double rc2,
                        int shifted, pbo_lj_params *p) {
  /* _setLJparams, kernelinline.hh:120-134 (here in FP64, plus the shift) */
  memset(p, 0, sizeof(*p));
  p->ntypes = ntypes;
  for (int i = 0; i < ntypes; ++i)
    for (int j = 0; j < ntypes; ++j) {
      const double s = sigma_eps[(i * ntypes + j) * 2];
      const double e = sigma_eps[(i * ntypes + j) * 2 + 1];
      const int t = i * PBO_MAXTYPES + j;
      p->s2[t] = s * s;
      p->e4[t] = 4.0 * e;
      if (shifted) {
        double sr = p->s2[t] / rc2;
        sr = sr * sr * sr;
        p->shift[t] = p->e4[t] * (sr * sr - sr);
      }
    }
}

/* One LJ pair: returns f/r^2 (force = fr * dx) and adds phi and f.r. */
static inline double lj_pair(double R2, int t, const pbo_lj_params *p,
                             double *phi) {
  double s6 = p->s2[t] / R2;   /* kernelinline.hh:90-95 */
  s6 = s6 * s6 * s6;
  *phi = p->e4[t] * s6 * (s6 - 1.0) - p->shift[t];
  return 6.0 * p->e4[t] * s6 * (2.0 * s6 - 1.0) / R2;
}

static inline double min_image_ref(double d, double cxh2, double L) {
  /* kerneltemplate.hh:541-543: one +-L correction when d^2 > (L/2)^2 */
  if (d * d > cxh2) d += (d < 0) ? L : -L;
  return d;
}

int pbo_lj_forces_cells(const double *r, const uint16_t *type, uint32_t n,
                        const double cell[6], const uint32_t blocknum[3],
                        const double blocklen[6], double rc2,
                        const pbo_lj_params *p, double *F, double *pe,
                        double *virial) {
  for (int a = 0; a < 3; ++a)
    if (blocknum[a] < 3 || blocklen[a] * blocklen[a] < rc2) return -1;
  const uint32_t ncell = blocknum[0] * blocknum[1] * blocknum[2];
  uint32_t *bid = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *pid = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *bs = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *bindex = (uint32_t *)malloc(sizeof(uint32_t) * (ncell + 1));
  uint32_t *table = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)ncell * 27);
  const double L[3] = {cell[1] - cell[0], cell[3] - cell[2], cell[5] - cell[4]};
  const double h2[3] = {0.25 * L[0] * L[0], 0.25 * L[1] * L[1], 0.25 * L[2] * L[2]};
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t d[3];
    for (int a = 0; a < 3; ++a) {
      double q = (r[3 * i + a] - blocklen[3 + a]) / blocklen[a];
      long k = (q > 0) ? (long)q : 0;
      if (k >= (long)blocknum[a]) k = blocknum[a] - 1;
      d[a] = (uint32_t)k;
    }
    bid[i] = d[2] + d[1] * blocknum[2] + d[0] * blocknum[1] * blocknum[2];
  }
  pbo_sort_bins(bid, n, ncell, pid, bs, bindex);
  /* cell ends: next non-empty start (kerneltemplate.hh:265-270) */
  uint32_t *bend = (uint32_t *)malloc(sizeof(uint32_t) * (ncell + 1));
  {
    uint32_t next = n;
    for (long c = (long)ncell - 1; c >= 0; --c) {
      bend[c] = next;
      if (bindex[c] != UINT32_MAX) next = bindex[c];
    }
  }
  pbo_neighbor_table(blocknum, 1, 1, table);
  double pe_s = 0.0, w_s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : pe_s, w_s)
  for (long I = 0; I < (long)ncell; ++I) {
    if (bindex[I] == UINT32_MAX) continue;
    for (uint32_t i0 = bindex[I]; i0 < bend[I]; ++i0) {
      const uint32_t i = pid[i0];
      double fx = 0, fy = 0, fz = 0, ph = 0, w = 0;
      for (int J27 = 0; J27 < 27; ++J27) {
        const uint32_t J = table[I * 27 + J27];
        if (bindex[J] == UINT32_MAX) continue;
        for (uint32_t j0 = bindex[J]; j0 < bend[J]; ++j0) {
          const uint32_t j = pid[j0];
          if (j == i) continue;
          const double dx = min_image_ref(r[3 * i] - r[3 * j], h2[0], L[0]);
          const double dy = min_image_ref(r[3 * i + 1] - r[3 * j + 1], h2[1], L[1]);
          const double dz = min_image_ref(r[3 * i + 2] - r[3 * j + 2], h2[2], L[2]);
          const double R2 = dx * dx + dy * dy + dz * dz;
          if (R2 < rc2) {
            double phi;
            const double fr = lj_pair(R2, type[i] * PBO_MAXTYPES + type[j], p, &phi);
            fx += fr * dx;
            fy += fr * dy;
            fz += fr * dz;
            ph += phi;
            w += fr * R2;
          }
        }
      }
      F[3 * i] = fx;
      F[3 * i + 1] = fy;
      F[3 * i + 2] = fz;
      pe_s += ph;
      w_s += w;
    }
  }
  if (pe) *pe = 0.5 * pe_s;
  if (virial) *virial = 0.5 * w_s;
  free(bid); free(pid); free(bs); free(bindex); free(table); free(bend);
  return 0;
}

void pbo_lj_forces_brute(const double *r, const uint16_t *type, uint32_t n,
                         const double cell[6], double rc2,
                         const pbo_lj_params *p, double *F, double *pe,
                         double *virial) {
  const double L[3] = {cell[1] - cell[0], cell[3] - cell[2], cell[5] - cell[4]};
  double pe_s = 0.0, w_s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : pe_s, w_s)
  for (long i = 0; i < (long)n; ++i) {
    double f[3] = {0, 0, 0}, ph = 0, w = 0;
    for (uint32_t j = 0; j < n; ++j) {
      if ((long)j == i) continue;
      double d[3];
      for (int a = 0; a < 3; ++a) {
        d[a] = r[3 * i + a] - r[3 * j + a];
        d[a] -= L[a] * nearbyint(d[a] / L[a]);
      }
      const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (R2 < rc2) {
        double phi;
        const double fr = lj_pair(R2, type[i] * PBO_MAXTYPES + type[j], p, &phi);
        for (int a = 0; a < 3; ++a) f[a] += fr * d[a];
        ph += phi;
        w += fr * R2;
      }
    }
    for (int a = 0; a < 3; ++a) F[3 * i + a] = f[a];
    pe_s += ph;
    w_s += w;
  }
  if (pe) *pe = 0.5 * pe_s;
  if (virial) *virial = 0.5 * w_s;
}

void pbo_vv_step(double *r, double *v, const double *F, double *Fold,
                 const double *minv, uint32_t n, double dt) {
  /* kernelfuncs.cu:347-365 */
  const double dt2 = dt * 0.5;
  for (uint32_t i = 0; i < n; ++i) {
    const double dtm2 = dt2 * minv[i];
    const double dt22 = dtm2 * dt;
    for (int a = 0; a < 3; ++a) {
      const double f = F[3 * i + a];
      v[3 * i + a] += (f + Fold[3 * i + a]) * dtm2;
      r[3 * i + a] += v[3 * i + a] * dt + f * dt22;
      Fold[3 * i + a] = f;
    }
  }
}

static inline double wrap1(double x, double c0, double c1) {
  /* kernelfuncs.cu:95-129 */
  const double L = c1 - c0;
  const double d0 = c0 - x;
  if (d0 > 0) x += (double)((int)(d0 / L) + 1) * L;
  const double d1 = x - c1;
  if (d1 >= 0) x -= (double)((int)(d1 / L) + 1) * L;
  if (x >= c1) x = c0;    /* rounding onto the upper face: see header */
  return x;
}

void pbo_wrap(double *r, uint32_t n, const double cell[6]) {
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a)
      r[3 * i + a] = wrap1(r[3 * i + a], cell[2 * a], cell[2 * a + 1]);
}

double pbo_mv2(const double *v, const double *minv, uint32_t n) {
  double s = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    const double v2 = v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] +
                      v[3 * i + 2] * v[3 * i + 2];
    s += v2 / minv[i];
  }
  return s;
}

int pbo_lj_run(double *r, double *v, double *Fold, const double *minv,
               const uint16_t *type, uint32_t n, const double cell[6],
               double rmax_cells, double rc2, const pbo_lj_params *p,
               double dt, uint32_t nsteps, double *energies, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  uint32_t bn[3];
  double bl[6];
  for (int a = 0; a < 3; ++a) {
    const double c = cell[2 * a + 1] - cell[2 * a];
    bn[a] = (uint32_t)(c / rmax_cells);
    if (bn[a] < 3) return -1;
    bl[a] = c / bn[a];
    bl[a + 3] = cell[2 * a];
  }
  double *F = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
  for (uint32_t s = 0; s < nsteps; ++s) {
    double pe = 0, w = 0;
    if (pbo_lj_forces_cells(r, type, n, cell, bn, bl, rc2, p, F, &pe, &w)) {
      free(F);
      return -1;
    }
    /* VV split into its velocity and position halves so that KE(v(t)) is
     * taken at the same t as PE(r(t)); arithmetic as pbo_vv_step. */
    double ke = 0.0;
    const double dt2 = dt * 0.5;
#pragma omp parallel for schedule(static) reduction(+ : ke)
    for (long i = 0; i < (long)n; ++i) {
      const double dtm2 = dt2 * minv[i];
      const double dt22 = dtm2 * dt;
      double v2 = 0;
      for (int a = 0; a < 3; ++a) {
        const double f = F[3 * i + a];
        v[3 * i + a] += (f + Fold[3 * i + a]) * dtm2;
        v2 += v[3 * i + a] * v[3 * i + a];
        r[3 * i + a] += v[3 * i + a] * dt + f * dt22;
        Fold[3 * i + a] = f;
        r[3 * i + a] = wrap1(r[3 * i + a], cell[2 * a], cell[2 * a + 1]);
      }
      ke += 0.5 * v2 / minv[i];
    }
    if (energies) {
      energies[3 * s] = ke;
      energies[3 * s + 1] = pe;
      energies[3 * s + 2] = w;
    }
  }
  free(F);
  return 0;
}

/* ------------------------------------------------- f32-exact integer paths */

float pbo_distance2p(float r1x, float r1y, float r1z, float r2x, float r2y,
                     float r2z, float cx, float cy, float cz) {
  /* kernelinline.hh:47-62 */
  r1x -= r2x;
  r1y -= r2y;
  r1z -= r2z;
  r1x *= r1x;
  r1y *= r1y;
  r1z *= r1z;
  if (r1x > cx) r1x += 4.0f * (cx - sqrtf(cx * r1x));
  if (r1y > cy) r1y += 4.0f * (cy - sqrtf(cy * r1y));
  if (r1z > cz) r1z += 4.0f * (cz - sqrtf(cz * r1z));
  return r1x + r1y + r1z;
}

static void half_cell2(const float cell[6], float c[3]) {
  /* cudaFrame::setCell, cudaFrame.cu:85-91 */
  for (int i = 0; i < 3; ++i) {
    c[i] = (cell[i * 2 + 1] - cell[i * 2]) / 2;
    c[i] *= c[i];
  }
}

int pbo_coord_counts_f32(const float *r, uint32_t n, const float cell[6],
                         float rmax_cells, float rmax2t, uint32_t *counts) {
  uint32_t bn[3];
  float bl[6];
  if (pbo_setup_cells(cell, rmax_cells, 1, bn, bl)) return -1;
  const uint32_t ncell = bn[0] * bn[1] * bn[2];
  uint32_t *bid = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *pid = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *bs = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *bindex = (uint32_t *)malloc(sizeof(uint32_t) * (ncell + 1));
  uint32_t *bend = (uint32_t *)malloc(sizeof(uint32_t) * (ncell + 1));
  uint32_t *table = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)ncell * 27);
  pbo_calc_bid(r, n, bl, bn, bid);
  pbo_sort_bins(bid, n, ncell, pid, bs, bindex);
  {
    uint32_t next = n;
    for (long c = (long)ncell - 1; c >= 0; --c) {
      bend[c] = next;
      if (bindex[c] != UINT32_MAX) next = bindex[c];
    }
  }
  pbo_neighbor_table(bn, 1, 1, table);
  float c[3];
  half_cell2(cell, c);
#pragma omp parallel for schedule(static)
  for (long I = 0; I < (long)ncell; ++I) {
    if (bindex[I] == UINT32_MAX) continue;
    for (uint32_t i0 = bindex[I]; i0 < bend[I]; ++i0) {
      const uint32_t i = pid[i0];
      uint32_t k = 0;
      for (int J27 = 0; J27 < 27; ++J27) {
        const uint32_t J = table[I * 27 + J27];
        if (bindex[J] == UINT32_MAX) continue;
        for (uint32_t j0 = bindex[J]; j0 < bend[J]; ++j0) {
          const uint32_t j = pid[j0];
          if (j == i) continue;
          const float R2 = pbo_distance2p(r[3 * i], r[3 * i + 1], r[3 * i + 2],
                                          r[3 * j], r[3 * j + 1], r[3 * j + 2],
                                          c[0], c[1], c[2]);
          if (R2 < rmax2t) ++k;
        }
      }
      counts[i] = k;
    }
  }
  free(bid); free(pid); free(bs); free(bindex); free(bend); free(table);
  return 0;
}

static int cmp_u32(const void *a, const void *b) {
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

int pbo_nlist_f32(const float *r, uint32_t n, const float cell[6], float rmax_cells,
                  float rmax2t, uint32_t *rowptr, uint32_t *colidx, uint32_t cap,
                  uint32_t *nnz) {
  /* makeNList, cudaParticleMD.cu:338-394: counts as pbo_coord_counts_f32,
   * then the same traversal fills each row (makeJlist_WithBlock_F4,
   * kerneltemplate.hh:725-775); rows sorted ascending. */
  uint32_t *cnt = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  if (pbo_coord_counts_f32(r, n, cell, rmax_cells, rmax2t, cnt)) {
    free(cnt);
    return -1;
  }
  rowptr[0] = 0;
  for (uint32_t i = 0; i < n; ++i) rowptr[i + 1] = rowptr[i] + cnt[i];
  free(cnt);
  *nnz = rowptr[n];
  if (!colidx || cap < rowptr[n]) return 0;
  uint32_t bn[3];
  float bl[6];
  pbo_setup_cells(cell, rmax_cells, 1, bn, bl);
  const uint32_t ncell = bn[0] * bn[1] * bn[2];
  uint32_t *bid = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *pid = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *bs = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *bindex = (uint32_t *)malloc(sizeof(uint32_t) * (ncell + 1));
  uint32_t *bend = (uint32_t *)malloc(sizeof(uint32_t) * (ncell + 1));
  uint32_t *table = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)ncell * 27);
  pbo_calc_bid(r, n, bl, bn, bid);
  pbo_sort_bins(bid, n, ncell, pid, bs, bindex);
  {
    uint32_t next = n;
    for (long c = (long)ncell - 1; c >= 0; --c) {
      bend[c] = next;
      if (bindex[c] != UINT32_MAX) next = bindex[c];
    }
  }
  pbo_neighbor_table(bn, 1, 1, table);
  float c[3];
  half_cell2(cell, c);
#pragma omp parallel for schedule(static)
  for (long I = 0; I < (long)ncell; ++I) {
    if (bindex[I] == UINT32_MAX) continue;
    for (uint32_t i0 = bindex[I]; i0 < bend[I]; ++i0) {
      const uint32_t i = pid[i0];
      uint32_t k = rowptr[i];
      for (int J27 = 0; J27 < 27; ++J27) {
        const uint32_t J = table[I * 27 + J27];
        if (bindex[J] == UINT32_MAX) continue;
        for (uint32_t j0 = bindex[J]; j0 < bend[J]; ++j0) {
          const uint32_t j = pid[j0];
          if (j == i) continue;
          const float R2 = pbo_distance2p(r[3 * i], r[3 * i + 1], r[3 * i + 2],
                                          r[3 * j], r[3 * j + 1], r[3 * j + 2],
                                          c[0], c[1], c[2]);
          if (R2 < rmax2t) colidx[k++] = j;
        }
      }
      qsort(colidx + rowptr[i], rowptr[i + 1] - rowptr[i], sizeof(uint32_t), cmp_u32);
    }
  }
  free(bid); free(pid); free(bs); free(bindex); free(bend); free(table);
  return 0;
}

void pbo_lj_forces_list(const double *r, const uint16_t *type, uint32_t n,
                        const double cell[6], double rc2, const pbo_lj_params *p,
                        const uint32_t *rowptr, const uint32_t *colidx, double *F,
                        double *pe, double *virial) {
  /* calcF_IJpairWithList_F4, kerneltemplate.hh:136-168, FP64 */
  const double L[3] = {cell[1] - cell[0], cell[3] - cell[2], cell[5] - cell[4]};
  double pe_s = 0.0, w_s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : pe_s, w_s)
  for (long i = 0; i < (long)n; ++i) {
    double f[3] = {0, 0, 0}, ph = 0, w = 0;
    for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const uint32_t j = colidx[k];
      double d[3];
      for (int a = 0; a < 3; ++a) {
        d[a] = r[3 * i + a] - r[3 * j + a];
        d[a] -= L[a] * nearbyint(d[a] / L[a]);
      }
      const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (R2 < rc2) {
        double phi;
        const double fr = lj_pair(R2, type[i] * PBO_MAXTYPES + type[j], p, &phi);
        for (int a = 0; a < 3; ++a) f[a] += fr * d[a];
        ph += phi;
        w += fr * R2;
      }
    }
    for (int a = 0; a < 3; ++a) F[3 * i + a] = f[a];
    pe_s += ph;
    w_s += w;
  }
  if (pe) *pe = 0.5 * pe_s;
  if (virial) *virial = 0.5 * w_s;
}

void pbo_softcore_forces(const double *r, uint32_t n, const double cell[6], double rc,
                         double f0, double rmax2, double *F) {
  /* potentialSC::F (kernelinline.hh:152-157) through MDpairForce
   * (kerneltemplate.hh:530-564), FP64, all pairs (minimum image) */
  const double L[3] = {cell[1] - cell[0], cell[3] - cell[2], cell[5] - cell[4]};
  const double pi2 = 0.5 * 3.14159265358979323846;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) {
    double f[3] = {0, 0, 0};
    for (uint32_t j = 0; j < n; ++j) {
      if ((long)j == i) continue;
      double d[3];
      for (int a = 0; a < 3; ++a) {
        d[a] = r[3 * i + a] - r[3 * j + a];
        d[a] -= L[a] * nearbyint(d[a] / L[a]);
      }
      const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (!(R2 < rmax2)) continue;
      const double R = sqrt(R2);
      if (R > rc) continue;
      const double fr = f0 * cos(R / rc * pi2) / R;
      for (int a = 0; a < 3; ++a) f[a] += fr * d[a];
    }
    for (int a = 0; a < 3; ++a) F[3 * i + a] = f[a];
  }
}

void pbo_euler_step(double *r, double *v, const double *a, const uint8_t *move, uint32_t n,
                    double dt) {
  /* propagateEuler_F4, kernelfuncs.cu:174-192 */
  for (uint32_t i = 0; i < n; ++i) {
    const double d = move[i] ? dt : 0.0;
    for (int c = 0; c < 3; ++c) {
      r[3 * i + c] += v[3 * i + c] * d;
      v[3 * i + c] += a[3 * i + c] * d;
    }
  }
}

void pbo_gr_setup(float rstep, float rmax, float *rstepinv, uint32_t *rnum) {
  /* cudaCalcGR::setup, cudaCalcGR.hh:25-34 */
  *rstepinv = (float)(1.0 / (double)rstep);
  *rnum = f2u_sat(rmax * *rstepinv);
}

static inline uint32_t gr_bin(float R2, float rstepinv, uint32_t rnum) {
  /* calcGR_F4, kernelinline.hh:442-445 */
  const float r0 = sqrtf(R2) * rstepinv;
  const double t = (double)r0 + 0.5;
  return (t < (double)rnum) ? (uint32_t)t : 0u;
}

void pbo_gr_hist_f32(const float *r, uint32_t n, const float cell[6],
                     float rstepinv, uint32_t rnum, uint64_t *hist) {
  float c[3];
  half_cell2(cell, c);
  memset(hist, 0, sizeof(uint64_t) * rnum);
#pragma omp parallel
  {
    uint64_t *h = (uint64_t *)calloc(rnum, sizeof(uint64_t));
#pragma omp for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; ++i)
      for (uint32_t j = (uint32_t)i + 1; j < n; ++j) {
        const float R2 = pbo_distance2p(r[3 * i], r[3 * i + 1], r[3 * i + 2],
                                        r[3 * j], r[3 * j + 1], r[3 * j + 2],
                                        c[0], c[1], c[2]);
        h[gr_bin(R2, rstepinv, rnum)]++;
      }
#pragma omp critical
    for (uint32_t b = 0; b < rnum; ++b) hist[b] += h[b];
    free(h);
  }
}

void pbo_gr_normalize(const uint64_t *igr, uint32_t rnum, float rstepinv,
                      const float cell[6], uint32_t n, float *gr) {
  /* igr2gr + kigr2gr, cudaCalcGR.cu:14-28, 115-132 */
  const float rhoinv = (cell[1] - cell[0]) * (cell[3] - cell[2]) *
                       (cell[5] - cell[4]) / (float)n;
  const float Vinv = (float)(M_1_PI * rstepinv * rstepinv * rstepinv);
  const float Q = (float)(2.0 / n * (double)Vinv * (double)rhoinv);
  const float onethird = (float)(1.0 / 3.0);
  for (uint32_t j = 1; j < rnum; ++j) {
    const float rinv2 = (float)(1.0 / (double)((float)(4u * j * j) + onethird));
    gr[j] = (float)igr[j] * Q * rinv2;
  }
  if (rnum) gr[0] = 0.0f;
}
