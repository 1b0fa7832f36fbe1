/*
 * pb_oracle.h -- CPU restatement of the cudaParticleBase short-range path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (particlebase_amd/,
 * include/) links, loads or calls this library.  It is used by tests/ as the
 * checker, by __graft_entry__.smoke() as the checker, and by bench.py's
 * cpu_baseline leg as the timed CPU path ("kind": "port").
 *
 * Two arithmetic modes are restated:
 *   - "f32-exact": the reference's FP32 operation order (no FMA contraction,
 *     correctly rounded div/sqrt), for the integer outputs that must match
 *     bit for bit (cell ids, sort, neighbour counts, g(r) histograms).
 *   - FP64: the tolerance oracle for forces, energies and trajectories.
 *
 * Every function names the reference file:line it restates.  The reference
 * ships no golden vectors, fixtures or tests; each function is pinned in
 * tests/test_oracle_*.py by closed forms, invariants, brute force on small
 * inputs or textbook routines (see the "pinned by" line of each function).
 */
#ifndef PB_ORACLE_H
#define PB_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBO_MAXTYPES 5   /* MAXELEMNUM, kernelinline.hh:68 */

/* ---------------------------------------------------------------- cells */

/* cudaCutoffBlock::setupCutoffBlock, cudaCutoffBlock.cu:97-113 (FP32).
 * periodic: nb = (uint)(c/rmax), len = c/(float)nb; else len = rmax,
 * nb = (uint)(c/rmax)+1.  blocklen[3..5] = box minima.
 * Returns 0, or -1 if some nb <= 2 (the reference asserts nb > 2).
 * pinned by: closed form in tests/test_oracle_cells.py */
int pbo_setup_cells(const float cell[6], float rmax, int periodic,
                    uint32_t blocknum[3], float blocklen[6]);

/* calc_BlockNeighbor, cudaCutoffBlock.cu:17-65: neighbour cell table,
 * (2w+1)^3 entries per cell in (i1,j1,k1) loop order, UINT32_MAX outside a
 * non-periodic box.  table has ncell*(2w+1)^3 entries.
 * pinned by: closed form + symmetry. */
void pbo_neighbor_table(const uint32_t blocknum[3], int w, int periodic,
                        uint32_t *table);

/* calcBID_F4, kernelfuncs.cu:807-841, f32-exact, in identity (pid=i) order.
 * r is xyz-interleaved FP32 (stride 3).
 * pinned by: edge cases (xmin, xmax, nextafter, negative) + cell-interval
 * containment. */
void pbo_calc_bid(const float *r, uint32_t n, const float blocklen[6],
                  const uint32_t blocknum[3], uint32_t *bid);

/* Canonical replacement for sortByBID_M1/sortByBID/makeBindex
 * (kernelfuncs.cu:844-932, cudaCutoffBlock.cu:241-256): stable sort of
 * (bid, pid) from identity order; bindex[c] = first sorted slot of cell c or
 * UINT32_MAX if c is empty (the reference's makeBindex leaves cell 0 at 0
 * when it is empty; the correct semantics are used), bindex[ncell] = n.
 * pinned by: numpy.argsort(kind="stable") and counts. */
void pbo_sort_bins(const uint32_t *bid, uint32_t n, uint32_t ncell,
                   uint32_t *pid, uint32_t *bid_sorted, uint32_t *bindex);

/* ------------------------------------------------------------------ LJ */

/* LJ parameters as the reference's setLJparams vector (kernelinline.hh:
 * 120-134): sigma_eps[(i*ntypes+j)*2+0] = sigma, [..+1] = epsilon. */
typedef struct {
  int ntypes;
  double s2[PBO_MAXTYPES * PBO_MAXTYPES];   /* sigma^2 */
  double e4[PBO_MAXTYPES * PBO_MAXTYPES];   /* 4 epsilon */
  double shift[PBO_MAXTYPES * PBO_MAXTYPES];/* phi(rc) per pair type (0 if unshifted) */
} pbo_lj_params;

void pbo_lj_make_params(const double *sigma_eps, int ntypes, double rc2,
                        int shifted, pbo_lj_params *p);

/* LJ forces (MDpairForce<potentialLJ>, kerneltemplate.hh:530-564;
 * potentialLJ::F kernelinline.hh:86-96) in FP64 over the 27-cell stencil of
 * calcF_IJpairWithBlock_F4 (kerneltemplate.hh:244-293), full neighbour sum
 * (no Newton's third law), minimum image by the reference's +-L rule
 * (kerneltemplate.hh:541-543), cutoff R2 < rc2.  F is written in original
 * particle order (as reduce27_F4 does, kernelfuncs.cu:935-949).
 * Also returns the correct PE = 1/2 sum phi (shifted if the params are) and
 * the virial W = 1/2 sum f.r (the reference's calcPotentialE is defective;
 * the correct quantity is returned).  r, F xyz-interleaved FP64.
 * Cells: the caller passes the cell grid (blocknum, blocklen) to use.
 * pinned by: two-body closed form, FCC F=0, sum F = 0, brute force below. */
int pbo_lj_forces_cells(const double *r, const uint16_t *type, uint32_t n,
                        const double cell[6], const uint32_t blocknum[3],
                        const double blocklen[6], double rc2,
                        const pbo_lj_params *p, double *F, double *pe,
                        double *virial);

/* O(N^2) minimum-image FP64 brute force: the independent check of the cell
 * path above (same physics, no cells). */
void pbo_lj_forces_brute(const double *r, const uint16_t *type, uint32_t n,
                         const double cell[6], double rc2,
                         const pbo_lj_params *p, double *F, double *pe,
                         double *virial);

/* propagateVelocityVerlet_F4, kernelfuncs.cu:333-367, FP64:
 * v += (F+Fold) dt/2m ; r += v dt + F dt^2/2m ; Fold = F. */
void pbo_vv_step(double *r, double *v, const double *F, double *Fold,
                 const double *minv, uint32_t n, double dt);

/* applyPeriodicCondition_F4, kernelfuncs.cu:87-139, FP64, with one defined
 * deviation: a coordinate that rounds onto the upper face after wrapping is
 * put on the lower face (the reference asserts there). */
void pbo_wrap(double *r, uint32_t n, const double cell[6]);

/* sum m v^2 in FP64 (calcMV2, cudaParticleMD.cu:124-133). */
double pbo_mv2(const double *v, const double *minv, uint32_t n);

/* The reference LJ loop of testMD1.cu:135-169 restated in FP64:
 * per step: rebin (rmax cells), forces, VV, wrap.  Fold must be zero at
 * start for the reference's half-kick start (cudaParticleVV.cu:10-15).
 * energies (if not NULL) receives per step {KE(v(t)), PE(r(t)), W(r(t))}
 * where v(t) is the VV velocity synchronous with the force at r(t).
 * nthreads <= 0 uses all OpenMP threads.  Returns 0 or -1 on bad cells. */
int pbo_lj_run(double *r, double *v, double *Fold, const double *minv,
               const uint16_t *type, uint32_t n, const double cell[6],
               double rmax_cells, double rc2, const pbo_lj_params *p,
               double dt, uint32_t nsteps, double *energies, int nthreads);

/* ------------------------------------------------- f32-exact integer paths */

/* calcCoordMD_F4 (kernelinline.hh:387-399) driven by
 * calcF_IJpairWithBlock_F4 over the 27-cell table (kerneltemplate.hh:
 * 244-293) as in cudaParticleMD::makeNList step 1 (cudaParticleMD.cu:343-366):
 * count j != i with distance2p (kernelinline.hh:47-62) < rmax2t, FP32.
 * rmax2t is given (the caller forms (sqrt(rmax2)+thickness)^2 in FP32).
 * counts in original order.  Returns -1 on bad cells.
 * pinned by: O(N^2) brute force and FCC shell counts. */
int pbo_coord_counts_f32(const float *r, uint32_t n, const float cell[6],
                         float rmax_cells, float rmax2t, uint32_t *counts);

/* distance2p, kernelinline.hh:47-62, FP32 exact op order. */
float pbo_distance2p(float r1x, float r1y, float r1z, float r2x, float r2y,
                     float r2z, float cx, float cy, float cz);

/* g(r) integer histogram: calcGR_F4 (kernelinline.hh:440-448) over every
 * unordered pair, the pair set of calcF_IJpairWithBlock5_F4 +
 * makePairInfo (kerneltemplate.hh:423-512, cudaCalcGR.cu:30-84).
 * bin = (uint)(sqrt(R2)*rstepinv + 0.5 [double]); bins >= rnum go to 0;
 * hist[0] is therefore the overflow count (the reference drops it).
 * rstepinv/rnum as cudaCalcGR::setup (cudaCalcGR.hh:25-34).
 * pinned by: FCC shell counts (12,6,24,12,24) and brute-force pair count. */
void pbo_gr_hist_f32(const float *r, uint32_t n, const float cell[6],
                     float rstepinv, uint32_t rnum, uint64_t *hist);

/* cudaCalcGR::setup scalars: rstepinv = (float)(1.0/rstep),
 * rnum = (uint)(rmax*rstepinv) in FP32. */
void pbo_gr_setup(float rstep, float rmax, float *rstepinv, uint32_t *rnum);

/* igr2gr / kigr2gr (cudaCalcGR.cu:14-28, 115-132) in FP32 op order. */
void pbo_gr_normalize(const uint64_t *igr, uint32_t rnum, float rstepinv,
                      const float cell[6], uint32_t n, float *gr);

/* ------------------------------------------------------------------ SPH */
/* Reference physics of cudaParticleSPH_NS (pb_oracle_sph.c): Lucy kernel of
 * support h, p = c^2 (rho - rho0) with rho0 = 1, physical viscosity.
 * pinned by: int W 4 pi r^2 dr = 1, single-particle rho = 105 m/(16 pi h^3),
 * two-particle closed forms, sum m_i a_i = 0 (tests/test_oracle_sph.py). */
double pbo_sph_w0(double h);
double pbo_sph_lucy_w(double r, double h, double w0);
double pbo_sph_lucy_dw(double r, double h, double w1);
void pbo_sph_density(const double *r, const double *m, const double *mu0, uint32_t n,
                     double h, double *num_out, double *rhoinv);
void pbo_sph_accel(const double *r, const double *v, const double *m, const double *c2,
                   const double *rhoinv, const double *mu, uint32_t n, double h,
                   double rho0, double *a);
/* Config physics of the dam-break workload (pb_oracle_sph.c): M4 cubic
 * spline of support 2h, summation density, Tait EOS clamped at pmin,
 * Monaghan artificial viscosity with mean sound speed c0; no counterpart in
 * the reference's kernels.  pinned by: int W 4 pi r^2 dr = 1, W and dW/dr
 * continuous at q = 1 and 2, dW/dr against a finite difference of W, the
 * lattice sum, single-particle rho = m/(pi h^3), gamma = 1 reducing to the
 * reference EOS c0^2 (rho - rho0), sum m_i a_i = 0, sum m_i r_i x a_i = 0,
 * viscous dissipation <= 0 and none between receding pairs
 * (tests/test_oracle_sph.py). */
double pbo_wc_w(double r, double h);
double pbo_wc_dw(double r, double h);
void pbo_wc_density(const double *r, const double *m, uint32_t n, double h, double rho0,
                    double c0, double gamma, double pmin, double *num_out, double *rhoinv);
void pbo_wc_accel(const double *r, const double *v, const double *m, const double *rho,
                  const double *p, uint32_t n, double h, double c0, double alpha, double *a);
/* propagateLeapFrog_F4 (kernelfuncs.cu:270-296), FP64 */
void pbo_lf_step(double *r, double *v, const double *a, const uint8_t *move, uint32_t n,
                 double dt);
/* rollbackLeapFrog_F4 (kernelfuncs.cu:298-327), FP64.
 * pinned by: two leap-frog steps under a constant acceleration followed by
 * the rollback give back r(0) and the first half-step velocity
 * (tests/test_oracle_dem.py). */
void pbo_lf_rollback(double *r, double *v, const double *a, const uint8_t *move, uint32_t n,
                     double dt);

/* ------------------------------------------------------------------ DEM */
/* cudaParticleDEM restated in FP64 (pb_oracle_dem.c).
 * pinned by: brute-force geometry, Hertz closed form on two spheres,
 * Coulomb cap, torque = Rij x Fs, scripted history sequences, rotation
 * closed form (tests/test_oracle_dem.py). */
uint32_t pbo_dem_contacts(const double *r, const double *rad, const uint8_t *move, uint32_t n,
                          uint32_t *rowptr, uint32_t *colidx);
void pbo_dem_carry(const uint32_t *prow, const uint32_t *pcol, const double *pval,
                   const uint32_t *crow, const uint32_t *ccol, double *cval, uint32_t n);
uint32_t pbo_dem_force(const double *r, const double *v, const double *w, const double *rad,
                       const double *minv, const uint8_t *move, uint32_t n,
                       const uint32_t *rowptr, const uint32_t *colidx, double *val, double E,
                       double mu, double s2, double gamma2, double mu_r, double dt, double *F,
                       double *T);
void pbo_rot_step(double *w, double *L, const double *T, const double *m0inv,
                  const uint8_t *move, uint32_t n, double dt);
/* Linear spring-dashpot + tangential spring + Coulomb law (published
 * method, no reference counterpart; hist = 3 doubles per contact).
 * pinned by: head-on collision time pi/omega_d and restitution e from a
 * two-sphere integration, e = 1 conserving energy, the static spring
 * force kn delta, stick (Ft = -kt sum vt dt) and slip (|Ft| = mu |Fn|,
 * opposite vt) regimes, torque = (-rad_i n) x Ft, Newton's third law,
 * the tangent-plane projection of the history, a fixed sphere acting as an
 * infinite mass (tests/test_oracle_dem.py). */
uint32_t pbo_dem_force_lsd(const double *r, const double *v, const double *w, const double *rad,
                           const double *minv, const uint8_t *move, uint32_t n,
                           const uint32_t *rowptr, const uint32_t *colidx, double *hist,
                           double kn, double zeta, double kt, double mu, double dt, double *F,
                           double *T);
/* pbo_dem_contacts over a cell grid for large N (same test, same CSR).
 * pinned by: equality with the brute-force pbo_dem_contacts on random
 * packings, cells at the box faces and centres outside the box included
 * (tests/test_oracle_dem.py). */
uint32_t pbo_dem_contacts_cells(const double *r, const double *rad, const uint8_t *move,
                                uint32_t n, const double cell[6], double cellsize,
                                uint32_t *rowptr, uint32_t *colidx);
/* pbo_dem_carry with hw doubles of history per contact.
 * pinned by: equality with pbo_dem_carry at hw = 1 and per-component
 * scripted sequences at hw = 3 (tests/test_oracle_dem.py). */
void pbo_dem_carry_w(const uint32_t *prow, const uint32_t *pcol, const double *pval,
                     const uint32_t *crow, const uint32_t *ccol, double *cval, uint32_t n,
                     uint32_t hw);

/* ------------------------------------------- MD neighbour list, annealing */
/* makeNList (cudaParticleMD.cu:338-394): CSR list by original index over the
 * periodic 27-cell grid of rmax_cells, j != i with FP32 distance2p < rmax2t,
 * rows ascending.  rowptr[n+1] always filled, *nnz = entries; colidx filled
 * when cap >= nnz.  Returns -1 on bad cells.
 * pinned by: numpy FP32 brute force over all pairs (tests/test_oracle_md.py). */
int pbo_nlist_f32(const float *r, uint32_t n, const float cell[6], float rmax_cells,
                  float rmax2t, uint32_t *rowptr, uint32_t *colidx, uint32_t cap,
                  uint32_t *nnz);
/* calcF_IJpairWithList_F4 in FP64: LJ forces of the list's pairs below rc2.
 * pinned by: equality with pbo_lj_forces_brute for a complete list, and a
 * list with pairs removed. */
void pbo_lj_forces_list(const double *r, const uint16_t *type, uint32_t n,
                        const double cell[6], double rc2, const pbo_lj_params *p,
                        const uint32_t *rowptr, const uint32_t *colidx, double *F,
                        double *pe, double *virial);
/* initialAnnealing's soft-core force (potentialSC), FP64, all pairs.
 * pinned by: two-particle closed form, zero beyond rc, sum F = 0. */
void pbo_softcore_forces(const double *r, uint32_t n, const double cell[6], double rc,
                         double f0, double rmax2, double *F);
/* propagateEuler_F4 (kernelfuncs.cu:174-192), FP64.
 * pinned by: constant-acceleration closed form after k steps. */
void pbo_euler_step(double *r, double *v, const double *a, const uint8_t *move, uint32_t n,
                    double dt);

/* ------------------------------------------------------------- thermostats */
/* GaussianThermo / adjustVelocities / constTemp restated in FP64
 * (pb_oracle_thermo.c); xyz-interleaved arrays.
 * pinned by: the closed-form Gaussian fixed point, the isokinetic property,
 * zero mean / target variance, sum F.v = 0 after constTemp
 * (tests/test_oracle_thermo.py). */
int pbo_gt_step(double *r, double *v, const double *F, double *Fold, const double *minv,
                uint32_t n, double dt, double *xi_io, double tol);
void pbo_adjust_velocities(double *v, uint32_t n, double v1, double *stats);
double pbo_const_temp(const double *v, double *F, const double *minv, uint32_t n,
                      double *lambda_out);

#ifdef __cplusplus
}
#endif
#endif
