/*
 * pbamd.h -- C ABI of the MI355X-native short-range particle engine.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (PB_OK = 0, negative on error; the message is in pb_ctx_last_error) except
 * the few that return a value.  One context = one HIP device + one stream;
 * calls on a context are serialised by the caller (the reference's objects
 * are not thread-safe either).  Device memory is owned by the library, host
 * buffers by the caller.  Per-particle host arrays are in the caller's
 * original particle order; vectors are xyz-interleaved (3 doubles a particle).
 *
 * Each declaration names the reference interface it replaces (file:line in
 * particlelover/ParticleBase).  A "system" is one cudaParticleBase object,
 * or a batch of independent periodic replicas that share one launch.
 */
#ifndef PBAMD_H
#define PBAMD_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- status codes */
#define PB_OK               0
#define PB_ERR_ARG         -1   /* bad argument (sizes, nulls, ranges) */
#define PB_ERR_HIP         -2   /* a HIP runtime call failed */
#define PB_ERR_STATE       -3   /* call out of order (e.g. no cells set up) */
#define PB_ERR_CELLS       -4   /* cell grid invalid (nb <= 2, len < cutoff) */
#define PB_ERR_NLIST       -5   /* neighbour list overflowed its capacity */
#define PB_ERR_SKIN        -6   /* a particle moved more than skin/2 between
                                   list rebuilds: pairs may have been missed */
#define PB_ERR_DT          -7   /* DEM time step above the Cundall-Strack bound
                                   (the reference asserts, kernelinline.hh:285-286) */

/* precision modes of the pair kernels */
#define PB_FP64            0    /* pair math, integration, reductions in FP64 */
#define PB_MIXED           1    /* FP32 pair math on FP64 displacements;
                                   FP64 accumulation, integration, reductions */

/* fields for pb_sys_get / pb_sys_set (original particle order) */
#define PB_FIELD_R         0    /* double[3n] position      (r,   cudaParticleBase.hh:33) */
#define PB_FIELD_V         1    /* double[3n] velocity      (v,   :35) */
#define PB_FIELD_A         2    /* double[3n] acceleration; Fold in VV (a, :37; cudaParticleVV.cu:7) */
#define PB_FIELD_F         3    /* double[3n] force         (F,   :41) */
#define PB_FIELD_MINV      4    /* double[n]  1/mass        (minv,:39) */
#define PB_FIELD_TYPE      5    /* uint16[n]  type id       (typeID, :45) */
#define PB_FIELD_MOVE      6    /* uint8[n]   movable flag  (move, :47) */
#define PB_FIELD_W         7    /* double[3n] angular velocity (cudaParticleRotation.hh:22) */
#define PB_FIELD_L         8    /* double[3n] angular momentum (:24) */
#define PB_FIELD_T         9    /* double[3n] torque           (:26) */
#define PB_FIELD_RADIUS   10    /* double[n]  radius (DEM r.w, cudaParticleDEM.cu:221-252) */
#define PB_FIELD_SPH_NUM  11    /* double[3n] number density, rho, mu (num, cudaParticleSPHBase.hh:24) */
#define PB_FIELD_SPH_RHOINV 12  /* double[n]  1/rho (rhoinv, cudaParticleSPHBase.hh:27) (get only) */

typedef struct pb_ctx pb_ctx;
typedef struct pb_sys pb_sys;

/* --------------------------------------------------------------- context */
/* Replaces CUDAenv<T>'s per-device setup (CUDAenv.hh:93-156) for one GPU. */
int         pb_ctx_create(int device, pb_ctx **out);
/* Releases the context.  Refused with PB_ERR_STATE (the context stays valid)
 * while systems created on it are alive: destroy those first. */
int         pb_ctx_destroy(pb_ctx *ctx);
/* Replaces cudaFrame::ErrorInfo (cudaFrame.cu:20-32): message of the last
 * failing call on this context ("" if none). */
const char *pb_ctx_last_error(const pb_ctx *ctx);
int         pb_ctx_sync(pb_ctx *ctx);
/* Device name and compute-unit count (CUDAenv::setThnum, CUDAenv.hh:36-50). */
int         pb_ctx_info(pb_ctx *ctx, char *name, size_t name_len, int *cu_count);
/* Device limits gathered by CUDAenv's constructor (CUDAenv.hh:110-118):
 * global memory in bytes (totalGlobalMem), threads per block, grid size in
 * x, LDS bytes per block (sharedMemPerBlock) and vector lanes per compute
 * unit (numberOfCores).  Null pointers are skipped. */
int         pb_ctx_limits(pb_ctx *ctx, uint64_t *global_mem, uint32_t *max_threads,
                          uint32_t *max_grid, uint32_t *lds_per_block,
                          uint32_t *lanes_per_cu);

/* ---------------------------------------------------------------- system */
/* cudaParticleBase::setup(int n) (cudaParticleBase.cu:18-66).  nrep = 1 is
 * the reference's single object; nrep > 1 is a batch of independent
 * replicas with counts[nrep] particles each, stored back to back. */
int  pb_sys_create(pb_ctx *ctx, uint32_t nrep, const uint32_t *counts,
                   pb_sys **out);
void pb_sys_destroy(pb_sys *sys);
uint32_t pb_sys_count(const pb_sys *sys);     /* total particles */
uint32_t pb_sys_replicas(const pb_sys *sys);

/* cudaFrame::setCell (cudaFrame.cu:85-91): cells[6*nrep] = xmin,xmax,ymin,
 * ymax,zmin,zmax per replica. */
int pb_sys_set_cell(pb_sys *sys, const double *cells);

/* cudaParticleBase::import(vector<ParticleBase>) (cudaParticleBase.cu:
 * 68-117), field by field of ParticleBase (ParticleBase.hh:10-18).
 * v, a may be NULL (zero); is_fixed may be NULL (all movable). */
int pb_sys_import(pb_sys *sys, const double *r, const double *v,
                  const double *a, const double *m, const uint32_t *type,
                  const uint8_t *is_fixed);

/* getPosition / getAcceleration / getForce / getTypeID
 * (cudaParticleBase.cu:119-149) and the matching uploads. */
int pb_sys_get(pb_sys *sys, int field, void *host);
int pb_sys_set(pb_sys *sys, int field, const void *host);

/* cudaParticleBase::clearForce + cudaParticleVV::clearForce
 * (cudaParticleBase.cu:158-163, cudaParticleVV.cu:10-15): F = 0 and
 * (with_fold) a = Fold = 0. */
int pb_clear_force(pb_sys *sys, int with_fold);
/* calcAcceleration: a = F * minv (cudaParticleBase.cu:170-173). */
int pb_calc_acceleration(pb_sys *sys);
/* addForceX/Y/Z, addAccelerationX/Y/Z (cudaParticleBase.cu:208-242). */
int pb_add_force(pb_sys *sys, int axis, double value);
int pb_add_acceleration(pb_sys *sys, int axis, double value);
/* treatPeriodicCondition / treatAbsoluteCondition / treatRefrectCondition
 * (cudaParticleBase.cu:175-206; kernelfuncs.cu:87-172). */
int pb_treat_periodic(pb_sys *sys);
int pb_treat_absolute(pb_sys *sys);
int pb_treat_reflect(pb_sys *sys);
/* cudaParticleVV::TimeEvolution (cudaParticleVV.cu:17-34). */
int pb_vv_step(pb_sys *sys, double dt);
/* cudaParticleLF::calcVinit / TimeEvolution / rollback
 * (cudaParticleLF.cu:4-22; kernelfuncs.cu:246-327). */
int pb_lf_vinit(pb_sys *sys, double dt);
int pb_lf_step(pb_sys *sys, double dt);
int pb_lf_rollback(pb_sys *sys, double dt);
/* cudaParticleMD::calcMV2 (cudaParticleMD.cu:124-133): per-replica
 * sum m v^2 in FP64 (replaces the cuBLAS FP32 dot). out[nrep]. */
int pb_mv2(pb_sys *sys, double *out);
/* mulArray on v (scaleTemp, cudaParticleMD.cu:168-175). */
int pb_scale_velocities(pb_sys *sys, double s);
/* cudaParticleBase::inspectVelocity (cudaParticleBase.cu:244-282,
 * inspectV_F4 kernelfuncs.cu:194-240).  vlim is truncated to uint32 as the
 * reference's kernel signature does (kernelfuncs.h:116).  *result = 1, 0 or
 * -1; vratio_max / imax (may be NULL) report the largest |v|/vlim. */
int pb_inspect_velocity(pb_sys *sys, double vlim, double lim_u,
                        double lim_l, int *result, double *vratio_max,
                        uint32_t *imax);

/* ------------------------------------------------------------ cell grid */
/* cudaCutoffBlock::setupCutoffBlock (cudaCutoffBlock.cu:75-231): FP32 cell
 * rules of the reference, single replica only.  single_particle selects the
 * 125-neighbour single-occupancy mode (:115-138). */
int pb_setup_cells(pb_sys *sys, float rmax, int periodic, int single_particle);
/* the blocknum / blocklen fields that setupCutoffBlock fills
 * (cudaCutoffBlock.hh:30-32). */
int pb_get_cell_info(pb_sys *sys, uint32_t blocknum[3], float blocklen[6]);
/* calcBID_F4 + sort + makeBindex, then cudaSortedPositions::calcBlockID
 * (cudaCutoffBlock.cu:233-258, kernelfuncs.cu:807-932,
 * cudaSortedPositions.cu:27-30): cell ids from FP32 positions exactly as
 * the reference computes them; canonical stable order (ascending original
 * index inside a cell). */
int pb_bin(pb_sys *sys);
/* download of the pid / bid / bindex arrays that calcBlockID leaves
 * (cudaCutoffBlock.cu:233-258): pid[n], bid (sorted)[n], bindex[ncell+1]
 * (UINT32_MAX = empty). */
int pb_get_bins(pb_sys *sys, uint32_t *pid, uint32_t *bid_sorted,
                uint32_t *bindex);
/* calc_BlockNeighbor table (cudaCutoffBlock.cu:17-65): ncell*27 or *125. */
int pb_get_neighbor_table(pb_sys *sys, uint32_t *table);

/* -------------------------------------------------------------------- LJ */
/* cudaParticleMD::setLJparams (cudaParticleMD.cu:200-202,
 * kernelinline.hh:120-134): sigma_eps[(i*ntypes+j)*2 + {0,1}] = sigma, eps. */
int pb_lj_set_params(pb_sys *sys, const double *sigma_eps, int ntypes);
/* cudaParticleMD::rmax2 (public field, cudaParticleMD.hh:160) plus the
 * energy shift the benchmark configs use (shifted = 1: phi(rc) subtracted
 * per pair in energies; forces are never shifted). */
int pb_lj_set_cutoff(pb_sys *sys, double rc2, int shifted);
/* cudaParticleMD::calcForce + a correct calcPotentialE
 * (cudaParticleMD.cu:24-122): F from the current positions; pe, virial
 * (FP64 per replica, may be NULL) = 1/2 sum phi, 1/2 sum f.r. */
int pb_lj_force(pb_sys *sys, int precision, double *pe, double *virial);
/* calcCoordMD_F4 counts via makeNList step 1 (cudaParticleMD.cu:343-366,
 * kernelinline.hh:387-399): FP32-exact number of j != i with distance2p <
 * rmax2t, over the 27 cells of the grid set by pb_setup_cells.
 * counts[n] in original order. */
int pb_coord_counts(pb_sys *sys, float rmax2t, uint32_t *counts);

/* Options of the fused LJ run (no reference counterpart as one call; it is
 * the loop of testMD1.cu:135-169 -- calcBlockID, calcForce, TimeEvolution,
 * treatPeriodicCondition -- fused on the GPU). */
typedef struct {
  double   dt;
  uint32_t nsteps;
  int      precision;      /* PB_FP64 or PB_MIXED */
  double   skin;           /* Verlet skin; list cutoff = rc + skin */
  uint32_t rebuild_every;  /* list rebuild + cell-order re-sort interval */
  uint32_t energy_every;   /* energies recorded every k steps (0 = never) */
  int      use_graph;      /* capture the step sequence in a hipGraph */
  uint32_t block_size;     /* threads per workgroup of the pair kernel (0 = auto) */
  int      timing;         /* k > 0: time every k-th step, the step after each
                              rebuild, and every rebuild with HIP events on the
                              context's stream (pb_lj_last_run_timing); a timed
                              step costs about 9 us of idle GPU before it */
  int      persistent;     /* 1: the step kernel runs as a persistent grid (two
                              workgroups per compute unit) draining the
                              (replica x tile) segments from per-XCD work queues;
                              0: one workgroup per segment */
} pb_run_opts;

/* Fused cudaParticleMD loop (the testMD1.cu:135-169 step: calcBlockID,
 * calcForce, TimeEvolution, treatPeriodicCondition; no single reference
 * function): nsteps of [force -> VV -> periodic wrap] with a
 * Verlet list re-built (and particles re-sorted into cell order) every
 * rebuild_every steps, or earlier when a particle may have moved half its
 * replica's skin.  With several replicas, replica q's skin is
 * skin * (T_q / T_min)^(1/4) from the temperatures at the start (capped
 * below half its box), so hot replicas do not force rebuilds of all.
 * Starts from and ends in the reference VV state (r(t), v(t-dt),
 * Fold = F(t-dt)); call pb_clear_force(sys, 1) first for the reference's
 * start.  energies (may be NULL) receives, for each recorded
 * step s (s % energy_every == 0), nrep x {KE(v(t)), PE(r(t)), W(r(t))}. */
int pb_lj_run(pb_sys *sys, const pb_run_opts *opts, double *energies);
/* (no reference counterpart) statistics of the last pb_lj_run: rebuild
 * count, max list length, and whether a skin violation occurred (1) --
 * never expected for sane skins. */
int pb_lj_last_run_stats(pb_sys *sys, uint32_t *rebuilds, uint32_t *max_nn,
                         int *skin_violation);

/* (no reference counterpart) GPU time of the last pb_lj_run with
 * opts.timing > 0: summed HIP-event time of the timed step kernels and of
 * the list rebuilds, and their counts. */
int pb_lj_last_run_timing(pb_sys *sys, double *step_ms, uint32_t *nstep,
                          double *rebuild_ms, uint32_t *nrebuild);

/* Tile-shape autotuner of the fused LJ run (no reference counterpart; the
 * reference's launch shapes come from CUDAenv::setThnum, CUDAenv.hh:36-50):
 * times `steps` fused steps with opts for each candidate tile shape (1 or 2
 * list-cell columns per tile side x 6 LDS fill targets, 0.55 to 1) from the same saved
 * state, keeps the fastest for later runs of this system and restores the
 * state.  Outputs (may be NULL): ms per step of the best candidate, its
 * columns per side and fill. */
int pb_lj_autotune(pb_sys *sys, const pb_run_opts *opts, uint32_t steps, double *best_ms,
                   uint32_t *best_w, double *best_fill);
/* Sets the tile shape of later fused runs directly (no reference
 * counterpart): tile_w = 1 or 2 list-cell columns per tile side, tile_fill
 * in [0.1, 1] the target fill of the LDS window stage (default 2, 0.72). */
int pb_lj_set_tiles(pb_sys *sys, uint32_t tile_w, double tile_fill);

/* ------------------------------------------------------------------ g(r) */
/* cudaCalcGR::setup + makePairInfo + calcgr (cudaCalcGR.hh:25-39,
 * cudaCalcGR.cu:30-113, kerneltemplate.hh:423-512, kernelinline.hh:440-448):
 * integer histogram over every unordered pair of each replica, FP32-exact
 * bins.  rstep[nrep] per replica, rnum bins; hist[nrep*rnum] (uint64).
 * hist[rep*rnum + 0] holds the pairs the reference drops (bin 0 and
 * overflow). */
int pb_gr_hist(pb_sys *sys, const float *rstep, uint32_t rnum,
               uint64_t *hist);

/* pb_gr_hist without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as pb_gr_hist; the replica table and
 * the histograms live in a static device array of the library (one per
 * device, about 1.1 MB, held until the process ends), so nrep <= 4096 and
 * nrep * rnum <= 131072 (PB_ERR_ARG beyond).  Every call on a device shares
 * that array: calls that overlap on one device (two host threads, two
 * contexts of one device) race and must be serialised by the caller. */
int pb_gr_hist_noalloc(pb_sys *sys, const float *rstep, uint32_t rnum,
                       uint64_t *hist);

/* ------------------------------------------------------------------- SPH */
/* Reference physics of cudaParticleSPH_NS: Lucy kernel of support h,
 * p = c^2 (rho - rho0) with rho0 = 1, physical viscosity mu; non-periodic
 * distances.  Single replica.  Cells: pb_setup_cells(sys, h, 0, 0) once and
 * pb_bin(sys) before every pb_sph_density (calcBlockID). */

/* cudaParticleSPH_NS::setSPHProperties (cudaParticleSPH_NS.cu:56-66) +
 * cudaParticleSPHBase::setSPHProperties (cudaParticleSPHBase.cu:119-124):
 * mu0[n] viscosity, c[n] sound speed (c^2 is kept), kernel radius h;
 * w0 = 105 / (16 pi h^7). */
int pb_sph_set_properties(pb_sys *sys, const double *mu0, const double *c, double h);

/* The dam-break workload's config physics behind the same SPH calls
 * (pb_sph_density, pb_sph_acceleration, pb_sph_inspect_density).  No
 * counterpart in the reference's kernels, whose SPH is the Lucy form above;
 * it follows the published weakly compressible method: M4 cubic spline of
 * support 2h (Monaghan & Lattanzio 1985), summation density with the self
 * term, Tait equation of state p = max(B ((rho/rho0)^gamma - 1), pmin) with
 * B = rho0 c0^2 / gamma, and Monaghan's artificial viscosity (alpha, beta = 0,
 * eps = 0.01, sound speed c0).  Masses are the particles' own.  Cells:
 * pb_setup_cells(sys, 2h, 0, 0).  The third component of PB_FIELD_SPH_NUM
 * is p here (mu under the reference physics).  pb_sph_set_properties
 * switches back to the reference physics. */
int pb_sph_set_wcsph(pb_sys *sys, double h, double rho0, double c0, double gamma,
                     double alpha, double pmin);

/* cudaParticleSPH_NS::calcDensity + calcDensityPost (cudaParticleSPH_NS.cu:
 * 136-207; SPHcalcDensity<SPHKernelLucy>, kerneltemplate.hh:636-656):
 * number density, rho = sum m_j W (self included), 1/rho and the field mu;
 * read back with PB_FIELD_SPH_NUM / PB_FIELD_SPH_RHOINV. */
int pb_sph_density(pb_sys *sys);

/* cudaParticleSPH_NS::calcAcceleration (cudaParticleSPH_NS.cu:71-134;
 * SPHNavierStokes<SPHKernelLucyDW>, kerneltemplate.hh:666-709): a is
 * overwritten with the pressure + viscous acceleration of the last
 * pb_sph_density's rho and mu at the current positions and velocities. */
int pb_sph_acceleration(pb_sys *sys);

/* cudaParticleSPH_NS::inspectDensity (cudaParticleSPH_NS.cu:209-217,
 * inspectDense_x kernelfuncs.cu:772-801): mean number density of the
 * movable particles. */
int pb_sph_inspect_density(pb_sys *sys, double *mean_num);

/* ------------------------------------------------------------------- DEM */
/* cudaParticleDEM::setDEMProperties, both overloads (cudaParticleDEM.cu:
 * 221-252): E_eff = E (1 - sigma^2), s2 = (2 - sigma)/2/(1 - sigma), gamma
 * (the reference's gamma2 = 2 gamma), mu, mu_r; radius r0[n] (original
 * order) or r0_all when r0 is NULL.  sigma must be < 0.5.  Allocates the
 * rotational state (w, L, T). */
int pb_dem_set_properties(pb_sys *sys, double E, double mu, double sigma, double gamma,
                          double mu_r, const double *r0, double r0_all);

/* Linear spring-dashpot contact law with a tangential spring and Coulomb
 * friction (Cundall & Strack 1979) behind the same DEM calls.  No
 * counterpart in the reference, which has only the Hertz law of
 * setDEMProperties (cudaParticleDEM.cu:221-252, kernelinline.hh:235-363):
 * it is the granular settle workload's law, set up in place of
 * setDEMProperties.  Normal force (kn delta - 2 zeta sqrt(meff kn) v_n) n
 * with zeta the damping ratio of restitution e (0 < e <= 1), a fixed
 * partner acting as an infinite mass; tangential spring kt on the
 * accumulated tangential displacement of the contact points (kept in the
 * tangent plane), capped at mu |Fn|; torque (-r_i n) x Ft.  Radius r0[n]
 * (original order) or r0_all.  The contact history becomes 3 doubles per
 * contact (pb_dem_history_width); switching the law drops the contact
 * matrices. */
int pb_dem_set_linear(pb_sys *sys, double kn, double e, double kt, double mu,
                      const double *r0, double r0_all);

/* History doubles per contact of the current law: 1 (the reference's Hertz
 * slip length, cudaSparseMat val) or 3 (the linear law's spring vector).
 * No reference counterpart (the reference has one law). */
int pb_dem_history_width(pb_sys *sys, uint32_t *hw);

/* cudaParticleRotation::setInertia, both overloads (cudaParticleRotation.cu:
 * 35-61): w = L = 0, m0inv = minv / (0.4 r0^2). */
int pb_rot_set_inertia(pb_sys *sys, const double *r0, double r0_all);

/* cudaParticleRotation::TimeEvolution (cudaParticleRotation.cu:29-33,
 * propagateEulerianRotation_F4 kernelfuncs.cu:966-986):
 * L += T dt move, w = L m0inv move. */
int pb_rot_step(pb_sys *sys, double dt);

/* cudaParticleDEM::calcForce(dt) (cudaParticleDEM.cu:29-213): contact CSR
 * of the movable particles (|r_ij| < r_i + r_j, rows and columns by
 * original index, columns ascending), tangential history carried over from
 * the previous call's matrix, Hertz contact force and torque into F and T
 * (overwritten; the linear law of pb_dem_set_linear when that was the
 * last set-up).  Needs pb_bin (calcBlockID) on the current positions with
 * cells from pb_setup_cells(sys, rmax, 0, single).  autotune = 0 makes a
 * step above the Cundall-Strack bound an error (PB_ERR_DT; the reference
 * asserts unless autotunetimestep).  nnz_out (may be NULL) gets nnz. */
int pb_dem_calc_force(pb_sys *sys, double dt, int autotune, uint32_t *nnz_out);

/* Download of the latest contact matrix (cudaSparseMat rowPtr / colIdx /
 * val, cudaSparseMat.hh): rowptr[n+1]; colidx and val need cap >= nnz
 * (either may be NULL); val holds pb_dem_history_width doubles per
 * contact. */
int pb_dem_get_contacts(pb_sys *sys, uint32_t *rowptr, uint32_t *colidx, double *val,
                        uint32_t cap, uint32_t *nnz);

/* Restores a contact matrix (restart from a checkpoint: the cudaSparseMat
 * fields of cudaParticleDEM's archive, cudaParticleDEM.hh:124-180,
 * cudaSparseMat.hh:70-122) as the latest one, so the next
 * pb_dem_calc_force carries its tangential history.  rowptr[n+1] over
 * original indices, columns ascending, val[nnz * history width]. */
int pb_dem_set_contacts(pb_sys *sys, const uint32_t *rowptr, const uint32_t *colidx,
                        const double *val, uint32_t nnz);

/* cudaParticleDEM::rollback's `_N -= 1` (cudaParticleDEM.cu:302-306): the
 * contact matrix before the latest becomes the latest again, so the next
 * pb_dem_calc_force (after the rollback) carries the tangential history of
 * the step before the discarded one.  A latest matrix with no predecessor
 * (the first call, or one restored by pb_dem_set_contacts) is dropped, and
 * the next call starts from zero history.  One step is kept: a second rewind
 * without a pb_dem_calc_force in between is PB_ERR_STATE. */
int pb_dem_rewind(pb_sys *sys);

/* Contacts of the last pb_dem_calc_force whose time step exceeded the
 * Cundall-Strack bound (counted even when autotune allowed the step). */
int pb_dem_last_violations(pb_sys *sys, uint32_t *count);

/* ------------------------------------------------------------ thermostats */
/* GaussianThermo::TimeEvolution (GaussianThermo.cu:29-83; kernels
 * kernelfuncs.cu:372-477): isokinetic velocity Verlet on (r, v, Fold = a)
 * with F(t) in F.  The implicit velocity equation the reference solves by
 * Newton-Raphson has the closed-form solution v(t) = A (1 - dt/2 g),
 * xi(t) = g / (1 - dt/2 g), g = sum F.A / sum m A.A per replica, which is
 * the loop's converged result.  tol is the loop's threshold (setE): when
 * its first pass, taken with the previous xi, moves v by an rms of at most
 * tol (always when that xi is 0) the reference stops after that pass, and
 * so does this step.  No host synchronisation.  xi of every replica
 * persists on the device (GaussianThermo::xi, starts at 0). */
int pb_gt_step(pb_sys *sys, double dt, double tol);
/* GaussianThermo::xi (GaussianThermo.hh:16): get_xi[nrep] (may be NULL)
 * receives the current values, then set_xi[nrep] (may be NULL) replaces
 * them (restart).  bad_steps (may be NULL) counts the steps since the last
 * set whose implicit equation had no solution (1 - dt/2 g <= 0; such a step
 * is taken with xi = 0, where the reference's Newton loop exits). */
int pb_gt_xi(pb_sys *sys, const double *set_xi, double *get_xi, uint32_t *bad_steps);
/* adjustVelocities (cudaParticleMD.cu:177-198, GaussianThermo.cu:98-119;
 * adjustVelocity_F4/_VD kernelfuncs.cu:546-699): per replica and axis,
 * v = (v - mean) v1 / sqrt(var); v1 = sqrt(kB T / m0) is the caller's.
 * stats (may be NULL) receives 6 per replica: the means and variances
 * before the adjustment.  A component with zero variance is an error. */
int pb_adjust_velocities(pb_sys *sys, double v1, double *stats);
/* cudaParticleMD::constTemp (cudaParticleMD.cu:140-155, correctConstTemp_F4
 * kernelfuncs.cu:533-544): lambda = sum F.v / sum m v^2 per replica,
 * F -= lambda m v.  mv2[nrep], lambda[nrep] (may be NULL) receive sum m v^2
 * and lambda (the reference returns T = mv2 / 3 N kB). */
int pb_const_temp(pb_sys *sys, double *mv2, double *lambda);

/* ------------------------------------------- MD neighbour list, annealing */
/* cudaParticleMD::makeNList (cudaParticleMD.cu:338-394; calcCoordMD_F4 and
 * makeJlist_WithBlock_F4 with MDNeighborRegister, kerneltemplate.hh:725-775,
 * kernelinline.hh:387-412): CSR list over the 27-cell grid of
 * pb_setup_cells after pb_bin, rows and columns by original index, every
 * j != i with FP32 distance2p < rmax2t (rmax2t = (sqrt(rmax2) + thickness)^2),
 * columns ascending.  nnz_out (may be NULL) receives the entry count. */
int pb_nlist_build(pb_sys *sys, float rmax2t, uint32_t *nnz_out);
/* cudaSparseMat rowPtr / colIdx of the last pb_nlist_build: rowptr[n+1];
 * colidx needs cap >= nnz (either may be NULL). */
int pb_nlist_get(pb_sys *sys, uint32_t *rowptr, uint32_t *colidx, uint32_t cap, uint32_t *nnz);
/* cudaParticleMD::calcForce with useNList (cudaParticleMD.cu:20-40,
 * calcF_IJpairWithList_F4 kerneltemplate.hh:136-168 with MDpairForce
 * kerneltemplate.hh:524-565): FP64 LJ forces (F overwritten) from the list's
 * pairs with r^2 below the cutoff of pb_lj_set_cutoff, minimum image;
 * pe, virial (may be NULL) = 1/2 sum phi, 1/2 sum f.r.  Single replica. */
int pb_lj_force_nlist(pb_sys *sys, double *pe, double *virial);
/* The soft-core force of cudaParticleMD::initialAnnealing
 * (cudaParticleMD.cu:205-283; potentialSC kernelinline.hh:140-158 through
 * MDpairForce): F_i = sum_j f0 cos(r/rc pi/2) (r_i - r_j)/r over pairs of the
 * 27-cell grid (after pb_bin) with r <= rc and r^2 < rmax2; F overwritten. */
int pb_softcore_force(pb_sys *sys, double rc, double f0, double rmax2);
/* propagateEuler_F4 (kernelfuncs.cu:174-192): r += v dt, v += a dt for
 * movable particles. */
int pb_euler_step(pb_sys *sys, double dt);

/* (no reference counterpart; diagnostic) Vector-ALU issue probe: kind 0, 1,
 * 2 or 3 selects FP32 FMAs, FP64 FMAs, 32-bit integer xor-adds or
 * compare-select-adds; eight
 * independent chains per lane for iters iterations of 64 operations, 8
 * wavefronts per SIMD.  Returns the launch's event time and the number of
 * wave64 VALU instructions it executed, so a caller (or rocprofv3 on the same
 * launch) calibrates the issue rate that counter-based VALU utilisations are
 * divided by. */
int pb_diag_valu_probe(pb_ctx *ctx, int kind, uint32_t iters, double *ms, double *wave_instr);

#ifdef __cplusplus
}
#endif
#endif
