// pb_internal.h -- shared types of the MI355X particle engine (host + device).
//
// Storage layout (HBM): every per-particle array is indexed by a storage
// slot s in [0, npad).  Each replica owns a 64-aligned slot range
// [atom0, atom0 + npad_r) holding its natoms real particles followed by
// padding slots, so every wavefront belongs to exactly one replica and all
// per-replica reductions are per-wave partials summed in a fixed order.
// Positions are double4 (x, y, z, w = type id), velocities double4
// (x, y, z, w = 1/mass): one 32-byte load brings everything the pair and
// integration kernels need about a particle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/pbamd.h"

#define PB_WAVE 64
#define PB_PAD_TYPE 0xFFFFu
#define PB_NONE 0xFFFFFFFFu
#define PB_MAXT 5

struct RepInfo {
  double lo[3], hi[3], L[3], Linv[3];
  uint32_t nb[3];      // list-cell grid of this replica (fused path)
  uint32_t cell0;      // first global list cell; the replica's pad cell is cell0 + nb0*nb1*nb2
  uint32_t atom0;      // first storage slot (multiple of 64)
  uint32_t natoms;     // real particles
  uint32_t npad;       // slots (multiple of 64)
  uint32_t wave0, nwave;
  uint32_t seg0, nseg;   // pair-kernel work units (segments) of this replica
  double rlist;          // Verlet list radius of this replica (rc + its skin)
  double half_skin2;     // (skin / 2)^2: a particle beyond it invalidates the list
  float rb2;             // FP32 list-build radius^2 (rlist + rounding margin)^2
  float inv_hs2;         // 1 / (skin / 2)^2 (0 without a skin): displacement scale
};

struct LJTable {
  int ntypes;
  double s2[PB_MAXT * PB_MAXT], e24[PB_MAXT * PB_MAXT], e4[PB_MAXT * PB_MAXT],
      shift[PB_MAXT * PB_MAXT];
  float s2f[PB_MAXT * PB_MAXT], e24f[PB_MAXT * PB_MAXT], e4f[PB_MAXT * PB_MAXT],
      shiftf[PB_MAXT * PB_MAXT];
};

struct pb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // device address of pb_bin.hip's static g(r) scratch array on this
  // context's device (pb_gr_hist_noalloc); set when the context is created
  void *gr_scratch = nullptr;
  std::string err;
  int cu_count = 0;
  std::string name;
  uint64_t global_mem = 0;
  uint32_t max_threads = 0, max_grid = 0, lds_per_block = 0;
  uint32_t nsys = 0;   // live systems created on this context (destroy refuses while > 0)
};

// device buffer helper
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
};

struct pb_sys {
  pb_ctx *ctx = nullptr;
  uint32_t nrep = 0, n = 0, npad = 0, nwave = 0;
  std::vector<uint32_t> counts, offsets;   // per replica (original order)
  std::vector<RepInfo> reps;               // host copy
  std::vector<double> cells;               // 6 per replica
  bool cell_set = false;

  // device state (storage order)
  double4 *r = nullptr, *v = nullptr, *a = nullptr, *F = nullptr;
  double4 *w = nullptr, *Lm = nullptr, *T = nullptr;   // rotation (lazy)
  double *radius = nullptr;                            // DEM (lazy)
  uint8_t *move = nullptr;
  uint16_t *type = nullptr;
  uint32_t *orig = nullptr;       // storage slot -> original index (PB_NONE: pad)
  uint32_t *waveRep = nullptr;    // wave -> replica
  RepInfo *d_reps = nullptr;
  double *partials = nullptr;     // per-wave partial sums, 4 per wave
  uint32_t *d_counter = nullptr;  // last-block counters / flags (16 words)
  double *d_out = nullptr;        // small reduction outputs (16 per replica)

  // reference cell grid (pb_setup_cells), single replica
  bool ref_cells = false;
  int ref_periodic = 1, ref_single = 0;
  uint32_t blocknum[3] = {0, 0, 0};
  float blocklen[6] = {0, 0, 0, 0, 0, 0};
  uint32_t ref_ncell = 0;
  uint32_t *ref_table = nullptr;  // ncell * 27 (or 125)
  uint32_t *ref_bid = nullptr, *ref_pid = nullptr, *ref_bindex = nullptr;
  bool ref_binned = false;

  // counting-sort workspace (shared by pb_bin and the list rebuild)
  uint32_t *cellid = nullptr, *rank = nullptr, *perm = nullptr;
  uint32_t *ccount = nullptr, *cstart = nullptr;  // ncell_cap + 1
  uint32_t ncell_cap = 0;
  uint32_t *scan_tmp = nullptr;
  size_t scan_tmp_n = 0;

  // LJ
  LJTable lj{};
  bool lj_set = false;
  double rc2 = 0.0;
  int shifted = 0;
  LJTable *d_lj = nullptr;

  // fused path: Verlet list
  uint16_t *nl = nullptr;          // window-local indices, entry k of slot s at nl[k * npad + s]
  uint32_t *nn = nullptr;          // list length per slot
  uint32_t maxnb = 0;              // list capacity
  float4 *rbuild = nullptr;
  double4 *r2buf = nullptr, *v2buf = nullptr;   // ping-pong partners
  uint32_t *orig2 = nullptr;
  uint16_t *type2 = nullptr;
  uint8_t *move2 = nullptr;
  double4 *a2buf = nullptr;
  uint32_t list_ncell = 0;
  double list_rlist = 0.0;         // largest list radius over the replicas
  uint32_t *list_segs = nullptr;   // list-build work units (8 words each)
  uint32_t list_nsegs = 0;
  uint4 *list_wtab = nullptr;      // per-segment window tables (cell order of the last sort)
  uint32_t list_wtab_cap = 0;      // segments list_wtab holds
  uint32_t tile_w = 2;             // list-cell columns per tile side (1 or 2)
  double tile_fill = 0.72;         // target fill of the LDS window stage
  // Continuation of the fused LJ run: the cell order, Verlet list and list
  // cells left by the last pb_lj_run stay valid for the next one when the
  // positions and the storage order are exactly those that run left (a
  // fingerprint of r and orig) and the list parameters are unchanged.  Any
  // sort (pb_counting_sort) clears lj_cont.
  bool lj_cont = false;
  uint64_t lj_cont_hash = 0;       // fingerprint of r, orig at the end of the run
  double lj_cont_rc2 = 0, lj_cont_fill = 0;
  uint32_t lj_cont_tile_w = 0;
  std::vector<double> lj_cont_skins;
  uint32_t lj_cont_since = 0;      // steps since the list was built
  uint32_t lj_cont_d2 = 0, lj_cont_v2 = 0;   // displacement / speed maxima (float bits)
  unsigned long long *d_hash = nullptr;
  uint32_t *lj_queue = nullptr;    // persistent step kernel: 8 work queues + exit counter
  uint32_t last_rebuilds = 0, last_maxnn = 0;
  int last_skin_violation = 0;
  double last_step_ms = 0.0, last_rebuild_ms = 0.0;
  uint32_t last_nstep_timed = 0, last_nrebuild_timed = 0;
  // host-mapped ring through which step kernels publish the previous step's
  // max displacement / speed (drives the rebuild schedule without a sync)
  uint32_t *hmon = nullptr, *dmon = nullptr;

  // Gaussian thermostat: per replica {xi, 1/c} of the last step, and the
  // count of steps whose implicit velocity equation had no solution
  double *gt = nullptr;
  uint32_t *gt_bad = nullptr;

  // MD neighbour list (cudaParticleMD::makeNList): CSR by original index
  uint32_t *md_row = nullptr, *md_col = nullptr, *md_inv = nullptr;
  uint32_t md_cap = 0, md_nnz = 0;
  bool md_valid = false;

  // g(r)
  uint64_t *gr_hist = nullptr;
  size_t gr_cap = 0;

  // SPH (cudaParticleSPH_NS), storage order unless noted
  double *sph_mu0 = nullptr, *sph_c2 = nullptr, *sph_rhoinv = nullptr;
  double4 *sph_num = nullptr;                  // (number density, rho, mu or p, 0)
  double4 *sph_rs = nullptr, *sph_vs = nullptr, *sph_ds = nullptr;   // cell-sorted views
  double sph_h = 0.0, sph_rho0 = 1.0;
  bool sph_set = false, sph_dens = false;
  // 0: reference physics (Lucy, linear EOS, mu); 1: weakly compressible
  // config physics (cubic spline, Tait, artificial viscosity)
  int sph_kind = 0;
  double sph_c0 = 0.0, sph_gamma = 7.0, sph_alpha = 0.0, sph_pmin = 0.0;

  // DEM (cudaParticleDEM, cudaParticleRotation)
  double *m0inv = nullptr;                     // minv / (0.4 r0^2), storage order
  double dem_E = 0, dem_mu = 0, dem_s2 = 0, dem_gamma2 = 0, dem_mu_r = 0;
  bool dem_set = false;
  // contact law: 0 = the reference's Hertz law (scalar history per contact),
  // 1 = linear spring-dashpot + tangential spring (3 doubles per contact)
  int dem_law = 0;
  uint32_t dem_hw = 1;                         // history doubles per contact
  double dem_kn = 0, dem_zeta = 0, dem_kt = 0;
  // contact CSR by original index, double-buffered for the tangential history
  uint32_t *cm_row[2] = {nullptr, nullptr}, *cm_col[2] = {nullptr, nullptr};
  double *cm_val[2] = {nullptr, nullptr};
  uint32_t cm_cap[2] = {0, 0}, cm_nnz[2] = {0, 0};
  int cm_cur = -1;                             // buffer of the latest CSR, -1 before any
  bool cm_prev = false;                        // the other buffer holds the matrix before it
  bool cm_rewound = false;                     // the last matrix operation was pb_dem_rewind
  uint32_t *dem_inv = nullptr;                 // original index -> storage slot
  double4 *dem_rs = nullptr;                   // cell order: (x, y, z, radius)
  uint32_t *dem_so = nullptr;                  // cell order: original index
  uint32_t *dem_rowbuf = nullptr;              // contacts kept by the counting scan
  uint32_t dem_bad = 0;                        // Cundall-Strack violations of the last call
};

// --------------------------------------------------------------- host util
int pb_fail(pb_ctx *ctx, int code, const std::string &msg);
int pb_hip_check(pb_ctx *ctx, hipError_t e, const char *what);
#define PB_HIP(ctx, call)                                              \
  do {                                                                 \
    hipError_t _e = (call);                                            \
    if (_e != hipSuccess) return pb_hip_check((ctx), _e, #call);       \
  } while (0)
#define PB_LAUNCH_CHECK(ctx, what)                                     \
  do {                                                                 \
    hipError_t _e = hipGetLastError();                                 \
    if (_e != hipSuccess) return pb_hip_check((ctx), _e, what);        \
  } while (0)

template <typename T>
int pb_alloc(pb_ctx *ctx, T **p, size_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
  if (e != hipSuccess) return pb_hip_check(ctx, e, "hipMalloc");
  return PB_OK;
}
template <typename T>
void pb_free(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static inline uint32_t pb_blocks(size_t n, uint32_t bs) {
  return (uint32_t)((n + bs - 1) / bs);
}

// counting sort helpers (pb_sort.hip)
int pb_ensure_sort_ws(pb_sys *sys, uint32_t ncell);
// exclusive scan of in[0..n) into out[0..n], out[n] = total
int pb_exclusive_scan(pb_sys *sys, const uint32_t *in, uint32_t *out, uint32_t n);
// cellid[s] given (PB_NONE never); builds perm (sorted slot -> storage slot)
// ordered by (cell, key[s]); counts/starts in sys->ccount / sys->cstart.
int pb_counting_sort(pb_sys *sys, uint32_t ncell, const uint32_t *key);
int pb_counting_sort_counted(pb_sys *sys, uint32_t ncell, const uint32_t *key);
// per-replica reduction of K per-wave partial components into out[rep*8+c]
int pb_reduce_waves(pb_sys *sys, int K);
// storage slot -> original index (PB_NONE for padding), downloaded
std::vector<uint32_t> pb_download_orig(pb_sys *sys, int *rc);
// rotation / radius state, allocated on first use
int pb_ensure_rotation(pb_sys *sys);
int pb_gr_scratch_init(pb_ctx *ctx);          // pb_bin.hip: device address of the g(r) scratch
int pb_ensure_radius(pb_sys *sys);

// ------------------------------------------------------------ device util
__device__ __forceinline__ uint32_t pb_lane() { return __lane_id(); }

// Rank of a slot among the slots of cell c counted so far: lanes holding the
// same cell as their left neighbour form a run, whose head does one
// atomicAdd on counts[c] for the whole run and broadcasts the base (one
// atomic per run when the input is nearly cell-ordered).  All lanes of the
// wavefront call it; invalid lanes pass valid = false.
__device__ __forceinline__ uint32_t pb_count_rank(uint32_t c, bool valid, uint32_t *counts) {
  const uint32_t lane = pb_lane();
  const uint32_t left = __shfl_up(c, 1, 64);
  const bool head = valid && (lane == 0 || left != c);
  const uint64_t heads = __ballot(head);
  const uint64_t validm = __ballot(valid);
  // my run head: highest head lane <= lane
  const uint64_t le = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1ull));
  const uint32_t hl = le ? (63u - (uint32_t)__clzll((long long)le)) : 0u;
  // run length (for heads): distance to the next head or the end of valid lanes
  uint32_t base = 0;
  if (head) {
    const uint64_t gt = heads & ~((lane == 63) ? ~0ull : ((2ull << lane) - 1ull));
    const uint32_t nextl = gt ? (uint32_t)__ffsll((long long)gt) - 1u : (uint32_t)__popcll(validm);
    base = atomicAdd(&counts[c], nextl - lane);
  }
  base = __shfl(base, (int)hl, 64);
  return base + (lane - hl);
}

__device__ __forceinline__ double pb_wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ uint32_t pb_wave_sum_u32(uint32_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
  return x;
}
__device__ __forceinline__ double pb_wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}

// XCD-aware logical block of a 1-D grid (MI355X dispatches workgroups
// round-robin over its 8 XCDs, each with its own L2): within every 64
// workgroups, XCD x takes 8 consecutive logical blocks (the tail past the
// last whole 64 keeps its order).  Neighbouring blocks share an XCD's L2 in eights while
// dense and sparse regions stay spread over all XCDs.  Bijective.
__device__ __forceinline__ uint32_t pb_xcd_chunk8() {
  const uint32_t b = blockIdx.x, full = gridDim.x & ~63u;
  return b < full ? (b & ~63u) + (b & 7u) * 8u + ((b >> 3) & 7u) : b;
}

__device__ __forceinline__ double pb_load_relaxed(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// sum of K per-wave partial components for every replica, executed by one
// block (used by the last-block reductions and by k_reduce_waves).
__device__ __forceinline__ void pb_block_reduce_reps(const double *partials, int K, const RepInfo *reps,
                                     uint32_t nrep, double *out, bool by_segment,
                                     double *scratch) {
  // Deterministic: every replica's partials (one per wave, or one per
  // pair-kernel segment when by_segment) are summed in a fixed order that
  // depends only on the launch shape of this block.  Callers that read
  // partials written by other workgroups issue a device-scope fence first
  // (it invalidates this CU's vector L1), so plain loads are coherent.
  // scratch: 4 x 256 doubles of the caller's LDS
  double (*sm)[256] = reinterpret_cast<double (*)[256]>(scratch);
  const uint32_t t = threadIdx.x, bs = blockDim.x;
  if (nrep >= 16) {
    for (uint32_t q = t; q < nrep; q += bs) {
      const RepInfo R = reps[q];
      const uint32_t w0 = by_segment ? R.seg0 : R.wave0, nw = by_segment ? R.nseg : R.nwave;
      double acc[4] = {0, 0, 0, 0};
      for (uint32_t w = w0; w < w0 + nw; ++w)
        for (int c = 0; c < K; ++c)
          acc[c] += partials[4 * w + c];
      for (int c = 0; c < K; ++c) out[8 * q + c] = acc[c];
    }
    return;
  }
  // the first 256 threads reduce (any block size >= 256 may call this)
  const uint32_t T = 256;
  for (uint32_t q = 0; q < nrep; ++q) {
    const uint32_t w0 = by_segment ? reps[q].seg0 : reps[q].wave0;
    const uint32_t nw = by_segment ? reps[q].nseg : reps[q].nwave;
    double acc[4] = {0, 0, 0, 0};
    if (t < T)
      for (uint32_t w = t; w < nw; w += T)
        for (int c = 0; c < K; ++c) acc[c] += partials[4 * (w0 + w) + c];
    if (t < T)
      for (int c = 0; c < K; ++c) sm[c][t] = acc[c];
    __syncthreads();
    for (uint32_t o = T / 2; o > 0; o >>= 1) {
      if (t < o)
        for (int c = 0; c < K; ++c) sm[c][t] += sm[c][t + o];
      __syncthreads();
    }
    if (t == 0)
      for (int c = 0; c < K; ++c) out[8 * q + c] = sm[c][0];
    __syncthreads();
  }
}

// applyPeriodicCondition_F4 (kernelfuncs.cu:95-129) in FP64; a value that
// rounds onto the upper face is put on the lower one.
__device__ __forceinline__ double wrap1(double x, double c0, double c1) {
  const double L = c1 - c0;
  const double d0 = c0 - x;
  if (d0 > 0) x += (double)((int)(d0 / L) + 1) * L;
  const double d1 = x - c1;
  if (d1 >= 0) x -= (double)((int)(d1 / L) + 1) * L;
  if (x >= c1) x = c0;
  return x;
}
