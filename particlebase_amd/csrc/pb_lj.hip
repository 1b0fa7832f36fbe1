// pb_lj.hip -- Lennard-Jones path of cudaParticleMD on MI355X.
//
// Reference path (cudaParticleMD.cu:24-69, kerneltemplate.hh:229-294,
// 523-565, kernelinline.hh:47-110, kernelfuncs.cu:333-367, 935-949): per
// step a cell sort, a (cell x 27) pair kernel writing 81 partial forces per
// particle to HBM, a 27-way reduction, then VV and the periodic wrap.
//
// Here: a Verlet list (cutoff rc + skin) is built from a cell grid whose
// cells are >= rc + skin; at every rebuild the particles are re-sorted into
// cell order so list neighbours are close in HBM.  Between rebuilds one
// kernel per step computes the pair forces from the list, integrates
// (velocity Verlet in its kick-drift form, FP64) and -- on energy steps --
// writes per-segment KE, PE and virial partials that a one-workgroup kernel
// sums per replica in FP64 in a fixed order (deterministic).  The 81N
// partial-force buffer, the reduce27 pass and every host synchronisation of
// the reference loop are gone.
#include "pb_internal.h"
#include <hip/hip_ext.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

int pb_ensure_sort_ws(pb_sys *s, uint32_t ncell);
int pb_counting_sort(pb_sys *s, uint32_t ncell, const uint32_t *key);
int pb_counting_sort_counted(pb_sys *s, uint32_t ncell, const uint32_t *key);
__global__ void k_fill_u32(uint32_t *x, uint32_t n, uint32_t val);

enum { CNT_LAST = 0, CNT_SKIN = 1, CNT_OVF = 2, CNT_MAXNN = 3, CNT_REC = 4, CNT_DISP = 5,
       CNT_VMAX = 6, CNT_WIN = 7 };

// ---------------------------------------------------------------- params
static void fill_table(pb_sys *s) {
  LJTable &t = s->lj;
  for (int k = 0; k < PB_MAXT * PB_MAXT; ++k) {
    t.shift[k] = 0.0;
    if (s->shifted && s->rc2 > 0 && t.s2[k] > 0) {
      double sr = t.s2[k] / s->rc2;
      sr = sr * sr * sr;
      t.shift[k] = t.e4[k] * (sr * sr - sr);
    }
    t.s2f[k] = (float)t.s2[k];
    t.e24f[k] = (float)t.e24[k];
    t.e4f[k] = (float)t.e4[k];
    t.shiftf[k] = (float)t.shift[k];
  }
}

static int upload_lj(pb_sys *s) {
  fill_table(s);
  PB_HIP(s->ctx, hipMemcpyAsync(s->d_lj, &s->lj, sizeof(LJTable), hipMemcpyHostToDevice,
                                s->ctx->stream));
  return PB_OK;
}

extern "C" int pb_lj_set_params(pb_sys *s, const double *sigma_eps, int ntypes) {
  if (!s || !sigma_eps || ntypes < 1 || ntypes > PB_MAXT) return PB_ERR_ARG;
  // _setLJparams, kernelinline.hh:120-134 (sigma^2, 24 eps), kept in FP64
  std::memset(&s->lj, 0, sizeof(LJTable));
  s->lj.ntypes = ntypes;
  for (int i = 0; i < ntypes; ++i)
    for (int j = 0; j < ntypes; ++j) {
      const double sg = sigma_eps[(i * ntypes + j) * 2], ep = sigma_eps[(i * ntypes + j) * 2 + 1];
      const int k = i * PB_MAXT + j;
      s->lj.s2[k] = sg * sg;
      s->lj.e24[k] = 24.0 * ep;
      s->lj.e4[k] = 4.0 * ep;
    }
  s->lj_set = true;
  return upload_lj(s);
}

extern "C" int pb_lj_set_cutoff(pb_sys *s, double rc2, int shifted) {
  if (!s || !(rc2 > 0)) return PB_ERR_ARG;
  s->rc2 = rc2;
  s->shifted = shifted ? 1 : 0;
  return upload_lj(s);
}

// ------------------------------------------------------------ list cells
__device__ __forceinline__ double wrap_pos(double x, double lo, double L, double Linv) {
  double t = x - lo;
  t -= L * floor(t * Linv);
  if (t >= L) t -= L;
  if (t < 0.0) t = 0.0;
  double y = lo + t;
  if (y >= lo + L) y = lo;
  return y;
}

__device__ __forceinline__ uint32_t cell_axis(double x, double lo, uint32_t nb, double Linv) {
  const double q = (x - lo) * ((double)nb * Linv);
  uint32_t d = (q > 0.0) ? (uint32_t)q : 0u;
  return d >= nb ? nb - 1 : d;
}

__global__ void k_list_cellid(const double4 *r, const uint32_t *orig, const uint32_t *waveRep,
                              const RepInfo *reps, uint32_t n, uint32_t *cellid,
                              uint32_t *counts, uint32_t *rank) {
  // the list cell of slot s (padding: the replica's pad cell), counted in
  // the same pass (the counting sort's k_count)
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = s < n;
  uint32_t c = PB_NONE;
  if (valid) {
    const RepInfo &R = reps[waveRep[s / PB_WAVE]];
    const uint32_t ncr = R.nb[0] * R.nb[1] * R.nb[2];
    if (orig[s] == PB_NONE) {
      c = R.cell0 + ncr;
    } else {
      const double4 p = r[s];
      const double x = wrap_pos(p.x, R.lo[0], R.L[0], R.Linv[0]);
      const double y = wrap_pos(p.y, R.lo[1], R.L[1], R.Linv[1]);
      const double z = wrap_pos(p.z, R.lo[2], R.L[2], R.Linv[2]);
      const uint32_t ci = cell_axis(x, R.lo[0], R.nb[0], R.Linv[0]);
      const uint32_t cj = cell_axis(y, R.lo[1], R.nb[1], R.Linv[1]);
      const uint32_t ck = cell_axis(z, R.lo[2], R.nb[2], R.Linv[2]);
      c = R.cell0 + ck + R.nb[2] * (cj + R.nb[1] * ci);
    }
    cellid[s] = c;
  }
  const uint32_t k = pb_count_rank(c, valid, counts);
  if (valid) rank[s] = k;
}

struct ReorderArgs {
  const uint32_t *perm;
  const double4 *r, *v, *a;
  const uint32_t *orig;
  const uint16_t *type;
  const uint8_t *move;
  double4 *r2, *v2, *a2;
  uint32_t *orig2;
  uint16_t *type2;
  uint8_t *move2;
  float4 *rbuild;
  uint32_t *nn;
  const uint32_t *waveRep;
  const RepInfo *reps;
  uint32_t n;
};

// gather every per-particle array into cell order (positions wrapped)
__global__ void k_reorder(ReorderArgs A) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  const uint32_t s = A.perm[k];
  double4 p = A.r[s];
  const uint32_t o = A.orig[s];
  if (o != PB_NONE) {
    const RepInfo &R = A.reps[A.waveRep[k / PB_WAVE]];
    p.x = wrap_pos(p.x, R.lo[0], R.L[0], R.Linv[0]);
    p.y = wrap_pos(p.y, R.lo[1], R.L[1], R.Linv[1]);
    p.z = wrap_pos(p.z, R.lo[2], R.L[2], R.Linv[2]);
  }
  A.r2[k] = p;
  A.v2[k] = A.v[s];
  if (A.a) A.a2[k] = A.a[s];
  A.orig2[k] = o;
  A.type2[k] = A.type[s];
  A.move2[k] = A.move[s];
  A.rbuild[k] = make_float4((float)p.x, (float)p.y, (float)p.z, 0.f);
  if (o == PB_NONE) A.nn[k] = 0;   // padding never gets a list
}

// ---------------------------------------------------------------- tiles
// Work unit of the list build and of the pair kernel ("segment"): a tile of
// tx x ty cell columns x nz cells along z of one replica.  Its window -- the
// tile grown by one list cell on every side -- is staged in LDS with the
// periodic image shift of every window cell already applied, so a pair is
// x_i - x_j with no minimum-image arithmetic.  The Verlet list stores
// window-local indices.  A replica with fewer than 3 list cells along an
// axis is one "whole" segment: every particle is staged once and pairs use
// the minimum image.
struct ListSeg {
  uint32_t rep, ci, cj, ck0, nz, tx, ty, whole;
};
// list-build workgroup: 8 waves.  A segment holds about as many tile
// particles as the step kernel's 512 lanes take in one pass (415 on the 1M
// fluid), one particle a lane here as well: with 384 lanes the build made a
// second pass for the last few particles of most segments (340 us at 1M
// atoms; 448 lanes: 326 us; 512: 288 us)
#ifndef TB
#define TB 512
#endif
#ifndef BUILD_GROUP
#define BUILD_GROUP 32u   // candidates tested per mask step of the list build (<= 32)
#endif
#ifndef BUILD_UNROLL
#define BUILD_UNROLL 4
#endif
#define TBS 512           // step workgroup (8 waves; two per CU at <= 128 VGPRs)
#ifndef WCAP
#define WCAP 2816         // staged window particles per workgroup (LDS)
#endif
#define MAXCOL 16         // (tx+2)*(ty+2) window columns, tx, ty <= 2
#define MAXZ 16           // nz <= MAXZ
#define MAXE (MAXCOL * (MAXZ + 2))

// entry table of a window: entry e = (window column, z slot) -> LDS offset,
// storage start, count, packed image shift; plus the runs of tile particles
struct alignas(16) WinTable {
  uint32_t eoff[MAXE + 1], eg0[MAXE], ec[MAXE], esh[MAXE];
  uint32_t nent, total, ncol, zs;
  uint32_t nrun, rloc[4], rpref[5], rg0[4], rcol[4];
  uint32_t wsum[TBS / 64];
};
// Two list formats, chosen per segment by the same test in the build and the
// step kernel: segments whose window is staged in LDS and that need no
// minimum image (the fast path) store 12-bit entries, five to a word; the
// others (windows larger than the stage, whole-replica segments) store
// 16-bit entries, four to a word.
// Padding entries: WCAP in 12-bit rows (the inert LDS slot, so the step
// kernel uses every entry as it stands), all ones in 16-bit rows (tested
// for).  A batch of the step kernel is one word per lane: 5 * LPA or 4 * LPA
// entries; rows are padded to whole batches.
#define NL_PAD 0xFFFFu     // 16-bit list padding entry
#define NL_PADW12 ((uint64_t)WCAP * 0x1001001001001ull)   // a 12-bit word of five WCAP entries
static_assert(WCAP < 0xFFF, "12-bit list entries address the staged window and its inert slot");
// Step kernel: lanes per particle, each taking every LPA-th list word.  One
// lane (64 particles a wavefront) measured 125 us a step at 1M atoms
// against 131 us with two lanes, 151 with four and 178 with eight: rows
// are padded to 5 * LPA entries, and with fewer lanes to a row less of each
// wavefront idles on padding, more than making up for the wavefront
// running its longest row.
#ifndef LPA
#define LPA 1u
#endif
#define NL_ROW (20u * LPA) // list capacity granule: a multiple of both batch sizes
#define PAD_FAR 1.0e10     // coordinate of that slot
struct Stage {              // pair kernel: FP64 positions with shifts applied
  double x[WCAP + 1], y[WCAP + 1], z[WCAP + 1];   // slot WCAP: padding target
  uint8_t t[WCAP + 1];
  alignas(16) WinTable W;
  unsigned long long hsum[16];  // last step of a run: fingerprint sums (lane mod 16, so
                                // four lanes of a wavefront share an LDS atomic address)
};
#define WCAPB WCAP          // build kernel stage (FP32, 12 B a particle): four workgroups per CU
struct StageB {             // list build: FP32 window-relative positions, 12 B a
  float2 xy[WCAPB + 32];     // particle (four workgroups per CU); 32 slack
  float z[WCAPB + 32];       // entries read by the last group of 32 tests
  alignas(16) WinTable W;
};

// exclusive block scan (blockDim multiple of 64, <= TBS)
__device__ __forceinline__ uint32_t tile_scan(uint32_t x, uint32_t *wsum, uint32_t *total) {
  const uint32_t lane = pb_lane(), w = threadIdx.x / 64, nw = blockDim.x / 64;
  uint32_t inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (uint32_t k = 0; k < nw; ++k) {
    if (k < w) off += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return off + inc - x;
}

// window entry e -> (cell, image shift); shift packed as 3 x 2 bits (+1 / -1)
__device__ __forceinline__ void entry_cell(const ListSeg &G, const RepInfo &R, uint32_t e,
                                           uint32_t zs, uint32_t &J, uint32_t &sh) {
  const int nb0 = (int)R.nb[0], nb1 = (int)R.nb[1], nb2 = (int)R.nb[2];
  const uint32_t col = e / zs, t = e % zs;
  const int wy = (int)G.ty + 2;
  const int ii = (int)G.ci - 1 + (int)(col / wy), jj = (int)G.cj - 1 + (int)(col % wy);
  const int kk = (int)G.ck0 - 1 + (int)t;
  const int sx = ii < 0 ? -1 : (ii >= nb0 ? 1 : 0);
  const int sy = jj < 0 ? -1 : (jj >= nb1 ? 1 : 0);
  const int sz = kk < 0 ? -1 : (kk >= nb2 ? 1 : 0);
  J = R.cell0 + (uint32_t)((kk - sz * nb2) + nb2 * ((jj - sy * nb1) + nb1 * (ii - sx * nb0)));
  sh = (uint32_t)(sx + 1) | ((uint32_t)(sy + 1) << 2) | ((uint32_t)(sz + 1) << 4);
}

// Entry table of the segment's window (all threads of the block call it).
__device__ void window_table(WinTable &W, const ListSeg &G, const RepInfo &R,
                             const uint32_t *cstart) {
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    W.ncol = G.whole ? 1u : (G.tx + 2) * (G.ty + 2);
    W.zs = G.whole ? 1u : G.nz + 2;
    W.nent = W.ncol * W.zs;
  }
  __syncthreads();
  uint32_t c = 0;
  if (tid < W.nent) {
    uint32_t g0, sh = 0x15;   // no shift
    if (G.whole) {
      g0 = R.atom0;
      c = R.natoms;
    } else {
      uint32_t J;
      entry_cell(G, R, tid, W.zs, J, sh);
      g0 = cstart[J];
      c = cstart[J + 1] - g0;
    }
    W.eg0[tid] = g0;
    W.ec[tid] = c;
    W.esh[tid] = sh;
  }
  uint32_t tot;
  const uint32_t off = tile_scan(c, W.wsum, &tot);
  if (tid < W.nent) W.eoff[tid] = off;
  __syncthreads();                   // thread 0 reads other threads' offsets below
  if (tid == 0) {
    W.eoff[W.nent] = tot;
    W.total = tot;
    // runs of the tile's own particles: inner columns, inner z slots
    if (G.whole) {
      W.nrun = 1;
      W.rloc[0] = 0;
      W.rg0[0] = R.atom0;
      W.rcol[0] = 0;
      W.rpref[0] = 0;
      W.rpref[1] = tot;
    } else {
      uint32_t n = 0, pref = 0;
      for (uint32_t a = 0; a < G.tx; ++a)
        for (uint32_t b = 0; b < G.ty; ++b) {
          const uint32_t col = (a + 1) * (G.ty + 2) + (b + 1);
          const uint32_t e0 = col * W.zs + 1;
          W.rloc[n] = W.eoff[e0];
          W.rg0[n] = W.eg0[e0];
          W.rcol[n] = col;
          W.rpref[n] = pref;
          pref += W.eoff[e0 + G.nz] - W.eoff[e0];
          ++n;
        }
      W.nrun = n;
      W.rpref[n] = pref;
    }
  }
  __syncthreads();
}

// The window tables depend only on the segments and the cell starts of the
// last sort, so the list build, which builds them, also stores them (one
// per segment) and each step workgroup copies its table into LDS with one
// coalesced load round instead of reading the segment descriptor and then
// its cells' starts (two dependent rounds) and scanning.
static_assert(sizeof(WinTable) % 16 == 0, "window tables are copied 16 bytes a thread");
constexpr uint32_t WTAB_U4 = sizeof(WinTable) / 16;

__device__ __forceinline__ void store_window_table(const WinTable &W, uint4 *tabs, uint32_t seg) {
  const uint4 *src = reinterpret_cast<const uint4 *>(&W);
  uint4 *dst = tabs + (size_t)seg * WTAB_U4;
  for (uint32_t i = threadIdx.x; i < WTAB_U4; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ void load_window_table(WinTable &W, const uint4 *tabs, uint32_t seg) {
  const uint4 *src = tabs + (size_t)seg * WTAB_U4;
  uint4 *dst = reinterpret_cast<uint4 *>(&W);
  for (uint32_t i = threadIdx.x; i < WTAB_U4; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

// entry holding window-local particle q (largest e with eoff[e] <= q)
__device__ __forceinline__ uint32_t window_entry(const WinTable &W, uint32_t q) {
  uint32_t lo = 0, hi = W.nent;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (W.eoff[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

// image-shifted FP64 position of window particle q read from global memory
__device__ __forceinline__ double4 window_global(const WinTable &W, const ListSeg &G,
                                                 const RepInfo &R, const double4 *r, uint32_t q) {
  const uint32_t e = window_entry(W, q);
  const uint32_t sh = W.esh[e];
  double4 p = r[W.eg0[e] + (q - W.eoff[e])];
  p.x += (double)((int)(sh & 3u) - 1) * R.L[0];
  p.y += (double)((int)((sh >> 2) & 3u) - 1) * R.L[1];
  p.z += (double)((int)((sh >> 4) & 3u) - 1) * R.L[2];
  return p;
}

// Copy the window into LDS (image shifts applied), calling f(q, p) for
// window-local particle q.  The copy is latency bound.  The cells of a
// window column are consecutive in storage except across the periodic wrap
// in z, so a column is one or two contiguous runs: with columns of
// STAGE_RUN_MIN particles or more on average, one wavefront copies a
// column, 64 particles a load and up to four loads of a lane in flight
// before their stores (the 1M-atom fluid's step kernel, 153 particles a
// column: 125.3 -> 121.5 us).  Windows of shorter columns (the 256 batched
// replicas, about 100 a column, whose columns also wrap more often) are
// copied entry by entry, 16 lanes to an entry and four entries a wavefront
// (the run form was 3 % slower there).  A whole-replica window (one entry)
// is copied by all threads.
#ifndef STAGE_RUN_MIN
#define STAGE_RUN_MIN 128u
#endif
template <typename F>
__device__ __forceinline__ void for_window_particles(const WinTable &W, const RepInfo &R,
                                                     const double4 *r, F &&f) {
  constexpr uint32_t U = 4;
  auto copy = [&](uint32_t src, uint32_t dst, uint32_t n, uint32_t sh, uint32_t t0,
                  uint32_t stride) {
    const double sx = (double)((int)(sh & 3u) - 1) * R.L[0];
    const double sy = (double)((int)((sh >> 2) & 3u) - 1) * R.L[1];
    const double sz = (double)((int)((sh >> 4) & 3u) - 1) * R.L[2];
    for (uint32_t k0 = t0; k0 < n; k0 += U * stride) {
      double4 p[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t k = k0 + u * stride;
        if (k < n) p[u] = r[src + k];
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t k = k0 + u * stride;
        if (k < n) f(dst + k, make_double4(p[u].x + sx, p[u].y + sy, p[u].z + sz, p[u].w));
      }
    }
  };
  if (W.nent == 1) {                       // whole replica
    copy(W.eg0[0], W.eoff[0], W.ec[0], W.esh[0], threadIdx.x, blockDim.x);
    return;
  }
  const uint32_t lane = pb_lane(), wave = threadIdx.x / 64, nwv = blockDim.x / 64;
  if (W.total < STAGE_RUN_MIN * W.ncol) {   // block-uniform: short columns, entry-wise
    const uint32_t g = lane >> 4, l16 = lane & 15u;
    for (uint32_t e0 = wave * 4; e0 < W.nent; e0 += nwv * 4) {
      const uint32_t e = e0 + g;
      if (e >= W.nent) continue;
      const uint32_t c = W.ec[e], src = W.eg0[e], dst = W.eoff[e], sh = W.esh[e];
      const double sx = (double)((int)(sh & 3u) - 1) * R.L[0];
      const double sy = (double)((int)((sh >> 2) & 3u) - 1) * R.L[1];
      const double sz = (double)((int)((sh >> 4) & 3u) - 1) * R.L[2];
      for (uint32_t k = l16; k < c; k += 16) {
        const double4 p = r[src + k];
        f(dst + k, make_double4(p.x + sx, p.y + sy, p.z + sz, p.w));
      }
    }
    return;
  }
  const uint32_t zs = W.zs;                // entries of a column (<= MAXZ + 2 < 64)
  for (uint32_t col = wave; col < W.ncol; col += nwv) {   // wave-uniform
    const uint32_t e0 = col * zs;
    // entry e0 + t starts a run: the column's first, or not contiguous in
    // storage with the entry before it, or another image shift
    bool start = false;
    if (lane < zs) {
      const uint32_t e = e0 + lane;
      start = lane == 0 || W.esh[e] != W.esh[e - 1] || W.eg0[e] != W.eg0[e - 1] + W.ec[e - 1];
    }
    uint64_t runs = __ballot(start) | (1ull << zs);
    while (true) {
      const uint32_t t0 = (uint32_t)__builtin_ctzll(runs);
      if (t0 >= zs) break;
      runs &= runs - 1ull;
      const uint32_t t1 = (uint32_t)__builtin_ctzll(runs);
      const uint32_t ea = e0 + t0, eb = e0 + t1;
      copy(W.eg0[ea], W.eoff[ea], W.eoff[eb] - W.eoff[ea], W.esh[ea], lane, 64u);
    }
  }
}

// Stage the window in FP64 (pair kernel) when it fits.
// (the window table is built before: the step kernel issues its first list
// and state loads between the two)
__device__ void stage_window(Stage &S, const RepInfo &R, const double4 *r) {
  if (S.W.total <= WCAP)
    for_window_particles(S.W, R, r, [&](uint32_t q, const double4 &p) {
      S.x[q] = p.x;
      S.y[q] = p.y;
      S.z[q] = p.z;
      S.t[q] = (uint8_t)p.w;
    });
  if (threadIdx.x == 0) {            // padding target: beyond any cutoff
    S.x[WCAP] = S.y[WCAP] = S.z[WCAP] = PAD_FAR;
    S.t[WCAP] = 0;
  }
  __syncthreads();
}

// position of window-local particle q (from LDS, or from global memory when
// the window did not fit)
__device__ __forceinline__ double4 window_pos(const Stage &S, const ListSeg &G, const RepInfo &R,
                                              const double4 *r, bool staged, uint32_t q) {
  if (staged) {
    q = q == NL_PAD ? WCAP : q;
    return make_double4(S.x[q], S.y[q], S.z[q], (double)S.t[q]);
  }
  if (q == NL_PAD) return make_double4(PAD_FAR, PAD_FAR, PAD_FAR, 0);
  return window_global(S.W, G, R, r, q);
}

template <uint32_t V> struct UConst { static constexpr uint32_t value = V; };

// tile particle u (0 <= u < rpref[nrun]) -> window-local index, storage slot, run
__device__ __forceinline__ void tile_particle(const WinTable &W, uint32_t u, uint32_t &li,
                                              uint32_t &gi, uint32_t &run) {
  uint32_t k = 0;
  while (k + 1 < W.nrun && W.rpref[k + 1] <= u) ++k;
  run = k;
  li = W.rloc[k] + (u - W.rpref[k]);
  gi = W.rg0[k] + (u - W.rpref[k]);
}

// Verlet-list build: the window is staged in LDS as FP32 coordinates
// relative to the window's lower corner (image shifts applied; |x| < 4 list
// cells, so FP32 keeps ~1e-6 absolute); every tile particle scans its 27
// neighbour cells (all staged particles in whole mode) and keeps
// window-local q with |r_iq| < rb, where rb = rc + skin + a margin covering
// the FP32 rounding: the list is a superset of the FP64 pairs within
// rc + skin.  A window larger than the stage reads FP64 from global memory;
// a window of more than 65534 particles raises CNT_WIN.
// Layout: the entries of a slot are grouped in words of 8 bytes (five 12-bit
// entries in staged segments, four 16-bit entries otherwise), and the words
// in pairs: words 2j and 2j+1 of slot gi are the 16 bytes at pair j * npad +
// gi, so a lane of the step kernel loads two words at a time and the
// particles of a wavefront read adjacent pairs.  Rows are padded with the
// format's inert entry to whole word pairs (nn holds the padded length).
// 12-bit entries move 20 % fewer list bytes per step than 16-bit ones
// (measured: the step kernel 159 -> 151 us at 1M atoms).
// eight wavefronts per SIMD (64 VGPRs; the few values spilled are per
// particle, outside the candidate loops): with the 38.6 KB stage, four
// 512-thread workgroups per CU (six per SIMD, three workgroups: 297 us)
#ifndef BUILD_WPE
#define BUILD_WPE 8
#endif
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(BUILD_WPE))) k_build_list(const double4 *r, const uint32_t *cstart,
                                                   const ListSeg *segs, const RepInfo *reps,
                                                   uint32_t npad, uint32_t maxnb, uint16_t *nl,
                                                   uint32_t *nn, uint32_t *counter, uint4 *wtab) {
  __shared__ StageB S;
  const uint32_t seg = pb_xcd_chunk8();
  const ListSeg G = segs[seg];
  const RepInfo &R = reps[G.rep];
  WinTable &W = S.W;
  window_table(W, G, R, cstart);
  store_window_table(W, wtab, seg);              // for the step kernels
  if (W.total >= NL_PAD) {                       // block-uniform
    if (threadIdx.x == 0) atomicOr(&counter[CNT_WIN], 1u);
    for (uint32_t u = threadIdx.x; u < W.rpref[W.nrun]; u += blockDim.x) {
      uint32_t li, gi, run;
      tile_particle(W, u, li, gi, run);
      nn[gi] = 0;                                  // no stale rows survive
    }
    return;
  }
  const bool staged = W.total <= WCAPB;
  const bool fmt12 = staged && !G.whole;        // the step kernel's fast path
  const double rlist2 = R.rlist * R.rlist;
  const float rb2 = R.rb2;
  // window origin (FP64): lower corner of the first window cell
  const double ox = R.lo[0] + ((double)G.ci - 1.0) * R.L[0] / R.nb[0];
  const double oy = R.lo[1] + ((double)G.cj - 1.0) * R.L[1] / R.nb[1];
  const double oz = R.lo[2] + ((double)G.ck0 - 1.0) * R.L[2] / R.nb[2];
  if (staged)
    for_window_particles(W, R, r, [&](uint32_t q, const double4 &p) {
      S.xy[q] = make_float2((float)(p.x - ox), (float)(p.y - oy));
      S.z[q] = (float)(p.z - oz);
    });
  __syncthreads();
  const uint32_t ntile = W.rpref[W.nrun];
  const float Lx = (float)R.L[0], Ly = (float)R.L[1], Lz = (float)R.L[2];
  const float Ix = (float)R.Linv[0], Iy = (float)R.Linv[1], Iz = (float)R.Linv[2];
  uint32_t mymax = 0;
  // the format as compile-time constants (entries per word, bits per
  // entry): the per-hit word arithmetic then has no division
  auto tile_rows = [&](auto fmt) {
  constexpr uint32_t epw = decltype(fmt)::value, eb = epw == 5u ? 12u : 16u;
  constexpr uint64_t padw = epw == 5u ? NL_PADW12 : ~0ull;   // a word of padding entries
  for (uint32_t u = threadIdx.x; u < ntile; u += blockDim.x) {
    uint32_t li, gi, run;
    tile_particle(W, u, li, gi, run);
    // Kept entries are packed epw to a 64-bit word in registers; word m of
    // slot gi is written whole at waddr(m), for m < maxnb / epw (the
    // allocation holds maxnb + 48 entries per slot: slack words for the
    // step kernel's read one word pair ahead).
    uint64_t *const wrow = reinterpret_cast<uint64_t *>(nl) + 2u * gi;
    auto waddr = [&](uint32_t m) -> uint64_t * {
      return wrow + (size_t)(m >> 1) * 2u * npad + (m & 1u);
    };
    const uint32_t wcap = maxnb / epw;
    // 12-bit format: an entry enters the word at bit 0 and the word moves up
    // 12 bits per append, behind a marker bit that starts at bit 0; after
    // five appends the marker reaches bit 60, above the five fields (readers
    // mask the fields), and the word is complete: one compare of the high
    // half instead of an entry counter.  16-bit format (no spare bit):
    // entries enter at the top, move down 16 bits, and a counter completes
    // the word.  The store pointer advances one word per completed word;
    // words past the capacity are not stored (an overflow is reported and
    // the list rebuilt larger).  cnt is advanced per group of tests
    // (popcount).
    uint64_t word = epw == 5u ? 1ull : 0ull;
    uint32_t cnt = 0, kw = 0, odd = 0;
    uint64_t *wp = wrow;
    const uint64_t *const wend = waddr(wcap);
    auto store_word = [&]() {
      if (wp < wend) *wp = word;
      wp += odd ? 2u * npad - 1u : 1u;
      odd ^= 1u;
    };
    auto append = [&](uint32_t q) {
      if constexpr (epw == 5u) {
        word = (word << 12) + q;                 // the low field is zero after the shift
        if ((uint32_t)(word >> 32) >= (1u << 28)) {
          store_word();
          word = 1ull;
        }
      } else {
        word = (word >> eb) | ((uint64_t)q << (eb * (epw - 1u)));
        if (++kw == epw) {
          store_word();
          kw = 0;
        }
      }
    };
    // staged candidates [q0, q1) except li: 32 distance tests into a bit
    // mask, then the hits appended in order (the append runs once per hit,
    // not once per candidate).  The last group reads up to 31 entries past
    // q1 (the stage's slack) and li's own test is kept; their bits are
    // cleared, so every group runs the same unrolled tests.
    auto scan_staged = [&](uint32_t q0, uint32_t q1, bool mi) {
      const float2 pixy = S.xy[li];
      const float piz = S.z[li];
      // |r_iq|^2 - rb2 as raw bits: the sign bit is set when q is near
      auto near = [&](uint32_t q) {
        const float2 pj = S.xy[q];
        float dx = pixy.x - pj.x, dy = pixy.y - pj.y, dz = piz - S.z[q];
        if (mi) {
          dx -= Lx * rintf(dx * Ix);
          dy -= Ly * rintf(dy * Iy);
          dz -= Lz * rintf(dz * Iz);
        }
        return __float_as_uint(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, -rb2))));
      };
      for (uint32_t c0 = q0; c0 < q1; c0 += BUILD_GROUP) {
        // test k shifts in at bit 0, so candidate c0 + k ends at bit
        // BUILD_GROUP - 1 - k: alignbit(m, t, 31) = (m << 1) | (t >> 31),
        // one instruction a candidate for the compare and the mask
        uint32_t m = 0;
#pragma unroll BUILD_UNROLL
        for (uint32_t k = 0; k < BUILD_GROUP; ++k) m = __builtin_amdgcn_alignbit(m, near(c0 + k), 31);
        constexpr uint32_t top = 1u << (BUILD_GROUP - 1u);
        if (q1 - c0 < BUILD_GROUP) m &= ~((top >> (q1 - c0 - 1u)) - 1u);
        if (li - c0 < BUILD_GROUP) m &= ~(top >> (li - c0));   // li < c0 wraps past the group
        cnt += __popc(m);
        while (m) {
          const uint32_t k = (uint32_t)__builtin_clz(m) - (32u - BUILD_GROUP);
          append(c0 + k);
          m &= ~(top >> k);
        }
      }
    };
    // window read from global memory (window larger than the stage)
    auto scan_global = [&](uint32_t q0, uint32_t q1, bool mi) {
      const double4 pi = window_global(W, G, R, r, li);
      for (uint32_t q = q0; q < q1; ++q) {
        const double4 pj = window_global(W, G, R, r, q);
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        if (mi) {
          dx -= R.L[0] * rint(dx * R.Linv[0]);
          dy -= R.L[1] * rint(dy * R.Linv[1]);
          dz -= R.L[2] * rint(dz * R.Linv[2]);
        }
        if (dx * dx + dy * dy + dz * dz < rlist2) {
          ++cnt;
          append(q);
        }
      }
    };
    // candidates [q0, q1) except the particle itself
    auto scan = [&](uint32_t q0, uint32_t q1, bool mi) {
      if (staged) {
        scan_staged(q0, q1, mi);
      } else if (li >= q0 && li < q1) {
        scan_global(q0, li, mi);
        scan_global(li + 1, q1, mi);
      } else {
        scan_global(q0, q1, mi);
      }
    };
    if (G.whole) {
      scan(0, W.total, true);
    } else {
      // z slot of this particle: the inner entry of its column holding li
      const uint32_t col = W.rcol[run];
      uint32_t ti = 1;
      while (ti < G.nz && W.eoff[col * W.zs + ti + 1] <= li) ++ti;
      const uint32_t wy = G.ty + 2;
      const uint32_t ca = col / wy, cb = col % wy;
      for (uint32_t a = ca - 1; a <= ca + 1; ++a)
        for (uint32_t b = cb - 1; b <= cb + 1; ++b) {
          const uint32_t e0 = (a * wy + b) * W.zs + ti - 1;
          scan(W.eoff[e0], W.eoff[e0 + 3], false);   // three z slots
        }
    }
    // close the partial word and pad to whole batches with the inert entry
    // (nn holds the padded length)
    const uint32_t used = cnt < maxnb ? cnt : maxnb;
    const uint32_t batch = 2u * epw * LPA;      // whole word pairs
    const uint32_t padded = (used + batch - 1u) / batch * batch;
    if (cnt % epw != 0u && cnt < maxnb) {
      // the partial word: its kr entries moved to the word's top fields, the
      // rest filled with the inert entry
      const uint32_t kr = cnt % epw;
      uint64_t last;
      if constexpr (epw == 5u) {
        const uint32_t sh = 12u * (5u - kr);
        last = (word << sh) | (padw & ((1ull << sh) - 1ull));
      } else {
        last = (word >> (eb * (epw - kr))) | (padw & (~0ull << (eb * kr)));
      }
      *waddr(cnt / epw) = last;
    }
    for (uint32_t m = (used + epw - 1u) / epw; m < padded / epw; ++m) *waddr(m) = padw;
    nn[gi] = padded;
    if (cnt > maxnb) atomicMax(&counter[CNT_OVF], cnt);
    mymax = max(mymax, cnt);
  }
  };
  if (fmt12) tile_rows(UConst<5u>{});
  else tile_rows(UConst<4u>{});
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mymax = max(mymax, (uint32_t)__shfl_xor((int)mymax, o, 64));
  if (pb_lane() == 0 && mymax) atomicMax(&counter[CNT_MAXNN], mymax);
}

// Fingerprint of the positions and the storage order (what the Verlet list
// and the cell order depend on): a wrapping sum over slots of a 64-bit mix of
// the slot index, the raw bits of r and orig; order-independent, so the
// atomics leave it deterministic.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t slot_hash(uint32_t s, uint32_t o, const double4 &p) {
  uint64_t h = mix64((uint64_t)s * 0x9E3779B97F4A7C15ull ^ (uint64_t)o);
  h = mix64(h ^ (uint64_t)__double_as_longlong(p.x));
  h = mix64(h ^ (uint64_t)__double_as_longlong(p.y));
  h = mix64(h ^ (uint64_t)__double_as_longlong(p.z));
  return mix64(h ^ (uint64_t)__double_as_longlong(p.w));
}
// The sum is spread over HASH_SLOTS addresses (workgroup b adds to slot
// b mod HASH_SLOTS) and the host adds the slots: atomics on one address
// serialise in its L2 channel.  Any workgroup size up to 1024.
#define HASH_SLOTS 64u
__device__ __forceinline__ void hash_flush(uint64_t acc, unsigned long long *out) {
  __shared__ uint64_t part[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += (uint64_t)__shfl_xor((long long)acc, o, 64);
  if (pb_lane() == 0) part[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += part[w];
    if (t) atomicAdd(out + blockIdx.x % HASH_SLOTS, (unsigned long long)t);
  }
}

// ------------------------------------------------------------- step kernel
struct StepArgs {
  const double4 *rin, *vin;
  double4 *rout, *vout, *Fout;
  const uint64_t *nl;       // 5 12-bit or 4 16-bit entries per word (see k_build_list)
  const uint32_t *nn;
  const float4 *rbuild;
  const uint32_t *cstart;
  const uint4 *wtab;        // per-segment window tables
  const ListSeg *segs;
  const RepInfo *reps;
  const LJTable *lj;
  double *partials;
  uint32_t *counter;
  double *energy;       // records of 8*nrep doubles
  uint32_t *mon;        // host-mapped ring (8 slots x 4 words) or null
  uint32_t *queue;      // persistent form: 8 work-queue counters + exit counter (else null)
  uint32_t nsegs;       // segments (work units) of the launch
  uint32_t step;        // index of this step inside the run
  uint32_t npad, nrep;
  double dt, rc2;
  // single-type parameters
  double s2, e24, e4, shift;
  float s2f, e24f, e4f, shiftf;
  // last step of a run
  double4 *Aout;            // F(t) again (the VV state's Fold)
  const uint32_t *orig;     // slot -> particle (fingerprint)
  unsigned long long *hash; // HASH_SLOTS fingerprint sums
};

__device__ __forceinline__ double rcp64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(fma(-x, y, 1.0), y, y);
  y = fma(fma(-x, y, 1.0), y, y);
  return y;
}

// Copy the displacement / speed maxima of the kernels so far (the running
// maxima since the last rebuild) into the host ring's slot step - 1, marked
// with `step`.
__device__ __forceinline__ void publish_maxima(const uint32_t *counter, uint32_t *mon,
                                               uint32_t step) {
  const uint32_t slot = ((step - 1) & 7u) * 4u;
  const uint32_t d = __hip_atomic_load(&counter[CNT_DISP], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t v = __hip_atomic_load(&counter[CNT_VMAX], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t k = __hip_atomic_load(&counter[CNT_SKIN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&mon[slot + 0], d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&mon[slot + 1], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&mon[slot + 2], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&mon[slot + 3], step, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Publish the previous step's maxima to the host ring.  Runs in workgroup 0
// when this kernel starts, i.e. after the previous step's kernel has
// finished (stream order): the host reads it without a synchronisation.
__device__ __forceinline__ void publish_prev(const StepArgs &A) {
  if (!A.mon || A.step == 0 || blockIdx.x != 0 || threadIdx.x != 0) return;
  publish_maxima(A.counter, A.mon, A.step);
}

// The same report on demand, launched behind step kernel `step - 1` when the
// host needs the displacement of the positions step `step` will use before
// it decides on a rebuild (the report the kernel of step `step` would make).
__global__ void k_publish_maxima(const uint32_t *counter, uint32_t *mon, uint32_t step) {
  if (threadIdx.x == 0) publish_maxima(counter, mon, step);
}

// Pairs and integration of one segment's tile particles: one lane a
// particle (LPA = 1), 64 particles a wavefront.  With FAST (window staged,
// no minimum image) a lane takes a word pair (ten pairs) a step, the next
// pair's 16-byte load in flight while they are evaluated; rows are padded
// with the inert LDS slot, so a pair is a straight-line sequence and the
// cutoff is a select.  Other segments take one word (four pairs) a step.
// (LPA > 1 lanes a particle would sum partial forces over the lanes and
// let the first lane integrate.)
template <bool FAST, bool MIXED, bool ENERGY, bool WRITE_F, bool ONE_TYPE, bool INTEGRATE>
__device__ __forceinline__ void lj_tiles(const StepArgs &A, Stage &S, const ListSeg &G,
                                         const RepInfo &R, bool staged, bool mi, double &ke, double &pe_b,
                                         double &vir_b, float &d2m, float &v2m,
                                         unsigned long long *hsum) {
  const uint32_t ntile = S.W.rpref[S.W.nrun];
  const double L0 = R.L[0], L1 = R.L[1], L2 = R.L[2];
  const double I0 = R.Linv[0], I1 = R.Linv[1], I2 = R.Linv[2];
  const uint32_t lane = pb_lane(), sl = lane & (LPA - 1u), grp = lane / LPA;
  constexpr uint32_t APW = 64u / LPA;            // particles per wavefront
  const uint32_t wave = threadIdx.x / 64, nwv = blockDim.x / 64;
  // replica constants read once, before any store (stores could alias them)
  const double half_skin2 = R.half_skin2;
  const float inv_hs2 = R.inv_hs2;
  const size_t ld = A.npad;
  // Software pipeline across particle groups: while group u is evaluated,
  // the next group's slot, list length, first list batch and integration
  // state are already in flight.  The first word pair is read without
  // knowing the list length: those words lie inside the list array for any
  // slot (capacity >= 20 entries) and are not used past the length.
  uint32_t nli = 0, ngi = 0, nnn = 0;
  ulonglong2 nqq = make_ulonglong2(0, 0);
  double4 np_own = make_double4(0, 0, 0, 0), nv_own = np_own;
  float4 nrb_own = make_float4(0, 0, 0, 0);
  // the next group's slot, length and first list word, needed when its
  // iteration starts: requested at the start of the current iteration
  auto fetch_list = [&](uint32_t ub) {
    const uint32_t u = ub + grp;
    uint32_t run;
    nli = 0;
    ngi = 0;
    nnn = 0;
    if (u < ntile) {
      tile_particle(S.W, u, nli, ngi, run);
      nnn = A.nn[ngi];
      nqq = reinterpret_cast<const ulonglong2 *>(A.nl)[ngi];
    }
  };
  // the next group's integration state, needed at the end of its iteration:
  // requested at the end of the current one (one set of registers live)
  uint32_t norig = 0;
  auto fetch_state = [&](uint32_t ub) {
    if (INTEGRATE && sl == 0 && ub + grp < ntile) {
      np_own = A.rin[ngi];
      nv_own = A.vin[ngi];
      nrb_own = A.rbuild[ngi];
      if (WRITE_F) norig = A.orig[ngi];
    }
  };
  // the first group's list and state loads are in flight while the window
  // is staged
  fetch_list(wave * APW);
  fetch_state(wave * APW);
  stage_window(S, R, A.rin);
  for (uint32_t ub = wave * APW; ub < ntile; ub += nwv * APW) {   // wave-uniform
    const uint32_t u = ub + grp;
    const bool act = u < ntile;
    const uint32_t li = nli, gi = ngi, nni = nnn;
    ulonglong2 qq = nqq;
    const double4 p_own = np_own, v_own = nv_own;
    const float4 rb_own = nrb_own;
    const uint32_t o_own = norig;
    fetch_list(ub + nwv * APW);
    double4 pi;
    if (FAST) pi = make_double4(S.x[li], S.y[li], S.z[li], (double)S.t[li]);
    else pi = act ? window_pos(S, G, R, A.rin, staged, li) : make_double4(0, 0, 0, 0);
    const uint32_t ti = ONE_TYPE ? 0u : (uint32_t)pi.w;
    // forces and energies accumulate in FP64 in both modes; MIXED evaluates
    // each pair in FP32 on the FP64 displacement after the FP64 cutoff test
    double fx = 0, fy = 0, fz = 0, pe = 0, vir = 0;
    float pef = 0.f, virf = 0.f;   // MIXED energies: per-lane FP32 sums of ~20 pairs
    auto pair = [&](uint32_t q) {
      double4 pj;
      if (FAST) pj = make_double4(S.x[q], S.y[q], S.z[q], ONE_TYPE ? 0.0 : (double)S.t[q]);
      else pj = window_pos(S, G, R, A.rin, staged, q);
      double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      if (!FAST && mi) {
        dx -= L0 * rint(dx * I0);
        dy -= L1 * rint(dy * I1);
        dz -= L2 * rint(dz * I2);
      }
      uint32_t t = 0;
      if (!ONE_TYPE) t = ti * PB_MAXT + (uint32_t)pj.w;
      const double r2d = dx * dx + dy * dy + dz * dz;
      const bool in = r2d < A.rc2 && (FAST || !mi || q != NL_PAD);
      if (MIXED) {
        // r2 of a padding entry is large but finite; zero r2 (an image of a
        // padding entry under the minimum image) gives inf / NaN that the
        // selects below discard
        const float r2 = (float)r2d;
        const float s2 = ONE_TYPE ? A.s2f : A.lj->s2f[t];
        const float e24 = ONE_TYPE ? A.e24f : A.lj->e24f[t];
        float ir2 = __builtin_amdgcn_rcpf(r2);
        ir2 = ir2 * (2.0f - r2 * ir2);               // one Newton step: ~0.5 ulp
        ir2 = in ? ir2 : 0.0f;                       // outside rc: f/r = 0 (one FP32 select)
        float s6 = s2 * ir2;
        s6 = s6 * s6 * s6;
        // f/r = 24 eps s6 (2 s6 - 1) / r^2 (kernelinline.hh:90-95), written as
        // ir2 s6 (48 eps s6 - 24 eps): one multiply fewer a pair
        const float e48 = e24 + e24;                 // exact; hoisted with one type
        const float fr = ir2 * s6 * fmaf(e48, s6, -e24);
        const double frd = (double)fr;              // f/r, applied to the FP64 displacement
        fx = fma(frd, dx, fx);
        fy = fma(frd, dy, fy);
        fz = fma(frd, dz, fz);
        if (ENERGY) {
          const float e4 = ONE_TYPE ? A.e4f : A.lj->e4f[t];
          const float sh = ONE_TYPE ? A.shiftf : A.lj->shiftf[t];
          const float ep = e4 * s6 * (s6 - 1.0f) - sh;
          pef += in ? ep : 0.0f;
          virf = fmaf(fr, r2, virf);
        }
      } else {
        const double r2 = r2d;
        const double s2 = ONE_TYPE ? A.s2 : A.lj->s2[t];
        const double e24 = ONE_TYPE ? A.e24 : A.lj->e24[t];
        const double ir2 = rcp64(r2);
        double s6 = s2 * ir2;
        s6 = s6 * s6 * s6;
        double fr = e24 * s6 * (2.0 * s6 - 1.0) * ir2;
        fr = in ? fr : 0.0;
        fx += fr * dx;
        fy += fr * dy;
        fz += fr * dz;
        if (ENERGY) {
          const double e4 = ONE_TYPE ? A.e4 : A.lj->e4[t];
          const double sh = ONE_TYPE ? A.shift : A.lj->shift[t];
          const double ep = e4 * s6 * (s6 - 1.0) - sh;
          pe += in ? ep : 0.0;
          vir += fr * r2;
        }
      }
    };
    // rows are padded with the inert entry to whole word pairs
    constexpr uint32_t EPW = FAST ? 5u : 4u, EB = FAST ? 12u : 16u;
    constexpr uint32_t EMASK = (1u << EB) - 1u;
    const uint32_t nbat = act ? nni / (EPW * LPA) : 0u;
    auto batch = [&](uint64_t w64) {
#pragma unroll
      for (uint32_t j = 0; j < EPW; ++j) {
        const uint32_t qj = (uint32_t)(w64 >> (EB * j)) & EMASK;
        pair(qj);
      }
    };
    static_assert(LPA == 1u, "word-pair rows assume one lane a particle");
    if (!FAST) {
      // one word a batch (the minimum-image and unstaged paths use more
      // registers a pair): word m of the row at 2 gi + (m >> 1) 2 npad + (m & 1)
      uint64_t qw = qq.x;
      const uint64_t *wp = A.nl + 2u * (size_t)gi + 1u;
      for (uint32_t b = 0; b < nbat; ++b) {
        const uint64_t w64 = qw;
        qw = *wp;                                 // one word ahead (slack words at the end)
        wp += (b & 1u) ? 1u : 2u * ld - 1u;
        batch(w64);
      }
    } else {
      // one 16-byte load brings the next two words while ten pairs are
      // evaluated
      const ulonglong2 *pp = reinterpret_cast<const ulonglong2 *>(A.nl) + ld + gi;
      const uint32_t npair = act ? nni / (2u * EPW) : 0u;
      for (uint32_t b = 0; b < npair; ++b) {
        // unpack both words first, so the load lands in their registers and
        // is waited for only when the next step unpacks it
        uint32_t e[2 * EPW];
#pragma unroll
        for (uint32_t j = 0; j < EPW; ++j) {
          e[j] = (uint32_t)(qq.x >> (EB * j)) & EMASK;
          e[EPW + j] = (uint32_t)(qq.y >> (EB * j)) & EMASK;
        }
        // the entries are computed here (an empty asm that reads and writes
        // them), so the old words are dead when the load is issued
#pragma unroll
        for (uint32_t j = 0; j < 2 * EPW; ++j) asm volatile("" : "+v"(e[j]));
        // the row's next pair; none past the row's end (its cache lines
        // would be fetched for nothing when all their lanes are done)
        if (b + 1u < npair) qq = *pp;
        pp += ld;
#pragma unroll
        for (uint32_t j = 0; j < EPW; ++j) pair(e[j]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (uint32_t j = 0; j < EPW; ++j) pair(e[EPW + j]);
      }
    }
    if (MIXED && ENERGY) {
      pe = (double)pef;
      vir = (double)virf;
    }
#pragma unroll
    for (int o = (int)LPA / 2; o > 0; o >>= 1) {
      fx += __shfl_xor(fx, o, 64);
      fy += __shfl_xor(fy, o, 64);
      fz += __shfl_xor(fz, o, 64);
      if (ENERGY) {
        pe += __shfl_xor(pe, o, 64);
        vir += __shfl_xor(vir, o, 64);
      }
    }
    fetch_state(ub + nwv * APW);
    if (!act || sl != 0) continue;

    if (INTEGRATE) {
      double4 p = p_own;
      double4 vi = v_own;
      const float4 rb = rb_own;
      // the list is valid for this step iff the positions the pair loop used
      // are within skin/2 of their positions at the last rebuild
      auto disp2 = [&](const double4 &x) {
        const float ex = (float)(x.x - (double)rb.x), ey = (float)(x.y - (double)rb.y),
                    ez = (float)(x.z - (double)rb.z);
        return ex * ex + ey * ey + ez * ez;
      };
      if ((double)disp2(p) > half_skin2) atomicOr(&A.counter[CNT_SKIN], 1u);
      // kick-drift form of propagateVelocityVerlet_F4 (kernelfuncs.cu:
      // 347-365): v holds v(t - dt/2) = v_ref(t-dt) + Fold dt/2m
      const double dtm2 = 0.5 * A.dt * vi.w;
      const double vx = vi.x + fx * dtm2, vy = vi.y + fy * dtm2, vz = vi.z + fz * dtm2;  // v(t)
      if (ENERGY) ke += 0.5 * (vx * vx + vy * vy + vz * vz) / vi.w;
      vi.x = vx + fx * dtm2;
      vi.y = vy + fy * dtm2;
      vi.z = vz + fz * dtm2;
      p.x += vi.x * A.dt;
      p.y += vi.y * A.dt;
      p.z += vi.z * A.dt;
      if (!WRITE_F) {
        A.rout[gi] = p;
        A.vout[gi] = vi;
      }
      // displacement of the new positions feeds the rebuild schedule
      // in units of the replica's half skin (replicas may have different skins)
      d2m = fmaxf(d2m, disp2(p) * inv_hs2);
      v2m = fmaxf(v2m, (float)(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z) * inv_hs2);
      if (WRITE_F) {
        // last step of a run: the reference VV state r(t+dt), v(t), Fold =
        // F(t), positions wrapped into the box (treatPeriodicCondition), and
        // the fingerprint of the slot that a next run compares against
        p.x = wrap1(p.x, R.lo[0], R.hi[0]);
        p.y = wrap1(p.y, R.lo[1], R.hi[1]);
        p.z = wrap1(p.z, R.lo[2], R.hi[2]);
        A.rout[gi] = p;
        A.vout[gi] = make_double4(vx, vy, vz, vi.w);
        A.Aout[gi] = make_double4(fx, fy, fz, 0.0);
        atomicAdd(hsum + (lane & 15u), (unsigned long long)slot_hash(gi, o_own, p));   // LDS
      }
    }
    if (WRITE_F) A.Fout[gi] = make_double4(fx, fy, fz, 0.0);
    if (ENERGY) {
      pe_b += 0.5 * pe;
      vir_b += 0.5 * vir;
    }
  }
}

// One segment: stage its window, pairs + integration of its tile
// particles, block maxima and energy partials.
template <bool MIXED, bool ENERGY, bool WRITE_F, bool ONE_TYPE, bool INTEGRATE>
__device__ __forceinline__ void lj_segment(const StepArgs &A, uint32_t seg) {
  __shared__ Stage S;
  const ListSeg G = A.segs[seg];
  const RepInfo &R = A.reps[G.rep];
  if (WRITE_F && INTEGRATE && threadIdx.x < 16) S.hsum[threadIdx.x] = 0;   // visible after the staging
  load_window_table(S.W, A.wtab, seg);
  const bool staged = S.W.total <= WCAP;
  const bool mi = G.whole != 0;

  double ke = 0.0, pe_b = 0.0, vir_b = 0.0;
  float d2m = 0.f, v2m = 0.f;
  if (staged && !mi)          // block-uniform: the common case, branch-free pairs
    lj_tiles<true, MIXED, ENERGY, WRITE_F, ONE_TYPE, INTEGRATE>(A, S, G, R, staged, mi, ke, pe_b, vir_b,
                                                                d2m, v2m, S.hsum);
  else
    lj_tiles<false, MIXED, ENERGY, WRITE_F, ONE_TYPE, INTEGRATE>(A, S, G, R, staged, mi, ke, pe_b,
                                                                 vir_b, d2m, v2m, S.hsum);

  __shared__ float smx[2][TBS / 64];
  __shared__ double sen[3][TBS / 64];
  if (INTEGRATE) {
    // block maxima of displacement^2 and speed^2; an atomic only when the
    // block beats the running maximum (non-negative floats order as uints)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      d2m = fmaxf(d2m, __shfl_xor(d2m, o, 64));
      v2m = fmaxf(v2m, __shfl_xor(v2m, o, 64));
    }
    if (pb_lane() == 0) {
      smx[0][threadIdx.x / 64] = d2m;
      smx[1][threadIdx.x / 64] = v2m;
    }
  }
  if (ENERGY) {
    const double kw = pb_wave_sum(ke), pw = pb_wave_sum(pe_b), ww = pb_wave_sum(vir_b);
    if (pb_lane() == 0) {
      sen[0][threadIdx.x / 64] = kw;
      sen[1][threadIdx.x / 64] = pw;
      sen[2][threadIdx.x / 64] = ww;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t nw = blockDim.x / 64;
    if (WRITE_F && INTEGRATE) {
      unsigned long long t = 0;
      for (int k = 0; k < 16; ++k) t += S.hsum[k];
      if (t) atomicAdd(A.hash + blockIdx.x % HASH_SLOTS, t);
    }
    if (INTEGRATE) {
      float bd = 0.f, bv = 0.f;
      for (uint32_t w = 0; w < nw; ++w) {
        bd = fmaxf(bd, smx[0][w]);
        bv = fmaxf(bv, smx[1][w]);
      }
      const uint32_t ud = __float_as_uint(bd), uv = __float_as_uint(bv);
      if (ud > __hip_atomic_load(&A.counter[CNT_DISP], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(&A.counter[CNT_DISP], ud);
      if (uv > __hip_atomic_load(&A.counter[CNT_VMAX], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(&A.counter[CNT_VMAX], uv);
    }
    if (ENERGY) {
      // per-segment partials (a segment never spans two replicas), summed
      // per replica in a fixed order by k_energy_reduce after this kernel
      double acc[3] = {0, 0, 0};
      for (uint32_t w = 0; w < nw; ++w)
        for (int c = 0; c < 3; ++c) acc[c] += sen[c][w];
      double *P = A.partials + 4 * (size_t)seg;
      P[0] = acc[0];
      P[1] = acc[1];
      P[2] = acc[2];
    }
  }
}

template <bool MIXED, bool ENERGY, bool WRITE_F, bool ONE_TYPE, bool INTEGRATE, bool PERSIST>
__device__ __forceinline__ void lj_step_impl(const StepArgs &A) {
  publish_prev(A);
  if constexpr (PERSIST) {
    // Persistent form: a grid of (compute units x resident workgroups) drains
    // the segments from eight work queues, one per slab of consecutive
    // segments (an eighth of them each); a workgroup takes
    // from the queue of the XCD it runs on first (workgroups are dispatched
    // round-robin over the 8 XCDs), then from the others.  The last
    // workgroup to finish resets the queues for the next launch (stream
    // order makes that visible), so the launch can be replayed in a graph.
    __shared__ uint32_t ticket;
    const uint32_t n = A.nsegs, per = n >> 3, rmd = n & 7u;
    const uint32_t home = blockIdx.x & 7u;
    uint32_t k = 0;
    for (;;) {
      // queue, slab start and segment index are block-uniform (scalar registers)
      const uint32_t q = __builtin_amdgcn_readfirstlane((home + k) & 7u);
      __syncthreads();                   // the last segment is done with LDS and `ticket`
      if (threadIdx.x == 0) ticket = atomicAdd(&A.queue[q], 1u);
      __syncthreads();
      const uint32_t t = __builtin_amdgcn_readfirstlane(ticket);
      const uint32_t cnt = per + (q < rmd ? 1u : 0u);
      if (t >= cnt) {                    // this queue is drained: the next one
        if (++k == 8) break;
        continue;
      }
      const uint32_t seg = __builtin_amdgcn_readfirstlane(q * per + (q < rmd ? q : rmd) + t);
      lj_segment<MIXED, ENERGY, WRITE_F, ONE_TYPE, INTEGRATE>(A, seg);
    }
    if (threadIdx.x == 0) {
      __threadfence();
      if (atomicAdd(&A.queue[8], 1u) == gridDim.x - 1) {
        // every other workgroup has finished taking work: reset
        for (int q = 0; q < 9; ++q) __hip_atomic_store(&A.queue[q], 0u, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    return;
  }
  // eight consecutive segments per XCD in every 64 (pb_xcd_chunk8): the same
  // as one slab per XCD on a uniform fluid, 3 % faster on the batched
  // replicas, whose segment sizes vary with the replica
  lj_segment<MIXED, ENERGY, WRITE_F, ONE_TYPE, INTEGRATE>(A, pb_xcd_chunk8());
}

// Two entry points over the same body: the common variants fit 128 VGPRs
// (4 wavefronts per SIMD, two 512-thread workgroups per CU); the FP64 and
// multi-type energy variants need more registers and run without the cap
// (no spills) at lower occupancy.
template <bool MIXED, bool ENERGY, bool WRITE_F, bool ONE_TYPE, bool INTEGRATE, bool PERSIST>
__global__ void __launch_bounds__(TBS) __attribute__((amdgpu_waves_per_eu(4))) k_lj_step(StepArgs A) {
  lj_step_impl<MIXED, ENERGY, WRITE_F, ONE_TYPE, INTEGRATE, PERSIST>(A);
}
template <bool MIXED, bool ENERGY, bool WRITE_F, bool ONE_TYPE, bool INTEGRATE, bool PERSIST>
__global__ void __launch_bounds__(TBS) k_lj_step_wide(StepArgs A) {
  lj_step_impl<MIXED, ENERGY, WRITE_F, ONE_TYPE, INTEGRATE, PERSIST>(A);
}

// ------------------------------------------------------------- host side
static uint32_t expected_maxnb(const pb_sys *s) {
  double mean = 0.0;
  for (uint32_t q = 0; q < s->nrep; ++q) {
    const RepInfo &R = s->reps[q];
    const double rho = (double)R.natoms / (R.L[0] * R.L[1] * R.L[2]);
    mean = std::max(mean, 4.0 / 3.0 * M_PI * R.rlist * R.rlist * R.rlist * rho);
  }
  uint32_t m = (uint32_t)std::ceil(1.5 * mean + 48.0);
  return (m + NL_ROW - 1u) / NL_ROW * NL_ROW;
}

// List cells, radii and segments for cutoff rc with skin[q] per replica.
static int setup_list_cells(pb_sys *s, double rc, const std::vector<double> &skin) {
  uint32_t cell0 = 0;
  double rmax_all = 0.0;
  for (uint32_t q = 0; q < s->nrep; ++q) {
    RepInfo &R = s->reps[q];
    const double rlist = rc + skin[q];
    rmax_all = std::max(rmax_all, rlist);
    R.rlist = rlist;
    // FP32 build radius: window-relative coordinates stay below ~4 list
    // cells plus the drift since the last wrap; 1e-4 covers their rounding
    R.rb2 = (float)((rlist + 1e-4) * (rlist + 1e-4));
    const double hs = 0.5 * skin[q];
    R.half_skin2 = hs * hs;
    R.inv_hs2 = hs > 0.0 ? (float)(1.0 / (hs * hs)) : 0.0f;
    for (int a = 0; a < 3; ++a) {
      if (!(2.0 * rlist < R.L[a]))
        return pb_fail(s->ctx, PB_ERR_CELLS, "list cutoff must be below half the box");
      uint32_t nb = (uint32_t)std::floor(R.L[a] / rlist);
      if (nb < 1) nb = 1;
      if (nb > 1024) nb = 1024;
      R.nb[a] = nb;
    }
    R.cell0 = cell0;
    cell0 += R.nb[0] * R.nb[1] * R.nb[2] + 1;   // + pad cell
  }
  s->list_ncell = cell0;
  s->list_rlist = rmax_all;
  // Tiles (segments) of tx x ty columns x nz cells.  nz is chosen so the
  // expected window (tx+2)(ty+2)(nz+2) cells fills tile_fill of the LDS
  // stage (0.72 unless pb_lj_autotune chose otherwise);
  // nz <= nb2 - 2 so a window never holds one cell twice.  A replica with
  // fewer than 3 list cells along an axis is one whole segment.
  std::vector<ListSeg> segs;
  for (uint32_t q = 0; q < s->nrep; ++q) {
    RepInfo &R = s->reps[q];
    R.seg0 = (uint32_t)segs.size();
    const uint32_t nc = R.nb[0] * R.nb[1] * R.nb[2];
    if (R.nb[0] < 3 || R.nb[1] < 3 || R.nb[2] < 3) {
      if (R.natoms > WCAP)
        return pb_fail(s->ctx, PB_ERR_CELLS,
                       "a replica thinner than 3 list cells must fit one LDS window");
      segs.push_back(ListSeg{q, 0, 0, 0, 1, 1, 1, 1});
    } else {
      const uint32_t tw = s->tile_w;
      const uint32_t tx = R.nb[0] >= tw + 2 ? tw : 1, ty = R.nb[1] >= tw + 2 ? tw : 1;
      const double per_cell = (double)R.natoms / nc;
      int nz = (int)std::floor(s->tile_fill * WCAP /
                               ((tx + 2) * (ty + 2) * std::max(per_cell, 1e-3))) - 2;
      nz = std::max(1, std::min<int>({nz, MAXZ, (int)R.nb[2] - 2}));
      // the column split into the fewest z runs of at most nz cells, of
      // equal length to within one cell (a short last run would be a
      // workgroup with a few busy wavefronts)
      const uint32_t nrun = (R.nb[2] + (uint32_t)nz - 1u) / (uint32_t)nz;
      const uint32_t zb = R.nb[2] / nrun, zr = R.nb[2] % nrun;
      for (uint32_t ci = 0; ci < R.nb[0]; ci += tx)
        for (uint32_t cj = 0; cj < R.nb[1]; cj += ty)
          for (uint32_t k = 0, ck = 0; k < nrun; ++k) {
            const uint32_t len = zb + (k < zr ? 1u : 0u);
            segs.push_back(ListSeg{q, ci, cj, ck, len, std::min(tx, R.nb[0] - ci),
                                   std::min(ty, R.nb[1] - cj), 0});
            ck += len;
          }
    }
    R.nseg = (uint32_t)segs.size() - R.seg0;
  }
  s->list_nsegs = (uint32_t)segs.size();
  if (int rc = pb_alloc(s->ctx, &s->list_segs, 8 * segs.size())) return rc;
  // the window tables persist while the segments do (a continued list uses
  // the ones its build stored)
  if (segs.size() > s->list_wtab_cap) {
    if (int rc = pb_alloc(s->ctx, &s->list_wtab, (size_t)WTAB_U4 * segs.size())) return rc;
    s->list_wtab_cap = (uint32_t)segs.size();
  }
  PB_HIP(s->ctx, hipMemcpyAsync(s->list_segs, segs.data(), sizeof(ListSeg) * segs.size(),
                                hipMemcpyHostToDevice, s->ctx->stream));
  // per-segment energy partials
  if (s->list_nsegs > s->nwave) {
    if (int rc = pb_alloc(s->ctx, &s->partials, 4 * (size_t)s->list_nsegs)) return rc;
  }
  PB_HIP(s->ctx, hipMemcpyAsync(s->d_reps, s->reps.data(), sizeof(RepInfo) * s->nrep,
                                hipMemcpyHostToDevice, s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));   // segs is a host temporary
  return pb_ensure_sort_ws(s, s->list_ncell);
}

static int ensure_list_buffers(pb_sys *s, uint32_t maxnb) {
  int rc;
  if (!s->r2buf) {
    if ((rc = pb_alloc(s->ctx, &s->r2buf, s->npad)) || (rc = pb_alloc(s->ctx, &s->v2buf, s->npad)) ||
        (rc = pb_alloc(s->ctx, &s->a2buf, s->npad)) || (rc = pb_alloc(s->ctx, &s->orig2, s->npad)) ||
        (rc = pb_alloc(s->ctx, &s->type2, s->npad)) || (rc = pb_alloc(s->ctx, &s->move2, s->npad)) ||
        (rc = pb_alloc(s->ctx, &s->rbuild, s->npad)) || (rc = pb_alloc(s->ctx, &s->nn, s->npad)))
      return rc;
  }
  if (s->maxnb < maxnb) {
    // + 48 entries per slot: the step kernel prefetches one batch (4 words)
    // past the end
    if ((rc = pb_alloc(s->ctx, &s->nl, (size_t)(maxnb + 48) * s->npad))) return rc;
    s->maxnb = maxnb;
  }
  return PB_OK;
}

// rebuild: cell ids -> counting sort -> reorder into cell order -> list.
// F is not carried: every caller rewrites it before it is read again.
// carry_a = false inside a fused run, where Fold (a) is dead between the
// conversion to the kick-drift state and the final one (which rewrites it).
static int rebuild(pb_sys *s, bool carry_a) {
  hipStream_t st = s->ctx->stream;
  const uint32_t n = s->npad;
  if (int rc = pb_ensure_sort_ws(s, s->list_ncell)) return rc;
  hipLaunchKernelGGL(k_fill_u32, dim3(pb_blocks(s->list_ncell + 1, 256)), dim3(256), 0, st,
                     s->ccount, s->list_ncell + 1, 0u);
  hipLaunchKernelGGL(k_list_cellid, dim3(pb_blocks(n, 256)), dim3(256), 0, st, s->r, s->orig,
                     s->waveRep, s->d_reps, n, s->cellid, s->ccount, s->rank);
  PB_LAUNCH_CHECK(s->ctx, "k_list_cellid");
  if (int rc = pb_counting_sort_counted(s, s->list_ncell, s->orig)) return rc;
  ReorderArgs A;
  A.perm = s->perm;
  A.r = s->r; A.v = s->v; A.a = carry_a ? s->a : nullptr;
  A.orig = s->orig; A.type = s->type; A.move = s->move;
  A.r2 = s->r2buf; A.v2 = s->v2buf; A.a2 = s->a2buf;
  A.orig2 = s->orig2; A.type2 = s->type2; A.move2 = s->move2;
  A.rbuild = s->rbuild;
  A.nn = s->nn;
  A.waveRep = s->waveRep; A.reps = s->d_reps; A.n = n;
  hipLaunchKernelGGL(k_reorder, dim3(pb_blocks(n, 256)), dim3(256), 0, st, A);
  PB_LAUNCH_CHECK(s->ctx, "k_reorder");
  std::swap(s->r, s->r2buf);
  std::swap(s->v, s->v2buf);
  if (carry_a) std::swap(s->a, s->a2buf);
  std::swap(s->orig, s->orig2);
  std::swap(s->type, s->type2);
  std::swap(s->move, s->move2);
  hipLaunchKernelGGL(k_build_list, dim3(s->list_nsegs), dim3(TB), 0, st, s->r, s->cstart,
                     (const ListSeg *)s->list_segs, s->d_reps, n, s->maxnb, s->nl, s->nn,
                     s->d_counter, s->list_wtab);
  PB_LAUNCH_CHECK(s->ctx, "k_build_list");
  s->ref_binned = false;
  return PB_OK;
}

static StepArgs make_args(pb_sys *s, double dt) {
  StepArgs A;
  std::memset(&A, 0, sizeof(A));
  A.nl = (const uint64_t *)s->nl;
  A.nn = s->nn;
  A.rbuild = s->rbuild;
  A.cstart = s->cstart;
  A.wtab = s->list_wtab;
  A.segs = (const ListSeg *)s->list_segs;
  A.reps = s->d_reps;
  A.lj = s->d_lj;
  A.partials = s->partials;
  A.counter = s->d_counter;
  A.nsegs = s->list_nsegs;
  A.npad = s->npad;
  A.nrep = s->nrep;
  A.dt = dt;
  A.rc2 = s->rc2;
  A.s2 = s->lj.s2[0];
  A.e24 = s->lj.e24[0];
  A.e4 = s->lj.e4[0];
  A.shift = s->lj.shift[0];
  A.s2f = s->lj.s2f[0];
  A.e24f = s->lj.e24f[0];
  A.e4f = s->lj.e4f[0];
  A.shiftf = s->lj.shiftf[0];
  return A;
}

// Energy record of a step: per-replica sums of the per-segment partials in a
// fixed order (deterministic), into record counter[CNT_REC], which is then
// advanced.  A separate launch: stream order makes the partials visible.
__global__ void __launch_bounds__(256) k_energy_reduce(const double *partials, const RepInfo *reps,
                                                       uint32_t nrep, double *energy,
                                                       uint32_t *counter) {
  __shared__ double sm[4 * 256];
  const uint32_t rec = counter[CNT_REC];
  pb_block_reduce_reps(partials, 3, reps, nrep, energy + (size_t)rec * 8 * nrep, true, sm);
  __syncthreads();
  if (threadIdx.x == 0) counter[CNT_REC] = rec + 1;
}

typedef void (*StepKernel)(StepArgs);

template <bool MIXED, bool ONE, bool E, bool W, bool I, bool P>
static StepKernel pick1() {
  // variants that do not fit 128 VGPRs without spilling
  if constexpr ((E && (!MIXED || !ONE)) || (!MIXED && !ONE))
    return k_lj_step_wide<MIXED, E, W, ONE, I, P>;
  else
    return k_lj_step<MIXED, E, W, ONE, I, P>;
}

template <bool MIXED, bool ONE, bool P>
static StepKernel pick3(bool energy, bool wf, bool integ) {
  if (integ) {
    if (energy) return wf ? pick1<MIXED, ONE, true, true, true, P>() : pick1<MIXED, ONE, true, false, true, P>();
    return wf ? pick1<MIXED, ONE, false, true, true, P>() : pick1<MIXED, ONE, false, false, true, P>();
  }
  return energy ? pick1<MIXED, ONE, true, true, false, false>() : pick1<MIXED, ONE, false, true, false, false>();
}

template <bool MIXED, bool ONE>
static StepKernel pick2(bool energy, bool wf, bool integ, bool persist) {
  return persist ? pick3<MIXED, ONE, true>(energy, wf, integ) : pick3<MIXED, ONE, false>(energy, wf, integ);
}

// persistent variants exist for the integrating steps of a fused run
static StepKernel pick_kernel(bool mixed, bool one, bool energy, bool wf, bool integ, bool persist) {
  if (mixed) return one ? pick2<true, true>(energy, wf, integ, persist) : pick2<true, false>(energy, wf, integ, persist);
  return one ? pick2<false, true>(energy, wf, integ, persist) : pick2<false, false>(energy, wf, integ, persist);
}

// With events, their timestamps are the step kernel's own start and end
// (taken by the dispatch, hipExtLaunchKernelGGL): a marker packet of
// hipEventRecord between kernels costs about 10 us of idle GPU per step.
static int launch_step(pb_sys *s, const StepArgs &A, bool mixed, bool energy, bool wf,
                       bool integ, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  const bool one = s->lj.ntypes == 1;
  StepKernel k = pick_kernel(mixed, one, energy, wf, integ, A.queue != nullptr);
  // persistent form: two 512-thread workgroups per compute unit (the step
  // kernel's LDS and register use), never more than there are segments
  const uint32_t grid = A.queue ? std::min<uint32_t>(s->list_nsegs, 2u * (uint32_t)std::max(s->ctx->cu_count, 1))
                                : s->list_nsegs;
  if (ev0 && ev1)
    hipExtLaunchKernelGGL(k, dim3(grid), dim3(TBS), 0, s->ctx->stream, ev0, ev1, 0u, A);
  else
    hipLaunchKernelGGL(k, dim3(grid), dim3(TBS), 0, s->ctx->stream, A);
  if (energy)
    hipLaunchKernelGGL(k_energy_reduce, dim3(1), dim3(256), 0, s->ctx->stream, A.partials, A.reps,
                       A.nrep, A.energy, A.counter);
  PB_LAUNCH_CHECK(s->ctx, "k_lj_step");
  return PB_OK;
}

static int check_lj_ready(pb_sys *s) {
  if (!s->lj_set) return pb_fail(s->ctx, PB_ERR_STATE, "LJ parameters not set");
  if (!(s->rc2 > 0)) return pb_fail(s->ctx, PB_ERR_STATE, "LJ cutoff not set");
  if (!s->cell_set) return pb_fail(s->ctx, PB_ERR_STATE, "cell not set");
  if (s->w || s->radius)
    return pb_fail(s->ctx, PB_ERR_STATE, "LJ path does not carry rotational/radius state");
  return PB_OK;
}

static int read_counters(pb_sys *s, uint32_t *c) {
  PB_HIP(s->ctx, hipMemcpyAsync(c, s->d_counter, 16 * 4, hipMemcpyDeviceToHost, s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  return PB_OK;
}

static int reset_counters(pb_sys *s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s->ctx->stream, s->d_counter, 16u, 0u);
  PB_LAUNCH_CHECK(s->ctx, "reset counters");
  return PB_OK;
}

// build (or grow) the list so that no particle overflows it
static int build_checked(pb_sys *s, bool carry_a) {
  uint32_t want = std::max(s->maxnb, expected_maxnb(s));
  for (int attempt = 0; attempt < 4; ++attempt) {
    if (int rc = ensure_list_buffers(s, want)) return rc;
    if (int rc = reset_counters(s)) return rc;
    if (int rc = rebuild(s, carry_a)) return rc;
    uint32_t c[16];
    if (int rc = read_counters(s, c)) return rc;
    if (c[CNT_WIN])
      return pb_fail(s->ctx, PB_ERR_CELLS, "a list window holds more than 65534 particles");
    if (c[CNT_OVF] == 0) {
      s->last_maxnn = c[CNT_MAXNN];
      return PB_OK;
    }
    want = (c[CNT_OVF] * 5 / 4 + 2u * NL_ROW - 1u) / NL_ROW * NL_ROW;
  }
  return pb_fail(s->ctx, PB_ERR_NLIST, "neighbour list keeps overflowing");
}

extern "C" int pb_lj_force(pb_sys *s, int precision, double *pe, double *virial) {
  if (!s) return PB_ERR_ARG;
  if (int rc = check_lj_ready(s)) return rc;
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  const double rc = std::sqrt(s->rc2);
  if (int e = setup_list_cells(s, rc, std::vector<double>(s->nrep, rc * 1e-12))) return e;
  if (int e = build_checked(s, true)) return e;
  StepArgs A = make_args(s, 0.0);
  A.rin = s->r;
  A.vin = s->v;
  A.Fout = s->F;
  const bool energy = (pe || virial);
  double *ebuf = nullptr;
  if (energy) {
    if (int e = pb_alloc(s->ctx, &ebuf, 8 * (size_t)s->nrep)) return e;
    A.energy = ebuf;
  }
  if (int e = launch_step(s, A, precision == PB_MIXED, energy, true, false)) {
    pb_free(ebuf);
    return e;
  }
  if (energy) {
    std::vector<double> h(8 * (size_t)s->nrep);
    hipError_t e = hipMemcpyAsync(h.data(), ebuf, 8 * sizeof(double) * s->nrep,
                                  hipMemcpyDeviceToHost, s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    pb_free(ebuf);
    if (e != hipSuccess) return pb_hip_check(s->ctx, e, "pb_lj_force energies");
    for (uint32_t q = 0; q < s->nrep; ++q) {
      if (pe) pe[q] = h[8 * q + 1];
      if (virial) virial[q] = h[8 * q + 2];
    }
  }
  return PB_OK;
}

// Sum over slots of a per-slot hash (order-free, so any grid gives the same
// value): a grid-stride loop and one atomic per workgroup, since atomics of
// every wavefront on one address serialise (one per wavefront took 200 us at
// 1M slots).
__global__ void __launch_bounds__(256) k_state_hash(const double4 *r, const uint32_t *orig,
                                                    uint32_t n, unsigned long long *out) {
  uint64_t acc = 0;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const uint32_t o = orig[s];
    if (o != PB_NONE) acc += slot_hash(s, o, r[s]);
  }
  hash_flush(acc, out);
}

// Start of a fused run: reference VV state -> kick-drift state, and for a
// continued list the unwrap: the periodic wrap at the end of the previous
// run moved particles that left the box by one box length, while the list
// and the window images expect every particle near its position at the
// last list build, so each is put back at the image nearest to that
// position.
__global__ void k_run_start(double4 *v, const double4 *Fold, double4 *r, const float4 *rbuild,
                            const uint32_t *orig, const uint32_t *waveRep, const RepInfo *reps,
                            uint32_t n, double dt, int unwrap) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double4 V = v[s];
  const double4 f = Fold[s];
  const double dtm2 = 0.5 * dt * V.w;
  V.x += f.x * dtm2;
  V.y += f.y * dtm2;
  V.z += f.z * dtm2;
  v[s] = V;
  if (!unwrap || orig[s] == PB_NONE) return;
  const RepInfo &R = reps[waveRep[s / PB_WAVE]];
  double4 p = r[s];
  const float4 b = rbuild[s];
  p.x -= R.L[0] * rint((p.x - (double)b.x) * R.Linv[0]);
  p.y -= R.L[1] * rint((p.y - (double)b.y) * R.Linv[1]);
  p.z -= R.L[2] * rint((p.z - (double)b.z) * R.Linv[2]);
  r[s] = p;
}

// grid of the hash pass: up to 4 workgroups of 256 per CU, a grid-stride
// loop (measured faster than one slot a thread: 10.4 against 16.4 us at 1M)
static uint32_t hash_blocks(const pb_sys *s) {
  return std::max<uint32_t>(
      1u, std::min<uint32_t>(pb_blocks(s->npad, 256), 4u * (uint32_t)s->ctx->cu_count));
}

static int hash_slots_ready(pb_sys *s) {
  if (!s->d_hash) {
    if (int rc = pb_alloc(s->ctx, &s->d_hash, HASH_SLOTS)) return rc;
  }
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(2 * HASH_SLOTS), 0, s->ctx->stream,
                     (uint32_t *)s->d_hash, 2u * HASH_SLOTS, 0u);
  return PB_OK;
}

static uint64_t hash_sum(const unsigned long long *slots) {
  uint64_t h = 0;
  for (uint32_t k = 0; k < HASH_SLOTS; ++k) h += slots[k];
  return h;
}

static int state_hash(pb_sys *s, uint64_t *h) {
  if (int rc = hash_slots_ready(s)) return rc;
  hipStream_t st = s->ctx->stream;
  hipLaunchKernelGGL(k_state_hash, dim3(hash_blocks(s)), dim3(256), 0, st, s->r, s->orig,
                     s->npad, s->d_hash);
  unsigned long long v[HASH_SLOTS];
  PB_HIP(s->ctx, hipMemcpyAsync(v, s->d_hash, sizeof(v), hipMemcpyDeviceToHost, st));
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  *h = hash_sum(v);
  return PB_OK;
}

extern "C" int pb_lj_run(pb_sys *s, const pb_run_opts *o, double *energies) {
  if (!s || !o) return PB_ERR_ARG;
  if (int rc = check_lj_ready(s)) return rc;
  if (o->nsteps == 0) return PB_OK;
  if (!(o->dt > 0) || o->skin < 0 || o->rebuild_every == 0)
    return pb_fail(s->ctx, PB_ERR_ARG, "pb_lj_run: bad options");
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const bool mixed = o->precision == PB_MIXED;
  // Per-replica skins: replica q's particles move with its thermal speed, so
  // a hot replica would reach a common skin first and force rebuilds of
  // all.  Its skin is skin * (T_q / T_min)^(1/4) (T from sum m v^2 per atom
  // at the start), capped below half its box.  The exponent 1/2 would make
  // every replica reach its skin after the same number of steps; 1/4 trades
  // a few more rebuilds for shorter lists of the hot replicas and measured
  // best on the 256-replica set (2.57e9 atom-steps/s against 2.48e9 with
  // 1/2, 2.54e9 with 1/3, 2.39e9 with 1/8 and 2.27e9 with one skin).  One
  // replica keeps the given skin.
  std::vector<double> skins(s->nrep, o->skin);
  if (s->nrep > 1 && o->skin > 0) {
    std::vector<double> mv2(s->nrep);
    if (int rc = pb_mv2(s, mv2.data())) return rc;
    double tmin = 0.0;
    std::vector<double> tq(s->nrep, 0.0);
    for (uint32_t q = 0; q < s->nrep; ++q) {
      if (!s->reps[q].natoms) continue;
      tq[q] = mv2[q] / s->reps[q].natoms;
      if (tq[q] > 0 && (tmin == 0.0 || tq[q] < tmin)) tmin = tq[q];
    }
    const double rc0 = std::sqrt(s->rc2);
    for (uint32_t q = 0; q < s->nrep && tmin > 0; ++q) {
      const RepInfo &R = s->reps[q];
      const double lmin = std::min(R.L[0], std::min(R.L[1], R.L[2]));
      const double cap = 0.5 * lmin * (1.0 - 1e-9) - rc0;
      const double sk = o->skin * std::pow(std::max(tq[q], tmin) / tmin, 0.25);
      skins[q] = std::max(o->skin, std::min(sk, cap));
    }
  }
  if (int rc = setup_list_cells(s, std::sqrt(s->rc2), skins)) return rc;
  const uint32_t ne = o->energy_every;
  const uint32_t nrec = ne ? (o->nsteps + ne - 1) / ne : 0;
  double *ebuf = nullptr;
  if (nrec) {
    if (int rc = pb_alloc(s->ctx, &ebuf, (size_t)nrec * 8 * s->nrep)) return rc;
  }
  int rc = PB_OK;
  uint32_t rebuilds = 0;
  int skin_bad = 0;
  bool cont = false;
  int64_t end_since = 0;             // steps since the last list build at the end
  std::vector<double> ehost;         // energy records read back at the end
  unsigned long long end_slots[HASH_SLOTS];   // fingerprint of the final state (slot sums)
  // optional HIP-event timing on the context's stream
  std::vector<hipEvent_t> ev_step, ev_graph;   // graph periods: R steps + rebuild
  // Eager steps: opts.timing = k times every k-th step and the step after
  // each rebuild, with the kernel's own timestamps; a rebuild is timed from
  // the end of the step before it (or a marker when that step was not
  // timed) to the start of the step after it.  A timed dispatch idles the
  // GPU several microseconds before it (about 2 us less with the
  // timing-only events below), so k > 1 keeps that cost off most steps.
  const uint32_t tstride = o->timing > 0 ? (uint32_t)o->timing : 0u;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> reb_spans;
  std::vector<hipEvent_t> ev_mark;                     // rebuild-start markers
  hipEvent_t reb_from = nullptr, prev_stop = nullptr;
  // timing-only events: no system-scope fence (no cache write-back and
  // invalidation around the timed work); they are read after the run's
  // synchronisation
  auto tev = [](hipEvent_t *e) {
    return hipEventCreateWithFlags(e, hipEventDisableSystemFence) == hipSuccess;
  };
  auto mk = [&](std::vector<hipEvent_t> &v) -> hipEvent_t {
    hipEvent_t e = nullptr;
    if (tev(&e)) v.push_back(e);
    return e;
  };
  auto rec = [&](std::vector<hipEvent_t> &v) {
    if (o->timing) {
      hipEvent_t e = mk(v);
      if (e) (void)hipEventRecord(e, st);
    }
  };
  // host-mapped ring for the lagged displacement reports
  if (!s->hmon) {
    if (hipHostMalloc((void **)&s->hmon, 32 * sizeof(uint32_t),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&s->dmon, s->hmon, 0) != hipSuccess) {
      rc = pb_fail(s->ctx, PB_ERR_HIP, "pb_lj_run: host-mapped monitor allocation failed");
      goto done;
    }
  }
  for (int k = 0; k < 32; ++k) ((volatile uint32_t *)s->hmon)[k] = 0u;
  // Continue with the list of the previous run when nothing has changed the
  // positions or the storage order since (in graph mode the first captured
  // period then starts where the eager schedule would rebuild)
  cont = false;
  if (s->lj_cont && s->lj_cont_rc2 == s->rc2 && s->lj_cont_skins == skins &&
      s->lj_cont_tile_w == s->tile_w && s->lj_cont_fill == s->tile_fill) {
    uint64_t h = 0;
    if ((rc = state_hash(s, &h))) goto done;
    cont = h == s->lj_cont_hash;
  }
  s->lj_cont = false;
  // the run's last step adds the fingerprint of the final state here
  if ((rc = hash_slots_ready(s))) goto done;
  // reference VV state -> kick-drift state (a continued list: particles
  // put back at the images nearest their list-build positions)
  hipLaunchKernelGGL(k_run_start, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, st, s->v, s->a,
                     s->r, s->rbuild, s->orig, s->waveRep, s->d_reps, s->npad, o->dt, cont ? 1 : 0);
  if (!cont) {
    if ((rc = build_checked(s, false))) goto done;
    ++rebuilds;
  }
  if ((rc = reset_counters(s))) goto done;
  if (o->persistent && !s->lj_queue) {
    // zeroed once; the persistent kernel resets the counters after each launch
    if ((rc = pb_alloc(s->ctx, &s->lj_queue, 16))) goto done;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, s->lj_queue, 16u, 0u);
  }
  {
    StepArgs A = make_args(s, o->dt);
    A.energy = ebuf;
    A.mon = s->dmon;
    A.queue = o->persistent ? s->lj_queue : nullptr;
    // steps are counted from the list build: a continued list is `since`
    // steps old at step 0
    const int64_t since0 = cont ? (int64_t)s->lj_cont_since : 0;
    int64_t last_build = -since0;
    uint32_t step = 0;
    // Graph mode: whole rebuild periods -- R steps followed by the list
    // rebuild that prepares the next period -- are captured once as a
    // hipGraph and replayed.  The schedule is the fixed interval R (no host
    // polling); a particle outside skin/2 is still caught by the step
    // kernel's check.  The step sequence of a period may leave the state in
    // the other ping-pong buffer, so up to two graphs (one per buffer
    // assignment) are kept.  The remaining steps (at least the last one,
    // which writes F) run eagerly.
    auto rebuild_now = [&]() -> int {
      // rebuilds after the first use the capacity found by the first one;
      // an overflow is reported at the end of the run
      if (tstride) {
        reb_from = prev_stop;
        if (!reb_from) {
          reb_from = mk(ev_mark);
          if (reb_from) (void)hipEventRecord(reb_from, st);
        }
      }
      if (int e = rebuild(s, false)) return e;
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, s->d_counter + CNT_DISP, 2u, 0u);
      ++rebuilds;
      last_build = step;
      A.nl = (const uint64_t *)s->nl;
      A.nn = s->nn;
      return PB_OK;
    };
    // eager steps [step, to) with the host-polled rebuild schedule
    auto eager = [&](uint32_t to) -> int {
    for (; step < to; ++step) {
      // Rebuild schedule: at most every rebuild_every steps, and earlier when
      // the particles may reach skin/2.  Kernel `step-1` published the
      // maxima after step `step-2` when it started, so reading them keeps one
      // kernel in flight; the one step of motion in between is bounded by
      // 1.5 * vmax * dt.  When that bound reaches skin/2 but one step at the
      // top speed seen so far does not, the maxima after step `step-1` --
      // the displacement of the positions kernel `step` uses -- are fetched
      // instead (one wait for the GPU) and decide alone.  The fastest
      // particles move ballistically over a rebuild period, so in about
      // half the periods this puts the rebuild one step later.
      // A continued list: the previous run's last kernel measured the
      // displacement of the positions kernel 0 uses (its maxima d0, v0), so
      // step 0 rebuilds only if they already reach skin/2, and step 1 decides
      // from d0, v0 as later steps do from the report of step `step-2`.
      bool reb = false;
      volatile uint32_t *mon = (volatile uint32_t *)s->hmon;
      if (step == 0 && cont) {
        float d2;
        std::memcpy(&d2, &s->lj_cont_d2, 4);
        // float rounding of the squared ratio: a margin of 1e-3
        reb = since0 >= (int64_t)o->rebuild_every || d2 >= 0.999f;
      } else if (step > 0) {
        if ((int64_t)step - last_build >= (int64_t)o->rebuild_every) {
          reb = true;
        } else if ((int64_t)step - last_build >= 2 && (step >= 2 || cont)) {
          float d2, v2;
          if (step >= 2) {
            const uint32_t slot = ((step - 2) & 7u) * 4u;
            const auto t0 = std::chrono::steady_clock::now();
            while (mon[slot + 3] != step - 1) {
              if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30)) {
                return pb_fail(s->ctx, PB_ERR_HIP, "pb_lj_run: step report timed out");
              }
              std::this_thread::yield();
            }
            std::atomic_thread_fence(std::memory_order_acquire);
            const uint32_t ud = mon[slot + 0], uv = mon[slot + 1];
            std::memcpy(&d2, &ud, 4);
            std::memcpy(&v2, &uv, 4);
          } else {
            std::memcpy(&d2, &s->lj_cont_d2, 4);
            std::memcpy(&v2, &s->lj_cont_v2, 4);
          }
          // maxima are in units of each replica's half skin
          const double dl = std::sqrt((double)d2), vdt = std::sqrt((double)v2) * o->dt;
          if (dl + vdt >= 1.0) {
            reb = true;
          } else if (dl + 1.5 * vdt >= 1.0) {
            const uint32_t eslot = ((step - 1) & 7u) * 4u;
            hipLaunchKernelGGL(k_publish_maxima, dim3(1), dim3(64), 0, st, s->d_counter, s->dmon,
                               step);
            PB_LAUNCH_CHECK(s->ctx, "k_publish_maxima");
            const auto t1 = std::chrono::steady_clock::now();
            while (mon[eslot + 3] != step) {
              if (std::chrono::steady_clock::now() - t1 > std::chrono::seconds(30)) {
                return pb_fail(s->ctx, PB_ERR_HIP, "pb_lj_run: step report timed out");
              }
              std::this_thread::yield();
            }
            std::atomic_thread_fence(std::memory_order_acquire);
            const uint32_t ue = mon[eslot + 0];
            float e2;
            std::memcpy(&e2, &ue, 4);
            if (e2 >= 0.999f) reb = true;
          }
        }
      }
      if (reb) {
        if (int e = rebuild_now()) return e;
      }
      const bool energy = ne && (step % ne == 0);
      const bool last = step + 1 == o->nsteps;
      A.rin = s->r;
      A.vin = s->v;
      A.rout = s->r2buf;
      A.vout = s->v2buf;
      A.Fout = s->F;
      A.Aout = s->a;
      A.orig = s->orig;
      A.hash = s->d_hash;
      A.step = step;
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (tstride && (reb_from || step % tstride == 0)) {
        if (tev(&e0) && tev(&e1)) {
          ev_step.push_back(e0);
          ev_step.push_back(e1);
          if (reb_from) reb_spans.emplace_back(reb_from, e0);
        } else {
          if (e0) (void)hipEventDestroy(e0);
          e0 = e1 = nullptr;
        }
        reb_from = nullptr;
      }
      if (int e = launch_step(s, A, mixed, energy, last, true, e0, e1)) return e;
      prev_stop = e1;
      std::swap(s->r, s->r2buf);
      std::swap(s->v, s->v2buf);
    }
    return PB_OK;
    };
    if (o->use_graph) {
      const uint32_t R = o->rebuild_every;
      if (ne && R % ne) {
        rc = pb_fail(s->ctx, PB_ERR_ARG, "pb_lj_run: with use_graph, energy_every must divide rebuild_every");
        goto done;
      }
      // a continued list: eager steps until it is R steps old (the eager
      // schedule's forced rebuild), then the rebuild, so every period starts
      // from a fresh list as in a run that built at its start
      if (cont) {
        const uint32_t prefix = (uint32_t)std::min<int64_t>(
            o->nsteps, std::max<int64_t>(0, (int64_t)R - since0));
        if ((rc = eager(prefix))) goto done;
        if (step < o->nsteps && (rc = rebuild_now())) goto done;
      }
      const uint32_t periods = step < o->nsteps ? (o->nsteps - step - 1) / R : 0u;
      // energies are recorded at steps s with s % ne == 0 in both schedules;
      // R % ne == 0, so every period records at the same offsets
      const uint32_t step0 = step;
      hipGraphExec_t gx[2] = {nullptr, nullptr};
      hipGraph_t gg[2] = {nullptr, nullptr};
      // a graph is valid for the buffer assignment it was captured with
      double4 *gkey[2] = {nullptr, nullptr};
      uint32_t *gkey2[2] = {nullptr, nullptr};
      A.mon = nullptr;
      auto capture = [&](int slot) -> int {
        PB_HIP(s->ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int e = PB_OK;
        for (uint32_t l = 0; l < R && e == PB_OK; ++l) {
          A.rin = s->r;
          A.vin = s->v;
          A.rout = s->r2buf;
          A.vout = s->v2buf;
          A.Fout = s->F;
          A.step = 0;
          e = launch_step(s, A, mixed, ne && ((step0 + l) % ne == 0), false, true);
          std::swap(s->r, s->r2buf);
          std::swap(s->v, s->v2buf);
        }
        if (e == PB_OK) e = rebuild(s, false);
        if (e == PB_OK)
          hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, s->d_counter + CNT_DISP, 2u, 0u);
        hipGraph_t g = nullptr;
        hipError_t he = hipStreamEndCapture(st, &g);
        if (e != PB_OK) {
          if (g) (void)hipGraphDestroy(g);
          return e;
        }
        if (he != hipSuccess) return pb_hip_check(s->ctx, he, "hipStreamEndCapture");
        gg[slot] = g;
        he = hipGraphInstantiate(&gx[slot], g, nullptr, nullptr, 0);
        if (he != hipSuccess) return pb_hip_check(s->ctx, he, "hipGraphInstantiate");
        return PB_OK;
      };
      for (uint32_t p = 0; p < periods && rc == PB_OK; ++p) {
        int slot = -1;
        for (int k = 0; k < 2; ++k)
          if (gx[k] && gkey[k] == s->r && gkey2[k] == s->orig) slot = k;
        if (slot < 0) {
          // capture records the launches without running them and applies
          // the period's buffer swaps to s: undo them before replaying
          double4 *r0 = s->r, *v0 = s->v, *a0 = s->a;
          uint32_t *o0 = s->orig;
          uint16_t *t0 = s->type;
          uint8_t *m0 = s->move;
          double4 *r1 = s->r2buf, *v1 = s->v2buf, *a1 = s->a2buf;
          uint32_t *o1 = s->orig2;
          uint16_t *t1 = s->type2;
          uint8_t *m1 = s->move2;
          slot = gx[0] ? 1 : 0;
          if ((rc = capture(slot))) break;
          s->r = r0; s->v = v0; s->a = a0; s->orig = o0; s->type = t0; s->move = m0;
          s->r2buf = r1; s->v2buf = v1; s->a2buf = a1; s->orig2 = o1; s->type2 = t1;
          s->move2 = m1;
          gkey[slot] = r0;
          gkey2[slot] = o0;
        }
        rec(ev_graph);
        hipError_t he = hipGraphLaunch(gx[slot], st);
        rec(ev_graph);
        if (he != hipSuccess) {
          rc = pb_hip_check(s->ctx, he, "hipGraphLaunch");
          break;
        }
        // the replayed period leaves the buffers as its capture did
        for (uint32_t l = 0; l < R; ++l) {
          std::swap(s->r, s->r2buf);
          std::swap(s->v, s->v2buf);
        }
        std::swap(s->r, s->r2buf);          // the rebuild's reorder (Fold is not carried)
        std::swap(s->v, s->v2buf);
        std::swap(s->orig, s->orig2);
        std::swap(s->type, s->type2);
        std::swap(s->move, s->move2);
        rebuilds += 1;
        step += R;
        last_build = step;
      }
      for (int k = 0; k < 2; ++k) {
        if (gx[k]) (void)hipGraphExecDestroy(gx[k]);
        if (gg[k]) (void)hipGraphDestroy(gg[k]);
      }
      if (rc) goto done;
      A.mon = s->dmon;
    }
    if ((rc = eager(o->nsteps))) goto done;
    end_since = (int64_t)o->nsteps - last_build;
  }
  // the last step left the reference VV state, the wrap and the
  // fingerprint; the fingerprint, the energies and the counters come back
  // with one synchronisation
  {
    hipError_t e = hipSuccess;
    if (energies && nrec) {
      ehost.resize((size_t)nrec * 8 * s->nrep);
      e = hipMemcpyAsync(ehost.data(), ebuf, sizeof(double) * ehost.size(), hipMemcpyDeviceToHost,
                         st);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(end_slots, s->d_hash, sizeof(end_slots), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) {
      rc = pb_hip_check(s->ctx, e, "pb_lj_run results");
      goto done;
    }
  }
  {
    uint32_t c[16];
    if ((rc = read_counters(s, c))) goto done;
    skin_bad = c[CNT_SKIN] ? 1 : 0;
    s->last_maxnn = std::max(s->last_maxnn, c[CNT_MAXNN]);
    if (c[CNT_WIN]) {
      rc = pb_fail(s->ctx, PB_ERR_CELLS, "a list window holds more than 65534 particles");
      goto done;
    }
    if (c[CNT_OVF]) {
      rc = pb_fail(s->ctx, PB_ERR_NLIST, "neighbour list overflow during the run");
      goto done;
    }
    if (!skin_bad) {
      // the list and cell order stay valid for a next run from this state
      s->lj_cont = true;
      s->lj_cont_hash = hash_sum(end_slots);
      s->lj_cont_rc2 = s->rc2;
      s->lj_cont_skins = skins;
      s->lj_cont_tile_w = s->tile_w;
      s->lj_cont_fill = s->tile_fill;
      s->lj_cont_since = (uint32_t)std::min<int64_t>(end_since, 0x7FFFFFFF);
      s->lj_cont_d2 = c[CNT_DISP];
      s->lj_cont_v2 = c[CNT_VMAX];
    }
  }
  if (energies && nrec) {   // copied before the counters' synchronisation
    for (uint32_t r = 0; r < nrec; ++r)
      for (uint32_t q = 0; q < s->nrep; ++q)
        for (int c = 0; c < 3; ++c)
          energies[((size_t)r * s->nrep + q) * 3 + c] = ehost[((size_t)r * s->nrep + q) * 8 + c];
  }
  s->last_rebuilds = rebuilds;
  s->last_skin_violation = skin_bad;
  s->last_step_ms = s->last_rebuild_ms = 0.0;
  s->last_nstep_timed = (uint32_t)(ev_step.size() / 2);
  s->last_nrebuild_timed = (uint32_t)reb_spans.size();
  for (size_t k = 0; k + 1 < ev_step.size(); k += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev_step[k], ev_step[k + 1]) == hipSuccess) s->last_step_ms += ms;
  }
  for (const auto &sp : reb_spans) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, sp.first, sp.second) == hipSuccess) s->last_rebuild_ms += ms;
  }
  // a replayed period counts as rebuild_every steps (its rebuild included)
  for (size_t k = 0; k + 1 < ev_graph.size(); k += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev_graph[k], ev_graph[k + 1]) == hipSuccess) s->last_step_ms += ms;
    s->last_nstep_timed += o->rebuild_every;
  }
  if (skin_bad) rc = pb_fail(s->ctx, PB_ERR_SKIN, "skin violated: increase skin or rebuild more often");
done:
  for (hipEvent_t e : ev_step) (void)hipEventDestroy(e);
  for (hipEvent_t e : ev_mark) (void)hipEventDestroy(e);
  for (hipEvent_t e : ev_graph) (void)hipEventDestroy(e);
  pb_free(ebuf);
  return rc;
}

extern "C" int pb_lj_last_run_timing(pb_sys *s, double *step_ms, uint32_t *nstep,
                                     double *rebuild_ms, uint32_t *nrebuild) {
  if (!s) return PB_ERR_ARG;
  if (step_ms) *step_ms = s->last_step_ms;
  if (nstep) *nstep = s->last_nstep_timed;
  if (rebuild_ms) *rebuild_ms = s->last_rebuild_ms;
  if (nrebuild) *nrebuild = s->last_nrebuild_timed;
  return PB_OK;
}

extern "C" int pb_lj_last_run_stats(pb_sys *s, uint32_t *rebuilds, uint32_t *max_nn,
                                    int *skin_violation) {
  if (!s) return PB_ERR_ARG;
  if (rebuilds) *rebuilds = s->last_rebuilds;
  if (max_nn) *max_nn = s->last_maxnn;
  if (skin_violation) *skin_violation = s->last_skin_violation;
  return PB_OK;
}

// ------------------------------------------------------------- autotuner
// Tile shape of the list build and the step kernel: columns per tile side
// (1 or 2) and the fill of the LDS stage that sets the tile depth nz.
// Every candidate runs `steps` fused steps from the same saved state, timed
// with HIP events on the context's stream; the fastest is kept and the
// state is restored, so the call leaves the simulation where it was.
extern "C" int pb_lj_set_tiles(pb_sys *s, uint32_t tile_w, double tile_fill) {
  if (!s) return PB_ERR_ARG;
  if ((tile_w != 1 && tile_w != 2) || !(tile_fill >= 0.1 && tile_fill <= 1.0))
    return pb_fail(s->ctx, PB_ERR_ARG, "pb_lj_set_tiles: tile_w must be 1 or 2, tile_fill in [0.1, 1]");
  s->tile_w = tile_w;
  s->tile_fill = tile_fill;
  return PB_OK;
}

extern "C" int pb_lj_autotune(pb_sys *s, const pb_run_opts *o, uint32_t steps, double *best_ms,
                              uint32_t *best_w, double *best_fill) {
  if (!s || !o || steps < 2) return PB_ERR_ARG;
  if (int rc = check_lj_ready(s)) return rc;
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const size_t np = s->npad;
  double4 *sr = nullptr, *sv = nullptr, *sa = nullptr, *sf = nullptr;
  uint32_t *so = nullptr;
  uint16_t *sty = nullptr;
  uint8_t *smv = nullptr;
  int rc = PB_OK;
  auto copy_all = [&](bool save) -> int {
    auto cp = [&](void *dst, const void *src, size_t bytes) {
      return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
    };
    hipError_t e = hipSuccess;
    if (save) {
      e = cp(sr, s->r, 32 * np);
      if (e == hipSuccess) e = cp(sv, s->v, 32 * np);
      if (e == hipSuccess) e = cp(sa, s->a, 32 * np);
      if (e == hipSuccess) e = cp(sf, s->F, 32 * np);
      if (e == hipSuccess) e = cp(so, s->orig, 4 * np);
      if (e == hipSuccess) e = cp(sty, s->type, 2 * np);
      if (e == hipSuccess) e = cp(smv, s->move, np);
    } else {
      e = cp(s->r, sr, 32 * np);
      if (e == hipSuccess) e = cp(s->v, sv, 32 * np);
      if (e == hipSuccess) e = cp(s->a, sa, 32 * np);
      if (e == hipSuccess) e = cp(s->F, sf, 32 * np);
      if (e == hipSuccess) e = cp(s->orig, so, 4 * np);
      if (e == hipSuccess) e = cp(s->type, sty, 2 * np);
      if (e == hipSuccess) e = cp(s->move, smv, np);
    }
    if (e != hipSuccess) return pb_hip_check(s->ctx, e, "pb_lj_autotune copy");
    return PB_OK;
  };
  const uint32_t ws[2] = {1, 2};
  const double fills[6] = {0.55, 0.72, 0.85, 0.9, 0.95, 1.0};
  const uint32_t w0 = s->tile_w;
  const double f0 = s->tile_fill;
  double best = 1e300;
  uint32_t bw = w0;
  double bf = f0;
  if ((rc = pb_alloc(s->ctx, &sr, np)) || (rc = pb_alloc(s->ctx, &sv, np)) ||
      (rc = pb_alloc(s->ctx, &sa, np)) || (rc = pb_alloc(s->ctx, &sf, np)) ||
      (rc = pb_alloc(s->ctx, &so, np)) || (rc = pb_alloc(s->ctx, &sty, np)) ||
      (rc = pb_alloc(s->ctx, &smv, np)))
    goto out;
  if ((rc = copy_all(true))) goto out;
  for (uint32_t w : ws)
    for (double f : fills) {
      s->tile_w = w;
      s->tile_fill = f;
      if ((rc = copy_all(false))) goto out;
      pb_run_opts t = *o;
      t.nsteps = steps;
      t.energy_every = 0;
      t.timing = 1;
      t.use_graph = 0;
      rc = pb_lj_run(s, &t, nullptr);
      if (rc == PB_ERR_SKIN || rc == PB_ERR_CELLS) {   // candidate unusable here
        rc = PB_OK;
        continue;
      }
      if (rc) goto out;
      const double ms = (s->last_step_ms + s->last_rebuild_ms) / steps;
      if (ms < best) {
        best = ms;
        bw = w;
        bf = f;
      }
    }
  if (!(best < 1e300)) {
    rc = pb_fail(s->ctx, PB_ERR_STATE, "pb_lj_autotune: no tile shape could run");
    goto out;
  }
  s->tile_w = bw;
  s->tile_fill = bf;
  if ((rc = copy_all(false))) goto out;
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  if (best_ms) *best_ms = best;
  if (best_w) *best_w = bw;
  if (best_fill) *best_fill = bf;
out:
  if (rc) {
    s->tile_w = w0;
    s->tile_fill = f0;
  }
  s->lj_cont = false;             // the restored state has no list of its own
  (void)hipStreamSynchronize(st);
  pb_free(sr); pb_free(sv); pb_free(sa); pb_free(sf); pb_free(so); pb_free(sty); pb_free(smv);
  return rc;
}
