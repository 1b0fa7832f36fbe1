// pb_bin.hip -- the integer-output paths that must match the reference bit
// for bit: cell ids (calcBID_F4), the cell sort (calcBlockID), neighbour
// counts (calcCoordMD_F4) and g(r) histograms (calcGR_F4).  This file is
// compiled with -ffp-contract=off: every FP32 operation is rounded exactly as
// the reference's source orders it (division and sqrt correctly rounded).
#include "pb_internal.h"
#include <cmath>
#include <cstring>

int pb_ensure_sort_ws(pb_sys *s, uint32_t ncell);
int pb_counting_sort(pb_sys *s, uint32_t ncell, const uint32_t *key);

// CUDA's float->uint conversion saturates; calcBID_F4 relies on it.
__device__ __host__ __forceinline__ uint32_t f2u_sat(float q) {
  if (!(q > 0.0f)) return 0u;
  if (q >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)q;
}
__device__ __forceinline__ uint32_t clamp_axis(uint32_t d, uint32_t c) {
  // kernelfuncs.cu:821-835
  return (d == c) ? c - 1 : ((d > c) ? 0u : d);
}

// distance2p, kernelinline.hh:47-62
__device__ __forceinline__ float distance2p(float r1x, float r1y, float r1z, float r2x,
                                            float r2y, float r2z, float cx, float cy, float cz) {
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

struct RefGrid {
  float len[3], lo[3];
  uint32_t nb[3];
};

// calcBID_F4, kernelfuncs.cu:807-841; padding slots go to the extra cell
// ncell so they sort after every real particle.
__global__ void k_ref_bid(const double4 *r, const uint32_t *orig, uint32_t n, RefGrid g,
                          uint32_t ncell, uint32_t *cellid) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (orig[s] == PB_NONE) {
    cellid[s] = ncell;
    return;
  }
  const double4 p = r[s];
  const float x = (float)p.x, y = (float)p.y, z = (float)p.z;
  const uint32_t d0 = clamp_axis(f2u_sat((x - g.lo[0]) / g.len[0]), g.nb[0]);
  const uint32_t d1 = clamp_axis(f2u_sat((y - g.lo[1]) / g.len[1]), g.nb[1]);
  const uint32_t d2 = clamp_axis(f2u_sat((z - g.lo[2]) / g.len[2]), g.nb[2]);
  cellid[s] = d2 + d1 * g.nb[2] + d0 * g.nb[1] * g.nb[2];
}

__global__ void k_ref_outputs(const uint32_t *perm, const uint32_t *cellid, const uint32_t *orig,
                              const uint32_t *count, const uint32_t *start, uint32_t nreal,
                              uint32_t ncell, uint32_t *pid, uint32_t *bid, uint32_t *bindex) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nreal) {
    const uint32_t s = perm[k];
    pid[k] = orig[s];
    bid[k] = cellid[s];
  }
  if (k < ncell) bindex[k] = count[k] ? start[k] : PB_NONE;
  if (k == 0) bindex[ncell] = nreal;
}

static void host_neighbor_table(const uint32_t bn[3], int w, int periodic, uint32_t *table) {
  // calc_BlockNeighbor, cudaCutoffBlock.cu:17-65
  const long bx = bn[0], by = bn[1], bz = bn[2];
  const int J = (2 * w + 1) * (2 * w + 1) * (2 * w + 1);
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
              uint32_t v = PB_NONE;
              if (periodic || ((i + i1) >= 0 && (i + i1) < bx && (j + j1) >= 0 && (j + j1) < by &&
                               (k + k1) >= 0 && (k + k1) < bz))
                v = (uint32_t)(k2 + j2 * bz + i2 * bz * by);
              table[I * J + L++] = v;
            }
      }
}

extern "C" int pb_setup_cells(pb_sys *s, float rmax, int periodic, int single_particle) {
  if (!s) return PB_ERR_ARG;
  if (s->nrep != 1) return pb_fail(s->ctx, PB_ERR_ARG, "pb_setup_cells: single replica only");
  if (!s->cell_set) return pb_fail(s->ctx, PB_ERR_STATE, "pb_setup_cells: cell not set");
  if (!(rmax > 0)) return pb_fail(s->ctx, PB_ERR_ARG, "pb_setup_cells: rmax <= 0");
  // cudaCutoffBlock::setupCutoffBlock, cudaCutoffBlock.cu:97-141 (FP32)
  float cell[6];
  for (int i = 0; i < 6; ++i) cell[i] = (float)s->cells[i];
  float c[3];
  for (int i = 0; i < 3; ++i) {
    c[i] = cell[i * 2 + 1] - cell[i * 2];
    if (periodic) {
      s->blocknum[i] = f2u_sat(c[i] / rmax);
      if (s->blocknum[i] == 0) return pb_fail(s->ctx, PB_ERR_CELLS, "pb_setup_cells: box smaller than rmax");
      s->blocklen[i] = c[i] / (float)s->blocknum[i];
    } else {
      s->blocklen[i] = rmax;
      s->blocknum[i] = f2u_sat(c[i] / rmax) + 1u;
    }
    if (s->blocknum[i] <= 2)
      return pb_fail(s->ctx, PB_ERR_CELLS, "pb_setup_cells: fewer than 3 cells along an axis");
    s->blocklen[i + 3] = cell[i * 2];
  }
  if (single_particle) {
    // cudaCutoffBlock.cu:115-138: shrink cells until the cell diagonal is
    // below the particle diameter (r0 = rmax*sqrt(3)/2)
    const double r0 = rmax * sqrt(3.0) / 2.0;
    const double R02 = (r0 * 2.0) * (r0 * 2.0);
    while ((double)(s->blocklen[0] * s->blocklen[0] + s->blocklen[1] * s->blocklen[1] +
                    s->blocklen[2] * s->blocklen[2]) > R02) {
      if ((s->blocklen[0] > s->blocklen[1]) && (s->blocklen[0] > s->blocklen[2])) {
        ++s->blocknum[0];
        s->blocklen[0] = c[0] / s->blocknum[0];
      } else if (s->blocklen[1] > s->blocklen[2]) {
        ++s->blocknum[1];
        s->blocklen[1] = c[1] / s->blocknum[1];
      } else {
        ++s->blocknum[2];
        s->blocklen[2] = c[2] / s->blocknum[2];
      }
    }
  }
  const uint64_t nc = (uint64_t)s->blocknum[0] * s->blocknum[1] * s->blocknum[2];
  if (nc >= 0x7FFFFFFFull) return pb_fail(s->ctx, PB_ERR_CELLS, "pb_setup_cells: too many cells");
  s->ref_ncell = (uint32_t)nc;
  s->ref_periodic = periodic ? 1 : 0;
  s->ref_single = single_particle ? 1 : 0;
  int rc;
  if ((rc = pb_alloc(s->ctx, &s->ref_bid, s->npad)) || (rc = pb_alloc(s->ctx, &s->ref_pid, s->npad)) ||
      (rc = pb_alloc(s->ctx, &s->ref_bindex, (size_t)s->ref_ncell + 1)))
    return rc;
  if ((rc = pb_ensure_sort_ws(s, s->ref_ncell + 1))) return rc;
  s->ref_cells = true;
  s->ref_binned = false;
  return PB_OK;
}

extern "C" int pb_get_cell_info(pb_sys *s, uint32_t blocknum[3], float blocklen[6]) {
  if (!s || !s->ref_cells) return pb_fail(s ? s->ctx : nullptr, PB_ERR_STATE, "no cells");
  if (blocknum) std::memcpy(blocknum, s->blocknum, sizeof(s->blocknum));
  if (blocklen) std::memcpy(blocklen, s->blocklen, sizeof(s->blocklen));
  return PB_OK;
}

extern "C" int pb_get_neighbor_table(pb_sys *s, uint32_t *table) {
  if (!s || !table) return PB_ERR_ARG;
  if (!s->ref_cells) return pb_fail(s->ctx, PB_ERR_STATE, "no cells");
  host_neighbor_table(s->blocknum, s->ref_single ? 2 : 1, s->ref_periodic, table);
  return PB_OK;
}

static RefGrid ref_grid(const pb_sys *s) {
  RefGrid g;
  for (int a = 0; a < 3; ++a) {
    g.len[a] = s->blocklen[a];
    g.lo[a] = s->blocklen[3 + a];
    g.nb[a] = s->blocknum[a];
  }
  return g;
}

extern "C" int pb_bin(pb_sys *s) {
  if (!s) return PB_ERR_ARG;
  if (!s->ref_cells) return pb_fail(s->ctx, PB_ERR_STATE, "pb_bin: pb_setup_cells first");
  hipStream_t st = s->ctx->stream;
  const uint32_t nc = s->ref_ncell;
  hipLaunchKernelGGL(k_ref_bid, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, st, s->r, s->orig,
                     s->npad, ref_grid(s), nc, s->cellid);
  PB_LAUNCH_CHECK(s->ctx, "k_ref_bid");
  if (int rc = pb_counting_sort(s, nc + 1, s->orig)) return rc;
  const uint32_t m = (s->n > nc + 1) ? s->n : nc + 1;
  hipLaunchKernelGGL(k_ref_outputs, dim3(pb_blocks(m, 256)), dim3(256), 0, st, s->perm, s->cellid,
                     s->orig, s->ccount, s->cstart, s->n, nc, s->ref_pid, s->ref_bid,
                     s->ref_bindex);
  PB_LAUNCH_CHECK(s->ctx, "k_ref_outputs");
  s->ref_binned = true;
  return PB_OK;
}

extern "C" int pb_get_bins(pb_sys *s, uint32_t *pid, uint32_t *bid_sorted, uint32_t *bindex) {
  if (!s) return PB_ERR_ARG;
  if (!s->ref_binned) return pb_fail(s->ctx, PB_ERR_STATE, "pb_get_bins: pb_bin first");
  hipStream_t st = s->ctx->stream;
  if (pid) PB_HIP(s->ctx, hipMemcpyAsync(pid, s->ref_pid, 4 * (size_t)s->n, hipMemcpyDeviceToHost, st));
  if (bid_sorted)
    PB_HIP(s->ctx, hipMemcpyAsync(bid_sorted, s->ref_bid, 4 * (size_t)s->n, hipMemcpyDeviceToHost, st));
  if (bindex)
    PB_HIP(s->ctx, hipMemcpyAsync(bindex, s->ref_bindex, 4 * ((size_t)s->ref_ncell + 1),
                                  hipMemcpyDeviceToHost, st));
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  return PB_OK;
}

// ------------------------------------------------------- neighbour counts
// calcCoordMD_F4 (kernelinline.hh:387-399) over the 27 cells of
// calcF_IJpairWithBlock_F4 (kerneltemplate.hh:244-293); cells computed on
// the fly with the wrap of calc_BlockNeighbor (periodic, nb >= 3).
__global__ void k_coord_counts(const double4 *r, const uint32_t *perm, const uint32_t *cellid,
                               const uint32_t *start, uint32_t nreal, RefGrid g, float cx,
                               float cy, float cz, float rmax2t, int periodic, uint32_t *out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nreal) return;
  const uint32_t si = perm[k];
  const uint32_t c = cellid[si];
  const uint32_t nyz = g.nb[1] * g.nb[2];
  const int i0 = (int)(c / nyz), j0 = (int)((c / g.nb[2]) % g.nb[1]), k0 = (int)(c % g.nb[2]);
  const double4 pi = r[si];
  const float xi = (float)pi.x, yi = (float)pi.y, zi = (float)pi.z;
  uint32_t cnt = 0;
  for (int a = -1; a <= 1; ++a) {
    int i2 = i0 + a;
    if (i2 < 0 || i2 >= (int)g.nb[0]) {
      if (!periodic) continue;
      i2 = (i2 + (int)g.nb[0]) % (int)g.nb[0];
    }
    for (int b = -1; b <= 1; ++b) {
      int j2 = j0 + b;
      if (j2 < 0 || j2 >= (int)g.nb[1]) {
        if (!periodic) continue;
        j2 = (j2 + (int)g.nb[1]) % (int)g.nb[1];
      }
      for (int d = -1; d <= 1; ++d) {
        int k2 = k0 + d;
        if (k2 < 0 || k2 >= (int)g.nb[2]) {
          if (!periodic) continue;
          k2 = (k2 + (int)g.nb[2]) % (int)g.nb[2];
        }
        const uint32_t J = (uint32_t)k2 + (uint32_t)j2 * g.nb[2] + (uint32_t)i2 * nyz;
        for (uint32_t jj = start[J]; jj < start[J + 1]; ++jj) {
          const uint32_t sj = perm[jj];
          if (sj == si) continue;
          const double4 pj = r[sj];
          const float R2 = distance2p(xi, yi, zi, (float)pj.x, (float)pj.y, (float)pj.z, cx, cy, cz);
          if (R2 < rmax2t) ++cnt;
        }
      }
    }
  }
  out[si] = cnt;
}

static void half_cell2(const pb_sys *s, uint32_t q, float c[3]) {
  // cudaFrame::setCell, cudaFrame.cu:85-91, on the FP32 cell values
  for (int a = 0; a < 3; ++a) {
    const float lo = (float)s->cells[6 * q + 2 * a], hi = (float)s->cells[6 * q + 2 * a + 1];
    c[a] = (hi - lo) / 2;
    c[a] *= c[a];
  }
}

extern "C" int pb_coord_counts(pb_sys *s, float rmax2t, uint32_t *counts) {
  if (!s || !counts) return PB_ERR_ARG;
  if (!s->ref_cells || s->ref_single)
    return pb_fail(s->ctx, PB_ERR_STATE, "pb_coord_counts: needs a 27-neighbour cell grid");
  float rc = sqrtf(rmax2t);
  for (int a = 0; a < 3; ++a)
    if (s->blocklen[a] < rc)
      return pb_fail(s->ctx, PB_ERR_CELLS, "pb_coord_counts: cell length below the count radius");
  if (int rc2 = pb_bin(s)) return rc2;
  float c[3];
  half_cell2(s, 0, c);
  uint32_t *dcnt = nullptr;
  if (int rc2 = pb_alloc(s->ctx, &dcnt, s->npad)) return rc2;
  hipStream_t st = s->ctx->stream;
  hipLaunchKernelGGL(k_coord_counts, dim3(pb_blocks(s->n, 256)), dim3(256), 0, st, s->r, s->perm,
                     s->cellid, s->cstart, s->n, ref_grid(s), c[0], c[1], c[2], rmax2t,
                     s->ref_periodic, dcnt);
  hipError_t e = hipGetLastError();
  std::vector<uint32_t> h(s->npad), o(s->npad);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dcnt, 4 * (size_t)s->npad, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(o.data(), s->orig, 4 * (size_t)s->npad, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  pb_free(dcnt);
  if (e != hipSuccess) return pb_hip_check(s->ctx, e, "pb_coord_counts");
  for (uint32_t k = 0; k < s->npad; ++k)
    if (o[k] != PB_NONE) counts[o[k]] = h[k];
  return PB_OK;
}

// ------------------------------------------------------------------ g(r)
// All unordered pairs of every replica (the pair set of
// calcF_IJpairWithBlock5_F4 over makePairInfo's tiles, kerneltemplate.hh:
// 423-512, cudaCalcGR.cu:30-84), binned by calcGR_F4 (kernelinline.hh:
// 440-448).  Work = triangular tile pairs of every replica, found by a
// binary search over the per-replica tile-pair prefix.  Each workgroup bins
// into an LDS histogram and flushes it with 64-bit atomics: integer sums, so
// the result is independent of the order.  Bin 0 (the pairs the reference
// drops: r < dr/2 and r beyond the last bin, about half of all pairs in a
// periodic box) is counted in registers and added once per tile pair, so
// it is not a serialising LDS atomic hot spot.
#define GR_T 256
// A tile pair adds at most GR_T * GR_T = 2^16 pairs to one bin of the
// workgroup's uint32 LDS histogram; flushing at least every 2^15 tile pairs
// keeps every bin below 2^31, whatever the replica's size.
#define GR_FLUSH_TILES 32768u
struct GrRep {
  uint32_t atom0, natoms, ntile;
  uint64_t pair0;      // first tile-pair index of this replica
  float cx, cy, cz, rstepinv;
};

__global__ void __launch_bounds__(GR_T) k_gr(const double4 *r, const GrRep *reps, uint32_t nrep,
                                             uint64_t npairs, uint32_t rnum, uint64_t *hist) {
  extern __shared__ uint32_t h[];
  __shared__ float4 tj[GR_T];
  for (uint32_t b = threadIdx.x; b < rnum; b += GR_T) h[b] = 0;
  uint32_t since = 0;                  // tile pairs in the LDS histogram since its last flush
  for (uint64_t p = blockIdx.x; p < npairs; p += gridDim.x) {
    uint32_t lo = 0, hi = nrep;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (reps[mid].pair0 <= p) lo = mid;
      else hi = mid;
    }
    const GrRep R = reps[lo];
    // triangular index -> (ti <= tj)
    uint64_t t = p - R.pair0;
    uint32_t ti = 0;
    uint32_t row = R.ntile;
    while (t >= row) {
      t -= row;
      ++ti;
      --row;
    }
    const uint32_t tjx = ti + (uint32_t)t;
    uint32_t zero = 0;                 // this thread's bin-0 pairs of the tile pair
    __syncthreads();
    const uint32_t j = tjx * GR_T + threadIdx.x;
    if (j < R.natoms) {
      const double4 q = r[R.atom0 + j];
      tj[threadIdx.x] = make_float4((float)q.x, (float)q.y, (float)q.z, 0.f);
    }
    __syncthreads();
    const uint32_t i = ti * GR_T + threadIdx.x;
    if (i < R.natoms) {
      const double4 q = r[R.atom0 + i];
      const float xi = (float)q.x, yi = (float)q.y, zi = (float)q.z;
      const uint32_t jn = min(GR_T, R.natoms - tjx * GR_T);
      const uint32_t jb = (ti == tjx) ? threadIdx.x + 1 : 0;
      for (uint32_t jj = jb; jj < jn; ++jj) {
        const float4 pj = tj[jj];
        const float R2 = distance2p(xi, yi, zi, pj.x, pj.y, pj.z, R.cx, R.cy, R.cz);
        const float r0 = sqrtf(R2) * R.rstepinv;
        const double tt = (double)r0 + 0.5;
        const uint32_t bin = (tt < (double)rnum) ? (uint32_t)tt : 0u;
        if (bin == 0u) ++zero;
        else atomicAdd(&h[bin], 1u);
      }
    }
    {
      const uint32_t wz = pb_wave_sum_u32(zero);
      if (pb_lane() == 0 && wz) atomicAdd(&h[0], wz);
      zero = 0;
    }
    // flush when the replica changes, at the end (histograms are per
    // replica) and every GR_FLUSH_TILES tile pairs (no uint32 bin can wrap)
    const uint64_t pn = p + gridDim.x;
    bool flush = pn >= npairs || ++since >= GR_FLUSH_TILES;
    if (!flush) {
      flush = !(pn >= R.pair0 && (lo + 1 >= nrep || pn < reps[lo + 1].pair0));
    }
    __syncthreads();
    if (flush) {
      for (uint32_t b = threadIdx.x; b < rnum; b += GR_T) {
        const uint32_t v = h[b];
        if (v) atomicAdd((unsigned long long *)&hist[(uint64_t)lo * rnum + b], (unsigned long long)v);
        h[b] = 0;
      }
      since = 0;
    }
  }
}

extern "C" int pb_gr_hist(pb_sys *s, const float *rstep, uint32_t rnum, uint64_t *hist) {
  if (!s || !rstep || !hist || rnum == 0) return PB_ERR_ARG;
  if (rnum > 16384) return pb_fail(s->ctx, PB_ERR_ARG, "pb_gr_hist: rnum > 16384");
  if (!s->cell_set) return pb_fail(s->ctx, PB_ERR_STATE, "pb_gr_hist: cell not set");
  std::vector<GrRep> hr(s->nrep);
  uint64_t tot = 0;
  for (uint32_t q = 0; q < s->nrep; ++q) {
    GrRep &g = hr[q];
    g.atom0 = s->reps[q].atom0;
    g.natoms = s->reps[q].natoms;
    g.ntile = (g.natoms + GR_T - 1) / GR_T;
    g.pair0 = tot;
    tot += (uint64_t)g.ntile * (g.ntile + 1) / 2;
    float c[3];
    half_cell2(s, q, c);
    g.cx = c[0];
    g.cy = c[1];
    g.cz = c[2];
    // cudaCalcGR::setup: rstepinv = 1.0 / rstep (double) stored as real
    g.rstepinv = (float)(1.0 / (double)rstep[q]);
  }
  // storage order must hold each replica's real particles at [atom0, atom0+natoms):
  // true at all times (padding sorts after real particles of its replica).
  GrRep *dreps = nullptr;
  const size_t hn = (size_t)s->nrep * rnum;
  if (s->gr_cap < hn) {
    if (int rc = pb_alloc(s->ctx, &s->gr_hist, hn)) return rc;
    s->gr_cap = hn;
  }
  if (int rc = pb_alloc(s->ctx, &dreps, s->nrep)) return rc;
  hipStream_t st = s->ctx->stream;
  hipError_t e = hipMemcpyAsync(dreps, hr.data(), sizeof(GrRep) * s->nrep, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(s->gr_hist, 0, sizeof(uint64_t) * hn, st);
  if (e == hipSuccess) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>(tot, 8192);
    hipLaunchKernelGGL(k_gr, dim3(grid), dim3(GR_T), sizeof(uint32_t) * rnum, st, s->r, dreps,
                       s->nrep, tot, rnum, s->gr_hist);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hist, s->gr_hist, sizeof(uint64_t) * hn, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  pb_free(dreps);
  if (e != hipSuccess) return pb_hip_check(s->ctx, e, "pb_gr_hist");
  return PB_OK;
}

// ---------------------------------------- g(r) without allocations
// pb_gr_hist_noalloc keeps its replica table and histograms in this
// file-scope device array instead of library-allocated memory: each device
// holds its own copy for the life of the process.  Host code never passes
// the array's name to a kernel (on the host it is only a stand-in symbol);
// it uses the device address that pb_gr_scratch_init stores in the context.
#define GR_NA_MAX_REPS 4096u                          // replicas per call
#define GR_NA_MAX_SLOTS 131072u                       // nrep * rnum histogram slots
#define GR_NA_REP_BYTES ((size_t)GR_NA_MAX_REPS * sizeof(GrRep))
#define GR_NA_RSTEP_BYTES ((size_t)GR_NA_MAX_REPS * sizeof(float))
#define GR_NA_BYTES (GR_NA_REP_BYTES + GR_NA_RSTEP_BYTES + (size_t)GR_NA_MAX_SLOTS * 8)
static __device__ __align__(256) unsigned char scratch[GR_NA_BYTES];

// the device address of `scratch` on the context's (current) device
int pb_gr_scratch_init(pb_ctx *c) {
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return pb_hip_check(c, e != hipSuccess ? e : hipErrorUnknown,
                                                   "hipGetSymbolAddress(g(r) scratch)");
  c->gr_scratch = ptr;
  return PB_OK;
}

// replica table of the g(r) kernel from the device replica records: one
// workgroup; per-replica tile-pair counts scanned in order (uint64)
__global__ void __launch_bounds__(256) k_gr_prep(const RepInfo *reps, const float *rstep,
                                                 uint32_t nrep, GrRep *out) {
  __shared__ unsigned long long part[256];
  const uint32_t per = (nrep + 255u) / 256u;
  const uint32_t q0 = threadIdx.x * per, q1 = min(nrep, q0 + per);
  unsigned long long mine = 0;
  for (uint32_t q = q0; q < q1; ++q) {
    const unsigned long long nt = (reps[q].natoms + GR_T - 1) / GR_T;
    mine += nt * (nt + 1) / 2;
  }
  part[threadIdx.x] = mine;
  __syncthreads();
  unsigned long long base = 0;
  for (uint32_t k = 0; k < threadIdx.x; ++k) base += part[k];
  for (uint32_t q = q0; q < q1; ++q) {
    const RepInfo &R = reps[q];
    GrRep g;
    g.atom0 = R.atom0;
    g.natoms = R.natoms;
    g.ntile = (R.natoms + GR_T - 1) / GR_T;
    g.pair0 = base;
    base += (unsigned long long)g.ntile * (g.ntile + 1) / 2;
    // cudaFrame::setCell (cudaFrame.cu:85-91) on the FP32 cell values, as
    // the host-side table of pb_gr_hist
    float c[3];
    for (int a = 0; a < 3; ++a) {
      const float lo = (float)R.lo[a], hi = (float)R.hi[a];
      c[a] = (hi - lo) * 0.5f;            // = (hi - lo) / 2 exactly
      c[a] *= c[a];
    }
    g.cx = c[0];
    g.cy = c[1];
    g.cz = c[2];
    g.rstepinv = (float)(1.0 / (double)rstep[q]);
    out[q] = g;
  }
}

__global__ void k_zero_u64(uint64_t *x, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

extern "C" int pb_gr_hist_noalloc(pb_sys *s, const float *rstep, uint32_t rnum, uint64_t *hist) {
  if (!s || !rstep || !hist || rnum == 0) return PB_ERR_ARG;
  if (rnum > 16384) return pb_fail(s->ctx, PB_ERR_ARG, "pb_gr_hist_noalloc: rnum > 16384");
  if (!s->cell_set) return pb_fail(s->ctx, PB_ERR_STATE, "pb_gr_hist_noalloc: cell not set");
  if (s->nrep > GR_NA_MAX_REPS)
    return pb_fail(s->ctx, PB_ERR_ARG, "pb_gr_hist_noalloc: more than 4096 replicas");
  if ((uint64_t)s->nrep * rnum > GR_NA_MAX_SLOTS)
    return pb_fail(s->ctx, PB_ERR_ARG,
                   "pb_gr_hist_noalloc: replicas x bins above the scratch limit of 131072");
  if (!s->ctx->gr_scratch)
    return pb_fail(s->ctx, PB_ERR_STATE, "pb_gr_hist_noalloc: no scratch address");
  for (uint32_t q = 0; q < s->nrep; ++q)
    if (!(rstep[q] > 0.0f)) return pb_fail(s->ctx, PB_ERR_ARG, "pb_gr_hist_noalloc: rstep <= 0");
  // tile pairs of all replicas (host replica records, no allocation)
  uint64_t tot = 0;
  for (uint32_t q = 0; q < s->nrep; ++q) {
    const uint64_t nt = (s->reps[q].natoms + GR_T - 1) / GR_T;
    tot += nt * (nt + 1) / 2;
  }
  unsigned char *base = static_cast<unsigned char *>(s->ctx->gr_scratch);
  GrRep *dreps = reinterpret_cast<GrRep *>(base);
  float *drstep = reinterpret_cast<float *>(base + GR_NA_REP_BYTES);
  uint64_t *dhist = reinterpret_cast<uint64_t *>(base + GR_NA_REP_BYTES + GR_NA_RSTEP_BYTES);
  const uint32_t hn = s->nrep * rnum;
  hipStream_t st = s->ctx->stream;
  hipError_t e = hipMemcpyAsync(drstep, rstep, sizeof(float) * s->nrep, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_gr_prep, dim3(1), dim3(256), 0, st, s->d_reps, drstep, s->nrep, dreps);
    hipLaunchKernelGGL(k_zero_u64, dim3(pb_blocks(hn, 256)), dim3(256), 0, st, dhist, hn);
    if (tot) {
      const uint32_t grid = (uint32_t)std::min<uint64_t>(tot, 8192);
      hipLaunchKernelGGL(k_gr, dim3(grid), dim3(GR_T), sizeof(uint32_t) * rnum, st, s->r, dreps,
                         s->nrep, tot, rnum, dhist);
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hist, dhist, sizeof(uint64_t) * hn, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return pb_hip_check(s->ctx, e, "pb_gr_hist_noalloc");
  return PB_OK;
}

__global__ void k_clear4_bin(double4 *x, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// ------------------------------------------- MD neighbour list (CSR) path
// cudaParticleMD::useNList / makeNList (cudaParticleMD.cu:338-394;
// calcCoordMD_F4 counts, makeJlist_WithBlock_F4 kerneltemplate.hh:725-775,
// MDNeighborRegister kernelinline.hh:404-412) and the forces of
// calcF_IJpairWithList_F4 (kerneltemplate.hh:136-168).  Rows and columns are
// original particle indices; a row holds every j != i of the 27 cells with
// FP32 distance2p < rmax2t, in ascending order (the reference leaves them
// in traversal order).

// visit the storage slots sj of the 27 cells around cell c (periodic wrap
// as calc_BlockNeighbor; non-periodic cells outside the box are skipped)
template <typename Fn>
__device__ __forceinline__ void for_27_cells(const RefGrid &g, uint32_t c, int periodic,
                                             const uint32_t *start, const uint32_t *perm, Fn &&fn) {
  const uint32_t nyz = g.nb[1] * g.nb[2];
  const int i0 = (int)(c / nyz), j0 = (int)((c / g.nb[2]) % g.nb[1]), k0 = (int)(c % g.nb[2]);
  for (int a = -1; a <= 1; ++a) {
    int i2 = i0 + a;
    if (i2 < 0 || i2 >= (int)g.nb[0]) {
      if (!periodic) continue;
      i2 = (i2 + (int)g.nb[0]) % (int)g.nb[0];
    }
    for (int b = -1; b <= 1; ++b) {
      int j2 = j0 + b;
      if (j2 < 0 || j2 >= (int)g.nb[1]) {
        if (!periodic) continue;
        j2 = (j2 + (int)g.nb[1]) % (int)g.nb[1];
      }
      for (int d = -1; d <= 1; ++d) {
        int k2 = k0 + d;
        if (k2 < 0 || k2 >= (int)g.nb[2]) {
          if (!periodic) continue;
          k2 = (k2 + (int)g.nb[2]) % (int)g.nb[2];
        }
        const uint32_t J = (uint32_t)k2 + (uint32_t)j2 * g.nb[2] + (uint32_t)i2 * nyz;
        for (uint32_t jj = start[J]; jj < start[J + 1]; ++jj) fn(perm[jj]);
      }
    }
  }
}

struct NlistArgs {
  const double4 *r;
  const uint32_t *perm, *cellid, *start, *orig;
  uint32_t nreal;
  RefGrid g;
  float cx, cy, cz, rmax2t;
  int periodic;
};

__global__ void k_nlist_count(NlistArgs A, uint32_t *rowcnt) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.nreal) return;
  const uint32_t si = A.perm[k];
  const double4 pi = A.r[si];
  const float xi = (float)pi.x, yi = (float)pi.y, zi = (float)pi.z;
  uint32_t cnt = 0;
  for_27_cells(A.g, A.cellid[si], A.periodic, A.start, A.perm, [&](uint32_t sj) {
    if (sj == si) return;
    const double4 pj = A.r[sj];
    if (distance2p(xi, yi, zi, (float)pj.x, (float)pj.y, (float)pj.z, A.cx, A.cy, A.cz) < A.rmax2t)
      ++cnt;
  });
  rowcnt[A.orig[si]] = cnt;
}

__global__ void k_nlist_fill(NlistArgs A, const uint32_t *rowptr, uint32_t *col) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.nreal) return;
  const uint32_t si = A.perm[k];
  const double4 pi = A.r[si];
  const float xi = (float)pi.x, yi = (float)pi.y, zi = (float)pi.z;
  const uint32_t b = rowptr[A.orig[si]];
  uint32_t e = b;
  for_27_cells(A.g, A.cellid[si], A.periodic, A.start, A.perm, [&](uint32_t sj) {
    if (sj == si) return;
    const double4 pj = A.r[sj];
    if (distance2p(xi, yi, zi, (float)pj.x, (float)pj.y, (float)pj.z, A.cx, A.cy, A.cz) < A.rmax2t)
      col[e++] = A.orig[sj];
  });
  // ascending columns (insertion sort of the row)
  for (uint32_t i = b + 1; i < e; ++i) {
    const uint32_t x = col[i];
    uint32_t j = i;
    while (j > b && col[j - 1] > x) {
      col[j] = col[j - 1];
      --j;
    }
    col[j] = x;
  }
}

__global__ void k_inverse_perm(const uint32_t *orig, uint32_t n, uint32_t *inv) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && orig[s] != PB_NONE) inv[orig[s]] = s;
}

// FP64 minimum-image displacement (the distance2p fold of MDpairForce,
// kerneltemplate.hh:538-543, without its FP32 rounding)
__device__ __forceinline__ double min_image(double d, double L, int periodic) {
  if (periodic && d * d > 0.25 * L * L) d += d < 0 ? L : -L;
  return d;
}

struct PairArgs {
  const double4 *r;
  const uint32_t *orig, *inv;
  const uint32_t *rowptr, *col;     // list path
  const uint32_t *perm, *cellid, *start;   // cell path
  uint32_t npad;
  RefGrid g;
  double L[3];
  int periodic;
  const LJTable *lj;
  double rc2;
  double sc_rc, sc_f0;              // soft core
  double4 *F;
  double *partials;
};

// LJ forces, PE and virial from the CSR list (pairs with r^2 < rc2)
__global__ void k_lj_force_nlist(PairArgs A) {
  const uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
  double pe = 0.0, vir = 0.0;
  if (si < A.npad && A.orig[si] != PB_NONE) {
    const uint32_t row = A.orig[si];
    const double4 pi = A.r[si];
    const uint32_t ti = (uint32_t)pi.w;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (uint32_t k = A.rowptr[row]; k < A.rowptr[row + 1]; ++k) {
      const double4 pj = A.r[A.inv[A.col[k]]];
      const double dx = min_image(pi.x - pj.x, A.L[0], A.periodic);
      const double dy = min_image(pi.y - pj.y, A.L[1], A.periodic);
      const double dz = min_image(pi.z - pj.z, A.L[2], A.periodic);
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (!(r2 < A.rc2)) continue;
      const uint32_t t = ti * PB_MAXT + (uint32_t)pj.w;
      double s6 = A.lj->s2[t] / r2;
      s6 = s6 * s6 * s6;
      const double fr = A.lj->e24[t] * s6 * (2.0 * s6 - 1.0) / r2;
      fx += fr * dx;
      fy += fr * dy;
      fz += fr * dz;
      pe += 0.5 * (A.lj->e4[t] * s6 * (s6 - 1.0) - A.lj->shift[t]);
      vir += 0.5 * fr * r2;
    }
    A.F[si] = make_double4(fx, fy, fz, 0.0);
  }
  pe = pb_wave_sum(pe);
  vir = pb_wave_sum(vir);
  if (pb_lane() == 0 && si < A.npad) {
    A.partials[4 * (si / PB_WAVE) + 0] = pe;
    A.partials[4 * (si / PB_WAVE) + 1] = vir;
  }
}

// Soft-core forces of initialAnnealing (potentialSC, kernelinline.hh:140-158,
// through MDpairForce, kerneltemplate.hh:524-565) over the 27 cells:
// F_i = sum_j f0 cos(r/rc pi/2) (r_i - r_j)/r for r <= rc and r^2 < rmax2.
__global__ void k_softcore_force(PairArgs A, uint32_t nreal) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nreal) return;
  const uint32_t si = A.perm[k];
  const double4 pi = A.r[si];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  const double hpi = 0.5 * M_PI / A.sc_rc;
  for_27_cells(A.g, A.cellid[si], A.periodic, A.start, A.perm, [&](uint32_t sj) {
    if (sj == si) return;
    const double4 pj = A.r[sj];
    const double dx = min_image(pi.x - pj.x, A.L[0], A.periodic);
    const double dy = min_image(pi.y - pj.y, A.L[1], A.periodic);
    const double dz = min_image(pi.z - pj.z, A.L[2], A.periodic);
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < A.rc2)) return;
    const double r = sqrt(r2);
    if (r > A.sc_rc) return;
    const double f = A.sc_f0 * cos(r * hpi) / r;
    fx += f * dx;
    fy += f * dy;
    fz += f * dz;
  });
  A.F[si] = make_double4(fx, fy, fz, 0.0);
}

static int need_binned_grid(pb_sys *s, const char *who) {
  if (!s->ref_cells || s->ref_single)
    return pb_fail(s->ctx, PB_ERR_STATE, std::string(who) + ": needs a 27-neighbour cell grid");
  if (!s->ref_binned) return pb_fail(s->ctx, PB_ERR_STATE, std::string(who) + ": pb_bin first");
  return PB_OK;
}

static NlistArgs nlist_args(pb_sys *s, float rmax2t) {
  NlistArgs A;
  A.r = s->r;
  A.perm = s->perm;
  A.cellid = s->cellid;
  A.start = s->cstart;
  A.orig = s->orig;
  A.nreal = s->n;
  A.g = ref_grid(s);
  float c[3];
  half_cell2(s, 0, c);
  A.cx = c[0];
  A.cy = c[1];
  A.cz = c[2];
  A.rmax2t = rmax2t;
  A.periodic = s->ref_periodic;
  return A;
}

extern "C" int pb_nlist_build(pb_sys *s, float rmax2t, uint32_t *nnz_out) {
  if (!s || !(rmax2t > 0.0f)) return PB_ERR_ARG;
  if (int rc = need_binned_grid(s, "pb_nlist_build")) return rc;
  const float rc = sqrtf(rmax2t);
  for (int a = 0; a < 3; ++a)
    if (s->blocklen[a] < rc)
      return pb_fail(s->ctx, PB_ERR_CELLS, "pb_nlist_build: cell length below the list radius");
  hipStream_t st = s->ctx->stream;
  const uint32_t n = s->n;
  int rc2;
  if (!s->md_row && (rc2 = pb_alloc(s->ctx, &s->md_row, (size_t)n + 1))) return rc2;
  uint32_t *cnt = nullptr;
  if ((rc2 = pb_alloc(s->ctx, &cnt, (size_t)n + 1))) return rc2;
  const NlistArgs A = nlist_args(s, rmax2t);
  hipLaunchKernelGGL(k_nlist_count, dim3(pb_blocks(n, 256)), dim3(256), 0, st, A, cnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pb_free(cnt);
    return pb_hip_check(s->ctx, e, "k_nlist_count");
  }
  rc2 = pb_exclusive_scan(s, cnt, s->md_row, n);
  pb_free(cnt);
  if (rc2) return rc2;
  uint32_t nnz = 0;
  PB_HIP(s->ctx, hipMemcpyAsync(&nnz, s->md_row + n, 4, hipMemcpyDeviceToHost, st));
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  if (s->md_cap < nnz || !s->md_col) {
    if ((rc2 = pb_alloc(s->ctx, &s->md_col, (size_t)nnz + 1))) return rc2;
    s->md_cap = nnz;
  }
  hipLaunchKernelGGL(k_nlist_fill, dim3(pb_blocks(n, 256)), dim3(256), 0, st, A, s->md_row, s->md_col);
  PB_LAUNCH_CHECK(s->ctx, "k_nlist_fill");
  s->md_nnz = nnz;
  s->md_valid = true;
  if (nnz_out) *nnz_out = nnz;
  return PB_OK;
}

extern "C" int pb_nlist_get(pb_sys *s, uint32_t *rowptr, uint32_t *colidx, uint32_t cap,
                            uint32_t *nnz) {
  if (!s) return PB_ERR_ARG;
  if (!s->md_valid) return pb_fail(s->ctx, PB_ERR_STATE, "pb_nlist_get: pb_nlist_build first");
  hipStream_t st = s->ctx->stream;
  if (nnz) *nnz = s->md_nnz;
  if (rowptr)
    PB_HIP(s->ctx, hipMemcpyAsync(rowptr, s->md_row, 4 * ((size_t)s->n + 1), hipMemcpyDeviceToHost, st));
  if (colidx) {
    if (cap < s->md_nnz) return pb_fail(s->ctx, PB_ERR_ARG, "pb_nlist_get: colidx too small");
    PB_HIP(s->ctx, hipMemcpyAsync(colidx, s->md_col, 4 * (size_t)s->md_nnz, hipMemcpyDeviceToHost, st));
  }
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  return PB_OK;
}

static PairArgs pair_args(pb_sys *s) {
  PairArgs A;
  std::memset(&A, 0, sizeof(A));
  A.r = s->r;
  A.orig = s->orig;
  A.npad = s->npad;
  A.g = ref_grid(s);
  for (int a = 0; a < 3; ++a) A.L[a] = s->cells[2 * a + 1] - s->cells[2 * a];
  A.periodic = s->ref_periodic;
  A.lj = s->d_lj;
  A.rc2 = s->rc2;
  A.F = s->F;
  A.partials = s->partials;
  A.perm = s->perm;
  A.cellid = s->cellid;
  A.start = s->cstart;
  return A;
}

extern "C" int pb_lj_force_nlist(pb_sys *s, double *pe, double *virial) {
  if (!s) return PB_ERR_ARG;
  if (!s->md_valid) return pb_fail(s->ctx, PB_ERR_STATE, "pb_lj_force_nlist: pb_nlist_build first");
  if (!s->lj_set || !(s->rc2 > 0)) return pb_fail(s->ctx, PB_ERR_STATE, "LJ parameters or cutoff not set");
  if (s->nrep != 1) return pb_fail(s->ctx, PB_ERR_STATE, "pb_lj_force_nlist: single replica only");
  hipStream_t st = s->ctx->stream;
  int rc;
  if (!s->md_inv && (rc = pb_alloc(s->ctx, &s->md_inv, (size_t)s->n))) return rc;
  hipLaunchKernelGGL(k_inverse_perm, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, st, s->orig,
                     s->npad, s->md_inv);
  PairArgs A = pair_args(s);
  A.inv = s->md_inv;
  A.rowptr = s->md_row;
  A.col = s->md_col;
  A.periodic = 1;       // MD boxes are periodic (MDpairForce folds every pair)
  hipLaunchKernelGGL(k_lj_force_nlist, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, st, A);
  PB_LAUNCH_CHECK(s->ctx, "k_lj_force_nlist");
  if (pe || virial) {
    if ((rc = pb_reduce_waves(s, 2))) return rc;
    double h[8];
    PB_HIP(s->ctx, hipMemcpyAsync(h, s->d_out, sizeof(h), hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    if (pe) *pe = h[0];
    if (virial) *virial = h[1];
  }
  return PB_OK;
}

extern "C" int pb_softcore_force(pb_sys *s, double rc, double f0, double rmax2) {
  if (!s || !(rc > 0) || !(rmax2 > 0)) return PB_ERR_ARG;
  if (int e = need_binned_grid(s, "pb_softcore_force")) return e;
  for (int a = 0; a < 3; ++a)
    if ((double)s->blocklen[a] < std::min(rc, std::sqrt(rmax2)))
      return pb_fail(s->ctx, PB_ERR_CELLS, "pb_softcore_force: cell length below the soft-core radius");
  PairArgs A = pair_args(s);
  A.rc2 = rmax2;
  A.sc_rc = rc;
  A.sc_f0 = f0;
  hipLaunchKernelGGL(k_clear4_bin, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, s->ctx->stream,
                     s->F, s->npad);
  hipLaunchKernelGGL(k_softcore_force, dim3(pb_blocks(s->n, 256)), dim3(256), 0, s->ctx->stream, A,
                     s->n);
  PB_LAUNCH_CHECK(s->ctx, "pb_softcore_force");
  return PB_OK;
}
