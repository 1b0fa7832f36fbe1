// pb_sort.hip -- deterministic counting sort of storage slots by cell.
//
// Replaces the reference's calcBlockID sort chain (sortByBID_M1 insertion
// pre-sort, single-block odd-even sortByBID and makeBindex,
// kernelfuncs.cu:844-932, cudaCutoffBlock.cu:241-256), whose in-cell order is
// unstable and timing dependent.  Here:
//   1. k_count   : wave64 atomics on per-cell counters; equal keys in
//                  neighbouring lanes are aggregated into one atomic.
//   2. scan      : exclusive prefix sum, LDS-tiled (block sums, then
//                  scan-and-add with each block summing the sums before it),
//                  any length.
//   3. k_scatter : slot = start[cell] + rank.
//   4. k_cell_sort: inside every cell, order by a key (the original particle
//                  index) -- the in-cell order produced by the atomics is
//                  discarded, so the permutation is the canonical stable one.
#include "pb_internal.h"

#define SCAN_BS 256
#define SCAN_PER 8
#define SCAN_CH (SCAN_BS * SCAN_PER)

__global__ void k_fill_u32(uint32_t *x, uint32_t n, uint32_t val) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = val;
}

// inclusive wave scan (64 lanes)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const uint32_t lane = pb_lane();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  return x;
}

// exclusive block scan of one value per thread; returns the exclusive prefix
// and the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *total) {
  __shared__ uint32_t wsum[SCAN_BS / 64];
  const uint32_t lane = pb_lane(), w = threadIdx.x / 64;
  const uint32_t inc = wave_incl_scan(x);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (uint32_t k = 0; k < blockDim.x / 64; ++k) {
    if (k < w) off += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return off + inc - x;
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_reduce(const uint32_t *in, uint32_t n,
                                                         uint32_t *bsum) {
  const uint32_t base = blockIdx.x * SCAN_CH + threadIdx.x * SCAN_PER;
  uint32_t x = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k)
    if (base + k < n) x += in[base + k];
  uint32_t tot;
  (void)block_excl_scan(x, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// scan-and-add: each block adds the sums of the blocks before it (summed by
// the block itself, so no separate pass scans the block sums); the last
// block writes the total to out[n]
__global__ void __launch_bounds__(SCAN_BS) k_scan_final(const uint32_t *in, uint32_t n,
                                                        const uint32_t *bsum, uint32_t *out) {
  uint32_t pre = 0;
  for (uint32_t k = threadIdx.x; k < blockIdx.x; k += blockDim.x) pre += bsum[k];
  uint32_t before;
  (void)block_excl_scan(pre, &before);
  const uint32_t base = blockIdx.x * SCAN_CH + threadIdx.x * SCAN_PER;
  uint32_t v[SCAN_PER];
  uint32_t x = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    x += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(x, &tot) + before;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k)
    if (base + k < n) {
      out[base + k] = run;
      run += v[k];
    }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = before + tot;
}

int pb_exclusive_scan(pb_sys *s, const uint32_t *in, uint32_t *out, uint32_t n) {
  const uint32_t nb = (n + SCAN_CH - 1) / SCAN_CH;
  const size_t need = (size_t)nb + 1;
  if (s->scan_tmp_n < need) {
    if (int rc = pb_alloc(s->ctx, &s->scan_tmp, need)) return rc;
    s->scan_tmp_n = need;
  }
  hipStream_t st = s->ctx->stream;
  if (n == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, out, 1u, 0u);
    PB_LAUNCH_CHECK(s->ctx, "scan(empty)");
    return PB_OK;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BS), 0, st, in, n, s->scan_tmp);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(SCAN_BS), 0, st, in, n, s->scan_tmp, out);
  PB_LAUNCH_CHECK(s->ctx, "pb_exclusive_scan");
  return PB_OK;
}

// Wave-aggregated counting: lanes holding the same cell as their left
// neighbour form a run; the run head does one atomicAdd for the whole run and
// broadcasts the base.  Correct for any input; one atomic per run when the
// input is already nearly cell-ordered (the steady state after a re-sort).
__global__ void k_count(const uint32_t *cellid, uint32_t n, uint32_t *counts, uint32_t *rank) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = s < n;
  const uint32_t c = valid ? cellid[s] : PB_NONE;
  const uint32_t r = pb_count_rank(c, valid, counts);
  if (valid) rank[s] = r;
}

__global__ void k_scatter(const uint32_t *cellid, const uint32_t *rank, const uint32_t *start,
                          uint32_t n, uint32_t *perm) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) perm[start[cellid[s]] + rank[s]] = s;
}

// CS_CELLS consecutive cells per workgroup own one contiguous range of perm
// (64 cells: about 900 particles at liquid density, so a 1M-particle sort
// is ~1200 workgroups and fills the chip).
// The range is sorted in passes: each pass stages the longest run of the
// remaining cells whose particles fit the LDS stage (at least one cell).
// Every element then finds its cell and its rank in it -- the number of
// elements of the cell with a smaller key (equal keys: smaller position) --
// and is written to cell start + rank.  The rank loops are independent LDS
// reads, so a crowded cell costs O(n) per element in parallel instead of a
// serial O(n^2) insertion sort.  A single cell larger than the stage (a
// pathological pile-up) is insertion-sorted in global memory by its thread.
#define CS_CELLS 64
#define CS_THREADS 256
#define CS_CAP 3072
__global__ void __launch_bounds__(CS_THREADS) k_cell_sort(const uint32_t *start, uint32_t ncell,
                                                        const uint32_t *key, uint32_t *perm) {
  __shared__ uint32_t sv[CS_CAP], sk[CS_CAP];
  __shared__ uint32_t cst[CS_CELLS + 1];
  __shared__ uint32_t s_next;
  const uint32_t c0 = blockIdx.x * CS_CELLS;
  const uint32_t c1 = min(c0 + CS_CELLS, ncell);
  const uint32_t c = c0 + threadIdx.x;
  // one pass when the whole range fits (the common case); otherwise runs of
  // cells found by thread 0
  const bool whole = start[c1] - start[c0] <= CS_CAP;
  for (uint32_t ca = c0; ca < c1;) {                 // block-uniform
    if (whole) {
      if (threadIdx.x == 0) s_next = c1;
    } else if (threadIdx.x == 0) {
      // largest cb in [ca + 1, c1] with start[cb] - start[ca] <= CS_CAP
      const uint32_t lim = start[ca] + CS_CAP;
      uint32_t lo = ca + 1, hi = c1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (start[mid] <= lim) lo = mid;
        else hi = mid - 1;
      }
      s_next = lo;
    }
    __syncthreads();
    const uint32_t cb = s_next;
    const uint32_t base = start[ca], len = start[cb] - base;
    if (len <= CS_CAP) {
      const uint32_t nc = cb - ca;
      for (uint32_t k = threadIdx.x; k <= nc; k += blockDim.x) cst[k] = start[ca + k] - base;
      for (uint32_t q = threadIdx.x; q < len; q += blockDim.x) {
        const uint32_t p = perm[base + q];
        sv[q] = p;
        sk[q] = key[p];
      }
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
        uint32_t lo = 0, hi = nc;                    // cst[lo] <= i < cst[hi]
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (cst[mid] <= i) lo = mid;
          else hi = mid;
        }
        const uint32_t b = cst[lo], e = cst[lo + 1];
        const uint32_t ki = sk[i];
        uint32_t rank = 0;
        for (uint32_t j = b; j < e; ++j) {
          const uint32_t kj = sk[j];
          rank += (kj < ki || (kj == ki && j < i)) ? 1u : 0u;
        }
        perm[base + b + rank] = sv[i];
      }
    } else if (c >= ca && c < cb) {
      const uint32_t b = start[c], e = start[c + 1];
      for (uint32_t i = b + 1; i < e; ++i) {
        const uint32_t p = perm[i];
        const uint32_t kp = key[p];
        uint32_t j = i;
        while (j > b) {
          const uint32_t q = perm[j - 1];
          if (key[q] <= kp) break;
          perm[j] = q;
          --j;
        }
        perm[j] = p;
      }
    }
    __syncthreads();                                 // LDS and s_next are reused
    ca = cb;
  }
}

int pb_ensure_sort_ws(pb_sys *s, uint32_t ncell) {
  int rc;
  if (!s->cellid) {
    if ((rc = pb_alloc(s->ctx, &s->cellid, s->npad)) || (rc = pb_alloc(s->ctx, &s->rank, s->npad)) ||
        (rc = pb_alloc(s->ctx, &s->perm, s->npad)))
      return rc;
  }
  if (s->ncell_cap < ncell) {
    if ((rc = pb_alloc(s->ctx, &s->ccount, (size_t)ncell + 1)) ||
        (rc = pb_alloc(s->ctx, &s->cstart, (size_t)ncell + 1)))
      return rc;
    s->ncell_cap = ncell;
  }
  return PB_OK;
}

int pb_counting_sort(pb_sys *s, uint32_t ncell, const uint32_t *key) {
  if (int rc = pb_ensure_sort_ws(s, ncell)) return rc;
  hipStream_t st = s->ctx->stream;
  const uint32_t n = s->npad;
  hipLaunchKernelGGL(k_fill_u32, dim3(pb_blocks(ncell + 1, 256)), dim3(256), 0, st, s->ccount,
                     ncell + 1, 0u);
  hipLaunchKernelGGL(k_count, dim3(pb_blocks(n, 256)), dim3(256), 0, st, s->cellid, n, s->ccount,
                     s->rank);
  PB_LAUNCH_CHECK(s->ctx, "k_count");
  return pb_counting_sort_counted(s, ncell, key);
}

// the sort after the counting: s->ccount and s->rank filled (by k_count or
// by a caller's kernel that computes the cell ids and counts in one pass)
int pb_counting_sort_counted(pb_sys *s, uint32_t ncell, const uint32_t *key) {
  s->lj_cont = false;             // the fused LJ run's cell order is gone
  hipStream_t st = s->ctx->stream;
  const uint32_t n = s->npad;
  if (int rc = pb_exclusive_scan(s, s->ccount, s->cstart, ncell)) return rc;
  hipLaunchKernelGGL(k_scatter, dim3(pb_blocks(n, 256)), dim3(256), 0, st, s->cellid, s->rank,
                     s->cstart, n, s->perm);
  hipLaunchKernelGGL(k_cell_sort, dim3(pb_blocks(ncell, CS_CELLS)), dim3(CS_THREADS), 0, st,
                     s->cstart, ncell, key, s->perm);
  PB_LAUNCH_CHECK(s->ctx, "pb_counting_sort");
  return PB_OK;
}
