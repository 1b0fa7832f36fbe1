// pb_sph.hip -- SPH Navier-Stokes path of cudaParticleSPH_NS on MI355X.
//
// Reference path (cudaParticleSPH_NS.cu:71-207, kerneltemplate.hh:636-709,
// kernelinline.hh:173-215): per step a cell sort, a (cell x 27) density
// kernel writing 81 partials per particle, a 27-way reduction and two
// element-wise passes (1/rho, mu), then the same for the acceleration.
//
// Here: after pb_bin (the counting sort of the reference cell grid) the
// particles are gathered into cell order once per call (positions with the
// mass in w, velocities, per-particle material terms); one kernel computes
// the number density, rho, 1/rho and the field mu in a single pass over
// the 27 cells, and one kernel computes the acceleration.  Both stage the
// cell-sorted views through LDS column by column (neighbour cells of a
// column are one contiguous range) and write straight to the particle's
// storage slot: no partial buffers, no reductions.  Distances are
// non-periodic (distance2,
// kernelinline.hh:19-29); the density sum includes the particle itself, the
// acceleration sum excludes it; rho0 = 1 (cudaParticleSPH_NS.cu:91).
//
// The same calls also run the dam-break workload's config physics, which the
// reference's kernels do not hold (weakly compressible SPH): the M4 cubic
// spline of support 2h, summation density, the Tait equation of state
// clamped at pmin, and Monaghan's artificial viscosity; selected by
// pb_sph_set_wcsph, with the same gathered views, cell grid and launches.
#include "pb_internal.h"
#include <cmath>
#include <vector>

struct SphGrid {
  uint32_t nb[3];
};

struct SphArgs {
  const uint32_t *perm;      // sorted slot -> storage slot
  const uint32_t *cellid;    // storage slot -> cell
  const uint32_t *cstart;    // cell -> first sorted slot (cstart[c+1] ends it)
  const double4 *rs;         // sorted (x, y, z, m)
  const double4 *vs;         // sorted (vx, vy, vz, mu0 or mu)
  double4 *ds;               // sorted (1/rho, P, mu, 0) from the density pass;
                             // (rho, p/rho^2, 0, 0) under the config physics
  SphGrid g;
  uint32_t n;
  double h, w0, w1, rho0;
  // config physics: support^2, 1/h, 1/(pi h^3), 1/(pi h^5), Tait B, 1/rho0,
  // gamma, pmin, alpha c0 h, 0.01 h^2
  double sup2, hinv, ws, wf, B, rho0inv, gamma, pmin, avh, eh2;
};

// positions (mass in w), velocities, mu0 in cell order (density call)
__global__ void k_sph_gather(const uint32_t *perm, const double4 *r, const double4 *v,
                             const double *mu0, uint32_t n, double4 *rs, double4 *vs) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = perm[k];
  const double4 p = r[s], u = v[s];
  rs[k] = make_double4(p.x, p.y, p.z, 1.0 / u.w);
  vs[k] = make_double4(u.x, u.y, u.z, mu0 ? mu0[s] : 0.0);
}

// Acceleration call: positions, velocities and the per-particle terms of the
// last density call, read from the storage-order outputs, so a pb_bin
// between the two calls is harmless.  Reference physics: ds = (1/rho,
// c2/rho (1 - rho0/rho), mu); config physics: ds = (rho, p/rho^2).
__global__ void k_sph_gather_acc(const uint32_t *perm, const double4 *r, const double4 *v,
                                 const double4 *num, const double *rhoinv, const double *c2,
                                 double rho0, int wc, uint32_t n, double4 *rs, double4 *vs,
                                 double4 *ds) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = perm[k];
  const double4 p = r[s], u = v[s], d = num[s];
  const double ri = rhoinv[s];
  rs[k] = make_double4(p.x, p.y, p.z, 1.0 / u.w);
  vs[k] = make_double4(u.x, u.y, u.z, 0.0);
  ds[k] = wc ? make_double4(d.y, d.z * ri * ri, 0.0, 0.0)
             : make_double4(ri, c2[s] * ri * (1.0 - rho0 * ri), d.z, 0.0);
}

// ---------------------------------------------------------------------------
// Pair kernels.  One workgroup per segment of zseg cells of a cell column
// (i, j): the segment's sorted slots are taken 256 at a time, one thread
// each.  For each of the 9 neighbour columns the candidates of cells
// z-1 .. z+1 of every thread in the chunk form one contiguous sorted range;
// it is staged into LDS (coalesced loads, sph_cap<M>() at a time) and each
// thread scans its own 3-cell window of it:
//   * a conservative FP32 filter, 32 candidates per step into a bit mask:
//     |ri - rj|^2 = |ri|^2 + |rj|^2 - 2 ri.rj in coordinates relative to
//     the segment's first particle, against r2max plus a margin that covers
//     the FP32 rounding of that form;
//   * the FP64 pair body per set bit, from the staged FP64 fields, which
//     applies the exact support test itself (so the margin changes nothing).
// Pairs are visited in the order of a per-particle 27-cell loop (neighbour
// columns in (a, b) order, slots ascending).
constexpr uint32_t SC_T = 256;     // threads per workgroup, one sorted slot each

// FP32 pre-filter bound r2 (r2max plus the rounding margin) and the squared
// coordinate range l2 that the margin covers
struct SphFilter {
  float r2, l2;
};

enum { SPH_LUCY_DENS = 0, SPH_LUCY_ACC = 1, SPH_WC_DENS = 2, SPH_WC_ACC = 3 };

// staged FP64 fields per candidate: x, y, z, m, then
//   Lucy density: mu0 | Lucy accel: vx, vy, vz, 1/rho, P, mu |
//   config density: - | config accel: vx, vy, vz, rho, P
template <int M>
constexpr int sph_nf() {
  return M == SPH_LUCY_DENS ? 5 : M == SPH_LUCY_ACC ? 10 : M == SPH_WC_DENS ? 4 : 9;
}
// candidates per batch: LDS of a workgroup stays below 40 KB (4 per CU)
template <int M>
constexpr uint32_t sph_cap() {
  return (M == SPH_LUCY_ACC || M == SPH_WC_ACC) ? 384u : 512u;
}

// 1/sqrt(x), x > 0: the hardware estimate and two Newton steps (about one
// ulp); and 1/x the same way
__device__ __forceinline__ double sph_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * fma(-hx * y, y, 1.5);
  y = y * fma(-hx * y, y, 1.5);
  return y;
}
__device__ __forceinline__ double sph_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}
// Tait pressure max(B ((rho/rho0)^gamma - 1), pmin); out of line, once per
// particle, so pow's registers do not weigh on the pair loop
__device__ __attribute__((noinline)) double sph_tait(double rho, double rho0inv, double gamma,
                                                     double B, double pmin) {
  return fmax(B * (pow(rho * rho0inv, gamma) - 1.0), pmin);
}

// |r| from r^2 (0 at 0)
__device__ __forceinline__ double sph_norm(double r2, double *rinv) {
  const double y = r2 > 0.0 ? sph_rsqrt(r2) : 0.0;
  *rinv = y;
  return r2 * y;
}

template <int M>
__device__ __forceinline__ void sph_pairs(const SphArgs &A, uint32_t zseg, SphFilter FB,
                                          double4 *out4, double *out1) {
  constexpr int NF = sph_nf<M>();
  constexpr uint32_t CAP = sph_cap<M>();
  __shared__ double L[NF][CAP];
  __shared__ float4 Fq[CAP + 64];             // filter view; 64 spare for the last step
  const uint32_t nz = A.g.nb[2], ny = A.g.nb[1];
  const uint32_t ncol = A.g.nb[0] * ny;
  // 1-D grid of (z segment, column), columns fastest.  Workgroups go to the
  // XCDs round-robin; within every 64 workgroups XCD x takes 8 consecutive
  // columns, which share neighbour columns in its L2, while the fine
  // interleave keeps dense and empty regions spread over all XCDs.
  const uint32_t lb = pb_xcd_chunk8();
  const uint32_t col = lb % ncol;
  const int ci = (int)(col / ny), cj = (int)(col % ny);
  // this workgroup's cells: z in [zs, ze) of column col
  const uint32_t zs = (lb / ncol) * zseg, ze = zs + zseg < nz ? zs + zseg : nz;
  const uint32_t s0 = A.cstart[col * nz + zs], s1 = A.cstart[col * nz + ze];
  if (s0 >= s1) return;
  const double4 O = A.rs[s0];                 // origin of the filter coordinates
  for (uint32_t c0 = s0; c0 < s1; c0 += SC_T) {
    const uint32_t c1 = c0 + SC_T < s1 ? c0 + SC_T : s1;
    const uint32_t k = c0 + threadIdx.x;
    const bool act = k < c1;
    uint32_t s = 0, kz = 0;
    double4 pi = O, vi = make_double4(0.0, 0.0, 0.0, 0.0), di = vi;
    if (act) {
      s = A.perm[k];
      kz = A.cellid[s] % nz;
      pi = A.rs[k];
      if constexpr (M == SPH_LUCY_ACC || M == SPH_WC_ACC) {
        vi = A.vs[k];
        di = A.ds[k];
      }
    }
    // filter constants: -2 ri and r2 margin - |ri|^2, FP32, relative to O;
    // a particle farther from O than the margin covers (outside the cell
    // grid) passes every candidate to the exact test
    const float xf = (float)(pi.x - O.x), yf = (float)(pi.y - O.y), zf = (float)(pi.z - O.z);
    const float fx = -2.0f * xf, fy = -2.0f * yf, fz = -2.0f * zf;
    const float bi = xf * xf + yf * yf + zf * zf;
    const float lim = bi <= FB.l2 ? FB.r2 - bi : INFINITY;
    // z range of the chunk (slots are sorted by cell, so by z in a column)
    const uint32_t za = A.cellid[A.perm[c0]] % nz, zb = A.cellid[A.perm[c1 - 1]] % nz;
    const uint32_t zlo = za > 0 ? za - 1 : 0, zhi = zb + 1 < nz ? zb + 1 : zb;
    const uint32_t klo = kz > 0 ? kz - 1 : 0, khi = kz + 1 < nz ? kz + 1 : kz;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    for (int a = -1; a <= 1; ++a) {
      const int i2 = ci + a;
      if (i2 < 0 || i2 >= (int)A.g.nb[0]) continue;
      for (int b = -1; b <= 1; ++b) {
        const int j2 = cj + b;
        if (j2 < 0 || j2 >= (int)ny) continue;
        const uint32_t cb = ((uint32_t)i2 * ny + (uint32_t)j2) * nz;
        const uint32_t r0 = A.cstart[cb + zlo], r1 = A.cstart[cb + zhi + 1];
        uint32_t w0 = 0, w1 = 0;          // this thread's window (empty if inactive)
        if (act) {
          w0 = A.cstart[cb + klo];
          w1 = A.cstart[cb + khi + 1];
        }
        for (uint32_t b0 = r0; b0 < r1; b0 += CAP) {
          const uint32_t b1 = b0 + CAP < r1 ? b0 + CAP : r1;
          __syncthreads();                // the previous batch is consumed
          for (uint32_t t = threadIdx.x; t < b1 - b0; t += SC_T) {
            const double4 pj = A.rs[b0 + t];
            L[0][t] = pj.x;
            L[1][t] = pj.y;
            L[2][t] = pj.z;
            L[3][t] = pj.w;
            const float xr = (float)(pj.x - O.x), yr = (float)(pj.y - O.y),
                        zr = (float)(pj.z - O.z);
            const float bj = xr * xr + yr * yr + zr * zr;
            // beyond the margin's range: a filter entry every thread passes
            Fq[t] = bj <= FB.l2 ? make_float4(xr, yr, zr, bj) : make_float4(0.f, 0.f, 0.f, -INFINITY);
            if constexpr (M == SPH_LUCY_DENS) {
              L[4][t] = A.vs[b0 + t].w;
            } else if constexpr (M == SPH_LUCY_ACC || M == SPH_WC_ACC) {
              const double4 vj = A.vs[b0 + t], dj = A.ds[b0 + t];
              L[4][t] = vj.x;
              L[5][t] = vj.y;
              L[6][t] = vj.z;
              L[7][t] = dj.x;
              L[8][t] = dj.y;
              if constexpr (M == SPH_LUCY_ACC) L[9][t] = dj.z;
            }
          }
          __syncthreads();
          const uint32_t q0 = w0 > b0 ? w0 : b0, q1 = w1 < b1 ? w1 : b1;
          for (uint32_t base = q0; base < q1; base += 64) {
            const uint32_t t0 = base - b0;
            // 64 tests, two 32-bit halves (entries past the window are
            // masked off below; Fq has 64 spare entries so the reads stay
            // inside the array).  A window of three cells usually fits in
            // one step, so the pair loop below runs the largest per-lane hit
            // count of the window once.
            uint32_t lo = 0, hi = 0;
#pragma unroll 1
            for (uint32_t g = 0; g < 32; g += 8) {
#pragma unroll
              for (uint32_t u = 0; u < 8; ++u) {
                const float4 f = Fq[t0 + g + u];
                if (fmaf(fx, f.x, fmaf(fy, f.y, fmaf(fz, f.z, f.w))) < lim) lo |= 1u << (g + u);
              }
            }
            if (q1 - base > 32u) {
#pragma unroll 1
              for (uint32_t g = 0; g < 32; g += 8) {
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                  const float4 f = Fq[t0 + 32 + g + u];
                  if (fmaf(fx, f.x, fmaf(fy, f.y, fmaf(fz, f.z, f.w))) < lim) hi |= 1u << (g + u);
                }
              }
            }
            uint64_t mask = ((uint64_t)hi << 32) | lo;
            if (q1 - base < 64u) mask &= (1ull << (q1 - base)) - 1ull;
            while (mask) {
              const uint32_t u = (uint32_t)__builtin_ctzll(mask);
              mask &= mask - 1;
              const uint32_t j = t0 + u, q = base + u;
              const double xj = L[0][j], yj = L[1][j], zj = L[2][j], mj = L[3][j];
              if constexpr (M == SPH_LUCY_DENS) {
                // SPHcalcDensity<SPHKernelLucy> (kerneltemplate.hh:636-656):
                // W = w0 (h + 3R)(h - R)^3, 0 beyond h (kernelinline.hh:182-188)
                const double dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                double ri;
                const double R = sph_norm(dx * dx + dy * dy + dz * dz, &ri);
                const double tt = fmax(A.h - R, 0.0);
                const double W = A.w0 * (A.h + 3.0 * R) * (tt * tt * tt);
                const double mW = mj * W;
                acc0 += W;
                acc1 += mW;
                acc2 = fma(L[4][j], mW, acc2);
              } else if constexpr (M == SPH_LUCY_ACC) {
                // SPHNavierStokes<SPHKernelLucyDW> (kerneltemplate.hh:666-709)
                const double dx = xj - pi.x, dy = yj - pi.y, dz = zj - pi.z;
                double ri;
                const double R = sph_norm(dx * dx + dy * dy + dz * dz, &ri);
                const double tt = A.h - R;
                if (tt > 0.0 && q != k) {                 // no self term, 0 beyond h
                  const double mw = mj * (A.w1 * tt * tt);  // m_j (dW/dr)/r, kernelinline.hh:209-214
                  const double p1 = di.y + L[8][j];
                  const double v1 = (di.z + L[9][j]) * di.x * L[7][j];
                  acc0 += mw * (dx * p1 - (L[4][j] - vi.x) * v1);
                  acc1 += mw * (dy * p1 - (L[5][j] - vi.y) * v1);
                  acc2 += mw * (dz * p1 - (L[6][j] - vi.z) * v1);
                }
              } else if constexpr (M == SPH_WC_DENS) {
                // M4 spline inside the support: W = ws (t2^3/4 - t1^3),
                // t1 = max(1 - q, 0), t2 = 2 - q, q = r/h
                const double dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                const double R2 = dx * dx + dy * dy + dz * dz;
                if (R2 < A.sup2) {
                  double ri;
                  const double qq = sph_norm(R2, &ri) * A.hinv;
                  const double t1 = fmax(1.0 - qq, 0.0), t2 = 2.0 - qq;
                  const double W = 0.25 * (t2 * t2 * t2) - t1 * t1 * t1;
                  acc0 += W;
                  acc1 = fma(mj, W, acc1);
                }
              } else {
                // a_i = -sum m_j (P_i + P_j + Pi_ij) F(r) r_ij, F = (dW/dr)/r
                // = wf (-3 + 9/4 q) for q < 1, wf (-3/4 (2 - q)^2 / q) below 2;
                // Pi_ij = -alpha c0 h (v_ij . r_ij) / ((r^2 + 0.01 h^2) rho_bar)
                // for approaching pairs
                const double dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                const double R2 = dx * dx + dy * dy + dz * dz;
                if (R2 < A.sup2 && q != k) {
                  double ri;
                  const double qq = sph_norm(R2, &ri) * A.hinv, t2 = 2.0 - qq;
                  const double g = qq < 1.0 ? fma(2.25, qq, -3.0) : -0.75 * t2 * t2 * (A.h * ri);
                  const double vr =
                      (vi.x - L[4][j]) * dx + (vi.y - L[5][j]) * dy + (vi.z - L[6][j]) * dz;
                  const double visc =
                      vr < 0.0 ? -A.avh * vr * sph_rcp((R2 + A.eh2) * (0.5 * (di.x + L[7][j])))
                               : 0.0;
                  const double f = -mj * (di.y + L[8][j] + visc) * g;
                  acc0 = fma(f, dx, acc0);
                  acc1 = fma(f, dy, acc1);
                  acc2 = fma(f, dz, acc2);
                }
              }
            }
          }
        }
      }
    }
    if (act) {
      if constexpr (M == SPH_LUCY_DENS) {
        // calcDensityPost (cudaParticleSPH_NS.cu:184-207)
        const double ri = 1.0 / acc1;
        out4[s] = make_double4(acc0, acc1, acc2 * ri, 0.0);
        out1[s] = ri;
      } else if constexpr (M == SPH_WC_DENS) {
        // Tait: p = max(B ((rho/rho0)^gamma - 1), pmin)
        const double n0 = acc0 * A.ws, rho = acc1 * A.ws;
        const double p = sph_tait(rho, A.rho0inv, A.gamma, A.B, A.pmin);
        out4[s] = make_double4(n0, rho, p, 0.0);
        out1[s] = 1.0 / rho;
      } else if constexpr (M == SPH_LUCY_ACC) {
        out4[s] = make_double4(acc0, acc1, acc2, 0.0);
      } else {
        out4[s] = make_double4(acc0 * A.wf, acc1 * A.wf, acc2 * A.wf, 0.0);
      }
    }
  }
}

// The config density fits six wavefronts per SIMD (80 VGPRs, 25 KB of
// LDS); the other passes need more registers and run at four or five.
template <int M>
__global__ void __launch_bounds__(SC_T) k_sph_pairs(SphArgs A, uint32_t zseg, SphFilter FB,
                                                    double4 *out4, double *out1) {
  sph_pairs<M>(A, zseg, FB, out4, out1);
}
__global__ void __launch_bounds__(SC_T) __attribute__((amdgpu_waves_per_eu(6)))
k_sph_pairs_wc_density(SphArgs A, uint32_t zseg, SphFilter FB, double4 *out4, double *out1) {
  sph_pairs<SPH_WC_DENS>(A, zseg, FB, out4, out1);
}

// inspectDense_x (kernelfuncs.cu:772-801): per-wave partial sums of the
// number density and the count of movable particles
__global__ void k_sph_inspect(const double4 *num, const uint8_t *move, const uint32_t *orig,
                              uint32_t n, double *partials) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double x = 0.0, c = 0.0;
  if (s < n && orig[s] != PB_NONE && move[s] > 0) {
    x = num[s].x;
    c = 1.0;
  }
  x = pb_wave_sum(x);
  c = pb_wave_sum(c);
  if (pb_lane() == 0 && s < n) {
    partials[4 * (s / PB_WAVE)] = x;
    partials[4 * (s / PB_WAVE) + 1] = c;
  }
}

// ------------------------------------------------------------------ host
static int sph_ready(pb_sys *s) {
  if (!s->sph_set) return pb_fail(s->ctx, PB_ERR_STATE, "SPH properties not set");
  if (!s->ref_binned) return pb_fail(s->ctx, PB_ERR_STATE, "SPH: pb_bin (calcBlockID) first");
  if (s->ref_periodic) return pb_fail(s->ctx, PB_ERR_CELLS, "SPH: cells must be non-periodic");
  const float hf = (float)(s->sph_kind == 1 ? 2.0 * s->sph_h : s->sph_h);
  if (s->blocklen[0] < hf || s->blocklen[1] < hf || s->blocklen[2] < hf)
    return pb_fail(s->ctx, PB_ERR_CELLS, "SPH: cell length below the kernel support");
  return PB_OK;
}

static SphArgs sph_args(pb_sys *s) {
  SphArgs A;
  A.perm = s->perm;
  A.cellid = s->cellid;
  A.cstart = s->cstart;
  A.rs = s->sph_rs;
  A.vs = s->sph_vs;
  A.ds = s->sph_ds;
  for (int a = 0; a < 3; ++a) A.g.nb[a] = s->blocknum[a];
  A.n = s->n;
  A.h = s->sph_h;
  A.w0 = 105.0 / (16.0 * M_PI * std::pow(s->sph_h, 7));   // cudaParticleSPHBase.cu:121
  A.w1 = -12.0 * A.w0;
  A.rho0 = s->sph_rho0;
  const double h = s->sph_h;
  A.sup2 = 4.0 * h * h;
  A.hinv = 1.0 / h;
  A.ws = 1.0 / (M_PI * h * h * h);
  A.wf = A.ws / (h * h);
  A.B = s->sph_rho0 * s->sph_c0 * s->sph_c0 / s->sph_gamma;
  A.rho0inv = 1.0 / s->sph_rho0;
  A.gamma = s->sph_gamma;
  A.pmin = s->sph_pmin;
  A.avh = s->sph_alpha * s->sph_c0 * h;
  A.eh2 = 0.01 * h * h;
  return A;
}

// Pair-kernel grid: (columns, z segments).  A segment of zseg cells holds
// 64 particles at the mean cell occupancy, and two to three times that in
// the filled part of a partly filled box, so dense columns are spread over
// several workgroups of one or two 256-slot chunks each.
static dim3 sph_pair_grid(pb_sys *s, uint32_t *zseg) {
  const uint32_t nz = s->blocknum[2];
  const double ncell = (double)s->blocknum[0] * s->blocknum[1] * nz;
  const double occ = s->n / (ncell > 0 ? ncell : 1.0);
  uint32_t z = (uint32_t)std::lround(64.0 / (occ > 1e-9 ? occ : 1e-9));
  z = z < 1 ? 1 : (z > nz ? nz : z);
  *zseg = z;
  return dim3(s->blocknum[0] * s->blocknum[1] * ((nz + z - 1) / z));
}

// FP32 filter bound: r2max plus 1e-6 Lmax^2, where Lmax bounds the filter
// coordinates' magnitude for particles inside the cell grid (two cells
// across x and y, the segment and one cell along z); the rounding of the
// expanded |ri - rj|^2 in FP32 stays below ten units of 2^-24 Lmax^2, so no
// pair within r2max is filtered out.  Coordinates beyond Lmax (particles
// outside the grid, binned into its edge cells) bypass the filter.
static SphFilter sph_filter(pb_sys *s, uint32_t zseg, double r2max) {
  const double cx = s->blocklen[0], cy = s->blocklen[1], cz = s->blocklen[2];
  const double l2 = 4.0 * cx * cx + 4.0 * cy * cy + (zseg + 1.0) * (zseg + 1.0) * cz * cz;
  SphFilter F;
  F.r2 = (float)(r2max * (1.0 + 1e-6) + 1e-6 * l2);
  F.l2 = (float)l2;
  return F;
}

// density, 1/rho and the cell-order views, shared by both physics
static int sph_alloc_views(pb_sys *s) {
  const size_t np = s->npad;
  int rc = PB_OK;
  if (!s->sph_rs) {
    if ((rc = pb_alloc(s->ctx, &s->sph_rhoinv, np)) || (rc = pb_alloc(s->ctx, &s->sph_num, np)) ||
        (rc = pb_alloc(s->ctx, &s->sph_rs, np)) || (rc = pb_alloc(s->ctx, &s->sph_vs, np)) ||
        (rc = pb_alloc(s->ctx, &s->sph_ds, np)))
      return rc;
  }
  hipStream_t st = s->ctx->stream;
  PB_HIP(s->ctx, hipMemsetAsync(s->sph_num, 0, sizeof(double4) * np, st));
  PB_HIP(s->ctx, hipMemsetAsync(s->sph_rhoinv, 0, 8 * np, st));
  return PB_OK;
}

extern "C" int pb_sph_set_properties(pb_sys *s, const double *mu0, const double *c, double h) {
  if (!s || !mu0 || !c || !(h > 0)) return PB_ERR_ARG;
  if (s->nrep != 1) return pb_fail(s->ctx, PB_ERR_ARG, "SPH: single replica only");
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  int rc = PB_OK;
  std::vector<uint32_t> o = pb_download_orig(s, &rc);
  if (rc) return rc;
  const size_t np = s->npad;
  std::vector<double> hm(np, 0.0), hc(np, 0.0);
  for (size_t k = 0; k < np; ++k)
    if (o[k] != PB_NONE) {
      hm[k] = mu0[o[k]];
      hc[k] = c[o[k]] * c[o[k]];                    // c^2, cudaParticleSPH_NS.cu:63-64
    }
  if (!s->sph_mu0) {
    if ((rc = pb_alloc(s->ctx, &s->sph_mu0, np)) || (rc = pb_alloc(s->ctx, &s->sph_c2, np)))
      return rc;
  }
  if ((rc = sph_alloc_views(s))) return rc;
  hipStream_t st = s->ctx->stream;
  PB_HIP(s->ctx, hipMemcpyAsync(s->sph_mu0, hm.data(), 8 * np, hipMemcpyHostToDevice, st));
  PB_HIP(s->ctx, hipMemcpyAsync(s->sph_c2, hc.data(), 8 * np, hipMemcpyHostToDevice, st));
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  s->sph_h = h;
  s->sph_rho0 = 1.0;
  s->sph_kind = 0;
  s->sph_set = true;
  s->sph_dens = false;
  return PB_OK;
}

extern "C" int pb_sph_set_wcsph(pb_sys *s, double h, double rho0, double c0, double gamma,
                                double alpha, double pmin) {
  if (!s || !(h > 0) || !(rho0 > 0) || !(c0 > 0) || !(gamma > 0) || !(alpha >= 0) ||
      std::isnan(pmin) || !std::isfinite(h * rho0 * c0 * gamma * alpha))
    return PB_ERR_ARG;
  if (s->nrep != 1) return pb_fail(s->ctx, PB_ERR_ARG, "SPH: single replica only");
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  if (int rc = sph_alloc_views(s)) return rc;
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  s->sph_h = h;
  s->sph_rho0 = rho0;
  s->sph_c0 = c0;
  s->sph_gamma = gamma;
  s->sph_alpha = alpha;
  s->sph_pmin = pmin;
  s->sph_kind = 1;
  s->sph_set = true;
  s->sph_dens = false;
  return PB_OK;
}

extern "C" int pb_sph_density(pb_sys *s) {
  if (!s) return PB_ERR_ARG;
  if (int rc = sph_ready(s)) return rc;
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const uint32_t n = s->n;
  if (n == 0) return PB_OK;
  SphArgs A = sph_args(s);
  const bool wc = s->sph_kind == 1;
  uint32_t zseg = 0;
  const dim3 grid = sph_pair_grid(s, &zseg);
  hipLaunchKernelGGL(k_sph_gather, dim3(pb_blocks(n, 256)), dim3(256), 0, st, s->perm, s->r, s->v,
                     wc ? (const double *)nullptr : s->sph_mu0, n, s->sph_rs, s->sph_vs);
  if (wc)
    hipLaunchKernelGGL(k_sph_pairs_wc_density, grid, dim3(SC_T), 0, st, A, zseg,
                       sph_filter(s, zseg, A.sup2), s->sph_num, s->sph_rhoinv);
  else
    hipLaunchKernelGGL(k_sph_pairs<SPH_LUCY_DENS>, grid, dim3(SC_T), 0, st, A, zseg,
                       sph_filter(s, zseg, A.h * A.h), s->sph_num, s->sph_rhoinv);
  PB_LAUNCH_CHECK(s->ctx, "k_sph_pairs (density)");
  s->sph_dens = true;
  return PB_OK;
}

extern "C" int pb_sph_acceleration(pb_sys *s) {
  if (!s) return PB_ERR_ARG;
  if (int rc = sph_ready(s)) return rc;
  if (!s->sph_dens) return pb_fail(s->ctx, PB_ERR_STATE, "SPH: pb_sph_density first");
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const uint32_t n = s->n;
  if (n == 0) return PB_OK;
  SphArgs A = sph_args(s);
  const bool wc = s->sph_kind == 1;
  uint32_t zseg = 0;
  const dim3 grid = sph_pair_grid(s, &zseg);
  // current positions and velocities, and the density call's rho, 1/rho,
  // mu or p, in the current cell order
  hipLaunchKernelGGL(k_sph_gather_acc, dim3(pb_blocks(n, 256)), dim3(256), 0, st, s->perm, s->r,
                     s->v, s->sph_num, s->sph_rhoinv,
                     s->sph_c2 ? s->sph_c2 : s->sph_rhoinv,   // not read under the config physics
                     s->sph_rho0, wc ? 1 : 0, n,
                     s->sph_rs, s->sph_vs, s->sph_ds);
  if (wc)
    hipLaunchKernelGGL(k_sph_pairs<SPH_WC_ACC>, grid, dim3(SC_T), 0, st, A, zseg,
                       sph_filter(s, zseg, A.sup2), s->a, (double *)nullptr);
  else
    hipLaunchKernelGGL(k_sph_pairs<SPH_LUCY_ACC>, grid, dim3(SC_T), 0, st, A, zseg,
                       sph_filter(s, zseg, A.h * A.h), s->a, (double *)nullptr);
  PB_LAUNCH_CHECK(s->ctx, "k_sph_pairs (acceleration)");
  return PB_OK;
}

extern "C" int pb_sph_inspect_density(pb_sys *s, double *mean_num) {
  if (!s || !mean_num) return PB_ERR_ARG;
  if (!s->sph_num) return pb_fail(s->ctx, PB_ERR_STATE, "SPH: no density yet");
  hipLaunchKernelGGL(k_sph_inspect, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, s->ctx->stream,
                     s->sph_num, s->move, s->orig, s->npad, s->partials);
  PB_LAUNCH_CHECK(s->ctx, "k_sph_inspect");
  if (int rc = pb_reduce_waves(s, 2)) return rc;
  double h[2];
  PB_HIP(s->ctx, hipMemcpyAsync(h, s->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  *mean_num = h[0] / h[1];     // 0/0 (no movable particle) is NaN, as the reference divides
  return PB_OK;
}
