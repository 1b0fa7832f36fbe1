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
// the 27 cells, and one kernel computes the acceleration.  Both read the
// cell-sorted views (neighbour cells are contiguous ranges, coalesced in
// HBM) and write straight to the particle's storage slot: no partial
// buffers, no reductions.  Distances are non-periodic (distance2,
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

// positions (mass in w), velocities, mu0 in cell order
__global__ void k_sph_gather(const uint32_t *perm, const double4 *r, const double4 *v,
                             const double *mu0, uint32_t n, double4 *rs, double4 *vs) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = perm[k];
  const double4 p = r[s], u = v[s];
  rs[k] = make_double4(p.x, p.y, p.z, 1.0 / u.w);
  vs[k] = make_double4(u.x, u.y, u.z, mu0 ? mu0[s] : 0.0);
}

// visit the sorted slots of the 27 cells around cell c (non-periodic)
template <typename F>
__device__ __forceinline__ void for_neighbors(const SphArgs &A, uint32_t c, F &&f) {
  const uint32_t nz = A.g.nb[2], nyz = A.g.nb[1] * nz;
  const int i0 = (int)(c / nyz), j0 = (int)((c / nz) % A.g.nb[1]), k0 = (int)(c % nz);
  for (int a = -1; a <= 1; ++a) {
    const int i2 = i0 + a;
    if (i2 < 0 || i2 >= (int)A.g.nb[0]) continue;
    for (int b = -1; b <= 1; ++b) {
      const int j2 = j0 + b;
      if (j2 < 0 || j2 >= (int)A.g.nb[1]) continue;
      // the three z cells of a column are adjacent in the sorted order
      const int kl = k0 > 0 ? k0 - 1 : 0, kh = k0 + 1 < (int)nz ? k0 + 1 : k0;
      const uint32_t cb = (uint32_t)i2 * nyz + (uint32_t)j2 * nz;
      const uint32_t q0 = A.cstart[cb + (uint32_t)kl], q1 = A.cstart[cb + (uint32_t)kh + 1];
      // four candidates per iteration so their loads are in flight together
      uint32_t q = q0;
      for (; q + 4 <= q1; q += 4) {
        f(q);
        f(q + 1);
        f(q + 2);
        f(q + 3);
      }
      for (; q < q1; ++q) f(q);
    }
  }
}

// SPHcalcDensity<SPHKernelLucy> (kerneltemplate.hh:636-656) + calcDensityPost
// (cudaParticleSPH_NS.cu:184-207), one thread per sorted slot.
__global__ void __launch_bounds__(256) k_sph_density(SphArgs A, double4 *num, double *rhoinv) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  const uint32_t s = A.perm[k];
  const double4 pi = A.rs[k];
  double n0 = 0.0, rho = 0.0, mr = 0.0;
  for_neighbors(A, A.cellid[s], [&](uint32_t q) {
    const double4 pj = A.rs[q];
    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const double R = sqrt(dx * dx + dy * dy + dz * dz);
    const double t = fmax(A.h - R, 0.0);            // W = 0 beyond h, kernelinline.hh:182-188
    const double W = A.w0 * (A.h + 3.0 * R) * (t * t * t);
    const double mW = pj.w * W;
    n0 += W;
    rho += mW;
    mr = fma(A.vs[q].w, mW, mr);
  });
  const double ri = 1.0 / rho;
  const double mu = mr * ri;
  num[s] = make_double4(n0, rho, mu, 0.0);
  rhoinv[s] = ri;
  A.ds[k] = make_double4(ri, 0.0, mu, 0.0);   // P filled by the acceleration call
}

// per-particle pressure term c2/rho (1 - rho0/rho) in cell order
__global__ void k_sph_pressure(SphArgs A, const double *c2) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  double4 d = A.ds[k];
  d.y = c2[A.perm[k]] * d.x * (1.0 - A.rho0 * d.x);
  A.ds[k] = d;
}

// SPHNavierStokes<SPHKernelLucyDW> (kerneltemplate.hh:666-709), one thread
// per sorted slot; a is overwritten (the reference clears it first).
__global__ void __launch_bounds__(256) k_sph_accel(SphArgs A, double4 *acc) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  const uint32_t s = A.perm[k];
  const double4 pi = A.rs[k], vi = A.vs[k], di = A.ds[k];
  double ax = 0.0, ay = 0.0, az = 0.0;
  for_neighbors(A, A.cellid[s], [&](uint32_t q) {
    const double4 pj = A.rs[q];
    const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
    const double R = sqrt(dx * dx + dy * dy + dz * dz);
    const double t = A.h - R;
    if (t > 0.0 && q != k) {                      // no self term, 0 beyond h
      const double4 vj = A.vs[q], dj = A.ds[q];
      const double mj = pj.w * (A.w1 * t * t);    // m_j (dW/dr)/r, kernelinline.hh:209-214
      const double p1 = di.y + dj.y;
      const double v1 = (di.z + dj.z) * di.x * dj.x;
      ax += mj * (dx * p1 - (vj.x - vi.x) * v1);
      ay += mj * (dy * p1 - (vj.y - vi.y) * v1);
      az += mj * (dz * p1 - (vj.z - vi.z) * v1);
    }
  });
  acc[s] = make_double4(ax, ay, az, 0.0);
}

// Config physics density: inside the support W = ws (t2^3/4 - t1^3) with
// t1 = max(1 - q, 0), t2 = 2 - q, q = r/h (the M4 spline in B-spline form),
// rho = sum m_j W (self included), p = max(B ((rho/rho0)^gamma - 1), pmin).
// num = (sum W, rho, p), ds = (rho, p/rho^2) for the acceleration pass.
__global__ void __launch_bounds__(256) k_wc_density(SphArgs A, double4 *num, double *rhoinv) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  const uint32_t s = A.perm[k];
  const double4 pi = A.rs[k];
  double n0 = 0.0, rho = 0.0;
  for_neighbors(A, A.cellid[s], [&](uint32_t q) {
    const double4 pj = A.rs[q];
    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const double R2 = dx * dx + dy * dy + dz * dz;
    if (R2 < A.sup2) {
      const double qq = sqrt(R2) * A.hinv;
      const double t1 = fmax(1.0 - qq, 0.0), t2 = 2.0 - qq;
      const double W = 0.25 * (t2 * t2 * t2) - t1 * t1 * t1;
      n0 += W;
      rho = fma(pj.w, W, rho);
    }
  });
  n0 *= A.ws;
  rho *= A.ws;
  const double p = fmax(A.B * (pow(rho * A.rho0inv, A.gamma) - 1.0), A.pmin);
  const double ri = 1.0 / rho;
  num[s] = make_double4(n0, rho, p, 0.0);
  rhoinv[s] = ri;
  A.ds[k] = make_double4(rho, p * ri * ri, 0.0, 0.0);
}

// Config physics acceleration (gravity excluded):
//   a_i = -sum_{j != i} m_j (p_i/rho_i^2 + p_j/rho_j^2 + Pi_ij) F(r) r_ij,
// F = (dW/dr)/r = wf (-3 + 9/4 q) for q < 1, wf (-3/4 (2 - q)^2 / q) below 2;
// Pi_ij = -alpha c0 h (v_ij . r_ij) / ((r^2 + 0.01 h^2) (rho_i + rho_j)/2)
// for approaching pairs, 0 otherwise.  a is overwritten.
__global__ void __launch_bounds__(256) k_wc_accel(SphArgs A, double4 *acc) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  const uint32_t s = A.perm[k];
  const double4 pi = A.rs[k], vi = A.vs[k], di = A.ds[k];
  double ax = 0.0, ay = 0.0, az = 0.0;
  for_neighbors(A, A.cellid[s], [&](uint32_t q) {
    const double4 pj = A.rs[q];
    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const double R2 = dx * dx + dy * dy + dz * dz;
    if (R2 < A.sup2 && q != k) {
      const double4 vj = A.vs[q], dj = A.ds[q];
      const double qq = sqrt(R2) * A.hinv, t2 = 2.0 - qq;
      const double g = qq < 1.0 ? fma(2.25, qq, -3.0) : -0.75 * t2 * t2 / qq;
      const double vr = (vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz;
      const double visc = vr < 0.0 ? -A.avh * vr / ((R2 + A.eh2) * (0.5 * (di.x + dj.x))) : 0.0;
      const double f = -pj.w * (di.y + dj.y + visc) * g;
      ax = fma(f, dx, ax);
      ay = fma(f, dy, ay);
      az = fma(f, dz, az);
    }
  });
  acc[s] = make_double4(ax * A.wf, ay * A.wf, az * A.wf, 0.0);
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
  hipLaunchKernelGGL(k_sph_gather, dim3(pb_blocks(n, 256)), dim3(256), 0, st, s->perm, s->r, s->v,
                     wc ? (const double *)nullptr : s->sph_mu0, n, s->sph_rs, s->sph_vs);
  if (wc)
    hipLaunchKernelGGL(k_wc_density, dim3(pb_blocks(n, 256)), dim3(256), 0, st, A, s->sph_num,
                       s->sph_rhoinv);
  else
    hipLaunchKernelGGL(k_sph_density, dim3(pb_blocks(n, 256)), dim3(256), 0, st, A, s->sph_num,
                       s->sph_rhoinv);
  PB_LAUNCH_CHECK(s->ctx, "k_sph_density");
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
  // current positions and velocities in cell order (1/rho and mu of the
  // density pass are already in ds)
  hipLaunchKernelGGL(k_sph_gather, dim3(pb_blocks(n, 256)), dim3(256), 0, st, s->perm, s->r, s->v,
                     (const double *)nullptr, n, s->sph_rs, s->sph_vs);
  if (s->sph_kind == 1) {
    hipLaunchKernelGGL(k_wc_accel, dim3(pb_blocks(n, 256)), dim3(256), 0, st, A, s->a);
  } else {
    hipLaunchKernelGGL(k_sph_pressure, dim3(pb_blocks(n, 256)), dim3(256), 0, st, A, s->sph_c2);
    hipLaunchKernelGGL(k_sph_accel, dim3(pb_blocks(n, 256)), dim3(256), 0, st, A, s->a);
  }
  PB_LAUNCH_CHECK(s->ctx, "k_sph_accel");
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
