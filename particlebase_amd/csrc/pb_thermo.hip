// pb_thermo.hip -- MD thermostat helpers on the storage layout:
//   * Gaussian isokinetic velocity Verlet (GaussianThermo::TimeEvolution,
//     GaussianThermo.cu:29-83),
//   * adjustVelocities (cudaParticleMD.cu:177-198, GaussianThermo.cu:98-119),
//   * constTemp (cudaParticleMD.cu:140-155).
//
// Gaussian step.  The reference solves the implicit velocity half of
//   v(t) = v(t-dt) + dt/2m (F(t) + F(t-dt)) - dt/2 (v(t-dt) xi(t-dt) + v(t) xi(t)),
//   xi(t) = sum F(t).v(t) / sum m v(t)^2
// by a Newton-Raphson loop with three host round trips per pass, stopping
// when the rms update of a pass is <= tol.  Writing
// A = v(t-dt) + dt/2m (F + Fold) - dt/2 v(t-dt) xi(t-dt), the system reads
// v = A / (1 + dt/2 xi(v)); every solution is v = A / c with one scalar c,
// and xi(A/c) = c (F.A)/(m A.A) gives c = 1 / (1 - dt/2 g), g = F.A / m A.A.
// So v(t) = A (1 - dt/2 g) and xi(t) = g / (1 - dt/2 g) is the fixed point
// the loop converges to.  The loop's first pass uses the previous step's
// xi; when that pass already moves v by no more than tol (always so when
// xi(t-dt) = 0, as on the first step) the reference stops there with the
// one-pass velocity and xi of it, so that pass is evaluated too and decides
// the branch exactly as the loop's test does.  Five launches, no host
// synchronisation:
//   k_gt_a       A into v, v' = v + dt/2m (F + Fold) into a; partials
//                F.A, m A.A, m v'.v'
//   k_gt_solve1  per replica (fixed order): 1/c, xi of the fixed point,
//                1 / sum m v'.v'
//   k_gt_pass1   the loop's first Newton pass from v' with the previous xi:
//                v1 = v' - f/f' into a; partials (f/f')^2, F.v1, m v1.v1
//   k_gt_solve2  rms of the pass vs tol: one-pass result or fixed point
//   k_gt_drift   v = v1 or A/c ; r += v (dt - xi dt^2/2) + F dt^2/2m ;
//                Fold = F
// Per-replica xi persists on the device between steps (GaussianThermo::xi).
#include "pb_internal.h"

#define GT_BLOCK 256

namespace {

// gt holds 4 doubles per replica: [0] xi (persistent), [1] 1/c of the
// fixed point or -1 when the one-pass result is taken, [2] 1 / sum m v'.v',
// [3] xi of the fixed point.

__global__ void __launch_bounds__(GT_BLOCK)
    k_gt_a(double4 *v, const double4 *F, double4 *Fold, const uint32_t *orig,
           const uint32_t *waveRep, const double *gt, uint32_t n, double dt, double *partials) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double fa = 0.0, ma2 = 0.0, mv2 = 0.0;
  if (s < n && orig[s] != PB_NONE) {
    const double xi = gt[4 * waveRep[s / PB_WAVE]];
    const double dt2 = 0.5 * dt;
    double4 V = v[s];
    const double4 f = F[s], fo = Fold[s];
    const double dtm2 = dt2 * V.w, xd = xi * dt2;
    const double px = V.x + (f.x + fo.x) * dtm2;     // v'
    const double py = V.y + (f.y + fo.y) * dtm2;
    const double pz = V.z + (f.z + fo.z) * dtm2;
    const double ax = px - V.x * xd, ay = py - V.y * xd, az = pz - V.z * xd;   // A
    fa = f.x * ax + f.y * ay + f.z * az;
    ma2 = (ax * ax + ay * ay + az * az) / V.w;
    mv2 = (px * px + py * py + pz * pz) / V.w;
    V.x = ax;
    V.y = ay;
    V.z = az;
    v[s] = V;
    Fold[s] = make_double4(px, py, pz, 0.0);
  }
  fa = pb_wave_sum(fa);
  ma2 = pb_wave_sum(ma2);
  mv2 = pb_wave_sum(mv2);
  if (pb_lane() == 0 && s < n) {
    double *P = partials + 4 * (s / PB_WAVE);
    P[0] = fa;
    P[1] = ma2;
    P[2] = mv2;
  }
}

// One workgroup.  A replica whose A vanishes (all particles at rest, no
// force) has no defined xi: g = 0.
__global__ void __launch_bounds__(256)
    k_gt_solve1(const double *partials, const RepInfo *reps, uint32_t nrep, double dt,
                double *out, double *gt) {
  __shared__ double sm[4 * 256];
  pb_block_reduce_reps(partials, 3, reps, nrep, out, false, sm);
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < nrep; q += blockDim.x) {
    const double fa = out[8 * q], ma2 = out[8 * q + 1], mv2 = out[8 * q + 2];
    const double g = ma2 > 0.0 ? fa / ma2 : 0.0;
    const double cinv = 1.0 - 0.5 * dt * g;
    gt[4 * q + 1] = cinv;
    gt[4 * q + 2] = mv2 > 0.0 ? 1.0 / mv2 : 0.0;
    gt[4 * q + 3] = g / cinv;
  }
}

// The first pass of calcGaussianThermoFoverF_F4 (kernelfuncs.cu:407-454):
// f = A - (dt/2 xi + 1) v', f' = -dt/2 (xi + (F - 2 xi m v') v' / sum m v'^2) - 1.
__global__ void __launch_bounds__(GT_BLOCK)
    k_gt_pass1(const double4 *v, const double4 *F, double4 *Fold, const uint32_t *orig,
               const uint32_t *waveRep, const double *gt, uint32_t n, double dt,
               double *partials) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double qq = 0.0, fv = 0.0, mv2 = 0.0;
  if (s < n && orig[s] != PB_NONE) {
    const double *G = gt + 4 * waveRep[s / PB_WAVE];
    const double xi = G[0], mv2inv = G[2], dt2 = 0.5 * dt;
    const double4 A = v[s], f = F[s], p = Fold[s];
    const double m = 1.0 / A.w, k1 = dt2 * xi + 1.0, k2 = 2.0 * xi * m;
    const double qx = (A.x - k1 * p.x) / (-dt2 * (xi + (f.x - k2 * p.x) * p.x * mv2inv) - 1.0);
    const double qy = (A.y - k1 * p.y) / (-dt2 * (xi + (f.y - k2 * p.y) * p.y * mv2inv) - 1.0);
    const double qz = (A.z - k1 * p.z) / (-dt2 * (xi + (f.z - k2 * p.z) * p.z * mv2inv) - 1.0);
    const double x = p.x - qx, y = p.y - qy, z = p.z - qz;
    qq = qx * qx + qy * qy + qz * qz;
    fv = f.x * x + f.y * y + f.z * z;
    mv2 = (x * x + y * y + z * z) * m;
    Fold[s] = make_double4(x, y, z, 0.0);
  }
  qq = pb_wave_sum(qq);
  fv = pb_wave_sum(fv);
  mv2 = pb_wave_sum(mv2);
  if (pb_lane() == 0 && s < n) {
    double *P = partials + 4 * (s / PB_WAVE);
    P[0] = qq;
    P[1] = fv;
    P[2] = mv2;
  }
}

// One workgroup: the loop's test after its first pass, v0 = sqrt(sum q^2 /
// 3N) <= tol, picks the one-pass result; otherwise the fixed point, which
// has no solution when 1/c <= 0 (dt too large for the forces; the
// reference's loop does not converge there): counted in bad[0] and taken
// as xi = 0, v = A.
__global__ void __launch_bounds__(256)
    k_gt_solve2(const double *partials, const RepInfo *reps, uint32_t nrep, double tol,
                double *out, double *gt, uint32_t *bad) {
  __shared__ double sm[4 * 256];
  pb_block_reduce_reps(partials, 3, reps, nrep, out, false, sm);
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < nrep; q += blockDim.x) {
    double *G = gt + 4 * q;
    const double v0 = sqrt(out[8 * q] / (3.0 * reps[q].natoms));
    if (!(v0 > tol)) {
      const double mv2 = out[8 * q + 2];
      G[0] = mv2 > 0.0 ? out[8 * q + 1] / mv2 : 0.0;
      G[1] = -1.0;
    } else if (G[1] > 0.0 && isfinite(G[3])) {
      G[0] = G[3];
    } else {
      atomicAdd(bad, 1u);
      G[0] = 0.0;
      G[1] = 1.0;
    }
  }
}

__global__ void __launch_bounds__(GT_BLOCK)
    k_gt_drift(double4 *r, double4 *v, const double4 *F, double4 *Fold, const uint32_t *orig,
               const uint32_t *waveRep, const double *gt, uint32_t n, double dt) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const double *G = gt + 4 * waveRep[s / PB_WAVE];
  const double xi = G[0], cinv = G[1];
  double4 V = v[s];
  const double4 f = F[s];
  if (cinv < 0.0) {
    const double4 p = Fold[s];
    V.x = p.x;
    V.y = p.y;
    V.z = p.z;
  } else {
    V.x *= cinv;
    V.y *= cinv;
    V.z *= cinv;
  }
  const double dt22 = 0.5 * dt * dt;
  const double cv = dt - xi * dt22, cf = dt22 * V.w;
  double4 R = r[s];
  R.x += V.x * cv + f.x * cf;
  R.y += V.y * cv + f.y * cf;
  R.z += V.z * cv + f.z * cf;
  r[s] = R;
  v[s] = V;
  Fold[s] = make_double4(f.x, f.y, f.z, 0.0);
}

// per-wave partials of sum v (K = 3 components) for the velocity adjustment
__global__ void __launch_bounds__(GT_BLOCK)
    k_vsum_partial(const double4 *v, const uint32_t *orig, uint32_t n, double *partials) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double x = 0.0, y = 0.0, z = 0.0;
  if (s < n && orig[s] != PB_NONE) {
    const double4 V = v[s];
    x = V.x;
    y = V.y;
    z = V.z;
  }
  x = pb_wave_sum(x);
  y = pb_wave_sum(y);
  z = pb_wave_sum(z);
  if (pb_lane() == 0 && s < n) {
    double *P = partials + 4 * (s / PB_WAVE);
    P[0] = x;
    P[1] = y;
    P[2] = z;
  }
}

// per-wave partials of sum (v - mean)^2 per axis; means at stats[6q + 0..2]
__global__ void __launch_bounds__(GT_BLOCK)
    k_vvar_partial(const double4 *v, const uint32_t *orig, const uint32_t *waveRep,
                   const double *stats, uint32_t n, double *partials) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double x = 0.0, y = 0.0, z = 0.0;
  if (s < n && orig[s] != PB_NONE) {
    const double *S = stats + 6 * waveRep[s / PB_WAVE];
    const double4 V = v[s];
    x = V.x - S[0];
    y = V.y - S[1];
    z = V.z - S[2];
    x *= x;
    y *= y;
    z *= z;
  }
  x = pb_wave_sum(x);
  y = pb_wave_sum(y);
  z = pb_wave_sum(z);
  if (pb_lane() == 0 && s < n) {
    double *P = partials + 4 * (s / PB_WAVE);
    P[0] = x;
    P[1] = y;
    P[2] = z;
  }
}

// v = (v - mean) * v1 / sqrt(var); mean and variance per replica in
// stats[6q + 0..5] (means, then variances).
__global__ void __launch_bounds__(GT_BLOCK)
    k_vadjust(double4 *v, const uint32_t *orig, const uint32_t *waveRep, const double *stats,
              uint32_t n, double v1) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const double *S = stats + 6 * waveRep[s / PB_WAVE];
  double4 V = v[s];
  V.x = (V.x - S[0]) * (v1 / sqrt(S[3]));
  V.y = (V.y - S[1]) * (v1 / sqrt(S[4]));
  V.z = (V.z - S[2]) * (v1 / sqrt(S[5]));
  v[s] = V;
}

// means / variances from the two passes of sums; one workgroup
__global__ void __launch_bounds__(256)
    k_vstats(const double *partials, const RepInfo *reps, uint32_t nrep, double *out,
             double *stats, int pass) {
  __shared__ double sm[4 * 256];
  pb_block_reduce_reps(partials, 3, reps, nrep, out, false, sm);
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < nrep; q += blockDim.x) {
    const double inv = 1.0 / (double)reps[q].natoms;
    for (int c = 0; c < 3; ++c) stats[6 * q + 3 * pass + c] = out[8 * q + c] * inv;
  }
}

// per-wave partials of sum F.v and sum m v^2 (constTemp)
__global__ void __launch_bounds__(GT_BLOCK)
    k_fv_partial(const double4 *v, const double4 *F, const uint32_t *orig, uint32_t n,
                 double *partials) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double fv = 0.0, mv2 = 0.0;
  if (s < n && orig[s] != PB_NONE) {
    const double4 V = v[s], f = F[s];
    fv = f.x * V.x + f.y * V.y + f.z * V.z;
    mv2 = (V.x * V.x + V.y * V.y + V.z * V.z) / V.w;
  }
  fv = pb_wave_sum(fv);
  mv2 = pb_wave_sum(mv2);
  if (pb_lane() == 0 && s < n) {
    partials[4 * (s / PB_WAVE) + 0] = fv;
    partials[4 * (s / PB_WAVE) + 1] = mv2;
  }
}

// lambda per replica into out[8q + 4] (0 when sum m v^2 vanishes)
__global__ void __launch_bounds__(256)
    k_ct_solve(const double *partials, const RepInfo *reps, uint32_t nrep, double *out) {
  __shared__ double sm[4 * 256];
  pb_block_reduce_reps(partials, 2, reps, nrep, out, false, sm);
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < nrep; q += blockDim.x) {
    const double fv = out[8 * q], mv2 = out[8 * q + 1];
    out[8 * q + 4] = mv2 > 0.0 ? fv / mv2 : 0.0;
  }
}

// F -= lambda m v (correctConstTemp_F4, kernelfuncs.cu:533-544)
__global__ void __launch_bounds__(GT_BLOCK)
    k_ct_apply(const double4 *v, double4 *F, const uint32_t *orig, const uint32_t *waveRep,
               const double *out, uint32_t n) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const double lam = out[8 * waveRep[s / PB_WAVE] + 4];
  if (lam == 0.0) return;
  const double4 V = v[s];
  double4 f = F[s];
  const double lm = lam / V.w;
  f.x -= lm * V.x;
  f.y -= lm * V.y;
  f.z -= lm * V.z;
  F[s] = f;
}

int ensure_gt(pb_sys *s) {
  if (s->gt) return PB_OK;
  if (int rc = pb_alloc(s->ctx, &s->gt_bad, 1)) return rc;
  if (int rc = pb_alloc(s->ctx, &s->gt, 4 * (size_t)s->nrep)) return rc;
  const uint32_t zero = 0;
  PB_HIP(s->ctx, hipMemcpyAsync(s->gt_bad, &zero, sizeof(uint32_t), hipMemcpyHostToDevice,
                                s->ctx->stream));
  std::vector<double> h(4 * (size_t)s->nrep, 0.0);   // xi(0) = 0 (GaussianThermo() : xi(0))
  PB_HIP(s->ctx, hipMemcpyAsync(s->gt, h.data(), sizeof(double) * h.size(),
                                hipMemcpyHostToDevice, s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  return PB_OK;
}

}  // namespace

#define GT_GRID(n) dim3(pb_blocks((n), GT_BLOCK)), dim3(GT_BLOCK), 0, s->ctx->stream

extern "C" int pb_gt_step(pb_sys *s, double dt, double tol) {
  if (!s) return PB_ERR_ARG;
  if (!(dt > 0.0)) return pb_fail(s->ctx, PB_ERR_ARG, "pb_gt_step: dt must be positive");
  if (!(tol >= 0.0)) return pb_fail(s->ctx, PB_ERR_ARG, "pb_gt_step: tol must be >= 0");
  if (int rc = ensure_gt(s)) return rc;
  hipStream_t st = s->ctx->stream;
  hipLaunchKernelGGL(k_gt_a, GT_GRID(s->npad), s->v, s->F, s->a, s->orig, s->waveRep, s->gt,
                     s->npad, dt, s->partials);
  hipLaunchKernelGGL(k_gt_solve1, dim3(1), dim3(256), 0, st, s->partials, s->d_reps, s->nrep, dt,
                     s->d_out, s->gt);
  hipLaunchKernelGGL(k_gt_pass1, GT_GRID(s->npad), s->v, s->F, s->a, s->orig, s->waveRep, s->gt,
                     s->npad, dt, s->partials);
  hipLaunchKernelGGL(k_gt_solve2, dim3(1), dim3(256), 0, st, s->partials, s->d_reps, s->nrep, tol,
                     s->d_out, s->gt, s->gt_bad);
  hipLaunchKernelGGL(k_gt_drift, GT_GRID(s->npad), s->r, s->v, s->F, s->a, s->orig, s->waveRep,
                     s->gt, s->npad, dt);
  PB_LAUNCH_CHECK(s->ctx, "pb_gt_step");
  return PB_OK;
}

extern "C" int pb_gt_xi(pb_sys *s, const double *set_xi, double *get_xi, uint32_t *bad_steps) {
  if (!s) return PB_ERR_ARG;
  if (int rc = ensure_gt(s)) return rc;
  std::vector<double> h(4 * (size_t)s->nrep);
  PB_HIP(s->ctx, hipMemcpyAsync(h.data(), s->gt, sizeof(double) * h.size(),
                                hipMemcpyDeviceToHost, s->ctx->stream));
  uint32_t nb = 0;
  PB_HIP(s->ctx, hipMemcpyAsync(&nb, s->gt_bad, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  if (get_xi)
    for (uint32_t q = 0; q < s->nrep; ++q) get_xi[q] = h[4 * q];
  if (bad_steps) *bad_steps = nb;
  if (set_xi) {
    for (uint32_t q = 0; q < s->nrep; ++q) h[4 * q] = set_xi[q];
    PB_HIP(s->ctx, hipMemcpyAsync(s->gt, h.data(), sizeof(double) * h.size(),
                                  hipMemcpyHostToDevice, s->ctx->stream));
    nb = 0;
    PB_HIP(s->ctx, hipMemcpyAsync(s->gt_bad, &nb, sizeof(uint32_t), hipMemcpyHostToDevice,
                                  s->ctx->stream));
    PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  }
  return PB_OK;
}

extern "C" int pb_adjust_velocities(pb_sys *s, double v1, double *stats) {
  if (!s) return PB_ERR_ARG;
  if (!(v1 >= 0.0)) return pb_fail(s->ctx, PB_ERR_ARG, "pb_adjust_velocities: v1 < 0");
  double *dstats = s->d_out + 8 * (size_t)s->nrep;   // 6 per replica after the 8-word records
  // pass 0: sums -> means; pass 1: centred squares -> variances (two-pass)
  hipLaunchKernelGGL(k_vsum_partial, GT_GRID(s->npad), s->v, s->orig, s->npad, s->partials);
  hipLaunchKernelGGL(k_vstats, dim3(1), dim3(256), 0, s->ctx->stream, s->partials, s->d_reps,
                     s->nrep, s->d_out, dstats, 0);
  hipLaunchKernelGGL(k_vvar_partial, GT_GRID(s->npad), s->v, s->orig, s->waveRep, dstats,
                     s->npad, s->partials);
  hipLaunchKernelGGL(k_vstats, dim3(1), dim3(256), 0, s->ctx->stream, s->partials, s->d_reps,
                     s->nrep, s->d_out, dstats, 1);
  PB_LAUNCH_CHECK(s->ctx, "pb_adjust_velocities (statistics)");
  std::vector<double> h(6 * (size_t)s->nrep);
  PB_HIP(s->ctx, hipMemcpyAsync(h.data(), dstats, sizeof(double) * h.size(),
                                hipMemcpyDeviceToHost, s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  for (uint32_t q = 0; q < s->nrep; ++q)
    for (int c = 0; c < 3; ++c)
      if (!(h[6 * q + 3 + c] > 0.0))
        return pb_fail(s->ctx, PB_ERR_ARG,
                       "pb_adjust_velocities: a velocity component has zero variance");
  if (stats) std::copy(h.begin(), h.end(), stats);
  hipLaunchKernelGGL(k_vadjust, GT_GRID(s->npad), s->v, s->orig, s->waveRep, dstats, s->npad, v1);
  PB_LAUNCH_CHECK(s->ctx, "pb_adjust_velocities");
  return PB_OK;
}

extern "C" int pb_const_temp(pb_sys *s, double *mv2, double *lambda) {
  if (!s) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_fv_partial, GT_GRID(s->npad), s->v, s->F, s->orig, s->npad, s->partials);
  hipLaunchKernelGGL(k_ct_solve, dim3(1), dim3(256), 0, s->ctx->stream, s->partials, s->d_reps,
                     s->nrep, s->d_out);
  hipLaunchKernelGGL(k_ct_apply, GT_GRID(s->npad), s->v, s->F, s->orig, s->waveRep, s->d_out,
                     s->npad);
  PB_LAUNCH_CHECK(s->ctx, "pb_const_temp");
  if (mv2 || lambda) {
    std::vector<double> h(8 * (size_t)s->nrep);
    PB_HIP(s->ctx, hipMemcpyAsync(h.data(), s->d_out, sizeof(double) * h.size(),
                                  hipMemcpyDeviceToHost, s->ctx->stream));
    PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    for (uint32_t q = 0; q < s->nrep; ++q) {
      if (mv2) mv2[q] = h[8 * q + 1];
      if (lambda) lambda[q] = h[8 * q + 4];
    }
  }
  return PB_OK;
}
