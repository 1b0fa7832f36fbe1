// pb_dem.hip -- DEM contact path of cudaParticleDEM (+ cudaParticleRotation)
// on MI355X.  Compiled without FP contraction: the contact decisions
// (|r_ij|^2 < (r_i + r_j)^2 in FP64) are the same IEEE operations as the
// FP64 oracle's, so the contact CSR is reproducible bit for bit.
//
// Reference path (cudaParticleDEM.cu:29-213, kernelfuncs.cu:1088-1309,
// kerneltemplate.hh:135-168, 311-418, kernelinline.hh:235-382): a
// coordination-number pass, a single-block partial sum for the row
// pointers, a host read of nnz, a column-index pass, a per-row sort, the
// history carry-over from the previous matrix and the contact-force pass.
//
// Here the same stages run as: a gather of (x, y, z, radius) and the
// original index into cell order -> count (one thread per particle in cell
// order, 27 or 125 neighbour cells as contiguous ranges of that view) ->
// LDS-tiled exclusive scan -> fill + in-row insertion sort (rows are short)
// -> merge carry of the tangential history (both rows sorted) -> contact
// force and torque.  The column arrays are sized from the previous step's
// count; the stages after the scan stop at that capacity and the host reads
// nnz and the time-step check together at the end, repeating the stages
// after the scan with larger arrays in the rare step where nnz outgrew them
// (one host synchronisation per call).  Rows are indexed by the original particle index and hold
// original indices in ascending order.  Contacts use the strict
// |r_ij| < r_i + r_j for both the count and the list (the reference lists
// with <=, which overruns a row on an exact touch, kernelfuncs.cu:1158).
#include "pb_internal.h"
#include <algorithm>
#include <cmath>
#include <vector>

__global__ void k_fill_u32(uint32_t *x, uint32_t n, uint32_t val);

struct DemGrid {
  const uint32_t *perm, *cellid, *cstart;
  const double4 *rs;        // cell order: (x, y, z, radius)
  const uint32_t *so;       // cell order: original index
  const uint8_t *move;
  uint32_t nb[3];
  int w;            // neighbour half-width: 1 (27 cells) or 2 (125, single-particle cells)
  uint32_t n;
  uint32_t cap;     // capacity of the column arrays
  uint32_t *rowbuf; // DEM_ROWBUF contacts per particle, by original index
};

// contacts a particle's row keeps from the counting scan; longer rows (more
// touching neighbours than any sphere packing gives) are scanned again
#define DEM_ROWBUF 16u

// positions with the radius in w, and original indices, in cell order
__global__ void k_dem_gather(const uint32_t *perm, const double4 *r, const double *rad,
                             const uint32_t *orig, uint32_t n, double4 *rs, uint32_t *so) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = perm[k];
  const double4 p = r[s];
  rs[k] = make_double4(p.x, p.y, p.z, rad[s]);
  so[k] = orig[s];
}

// visit the sorted slots of the (2w+1)^3 cells around cell c (non-periodic:
// DEM distances never use an image, kernelinline.hh:248)
template <typename F>
__device__ __forceinline__ void dem_cells(const DemGrid &G, uint32_t c, F &&f) {
  const uint32_t nz = G.nb[2], nyz = G.nb[1] * nz;
  const int i0 = (int)(c / nyz), j0 = (int)((c / nz) % G.nb[1]), k0 = (int)(c % nz);
  const int kl = max(k0 - G.w, 0), kh = min(k0 + G.w, (int)nz - 1);
  for (int a = -G.w; a <= G.w; ++a) {
    const int i2 = i0 + a;
    if (i2 < 0 || i2 >= (int)G.nb[0]) continue;
    for (int b = -G.w; b <= G.w; ++b) {
      const int j2 = j0 + b;
      if (j2 < 0 || j2 >= (int)G.nb[1]) continue;
      const uint32_t cb = (uint32_t)i2 * nyz + (uint32_t)j2 * nz;
      const uint32_t q0 = G.cstart[cb + (uint32_t)kl], q1 = G.cstart[cb + (uint32_t)kh + 1];
      for (uint32_t q = q0; q < q1; ++q) f(q);
    }
  }
}

// (x, y, z, radius) of both particles
__device__ __forceinline__ bool in_contact(const double4 &pi, const double4 &pj) {
  const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
  const double R2 = dx * dx + dy * dy + dz * dz;
  const double rc = pi.w + pj.w;
  return R2 < rc * rc;
}

// calcCoord_F4 (kernelinline.hh:371-382): contacts of movable particles
// The same scan keeps the first DEM_ROWBUF contacts (original indices, scan
// order) in the row buffer for the fill.
__global__ void k_dem_count(DemGrid G, uint32_t *cnt) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G.n) return;
  const uint32_t s = G.perm[k];
  const uint32_t i = G.so[k];
  uint32_t c = 0;
  if (G.move[s]) {
    const double4 pi = G.rs[k];
    uint32_t *buf = G.rowbuf + (size_t)i * DEM_ROWBUF;
    dem_cells(G, G.cellid[s], [&](uint32_t q) {
      if (q != k && in_contact(pi, G.rs[q])) {
        if (c < DEM_ROWBUF) buf[c] = G.so[q];
        ++c;
      }
    });
  }
  cnt[i] = c;
}

// makeJlist_WithBlock2/4_F4 (kernelfuncs.cu:1120-1216) + sortColIdx
// (:1262-1290): the thread that owns a row fills it (from the counting
// scan's row buffer, or by scanning again when the row is longer) and
// sorts it.
__global__ void k_dem_fill(DemGrid G, const uint32_t *row, uint32_t *col) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G.n) return;
  const uint32_t s = G.perm[k];
  if (!G.move[s]) return;
  const uint32_t i = G.so[k];
  const uint32_t b = row[i], e = min(row[i + 1], G.cap);   // rows past the capacity: redone
  if (b >= e) return;
  uint32_t p = b;
  if (row[i + 1] - b <= DEM_ROWBUF) {     // the counting scan kept the whole row
    const uint32_t *buf = G.rowbuf + (size_t)i * DEM_ROWBUF;
    for (; p < e; ++p) col[p] = buf[p - b];
  } else {
    const double4 pi = G.rs[k];
    dem_cells(G, G.cellid[s], [&](uint32_t q) {
      if (q != k && in_contact(pi, G.rs[q]) && p < e) col[p++] = G.so[q];
    });
  }
  for (uint32_t x = b + 1; x < e; ++x) {
    const uint32_t v = col[x];
    uint32_t y = x;
    while (y > b && col[y - 1] > v) {
      col[y] = col[y - 1];
      --y;
    }
    col[y] = v;
  }
}

// succeedPrevState (kernelfuncs.cu:1292-1309): both rows are sorted, so a
// merge finds the contacts present in both matrices; new contacts start at 0.
// hw history doubles per contact (1: Hertz slip length, 3: tangential spring).
__global__ void k_dem_carry(const uint32_t *prow, const uint32_t *pcol, const double *pval,
                            const uint32_t *crow, const uint32_t *ccol, double *cval,
                            uint32_t n, uint32_t cap, uint32_t hw) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t p = prow ? prow[i] : 0u;
  const uint32_t pe = prow ? prow[i + 1] : 0u;
  const uint32_t qe = min(crow[i + 1], cap);
  for (uint32_t q = crow[i]; q < qe; ++q) {
    const uint32_t j = ccol[q];
    while (p < pe && pcol[p] < j) ++p;
    const bool hit = p < pe && pcol[p] == j;
    for (uint32_t k = 0; k < hw; ++k)
      cval[(size_t)q * hw + k] = hit ? pval[(size_t)p * hw + k] : 0.0;
  }
}

struct DemForceArgs {
  const uint32_t *orig, *inv;          // storage -> original, original -> storage
  const uint32_t *row, *col;
  double *val;
  const double4 *r, *v, *w;
  const double *rad;
  const uint8_t *move;
  double4 *F, *T;
  uint32_t *bad;                         // Cundall-Strack violations
  uint32_t n, npad;
  uint32_t cap;                          // capacity of col / val
  double E, mu, s2, gamma2, mu_r, dt;
  double kn, zeta, kt;                   // linear law
};

// DEMContact_F4 (kernelinline.hh:235-363) over the CSR rows
// (calcF_IJpairWithList_F4, kerneltemplate.hh:135-168), FP64.  F and T are
// overwritten (the reference clears both first, cudaParticleDEM.cu:168-169).
// One thread per storage slot, so the particle's own state and its F and T
// are coalesced; padding slots are skipped (they stay zero).
__global__ void __launch_bounds__(256) k_dem_force(DemForceArgs A) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.npad) return;
  const uint32_t i = A.orig[s];
  if (i == PB_NONE) return;
  const double4 pi = A.r[s], vi = A.v[s], wi = A.w[s];
  const double ri = A.rad[s];
  double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
  uint32_t bad = 0;
  const bool wnz = wi.x != 0 && wi.y != 0 && wi.z != 0;
  const uint32_t qe = min(A.row[i + 1], A.cap);
  for (uint32_t q = A.row[i]; q < qe; ++q) {
    const uint32_t sj = A.inv[A.col[q]];
    const double4 pj = A.r[sj], vj = A.v[sj], wj = A.w[sj];
    const double rj = A.rad[sj];
    const double d[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
    const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double rs = ri + rj;
    if (!(R2 < rs * rs)) continue;
    const double R = sqrt(R2);
    const double nv[3] = {d[0] / R, d[1] / R, d[2] / R};
    const double c = -ri / rs;
    const double Rij[3] = {c * d[0], c * d[1], c * d[2]};
    const double vij[3] = {vi.x - vj.x, vi.y - vj.y, vi.z - vj.z};
    const double xi = rs - R;
    const double C3 = vij[0] * nv[0] + vij[1] * nv[1] + vij[2] * nv[2];
    const double Rbar = ri * rj / rs;
    const double mbar = 1.0 / (vi.w + vj.w);
    const double Kn = 2.0 / 3.0 * A.E * sqrt(Rbar * xi);
    const double C4 = Kn * xi - A.gamma2 * sqrt(mbar * Kn) * C3;
    const double Fn[3] = {nv[0] * C4, nv[1] * C4, nv[2] * C4};
    if (!(A.dt < 2.0 * sqrt(mbar / Kn))) ++bad;    // Cundall-Strack, kernelinline.hh:285-286
    const double F0 = sqrt(Fn[0] * Fn[0] + Fn[1] * Fn[1] + Fn[2] * Fn[2]);
    const double ws[3] = {wi.x + wj.x, wi.y + wj.y, wi.z + wj.z};
    double vt[3] = {vij[0] - C3 * nv[0], vij[1] - C3 * nv[1], vij[2] - C3 * nv[2]};
    vt[0] += ws[1] * Rij[2] - ws[2] * Rij[1];
    vt[1] += ws[2] * Rij[0] - ws[0] * Rij[2];
    vt[2] += ws[0] * Rij[1] - ws[1] * Rij[0];
    double Fs[3] = {0, 0, 0};
    if (vt[0] != 0 && vt[1] != 0 && vt[2] != 0) {   // as the reference, kernelinline.hh:296
      const double xsm = A.mu * A.s2 * xi;
      const double vn = sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
      const double xs = A.val[q];
      const double v1 = -A.mu * F0 * ((xs < xsm) ? (xs / xsm) : 1.0) / vn;
      Fs[0] = v1 * vt[0];
      Fs[1] = v1 * vt[1];
      Fs[2] = v1 * vt[2];
      A.val[q] = xs + vn * A.dt;                    // history, owned by this row
    }
    const double Ts[3] = {Rij[1] * Fs[2] - Rij[2] * Fs[1], Rij[2] * Fs[0] - Rij[0] * Fs[2],
                          Rij[0] * Fs[1] - Rij[1] * Fs[0]};
    double Tr[3] = {0, 0, 0};
    if (wnz) {
      const double wn = sqrt(wi.x * wi.x + wi.y * wi.y + wi.z * wi.z);
      const double cr = -A.mu_r * ri * F0 / wn;
      Tr[0] = cr * wi.x;
      Tr[1] = cr * wi.y;
      Tr[2] = cr * wi.z;
    }
    const double Q = 2.0 - (double)A.move[sj];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      f[x] += (Fn[x] + Fs[x]) * Q;
      t[x] += (Ts[x] + Tr[x]) * Q;
    }
  }
  A.F[s] = make_double4(f[0], f[1], f[2], 0.0);
  A.T[s] = make_double4(t[0], t[1], t[2], 0.0);
  if (bad) atomicAdd(A.bad, bad);
}

// Linear spring-dashpot normal force with a tangential spring and a Coulomb
// cap (Cundall & Strack, Geotechnique 29 (1979) 47-65; no reference
// counterpart: the granular settle workload's law, behind the same calls as
// the Hertz kernel above).  Per contact, n = (r_i - r_j)/|r_ij|,
// delta = rad_i + rad_j - |r_ij|, 1/meff = minv_i + minv_j (0 for a fixed j):
//   Fn = (kn delta - 2 zeta sqrt(meff kn) v_ij.n) n
//   vt = tangential part of v_ij - (rad_i w_i + rad_j w_j) x n
//   s  = history projected on the tangent plane + vt dt, Ft = -kt s,
//        capped at mu |Fn| (the spring then follows the cap, s = -Ft/kt)
//   F_i += Fn + Ft, T_i += (-rad_i n) x Ft
// The history (3 doubles per contact) belongs to row i and is updated in
// place.  FP64, one thread per storage slot as k_dem_force.
__global__ void __launch_bounds__(256) k_dem_force_lsd(DemForceArgs A) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.npad) return;
  const uint32_t i = A.orig[s];
  if (i == PB_NONE) return;
  const double4 pi = A.r[s], vi = A.v[s], wi = A.w[s];
  const double ri = A.rad[s];
  double f[3] = {0, 0, 0}, t[3] = {0, 0, 0};
  uint32_t bad = 0;
  const uint32_t qe = min(A.row[i + 1], A.cap);
  for (uint32_t q = A.row[i]; q < qe; ++q) {
    const uint32_t sj = A.inv[A.col[q]];
    const double4 pj = A.r[sj], vj = A.v[sj], wj = A.w[sj];
    const double rj = A.rad[sj];
    const double d[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
    const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double rs = ri + rj;
    if (!(R2 < rs * rs)) continue;
    const double R = sqrt(R2);
    const double nv[3] = {d[0] / R, d[1] / R, d[2] / R};
    const double delta = rs - R;
    const double meff = 1.0 / (vi.w + (A.move[sj] ? vj.w : 0.0));
    const double gn = 2.0 * A.zeta * sqrt(meff * A.kn);
    if (!(A.dt < 2.0 * sqrt(meff / A.kn))) ++bad;
    const double vij[3] = {vi.x - vj.x, vi.y - vj.y, vi.z - vj.z};
    const double vn = vij[0] * nv[0] + vij[1] * nv[1] + vij[2] * nv[2];
    const double fn = A.kn * delta - gn * vn;
    const double W[3] = {ri * wi.x + rj * wj.x, ri * wi.y + rj * wj.y, ri * wi.z + rj * wj.z};
    const double Wxn[3] = {W[1] * nv[2] - W[2] * nv[1], W[2] * nv[0] - W[0] * nv[2],
                           W[0] * nv[1] - W[1] * nv[0]};
    double *h = A.val + (size_t)q * 3;
    const double h0 = h[0], h1 = h[1], h2 = h[2];
    const double hn = h0 * nv[0] + h1 * nv[1] + h2 * nv[2];
    double sp[3], ft[3];
    sp[0] = h0 - hn * nv[0] + (vij[0] - Wxn[0] - vn * nv[0]) * A.dt;
    sp[1] = h1 - hn * nv[1] + (vij[1] - Wxn[1] - vn * nv[1]) * A.dt;
    sp[2] = h2 - hn * nv[2] + (vij[2] - Wxn[2] - vn * nv[2]) * A.dt;
#pragma unroll
    for (int k = 0; k < 3; ++k) ft[k] = -A.kt * sp[k];
    const double ftm = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
    const double cap = A.mu * fabs(fn);
    if (ftm > cap) {
      const double c = cap / ftm;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        ft[k] *= c;
        sp[k] = -ft[k] / A.kt;
      }
    }
    h[0] = sp[0];
    h[1] = sp[1];
    h[2] = sp[2];
    const double nxf[3] = {nv[1] * ft[2] - nv[2] * ft[1], nv[2] * ft[0] - nv[0] * ft[2],
                           nv[0] * ft[1] - nv[1] * ft[0]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      f[k] += fn * nv[k] + ft[k];
      t[k] += -ri * nxf[k];
    }
  }
  A.F[s] = make_double4(f[0], f[1], f[2], 0.0);
  A.T[s] = make_double4(t[0], t[1], t[2], 0.0);
  if (bad) atomicAdd(A.bad, bad);
}

__global__ void k_dem_inv(const uint32_t *orig, uint32_t n, uint32_t *inv) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && orig[s] != PB_NONE) inv[orig[s]] = s;
}

// propagateEulerianRotation_F4, kernelfuncs.cu:966-986
__global__ void k_rot_step(double4 *w, double4 *L, const double4 *T, const double *m0inv,
                           const uint8_t *move, uint32_t n, double dt) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double mv = (double)move[s];
  const double d = dt * mv, m0 = m0inv[s] * mv;
  double4 l = L[s];
  const double4 t = T[s];
  l.x += t.x * d;
  l.y += t.y * d;
  l.z += t.z * d;
  L[s] = l;
  w[s] = make_double4(l.x * m0, l.y * m0, l.z * m0, 0.0);
}

// m0inv = minv / (0.4 r0^2), cudaParticleRotation::setInertia (.cu:35-61)
__global__ void k_set_inertia(const double4 *v, const double *r0, uint32_t n, double *m0inv) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double r = r0[s];
  m0inv[s] = r > 0 ? v[s].w * (1.0 / (0.4 * r * r)) : 0.0;
}

// ------------------------------------------------------------------ host
static int upload_radius(pb_sys *s, const double *r0, double r0_all) {
  int rc = pb_ensure_radius(s);
  if (rc) return rc;
  std::vector<uint32_t> o = pb_download_orig(s, &rc);
  if (rc) return rc;
  std::vector<double> h(s->npad, 0.0);
  for (size_t k = 0; k < s->npad; ++k)
    if (o[k] != PB_NONE) {
      const double x = r0 ? r0[o[k]] : r0_all;
      if (!(x > 0)) return pb_fail(s->ctx, PB_ERR_ARG, "DEM: radius must be > 0");
      h[k] = x;
    }
  PB_HIP(s->ctx, hipMemcpyAsync(s->radius, h.data(), 8 * (size_t)s->npad, hipMemcpyHostToDevice,
                                s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  return PB_OK;
}

// a change of contact law changes the history width: the contact matrices
// start again (as after the first set-up)
static void set_law(pb_sys *s, int law, uint32_t hw) {
  if (s->dem_hw != hw) {
    for (int b = 0; b < 2; ++b) {
      pb_free(s->cm_val[b]);
      s->cm_val[b] = nullptr;
      s->cm_cap[b] = 0;
      s->cm_nnz[b] = 0;
    }
    s->cm_cur = -1;
    s->cm_prev = false;
    s->cm_rewound = false;
  }
  s->dem_law = law;
  s->dem_hw = hw;
}

extern "C" int pb_dem_set_properties(pb_sys *s, double E, double mu, double sigma, double gamma,
                                     double mu_r, const double *r0, double r0_all) {
  if (!s) return PB_ERR_ARG;
  if (s->nrep != 1) return pb_fail(s->ctx, PB_ERR_ARG, "DEM: single replica only");
  if (!(sigma < 0.5)) return pb_fail(s->ctx, PB_ERR_ARG, "DEM: Poisson ratio must be < 0.5");
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  // cudaParticleDEM::setDEMProperties, cudaParticleDEM.cu:221-252
  s->dem_E = E * (1.0 - sigma * sigma);
  s->dem_mu = mu;
  s->dem_s2 = (2.0 - sigma) / 2.0 / (1.0 - sigma);
  s->dem_gamma2 = gamma;
  s->dem_mu_r = mu_r;
  if (int rc = upload_radius(s, r0, r0_all)) return rc;
  if (int rc = pb_ensure_rotation(s)) return rc;
  set_law(s, 0, 1);
  s->dem_set = true;
  return PB_OK;
}

extern "C" int pb_dem_set_linear(pb_sys *s, double kn, double e, double kt, double mu,
                                 const double *r0, double r0_all) {
  if (!s) return PB_ERR_ARG;
  if (s->nrep != 1) return pb_fail(s->ctx, PB_ERR_ARG, "DEM: single replica only");
  if (!(kn > 0) || !(kt > 0) || !(mu >= 0))
    return pb_fail(s->ctx, PB_ERR_ARG, "pb_dem_set_linear: need kn > 0, kt > 0, mu >= 0");
  if (!(e > 0 && e <= 1))
    return pb_fail(s->ctx, PB_ERR_ARG, "pb_dem_set_linear: restitution must be in (0, 1]");
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  const double le = std::log(e);
  s->dem_kn = kn;
  s->dem_kt = kt;
  s->dem_zeta = -le / std::sqrt(M_PI * M_PI + le * le);   // damping ratio of restitution e
  s->dem_mu = mu;
  if (int rc = upload_radius(s, r0, r0_all)) return rc;
  if (int rc = pb_ensure_rotation(s)) return rc;
  set_law(s, 1, 3);
  s->dem_set = true;
  return PB_OK;
}

extern "C" int pb_rot_set_inertia(pb_sys *s, const double *r0, double r0_all) {
  if (!s) return PB_ERR_ARG;
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  int rc = pb_ensure_rotation(s);
  if (rc) return rc;
  // the inertia radius may differ from the contact radius: stage it in a
  // temporary device array in storage order
  std::vector<uint32_t> o = pb_download_orig(s, &rc);
  if (rc) return rc;
  std::vector<double> h(s->npad, 0.0);
  for (size_t k = 0; k < s->npad; ++k)
    if (o[k] != PB_NONE) {
      const double x = r0 ? r0[o[k]] : r0_all;
      if (!(x > 0)) return pb_fail(s->ctx, PB_ERR_ARG, "setInertia: radius must be > 0");
      h[k] = x;
    }
  double *d = nullptr;
  if ((rc = pb_alloc(s->ctx, &d, s->npad))) return rc;
  if (!s->m0inv && (rc = pb_alloc(s->ctx, &s->m0inv, s->npad))) {
    pb_free(d);
    return rc;
  }
  hipStream_t st = s->ctx->stream;
  hipError_t e = hipMemcpyAsync(d, h.data(), 8 * (size_t)s->npad, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_set_inertia, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, st, s->v, d,
                       s->npad, s->m0inv);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemsetAsync(s->w, 0, sizeof(double4) * s->npad, st);
  if (e == hipSuccess) e = hipMemsetAsync(s->Lm, 0, sizeof(double4) * s->npad, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  pb_free(d);
  if (e != hipSuccess) return pb_hip_check(s->ctx, e, "pb_rot_set_inertia");
  return PB_OK;
}

extern "C" int pb_rot_step(pb_sys *s, double dt) {
  if (!s) return PB_ERR_ARG;
  if (!s->m0inv) return pb_fail(s->ctx, PB_ERR_STATE, "rotation: pb_rot_set_inertia first");
  hipLaunchKernelGGL(k_rot_step, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, s->ctx->stream,
                     s->w, s->Lm, s->T, s->m0inv, s->move, s->npad, dt);
  PB_LAUNCH_CHECK(s->ctx, "k_rot_step");
  return PB_OK;
}

static int ensure_csr(pb_sys *s, int b, uint32_t nnz) {
  int rc;
  if (!s->cm_row[b] && (rc = pb_alloc(s->ctx, &s->cm_row[b], (size_t)s->n + 1))) return rc;
  if (s->cm_cap[b] < nnz || !s->cm_col[b]) {
    const uint32_t cap = (uint32_t)std::min<uint64_t>(((uint64_t)nnz * 5 / 4 + 512), 0xFFFFFFFFull);
    if ((rc = pb_alloc(s->ctx, &s->cm_col[b], cap)) ||
        (rc = pb_alloc(s->ctx, &s->cm_val[b], (size_t)cap * s->dem_hw)))
      return rc;
    s->cm_cap[b] = cap;
  }
  return PB_OK;
}

extern "C" int pb_dem_calc_force(pb_sys *s, double dt, int autotune, uint32_t *nnz_out) {
  if (!s) return PB_ERR_ARG;
  if (!s->dem_set) return pb_fail(s->ctx, PB_ERR_STATE, "DEM properties not set");
  if (!s->ref_binned) return pb_fail(s->ctx, PB_ERR_STATE, "DEM: pb_bin (calcBlockID) first");
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const uint32_t n = s->n;
  int rc;
  if (!s->dem_inv) {
    if ((rc = pb_alloc(s->ctx, &s->dem_inv, (size_t)n + 1))) return rc;
  }
  hipLaunchKernelGGL(k_dem_inv, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, st, s->orig, s->npad,
                     s->dem_inv);
  if (!s->dem_rs) {
    if ((rc = pb_alloc(s->ctx, &s->dem_rs, s->npad)) || (rc = pb_alloc(s->ctx, &s->dem_so, s->npad)) ||
        (rc = pb_alloc(s->ctx, &s->dem_rowbuf, (size_t)s->npad * DEM_ROWBUF)))
      return rc;
  }
  if (n)
    hipLaunchKernelGGL(k_dem_gather, dim3(pb_blocks(n, 256)), dim3(256), 0, st, s->perm, s->r,
                       s->radius, s->orig, n, s->dem_rs, s->dem_so);
  DemGrid G;
  G.perm = s->perm;
  G.cellid = s->cellid;
  G.cstart = s->cstart;
  G.rs = s->dem_rs;
  G.so = s->dem_so;
  G.rowbuf = s->dem_rowbuf;
  G.move = s->move;
  for (int a = 0; a < 3; ++a) G.nb[a] = s->blocknum[a];
  G.w = s->ref_single ? 2 : 1;
  G.n = n;
  const int b = (s->cm_cur + 1) & 1;         // the other buffer keeps the previous matrix
  // capacity from the previous matrix (ensure_csr adds a quarter)
  if ((rc = ensure_csr(s, b, s->cm_cur >= 0 ? s->cm_nnz[s->cm_cur] : 0))) return rc;
  // per-row counts go to the sort workspace's rank array (free once pb_bin
  // has produced perm / cstart), their exclusive scan to the new row array
  uint32_t *cnt = s->rank;
  const uint32_t grid = pb_blocks(n ? n : 1, 256);
  if (n) {
    hipLaunchKernelGGL(k_dem_count, dim3(grid), dim3(256), 0, st, G, cnt);
    PB_LAUNCH_CHECK(s->ctx, "k_dem_count");
  }
  if ((rc = pb_exclusive_scan(s, cnt, s->cm_row[b], n))) return rc;
  DemForceArgs A;
  A.orig = s->orig;
  A.inv = s->dem_inv;
  A.r = s->r;
  A.v = s->v;
  A.w = s->w;
  A.rad = s->radius;
  A.move = s->move;
  A.F = s->F;
  A.T = s->T;
  A.bad = s->d_counter + 8;
  A.n = n;
  A.npad = s->npad;
  A.E = s->dem_E;
  A.mu = s->dem_mu;
  A.s2 = s->dem_s2;
  A.gamma2 = s->dem_gamma2;
  A.mu_r = s->dem_mu_r;
  A.dt = dt;
  A.kn = s->dem_kn;
  A.zeta = s->dem_zeta;
  A.kt = s->dem_kt;
  const int p = s->cm_cur;
  // fill + sort, history carry, contact forces and torques up to the
  // capacity; then nnz and the time-step check in one read
  auto contacts_and_force = [&]() -> int {
    G.cap = A.cap = s->cm_cap[b];
    A.row = s->cm_row[b];
    A.col = s->cm_col[b];
    A.val = s->cm_val[b];
    if (n) {
      hipLaunchKernelGGL(k_dem_fill, dim3(grid), dim3(256), 0, st, G, s->cm_row[b], s->cm_col[b]);
      hipLaunchKernelGGL(k_dem_carry, dim3(grid), dim3(256), 0, st,
                         p >= 0 ? s->cm_row[p] : (const uint32_t *)nullptr,
                         p >= 0 ? s->cm_col[p] : (const uint32_t *)nullptr,
                         p >= 0 ? s->cm_val[p] : (const double *)nullptr, s->cm_row[b],
                         s->cm_col[b], s->cm_val[b], n, s->cm_cap[b], s->dem_hw);
      PB_LAUNCH_CHECK(s->ctx, "k_dem_fill/carry");
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, s->d_counter + 8, 1u, 0u);
    // every real particle's F and T is written (empty rows give zero);
    // padding slots are never written and stay zero
    if (n) {
      if (s->dem_law == 1)
        hipLaunchKernelGGL(k_dem_force_lsd, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, st, A);
      else
        hipLaunchKernelGGL(k_dem_force, dim3(pb_blocks(s->npad, 256)), dim3(256), 0, st, A);
    }
    PB_LAUNCH_CHECK(s->ctx, "k_dem_force");
    return PB_OK;
  };
  if ((rc = contacts_and_force())) return rc;
  uint32_t nnz = 0, bad = 0;
  PB_HIP(s->ctx, hipMemcpyAsync(&nnz, s->cm_row[b] + n, 4, hipMemcpyDeviceToHost, st));
  PB_HIP(s->ctx, hipMemcpyAsync(&bad, s->d_counter + 8, 4, hipMemcpyDeviceToHost, st));
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  if (nnz > s->cm_cap[b]) {                  // the matrix outgrew the arrays: again, larger
    if ((rc = ensure_csr(s, b, nnz)) || (rc = contacts_and_force())) return rc;
    PB_HIP(s->ctx, hipMemcpyAsync(&bad, s->d_counter + 8, 4, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
  }
  s->cm_nnz[b] = nnz;
  s->cm_prev = p >= 0;
  s->cm_rewound = false;
  s->cm_cur = b;
  s->dem_bad = bad;
  if (nnz_out) *nnz_out = nnz;
  if (bad && !autotune)
    return pb_fail(s->ctx, PB_ERR_DT, "DEM: time step above the Cundall-Strack bound 2 sqrt(m/Kn)");
  return PB_OK;
}

extern "C" int pb_dem_get_contacts(pb_sys *s, uint32_t *rowptr, uint32_t *colidx, double *val,
                                   uint32_t cap, uint32_t *nnz) {
  if (!s) return PB_ERR_ARG;
  if (s->cm_cur < 0) return pb_fail(s->ctx, PB_ERR_STATE, "DEM: no contact matrix yet");
  const int b = s->cm_cur;
  const uint32_t z = s->cm_nnz[b];
  if (nnz) *nnz = z;
  hipStream_t st = s->ctx->stream;
  if (rowptr)
    PB_HIP(s->ctx, hipMemcpyAsync(rowptr, s->cm_row[b], 4 * ((size_t)s->n + 1),
                                  hipMemcpyDeviceToHost, st));
  if ((colidx || val) && z > cap)
    return pb_fail(s->ctx, PB_ERR_ARG, "pb_dem_get_contacts: capacity below nnz");
  if (colidx && z)
    PB_HIP(s->ctx, hipMemcpyAsync(colidx, s->cm_col[b], 4 * (size_t)z, hipMemcpyDeviceToHost, st));
  if (val && z)
    PB_HIP(s->ctx, hipMemcpyAsync(val, s->cm_val[b], 8 * (size_t)z * s->dem_hw,
                                  hipMemcpyDeviceToHost, st));
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  return PB_OK;
}

extern "C" int pb_dem_set_contacts(pb_sys *s, const uint32_t *rowptr, const uint32_t *colidx,
                                   const double *val, uint32_t nnz) {
  if (!s || !rowptr || (nnz && (!colidx || !val))) return PB_ERR_ARG;
  const uint32_t n = s->n;
  // a CSR the history merge can read: rows of original indices, columns
  // ascending and in range
  if (rowptr[0] != 0 || rowptr[n] != nnz)
    return pb_fail(s->ctx, PB_ERR_ARG, "pb_dem_set_contacts: rowptr does not span nnz");
  for (uint32_t i = 0; i < n; ++i) {
    if (rowptr[i + 1] < rowptr[i])
      return pb_fail(s->ctx, PB_ERR_ARG, "pb_dem_set_contacts: rowptr decreases");
    for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (colidx[k] >= n || (k > rowptr[i] && colidx[k] <= colidx[k - 1]))
        return pb_fail(s->ctx, PB_ERR_ARG, "pb_dem_set_contacts: columns out of range or unsorted");
  }
  const int b = s->cm_cur < 0 ? 0 : s->cm_cur;
  if (int rc = ensure_csr(s, b, nnz)) return rc;
  hipStream_t st = s->ctx->stream;
  PB_HIP(s->ctx, hipMemcpyAsync(s->cm_row[b], rowptr, 4 * ((size_t)n + 1), hipMemcpyHostToDevice, st));
  if (nnz) {
    PB_HIP(s->ctx, hipMemcpyAsync(s->cm_col[b], colidx, 4 * (size_t)nnz, hipMemcpyHostToDevice, st));
    PB_HIP(s->ctx, hipMemcpyAsync(s->cm_val[b], val, 8 * (size_t)nnz * s->dem_hw,
                                  hipMemcpyHostToDevice, st));
  }
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  s->cm_nnz[b] = nnz;
  s->cm_cur = b;
  s->cm_prev = false;                        // the other buffer is not its predecessor
  s->cm_rewound = false;
  return PB_OK;
}

// cudaParticleDEM::rollback's `_N -= 1` (cudaParticleDEM.cu:302-306): the
// matrix before the latest becomes the latest again, so the next
// pb_dem_calc_force overwrites the discarded one and carries the history
// of the step before it.  When the latest matrix has no predecessor (the
// first call, or a matrix restored by pb_dem_set_contacts) the next call
// starts every contact at zero history, as the reference's cleared second
// matrix gives.  Only one earlier step is kept, so a second rewind without a
// pb_dem_calc_force in between is refused.
extern "C" int pb_dem_rewind(pb_sys *s) {
  if (!s) return PB_ERR_ARG;
  if (s->cm_rewound)
    return pb_fail(s->ctx, PB_ERR_STATE, "pb_dem_rewind: already rewound (one step is kept)");
  if (s->cm_cur < 0)
    return pb_fail(s->ctx, PB_ERR_STATE, "pb_dem_rewind: no contact matrix to rewind");
  s->cm_cur = s->cm_prev ? (s->cm_cur ^ 1) : -1;
  s->cm_prev = false;
  s->cm_rewound = true;
  return PB_OK;
}

extern "C" int pb_dem_history_width(pb_sys *s, uint32_t *hw) {
  if (!s || !hw) return PB_ERR_ARG;
  *hw = s->dem_hw;
  return PB_OK;
}

extern "C" int pb_dem_last_violations(pb_sys *s, uint32_t *count) {
  if (!s || !count) return PB_ERR_ARG;
  *count = s->dem_bad;
  return PB_OK;
}
