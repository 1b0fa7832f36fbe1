// pb_core.hip -- context, particle storage, import/export, integrators,
// boundaries and FP64 reductions of the MI355X particle engine.
//
// Reference counterparts: cudaFrame (cudaFrame.cu), cudaParticleBase
// (cudaParticleBase.cu), cudaParticleVV (cudaParticleVV.cu), cudaParticleLF
// (cudaParticleLF.cu) and the element-wise kernels of kernelfuncs.cu.
#include "pb_internal.h"
#include <algorithm>
#include <cmath>
#include <cstring>

int pb_fail(pb_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}
int pb_hip_check(pb_ctx *ctx, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return pb_fail(ctx, PB_ERR_HIP, m);
}

// ================================================================= context
extern "C" int pb_ctx_create(int device, pb_ctx **out) {
  if (!out) return PB_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PB_ERR_HIP;
  if (device < 0 || device >= ndev) return PB_ERR_ARG;
  pb_ctx *c = new pb_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return PB_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->cu_count = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    c->global_mem = (uint64_t)prop.totalGlobalMem;
    c->max_threads = (uint32_t)prop.maxThreadsPerBlock;
    c->max_grid = (uint32_t)prop.maxGridSize[0];
    c->lds_per_block = (uint32_t)prop.sharedMemPerBlock;
  }
  // default-priority stream (never a different priority)
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PB_ERR_HIP;
  }
  if (pb_gr_scratch_init(c) != PB_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PB_ERR_HIP;
  }
  *out = c;
  return PB_OK;
}

extern "C" int pb_ctx_destroy(pb_ctx *ctx) {
  if (!ctx) return PB_ERR_ARG;
  // a system keeps its context's device and stream: destroying the context
  // first would leave every later call on the system (its destroy included)
  // with a freed context
  if (ctx->nsys)
    return pb_fail(ctx, PB_ERR_STATE,
                   "pb_ctx_destroy: " + std::to_string(ctx->nsys) +
                       " system(s) still alive on this context; pb_sys_destroy them first");
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
  return PB_OK;
}

extern "C" const char *pb_ctx_last_error(const pb_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int pb_ctx_sync(pb_ctx *ctx) {
  if (!ctx) return PB_ERR_ARG;
  PB_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PB_OK;
}

extern "C" int pb_ctx_info(pb_ctx *ctx, char *name, size_t name_len, int *cu_count) {
  if (!ctx) return PB_ERR_ARG;
  if (name && name_len) {
    std::strncpy(name, ctx->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = ctx->cu_count;
  return PB_OK;
}

extern "C" int pb_ctx_limits(pb_ctx *ctx, uint64_t *global_mem, uint32_t *max_threads,
                             uint32_t *max_grid, uint32_t *lds_per_block,
                             uint32_t *lanes_per_cu) {
  if (!ctx) return PB_ERR_ARG;
  if (global_mem) *global_mem = ctx->global_mem;
  if (max_threads) *max_threads = ctx->max_threads;
  if (max_grid) *max_grid = ctx->max_grid;
  if (lds_per_block) *lds_per_block = ctx->lds_per_block;
  // a CDNA compute unit issues to 4 SIMDs of 16 lanes each
  if (lanes_per_cu) *lanes_per_cu = 64;
  return PB_OK;
}

// ================================================================== system
extern "C" int pb_sys_create(pb_ctx *ctx, uint32_t nrep, const uint32_t *counts,
                             pb_sys **out) {
  if (!ctx || !out || nrep == 0 || !counts) return pb_fail(ctx, PB_ERR_ARG, "pb_sys_create: bad args");
  *out = nullptr;
  uint64_t total = 0, totpad = 0;
  for (uint32_t q = 0; q < nrep; ++q) {
    if (counts[q] == 0) return pb_fail(ctx, PB_ERR_ARG, "pb_sys_create: empty replica");
    total += counts[q];
    totpad += (counts[q] + PB_WAVE - 1) / PB_WAVE * PB_WAVE;
  }
  if (totpad >= (1ull << 31)) return pb_fail(ctx, PB_ERR_ARG, "pb_sys_create: too many particles");
  PB_HIP(ctx, hipSetDevice(ctx->device));
  pb_sys *s = new pb_sys();
  s->ctx = ctx;
  ++ctx->nsys;
  s->nrep = nrep;
  s->n = (uint32_t)total;
  s->npad = (uint32_t)totpad;
  s->nwave = s->npad / PB_WAVE;
  s->counts.assign(counts, counts + nrep);
  s->offsets.resize(nrep);
  s->reps.resize(nrep);
  s->cells.assign(6 * (size_t)nrep, 0.0);
  uint32_t off = 0, slot = 0;
  for (uint32_t q = 0; q < nrep; ++q) {
    RepInfo &R = s->reps[q];
    std::memset(&R, 0, sizeof(R));
    s->offsets[q] = off;
    R.atom0 = slot;
    R.natoms = counts[q];
    R.npad = (counts[q] + PB_WAVE - 1) / PB_WAVE * PB_WAVE;
    R.wave0 = slot / PB_WAVE;
    R.nwave = R.npad / PB_WAVE;
    off += counts[q];
    slot += R.npad;
  }
  int rc = PB_OK;
  const size_t np = s->npad;
  if ((rc = pb_alloc(ctx, &s->r, np)) || (rc = pb_alloc(ctx, &s->v, np)) ||
      (rc = pb_alloc(ctx, &s->a, np)) || (rc = pb_alloc(ctx, &s->F, np)) ||
      (rc = pb_alloc(ctx, &s->move, np)) || (rc = pb_alloc(ctx, &s->type, np)) ||
      (rc = pb_alloc(ctx, &s->orig, np)) || (rc = pb_alloc(ctx, &s->waveRep, s->nwave)) ||
      (rc = pb_alloc(ctx, &s->d_reps, nrep)) ||
      (rc = pb_alloc(ctx, &s->partials, 4 * (size_t)s->nwave)) ||
      (rc = pb_alloc(ctx, &s->d_counter, 16)) || (rc = pb_alloc(ctx, &s->d_out, 16 * (size_t)nrep)) ||
      (rc = pb_alloc(ctx, &s->d_lj, 1))) {
    pb_sys_destroy(s);
    return rc;
  }
  std::vector<uint32_t> wr(s->nwave);
  for (uint32_t q = 0; q < nrep; ++q)
    for (uint32_t k = 0; k < s->reps[q].nwave; ++k) wr[s->reps[q].wave0 + k] = q;
  std::vector<uint32_t> orig(np, PB_NONE);
  for (uint32_t q = 0; q < nrep; ++q)
    for (uint32_t i = 0; i < counts[q]; ++i) orig[s->reps[q].atom0 + i] = s->offsets[q] + i;
  std::vector<uint16_t> ty(np, (uint16_t)PB_PAD_TYPE);
  for (uint32_t q = 0; q < nrep; ++q)
    for (uint32_t i = 0; i < counts[q]; ++i) ty[s->reps[q].atom0 + i] = 0;
  hipStream_t st = ctx->stream;
  if (hipMemcpyAsync(s->waveRep, wr.data(), 4 * wr.size(), hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(s->orig, orig.data(), 4 * np, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(s->type, ty.data(), 2 * np, hipMemcpyHostToDevice, st) ||
      hipMemsetAsync(s->r, 0, sizeof(double4) * np, st) ||
      hipMemsetAsync(s->v, 0, sizeof(double4) * np, st) ||
      hipMemsetAsync(s->a, 0, sizeof(double4) * np, st) ||
      hipMemsetAsync(s->F, 0, sizeof(double4) * np, st) ||
      hipMemsetAsync(s->move, 0, np, st) ||
      hipMemsetAsync(s->d_counter, 0, 16 * 4, st) ||
      hipStreamSynchronize(st)) {
    pb_sys_destroy(s);
    return pb_fail(ctx, PB_ERR_HIP, "pb_sys_create: upload failed");
  }
  *out = s;
  return PB_OK;
}

extern "C" void pb_sys_destroy(pb_sys *s) {
  if (!s) return;
  if (s->ctx) {
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
  }
  pb_free(s->r); pb_free(s->v); pb_free(s->a); pb_free(s->F);
  pb_free(s->w); pb_free(s->Lm); pb_free(s->T); pb_free(s->radius);
  pb_free(s->move); pb_free(s->type); pb_free(s->orig); pb_free(s->waveRep);
  pb_free(s->d_reps); pb_free(s->partials); pb_free(s->d_counter); pb_free(s->d_out);
  pb_free(s->d_hash); pb_free(s->lj_queue);
  pb_free(s->gt); pb_free(s->gt_bad);
  pb_free(s->md_row); pb_free(s->md_col); pb_free(s->md_inv);
  pb_free(s->ref_table); pb_free(s->ref_bid); pb_free(s->ref_pid); pb_free(s->ref_bindex);
  pb_free(s->cellid); pb_free(s->rank); pb_free(s->perm); pb_free(s->ccount);
  pb_free(s->cstart); pb_free(s->scan_tmp); pb_free(s->d_lj);
  pb_free(s->nl); pb_free(s->nn); pb_free(s->rbuild); pb_free(s->r2buf); pb_free(s->v2buf);
  pb_free(s->orig2); pb_free(s->type2); pb_free(s->move2); pb_free(s->a2buf);
  pb_free(s->sph_mu0); pb_free(s->sph_c2); pb_free(s->sph_rhoinv); pb_free(s->sph_num);
  pb_free(s->sph_rs); pb_free(s->sph_vs); pb_free(s->sph_ds); pb_free(s->m0inv);
  pb_free(s->dem_inv); pb_free(s->dem_rs); pb_free(s->dem_so); pb_free(s->dem_rowbuf);
  for (int b = 0; b < 2; ++b) {
    pb_free(s->cm_row[b]); pb_free(s->cm_col[b]); pb_free(s->cm_val[b]);
  }
  pb_free(s->gr_hist);
  pb_free(s->list_segs); pb_free(s->list_wtab);
  if (s->hmon) (void)hipHostFree(s->hmon);
  if (s->ctx) --s->ctx->nsys;
  delete s;
}

extern "C" uint32_t pb_sys_count(const pb_sys *s) { return s ? s->n : 0; }
extern "C" uint32_t pb_sys_replicas(const pb_sys *s) { return s ? s->nrep : 0; }

static int upload_reps(pb_sys *s) {
  PB_HIP(s->ctx, hipMemcpyAsync(s->d_reps, s->reps.data(), sizeof(RepInfo) * s->nrep,
                                hipMemcpyHostToDevice, s->ctx->stream));
  return PB_OK;
}

extern "C" int pb_sys_set_cell(pb_sys *s, const double *cells) {
  if (!s || !cells) return PB_ERR_ARG;
  for (uint32_t q = 0; q < s->nrep; ++q) {
    RepInfo &R = s->reps[q];
    for (int a = 0; a < 3; ++a) {
      const double lo = cells[6 * q + 2 * a], hi = cells[6 * q + 2 * a + 1];
      if (!(hi > lo)) return pb_fail(s->ctx, PB_ERR_ARG, "pb_sys_set_cell: empty box");
      R.lo[a] = lo;
      R.hi[a] = hi;
      R.L[a] = hi - lo;
      R.Linv[a] = 1.0 / (hi - lo);
    }
  }
  std::memcpy(s->cells.data(), cells, sizeof(double) * 6 * s->nrep);
  s->cell_set = true;
  s->lj_cont = false;             // the fused LJ run's list cells follow the box
  return upload_reps(s);
}

// host-side packing: original order <-> storage order
std::vector<uint32_t> pb_download_orig(pb_sys *s, int *rc) {
  std::vector<uint32_t> o(s->npad);
  *rc = PB_OK;
  if (hipMemcpyAsync(o.data(), s->orig, 4 * (size_t)s->npad, hipMemcpyDeviceToHost,
                     s->ctx->stream) != hipSuccess ||
      hipStreamSynchronize(s->ctx->stream) != hipSuccess)
    *rc = pb_fail(s->ctx, PB_ERR_HIP, "download orig");
  return o;
}

extern "C" int pb_sys_import(pb_sys *s, const double *r, const double *v, const double *a,
                             const double *m, const uint32_t *type, const uint8_t *is_fixed) {
  if (!s || !r || !m) return pb_fail(s ? s->ctx : nullptr, PB_ERR_ARG, "pb_sys_import: r and m required");
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  int rc = PB_OK;
  std::vector<uint32_t> o = pb_download_orig(s, &rc);
  if (rc) return rc;
  const size_t np = s->npad;
  std::vector<double4> hr(np), hv(np), ha(np);
  std::vector<uint16_t> ht(np);
  std::vector<uint8_t> hm(np);
  for (size_t k = 0; k < np; ++k) {
    const uint32_t g = o[k];
    if (g == PB_NONE) {
      hr[k] = make_double4(0, 0, 0, (double)PB_PAD_TYPE);
      hv[k] = make_double4(0, 0, 0, 0);
      ha[k] = make_double4(0, 0, 0, 0);
      ht[k] = (uint16_t)PB_PAD_TYPE;
      hm[k] = 0;
      continue;
    }
    if (type && type[g] >= PB_MAXT) return pb_fail(s->ctx, PB_ERR_ARG, "pb_sys_import: type >= 5");
    if (!(m[g] > 0)) return pb_fail(s->ctx, PB_ERR_ARG, "pb_sys_import: mass must be > 0");
    const double t = type ? (double)type[g] : 0.0;
    hr[k] = make_double4(r[3 * g], r[3 * g + 1], r[3 * g + 2], t);
    hv[k] = v ? make_double4(v[3 * g], v[3 * g + 1], v[3 * g + 2], 1.0 / m[g])
              : make_double4(0, 0, 0, 1.0 / m[g]);
    ha[k] = a ? make_double4(a[3 * g], a[3 * g + 1], a[3 * g + 2], 0) : make_double4(0, 0, 0, 0);
    ht[k] = type ? (uint16_t)type[g] : 0;
    hm[k] = is_fixed ? (is_fixed[g] ? 0 : 1) : 1;
  }
  hipStream_t st = s->ctx->stream;
  PB_HIP(s->ctx, hipMemcpyAsync(s->r, hr.data(), sizeof(double4) * np, hipMemcpyHostToDevice, st));
  PB_HIP(s->ctx, hipMemcpyAsync(s->v, hv.data(), sizeof(double4) * np, hipMemcpyHostToDevice, st));
  PB_HIP(s->ctx, hipMemcpyAsync(s->a, ha.data(), sizeof(double4) * np, hipMemcpyHostToDevice, st));
  PB_HIP(s->ctx, hipMemcpyAsync(s->type, ht.data(), 2 * np, hipMemcpyHostToDevice, st));
  PB_HIP(s->ctx, hipMemcpyAsync(s->move, hm.data(), np, hipMemcpyHostToDevice, st));
  PB_HIP(s->ctx, hipMemsetAsync(s->F, 0, sizeof(double4) * np, st));
  PB_HIP(s->ctx, hipStreamSynchronize(st));
  s->ref_binned = false;
  return PB_OK;
}

static double4 *vec_field(pb_sys *s, int field) {
  switch (field) {
    case PB_FIELD_R: return s->r;
    case PB_FIELD_V: return s->v;
    case PB_FIELD_A: return s->a;
    case PB_FIELD_F: return s->F;
    case PB_FIELD_W: return s->w;
    case PB_FIELD_L: return s->Lm;
    case PB_FIELD_T: return s->T;
    case PB_FIELD_SPH_NUM: return s->sph_num;
    default: return nullptr;
  }
}

int pb_ensure_rotation(pb_sys *s) {
  if (s->w) return PB_OK;
  int rc;
  if ((rc = pb_alloc(s->ctx, &s->w, s->npad)) || (rc = pb_alloc(s->ctx, &s->Lm, s->npad)) ||
      (rc = pb_alloc(s->ctx, &s->T, s->npad)))
    return rc;
  hipStream_t st = s->ctx->stream;
  PB_HIP(s->ctx, hipMemsetAsync(s->w, 0, sizeof(double4) * s->npad, st));
  PB_HIP(s->ctx, hipMemsetAsync(s->Lm, 0, sizeof(double4) * s->npad, st));
  PB_HIP(s->ctx, hipMemsetAsync(s->T, 0, sizeof(double4) * s->npad, st));
  return PB_OK;
}
int pb_ensure_radius(pb_sys *s) {
  if (s->radius) return PB_OK;
  int rc = pb_alloc(s->ctx, &s->radius, s->npad);
  if (rc) return rc;
  PB_HIP(s->ctx, hipMemsetAsync(s->radius, 0, sizeof(double) * s->npad, s->ctx->stream));
  return PB_OK;
}

extern "C" int pb_sys_get(pb_sys *s, int field, void *host) {
  if (!s || !host) return PB_ERR_ARG;
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  int rc = PB_OK;
  std::vector<uint32_t> o = pb_download_orig(s, &rc);
  if (rc) return rc;
  const size_t np = s->npad;
  hipStream_t st = s->ctx->stream;
  if (double4 *src = vec_field(s, field)) {
    std::vector<double4> h(np);
    PB_HIP(s->ctx, hipMemcpyAsync(h.data(), src, sizeof(double4) * np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    double *out = (double *)host;
    for (size_t k = 0; k < np; ++k)
      if (o[k] != PB_NONE) {
        out[3 * (size_t)o[k]] = h[k].x;
        out[3 * (size_t)o[k] + 1] = h[k].y;
        out[3 * (size_t)o[k] + 2] = h[k].z;
      }
    return PB_OK;
  }
  if (field == PB_FIELD_W || field == PB_FIELD_L || field == PB_FIELD_T)
    return pb_fail(s->ctx, PB_ERR_STATE, "pb_sys_get: rotation state not allocated");
  if (field == PB_FIELD_SPH_NUM || field == PB_FIELD_SPH_RHOINV) {
    if (!s->sph_num) return pb_fail(s->ctx, PB_ERR_STATE, "pb_sys_get: no SPH density yet");
    if (field == PB_FIELD_SPH_RHOINV) {
      std::vector<double> h(np);
      PB_HIP(s->ctx, hipMemcpyAsync(h.data(), s->sph_rhoinv, 8 * np, hipMemcpyDeviceToHost, st));
      PB_HIP(s->ctx, hipStreamSynchronize(st));
      double *out = (double *)host;
      for (size_t k = 0; k < np; ++k)
        if (o[k] != PB_NONE) out[o[k]] = h[k];
      return PB_OK;
    }
  }
  if (field == PB_FIELD_MINV) {
    std::vector<double4> h(np);
    PB_HIP(s->ctx, hipMemcpyAsync(h.data(), s->v, sizeof(double4) * np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    double *out = (double *)host;
    for (size_t k = 0; k < np; ++k)
      if (o[k] != PB_NONE) out[o[k]] = h[k].w;
    return PB_OK;
  }
  if (field == PB_FIELD_TYPE) {
    std::vector<uint16_t> h(np);
    PB_HIP(s->ctx, hipMemcpyAsync(h.data(), s->type, 2 * np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    uint16_t *out = (uint16_t *)host;
    for (size_t k = 0; k < np; ++k)
      if (o[k] != PB_NONE) out[o[k]] = h[k];
    return PB_OK;
  }
  if (field == PB_FIELD_MOVE) {
    std::vector<uint8_t> h(np);
    PB_HIP(s->ctx, hipMemcpyAsync(h.data(), s->move, np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    uint8_t *out = (uint8_t *)host;
    for (size_t k = 0; k < np; ++k)
      if (o[k] != PB_NONE) out[o[k]] = h[k];
    return PB_OK;
  }
  if (field == PB_FIELD_RADIUS) {
    if (!s->radius) return pb_fail(s->ctx, PB_ERR_STATE, "pb_sys_get: radius not set");
    std::vector<double> h(np);
    PB_HIP(s->ctx, hipMemcpyAsync(h.data(), s->radius, 8 * np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    double *out = (double *)host;
    for (size_t k = 0; k < np; ++k)
      if (o[k] != PB_NONE) out[o[k]] = h[k];
    return PB_OK;
  }
  return pb_fail(s->ctx, PB_ERR_ARG, "pb_sys_get: unknown field");
}

extern "C" int pb_sys_set(pb_sys *s, int field, const void *host) {
  if (!s || !host) return PB_ERR_ARG;
  PB_HIP(s->ctx, hipSetDevice(s->ctx->device));
  int rc = PB_OK;
  if ((field == PB_FIELD_W || field == PB_FIELD_L || field == PB_FIELD_T) && (rc = pb_ensure_rotation(s)))
    return rc;
  if (field == PB_FIELD_RADIUS && (rc = pb_ensure_radius(s))) return rc;
  if (field == PB_FIELD_SPH_NUM && !s->sph_num)
    return pb_fail(s->ctx, PB_ERR_STATE, "pb_sys_set: no SPH density yet");
  std::vector<uint32_t> o = pb_download_orig(s, &rc);
  if (rc) return rc;
  const size_t np = s->npad;
  hipStream_t st = s->ctx->stream;
  if (double4 *dst = vec_field(s, field)) {
    std::vector<double4> h(np);
    PB_HIP(s->ctx, hipMemcpyAsync(h.data(), dst, sizeof(double4) * np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    const double *in = (const double *)host;
    for (size_t k = 0; k < np; ++k)
      if (o[k] != PB_NONE) {
        h[k].x = in[3 * (size_t)o[k]];
        h[k].y = in[3 * (size_t)o[k] + 1];
        h[k].z = in[3 * (size_t)o[k] + 2];
      }
    PB_HIP(s->ctx, hipMemcpyAsync(dst, h.data(), sizeof(double4) * np, hipMemcpyHostToDevice, st));
    if (field == PB_FIELD_SPH_NUM && s->sph_rhoinv) {
      // 1/rho follows the imported rho (calcDensityPost's reciprocal,
      // cudaParticleSPH_NS.cu:190-195), so an acceleration after a density
      // import uses the imported densities throughout
      std::vector<double> hr(np, 0.0);
      for (size_t k = 0; k < np; ++k)
        if (o[k] != PB_NONE) hr[k] = 1.0 / in[3 * (size_t)o[k] + 1];
      PB_HIP(s->ctx, hipMemcpyAsync(s->sph_rhoinv, hr.data(), 8 * np, hipMemcpyHostToDevice, st));
    }
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    return PB_OK;
  }
  if (field == PB_FIELD_MINV || field == PB_FIELD_TYPE) {
    std::vector<double4> hr(np), hv(np);
    std::vector<uint16_t> ht(np);
    PB_HIP(s->ctx, hipMemcpyAsync(hv.data(), s->v, sizeof(double4) * np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipMemcpyAsync(hr.data(), s->r, sizeof(double4) * np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipMemcpyAsync(ht.data(), s->type, 2 * np, hipMemcpyDeviceToHost, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    for (size_t k = 0; k < np; ++k)
      if (o[k] != PB_NONE) {
        if (field == PB_FIELD_MINV) {
          hv[k].w = ((const double *)host)[o[k]];
        } else {
          const uint16_t t = ((const uint16_t *)host)[o[k]];
          if (t >= PB_MAXT) return pb_fail(s->ctx, PB_ERR_ARG, "pb_sys_set: type >= 5");
          ht[k] = t;
          hr[k].w = (double)t;
        }
      }
    PB_HIP(s->ctx, hipMemcpyAsync(s->v, hv.data(), sizeof(double4) * np, hipMemcpyHostToDevice, st));
    PB_HIP(s->ctx, hipMemcpyAsync(s->r, hr.data(), sizeof(double4) * np, hipMemcpyHostToDevice, st));
    PB_HIP(s->ctx, hipMemcpyAsync(s->type, ht.data(), 2 * np, hipMemcpyHostToDevice, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    return PB_OK;
  }
  if (field == PB_FIELD_MOVE || field == PB_FIELD_RADIUS) {
    const bool isr = field == PB_FIELD_RADIUS;
    std::vector<uint8_t> hm(np, 0);
    std::vector<double> hrad(np, 0.0);
    for (size_t k = 0; k < np; ++k)
      if (o[k] != PB_NONE) {
        if (isr) hrad[k] = ((const double *)host)[o[k]];
        else hm[k] = ((const uint8_t *)host)[o[k]];
      }
    if (isr)
      PB_HIP(s->ctx, hipMemcpyAsync(s->radius, hrad.data(), 8 * np, hipMemcpyHostToDevice, st));
    else
      PB_HIP(s->ctx, hipMemcpyAsync(s->move, hm.data(), np, hipMemcpyHostToDevice, st));
    PB_HIP(s->ctx, hipStreamSynchronize(st));
    return PB_OK;
  }
  return pb_fail(s->ctx, PB_ERR_ARG, "pb_sys_set: unknown field");
}

// ======================================================= element-wise kernels
// All take the storage slot range [0, npad) and skip padding slots via
// orig[s] == PB_NONE.

__global__ void k_clear4(double4 *x, uint32_t n) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) x[s] = make_double4(0, 0, 0, 0);
}

__global__ void k_add_axis(double4 *x, const uint32_t *orig, uint32_t n, int axis, double val) {
  // addArray_F4 (value), kernelfuncs.cu:27-38
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  double4 t = x[s];
  if (axis == 0) t.x += val;
  else if (axis == 1) t.y += val;
  else t.z += val;
  x[s] = t;
}

__global__ void k_calc_acc(double4 *a, const double4 *F, const double4 *v,
                           const uint32_t *orig, uint32_t n) {
  // calcA_F4, kernelfuncs.cu:70-82
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const double mi = v[s].w;
  const double4 f = F[s];
  double4 A = a[s];
  A.x = f.x * mi;
  A.y = f.y * mi;
  A.z = f.z * mi;
  a[s] = A;
}

__global__ void k_vv_step(double4 *r, double4 *v, const double4 *F, double4 *Fold,
                          const uint32_t *orig, uint32_t n, double dt) {
  // propagateVelocityVerlet_F4, kernelfuncs.cu:333-367 (FP64)
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const double dt2 = dt * 0.5;
  double4 V = v[s];
  const double dtm2 = dt2 * V.w;
  const double4 f = F[s];
  const double4 fo = Fold[s];
  V.x += (f.x + fo.x) * dtm2;
  V.y += (f.y + fo.y) * dtm2;
  V.z += (f.z + fo.z) * dtm2;
  const double dt22 = dtm2 * dt;
  double4 R = r[s];
  R.x += V.x * dt + f.x * dt22;
  R.y += V.y * dt + f.y * dt22;
  R.z += V.z * dt + f.z * dt22;
  Fold[s] = make_double4(f.x, f.y, f.z, 0.0);
  v[s] = V;
  r[s] = R;
}

__global__ void k_euler_step(double4 *r, double4 *v, const double4 *a, const uint8_t *move,
                             const uint32_t *orig, uint32_t n, double dt) {
  // propagateEuler_F4, kernelfuncs.cu:174-192 (FP64): r += v dt, v += a dt
  // for movable particles
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const double d = move[s] ? dt : 0.0;
  double4 R = r[s], V = v[s];
  const double4 A = a[s];
  R.x += V.x * d;
  R.y += V.y * d;
  R.z += V.z * d;
  V.x += A.x * d;
  V.y += A.y * d;
  V.z += A.z * d;
  r[s] = R;
  v[s] = V;
}

__global__ void k_wrap(double4 *r, const uint32_t *orig, const uint32_t *waveRep,
                       const RepInfo *reps, uint32_t n) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const RepInfo &R = reps[waveRep[s / PB_WAVE]];
  double4 p = r[s];
  p.x = wrap1(p.x, R.lo[0], R.hi[0]);
  p.y = wrap1(p.y, R.lo[1], R.hi[1]);
  p.z = wrap1(p.z, R.lo[2], R.hi[2]);
  r[s] = p;
}

__global__ void k_absolute(double4 *r, const uint32_t *orig, const uint32_t *waveRep,
                           const RepInfo *reps, uint32_t n) {
  // treatAbsoluteBoundary_F4, kernelfuncs.cu:141-155
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const RepInfo &R = reps[waveRep[s / PB_WAVE]];
  double4 p = r[s];
  p.x = fmin(fmax(p.x, R.lo[0]), R.hi[0]);
  p.y = fmin(fmax(p.y, R.lo[1]), R.hi[1]);
  p.z = fmin(fmax(p.z, R.lo[2]), R.hi[2]);
  r[s] = p;
}

__global__ void k_reflect(const double4 *r, double4 *v, const uint32_t *orig,
                          const uint32_t *waveRep, const RepInfo *reps, uint32_t n) {
  // treatRefrectBoundary_F4, kernelfuncs.cu:157-172
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const RepInfo &R = reps[waveRep[s / PB_WAVE]];
  const double4 p = r[s];
  double4 V = v[s];
  if (p.x < R.lo[0]) V.x = fabs(V.x);
  if (p.x > R.hi[0]) V.x = -fabs(V.x);
  if (p.y < R.lo[1]) V.y = fabs(V.y);
  if (p.y > R.hi[1]) V.y = -fabs(V.y);
  if (p.z < R.lo[2]) V.z = fabs(V.z);
  if (p.z > R.hi[2]) V.z = -fabs(V.z);
  v[s] = V;
}

__global__ void k_lf_vinit(double4 *v, const double4 *F, const uint32_t *orig, uint32_t n,
                           double dt_2) {
  // calcLFVinit_F4, kernelfuncs.cu:246-258
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  double4 V = v[s];
  const double mi = V.w * dt_2;
  const double4 f = F[s];
  V.x -= f.x * mi;
  V.y -= f.y * mi;
  V.z -= f.z * mi;
  v[s] = V;
}

__global__ void k_lf_step(double4 *r, double4 *v, const double4 *a, const uint8_t *move,
                          const uint32_t *orig, uint32_t n, double dt) {
  // propagateLeapFrog_F4, kernelfuncs.cu:260-284
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const double dt2 = dt * (double)move[s];
  double4 V = v[s];
  const double4 A = a[s];
  V.x += A.x * dt2;
  V.y += A.y * dt2;
  V.z += A.z * dt2;
  v[s] = V;
  double4 R = r[s];
  R.x += V.x * dt2;
  R.y += V.y * dt2;
  R.z += V.z * dt2;
  r[s] = R;
}

__global__ void k_lf_rollback(double4 *r, double4 *v, const double4 *a, const uint8_t *move,
                              const uint32_t *orig, uint32_t n, double dt) {
  // rollbackLeapFrog_F4, kernelfuncs.cu:286-327
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  const double dt2 = dt * (double)move[s];
  double4 V = v[s];
  double4 R = r[s];
  const double4 A = a[s];
  const double rx = R.x - V.x * dt2, ry = R.y - V.y * dt2, rz = R.z - V.z * dt2;
  V.x -= A.x * dt2;
  V.y -= A.y * dt2;
  V.z -= A.z * dt2;
  v[s] = V;
  R.x = rx - V.x * dt2;
  R.y = ry - V.y * dt2;
  R.z = rz - V.z * dt2;
  r[s] = R;
}

__global__ void k_scale_v(double4 *v, const uint32_t *orig, uint32_t n, double sc) {
  // mulArray on v, kerneltemplate.hh:29-35
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || orig[s] == PB_NONE) return;
  double4 V = v[s];
  V.x *= sc;
  V.y *= sc;
  V.z *= sc;
  v[s] = V;
}

// per-wave partial of m v^2 (calcV2_F4 + DOT, kernelfuncs.cu:520-530,
// cudaParticleMD.cu:124-133)
__global__ void k_mv2_partial(const double4 *v, const uint32_t *orig, uint32_t n,
                              double *partials) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double x = 0.0;
  if (s < n && orig[s] != PB_NONE) {
    const double4 V = v[s];
    x = (V.x * V.x + V.y * V.y + V.z * V.z) / V.w;
  }
  x = pb_wave_sum(x);
  if (pb_lane() == 0 && s < n) partials[4 * (s / PB_WAVE)] = x;
}

// inspectV_F4 (kernelfuncs.cu:194-240): per-wave max of |v|/vlim with its
// slot; vlim already truncated to an integer.
__global__ void k_inspect_partial(const double4 *v, const uint32_t *orig, uint32_t n,
                                  double vlim, double *partials) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double x = -1.0;
  if (s < n && orig[s] != PB_NONE) {
    const double4 V = v[s];
    x = sqrt(V.x * V.x + V.y * V.y + V.z * V.z) / vlim;
  }
  double m = pb_wave_max(x);
  // lowest original index attaining the max inside the wave
  double idx = (x == m && s < n && orig[s] != PB_NONE) ? (double)orig[s] : 1e300;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) idx = fmin(idx, __shfl_xor(idx, o, 64));
  if (pb_lane() == 0 && s < n) {
    partials[4 * (s / PB_WAVE)] = m;
    partials[4 * (s / PB_WAVE) + 1] = idx;
  }
}

__global__ void __launch_bounds__(1024) k_inspect_final(const double *partials, uint32_t nwave,
                                                       double *out) {
  // one block: max over waves, ties -> lowest index (a total order, so the
  // result does not depend on the visiting order)
  __shared__ double sm[1024], si[1024];
  double m = -1.0, id = 1e300;
  auto take = [&](double x, double i) {
    if (x > m || (x == m && i < id)) {
      m = x;
      id = i;
    }
  };
  const uint32_t T = blockDim.x;
  uint32_t w = threadIdx.x;
  // four partials per step, all loads issued before the compares
  for (; w + 3 * T < nwave; w += 4 * T) {
    double x[4], i[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      x[u] = partials[4 * (size_t)(w + u * T)];
      i[u] = partials[4 * (size_t)(w + u * T) + 1];
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) take(x[u], i[u]);
  }
  for (; w < nwave; w += T) take(partials[4 * (size_t)w], partials[4 * (size_t)w + 1]);
  sm[threadIdx.x] = m;
  si[threadIdx.x] = id;
  __syncthreads();
  for (int o = T / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const double x = sm[threadIdx.x + o], i = si[threadIdx.x + o];
      if (x > sm[threadIdx.x] || (x == sm[threadIdx.x] && i < si[threadIdx.x])) {
        sm[threadIdx.x] = x;
        si[threadIdx.x] = i;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sm[0];
    out[1] = si[0];
  }
}

// ============================================================ reductions
__global__ void __launch_bounds__(256) k_reduce_waves(const double *partials, int K,
                                                      const RepInfo *reps, uint32_t nrep,
                                                      double *out) {
  __shared__ double sm[4 * 256];
  pb_block_reduce_reps(partials, K, reps, nrep, out, false, sm);
}

int pb_reduce_waves(pb_sys *s, int K) {
  hipLaunchKernelGGL(k_reduce_waves, dim3(1), dim3(256), 0, s->ctx->stream, s->partials, K,
                     s->d_reps, s->nrep, s->d_out);
  PB_LAUNCH_CHECK(s->ctx, "k_reduce_waves");
  return PB_OK;
}

// ================================================================ wrappers
#define PB_GRID(n) dim3(pb_blocks((n), 256)), dim3(256), 0, s->ctx->stream

extern "C" int pb_clear_force(pb_sys *s, int with_fold) {
  if (!s) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_clear4, PB_GRID(s->npad), s->F, s->npad);
  if (with_fold) hipLaunchKernelGGL(k_clear4, PB_GRID(s->npad), s->a, s->npad);
  PB_LAUNCH_CHECK(s->ctx, "pb_clear_force");
  return PB_OK;
}

extern "C" int pb_calc_acceleration(pb_sys *s) {
  if (!s) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_calc_acc, PB_GRID(s->npad), s->a, s->F, s->v, s->orig, s->npad);
  PB_LAUNCH_CHECK(s->ctx, "pb_calc_acceleration");
  return PB_OK;
}

extern "C" int pb_add_force(pb_sys *s, int axis, double value) {
  if (!s || axis < 0 || axis > 2) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_add_axis, PB_GRID(s->npad), s->F, s->orig, s->npad, axis, value);
  PB_LAUNCH_CHECK(s->ctx, "pb_add_force");
  return PB_OK;
}

extern "C" int pb_add_acceleration(pb_sys *s, int axis, double value) {
  if (!s || axis < 0 || axis > 2) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_add_axis, PB_GRID(s->npad), s->a, s->orig, s->npad, axis, value);
  PB_LAUNCH_CHECK(s->ctx, "pb_add_acceleration");
  return PB_OK;
}

static int need_cell(pb_sys *s) {
  return s->cell_set ? PB_OK : pb_fail(s->ctx, PB_ERR_STATE, "cell not set (pb_sys_set_cell)");
}

extern "C" int pb_treat_periodic(pb_sys *s) {
  if (!s) return PB_ERR_ARG;
  if (int rc = need_cell(s)) return rc;
  hipLaunchKernelGGL(k_wrap, PB_GRID(s->npad), s->r, s->orig, s->waveRep, s->d_reps, s->npad);
  PB_LAUNCH_CHECK(s->ctx, "pb_treat_periodic");
  return PB_OK;
}

extern "C" int pb_treat_absolute(pb_sys *s) {
  if (!s) return PB_ERR_ARG;
  if (int rc = need_cell(s)) return rc;
  hipLaunchKernelGGL(k_absolute, PB_GRID(s->npad), s->r, s->orig, s->waveRep, s->d_reps, s->npad);
  PB_LAUNCH_CHECK(s->ctx, "pb_treat_absolute");
  return PB_OK;
}

extern "C" int pb_treat_reflect(pb_sys *s) {
  if (!s) return PB_ERR_ARG;
  if (int rc = need_cell(s)) return rc;
  hipLaunchKernelGGL(k_reflect, PB_GRID(s->npad), s->r, s->v, s->orig, s->waveRep, s->d_reps,
                     s->npad);
  PB_LAUNCH_CHECK(s->ctx, "pb_treat_reflect");
  return PB_OK;
}

extern "C" int pb_vv_step(pb_sys *s, double dt) {
  if (!s) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_vv_step, PB_GRID(s->npad), s->r, s->v, s->F, s->a, s->orig, s->npad, dt);
  PB_LAUNCH_CHECK(s->ctx, "pb_vv_step");
  return PB_OK;
}

extern "C" int pb_euler_step(pb_sys *s, double dt) {
  if (!s) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_euler_step, PB_GRID(s->npad), s->r, s->v, s->a, s->move, s->orig, s->npad, dt);
  PB_LAUNCH_CHECK(s->ctx, "pb_euler_step");
  return PB_OK;
}

extern "C" int pb_lf_vinit(pb_sys *s, double dt) {
  if (!s) return PB_ERR_ARG;
  // cudaParticleLF::calcVinit passes dt/2 (cudaParticleLF.cu:4-7)
  hipLaunchKernelGGL(k_lf_vinit, PB_GRID(s->npad), s->v, s->F, s->orig, s->npad, dt * 0.5);
  PB_LAUNCH_CHECK(s->ctx, "pb_lf_vinit");
  return PB_OK;
}

extern "C" int pb_lf_step(pb_sys *s, double dt) {
  if (!s) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_lf_step, PB_GRID(s->npad), s->r, s->v, s->a, s->move, s->orig, s->npad, dt);
  PB_LAUNCH_CHECK(s->ctx, "pb_lf_step");
  return PB_OK;
}

extern "C" int pb_lf_rollback(pb_sys *s, double dt) {
  if (!s) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_lf_rollback, PB_GRID(s->npad), s->r, s->v, s->a, s->move, s->orig,
                     s->npad, dt);
  PB_LAUNCH_CHECK(s->ctx, "pb_lf_rollback");
  return PB_OK;
}

extern "C" int pb_scale_velocities(pb_sys *s, double sc) {
  if (!s) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_scale_v, PB_GRID(s->npad), s->v, s->orig, s->npad, sc);
  PB_LAUNCH_CHECK(s->ctx, "pb_scale_velocities");
  return PB_OK;
}

extern "C" int pb_mv2(pb_sys *s, double *out) {
  if (!s || !out) return PB_ERR_ARG;
  hipLaunchKernelGGL(k_mv2_partial, PB_GRID(s->npad), s->v, s->orig, s->npad, s->partials);
  PB_LAUNCH_CHECK(s->ctx, "k_mv2_partial");
  if (int rc = pb_reduce_waves(s, 1)) return rc;
  std::vector<double> h(8 * (size_t)s->nrep);
  PB_HIP(s->ctx, hipMemcpyAsync(h.data(), s->d_out, 8 * sizeof(double) * s->nrep,
                                hipMemcpyDeviceToHost, s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  for (uint32_t q = 0; q < s->nrep; ++q) out[q] = h[8 * q];
  return PB_OK;
}

extern "C" int pb_inspect_velocity(pb_sys *s, double vlim, double lim_u, double lim_l,
                                   int *result, double *vratio_max, uint32_t *imax) {
  if (!s || !result) return PB_ERR_ARG;
  // inspectV_F4 takes uint32_t vlim (kernelfuncs.h:116): the float limit is
  // truncated before the division.
  const double vl = (double)(uint32_t)(vlim > 0 ? vlim : 0);
  if (vl <= 0) return pb_fail(s->ctx, PB_ERR_ARG, "pb_inspect_velocity: vlim truncates to 0");
  hipLaunchKernelGGL(k_inspect_partial, PB_GRID(s->npad), s->v, s->orig, s->npad, vl, s->partials);
  hipLaunchKernelGGL(k_inspect_final, dim3(1), dim3(1024), 0, s->ctx->stream, s->partials,
                     s->nwave, s->d_out);
  PB_LAUNCH_CHECK(s->ctx, "pb_inspect_velocity");
  double h[2];
  PB_HIP(s->ctx, hipMemcpyAsync(h, s->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                s->ctx->stream));
  PB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  // float thresholds as the reference's float2 thresh (cudaParticleBase.cu:248-250)
  const double mx = h[0];
  if (mx >= (double)(float)lim_u) *result = 1;
  else if (mx >= (double)(float)lim_l) *result = 0;
  else *result = -1;
  if (vratio_max) *vratio_max = mx;
  if (imax) *imax = (uint32_t)h[1];
  return PB_OK;
}
