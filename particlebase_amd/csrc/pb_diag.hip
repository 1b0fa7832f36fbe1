// Issue-rate probe of the vector ALU (no reference counterpart).
//
// The counter-based VALU figures in the profiles divide SQ_INSTS_VALU (wave
// instructions issued) by the launch's CU-cycles.  How many wave64 VALU
// instructions a CU can issue per cycle depends on the operation, so the
// capacity is measured here as the traffic figures are calibrated by a copy
// of known size: every lane runs eight independent chains of FMAs (FP32 or
// FP64) or of integer xor-adds, 64 an iteration with no other vector
// instruction in the loop,
// on enough wavefronts to fill every SIMD.  Elapsed time and the known
// instruction count give wave instructions per second; under rocprofv3 the
// same launch gives SQ_INSTS_VALU per CU-cycle at full issue.
#include "pb_internal.h"
#include <hip/hip_ext.h>
#include <algorithm>

#define PROBE_TB 256u
#define PROBE_CHAINS 8
#define PROBE_UNROLL 8

__device__ __forceinline__ float probe_op(float a, float m, float c) { return fmaf(a, m, c); }
__device__ __forceinline__ double probe_op(double a, double m, double c) { return fma(a, m, c); }
// (a ^ m) + c: an integer chain the compiler cannot fold (v_xor + v_add)
__device__ __forceinline__ uint32_t probe_op(uint32_t a, uint32_t m, uint32_t c) {
  return (a ^ m) + c;
}

// compare + select chain: v_cmp, v_sub, v_add, v_cndmask (the integer
// bookkeeping mix of the pair kernels)
__device__ __forceinline__ uint32_t probe_sel(uint32_t a, uint32_t m, uint32_t c) {
  return a > m ? a - c : a + c;
}

template <typename T, bool SEL>
__global__ void __launch_bounds__(PROBE_TB) k_valu_probe(T *out, uint32_t iters, T m, T c) {
  T a[PROBE_CHAINS];
#pragma unroll
  for (int k = 0; k < PROBE_CHAINS; ++k) a[k] = (T)(threadIdx.x + k) * (T)3;
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < PROBE_UNROLL; ++u)
#pragma unroll
      for (int k = 0; k < PROBE_CHAINS; ++k) {
        if constexpr (SEL) a[k] = probe_sel(a[k], m, c);
        else a[k] = probe_op(a[k], m, c);
      }
  }
  T s = 0;
#pragma unroll
  for (int k = 0; k < PROBE_CHAINS; ++k) s += a[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;   // keeps the chains live
}

extern "C" int pb_diag_valu_probe(pb_ctx *ctx, int kind, uint32_t iters, double *ms,
                                  double *wave_instr) {
  if (!ctx || !ms || !wave_instr || iters == 0 || iters > (1u << 20) || kind < 0 || kind > 3)
    return PB_ERR_ARG;
  PB_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t blocks = 8u * (uint32_t)std::max(ctx->cu_count, 1);   // 8 waves per SIMD
  uint8_t *out = nullptr;
  const size_t bytes = (size_t)blocks * PROBE_TB * (kind == 1 ? 8 : 4);
  if (int rc = pb_alloc(ctx, &out, bytes)) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  if (he == hipSuccess) {
    if (kind == 1)
      hipExtLaunchKernelGGL(k_valu_probe<double, false>, dim3(blocks), dim3(PROBE_TB), 0, ctx->stream, e0,
                            e1, 0u, (double *)out, iters, 0.999999, 1e-7);
    else if (kind == 3)
      hipExtLaunchKernelGGL(k_valu_probe<uint32_t, true>, dim3(blocks), dim3(PROBE_TB), 0,
                            ctx->stream, e0, e1, 0u, (uint32_t *)out, iters, 0x80000000u,
                            0x9E3779B9u);
    else if (kind == 2)
      hipExtLaunchKernelGGL(k_valu_probe<uint32_t, false>, dim3(blocks), dim3(PROBE_TB), 0, ctx->stream,
                            e0, e1, 0u, (uint32_t *)out, iters, 0x9E3779B9u, 0x7F4A7C15u);
    else
      hipExtLaunchKernelGGL(k_valu_probe<float, false>, dim3(blocks), dim3(PROBE_TB), 0, ctx->stream, e0,
                            e1, 0u, (float *)out, iters, 0.9999f, 1e-4f);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
  float t = 0.f;
  if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  pb_free(out);
  if (he != hipSuccess) return pb_hip_check(ctx, he, "pb_diag_valu_probe");
  *ms = t;
  // instructions a chain step compiles to: the integer chain v_xor_b32 +
  // v_add_u32, the select chain v_cmp_lt_u32 + v_cndmask_b32 + v_add_u32
  *wave_instr = (double)blocks * (PROBE_TB / 64) * iters * PROBE_UNROLL * PROBE_CHAINS *
                (kind == 2 ? 2.0 : kind == 3 ? 3.0 : 1.0);
  return PB_OK;
}
