// probe.hip -- measurement support for bench.py, built into liboptphot_probe.so
// (not part of the transport library and not declared in include/).
//
// Calibration of the VALU counters and issue rate: 8 independent FMA chains
// per lane, `iters` x 8 v_fma_f32 per wave (the loop is unrolled by 4; its
// control is scalar). With `half` only odd lanes run the loop, so a profiler's
// active-lane figure for this kernel must read 32 of 64.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__global__ __launch_bounds__(256) void valu_probe_kernel(float* out, uint32_t iters,
                                                         uint32_t half) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  float a0 = gid * 1e-9f, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
  float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
  if (!half || (lane & 1u)) {
#pragma unroll 4
    for (uint32_t i = 0; i < iters; ++i) {
      a0 = __builtin_fmaf(a0, 0.999999f, 1e-7f);
      a1 = __builtin_fmaf(a1, 0.999999f, 1e-7f);
      a2 = __builtin_fmaf(a2, 0.999999f, 1e-7f);
      a3 = __builtin_fmaf(a3, 0.999999f, 1e-7f);
      a4 = __builtin_fmaf(a4, 0.999999f, 1e-7f);
      a5 = __builtin_fmaf(a5, 0.999999f, 1e-7f);
      a6 = __builtin_fmaf(a6, 0.999999f, 1e-7f);
      a7 = __builtin_fmaf(a7, 0.999999f, 1e-7f);
    }
  }
  out[gid] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

}  // namespace

// One launch of `blocks` x 256 lanes on a stream of its own, timed with HIP
// events. Returns 0, or -1 for invalid arguments, or the HIP error code.
extern "C" __attribute__((visibility("default"))) int ophip_valu_probe(uint32_t blocks,
                                                                       uint32_t iters,
                                                                       int half_lanes,
                                                                       double* kernel_ms) {
  if (blocks == 0 || iters == 0 || kernel_ms == nullptr) return -1;
  float* out = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipMalloc(&out, static_cast<size_t>(blocks) * 256 * sizeof(float));
  if (e == hipSuccess) e = hipStreamCreate(&stream);
  if (e == hipSuccess) e = hipEventCreate(&ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&ev[1]);
  if (e == hipSuccess) e = hipEventRecord(ev[0], stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, stream, out, iters,
                       half_lanes != 0 ? 1u : 0u);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(ev[1], stream);
  if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
  if (e == hipSuccess) *kernel_ms = ms;
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (stream) (void)hipStreamDestroy(stream);
  if (out) (void)hipFree(out);
  return e == hipSuccess ? 0 : static_cast<int>(e);
}
