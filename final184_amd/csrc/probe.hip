// probe.hip — calibration kernels for the counter-based rooflines of the path's kernels.
//
// The cone march is bound by VALU issue, the voxelize front end by returning atomics and
// dependent loads, the mip chain and the screen passes by HBM. Their ceilings on this chip
// are measured, not assumed: a kernel of independent FP32 fused multiply-add chains gives
// the VALU issue rate a kernel can reach, and kernels of 32-bit atomic adds, each lane on
// its own 64-byte line (no two lanes contend), give the rate of returning and of
// non-returning atomics; a kernel of independent 32-bit integer chains gives the integer
// VALU issue rate. HBM bytes are calibrated with the plain copy kernel k_copy
// (vgi_memcpy_d2d). No reference counterpart: the reference has no measurement code.
#include "vgi_internal.hpp"

namespace vgi {

// Eight independent chains a = a * m + c per lane: each round is eight v_fma_f32 that do
// not wait on each other, so issue, not latency, bounds the kernel. m and c come from the
// arguments (the compiler cannot fold them) and the sums are stored, so no chain is dead.
__global__ __launch_bounds__(256) void k_probe_fma(float* __restrict__ out, uint32_t iters, float m, float c) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float a0 = (float)i, a1 = a0 + 1.f, a2 = a0 + 2.f, a3 = a0 + 3.f, a4 = a0 + 4.f, a5 = a0 + 5.f,
        a6 = a0 + 6.f, a7 = a0 + 7.f;
  for (uint32_t k = 0; k < iters; k++) {
    a0 = fmaf(a0, m, c);
    a1 = fmaf(a1, m, c);
    a2 = fmaf(a2, m, c);
    a3 = fmaf(a3, m, c);
    a4 = fmaf(a4, m, c);
    a5 = fmaf(a5, m, c);
    a6 = fmaf(a6, m, c);
    a7 = fmaf(a7, m, c);
  }
  gst(out + i, ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7)));
}

// The same shape with 32-bit integer operations (exclusive-or with a value the loop
// changes, so that no two rounds cancel and the compiler cannot fold them): the issue rate
// of the integer index and mask work that a gather kernel mixes with its arithmetic.
__global__ __launch_bounds__(256) void k_probe_int(uint32_t* __restrict__ out, uint32_t iters, uint32_t m) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a0 = i, a1 = i + 1, a2 = i + 2, a3 = i + 3, a4 = i + 4, a5 = i + 5, a6 = i + 6, a7 = i + 7;
  for (uint32_t k = 0; k < iters; k++) {
    a0 ^= m;
    a1 ^= m;
    a2 ^= m;
    a3 ^= m;
    a4 ^= m;
    a5 ^= m;
    a6 ^= m;
    a7 ^= m;
    m = m * 3u + 1u;  // keeps the loop from collapsing (even iteration counts would cancel)
  }
  gst(out + i, ((a0 ^ a1) ^ (a2 ^ a3)) ^ ((a4 ^ a5) ^ (a6 ^ a7)));
}

// Each lane adds 1 to the first word of its own 64-byte line `iters` times. RETURN: the
// returned values are summed and stored (the atomic must return its result).
template <bool RETURN>
__global__ __launch_bounds__(256) void k_probe_atomic(uint32_t* __restrict__ lines, uint32_t iters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t* p = lines + (size_t)i * 16u;
  uint32_t sum = 0;
  for (uint32_t k = 0; k < iters; k++) {
    if constexpr (RETURN) sum += gatomic_add(p, 1u);
    else __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if constexpr (RETURN) gst(p + 1, sum);
}

hipError_t launch_probe(uint32_t kind, void* buf, uint32_t blocks, uint32_t iters, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  switch (kind) {
    case VGI_PROBE_FMA:
      hipLaunchKernelGGL(k_probe_fma, dim3(blocks), dim3(256), 0, s, (float*)buf, iters, 0.999f, 1e-3f);
      break;
    case VGI_PROBE_ATOMIC_RETURN:
      hipLaunchKernelGGL(k_probe_atomic<true>, dim3(blocks), dim3(256), 0, s, (uint32_t*)buf, iters);
      break;
    case VGI_PROBE_ATOMIC_NORETURN:
      hipLaunchKernelGGL(k_probe_atomic<false>, dim3(blocks), dim3(256), 0, s, (uint32_t*)buf, iters);
      break;
    case VGI_PROBE_INT:
      hipLaunchKernelGGL(k_probe_int, dim3(blocks), dim3(256), 0, s, (uint32_t*)buf, iters, 0x5A5A5A5Bu);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace vgi
