// Achievable-rate probes for the measurement (libpbf_peaks.so, not part of the
// pbf_* interface): a float4 device copy (HBM read + write bandwidth) and a
// chain of independent FP32 fused multiply-adds (vector FP32 rate; built with
// packed FP32 off, as libpbf.so is, so this is the non-packed rate).
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

// Each block copies 1,024 consecutive float4, four per thread (four loads in
// flight, each coalesced over the block).
__global__ __launch_bounds__(256) void k_peak_copy(const float4* __restrict__ in, float4* __restrict__ out,
                                                   size_t n4) {
  const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + 256 * k < n4) v[k] = in[base + 256 * k];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + 256 * k < n4) out[base + 256 * k] = v[k];
}

// Eight independent chains per thread; `iters` rounds of 8 FMAs. The result
// depends on every FMA and is stored, so nothing is removed.
__global__ __launch_bounds__(256) void k_peak_fma(float* __restrict__ out, int iters, float b, float c) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  float a0 = t * 1e-9f, a1 = a0 + 1e-3f, a2 = a0 + 2e-3f, a3 = a0 + 3e-3f;
  float a4 = a0 + 4e-3f, a5 = a0 + 5e-3f, a6 = a0 + 6e-3f, a7 = a0 + 7e-3f;
  for (int k = 0; k < iters; ++k) {
    a0 = fmaf(a0, b, c);
    a1 = fmaf(a1, b, c);
    a2 = fmaf(a2, b, c);
    a3 = fmaf(a3, b, c);
    a4 = fmaf(a4, b, c);
    a5 = fmaf(a5, b, c);
    a6 = fmaf(a6, b, c);
    a7 = fmaf(a7, b, c);
  }
  out[t] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

}  // namespace

extern "C" {

// Copies n4 float4 from in to out on `stream` (device pointers). Returns 0 or
// the HIP error of the launch.
int pbf_peak_copy(const void* in, void* out, long long n4, void* stream) {
  if (n4 <= 0 || !in || !out) return 1;
  const long long blocks = (n4 + 1023) / 1024;
  if (blocks > 0x7fffffffLL) return 1;
  k_peak_copy<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>((const float4*)in, (float4*)out, (size_t)n4);
  return (int)hipGetLastError();
}

// blocks x 256 threads, each 8 * iters FMAs (16 * iters flops); out holds
// blocks * 256 floats.
int pbf_peak_fma(float* out, int blocks, int iters, void* stream) {
  if (blocks <= 0 || iters <= 0 || !out) return 1;
  k_peak_fma<<<blocks, 256, 0, (hipStream_t)stream>>>(out, iters, 0.999999f, 1e-7f);
  return (int)hipGetLastError();
}

}  // extern "C"
