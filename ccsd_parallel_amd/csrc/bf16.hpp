// BF16 hi/lo split shared by the conversion kernels, the unpack of the ring
// amplitudes and the GEMM epilogues: x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi), both rounded to nearest even.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ccsd {

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ void bf16_split(double x, uint16_t& hi, uint16_t& lo) {
  hi = bf16_rne((float)x);
  lo = bf16_rne((float)(x - (double)bf16_to_f(hi)));
}

}  // namespace ccsd
