// Shared helpers for the somax HIP library (gfx950 / MI355X only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace somax {

// 8 x bf16 MFMA operand and the accumulator vectors of the gfx950 MFMA builtins.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Thread-local last error text for somaxhip_last_error().
void set_error(const std::string& msg);

// Round-to-nearest-even float -> bf16 bits (finite inputs; NaN kept quiet).
__host__ __device__ inline uint16_t f32_to_bf16_bits(float x) {
  union {
    float f;
    uint32_t u;
  } v;
  v.f = x;
  if ((v.u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((v.u >> 16) | ((v.u & 0xffffu) ? 0x40u : 0u));
  uint32_t r = v.u + 0x7fffu + ((v.u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

// Device form: the hardware round-to-nearest-even conversion
// (v_cvt_pk_bf16_f32 on gfx950), branch free; equal to f32_to_bf16_bits
// for every finite input.
__device__ inline uint16_t bf16_rne(float x) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x)); }

__host__ __device__ inline float bf16_bits_to_f32(uint16_t b) {
  union {
    float f;
    uint32_t u;
  } v;
  v.u = ((uint32_t)b) << 16;
  return v.f;
}

}  // namespace somax

#define SOMAX_HIP_CHECK(expr)                                                        \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      somax::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));           \
      return SOMAXHIP_ERR_HIP;                                                       \
    }                                                                                \
  } while (0)
