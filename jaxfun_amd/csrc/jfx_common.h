// Shared helpers for the jfx HIP kernels (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short v8u16 __attribute__((ext_vector_type(8)));

// Status codes returned by every C-ABI entry point.
enum {
  JFX_OK = 0,
  JFX_EHIP = -1,       // a HIP runtime call failed
  JFX_ESHAPE = -2,     // unsupported or inconsistent sizes
  JFX_ESINGULAR = -3,  // zero / non-finite pivot in a banded factorisation
  JFX_EARG = -4,       // invalid argument (null pointer, bad enum)
};

namespace jfx {

void set_error(const char* fmt, ...);

#define JFX_CHECK_HIP(expr)                                              \
  do {                                                                   \
    hipError_t _e = (expr);                                              \
    if (_e != hipSuccess) {                                              \
      ::jfx::set_error("%s failed: %s", #expr, hipGetErrorString(_e));   \
      return JFX_EHIP;                                                   \
    }                                                                    \
  } while (0)

#define JFX_LAUNCH_CHECK()                                               \
  do {                                                                   \
    hipError_t _e = hipGetLastError();                                   \
    if (_e != hipSuccess) {                                              \
      ::jfx::set_error("kernel launch failed: %s", hipGetErrorString(_e)); \
      return JFX_EHIP;                                                   \
    }                                                                    \
  } while (0)

// float -> bf16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) {
  return __uint_as_float(((unsigned int)b) << 16);
}

// Split 8 fp32 values into 3 bf16 terms x = h + m + l (each round to nearest
// even, the remainders exact), packed for an MFMA operand.  Uses the packed
// hardware conversion (two values per instruction); results equal f2bf-based
// splitting bit for bit on finite inputs.
__device__ __forceinline__ void split3_pk8(const float (&x)[8], v8u16& H, v8u16& M, v8u16& L) {
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  u4 hw, mw, lw;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float x0 = x[2 * e], x1 = x[2 * e + 1];
    bf2 h = __builtin_convertvector((f2){x0, x1}, bf2);
    const unsigned hu = __builtin_bit_cast(unsigned, h);
    const float r0 = x0 - __uint_as_float(hu << 16), r1 = x1 - __uint_as_float(hu & 0xFFFF0000u);
    bf2 m = __builtin_convertvector((f2){r0, r1}, bf2);
    const unsigned mu = __builtin_bit_cast(unsigned, m);
    const float s0 = r0 - __uint_as_float(mu << 16), s1 = r1 - __uint_as_float(mu & 0xFFFF0000u);
    bf2 l = __builtin_convertvector((f2){s0, s1}, bf2);
    hw[e] = hu;
    mw[e] = mu;
    lw[e] = __builtin_bit_cast(unsigned, l);
  }
  H = __builtin_bit_cast(v8u16, hw);
  M = __builtin_bit_cast(v8u16, mw);
  L = __builtin_bit_cast(v8u16, lw);
}

// Split 8 fp32 values into two fp16 terms x = h + l + r (packed hardware
// conversions, round to nearest even; x - h is exact in fp32):
// |r| <= max(2^-22 |x|, 2^-25), the second bound where x - h falls below
// fp16's normal range.  The caller scales x so that |x| < 2^15.
__device__ __forceinline__ void split2_f16_pk8(const float (&x)[8], v8u16& H, v8u16& L) {
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  u4 hw, lw;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float x0 = x[2 * e], x1 = x[2 * e + 1];
    const h2 h = __builtin_convertvector((f2){x0, x1}, h2);
    const f2 hb = __builtin_convertvector(h, f2);
    const h2 l = __builtin_convertvector((f2){x0 - hb[0], x1 - hb[1]}, h2);
    hw[e] = __builtin_bit_cast(unsigned, h);
    lw[e] = __builtin_bit_cast(unsigned, l);
  }
  H = __builtin_bit_cast(v8u16, hw);
  L = __builtin_bit_cast(v8u16, lw);
}

// Map a linear workgroup id so that consecutive logical tiles land on the
// same XCD (workgroups are dispatched round-robin over the 8 XCDs).
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nblk) {
  const unsigned nx = 8;
  if (nblk % nx != 0) return bid;
  return (bid % nx) * (nblk / nx) + bid / nx;
}

// Output row of accumulator element r held by `lane` for a 16x16 MFMA block
// (column = lane % 16).  Measured on gfx950: f32/bf16/f16 16x16 outputs hold
// rows 4*(lane/16) + r; v_mfma_f64_16x16x4f64 holds rows lane/16 + 4*r.
__device__ __forceinline__ int mfma16_row_f32(int lane, int r) { return 4 * (lane >> 4) + r; }
__device__ __forceinline__ int mfma16_row_f64(int lane, int r) { return (lane >> 4) + 4 * r; }

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma_f32(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma_f16(v8u16 a, v8u16 b, v4f c) {
  typedef _Float16 v8h __attribute__((ext_vector_type(8)));
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0, 0,
                                                0);
}
__device__ __forceinline__ v4f mfma_bf16(v8u16 a, v8u16 b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                 __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

}  // namespace jfx
