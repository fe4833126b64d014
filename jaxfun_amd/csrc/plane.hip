// Fused two-axis contraction of 2D slabs:
//     Y[b] = A0 (M0 x K0) . X[b] (K0 x K1) . A1^T (K1 x M1)     (all dims <= 128)
// optionally followed by a pointwise epilogue E (M0 x M1):  Y /= E  or  Y *= E.
// The intermediate T = X[b] A1^T (K0 x M1) stays in LDS, so a slab is read
// from HBM once and written once: a 2D TensorProductSpace transform, the two
// eigen-transforms + eigenvalue division of a 2D fast-diagonalisation solve,
// or the last two axes of a 3D transform, in one pass.
// One persistent workgroup of 8 waves per CU walks the slabs; wave w owns
// output row block w (16 rows) of each phase and all of its column blocks.
//   POL 0: FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 I/O
//   POL 3: bf16 3-term float split (6 products), FP32 accumulation, FP32 I/O
#include <type_traits>

#include "jfx_common.h"

namespace jfx {

constexpr int QMAX = 128, QNT = 512;

template <int POL>
struct QT;
template <>
struct QT<0> {
  typedef double TX;
  static constexpr int LDT = 129;
};
template <>
struct QT<3> {
  typedef float TX;
  static constexpr int LDT = 130;
};

__device__ __forceinline__ void split3q(float x, unsigned short& h, unsigned short& m, unsigned short& l) {
  h = f2bf(x);
  float r = x - bf2f(h);
  m = f2bf(r);
  r -= bf2f(m);
  l = f2bf(r);
}

// C[r][c] = sum_k L[r][k] R[k][c] for the wave's 16-row block, up to 8 col blocks.
// L rows come from global (row-major, ldl) or slices; R from a callback.
template <int POL>
struct QAcc {
  typedef typename std::conditional<POL == 0, v4d, v4f>::type V;
  V v[8];
};

template <int POL>
__global__ __launch_bounds__(QNT) void plane_kernel(const double* __restrict__ A0d, const double* __restrict__ A1d,
                                                    const unsigned short* __restrict__ A0s,
                                                    const unsigned short* __restrict__ A1s,
                                                    const typename QT<POL>::TX* __restrict__ X,
                                                    typename QT<POL>::TX* __restrict__ Y,
                                                    const typename QT<POL>::TX* __restrict__ E, int epi,
                                                    long long B, int M0, int K0, int M1, int K1) {
  typedef typename QT<POL>::TX TX;
  constexpr int LDT = QT<POL>::LDT;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  TX* T = reinterpret_cast<TX*>(smem);  // T[i][m], i < K0, m < M1
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int K0p = (K0 + 31) & ~31, K1p = (K1 + 31) & ~31;  // slice row strides
  const int ncb1 = (M1 + 15) >> 4;

  for (long long b = blockIdx.x; b < B; b += gridDim.x) {
    const TX* Xb = X + b * (long long)K0 * K1;
    TX* Yb = Y + b * (long long)M0 * M1;
    // ---------------- phase A: T (K0 x M1) = X_b A1^T ; wave w -> rows 16w..16w+15 of T
    if (16 * w < K0) {
      typename QAcc<POL>::V acc[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = typename QAcc<POL>::V{0, 0, 0, 0};
      const int i = 16 * w + (lane & 15);
      if constexpr (POL == 0) {
        for (int k0 = 0; k0 < K1; k0 += 4) {
          const int k = k0 + (lane >> 4);
          const double a = (i < K0 && k < K1) ? Xb[(long long)i * K1 + k] : 0.0;
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            if (c < ncb1) {
              const int m = 16 * c + (lane & 15);
              const double bv = (m < M1 && k < K1) ? A1d[(long long)m * K1 + k] : 0.0;
              acc[c] = mfma_f64(a, bv, acc[c]);
            }
          }
        }
      } else {
        for (int k0 = 0; k0 < K1; k0 += 32) {
          v8u16 ah, am, al;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k = k0 + 8 * (lane >> 4) + e;
            float x = (i < K0 && k < K1) ? (float)Xb[(long long)i * K1 + k] : 0.f;
            unsigned short h, m_, l;
            split3q(x, h, m_, l);
            ah[e] = h;
            am[e] = m_;
            al[e] = l;
          }
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            if (c < ncb1) {
              const int m = 16 * c + (lane & 15);
              v8u16 bh = {0, 0, 0, 0, 0, 0, 0, 0}, bm = bh, bl = bh;
              if (m < M1) {
                const long long o = (long long)m * K1p + k0 + 8 * (lane >> 4);
                const long long sl = (long long)M1 * K1p;
                bh = *(const v8u16*)(A1s + o);
                bm = *(const v8u16*)(A1s + o + sl);
                bl = *(const v8u16*)(A1s + o + 2 * sl);
              }
              acc[c] = mfma_bf16(al, bh, acc[c]);
              acc[c] = mfma_bf16(am, bm, acc[c]);
              acc[c] = mfma_bf16(ah, bl, acc[c]);
              acc[c] = mfma_bf16(am, bh, acc[c]);
              acc[c] = mfma_bf16(ah, bm, acc[c]);
              acc[c] = mfma_bf16(ah, bh, acc[c]);
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (c >= ncb1) continue;
        const int m = 16 * c + (lane & 15);
        if (m >= M1) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ii = 16 * w + ((POL == 0) ? mfma16_row_f64(lane, r) : mfma16_row_f32(lane, r));
          if (ii < K0) T[ii * LDT + m] = (TX)acc[c][r];
        }
      }
    }
    __syncthreads();
    // ---------------- phase B: Y (M0 x M1) = A0 T ; wave w -> rows 16w..16w+15 of Y
    if (16 * w < M0) {
      typename QAcc<POL>::V acc[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = typename QAcc<POL>::V{0, 0, 0, 0};
      const int mrow = 16 * w + (lane & 15);
      if constexpr (POL == 0) {
        for (int k0 = 0; k0 < K0; k0 += 4) {
          const int k = k0 + (lane >> 4);
          const double a = (mrow < M0 && k < K0) ? A0d[(long long)mrow * K0 + k] : 0.0;
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            if (c < ncb1) {
              const int j = 16 * c + (lane & 15);
              const double bv = (j < M1 && k < K0) ? T[k * LDT + j] : 0.0;
              acc[c] = mfma_f64(a, bv, acc[c]);
            }
          }
        }
      } else {
        for (int k0 = 0; k0 < K0; k0 += 32) {
          v8u16 ah = {0, 0, 0, 0, 0, 0, 0, 0}, am = ah, al = ah;
          if (mrow < M0) {
            const long long o = (long long)mrow * K0p + k0 + 8 * (lane >> 4);
            const long long sl = (long long)M0 * K0p;
            ah = *(const v8u16*)(A0s + o);
            am = *(const v8u16*)(A0s + o + sl);
            al = *(const v8u16*)(A0s + o + 2 * sl);
          }
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            if (c < ncb1) {
              const int j = 16 * c + (lane & 15);
              v8u16 bh, bm, bl;
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                const int k = k0 + 8 * (lane >> 4) + e;
                float x = (j < M1 && k < K0) ? (float)T[k * LDT + j] : 0.f;
                unsigned short h, m_, l;
                split3q(x, h, m_, l);
                bh[e] = h;
                bm[e] = m_;
                bl[e] = l;
              }
              acc[c] = mfma_bf16(al, bh, acc[c]);
              acc[c] = mfma_bf16(am, bm, acc[c]);
              acc[c] = mfma_bf16(ah, bl, acc[c]);
              acc[c] = mfma_bf16(am, bh, acc[c]);
              acc[c] = mfma_bf16(ah, bm, acc[c]);
              acc[c] = mfma_bf16(ah, bh, acc[c]);
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (c >= ncb1) continue;
        const int j = 16 * c + (lane & 15);
        if (j >= M1) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = 16 * w + ((POL == 0) ? mfma16_row_f64(lane, r) : mfma16_row_f32(lane, r));
          if (m >= M0) continue;
          TX v = (TX)acc[c][r];
          if (epi == 1) v = v / E[(long long)m * M1 + j];
          if (epi == 2) v = v * E[(long long)m * M1 + j];
          Yb[(long long)m * M1 + j] = v;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace jfx

using namespace jfx;

extern "C" {

// Y[b] = A0 X[b] A1^T for b < B (X: B x K0 x K1, Y: B x M0 x M1), dims <= 128.
// prec 0: FP64 (A0d/A1d FP64 row-major); prec 1: 3-term bf16 split on FP32
// data (A0s/A1s: [3][M][Kp] slices, Kp = 32*ceil(K/32)).  epi: 0 none,
// 1 divide by E (M0 x M1), 2 multiply by E.
int jfx_plane_contract(int prec, const double* A0d, const double* A1d, const unsigned short* A0s,
                       const unsigned short* A1s, const void* X, void* Y, const void* E, int epi, long long B,
                       int M0, int K0, int M1, int K1, hipStream_t stream) {
  if (!X || !Y || B < 0 || M0 <= 0 || K0 <= 0 || M1 <= 0 || K1 <= 0 || epi < 0 || epi > 2 ||
      (epi && !E))
    return set_error("plane_contract: bad arguments"), JFX_EARG;
  if (M0 > QMAX || K0 > QMAX || M1 > QMAX || K1 > QMAX)
    return set_error("plane_contract: dims must be <= %d", QMAX), JFX_ESHAPE;
  if (B == 0) return JFX_OK;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  long long grid = ncu > 0 ? ncu : 256;
  if (prec == 1) grid *= 2;
  if (grid > B) grid = B;
  if (prec == 0) {
    if (!A0d || !A1d) return set_error("plane_contract: FP64 operators missing"), JFX_EARG;
    const size_t shm = (size_t)QMAX * QT<0>::LDT * 8;
    hipError_t e = hipFuncSetAttribute((const void*)plane_kernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(plane_kernel<0>, dim3((unsigned)grid), dim3(QNT), shm, stream, A0d, A1d, A0s, A1s,
                       (const double*)X, (double*)Y, (const double*)E, epi, B, M0, K0, M1, K1);
  } else if (prec == 1) {
    if (!A0s || !A1s) return set_error("plane_contract: bf16 slices missing"), JFX_EARG;
    const size_t shm = (size_t)QMAX * QT<3>::LDT * 4;
    hipError_t e = hipFuncSetAttribute((const void*)plane_kernel<3>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(plane_kernel<3>, dim3((unsigned)grid), dim3(QNT), shm, stream, A0d, A1d, A0s, A1s,
                       (const float*)X, (float*)Y, (const float*)E, epi, B, M0, K0, M1, K1);
  } else {
    return set_error("plane_contract: bad prec"), JFX_EARG;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // extern "C"
