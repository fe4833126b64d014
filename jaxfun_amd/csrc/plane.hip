// Fused two-axis contraction of 2D slabs:
//     Y[b] = A0 (M0 x K0) . X[b] (K0 x K1) . A1^T (K1 x M1)     (all dims <= 128)
// optionally followed by a pointwise epilogue E (M0 x M1):  Y /= E  or  Y *= E.
// The intermediate T = X[b] A1^T (K0 x M1) stays in LDS, so a slab is read
// from HBM once and written once: a 2D TensorProductSpace transform, the two
// eigen-transforms + eigenvalue division of a 2D fast-diagonalisation solve,
// or the last two axes of a 3D transform, in one pass.
// Persistent workgroups of 8 waves walk the slabs: two per CU when the LDS
// buffer (K0 rows rounded up to 32) fits in half the CU's LDS and a wave holds
// at most 2 column blocks.  The work split inside a slab is described at the
// kernel.
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
};

// CBW: 16-column blocks per wave (1, 2, 4 or 8).  Each 16-row block of a
// phase is shared by wpr waves (8 / row blocks, at most 8), which take
// consecutive groups of CBW column blocks; every wave runs CBW MFMA chains per
// k-step unconditionally, so the loop has no divergent branches.  Operands
// are loaded from clamped, always-valid addresses: rows and columns past the
// array edges then hold copies of edge values, which reach only output rows
// and columns that are never stored, and the contraction index is cut off by
// zeros (X times a mask factor for k >= K1; T's zeroed rows >= K0; the bf16
// slices are zero-padded to Kp).  FP64 operands of the next k-step are loaded
// while the current step's MFMAs run.
// One slab: T = Xb A1^T into LDS, then Yb = A0 T (with the epilogue).  Xb
// rows are ldx apart and Yb rows ldy apart, so either may live in LDS (the
// fused solve passes its intermediate that way).  Ends with a barrier.
template <int POL, int CBW>
__device__ __forceinline__ void plane_pass(const double* __restrict__ A0d, const double* __restrict__ A1d,
                                           const unsigned short* __restrict__ A0s,
                                           const unsigned short* __restrict__ A1s,
                                           const typename QT<POL>::TX* __restrict__ Xb, int ldx,
                                           typename QT<POL>::TX* __restrict__ Yb, int ldy,
                                           const typename QT<POL>::TX* __restrict__ E, int epi, int M0, int K0,
                                           int M1, int K1, int wpr, typename QT<POL>::TX* __restrict__ T) {
  typedef typename QT<POL>::TX TX;
  typedef typename QAcc<POL>::V V;
  constexpr int LDT = QT<POL>::LDT;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int K0p = (K0 + 31) & ~31, K1p = (K1 + 31) & ~31;  // slice row strides
  const int ncb1 = (M1 + 15) >> 4;
  const int rb = w / wpr, c0 = (w % wpr) * CBW;
  const bool actA = 16 * rb < K0 && c0 < ncb1, actB = 16 * rb < M0 && c0 < ncb1;
  const int l16 = lane & 15, kl = lane >> 4;
  // ---------------- phase A: T (K0 x M1) = X_b A1^T
  if (actA) {
    V acc[CBW];
#pragma unroll
    for (int t = 0; t < CBW; ++t) acc[t] = V{0, 0, 0, 0};
    const int i = 16 * rb + l16;
    const int ic = i < K0 ? i : K0 - 1;
    if constexpr (POL == 0) {
      double a0, a1, b0[CBW], b1[CBW];
      // a mask factor (not a select) zeroes k >= K1, so that no load is
      // predicated and the loads of a k-step issue back to back
      auto load = [&](int k0, double& a, double(&bv)[CBW]) {
        const int k = k0 + kl, kc = k < K1 ? k : K1 - 1;
        a = Xb[(long long)ic * ldx + kc] * (k < K1 ? 1.0 : 0.0);
#pragma unroll
        for (int t = 0; t < CBW; ++t) {
          const int m = 16 * (c0 + t) + l16, mc = m < M1 ? m : M1 - 1;
          bv[t] = A1d[(long long)mc * K1 + kc];
        }
      };
      load(0, a0, b0);
      for (int k0 = 0; k0 < K1; k0 += 8) {
        load(k0 + 4, a1, b1);
#pragma unroll
        for (int t = 0; t < CBW; ++t) acc[t] = mfma_f64(a0, b0[t], acc[t]);
        load(k0 + 8, a0, b0);
#pragma unroll
        for (int t = 0; t < CBW; ++t) acc[t] = mfma_f64(a1, b1[t], acc[t]);
      }
    } else {
      for (int k0 = 0; k0 < K1; k0 += 32) {
        v8u16 ah, am, al;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = k0 + 8 * kl + e, kc = k < K1 ? k : K1 - 1;
          const float x = (float)Xb[(long long)ic * ldx + kc] * (k < K1 ? 1.f : 0.f);
          unsigned short h, m_, l;
          split3q(x, h, m_, l);
          ah[e] = h;
          am[e] = m_;
          al[e] = l;
        }
#pragma unroll
        for (int t = 0; t < CBW; ++t) {
          const int m = 16 * (c0 + t) + l16, mc = m < M1 ? m : M1 - 1;
          const long long o = (long long)mc * K1p + k0 + 8 * kl;
          const long long sl = (long long)M1 * K1p;
          const v8u16 bh = *(const v8u16*)(A1s + o), bm = *(const v8u16*)(A1s + o + sl),
                      bl = *(const v8u16*)(A1s + o + 2 * sl);
          acc[t] = mfma_bf16(al, bh, acc[t]);
          acc[t] = mfma_bf16(am, bm, acc[t]);
          acc[t] = mfma_bf16(ah, bl, acc[t]);
          acc[t] = mfma_bf16(am, bh, acc[t]);
          acc[t] = mfma_bf16(ah, bm, acc[t]);
          acc[t] = mfma_bf16(ah, bh, acc[t]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < CBW; ++t) {
      const int m = 16 * (c0 + t) + l16;
      if (m >= M1) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ii = 16 * rb + ((POL == 0) ? mfma16_row_f64(lane, r) : mfma16_row_f32(lane, r));
        if (ii < K0) T[ii * LDT + m] = (TX)acc[t][r];
      }
    }
  }
  __syncthreads();
  // ---------------- phase B: Y (M0 x M1) = A0 T
  if (actB) {
    V acc[CBW];
#pragma unroll
    for (int t = 0; t < CBW; ++t) acc[t] = V{0, 0, 0, 0};
    const int mrow = 16 * rb + l16, mrc = mrow < M0 ? mrow : M0 - 1;
    if constexpr (POL == 0) {
      auto lda = [&](int k) { return A0d[(long long)mrc * K0 + (k < K0 ? k : K0 - 1)]; };
      double an = lda(kl);
      for (int k0 = 0; k0 < K0; k0 += 4) {
        const double a = an;
        const int k = k0 + kl;
        an = lda(k + 4);
#pragma unroll
        for (int t = 0; t < CBW; ++t) acc[t] = mfma_f64(a, T[k * LDT + 16 * (c0 + t) + l16], acc[t]);
      }
    } else {
      for (int k0 = 0; k0 < K0; k0 += 32) {
        const long long o = (long long)mrc * K0p + k0 + 8 * kl;
        const long long sl = (long long)M0 * K0p;
        const v8u16 ah = *(const v8u16*)(A0s + o), am = *(const v8u16*)(A0s + o + sl),
                    al = *(const v8u16*)(A0s + o + 2 * sl);
#pragma unroll
        for (int t = 0; t < CBW; ++t) {
          const int j = 16 * (c0 + t) + l16;
          v8u16 bh, bm, bl;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const float x = (float)T[(k0 + 8 * kl + e) * LDT + j];
            unsigned short h, m_, l;
            split3q(x, h, m_, l);
            bh[e] = h;
            bm[e] = m_;
            bl[e] = l;
          }
          acc[t] = mfma_bf16(al, bh, acc[t]);
          acc[t] = mfma_bf16(am, bm, acc[t]);
          acc[t] = mfma_bf16(ah, bl, acc[t]);
          acc[t] = mfma_bf16(am, bh, acc[t]);
          acc[t] = mfma_bf16(ah, bm, acc[t]);
          acc[t] = mfma_bf16(ah, bh, acc[t]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < CBW; ++t) {
      const int j = 16 * (c0 + t) + l16;
      if (j >= M1) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 16 * rb + ((POL == 0) ? mfma16_row_f64(lane, r) : mfma16_row_f32(lane, r));
        if (m >= M0) continue;
        TX v = (TX)acc[t][r];
        if (epi == 1) v = v / E[(long long)m * M1 + j];
        if (epi == 2) v = v * E[(long long)m * M1 + j];
        Yb[(long long)m * ldy + j] = v;
      }
    }
  }
  __syncthreads();
}

template <int POL, int CBW>
__global__ __launch_bounds__(QNT, CBW <= 2 ? 4 : 2) void plane_kernel(
    const double* __restrict__ A0d, const double* __restrict__ A1d, const unsigned short* __restrict__ A0s,
    const unsigned short* __restrict__ A1s, const typename QT<POL>::TX* __restrict__ X,
    typename QT<POL>::TX* __restrict__ Y, const typename QT<POL>::TX* __restrict__ E, int epi, long long B,
    int M0, int K0, int M1, int K1, int wpr) {
  typedef typename QT<POL>::TX TX;
  constexpr int LDT = QT<POL>::LDT;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  TX* T = reinterpret_cast<TX*>(smem);  // T[i][m], i < K0p (rows >= K0 zero), m < LDT
  const int K0p = (K0 + 31) & ~31;
  // rows >= K0 and columns >= M1 of T are zeroed once and never written, so
  // phase B reads T without masks
  for (int e = threadIdx.x; e < K0p * LDT; e += QNT) T[e] = (TX)0;
  __syncthreads();
  for (long long b = blockIdx.x; b < B; b += gridDim.x)
    plane_pass<POL, CBW>(A0d, A1d, A0s, A1s, X + b * (long long)K0 * K1, K1, Y + b * (long long)M0 * M1, M1, E,
                         epi, M0, K0, M1, K1, wpr, T);
}

// Two-axis fast-diagonalisation solve of n x n slabs in one launch:
//     Y[b] = V0 ((W0 X[b] W1^T) / E) V1^T
// (the forward eigen-transform with the eigenvalue division, then the back
// transform); the intermediate stays in a second LDS tile whose padding is
// zero, so each slab is read from and written to HBM once.
template <int POL, int CBW>
__global__ __launch_bounds__(QNT, 2) void plane_fd_kernel(
    const double* __restrict__ W0d, const double* __restrict__ W1d, const double* __restrict__ V0d,
    const double* __restrict__ V1d, const unsigned short* __restrict__ W0s, const unsigned short* __restrict__ W1s,
    const unsigned short* __restrict__ V0s, const unsigned short* __restrict__ V1s,
    const typename QT<POL>::TX* __restrict__ X, typename QT<POL>::TX* __restrict__ Y,
    const typename QT<POL>::TX* __restrict__ E, long long B, int n, int wpr) {
  typedef typename QT<POL>::TX TX;
  constexpr int LDT = QT<POL>::LDT;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int np = (n + 31) & ~31;
  TX* T = reinterpret_cast<TX*>(smem);  // phase tile, np x LDT
  TX* Z = T + np * LDT;                 // intermediate, np x LDT
  for (int e = threadIdx.x; e < 2 * np * LDT; e += QNT) T[e] = (TX)0;
  __syncthreads();
  for (long long b = blockIdx.x; b < B; b += gridDim.x) {
    plane_pass<POL, CBW>(W0d, W1d, W0s, W1s, X + b * (long long)n * n, n, Z, LDT, E, 1, n, n, n, n, wpr, T);
    plane_pass<POL, CBW>(V0d, V1d, V0s, V1s, Z, LDT, Y + b * (long long)n * n, n, nullptr, 0, n, n, n, n, wpr,
                         T);
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
  if (ncu <= 0) ncu = 256;
  // waves per 16-row block and column blocks per wave
  const int nrb = ((K0 > M0 ? K0 : M0) + 15) >> 4, ncb = (M1 + 15) >> 4;
  const int wpr = nrb == 1 ? 8 : nrb == 2 ? 4 : nrb <= 4 ? 2 : 1;
  const int nb = (ncb + wpr - 1) / wpr, cbw = nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : 8;
  // T holds K0p rows; two workgroups share a CU when the LDS and registers allow
  const int K0p = (K0 + 31) & ~31;
  const size_t shm = (size_t)K0p * (prec == 0 ? QT<0>::LDT * 8 : QT<3>::LDT * 4);
  long long grid = (long long)ncu * ((shm <= 80 * 1024 && cbw <= 2) ? 2 : 1);
  if (grid > B) grid = B;
  typedef void (*KF64)(const double*, const double*, const unsigned short*, const unsigned short*, const double*,
                       double*, const double*, int, long long, int, int, int, int, int);
  typedef void (*KF32)(const double*, const double*, const unsigned short*, const unsigned short*, const float*,
                       float*, const float*, int, long long, int, int, int, int, int);
  if (prec == 0) {
    if (!A0d || !A1d) return set_error("plane_contract: FP64 operators missing"), JFX_EARG;
    KF64 kern = cbw == 1 ? plane_kernel<0, 1> : cbw == 2 ? plane_kernel<0, 2> : cbw == 4 ? plane_kernel<0, 4>
                                                                                         : plane_kernel<0, 8>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(QNT), shm, stream, A0d, A1d, A0s, A1s,
                       (const double*)X, (double*)Y, (const double*)E, epi, B, M0, K0, M1, K1, wpr);
  } else if (prec == 1) {
    if (!A0s || !A1s) return set_error("plane_contract: bf16 slices missing"), JFX_EARG;
    KF32 kern = cbw == 1 ? plane_kernel<3, 1> : cbw == 2 ? plane_kernel<3, 2> : cbw == 4 ? plane_kernel<3, 4>
                                                                                         : plane_kernel<3, 8>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(QNT), shm, stream, A0d, A1d, A0s, A1s,
                       (const float*)X, (float*)Y, (const float*)E, epi, B, M0, K0, M1, K1, wpr);
  } else {
    return set_error("plane_contract: bad prec"), JFX_EARG;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

// Y[b] = V0 ((W0 X[b] W1^T) / E) V1^T for n x n slabs (X, Y: B x n x n), one
// launch.  prec 0: FP64 operators, n <= 64; prec 1: bf16 3-term slices
// ([3][n][np]) on FP32 data, n <= 128.
int jfx_plane_fd(int prec, const double* W0d, const double* W1d, const double* V0d, const double* V1d,
                 const unsigned short* W0s, const unsigned short* W1s, const unsigned short* V0s,
                 const unsigned short* V1s, const void* X, void* Y, const void* E, long long B, int n,
                 hipStream_t stream) {
  if (!X || !Y || !E || B < 0 || n <= 0) return set_error("plane_fd: bad arguments"), JFX_EARG;
  if (n > (prec == 0 ? 64 : QMAX)) return set_error("plane_fd: n = %d too large for the fused solve", n), JFX_ESHAPE;
  if (B == 0) return JFX_OK;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (ncu <= 0) ncu = 256;
  const int nrb = (n + 15) >> 4;
  const int wpr = nrb == 1 ? 8 : nrb == 2 ? 4 : nrb <= 4 ? 2 : 1;
  const int nb = (nrb + wpr - 1) / wpr, cbw = nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : 8;
  const int np = (n + 31) & ~31;
  const size_t shm = (size_t)2 * np * (prec == 0 ? QT<0>::LDT * 8 : QT<3>::LDT * 4);
  const long long grid = (long long)ncu < B ? ncu : B;
  typedef void (*KF64)(const double*, const double*, const double*, const double*, const unsigned short*,
                       const unsigned short*, const unsigned short*, const unsigned short*, const double*, double*,
                       const double*, long long, int, int);
  typedef void (*KF32)(const double*, const double*, const double*, const double*, const unsigned short*,
                       const unsigned short*, const unsigned short*, const unsigned short*, const float*, float*,
                       const float*, long long, int, int);
  hipError_t e;
  if (prec == 0) {
    if (!W0d || !W1d || !V0d || !V1d) return set_error("plane_fd: FP64 operators missing"), JFX_EARG;
    KF64 kern = cbw == 1 ? plane_fd_kernel<0, 1> : cbw == 2 ? plane_fd_kernel<0, 2> : plane_fd_kernel<0, 4>;
    e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane_fd: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(QNT), shm, stream, W0d, W1d, V0d, V1d, W0s, W1s, V0s, V1s,
                       (const double*)X, (double*)Y, (const double*)E, B, n, wpr);
  } else if (prec == 1) {
    if (!W0s || !W1s || !V0s || !V1s) return set_error("plane_fd: bf16 slices missing"), JFX_EARG;
    KF32 kern = cbw == 1 ? plane_fd_kernel<3, 1>
                : cbw == 2 ? plane_fd_kernel<3, 2>
                : cbw == 4 ? plane_fd_kernel<3, 4>
                           : plane_fd_kernel<3, 8>;
    e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane_fd: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(QNT), shm, stream, W0d, W1d, V0d, V1d, W0s, W1s, V0s, V1s,
                       (const float*)X, (float*)Y, (const float*)E, B, n, wpr);
  } else {
    return set_error("plane_fd: bad prec"), JFX_EARG;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // extern "C"
