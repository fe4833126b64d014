// Persistent per-axis contraction for operators with M, K <= 128 (every
// 1D operator of the benchmark configurations except padded backward
// transforms): Y[p, m, q] = sum_k A[m, k] X[p, k, q].
//
// One workgroup per CU stays resident and walks 64-column tiles
// (tile += gridDim.x).  The operator (or its bf16 slices) is loaded into LDS
// once per workgroup; X streams through a single LDS stage with the next
// chunk prefetched into registers while the current one is multiplied.
// 8 waves as 4 (m) x 2 (n), each owning a 32 x 32 block of 16x16 MFMA tiles.
//   POL 0: FP64 (v_mfma_f64_16x16x4f64)   POL 1: FP32 (v_mfma_f32_16x16x4f32)
//   POL 2/3: bf16 float split, 2/3 slices (3/6 products), FP32 accumulation
//   POL 4: fixed-point 3-digit slices on FP64 data (exact accumulation)
#include <type_traits>

#include "jfx_common.h"

namespace jfx {

constexpr int PM = 128, PK = 128, PTN = 64, PNT = 512;

template <int POL>
struct PTraits;
template <>
struct PTraits<0> {
  typedef double TX;
  static constexpr int S = 0, KT = 16;
};
template <>
struct PTraits<1> {
  typedef float TX;
  static constexpr int S = 0, KT = 16;
};
template <>
struct PTraits<2> {
  typedef float TX;
  static constexpr int S = 2, KT = 32;
};
template <>
struct PTraits<3> {
  typedef float TX;
  static constexpr int S = 3, KT = 32;
};
template <>
struct PTraits<4> {
  typedef double TX;
  static constexpr int S = 3, KT = 32;
};

struct PMap {
  long long P, Q, ncols, ntiles;
  int M, K;
};

// LDS layout sizes (bytes)
template <int POL>
struct PLds {
  typedef typename PTraits<POL>::TX TX;
  static constexpr int S = PTraits<POL>::S, KT = PTraits<POL>::KT;
  static constexpr int ALD = (S == 0) ? (PM + 1) : (PK + 8);
  static constexpr size_t A = (S == 0) ? (size_t)PK * ALD * sizeof(TX) : (size_t)S * PM * ALD * 2;
  static constexpr int BLD = (S == 0) ? (PTN + 1) : (KT + 8);
  static constexpr size_t B = (S == 0) ? (size_t)KT * BLD * sizeof(TX) : (size_t)S * PTN * BLD * 2;
  static constexpr size_t TOTAL = A + B;
};

template <int POL>
__global__ __launch_bounds__(PNT) void contract_persist_kernel(const void* __restrict__ Aptr,
                                                               const double* __restrict__ Arow,
                                                               const typename PTraits<POL>::TX* __restrict__ X,
                                                               typename PTraits<POL>::TX* __restrict__ Y,
                                                               PMap pm) {
  typedef typename PTraits<POL>::TX TX;
  constexpr int S = PTraits<POL>::S, KT = PTraits<POL>::KT;
  constexpr int ALD = PLds<POL>::ALD, BLD = PLds<POL>::BLD;
  constexpr int PER = KT * PTN / PNT;  // X elements per thread per stage
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  unsigned char* Abuf = smem;
  unsigned char* Bbuf = smem + PLds<POL>::A;
  __shared__ long long xb[PTN], yb[PTN];
  __shared__ double colscale[PTN];
  __shared__ int colexp[PTN];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w & 3) * 32, wn = (w >> 2) * 32;
  const int M = pm.M, K = pm.K;
  const long long Q = pm.Q;

  // ---- resident operator
  if constexpr (S == 0) {
    TX* As = reinterpret_cast<TX*>(Abuf);  // As[k][m]
    const TX* A = reinterpret_cast<const TX*>(Aptr);
    for (int e = tid; e < PK * PM; e += PNT) {
      int m = e / PK, k = e - m * PK;
      As[k * ALD + m] = (m < M && k < K) ? A[(long long)m * K + k] : TX(0);
    }
  } else {
    unsigned short* As = reinterpret_cast<unsigned short*>(Abuf);  // As[s][m][k]
    const unsigned short* Asl = reinterpret_cast<const unsigned short*>(Aptr);
    const int Kp = (K + 31) / 32 * 32;
    for (int ch = tid; ch < S * PM * (PK / 8); ch += PNT) {
      int sl = ch / (PM * (PK / 8)), rem = ch - sl * PM * (PK / 8), m = rem / (PK / 8), part = rem - m * (PK / 8);
      v8u16 v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (m < M && part * 8 < Kp) v = *(const v8u16*)(Asl + ((long long)sl * M + m) * Kp + part * 8);
      *(v8u16*)&As[(sl * PM + m) * ALD + part * 8] = v;
    }
  }

  typedef typename std::conditional<POL == 0, v4d, v4f>::type vacc;
  constexpr int NL = (POL == 4) ? S : 1;

  for (long long tile = blockIdx.x; tile < pm.ntiles; tile += gridDim.x) {
    const long long c0 = tile * PTN;
    __syncthreads();  // previous tile's readers of xb/yb/B are done (and A is loaded)
    if (tid < PTN) {
      long long c = c0 + tid;
      if (c >= pm.ncols) c = pm.ncols - 1;
      long long p = c / Q, q = c - p * Q;
      xb[tid] = p * K * Q + q;
      yb[tid] = p * (long long)M * Q + q;
    }
    __syncthreads();
    if constexpr (POL == 4) {
      const int nn = tid >> 3, part = tid & 7;
      double mx = 0.0;
      if (c0 + nn < pm.ncols)
        for (int k = part; k < K; k += 8) mx = fmax(mx, fabs((double)X[xb[nn] + k * Q]));
      mx = fmax(mx, __shfl_xor(mx, 1));
      mx = fmax(mx, __shfl_xor(mx, 2));
      mx = fmax(mx, __shfl_xor(mx, 4));
      if (part == 0) {
        int e = 0;
        if (mx > 0.0) frexp(mx, &e);
        colexp[nn] = e + 1;
        colscale[nn] = ldexp(1.0, e + 1);
      }
      __syncthreads();
    }

    vacc acc[NL][2][2];
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[l][i][j] = vacc{0, 0, 0, 0};

    // stage loader: element e -> (kk, nn); for the last axis (Q == 1) k is
    // contiguous in memory, otherwise n is.
    auto coords = [&](int e, int& kk, int& nn) {
      if (Q == 1) {
        nn = e / KT;
        kk = e - nn * KT;
      } else {
        kk = e / PTN;
        nn = e - kk * PTN;
      }
    };
    TX pre[PER];
    auto prefetch = [&](int k0) {
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        int kk, nn;
        coords(tid + u * PNT, kk, nn);
        const int k = k0 + kk;
        pre[u] = (c0 + nn < pm.ncols && k < K) ? X[xb[nn] + k * Q] : TX(0);
      }
    };
    prefetch(0);
    for (int k0 = 0; k0 < K; k0 += KT) {
      // registers -> LDS stage
      if constexpr (S == 0) {
        TX* Bs = reinterpret_cast<TX*>(Bbuf);  // Bs[kk][nn]
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          int kk, nn;
          coords(tid + u * PNT, kk, nn);
          Bs[kk * BLD + nn] = pre[u];
        }
      } else {
        unsigned short* Bs = reinterpret_cast<unsigned short*>(Bbuf);  // Bs[s][nn][kk]
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          int kk, nn;
          coords(tid + u * PNT, kk, nn);
          if constexpr (POL == 4) {
            double r = ldexp((double)pre[u], 8 - colexp[nn]);
#pragma unroll
            for (int s = 0; s < S; ++s) {
              double d = rint(r);
              Bs[(s * PTN + nn) * BLD + kk] = f2bf((float)d);
              r = (r - d) * 256.0;
            }
          } else {
            float r = (float)pre[u];
#pragma unroll
            for (int s = 0; s < S; ++s) {
              unsigned short b = f2bf(r);
              Bs[(s * PTN + nn) * BLD + kk] = b;
              r -= bf2f(b);
            }
          }
        }
      }
      __syncthreads();
      if (k0 + KT < K) prefetch(k0 + KT);  // overlaps the MFMAs below
      if constexpr (S == 0) {
        const TX* As = reinterpret_cast<const TX*>(Abuf);
        const TX* Bs = reinterpret_cast<const TX*>(Bbuf);
#pragma unroll
        for (int ks = 0; ks < KT; ks += 4) {
          const int kl = ks + (lane >> 4);
          TX a[2], b[2];
#pragma unroll
          for (int i = 0; i < 2; ++i) a[i] = As[(k0 + kl) * ALD + wm + 16 * i + (lane & 15)];
#pragma unroll
          for (int j = 0; j < 2; ++j) b[j] = Bs[kl * BLD + wn + 16 * j + (lane & 15)];
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              if constexpr (POL == 0)
                acc[0][i][j] = mfma_f64(a[i], b[j], acc[0][i][j]);
              else
                acc[0][i][j] = mfma_f32(a[i], b[j], acc[0][i][j]);
            }
        }
      } else {
        const unsigned short* As = reinterpret_cast<const unsigned short*>(Abuf);
        const unsigned short* Bs = reinterpret_cast<const unsigned short*>(Bbuf);
        v8u16 a[S][2], b[S][2];
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
          for (int i = 0; i < 2; ++i)
            a[s][i] = *(const v8u16*)&As[(s * PM + wm + 16 * i + (lane & 15)) * ALD + k0 + 8 * (lane >> 4)];
#pragma unroll
          for (int j = 0; j < 2; ++j)
            b[s][j] = *(const v8u16*)&Bs[(s * PTN + wn + 16 * j + (lane & 15)) * BLD + 8 * (lane >> 4)];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int s = S - 1; s >= 0; --s)
#pragma unroll
              for (int t = S - 1 - s; t >= 0; --t) {
                const int l = (POL == 4) ? (s + t) : 0;
                acc[l][i][j] = mfma_bf16(a[s][i], b[t][j], acc[l][i][j]);
              }
      }
      __syncthreads();
    }
    // epilogue
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int nn = wn + 16 * j + (lane & 15);
        if (c0 + nn >= pm.ncols) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = wm + 16 * i + ((POL == 0) ? mfma16_row_f64(lane, r) : mfma16_row_f32(lane, r));
          if (m >= M) continue;
          if constexpr (POL == 4) {
            double y = 0.0;
#pragma unroll
            for (int l = NL - 1; l >= 0; --l) y = y * (1.0 / 256.0) + (double)acc[l][i][j][r];
            Y[yb[nn] + m * Q] = (TX)(y * (1.0 / 65536.0) * Arow[m] * colscale[nn]);
          } else {
            Y[yb[nn] + m * Q] = (TX)acc[0][i][j][r];
          }
        }
      }
  }
}

static int g_ncu = 0;

int persist_grid() {
  if (g_ncu == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&g_ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || g_ncu <= 0)
      g_ncu = 256;
  }
  return g_ncu;
}

// Returns 1 if launched, 0 if the shape is not handled (caller falls back).
template <int POL>
int launch_persist(const void* A, const double* Arow, int M, int K, const void* X, void* Y, long long P,
                   long long Q, hipStream_t st, int* rc) {
  if (M > PM || K > PK || M <= 0 || K <= 0 || Q <= 0 || P < 0) return 0;
  typedef typename PTraits<POL>::TX TX;
  PMap pm;
  pm.P = P;
  pm.Q = Q;
  pm.ncols = P * Q;
  pm.M = M;
  pm.K = K;
  pm.ntiles = (pm.ncols + PTN - 1) / PTN;
  if (pm.ncols == 0) {
    *rc = JFX_OK;
    return 1;
  }
  auto kern = contract_persist_kernel<POL>;
  const size_t shm = PLds<POL>::TOTAL;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) {
    set_error("contract_persist: hipFuncSetAttribute: %s", hipGetErrorString(e));
    *rc = JFX_EHIP;
    return 1;
  }
  long long grid = persist_grid();
  if (grid > pm.ntiles) grid = pm.ntiles;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(PNT), shm, st, A, Arow, (const TX*)X, (TX*)Y, pm);
  e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("contract_persist launch: %s", hipGetErrorString(e));
    *rc = JFX_EHIP;
    return 1;
  }
  *rc = JFX_OK;
  return 1;
}

template int launch_persist<0>(const void*, const double*, int, int, const void*, void*, long long, long long,
                               hipStream_t, int*);
template int launch_persist<1>(const void*, const double*, int, int, const void*, void*, long long, long long,
                               hipStream_t, int*);
template int launch_persist<2>(const void*, const double*, int, int, const void*, void*, long long, long long,
                               hipStream_t, int*);
template int launch_persist<3>(const void*, const double*, int, int, const void*, void*, long long, long long,
                               hipStream_t, int*);
template int launch_persist<4>(const void*, const double*, int, int, const void*, void*, long long, long long,
                               hipStream_t, int*);

}  // namespace jfx
