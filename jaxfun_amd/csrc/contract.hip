// Per-axis contraction  Y[p, m, q] = sum_k A[m, k] X[p, k, q]  on MFMA.
//
// This is the one primitive behind TensorProductSpace.forward / backward /
// scalar_product and TPMatrix application: a small constant matrix A
// (M x K, M, K <= 256) applied along one axis of a C-order batch X viewed as
// (P, K, Q).  It is a GEMM C (M x P*Q) = A (M x K) * B (K x P*Q) whose column
// c maps to (p, q) = (c / Q, c % Q).
//
// Workgroup tile: TM x 64 outputs (TM = 64 or 128), 256 threads = 4 waves
// arranged 2 x 2; each wave owns (TM/2) x 32 outputs as 16 x 16 MFMA blocks.
// A tiles (constant, L2 resident) and X tiles stream through LDS; X is read
// from HBM once per m-tile (once in total when M <= TM).
//
// Precision variants:
//   f64    v_mfma_f64_16x16x4f64
//   f32    v_mfma_f32_16x16x4f32
//   bf16x3 fp32 operands split into 2 bf16 terms, 3 products, fp32 accumulate
//   bf16x6 fp32 operands split into 3 bf16 terms, 6 products, fp32 accumulate
//   ozaki  fp64 operands cut into S fixed-point 8-bit digit slices held in
//          bf16 (A per row, X per column, power-of-two scales); each bf16
//          product is an integer and every partial sum stays below 2^24, so
//          the fp32 MFMA accumulation is exact; levels are combined in fp64.
#include "jfx_common.h"

namespace jfx {

constexpr int TN = 64;

// Column-address modes for B/C tiles.
enum { COLS_GENERAL = 0, COLS_QONE = 1, COLS_INTILE = 2 };

struct ColMap {
  long long P, Q, ncols;
  int K, M, mode;
};

// Per-tile column bases: X address of (k, c0+nn) is xb[nn] + k*Q, Y address
// of (m, c0+nn) is yb[nn] + m*Q.  One 64-bit division per column per tile.
__device__ __forceinline__ void col_bases(const ColMap& cm, long long c0, long long* xb,
                                          long long* yb) {
  int nn = threadIdx.x;
  if (nn < TN) {
    long long c = c0 + nn;
    if (c >= cm.ncols) c = cm.ncols - 1;
    long long p = c / cm.Q, q = c - p * cm.Q;
    xb[nn] = p * cm.K * cm.Q + q;
    yb[nn] = p * cm.M * cm.Q + q;
  }
  __syncthreads();
}

__device__ __forceinline__ void tile_ids(int nmt, long long& m0, long long& c0, int TM) {
  unsigned nblk = gridDim.x;
  unsigned bid = xcd_remap(blockIdx.x, nblk);
  m0 = (long long)(bid % nmt) * TM;
  c0 = (long long)(bid / nmt) * TN;
}

// ----------------------------------------------------------------------------
// FP64 / FP32 MFMA (16x16x4): one element of A and B per lane per MFMA.
// ----------------------------------------------------------------------------
template <typename T, int TM>
__global__ __launch_bounds__(256) void contract_plain(const T* __restrict__ A,
                                                      const T* __restrict__ X,
                                                      T* __restrict__ Y, ColMap cm,
                                                      int nmt) {
  constexpr int KT = 16;
  constexpr int WM = TM / 2;        // rows per wave
  constexpr int BI = WM / 16;       // 16x16 blocks per wave along m
  constexpr int BJ = 2;             // 32 columns per wave
  __shared__ T As[KT][TM + 1];
  __shared__ T Bs[KT][TN + 1];
  __shared__ long long xb[TN], yb[TN];
  typedef T vacc __attribute__((ext_vector_type(4)));

  long long m0, c0;
  tile_ids(nmt, m0, c0, TM);
  col_bases(cm, c0, xb, yb);
  const long long Q = cm.Q;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w & 1) * WM, wn = (w >> 1) * 32;
  const int K = cm.K, M = cm.M;

  vacc acc[BI][BJ];
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) acc[i][j] = vacc{0, 0, 0, 0};

  for (int k0 = 0; k0 < K; k0 += KT) {
    for (int e = tid; e < TM * KT; e += 256) {
      int mm = e / KT, kk = e - mm * KT;
      long long m = m0 + mm;
      int k = k0 + kk;
      As[kk][mm] = (m < M && k < K) ? A[m * K + k] : T(0);
    }
    if (cm.mode == COLS_QONE) {
      for (int e = tid; e < KT * TN; e += 256) {
        int nn = e / KT, kk = e - nn * KT;
        long long c = c0 + nn;
        int k = k0 + kk;
        Bs[kk][nn] = (c < cm.ncols && k < K) ? X[xb[nn] + k * Q] : T(0);
      }
    } else {
      for (int e = tid; e < KT * TN; e += 256) {
        int kk = e / TN, nn = e - kk * TN;
        long long c = c0 + nn;
        int k = k0 + kk;
        Bs[kk][nn] = (c < cm.ncols && k < K) ? X[xb[nn] + k * Q] : T(0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4) {
      T a[BI], b[BJ];
#pragma unroll
      for (int i = 0; i < BI; ++i) a[i] = As[kk + (lane >> 4)][wm + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < BJ; ++j) b[j] = Bs[kk + (lane >> 4)][wn + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) {
          if constexpr (sizeof(T) == 8)
            acc[i][j] = mfma_f64(a[i], b[j], acc[i][j]);
          else
            acc[i][j] = mfma_f32(a[i], b[j], acc[i][j]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      int nn = wn + 16 * j + (lane & 15);
      long long c = c0 + nn;
      if (c >= cm.ncols) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = (sizeof(T) == 8) ? mfma16_row_f64(lane, r) : mfma16_row_f32(lane, r);
        long long m = m0 + wm + 16 * i + rr;
        if (m < M) Y[yb[nn] + m * Q] = acc[i][j][r];
      }
    }
}

// ----------------------------------------------------------------------------
// BF16 slice kernels (float split and fixed-point "ozaki" digits).
// ----------------------------------------------------------------------------
constexpr int KTB = 32;       // k per LDS stage (one 16x16x32 MFMA)
constexpr int LDB = KTB + 8;  // padded row (ushort) -> 80 B stride

// S slices of A (pre-split, [S][M][K] bf16 bits) and of X (split here).
// MODE 0: float split, one fp32 accumulator, pairs with s + t < S.
// MODE 1: fixed-point digits, one exact fp32 accumulator per level s + t.
template <int S, int MODE, int TM, typename TX>
__global__ __launch_bounds__(256) void contract_slices(const unsigned short* __restrict__ Asl,
                                                       const double* __restrict__ Arow,
                                                       const TX* __restrict__ X,
                                                       TX* __restrict__ Y, ColMap cm, int nmt) {
  const int Kp = (cm.K + KTB - 1) / KTB * KTB;  // slice row stride (zero padded)
  constexpr int WM = TM / 2;
  constexpr int BI = WM / 16;
  constexpr int BJ = 2;
  constexpr int NL = (MODE == 1) ? S : 1;  // accumulator levels
  __shared__ __attribute__((aligned(16))) unsigned short As[S][TM][LDB];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[S][TN][LDB];
  __shared__ double colscale[TN];  // 2^(e_c+1) per column (MODE 1)
  __shared__ int colexp[TN];
  __shared__ long long xb[TN], yb[TN];

  long long m0, c0;
  tile_ids(nmt, m0, c0, TM);
  col_bases(cm, c0, xb, yb);
  const long long Q = cm.Q;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w & 1) * WM, wn = (w >> 1) * 32;
  const int K = cm.K, M = cm.M;

  if constexpr (MODE == 1) {
    // per-column exponent: 4 threads per column scan K/4 entries each
    int nn = tid >> 2, part = tid & 3;
    long long c = c0 + nn;
    double mx = 0.0;
    if (c < cm.ncols)
      for (int k = part; k < K; k += 4) mx = fmax(mx, fabs((double)X[xb[nn] + k * Q]));
    mx = fmax(mx, __shfl_xor(mx, 1));
    mx = fmax(mx, __shfl_xor(mx, 2));
    if (part == 0) {
      int e = 0;
      if (mx > 0.0) frexp(mx, &e);
      colexp[nn] = e + 1;  // |x| * 2^-(e+1) < 0.5
      colscale[nn] = ldexp(1.0, e + 1);
    }
    __syncthreads();
  }

  v4f acc[NL][BI][BJ];
#pragma unroll
  for (int l = 0; l < NL; ++l)
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) acc[l][i][j] = v4f{0, 0, 0, 0};

  for (int k0 = 0; k0 < K; k0 += KTB) {
    // A slices: TM x KTB per slice, 16-byte chunks (slice rows padded to Kp)
    for (int ch = tid; ch < S * TM * (KTB / 8); ch += 256) {
      const int sl = ch / (TM * (KTB / 8)), rem = ch - sl * TM * (KTB / 8);
      const int mm = rem / (KTB / 8), part = rem - mm * (KTB / 8);
      const long long m = m0 + mm;
      v8u16 v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (m < M) v = *(const v8u16*)(Asl + ((long long)sl * M + m) * Kp + k0 + part * 8);
      *(v8u16*)&As[sl][mm][part * 8] = v;
    }
    // X tile: split each element into S slices, store [s][n][k]
    auto put = [&](int nn, int kk, double x) {
      if constexpr (MODE == 0) {
        float r = (float)x;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          unsigned short b = f2bf(r);
          Bs[s][nn][kk] = b;
          r -= bf2f(b);
        }
      } else {
        double r = ldexp(x, 8 - colexp[nn]);
#pragma unroll
        for (int s = 0; s < S; ++s) {
          double d = rint(r);
          Bs[s][nn][kk] = f2bf((float)d);  // |d| <= 128: exact in bf16
          r = (r - d) * 256.0;
        }
      }
    };
    constexpr int VW = 16 / sizeof(TX);  // elements per 16-byte load
    if (cm.mode == COLS_INTILE && c0 + TN <= cm.ncols && k0 + KTB <= K) {
      // a k-row of the tile is TN contiguous elements
      for (int ch = tid; ch < KTB * (TN / VW); ch += 256) {
        const int kk = ch / (TN / VW), nn = (ch - kk * (TN / VW)) * VW;
        const TX* src = X + xb[0] + (long long)(k0 + kk) * Q + nn;
        TX v[VW];
        *(uint4*)v = *(const uint4*)src;
#pragma unroll
        for (int i = 0; i < VW; ++i) put(nn + i, kk, (double)v[i]);
      }
    } else if (cm.mode == COLS_QONE && c0 + TN <= cm.ncols && k0 + KTB <= K && (K % VW) == 0) {
      // a column of the tile is KTB contiguous elements
      for (int ch = tid; ch < TN * (KTB / VW); ch += 256) {
        const int nn = ch / (KTB / VW), kk = (ch - nn * (KTB / VW)) * VW;
        const TX* src = X + xb[nn] + k0 + kk;
        TX v[VW];
        *(uint4*)v = *(const uint4*)src;
#pragma unroll
        for (int i = 0; i < VW; ++i) put(nn, kk + i, (double)v[i]);
      }
    } else {
      for (int e = tid; e < KTB * TN; e += 256) {
        int kk, nn;
        if (cm.mode == COLS_QONE) {
          nn = e / KTB;
          kk = e - nn * KTB;
        } else {
          kk = e / TN;
          nn = e - kk * TN;
        }
        long long c = c0 + nn;
        int k = k0 + kk;
        put(nn, kk, (c < cm.ncols && k < K) ? (double)X[xb[nn] + k * Q] : 0.0);
      }
    }
    __syncthreads();
    v8u16 a[S][BI], b[S][BJ];
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int i = 0; i < BI; ++i)
        a[s][i] = *(const v8u16*)&As[s][wm + 16 * i + (lane & 15)][8 * (lane >> 4)];
#pragma unroll
      for (int j = 0; j < BJ; ++j)
        b[s][j] = *(const v8u16*)&Bs[s][wn + 16 * j + (lane & 15)][8 * (lane >> 4)];
    }
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j)
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
          for (int t = 0; t < S - s; ++t) {
            const int l = (MODE == 1) ? (s + t) : 0;
            acc[l][i][j] = mfma_bf16(a[s][i], b[t][j], acc[l][i][j]);
          }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      int nn = wn + 16 * j + (lane & 15);
      long long c = c0 + nn;
      if (c >= cm.ncols) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        long long m = m0 + wm + 16 * i + mfma16_row_f32(lane, r);
        if (m >= M) continue;
        if constexpr (MODE == 0) {
          Y[yb[nn] + m * Q] = (TX)acc[0][i][j][r];
        } else {
          double y = 0.0;
#pragma unroll
          for (int l = NL - 1; l >= 0; --l) y = y * (1.0 / 256.0) + (double)acc[l][i][j][r];
          // digits carry 2^-8 per slice: value = sum_l acc_l 2^-8(l+2) * scales
          y *= (1.0 / 65536.0) * Arow[m] * colscale[nn];
          Y[yb[nn] + m * Q] = (TX)y;
        }
      }
    }
}

// Split a row-major FP64 matrix into S bf16 slices.
// mode 0: float split of fp32(A): A ~ sum_s bf16_s
// mode 1: fixed-point digits per row: A[m,k] = 2^(e_m) sum_s d_s 2^-8(s+1),
//         |d_s| <= 128, rowscale[m] = 2^(e_m).
__global__ void split_matrix_kernel(const double* __restrict__ A, int M, int K, int mode, int S,
                                    unsigned short* __restrict__ out, double* __restrict__ rowscale) {
  int m = blockIdx.x;
  __shared__ double red[256];
  double mx = 0.0;
  for (int k = threadIdx.x; k < K; k += blockDim.x) mx = fmax(mx, fabs(A[(long long)m * K + k]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  mx = red[0];
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);
  e += 1;
  if (threadIdx.x == 0 && rowscale) rowscale[m] = ldexp(1.0, e);
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    double a = A[(long long)m * K + k];
    if (mode == 0) {
      float r = (float)a;
      for (int s = 0; s < S; ++s) {
        unsigned short b = f2bf(r);
        out[((long long)s * M + m) * K + k] = b;
        r -= bf2f(b);
      }
    } else {
      double r = ldexp(a, 8 - e);
      for (int s = 0; s < S; ++s) {
        double d = rint(r);
        out[((long long)s * M + m) * K + k] = f2bf((float)d);
        r = (r - d) * 256.0;
      }
    }
  }
}

static int make_colmap(int M, int K, long long P, long long Q, ColMap& cm) {
  if (M <= 0 || K <= 0 || P < 0 || Q <= 0) {
    set_error("contract: bad sizes M=%d K=%d P=%lld Q=%lld", M, K, P, Q);
    return JFX_ESHAPE;
  }
  cm.P = P;
  cm.Q = Q;
  cm.ncols = P * Q;
  cm.K = K;
  cm.M = M;
  cm.mode = (Q == 1) ? COLS_QONE : ((Q % TN == 0) ? COLS_INTILE : COLS_GENERAL);
  return JFX_OK;
}

template <int POL>
int launch_persist(const void* A, const double* Arow, int M, int K, const void* X, void* Y, long long P,
                   long long Q, hipStream_t st, int* rc);
template <int POL>
int launch_ra(const void* A, int M, int K, const void* X, void* Y, long long P, long long Q, hipStream_t st,
              int* rc, int Mtot = 0, int moff = 0, int lda = 0, long long xps = 0, int accum = 0);

template <typename T>
static int launch_plain(const T* A, int M, int K, const T* X, T* Y, long long P, long long Q,
                        hipStream_t st) {
  {
    int rc = 0;
    if (launch_ra<(sizeof(T) == 8) ? 0 : 1>(A, M, K, X, Y, P, Q, st, &rc)) return rc;
    // 128 < M <= 256 (the padded backward transforms): two register-resident
    // launches over row halves of A, each writing its rows of Y
    if (M > 128 && M <= 256 && M % 4 == 0 && K <= 128) {
      const int M0 = ((M + 1) / 2 + 3) & ~3, M1 = M - M0;
      int rc1 = 0;
      if (M0 <= 128 && M1 > 0 && launch_ra<(sizeof(T) == 8) ? 0 : 1>(A, M0, K, X, Y, P, Q, st, &rc, M, 0)) {
        if (rc) return rc;
        if (launch_ra<(sizeof(T) == 8) ? 0 : 1>(A + (long long)M0 * K, M1, K, X, Y, P, Q, st, &rc1, M, M0))
          return rc1;
        return set_error("contract: second row half rejected"), JFX_ESHAPE;
      }
    }
    // 128 < K <= 256 with Q % 64 == 0 or Q == 1 (the padded forward
    // transforms): the columns of A in two halves, each half's rows in one
    // launch (M <= 128) or two; the second column half adds its partial sums
    // to Y.  With Q == 1 a half reads a strided slice of each row of X.
    if (K > 128 && K <= 256 && (Q % 64 == 0 || (Q == 1 && K % (16 / (int)sizeof(T)) == 0)) && M <= 256 &&
        (M <= 128 || M % 4 == 0)) {
      const int K0 = ((K + 1) / 2 + 3) & ~3;
      const int Mh = M <= 128 ? M : (((M + 1) / 2 + 3) & ~3);
      int n = 0;
      for (int kh = 0; kh < 2; ++kh)
        for (int m0 = 0; m0 < M; m0 += Mh) {
          const int k0 = kh ? K0 : 0, kn = kh ? K - K0 : K0, mr = M - m0 < Mh ? M - m0 : Mh;
          if (!launch_ra<(sizeof(T) == 8) ? 0 : 1>(A + (long long)m0 * K + k0, mr, kn, X + (long long)k0 * Q, Y, P,
                                                   Q, st, &rc, M, m0, K, (long long)K * Q, kh)) {
            if (n) return set_error("contract: K-split launch rejected"), JFX_ESHAPE;
            kh = 2;
            break;
          }
          if (rc) return rc;
          ++n;
        }
      if (n) return JFX_OK;
    }
    if (launch_persist<(sizeof(T) == 8) ? 0 : 1>(A, nullptr, M, K, X, Y, P, Q, st, &rc)) return rc;
  }
  ColMap cm;
  int rc = make_colmap(M, K, P, Q, cm);
  if (rc) return rc;
  if (cm.ncols == 0) return JFX_OK;
  const int TM = (M <= 64) ? 64 : 128;
  int nmt = (M + TM - 1) / TM;
  long long nct = (cm.ncols + TN - 1) / TN;
  long long nblk = nct * nmt;
  if (nblk > 0x7fffffffLL) return set_error("contract: grid too large"), JFX_ESHAPE;
  if (TM == 64)
    hipLaunchKernelGGL((contract_plain<T, 64>), dim3((unsigned)nblk), dim3(256), 0, st, A, X, Y, cm, nmt);
  else
    hipLaunchKernelGGL((contract_plain<T, 128>), dim3((unsigned)nblk), dim3(256), 0, st, A, X, Y, cm, nmt);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

template <int S, int MODE, typename TX>
static int launch_slices(const unsigned short* Asl, const double* Arow, int M, int K, const TX* X,
                         TX* Y, long long P, long long Q, hipStream_t st) {
  if (S == 2 || S == 3) {
    constexpr int POL = (MODE == 1) ? 4 : (S == 2 ? 2 : 3);
    int rc = 0;
    if constexpr (MODE == 0 && S == 3) {
      if (launch_ra<3>(Asl, M, K, X, Y, P, Q, st, &rc)) return rc;
      // 128 < M <= 256: two launches over row halves of the slices (rows of
      // Kp = K rounded up to 32 bf16 values)
      if (M > 128 && M <= 256 && M % 4 == 0 && K <= 128) {
        const int M0 = ((M + 1) / 2 + 3) & ~3, M1 = M - M0, Kp = (K + KTB - 1) / KTB * KTB;
        int rc1 = 0;
        if (M0 <= 128 && M1 > 0 && launch_ra<3>(Asl, M0, K, X, Y, P, Q, st, &rc, M, 0)) {
          if (rc) return rc;
          if (launch_ra<3>(Asl + (long long)M0 * Kp, M1, K, X, Y, P, Q, st, &rc1, M, M0)) return rc1;
          return set_error("contract: second row half rejected"), JFX_ESHAPE;
        }
      }
      // 128 < K <= 256: two column halves split at a multiple of 32 (the
      // first half's padded k-steps end where the second half starts), the
      // second accumulating into Y; rows in one launch or two halves
      if (K > 128 && K <= 256 && (Q % 64 == 0 || (Q == 1 && K % 8 == 0)) && M <= 256 &&
          (M <= 128 || M % 4 == 0)) {
        const int K0 = (K + 63) / 64 * 32, Kp = (K + KTB - 1) / KTB * KTB;
        const int Mh = M <= 128 ? M : (((M + 1) / 2 + 3) & ~3);
        int n = 0;
        for (int kh = 0; kh < 2; ++kh)
          for (int m0 = 0; m0 < M; m0 += Mh) {
            const int k0 = kh ? K0 : 0, kn = kh ? K - K0 : K0, mr = M - m0 < Mh ? M - m0 : Mh;
            if (!launch_ra<3>(Asl + (long long)m0 * Kp + k0, mr, kn, X + (long long)k0 * Q, Y, P, Q, st, &rc, M, m0,
                              Kp, (long long)K * Q, kh)) {
              if (n) return set_error("contract: K-split launch rejected"), JFX_ESHAPE;
              kh = 2;
              break;
            }
            if (rc) return rc;
            ++n;
          }
        if (n) return JFX_OK;
      }
    }
    if ((MODE == 0 || S == 3) && launch_persist<POL>(Asl, Arow, M, K, X, Y, P, Q, st, &rc)) return rc;
  }
  ColMap cm;
  int rc = make_colmap(M, K, P, Q, cm);
  if (rc) return rc;
  if (cm.ncols == 0) return JFX_OK;
  const int TM = (M <= 64) ? 64 : 128;
  int nmt = (M + TM - 1) / TM;
  long long nblk = ((cm.ncols + TN - 1) / TN) * nmt;
  if (nblk > 0x7fffffffLL) return set_error("contract: grid too large"), JFX_ESHAPE;
  if (TM == 64)
    hipLaunchKernelGGL((contract_slices<S, MODE, 64, TX>), dim3((unsigned)nblk), dim3(256), 0, st,
                       Asl, Arow, X, Y, cm, nmt);
  else
    hipLaunchKernelGGL((contract_slices<S, MODE, 128, TX>), dim3((unsigned)nblk), dim3(256), 0, st,
                       Asl, Arow, X, Y, cm, nmt);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // namespace jfx

using namespace jfx;

extern "C" {

int jfx_contract_f64(const double* A, int M, int K, const double* X, double* Y, long long P,
                     long long Q, hipStream_t stream) {
  if (!A || !X || !Y) return set_error("contract_f64: null pointer"), JFX_EARG;
  return launch_plain<double>(A, M, K, X, Y, P, Q, stream);
}

int jfx_contract_f32(const float* A, int M, int K, const float* X, float* Y, long long P,
                     long long Q, hipStream_t stream) {
  if (!A || !X || !Y) return set_error("contract_f32: null pointer"), JFX_EARG;
  return launch_plain<float>(A, M, K, X, Y, P, Q, stream);
}

int jfx_split_matrix(const double* A, int M, int K, int mode, int nslices, unsigned short* slices,
                     double* rowscale, hipStream_t stream) {
  if (!A || !slices || (mode == 1 && !rowscale)) return set_error("split_matrix: null pointer"), JFX_EARG;
  if (M <= 0 || K <= 0 || nslices < 1 || nslices > 6 || (mode != 0 && mode != 1))
    return set_error("split_matrix: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(split_matrix_kernel, dim3(M), dim3(256), 0, stream, A, M, K, mode, nslices, slices,
                     rowscale);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

// Float-split bf16 contraction on fp32 data; nslices 2 (3 products) or 3 (6 products).
int jfx_contract_bf16x(const unsigned short* Aslices, int nslices, int M, int K, const float* X,
                       float* Y, long long P, long long Q, hipStream_t stream) {
  if (!Aslices || !X || !Y) return set_error("contract_bf16x: null pointer"), JFX_EARG;
  if (nslices == 2) return launch_slices<2, 0, float>(Aslices, nullptr, M, K, X, Y, P, Q, stream);
  if (nslices == 3) return launch_slices<3, 0, float>(Aslices, nullptr, M, K, X, Y, P, Q, stream);
  return set_error("contract_bf16x: nslices must be 2 or 3"), JFX_EARG;
}

// Fixed-point-slice contraction on fp64 data (exact fp32 MFMA accumulation).
int jfx_contract_ozaki(const unsigned short* Aslices, const double* Arowscale, int nslices, int M,
                       int K, const double* X, double* Y, long long P, long long Q,
                       hipStream_t stream) {
  if (!Aslices || !Arowscale || !X || !Y) return set_error("contract_ozaki: null pointer"), JFX_EARG;
  // every level l = s + t sums (l+1) pairs of products |d| <= 128: exact if (l+1) K 2^14 <= 2^24
  if ((long long)nslices * K > 1024)
    return set_error("contract_ozaki: nslices*K=%lld exceeds 1024 (fp32 exactness)",
                     (long long)nslices * K), JFX_ESHAPE;
  switch (nslices) {
    case 3: return launch_slices<3, 1, double>(Aslices, Arowscale, M, K, X, Y, P, Q, stream);
    case 4: return launch_slices<4, 1, double>(Aslices, Arowscale, M, K, X, Y, P, Q, stream);
    case 5: return launch_slices<5, 1, double>(Aslices, Arowscale, M, K, X, Y, P, Q, stream);
  }
  return set_error("contract_ozaki: nslices must be 3, 4 or 5"), JFX_EARG;
}

}  // extern "C"
