// FP64 per-axis contraction on int8 MFMA with exact integer accumulation:
//     Y[p, m, q] = sum_k A[m, k] X[p, k, q],   M, K <= 128, X and Y FP64,
// for the two column layouts of every axis of a C-order array (Q % 64 == 0:
// a tile is 64 consecutive q of one p; Q == 1: a tile is 64 consecutive rows
// p, each K contiguous values).
//
// Both operands are cut into S balanced base-128 digits (values in [-64, 64])
// under a power-of-two scale: A once per row (jfx_split_matrix_i8), X per
// column of each staged tile (in the kernel), so that
//     A[m, k] = 2^ea_m sum_s a_s[m, k] 2^(-6-7s),   X[k, c] = 2^ex_c sum_t x_t[k, c] 2^(-6-7t)
// up to 2^(-6-7(S-1)) of the row / column maximum.  The S(S+1)/2 digit
// products with s + t < S run on v_mfma_i32_16x16x64_i8; every product and
// every sum of one level s + t = l is an integer below (l+1) K 64^2 < 2^23,
// so the int32 accumulators are exact, and the levels are combined in FP64:
//     Y[m, c] = 2^(ea_m + ex_c - 12) sum_l 2^(-7l) acc_l[m, c].
//
// One persistent workgroup of 8 waves per CU; wave w keeps the digit slices
// of operator rows 16w..16w+15 in registers and splits columns 8w..8w+7 of
// every tile (lane l: column 8w + l/8, rows 16 (l % 8) .. + 15; the column
// maxima are combined with lane shuffles), straight from the registers the
// tile was loaded into.  Digit tiles are double buffered, so one barrier per
// tile separates the split of tile t+1 from the products of tile t, and the
// two waves sharing a SIMD run the two in opposite order so that one's VALU
// work overlaps the other's MFMAs.
#include "jfx_common.h"

namespace jfx {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int I8T = 64, I8NT = 512, I8KMAX = 128;
constexpr int I8DLD = I8KMAX + 16;  // digit tile column stride (bytes): Dg[buf][s][c][k]

struct I8Map {
  long long P, Q, ncols, ntiles, nqt;  // nqt: 64-column tiles per p (Q % 64 == 0)
  int M, K, Kp, qone;
};

template <int S>
struct I8Lds {
  static constexpr size_t DG = (size_t)S * I8T * I8DLD;  // one digit tile
  static constexpr size_t CE = (size_t)I8T * 4;          // its column exponents
  static constexpr size_t TOTAL = 2 * (DG + CE);
};

// Balanced base-128 digits of the S-digit fixed-point value of xs (|xs| < 1),
// most significant first: xs ~ sum_s d[s] 2^(-6-7s) (the last digit from a
// truncated value).  y = xs 2^20 is cut into hi = trunc(y) and the residual
// lo = trunc((y - hi) 2^(7(S-1)-14)); digits are sign-extended 7-bit fields
// with the carry taken out, so every digit lies in [-64, 64].
template <int S>
__device__ __forceinline__ void i8_digits(double y, int (&d)[S]) {
  constexpr int LB = 7 * (S - 1) - 14;  // bits of lo
  const double h = trunc(y);            // |h| < 2^20
  int lo = (int)ldexp(y - h, LB);       // |lo| < 2^LB
  int hi = (int)h;
#pragma unroll
  for (int s = S - 1; s >= 3; --s) {  // the digits of lo
    const int dd = __builtin_amdgcn_sbfe(lo, 0, 7);
    d[s] = dd;
    lo = (lo - dd) >> 7;
  }
  hi += lo;  // carry into the 2^-20 part
#pragma unroll
  for (int s = 2; s >= 1; --s) {
    const int dd = __builtin_amdgcn_sbfe(hi, 0, 7);
    d[s] = dd;
    hi = (hi - dd) >> 7;
  }
  d[0] = hi;  // |hi| <= 64
}

// Low bytes of a, b, c, d packed into one dword (a lowest).
__device__ __forceinline__ int pack4(int a, int b, int c, int d) {
  const unsigned ab = __builtin_amdgcn_perm((unsigned)b, (unsigned)a, 0x0c0c0400u);
  const unsigned cd = __builtin_amdgcn_perm((unsigned)d, (unsigned)c, 0x0c0c0400u);
  return (int)__builtin_amdgcn_perm(cd, ab, 0x05040100u);
}

// FULL: K == 128 and M % 16 == 0, so no row or k bound needs a test.
template <int S, bool FULL>
__global__ __launch_bounds__(I8NT) void contract_i8_kernel(const signed char* __restrict__ Asl,
                                                           const int* __restrict__ Aexp,
                                                           const double* __restrict__ X, double* __restrict__ Y,
                                                           I8Map mp) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  signed char* Dg0 = reinterpret_cast<signed char*>(smem);                       // [2][S][64][I8DLD]
  int* ce0 = reinterpret_cast<int*>(smem + 2 * I8Lds<S>::DG);                    // [2][64]

  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int M = mp.M, K = mp.K, Kp = mp.Kp, nks = Kp >> 6;
  const bool wact = 16 * w < M;
  const int cw = 8 * w + (lane >> 3), kg = lane & 7;  // split: column and 16-row group of this lane
  const bool kact = 16 * kg < Kp;

  // ---- operator digits of rows 16w..16w+15 (zero past M), resident for the launch
  v4i ad[S][2];
  {
    const int row = 16 * w + (lane & 15);
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        v4i v = {0, 0, 0, 0};
        if (row < M && ks < nks)
          v = *(const v4i*)(Asl + ((long long)s * M + row) * Kp + 64 * ks + 16 * (lane >> 4));
        ad[s][ks] = v;
      }
  }
  int rexp[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = 16 * w + 4 * (lane >> 4) + r;
    rexp[r] = m < M ? Aexp[m] : 0;
  }

  // ---- this lane's 16 values of a tile: column cw, rows 16 kg + e
  double v[16];
  auto load_tile = [&](long long t) {
    if (mp.qone) {
      const long long col = t * I8T + cw;
      const bool ok = col < mp.ncols && kact;
      const double* src = X + col * K + 16 * kg;
#pragma unroll
      for (int e = 0; e < 16; e += 2) {
        double2 d = {0.0, 0.0};
        if (ok && (FULL || 16 * kg + e < K)) d = *(const double2*)(src + e);
        v[e] = d.x;
        v[e + 1] = d.y;
      }
    } else {
      // uniform slab base and 32-bit lane offsets (a slab K Q is below 2^29 values)
      const long long p = t / mp.nqt;
      const double* base = X + p * (long long)K * mp.Q + (t - p * mp.nqt) * I8T;
      const unsigned Qu = (unsigned)mp.Q;
      unsigned off = (unsigned)(16 * kg) * Qu + (unsigned)cw;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        if (FULL)
          v[e] = base[off];
        else
          v[e] = (kact && 16 * kg + e < K) ? base[off] : 0.0;
        off += Qu;
      }
    }
  };
  // split of the tile held in v, in four quarters of 4 values (quarter 0
  // also reduces the column maximum); the packed digits of the 16 values go
  // to LDS as one 16-byte store per slice after quarter 3
  int ex = 0;
  int packed[S][4];
  auto split_part = [&](int buf, int j) {
    if (j == 0) {
      double mx = 0.0;
#pragma unroll
      for (int e = 0; e < 16; ++e) mx = fmax(mx, fabs(v[e]));
#pragma unroll
      for (int o = 1; o < 8; o <<= 1) mx = fmax(mx, __shfl_xor(mx, o));
      ex = mx > 0.0 ? ilogb(mx) + 1 : 0;  // |X[:, cw]| < 2^ex
      if (kg == 0) ce0[buf * I8T + cw] = ex;
    }
    int dq[4][S];
#pragma unroll
    for (int e = 0; e < 4; ++e) i8_digits<S>(ldexp(v[4 * j + e], 20 - ex), dq[e]);
#pragma unroll
    for (int s = 0; s < S; ++s) packed[s][j] = pack4(dq[0][s], dq[1][s], dq[2][s], dq[3][s]);
    if (j == 3 && kact) {
      signed char* dst = Dg0 + (size_t)buf * I8Lds<S>::DG + (size_t)cw * I8DLD + 16 * kg;
#pragma unroll
      for (int s = 0; s < S; ++s)
        *(v4i*)(dst + (size_t)s * I8T * I8DLD) = v4i{packed[s][0], packed[s][1], packed[s][2], packed[s][3]};
    }
  };
  // products of tile t (digits in buf) by column block; after block j the
  // quarter j of the next tile's split (values in v) runs, and once v is
  // consumed the tile after that (tload >= 0) is loaded into it
  auto products = [&](long long t, int buf, bool split_next, long long tload) {
    const signed char* Dg = Dg0 + (size_t)buf * I8Lds<S>::DG;
    const int* ce = ce0 + buf * I8T;
    long long c0, ybase;
    if (mp.qone) {
      c0 = t * I8T;
      ybase = 0;
    } else {
      const long long p = t / mp.nqt;
      c0 = (t - p * mp.nqt) * I8T;
      ybase = p * (long long)M * mp.Q + c0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (wact) {
      v4i acc[S];
#pragma unroll
      for (int l = 0; l < S; ++l) acc[l] = v4i{0, 0, 0, 0};
      const int col = 16 * j + (lane & 15);
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        if (ks < nks) {
          v4i bd[S];
#pragma unroll
          for (int s = 0; s < S; ++s)
            bd[s] = *(const v4i*)(Dg + ((size_t)s * I8T + col) * I8DLD + 64 * ks + 16 * (lane >> 4));
          // operator slice s against data slice l - s for every level l >= s:
          // consecutive MFMAs go to different accumulators, so none waits on the last
#pragma unroll
          for (int s = 0; s < S; ++s)
#pragma unroll
            for (int l = s; l < S; ++l)
              acc[l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ad[s][ks], bd[l - s], acc[l], 0, 0, 0);
        }
      }
      const int ec = ce[col];
      double yv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // adjacent levels pair exactly in int32 (|acc_l| < 2^22): acc_l 2^7 + acc_{l+1}
        // terms from the least significant: pairs (l-1, l) in units of 2^(-7l),
        // each 2^14 above the previous; an odd S ends with level 0 alone
        double y = 0.0;
#pragma unroll
        for (int l = S - 1; l >= 0; l -= 2) {
          const int pr = l >= 1 ? (acc[l - 1][r] << 7) + acc[l][r] : acc[l][r];
          y = fma(y, 0x1p-14, (double)pr);
        }
        yv[r] = ldexp(y, rexp[r] + ec - 12 - ((S & 1) ? 0 : 7));
      }
      const int m0 = 16 * w + 4 * (lane >> 4);
      if (mp.qone) {
        const long long col_g = c0 + col;
        if (col_g < mp.ncols) {
          double* dst = Y + col_g * M + m0;
          if (FULL) {  // M even: 16-byte aligned pairs
            *(double2*)dst = double2{yv[0], yv[1]};
            *(double2*)(dst + 2) = double2{yv[2], yv[3]};
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (m0 + r < M) dst[r] = yv[r];
          }
        }
      } else {
        double* yb = Y + ybase;  // uniform; 32-bit lane offsets
        const unsigned Qu = (unsigned)mp.Q;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (FULL || m0 + r < M) yb[(unsigned)(m0 + r) * Qu + (unsigned)col] = yv[r];
      }
      }
      if (split_next) split_part(buf ^ 1, j);  // VALU work of the next tile between this tile's MFMA blocks
    }
    if (tload >= 0) load_tile(tload);
  };

  const long long nt = mp.ntiles, G = gridDim.x;
  long long t = blockIdx.x;
  if (t >= nt) return;
  load_tile(t);
#pragma unroll
  for (int j = 0; j < 4; ++j) split_part(0, j);
  if (t + G < nt) load_tile(t + G);
  __syncthreads();
  int buf = 0;
  for (; t < nt; t += G) {
    products(t, buf, t + G < nt, t + 2 * G < nt ? t + 2 * G : -1);
    __syncthreads();
    buf ^= 1;
  }
}

// A (M x K) -> S balanced int8 digit slices [s][M][Kp] (Kp = 64 ceil(K / 64),
// zero padded) and row exponents: |A[m, :]| < 2^aexp[m].
template <int S>
__global__ void split_i8_kernel(const double* __restrict__ A, int M, int K, int Kp, signed char* __restrict__ sl,
                                int* __restrict__ aexp) {
  const int m = blockIdx.x;
  __shared__ double red[64];
  double mx = 0.0;
  for (int k = threadIdx.x; k < K; k += 64) mx = fmax(mx, fabs(A[(long long)m * K + k]));
  red[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int i = 0; i < 64; ++i) r = fmax(r, red[i]);
    red[0] = r;
  }
  __syncthreads();
  const double rm = red[0];
  const int e = rm > 0.0 ? ilogb(rm) + 1 : 0;
  if (threadIdx.x == 0) aexp[m] = e;
  for (int k = threadIdx.x; k < Kp; k += 64) {
    int d[S];
    i8_digits<S>(k < K ? ldexp(A[(long long)m * K + k], 20 - e) : 0.0, d);
#pragma unroll
    for (int s = 0; s < S; ++s) sl[((long long)s * M + m) * Kp + k] = (signed char)d[s];
  }
}

template <int S>
static int launch_contract_i8(const signed char* Asl, const int* Aexp, int M, int K, const double* X, double* Y,
                              long long P, long long Q, hipStream_t st) {
  I8Map mp{};
  mp.P = P;
  mp.Q = Q;
  mp.M = M;
  mp.K = K;
  mp.Kp = (K + 63) / 64 * 64;
  if (Q == 1) {
    mp.qone = 1;
    mp.ncols = P;
    mp.ntiles = (P + I8T - 1) / I8T;
    mp.nqt = 1;
  } else {
    mp.qone = 0;
    mp.nqt = Q / I8T;
    mp.ncols = P * Q;
    mp.ntiles = P * mp.nqt;
  }
  if (mp.ntiles == 0) return JFX_OK;
  int dev = 0, cus = 0;
  JFX_CHECK_HIP(hipGetDevice(&dev));
  JFX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const long long nb = mp.ntiles < cus ? mp.ntiles : cus;
  const size_t shm = I8Lds<S>::TOTAL;
  const bool full = K == I8KMAX && M % 16 == 0;
  auto kern = full ? contract_i8_kernel<S, true> : contract_i8_kernel<S, false>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) return set_error("contract_i8: hipFuncSetAttribute: %s", hipGetErrorString(e)), JFX_EHIP;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(I8NT), shm, st, Asl, Aexp, X, Y, mp);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // namespace jfx

using namespace jfx;

extern "C" {

int jfx_split_matrix_i8(const double* A, int M, int K, int nslices, signed char* slices, int* rowexp,
                        hipStream_t stream) {
  if (!A || !slices || !rowexp || M <= 0 || K <= 0) return set_error("split_matrix_i8: bad arguments"), JFX_EARG;
  const int Kp = (K + 63) / 64 * 64;
  switch (nslices) {
    case 5:
      hipLaunchKernelGGL(split_i8_kernel<5>, dim3(M), dim3(64), 0, stream, A, M, K, Kp, slices, rowexp);
      break;
    case 6:
      hipLaunchKernelGGL(split_i8_kernel<6>, dim3(M), dim3(64), 0, stream, A, M, K, Kp, slices, rowexp);
      break;
    case 7:
      hipLaunchKernelGGL(split_i8_kernel<7>, dim3(M), dim3(64), 0, stream, A, M, K, Kp, slices, rowexp);
      break;
    default:
      return set_error("split_matrix_i8: nslices must be 5, 6 or 7"), JFX_EARG;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_contract_i8(const signed char* Aslices, const int* Arowexp, int nslices, int M, int K, const double* X,
                    double* Y, long long P, long long Q, hipStream_t stream) {
  if (!Aslices || !Arowexp || !X || !Y || P < 0 || Q < 1) return set_error("contract_i8: bad arguments"), JFX_EARG;
  if (M < 1 || M > I8KMAX || K < 1 || K > I8KMAX)
    return set_error("contract_i8: M = %d, K = %d outside [1, %d]", M, K, I8KMAX), JFX_ESHAPE;
  if (Q != 1 && Q % I8T != 0)
    return set_error("contract_i8: the trailing size Q = %lld must be 1 or a multiple of %d", Q, I8T), JFX_ESHAPE;
  if (Q == 1 && (K & 1)) return set_error("contract_i8: K must be even when Q == 1"), JFX_ESHAPE;
  if ((double)K * (double)Q >= 536870912.0 || (double)M * (double)Q >= 536870912.0)
    return set_error("contract_i8: one slab K Q or M Q must stay below 2^29 values"), JFX_ESHAPE;
  if (((uintptr_t)X & 15) || ((uintptr_t)Y & 15))
    return set_error("contract_i8: X and Y must be 16-byte aligned"), JFX_EARG;
  switch (nslices) {
    case 5:
      return launch_contract_i8<5>(Aslices, Arowexp, M, K, X, Y, P, Q, stream);
    case 6:
      return launch_contract_i8<6>(Aslices, Arowexp, M, K, X, Y, P, Q, stream);
    case 7:
      return launch_contract_i8<7>(Aslices, Arowexp, M, K, X, Y, P, Q, stream);
    default:
      return set_error("contract_i8: nslices must be 5, 6 or 7"), JFX_EARG;
  }
}

}  // extern "C"
