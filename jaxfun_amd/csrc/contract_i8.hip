// FP64 per-axis contraction on int8 MFMA with exact integer accumulation:
//     Y[p, m, q] = sum_k A[m, k] X[p, k, q],   M, K <= 128, X and Y FP64,
// for the two column layouts of every axis of a C-order array (Q % 64 == 0:
// a tile is 32 consecutive q of one p; Q == 1: a tile is 32 consecutive rows
// p, each K contiguous values).
//
// Both operands are cut into S balanced base-128 digits (values in [-64, 64])
// under a power-of-two scale: A once per row (jfx_split_matrix_i8), X per
// column of each tile (in the kernel), so that
//     A[m, k] = 2^ea_m sum_s a_s[m, k] 2^(-6-7s),   X[k, c] = 2^ex_c sum_t x_t[k, c] 2^(-6-7t)
// up to 2^(-6-7(S-1)) of the row / column maximum.  The S(S+1)/2 digit
// products with s + t < S run on v_mfma_i32_16x16x64_i8; every product and
// every sum of one level s + t = l is an integer below (l+1) K 64^2 < 2^23,
// so the int32 accumulators are exact, and the levels are combined in FP64:
//     Y[m, c] = 2^(ea_m + ex_c - 12) sum_l 2^(-7l) acc_l[m, c].
//
// One persistent workgroup of 8 waves per CU; wave w keeps the digit slices
// of operator rows 16w..16w+15 in registers.  Tiles stream from global memory
// straight into two LDS staging buffers two tiles ahead; every wave splits
// four columns of a staged tile (lane l: column 4w + l/16, rows 8 (l % 16) ..
// + 7; column maxima by lane shuffles) into one of two digit buffers, between
// the MFMA blocks of the previous tile.
#include "jfx_common.h"

namespace jfx {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int I8T = 64, I8NT = 512, I8KMAX = 128;  // I8T: the trailing-size granule (Q % 64 == 0)
constexpr int I8DLD = I8KMAX + 16;  // digit tile column stride (bytes): Dg[buf][s][c][k]

struct I8Map {
  long long P, Q, ncols, ntiles, nqt;  // nqt: 32-column tiles per p (Q % 64 == 0)
  int M, K, Kp, qone;
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

// ---------------------------------------------------------------------------
// The FP64 values of a 32-column tile go from global memory straight into one
// of two LDS staging buffers (global_load_lds, 16 bytes per lane), issued two
// tiles ahead, so no register holds a tile in flight.
// Per tile: wait for the tile's own loads, barrier, then the split of tile
// t + 1 (from LDS) runs between the MFMA blocks of tile t.
//   Q % 64 == 0 staging: row k of the tile (32 doubles) at
//     (k / 4) * D32G + (k % 4) * 256 bytes (a 16-byte pad after every 4 rows);
//   Q == 1 staging: row c of the tile (K doubles) at c * D32G.
constexpr int I8T2 = 32;               // columns per streamed tile
constexpr int D32G = 1024 + 16;        // bytes per staged group (4 rows of 32 / one row of 128) + pad
constexpr int D32B = 32 * D32G;        // one staging buffer

template <int S>
struct I8DLds {
  static constexpr size_t DG = (size_t)S * I8T2 * I8DLD;  // one digit tile
};

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// Workgroup barrier over LDS only: this wave's LDS accesses complete, then
// s_barrier.  The staged tiles' direct loads are waited for explicitly
// (wait_vmcnt) by every wave before it, so the barrier need not also wait for
// the loads of the tile after next, which __syncthreads' fence would.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

typedef double v2d __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned lds_off(const void* p) {
  return (unsigned)(uintptr_t)(const __attribute__((address_space(3))) char*)p;
}

// Eight staged values of one column (Q % 64 == 0 layout): rows r..r+3 of one
// group and the next, at byte offsets 0, 256, 512, 768, D32G + (0..3) 256.
// Issued and waited for in one block, so no load is left in flight; read
// through asm because the compiler cannot tell these LDS reads from the
// direct loads still streaming into the other staging buffer.
__device__ __forceinline__ void lds_read8_col(unsigned a, double (&v)[8]) {
  static_assert(D32G == 1040, "offsets below assume 1040-byte groups");
  asm volatile(
      "ds_read_b64 %0, %8 offset:0\n\t"
      "ds_read_b64 %1, %8 offset:256\n\t"
      "ds_read_b64 %2, %8 offset:512\n\t"
      "ds_read_b64 %3, %8 offset:768\n\t"
      "ds_read_b64 %4, %8 offset:1040\n\t"
      "ds_read_b64 %5, %8 offset:1296\n\t"
      "ds_read_b64 %6, %8 offset:1552\n\t"
      "ds_read_b64 %7, %8 offset:1808\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]), "=&v"(v[7])
      : "v"(a)
      : "memory");
}

// Eight consecutive staged values of one row (Q == 1 layout).
__device__ __forceinline__ void lds_read8_row(unsigned a, double (&v)[8]) {
  v2d r0, r1, r2, r3;
  asm volatile(
      "ds_read_b128 %0, %4 offset:0\n\t"
      "ds_read_b128 %1, %4 offset:16\n\t"
      "ds_read_b128 %2, %4 offset:32\n\t"
      "ds_read_b128 %3, %4 offset:48\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3)
      : "v"(a)
      : "memory");
  v[0] = r0.x; v[1] = r0.y; v[2] = r1.x; v[3] = r1.y;
  v[4] = r2.x; v[5] = r2.y; v[6] = r3.x; v[7] = r3.y;
}

// The S digit fragments of one (column block, k-step) for this lane: slices
// 32 * I8DLD bytes apart, read and waited for in one block (same reason as above).
template <int S>
__device__ __forceinline__ void lds_read_frags(unsigned a, v4i (&b)[S]);
template <>
__device__ __forceinline__ void lds_read_frags<5>(unsigned a, v4i (&b)[5]) {
  static_assert(I8T2 * I8DLD == 4608, "slice stride");
  asm volatile(
      "ds_read_b128 %0, %5 offset:0\n\tds_read_b128 %1, %5 offset:4608\n\t"
      "ds_read_b128 %2, %5 offset:9216\n\tds_read_b128 %3, %5 offset:13824\n\t"
      "ds_read_b128 %4, %5 offset:18432\n\ts_waitcnt lgkmcnt(0)"
      : "=&v"(b[0]), "=&v"(b[1]), "=&v"(b[2]), "=&v"(b[3]), "=&v"(b[4])
      : "v"(a)
      : "memory");
}
template <>
__device__ __forceinline__ void lds_read_frags<6>(unsigned a, v4i (&b)[6]) {
  asm volatile(
      "ds_read_b128 %0, %6 offset:0\n\tds_read_b128 %1, %6 offset:4608\n\t"
      "ds_read_b128 %2, %6 offset:9216\n\tds_read_b128 %3, %6 offset:13824\n\t"
      "ds_read_b128 %4, %6 offset:18432\n\tds_read_b128 %5, %6 offset:23040\n\ts_waitcnt lgkmcnt(0)"
      : "=&v"(b[0]), "=&v"(b[1]), "=&v"(b[2]), "=&v"(b[3]), "=&v"(b[4]), "=&v"(b[5])
      : "v"(a)
      : "memory");
}
template <>
__device__ __forceinline__ void lds_read_frags<7>(unsigned a, v4i (&b)[7]) {
  asm volatile(
      "ds_read_b128 %0, %7 offset:0\n\tds_read_b128 %1, %7 offset:4608\n\t"
      "ds_read_b128 %2, %7 offset:9216\n\tds_read_b128 %3, %7 offset:13824\n\t"
      "ds_read_b128 %4, %7 offset:18432\n\tds_read_b128 %5, %7 offset:23040\n\t"
      "ds_read_b128 %6, %7 offset:27648\n\ts_waitcnt lgkmcnt(0)"
      : "=&v"(b[0]), "=&v"(b[1]), "=&v"(b[2]), "=&v"(b[3]), "=&v"(b[4]), "=&v"(b[5]), "=&v"(b[6])
      : "v"(a)
      : "memory");
}

// FULL: K == 128 and M % 16 == 0, so no row or k bound needs a test.
template <int S, bool FULL>
__global__ __launch_bounds__(I8NT) void contract_i8s_kernel(const signed char* __restrict__ Asl,
                                                            const int* __restrict__ Aexp,
                                                            const double* __restrict__ X, double* __restrict__ Y,
                                                            I8Map mp) {
  __shared__ __attribute__((aligned(16))) unsigned char stage0[D32B];
  __shared__ __attribute__((aligned(16))) unsigned char stage1[D32B];
  __shared__ __attribute__((aligned(16))) signed char Dg0[2 * I8DLds<S>::DG];
  __shared__ int ce0[2 * I8T2];

  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int M = mp.M, K = mp.K, Kp = mp.Kp, nks = Kp >> 6;
  const bool wact = 16 * w < M;
  const int cs = 4 * w + (lane >> 4), kg = lane & 15;  // split: column and 8-row group of this lane
  const bool kact = 8 * kg < Kp;

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

  // ---- four 16-byte direct loads per lane per tile (wave w: staged groups w + 8i)
  auto issue_loads = [&](long long t, unsigned char* stage) {
    if (mp.qone) {
      const long long last = mp.ncols - 1;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = w + 8 * i;                        // tile row (output column)
        const long long rowg = min(t * I8T2 + c, last);  // tail rows re-read the last row (zeroed in the split)
        const int ch = min(lane, (K >> 1) - 1);         // 16-byte chunk of the row (K even)
        const double* src = X + rowg * K + 2 * ch;
        __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(stage + c * D32G), 16, 0, 0);
      }
    } else {
      const long long p = t / mp.nqt;
      const double* base = X + p * (long long)K * mp.Q + (t - p * mp.nqt) * I8T2;
      const unsigned Qu = (unsigned)mp.Q;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int g = w + 8 * i;                              // group of 4 rows
        const int k = min(4 * g + (lane >> 4), K - 1);       // rows past K re-read row K-1 (zeroed in the split)
        const double* src = base + (unsigned)k * Qu + 2 * (lane & 15);
        __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(stage + g * D32G), 16, 0, 0);
      }
    }
  };

  // ---- split of one staged tile into digit buffer `buf`, in two halves of 4 values
  double v[8];
  int ex = 0;
  int packed[S][2];
  auto split_part = [&](long long t, const unsigned char* stage, int buf, int h) {
    if (h == 0) {
      const bool colok = !mp.qone || t * I8T2 + cs < mp.ncols;
      if (mp.qone)
        lds_read8_row(lds_off(stage + cs * D32G + 64 * kg), v);
      else
        lds_read8_col(lds_off(stage + 2 * kg * D32G + cs * 8), v);  // rows 8 kg .. 8 kg + 7
      double mx = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        if (!(FULL || (kact && 8 * kg + e < K)) || !colok) v[e] = 0.0;
        mx = fmax(mx, fabs(v[e]));
      }
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) mx = fmax(mx, __shfl_xor(mx, o));
      ex = mx > 0.0 ? ilogb(mx) + 1 : 0;
      if (kg == 0) ce0[buf * I8T2 + cs] = ex;
    }
    int dq[4][S];
#pragma unroll
    for (int e = 0; e < 4; ++e) i8_digits<S>(ldexp(v[4 * h + e], 20 - ex), dq[e]);
#pragma unroll
    for (int s = 0; s < S; ++s) packed[s][h] = pack4(dq[0][s], dq[1][s], dq[2][s], dq[3][s]);
    if (h == 1 && kact) {
      signed char* dst = Dg0 + (size_t)buf * I8DLds<S>::DG + (size_t)cs * I8DLD + 8 * kg;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        typedef int v2i __attribute__((ext_vector_type(2)));
        *(v2i*)(dst + (size_t)s * I8T2 * I8DLD) = v2i{packed[s][0], packed[s][1]};
      }
    }
  };

  auto products = [&](long long t, int buf, bool split_next, long long tn, const unsigned char* stage_n) {
    const signed char* Dg = Dg0 + (size_t)buf * I8DLds<S>::DG;
    const int* ce = ce0 + buf * I8T2;
    long long c0, ybase;
    if (mp.qone) {
      c0 = t * I8T2;
      ybase = 0;
    } else {
      const long long p = t / mp.nqt;
      c0 = (t - p * mp.nqt) * I8T2;
      ybase = p * (long long)M * mp.Q + c0;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (wact) {
        v4i acc[S];
#pragma unroll
        for (int l = 0; l < S; ++l) acc[l] = v4i{0, 0, 0, 0};
        const int col = 16 * j + (lane & 15);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          if (ks < nks) {
            v4i bd[S];
            lds_read_frags<S>(lds_off(Dg + (size_t)col * I8DLD + 64 * ks + 16 * (lane >> 4)), bd);
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
            if (FULL) {
              *(double2*)dst = double2{yv[0], yv[1]};
              *(double2*)(dst + 2) = double2{yv[2], yv[3]};
            } else {
#pragma unroll
              for (int r = 0; r < 4; ++r)
                if (m0 + r < M) dst[r] = yv[r];
            }
          }
        } else {
          double* yb = Y + ybase;
          const unsigned Qu = (unsigned)mp.Q;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (FULL || m0 + r < M) yb[(unsigned)(m0 + r) * Qu + (unsigned)col] = yv[r];
        }
      }
      if (split_next) split_part(tn, stage_n, buf ^ 1, j);
    }
  };

  const long long nt = mp.ntiles, G = gridDim.x;
  long long t = blockIdx.x;
  if (t >= nt) return;
  issue_loads(t, stage0);
  const bool has1 = t + G < nt;
  if (has1) issue_loads(t + G, stage1);
  if (has1)
    wait_vmcnt<4>();  // tile t's four loads done, tile t + G's may fly
  else
    wait_vmcnt<0>();
  lds_barrier();
  split_part(t, stage0, 0, 0);
  split_part(t, stage0, 0, 1);
  lds_barrier();
  // one step per tile: sn holds tile t + G, snn (split in the previous step)
  // receives tile t + 2G; the loop is unrolled by two so that every access
  // names one staging buffer, and the waits need not cover the other's loads
  auto step = [&](long long tt, int buf, const unsigned char* sn, unsigned char* snn) -> bool {
    const bool hasn = tt + G < nt, hasnn = tt + 2 * G < nt;
    if (hasnn) issue_loads(tt + 2 * G, snn);
    if (hasnn)
      wait_vmcnt<4>();
    else
      wait_vmcnt<0>();
    lds_barrier();
    products(tt, buf, hasn, tt + G, sn);
    lds_barrier();
    return hasn;
  };
  for (;;) {
    if (!step(t, 0, stage1, stage0)) break;
    t += G;
    if (!step(t, 1, stage0, stage1)) break;
    t += G;
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
    mp.ntiles = (P + I8T2 - 1) / I8T2;
    mp.nqt = 1;
  } else {
    mp.qone = 0;
    mp.nqt = Q / I8T2;
    mp.ncols = P * Q;
    mp.ntiles = P * mp.nqt;
  }
  if (mp.ntiles == 0) return JFX_OK;
  int dev = 0, cus = 0;
  JFX_CHECK_HIP(hipGetDevice(&dev));
  JFX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const long long nb = mp.ntiles < cus ? mp.ntiles : cus;
  const bool full = K == I8KMAX && M % 16 == 0;
  auto kern = full ? contract_i8s_kernel<S, true> : contract_i8s_kernel<S, false>;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(I8NT), 0, st, Asl, Aexp, X, Y, mp);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

// ---------------------------------------------------------------------------
// Fused two-axis FP64 contraction of 2D slabs on int8 MFMA:
//     Y[b] = A0 X[b] A1^T,   X[b]: K0 x K1, Y[b]: M0 x M1 (all <= 128, K1 even),
// one HBM read and one write per slab instead of two of each.  With the
// digit scheme above, per slab:
//   * the rows of X are staged in two halves of 64: each row is scaled by a
//     power of two from its maximum and cut into S digit slices in LDS (one
//     split per element, by the staging thread);
//   * phase A (wave w = column block w of T = X A1^T, A1's digit fragment of
//     rows 16w.. resident in registers for every slab) sums the levels of
//     T's rows of the half in int32 and combines them in FP64; T's column
//     j is then scaled per half by its maximum over the half's 64 rows
//     (lane shuffles) and cut into digits, stored transposed (T^T[j][i]);
//   * phase B (wave w = row block w of Y, A0's fragment read from L2) contracts
//     over i = T's rows: one 64-deep k-step per half, each combined in FP64
//     with its half's column scale, so the halves need no common scale.
// LDS rows are 128 bytes with 16-byte chunks XOR-swizzled by the row's low
// three bits (conflict-free 16-byte fragment reads without padding):
// S X-half slices of 64 rows and S T^T slices of 128 rows; S <= 6 fits.
// The next half (or slab) is loaded into registers during the current phase.
constexpr int PI8_XS = 64 * I8KMAX, PI8_TS = 128 * I8KMAX;  // bytes per digit slice (X half, T^T)

template <int S>
struct PI8Lds {
  static constexpr size_t X = (size_t)S * PI8_XS, T = (size_t)S * PI8_TS;
  static constexpr size_t TOTAL = X + T + (64 + 2 * 128) * sizeof(int);
};

__device__ __forceinline__ int pi8_addr(int row, int k) {  // byte k of a 128-byte swizzled row
  return row * I8KMAX + ((((k >> 4) ^ (row & 7))) << 4) + (k & 15);
}

// rows 64 h .. 64 h + 63 of one slab: thread q takes rows 64 h + q / 16 and
// + 32, values 8 (q % 16) .. + 7 (pairs; K1 even, the slab 16-byte aligned);
// zero outside the slab.
__device__ __forceinline__ void pi8_load(const double* __restrict__ Xb, int K0, int K1, int h, double (&v)[2][8]) {
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int i = 64 * h + (threadIdx.x >> 4) + 32 * u, k0 = 8 * (threadIdx.x & 15);
    const double* p = Xb + (long long)i * K1 + k0;
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      const double2 d = (i < K0 && k0 + e < K1) ? *(const double2*)(p + e) : double2{0.0, 0.0};
      v[u][e] = d.x;
      v[u][e + 1] = d.y;
    }
  }
}

template <int S>
__device__ __forceinline__ void pi8_store(signed char* Xd, int* xe, const double (&v)[2][8]) {
  typedef int v2i __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int il = (threadIdx.x >> 4) + 32 * u, kg = threadIdx.x & 15;
    double mx = 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e) mx = fmax(mx, fabs(v[u][e]));
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) mx = fmax(mx, __shfl_xor(mx, o));  // the row's 16 lanes
    const int ex = mx > 0.0 ? ilogb(mx) + 1 : 0;
    if (kg == 0) xe[il] = ex;
    int d[8][S];
#pragma unroll
    for (int e = 0; e < 8; ++e) i8_digits<S>(ldexp(v[u][e], 20 - ex), d[e]);
    signed char* dst = Xd + pi8_addr(il, 8 * kg);
#pragma unroll
    for (int s = 0; s < S; ++s)
      *(v2i*)(dst + s * PI8_XS) = v2i{pack4(d[0][s], d[1][s], d[2][s], d[3][s]), pack4(d[4][s], d[5][s], d[6][s], d[7][s])};
  }
}

// sum_l 2^(-7l) acc_l for element r (adjacent levels paired in int32; an
// even S leaves a factor 2^7, taken out by the caller's exponent)
template <int S>
__device__ __forceinline__ double pi8_levels(const v4i (&acc)[S], int r) {
  double y = 0.0;
#pragma unroll
  for (int l = S - 1; l >= 0; l -= 2) {
    const int pr = l >= 1 ? (acc[l - 1][r] << 7) + acc[l][r] : acc[l][r];
    y = fma(y, 0x1p-14, (double)pr);
  }
  return y;
}

template <int S>
__global__ __launch_bounds__(I8NT) void plane_i8_kernel(const signed char* __restrict__ A0sl,
                                                        const int* __restrict__ A0exp,
                                                        const signed char* __restrict__ A1sl,
                                                        const int* __restrict__ A1exp, const double* __restrict__ X,
                                                        double* __restrict__ Y, long long B, int M0, int K0, int M1,
                                                        int K1) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  signed char* Xd = reinterpret_cast<signed char*>(smem);  // [S][64 rows i][128 k]
  signed char* Td = Xd + PI8Lds<S>::X;                      // [S][128 columns j][128 rows i]
  int* xe = reinterpret_cast<int*>(Td + PI8Lds<S>::T);      // row exponents of the staged half
  int* te = xe + 64;                                         // [half][column] exponents of T
  const int lane = threadIdx.x & 63, l16 = lane & 15, kq = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int Kp0 = (K0 + 63) / 64 * 64, Kp1 = (K1 + 63) / 64 * 64;
  const int nh = Kp0 >> 6, nks1 = Kp1 >> 6;
  const bool actA = 16 * w < M1, actB = 16 * w < M0;
  constexpr int CORR = 12 + ((S & 1) ? 0 : 7);
  const long long slab = (long long)K0 * K1;

  // A1 digits of T's column block w (all k, resident); A0 digits of Y's row
  // block w are read from L2 at each phase B (registers)
  v4i b1[S][2];
  int e1, e0[4];
  const int mrow = 16 * w + l16;
  {
    const int m = mrow;
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
        b1[s][ks] = (m < M1 && ks < nks1) ? *(const v4i*)(A1sl + ((long long)s * M1 + m) * Kp1 + 64 * ks + 16 * kq)
                                          : v4i{0, 0, 0, 0};
    e1 = m < M1 ? A1exp[m] : 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mo = 16 * w + 4 * kq + r;
      e0[r] = mo < M0 ? A0exp[mo] : 0;
    }
  }

  long long b = blockIdx.x;
  if (b >= B) return;
  double v[2][8];
  pi8_load(X + b * slab, K0, K1, 0, v);
  pi8_store<S>(Xd, xe, v);
  __syncthreads();
  for (;;) {
    for (int h = 0; h < nh; ++h) {
      const bool nexth = h + 1 < nh;
      if (nexth) pi8_load(X + b * slab, K0, K1, h + 1, v);
      // ---- phase A: T rows 64 h .. 64 h + 63, column block w
      if (actA) {
        double tv[4][4];
        double mx = 0.0;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          v4i acc[S];
#pragma unroll
          for (int l = 0; l < S; ++l) acc[l] = v4i{0, 0, 0, 0};
#pragma unroll
          for (int ks = 0; ks < 2; ++ks) {
            if (ks < nks1) {
              const signed char* xr = Xd + pi8_addr(16 * rb + l16, 64 * ks + 16 * kq);
              v4i xd[S];
#pragma unroll
              for (int s = 0; s < S; ++s) xd[s] = *(const v4i*)(xr + s * PI8_XS);
              // level l = operator digit s + data digit l - s
#pragma unroll
              for (int s = 0; s < S; ++s)
#pragma unroll
                for (int l = s; l < S; ++l)
                  acc[l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(xd[l - s], b1[s][ks], acc[l], 0, 0, 0);
            }
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int il = 16 * rb + 4 * kq + r;  // T row within the half (mfma16_row_f32)
            tv[rb][r] = ldexp(pi8_levels<S>(acc, r), xe[il] + e1 - CORR);
            mx = fmax(mx, fabs(tv[rb][r]));
          }
        }
        // the column's maximum over the half: the four lanes holding column l16
        mx = fmax(mx, __shfl_xor(mx, 16));
        mx = fmax(mx, __shfl_xor(mx, 32));
        const int et = mx > 0.0 ? ilogb(mx) + 1 : 0;
        const int j = 16 * w + l16;
        if (kq == 0) te[128 * h + j] = et;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          int d[4][S];
#pragma unroll
          for (int r = 0; r < 4; ++r) i8_digits<S>(ldexp(tv[rb][r], 20 - et), d[r]);
          signed char* dst = Td + pi8_addr(j, 64 * h + 16 * rb + 4 * kq);
#pragma unroll
          for (int s = 0; s < S; ++s) *(int*)(dst + s * PI8_TS) = pack4(d[0][s], d[1][s], d[2][s], d[3][s]);
        }
      }
      __syncthreads();
      if (nexth) {
        pi8_store<S>(Xd, xe, v);
        __syncthreads();
      }
    }
    // ---- phase B: Y row block w; the next slab's first half loads meanwhile
    const long long nb = b + gridDim.x;
    const bool more = nb < B;
    if (more) pi8_load(X + nb * slab, K0, K1, 0, v);
    if (actB) {
      v4i a0[S][2];
#pragma unroll
      for (int s = 0; s < S; ++s)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
          a0[s][ks] = (mrow < M0 && ks < nh)
                          ? *(const v4i*)(A0sl + ((long long)s * M0 + mrow) * Kp0 + 64 * ks + 16 * kq)
                          : v4i{0, 0, 0, 0};
      double* Yb = Y + b * (long long)M0 * M1;
      const int ncb = (M1 + 15) >> 4;
#pragma unroll 1
      for (int t = 0; t < ncb; ++t) {
        const int j = 16 * t + l16;
        double y[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          if (ks < nh) {
            const signed char* tr = Td + pi8_addr(j, 64 * ks + 16 * kq);
            v4i td[S], acc[S];
#pragma unroll
            for (int s = 0; s < S; ++s) td[s] = *(const v4i*)(tr + s * PI8_TS);
#pragma unroll
            for (int l = 0; l < S; ++l) acc[l] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
              for (int l = s; l < S; ++l)
                acc[l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0[s][ks], td[l - s], acc[l], 0, 0, 0);
            const int ej = te[128 * ks + j];
#pragma unroll
            for (int r = 0; r < 4; ++r) y[r] += ldexp(pi8_levels<S>(acc, r), e0[r] + ej - CORR);
          }
        }
        if (j < M1) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int mo = 16 * w + 4 * kq + r;
            if (mo < M0) Yb[(long long)mo * M1 + j] = y[r];
          }
        }
      }
    }
    __syncthreads();
    if (!more) break;
    pi8_store<S>(Xd, xe, v);
    __syncthreads();
    b = nb;
  }
}

template <int S>
static int launch_plane_i8(const signed char* A0sl, const int* A0exp, const signed char* A1sl, const int* A1exp,
                           const double* X, double* Y, long long B, int M0, int K0, int M1, int K1, hipStream_t st) {
  int dev = 0, cus = 0;
  JFX_CHECK_HIP(hipGetDevice(&dev));
  JFX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const size_t shm = PI8Lds<S>::TOTAL;
  JFX_CHECK_HIP(hipFuncSetAttribute((const void*)plane_i8_kernel<S>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)shm));
  const long long nb = B < cus ? B : cus;
  hipLaunchKernelGGL(plane_i8_kernel<S>, dim3((unsigned)nb), dim3(I8NT), shm, st, A0sl, A0exp, A1sl, A1exp, X, Y, B,
                     M0, K0, M1, K1);
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

int jfx_plane_i8(const signed char* A0slices, const int* A0rowexp, const signed char* A1slices, const int* A1rowexp,
                 int nslices, const double* X, double* Y, long long B, int M0, int K0, int M1, int K1,
                 hipStream_t stream) {
  if (!A0slices || !A0rowexp || !A1slices || !A1rowexp || !X || !Y || B < 0)
    return set_error("plane_i8: bad arguments"), JFX_EARG;
  if (M0 < 1 || M0 > I8KMAX || K0 < 1 || K0 > I8KMAX || M1 < 1 || M1 > I8KMAX || K1 < 1 || K1 > I8KMAX)
    return set_error("plane_i8: dims must lie in [1, %d]", I8KMAX), JFX_ESHAPE;
  if (K1 & 1) return set_error("plane_i8: K1 must be even"), JFX_ESHAPE;
  if (((uintptr_t)X & 15) || ((uintptr_t)Y & 15)) return set_error("plane_i8: X and Y must be 16-byte aligned"), JFX_EARG;
  if (B == 0) return JFX_OK;
  switch (nslices) {
    case 5:
      return launch_plane_i8<5>(A0slices, A0rowexp, A1slices, A1rowexp, X, Y, B, M0, K0, M1, K1, stream);
    case 6:
      return launch_plane_i8<6>(A0slices, A0rowexp, A1slices, A1rowexp, X, Y, B, M0, K0, M1, K1, stream);
    default:
      return set_error("plane_i8: nslices must be 5 or 6 (the LDS tiles of 7 do not fit)"), JFX_EARG;
  }
}

}  // extern "C"
