// Dense LU with partial pivoting and the matching solve, FP64 real or
// complex (re/im interleaved), row-major n x n, in place.
//
// Blocked right-looking factorisation, panels of NB = 32 columns (16 complex):
//   1. panel:  one 1024-thread workgroup factors columns k0 .. k0+NB-1 of rows
//              k0 .. n-1 (pivot = first row of largest |re| + |im| in the
//              column, row swap inside the panel, multipliers, rank-1 update
//              of the panel's remaining columns);
//   2. rows:   every column outside the panel applies the panel's row swaps;
//              the columns right of it then solve L11 U12 = A12 (L11 unit
//              lower, staged in LDS);
//   3. update: A22 -= L21 U12 by a 64 x 64-tiled FP64 GEMM.
// Pivots are recorded as ipiv[j] (row swapped with row j at step j) and,
// at the end, as the permutation perm (row i of P A is row perm[i] of A).
// The solve gathers P b, then runs blocked forward (unit L) and backward (U)
// substitution, one launch per 32-row block (16 complex; the block's triangular solve
// and the update of the remaining rows fused, with a work buffer in between).
// Replaces jax.scipy.linalg.lu_factor / lu_solve as used by Matrix.lu_factor
// (la/matrix.py:138-214), TPMatricesDenseLUFactors (la/tpmatrix.py:1016-1086)
// and TensorMatrix.solve (la/tensormatrix.py:76-80).
#include <type_traits>

#include "jfx_common.h"

namespace jfx {

struct zd {
  double re, im;
};

__device__ __forceinline__ double dl_abs1(double a) { return fabs(a); }
__device__ __forceinline__ double dl_abs1(zd a) { return fabs(a.re) + fabs(a.im); }
__device__ __forceinline__ double dl_fms(double c, double a, double b) { return fma(-a, b, c); }  // c - a b
__device__ __forceinline__ zd dl_fms(zd c, zd a, zd b) {
  return {fma(-a.re, b.re, fma(a.im, b.im, c.re)), fma(-a.re, b.im, fma(-a.im, b.re, c.im))};
}
__device__ __forceinline__ double dl_div(double a, double b) { return a / b; }
__device__ __forceinline__ zd dl_div(zd a, zd b) {
  // Smith's algorithm (no overflow for |b| near the range limits)
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ __forceinline__ double dl_mul(double a, double b) { return a * b; }
__device__ __forceinline__ zd dl_mul(zd a, zd b) {
  return {fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ double dl_rcp(double a) { return 1.0 / a; }
__device__ __forceinline__ zd dl_rcp(zd a) { return dl_div(zd{1.0, 0.0}, a); }
template <typename T>
__device__ __forceinline__ T dl_zero();
template <>
__device__ __forceinline__ double dl_zero<double>() { return 0.0; }
template <>
__device__ __forceinline__ zd dl_zero<zd>() { return {0.0, 0.0}; }
__device__ __forceinline__ bool dl_ok(double a) { return a != 0.0 && isfinite(a); }
__device__ __forceinline__ bool dl_ok(zd a) { return (a.re != 0.0 || a.im != 0.0) && isfinite(a.re) && isfinite(a.im); }

constexpr int DL_NB = 32;     // largest panel width
constexpr int DL_PT = 1024;   // panel workgroup
// Panel width and solve block: a column's (right-hand side's) entries of one
// block live in registers (32 doubles or 16 complex values).
template <typename T>
constexpr int dl_nb() {
  return sizeof(T) == sizeof(double) ? 32 : 16;
}
template <typename T>
constexpr int dl_nbs() {  // the fused solve step's LDS (3 tiles) stays under 64 KB
  return sizeof(T) == sizeof(double) ? 32 : 16;
}
constexpr int DL_MAXN = 16384;

// ---- 1. panel factorisation (one workgroup).
template <typename T>
__global__ __launch_bounds__(DL_PT) void dl_panel_kernel(T* __restrict__ A, int n, int k0, int nb,
                                                          int* __restrict__ ipiv, int* __restrict__ status) {
  __shared__ double wv[DL_PT / 64];
  __shared__ int wi[DL_PT / 64];
  __shared__ int piv_row;
  __shared__ T prow[DL_NB];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int j = k0; j < k0 + nb; ++j) {
    // pivot: the first row of largest |a_ij| over i >= j
    double best = -1.0;
    int bi = n;
    for (int i = j + tid; i < n; i += DL_PT) {
      const double v = dl_abs1(A[(long long)i * n + j]);
      if (v > best) best = v, bi = i;  // rows ascend within a thread: ties keep the first
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o);
      const int oi = __shfl_xor(bi, o);
      if (ov > best || (ov == best && oi < bi)) best = ov, bi = oi;
    }
    if (lane == 0) wv[w] = best, wi[w] = bi;
    __syncthreads();
    if (tid == 0) {
      double b = -1.0;
      int r = n;
      for (int k = 0; k < DL_PT / 64; ++k)
        if (wv[k] > b || (wv[k] == b && wi[k] < r)) b = wv[k], r = wi[k];
      r = r < n ? r : j;
      piv_row = r;
      ipiv[j] = r;
      if (!dl_ok(A[(long long)r * n + j]) && *status == 0) *status = j + 1;
    }
    __syncthreads();
    const int r = piv_row;
    // swap rows j and r over the panel's columns, keep the pivot row in LDS
    if (tid < nb) {
      const int c = k0 + tid;
      const T a = A[(long long)j * n + c], b = A[(long long)r * n + c];
      if (r != j) A[(long long)r * n + c] = a;
      A[(long long)j * n + c] = b;
      prow[tid] = b;
    }
    __syncthreads();
    const T p = prow[j - k0];
    if (dl_ok(p)) {
      // multipliers and the rank-1 update of the panel's remaining columns
      for (int i = j + 1 + tid; i < n; i += DL_PT) {
        T* row = A + (long long)i * n;
        const T l = dl_div(row[j], p);
        row[j] = l;
        for (int c = j + 1; c < k0 + nb; ++c) row[c] = dl_fms(row[c], l, prow[c - k0]);
      }
    }
    __syncthreads();
  }
}

// ---- 1'. the same panel with its rows in registers (n - k0 <= NT R).
// Thread t owns panel rows k0 + t + NT m (m < R) for the whole panel.  The
// panel goes in sub-panels of SB columns: a sub-panel's entries of the owned
// rows are held in registers while its SB columns are pivoted and
// eliminated (pivot search, row exchange through LDS, rank-1 updates all in
// registers; only the panel's other columns are swapped in memory), then
// written back; the rest of the panel is updated once per sub-panel (unit
// lower solve of the sub-panel's rows, then a rank-SB update of the rows
// below).  Memory traffic per panel falls from (n - k0) NB^2 to about
// (n - k0) (2 NB + NB^2 / SB) values, and no column step waits on memory.
template <typename T, int R, int NT>
__global__ __launch_bounds__(NT) void dl_panel_reg_kernel(T* __restrict__ A, int n, int k0, int nb,
                                                              int* __restrict__ ipiv, int* __restrict__ status,
                                                              T* __restrict__ S) {
  constexpr int SB = sizeof(T) == sizeof(double) ? 8 : 4;
  constexpr int TC = 64 / sizeof(T);  // columns per transpose tile (64-byte row pieces)
  __shared__ T tt[NT / 64][32][TC + 1];  // per-wave transpose tile: 32 rows x TC columns
  __shared__ double wv[NT / 64];
  __shared__ int wi[NT / 64];
  __shared__ int piv_sh[2];
  __shared__ int pvs[DL_NB];  // the panel's pivot rows, stored to ipiv at the end
  __shared__ T srow[2][2][SB];  // [column parity][row j, row r] sub-panel entries
  __shared__ T Lss[SB][SB];
  __shared__ T Us[SB][DL_NB];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // The panel A[k0.., k0..k0+nb) is worked on column-major in S (ld rows per
  // column): a thread's rows are then consecutive with its neighbours', so
  // every access below coalesces, where the row-major matrix would put each
  // lane on its own cache line.  It is copied in and out through per-wave
  // LDS tiles (32 rows x one 64-byte piece of each row), coalesced both ways.
  const int ld = n - k0;
  // (c - k0) * ld + (i - k0) < 32 n <= 2^19: 32-bit offsets
  auto SP = [&](int i, int c) -> T& { return S[(c - k0) * ld + (i - k0)]; };
  // TB tiles' loads are issued before any of them is used (one memory round
  // trip per TB tiles, not per tile; complex panels keep one tile, their
  // registers are full)
  constexpr int TB = sizeof(T) != sizeof(double) ? 1 : (NT >= 1024 ? 2 : 8), TV = 32 * TC / 64;
  for (int rb0 = k0 + 32 * w; rb0 < n; rb0 += 32 * (NT / 64) * TB) {
    for (int cg = 0; cg < nb; cg += TC) {
      if constexpr (TB == 1) {
#pragma unroll
        for (int t = 0; t < TV; ++t) {
          const int e = lane + 64 * t, rr = e / TC, cc = e % TC;
          if (rb0 + rr < n && cg + cc < nb) tt[w][rr][cc] = A[(long long)(rb0 + rr) * n + k0 + cg + cc];
        }
#pragma unroll
        for (int t = 0; t < TV; ++t) {
          const int e = lane + 64 * t, cc = e / 32, rr = e % 32;
          if (rb0 + rr < n && cg + cc < nb) SP(rb0 + rr, k0 + cg + cc) = tt[w][rr][cc];
        }
      } else {
        T ld_[TB][TV];
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
          const int rb = rb0 + 32 * (NT / 64) * tb;
#pragma unroll
          for (int t = 0; t < TV; ++t) {
            const int e = lane + 64 * t, rr = e / TC, cc = e % TC;
            const bool in = rb + rr < n && cg + cc < nb;
            ld_[tb][t] = A[in ? (long long)(rb + rr) * n + k0 + cg + cc : 0];
          }
        }
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
          const int rb = rb0 + 32 * (NT / 64) * tb;
          if (rb >= n) break;  // uniform
#pragma unroll
          for (int t = 0; t < TV; ++t) {
            const int e = lane + 64 * t, rr = e / TC, cc = e % TC;
            tt[w][rr][cc] = ld_[tb][t];
          }
#pragma unroll
          for (int t = 0; t < TV; ++t) {
            const int e = lane + 64 * t, cc = e / 32, rr = e % 32;
            if (rb + rr < n && cg + cc < nb) SP(rb + rr, k0 + cg + cc) = tt[w][rr][cc];
          }
        }
      }
    }
  }
  __syncthreads();
  int bad = 0;  // thread 0: 1 + the first column with a zero pivot
  for (int s0 = k0; s0 < k0 + nb; s0 += SB) {
    const int sb = min(SB, k0 + nb - s0);
    T v[R][SB];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int i = k0 + tid + NT * m;
#pragma unroll
      for (int c = 0; c < SB; ++c) v[m][c] = (i < n && i >= s0 && c < sb) ? SP(i, s0 + c) : dl_zero<T>();
    }
    // the column index within the sub-panel is a compile-time constant, so
    // the register slots it selects need no runtime selects
#pragma unroll
    for (int jj = 0; jj < SB; ++jj) {
      if (jj >= sb) break;  // uniform
      const int j = s0 + jj, par = jj & 1;
      // pivot: the first row of largest |a_ij| over i >= j
      double best = -1.0;
      int bi = n;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int i = k0 + tid + NT * m;
        if (i >= j && i < n) {
          T x = v[m][0];
#pragma unroll
          for (int c = 1; c < SB; ++c)
            if (c == jj) x = v[m][c];
          const double a = dl_abs1(x);
          if (a > best) best = a, bi = i;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        if (ov > best || (ov == best && oi < bi)) best = ov, bi = oi;
      }
      if (lane == 0) wv[w] = best, wi[w] = bi;
      __syncthreads();
      if (tid == 0) {
        double b = -1.0;
        int r = n;
        for (int k = 0; k < NT / 64; ++k)
          if (wv[k] > b || (wv[k] == b && wi[k] < r)) b = wv[k], r = wi[k];
        r = r < n ? r : j;
        piv_sh[par] = r;
        pvs[j - k0] = r;
      }
      __syncthreads();
      const int r = piv_sh[par];
      // rows j and r of the sub-panel through LDS; the panel's other columns in memory
      const int oj = (j - k0) % NT, mj = (j - k0) / NT, orr = (r - k0) % NT, mr = (r - k0) / NT;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        if (tid == oj && m == mj)
#pragma unroll
          for (int c = 0; c < SB; ++c) srow[par][0][c] = v[m][c];
        if (tid == orr && m == mr)
#pragma unroll
          for (int c = 0; c < SB; ++c) srow[par][1][c] = v[m][c];
      }
      __syncthreads();
      if (r != j) {
#pragma unroll
        for (int m = 0; m < R; ++m) {
          if (tid == oj && m == mj)
#pragma unroll
            for (int c = 0; c < SB; ++c) v[m][c] = srow[par][1][c];
          if (tid == orr && m == mr)
#pragma unroll
            for (int c = 0; c < SB; ++c) v[m][c] = srow[par][0][c];
        }
      }
      // srow[par][1] is the pivot row now
      T p = srow[par][1][0];
#pragma unroll
      for (int c = 1; c < SB; ++c)
        if (c == jj) p = srow[par][1][c];
      if (tid == 0 && !dl_ok(p) && !bad) bad = j + 1;
      if (dl_ok(p)) {
        // multipliers by one reciprocal of the pivot (as LAPACK's getf2
        // scales the column), one division per column instead of one per
        // register slot the column index could select
        const T rp = dl_rcp(p);
#pragma unroll
        for (int m = 0; m < R; ++m) {
          const int i = k0 + tid + NT * m;
          if (i > j && i < n) {
            T x = v[m][0];
#pragma unroll
            for (int c = 1; c < SB; ++c)
              if (c == jj) x = v[m][c];
            const T l = dl_mul(x, rp);
#pragma unroll
            for (int c = 0; c < SB; ++c) {
              if (c == jj) v[m][c] = l;
              if (c > jj) v[m][c] = dl_fms(v[m][c], l, srow[par][1][c]);
            }
          }
        }
      }
      // srow[par] is rewritten two columns later, after two more barriers
    }
    // the sub-panel's row exchanges on the panel's other columns, in order
    // (kept out of the column loop, where each would cost a memory round
    // trip before a barrier)
    if (tid < nb) {
      const int c = k0 + tid;
      if (c < s0 || c >= s0 + sb) {
        for (int jj = 0; jj < sb; ++jj) {
          const int j = s0 + jj, r = pvs[j - k0];
          if (r != j) {
            const T a = SP(j, c);
            SP(j, c) = SP(r, c);
            SP(r, c) = a;
          }
        }
      }
    }
    // the sub-panel back to memory; L of the sub-panel's own rows to LDS
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int i = k0 + tid + NT * m;
      if (i >= s0 && i < n) {
#pragma unroll
        for (int c = 0; c < SB; ++c)
          if (c < sb) SP(i, s0 + c) = v[m][c];
      }
      if (i >= s0 && i < s0 + sb) {
#pragma unroll
        for (int c = 0; c < SB; ++c) Lss[i - s0][c] = v[m][c];
      }
    }
    __syncthreads();
    const int rc = k0 + nb - s0 - sb;  // panel columns right of the sub-panel
    if (rc > 0) {
      if (tid < rc) {  // U rows of the sub-panel: L_ss^-1 A[s0 .. s0+sb-1, c]
        const int c = s0 + sb + tid;
        T x[SB];
#pragma unroll
        for (int a = 0; a < SB; ++a) x[a] = a < sb ? SP(s0 + a, c) : dl_zero<T>();
#pragma unroll
        for (int a = 1; a < SB; ++a)
#pragma unroll
          for (int t = 0; t < SB; ++t)
            if (t < a) x[a] = dl_fms(x[a], Lss[a][t], x[t]);
#pragma unroll
        for (int a = 0; a < SB; ++a)
          if (a < sb) {
            SP(s0 + a, c) = x[a];
            Us[a][tid] = x[a];
          }
      }
      __syncthreads();
      // rank-sb update of the rows below, multipliers from registers; a
      // row's remaining entries are loaded together, updated, then stored
      // (one memory round trip per row, not one per entry)
      constexpr int RCM = dl_nb<T>() - SB;  // the panel columns right of a sub-panel, at most
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int i = k0 + tid + NT * m;
        if (i >= s0 + sb && i < n) {
          T a[RCM];
#pragma unroll
          for (int cc = 0; cc < RCM; ++cc) a[cc] = cc < rc ? SP(i, s0 + sb + cc) : dl_zero<T>();
#pragma unroll
          for (int t = 0; t < SB; ++t) {
            if (t < sb) {
#pragma unroll
              for (int cc = 0; cc < RCM; ++cc) a[cc] = dl_fms(a[cc], v[m][t], Us[t][cc]);
            }
            asm volatile("" ::: "memory");  // one row of Us in registers at a time
          }
#pragma unroll
          for (int cc = 0; cc < RCM; ++cc)
            if (cc < rc) SP(i, s0 + sb + cc) = a[cc];
        }
        asm volatile("" ::: "memory");  // one row's entries in registers at a time
      }
    }
    __syncthreads();
  }
  // the factored panel back to A (the barrier above ends the last sub-panel)
  for (int rb0 = k0 + 32 * w; rb0 < n; rb0 += 32 * (NT / 64) * TB) {
    for (int cg = 0; cg < nb; cg += TC) {
      if constexpr (TB == 1) {
#pragma unroll
        for (int t = 0; t < TV; ++t) {
          const int e = lane + 64 * t, cc = e / 32, rr = e % 32;
          if (rb0 + rr < n && cg + cc < nb) tt[w][rr][cc] = SP(rb0 + rr, k0 + cg + cc);
        }
#pragma unroll
        for (int t = 0; t < TV; ++t) {
          const int e = lane + 64 * t, rr = e / TC, cc = e % TC;
          if (rb0 + rr < n && cg + cc < nb) A[(long long)(rb0 + rr) * n + k0 + cg + cc] = tt[w][rr][cc];
        }
      } else {
        T ld_[TB][TV];
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
          const int rb = rb0 + 32 * (NT / 64) * tb;
#pragma unroll
          for (int t = 0; t < TV; ++t) {
            const int e = lane + 64 * t, cc = e / 32, rr = e % 32;
            const bool in = rb + rr < n && cg + cc < nb;
            ld_[tb][t] = in ? SP(rb + rr, k0 + cg + cc) : SP(k0, k0);
          }
        }
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
          const int rb = rb0 + 32 * (NT / 64) * tb;
          if (rb >= n) break;  // uniform
#pragma unroll
          for (int t = 0; t < TV; ++t) {
            const int e = lane + 64 * t, cc = e / 32, rr = e % 32;
            tt[w][rr][cc] = ld_[tb][t];
          }
#pragma unroll
          for (int t = 0; t < TV; ++t) {
            const int e = lane + 64 * t, rr = e / TC, cc = e % TC;
            if (rb + rr < n && cg + cc < nb) A[(long long)(rb + rr) * n + k0 + cg + cc] = tt[w][rr][cc];
          }
        }
      }
    }
  }
  if (tid < nb) ipiv[k0 + tid] = pvs[tid];
  if (tid == 0 && bad && *status == 0) *status = bad;
}

// ---- 2. row swaps outside the panel, then U12 = L11^-1 A12 right of it.
// One thread per column; L11 staged in LDS.
template <typename T>
__global__ __launch_bounds__(256) void dl_rows_kernel(T* __restrict__ A, int n, int k0, int nb,
                                                       const int* __restrict__ ipiv) {
  constexpr int NB = dl_nb<T>();
  __shared__ T L[NB][NB + 1];
  __shared__ int pv[NB];
  const int tid = threadIdx.x;
  for (int e = tid; e < nb * nb; e += 256) L[e / nb][e % nb] = A[(long long)(k0 + e / nb) * n + k0 + e % nb];
  if (tid < nb) pv[tid] = ipiv[k0 + tid];
  __syncthreads();
  // columns outside [k0, k0 + nb): c = idx for idx < k0, else idx + nb
  const int idx = blockIdx.x * 256 + tid;
  if (idx >= n - nb) return;
  const int c = idx < k0 ? idx : idx + nb;
  for (int t = 0; t < nb; ++t) {
    const int j = k0 + t, r = pv[t];
    if (r != j) {
      const T a = A[(long long)j * n + c];
      A[(long long)j * n + c] = A[(long long)r * n + c];
      A[(long long)r * n + c] = a;
    }
  }
  if (c < k0) return;
  T x[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) x[t] = t < nb ? A[(long long)(k0 + t) * n + c] : dl_zero<T>();
  // constant trip counts (both loops unroll, x stays in registers); the
  // per-row branch keeps each row's LDS reads next to their use
#pragma unroll
  for (int i = 1; i < NB; ++i) {
    if (i < nb) {
#pragma unroll
      for (int t = 0; t < NB; ++t)
        if (t < i) x[i] = dl_fms(x[i], L[i][t], x[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NB; ++t)
    if (t < nb) A[(long long)(k0 + t) * n + c] = x[t];
}

// ---- 3. C (m x nc, ldc) -= A (m x k, lda) B (k x nc, ldb): 64 x 64 output
// tile per 256-thread workgroup, 4 x 4 per thread, k in LDS slices of 16.
constexpr int DL_KS = 16;
template <typename T>
__global__ __launch_bounds__(256) void dl_gemm_sub_kernel(const T* __restrict__ A, long long lda,
                                                           const T* __restrict__ B, long long ldb, T* __restrict__ C,
                                                           long long ldc, int m, int nc, int k) {
  __shared__ T As[DL_KS][65];  // [k][row]
  __shared__ T Bs[DL_KS][65];  // [k][col]
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = dl_zero<T>();
  for (int k0 = 0; k0 < k; k0 += DL_KS) {
    __syncthreads();  // the previous slice is consumed
    for (int e = tid; e < 64 * DL_KS; e += 256) {
      const int rr = e / DL_KS, kk = e % DL_KS;  // A: row rr, k index kk (k contiguous in memory)
      As[kk][rr] = (r0 + rr < m && k0 + kk < k) ? A[(long long)(r0 + rr) * lda + k0 + kk] : dl_zero<T>();
      const int k2 = e >> 6, cc = e & 63;  // B: row k2, column cc
      Bs[k2][cc] = (k0 + k2 < k && c0 + cc < nc) ? B[(long long)(k0 + k2) * ldb + c0 + cc] : dl_zero<T>();
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < DL_KS; ++kk) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[kk][tr + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][tc + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = dl_fms(acc[a][b], av[a], bv[b]);  // acc -= a b
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int r = r0 + tr + 16 * a;
    if (r >= m) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int c = c0 + tc + 16 * b;
      if (c < nc) {
        T* p = C + (long long)r * ldc + c;
        if constexpr (sizeof(T) == sizeof(double))
          *p = *p + acc[a][b];
        else
          *p = zd{p->re + acc[a][b].re, p->im + acc[a][b].im};
      }
    }
  }
}

// ---- permutation from the recorded swaps (one thread, LDS)
__global__ void dl_perm_kernel(const int* __restrict__ ipiv, int n, int* __restrict__ perm) {
  __shared__ int pm[DL_MAXN];
  for (int i = threadIdx.x; i < n; i += blockDim.x) pm[i] = i;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int j = 0; j < n; ++j) {
      const int r = ipiv[j];
      const int t = pm[j];
      pm[j] = pm[r];
      pm[r] = t;
    }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = pm[i];
}

// ---- solve
template <typename T>
__global__ void dl_gather_kernel(const T* __restrict__ B, const int* __restrict__ perm, T* __restrict__ X, int n,
                                 long long nrhs, int* __restrict__ flags, int nflags) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nflags) flags[e] = 0;  // the single-launch solve's block flags (nflags <= n)
  if (e >= (long long)n * nrhs) return;
  const long long i = e / nrhs, r = e - i * nrhs;
  X[e] = B[(long long)perm[i] * nrhs + r];
}

// ---------------------------------------------------------------------------
// Single-launch triangular solve for up to 16 right-hand sides (FP64; B and
// X are n x nc row-major, nc <= NR).  Workgroup g owns the 64-row block blk
// (blk = g forward, nb - 1 - g backward, so that workgroups are dispatched in
// dependency order and a waiting workgroup only waits for workgroups
// dispatched before it).  It subtracts T[block, block j] x_j for every block
// j it depends on (j < blk forward, j > blk backward) as soon as workgroup j
// has published x_j (flags[j], release / acquire at agent scope), wave w
// taking columns 16w .. 16w + 15 of each block, then solves the diagonal
// block in registers (unit lower L forward; U with reciprocal pivots
// backward; wave 0 alone for NR <= 2, else each wave NR / 4 right-hand
// sides), stores x_blk and sets flags[blk].  The wait for a flag is bounded:
// a workgroup that gives up writes NaN to its block (and still sets its
// flag), so a failure shows in the result instead of hanging.
constexpr int DL_SF_BS = 64;
constexpr int DL_SF_MAXR = 16;
constexpr int DL_SF_SPIN = 1 << 20;  // flag polls (~1 us each with the sleep) before giving up

// v broadcast from lane t (t a compile-time constant after unrolling)
__device__ __forceinline__ double dl_bcast(double v, int t) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)b, t);
  const int hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), t);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ bool dl_sf_wait(const int* f) {
  int it = 0;
  while (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    if (++it >= DL_SF_SPIN) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  return true;
}

// LDS (doubles): Ld[64][64] (the diagonal block, transposed: Ld[t][r] =
// T[r0 + r][r0 + t]), xs[4][64 NR] (per wave: the block being applied, then
// the wave's partial sums), rdg[64] (1 / U[r][r], backward).
template <int NR>
constexpr size_t dl_sf_lds() {
  return (size_t)(DL_SF_BS * DL_SF_BS + 4 * DL_SF_BS * NR + DL_SF_BS) * sizeof(double);
}

template <int NR, bool LOWER>
__global__ __launch_bounds__(256) void dl_trsv_sf_kernel(const double* __restrict__ LU, int n, int nc,
                                                          const double* __restrict__ Bv, double* Xv, int* flags) {
  constexpr int BS = DL_SF_BS;
  constexpr int CW = NR <= 2 ? NR : NR / 4;  // right-hand sides per solving wave
  extern __shared__ double sm_sf[];
  double* Ld = sm_sf;                         // [t * BS + r]
  double* xs = Ld + BS * BS;                  // [w][BS * NR]
  double* rdg = xs + 4 * BS * NR;             // [BS]
  __shared__ int sbad;
  const int nb = (n + BS - 1) / BS;
  const int blk = LOWER ? (int)blockIdx.x : nb - 1 - (int)blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r0 = blk * BS;
  const int row = min(r0 + lane, n - 1);  // rows past n reread row n - 1 (results unused)
  double* xw = xs + w * BS * NR;
  if (tid == 0) sbad = 0;
  for (int e = tid; e < BS * BS; e += 256) {
    const int r = e / BS, t = e % BS;
    const bool in = r0 + r < n && r0 + t < n;
    Ld[t * BS + r] = in ? LU[(long long)(r0 + r) * n + r0 + t] : 0.0;
  }
  if (!LOWER && tid < BS) rdg[tid] = r0 + tid < n ? 1.0 / LU[(long long)(r0 + tid) * n + r0 + tid] : 1.0;
  __syncthreads();  // sbad is initialised before any wave can set it
  double acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = 0.0;
  v4d cacc[4];  // NR = 16: the wave's partial sums as four 16 x 16 MFMA tiles (rows 16 rt ..)
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) cacc[rt] = v4d{0.0, 0.0, 0.0, 0.0};
  bool bad = false;
  const double* Lrow = LU + (long long)row * n;
  const int ndep = LOWER ? blk : nb - 1 - blk;
  for (int s = 0; s < ndep; ++s) {
    const int j = LOWER ? s : nb - 1 - s;  // blocks in the order they are published
    if (!dl_sf_wait(&flags[j])) {
      bad = true;
      break;
    }
    const int jr = j * BS + lane;
#pragma unroll
    for (int c = 0; c < NR; ++c) xw[lane * NR + c] = (jr < n && c < nc) ? Xv[(long long)jr * nc + c] : 0.0;
    if constexpr (NR == 16) {
      // -L[block rows, this wave's 16 columns] x_j on FP64 MFMA (16 x 16 x 4):
      // A = -L (lane: row lane % 16 of tile rt, column lane / 16 of k-step
      // s4), B = x_j (row lane / 16 of the k-step, right-hand side lane % 16)
      const int ar = lane & 15, ak = lane >> 4;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const int kl = 16 * w + 4 * s4 + ak;
        const int kc = min(j * BS + kl, n - 1);  // columns past n meet zero rows of x_j
        const double bx = xw[kl * 16 + ar];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          const double a = -LU[(long long)min(r0 + 16 * rt + ar, n - 1) * n + kc];
          cacc[rt] = mfma_f64(a, bx, cacc[rt]);
        }
      }
    } else {
      double lv[16];
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) lv[kk] = Lrow[min(j * BS + 16 * w + kk, n - 1)];
#pragma unroll
      for (int kk = 0; kk < 16; ++kk)
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] = fma(-lv[kk], xw[(16 * w + kk) * NR + c], acc[c]);
    }
  }
  if (bad) sbad = 1;
  // the wave's partial sums replace its staged block (same wave: in order)
  if constexpr (NR == 16) {
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) xw[(16 * rt + mfma16_row_f64(lane, r)) * 16 + (lane & 15)] = cacc[rt][r];
  } else {
#pragma unroll
    for (int c = 0; c < NR; ++c) xw[lane * NR + c] = acc[c];
  }
  __syncthreads();
  if (NR <= 2 && w != 0) return;
  const int cb = NR <= 2 ? 0 : w * CW;  // this wave's first right-hand side
  double x[CW];
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    const int cc = cb + c;
    const double b = cc < nc ? Bv[(long long)row * nc + cc] : 0.0;
    x[c] = b + xs[lane * NR + cc] + xs[BS * NR + lane * NR + cc] + xs[2 * BS * NR + lane * NR + cc] +
           xs[3 * BS * NR + lane * NR + cc];
  }
  if (LOWER) {
#pragma unroll
    for (int t = 0; t < BS - 1; ++t) {
      const double l = Ld[t * BS + lane];
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        const double xt = dl_bcast(x[c], t);
        if (lane > t) x[c] = fma(-l, xt, x[c]);
      }
    }
  } else {
#pragma unroll
    for (int t = BS - 1; t >= 0; --t) {
      const double u = Ld[t * BS + lane], rt = rdg[t];
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        const double xt = dl_bcast(x[c], t) * rt;
        if (lane == t) x[c] = xt;
        else if (lane < t) x[c] = fma(-u, xt, x[c]);
      }
    }
  }
  const bool failed = sbad != 0;
  if (r0 + lane < n) {
#pragma unroll
    for (int c = 0; c < CW; ++c)
      if (cb + c < nc) Xv[(long long)(r0 + lane) * nc + cb + c] = failed ? __builtin_nan("") : x[c];
  }
  // publish x_blk once every solving wave's stores are done: each such wave
  // releases its stores, then the barrier, then one flag store (release)
  if constexpr (NR > 2) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
  }
  if (tid == 0) __hip_atomic_store(&flags[blk], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NR>
static int dl_trsv_sf(const double* LU, int n, int nc, double* X, double* W, int* flags, hipStream_t st) {
  const int nb = (n + DL_SF_BS - 1) / DL_SF_BS;
  const size_t shm = dl_sf_lds<NR>();
  auto kf = dl_trsv_sf_kernel<NR, true>;
  auto kb = dl_trsv_sf_kernel<NR, false>;
  for (auto k : {kf, kb}) {
    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("dense_lu_solve: hipFuncSetAttribute: %s", hipGetErrorString(e)), JFX_EHIP;
  }
  hipLaunchKernelGGL(kf, dim3(nb), dim3(256), shm, st, LU, n, nc, X, W, flags);
  JFX_LAUNCH_CHECK();
  hipLaunchKernelGGL(kb, dim3(nb), dim3(256), shm, st, LU, n, nc, W, X, flags + nb);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

// Rows kb .. kb + nbs - 1 of X: X_k <- T_kk^-1 X_k with T_kk the unit lower
// (LOWER) or upper (else) diagonal block of LU; one thread per right-hand
// side, the block in LDS.
// One block step of the substitution, fused: every workgroup solves the
// diagonal block T_kk x_k = Tb[block k] for its 64 right-hand sides (in
// registers; the unit lower L_kk (LOWER) or U_kk), workgroup row 0 writes x_k
// to Out, and workgroup row r >= 1 updates its 64 rows of Tb below (LOWER)
// or above the block: Tb[rows] -= T[rows, block k] x_k.  Tb's block-k rows
// are only read in this launch, so the redundant solves agree and no
// workgroup waits for another.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void dl_solve_step_kernel(const T* __restrict__ LU, int n, int kb, int nbs,
                                                             T* __restrict__ Tb, T* __restrict__ Out,
                                                             long long nrhs) {
  constexpr int NBS = dl_nbs<T>();
  __shared__ T D[NBS][NBS + 1];
  __shared__ T ys[NBS][65];
  __shared__ T Ls[64][NBS + 1];
  const int tid = threadIdx.x;
  const long long c0 = (long long)blockIdx.x * 64;
  const int rt = blockIdx.y;
  for (int e = tid; e < nbs * nbs; e += 256)
    D[e / nbs][e % nbs] = LU[(long long)(kb + e / nbs) * n + kb + e % nbs];
  // rows this workgroup updates: below the block (LOWER) or above it
  const int r0 = LOWER ? kb + nbs + (rt - 1) * 64 : (rt - 1) * 64;
  const int rend = LOWER ? n : kb;
  if (rt > 0)
    for (int e = tid; e < 64 * nbs; e += 256) {
      const int r = e / nbs, k = e % nbs;
      Ls[r][k] = r0 + r < rend ? LU[(long long)(r0 + r) * n + kb + k] : dl_zero<T>();
    }
  __syncthreads();
  if (tid < 64) {
    const long long c = c0 + tid;
    T x[NBS];
#pragma unroll
    for (int t = 0; t < NBS; ++t) x[t] = (t < nbs && c < nrhs) ? Tb[(long long)(kb + t) * nrhs + c] : dl_zero<T>();
    if (LOWER) {
#pragma unroll
      for (int i = 1; i < NBS; ++i) {
        if (i < nbs) {
#pragma unroll
          for (int t = 0; t < NBS; ++t)
            if (t < i) x[i] = dl_fms(x[i], D[i][t], x[t]);
        }
      }
    } else {
#pragma unroll
      for (int i = NBS - 1; i >= 0; --i) {
        if (i < nbs) {
#pragma unroll
          for (int t = 0; t < NBS; ++t)
            if (t > i && t < nbs) x[i] = dl_fms(x[i], D[i][t], x[t]);
          x[i] = dl_div(x[i], D[i][i]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NBS; ++t) {
      ys[t][tid] = x[t];
      if (rt == 0 && t < nbs && c < nrhs) Out[(long long)(kb + t) * nrhs + c] = x[t];
    }
  }
  if (rt == 0) return;
  __syncthreads();
  // Tb[r0 + r][c0 + c] -= sum_k Ls[r][k] ys[k][c]: 4 rows x 4 columns per thread
  const int tr = tid >> 4, tc = tid & 15;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = dl_zero<T>();
  for (int k = 0; k < nbs; ++k) {
    T av[4], bv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) av[a] = Ls[tr + 16 * a][k];
#pragma unroll
    for (int b = 0; b < 4; ++b) bv[b] = ys[k][tc + 16 * b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = dl_fms(acc[a][b], av[a], bv[b]);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int r = r0 + tr + 16 * a;
    if (r >= rend) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const long long c = c0 + tc + 16 * b;
      if (c < nrhs) {
        T* p = Tb + (long long)r * nrhs + c;
        if constexpr (sizeof(T) == sizeof(double))
          *p = *p + acc[a][b];
        else
          *p = zd{p->re + acc[a][b].re, p->im + acc[a][b].im};
      }
    }
  }
}

template <typename T>
static int dl_gemm_sub(const T* A, long long lda, const T* B, long long ldb, T* C, long long ldc, int m, int nc,
                       int k, hipStream_t st) {
  if (m <= 0 || nc <= 0 || k <= 0) return JFX_OK;
  const dim3 grid((unsigned)((nc + 63) / 64), (unsigned)((m + 63) / 64));
  hipLaunchKernelGGL(dl_gemm_sub_kernel<T>, grid, dim3(256), 0, st, A, lda, B, ldb, C, ldc, m, nc, k);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

template <typename T>
static int dl_factor(T* A, int n, int* ipiv, int* perm, int* status, T* S, hipStream_t st) {
  constexpr int NB = dl_nb<T>();
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = n - k0 < NB ? n - k0 : NB;
    const int rows = n - k0;
    if (rows <= 1024)
      hipLaunchKernelGGL((dl_panel_reg_kernel<T, 1, 1024>), dim3(1), dim3(1024), 0, st, A, n, k0, nb, ipiv, status,
                         S);
    else if (rows <= 2048)
      hipLaunchKernelGGL((dl_panel_reg_kernel<T, 2, 1024>), dim3(1), dim3(1024), 0, st, A, n, k0, nb, ipiv, status,
                         S);
    else if (rows <= 4096)  // 512 threads: room for 8 rows of registers each
      hipLaunchKernelGGL((dl_panel_reg_kernel<T, 8, 512>), dim3(1), dim3(512), 0, st, A, n, k0, nb, ipiv, status,
                         S);
    else
      hipLaunchKernelGGL(dl_panel_kernel<T>, dim3(1), dim3(DL_PT), 0, st, A, n, k0, nb, ipiv, status);
    JFX_LAUNCH_CHECK();
    if (n > nb) {
      hipLaunchKernelGGL(dl_rows_kernel<T>, dim3((unsigned)((n - nb + 255) / 256)), dim3(256), 0, st, A, n, k0, nb,
                         ipiv);
      JFX_LAUNCH_CHECK();
    }
    const int m = n - k0 - nb;
    const long long base = (long long)(k0 + nb) * n;
    const int r = dl_gemm_sub<T>(A + base + k0, n, A + (long long)k0 * n + k0 + nb, n, A + base + k0 + nb, n, m, m,
                                 nb, st);
    if (r) return r;
  }
  hipLaunchKernelGGL(dl_perm_kernel, dim3(1), dim3(1024), 0, st, ipiv, n, perm);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

template <typename T>
static int dl_solve(const T* LU, int n, const int* perm, const T* B, T* X, T* W, long long nrhs, hipStream_t st) {
  constexpr int NBS = dl_nbs<T>();
  const long long tot = (long long)n * nrhs;
  const int nb = (n + DL_SF_BS - 1) / DL_SF_BS;
  // flags of the single-launch solve: 2 nb ints after the n x nrhs work values
  int* flags = reinterpret_cast<int*>(W + tot);
  const bool single = std::is_same<T, double>::value && nrhs <= DL_SF_MAXR;
  const long long nthreads = tot > 2LL * nb ? tot : 2LL * nb;
  hipLaunchKernelGGL(dl_gather_kernel<T>, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, st, B, perm, X,
                     n, nrhs, flags, single ? 2 * nb : 0);
  JFX_LAUNCH_CHECK();
  if constexpr (std::is_same<T, double>::value) {
    if (single) {
      const int nc = (int)nrhs;
      if (nc == 1) return dl_trsv_sf<1>(LU, n, nc, X, W, flags, st);
      if (nc == 2) return dl_trsv_sf<2>(LU, n, nc, X, W, flags, st);
      return dl_trsv_sf<16>(LU, n, nc, X, W, flags, st);  // 3..16 columns: the MFMA update
    }
  }
  const unsigned gx = (unsigned)((nrhs + 63) / 64);
  for (int kb = 0; kb < n; kb += NBS) {  // forward, L y = P b: X (updated) -> W
    const int nbs = n - kb < NBS ? n - kb : NBS;
    const unsigned gy = 1 + (unsigned)((n - kb - nbs + 63) / 64);
    hipLaunchKernelGGL((dl_solve_step_kernel<T, true>), dim3(gx, gy), dim3(256), 0, st, LU, n, kb, nbs, X, W, nrhs);
    JFX_LAUNCH_CHECK();
  }
  const int last = (n - 1) / NBS * NBS;
  for (int kb = last; kb >= 0; kb -= NBS) {  // backward, U x = y: W (updated) -> X
    const int nbs = n - kb < NBS ? n - kb : NBS;
    const unsigned gy = 1 + (unsigned)((kb + 63) / 64);
    hipLaunchKernelGGL((dl_solve_step_kernel<T, false>), dim3(gx, gy), dim3(256), 0, st, LU, n, kb, nbs, W, X, nrhs);
    JFX_LAUNCH_CHECK();
  }
  return JFX_OK;
}

}  // namespace jfx

using namespace jfx;

extern "C" {

int jfx_dense_lu_factor(int cplx, void* A, int n, int* ipiv, int* perm, int* status_dev, void* work,
                        hipStream_t stream) {
  if (!A || !ipiv || !perm || !status_dev || !work || work == A || n <= 0 || (cplx != 0 && cplx != 1))
    return set_error("dense_lu_factor: bad arguments"), JFX_EARG;
  if (n > DL_MAXN) return set_error("dense_lu_factor: n = %d exceeds %d", n, DL_MAXN), JFX_ESHAPE;
  return cplx ? dl_factor<zd>((zd*)A, n, ipiv, perm, status_dev, (zd*)work, stream)
              : dl_factor<double>((double*)A, n, ipiv, perm, status_dev, (double*)work, stream);
}

int jfx_dense_lu_solve(int cplx, const void* LU, int n, const int* perm, const void* B, void* X, long long nrhs,
                       void* work, hipStream_t stream) {
  if (!LU || !perm || !B || !X || !work || n <= 0 || nrhs < 0 || B == X || work == X || work == B ||
      (cplx != 0 && cplx != 1))
    return set_error("dense_lu_solve: bad arguments (B, X and work must be distinct)"), JFX_EARG;
  if (n > DL_MAXN) return set_error("dense_lu_solve: n = %d exceeds %d", n, DL_MAXN), JFX_ESHAPE;
  if (nrhs == 0) return JFX_OK;
  if (nrhs > 0x7fffffffLL) return set_error("dense_lu_solve: too many right-hand sides"), JFX_ESHAPE;
  return cplx ? dl_solve<zd>((const zd*)LU, n, perm, (const zd*)B, (zd*)X, (zd*)work, nrhs, stream)
              : dl_solve<double>((const double*)LU, n, perm, (const double*)B, (double*)X, (double*)work, nrhs,
                                 stream);
}

}  // extern "C"
