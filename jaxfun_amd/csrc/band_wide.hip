// One large banded system (the Kronecker matrix of banded factors) and the
// dense Kronecker assembly:
//   * band_lu_factor_wide: Doolittle LU without pivoting of one (n x n) band
//     matrix (band storage by rows, as in banded.hip: band[i W + o + p] =
//     A[i, i + o], W = p + q + 1) by one 1024-thread workgroup: step k
//     stages row k's U part and the p multipliers of column k in LDS, then
//     the p x q trailing window is updated in parallel;
//   * band_lu_solve_wide: forward and back substitution, one wave per
//     right-hand side with the vector in LDS; step k updates the next p (q)
//     entries in parallel, and the factor entries of step k + 1 are fetched
//     during step k;
//   * kron_sum: out = sum_t s_t kron(X_t, Y_t) for square X_t, Y_t;
//   * gather_rows: Y[b, i] = X[b, perm[i]] (reordered right-hand sides).
#include "jfx_common.h"

namespace jfx {

constexpr int BWT = 1024;        // factor threads
constexpr int BW_MAX = 1024;     // p, q limit (LDS staging of one step)

__global__ __launch_bounds__(BWT) void band_lu_factor_wide_kernel(const double* __restrict__ band, int n, int p, int q,
                                                                  double* __restrict__ lu, int* __restrict__ status) {
  __shared__ double urow[BW_MAX];  // U[k, k + 1 .. k + q]
  __shared__ double lcol[BW_MAX];  // L[k + 1 .. k + p, k]
  __shared__ int s_fail;
  const int tid = threadIdx.x;
  const int W = p + q + 1;
  const long long tot = (long long)n * W;
  for (long long e = tid; e < tot; e += BWT) lu[e] = band[e];
  if (tid == 0) s_fail = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double piv = lu[(long long)k * W + p];
    if (piv == 0.0 || !isfinite(piv)) {  // every thread reads the same pivot: uniform exit
      if (tid == 0) *status = k + 1;
      return;
    }
    const int pe = min(p, n - 1 - k), qe = min(q, n - 1 - k);
    for (int j = tid; j < qe; j += BWT) urow[j] = lu[(long long)k * W + p + 1 + j];
    for (int i = tid; i < pe; i += BWT) {
      double* l = lu + (long long)(k + 1 + i) * W + (p - 1 - i);  // A[k + 1 + i, k]
      const double v = *l / piv;
      *l = v;
      lcol[i] = v;
    }
    __syncthreads();
    const int cnt = pe * qe;
    for (int e = tid; e < cnt; e += BWT) {
      const int i = e / qe, j = e - i * qe;  // row k + 1 + i, column k + 1 + j
      lu[(long long)(k + 1 + i) * W + (p + j - i)] -= lcol[i] * urow[j];
    }
    __syncthreads();
  }
  if (tid == 0) *status = 0;
}

// One wave per right-hand side; X is (nrhs, n) with rows contiguous.
__global__ __launch_bounds__(64) void band_lu_solve_wide_kernel(const double* __restrict__ lu, int n, int p, int q,
                                                                double* __restrict__ X) {
  extern __shared__ double xs[];
  const int lane = threadIdx.x;
  const int W = p + q + 1;
  double* x = X + (long long)blockIdx.x * n;
  for (int i = lane; i < n; i += 64) xs[i] = x[i];
  __syncthreads();
  // forward: x[k + i] -= L[k + i, k] x[k], i = 1..p  (lanes take i = lane + 1 + 64 r)
  const int R = (p + 63) / 64;
  double lnext[16];
  auto lfetch = [&](int k, double(&dst)[16]) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = lane + 1 + 64 * r;
      dst[r] = (r < R && i <= p && k + i < n) ? lu[(long long)(k + i) * W + (p - i)] : 0.0;
    }
  };
  lfetch(0, lnext);
  for (int k = 0; k < n; ++k) {
    double lc[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) lc[r] = lnext[r];
    if (k + 1 < n) lfetch(k + 1, lnext);  // independent of x: in flight during this step
    const double yk = xs[k];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = lane + 1 + 64 * r;
      if (r < R && i <= p && k + i < n) xs[k + i] -= lc[r] * yk;
    }
    __syncthreads();
  }
  // backward: u[k] = x[k] / U[k, k]; x[k - i] -= U[k - i, k] u[k], i = 1..q
  const int Rq = (q + 63) / 64;
  double unext[16], dnext;
  auto ufetch = [&](int k, double(&dst)[16], double& d) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = lane + 1 + 64 * r;
      dst[r] = (r < Rq && i <= q && k - i >= 0) ? lu[(long long)(k - i) * W + (p + i)] : 0.0;
    }
    d = lu[(long long)k * W + p];
  };
  ufetch(n - 1, unext, dnext);
  for (int k = n - 1; k >= 0; --k) {
    double uc[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) uc[r] = unext[r];
    const double dk = dnext;
    if (k > 0) ufetch(k - 1, unext, dnext);
    const double uk = xs[k] / dk;
    __syncthreads();  // every lane has read xs[k] before lane 0 overwrites it
    if (lane == 0) xs[k] = uk;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = lane + 1 + 64 * r;
      if (r < Rq && i <= q && k - i >= 0) xs[k - i] -= uc[r] * uk;
    }
    __syncthreads();
  }
  for (int i = lane; i < n; i += 64) x[i] = xs[i];
}

// out (r m x r m) = sum_t s[t] kron(X_t, Y_t), X (nt, r, r), Y (nt, m, m), s on the device.
__global__ void kron_sum_kernel(const double* __restrict__ Xs, const double* __restrict__ Ys,
                                const double* __restrict__ s, int nt, int r, int m, double* __restrict__ out) {
  const long long N = (long long)r * m;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * N) return;
  const long long row = e / N, col = e - row * N;
  const int i0 = (int)(row / m), i1 = (int)(row - (long long)i0 * m);
  const int j0 = (int)(col / m), j1 = (int)(col - (long long)j0 * m);
  double acc = 0.0;
  for (int t = 0; t < nt; ++t)
    acc = fma(s[t] * Xs[((long long)t * r + i0) * r + j0], Ys[((long long)t * m + i1) * m + j1], acc);
  out[e] = acc;
}

__global__ void gather_rows_kernel(const double* __restrict__ X, const long long* __restrict__ perm, int n,
                                   long long nrhs, double* __restrict__ Y) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * nrhs) return;
  const long long b = e / n, i = e - b * n;
  Y[e] = X[b * n + perm[i]];
}

}  // namespace jfx

using namespace jfx;

extern "C" {

int jfx_band_lu_factor_wide(const double* band, int n, int p, int q, double* lu, int* status_dev,
                            hipStream_t stream) {
  if (!band || !lu || !status_dev || n <= 0 || p < 0 || q < 0)
    return set_error("band_lu_factor_wide: bad arguments"), JFX_EARG;
  if (p > BW_MAX || q > BW_MAX) return set_error("band_lu_factor_wide: bandwidths above %d", BW_MAX), JFX_ESHAPE;
  hipLaunchKernelGGL(band_lu_factor_wide_kernel, dim3(1), dim3(BWT), 0, stream, band, n, p, q, lu, status_dev);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_solve_wide(const double* lu, int n, int p, int q, double* X, long long nrhs, hipStream_t stream) {
  if (!lu || !X || n <= 0 || p < 0 || q < 0 || nrhs < 0) return set_error("band_lu_solve_wide: bad arguments"), JFX_EARG;
  if (p > 16 * 64 || q > 16 * 64) return set_error("band_lu_solve_wide: bandwidths above 1024"), JFX_ESHAPE;
  const size_t shm = (size_t)n * sizeof(double);
  if (shm > 160 * 1024) return set_error("band_lu_solve_wide: n = %d exceeds the LDS vector (20480)", n), JFX_ESHAPE;
  if (nrhs == 0) return JFX_OK;
  hipError_t e = hipFuncSetAttribute((const void*)band_lu_solve_wide_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) return set_error("band_lu_solve_wide: hipFuncSetAttribute: %s", hipGetErrorString(e)), JFX_EHIP;
  hipLaunchKernelGGL(band_lu_solve_wide_kernel, dim3((unsigned)nrhs), dim3(64), shm, stream, lu, n, p, q, X);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_kron_sum(const double* Xs, const double* Ys, const double* s, int nt, int r, int m, double* out,
                 hipStream_t stream) {
  if (!Xs || !Ys || !s || !out || nt <= 0 || r <= 0 || m <= 0) return set_error("kron_sum: bad arguments"), JFX_EARG;
  const long long N = (long long)r * m, tot = N * N;
  if (tot == 0) return JFX_OK;
  hipLaunchKernelGGL(kron_sum_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, Xs, Ys, s, nt, r, m,
                     out);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_gather_rows(const double* X, const long long* perm, int n, long long nrhs, double* Y, hipStream_t stream) {
  if (!X || !perm || !Y || n <= 0 || nrhs < 0) return set_error("gather_rows: bad arguments"), JFX_EARG;
  const long long tot = (long long)n * nrhs;
  if (tot == 0) return JFX_OK;
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, X, perm, n, nrhs,
                     Y);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // extern "C"
