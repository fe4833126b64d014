// Banded LU without pivoting (DIA matrices) and batched FP64 solves along an
// axis, plus the elementwise kernels the solvers and integrators need.
//
// Band storage ("by row"): band[i * W + (o + p)] = A[i, i + o] for offsets
// o in [-p, q], W = p + q + 1.  The factorisation overwrites a copy with L
// (unit lower, multipliers at o < 0) and U (o >= 0) in the same layout.
#include "jfx_common.h"

namespace jfx {

// Doolittle factorisation of one banded matrix by one lane; returns 0 or
// 1 + the row of the first zero / non-finite pivot.
__device__ __forceinline__ int band_lu_factor_one(const double* __restrict__ band, int n, int p, int q,
                                  double* __restrict__ lu) {
  const int W = p + q + 1;
  for (long long i = 0; i < (long long)n * W; ++i) lu[i] = band[i];
  for (int k = 0; k < n - 1; ++k) {
    double piv = lu[(long long)k * W + p];
    if (piv == 0.0 || !isfinite(piv)) return k + 1;
    int iend = min(n - 1, k + p);
    for (int i = k + 1; i <= iend; ++i) {
      double& lik = lu[(long long)i * W + (k - i + p)];
      lik /= piv;
      int jend = min(n - 1, k + q);
      for (int j = k + 1; j <= jend; ++j)
        lu[(long long)i * W + (j - i + p)] -= lik * lu[(long long)k * W + (j - k + p)];
    }
  }
  double piv = lu[(long long)(n - 1) * W + p];
  return (piv == 0.0 || !isfinite(piv)) ? n : 0;
}

// One small banded matrix (setup).
__global__ void band_lu_factor_kernel(const double* __restrict__ band, int n, int p, int q,
                                      double* __restrict__ lu, int* __restrict__ status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  *status = band_lu_factor_one(band, n, p, q, lu);
}

// nsys independent banded matrices (band (nsys, n, W)), one lane each; the
// first failure's 1 + row lands in *status.
__global__ void band_lu_factor_batched_kernel(const double* __restrict__ band, int nsys, int n, int p, int q,
                                              double* __restrict__ lu, int* __restrict__ status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const long long off = (long long)s * n * (p + q + 1);
  const int st = band_lu_factor_one(band + off, n, p, q, lu + off);
  if (st) atomicCAS(status, 0, st);
}

// Forward / back substitution of one line x (stride Q) with one factor.
// For bandwidths <= RW the last solved values stay in a register window, so
// the recurrence carries no memory round trip (the right-hand side and factor
// loads do not depend on it and are issued ahead).
constexpr int RW = 8;
__device__ __forceinline__ void band_lu_solve_one(const double* __restrict__ lu, int n, int p, int q, double* __restrict__ x,
                                  long long Q) {
  const int W = p + q + 1;
  if (p <= RW && q <= RW) {
    // rows go in chunks of RC: the chunk's right-hand sides and factor rows
    // are loaded together (one memory latency per chunk), then the chunk's
    // recurrence runs from registers
    constexpr int RC = 8;
    double w[RW];
#pragma unroll
    for (int t = 0; t < RW; ++t) w[t] = 0.0;
    for (int i0 = 0; i0 < n; i0 += RC) {
      double xv[RC], cf[RC][RW];
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const int i = i0 + c;
        xv[c] = 0.0;
#pragma unroll
        for (int t = 0; t < RW; ++t) cf[c][t] = 0.0;
        if (i < n) {
          const double* row = lu + (long long)i * W;
          xv[c] = x[(long long)i * Q];
#pragma unroll
          for (int t = 0; t < RW; ++t)
            if (t < p && i - 1 - t >= 0) cf[c][t] = row[p - 1 - t];
        }
      }
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const int i = i0 + c;
        if (i < n) {
          double s = xv[c];
#pragma unroll
          for (int t = 0; t < RW; ++t) s -= cf[c][t] * w[t];
#pragma unroll
          for (int t = RW - 1; t > 0; --t) w[t] = w[t - 1];
          w[0] = s;
          x[(long long)i * Q] = s;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < RW; ++t) w[t] = 0.0;
    for (int i0 = n - 1; i0 >= 0; i0 -= RC) {
      double xv[RC], dg[RC], cf[RC][RW];
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const int i = i0 - c;
        xv[c] = 0.0;
        dg[c] = 1.0;
#pragma unroll
        for (int t = 0; t < RW; ++t) cf[c][t] = 0.0;
        if (i >= 0) {
          const double* row = lu + (long long)i * W;
          xv[c] = x[(long long)i * Q];
          dg[c] = row[p];
#pragma unroll
          for (int t = 0; t < RW; ++t)
            if (t < q && i + 1 + t < n) cf[c][t] = row[p + 1 + t];
        }
      }
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const int i = i0 - c;
        if (i >= 0) {
          double s = xv[c];
#pragma unroll
          for (int t = 0; t < RW; ++t) s -= cf[c][t] * w[t];
          s /= dg[c];
#pragma unroll
          for (int t = RW - 1; t > 0; --t) w[t] = w[t - 1];
          w[0] = s;
          x[(long long)i * Q] = s;
        }
      }
    }
    return;
  }
  for (int i = 1; i < n; ++i) {
    double s = x[(long long)i * Q];
    int j0 = max(0, i - p);
    for (int j = j0; j < i; ++j) s -= lu[(long long)i * W + (j - i + p)] * x[(long long)j * Q];
    x[(long long)i * Q] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[(long long)i * Q];
    int j1 = min(n - 1, i + q);
    for (int j = i + 1; j <= j1; ++j) s -= lu[(long long)i * W + (j - i + p)] * x[(long long)j * Q];
    x[(long long)i * Q] = s / lu[(long long)i * W + p];
  }
}

// Solve L U x = b in place along axis of length n of a (P, n, Q) batch; one
// thread per line (coalesced over q).
__global__ __launch_bounds__(256) void band_lu_solve_kernel(const double* __restrict__ lu, int n, int p, int q,
                                     double* __restrict__ X, long long P, long long Q) {
  long long line = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= P * Q) return;
  long long pp = line / Q, qq = line - pp * Q;
  band_lu_solve_one(lu, n, p, q, X + pp * (long long)n * Q + qq, Q);
}

// As above with one factor per line group: line (pp, qq) uses system
// (pp * (Q / qdiv) + qq / qdiv) mod nsys (qdiv = 2 keeps the re/im pair of a
// complex value on one system; the modulus repeats the systems over leading
// batch dimensions).
__global__ __launch_bounds__(256) void band_lu_solve_batched_kernel(const double* __restrict__ lu, int nsys, int n, int p, int q,
                                             double* __restrict__ X, long long P, long long Q, int qdiv) {
  long long line = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= P * Q) return;
  long long pp = line / Q, qq = line - pp * Q;
  const long long sys = (pp * (Q / qdiv) + qq / qdiv) % nsys;
  band_lu_solve_one(lu + sys * (long long)n * (p + q + 1), n, p, q, X + pp * (long long)n * Q + qq, Q);
}

// ---- complex (re/im interleaved) batched banded LU for per-wavenumber
// systems with complex weights (odd derivative orders on a Fourier axis)
struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  const double d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

__global__ void band_lu_factor_batched_z_kernel(const cplx* __restrict__ band, int nsys, int n, int p, int q,
                                                cplx* __restrict__ lu, int* __restrict__ status) {
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nsys) return;
  const int W = p + q + 1;
  const long long off = (long long)sidx * n * W;
  const cplx* b = band + off;
  cplx* a = lu + off;
  for (long long i = 0; i < (long long)n * W; ++i) a[i] = b[i];
  int st = 0;
  for (int k = 0; k < n && st == 0; ++k) {
    const cplx piv = a[(long long)k * W + p];
    if ((piv.re == 0.0 && piv.im == 0.0) || !isfinite(piv.re) || !isfinite(piv.im)) {
      st = k + 1;
      break;
    }
    const int iend = min(n - 1, k + p);
    for (int i = k + 1; i <= iend; ++i) {
      cplx& lik = a[(long long)i * W + (k - i + p)];
      lik = cdiv(lik, piv);
      const int jend = min(n - 1, k + q);
      for (int j = k + 1; j <= jend; ++j) {
        const cplx t = cmul(lik, a[(long long)k * W + (j - k + p)]);
        cplx& aij = a[(long long)i * W + (j - i + p)];
        aij.re -= t.re;
        aij.im -= t.im;
      }
    }
  }
  if (st) atomicCAS(status, 0, st);
}

// Line (pp, qq) of X (P, n, Q) complex uses factor (pp * Q + qq) mod nsys.
__global__ __launch_bounds__(256) void band_lu_solve_batched_z_kernel(const cplx* __restrict__ lu, int nsys, int n,
                                                                      int p, int q, cplx* __restrict__ X, long long P,
                                                                      long long Q) {
  const long long line = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= P * Q) return;
  const long long pp = line / Q, qq = line - pp * Q;
  const int W = p + q + 1;
  const cplx* a = lu + ((pp * Q + qq) % nsys) * (long long)n * W;
  cplx* x = X + pp * (long long)n * Q + qq;
  for (int i = 1; i < n; ++i) {
    cplx s = x[(long long)i * Q];
    for (int j = max(0, i - p); j < i; ++j) {
      const cplx t = cmul(a[(long long)i * W + (j - i + p)], x[(long long)j * Q]);
      s.re -= t.re;
      s.im -= t.im;
    }
    x[(long long)i * Q] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    cplx s = x[(long long)i * Q];
    for (int j = i + 1; j <= min(n - 1, i + q); ++j) {
      const cplx t = cmul(a[(long long)i * W + (j - i + p)], x[(long long)j * Q]);
      s.re -= t.re;
      s.im -= t.im;
    }
    x[(long long)i * Q] = cdiv(s, a[(long long)i * W + p]);
  }
}

// Y[b, i] = X[b, i] * D[i], complex, D broadcast over the leading batch
__global__ void cmul_bcast_kernel(const cplx* __restrict__ X, const cplx* __restrict__ D, cplx* __restrict__ Y,
                                  long long total, long long size) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < total; i += stride) Y[i] = cmul(X[i], D[i % size]);
}

// Y[b, i] = X[b, i] / D[i]  (D broadcast over the leading batch)
template <typename T>
__global__ void div_bcast_kernel(const T* __restrict__ X, const T* __restrict__ D, T* __restrict__ Y,
                                 long long total, long long size) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < total; i += stride) Y[i] = X[i] / D[i % size];
}

template <typename T>
__global__ void mul_bcast_kernel(const T* __restrict__ X, const T* __restrict__ C, T* __restrict__ Y,
                                 long long total, long long size) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < total; i += stride) Y[i] = X[i] * C[i % size];
}

// Y[b, i] = X[b, i] + alpha * D[i]  (D broadcast over the leading batch)
__global__ void axpy_bcast_kernel(const double* __restrict__ X, const double* __restrict__ D, double alpha,
                                  double* __restrict__ Y, long long total, long long size) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < total; i += stride) Y[i] = X[i] + alpha * D[i % size];
}

// Complex result of a stacked real contraction.  Z = [Ar; Ai] applied along
// an axis: Z (P, 2M, Qr) with Qr = 2Q (complex input, re/im interleaved along
// the last axis) or Q (real input); Y (P, M, Q) complex, re/im interleaved.
__global__ void cplx_combine_kernel(const double* __restrict__ Z, double* __restrict__ Y, long long P,
                                    long long M, long long Q, int real_input) {
  const long long total = P * M * Q;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    const long long q = i % Q, pm = i / Q, m = pm % M, p = pm / M;
    double re, im;
    if (real_input) {
      const double* zr = Z + (p * 2 * M + m) * Q + q;
      re = zr[0];
      im = zr[M * Q];
    } else {
      const double* zr = Z + (p * 2 * M + m) * 2 * Q + 2 * q;
      const double* zi = zr + M * 2 * Q;
      re = zr[0] - zi[1];
      im = zr[1] + zi[0];
    }
    Y[2 * i] = re;
    Y[2 * i + 1] = im;
  }
}

// out = sum_t c_t * A_t * B_t  (pointwise, nterms <= 4)
template <typename T>
struct SumProd {
  const T* a[4];
  const T* b[4];
  double c[4];
};
template <typename T>
__global__ void sum_products_kernel(SumProd<T> sp, int nterms, T* __restrict__ out, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    T s = 0;
    for (int t = 0; t < nterms; ++t) s += (T)sp.c[t] * sp.a[t][i] * sp.b[t][i];
    out[i] = s;
  }
}

// y = sum_j c_j x_j (n <= 8 inputs), optionally accumulating into y.
template <typename T>
struct LinComb {
  const T* x[8];
  double c[8];
};
template <typename T>
__global__ void lincomb_kernel(LinComb<T> lc, int nin, T* __restrict__ y, long long n, int accumulate) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    double s = accumulate ? (double)y[i] : 0.0;
    for (int j = 0; j < nin; ++j) s += lc.c[j] * (double)lc.x[j][i];
    y[i] = (T)s;
  }
}

template <typename TI, typename TO>
__global__ void cast_kernel(const TI* __restrict__ x, TO* __restrict__ y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = (TO)x[i];
}

static unsigned grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace jfx

using namespace jfx;

extern "C" {

int jfx_band_lu_factor(const double* band, int n, int p, int q, double* lu, int* status_dev,
                       hipStream_t stream) {
  if (!band || !lu || !status_dev || n <= 0 || p < 0 || q < 0)
    return set_error("band_lu_factor: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(band_lu_factor_kernel, dim3(1), dim3(64), 0, stream, band, n, p, q, lu, status_dev);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_solve(const double* lu, int n, int p, int q, double* X, long long P, long long Q,
                      hipStream_t stream) {
  if (!lu || !X || n <= 0 || p < 0 || q < 0 || P < 0 || Q <= 0)
    return set_error("band_lu_solve: bad arguments"), JFX_EARG;
  long long lines = P * Q;
  if (lines == 0) return JFX_OK;
  hipLaunchKernelGGL(band_lu_solve_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, stream, lu,
                     n, p, q, X, P, Q);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_factor_batched(const double* band, int nsys, int n, int p, int q, double* lu, int* status_dev,
                               hipStream_t stream) {
  if (!band || !lu || !status_dev || nsys <= 0 || n <= 0 || p < 0 || q < 0)
    return set_error("band_lu_factor_batched: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(band_lu_factor_batched_kernel, dim3((unsigned)((nsys + 63) / 64)), dim3(64), 0, stream, band,
                     nsys, n, p, q, lu, status_dev);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_solve_batched(const double* lu, int nsys, int n, int p, int q, double* X, long long P, long long Q,
                              int qdiv, hipStream_t stream) {
  if (!lu || !X || nsys <= 0 || n <= 0 || p < 0 || q < 0 || P < 0 || Q <= 0 || qdiv < 1 || Q % qdiv ||
      (P * (Q / qdiv)) % nsys)
    return set_error("band_lu_solve_batched: bad arguments (%d systems for P=%lld, Q=%lld, qdiv=%d)", nsys, P, Q,
                     qdiv),
           JFX_EARG;
  long long lines = P * Q;
  if (lines == 0) return JFX_OK;
  hipLaunchKernelGGL(band_lu_solve_batched_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, stream,
                     lu, nsys, n, p, q, X, P, Q, qdiv);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_factor_batched_z(const double* band, int nsys, int n, int p, int q, double* lu, int* status_dev,
                                 hipStream_t stream) {
  if (!band || !lu || !status_dev || nsys <= 0 || n <= 0 || p < 0 || q < 0)
    return set_error("band_lu_factor_batched_z: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(band_lu_factor_batched_z_kernel, dim3((unsigned)((nsys + 63) / 64)), dim3(64), 0, stream,
                     (const cplx*)band, nsys, n, p, q, (cplx*)lu, status_dev);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_solve_batched_z(const double* lu, int nsys, int n, int p, int q, double* X, long long P, long long Q,
                                hipStream_t stream) {
  if (!lu || !X || nsys <= 0 || n <= 0 || p < 0 || q < 0 || P < 0 || Q <= 0 || (P * Q) % nsys)
    return set_error("band_lu_solve_batched_z: bad arguments"), JFX_EARG;
  const long long lines = P * Q;
  if (lines == 0) return JFX_OK;
  hipLaunchKernelGGL(band_lu_solve_batched_z_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, stream,
                     (const cplx*)lu, nsys, n, p, q, (cplx*)X, P, Q);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_cmul_bcast(const double* X, const double* D, double* Y, long long total, long long size,
                   hipStream_t stream) {
  if (!X || !D || !Y || size <= 0 || total % size) return set_error("cmul_bcast: bad arguments"), JFX_EARG;
  if (total == 0) return JFX_OK;
  hipLaunchKernelGGL(cmul_bcast_kernel, dim3(grid_for(total, 256)), dim3(256), 0, stream, (const cplx*)X,
                     (const cplx*)D, (cplx*)Y, total, size);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_div_bcast_f64(const double* X, const double* D, double* Y, long long total, long long size,
                      hipStream_t stream) {
  if (!X || !D || !Y || size <= 0 || total % size) return set_error("div_bcast: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(div_bcast_kernel<double>, dim3(grid_for(total, 256)), dim3(256), 0, stream, X, D, Y,
                     total, size);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_div_bcast_f32(const float* X, const float* D, float* Y, long long total, long long size,
                      hipStream_t stream) {
  if (!X || !D || !Y || size <= 0 || total % size) return set_error("div_bcast: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(div_bcast_kernel<float>, dim3(grid_for(total, 256)), dim3(256), 0, stream, X, D, Y,
                     total, size);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_mul_bcast_f64(const double* X, const double* C, double* Y, long long total, long long size,
                      hipStream_t stream) {
  if (!X || !C || !Y || size <= 0 || total % size) return set_error("mul_bcast: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(mul_bcast_kernel<double>, dim3(grid_for(total, 256)), dim3(256), 0, stream, X, C, Y,
                     total, size);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_axpy_bcast_f64(const double* X, const double* D, double alpha, double* Y, long long total,
                       long long size, hipStream_t stream) {
  if (!X || !D || !Y || size <= 0 || total % size) return set_error("axpy_bcast: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(axpy_bcast_kernel, dim3(grid_for(total, 256)), dim3(256), 0, stream, X, D, alpha, Y,
                     total, size);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_cplx_combine(const double* Z, double* Y, long long P, long long M, long long Q, int real_input,
                     hipStream_t stream) {
  if (!Z || !Y || P < 0 || M < 0 || Q < 0 || (real_input != 0 && real_input != 1))
    return set_error("cplx_combine: bad arguments"), JFX_EARG;
  if (P * M * Q == 0) return JFX_OK;
  hipLaunchKernelGGL(cplx_combine_kernel, dim3(grid_for(P * M * Q, 256)), dim3(256), 0, stream, Z, Y, P, M, Q,
                     real_input);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

// out = sum_t coeffs[t] * As[t] * Bs[t]; As, Bs, coeffs are HOST arrays (nterms <= 4)
int jfx_sum_products(int dtype, int nterms, const void* const* As, const void* const* Bs,
                     const double* coeffs, void* out, long long n, hipStream_t stream) {
  if (nterms < 1 || nterms > 4 || !As || !Bs || !coeffs || !out) return set_error("sum_products: bad arguments"), JFX_EARG;
  if (n == 0) return JFX_OK;
  if (dtype == 0) {
    SumProd<double> sp{};
    for (int t = 0; t < nterms; ++t) {
      sp.a[t] = (const double*)As[t];
      sp.b[t] = (const double*)Bs[t];
      sp.c[t] = coeffs[t];
    }
    hipLaunchKernelGGL(sum_products_kernel<double>, dim3(grid_for(n, 256)), dim3(256), 0, stream, sp, nterms,
                       (double*)out, n);
  } else if (dtype == 1) {
    SumProd<float> sp{};
    for (int t = 0; t < nterms; ++t) {
      sp.a[t] = (const float*)As[t];
      sp.b[t] = (const float*)Bs[t];
      sp.c[t] = coeffs[t];
    }
    hipLaunchKernelGGL(sum_products_kernel<float>, dim3(grid_for(n, 256)), dim3(256), 0, stream, sp, nterms,
                       (float*)out, n);
  } else {
    return set_error("sum_products: bad dtype"), JFX_EARG;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

// direction 0: f64 -> f32 (round to nearest), 1: f32 -> f64
int jfx_cast(int direction, const void* x, void* y, long long n, hipStream_t stream) {
  if (!x || !y || n < 0) return set_error("cast: bad arguments"), JFX_EARG;
  if (n == 0) return JFX_OK;
  if (direction == 0)
    hipLaunchKernelGGL((cast_kernel<double, float>), dim3(grid_for(n, 256)), dim3(256), 0, stream,
                       (const double*)x, (float*)y, n);
  else if (direction == 1)
    hipLaunchKernelGGL((cast_kernel<float, double>), dim3(grid_for(n, 256)), dim3(256), 0, stream,
                       (const float*)x, (double*)y, n);
  else
    return set_error("cast: bad direction"), JFX_EARG;
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

// dtype: 0 = f64, 1 = f32
int jfx_lincomb(int dtype, int nin, const void* const* xs, const double* coeffs, void* y, long long n,
                int accumulate, hipStream_t stream) {
  if (nin < 0 || nin > 8 || !y || (nin > 0 && (!xs || !coeffs)))
    return set_error("lincomb: bad arguments"), JFX_EARG;
  if (n == 0) return JFX_OK;
  if (dtype == 0) {
    LinComb<double> lc{};
    for (int j = 0; j < nin; ++j) {
      lc.x[j] = (const double*)xs[j];
      lc.c[j] = coeffs[j];
    }
    hipLaunchKernelGGL(lincomb_kernel<double>, dim3(grid_for(n, 256)), dim3(256), 0, stream, lc, nin,
                       (double*)y, n, accumulate);
  } else if (dtype == 1) {
    LinComb<float> lc{};
    for (int j = 0; j < nin; ++j) {
      lc.x[j] = (const float*)xs[j];
      lc.c[j] = coeffs[j];
    }
    hipLaunchKernelGGL(lincomb_kernel<float>, dim3(grid_for(n, 256)), dim3(256), 0, stream, lc, nin,
                       (float*)y, n, accumulate);
  } else {
    return set_error("lincomb: bad dtype"), JFX_EARG;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // extern "C"
