// Kernels of the learned-preconditioner path:
//   * a BF16 MFMA GEMM with FP32 accumulation and a fused bias + tanh
//     epilogue (forward and backward of the coefficient-space MLP),
//   * tanh backward, column sums, Adam,
//   * batched PCG reductions and updates (one system per batch entry,
//     per-system scalars kept on the device), elementwise exp.
#include "jfx_common.h"

namespace jfx {

// C (M x N) = alpha * op(A) op(B) + beta * C (+ bias[n]) (then tanh if act == 1)
// op(A) = A (M x K, lda) if ta == 0 else A^T with A stored (K x M, lda)
// op(B) = B (K x N, ldb) if tb == 0 else B^T with B stored (N x K, ldb)
// 64 x 64 tile per workgroup, 4 waves in 2 x 2, each 32 x 32 = 2 x 2 MFMA
// 16x16x32 blocks; operands rounded to bf16 while staged in LDS.
constexpr int GT = 64, GK = 32, GLD = GK + 8;

__global__ __launch_bounds__(256) void gemm_bf16_kernel(int ta, int tb, int M, int N, int K, float alpha,
                                                        const float* __restrict__ A, int lda,
                                                        const float* __restrict__ B, int ldb, float beta,
                                                        float* __restrict__ C, int ldc,
                                                        const float* __restrict__ bias, int act) {
  __shared__ __attribute__((aligned(16))) unsigned short As[GT][GLD];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[GT][GLD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0, 0, 0, 0};

  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      int r, k;
      if (ta == 0) {  // A row-major M x K: k fastest for coalescing
        r = e / GK;
        k = e - r * GK;
      } else {        // A stored K x M: m fastest
        k = e / GT;
        r = e - k * GT;
      }
      int m = m0 + r, kk = k0 + k;
      float v = 0.f;
      if (m < M && kk < K) v = (ta == 0) ? A[(long long)m * lda + kk] : A[(long long)kk * lda + m];
      As[r][k] = f2bf(v);
    }
    for (int e = tid; e < GT * GK; e += 256) {
      int c, k;
      if (tb == 0) {  // B row-major K x N: n fastest
        k = e / GT;
        c = e - k * GT;
      } else {        // B stored N x K: k fastest
        c = e / GK;
        k = e - c * GK;
      }
      int n = n0 + c, kk = k0 + k;
      float v = 0.f;
      if (n < N && kk < K) v = (tb == 0) ? B[(long long)kk * ldb + n] : B[(long long)n * ldb + kk];
      Bs[c][k] = f2bf(v);
    }
    __syncthreads();
    v8u16 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *(const v8u16*)&As[wm + 16 * i + (lane & 15)][8 * (lane >> 4)];
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *(const v8u16*)&Bs[wn + 16 * j + (lane & 15)][8 * (lane >> 4)];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma_bf16(a[i], b[j], acc[i][j]);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      int n = n0 + wn + 16 * j + (lane & 15);
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = m0 + wm + 16 * i + mfma16_row_f32(lane, r);
        if (m >= M) continue;
        float v = alpha * acc[i][j][r];
        if (beta != 0.f) v += beta * C[(long long)m * ldc + n];
        if (bias) v += bias[n];
        if (act == 1) v = tanhf(v);
        C[(long long)m * ldc + n] = v;
      }
    }
}

// dZ = dY * (1 - H^2)
__global__ void tanh_backward_kernel(const float* __restrict__ dY, const float* __restrict__ H,
                                     float* __restrict__ dZ, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) dZ[i] = dY[i] * (1.f - H[i] * H[i]);
}

// out[j] = sum_i X[i, j]   (rows x cols), one thread per column
__global__ void colsum_kernel(const float* __restrict__ X, float* __restrict__ out, int rows, int cols) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (int i = 0; i < rows; ++i) s += X[(long long)i * cols + j];
  out[j] = s;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) {
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// out[b] = sum_i x[b, i] y[b, i]   (FP64), one workgroup per batch entry
__global__ void batched_dot_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                   double* __restrict__ out, long long n) {
  __shared__ double red[256];
  const long long b = blockIdx.x;
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) s += x[b * n + i] * y[b * n + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = red[0];
}

// x += a p, r -= a q with a = num[b] / den[b] for active systems
__global__ void pcg_xr_kernel(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                              const double* __restrict__ q, const double* __restrict__ num,
                              const double* __restrict__ den, const int* __restrict__ active, long long n) {
  const long long b = blockIdx.y;
  if (!active[b]) return;
  const double a = num[b] / den[b];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[b * n + i] += a * p[b * n + i];
    r[b * n + i] -= a * q[b * n + i];
  }
}

// p = z + (num[b] / den[b]) p for active systems
__global__ void pcg_p_kernel(double* __restrict__ p, const double* __restrict__ z, const double* __restrict__ num,
                             const double* __restrict__ den, const int* __restrict__ active, long long n) {
  const long long b = blockIdx.y;
  if (!active[b]) return;
  const double beta = num[b] / den[b];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[b * n + i] = z[b * n + i] + beta * p[b * n + i];
}

// active[b] &= rr[b] > tol2 * bb[b]; iters[b] += (was active)
__global__ void pcg_check_kernel(const double* __restrict__ rr, const double* __restrict__ bb, double tol2,
                                 int* __restrict__ active, int* __restrict__ iters, int batch, int count) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  if (!active[b]) return;
  if (count) iters[b] += 1;
  if (!(rr[b] > tol2 * bb[b])) active[b] = 0;
}

// y[b, i] = x[b, i] * s[b]
__global__ void scale_batch_kernel(const double* __restrict__ x, const double* __restrict__ s,
                                   double* __restrict__ y, long long n) {
  const long long b = blockIdx.y;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[b * n + i] = x[b * n + i] * s[b];
}

__global__ void exp_kernel(const double* __restrict__ x, double* __restrict__ y, double scale, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) y[i] = exp(scale * x[i]);
}

static unsigned g1(long long n) {
  long long g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace jfx

using namespace jfx;

extern "C" {

int jfx_gemm_bf16(int ta, int tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B,
                  int ldb, float beta, float* C, int ldc, const float* bias, int act, hipStream_t stream) {
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0 || act < 0 || act > 1)
    return set_error("gemm_bf16: bad arguments"), JFX_EARG;
  if ((ta == 0 && lda < K) || (ta == 1 && lda < M) || (tb == 0 && ldb < N) || (tb == 1 && ldb < K) || ldc < N)
    return set_error("gemm_bf16: bad leading dimension"), JFX_ESHAPE;
  dim3 g((N + GT - 1) / GT, (M + GT - 1) / GT);
  hipLaunchKernelGGL(gemm_bf16_kernel, g, dim3(256), 0, stream, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C,
                     ldc, bias, act);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_tanh_backward(const float* dY, const float* H, float* dZ, long long n, hipStream_t stream) {
  if (!dY || !H || !dZ || n < 0) return set_error("tanh_backward: bad arguments"), JFX_EARG;
  if (n == 0) return JFX_OK;
  hipLaunchKernelGGL(tanh_backward_kernel, dim3(g1(n)), dim3(256), 0, stream, dY, H, dZ, n);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_colsum(const float* X, float* out, int rows, int cols, hipStream_t stream) {
  if (!X || !out || rows < 0 || cols <= 0) return set_error("colsum: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(colsum_kernel, dim3((cols + 255) / 256), dim3(256), 0, stream, X, out, rows, cols);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
             int step, hipStream_t stream) {
  if (!p || !g || !m || !v || n < 0 || step < 1) return set_error("adam: bad arguments"), JFX_EARG;
  if (n == 0) return JFX_OK;
  float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(g1(n)), dim3(256), 0, stream, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_batched_dot(const double* x, const double* y, double* out, int batch, long long n, hipStream_t stream) {
  if (!x || !y || !out || batch < 0 || n < 0) return set_error("batched_dot: bad arguments"), JFX_EARG;
  if (batch == 0) return JFX_OK;
  hipLaunchKernelGGL(batched_dot_kernel, dim3(batch), dim3(256), 0, stream, x, y, out, n);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_pcg_update_xr(double* x, double* r, const double* p, const double* q, const double* num,
                      const double* den, const int* active, int batch, long long n, hipStream_t stream) {
  if (!x || !r || !p || !q || !num || !den || !active || batch < 0 || batch > 65535)
    return set_error("pcg_update_xr: bad arguments"), JFX_EARG;
  if (batch == 0 || n == 0) return JFX_OK;
  dim3 g((unsigned)((n + 255) / 256 > 64 ? 64 : (n + 255) / 256), batch);
  hipLaunchKernelGGL(pcg_xr_kernel, g, dim3(256), 0, stream, x, r, p, q, num, den, active, n);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_pcg_update_p(double* p, const double* z, const double* num, const double* den, const int* active,
                     int batch, long long n, hipStream_t stream) {
  if (!p || !z || !num || !den || !active || batch < 0 || batch > 65535)
    return set_error("pcg_update_p: bad arguments"), JFX_EARG;
  if (batch == 0 || n == 0) return JFX_OK;
  dim3 g((unsigned)((n + 255) / 256 > 64 ? 64 : (n + 255) / 256), batch);
  hipLaunchKernelGGL(pcg_p_kernel, g, dim3(256), 0, stream, p, z, num, den, active, n);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_pcg_check(const double* rr, const double* bb, double tol2, int* active, int* iters, int batch, int count,
                  hipStream_t stream) {
  if (!rr || !bb || !active || !iters || batch < 0) return set_error("pcg_check: bad arguments"), JFX_EARG;
  if (batch == 0) return JFX_OK;
  hipLaunchKernelGGL(pcg_check_kernel, dim3((batch + 255) / 256), dim3(256), 0, stream, rr, bb, tol2, active, iters,
                     batch, count);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_scale_batch(const double* x, const double* s, double* y, int batch, long long n, hipStream_t stream) {
  if (!x || !s || !y || batch < 0 || batch > 65535) return set_error("scale_batch: bad arguments"), JFX_EARG;
  if (batch == 0 || n == 0) return JFX_OK;
  dim3 g((unsigned)((n + 255) / 256 > 64 ? 64 : (n + 255) / 256), batch);
  hipLaunchKernelGGL(scale_batch_kernel, g, dim3(256), 0, stream, x, s, y, n);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_exp(const double* x, double* y, double scale, long long n, hipStream_t stream) {
  if (!x || !y || n < 0) return set_error("exp: bad arguments"), JFX_EARG;
  if (n == 0) return JFX_OK;
  hipLaunchKernelGGL(exp_kernel, dim3(g1(n)), dim3(256), 0, stream, x, y, scale, n);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // extern "C"
