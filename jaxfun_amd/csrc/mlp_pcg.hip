// Kernels of the learned-preconditioner path:
//   * a BF16 MFMA GEMM with FP32 accumulation and a fused bias + tanh
//     epilogue (forward and backward of the coefficient-space MLP),
//   * tanh backward, column sums, Adam,
//   * batched PCG reductions and updates (one system per batch entry,
//     per-system scalars kept on the device), elementwise exp.
#include "jfx_common.h"

namespace jfx {

// C (M x N) = alpha * op(A) op(B) + beta * C (+ bias[n]) (then tanh if act == 1)
// op(A) = A (M x K, lda) if TA == 0 else A^T with A stored (K x M, lda)
// op(B) = B (K x N, ldb) if TB == 0 else B^T with B stored (N x K, ldb)
// The products are small (M, N, K of a few hundred: the MLP layers), so the
// kernel is built for latency and fill rather than operand reuse: one
// 32 x 32 output tile per workgroup, its four waves splitting K into
// interleaved 32-wide chunks.  Each wave loads its MFMA fragments straight
// from global memory (k-contiguous operands as 16-byte loads, the others as
// 64-byte rows across 16 lanes) with up to four chunks in flight before the
// first MFMA, rounds them to bf16 in registers and accumulates 2 x 2 blocks
// of 16x16x32 in FP32.  The four partial tiles are summed through LDS in a
// fixed order and each wave finishes one 16 x 16 block with the epilogue.
constexpr int GT = 32, GK = 32, GCH = 4;  // tile, k chunk, chunks in flight per wave

// Eight consecutive k of row r of an operand viewed as (rows x K):
// TR == false: X[r ld + k] (k contiguous), TR == true: X[k ld + r].
// Entries outside (rows x K) read as zero.  VEC (TR == false only): K % 8 == 0,
// ld % 4 == 0 and X 16-byte aligned, so a fragment is all in or all out.
template <bool TR, bool VEC>
__device__ __forceinline__ void gemm_frag(const float* __restrict__ X, int ld, int rows, int K, int r, int k,
                                          float (&v)[8]) {
  if (!TR && VEC) {
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
    if (r < rows && k < K) {
      const float4* p = reinterpret_cast<const float4*>(X + (long long)r * ld + k);
      a = p[0];
      b = p[1];
    }
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const bool in = r < rows && k + e < K;
      v[e] = in ? (TR ? X[(long long)(k + e) * ld + r] : X[(long long)r * ld + k + e]) : 0.f;
    }
  }
}

// 8 fp32 -> 8 bf16 (round to nearest even, hardware packed conversion).
__device__ __forceinline__ v8u16 pack_bf16x8(const float (&v)[8]) {
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  u4 w;
#pragma unroll
  for (int e = 0; e < 4; ++e) w[e] = __builtin_bit_cast(unsigned, __builtin_convertvector((f2){v[2 * e], v[2 * e + 1]}, bf2));
  return __builtin_bit_cast(v8u16, w);
}

template <int TA, int TB, bool VEC>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, float alpha,
                                                        const float* __restrict__ A, int lda,
                                                        const float* __restrict__ B, int ldb, float beta,
                                                        float* __restrict__ C, int ldc,
                                                        const float* __restrict__ bias, int act) {
  __shared__ v4f red[4][4][64];  // [wave][16 x 16 block][lane]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int rl = lane & 15, kq = 8 * (lane >> 4);
  const int nch = (K + GK - 1) / GK;
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0, 0, 0, 0};

  for (int c0 = w; c0 < nch; c0 += 4 * GCH) {
    float av[GCH][2][8], bv[GCH][2][8];
#pragma unroll
    for (int u = 0; u < GCH; ++u) {
      const int k = (c0 + 4 * u) * GK + kq;  // chunks past K load zeros
#pragma unroll
      for (int i = 0; i < 2; ++i) gemm_frag<TA == 1, VEC>(A, lda, M, K, m0 + 16 * i + rl, k, av[u][i]);
#pragma unroll
      for (int j = 0; j < 2; ++j) gemm_frag<TB == 0, VEC>(B, ldb, N, K, n0 + 16 * j + rl, k, bv[u][j]);
    }
#pragma unroll
    for (int u = 0; u < GCH; ++u) {
      if (c0 + 4 * u < nch) {  // uniform per wave
        v8u16 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = pack_bf16x8(av[u][i]);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = pack_bf16x8(bv[u][j]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = mfma_bf16(a[i], b[j], acc[i][j]);
      }
    }
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) red[w][b][lane] = acc[b >> 1][b & 1];
  __syncthreads();
  v4f s = red[0][w][lane];
#pragma unroll
  for (int v = 1; v < 4; ++v) s += red[v][w][lane];
  const int n = n0 + 16 * (w & 1) + rl;
  if (n >= N) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = m0 + 16 * (w >> 1) + mfma16_row_f32(lane, r);
    if (m >= M) continue;
    float v = alpha * s[r];
    if (beta != 0.f) v += beta * C[(long long)m * ldc + n];
    if (bias) v += bias[n];
    if (act == 1) v = tanhf(v);
    C[(long long)m * ldc + n] = v;
  }
}

// dZ = dY * (1 - H^2)
__global__ void tanh_backward_kernel(const float* __restrict__ dY, const float* __restrict__ H,
                                     float* __restrict__ dZ, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) dZ[i] = dY[i] * (1.f - H[i] * H[i]);
}

// out[j] = sum_i X[i, j]   (rows x cols): 64 columns per workgroup, 16 waves
// over interleaved rows, partial sums combined in LDS in a fixed order.
__global__ __launch_bounds__(1024) void colsum_kernel(const float* __restrict__ X, float* __restrict__ out,
                                                      int rows, int cols) {
  __shared__ float part[16][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + tx;
  float s = 0.f;
  if (j < cols) {
#pragma unroll 8
    for (int i = ty; i < rows; i += 16) s += X[(long long)i * cols + j];
  }
  part[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && j < cols) {
    float t = part[0][tx];
#pragma unroll
    for (int r = 1; r < 16; ++r) t += part[r][tx];
    out[j] = t;
  }
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
  const double r = rr[b];
  if (!isfinite(r)) {  // breakdown: the system stops and is marked with iters = -1
    active[b] = 0;
    iters[b] = -1;
    return;
  }
  if (count) iters[b] += 1;
  if (!(r > tol2 * bb[b])) active[b] = 0;
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
  if (ta < 0 || ta > 1 || tb < 0 || tb > 1) return set_error("gemm_bf16: ta, tb must be 0 or 1"), JFX_EARG;
  // 16-byte fragment loads of the k-contiguous operands
  const bool va = ta == 1 || (lda % 4 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0);
  const bool vb = tb == 0 || (ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0);
  const bool vec = K % 8 == 0 && va && vb;
  dim3 g((N + GT - 1) / GT, (M + GT - 1) / GT);
#define JFX_GEMM(TA, TB, V)                                                                                 \
  hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, V>), g, dim3(256), 0, stream, M, N, K, alpha, A, lda, B, ldb, \
                     beta, C, ldc, bias, act)
  const int sel = ta * 4 + tb * 2 + (vec ? 1 : 0);
  switch (sel) {
    case 0: JFX_GEMM(0, 0, false); break;
    case 1: JFX_GEMM(0, 0, true); break;
    case 2: JFX_GEMM(0, 1, false); break;
    case 3: JFX_GEMM(0, 1, true); break;
    case 4: JFX_GEMM(1, 0, false); break;
    case 5: JFX_GEMM(1, 0, true); break;
    case 6: JFX_GEMM(1, 1, false); break;
    default: JFX_GEMM(1, 1, true); break;
  }
#undef JFX_GEMM
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
  hipLaunchKernelGGL(colsum_kernel, dim3((cols + 63) / 64), dim3(1024), 0, stream, X, out, rows, cols);
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
