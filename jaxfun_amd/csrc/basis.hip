// Basis evaluation in FP64 on the GPU: Gauss nodes/weights and the
// (derivative) pseudo-Vandermonde matrices of Legendre / Chebyshev bases, of
// any basis defined by a three-term recurrence (the Jacobi family) and of the
// Fourier exponentials.
#include "jfx_common.h"

namespace jfx {

enum { FAM_LEGENDRE = 0, FAM_CHEBYSHEV = 1 };

// One thread per node; Newton on P_N with the three-term recurrence.
// Nodes ascending; weights 2 / ((1-x)(1+x) P_N'(x)^2).
__global__ void gauss_legendre_kernel(int N, double* __restrict__ x, double* __restrict__ w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double pi = 3.14159265358979323846;
  // i-th node in ascending order
  double xi = -cos(pi * (i + 0.75) / (N + 0.5));
  double p0 = 1.0, p1 = xi;
  for (int it = 0; it < 100; ++it) {
    p0 = 1.0;
    p1 = xi;
    for (int n = 1; n < N; ++n) {
      double p2 = ((2 * n + 1) * xi * p1 - n * p0) / (n + 1);
      p0 = p1;
      p1 = p2;
    }
    double dp = N * (xi * p1 - p0) / (xi * xi - 1.0);
    double dx = p1 / dp;
    xi -= dx;
    if (fabs(dx) < 1e-16) break;
  }
  p0 = 1.0;
  p1 = xi;
  for (int n = 1; n < N; ++n) {
    double p2 = ((2 * n + 1) * xi * p1 - n * p0) / (n + 1);
    p0 = p1;
    p1 = p2;
  }
  double omx2 = (1.0 - xi) * (1.0 + xi);
  double dp = N * (p0 - xi * p1) / omx2;
  x[i] = xi;
  w[i] = 2.0 / (omx2 * dp * dp);
}

__global__ void gauss_chebyshev_kernel(int N, double* __restrict__ x, double* __restrict__ w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double pi = 3.14159265358979323846;
  x[i] = cos(pi + (2 * i + 1) * pi / (2 * N));
  w[i] = pi / N;
}

// V[q, n] = d^k/dx^k psi_n(x_q); one thread per point, k <= 4.
// Differentiated recurrences:
//   Legendre   P^(d)_{n+1} = P^(d)_{n-1} + (2n+1) P^(d-1)_n  (d >= 1)
//   Chebyshev  T^(d)_{n+1} = 2x T^(d)_n + 2d T^(d-1)_n - T^(d)_{n-1}
__global__ void vandermonde_kernel(int family, const double* __restrict__ x, int Q, int N, int k,
                                   double scale, double* __restrict__ V) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  const double X = x[q];
  double prev[5], cur[5];
  for (int d = 0; d <= 4; ++d) {
    prev[d] = (d == 0) ? 1.0 : 0.0;  // psi_0
    cur[d] = (d == 0) ? X : (d == 1 ? 1.0 : 0.0);  // psi_1
  }
  double* row = V + (long long)q * N;
  row[0] = prev[k] * scale;
  if (N > 1) row[1] = cur[k] * scale;
  for (int n = 1; n < N - 1; ++n) {
    double nxt[5];
    if (family == FAM_LEGENDRE) {
      nxt[0] = ((2 * n + 1) * X * cur[0] - n * prev[0]) / (n + 1);
      for (int d = 1; d <= k; ++d) nxt[d] = prev[d] + (2 * n + 1) * cur[d - 1];
    } else {
      nxt[0] = 2.0 * X * cur[0] - prev[0];
      for (int d = 1; d <= k; ++d) nxt[d] = 2.0 * X * cur[d] + 2.0 * d * cur[d - 1] - prev[d];
    }
    for (int d = 0; d <= k; ++d) {
      prev[d] = cur[d];
      cur[d] = nxt[d];
    }
    row[n + 1] = cur[k] * scale;
  }
}

// V[q, n] = scale * d^k/dx^k Q_n(x_q) for any basis given by its three-term
// recurrence (Jacobi P^(a,b)_n scaled by g_n: Jacobi, ChebyshevU,
// Ultraspherical): Q_0 = 1 and, for n >= 1,
//     Q_n = r0[n] ((x - r1[n]) Q_{n-1} - r2[n] Q_{n-2})      (r2[1] unused)
// with rec = [r0 | r1 | r2], each N long.  Differentiating d times:
//     Q^(d)_n = r0[n] ((x - r1[n]) Q^(d)_{n-1} + d Q^(d-1)_{n-1} - r2[n] Q^(d)_{n-2}).
// One thread per point, k <= 4.
__global__ void recurrence_vandermonde_kernel(const double* __restrict__ x, int Q, int N, int k,
                                              const double* __restrict__ rec, double scale,
                                              double* __restrict__ V) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  const double X = x[q];
  const double* r0 = rec;
  const double* r1 = rec + N;
  const double* r2 = rec + 2 * N;
  double prev[5], cur[5];
  for (int d = 0; d <= 4; ++d) {
    prev[d] = 0.0;
    cur[d] = (d == 0) ? 1.0 : 0.0;  // Q_0
  }
  double* row = V + (long long)q * N;
  row[0] = cur[k] * scale;
  for (int n = 1; n < N; ++n) {
    const double a = r0[n], b = X - r1[n], c = (n >= 2) ? r2[n] : 0.0;
    double nxt[5];
    nxt[0] = a * (b * cur[0] - c * prev[0]);
    for (int d = 1; d <= k; ++d) nxt[d] = a * (b * cur[d] + d * cur[d - 1] - c * prev[d]);
    for (int d = 0; d <= k; ++d) {
      prev[d] = cur[d];
      cur[d] = nxt[d];
    }
    row[n] = cur[k] * scale;
  }
}

// out[c, r] = V[r, c] * (rs ? rs[r] : 1) * (cs ? cs[c] : 1) * s   (V is R x C)
__global__ void scale_transpose_kernel(const double* __restrict__ V, int R, int C,
                                       const double* __restrict__ rs, const double* __restrict__ cs,
                                       double s, double* __restrict__ out) {
  __shared__ double tile[32][33];
  int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    int r = r0 + j, c = c0 + tx;
    tile[j][tx] = (r < R && c < C) ? V[(long long)r * C + c] * (rs ? rs[r] : 1.0) : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    int c = c0 + j, r = r0 + tx;
    if (r < R && c < C) out[(long long)c * R + r] = tile[tx][j] * (cs ? cs[c] : 1.0) * s;
  }
}

}  // namespace jfx

using namespace jfx;

// Complex exponential basis of the N-mode Fourier space, wavenumbers in FFT
// order (k_j = j for j < (N+1)/2, j - N otherwise; for odd derivative orders
// of even N the Nyquist mode is zeroed).  Element (q, j):
//     (i k_j s)^d exp(i k_j x_q) * post,
// with x_q = 2 pi q / Q (grid != 0; the phase k_j q is reduced exactly mod Q)
// or x_q = x[q].  adjoint == 0 writes out (Q, N); adjoint == 1 writes the
// conjugate transpose (N, Q).  out is complex, re/im interleaved.
__global__ void fourier_basis_kernel(const double* __restrict__ x, int Q, int N, int d, double s, int grid,
                                     int adjoint, double post, double* __restrict__ out) {
  const long long total = (long long)Q * N;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  int q, j;
  if (adjoint) {
    j = (int)(idx / Q);
    q = (int)(idx - (long long)j * Q);
  } else {
    q = (int)(idx / N);
    j = (int)(idx - (long long)q * N);
  }
  const int kj = j < (N + 1) / 2 ? j : j - N;
  double kd = (double)kj;
  if ((d & 1) && (N % 2 == 0) && j == N / 2) kd = 0.0;
  double ang;
  if (grid) {
    long long r = ((long long)kj * q) % Q;
    if (r < 0) r += Q;
    ang = 2.0 * M_PI * (double)r / (double)Q;
  } else {
    ang = (double)kj * x[q];
  }
  double sn, cs;
  sincos(ang, &sn, &cs);
  double mag = post;
  for (int e = 0; e < d; ++e) mag *= kd * s;
  double fr = 0.0, fi = 0.0;
  switch (d & 3) {
    case 0: fr = mag; break;
    case 1: fi = mag; break;
    case 2: fr = -mag; break;
    default: fi = -mag; break;
  }
  const double re = fr * cs - fi * sn, im = fr * sn + fi * cs;
  out[2 * idx] = re;
  out[2 * idx + 1] = adjoint ? -im : im;
}

extern "C" {

int jfx_gauss_nodes(int family, int N, double* x, double* w, hipStream_t stream) {
  if (!x || !w || N <= 0) return set_error("gauss_nodes: bad arguments"), JFX_EARG;
  dim3 g((N + 255) / 256), b(256);
  if (family == FAM_LEGENDRE)
    hipLaunchKernelGGL(gauss_legendre_kernel, g, b, 0, stream, N, x, w);
  else if (family == FAM_CHEBYSHEV)
    hipLaunchKernelGGL(gauss_chebyshev_kernel, g, b, 0, stream, N, x, w);
  else
    return set_error("gauss_nodes: unknown family %d", family), JFX_EARG;
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_vandermonde_f64(int family, const double* x, int Q, int N, int k, double scale, double* V,
                        hipStream_t stream) {
  if (!x || !V || Q <= 0 || N <= 0 || k < 0 || k > 4 || (family != 0 && family != 1))
    return set_error("vandermonde: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(vandermonde_kernel, dim3((Q + 127) / 128), dim3(128), 0, stream, family, x, Q, N,
                     k, scale, V);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_vandermonde_rec_f64(const double* x, int Q, int N, int k, const double* rec, double scale, double* V,
                            hipStream_t stream) {
  if (!x || !rec || !V || Q <= 0 || N <= 0 || k < 0 || k > 4)
    return set_error("vandermonde_rec: bad arguments"), JFX_EARG;
  hipLaunchKernelGGL(recurrence_vandermonde_kernel, dim3((Q + 127) / 128), dim3(128), 0, stream, x, Q, N, k,
                     rec, scale, V);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_fourier_basis(const double* x, int Q, int N, int d, double s, int grid, int adjoint, double post,
                      double* out, hipStream_t stream) {
  if (!out || Q <= 0 || N <= 0 || d < 0 || d > 8 || (!grid && !x)) return set_error("fourier_basis: bad arguments"), JFX_EARG;
  const long long total = (long long)Q * N;
  hipLaunchKernelGGL(fourier_basis_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, x, Q, N, d, s,
                     grid, adjoint, post, out);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_scale_transpose(const double* V, int R, int C, const double* rowscale, const double* colscale,
                        double s, double* out, hipStream_t stream) {
  if (!V || !out || R <= 0 || C <= 0) return set_error("scale_transpose: bad arguments"), JFX_EARG;
  dim3 g((C + 31) / 32, (R + 31) / 32), b(32, 8);
  hipLaunchKernelGGL(scale_transpose_kernel, g, b, 0, stream, V, R, C, rowscale, colscale, s, out);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // extern "C"
