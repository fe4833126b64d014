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

// The same factorisation for bandwidths p, q <= B with the active part of the
// elimination held in registers: at step k, w[r][c] = A_k[k + r, k + c] for
// r, c in [0, B] (rows below k + p and columns beyond k + q are zero and stay
// zero).  Each step divides out column 0, updates the window, stores row k of
// U and the multipliers of column k, and slides the window by one row and
// column; the entries entering it (column k + 1 + B of rows k + 1 .. k + B and
// row k + 1 + B) have not been touched by any earlier step, so they are read
// from the input.  Those reads are issued one block of FD steps ahead, so the
// loop-carried chain is register arithmetic only (the plain version above
// round-trips every update through memory).  The arithmetic and its order per
// entry are the plain version's.
__device__ __forceinline__ double band_entry(const double* __restrict__ band, int n, int p, int q, int W, int i,
                                             int j) {
  // the load is unconditional (an in-bounds address when the entry is outside
  // the band) so that the memory-counter waits stay exact across the loop
  const int o = j - i;
  const bool in = i < n && j < n && o >= -p && o <= q;
  const double v = band[in ? (long long)i * W + o + p : 0];
  return in ? v : 0.0;
}

template <int B>
__device__ int band_lu_factor_win(const double* __restrict__ band, int n, int p, int q, double* __restrict__ lu) {
  constexpr int FD = B == 2 ? 8 : 4;  // steps per prefetch block (register budget)
  constexpr int NE = 2 * B + 1;       // entries entering the window per step
  const int W = p + q + 1;
  // band-storage slots outside the matrix (first p rows, last q rows) are copied
  for (int i = 0; i < min(n, p); ++i)
    for (int o = -p; o < -i; ++o) lu[(long long)i * W + o + p] = band[(long long)i * W + o + p];
  for (int i = max(0, n - q); i < n; ++i)
    for (int o = n - i; o <= q; ++o) lu[(long long)i * W + o + p] = band[(long long)i * W + o + p];
  double w[B + 1][B + 1];
#pragma unroll
  for (int r = 0; r <= B; ++r)
#pragma unroll
    for (int c = 0; c <= B; ++c) w[r][c] = band_entry(band, n, p, q, W, r, c);
  // e[d]: what the slide after step k0 + d brings in: column k + 1 + B of rows
  // k + 1 .. k + B (e[d][0 .. B-1]), then row k + 1 + B (e[d][B .. 2B])
  double e[FD][NE];
#pragma unroll
  for (int d = 0; d < FD; ++d) {
    const int k = d;
#pragma unroll
    for (int r = 0; r < B; ++r) e[d][r] = band_entry(band, n, p, q, W, k + 1 + r, k + 1 + B);
#pragma unroll
    for (int c = 0; c <= B; ++c) e[d][B + c] = band_entry(band, n, p, q, W, k + 1 + B, k + 1 + c);
  }
  for (int k0 = 0; k0 < n; k0 += FD) {
    double en[FD][NE];
#pragma unroll
    for (int d = 0; d < FD; ++d) {
      const int k = k0 + FD + d;
#pragma unroll
      for (int r = 0; r < B; ++r) en[d][r] = band_entry(band, n, p, q, W, k + 1 + r, k + 1 + B);
#pragma unroll
      for (int c = 0; c <= B; ++c) en[d][B + c] = band_entry(band, n, p, q, W, k + 1 + B, k + 1 + c);
    }
#pragma unroll
    for (int d = 0; d < FD; ++d) {
      const int k = k0 + d;
      if (k >= n) break;
      const double piv = w[0][0];
      if (piv == 0.0 || !isfinite(piv)) return k + 1;
      // Every store is unconditional: one that has no slot (a multiplier row
      // below the band or past n, a U entry past the band or past n) goes to
      // row k's pivot slot, which the last store of the step (c = 0)
      // overwrites with the pivot.
      const long long kp = (long long)k * W + p;
#pragma unroll
      for (int r = 1; r <= B; ++r) {
        const double l = w[r][0] / piv;
        lu[(r <= p && k + r < n) ? (long long)(k + r) * W + p - r : kp] = l;
#pragma unroll
        for (int c = 1; c <= B; ++c) w[r][c] -= l * w[0][c];
      }
#pragma unroll
      for (int c = B; c >= 0; --c) lu[(c <= q && k + c < n) ? kp + c : kp] = w[0][c];
#pragma unroll
      for (int r = 0; r < B; ++r) {
#pragma unroll
        for (int c = 0; c < B; ++c) w[r][c] = w[r + 1][c + 1];
        w[r][B] = e[d][r];
      }
#pragma unroll
      for (int c = 0; c <= B; ++c) w[B][c] = e[d][B + c];
    }
#pragma unroll
    for (int d = 0; d < FD; ++d)
#pragma unroll
      for (int t = 0; t < NE; ++t) e[d][t] = en[d][t];
  }
  return 0;
}

__device__ __forceinline__ int band_lu_factor_any(const double* __restrict__ band, int n, int p, int q,
                                                  double* __restrict__ lu) {
  if (p <= 2 && q <= 2) return band_lu_factor_win<2>(band, n, p, q, lu);
  if (p <= 4 && q <= 4) return band_lu_factor_win<4>(band, n, p, q, lu);
  return band_lu_factor_one(band, n, p, q, lu);
}

// One banded matrix, one lane.  The lane index enters the addresses (it is
// zero) so that the factor's loads stay vector loads, whose completion
// counter is in order: a uniform address would become a scalar load, and the
// scalar counter has to be drained completely before each use.
__global__ __launch_bounds__(64) void band_lu_factor_kernel(const double* __restrict__ band, int n, int p, int q,
                                      double* __restrict__ lu, int* __restrict__ status, int nlanes) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlanes) return;
  const long long off = (long long)t * n * (p + q + 1);
  *status = band_lu_factor_any(band + off, n, p, q, lu + off);
}

// One banded matrix (p, q <= B) by one wave: lane 0 runs the elimination
// window of band_lu_factor_win (with one reciprocal of the pivot per step
// instead of a division per multiplier), but reads the entries entering the
// window from LDS and writes its results to an LDS ring of output rows; the wave
// streams the input in and the results out in chunks of FCH rows with
// consecutive lanes on consecutive addresses (the next chunk's input is
// loaded into registers while lane 0 works).  A single lane's own loads and
// stores would be bound by its memory counter, not by the recurrence.
// FULL: p = q = B, so every window entry lies in the band (no shape masks).
// The reciprocal is the hardware estimate refined by two Newton steps (within
// an ulp of 1 / piv; a correctly rounded division costs about twice the
// instructions on the serial chain).
constexpr int FCH = 64;

__device__ __forceinline__ double band_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

template <int B, bool FULL>
__global__ __launch_bounds__(64) void band_lu_factor_wave_kernel(const double* __restrict__ band, int n, int p,
                                                                 int q, double* __restrict__ lu,
                                                                 int* __restrict__ status) {
  constexpr int WM = 2 * B + 1;            // widest band row
  constexpr int IR = FCH + B + 1;          // input rows per chunk
  constexpr int NIN = (IR * WM + 63) / 64;  // input values per lane and chunk
  __shared__ double inb[2][IR * WM];
  __shared__ double outb[2 * FCH * WM];     // ring: row i at slot i mod 2 FCH
  __shared__ int fail;
  const int lane = threadIdx.x;
  const int W = p + q + 1;
  const long long nW = (long long)n * W;
  // band-storage slots outside the matrix (first p rows, last q rows) are copied
  for (int e = lane; e < p * W; e += 64) {
    const int i = e / W, o = e % W - p;
    if (i < n && i + o < 0) lu[(long long)i * W + o + p] = band[(long long)i * W + o + p];
  }
  for (int e = lane; e < q * W; e += 64) {
    const int i = n - 1 - e / W, o = e % W - p;
    if (i >= 0 && i + o >= n) lu[(long long)i * W + o + p] = band[(long long)i * W + o + p];
  }
  if (lane == 0) fail = 0;
  // input of chunk c: band rows c FCH + 1 .. c FCH + IR (flat, contiguous)
  auto load_in = [&](int c, double (&r)[NIN]) {
    const long long base = ((long long)c * FCH + 1) * W;
#pragma unroll
    for (int t = 0; t < NIN; ++t) {
      const long long e = base + lane + 64 * t;
      const bool ok = lane + 64 * t < IR * W && e < nW;
      const double v = band[ok ? e : 0];
      r[t] = ok ? v : 0.0;
    }
  };
  auto store_in = [&](int buf, const double (&r)[NIN]) {
#pragma unroll
    for (int t = 0; t < NIN; ++t)
      if (lane + 64 * t < IR * W) inb[buf][lane + 64 * t] = r[t];
  };
  double rin[NIN];
  load_in(0, rin);
  store_in(0, rin);
  // lane 0's elimination window: w[r][c] = A_k[k + r, k + c]
  double w[B + 1][B + 1];
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r <= B; ++r)
#pragma unroll
      for (int c = 0; c <= B; ++c) w[r][c] = band_entry(band, n, p, q, W, r, c);
  }
  __syncthreads();
  const int nch = (n + FCH - 1) / FCH;
  for (int c = 0; c < nch; ++c) {
    const int buf = c & 1;
    if (c + 1 < nch) load_in(c + 1, rin);
    if (lane == 0 && !fail) {
      // row k + 1 of the matrix is local row k - c FCH of the chunk's input;
      // input rows past n are zero (load_in), the other entries outside the
      // band are masked here
      const double* pr = inb[buf];
      const int kend = min(n, (c + 1) * FCH);
      // no branch per step: the first zero / non-finite pivot of the chunk is
      // recorded and reported after it (the steps after it only produce
      // non-finite values that are never used), so the loop unrolls and the
      // window's slide becomes register renaming
      int firstbad = 0;
#pragma unroll 3
      for (int k = c * FCH; k < kend; ++k, pr += W) {
        const double piv = w[0][0];
        const bool badp = piv == 0.0 || !isfinite(piv);
        firstbad = (badp && firstbad == 0) ? k + 1 : firstbad;
        // the entering entries, read before the arithmetic
        double ecol[B], erow[B + 1];
        const bool colin = k + 1 + B < n;
#pragma unroll
        for (int r = 0; r < B; ++r) {
          const bool inb_ = FULL || B - r <= q;
          const double v = pr[r * W + (inb_ ? B - r + p : 0)];
          ecol[r] = (inb_ && colin) ? v : 0.0;
        }
#pragma unroll
        for (int cc = 0; cc <= B; ++cc) {
          const bool inb_ = FULL || cc - B >= -p;
          const double v = pr[B * W + (inb_ ? cc - B + p : 0)];
          erow[cc] = inb_ ? v : 0.0;
        }
        const double rp = band_rcp(piv);  // one reciprocal per step; multipliers by it
        double* ob = outb + (k & (2 * FCH - 1)) * WM;
        // ring slots of rows past n are written but never flushed
#pragma unroll
        for (int r = 1; r <= B; ++r) {
          const double l = w[r][0] * rp;
          if (FULL || r <= p) outb[((k + r) & (2 * FCH - 1)) * WM + p - r] = l;
#pragma unroll
          for (int cc = 1; cc <= B; ++cc) w[r][cc] -= l * w[0][cc];
        }
#pragma unroll
        for (int cc = 0; cc <= B; ++cc)
          if (FULL || cc <= q) ob[p + cc] = w[0][cc];
#pragma unroll
        for (int r = 0; r < B; ++r) {
#pragma unroll
          for (int cc = 0; cc < B; ++cc) w[r][cc] = w[r + 1][cc + 1];
          w[r][B] = ecol[r];
        }
#pragma unroll
        for (int cc = 0; cc <= B; ++cc) w[B][cc] = erow[cc];
      }
      if (firstbad) {
        *status = firstbad;
        fail = 1;
      }
    }
    __syncthreads();
    // rows of chunk c are complete: out to memory, in-matrix entries only
    {
      const long long row0 = (long long)c * FCH;
      const int nrow = (int)min((long long)FCH, n - row0);
      for (int e = lane; e < nrow * W; e += 64) {
        const int rr = e / W, o = e % W - p;
        const long long i = row0 + rr;
        if (i + o >= 0 && i + o < n) lu[i * W + o + p] = outb[(i % (2 * FCH)) * WM + o + p];
      }
    }
    if (c + 1 < nch) store_in(buf ^ 1, rin);
    __syncthreads();
    if (fail) break;  // uniform: read after the barrier
  }
  if (lane == 0 && !fail) *status = 0;
}

// nsys independent banded matrices (band (nsys, n, W)), one lane each; the
// first failure's 1 + row lands in *status.
__global__ __launch_bounds__(64) void band_lu_factor_batched_kernel(const double* __restrict__ band, int nsys, int n, int p, int q,
                                              double* __restrict__ lu, int* __restrict__ status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const long long off = (long long)s * n * (p + q + 1);
  const int st = band_lu_factor_any(band + off, n, p, q, lu + off);
  if (st) atomicCAS(status, 0, st);
}

// Forward / back substitution of one line x (stride Q) with one factor.
// For bandwidths <= RWT the last solved values stay in a register window, so
// the recurrence carries no memory round trip.  Rows go in batches of RC: a
// batch's right-hand sides and factor entries are loaded together, from
// clamped in-bounds rows (masked afterwards) so that every load is
// unconditional, and the next batch is loaded before the current one is
// solved, so one memory latency is spread over RC rows and overlapped.
template <int RWT, int RC>
struct LaneRows {
  double xv[RC], dg[RC], cf[RC][RWT];
};

template <int RWT, int RC, bool FWD>
__device__ __forceinline__ void lane_load(const double* __restrict__ lu, const double* __restrict__ x, long long Q,
                                          int n, int W, int p, int q, int i0, LaneRows<RWT, RC>& r) {
#pragma unroll
  for (int c = 0; c < RC; ++c) {
    const int i = FWD ? min(i0 + c, n - 1) : max(i0 - c, 0);
    const double* row = lu + (long long)i * W;
    r.xv[c] = x[(long long)i * Q];
    r.dg[c] = FWD ? 1.0 : row[p];
#pragma unroll
    for (int t = 0; t < RWT; ++t) {
      const bool in = FWD ? (t < p && i - 1 - t >= 0) : (t < q && i + 1 + t < n);
      const double v = row[in ? (FWD ? p - 1 - t : p + 1 + t) : 0];
      r.cf[c][t] = in ? v : 0.0;
    }
  }
}

template <int RWT, int RC, bool FWD>
__device__ __forceinline__ void lane_sweep(const double* __restrict__ lu, int n, int p, int q, double* __restrict__ x,
                                           long long Q) {
  const int W = p + q + 1;
  double w[RWT];
#pragma unroll
  for (int t = 0; t < RWT; ++t) w[t] = 0.0;
  LaneRows<RWT, RC> cur, nxt;
  lane_load<RWT, RC, FWD>(lu, x, Q, n, W, p, q, FWD ? 0 : n - 1, cur);
  for (int b = 0; b < n; b += RC) {
    lane_load<RWT, RC, FWD>(lu, x, Q, n, W, p, q, FWD ? b + RC : n - 1 - b - RC, nxt);
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      if (b + c >= n) break;
      double s = cur.xv[c];
#pragma unroll
      for (int t = 0; t < RWT; ++t) s -= cur.cf[c][t] * w[t];
      if (!FWD) s /= cur.dg[c];
#pragma unroll
      for (int t = RWT - 1; t > 0; --t) w[t] = w[t - 1];
      w[0] = s;
      x[(long long)(FWD ? b + c : n - 1 - b - c) * Q] = s;
    }
    cur = nxt;
  }
}

// Any bandwidth: the plain loops.
__device__ __forceinline__ void band_lu_solve_general(const double* __restrict__ lu, int n, int p, int q,
                                                      double* __restrict__ x, long long Q) {
  const int W = p + q + 1;
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

__device__ __forceinline__ void band_lu_solve_one(const double* __restrict__ lu, int n, int p, int q,
                                                  double* __restrict__ x, long long Q) {
  if (p <= 2 && q <= 2) {
    lane_sweep<2, 8, true>(lu, n, p, q, x, Q);
    lane_sweep<2, 8, false>(lu, n, p, q, x, Q);
  } else if (p <= 4 && q <= 4) {
    lane_sweep<4, 8, true>(lu, n, p, q, x, Q);
    lane_sweep<4, 8, false>(lu, n, p, q, x, Q);
  } else if (p <= 8 && q <= 8) {
    lane_sweep<8, 4, true>(lu, n, p, q, x, Q);
    lane_sweep<8, 4, false>(lu, n, p, q, x, Q);
  } else {
    band_lu_solve_general(lu, n, p, q, x, Q);
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

// ---- wavefronts per line: partitioned substitution
// The 64 * NW lanes of a workgroup split the n rows of one line into
// contiguous chunks of c = ceil(n / (64 NW)).  Forward substitution with the
// unit lower factor and back substitution with the upper factor are linear
// recurrences over a window of the last B solved values (bandwidths p, q <=
// B), so a chunk maps the window entering it to the window leaving it by an
// affine map w -> M w + v.  Pass 1: each lane builds its chunk's map (v from a
// zero window, column j of M from the unit window e_j).  A Kogge-Stone scan
// over the wave composes the maps (six shuffle steps); with several waves, the
// waves' total maps are chained through LDS.  That gives every lane the window
// entering its chunk; pass 2 re-runs the chunk from it and stores the
// solution.  The work is about (B + 2) times the sequential sweep's, spread
// over 64 NW lanes; a lane whose chunk is empty carries the identity map.
template <int B>
__device__ __forceinline__ void wave_push(double (&w)[B], double s) {
#pragma unroll
  for (int t = B - 1; t > 0; --t) w[t] = w[t - 1];
  w[0] = s;
}

template <int B, bool UP>
__device__ __forceinline__ void shfl_map(const double (&M)[B][B], const double (&v)[B], int d, double (&Mp)[B][B],
                                         double (&vp)[B]) {
#pragma unroll
  for (int r = 0; r < B; ++r) {
    vp[r] = UP ? __shfl_up(v[r], d) : __shfl_down(v[r], d);
#pragma unroll
    for (int k = 0; k < B; ++k) Mp[r][k] = UP ? __shfl_up(M[r][k], d) : __shfl_down(M[r][k], d);
  }
}

// (M, v) <- (M, v) o (Mp, vp) = (M Mp, M vp + v): the chunk of (Mp, vp) is
// traversed first.
template <int B>
__device__ __forceinline__ void compose_map(double (&M)[B][B], double (&v)[B], const double (&Mp)[B][B],
                                            const double (&vp)[B]) {
  double nv[B], nM[B][B];
#pragma unroll
  for (int r = 0; r < B; ++r) {
    double s = v[r];
#pragma unroll
    for (int k = 0; k < B; ++k) s += M[r][k] * vp[k];
    nv[r] = s;
#pragma unroll
    for (int k = 0; k < B; ++k) {
      double a = 0.0;
#pragma unroll
      for (int m = 0; m < B; ++m) a += M[r][m] * Mp[m][k];
      nM[r][k] = a;
    }
  }
#pragma unroll
  for (int r = 0; r < B; ++r) {
    v[r] = nv[r];
#pragma unroll
    for (int k = 0; k < B; ++k) M[r][k] = nM[r][k];
  }
}

// Window entering this lane's chunk.  UP: lanes (and waves) are traversed in
// ascending order, else descending.  sm / ss: NW * (B * B + B) and NW * B
// doubles of LDS (unused for NW = 1).
// Chunks of one line spread over several workgroups (the grid solve): the
// window entering this workgroup's rows is the state its predecessor in
// traversal order published (pred_flag null for the line's first chunk), and
// the window leaving them is published for the successor.
struct GridLink {
  const int* pred_flag;
  const double* pred_state;
  int* own_flag;
  double* own_state;
};

constexpr int GRID_SPIN = 1 << 20;  // flag polls (~1 us each with the sleep) before giving up

// Waits (bounded) for a predecessor's flag with acquire ordering at agent scope.
__device__ __forceinline__ bool grid_wait(const int* f) {
  int it = 0;
  while (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    if (++it >= GRID_SPIN) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  return true;
}

template <int B, bool UP, int NW>
__device__ __forceinline__ void line_scan(double (&M)[B][B], double (&v)[B], int lane, int wv, double* sm, double* ss,
                                          double (&win)[B], const GridLink* gl = nullptr) {
  // inclusive scan within the wave
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    double Mp[B][B], vp[B];
    shfl_map<B, UP>(M, v, d, Mp, vp);
    if (UP ? lane >= d : lane + d < 64) compose_map<B>(M, v, Mp, vp);
  }
  // exclusive: the previous lane's inclusive map (identity at the first lane)
  double Me[B][B], ve[B];
  shfl_map<B, UP>(M, v, 1, Me, ve);
  const bool first = UP ? lane == 0 : lane == 63;
  if (first) {
#pragma unroll
    for (int r = 0; r < B; ++r) {
      ve[r] = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) Me[r][k] = r == k ? 1.0 : 0.0;
    }
  }
  double S[B];
#pragma unroll
  for (int r = 0; r < B; ++r) S[r] = 0.0;
  if constexpr (NW > 1) {
    // the wave's last lane in traversal order holds the wave's total map
    if (UP ? lane == 63 : lane == 0) {
#pragma unroll
      for (int r = 0; r < B; ++r) {
        sm[wv * (B * B + B) + B * B + r] = v[r];
#pragma unroll
        for (int k = 0; k < B; ++k) sm[wv * (B * B + B) + r * B + k] = M[r][k];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double st[B];
#pragma unroll
      for (int r = 0; r < B; ++r) st[r] = 0.0;
      if (gl && gl->pred_flag) {
        // a predecessor that never publishes leaves NaN, which the rest of
        // the line inherits, instead of a hang
        const bool ok = grid_wait(gl->pred_flag);
#pragma unroll
        for (int r = 0; r < B; ++r) st[r] = ok ? gl->pred_state[r] : __builtin_nan("");
      }
      for (int t = 0; t < NW; ++t) {
        const int k = UP ? t : NW - 1 - t;
        const double* m = sm + k * (B * B + B);
        double nt[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
          ss[k * B + r] = st[r];
          double a = m[B * B + r];
#pragma unroll
          for (int c = 0; c < B; ++c) a += m[r * B + c] * st[c];
          nt[r] = a;
        }
#pragma unroll
        for (int r = 0; r < B; ++r) st[r] = nt[r];
      }
      if (gl) {
#pragma unroll
        for (int r = 0; r < B; ++r) gl->own_state[r] = st[r];
        __hip_atomic_store(gl->own_flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < B; ++r) S[r] = ss[wv * B + r];
  }
#pragma unroll
  for (int r = 0; r < B; ++r) {
    double a = ve[r];
#pragma unroll
    for (int k = 0; k < B; ++k) a += Me[r][k] * S[k];
    win[r] = a;
  }
}

// A batch of RC consecutive rows of one lane's chunk (in traversal order:
// ascending for the forward pass, descending for the backward pass) with their
// right-hand sides and factor entries.  Every load is issued unconditionally
// from an in-bounds address (rows past the chunk read a clamped row and are
// masked afterwards), so the memory-counter waits stay exact and a batch
// prefetched one step ahead overlaps the current batch.
template <int B, int RC>
struct WaveRows {
  double xv[RC], cf[RC][B], dg[RC];
  bool ok[RC];
};

template <int B, int RC, bool FWD>
__device__ __forceinline__ void wave_load(const double* __restrict__ lu, const double* __restrict__ x, long long Q,
                                          int n, int W, int p, int q, int i0, int i1, int ib, WaveRows<B, RC>& r) {
#pragma unroll
  for (int u = 0; u < RC; ++u) {
    const int i = FWD ? i0 + ib + u : i1 - 1 - ib - u;
    const bool ok = FWD ? i < i1 : i >= i0;
    const int ic = min(max(i, 0), n - 1);
    const double* row = lu + (long long)ic * W;
    r.ok[u] = ok;
    r.xv[u] = x[(long long)ic * Q];
#pragma unroll
    for (int t = 0; t < B; ++t) {
      const bool in = FWD ? (t < p && ic - 1 - t >= 0) : (t < q && ic + 1 + t < n);
      const double v = row[in ? (FWD ? p - 1 - t : p + 1 + t) : 0];
      r.cf[u][t] = in ? v : 0.0;
    }
    r.dg[u] = FWD ? 1.0 : row[p];
  }
}

template <int B>
__device__ __forceinline__ double wave_step(const double (&cf)[B], const double (&w)[B], double rhs, double dg,
                                            bool FWD) {
  double s = rhs;
#pragma unroll
  for (int t = 0; t < B; ++t) s -= cf[t] * w[t];
  return FWD ? s : s / dg;
}

// One direction of the partitioned substitution on this lane's chunk
// [i0, i1): pass 1 (chunk map), scan, pass 2 (solution stored to x).
template <int B, int RC, int NW, bool FWD>
__device__ __forceinline__ void wave_sweep(const double* __restrict__ lu, double* __restrict__ x, long long Q, int n,
                                           int W, int p, int q, int i0, int i1, int lane, int wv, double* sm,
                                           double* ss, const GridLink* gl = nullptr) {
  const int len = i1 - i0;
  double v[B], h[B][B], M[B][B], win[B];
#pragma unroll
  for (int t = 0; t < B; ++t) {
    v[t] = 0.0;
#pragma unroll
    for (int j = 0; j < B; ++j) h[j][t] = j == t ? 1.0 : 0.0;
  }
  WaveRows<B, RC> cur, nxt;
  wave_load<B, RC, FWD>(lu, x, Q, n, W, p, q, i0, i1, 0, cur);
  for (int ib = 0; ib < len; ib += RC) {
    wave_load<B, RC, FWD>(lu, x, Q, n, W, p, q, i0, i1, ib + RC, nxt);
#pragma unroll
    for (int u = 0; u < RC; ++u) {
      if (!cur.ok[u]) break;
      wave_push<B>(v, wave_step<B>(cur.cf[u], v, cur.xv[u], cur.dg[u], FWD));
#pragma unroll
      for (int j = 0; j < B; ++j) wave_push<B>(h[j], wave_step<B>(cur.cf[u], h[j], 0.0, cur.dg[u], FWD));
    }
    cur = nxt;
  }
#pragma unroll
  for (int t = 0; t < B; ++t)
#pragma unroll
    for (int j = 0; j < B; ++j) M[t][j] = h[j][t];
  line_scan<B, FWD, NW>(M, v, lane, wv, sm, ss, win, gl);
  wave_load<B, RC, FWD>(lu, x, Q, n, W, p, q, i0, i1, 0, cur);
  for (int ib = 0; ib < len; ib += RC) {
    wave_load<B, RC, FWD>(lu, x, Q, n, W, p, q, i0, i1, ib + RC, nxt);
#pragma unroll
    for (int u = 0; u < RC; ++u) {
      if (!cur.ok[u]) break;
      const double s = wave_step<B>(cur.cf[u], win, cur.xv[u], cur.dg[u], FWD);
      wave_push<B>(win, s);
      x[(long long)(FWD ? i0 + ib + u : i1 - 1 - ib - u) * Q] = s;
    }
    cur = nxt;
  }
}

// Solve L U x = b in place along the axis of length n of a (P, n, Q) batch;
// NW wavefronts per line, one line per workgroup (grid = P * Q).  p, q <= B.
// With nsys > 1 line (pp, qq) uses factor (pp * (Q / qdiv) + qq / qdiv) mod
// nsys of lu (nsys, n, p + q + 1), as the per-lane batched solve does.
// Each lane streams its own chunk, so the lanes' accesses do not coalesce;
// the chunks are short (n / (64 NW) rows) and the rows stay in L1/L2 between
// the passes.
template <int B, int NW>
__global__ __launch_bounds__(64 * NW) void band_lu_solve_wave_kernel(const double* __restrict__ lu, int n, int p,
                                                                     int q, double* __restrict__ X, long long Q,
                                                                     int nsys, int qdiv) {
  constexpr int RC = (B == 4 && NW > 1) ? 4 : 8;  // 256 VGPRs at 512 threads
  __shared__ double sm[NW * (B * B + B)];
  __shared__ double ss[NW * B];
  const long long line = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long pp = line / Q, qq = line - pp * Q;
  double* __restrict__ x = X + pp * (long long)n * Q + qq;
  const int W = p + q + 1;
  if (nsys > 1) lu += ((pp * (Q / qdiv) + qq / qdiv) % nsys) * (long long)n * W;
  const int c = (n + 64 * NW - 1) / (64 * NW);
  const int i0 = (int)min((long long)n, (long long)threadIdx.x * c);
  const int i1 = (int)min((long long)n, (long long)i0 + c);
  // forward: y_i = b_i - sum_t L[i, i-1-t] y_{i-1-t}
  wave_sweep<B, RC, NW, true>(lu, x, Q, n, W, p, q, i0, i1, lane, wv, sm, ss);
  if constexpr (NW > 1) __syncthreads();  // sm / ss are reused
  // backward: x_i = (y_i - sum_t U[i, i+1+t] x_{i+1+t}) / U[i, i]
  wave_sweep<B, RC, NW, false>(lu, x, Q, n, W, p, q, i0, i1, lane, wv, sm, ss);
}

// The same solve with each line's rows split over G workgroups of 8 waves
// (rows [g R, (g + 1) R) for chunk g), one launch per direction: each
// workgroup runs both passes of its rows and takes the window entering them
// from its predecessor in traversal order through `flags` / `states` (one
// flag and B doubles per line and chunk, flags zeroed before the launch).
// Workgroups are numbered in traversal order (the backward launch numbers the
// chunks from the end), so a workgroup only waits for one dispatched before
// it.  The forward launch writes y into X, the backward launch x.
template <int B, bool FWD>
__global__ __launch_bounds__(512) void band_lu_solve_grid_kernel(const double* __restrict__ lu, int n, int p, int q,
                                                                 double* __restrict__ X, long long Q, int G, int R,
                                                                 int* __restrict__ flags,
                                                                 double* __restrict__ states) {
  constexpr int NW = 8, RC = B == 4 ? 4 : 8;
  __shared__ double sm[NW * (B * B + B)];
  __shared__ double ss[NW * B];
  const long long bid = blockIdx.x;
  const long long line = bid / G;
  const int gi = (int)(bid - line * G), g = FWD ? gi : G - 1 - gi;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long pp = line / Q, qq = line - pp * Q;
  double* __restrict__ x = X + pp * (long long)n * Q + qq;
  const int W = p + q + 1;
  const int lo = g * R, hi = min(n, lo + R);
  const int c = (hi - lo + 64 * NW - 1) / (64 * NW);
  const int i0 = min(hi, lo + (int)threadIdx.x * c), i1 = min(hi, i0 + c);
  int* lf = flags + line * G;
  double* lst = states + line * G * B;
  const int pred = FWD ? g - 1 : g + 1;
  const bool haspred = pred >= 0 && pred < G;
  const GridLink gl{haspred ? lf + pred : nullptr, haspred ? lst + (long long)pred * B : nullptr, lf + g,
                    lst + (long long)g * B};
  wave_sweep<B, RC, NW, FWD>(lu, x, Q, n, W, p, q, i0, i1, lane, wv, sm, ss, &gl);
}

__global__ void band_grid_reset(int* __restrict__ flags, long long nflags) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nflags; i += (long long)gridDim.x * blockDim.x)
    flags[i] = 0;
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
// NT terms (compile time, so every load of an element issues before its
// arithmetic).  COMMON: every term has the same first factor (u u_x + u u_y):
// it is read once and the sum becomes a[i] * sum_t c_t b_t[i].
template <typename T, int NT, bool COMMON>
__global__ void sum_products_kernel(SumProd<T> sp, T* __restrict__ out, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    T a[NT], b[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (!COMMON || t == 0) a[t] = sp.a[t][i];
      b[t] = sp.b[t][i];
    }
    T s = 0;
    if constexpr (COMMON) {
#pragma unroll
      for (int t = 0; t < NT; ++t) s += (T)sp.c[t] * b[t];
      s *= a[0];
    } else {
#pragma unroll
      for (int t = 0; t < NT; ++t) s += (T)sp.c[t] * a[t] * b[t];
    }
    out[i] = s;
  }
}

// y = sum_j c_j x_j (n <= 8 inputs), optionally accumulating into y.
template <typename T>
struct LinComb {
  const T* x[8];
  double c[8];
};
// NIN inputs (compile time, so an element's loads all issue before its sum).
template <typename T, int NIN>
__global__ void lincomb_kernel(LinComb<T> lc, T* __restrict__ y, long long n, int accumulate) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    T x[NIN > 0 ? NIN : 1];
#pragma unroll
    for (int j = 0; j < NIN; ++j) x[j] = lc.x[j][i];
    double s = accumulate ? (double)y[i] : 0.0;
#pragma unroll
    for (int j = 0; j < NIN; ++j) s += lc.c[j] * (double)x[j];
    y[i] = (T)s;
  }
}

template <typename TI, typename TO>
__global__ void cast_kernel(const TI* __restrict__ x, TO* __restrict__ y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = (TO)x[i];
}

// Automatic choice between one lane per line and one wavefront per line: the
// wavefront solve pays (B + 2) x the arithmetic to cut the sequential depth
// 64-fold, which wins while the lines alone leave most of the GPU's lanes idle.
constexpr int WAVE_MAXB = 4;
constexpr int WAVE_MIN_N = 32;
constexpr long long WAVE_MAX_LINES = 2048;
constexpr long long WAVE_WIDE_LINES = 256;  // 8 waves per line: 2048 waves, two per SIMD
constexpr int GRID_ROWS = 4096;             // rows per workgroup of the grid solve (at least)
constexpr int GRID_MAX_CHUNKS = 64;         // workgroups per line
// lines x chunks.  Beyond the workgroups the GPU holds at once, the rest wait
// for slots; no workgroup can block one it depends on (each waits only for a
// lower-numbered one, and a slot's holders with lower numbers keep running).
constexpr long long GRID_MAX_WGS = 1024;

// Chunks per line and rows per chunk of the grid solve (G = 0: not applicable).
static void grid_shape(int n, long long lines, int& G, int& R) {
  G = 0;
  R = GRID_ROWS;
  if (n < 2 * GRID_ROWS) return;
  const long long per = ((long long)n + GRID_MAX_CHUNKS - 1) / GRID_MAX_CHUNKS;
  if (per > R) R = (int)((per + 511) / 512 * 512);
  const int g = (n + R - 1) / R;
  if (lines * g > GRID_MAX_WGS) return;
  G = g;
}

static unsigned grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <typename T, int NT>
static void launch_sum_products_nt(const SumProd<T>& sp, bool common, T* out, long long n, hipStream_t stream) {
  if (common)
    hipLaunchKernelGGL((sum_products_kernel<T, NT, true>), dim3(grid_for(n, 256)), dim3(256), 0, stream, sp, out, n);
  else
    hipLaunchKernelGGL((sum_products_kernel<T, NT, false>), dim3(grid_for(n, 256)), dim3(256), 0, stream, sp, out, n);
}

template <typename T>
static int launch_sum_products(int nterms, const void* const* As, const void* const* Bs, const double* coeffs,
                               void* out, long long n, hipStream_t stream) {
  SumProd<T> sp{};
  bool common = nterms > 1;
  for (int t = 0; t < nterms; ++t) {
    sp.a[t] = (const T*)As[t];
    sp.b[t] = (const T*)Bs[t];
    sp.c[t] = coeffs[t];
    common = common && As[t] == As[0];
  }
  switch (nterms) {
    case 1: launch_sum_products_nt<T, 1>(sp, false, (T*)out, n, stream); break;
    case 2: launch_sum_products_nt<T, 2>(sp, common, (T*)out, n, stream); break;
    case 3: launch_sum_products_nt<T, 3>(sp, common, (T*)out, n, stream); break;
    default: launch_sum_products_nt<T, 4>(sp, common, (T*)out, n, stream); break;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

template <typename T, int NIN>
static void launch_lincomb_n(const LinComb<T>& lc, T* y, long long n, int accumulate, hipStream_t stream) {
  hipLaunchKernelGGL((lincomb_kernel<T, NIN>), dim3(grid_for(n, 256)), dim3(256), 0, stream, lc, y, n, accumulate);
}

template <typename T>
static int launch_lincomb(int nin, const void* const* xs, const double* coeffs, void* y, long long n, int accumulate,
                          hipStream_t stream) {
  LinComb<T> lc{};
  for (int j = 0; j < nin; ++j) {
    lc.x[j] = (const T*)xs[j];
    lc.c[j] = coeffs[j];
  }
  T* yo = (T*)y;
  switch (nin) {
    case 0: launch_lincomb_n<T, 0>(lc, yo, n, accumulate, stream); break;
    case 1: launch_lincomb_n<T, 1>(lc, yo, n, accumulate, stream); break;
    case 2: launch_lincomb_n<T, 2>(lc, yo, n, accumulate, stream); break;
    case 3: launch_lincomb_n<T, 3>(lc, yo, n, accumulate, stream); break;
    case 4: launch_lincomb_n<T, 4>(lc, yo, n, accumulate, stream); break;
    case 5: launch_lincomb_n<T, 5>(lc, yo, n, accumulate, stream); break;
    case 6: launch_lincomb_n<T, 6>(lc, yo, n, accumulate, stream); break;
    case 7: launch_lincomb_n<T, 7>(lc, yo, n, accumulate, stream); break;
    default: launch_lincomb_n<T, 8>(lc, yo, n, accumulate, stream); break;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // namespace jfx

using namespace jfx;

// Wavefronts per line (p, q <= WAVE_MAXB): eight waves per line while the
// lines alone leave most of the GPU idle and the chunks stay at least a batch
// of rows long, else one.
static int launch_band_wave(const double* lu, int n, int p, int q, double* X, long long lines, long long Q, int nsys,
                            int qdiv, hipStream_t stream) {
  if (lines > 0x7fffffffLL) return set_error("band_lu_solve: too many lines for the wavefront solve"), JFX_EARG;
  const bool wide = lines <= WAVE_WIDE_LINES && n >= 64 * 8 * 8;
  const dim3 grid((unsigned)lines);
  if (p <= 2 && q <= 2) {
    if (wide)
      hipLaunchKernelGGL((band_lu_solve_wave_kernel<2, 8>), grid, dim3(512), 0, stream, lu, n, p, q, X, Q, nsys, qdiv);
    else
      hipLaunchKernelGGL((band_lu_solve_wave_kernel<2, 1>), grid, dim3(64), 0, stream, lu, n, p, q, X, Q, nsys, qdiv);
  } else {
    if (wide)
      hipLaunchKernelGGL((band_lu_solve_wave_kernel<4, 8>), grid, dim3(512), 0, stream, lu, n, p, q, X, Q, nsys, qdiv);
    else
      hipLaunchKernelGGL((band_lu_solve_wave_kernel<4, 1>), grid, dim3(64), 0, stream, lu, n, p, q, X, Q, nsys, qdiv);
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

extern "C" {

int jfx_band_lu_factor(const double* band, int n, int p, int q, double* lu, int* status_dev,
                       hipStream_t stream) {
  if (!band || !lu || !status_dev || n <= 0 || p < 0 || q < 0)
    return set_error("band_lu_factor: bad arguments"), JFX_EARG;
  if (p <= 2 && q <= 2)
    if (p == 2 && q == 2)
      hipLaunchKernelGGL((band_lu_factor_wave_kernel<2, true>), dim3(1), dim3(64), 0, stream, band, n, p, q, lu,
                         status_dev);
    else
      hipLaunchKernelGGL((band_lu_factor_wave_kernel<2, false>), dim3(1), dim3(64), 0, stream, band, n, p, q, lu,
                         status_dev);
  else if (p <= 4 && q <= 4)
    if (p == 4 && q == 4)
      hipLaunchKernelGGL((band_lu_factor_wave_kernel<4, true>), dim3(1), dim3(64), 0, stream, band, n, p, q, lu,
                         status_dev);
    else
      hipLaunchKernelGGL((band_lu_factor_wave_kernel<4, false>), dim3(1), dim3(64), 0, stream, band, n, p, q, lu,
                         status_dev);
  else
    hipLaunchKernelGGL(band_lu_factor_kernel, dim3(1), dim3(64), 0, stream, band, n, p, q, lu, status_dev, 1);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_solve_mode(const double* lu, int n, int p, int q, double* X, long long P, long long Q, int mode,
                           hipStream_t stream) {
  if (!lu || !X || n <= 0 || p < 0 || q < 0 || P < 0 || Q <= 0 || mode < 0 || mode > 2)
    return set_error("band_lu_solve: bad arguments"), JFX_EARG;
  const long long lines = P * Q;
  const bool narrow = p <= WAVE_MAXB && q <= WAVE_MAXB;
  if (mode == 2 && !narrow)
    return set_error("band_lu_solve: the wavefront solve takes bandwidths <= %d (p=%d, q=%d)", WAVE_MAXB, p, q),
           JFX_EARG;
  if (lines == 0) return JFX_OK;
  if (mode == 0)
    mode = (narrow && n >= WAVE_MIN_N && lines <= WAVE_MAX_LINES) ? 2 : 1;
  if (mode == 2) {
    const int rc = launch_band_wave(lu, n, p, q, X, lines, Q, 1, 1, stream);
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(band_lu_solve_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, stream, lu, n, p,
                       q, X, P, Q);
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_solve(const double* lu, int n, int p, int q, double* X, long long P, long long Q,
                      hipStream_t stream) {
  return jfx_band_lu_solve_mode(lu, n, p, q, X, P, Q, 0, stream);
}

long long jfx_band_lu_solve_grid_work_bytes(int n, int p, int q, long long P, long long Q) {
  if (n <= 0 || p < 0 || q < 0 || P < 0 || Q <= 0 || p > WAVE_MAXB || q > WAVE_MAXB) return 0;
  int G, R;
  grid_shape(n, P * Q, G, R);
  if (G == 0) return 0;
  const long long chunks = 2 * P * Q * G;  // forward and backward
  return (chunks * 4 + 15) / 16 * 16 + chunks * WAVE_MAXB * 8;
}

int jfx_band_lu_solve_grid(const double* lu, int n, int p, int q, double* X, long long P, long long Q, void* work,
                           long long work_bytes, hipStream_t stream) {
  if (!lu || !X || !work || n <= 0 || p < 0 || q < 0 || P < 0 || Q <= 0)
    return set_error("band_lu_solve_grid: bad arguments"), JFX_EARG;
  if (p > WAVE_MAXB || q > WAVE_MAXB)
    return set_error("band_lu_solve_grid: bandwidths must be <= %d (p=%d, q=%d)", WAVE_MAXB, p, q), JFX_EARG;
  const long long lines = P * Q;
  if (lines == 0) return JFX_OK;
  int G, R;
  grid_shape(n, lines, G, R);
  if (G == 0)
    return set_error("band_lu_solve_grid: %lld lines of %d rows do not suit the grid solve (n >= %d, lines x "
                     "chunks <= %lld)", lines, n, 2 * GRID_ROWS, GRID_MAX_WGS),
           JFX_ESHAPE;
  const long long need = jfx_band_lu_solve_grid_work_bytes(n, p, q, P, Q);
  if (work_bytes < need)
    return set_error("band_lu_solve_grid: work holds %lld bytes, %lld needed", work_bytes, need), JFX_ESHAPE;
  const long long chunks = lines * G;
  int* flags = static_cast<int*>(work);
  double* states = reinterpret_cast<double*>(static_cast<char*>(work) + (2 * chunks * 4 + 15) / 16 * 16);
  hipLaunchKernelGGL(band_grid_reset, dim3(1), dim3(256), 0, stream, flags, 2 * chunks);
  JFX_LAUNCH_CHECK();
  const dim3 grid((unsigned)chunks);
  if (p <= 2 && q <= 2) {
    hipLaunchKernelGGL((band_lu_solve_grid_kernel<2, true>), grid, dim3(512), 0, stream, lu, n, p, q, X, Q, G, R,
                       flags, states);
    JFX_LAUNCH_CHECK();
    hipLaunchKernelGGL((band_lu_solve_grid_kernel<2, false>), grid, dim3(512), 0, stream, lu, n, p, q, X, Q, G, R,
                       flags + chunks, states + chunks * WAVE_MAXB);
  } else {
    hipLaunchKernelGGL((band_lu_solve_grid_kernel<4, true>), grid, dim3(512), 0, stream, lu, n, p, q, X, Q, G, R,
                       flags, states);
    JFX_LAUNCH_CHECK();
    hipLaunchKernelGGL((band_lu_solve_grid_kernel<4, false>), grid, dim3(512), 0, stream, lu, n, p, q, X, Q, G, R,
                       flags + chunks, states + chunks * WAVE_MAXB);
  }
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

int jfx_band_lu_solve_batched_mode(const double* lu, int nsys, int n, int p, int q, double* X, long long P,
                                   long long Q, int qdiv, int mode, hipStream_t stream) {
  if (!lu || !X || nsys <= 0 || n <= 0 || p < 0 || q < 0 || P < 0 || Q <= 0 || qdiv < 1 || Q % qdiv ||
      (P * (Q / qdiv)) % nsys || mode < 0 || mode > 2)
    return set_error("band_lu_solve_batched: bad arguments (%d systems for P=%lld, Q=%lld, qdiv=%d)", nsys, P, Q,
                     qdiv),
           JFX_EARG;
  const bool narrow = p <= WAVE_MAXB && q <= WAVE_MAXB;
  if (mode == 2 && !narrow)
    return set_error("band_lu_solve_batched: the wavefront solve takes bandwidths <= %d (p=%d, q=%d)", WAVE_MAXB,
                     p, q),
           JFX_EARG;
  long long lines = P * Q;
  if (lines == 0) return JFX_OK;
  if (mode == 0) mode = (narrow && n >= WAVE_MIN_N && lines <= WAVE_MAX_LINES) ? 2 : 1;
  if (mode == 2) return launch_band_wave(lu, n, p, q, X, lines, Q, nsys, qdiv, stream);
  hipLaunchKernelGGL(band_lu_solve_batched_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, stream,
                     lu, nsys, n, p, q, X, P, Q, qdiv);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

int jfx_band_lu_solve_batched(const double* lu, int nsys, int n, int p, int q, double* X, long long P, long long Q,
                              int qdiv, hipStream_t stream) {
  return jfx_band_lu_solve_batched_mode(lu, nsys, n, p, q, X, P, Q, qdiv, 0, stream);
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
  if (dtype == 0) return launch_sum_products<double>(nterms, As, Bs, coeffs, out, n, stream);
  if (dtype == 1) return launch_sum_products<float>(nterms, As, Bs, coeffs, out, n, stream);
  return set_error("sum_products: bad dtype"), JFX_EARG;
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
  if (dtype == 0) return launch_lincomb<double>(nin, xs, coeffs, y, n, accumulate, stream);
  if (dtype == 1) return launch_lincomb<float>(nin, xs, coeffs, y, n, accumulate, stream);
  return set_error("lincomb: bad dtype"), JFX_EARG;
}

}  // extern "C"
