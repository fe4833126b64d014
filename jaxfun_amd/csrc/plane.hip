// Fused two-axis contraction of 2D slabs:
//     Y[b] = A0 (M0 x K0) . X[b] (K0 x K1) . A1^T (K1 x M1)     (all dims <= 128)
// optionally followed by a pointwise epilogue E (M0 x M1):  Y /= E  or  Y *= E.
// The intermediate T = X[b] A1^T (K0 x M1) stays in LDS, so a slab is read
// from HBM once and written once: a 2D TensorProductSpace transform, the two
// eigen-transforms + eigenvalue division of a 2D fast-diagonalisation solve,
// or the last two axes of a 3D transform, in one pass.
// Persistent workgroups of 8 waves walk the slabs: two per CU when the LDS
// buffer (K0 rows rounded up to 32) fits in half the CU's LDS and the
// registers allow.  The work split inside a slab is described at each pass.
//   FP64: v_mfma_f64_16x16x4f64, FP64 I/O (plane_pass64)
//   bf16x6: bf16 3-term float split (6 products), FP32 accumulation, FP32 I/O
//           (plane_half_kernel; the fused solve uses plane_pass3)
#include "jfx_common.h"

namespace jfx {

constexpr int QMAX = 128, QNT = 512;
constexpr int LDT64 = 129;  // LDS row stride of the FP64 T tile

// FP64 pass.  CBW: 16-column blocks per wave (1, 2, 4 or 8).  Each 16-row
// block of a phase is shared by wpr waves (8 / row blocks, at most 8), which
// take consecutive groups of CBW column blocks; every wave runs CBW MFMA
// chains per k-step unconditionally, so the loop has no divergent branches.
// Operands are loaded from clamped, always-valid addresses: rows and columns
// past the array edges then hold copies of edge values, which reach only
// output rows and columns that are never stored, and the contraction index is
// cut off by zeros (X times a mask factor for k >= K1; T's zeroed rows >= K0).
// Operands of the next k-step are loaded while the current step's MFMAs run.
// One slab: T = Xb A1^T into LDS, then Yb = alpha A0 T (with the epilogue,
// then + Yb when accum is set).  Xb
// rows are ldx apart and Yb rows ldy apart, so either may live in LDS (the
// fused solve passes its intermediate that way).  Ends with a barrier.
template <int CBW>
__device__ __forceinline__ void plane_pass64(const double* __restrict__ A0d, const double* __restrict__ A1d,
                                             const double* __restrict__ Xb, int ldx, double* __restrict__ Yb,
                                             int ldy, const double* __restrict__ E, int epi, double alpha, int accum,
                                             int M0, int K0, int M1, int K1, int wpr, double* __restrict__ T) {
  constexpr int LDT = LDT64;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ncb1 = (M1 + 15) >> 4;
  const int rb = w / wpr, c0 = (w % wpr) * CBW;
  const bool actA = 16 * rb < K0 && c0 < ncb1, actB = 16 * rb < M0 && c0 < ncb1;
  const int l16 = lane & 15, kl = lane >> 4;
  // ---------------- phase A: T (K0 x M1) = X_b A1^T
  if (actA) {
    v4d acc[CBW];
#pragma unroll
    for (int t = 0; t < CBW; ++t) acc[t] = v4d{0, 0, 0, 0};
    const int i = 16 * rb + l16;
    const int ic = i < K0 ? i : K0 - 1;
    double a0, a1, b0[CBW], b1[CBW];
    // a mask factor (not a select) zeroes k >= K1, so that no load is
    // predicated and the loads of a k-step issue back to back
    auto load = [&](int k0, double& a, double(&bv)[CBW]) {
      const int k = k0 + kl, kc = k < K1 ? k : K1 - 1;
      a = Xb[(long long)ic * ldx + kc] * (k < K1 ? 1.0 : 0.0);
#pragma unroll
      for (int t = 0; t < CBW; ++t) {
        const int m = 16 * (c0 + t) + l16, mc = m < M1 ? m : M1 - 1;
        bv[t] = A1d[(long long)mc * K1 + kc];
      }
    };
    load(0, a0, b0);
    for (int k0 = 0; k0 < K1; k0 += 8) {
      load(k0 + 4, a1, b1);
#pragma unroll
      for (int t = 0; t < CBW; ++t) acc[t] = mfma_f64(a0, b0[t], acc[t]);
      load(k0 + 8, a0, b0);
#pragma unroll
      for (int t = 0; t < CBW; ++t) acc[t] = mfma_f64(a1, b1[t], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < CBW; ++t) {
      const int m = 16 * (c0 + t) + l16;
      if (m >= M1) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ii = 16 * rb + mfma16_row_f64(lane, r);
        if (ii < K0) T[ii * LDT + m] = acc[t][r];
      }
    }
  }
  __syncthreads();
  // ---------------- phase B: Y (M0 x M1) = A0 T
  if (actB) {
    v4d acc[CBW];
#pragma unroll
    for (int t = 0; t < CBW; ++t) acc[t] = v4d{0, 0, 0, 0};
    const int mrow = 16 * rb + l16, mrc = mrow < M0 ? mrow : M0 - 1;
    auto lda = [&](int k) { return A0d[(long long)mrc * K0 + (k < K0 ? k : K0 - 1)]; };
    double an = lda(kl);
    for (int k0 = 0; k0 < K0; k0 += 4) {
      const double a = an;
      const int k = k0 + kl;
      an = lda(k + 4);
#pragma unroll
      for (int t = 0; t < CBW; ++t) acc[t] = mfma_f64(a, T[k * LDT + 16 * (c0 + t) + l16], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < CBW; ++t) {
      const int j = 16 * (c0 + t) + l16;
      if (j >= M1) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 16 * rb + mfma16_row_f64(lane, r);
        if (m >= M0) continue;
        double v = alpha * acc[t][r];
        if (epi == 1) v = v / E[(long long)m * M1 + j];
        if (epi == 2) v = v * E[(long long)m * M1 + j];
        if (accum) v += Yb[(long long)m * ldy + j];
        Yb[(long long)m * ldy + j] = v;
      }
    }
  }
  __syncthreads();
}

// bf16x6 pass (FP32 data; operators as [3][M][Kp] bf16 slices, Kp = 32
// ceil(K / 32), zero past K).  NK: k-steps of 32 held in registers (2 when
// every contracted dimension is <= 64, else 4).  Operator-stationary: each
// wave loads one 16-row fragment of the phase's operator (all k, 3 slices)
// once per phase and keeps it in registers, so the operators cross L2 once
// per phase and workgroup; the slab moves through LDS.
//   Phase A (T = Xs A1^T): wave w owns column block w of T (rows 16w.. of
//   A1) and walks the row blocks of Xs, splitting each X fragment (two
//   16-byte LDS reads) into bf16 in registers.  T is stored transposed,
//   Tt[m][i], four consecutive i per 16-byte write, rows i >= K0 as zeros.
//   Phase B (Y = A0 T): wave w owns row block w of Y (rows 16w.. of A0) and
//   walks the column blocks of T, reading its fragments from Tt.  When `pf`
//   is given, the next slab's X is loaded into registers during this phase
//   and written into Xs after it.
// The 6 products of a k-step run smallest first (lo*hi, mid*mid, hi*lo,
// mid*hi, hi*mid, hi*hi) into one FP32 accumulator.  Xs rows >= K0 and
// columns >= K1 are zero (filled once, never written); operator rows past
// the edge are clamped copies that reach only outputs that are not stored.
// Xs and Tt rows are LDX floats apart.  Ends with a barrier.
constexpr int LDX = 132;  // 16-byte LDS rows
constexpr int SREG = 32;  // staged X values per thread (K0 K1 <= 128 * 128 = 32 * QNT)

template <int NK>
__device__ __forceinline__ void load_frag3(const unsigned short* __restrict__ S, int M, int Kp, int nk, int r0,
                                           int lane, v8u16 (&f)[NK][3]) {
  const int r = r0 + (lane & 15), rc = r < M ? r : M - 1;
  const long long sl = (long long)M * Kp;
  const unsigned off = (unsigned)(rc * Kp + 8 * (lane >> 4));  // uniform bases + one 32-bit lane offset
#pragma unroll
  for (int ks = 0; ks < NK; ++ks) {
#pragma unroll
    for (int s3 = 0; s3 < 3; ++s3) f[ks][s3] = v8u16{0, 0, 0, 0, 0, 0, 0, 0};  // k-steps past Kp stay zero
    if (ks < nk) {
#pragma unroll
      for (int s3 = 0; s3 < 3; ++s3) f[ks][s3] = *(const v8u16*)(S + s3 * sl + off + 32 * ks);
    }
  }
}

__device__ __forceinline__ void lds_split8(const float* p, v8u16& h, v8u16& m, v8u16& l) {
  const float4 a = *(const float4*)p, b = *(const float4*)(p + 4);
  const float x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  split3_pk8(x, h, m, l);
}

// X slab (K0 x K1, contiguous) -> registers -> Xs[i][k]
__device__ __forceinline__ void stage_load(const float* __restrict__ Xg, int cnt, float (&v)[SREG]) {
#pragma unroll
  for (int q = 0; q < SREG; ++q) {
    const int f = threadIdx.x + QNT * q;
    v[q] = f < cnt ? Xg[f] : 0.f;
  }
}
__device__ __forceinline__ void stage_store(float* Xs, int cnt, int K1, const float (&v)[SREG]) {
  int i = threadIdx.x / K1, k = threadIdx.x - i * K1;
  const int di = QNT / K1, dk = QNT - di * K1;
#pragma unroll
  for (int q = 0; q < SREG; ++q) {
    if (threadIdx.x + QNT * q < cnt) Xs[i * LDX + k] = v[q];
    i += di;
    k += dk;
    if (k >= K1) {
      k -= K1;
      ++i;
    }
  }
}

template <int NK>
__device__ __forceinline__ void plane_pass3(const unsigned short* __restrict__ A0s,
                                            const unsigned short* __restrict__ A1s, float* Xs, float* Yb, int ldy,
                                            const float* __restrict__ E, int epi, int M0, int K0, int M1, int K1,
                                            float* Tt, const float* __restrict__ pf, int pfcnt) {
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l16 = lane & 15, kl = lane >> 4;
  const int K0p = (K0 + 31) & ~31, K1p = (K1 + 31) & ~31;
  const int nkA = K1p >> 5, nkB = K0p >> 5;  // <= NK
  // ---------------- phase A: T (K0 x M1) = X A1^T, stored as Tt (M1 x K0)
  // Every k-step up to NK runs: Xs columns past K1 and fragments past nk are zero.
  {
    const int nrb = (K0 + 15) >> 4, ncb = (M1 + 15) >> 4;
    if (w < ncb) {
      v8u16 bf[NK][3];
      load_frag3<NK>(A1s, M1, K1p, nkA, 16 * w, lane, bf);
      float* trow = Tt + (16 * w + l16) * LDX + 4 * kl;
      const float* xbase = Xs + l16 * LDX + 8 * kl;
      float4 xa[NK][2], xb[NK][2];
      auto ldx = [&](int rb, float4(&x)[NK][2]) {
        const float* p = xbase + 16 * rb * LDX;
#pragma unroll
        for (int ks = 0; ks < NK; ++ks) {
          x[ks][0] = *(const float4*)(p + 32 * ks);
          x[ks][1] = *(const float4*)(p + 32 * ks + 4);
        }
      };
      auto blk = [&](int rb, const float4(&x)[NK][2]) {
        v4f acc = v4f{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < NK; ++ks) {
          const float xv[8] = {x[ks][0].x, x[ks][0].y, x[ks][0].z, x[ks][0].w,
                               x[ks][1].x, x[ks][1].y, x[ks][1].z, x[ks][1].w};
          v8u16 xh, xm, xl;
          split3_pk8(xv, xh, xm, xl);
          acc = mfma_bf16(xl, bf[ks][0], acc);
          acc = mfma_bf16(xm, bf[ks][1], acc);
          acc = mfma_bf16(xh, bf[ks][2], acc);
          acc = mfma_bf16(xm, bf[ks][0], acc);
          acc = mfma_bf16(xh, bf[ks][1], acc);
          acc = mfma_bf16(xh, bf[ks][0], acc);
        }
        const int ii = 16 * rb + 4 * kl;  // rows ii .. ii + 3 of T (mfma16_row_f32)
        float4 v;
        v.x = ii + 0 < K0 ? acc[0] : 0.f;
        v.y = ii + 1 < K0 ? acc[1] : 0.f;
        v.z = ii + 2 < K0 ? acc[2] : 0.f;
        v.w = ii + 3 < K0 ? acc[3] : 0.f;
        *(float4*)(trow + 16 * rb) = v;
      };
      ldx(0, xa);
      for (int rb = 0; rb < nrb; rb += 2) {
        if (rb + 1 < nrb) ldx(rb + 1, xb);
        blk(rb, xa);
        if (rb + 2 < nrb) ldx(rb + 2, xa);
        if (rb + 1 < nrb) blk(rb + 1, xb);
      }
    }
  }
  __syncthreads();
  // ---------------- phase B: Y (M0 x M1) = A0 T
  {
    const int nrb = (M0 + 15) >> 4, ncb = (M1 + 15) >> 4;
    v8u16 af[NK][3];
    // the operator fragment is issued before the next slab's loads, so the
    // first MFMAs wait for it alone
    if (w < nrb) load_frag3<NK>(A0s, M0, K0p, nkB, 16 * w, lane, af);
    float pv[SREG];
    if (pf) stage_load(pf, pfcnt, pv);
    if (w < nrb) {
      const float* tbase = Tt + l16 * LDX + 8 * kl;
      float4 ta[NK][2], tb[NK][2];
      auto ldt = [&](int t, float4(&x)[NK][2]) {
        const float* p = tbase + 16 * t * LDX;
#pragma unroll
        for (int ks = 0; ks < NK; ++ks) {
          x[ks][0] = *(const float4*)(p + 32 * ks);
          x[ks][1] = *(const float4*)(p + 32 * ks + 4);
        }
      };
      auto blk = [&](int t, const float4(&x)[NK][2]) {
        v4f acc = v4f{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < NK; ++ks) {
          const float xv[8] = {x[ks][0].x, x[ks][0].y, x[ks][0].z, x[ks][0].w,
                               x[ks][1].x, x[ks][1].y, x[ks][1].z, x[ks][1].w};
          v8u16 th, tm, tl;
          split3_pk8(xv, th, tm, tl);
          acc = mfma_bf16(af[ks][2], th, acc);
          acc = mfma_bf16(af[ks][1], tm, acc);
          acc = mfma_bf16(af[ks][0], tl, acc);
          acc = mfma_bf16(af[ks][1], th, acc);
          acc = mfma_bf16(af[ks][0], tm, acc);
          acc = mfma_bf16(af[ks][0], th, acc);
        }
        const int j = 16 * t + l16;
        if (j < M1) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int m = 16 * w + mfma16_row_f32(lane, r);
            if (m >= M0) continue;
            float v = acc[r];
            if (epi == 1) v = v / E[(long long)m * M1 + j];
            if (epi == 2) v = v * E[(long long)m * M1 + j];
            Yb[(long long)m * ldy + j] = v;
          }
        }
      };
      ldt(0, ta);
      for (int t = 0; t < ncb; t += 2) {
        if (t + 1 < ncb) ldt(t + 1, tb);
        blk(t, ta);
        if (t + 2 < ncb) ldt(t + 2, ta);
        if (t + 1 < ncb) blk(t + 1, tb);
      }
    }
    __syncthreads();
    if (pf) {
      stage_store(Xs, pfcnt, K1, pv);
      __syncthreads();
    }
  }
}

template <int CBW>
__global__ __launch_bounds__(QNT, CBW <= 2 ? 4 : 2) void plane_kernel64(const double* __restrict__ A0d,
                                                                        const double* __restrict__ A1d,
                                                                        const double* __restrict__ X,
                                                                        double* __restrict__ Y,
                                                                        const double* __restrict__ E, int epi,
                                                                        double alpha, int accum, long long B, int M0,
                                                                        int K0, int M1, int K1, int wpr) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* T = reinterpret_cast<double*>(smem);  // T[i][m], i < K0p (rows >= K0 zero), m < LDT
  const int K0p = (K0 + 31) & ~31;
  // rows >= K0 and columns >= M1 of T are zeroed once and never written, so
  // phase B reads T without masks
  for (int e = threadIdx.x; e < K0p * LDT64; e += QNT) T[e] = 0.0;
  __syncthreads();
  for (long long b = blockIdx.x; b < B; b += gridDim.x)
    plane_pass64<CBW>(A0d, A1d, X + b * (long long)K0 * K1, K1, Y + b * (long long)M0 * M1, M1, E, epi, alpha,
                      accum, M0, K0, M1, K1, wpr, T);
}

template <int NK>
__global__ __launch_bounds__(QNT, 2) void plane_kernel3(const unsigned short* __restrict__ A0s,
                                                        const unsigned short* __restrict__ A1s,
                                                        const float* __restrict__ X, float* __restrict__ Y,
                                                        const float* __restrict__ E, int epi, long long B, int M0,
                                                        int K0, int M1, int K1) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int K0p = (K0 + 31) & ~31;
  float* Xs = reinterpret_cast<float*>(smem);  // K0p x LDX
  float* Tt = Xs + K0p * LDX;                  // 128 x LDX
  for (int e = threadIdx.x; e < (K0p + 128) * LDX; e += QNT) Xs[e] = 0.f;
  __syncthreads();
  const long long slab = (long long)K0 * K1;
  const int cnt = K0 * K1;
  long long b = blockIdx.x;
  if (b < B) {
    float v[SREG];
    stage_load(X + b * slab, cnt, v);
    stage_store(Xs, cnt, K1, v);
  }
  __syncthreads();
  for (; b < B; b += gridDim.x) {
    const long long nb = b + gridDim.x;
    plane_pass3<NK>(A0s, A1s, Xs, Y + b * (long long)M0 * M1, M1, E, epi, M0, K0, M1, K1, Tt,
                    nb < B ? X + nb * slab : nullptr, cnt);
  }
}

// ---------------------------------------------------------------------------
// bf16x6 plane contraction with every FP32 element split into bf16 slices
// once.  The contraction index of phase B (the rows of X, i.e. of T) runs in
// halves of 64 rows.  Per half: the 64 rows of X are staged into LDS as three
// bf16 slices (split by the staging thread); phase A (T = X A1^T for those
// rows; wave w = column block w of T) stores T transposed and split (each
// value split by the lane that produced it); phase B (wave w = row block w
// of Y) accumulates A0[:, half] T[half, :] into registers that persist over
// the halves.  The operator fragments (A1 rows 16w.. for phase A, A0 rows
// 16w.. for phase B; all k, 3 slices) are loaded once per workgroup and stay
// in registers for every slab.  The next half's (or slab's) rows are loaded
// into registers during phase B and stored after it.  Products and their
// order per k-step are those of the FP32-tile pass above, so results match it
// bit for bit.
constexpr int LDK = 136;  // bf16 row stride of the staged X half (272 B)
constexpr int LDH = 72;   // bf16 row stride of the transposed T half (144 B)
constexpr int XH_SLICE = 64 * LDK, TH_SLICE = 128 * LDH;  // elements per slice
constexpr size_t HALF_LDS = (size_t)(3 * XH_SLICE + 3 * TH_SLICE) * 2;

// 4 fp32 -> 3 bf16 slices of two packed dwords (the split of split3_pk8)
__device__ __forceinline__ void split3_pk4(const float (&x)[4], uint2& H, uint2& M, uint2& L) {
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  unsigned hw[2], mw[2], lw[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const float x0 = x[2 * e], x1 = x[2 * e + 1];
    const unsigned hu = __builtin_bit_cast(unsigned, __builtin_convertvector((f2){x0, x1}, bf2));
    const float r0 = x0 - __uint_as_float(hu << 16), r1 = x1 - __uint_as_float(hu & 0xFFFF0000u);
    const unsigned mu = __builtin_bit_cast(unsigned, __builtin_convertvector((f2){r0, r1}, bf2));
    const float s0 = r0 - __uint_as_float(mu << 16), s1 = r1 - __uint_as_float(mu & 0xFFFF0000u);
    hw[e] = hu;
    mw[e] = mu;
    lw[e] = __builtin_bit_cast(unsigned, __builtin_convertvector((f2){s0, s1}, bf2));
  }
  H = make_uint2(hw[0], hw[1]);
  M = make_uint2(mw[0], mw[1]);
  L = make_uint2(lw[0], lw[1]);
}

// rows 64 h .. 64 h + 63 of the slab (K0 x K1, contiguous): two 8-wide
// fragments per thread, zero outside the slab
__device__ __forceinline__ void half_load(const float* __restrict__ Xb, int K0, int K1, int h, float (&v)[2][8]) {
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = threadIdx.x + QNT * u, i = 64 * h + (q >> 4), k0 = 8 * (q & 15);
    const float* p = Xb + (long long)i * K1 + k0;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[u][e] = (i < K0 && k0 + e < K1) ? p[e] : 0.f;
  }
}
__device__ __forceinline__ void half_store(unsigned short* Xh, const float (&v)[2][8]) {
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = threadIdx.x + QNT * u;
    unsigned short* d = Xh + (q >> 4) * LDK + 8 * (q & 15);
    v8u16 hh, mm, ll;
    split3_pk8(v[u], hh, mm, ll);
    *(v8u16*)d = hh;
    *(v8u16*)(d + XH_SLICE) = mm;
    *(v8u16*)(d + 2 * XH_SLICE) = ll;
  }
}

// phase B over one half H of the contraction: acc[t] += A0[rows 16w.., half] T[half, cols 16t..]
template <int NK, int H>
__device__ __forceinline__ void half_phase_b(const unsigned short* Tt, const v8u16 (&af)[NK][3], int ncb, int lane,
                                             v4f (&acc)[8]) {
  if constexpr (2 * H < NK) {
    const unsigned short* tb = Tt + (lane & 15) * LDH + 8 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (t < ncb) {
        const unsigned short* tr = tb + 16 * t * LDH;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          const int ks = 2 * H + kk;
          if (ks < NK) {
            const v8u16 th = *(const v8u16*)(tr + 32 * kk), tm = *(const v8u16*)(tr + TH_SLICE + 32 * kk),
                        tl = *(const v8u16*)(tr + 2 * TH_SLICE + 32 * kk);
            acc[t] = mfma_bf16(af[ks][2], th, acc[t]);
            acc[t] = mfma_bf16(af[ks][1], tm, acc[t]);
            acc[t] = mfma_bf16(af[ks][0], tl, acc[t]);
            acc[t] = mfma_bf16(af[ks][1], th, acc[t]);
            acc[t] = mfma_bf16(af[ks][0], tm, acc[t]);
            acc[t] = mfma_bf16(af[ks][0], th, acc[t]);
          }
        }
      }
    }
  }
}

// One workgroup, 8 waves, per slab.  Phase A: wave w = column block w of T;
// phase B: wave w = row block w of Y, its 8 accumulators persisting over
// the halves.
template <int NK>
__global__ __launch_bounds__(QNT, 2) void plane_half_kernel(const unsigned short* __restrict__ A0s,
                                                            const unsigned short* __restrict__ A1s,
                                                            const float* __restrict__ X, float* __restrict__ Y,
                                                            const float* __restrict__ E, int epi, float alpha,
                                                            int accum, long long B, int M0, int K0, int M1, int K1) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  unsigned short* Xh = reinterpret_cast<unsigned short*>(smem);  // [3][64][LDK]
  unsigned short* Tt = Xh + 3 * XH_SLICE;                         // [3][128][LDH]
  const int lane = threadIdx.x & 63, l16 = lane & 15, kl = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int K0p = (K0 + 31) & ~31, K1p = (K1 + 31) & ~31;
  const int nh = (K0 + 63) >> 6;
  const int ncbT = (M1 + 15) >> 4, nrbY = (M0 + 15) >> 4;
  const long long slab = (long long)K0 * K1;
  v8u16 bf[NK][3], af[NK][3];
  if (w < ncbT) load_frag3<NK>(A1s, M1, K1p, K1p >> 5, 16 * w, lane, bf);
  if (w < nrbY) load_frag3<NK>(A0s, M0, K0p, K0p >> 5, 16 * w, lane, af);
  long long b = blockIdx.x;
  if (b >= B) return;
  {
    float v[2][8];
    half_load(X + b * slab, K0, K1, 0, v);
    half_store(Xh, v);
  }
  __syncthreads();
  v4f acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = v4f{0, 0, 0, 0};
  for (;;) {
    for (int h = 0; h < nh; ++h) {
      // ---- phase A: rows 64 h .. of T, transposed and split into Tt
      if (w < ncbT) {
        unsigned short* trow = Tt + (16 * w + l16) * LDH + 4 * kl;
        const unsigned short* xb = Xh + l16 * LDK + 8 * kl;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          v4f a = v4f{0, 0, 0, 0};
          const unsigned short* xr = xb + 16 * rb * LDK;
#pragma unroll
          for (int ks = 0; ks < NK; ++ks) {
            const v8u16 xh = *(const v8u16*)(xr + 32 * ks), xm = *(const v8u16*)(xr + XH_SLICE + 32 * ks),
                        xl = *(const v8u16*)(xr + 2 * XH_SLICE + 32 * ks);
            a = mfma_bf16(xl, bf[ks][0], a);
            a = mfma_bf16(xm, bf[ks][1], a);
            a = mfma_bf16(xh, bf[ks][2], a);
            a = mfma_bf16(xm, bf[ks][0], a);
            a = mfma_bf16(xh, bf[ks][1], a);
            a = mfma_bf16(xh, bf[ks][0], a);
          }
          const int i0 = 64 * h + 16 * rb + 4 * kl;  // T rows of a[0..3] (mfma16_row_f32)
          const float t4[4] = {i0 < K0 ? a[0] : 0.f, i0 + 1 < K0 ? a[1] : 0.f, i0 + 2 < K0 ? a[2] : 0.f,
                               i0 + 3 < K0 ? a[3] : 0.f};
          uint2 H, Mm, L;
          split3_pk4(t4, H, Mm, L);
          unsigned short* d = trow + 16 * rb;
          *(uint2*)d = H;
          *(uint2*)(d + TH_SLICE) = Mm;
          *(uint2*)(d + 2 * TH_SLICE) = L;
        }
      }
      __syncthreads();
      // ---- next half or slab into registers, phase B, then into Xh
      const bool last = h + 1 == nh;
      const long long nb = last ? b + gridDim.x : b;
      const bool more = nb < B;
      float v[2][8];
      if (more) half_load(X + nb * slab, K0, K1, last ? 0 : h + 1, v);
      if (w < nrbY) {
        if (h == 0)
          half_phase_b<NK, 0>(Tt, af, ncbT, lane, acc);
        else
          half_phase_b<NK, 1>(Tt, af, ncbT, lane, acc);
      }
      if (last && w < nrbY) {
        float* Yb = Y + b * (long long)M0 * M1;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int j = 16 * t + l16;
          if (t < ncbT && j < M1) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int m = 16 * w + mfma16_row_f32(lane, r);
              if (m >= M0) continue;
              float val = alpha * acc[t][r];
              if (epi == 1) val = val / E[(long long)m * M1 + j];
              if (epi == 2) val = val * E[(long long)m * M1 + j];
              if (accum) val += Yb[(long long)m * M1 + j];
              Yb[(long long)m * M1 + j] = val;
            }
          }
          acc[t] = v4f{0, 0, 0, 0};
        }
      }
      __syncthreads();
      if (more) half_store(Xh, v);
      __syncthreads();
    }
    b += gridDim.x;
    if (b >= B) break;
  }
}

// Two-axis fast-diagonalisation solve of n x n slabs in one launch:
//     Y[b] = V0 ((W0 X[b] W1^T) / E) V1^T
// (the forward eigen-transform with the eigenvalue division, then the back
// transform); the intermediate stays in a second LDS tile whose padding is
// zero, so each slab is read from and written to HBM once.
template <int CBW>
__global__ __launch_bounds__(QNT, 2) void plane_fd_kernel64(const double* __restrict__ W0d,
                                                            const double* __restrict__ W1d,
                                                            const double* __restrict__ V0d,
                                                            const double* __restrict__ V1d,
                                                            const double* __restrict__ X, double* __restrict__ Y,
                                                            const double* __restrict__ E, long long B, int n,
                                                            int wpr) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int np = (n + 31) & ~31;
  double* T = reinterpret_cast<double*>(smem);  // phase tile, np x LDT
  double* Z = T + np * LDT64;                   // intermediate, np x LDT
  for (int e = threadIdx.x; e < 2 * np * LDT64; e += QNT) T[e] = 0.0;
  __syncthreads();
  for (long long b = blockIdx.x; b < B; b += gridDim.x) {
    plane_pass64<CBW>(W0d, W1d, X + b * (long long)n * n, n, Z, LDT64, E, 1, 1.0, 0, n, n, n, n, wpr, T);
    plane_pass64<CBW>(V0d, V1d, Z, LDT64, Y + b * (long long)n * n, n, nullptr, 0, 1.0, 0, n, n, n, n, wpr, T);
  }
}

template <int NK>
__global__ __launch_bounds__(QNT, 2) void plane_fd_kernel3(const unsigned short* __restrict__ W0s,
                                                           const unsigned short* __restrict__ W1s,
                                                           const unsigned short* __restrict__ V0s,
                                                           const unsigned short* __restrict__ V1s,
                                                           const float* __restrict__ X, float* __restrict__ Y,
                                                           const float* __restrict__ E, long long B, int n) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int np = (n + 31) & ~31;
  float* Xs = reinterpret_cast<float*>(smem);  // slab, then the intermediate Z: np x LDX
  float* Tt = Xs + np * LDX;                   // 128 x LDX
  for (int e = threadIdx.x; e < (np + 128) * LDX; e += QNT) Xs[e] = 0.f;
  __syncthreads();
  const long long slab = (long long)n * n;
  const int cnt = n * n;
  long long b = blockIdx.x;
  if (b < B) {
    float v[SREG];
    stage_load(X + b * slab, cnt, v);
    stage_store(Xs, cnt, n, v);
  }
  __syncthreads();
  for (; b < B; b += gridDim.x) {
    const long long nb = b + gridDim.x;
    // Z = (W0 X W1^T) / E back into Xs (X is consumed by phase A), then Y = V0 Z V1^T
    plane_pass3<NK>(W0s, W1s, Xs, Xs, LDX, E, 1, n, n, n, n, Tt, nullptr, 0);
    plane_pass3<NK>(V0s, V1s, Xs, Y + b * slab, n, nullptr, 0, n, n, n, n, Tt, nb < B ? X + nb * slab : nullptr,
                    cnt);
  }
}

static int cu_count() {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
    ncu = 256;
  return ncu;
}

}  // namespace jfx

using namespace jfx;

extern "C" {

// Y[b] = alpha A0 X[b] A1^T for b < B (X: B x K0 x K1, Y: B x M0 x M1), dims
// <= 128.  prec 0: FP64 (A0d/A1d FP64 row-major); prec 1: 3-term bf16 split
// on FP32 data (A0s/A1s: [3][M][Kp] slices, Kp = 32*ceil(K/32)).  epi: 0
// none, 1 divide by E (M0 x M1), 2 multiply by E; accumulate = 1 then adds
// the previous Y (a sum of Kronecker-product terms without a separate pass).
int jfx_plane_contract(int prec, const double* A0d, const double* A1d, const unsigned short* A0s,
                       const unsigned short* A1s, const void* X, void* Y, const void* E, int epi, double alpha,
                       int accumulate, long long B, int M0, int K0, int M1, int K1, hipStream_t stream) {
  if (!X || !Y || B < 0 || M0 <= 0 || K0 <= 0 || M1 <= 0 || K1 <= 0 || epi < 0 || epi > 2 ||
      (epi && !E) || accumulate < 0 || accumulate > 1)
    return set_error("plane_contract: bad arguments"), JFX_EARG;
  if (accumulate && X == (const void*)Y) return set_error("plane_contract: accumulate needs Y distinct from X"), JFX_EARG;
  if (M0 > QMAX || K0 > QMAX || M1 > QMAX || K1 > QMAX)
    return set_error("plane_contract: dims must be <= %d", QMAX), JFX_ESHAPE;
  if (B == 0) return JFX_OK;
  const int ncu = cu_count();
  const int K0p = (K0 + 31) & ~31;
  hipError_t e;
  if (prec == 0) {
    if (!A0d || !A1d) return set_error("plane_contract: FP64 operators missing"), JFX_EARG;
    // waves per 16-row block and column blocks per wave
    const int nrb = ((K0 > M0 ? K0 : M0) + 15) >> 4, ncb = (M1 + 15) >> 4;
    const int wpr = nrb == 1 ? 8 : nrb == 2 ? 4 : nrb <= 4 ? 2 : 1;
    const int nb = (ncb + wpr - 1) / wpr, cbw = nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : 8;
    // T holds K0p rows; two workgroups share a CU when the LDS and registers allow
    const size_t shm = (size_t)K0p * LDT64 * 8;
    long long grid = (long long)ncu * ((shm <= 80 * 1024 && cbw <= 2) ? 2 : 1);
    if (grid > B) grid = B;
    typedef void (*KF)(const double*, const double*, const double*, double*, const double*, int, double, int,
                       long long, int, int, int, int, int);
    KF kern = cbw == 1 ? plane_kernel64<1> : cbw == 2 ? plane_kernel64<2> : cbw == 4 ? plane_kernel64<4>
                                                                                     : plane_kernel64<8>;
    e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(QNT), shm, stream, A0d, A1d, (const double*)X, (double*)Y,
                       (const double*)E, epi, alpha, accumulate, B, M0, K0, M1, K1, wpr);
  } else if (prec == 1) {
    if (!A0s || !A1s) return set_error("plane_contract: bf16 slices missing"), JFX_EARG;
    const int K1p = (K1 + 31) & ~31;
    const int nk = (K0p > K1p ? K0p : K1p) >> 5;
    const size_t shm = HALF_LDS;
    const long long grid = (long long)ncu < B ? ncu : B;  // one workgroup per CU (LDS)
    typedef void (*KF)(const unsigned short*, const unsigned short*, const float*, float*, const float*, int, float,
                       int, long long, int, int, int, int);
    KF kern = nk <= 2 ? plane_half_kernel<2> : plane_half_kernel<4>;
    e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(QNT), shm, stream, A0s, A1s, (const float*)X, (float*)Y,
                       (const float*)E, epi, (float)alpha, accumulate, B, M0, K0, M1, K1);
  } else {
    return set_error("plane_contract: bad prec"), JFX_EARG;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

// Y[b] = V0 ((W0 X[b] W1^T) / E) V1^T for n x n slabs (X, Y: B x n x n), one
// launch.  prec 0: FP64 operators, n <= 64; prec 1: bf16 3-term slices
// ([3][n][np]) on FP32 data, n <= 128.
int jfx_plane_fd(int prec, const double* W0d, const double* W1d, const double* V0d, const double* V1d,
                 const unsigned short* W0s, const unsigned short* W1s, const unsigned short* V0s,
                 const unsigned short* V1s, const void* X, void* Y, const void* E, long long B, int n,
                 hipStream_t stream) {
  if (!X || !Y || !E || B < 0 || n <= 0) return set_error("plane_fd: bad arguments"), JFX_EARG;
  if (n > (prec == 0 ? 64 : QMAX)) return set_error("plane_fd: n = %d too large for the fused solve", n), JFX_ESHAPE;
  if (B == 0) return JFX_OK;
  const int ncu = cu_count();
  const int np = (n + 31) & ~31;
  hipError_t e;
  if (prec == 0) {
    if (!W0d || !W1d || !V0d || !V1d) return set_error("plane_fd: FP64 operators missing"), JFX_EARG;
    const int nrb = (n + 15) >> 4;
    const int wpr = nrb == 1 ? 8 : nrb == 2 ? 4 : nrb <= 4 ? 2 : 1;
    const int nb = (nrb + wpr - 1) / wpr, cbw = nb <= 1 ? 1 : nb <= 2 ? 2 : 4;
    const size_t shm = (size_t)2 * np * LDT64 * 8;
    const long long grid = (long long)ncu < B ? ncu : B;
    typedef void (*KF)(const double*, const double*, const double*, const double*, const double*, double*,
                       const double*, long long, int, int);
    KF kern = cbw == 1 ? plane_fd_kernel64<1> : cbw == 2 ? plane_fd_kernel64<2> : plane_fd_kernel64<4>;
    e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane_fd: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(QNT), shm, stream, W0d, W1d, V0d, V1d, (const double*)X,
                       (double*)Y, (const double*)E, B, n, wpr);
  } else if (prec == 1) {
    if (!W0s || !W1s || !V0s || !V1s) return set_error("plane_fd: bf16 slices missing"), JFX_EARG;
    const int nk = np >> 5;
    const size_t shm = (size_t)(np + 128) * LDX * 4;
    const long long grid = (long long)ncu < B ? ncu : B;  // one workgroup per CU (LDS)
    typedef void (*KF)(const unsigned short*, const unsigned short*, const unsigned short*, const unsigned short*,
                       const float*, float*, const float*, long long, int);
    KF kern = nk <= 2 ? plane_fd_kernel3<2> : plane_fd_kernel3<4>;
    e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return set_error("plane_fd: %s", hipGetErrorString(e)), JFX_EHIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(QNT), shm, stream, W0s, W1s, V0s, V1s, (const float*)X,
                       (float*)Y, (const float*)E, B, n);
  } else {
    return set_error("plane_fd: bad prec"), JFX_EARG;
  }
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // extern "C"
