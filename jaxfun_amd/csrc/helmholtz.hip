// Ragged batch of 2D Legendre-Dirichlet Helmholtz solves in one persistent
// launch:   (alpha M (x) M - K (x) M - M (x) K) U = F,   U, F (n x n), n = N - 2.
//
// M (Dirichlet mass, bands 0 and +-2) and K = -diag(4j + 6) never couple
// indices of different parity, so every solve splits into four independent
// parity blocks F[p::2, q::2].  Each block is one work item:
//   1. G  = V_p^T F_pq             (MFMA; V_p: eigenvectors of K_p V = M_p V L,
//                                    V_p^T M_p V_p = I, FP64 tables built once per N)
//   2. row r of G:  ((alpha - l_r) M_q - K_q) u_r = g_r    (FP64 tridiagonal)
//   3. U_pq = V_p U~                (MFMA)
// which is partial fast diagonalisation (la/tpmatrix.py:1089-1233 for the
// eigenpairs, la/diamatrix.py:394-504 + 2079-2195 for the banded solve).
//
// Work items are taken from a device counter in the host's order (largest
// first); every workgroup loops until the counter passes nitems.
//   MODE 0: FP64 MFMA, FP64 storage
//   MODE 1: BF16 3-term float split (6 products) MFMA, FP32 accumulation and
//           storage, FP64 tridiagonal solve.
#include <type_traits>

#include "jfx_common.h"

namespace jfx {

struct HelmItem {
  long long foff;   // offset of this solve's F / U (n*n, row-major)
  long long voff;   // offset of the (AP x AP) eigenvector table (parity p)
  double alpha;
  int n, p, q, a, b, AP, loff, pad;
};

constexpr int HMAXA = 128;  // a, b <= 128  (n <= 256)

template <int MODE>
struct HelmTypes;
template <>
struct HelmTypes<0> {
  typedef double TS;
};
template <>
struct HelmTypes<1> {
  typedef float TS;
};

__global__ void helm_reset(int* counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

// Split an fp32 value into 3 bf16 terms.
__device__ __forceinline__ void split3(float x, unsigned short& h, unsigned short& m, unsigned short& l) {
  h = f2bf(x);
  float r = x - bf2f(h);
  m = f2bf(r);
  r -= bf2f(m);
  l = f2bf(r);
}

// FP64 reciprocal: hardware estimate refined by two Newton steps (x > 0, normal).
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

template <int MODE>
struct HelmLDS;
template <>
struct HelmLDS<0> {  // FP64: work block 128 x 129 doubles, V slab 16 x 129 doubles
  static constexpr int LDW = 129, KS = 16;
  static constexpr size_t WK = (size_t)HMAXA * LDW * 8, VS = (size_t)KS * (HMAXA + 1) * 8;
};
template <>
struct HelmLDS<1> {  // bf16 split: work block 128 x 130 floats, 3 bf16 slices of a 128 x 32 V slab
  static constexpr int LDW = 130, KS = 32, VLD = 40;
  static constexpr size_t WK = (size_t)HMAXA * LDW * 4, VS = (size_t)3 * HMAXA * VLD * 2;
};
constexpr size_t HTAB = 3 * HMAXA * 8;

// One GEMM phase: C (a x b, rows of V) = A (AP x AP table, applied as rows) * Wk
//   phase 1: A = V^T (tabT), result written back into Wk (column-owned panels)
//   phase 3: A = V   (tab),  result written to global U
template <int MODE, int NW, typename TIO, bool TO_GLOBAL>
__device__ __forceinline__ void helm_gemm_phase(const HelmItem& I, const double* __restrict__ tabd,
                                                const unsigned short* __restrict__ tabs,
                                                typename HelmTypes<MODE>::TS* Wk, unsigned char* vslab,
                                                TIO* __restrict__ Ub) {
  typedef typename HelmTypes<MODE>::TS TS;
  constexpr int LDW = HelmLDS<MODE>::LDW, KS = HelmLDS<MODE>::KS;
  constexpr int PPW = 8 / NW;  // column panels per wave (b <= 128 -> <= 8 panels)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int a = I.a, b = I.b, AP = I.AP, n = I.n;
  const int npanel = (b + 15) >> 4, nrb = AP >> 4;
  typedef typename std::conditional<MODE == 0, v4d, v4f>::type vacc;
  vacc acc[PPW][8];
#pragma unroll
  for (int pp = 0; pp < PPW; ++pp)
#pragma unroll
    for (int rb = 0; rb < 8; ++rb) acc[pp][rb] = vacc{0, 0, 0, 0};

  for (int k0 = 0; k0 < AP; k0 += KS) {
    if constexpr (MODE == 0) {
      double* Vs0 = reinterpret_cast<double*>(vslab);  // [KS][HMAXA + 1]: Vs0[kk][m] = A[m][k0 + kk]
      for (int e = tid; e < KS * AP; e += NW * 64) {
        int kk = e / AP, m = e - kk * AP;
        Vs0[kk * (HMAXA + 1) + m] = tabd[I.voff + (long long)(k0 + kk) * AP + m];
      }
    } else {
      constexpr int VLD = HelmLDS<1>::VLD;
      unsigned short* Vsl = reinterpret_cast<unsigned short*>(vslab);  // [3][HMAXA][VLD]
      for (int ch = tid; ch < 3 * AP * 4; ch += NW * 64) {
        int sl = ch / (AP * 4), rem = ch - sl * AP * 4, m = rem >> 2, part = rem & 3;
        *(v8u16*)&Vsl[(sl * HMAXA + m) * VLD + part * 8] =
            *(const v8u16*)(tabs + 3 * I.voff + (long long)sl * AP * AP + (long long)m * AP + k0 + part * 8);
      }
    }
    __syncthreads();
#pragma unroll
    for (int pp = 0; pp < PPW; ++pp) {
      const int pc = w + pp * NW;
      if (pc < npanel) {
        const int col = pc * 16 + (lane & 15);
        const bool cok = col < b;
        if constexpr (MODE == 0) {
          const double* Vs0 = reinterpret_cast<const double*>(vslab);
#pragma unroll
          for (int ks = 0; ks < KS; ks += 4) {
            const int k = k0 + ks + (lane >> 4);
            double bv = (cok && k < a) ? (double)Wk[k * LDW + col] : 0.0;
#pragma unroll
            for (int rb = 0; rb < 8; ++rb)
              if (rb < nrb)
                acc[pp][rb] = mfma_f64(Vs0[(ks + (lane >> 4)) * (HMAXA + 1) + rb * 16 + (lane & 15)], bv,
                                       acc[pp][rb]);
          }
        } else {
          constexpr int VLD = HelmLDS<1>::VLD;
          const unsigned short* Vsl = reinterpret_cast<const unsigned short*>(vslab);
          v8u16 bh, bm, bl;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k = k0 + 8 * (lane >> 4) + e;
            float x = (cok && k < a) ? (float)Wk[k * LDW + col] : 0.f;
            unsigned short h, m_, l;
            split3(x, h, m_, l);
            bh[e] = h;
            bm[e] = m_;
            bl[e] = l;
          }
#pragma unroll
          for (int rb = 0; rb < 8; ++rb) {
            if (rb < nrb) {
              const int row = rb * 16 + (lane & 15), ko = 8 * (lane >> 4);
              v8u16 ah = *(const v8u16*)&Vsl[(0 * HMAXA + row) * VLD + ko];
              v8u16 am = *(const v8u16*)&Vsl[(1 * HMAXA + row) * VLD + ko];
              v8u16 al = *(const v8u16*)&Vsl[(2 * HMAXA + row) * VLD + ko];
              acc[pp][rb] = mfma_bf16(al, bh, acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(am, bm, acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(ah, bl, acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(am, bh, acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(ah, bm, acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(ah, bh, acc[pp][rb]);
            }
          }
        }
      }
    }
    __syncthreads();
  }
  // each wave writes only its own columns
#pragma unroll
  for (int pp = 0; pp < PPW; ++pp) {
    const int pc = w + pp * NW;
    if (pc >= npanel) continue;
    const int col = pc * 16 + (lane & 15);
    if (col >= b) continue;
#pragma unroll
    for (int rb = 0; rb < 8; ++rb) {
      if (rb >= nrb) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = rb * 16 + ((MODE == 0) ? mfma16_row_f64(lane, r) : mfma16_row_f32(lane, r));
        if (m >= a) continue;
        if constexpr (TO_GLOBAL)
          Ub[(long long)(I.p + 2 * m) * n + I.q + 2 * col] = (TIO)acc[pp][rb][r];
        else
          Wk[m * LDW + col] = (TS)acc[pp][rb][r];
      }
    }
  }
}

template <int MODE, int NW, typename TIO>
__global__ __launch_bounds__(NW * 64) void helm_ragged_kernel(
    const HelmItem* __restrict__ items, int nitems, int* __restrict__ counter,
    const double* __restrict__ Vd, const double* __restrict__ VTd,
    const unsigned short* __restrict__ Vs, const unsigned short* __restrict__ VTs,
    const double* __restrict__ lam, const TIO* __restrict__ F, TIO* __restrict__ U,
    double* __restrict__ scratch) {
  typedef typename HelmTypes<MODE>::TS TS;
  constexpr int LDW = HelmLDS<MODE>::LDW;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  TS* Wk = reinterpret_cast<TS*>(smem_raw);
  unsigned char* vslab = smem_raw + HelmLDS<MODE>::WK;
  double* Hs = reinterpret_cast<double*>(smem_raw + HelmLDS<MODE>::WK + HelmLDS<MODE>::VS);  // h_c + h_{c+2}
  double* H2 = Hs + HMAXA;                                                                 // h_{c+2}
  double* Dg = H2 + HMAXA;                                                                // 4c + 6
  __shared__ int s_item;
  const int tid = threadIdx.x;
  double* cp = scratch + (long long)blockIdx.x * HMAXA * HMAXA;

  for (;;) {
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= nitems) break;
    const HelmItem I = items[it];
    const int n = I.n, p = I.p, q = I.q, a = I.a, b = I.b;
    const TIO* Fb = F + I.foff;

    // ---------------- phase 0: F_pq -> Wk, per-column constants -> LDS
    for (int e = tid; e < a * b; e += NW * 64) {
      const int r = e / b, c = e - r * b;
      Wk[r * LDW + c] = (TS)Fb[(long long)(p + 2 * r) * n + q + 2 * c];
    }
    for (int j = tid; j < b; j += NW * 64) {
      const int c = q + 2 * j;
      const double hc = 2.0 / (2 * c + 1), hc2 = 2.0 / (2 * c + 5);
      Hs[j] = hc + hc2;
      H2[j] = hc2;
      Dg[j] = 4.0 * c + 6.0;
    }
    __syncthreads();

    // ---------------- phase 1: G = V_p^T F_pq (in place, column-owned)
    helm_gemm_phase<MODE, NW, TIO, false>(I, Vd, VTs, Wk, vslab, nullptr);
    __syncthreads();

    // ---------------- phase 2: FP64 tridiagonal solves, two lanes per row
    // row r: T u = g with T = s M_q + diag(4c + 6), s = alpha - lambda_r:
    //   T[j][j] = s (h_c + h_{c+2}) + 4c + 6,  T[j][j+1] = T[j+1][j] = -s h_{c+2}
    // lane pair (r, 0) eliminates rows 0..m-1 downwards, (r, 1) rows b-1..m+1
    // upwards (twisted factorisation); both meet at m = b / 2.
    for (int t = tid; t < 2 * HMAXA; t += NW * 64) {
      const int r = t >> 1, side = t & 1;
      const bool act = r < a;
      const double s = act ? I.alpha - lam[I.loff + r] : 1.0;
      TS* row = Wk + (act ? r : 0) * LDW;
      const int m = b >> 1;
      double cl = 0.0, dl = 0.0;  // multiplier / rhs of the last eliminated row next to m
      if (act) {
        if (side == 0) {
          double cprev = 0.0, dprev = 0.0, offprev = 0.0;
          for (int j = 0; j < m; ++j) {
            const double off = -s * H2[j];
            const double den = fma(s, Hs[j], Dg[j]) - offprev * cprev;
            const double inv = rcp64(den);
            const double cj = off * inv;
            const double dj = ((double)row[j] - offprev * dprev) * inv;
            cp[(long long)j * HMAXA + r] = cj;
            row[j] = (TS)dj;
            cprev = cj;
            dprev = dj;
            offprev = off;
          }
          cl = cprev;
          dl = dprev;
        } else {
          double fnext = 0.0, enext = 0.0;
          for (int j = b - 1; j > m; --j) {
            const double offdn = -s * H2[j - 1];                    // couples j and j-1
            const double offup = (j < b - 1) ? -s * H2[j] : 0.0;    // couples j and j+1
            const double den = fma(s, Hs[j], Dg[j]) - offup * fnext;
            const double inv = rcp64(den);
            const double fj = offdn * inv;
            const double ej = ((double)row[j] - offup * enext) * inv;
            cp[(long long)j * HMAXA + r] = fj;
            row[j] = (TS)ej;
            fnext = fj;
            enext = ej;
          }
          cl = fnext;
          dl = enext;
        }
      }
      // exchange with the partner lane (adjacent lane of the same wave)
      const double co = __shfl_xor(cl, 1), dO = __shfl_xor(dl, 1);
      if (act) {
        const double ct = side == 0 ? cl : co, dt = side == 0 ? dl : dO;  // top part (rows < m)
        const double fb = side == 0 ? co : cl, eb = side == 0 ? dO : dl;  // bottom part (rows > m)
        const double am = (m > 0) ? -s * H2[m - 1] : 0.0;
        const double cm = (m < b - 1) ? -s * H2[m] : 0.0;
        const double um = ((double)row[m] - am * dt - cm * eb) / (fma(s, Hs[m], Dg[m]) - am * ct - cm * fb);
        if (side == 0) {
          double unext = um;
          int j = m - 1;
          for (; j >= 7; j -= 8) {  // prefetch 8 multipliers per chunk
            double c8[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) c8[u] = cp[(long long)(j - u) * HMAXA + r];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const double v = (double)row[j - u] - c8[u] * unext;
              row[j - u] = (TS)v;
              unext = v;
            }
          }
          for (; j >= 0; --j) {
            const double v = (double)row[j] - cp[(long long)j * HMAXA + r] * unext;
            row[j] = (TS)v;
            unext = v;
          }
          row[m] = (TS)um;
        } else {
          double uprev = um;
          int j = m + 1;
          for (; j + 7 < b; j += 8) {
            double f8[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) f8[u] = cp[(long long)(j + u) * HMAXA + r];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const double v = (double)row[j + u] - f8[u] * uprev;
              row[j + u] = (TS)v;
              uprev = v;
            }
          }
          for (; j < b; ++j) {
            const double v = (double)row[j] - cp[(long long)j * HMAXA + r] * uprev;
            row[j] = (TS)v;
            uprev = v;
          }
        }
      }
    }
    __syncthreads();

    // ---------------- phase 3: U_pq = V_p U~ -> global
    helm_gemm_phase<MODE, NW, TIO, true>(I, VTd, Vs, Wk, vslab, U + I.foff);
    __syncthreads();
  }
}

template <int MODE, int NW, typename TIO>
static int launch_helm(const HelmItem* items, int nitems, int* counter, const double* Vd,
                       const double* VTd, const unsigned short* Vs, const unsigned short* VTs,
                       const double* lam, const TIO* F, TIO* U, double* scratch, int nblocks,
                       hipStream_t st) {
  size_t shm = HelmLDS<MODE>::WK + HelmLDS<MODE>::VS + HTAB;
  auto kern = helm_ragged_kernel<MODE, NW, TIO>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) return set_error("helm: hipFuncSetAttribute: %s", hipGetErrorString(e)), JFX_EHIP;
  hipLaunchKernelGGL(helm_reset, dim3(1), dim3(64), 0, st, counter);
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(NW * 64), shm, st, items, nitems, counter, Vd, VTd, Vs, VTs,
                     lam, F, U, scratch);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // namespace jfx

using namespace jfx;

extern "C" {

int jfx_helm_item_size(void) { return (int)sizeof(HelmItem); }

int jfx_helm_scratch_doubles_per_block(void) { return HMAXA * HMAXA; }

// mode 0: FP64 (F, U double; Vd/VTd used)
// mode 1: BF16 3-term split, FP32 storage (F, U float if io_f32 else double; Vs/VTs used)
// nwaves: 4 or 8 waves per workgroup; nblocks: persistent grid size
//   (scratch must hold nblocks * jfx_helm_scratch_doubles_per_block() doubles).
int jfx_helm_ragged(int mode, int io_f32, int nwaves, const void* items, int nitems, int* counter,
                    const double* Vd, const double* VTd, const unsigned short* Vs,
                    const unsigned short* VTs, const double* lam, const void* F, void* U,
                    double* scratch, int nblocks, hipStream_t stream) {
  if (!items || !counter || !lam || !F || !U || !scratch || nitems < 0 || nblocks <= 0)
    return set_error("helm_ragged: bad arguments"), JFX_EARG;
  if (nitems == 0) return JFX_OK;
  const HelmItem* it = (const HelmItem*)items;
  if (mode == 0) {
    if (!Vd || !VTd || io_f32) return set_error("helm_ragged: mode 0 needs fp64 tables and fp64 I/O"), JFX_EARG;
    if (nwaves == 4)
      return launch_helm<0, 4, double>(it, nitems, counter, Vd, VTd, Vs, VTs, lam, (const double*)F,
                                       (double*)U, scratch, nblocks, stream);
    if (nwaves == 8)
      return launch_helm<0, 8, double>(it, nitems, counter, Vd, VTd, Vs, VTs, lam, (const double*)F,
                                       (double*)U, scratch, nblocks, stream);
  } else if (mode == 1) {
    if (!Vs || !VTs) return set_error("helm_ragged: mode 1 needs bf16 slice tables"), JFX_EARG;
    if (io_f32) {
      if (nwaves == 4)
        return launch_helm<1, 4, float>(it, nitems, counter, Vd, VTd, Vs, VTs, lam, (const float*)F,
                                        (float*)U, scratch, nblocks, stream);
      if (nwaves == 8)
        return launch_helm<1, 8, float>(it, nitems, counter, Vd, VTd, Vs, VTs, lam, (const float*)F,
                                        (float*)U, scratch, nblocks, stream);
    } else {
      if (nwaves == 4)
        return launch_helm<1, 4, double>(it, nitems, counter, Vd, VTd, Vs, VTs, lam, (const double*)F,
                                         (double*)U, scratch, nblocks, stream);
      if (nwaves == 8)
        return launch_helm<1, 8, double>(it, nitems, counter, Vd, VTd, Vs, VTs, lam, (const double*)F,
                                         (double*)U, scratch, nblocks, stream);
    }
  }
  return set_error("helm_ragged: bad mode/nwaves"), JFX_EARG;
}

}  // extern "C"
