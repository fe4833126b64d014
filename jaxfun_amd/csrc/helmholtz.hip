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
#include "jfx_common.h"

namespace jfx {

struct HelmItem {
  long long foff;   // offset of this solve's F / U (n*n, row-major)
  long long voff;   // offset of the (AP x AP) eigenvector table (parity p)
  double alpha;
  int n, p, q, a, b, AP, loff, pad;
};

constexpr int HMAXA = 128;  // a, b <= 128  (n <= 256)
constexpr int LDW = 129;    // LDS row stride of the work block

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

template <int MODE, int NW, typename TIO>
__global__ __launch_bounds__(NW * 64) void helm_ragged_kernel(
    const HelmItem* __restrict__ items, int nitems, int* __restrict__ counter,
    const double* __restrict__ Vd, const double* __restrict__ VTd,
    const unsigned short* __restrict__ Vs, const unsigned short* __restrict__ VTs,
    const double* __restrict__ lam, const TIO* __restrict__ F, TIO* __restrict__ U,
    double* __restrict__ scratch) {
  typedef typename HelmTypes<MODE>::TS TS;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  TS* Wk = reinterpret_cast<TS*>(smem_raw);
  __shared__ int s_item;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double* cp = scratch + (long long)blockIdx.x * HMAXA * HMAXA;  // c'_j of the row solves

  for (;;) {
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= nitems) break;
    const HelmItem I = items[it];
    const int n = I.n, p = I.p, q = I.q, a = I.a, b = I.b, AP = I.AP;
    const TIO* Fb = F + I.foff;
    TIO* Ub = U + I.foff;
    const int npanel = (b + 15) >> 4;
    const int nrb = AP >> 4;

    // ---------------- phase 1: G = V_p^T F_pq  -> Wk (a x b)
    for (int pc = w; pc < npanel; pc += NW) {
      const int col = pc * 16 + (lane & 15);
      const bool cok = col < b;
      if constexpr (MODE == 0) {
        v4d acc[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = v4d{0, 0, 0, 0};
        for (int k0 = 0; k0 < AP; k0 += 4) {
          const int k = k0 + (lane >> 4);
          double bv = (cok && k < a) ? (double)Fb[(long long)(p + 2 * k) * n + q + 2 * col] : 0.0;
#pragma unroll
          for (int rb = 0; rb < 8; ++rb) {
            if (rb < nrb) {
              double av = Vd[I.voff + (long long)k * AP + rb * 16 + (lane & 15)];  // V[k][m]
              acc[rb] = mfma_f64(av, bv, acc[rb]);
            }
          }
        }
#pragma unroll
        for (int rb = 0; rb < 8; ++rb)
          if (rb < nrb && cok)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              int m = rb * 16 + mfma16_row_f64(lane, r);
              if (m < a) Wk[m * LDW + col] = (TS)acc[rb][r];
            }
      } else {
        v4f acc[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = v4f{0, 0, 0, 0};
        for (int k0 = 0; k0 < AP; k0 += 32) {
          v8u16 bh, bm, bl;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k = k0 + 8 * (lane >> 4) + e;
            float x = (cok && k < a) ? (float)Fb[(long long)(p + 2 * k) * n + q + 2 * col] : 0.f;
            unsigned short h, m_, l;
            split3(x, h, m_, l);
            bh[e] = h;
            bm[e] = m_;
            bl[e] = l;
          }
          const long long sl = (long long)AP * AP;
#pragma unroll
          for (int rb = 0; rb < 8; ++rb) {
            if (rb < nrb) {
              const long long base = 3 * I.voff + (long long)(rb * 16 + (lane & 15)) * AP + k0 + 8 * (lane >> 4);
              v8u16 ah = *(const v8u16*)(VTs + base);
              v8u16 am = *(const v8u16*)(VTs + base + sl);
              v8u16 al = *(const v8u16*)(VTs + base + 2 * sl);
              acc[rb] = mfma_bf16(al, bh, acc[rb]);
              acc[rb] = mfma_bf16(am, bm, acc[rb]);
              acc[rb] = mfma_bf16(ah, bl, acc[rb]);
              acc[rb] = mfma_bf16(am, bh, acc[rb]);
              acc[rb] = mfma_bf16(ah, bm, acc[rb]);
              acc[rb] = mfma_bf16(ah, bh, acc[rb]);
            }
          }
        }
#pragma unroll
        for (int rb = 0; rb < 8; ++rb)
          if (rb < nrb && cok)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              int m = rb * 16 + mfma16_row_f32(lane, r);
              if (m < a) Wk[m * LDW + col] = (TS)acc[rb][r];
            }
      }
    }
    __syncthreads();

    // ---------------- phase 2: FP64 tridiagonal solve along every row
    // row r: (s M_q + diag(4c + 6)) u = g,  s = alpha - lambda_r, c = q + 2j,
    // M_q[j][j] = h_c + h_{c+2}, M_q[j][j+1] = -h_{c+2}, h_c = 2 / (2c + 1).
    for (int r = tid; r < a; r += NW * 64) {
      const double s = I.alpha - lam[I.loff + r];
      TS* row = Wk + r * LDW;
      double wprev = 1.0, gprev = 0.0, offprev = 0.0;
      for (int j = 0; j < b; ++j) {
        const int c = q + 2 * j;
        const double hc = 2.0 / (2 * c + 1), hc2 = 2.0 / (2 * c + 5);
        const double diag = s * (hc + hc2) + (4.0 * c + 6.0);
        const double off = -s * hc2;  // couples j and j+1
        const double mlt = (j > 0) ? offprev / wprev : 0.0;
        const double wj = diag - mlt * offprev;
        const double gj = (double)row[j] - mlt * gprev;
        const double inv = 1.0 / wj;
        cp[(long long)j * HMAXA + r] = off * inv;  // c'_j
        row[j] = (TS)(gj * inv);                    // d'_j
        wprev = wj;
        gprev = gj;
        offprev = off;
      }
      double unext = 0.0;
      for (int j = b - 1; j >= 0; --j) {
        double u = (double)row[j] - ((j < b - 1) ? cp[(long long)j * HMAXA + r] * unext : 0.0);
        row[j] = (TS)u;
        unext = u;
      }
    }
    __syncthreads();

    // ---------------- phase 3: U_pq = V_p U~   (Wk -> global)
    for (int pc = w; pc < npanel; pc += NW) {
      const int col = pc * 16 + (lane & 15);
      const bool cok = col < b;
      if constexpr (MODE == 0) {
        v4d acc[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = v4d{0, 0, 0, 0};
        for (int k0 = 0; k0 < AP; k0 += 4) {
          const int k = k0 + (lane >> 4);
          double bv = (cok && k < a) ? (double)Wk[k * LDW + col] : 0.0;
#pragma unroll
          for (int rb = 0; rb < 8; ++rb) {
            if (rb < nrb) {
              double av = VTd[I.voff + (long long)k * AP + rb * 16 + (lane & 15)];  // V[m][k]
              acc[rb] = mfma_f64(av, bv, acc[rb]);
            }
          }
        }
#pragma unroll
        for (int rb = 0; rb < 8; ++rb)
          if (rb < nrb && cok)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              int m = rb * 16 + mfma16_row_f64(lane, r);
              if (m < a) Ub[(long long)(p + 2 * m) * n + q + 2 * col] = (TIO)acc[rb][r];
            }
      } else {
        v4f acc[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = v4f{0, 0, 0, 0};
        for (int k0 = 0; k0 < AP; k0 += 32) {
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
          const long long sl = (long long)AP * AP;
#pragma unroll
          for (int rb = 0; rb < 8; ++rb) {
            if (rb < nrb) {
              const long long base = 3 * I.voff + (long long)(rb * 16 + (lane & 15)) * AP + k0 + 8 * (lane >> 4);
              v8u16 ah = *(const v8u16*)(Vs + base);
              v8u16 am = *(const v8u16*)(Vs + base + sl);
              v8u16 al = *(const v8u16*)(Vs + base + 2 * sl);
              acc[rb] = mfma_bf16(al, bh, acc[rb]);
              acc[rb] = mfma_bf16(am, bm, acc[rb]);
              acc[rb] = mfma_bf16(ah, bl, acc[rb]);
              acc[rb] = mfma_bf16(am, bh, acc[rb]);
              acc[rb] = mfma_bf16(ah, bm, acc[rb]);
              acc[rb] = mfma_bf16(ah, bh, acc[rb]);
            }
          }
        }
#pragma unroll
        for (int rb = 0; rb < 8; ++rb)
          if (rb < nrb && cok)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              int m = rb * 16 + mfma16_row_f32(lane, r);
              if (m < a) Ub[(long long)(p + 2 * m) * n + q + 2 * col] = (TIO)acc[rb][r];
            }
      }
    }
    __syncthreads();
  }
}

template <int MODE, int NW, typename TIO>
static int launch_helm(const HelmItem* items, int nitems, int* counter, const double* Vd,
                       const double* VTd, const unsigned short* Vs, const unsigned short* VTs,
                       const double* lam, const TIO* F, TIO* U, double* scratch, int nblocks,
                       hipStream_t st) {
  typedef typename HelmTypes<MODE>::TS TS;
  size_t shm = (size_t)HMAXA * LDW * sizeof(TS);
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
