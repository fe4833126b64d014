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
// Work items sit in nq queues (the host groups items that share eigenvector
// tables into one queue, largest first within it).  Workgroup g starts on
// queue g mod nq -- workgroups are dispatched round-robin over the XCDs, so a
// queue's tables stay in one XCD's L2 -- and moves on to the next queues when
// its own is drained; every workgroup loops until all queues are empty.
//   MODE 0: FP64 MFMA, FP64 storage
//   MODE 1: BF16 3-term float split (6 products) MFMA, FP32 accumulation and
//           storage, FP64 tridiagonal solve.
//   mode 2 (compact kernel only): FP16 2-term split (3 products) of operands
//           scaled by powers of two, otherwise as MODE 1.
#include <mutex>
#include <type_traits>

#include "jfx_common.h"

namespace jfx {

struct HelmItem {
  long long foff;   // offset of this solve's F / U (n*n, row-major)
  long long voff;   // offset of the (AP x AP) eigenvector table (parity p)
  double alpha;
  int n, p, q, a, b, AP, loff;
  int texp;         // fp16 split: the table holds V * 2^texp (else 0)
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

__global__ void helm_reset(int* counters, int nq) {
  if ((int)threadIdx.x < nq) counters[threadIdx.x] = 0;
}

// Work queues: queue g holds items[bounds[g] .. bounds[g+1]), claimed through counters[g].
constexpr int HMAXQ = 16;
struct HelmQueues {
  int* counters;
  const int* bounds;
  int nq;
  // Next item index, claimed from queue qi by lane 0 of the calling wave (all
  // 64 lanes call; every lane gets the result).  Once queue qi is drained the
  // lanes read all counters at once and the next queue (cyclically) that
  // still looks non-empty is tried; qi follows.  Returns bounds[nq] when every
  // queue is empty.  Counters only grow, so the loop ends.
  __device__ __forceinline__ int claim(int& qi) const {
    const int lane = threadIdx.x & 63;
    int c = 0;
    if (lane == 0) c = atomicAdd(&counters[qi], 1);
    c = __shfl(c, 0);
    if (bounds[qi] + c < bounds[qi + 1]) return bounds[qi] + c;
    for (;;) {
      const int cg = lane < nq ? __hip_atomic_load(&counters[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
      const unsigned long long m = __ballot(lane < nq && bounds[lane] + cg < bounds[lane + 1]);
      if (m == 0) return bounds[nq];
      const unsigned long long after = m & ~((2ull << qi) - 1);  // non-empty queues past qi
      qi = __ffsll((long long)(after ? after : m)) - 1;
      if (lane == 0) c = atomicAdd(&counters[qi], 1);
      c = __shfl(c, 0);
      if (bounds[qi] + c < bounds[qi + 1]) return bounds[qi] + c;
    }
  }
};

// FP64 reciprocal refined by one Newton step from the hardware estimate:
// ~2^-46 relative, ample where the results are stored in FP32 (the mixed
// kernels' multipliers and sweep values).
__device__ __forceinline__ double rcp64_1(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
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
struct HelmLDS<0> {  // FP64: work block 128 x 129 doubles, V slab 16 x 129 doubles; multipliers in global scratch
  static constexpr int LDW = 129, KS = 16;
  static constexpr size_t WK = (size_t)HMAXA * LDW * 8, U1 = (size_t)KS * (HMAXA + 1) * 8;
};
template <>
struct HelmLDS<1> {  // bf16 split: work block 128 x 130 floats; V slab (3 bf16 slices of 128 x 32) shares
                     // its space with the 128 x 128 FP32 tridiagonal multipliers
  static constexpr int LDW = 130, KS = 32, VLD = 40;
  static constexpr size_t WK = (size_t)HMAXA * LDW * 4, U1 = (size_t)HMAXA * HMAXA * 4;
};
constexpr size_t HTAB = 3 * HMAXA * 8;

// One GEMM phase (a x b block):   C = A * Wk   with A an (AP x AP) table applied by rows
//   phase 1: A = V^T, result back into Wk (each wave owns its column panels)
//   phase 3: A = V,   result scattered to U
// The next k-slab of A is prefetched into registers while the current one is multiplied.
template <int MODE, int NW, typename TIO, bool TO_GLOBAL>
__device__ __forceinline__ void helm_gemm_phase(const HelmItem& I, int q, int b, const double* __restrict__ tabd,
                                                const unsigned short* __restrict__ tabs,
                                                typename HelmTypes<MODE>::TS* Wk, unsigned char* vslab,
                                                TIO* __restrict__ Ub) {
  typedef typename HelmTypes<MODE>::TS TS;
  constexpr int LDW = HelmLDS<MODE>::LDW, KS = HelmLDS<MODE>::KS, NT = NW * 64;
  constexpr int PPW = 8 / NW;  // column panels per wave (b <= 128 -> <= 8 panels)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int a = I.a, AP = I.AP, n = I.n;
  const int npanel = (b + 15) >> 4, nrb = (a + 15) >> 4;  // output row blocks
  typedef typename std::conditional<MODE == 0, v4d, v4f>::type vacc;
  vacc acc[PPW][8];
#pragma unroll
  for (int pp = 0; pp < PPW; ++pp)
#pragma unroll
    for (int rb = 0; rb < 8; ++rb) acc[pp][rb] = vacc{0, 0, 0, 0};

  // register prefetch of one k-slab
  constexpr int NPRE = (MODE == 0) ? (KS * HMAXA / NT) : (3 * HMAXA * 4 / NT);
  typedef typename std::conditional<MODE == 0, double, v8u16>::type PT;
  PT pre[NPRE];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NPRE; ++u) {
      const int e = tid + u * NT;
      if constexpr (MODE == 0) {
        if (e < KS * AP) {
          const int kk = e / AP, m = e - kk * AP;
          pre[u] = tabd[I.voff + (long long)(k0 + kk) * AP + m];
        }
      } else {
        if (e < 3 * AP * 4) {
          const int sl = e / (AP * 4), rem = e - sl * AP * 4, m = rem >> 2, part = rem & 3;
          pre[u] = *(const v8u16*)(tabs + 3 * I.voff + (long long)sl * AP * AP + (long long)m * AP + k0 + part * 8);
        }
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int u = 0; u < NPRE; ++u) {
      const int e = tid + u * NT;
      if constexpr (MODE == 0) {
        if (e < KS * AP) {
          const int kk = e / AP, m = e - kk * AP;
          reinterpret_cast<double*>(vslab)[kk * (HMAXA + 1) + m] = pre[u];
        }
      } else {
        constexpr int VLD = HelmLDS<1>::VLD;
        if (e < 3 * AP * 4) {
          const int sl = e / (AP * 4), rem = e - sl * AP * 4, m = rem >> 2, part = rem & 3;
          *(v8u16*)&reinterpret_cast<unsigned short*>(vslab)[(sl * HMAXA + m) * VLD + part * 8] = pre[u];
        }
      }
    }
  };

  fetch(0);
  for (int k0 = 0; k0 < AP; k0 += KS) {
    store();
    __syncthreads();
    if (k0 + KS < AP) fetch(k0 + KS);
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
          float xv[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k = k0 + 8 * (lane >> 4) + e;
            xv[e] = (cok && k < a) ? (float)Wk[k * LDW + col] : 0.f;
          }
          split3_pk8(xv, bh, bm, bl);
#pragma unroll
          for (int rb = 0; rb < 8; ++rb) {
            if (rb < nrb) {
              const int row = rb * 16 + (lane & 15), ko = 8 * (lane >> 4);
              v8u16 ah = *(const v8u16*)&Vsl[(0 * HMAXA + row) * VLD + ko];
              v8u16 am = *(const v8u16*)&Vsl[(1 * HMAXA + row) * VLD + ko];
              v8u16 al = *(const v8u16*)&Vsl[(2 * HMAXA + row) * VLD + ko];
              // products ordered small to large
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
          Ub[(long long)(I.p + 2 * m) * n + q + 2 * col] = (TIO)acc[pp][rb][r];
        else
          Wk[m * LDW + col] = (TS)acc[pp][rb][r];
      }
    }
  }
}

// Multiplier storage of the row solves: FP32 LDS (mixed) or FP64 global scratch.
template <int MODE>
struct CPStore {
  float* lds;
  double* glob;
  __device__ __forceinline__ void put(int j, int r, double v) const {
    if constexpr (MODE == 1)
      lds[j * HMAXA + r] = (float)v;
    else
      glob[(long long)j * HMAXA + r] = v;
  }
  __device__ __forceinline__ double get(int j, int r) const {
    if constexpr (MODE == 1)
      return (double)lds[j * HMAXA + r];
    else
      return glob[(long long)j * HMAXA + r];
  }
};

// Rescale a continuant pair by the same power of two (exact).
__device__ __forceinline__ void renorm(double& P1, double& P2) {
  const int e = ilogb(P1);
  P1 = ldexp(P1, -e);
  P2 = ldexp(P2, -e);
}

// Gather the parity block F[p::2, q::2] (a x b) into the LDS work block:
// lane l of wave w takes columns l and l + 64 of rows w, w + NW, ...; a
// round's loads are all issued before its LDS stores, and no index needs an
// integer division.
template <int NW, int LDW, typename TIO, typename TS>
__device__ __forceinline__ void helm_gather(const TIO* __restrict__ Fb, int n, int p, int q, int a, int b,
                                            TS* Wk) {
  constexpr int R = sizeof(TIO) == 4 ? 16 : 8;  // rows per wave per round
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool ok0 = lane < b, ok1 = lane + 64 < b;
  const long long rs = 2LL * n;  // one parity row
  for (int r0 = w; r0 < a; r0 += R * NW) {
    const TIO* src = Fb + (long long)(p + 2 * r0) * n + q + 2 * lane;
    TIO v0[R], v1[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if (r0 + i * NW < a) {
        const TIO* s = src + i * NW * rs;
        if (ok0) v0[i] = s[0];
        if (ok1) v1[i] = s[128];
      }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int r = r0 + i * NW;
      if (r < a) {
        if (ok0) Wk[r * LDW + lane] = (TS)v0[i];
        if (ok1) Wk[r * LDW + lane + 64] = (TS)v1[i];
      }
    }
  }
}

template <int MODE, int NW, typename TIO>
__global__ __launch_bounds__(NW * 64) void helm_ragged_kernel(
    const HelmItem* __restrict__ items, int nitems, HelmQueues Q,
    const double* __restrict__ Vd, const double* __restrict__ VTd,
    const unsigned short* __restrict__ Vs, const unsigned short* __restrict__ VTs,
    const double* __restrict__ lam, const TIO* __restrict__ F, TIO* __restrict__ U,
    double* __restrict__ mult_scratch) {
  typedef typename HelmTypes<MODE>::TS TS;
  constexpr int LDW = HelmLDS<MODE>::LDW, NT = NW * 64;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  TS* Wk = reinterpret_cast<TS*>(smem_raw);
  unsigned char* u1 = smem_raw + HelmLDS<MODE>::WK;
  double* Hs = reinterpret_cast<double*>(smem_raw + HelmLDS<MODE>::WK + HelmLDS<MODE>::U1);  // h_c + h_{c+2}
  double* H2 = Hs + HMAXA;                                                                 // h_{c+2}
  double* Dg = H2 + HMAXA;                                                                 // 4c + 6
  __shared__ int s_nx[2];
  const int tid = threadIdx.x;
  CPStore<MODE> cps{reinterpret_cast<float*>(u1), mult_scratch + (long long)blockIdx.x * HMAXA * HMAXA};

  // the next item's index is claimed at the start of the current one, so the
  // counter round trip overlaps the work (two slots: slot k is rewritten only
  // after every thread has passed the barriers of the item in between)
  int qi = blockIdx.x % Q.nq;
  if (tid < 64) {
    const int first = Q.claim(qi);
    if (tid == 0) s_nx[0] = first;
  }
  __syncthreads();
  int it = s_nx[0], slot = 1;
  while (it < nitems) {
    const HelmItem I = items[it];
    int nxt = 0;
    if (tid < 64) nxt = Q.claim(qi);
    const int n = I.n, p = I.p, a = I.a;
    const TIO* Fb = F + I.foff;

    // both column parities of this (solve, row parity) in turn: the second
    // pass re-reads / re-writes the other half of the same lines from L2
    for (int q = 0; q < 2; ++q) {
      const int b = (n - q + 1) >> 1;
      if (b == 0) continue;
      // ---------------- phase 0: F_pq -> Wk, column constants
      helm_gather<NW, LDW>(Fb, n, p, q, a, b, Wk);
      for (int j = tid; j < b; j += NT) {
        const int c = q + 2 * j;
        const double hc = 2.0 / (2 * c + 1), hc2 = 2.0 / (2 * c + 5);
        Hs[j] = hc + hc2;
        H2[j] = hc2;
        Dg[j] = 4.0 * c + 6.0;
      }
      __syncthreads();

      // ---------------- phase 1: G = V_p^T F_pq (in place, column-owned)
      helm_gemm_phase<MODE, NW, TIO, false>(I, q, b, Vd, VTs, Wk, u1, nullptr);
      __syncthreads();

      // ---------------- phase 2: FP64 tridiagonal solves, two lanes per row
      // row r: T u = g, T = s M_q + diag(4c + 6), s = alpha - lambda_r, c = q + 2j:
      //   d_j = T[j][j] = s (h_c + h_{c+2}) + 4c + 6,  o_j = T[j][j+1] = T[j+1][j] = -s h_{c+2}
      // Twisted elimination: lane (r, 0) sweeps rows 0..m-1 down, lane (r, 1)
      // rows b-1..m+1 up, m = b / 2.  The pivots are ratios of continuants
      // P_j = d_j P_{j-1} - o_{j-1}^2 P_{j-2}  (P_{-1} = 1), so the serial
      // dependence per step is one FMA; reciprocals run off that chain.
      for (int t = tid; t < 2 * HMAXA; t += NT) {
        const int r = t >> 1, side = t & 1;
        const bool act = r < a;
        const double s = act ? I.alpha - lam[I.loff + r] : 1.0;
        TS* row = Wk + (act ? r : 0) * LDW;
        const int m = b >> 1;
        // both lanes of a pair run the same instruction stream; `side` only
        // selects the direction: step i visits j = i (down) or j = b-1-i (up)
        const int len = act ? (side ? b - 1 - m : m) : 0;
        const int maxlen = b - m - 1 > m ? b - m - 1 : m;
        double P1 = 1.0, P2 = 0.0, yprev = 0.0, oprev = 0.0, cl = 0.0;
        for (int i = 0; i < maxlen; ++i) {
          if (i < len) {
            const int j = side ? b - 1 - i : i;
            const int inext = side ? j - 1 : j;  // coupling towards the middle
            const double on = -s * H2[inext];
            const double Pj = fma(fma(s, Hs[j], Dg[j]), P1, -(oprev * oprev) * P2);
            const double rj = P1 * rcp64(Pj);
            const double yj = fma(-oprev, yprev, (double)row[j]) * rj;
            cl = on * rj;
            cps.put(j, r, cl);
            row[j] = (TS)yj;
            P2 = P1;
            P1 = Pj;
            yprev = yj;
            oprev = on;
            if ((i & 7) == 7) renorm(P1, P2);
          }
        }
        const double dl = yprev;
        // exchange with the partner lane (adjacent lane of the same wave)
        const double co = __shfl_xor(cl, 1), dO = __shfl_xor(dl, 1);
        if (act) {
          const double ct = side == 0 ? cl : co, dt = side == 0 ? dl : dO;  // rows < m
          const double fb = side == 0 ? co : cl, eb = side == 0 ? dO : dl;  // rows > m
          const double am = (m > 0) ? -s * H2[m - 1] : 0.0;
          const double cm = (m < b - 1) ? -s * H2[m] : 0.0;
          const double um = ((double)row[m] - am * dt - cm * eb) / (fma(s, Hs[m], Dg[m]) - am * ct - cm * fb);
          // back substitution outwards from m: j = m-1-i (down) or m+1+i (up)
          double uprev = um;
          int i = 0;
          for (; i + 8 <= len; i += 8) {
            double c8[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) c8[u] = cps.get(side ? m + 1 + i + u : m - 1 - i - u, r);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int j = side ? m + 1 + i + u : m - 1 - i - u;
              const double v = (double)row[j] - c8[u] * uprev;
              row[j] = (TS)v;
              uprev = v;
            }
          }
          for (; i < len; ++i) {
            const int j = side ? m + 1 + i : m - 1 - i;
            const double v = (double)row[j] - cps.get(j, r) * uprev;
            row[j] = (TS)v;
            uprev = v;
          }
          if (side == 0) row[m] = (TS)um;
        }
      }
      __syncthreads();

      // ---------------- phase 3: U_pq = V_p U~ -> global
      helm_gemm_phase<MODE, NW, TIO, true>(I, q, b, VTd, Vs, Wk, u1, U + I.foff);
      __syncthreads();
    }
    if (tid == 0) s_nx[slot] = nxt;
    __syncthreads();
    it = s_nx[slot];
    slot ^= 1;
  }
}

// ---------------------------------------------------------------------------
// Compact mixed-precision variant: 4 waves and under 80 KB of LDS, so two
// workgroups share a CU and one's MFMA phases overlap the other's row solves.
//   Wk   : 128 x 130 FP32 work block (as above)
//   slab : 3 bf16 slices x 64 table rows x 32 k, 16-byte chunks swizzled so a
//          16-lane group of b128 reads covers all banks
//   H    : h_{q+2j} = 2 / (2(q+2j) + 1), j = 0..b, and h_{q+2j} + h_{q+2j+2}   (FP64)
// The tridiagonal multipliers live in registers (at most 64 per lane).
struct HelmCompact {
  static constexpr int LDW = 130, NT = 256;
  static constexpr size_t WK = (size_t)HMAXA * LDW * 4;
  static constexpr size_t SLAB = 3 * 64 * 32 * 2;
  static constexpr size_t HT = 2 * (HMAXA + 2) * 8;  // H_j and H_j + H_{j+1}
  static constexpr size_t BYTES = WK + SLAB + HT;
};
static_assert(HelmCompact::BYTES + 64 <= 80 * 1024, "two compact workgroups must fit one CU's LDS");

__device__ __forceinline__ int slab_off(int sl, int row, int ck) {  // in bf16 units
  return ((sl * 64 + row) * 4 + (ck ^ ((row >> 2) & 3))) * 8;
}

// One k-slab of a 3-slice table in flight in registers (3 x 16 bytes per thread).
struct SlabRegs {
  v8u16 v[3];
};

// Issue the loads of slab (k0, half h) of table T (NS AP x AP 16-bit slices:
// three bf16 terms, or the fp16 high and low terms).
template <int NS>
__device__ __forceinline__ void slab_fetch(SlabRegs& P, const unsigned short* __restrict__ T, int AP, int k0, int h,
                                           int tid) {
  const int row = (tid & 255) >> 2, ck = tid & 3, gr = h * 64 + row;
#pragma unroll
  for (int u = 0; u < NS; ++u)
    if (gr < AP) P.v[u] = *(const v8u16*)(T + (long long)u * AP * AP + (long long)gr * AP + k0 + ck * 8);
}

template <int NS>
__device__ __forceinline__ void slab_store(const SlabRegs& P, unsigned short* slab, int AP, int h, int tid) {
  const int row = (tid & 255) >> 2, ck = tid & 3;
#pragma unroll
  for (int u = 0; u < NS; ++u)
    if (h * 64 + row < AP) *(v8u16*)&slab[slab_off(u, row, ck)] = P.v[u];
}

// The first two slabs of a GEMM phase, issued one phase ahead.  NH = number of
// 64-row halves of the table (AP > 64 -> 2); with NH = 1 the two slabs are
// the table's only k-slabs (AP <= 64).
template <int NH, int NS>
__device__ __forceinline__ void slab_prefetch2(SlabRegs (&P)[2], const unsigned short* __restrict__ T, int AP,
                                               int tid) {
  if constexpr (NH == 2) {
    slab_fetch<NS>(P[0], T, AP, 0, 0, tid);
    slab_fetch<NS>(P[1], T, AP, 0, 1, tid);
  } else {
    slab_fetch<NS>(P[0], T, AP, 0, 0, tid);
    if (AP > 32) slab_fetch<NS>(P[1], T, AP, 32, 0, tid);
  }
}

// C = A * Wk for the a x b block; A is the (AP x AP) 3-slice bf16 table T
// (row-major, k contiguous) whose first two slabs are already in flight in P.
// Wave w owns column panels w and w + 4; the B fragments of a k-slab are
// split once and reused over both row halves.  Each slab is refetched two
// steps ahead of its use, and the A fragments of the next row block are read
// from LDS before the current block's MFMAs issue.  Rows a..AP-1 of Wk hold
// zeros and columns >= b only feed discarded outputs, so B is read unmasked.
// `between` runs after the last k-slab (Wk and P are free from then on) and
// before the epilogue; `after` runs after it.
// SP 0: T holds three bf16 terms, B is split into three (six products).
// SP 1: T holds the fp16 high and low terms of V * 2^eA, B is scaled by sb =
//       2^eB and split into two fp16 terms (three products, the low x low
//       product dropped); the outputs are scaled back by 2^-es, es = eA + eB.
template <typename TIO, bool TO_GLOBAL, int NH, int SP, typename Between, typename After>
__device__ __forceinline__ void helmc_gemm(const HelmItem& I, int q, int b, const unsigned short* __restrict__ T,
                                           float* Wk, unsigned short* slab, TIO* __restrict__ Ub,
                                           SlabRegs (&P)[2], float sb, int es, Between&& between, After&& after) {
  constexpr int NS = SP ? 2 : 3;
  constexpr int LDW = HelmCompact::LDW;
  // opaque copy of the thread id: keeps the compiler from hoisting every
  // lane-dependent address of the epilogue out of the item loop into registers
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int a = I.a, AP = I.AP, n = I.n;
  const int npanel = (b + 15) >> 4, nrb = (a + 15) >> 4;  // output row blocks
  const int npp = (w + 4 < npanel) ? 2 : (w < npanel ? 1 : 0);  // panels of this wave
  v4f acc[2][8];
#pragma unroll
  for (int pp = 0; pp < 2; ++pp)
#pragma unroll
    for (int rb = 0; rb < 8; ++rb) acc[pp][rb] = v4f{0, 0, 0, 0};

  v8u16 bh[2], bm[2], bl[2];
  auto bfrag = [&](int k0) {
#pragma unroll
    for (int pp = 0; pp < 2; ++pp) {
      const float* src = Wk + (k0 + 8 * (lane >> 4)) * LDW + (w + 4 * pp) * 16 + (lane & 15);
      float xv[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) xv[e] = src[e * LDW];
      if constexpr (SP) {
#pragma unroll
        for (int e = 0; e < 8; ++e) xv[e] *= sb;
        split2_f16_pk8(xv, bh[pp], bl[pp]);
      } else {
        split3_pk8(xv, bh[pp], bm[pp], bl[pp]);
      }
    }
  };
  auto mult = [&](auto hc) {
    constexpr int h = decltype(hc)::value;
    const int off0 = slab_off(0, lane & 15, lane >> 4);
    v8u16 A[2][3];
    auto aload = [&](v8u16(&dst)[3], int rl) {
      const int off = off0 + rl * 16 * 4 * 8;  // 16 rows further; the swizzle depends on row bits 2-3 only
#pragma unroll
      for (int u = 0; u < NS; ++u) dst[u] = *(const v8u16*)&slab[off + u * 64 * 32];
    };
    aload(A[0], 0);
#pragma unroll
    for (int rl = 0; rl < 4; ++rl) {
      if (rl + 1 < 4) aload(A[(rl + 1) & 1], rl + 1);
      const int rb = h * 4 + rl;
      if (rb < nrb) {
#pragma unroll
        for (int pp = 0; pp < 2; ++pp) {
          if (pp < npp) {
            // products ordered small to large
            if constexpr (SP) {
              const v8u16 ah = A[rl & 1][0], al = A[rl & 1][1];
              acc[pp][rb] = mfma_f16(al, bh[pp], acc[pp][rb]);
              acc[pp][rb] = mfma_f16(ah, bl[pp], acc[pp][rb]);
              acc[pp][rb] = mfma_f16(ah, bh[pp], acc[pp][rb]);
            } else {
              const v8u16 ah = A[rl & 1][0], am = A[rl & 1][1], al = A[rl & 1][2];
              acc[pp][rb] = mfma_bf16(al, bh[pp], acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(am, bm[pp], acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(ah, bl[pp], acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(am, bh[pp], acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(ah, bm[pp], acc[pp][rb]);
              acc[pp][rb] = mfma_bf16(ah, bh[pp], acc[pp][rb]);
            }
          }
        }
      }
    }
  };

  if constexpr (NH == 2) {
    for (int k0 = 0; k0 < AP; k0 += 32) {
      bfrag(k0);
      slab_store<NS>(P[0], slab, AP, 0, tid);
      __syncthreads();
      if (k0 + 32 < AP) slab_fetch<NS>(P[0], T, AP, k0 + 32, 0, tid);
      mult(std::integral_constant<int, 0>());
      __syncthreads();
      slab_store<NS>(P[1], slab, AP, 1, tid);
      __syncthreads();
      if (k0 + 32 < AP) slab_fetch<NS>(P[1], T, AP, k0 + 32, 1, tid);
      mult(std::integral_constant<int, 1>());
      __syncthreads();
    }
  } else {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (32 * s < AP) {
        bfrag(32 * s);
        slab_store<NS>(P[s], slab, AP, 0, tid);
        __syncthreads();
        mult(std::integral_constant<int, 0>());
        __syncthreads();
      }
    }
  }
  between();
  // outputs scaled back by 2^-es (exact) in the fp16 split
  auto outv = [&](float v) -> TIO {
    if constexpr (!SP) return (TIO)v;
    else if constexpr (sizeof(TIO) == 4) return __builtin_ldexpf(v, -es);
    else return __builtin_ldexp((double)v, -es);
  };
  // epilogue: output row m = rb*16 + lr + r, lr = 4 (lane / 16)
  const int lr = 4 * (lane >> 4), lim = a - lr;
#pragma unroll
  for (int pp = 0; pp < 2; ++pp) {
    if (pp >= npp) continue;
    const int col = (w + 4 * pp) * 16 + (lane & 15);
    const bool cok = col < b;
    if constexpr (TO_GLOBAL) {
      TIO* dst = Ub + ((I.p + 2 * lr) * n + q + 2 * col);  // within one solve: 32-bit offsets
      const int rs = 2 * n;                                  // one parity row
#pragma unroll
      for (int rb = 0; rb < 8; ++rb) {
        if (rb >= nrb) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (cok && rb * 16 + r < lim) dst[0] = outv(acc[pp][rb][r]);
          dst += rs;
        }
        dst += 12 * rs;
      }
    } else {
      float* dst = Wk + lr * LDW + col;
#pragma unroll
      for (int rb = 0; rb < 8; ++rb) {
        if (rb >= nrb) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (cok && rb * 16 + r < lim)
            dst[(rb * 16 + r) * LDW] = SP ? __builtin_ldexpf(acc[pp][rb][r], -es) : acc[pp][rb][r];
      }
    }
  }
  after();
}

// Registers of a gather issued one pass ahead: rows w + 4i (i < 32) of the
// next parity block, columns lane and lane + 64 (FP32 I/O only; with FP64
// I/O the next block is gathered after the epilogue instead).
template <typename TIO>
struct GatherRegs {
  float v[sizeof(TIO) == 4 ? 32 : 1][2];
};

template <typename TIO>
__device__ __forceinline__ void gather_issue(GatherRegs<TIO>& G, const TIO* __restrict__ Fb, int n, int p, int q,
                                             int a, int b) {
  if constexpr (sizeof(TIO) == 4) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool ok0 = lane < b, ok1 = lane + 64 < b;
    // running element offset (an integer keeps the global address space)
    int off = (p + 2 * w) * n + q + 2 * lane;  // < n^2 <= 65536
    const int rs = 8 * n;                      // 4 parity rows further per i
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      if (w + 4 * i < a) {
        if (ok0) G.v[i][0] = Fb[off];
        if (ok1) G.v[i][1] = Fb[off + 128];
      }
      off += rs;
      asm volatile("" : "+v"(off));
    }
  }
}

// Largest |entry| of the block entries this thread gathered (rows w + 4i,
// columns lane and lane + 64), read back from LDS (its own stores: no barrier).
__device__ __forceinline__ float own_block_absmax(const float* Wk, int a, int b) {
  constexpr int LDW = HelmCompact::LDW;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool ok0 = lane < b, ok1 = lane + 64 < b;
  float m = 0.f;
  for (int r = w; r < a; r += 4) {
    if (ok0) m = fmaxf(m, fabsf(Wk[r * LDW + lane]));
    if (ok1) m = fmaxf(m, fabsf(Wk[r * LDW + lane + 64]));
  }
  return m;
}

// Stores the next block's gathered registers into the work block; with MAX,
// returns the largest |entry| this thread stored (else 0).
template <typename TIO, bool MAX>
__device__ __forceinline__ float gather_store(const GatherRegs<TIO>& G, const TIO* __restrict__ Fb, int n, int p,
                                              int q, int a, int b, float* Wk) {
  constexpr int LDW = HelmCompact::LDW;
  float m = 0.f;
  if constexpr (sizeof(TIO) == 4) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool ok0 = lane < b, ok1 = lane + 64 < b;
    int o = w * LDW + lane;  // running LDS offset of row w + 4i (one register, not 32)
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      if (w + 4 * i < a) {
        if (ok0) {
          Wk[o] = G.v[i][0];
          if constexpr (MAX) m = fmaxf(m, fabsf(G.v[i][0]));
        }
        if (ok1) {
          Wk[o + 64] = G.v[i][1];
          if constexpr (MAX) m = fmaxf(m, fabsf(G.v[i][1]));
        }
      }
      o += 4 * LDW;
      asm volatile("" : "+v"(o));
    }
  } else {
    helm_gather<4, LDW>(Fb, n, p, q, a, b, Wk);
    if constexpr (MAX) m = own_block_absmax(Wk, a, b);
  }
  return m;
}

// Wave-wide maximum of m, written by lane 0 to slot[wave].
__device__ __forceinline__ void publish_wave_max(float m, float* slot) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) slot[threadIdx.x >> 6] = m;
}

// Exponent e with 2^e * max(slot[0..3]) < 2^14 (0 for an all-zero block; at
// most 100, so that tiny blocks do not overflow the scale).
__device__ __forceinline__ int block_split_exp(const float* slot) {
  const float m = fmaxf(fmaxf(slot[0], slot[1]), fmaxf(slot[2], slot[3]));
  return m > 0.f ? min(14 - __builtin_amdgcn_frexp_expf(m), 100) : 0;
}

// One pass (column parity q) of work item I.  On entry Wk rows < a hold F_pq
// and P holds phase 1's first slabs.  The next pass -- (In, qn) if has_next --
// is set up during this pass's last epilogue: its F block is loaded into
// registers before the output stores and written to Wk after them, and its
// first table slabs are fetched into P.
// SP selects the split (see helmc_gemm); with SP 1, smx[0..3] hold the wave
// maxima of |F_pq| on entry (published before the pass's first barrier) and
// smx[4..7] receive those of U~; the next pass's F maxima go to smx[0..3].
template <typename TIO, int NH, int SP>
__device__ __forceinline__ void helmc_pass(const HelmItem& I, int q, const HelmItem& In, int qn, bool has_next,
                                           const unsigned short* __restrict__ Vs,
                                           const unsigned short* __restrict__ VTs, const double* __restrict__ lam,
                                           const TIO* __restrict__ F, TIO* __restrict__ U, float* Wk,
                                           unsigned short* slab, double* H, SlabRegs (&P)[2], float* smx) {
  constexpr int LDW = HelmCompact::LDW, NT = HelmCompact::NT, NS = SP ? 2 : 3;
  const int tid = threadIdx.x, wu = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int a = I.a, AP = I.AP;
  double* Hs = H + HMAXA + 2;                     // h_c + h_{c+2}
  const unsigned short* TV = Vs + NS * I.voff;    // V_p slices    (phase 3)
  const unsigned short* TVT = VTs + NS * I.voff;  // V_p^T slices  (phase 1)
  const int b = (I.n - q + 1) >> 1;
  {
    // ---------------- phase 0: pad rows, h tables (F_pq is already in Wk)
    for (int e = tid; e < (AP - a) * 128; e += NT) Wk[(a + (e >> 7)) * LDW + (e & 127)] = 0.f;
    for (int j = tid; j <= b; j += NT) {
      const double hj = 2.0 / (2 * (q + 2 * j) + 1);
      H[j] = hj;
      Hs[j] = hj + 2.0 / (2 * (q + 2 * j) + 5);
    }
    __syncthreads();

    // ---------------- phase 1: G = V_p^T F_pq (in place)
    int e1 = 0;
    if constexpr (SP) e1 = block_split_exp(smx);
    helmc_gemm<TIO, false, NH, SP>(I, q, b, TVT, Wk, slab, nullptr, P, __builtin_ldexpf(1.f, e1), I.texp + e1,
                                   [] {}, [] {});
    slab_prefetch2<NH, NS>(P, TV, AP, tid);  // phase 3's first slabs fly during the row solves
    __syncthreads();

    // ---------------- phase 2: FP64 row solves, two lanes per row (twisted)
    // T = s M_q + diag(4c + 6), c = q + 2j:  d_j = s Hs_j + 4q + 6 + 8j, Hs_j = H_j + H_{j+1},
    // o_j = T[j][j+1] = -s H_{j+1}.  Lane (r, 0) eliminates rows 0..len-1 downwards,
    // lane (r, 1) rows b-1..b-len upwards (step i visits j = i or b-1-i), len = b/2,
    // so both lanes run the same number of steps.  Pivots are continuant ratios
    // P_{j-1} / P_j with P_j = d_j P_{j-1} - o^2 P_{j-2}.  Odd b leaves the middle
    // row, solved from both sides' last multipliers; even b leaves a 2 x 2 system
    // between the two sides' last rows.  Full chunks of 8 steps run without
    // branches so their loads and reciprocals overlap.
    float umax = 0.f;  // largest |u~| this lane wrote (fp16 split only)
    if (wu * 32 < a) {
      // opaque lane index: stops per-step offsets (multiples of the lane's
      // direction) from being hoisted out of the item loop into registers
      int t2 = tid;
      asm volatile("" : "+v"(t2));
      const int r = t2 >> 1, side = t2 & 1;
      const bool act = r < a;
      const double s = act ? I.alpha - lam[I.loff + r] : 1.0;
      // rows >= a hold the zero padding the GEMMs read: idle lanes work on the slab instead
      float* row = act ? Wk + r * LDW : reinterpret_cast<float*>(slab);
      const int len = b >> 1;
      const int j0 = side ? b - 1 : 0, dj = side ? -1 : 1;
      const double* Ho = side ? H : H + 1;  // o towards the middle = -s Ho[j]
      const double dg0 = fma(8.0, (double)j0, 4.0 * q + 6.0), ddg = 8.0 * dj;
      float cm[64];
      double P1 = 1.0, P2 = 0.0, yprev = 0.0, oprev = 0.0, cl = 0.0;
      auto fwd_v = [&](int i, double g, double hs, double ho) {
        const int j = j0 + dj * i;
        const double d = fma(s, hs, fma((double)i, ddg, dg0));
        const double on = -s * ho;
        const double Pj = fma(d, P1, -(oprev * oprev) * P2);
        const double rj = P1 * rcp64_1(Pj);
        const double yj = fma(-oprev, yprev, g) * rj;
        cl = on * rj;
        cm[i] = (float)cl;
        row[j] = (float)yj;
        P2 = P1;
        P1 = Pj;
        yprev = yj;
        oprev = on;
      };
      auto fwd = [&](int i) {
        const int j = j0 + dj * i;
        fwd_v(i, row[j], Hs[j], Ho[j]);
      };
#pragma unroll
      for (int ch = 0; ch < 8; ++ch) {
        if (ch * 8 + 8 <= len) {
          // the chunk's loads all issue before its arithmetic (each step reads
          // its own row entry before writing it)
          float g[8];
          double hs[8], ho[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int j = j0 + dj * (ch * 8 + u);
            g[u] = row[j];
            hs[u] = Hs[j];
            ho[u] = Ho[j];
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < 8; ++u) fwd_v(ch * 8 + u, g[u], hs[u], ho[u]);
          renorm(P1, P2);
        } else if (ch * 8 < len) {
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (ch * 8 + u < len) fwd(ch * 8 + u);
        }
      }
      const double co = __shfl_xor(cl, 1), yo = __shfl_xor(yprev, 1);  // the partner's last step
      double own;
      int start;
      if (b & 1) {
        const int m = len;
        const double ct = side == 0 ? cl : co, yt = side == 0 ? yprev : yo;  // rows < m
        const double cb = side == 0 ? co : cl, yb = side == 0 ? yo : yprev;  // rows > m
        const double am = (m > 0) ? -s * H[m] : 0.0;
        const double cmm = (m < b - 1) ? -s * H[m + 1] : 0.0;
        const double dm = fma(s, Hs[m], fma(8.0, (double)m, 4.0 * q + 6.0));
        own = ((double)row[m] - am * yt - cmm * yb) / (dm - am * ct - cmm * cb);
        if constexpr (SP) umax = fabsf((float)own);
        start = len - 1;
        if (side == 0) row[m] = (float)own;  // both lanes have read row[m] above
      } else {
        // u_t = y_t - c_t u_b (row len-1),  u_b = y_b - c_b u_t (row len)
        const double ct = side == 0 ? cl : co, yt = side == 0 ? yprev : yo;
        const double cb = side == 0 ? co : cl, yb = side == 0 ? yo : yprev;
        const double ut = (yt - ct * yb) / (1.0 - ct * cb);
        own = side == 0 ? ut : yb - cb * ut;
        if constexpr (SP) umax = fabsf((float)own);
        start = len - 2;
        row[j0 + dj * (len - 1)] = (float)own;
      }
      // back substitution outwards, i = start .. 0
      double uprev = own;
      auto bwd = [&](int i, float y) {
        const double v = fma(-(double)cm[i], uprev, (double)y);
        row[j0 + dj * i] = (float)v;
        if constexpr (SP) umax = fmaxf(umax, fabsf((float)v));
        uprev = v;
      };
#pragma unroll
      for (int ch = 7; ch >= 0; --ch) {
        if (ch * 8 + 7 <= start) {
          float y[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) y[u] = row[j0 + dj * (ch * 8 + u)];
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 7; u >= 0; --u) bwd(ch * 8 + u, y[u]);
        } else if (ch * 8 <= start) {
#pragma unroll
          for (int u = 7; u >= 0; --u)
            if (ch * 8 + u <= start) bwd(ch * 8 + u, row[j0 + dj * (ch * 8 + u)]);
        }
      }
      if (!act) umax = 0.f;
    }
    if constexpr (SP) publish_wave_max(umax, smx + 4);
    __syncthreads();

    // ---------------- phase 3: U_pq = V_p U~ -> global, next pass set up around the epilogue
    GatherRegs<TIO> G;
    const int bn = (In.n - qn + 1) >> 1;
    int e2 = 0;
    if constexpr (SP) e2 = block_split_exp(smx + 4);
    helmc_gemm<TIO, true, NH, SP>(
        I, q, b, TV, Wk, slab, U + I.foff, P, __builtin_ldexpf(1.f, e2), I.texp + e2,
        [&] {
          if (has_next) {
            if (In.AP > 64)
              slab_prefetch2<2, NS>(P, VTs + NS * In.voff, In.AP, tid);
            else
              slab_prefetch2<1, NS>(P, VTs + NS * In.voff, In.AP, tid);
            gather_issue(G, F + In.foff, In.n, In.p, qn, In.a, bn);
          }
        },
        [&] {
          if (has_next) {
            const float m = gather_store<TIO, SP == 1>(G, F + In.foff, In.n, In.p, qn, In.a, bn, Wk);
            if constexpr (SP) publish_wave_max(m, smx);
          }
        });
  }
}

template <typename TIO, int SP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void helm_compact_kernel(
    const HelmItem* __restrict__ items, int nitems, HelmQueues Q, const unsigned short* __restrict__ Vs,
    const unsigned short* __restrict__ VTs, const double* __restrict__ lam, const TIO* __restrict__ F,
    TIO* __restrict__ U) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  float* Wk = reinterpret_cast<float*>(smem_raw);
  unsigned short* slab = reinterpret_cast<unsigned short*>(smem_raw + HelmCompact::WK);
  double* H = reinterpret_cast<double*>(smem_raw + HelmCompact::WK + HelmCompact::SLAB);
  __shared__ int s_nx[2];
  __shared__ float s_mx[8];  // wave maxima of |F_pq| and |U~| (fp16 split)
  const int tid = threadIdx.x;
  constexpr int NS = SP ? 2 : 3;

  // the next item's index is claimed at the start of the current one (two
  // slots: slot k is rewritten only after every thread has passed the
  // barriers of the item in between)
  int qi = blockIdx.x % Q.nq;
  if (tid < 64) {
    const int first = Q.claim(qi);
    if (tid == 0) s_nx[0] = first;
  }
  __syncthreads();
  const int it0 = __builtin_amdgcn_readfirstlane(s_nx[0]);
  if (it0 >= nitems) return;
  HelmItem I = items[it0];
  helm_gather<4, HelmCompact::LDW>(F + I.foff, I.n, I.p, 0, I.a, (I.n + 1) >> 1, Wk);
  if constexpr (SP) publish_wave_max(own_block_absmax(Wk, I.a, (I.n + 1) >> 1), s_mx);
  SlabRegs P[2];
  if (I.AP > 64)
    slab_prefetch2<2, NS>(P, VTs + NS * I.voff, I.AP, tid);
  else
    slab_prefetch2<1, NS>(P, VTs + NS * I.voff, I.AP, tid);
  int slot = 1;
  for (;;) {
    // the next item is claimed now; its index is published after this item's
    // first pass (two slots: slot k is rewritten only after every thread has
    // passed the barriers of the passes in between)
    int nxt = 0;
    if (tid < 64) nxt = Q.claim(qi);
    if (I.AP > 64)
      helmc_pass<TIO, 2, SP>(I, 0, I, 1, true, Vs, VTs, lam, F, U, Wk, slab, H, P, s_mx);
    else
      helmc_pass<TIO, 1, SP>(I, 0, I, 1, true, Vs, VTs, lam, F, U, Wk, slab, H, P, s_mx);
    if (tid == 0) s_nx[slot] = nxt;
    __syncthreads();
    const int itn = __builtin_amdgcn_readfirstlane(s_nx[slot]);
    slot ^= 1;
    const bool has = itn < nitems;
    const HelmItem In = items[has ? itn : it0];
    if (I.AP > 64)
      helmc_pass<TIO, 2, SP>(I, 1, In, 0, has, Vs, VTs, lam, F, U, Wk, slab, H, P, s_mx);
    else
      helmc_pass<TIO, 1, SP>(I, 1, In, 0, has, Vs, VTs, lam, F, U, Wk, slab, H, P, s_mx);
    if (!has) break;
    I = In;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Pair variant (mixed precision): one 8-wave workgroup per CU runs both column
// parities of a work item (solve, row parity p) at once -- waves 0-3 take
// q = 0, waves 4-7 take q = 1 -- so the rows of F[p::2, :] are read whole and
// once, both halves share every table slab, and the rows of U[p::2, :] are
// assembled in LDS and written whole and once.
//   Wk0, Wk1 : 128 x 130 FP32 parity blocks F_p0, F_p1 -> G_pq -> U~_pq; after
//              the second GEMM they hold the interleaved output rows
//              (128 x 257 FP32 staging, row m = row p + 2m of U)
//   slab     : 3 bf16 slices x 128 table rows x 32 k (16-byte chunks swizzled)
//   H        : per parity h_{q+2j} and h_{q+2j} + h_{q+2j+2}   (FP64)
struct HelmPair {
  static constexpr int LDW = 130, NT = 512, LDU = 257;
  static constexpr int WK1 = HMAXA * LDW + 32;  // offset of Wk1 (floats): shifted by 32 banks from Wk0
  static constexpr size_t WK = (size_t)(WK1 + HMAXA * LDW) * 4;
  static constexpr size_t SLAB = 3 * HMAXA * 32 * 2;
  static constexpr size_t HT = 2 * (HMAXA + 2) * 8;
  static constexpr size_t BYTES = WK + SLAB + 2 * HT;
};
static_assert(HelmPair::BYTES + 64 <= 160 * 1024, "the pair workgroup's LDS must fit one CU");
static_assert((size_t)HMAXA * HelmPair::LDU * 4 <= HelmPair::WK, "the output staging must fit the work blocks");

__device__ __forceinline__ int pslab_off(int sl, int row, int ck) {  // in bf16 units
  return ((sl * HMAXA + row) * 4 + (ck ^ ((row >> 2) & 3))) * 8;
}

struct PairSlab {
  v8u16 v[3];
};

// Loads of k-slab k0 of a 3-slice (AP x AP) bf16 table: 3 x AP rows x 4 chunks of 8.
__device__ __forceinline__ void pslab_fetch(PairSlab& P, const unsigned short* __restrict__ T, int AP, int k0,
                                            int tid) {
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int e = tid + u * 512, sl = e >> 9, row = (e & 511) >> 2, ck = e & 3;
    if (row < AP) P.v[u] = *(const v8u16*)(T + (long long)sl * AP * AP + (long long)row * AP + k0 + ck * 8);
  }
}

__device__ __forceinline__ void pslab_store(const PairSlab& P, unsigned short* slab, int AP, int tid) {
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int e = tid + u * 512, sl = e >> 9, row = (e & 511) >> 2, ck = e & 3;
    if (row < AP) *(v8u16*)&slab[pslab_off(sl, row, ck)] = P.v[u];
  }
}

// Rows of F[p::2, :] in registers: wave w takes rows w + 8i for i in
// [i0, i0 + 8), lane l columns l + 64c (c < 4; n <= 256).  FP32 I/O only.
// Columns past the row's end are loaded from its last column and land in
// parity-block columns >= b, which only feed output columns that are never
// stored, so the loads need no per-lane branches.
struct PairGather {
  float v[8][4];
};

__device__ __forceinline__ void pair_gather_issue(PairGather& G, const float* __restrict__ Fb, int n, int p, int a,
                                                  int i0, int lane) {
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned col[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) col[c] = (unsigned)min(lane + 64 * c, n - 1);
  unsigned off = (unsigned)((p + 2 * (w + 8 * i0)) * n);  // < n^2 <= 65536
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (w + 8 * (i0 + i) < a) {
#pragma unroll
      for (int c = 0; c < 4; ++c) G.v[i][c] = Fb[off + col[c]];
    }
    off += 16u * (unsigned)n;
  }
}

// Column j = l + 64c of row m goes to parity block j & 1 = l & 1 at column j >> 1 (<= 127).
__device__ __forceinline__ void pair_gather_store(const PairGather& G, int a, int i0, float* Wk, int lane) {
  constexpr int LDW = HelmPair::LDW;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* dst = Wk + (lane & 1) * HelmPair::WK1 + (w + 8 * i0) * LDW + (lane >> 1);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (w + 8 * (i0 + i) < a) {
#pragma unroll
      for (int c = 0; c < 4; ++c) dst[32 * c] = G.v[i][c];
    }
    dst += 8 * LDW;
    asm volatile("" : "+v"(dst));  // one running address, not 8 hoisted ones
  }
}

// All rows of one item (FP32 I/O): rows w + 8i, i < 8, through G, then the rest.
__device__ __forceinline__ void pair_gather_all(PairGather& G, const float* __restrict__ Fb, int n, int p, int a,
                                                float* Wk, int lane) {
  pair_gather_issue(G, Fb, n, p, a, 0, lane);
  pair_gather_store(G, a, 0, Wk, lane);
  if (a > 64) {
    pair_gather_issue(G, Fb, n, p, a, 8, lane);
    pair_gather_store(G, a, 8, Wk, lane);
  }
}

// FP64 I/O: the same gather without the register prefetch (loads then stores, 4 rows at a time).
__device__ __forceinline__ void pair_gather_direct(const double* __restrict__ Fb, int n, int p, int a, float* Wk) {
  constexpr int LDW = HelmPair::LDW;
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int col[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) col[c] = min(lane + 64 * c, n - 1);
  float* dst0 = Wk + (lane & 1) * HelmPair::WK1 + (lane >> 1);
  for (int m0 = w; m0 < a; m0 += 32) {
    double v[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + 8 * i;
      if (m < a) {
#pragma unroll
        for (int c = 0; c < 4; ++c) v[i][c] = Fb[(p + 2 * m) * n + col[c]];
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + 8 * i;
      if (m < a) {
#pragma unroll
        for (int c = 0; c < 4; ++c) dst0[m * LDW + 32 * c] = (float)v[i][c];
      }
    }
  }
}

// acc = T * Wkq over k < AP for this wave's two column panels (wq, wq + 4) of
// parity block Wkq (a x b); T is the 3-slice (AP x AP) table whose slab k0 = 0
// is already in flight in P.  Rows a..AP-1 of Wkq hold zeros.  On return P
// holds nothing useful and the slab area is free.
__device__ __forceinline__ void pair_gemm(int a, int AP, int b, const unsigned short* __restrict__ T,
                                          const float* Wkq, unsigned short* slab, PairSlab& P, v4f (&acc)[2][8],
                                          int wq, int lane, int tid) {
  constexpr int LDW = HelmPair::LDW;
  const int npanel = (b + 15) >> 4, nrb = (a + 15) >> 4;
  const int npp = (wq + 4 < npanel) ? 2 : (wq < npanel ? 1 : 0);
#pragma unroll
  for (int pp = 0; pp < 2; ++pp)
#pragma unroll
    for (int rb = 0; rb < 8; ++rb) acc[pp][rb] = v4f{0, 0, 0, 0};
  const int off0 = pslab_off(0, lane & 15, lane >> 4);
  for (int k0 = 0; k0 < AP; k0 += 32) {
    v8u16 bh[2], bm[2], bl[2];
#pragma unroll
    for (int pp = 0; pp < 2; ++pp) {
      const float* src = Wkq + (k0 + 8 * (lane >> 4)) * LDW + (wq + 4 * pp) * 16 + (lane & 15);
      float xv[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) xv[e] = src[e * LDW];
      split3_pk8(xv, bh[pp], bm[pp], bl[pp]);
    }
    pslab_store(P, slab, AP, tid);
    __syncthreads();
    if (k0 + 32 < AP) pslab_fetch(P, T, AP, k0 + 32, tid);
    v8u16 A[2][3];
    auto aload = [&](v8u16(&dst)[3], int rb) {
      const int off = off0 + rb * 16 * 4 * 8;  // 16 rows further; the swizzle depends on row bits 2-3 only
      dst[0] = *(const v8u16*)&slab[off];
      dst[1] = *(const v8u16*)&slab[off + HMAXA * 32];
      dst[2] = *(const v8u16*)&slab[off + 2 * HMAXA * 32];
    };
    aload(A[0], 0);
#pragma unroll
    for (int rb = 0; rb < 8; ++rb) {
      if (rb + 1 < 8 && rb + 1 < nrb) aload(A[(rb + 1) & 1], rb + 1);
      if (rb < nrb) {
        const v8u16 ah = A[rb & 1][0], am = A[rb & 1][1], al = A[rb & 1][2];
#pragma unroll
        for (int pp = 0; pp < 2; ++pp) {
          if (pp < npp) {
            // products ordered small to large
            acc[pp][rb] = mfma_bf16(al, bh[pp], acc[pp][rb]);
            acc[pp][rb] = mfma_bf16(am, bm[pp], acc[pp][rb]);
            acc[pp][rb] = mfma_bf16(ah, bl[pp], acc[pp][rb]);
            acc[pp][rb] = mfma_bf16(am, bh[pp], acc[pp][rb]);
            acc[pp][rb] = mfma_bf16(ah, bm[pp], acc[pp][rb]);
            acc[pp][rb] = mfma_bf16(ah, bh[pp], acc[pp][rb]);
          }
        }
      }
    }
    __syncthreads();
  }
}

// FP64 row solves of one parity block, two lanes per row (twisted
// factorisation, as in helmc_pass): tq in [0, 256) is the thread's index
// within its parity half.  Rows >= a hold the GEMMs' zero padding, so their
// idle lanes work on `dummy` instead.
__device__ __forceinline__ void pair_row_solves(const HelmItem& I, int q, int b, int tq, float* Wkq,
                                                const double* __restrict__ lam, const double* H, float* dummy) {
  constexpr int LDW = HelmPair::LDW;
  const int a = I.a;
  const double* Hs = H + HMAXA + 2;
  int t2 = tq;
  asm volatile("" : "+v"(t2));
  const int r = t2 >> 1, side = t2 & 1;
  const bool act = r < a;
  const double s = act ? I.alpha - lam[I.loff + r] : 1.0;
  float* row = act ? Wkq + r * LDW : dummy + side * LDW;
  const int len = b >> 1;
  const int j0 = side ? b - 1 : 0, dj = side ? -1 : 1;
  const double* Ho = side ? H : H + 1;  // o towards the middle = -s Ho[j]
  const double dg0 = fma(8.0, (double)j0, 4.0 * q + 6.0), ddg = 8.0 * dj;
  float cm[64];
  double P1 = 1.0, P2 = 0.0, yprev = 0.0, oprev = 0.0, cl = 0.0;
  auto fwd_v = [&](int i, double g, double hs, double ho) {
    const int j = j0 + dj * i;
    const double d = fma(s, hs, fma((double)i, ddg, dg0));
    const double on = -s * ho;
    const double Pj = fma(d, P1, -(oprev * oprev) * P2);
    const double rj = P1 * rcp64_1(Pj);
    const double yj = fma(-oprev, yprev, g) * rj;
    cl = on * rj;
    cm[i] = (float)cl;
    row[j] = (float)yj;
    P2 = P1;
    P1 = Pj;
    yprev = yj;
    oprev = on;
  };
  auto fwd = [&](int i) {
    const int j = j0 + dj * i;
    fwd_v(i, row[j], Hs[j], Ho[j]);
  };
#pragma unroll
  for (int ch = 0; ch < 8; ++ch) {
    if (ch * 8 + 8 <= len) {
      float g[8];
      double hs[8], ho[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = j0 + dj * (ch * 8 + u);
        g[u] = row[j];
        hs[u] = Hs[j];
        ho[u] = Ho[j];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < 8; ++u) fwd_v(ch * 8 + u, g[u], hs[u], ho[u]);
      renorm(P1, P2);
    } else if (ch * 8 < len) {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (ch * 8 + u < len) fwd(ch * 8 + u);
    }
  }
  const double co = __shfl_xor(cl, 1), yo = __shfl_xor(yprev, 1);  // the partner's last step
  double own;
  int start;
  if (b & 1) {
    const int m = len;
    const double ct = side == 0 ? cl : co, yt = side == 0 ? yprev : yo;  // rows < m
    const double cb = side == 0 ? co : cl, yb = side == 0 ? yo : yprev;  // rows > m
    const double am = (m > 0) ? -s * H[m] : 0.0;
    const double cmm = (m < b - 1) ? -s * H[m + 1] : 0.0;
    const double dm = fma(s, Hs[m], fma(8.0, (double)m, 4.0 * q + 6.0));
    own = ((double)row[m] - am * yt - cmm * yb) / (dm - am * ct - cmm * cb);
    start = len - 1;
    if (side == 0) row[m] = (float)own;  // both lanes have read row[m] above
  } else {
    const double ct = side == 0 ? cl : co, yt = side == 0 ? yprev : yo;
    const double cb = side == 0 ? co : cl, yb = side == 0 ? yo : yprev;
    const double ut = (yt - ct * yb) / (1.0 - ct * cb);
    own = side == 0 ? ut : yb - cb * ut;
    start = len - 2;
    row[j0 + dj * (len - 1)] = (float)own;
  }
  double uprev = own;
  auto bwd = [&](int i, float y) {
    const double v = fma(-(double)cm[i], uprev, (double)y);
    row[j0 + dj * i] = (float)v;
    uprev = v;
  };
#pragma unroll
  for (int ch = 7; ch >= 0; --ch) {
    if (ch * 8 + 7 <= start) {
      float y[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) y[u] = row[j0 + dj * (ch * 8 + u)];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 7; u >= 0; --u) bwd(ch * 8 + u, y[u]);
    } else if (ch * 8 <= start) {
#pragma unroll
      for (int u = 7; u >= 0; --u)
        if (ch * 8 + u <= start) bwd(ch * 8 + u, row[j0 + dj * (ch * 8 + u)]);
    }
  }
}

template <typename TIO>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void helm_pair_kernel(
    const HelmItem* __restrict__ items, int nitems, HelmQueues Q, const unsigned short* __restrict__ Vs,
    const unsigned short* __restrict__ VTs, const double* __restrict__ lam, const TIO* __restrict__ F,
    TIO* __restrict__ U) {
  constexpr int LDW = HelmPair::LDW, LDU = HelmPair::LDU, NT = HelmPair::NT;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  float* Wk = reinterpret_cast<float*>(smem_raw);
  unsigned short* slab = reinterpret_cast<unsigned short*>(smem_raw + HelmPair::WK);
  double* Htab = reinterpret_cast<double*>(smem_raw + HelmPair::WK + HelmPair::SLAB);  // [q][2][HMAXA + 2]
  __shared__ int s_nx[2];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int q = w >> 2, wq = w & 3;  // this wave's column parity and its index within the half
  float* Wkq = Wk + q * HelmPair::WK1;
  double* Hq = Htab + q * 2 * (HMAXA + 2);

  int qi = blockIdx.x % Q.nq;
  if (tid < 64) {
    const int first = Q.claim(qi);
    if (tid == 0) s_nx[0] = first;
  }
  __syncthreads();
  const int it0 = __builtin_amdgcn_readfirstlane(s_nx[0]);
  if (it0 >= nitems) return;
  HelmItem I = items[it0];
  if constexpr (sizeof(TIO) == 4) {
    PairGather G;
    pair_gather_all(G, reinterpret_cast<const float*>(F) + I.foff, I.n, I.p, I.a, Wk, lane);
  } else {
    pair_gather_direct(reinterpret_cast<const double*>(F) + I.foff, I.n, I.p, I.a, Wk);
  }
  PairSlab P;
  pslab_fetch(P, VTs + 3 * I.voff, I.AP, 0, tid);
  int slot = 1;
  for (;;) {
    int nxt = 0;
    if (tid < 64) nxt = Q.claim(qi);
    // opaque per-item copies of the thread indices: every lane-dependent
    // address is recomputed per item instead of being hoisted out of the
    // item loop into registers
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int a = I.a, AP = I.AP, n = I.n;
    const int b = (n - q + 1) >> 1;
    // ---------------- phase 0: zero pad rows (both blocks), h tables of this wave's parity
    for (int e = tid; e < (AP - a) * 128; e += NT) {
      const int r = a + (e >> 7), c = e & 127;
      Wk[r * LDW + c] = 0.f;
      Wk[HelmPair::WK1 + r * LDW + c] = 0.f;
    }
    for (int j = tid & 255; j <= b; j += 256) {
      const double hj = 2.0 / (2 * (q + 2 * j) + 1);
      Hq[j] = hj;
      Hq[HMAXA + 2 + j] = hj + 2.0 / (2 * (q + 2 * j) + 5);
    }
    __syncthreads();

    // ---------------- phase 1: G_pq = V_p^T F_pq, in place
    {
      v4f acc[2][8];
      pair_gemm(a, AP, b, VTs + 3 * I.voff, Wkq, slab, P, acc, wq, lane, tid);
      pslab_fetch(P, Vs + 3 * I.voff, AP, 0, tid);  // phase 3's first slab flies during the row solves
      int ln = lane;
      asm volatile("" : "+v"(ln));  // keeps the lane-dependent addresses inside the item loop
      const int npanel = (b + 15) >> 4, npp = (wq + 4 < npanel) ? 2 : (wq < npanel ? 1 : 0);
      const int nrb = (a + 15) >> 4, lr = 4 * (ln >> 4), lim = a - lr;
#pragma unroll
      for (int pp = 0; pp < 2; ++pp) {
        if (pp >= npp) continue;
        const int col = (wq + 4 * pp) * 16 + (ln & 15);
        float* dst = Wkq + lr * LDW + col;
#pragma unroll
        for (int rb = 0; rb < 8; ++rb) {
          if (rb >= nrb) continue;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (rb * 16 + r < lim) dst[r * LDW] = acc[pp][rb][r];
          dst += 16 * LDW;
          asm volatile("" : "+v"(dst));
        }
      }
    }
    __syncthreads();

    // ---------------- phase 2: FP64 row solves of both parities
    if (wq * 32 < a)
      pair_row_solves(I, q, b, tid & 255, Wkq, lam, Hq, reinterpret_cast<float*>(slab) + q * 2 * LDW);
    __syncthreads();

    // ---------------- phase 3: U_pq = V_p U~_pq; next item's loads in flight around the epilogue
    if (tid == 0) s_nx[slot] = nxt;
    v4f acc[2][8];
    pair_gemm(a, AP, b, Vs + 3 * I.voff, Wkq, slab, P, acc, wq, lane, tid);
    const int itn = __builtin_amdgcn_readfirstlane(s_nx[slot]);
    slot ^= 1;
    const bool has = itn < nitems;
    const HelmItem In = items[has ? itn : it0];
    if (has) pslab_fetch(P, VTs + 3 * In.voff, In.AP, 0, tid);
    // interleaved staging: row m, column 2c + q
    {
      int ln = lane;
      asm volatile("" : "+v"(ln));
      const int npanel = (b + 15) >> 4, npp = (wq + 4 < npanel) ? 2 : (wq < npanel ? 1 : 0);
      const int nrb = (a + 15) >> 4, lr = 4 * (ln >> 4), lim = a - lr;
#pragma unroll
      for (int pp = 0; pp < 2; ++pp) {
        if (pp >= npp) continue;
        const int col = (wq + 4 * pp) * 16 + (ln & 15);
        const bool cok = col < b;
        float* dst = Wk + lr * LDU + 2 * col + q;
#pragma unroll
        for (int rb = 0; rb < 8; ++rb) {
          if (rb >= nrb) continue;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (cok && rb * 16 + r < lim) dst[r * LDU] = acc[pp][rb][r];
          dst += 16 * LDU;
          asm volatile("" : "+v"(dst));
        }
      }
    }
    // the next item's first rows fly while this item's rows are stored
    PairGather G;
    if constexpr (sizeof(TIO) == 4)
      if (has) pair_gather_issue(G, reinterpret_cast<const float*>(F) + In.foff, In.n, In.p, In.a, 0, lane);
    __syncthreads();
    // whole rows of U[p::2, :]: wave w rows w, w + 8, ..., lane l columns l + 64c
    {
      TIO* Ub = U + I.foff;
      int ln = lane;
      asm volatile("" : "+v"(ln));
      for (int m = w; m < a; m += 8) {
        TIO* dst = Ub + (unsigned)((I.p + 2 * m) * n);
        const float* src = Wk + m * LDU;
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (ln + 64 * c < n) __builtin_nontemporal_store((TIO)src[ln + 64 * c], dst + (unsigned)(ln + 64 * c));
      }
    }
    __syncthreads();
    if (!has) break;
    if constexpr (sizeof(TIO) == 4) {
      pair_gather_store(G, In.a, 0, Wk, lane);
      if (In.a > 64) {
        pair_gather_issue(G, reinterpret_cast<const float*>(F) + In.foff, In.n, In.p, In.a, 8, lane);
        pair_gather_store(G, In.a, 8, Wk, lane);
      }
    } else
      pair_gather_direct(reinterpret_cast<const double*>(F) + In.foff, In.n, In.p, In.a, Wk);
    I = In;
    // (phase 0 of the next item writes only rows >= a and the h tables; its barrier publishes everything)
  }
}

template <typename TIO>
static int launch_helm_pair(const HelmItem* items, int nitems, HelmQueues Q, const unsigned short* Vs,
                            const unsigned short* VTs, const double* lam, const TIO* F, TIO* U, int nblocks,
                            hipStream_t st) {
  const size_t shm = HelmPair::BYTES;
  auto kern = helm_pair_kernel<TIO>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) return set_error("helm: hipFuncSetAttribute: %s", hipGetErrorString(e)), JFX_EHIP;
  hipLaunchKernelGGL(helm_reset, dim3(1), dim3(64), 0, st, Q.counters, Q.nq);
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(HelmPair::NT), shm, st, items, nitems, Q, Vs, VTs, lam, F, U);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

template <typename TIO, int SP = 0>
static int launch_helm_compact(const HelmItem* items, int nitems, HelmQueues Q, const unsigned short* Vs,
                               const unsigned short* VTs, const double* lam, const TIO* F, TIO* U, int nblocks,
                               hipStream_t st) {
  const size_t shm = HelmCompact::BYTES;
  auto kern = helm_compact_kernel<TIO, SP>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) return set_error("helm: hipFuncSetAttribute: %s", hipGetErrorString(e)), JFX_EHIP;
  hipLaunchKernelGGL(helm_reset, dim3(1), dim3(64), 0, st, Q.counters, Q.nq);
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(HelmCompact::NT), shm, st, items, nitems, Q, Vs, VTs, lam,
                     F, U);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

template <int MODE, int NW, typename TIO>
static int launch_helm(const HelmItem* items, int nitems, HelmQueues Q, const double* Vd,
                       const double* VTd, const unsigned short* Vs, const unsigned short* VTs,
                       const double* lam, const TIO* F, TIO* U, double* mult_scratch, int nblocks,
                       hipStream_t st) {
  size_t shm = HelmLDS<MODE>::WK + HelmLDS<MODE>::U1 + HTAB;
  auto kern = helm_ragged_kernel<MODE, NW, TIO>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) return set_error("helm: hipFuncSetAttribute: %s", hipGetErrorString(e)), JFX_EHIP;
  hipLaunchKernelGGL(helm_reset, dim3(1), dim3(64), 0, st, Q.counters, Q.nq);
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(NW * 64), shm, st, items, nitems, Q, Vd, VTd, Vs, VTs,
                     lam, F, U, mult_scratch);
  JFX_LAUNCH_CHECK();
  return JFX_OK;
}

}  // namespace jfx

using namespace jfx;

// ---------------------------------------------------------------------------
// Allocation-free variant of the mixed-precision ragged batch.  Its only
// scratch memory is the queue counters: HMAXQ ints in this file-scope device
// array (one copy per GPU, held until the process ends).  Host code never
// passes the array's name to a kernel: the device address of each GPU's copy
// is looked up once with hipGetSymbolAddress and kept per device.
static __device__ __align__(256) int scratch[HMAXQ];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current GPU's copy of `scratch` (looked up on the
// first call on that device), or nullptr with the error set.
int* noalloc_counters() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error("helm_ragged_noalloc: hipGetDevice: %s", hipGetErrorString(e)), nullptr;
  if (dev < 0 || dev >= kMaxDevices) return set_error("helm_ragged_noalloc: device %d out of range", dev), nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error("helm_ragged_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e)), nullptr;
    g_scratch_ptr[dev] = static_cast<int*>(ptr);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int jfx_helm_noalloc_max_queues(void) { return HMAXQ; }

// Same work as jfx_helm_ragged(mode 1, nwaves): the queue counters live in
// the static `scratch` array (cleared by the launch's own reset kernel), so
// nothing is allocated.  Calls that overlap on one GPU share the counters and
// race; calls on one stream are ordered.
int jfx_helm_ragged_noalloc(int io_f32, int nwaves, const void* items, int nitems, const int* qbounds, int nqueues,
                            const unsigned short* Vs, const unsigned short* VTs, const double* lam, const void* F,
                            void* U, int nblocks, hipStream_t stream) {
  if (!items || !qbounds || !lam || !F || !U || !Vs || !VTs || nitems < 0 || nblocks <= 0)
    return set_error("helm_ragged_noalloc: bad arguments"), JFX_EARG;
  if (nwaves != 4 && nwaves != 8) return set_error("helm_ragged_noalloc: nwaves must be 4 or 8"), JFX_EARG;
  if (nqueues < 1 || nqueues > HMAXQ)
    return set_error("helm_ragged_noalloc: nqueues %d exceeds the %d counters of the static scratch", nqueues, HMAXQ),
           JFX_ESHAPE;
  if (nitems == 0) return JFX_OK;
  int* counters = noalloc_counters();
  if (!counters) return JFX_EHIP;
  const HelmItem* it = (const HelmItem*)items;
  const HelmQueues queues{counters, qbounds, nqueues};
  if (nwaves == 8) {
    if (io_f32)
      return launch_helm_pair<float>(it, nitems, queues, Vs, VTs, lam, (const float*)F, (float*)U, nblocks, stream);
    return launch_helm_pair<double>(it, nitems, queues, Vs, VTs, lam, (const double*)F, (double*)U, nblocks,
                                    stream);
  }
  if (io_f32)
    return launch_helm_compact<float>(it, nitems, queues, Vs, VTs, lam, (const float*)F, (float*)U, nblocks, stream);
  return launch_helm_compact<double>(it, nitems, queues, Vs, VTs, lam, (const double*)F, (double*)U, nblocks,
                                     stream);
}

int jfx_helm_item_size(void) { return (int)sizeof(HelmItem); }

int jfx_helm_scratch_doubles_per_block(void) { return HMAXA * HMAXA; }

// mode 0: FP64 (F, U double; Vd/VTd used)
// mode 1: BF16 3-term split, FP32 storage (F, U float if io_f32 else double; Vs/VTs used)
// mode 2: FP16 2-term split (tables V 2^eA in two fp16 slices, eA in the
//         item's pad field; blocks scaled by a power of two), nwaves 4
// nwaves: 4 or 8 waves per workgroup; nblocks: persistent grid size
//   (mult_scratch must hold nblocks * jfx_helm_scratch_doubles_per_block() doubles).
int jfx_helm_ragged(int mode, int io_f32, int nwaves, const void* items, int nitems, int* counters,
                    const int* qbounds, int nqueues, const double* Vd, const double* VTd, const unsigned short* Vs,
                    const unsigned short* VTs, const double* lam, const void* F, void* U,
                    double* mult_scratch, int nblocks, hipStream_t stream) {
  if (!items || !counters || !qbounds || !lam || !F || !U || !mult_scratch || nitems < 0 || nblocks <= 0)
    return set_error("helm_ragged: bad arguments"), JFX_EARG;
  if (nqueues < 1 || nqueues > HMAXQ) return set_error("helm_ragged: nqueues must be in [1, 16]"), JFX_EARG;
  if (nitems == 0) return JFX_OK;
  const HelmItem* it = (const HelmItem*)items;
  const HelmQueues queues{counters, qbounds, nqueues};
  if (mode == 0) {
    if (!Vd || !VTd || io_f32) return set_error("helm_ragged: mode 0 needs fp64 tables and fp64 I/O"), JFX_EARG;
    if (nwaves == 4)
      return launch_helm<0, 4, double>(it, nitems, queues, Vd, VTd, Vs, VTs, lam, (const double*)F,
                                       (double*)U, mult_scratch, nblocks, stream);
    if (nwaves == 8)
      return launch_helm<0, 8, double>(it, nitems, queues, Vd, VTd, Vs, VTs, lam, (const double*)F,
                                       (double*)U, mult_scratch, nblocks, stream);
  } else if (mode == 1) {
    if (!Vs || !VTs) return set_error("helm_ragged: mode 1 needs bf16 slice tables"), JFX_EARG;
    if (io_f32) {
      if (nwaves == 4)
        return launch_helm_compact<float>(it, nitems, queues, Vs, VTs, lam, (const float*)F, (float*)U, nblocks,
                                          stream);
      if (nwaves == 8)
        return launch_helm_pair<float>(it, nitems, queues, Vs, VTs, lam, (const float*)F, (float*)U, nblocks,
                                       stream);
    } else {
      if (nwaves == 4)
        return launch_helm_compact<double>(it, nitems, queues, Vs, VTs, lam, (const double*)F, (double*)U,
                                           nblocks, stream);
      if (nwaves == 8)
        return launch_helm_pair<double>(it, nitems, queues, Vs, VTs, lam, (const double*)F, (double*)U, nblocks,
                                        stream);
    }
  } else if (mode == 2) {
    if (!Vs || !VTs) return set_error("helm_ragged: mode 2 needs fp16 slice tables"), JFX_EARG;
    if (nwaves != 4) return set_error("helm_ragged: mode 2 runs with nwaves 4 only"), JFX_EARG;
    if (io_f32)
      return launch_helm_compact<float, 1>(it, nitems, queues, Vs, VTs, lam, (const float*)F, (float*)U, nblocks,
                                           stream);
    return launch_helm_compact<double, 1>(it, nitems, queues, Vs, VTs, lam, (const double*)F, (double*)U, nblocks,
                                          stream);
  }
  return set_error("helm_ragged: bad mode/nwaves"), JFX_EARG;
}

}  // extern "C"
