// Register-resident-operator contraction for M <= 128, K <= 128 (bf16x6:
// K <= 192):
//     Y[p, m, q] = sum_k A[m, k] X[p, k, q],   X viewed as (P, K, Q)
// for the two column layouts that cover every axis of a C-order array:
// Q % 64 == 0 (a 64-column tile is 64 consecutive q of one p) and Q == 1
// (a tile is 64 consecutive rows p, each K contiguous values).
//
// One persistent workgroup of 8 waves per CU (bf16x6: 16 waves in two column
// groups, each computing half of the tile's 64 columns, so that more waves
// per SIMD hide the staging).  Wave w keeps rows 16w..16w+15 of the operator
// in registers for the whole launch (FP64/FP32: one element per MFMA k-step;
// bf16x6: three bf16 slices per 32-k step), so the only LDS operand traffic
// is the X tile.  Each 64-column tile of X (all
// K rows) is staged through LDS once -- already split into three bf16 terms
// for bf16x6 -- while the next tile is loaded into registers, so one barrier
// pair covers a whole tile (up to 128 MFMAs per wave).  With Q == 1 the output
// tile is transposed through LDS and written as one contiguous block.
//   POL 0: FP64 v_mfma_f64_16x16x4f64       POL 1: FP32 v_mfma_f32_16x16x4f32
//   POL 3: bf16 3-term float split, 6 products, FP32 accumulation (FP32 I/O)
#include <type_traits>

#include "jfx_common.h"

namespace jfx {

constexpr int RTN = 64, RNT = 512, RMAX = 128;

template <int POL>
struct RT;
template <>
struct RT<0> {
  typedef double TX;
  typedef v4d V;
};
template <>
struct RT<1> {
  typedef float TX;
  typedef v4f V;
};
template <>
struct RT<3> {
  typedef float TX;
  typedef v4f V;
};

// bf16x6 stages 8-wide k groups: 16 per column (K <= 128), 24 with six
// k-steps (K <= 192, one column group of 8 waves).
constexpr int RMAX3 = 192;
template <int POL, int NS>
constexpr int ra_kgroups() { return (POL == 3 && NS > 4) ? 4 * NS : 16; }

// Two X-tile buffers: plain Bs[k][c] (row stride 72 doubles / 80 floats),
// bf16x6 Bsl[3][c][k] (row stride 8 KG + 8 bf16: 136 for K <= 128, 200 for 192).
template <int POL, int NS>
struct RLds {
  typedef typename RT<POL>::TX TX;
  static constexpr int BLD = (POL == 3) ? 8 * ra_kgroups<POL, NS>() + 8 : (POL == 0 ? 72 : 80);
  static constexpr size_t B = (POL == 3) ? (size_t)3 * RTN * BLD * 2 : (size_t)RMAX * BLD * sizeof(TX);
  static constexpr size_t TOTAL = 2 * B;
};

struct RMap {
  long long P, Q, ncols, ntiles, nqt;  // nqt: 64-column tiles per p (Q % 64 == 0)
  int M, K, qone;
  int Mtot, moff;  // the output has Mtot rows along the axis; this launch writes rows moff..moff+M-1
  // A row stride lda (bf16x6: of the slices), X stride xps between
  // consecutive p (the row stride of X when Q == 1), and accum = 1 adds this
  // launch's sums to Y (a split of K over launches)
  int lda, accum;
  long long xps;
  int x16;  // bf16x6 with Q == 1: rows start 16-byte aligned and hold whole 8-value groups
  int o16;  // FP32 / bf16x6 with Q == 1: output columns in 16-byte chunks (else 8-byte pairs)
};

// RB: row blocks of 16 per wave (2: operators of up to 256 rows, bf16x6
// with one column group).
// NS: k-steps run unconditionally (4 deep for FP64/FP32, 32 deep for bf16x6);
// the operator registers and the staged rows past K are zero.  CG: column
// groups of 8 waves (1 or 2).
template <int POL, int NS, int CG, int RB = 1>
__global__ __launch_bounds__(RNT * CG) void contract_ra_kernel(const void* __restrict__ Aptr,
                                                          const typename RT<POL>::TX* __restrict__ X,
                                                          typename RT<POL>::TX* __restrict__ Y, RMap rm) {
  typedef typename RT<POL>::TX TX;
  typedef typename RT<POL>::V V;
  constexpr int BLD = RLds<POL, NS>::BLD;
  constexpr int KG = ra_kgroups<POL, NS>();        // bf16x6: 8-wide k groups staged per column
  constexpr int EPC = 16 / sizeof(TX);            // elements per 16-byte chunk
  constexpr int NT = RNT * CG, JB = 4 / CG;        // threads; 16-column blocks per wave
  constexpr int NCH = RMAX * RTN / EPC / NT;      // plain: chunks per thread per tile
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  TX* Bs0 = reinterpret_cast<TX*>(smem);  // buffer b at Bs0 + b * BSZ
  unsigned short* Bsl0 = reinterpret_cast<unsigned short*>(smem);
  constexpr int BSZ = (int)(RLds<POL, NS>::B / sizeof(TX)), BSL = (int)(RLds<POL, NS>::B / 2);

  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int w = wv & 7, cg = wv >> 3;  // row block, column group
  const int M = rm.M, K = rm.K;
  const long long Q = rm.Q;
  static_assert(RB == 1 || (POL == 3 && CG == 1), "two row blocks per wave: bf16x6, one column group");
  const bool wact = 16 * w < M;  // this wave owns output rows (row block 0)

  // ---- operator rows of this wave, resident in registers: row block rb
  // holds rows 16 w + 128 rb .. + 15
  typedef typename std::conditional<POL == 3, v8u16, TX>::type AR;
  AR ar[RB][POL == 3 ? 3 * NS : NS];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    const int row = 16 * w + RMAX * rb + (lane & 15);
    if constexpr (POL == 3) {
      // slices of Mtot rows of lda bf16 values each (a row split passes Aptr
      // at its first row, a column split at its first column)
      const unsigned short* Asl = reinterpret_cast<const unsigned short*>(Aptr);
      const int Kp = (K + 31) & ~31;
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int sl = 0; sl < 3; ++sl) {
          v8u16 v = {0, 0, 0, 0, 0, 0, 0, 0};
          if (row < M && 32 * s < Kp) v = *(const v8u16*)(Asl + ((long long)sl * rm.Mtot + row) * rm.lda + 32 * s + 8 * (lane >> 4));
          ar[rb][3 * s + sl] = v;
        }
    } else {
      const TX* A = reinterpret_cast<const TX*>(Aptr);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int k = 4 * s + (lane >> 4);
        ar[rb][s] = (row < M && k < K) ? A[(long long)row * rm.lda + k] : TX(0);
      }
    }
  }

  // ---- X tile loads (registers) and LDS staging
  // plain: NCH 16-byte chunks per thread; bf16x6: 2 groups of 8 consecutive k of one column
  typedef typename std::conditional<POL == 3, float, TX>::type RV;
  constexpr int GPT = KG * RTN / NT;  // bf16x6: k groups per thread per tile
  constexpr int NREG = (POL == 3) ? 8 * GPT : NCH * EPC;
  RV xr[NREG];
  auto tile_cols = [&](long long t, long long& xbase, long long& c0, int& ncol) {
    if (rm.qone) {
      c0 = t * RTN;
      xbase = c0 * rm.xps;
      const long long left = rm.ncols - c0;
      ncol = left < RTN ? (int)left : RTN;
    } else {
      const long long p = t / rm.nqt, q0 = (t - p * rm.nqt) * RTN;
      c0 = q0;
      xbase = p * rm.xps + q0;
      ncol = RTN;
    }
  };
  auto load_tile = [&](long long t) {
    long long xbase, c0;
    int ncol;
    tile_cols(t, xbase, c0, ncol);
    if constexpr (POL == 3) {
#pragma unroll
      for (int g = 0; g < GPT; ++g) {
        const int gi = tid + NT * g;
        if (rm.qone) {
          const int c = gi / KG, k8 = (gi - c * KG) * 8;
          const bool ok = c < ncol && k8 < K;
          const float* src = X + xbase + (long long)c * rm.xps + k8;
          if (rm.x16) {
            float4 v0 = ok ? *(const float4*)src : float4{0, 0, 0, 0};
            float4 v1 = ok ? *(const float4*)(src + 4) : float4{0, 0, 0, 0};
            xr[8 * g + 0] = v0.x; xr[8 * g + 1] = v0.y; xr[8 * g + 2] = v0.z; xr[8 * g + 3] = v0.w;
            xr[8 * g + 4] = v1.x; xr[8 * g + 5] = v1.y; xr[8 * g + 6] = v1.z; xr[8 * g + 7] = v1.w;
          } else {
            // even K, rows 8-byte aligned: pairs, each inside the row or zero
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
              float2 v = (ok && k8 + e < K) ? *(const float2*)(src + e) : float2{0, 0};
              xr[8 * g + e] = v.x;
              xr[8 * g + e + 1] = v.y;
            }
          }
        } else {
          const int c = gi & 63, k8 = (gi >> 6) * 8;
#pragma unroll
          for (int e = 0; e < 8; ++e) xr[8 * g + e] = (k8 + e < K) ? X[xbase + (long long)(k8 + e) * Q + c] : 0.f;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < NCH; ++i) {
        const int j = tid + NT * i;
        TX v[EPC];
        if (rm.qone) {
          const int e0 = j * EPC, c = e0 / K;
          const bool ok = c < ncol;
          if (ok) {
            const TX* src = X + xbase + (long long)c * rm.xps + (e0 - c * K);
            if constexpr (EPC == 2) {
              const double2 d = *(const double2*)src;
              v[0] = d.x;
              v[1] = d.y;
            } else {
              const float4 d = *(const float4*)src;
              v[0] = d.x; v[1] = d.y; v[2] = d.z; v[3] = d.w;
            }
          } else {
#pragma unroll
            for (int u = 0; u < EPC; ++u) v[u] = TX(0);
          }
        } else {
          constexpr int CPR = RTN / EPC;
          const int k = j / CPR, col = (j - k * CPR) * EPC;
          if (k < K) {
            const TX* src = X + xbase + (long long)k * Q + col;
            if constexpr (EPC == 2) {
              const double2 d = *(const double2*)src;
              v[0] = d.x;
              v[1] = d.y;
            } else {
              const float4 d = *(const float4*)src;
              v[0] = d.x; v[1] = d.y; v[2] = d.z; v[3] = d.w;
            }
          } else {
#pragma unroll
            for (int u = 0; u < EPC; ++u) v[u] = TX(0);
          }
        }
#pragma unroll
        for (int u = 0; u < EPC; ++u) xr[EPC * i + u] = v[u];
      }
    }
  };
  auto store_tile = [&](int buf) {
    TX* Bs = Bs0 + buf * BSZ;
    unsigned short* Bsl = Bsl0 + buf * BSL;
    if constexpr (POL == 3) {
#pragma unroll
      for (int g = 0; g < GPT; ++g) {
        const int gi = tid + NT * g;
        const int c = rm.qone ? (gi / KG) : (gi & 63), k8 = rm.qone ? (gi - (gi / KG) * KG) * 8 : (gi >> 6) * 8;
        float xv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) xv[e] = xr[8 * g + e];
        v8u16 h, m_, l;
        split3_pk8(xv, h, m_, l);
        *(v8u16*)&Bsl[(0 * RTN + c) * BLD + k8] = h;
        *(v8u16*)&Bsl[(1 * RTN + c) * BLD + k8] = m_;
        *(v8u16*)&Bsl[(2 * RTN + c) * BLD + k8] = l;
      }
    } else {
#pragma unroll
      for (int i = 0; i < NCH; ++i) {
        const int j = tid + NT * i;
        if (rm.qone) {
          const int e0 = j * EPC, c = e0 / K, k = e0 - c * K;
          if (c < RTN) {
#pragma unroll
            for (int u = 0; u < EPC; ++u) Bs[(k + u) * BLD + c] = xr[EPC * i + u];
          }
        } else {
          constexpr int CPR = RTN / EPC;
          const int k = j / CPR, col = (j - k * CPR) * EPC;
#pragma unroll
          for (int u = 0; u < EPC; ++u) Bs[k * BLD + col + u] = xr[EPC * i + u];
        }
      }
    }
  };

  if constexpr (POL != 3) {
    // with Q == 1 only rows < K are staged: the rest of the k-steps reads zeros
    for (int e = tid; e < 2 * BSZ; e += NT) Bs0[e] = TX(0);
  }
  typedef V AccT[RB][JB];
  auto compute = [&](AccT& acc, int buf) {
    const TX* Bs = Bs0 + buf * BSZ;
    const unsigned short* Bsl = Bsl0 + buf * BSL;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int j = 0; j < JB; ++j) acc[rb][j] = V{0, 0, 0, 0};
      if (wact) {
        // B fragments of step s + 1 are read from LDS before step s's MFMAs issue
        if constexpr (POL == 3 && RB > 1) {
          // two row blocks: each B fragment feeds 12 MFMAs, so fragments are
          // read one (k-step, column block) ahead instead of a whole k-step
          v8u16 bq[2][3];
          auto bl1 = [&](int idx, v8u16(&dst)[3]) {
            const int s = idx / JB, j = idx - s * JB;
            const int off = (16 * j + (lane & 15)) * BLD + 32 * s + 8 * (lane >> 4);
            dst[0] = *(const v8u16*)&Bsl[off];
            dst[1] = *(const v8u16*)&Bsl[RTN * BLD + off];
            dst[2] = *(const v8u16*)&Bsl[2 * RTN * BLD + off];
          };
          bl1(0, bq[0]);
  #pragma unroll
          for (int idx = 0; idx < NS * JB; ++idx) {
            const int s = idx / JB, j = idx - s * JB;
            if (idx + 1 < NS * JB) bl1(idx + 1, bq[(idx + 1) & 1]);
            __builtin_amdgcn_sched_barrier(0);
            const v8u16 bh = bq[idx & 1][0], bm = bq[idx & 1][1], bl = bq[idx & 1][2];
  #pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
              if (rb > 0 && 16 * w + RMAX * rb >= M) break;  // wave-uniform
              const v8u16 ah = ar[rb][3 * s], am = ar[rb][3 * s + 1], al = ar[rb][3 * s + 2];
              acc[rb][j] = mfma_bf16(al, bh, acc[rb][j]);
              acc[rb][j] = mfma_bf16(am, bm, acc[rb][j]);
              acc[rb][j] = mfma_bf16(ah, bl, acc[rb][j]);
              acc[rb][j] = mfma_bf16(am, bh, acc[rb][j]);
              acc[rb][j] = mfma_bf16(ah, bm, acc[rb][j]);
              acc[rb][j] = mfma_bf16(ah, bh, acc[rb][j]);
            }
          }
        } else if constexpr (POL == 3) {
          // B fragments double-buffered when the registers allow (not with six k-steps)
          constexpr int NBB = (CG == 2 || NS > 4 || RB > 1) ? 1 : 2;
          v8u16 bf[NBB][JB][3];
          auto bload = [&](int s, v8u16(&dst)[JB][3]) {
  #pragma unroll
            for (int j = 0; j < JB; ++j) {
              const int off = (16 * (cg * JB + j) + (lane & 15)) * BLD + 32 * s + 8 * (lane >> 4);
              dst[j][0] = *(const v8u16*)&Bsl[off];
              dst[j][1] = *(const v8u16*)&Bsl[RTN * BLD + off];
              dst[j][2] = *(const v8u16*)&Bsl[2 * RTN * BLD + off];
            }
          };
          if constexpr (NBB == 2) bload(0, bf[0]);
  #pragma unroll
          for (int s = 0; s < NS; ++s) {
            if constexpr (NBB == 2) {
              if (s + 1 < NS) bload(s + 1, bf[(s + 1) & 1]);
            } else {
              bload(s, bf[0]);
            }
            __builtin_amdgcn_sched_barrier(0);  // the next step's reads issue before this step's MFMAs
  #pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
              if (rb > 0 && 16 * w + RMAX * rb >= M) break;  // wave-uniform
              const v8u16 ah = ar[rb][3 * s], am = ar[rb][3 * s + 1], al = ar[rb][3 * s + 2];
  #pragma unroll
              for (int j = 0; j < JB; ++j) {
                const int sb = NBB == 2 ? (s & 1) : 0;
                const v8u16 bh = bf[sb][j][0], bm = bf[sb][j][1], bl = bf[sb][j][2];
                // products ordered small to large
                acc[rb][j] = mfma_bf16(al, bh, acc[rb][j]);
                acc[rb][j] = mfma_bf16(am, bm, acc[rb][j]);
                acc[rb][j] = mfma_bf16(ah, bl, acc[rb][j]);
                acc[rb][j] = mfma_bf16(am, bh, acc[rb][j]);
                acc[rb][j] = mfma_bf16(ah, bm, acc[rb][j]);
                acc[rb][j] = mfma_bf16(ah, bh, acc[rb][j]);
              }
            }
          }
        } else {
          // B fragments are read T k-steps at a time, so one LDS round trip
          // covers 4T MFMAs; a running offset keeps one address register
          constexpr int T = CG == 2 ? 2 : 4;
          int bo = (lane >> 4) * BLD + (lane & 15);
  #pragma unroll
          for (int s0 = 0; s0 < NS; s0 += T) {
            TX bv[T][JB];
  #pragma unroll
            for (int u = 0; u < T; ++u)
  #pragma unroll
              for (int j = 0; j < JB; ++j) bv[u][j] = Bs[bo + u * 4 * BLD + 16 * (cg * JB + j)];
            __builtin_amdgcn_sched_barrier(0);  // all T steps' reads issue before their MFMAs
            bo += T * 4 * BLD;
            asm volatile("" : "+v"(bo));
  #pragma unroll
            for (int u = 0; u < T; ++u)
  #pragma unroll
              for (int j = 0; j < JB; ++j) {
                if constexpr (POL == 0)
                  acc[0][j] = mfma_f64(ar[0][s0 + u], bv[u][j], acc[0][j]);
                else
                  acc[0][j] = mfma_f32(ar[0][s0 + u], bv[u][j], acc[0][j]);
              }
          }
        }
      }
  };
  auto orow = [&](int r) { return 16 * w + ((POL == 0) ? mfma16_row_f64(lane, r) : mfma16_row_f32(lane, r)); };
  auto epi_out = [&](long long tt, const AccT& acc) {
    long long xbase, c0;
    int ncol;
    tile_cols(tt, xbase, c0, ncol);
    if (!wact) return;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const int mb = RMAX * rb;  // row offset of this row block
      if (rb > 0 && 16 * w + mb >= M) break;
      if (!rm.qone) {
        TX* dst = Y + (tt / rm.nqt) * (long long)rm.Mtot * Q + (long long)rm.moff * Q + c0 + (lane & 15);
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int m = orow(r) + mb;
            if (m < M) {
              TX& d = dst[(long long)m * Q + 16 * (cg * JB + j)];
              d = rm.accum ? d + (TX)acc[rb][j][r] : (TX)acc[rb][j][r];
            }
          }
      } else {
        // output column c is the contiguous M-vector Y[(c0 + c) M ...]; with FP32
        // accumulators a lane holds 4 consecutive rows of it (one 16-byte store)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
          const int c = 16 * (cg * JB + j) + (lane & 15);
          if (c < ncol) {
            TX* dst = Y + (c0 + c) * rm.Mtot + rm.moff;
            if constexpr (POL == 0) {
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int m = orow(r) + mb;
                if (m < M) dst[m] = rm.accum ? dst[m] + acc[rb][j][r] : acc[rb][j][r];
              }
            } else if (rm.o16) {
              const int m = orow(0) + mb;  // rows m .. m+3, M % 4 == 0
              if (m < M) {
                float4 o = float4{acc[rb][j][0], acc[rb][j][1], acc[rb][j][2], acc[rb][j][3]};
                if (rm.accum) {
                  const float4 d = *(const float4*)(dst + m);
                  o.x += d.x; o.y += d.y; o.z += d.z; o.w += d.w;
                }
                *(float4*)(dst + m) = o;
              }
            } else {
              // even M (Mtot, moff even): two 8-byte pairs, each inside the column or skipped
#pragma unroll
              for (int h2 = 0; h2 < 2; ++h2) {
                const int m = orow(2 * h2) + mb;
                if (m < M) {
                  float2 o = float2{acc[rb][j][2 * h2], acc[rb][j][2 * h2 + 1]};
                  if (rm.accum) {
                    const float2 d = *(const float2*)(dst + m);
                    o.x += d.x;
                    o.y += d.y;
                  }
                  *(float2*)(dst + m) = o;
                }
              }
            }
          }
        }
      }
    }
  };

  // Pipeline: tile i is multiplied from LDS buffer i % 2 while tile i + 1 is
  // staged into the other buffer from registers loaded one tile earlier; one
  // barrier per tile separates the staging of a buffer from its readers.
  long long t = blockIdx.x;
  if (t >= rm.ntiles) return;
  load_tile(t);
  __syncthreads();  // the zero fill above is complete
  store_tile(0);
  if (t + gridDim.x < rm.ntiles) load_tile(t + gridDim.x);
  int buf = 0;
  AccT acc;
  for (;;) {
    __syncthreads();  // buffer buf is staged; the other buffer's readers are done
    compute(acc, buf);
    const long long tn = t + gridDim.x;
    if (tn < rm.ntiles) store_tile(buf ^ 1);
    epi_out(t, acc);
    if (tn >= rm.ntiles) break;
    if (tn + gridDim.x < rm.ntiles) load_tile(tn + gridDim.x);
    t = tn;
    buf ^= 1;
  }
}

static long long ra_grid() {
  static long long g = 0;
  if (!g) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
      ncu = 256;
    g = ncu;
  }
  return g;
}

// Returns 1 (and the status in *rc) if the register-resident kernel handled
// the call, 0 if the shape or alignment needs the general kernels.
template <int POL>
int launch_ra(const void* A, int M, int K, const void* X, void* Y, long long P, long long Q, hipStream_t st,
              int* rc, int Mtot, int moff, int lda, long long xps, int accum) {
  typedef typename RT<POL>::TX TX;
  constexpr int EPC = 16 / sizeof(TX);
  // bf16x6 with K <= 128 takes M <= 256 in one launch (two row blocks per wave)
  const int mmax = (POL == 3 && K <= RMAX) ? 2 * RMAX : RMAX;
  if (M <= 0 || K <= 0 || M > mmax || K > (POL == 3 ? RMAX3 : RMAX) || P <= 0 || Q <= 0) return 0;
  if (((uintptr_t)X & 15) || ((uintptr_t)Y & 15)) return 0;
  RMap rm;
  rm.P = P;
  rm.Q = Q;
  rm.ncols = P * Q;
  rm.M = M;
  rm.Mtot = Mtot > 0 ? Mtot : M;
  rm.moff = moff;
  if (rm.moff < 0 || rm.moff + M > rm.Mtot) return 0;
  rm.K = K;
  // bf16x6 slice rows are zero-padded to a multiple of 32 values
  const int lda0 = POL == 3 ? (K + 31) & ~31 : K;
  rm.lda = lda > 0 ? lda : lda0;
  rm.xps = xps > 0 ? xps : (long long)K * Q;
  rm.accum = accum;
  rm.x16 = (K % 8 == 0 && rm.xps % 4 == 0) ? 1 : 0;
  rm.o16 = (M % 4 == 0 && rm.Mtot % 4 == 0 && rm.moff % 4 == 0) ? 1 : 0;
  const bool ksplit = rm.lda != lda0 || rm.xps != (long long)K * Q || accum;
  if (ksplit && ((Q != 1 && Q % RTN) || rm.lda < lda0 || rm.xps < (long long)K * Q ||
                 (Q == 1 && rm.xps % (POL == 3 ? 8 : EPC))))
    return 0;
  if (Q == 1) {
    // 16-byte chunks never straddle rows (K, M multiples of the chunk; bf16:
    // K % 8 == 0 for its 8-wide k groups).  Staged rows K..4*NS-1 of a k-step
    // keep the zero fill (stores write rows < K only) and the operator
    // registers past K are zero, so FP64 takes any even K (126: a 128-point
    // Dirichlet composite)
    // bf16x6 takes any even K: rows that are not 16-byte aligned or end inside
    // an 8-value group are read in 8-byte pairs
    // bf16x6 also writes any even M (8-byte output pairs)
    if (POL == 3 ? (K % 2 || M % 2 || rm.Mtot % 2 || rm.moff % 2)
                 : (K % EPC || M % EPC || rm.Mtot % EPC || rm.moff % EPC))
      return 0;
    if (POL == 3 && rm.xps % 2) return 0;
    rm.qone = 1;
    rm.nqt = 0;
    rm.ntiles = (P + RTN - 1) / RTN;
  } else if (Q % RTN == 0) {
    rm.qone = 0;
    rm.nqt = Q / RTN;
    rm.ntiles = P * rm.nqt;
  } else {
    return 0;
  }
  typedef void (*KernT)(const void*, const TX*, TX*, RMap);
  // bf16x6: 16 waves in two column groups (more waves per SIMD to hide the
  // staging; 128 VGPRs fit); FP64/FP32 keep 8 waves (their operator rows and
  // B fragments need more registers than 16 waves allow)
  // K > 128 (six k-steps): one column group, whose 8 waves have the
  // registers for the longer operator rows
  int cgs = (POL == 3) ? 2 : 1;
  KernT kern;
  size_t shm;
  if constexpr (POL == 3) {
    const int kst = (K + 31) / 32;  // 32-deep k-steps needed
    if (kst > 4) {
      cgs = 1;
      kern = contract_ra_kernel<POL, 6, 1>;
      shm = RLds<POL, 6>::TOTAL;
    } else if (M > RMAX) {
      // one launch reads X once for all M rows: 8 waves, each with row
      // blocks w and w + 8 sharing every staged B fragment
      cgs = 1;
      kern = kst <= 1 ? contract_ra_kernel<POL, 1, 1, 2> : kst <= 2 ? contract_ra_kernel<POL, 2, 1, 2>
                                                           : contract_ra_kernel<POL, 4, 1, 2>;
      shm = RLds<POL, 4>::TOTAL;
    } else {
      kern = kst <= 1 ? contract_ra_kernel<POL, 1, 2> : kst <= 2 ? contract_ra_kernel<POL, 2, 2>
                                                        : contract_ra_kernel<POL, 4, 2>;
      shm = RLds<POL, 4>::TOTAL;
    }
  } else {
    const int kst = (K + 3) / 4;
    kern = kst <= 8 ? contract_ra_kernel<POL, 8, 1> : kst <= 16 ? contract_ra_kernel<POL, 16, 1>
                                                    : contract_ra_kernel<POL, 32, 1>;
    shm = RLds<POL, 8>::TOTAL;
  }
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
  if (e != hipSuccess) {
    set_error("contract_ra: hipFuncSetAttribute: %s", hipGetErrorString(e));
    *rc = JFX_EHIP;
    return 1;
  }
  long long grid = ra_grid();
  if (grid > rm.ntiles) grid = rm.ntiles;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(RNT * cgs), shm, st, A, (const TX*)X, (TX*)Y, rm);
  e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("contract_ra launch: %s", hipGetErrorString(e));
    *rc = JFX_EHIP;
    return 1;
  }
  *rc = JFX_OK;
  return 1;
}

template int launch_ra<0>(const void*, int, int, const void*, void*, long long, long long, hipStream_t, int*, int,
                          int, int, long long, int);
template int launch_ra<1>(const void*, int, int, const void*, void*, long long, long long, hipStream_t, int*, int,
                          int, int, long long, int);
template int launch_ra<3>(const void*, int, int, const void*, void*, long long, long long, hipStream_t, int*, int,
                          int, int, long long, int);

}  // namespace jfx
