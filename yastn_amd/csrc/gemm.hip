// Grouped FP64 / complex128 GEMM on CDNA4 FP64 matrix cores.
//
// One launch computes every output block of a plan:
//   C_g (=|+=) sum_t op(A_t) op(B_t)
// where op() reads an operand through (row, col) strides with optional
// conjugation, so A^H / B^H of the dot adjoint need no copies.  Output tiles are
// owned by exactly one workgroup and terms are summed in a fixed order, so results
// are bitwise reproducible (no atomics).
//
// Tile: BM x BN per workgroup of 4 waves (WM x WN); BK-deep slices staged through two LDS buffers
// k-major; each wave issues v_mfma_f64_16x16x4f64 on FM x FN 16x16 fragments.
// Complex data is held as separate real / imaginary LDS planes and multiplied
// with four real MFMAs per fragment.
#include <mutex>

#include "common.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int BK = 16;
constexpr int PAD = 2;

struct Bufs {
  const void* in[YH_MAX_BUFS];
  void* out[YH_MAX_BUFS];
};

template <bool CPLX>
struct Elt;
template <>
struct Elt<false> {
  typedef double T;
};
template <>
struct Elt<true> {
  typedef double2 T;
};

// Zero elements (never written), read in place of operand elements outside a block; thread t reads
// element t mod kZeroSpan, so that the fallback loads of a wave spread over cache lines like ordinary
// coalesced loads instead of all hitting one line.
constexpr int kZeroSpan = 256;
__device__ double2 kZeroElt[kZeroSpan];

// Workgroups per CU the register allocation must allow: 4 for 32x32 tiles (<= 128 VGPRs, no spills;
// the short-K launches of many tiles are bound by how many tiles a CU holds at once), else no bound.
#define GEMM_MINB(CPLX, BM, BN) ((BM) == 32 && (BN) == 32 ? 4 : 1)

__device__ inline d4 mfma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// One workgroup per output tile.  Per term, the k loop is software-pipelined: the next BK slice of
// op(A) and op(B) is loaded into registers while the MFMAs consume the current slice from one of two
// LDS buffers, then stored into the other buffer; one barrier per slice.  Tiles with a single 16x16
// fragment per wave keep NS partial accumulators (k sub-steps rotate over them, summed at the end)
// so that consecutive MFMAs do not wait on each other.
template <bool CPLX, int BM, int BN, int WM, int WN, int BK>
__global__ void __launch_bounds__(kThreads, GEMM_MINB(CPLX, BM, BN)) gemm_kernel(Bufs bufs, const int64_t* __restrict__ gemms,
                                                        const int64_t* __restrict__ terms,
                                                        const int32_t* __restrict__ tiles) {
  typedef typename Elt<CPLX>::T T;
  constexpr int FM = BM / WM / 16;
  constexpr int FN = BN / WN / 16;
  constexpr int NPL = CPLX ? 2 : 1;  // LDS planes
  constexpr int NS = (FM * FN >= 4) ? 1 : 4 / (FM * FN);
  constexpr int KSUB = BK / 4;
  constexpr int NSU = NS < KSUB ? NS : KSUB;  // partial accumulator sets actually used
  constexpr int A_PER = (BM * BK + kThreads - 1) / kThreads;
  constexpr int B_PER = (BN * BK + kThreads - 1) / kThreads;
  constexpr bool A_FULL = (BM * BK) % kThreads == 0;  // every thread has all A_PER elements
  constexpr bool B_FULL = (BN * BK) % kThreads == 0;
  static_assert(BK % 4 == 0, "BK");
  __shared__ double As[2][NPL][BK][BM + PAD];
  __shared__ double Bs[2][NPL][BK][BN + PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;

  const int32_t* tl = tiles + 3 * (int64_t)blockIdx.x;
  const int64_t g = tl[0];
  const int64_t m0 = tl[1], n0 = tl[2];
  const int64_t* gw = gemms + g * YH_GEMM_WORDS;
  const int64_t M = gw[0], N = gw[1], c_off = gw[2], c_rs = gw[3], c_cs = gw[4];
  const int64_t tb = gw[5], tc = gw[6], gflags = gw[7];

  d4 accR[NSU][FM][FN], accI[NSU][FM][FN];
#pragma unroll
  for (int s = 0; s < NSU; ++s)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        accR[s][i][j] = d4{0, 0, 0, 0};
        accI[s][i][j] = d4{0, 0, 0, 0};
      }

  for (int64_t q = tb; q < tb + tc; ++q) {
    const int64_t* tw = terms + q * YH_TERM_WORDS;
    const int64_t a_off = tw[0], a_rs = tw[1], a_cs = tw[2];
    const int64_t b_off = tw[3], b_rs = tw[4], b_cs = tw[5], K = tw[6], tf = tw[7];
    const bool conjA = tf & 1, conjB = (tf >> 1) & 1;
    const T* __restrict__ A = (const T*)bufs.in[(tf >> 8) & 0xff];
    const T* __restrict__ B = (const T*)bufs.in[(tf >> 16) & 0xff];
    const bool a_kfast = (a_cs == 1);
    const bool b_nfast = (b_cs == 1);
    if (K <= 0) continue;

    // per-thread slice coordinates (fixed for the term); the k coordinate advances by BK per slice
    int a_i[A_PER], a_k[A_PER], b_j[B_PER], b_k[B_PER];
    bool a_in[A_PER], b_in[B_PER];
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int e = tid + p * kThreads;
      if (a_kfast) {
        a_k[p] = e % BK;
        a_i[p] = e / BK;
      } else {
        a_i[p] = e % BM;
        a_k[p] = e / BM;
      }
      a_in[p] = (A_FULL || e < BM * BK) && (m0 + a_i[p] < M);
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = tid + p * kThreads;
      if (b_nfast) {
        b_j[p] = e % BN;
        b_k[p] = e / BN;
      } else {
        b_k[p] = e % BK;
        b_j[p] = e / BK;
      }
      b_in[p] = (B_FULL || e < BN * BK) && (n0 + b_j[p] < N);
    }

    // Two register sets: while slice s is consumed from LDS, slice s+1 waits in one set (loaded an
    // iteration earlier) and slice s+2 is being loaded into the other, so two slices of global loads
    // are in flight; slice s+1 goes to the other LDS buffer at the end of the iteration.
    T ra0[A_PER], rb0[B_PER], ra1[A_PER], rb1[B_PER];
    // Every lane issues every load; an element outside the operand is read from kZeroElt.  (A select
    // after the load would be hoisted by the scheduler and wait for loads still meant to be in flight.)
    const T* zero = (const T*)kZeroElt + (tid & (kZeroSpan - 1));
    const T* a_row[A_PER];
    const T* b_col[B_PER];
#pragma unroll
    for (int p = 0; p < A_PER; ++p) a_row[p] = A + a_off + (m0 + a_i[p]) * a_rs + a_k[p] * a_cs;
#pragma unroll
    for (int p = 0; p < B_PER; ++p) b_col[p] = B + b_off + b_k[p] * b_rs + (n0 + b_j[p]) * b_cs;
    auto load = [&](int64_t k0, T* ra, T* rb) {
      const int64_t da = k0 * a_cs, db = k0 * b_rs;
#pragma unroll
      for (int p = 0; p < A_PER; ++p) ra[p] = *((a_in[p] & (k0 + a_k[p] < K)) ? a_row[p] + da : zero);
#pragma unroll
      for (int p = 0; p < B_PER; ++p) rb[p] = *((b_in[p] & (k0 + b_k[p] < K)) ? b_col[p] + db : zero);
    };
    auto store = [&](int buf, const T* ra, const T* rb) {
#pragma unroll
      for (int p = 0; p < A_PER; ++p) {
        if (A_FULL || tid + p * kThreads < BM * BK) {
          if constexpr (CPLX) {
            As[buf][0][a_k[p]][a_i[p]] = ra[p].x;
            As[buf][1][a_k[p]][a_i[p]] = conjA ? -ra[p].y : ra[p].y;
          } else {
            As[buf][0][a_k[p]][a_i[p]] = ra[p];
          }
        }
      }
#pragma unroll
      for (int p = 0; p < B_PER; ++p) {
        if (B_FULL || tid + p * kThreads < BN * BK) {
          if constexpr (CPLX) {
            Bs[buf][0][b_k[p]][b_j[p]] = rb[p].x;
            Bs[buf][1][b_k[p]][b_j[p]] = conjB ? -rb[p].y : rb[p].y;
          } else {
            Bs[buf][0][b_k[p]][b_j[p]] = rb[p];
          }
        }
      }
    };
    // ---- One BK slice in two steps: read_frag copies the wave's operand fragments of LDS buffer `buf`
    // into registers, mma runs the MFMAs on them (4 k per instruction).  The fragments of slice s+1 are
    // read while the MFMAs of slice s run, so the LDS round trip after each barrier is hidden.
    typedef double FragA[KSUB][FM];
    typedef double FragB[KSUB][FN];
    FragA far0, fai0, far1, fai1;
    FragB fbr0, fbi0, fbr1, fbi1;
    auto read_frag = [&](int buf, FragA& ar, FragA& ai, FragB& br, FragB& bi) {
#pragma unroll
      for (int ks = 0; ks < KSUB; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int col = wm * (BM / WM) + i * 16 + (lane & 15);
          ar[ks][i] = As[buf][0][kr][col];
          if constexpr (CPLX) ai[ks][i] = As[buf][1][kr][col];
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int col = wn * (BN / WN) + j * 16 + (lane & 15);
          br[ks][j] = Bs[buf][0][kr][col];
          if constexpr (CPLX) bi[ks][j] = Bs[buf][1][kr][col];
        }
      }
    };
    auto mma = [&](const FragA& ar, const FragA& ai, const FragB& br, const FragB& bi) {
#pragma unroll
      for (int ks = 0; ks < KSUB; ++ks) {
        const int s = ks % NSU;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            accR[s][i][j] = mfma(ar[ks][i], br[ks][j], accR[s][i][j]);
            if constexpr (CPLX) {
              accR[s][i][j] = mfma(-ai[ks][i], bi[ks][j], accR[s][i][j]);
              accI[s][i][j] = mfma(ar[ks][i], bi[ks][j], accI[s][i][j]);
              accI[s][i][j] = mfma(ai[ks][i], br[ks][j], accI[s][i][j]);
            }
          }
      }
    };

    const int64_t ns = (K + BK - 1) / BK;  // slices
    // The loads are issued unconditionally (slices past the end read only in-range fallbacks): a
    // load under a branch leaves the compiler unsure how many are in flight, and it then waits for
    // all of them before the next LDS store.  A buffer is overwritten only after a barrier that
    // follows the completion of every wave's fragment reads of it (__syncthreads waits for them).
    load(0, ra0, rb0);
    store(0, ra0, rb0);
    load(BK, ra0, rb0);
    load(2 * BK, ra1, rb1);
    __syncthreads();
    read_frag(0, far0, fai0, fbr0, fbi0);
    int64_t s = 0;
    for (; s + 1 < ns; s += 2) {
      // fragments 0 hold slice s; register set 0 slice s+1, set 1 slice s+2
      store(1, ra0, rb0);
      load((s + 3) * BK, ra0, rb0);
      __syncthreads();
      read_frag(1, far1, fai1, fbr1, fbi1);
      mma(far0, fai0, fbr0, fbi0);
      // fragments 1 hold slice s+1; set 1 slice s+2 (all zeros past the end), set 0 slice s+3
      store(0, ra1, rb1);
      load((s + 4) * BK, ra1, rb1);
      __syncthreads();
      read_frag(0, far0, fai0, fbr0, fbi0);
      mma(far1, fai1, fbr1, fbi1);
    }
    if (s < ns) mma(far0, fai0, fbr0, fbi0);  // odd slice count: fragments 0 hold the last slice
    __syncthreads();  // the next term's first store reuses buffer 0
  }

  // ---- epilogue: v_mfma_f64_16x16x4f64 leaves D[lane/16 + 4 v][lane % 16] in item v of each lane
  T* __restrict__ C = (T*)bufs.out[(gflags >> 8) & 0xff];
  const bool accumulate = gflags & 1;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      d4 sr = accR[0][i][j], si = accI[0][i][j];
#pragma unroll
      for (int s = 1; s < NSU; ++s) {
        sr += accR[s][i][j];
        si += accI[s][i][j];
      }
      const int64_t col = n0 + wn * (BN / WN) + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = m0 + wm * (BM / WM) + i * 16 + (lane >> 4) + 4 * v;
        if (row < M && col < N) {
          T* p = C + c_off + row * c_rs + col * c_cs;
          if constexpr (CPLX) {
            double2 r = make_double2(sr[v], si[v]);
            if (accumulate) {
              const double2 o = *p;
              r.x += o.x;
              r.y += o.y;
            }
            *p = r;
          } else {
            double r = sr[v];
            if (accumulate) r += *p;
            *p = r;
          }
        }
      }
    }
}


// Register-direct 32x32 tile (YH_TILE_RD32): no LDS in the k loop and no barrier.  The four waves split
// each term's k-steps (4 k each) into quarters; every wave computes the whole tile (2 x 2 fragments of
// 16 x 16) over its quarter, lane l supplying A(row l % 16, k l / 16) and B(k l / 16, col l % 16) of each
// step straight from memory, kRdDepth steps ahead of the MFMAs (an element outside the operand, or a
// step past the wave's quarter, is read from kZeroElt and adds zero).  The quarters are summed through
// LDS in wave order, one fragment per wave, so the result does not depend on timing.
constexpr int kRdDepth = 4;
template <bool CPLX>
__global__ void __launch_bounds__(kThreads, CPLX ? 2 : 4) gemm_rd_kernel(Bufs bufs, const int64_t* __restrict__ gemms,
                                                             const int64_t* __restrict__ terms,
                                                             const int32_t* __restrict__ tiles) {
  typedef typename Elt<CPLX>::T T;
  constexpr int NPL = CPLX ? 2 : 1;
  __shared__ double red[4][NPL][4][4][64];  // [wave][plane][fragment][item][lane]
  // the wave index as a scalar: the k loop bounds depend on it, and a loop the compiler thinks divergent
  // ends each trip waiting for every load in flight
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r16 = lane & 15, kq = lane >> 4;
  const int32_t* tl = tiles + 3 * (int64_t)blockIdx.x;
  const int64_t g = tl[0];
  const int64_t m0 = tl[1], n0 = tl[2];
  const int64_t* gw = gemms + g * YH_GEMM_WORDS;
  const int64_t M = gw[0], N = gw[1], c_off = gw[2], c_rs = gw[3], c_cs = gw[4];
  const int64_t tb = gw[5], tc = gw[6], gflags = gw[7];
  const T* zero = (const T*)kZeroElt + (threadIdx.x & (kZeroSpan - 1));
  d4 accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = d4{0, 0, 0, 0};
      accI[i][j] = d4{0, 0, 0, 0};
    }
  const bool oka0 = m0 + r16 < M, oka1 = m0 + 16 + r16 < M;
  const bool okb0 = n0 + r16 < N, okb1 = n0 + 16 + r16 < N;
  for (int64_t q = tb; q < tb + tc; ++q) {
    const int64_t* tw = terms + q * YH_TERM_WORDS;
    const int64_t a_off = tw[0], a_rs = tw[1], a_cs = tw[2];
    const int64_t b_off = tw[3], b_rs = tw[4], b_cs = tw[5], K = tw[6], tf = tw[7];
    if (K <= 0) continue;
    const bool conjA = tf & 1, conjB = (tf >> 1) & 1;
    const T* __restrict__ A = (const T*)bufs.in[(tf >> 8) & 0xff];
    const T* __restrict__ B = (const T*)bufs.in[(tf >> 16) & 0xff];
    const T* pa0 = A + a_off + (m0 + r16) * a_rs;
    const T* pa1 = pa0 + 16 * a_rs;
    const T* pb0 = B + b_off + (n0 + r16) * b_cs;
    const T* pb1 = pb0 + 16 * b_cs;
    const int64_t nst = (K + 3) / 4, per = (nst + 3) / 4;
    const int64_t s0 = wave * per, s1 = min(nst, s0 + per);
    // step s of this lane: k = 4 s + kq; values a0, a1, b0, b1
    auto load = [&](int64_t s, T* v) {
      const int64_t k = 4 * s + kq;
      const bool ok = (s < s1) & (k < K);
      v[0] = *((ok & oka0) ? pa0 + k * a_cs : zero);
      v[1] = *((ok & oka1) ? pa1 + k * a_cs : zero);
      v[2] = *((ok & okb0) ? pb0 + k * b_rs : zero);
      v[3] = *((ok & okb1) ? pb1 + k * b_rs : zero);
    };
    T buf[kRdDepth][4];
#pragma unroll
    for (int d = 0; d < kRdDepth; ++d) load(s0 + d, buf[d]);
    for (int64_t s = s0; s < s1; s += kRdDepth) {
#pragma unroll
      for (int d = 0; d < kRdDepth; ++d) {
        double ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          if constexpr (CPLX) {
            ar[i] = buf[d][i].x;
            ai[i] = conjA ? -buf[d][i].y : buf[d][i].y;
            br[i] = buf[d][2 + i].x;
            bi[i] = conjB ? -buf[d][2 + i].y : buf[d][2 + i].y;
          } else {
            ar[i] = buf[d][i];
            br[i] = buf[d][2 + i];
          }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            accR[i][j] = mfma(ar[i], br[j], accR[i][j]);
            if constexpr (CPLX) {
              accR[i][j] = mfma(-ai[i], bi[j], accR[i][j]);
              accI[i][j] = mfma(ar[i], bi[j], accI[i][j]);
              accI[i][j] = mfma(ai[i], br[j], accI[i][j]);
            }
          }
        // refill the slot after its last use, so that the slot keeps its registers from trip to trip (an
        // overlap of old and new values would need copies, and a copy waits for its load); the program
        // order is kept: the scheduler would otherwise hoist consumers next to their loads
        __builtin_amdgcn_sched_barrier(0);
        load(s + d + kRdDepth, buf[d]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // quarters -> LDS; wave w sums fragment w over the waves in order and writes it
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        red[wave][0][i * 2 + j][v][lane] = accR[i][j][v];
        if constexpr (CPLX) red[wave][NPL - 1][i * 2 + j][v][lane] = accI[i][j][v];
      }
  __syncthreads();
  T* __restrict__ C = (T*)bufs.out[(gflags >> 8) & 0xff];
  const bool accumulate = gflags & 1;
  const int fi = wave >> 1, fj = wave & 1;
  const int64_t col = n0 + fj * 16 + r16;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double sr = red[0][0][wave][v][lane], si = CPLX ? red[0][NPL - 1][wave][v][lane] : 0.0;
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      sr += red[w][0][wave][v][lane];
      if constexpr (CPLX) si += red[w][NPL - 1][wave][v][lane];
    }
    const int64_t row = m0 + fi * 16 + kq + 4 * v;
    if (row < M && col < N) {
      T* p = C + c_off + row * c_rs + col * c_cs;
      if constexpr (CPLX) {
        double2 r = make_double2(sr, si);
        if (accumulate) {
          const double2 o = *p;
          r.x += o.x;
          r.y += o.y;
        }
        *p = r;
      } else {
        double r = sr;
        if (accumulate) r += *p;
        *p = r;
      }
    }
  }
}

// Split-K reduction: for each record, C (=|+=) sum_s P_s in the order s = 0, 1, ... (deterministic).
constexpr unsigned kSplitKBlocks = 32;
// Record words (YH_SPLITK_WORDS): 0 M, 1 N, 2 c_off, 3 c_rs, 4 c_cs, 5 p_off (partial s at
// p_off + s M N, row-major M x N), 6 S, 7 flags (bit 0 accumulate, bits 8-15 output buffer).
template <bool CPLX>
__global__ void __launch_bounds__(kThreads) splitk_reduce_kernel(Bufs bufs, const typename Elt<CPLX>::T* __restrict__ P,
                                                                 const int64_t* __restrict__ recs) {
  typedef typename Elt<CPLX>::T T;
  const int64_t* w = recs + (int64_t)blockIdx.y * YH_SPLITK_WORDS;
  const int64_t M = w[0], N = w[1], c_off = w[2], c_rs = w[3], c_cs = w[4], p_off = w[5], S = w[6], fl = w[7];
  T* __restrict__ C = (T*)bufs.out[(fl >> 8) & 0xff];
  // kSplitKBlocks workgroups per record stride over its M N elements: a record's partials are read by
  // many CUs at once (one workgroup per record was bound by a single CU's memory throughput)
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < M * N; e += (int64_t)gridDim.x * kThreads) {
    const int64_t i = e / N, j = e - i * N;
    T acc = P[p_off + e];
    for (int64_t s = 1; s < S; ++s) {
      const T v = P[p_off + s * M * N + e];
      if constexpr (CPLX) {
        acc.x += v.x;
        acc.y += v.y;
      } else {
        acc += v;
      }
    }
    T* c = C + c_off + i * c_rs + j * c_cs;
    if (fl & 1) {
      if constexpr (CPLX) {
        acc.x += (*c).x;
        acc.y += (*c).y;
      } else {
        acc += *c;
      }
    }
    *c = acc;
  }
}


// Streaming kernel for GEMMs whose M is at most 16 (e.g. contractions with a small MPO leg): a tile
// is all M rows times 256 columns, one thread per column; the A entries are wave-uniform (scalar
// loads), B and C are read / written along the columns, so the kernel runs at the HBM rate.
constexpr int kSmallM = 16, kSmallN = 256;

template <bool CPLX>
__global__ void __launch_bounds__(kThreads) gemm_smallm_kernel(Bufs bufs, const int64_t* __restrict__ gemms,
                                                               const int64_t* __restrict__ terms,
                                                               const int32_t* __restrict__ tiles) {
  typedef typename Elt<CPLX>::T T;
  const int32_t* tl = tiles + 3 * (int64_t)blockIdx.x;
  const int64_t* gw = gemms + (int64_t)tl[0] * YH_GEMM_WORDS;
  const int64_t M = gw[0], N = gw[1], c_off = gw[2], c_rs = gw[3], c_cs = gw[4];
  const int64_t tb = gw[5], tc = gw[6], gflags = gw[7];
  const int64_t j = (int64_t)tl[2] + threadIdx.x;
  double accr[kSmallM], acci[kSmallM];
#pragma unroll
  for (int i = 0; i < kSmallM; ++i) {
    accr[i] = 0.0;
    acci[i] = 0.0;
  }
  const bool live = j < N;
  for (int64_t q = tb; q < tb + tc; ++q) {
    const int64_t* tw = terms + q * YH_TERM_WORDS;
    const int64_t a_off = tw[0], a_rs = tw[1], a_cs = tw[2];
    const int64_t b_off = tw[3], b_rs = tw[4], b_cs = tw[5], K = tw[6], tf = tw[7];
    const double sa = (tf & 1) ? -1.0 : 1.0, sb = (tf & 2) ? -1.0 : 1.0;  // conjugation signs
    const T* __restrict__ A = (const T*)bufs.in[(tf >> 8) & 0xff];
    const T* __restrict__ B = (const T*)bufs.in[(tf >> 16) & 0xff];
    for (int64_t k = 0; k < K; ++k) {
      T b;
      if (live)
        b = B[b_off + k * b_rs + j * b_cs];
      else
        b = T{};
#pragma unroll
      for (int i = 0; i < kSmallM; ++i) {
        if (i < M) {
          const T a = A[a_off + i * a_rs + k * a_cs];
          if constexpr (CPLX) {
            const double ar = a.x, ai = sa * a.y, br = b.x, bi = sb * b.y;
            accr[i] += ar * br - ai * bi;
            acci[i] += ar * bi + ai * br;
          } else {
            accr[i] += a * b;
          }
        }
      }
    }
  }
  if (!live) return;
  T* __restrict__ C = (T*)bufs.out[(gflags >> 8) & 0xff];
  const bool accumulate = gflags & 1;
#pragma unroll
  for (int i = 0; i < kSmallM; ++i) {
    if (i < M) {
      T* p = C + c_off + i * c_rs + j * c_cs;
      if constexpr (CPLX) {
        double2 r = make_double2(accr[i], acci[i]);
        if (accumulate) {
          const double2 o = *p;
          r.x += o.x;
          r.y += o.y;
        }
        *p = r;
      } else {
        double r = accr[i];
        if (accumulate) r += *p;
        *p = r;
      }
    }
  }
}

// BK per configuration (measured on the Heff2 chi=1024 and U(1)xU(1) chi=2048 sector GEMMs): shorter
// slices are faster now that two slices of global loads really are in flight; the double-buffered
// LDS tiles stay within 64 KiB.
template <bool CPLX>
int launch(const Bufs& b, const int64_t* gemms, const int64_t* terms, const int32_t* tiles, int64_t ntiles,
           int cfg, hipStream_t s) {
  const dim3 grid((unsigned)ntiles), block(kThreads);
  constexpr int BKW = 8;              // 64 / 128-wide tiles
  constexpr int BKH = CPLX ? 4 : 8;   // 128 x 128
  constexpr int BKS = CPLX ? 8 : 16;  // 32-wide and skinny tiles
  switch (cfg) {
    case YH_TILE_32x32:
      hipLaunchKernelGGL((gemm_kernel<CPLX, 32, 32, 2, 2, BKS>), grid, block, 0, s, b, gemms, terms, tiles);
      break;
    case YH_TILE_64x64:
      hipLaunchKernelGGL((gemm_kernel<CPLX, 64, 64, 2, 2, BKW>), grid, block, 0, s, b, gemms, terms, tiles);
      break;
    case YH_TILE_128x64:
      hipLaunchKernelGGL((gemm_kernel<CPLX, 128, 64, 2, 2, BKW>), grid, block, 0, s, b, gemms, terms, tiles);
      break;
    case YH_TILE_16x64:
      hipLaunchKernelGGL((gemm_kernel<CPLX, 16, 64, 1, 4, BKS>), grid, block, 0, s, b, gemms, terms, tiles);
      break;
    case YH_TILE_64x16:
      hipLaunchKernelGGL((gemm_kernel<CPLX, 64, 16, 4, 1, BKS>), grid, block, 0, s, b, gemms, terms, tiles);
      break;
    case YH_TILE_128x128:
      hipLaunchKernelGGL((gemm_kernel<CPLX, 128, 128, 2, 2, BKH>), grid, block, 0, s, b, gemms, terms, tiles);
      break;
    case YH_TILE_RD32:
      hipLaunchKernelGGL(gemm_rd_kernel<CPLX>, grid, block, 0, s, b, gemms, terms, tiles);
      break;
    case YH_TILE_SMALLM:
      hipLaunchKernelGGL((gemm_smallm_kernel<CPLX>), grid, block, 0, s, b, gemms, terms, tiles);
      break;
    default:
      yh::set_error("unknown tile config %d", cfg);
      return YH_ERR_ARG;
  }
  return yh::hip_status(hipGetLastError(), "grouped_gemm launch");
}

void tile_dims(int cfg, int* bm, int* bn) {
  switch (cfg) {
    case YH_TILE_32x32: *bm = 32; *bn = 32; break;
    case YH_TILE_64x64: *bm = 64; *bn = 64; break;
    case YH_TILE_128x64: *bm = 128; *bn = 64; break;
    case YH_TILE_16x64: *bm = 16; *bn = 64; break;
    case YH_TILE_128x128: *bm = 128; *bn = 128; break;
    case YH_TILE_SMALLM: *bm = kSmallM; *bn = kSmallN; break;
    case YH_TILE_RD32: *bm = 32; *bn = 32; break;
    default: *bm = 64; *bn = 16; break;
  }
}

}  // namespace

extern "C" int yastn_hip_check_gemm_plan(const int64_t* h_gemms, int64_t ngemms, const int64_t* h_terms,
                                         int64_t nterms, const int32_t* h_tiles, int64_t ntiles, int cfg,
                                         const int64_t* in_numel, int nin, const int64_t* out_numel, int nout) {
  YH_CHECK_ARG(cfg >= YH_TILE_32x32 && cfg <= YH_TILE_RD32, "tile config %d", cfg);
  YH_CHECK_ARG(nin <= YH_MAX_BUFS && nout <= YH_MAX_BUFS, "too many buffers");
  int bm, bn;
  tile_dims(cfg, &bm, &bn);
  for (int64_t g = 0; g < ngemms; ++g) {
    const int64_t* w = h_gemms + g * YH_GEMM_WORDS;
    const int64_t M = w[0], N = w[1], c_off = w[2], c_rs = w[3], c_cs = w[4], tb = w[5], tc = w[6];
    const int out = (int)((w[7] >> 8) & 0xff);
    YH_CHECK_ARG(M >= 0 && N >= 0 && c_rs >= 0 && c_cs >= 0, "gemm %lld: negative dims", (long long)g);
    YH_CHECK_ARG(cfg != YH_TILE_SMALLM || M <= bm, "gemm %lld: M = %lld exceeds the small-M tile", (long long)g,
                 (long long)M);
    YH_CHECK_ARG(out < nout, "gemm %lld: output buffer %d", (long long)g, out);
    if (M > 0 && N > 0)
      YH_CHECK_ARG(c_off >= 0 && c_off + (M - 1) * c_rs + (N - 1) * c_cs < out_numel[out],
                   "gemm %lld: output out of bounds", (long long)g);
    YH_CHECK_ARG(tb >= 0 && tc >= 0 && tb + tc <= nterms, "gemm %lld: term range", (long long)g);
    for (int64_t q = tb; q < tb + tc; ++q) {
      const int64_t* t = h_terms + q * YH_TERM_WORDS;
      const int64_t K = t[6];
      const int ia = (int)((t[7] >> 8) & 0xff), ib = (int)((t[7] >> 16) & 0xff);
      YH_CHECK_ARG(ia < nin && ib < nin, "term %lld: input buffer", (long long)q);
      YH_CHECK_ARG(K >= 0 && t[1] >= 0 && t[2] >= 0 && t[4] >= 0 && t[5] >= 0, "term %lld: negative", (long long)q);
      if (M > 0 && K > 0)
        YH_CHECK_ARG(t[0] >= 0 && t[0] + (M - 1) * t[1] + (K - 1) * t[2] < in_numel[ia],
                     "term %lld: A out of bounds", (long long)q);
      if (N > 0 && K > 0)
        YH_CHECK_ARG(t[3] >= 0 && t[3] + (K - 1) * t[4] + (N - 1) * t[5] < in_numel[ib],
                     "term %lld: B out of bounds", (long long)q);
    }
  }
  for (int64_t t = 0; t < ntiles; ++t) {
    const int32_t* x = h_tiles + 3 * t;
    YH_CHECK_ARG(x[0] >= 0 && x[0] < ngemms, "tile %lld: gemm index", (long long)t);
    const int64_t* w = h_gemms + (int64_t)x[0] * YH_GEMM_WORDS;
    YH_CHECK_ARG(x[1] >= 0 && x[1] < w[0] && x[1] % bm == 0 && x[2] >= 0 && x[2] < w[1] && x[2] % bn == 0,
                 "tile %lld: origin", (long long)t);
  }
  return YH_OK;
}

extern "C" int yastn_hip_grouped_gemm(void* const* outs, const void* const* ins, int dtype, const int64_t* d_gemms,
                                      int64_t ngemms, const int64_t* d_terms, int64_t nterms,
                                      const int32_t* d_tiles, int64_t ntiles, int cfg, hipStream_t stream) {
  (void)ngemms;
  (void)nterms;
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(ntiles >= 0 && ntiles < (int64_t)INT32_MAX, "tile count");
  if (ntiles == 0) return YH_OK;
  Bufs b;
  for (int i = 0; i < YH_MAX_BUFS; ++i) {
    b.in[i] = ins[i];
    b.out[i] = outs[i];
  }
  return dtype == YH_F64 ? launch<false>(b, d_gemms, d_terms, d_tiles, ntiles, cfg, stream)
                         : launch<true>(b, d_gemms, d_terms, d_tiles, ntiles, cfg, stream);
}

extern "C" int yastn_hip_dot(void* C, const void* A, const void* B, int dtype, const int64_t* d_gemms,
                             int64_t ngemms, const int64_t* d_terms, int64_t nterms, const int32_t* d_tiles,
                             int64_t ntiles, int cfg, hipStream_t stream) {
  void* outs[YH_MAX_BUFS] = {C, nullptr, nullptr, nullptr};
  const void* ins[YH_MAX_BUFS] = {A, B, nullptr, nullptr};
  return yastn_hip_grouped_gemm(outs, ins, dtype, d_gemms, ngemms, d_terms, nterms, d_tiles, ntiles, cfg, stream);
}

extern "C" int yastn_hip_dot_adjoint(void* gA, void* gB, const void* gC, const void* A, const void* B, int dtype,
                                     const int64_t* d_gemms, int64_t ngemms, const int64_t* d_terms,
                                     int64_t nterms, const int32_t* d_tiles, int64_t ntiles, int cfg,
                                     hipStream_t stream) {
  void* outs[YH_MAX_BUFS] = {gA, gB, nullptr, nullptr};
  const void* ins[YH_MAX_BUFS] = {A, B, gC, nullptr};
  return yastn_hip_grouped_gemm(outs, ins, dtype, d_gemms, ngemms, d_terms, nterms, d_tiles, ntiles, cfg, stream);
}

extern "C" int yastn_hip_check_splitk_plan(const int64_t* h_recs, int64_t nrec, const int64_t* out_numel, int nout,
                                           int64_t part_numel) {
  YH_CHECK_ARG(nout <= YH_MAX_BUFS, "too many buffers");
  for (int64_t r = 0; r < nrec; ++r) {
    const int64_t* w = h_recs + r * YH_SPLITK_WORDS;
    const int64_t M = w[0], N = w[1], c_off = w[2], c_rs = w[3], c_cs = w[4], p_off = w[5], S = w[6];
    const int out = (int)((w[7] >> 8) & 0xff);
    YH_CHECK_ARG(M >= 1 && N >= 1 && S >= 1 && c_rs >= 0 && c_cs >= 0 && out < nout, "split-k record %lld",
                 (long long)r);
    YH_CHECK_ARG(c_off >= 0 && c_off + (M - 1) * c_rs + (N - 1) * c_cs < out_numel[out],
                 "split-k record %lld: output out of bounds", (long long)r);
    YH_CHECK_ARG(p_off >= 0 && p_off + S * M * N <= part_numel, "split-k record %lld: partials out of bounds",
                 (long long)r);
  }
  return YH_OK;
}

extern "C" int yastn_hip_splitk_reduce(void* const* outs, const void* partials, int dtype, const int64_t* d_recs,
                                       int64_t nrec, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(nrec >= 0 && nrec <= 65535, "split-k records %lld", (long long)nrec);
  if (nrec == 0) return YH_OK;
  const dim3 grid(kSplitKBlocks, (unsigned)nrec);
  Bufs b;
  for (int i = 0; i < YH_MAX_BUFS; ++i) {
    b.in[i] = nullptr;
    b.out[i] = outs[i];
  }
  if (dtype == YH_F64)
    hipLaunchKernelGGL(splitk_reduce_kernel<false>, grid, dim3(kThreads), 0, stream, b, (const double*)partials,
                       d_recs);
  else
    hipLaunchKernelGGL(splitk_reduce_kernel<true>, grid, dim3(kThreads), 0, stream, b, (const double2*)partials,
                       d_recs);
  return yh::hip_status(hipGetLastError(), "split-k reduce launch");
}

// ---- the split-K GEMM without allocation ----------------------------------------------------------
// Scratch of yastn_hip_grouped_gemm_noalloc: the split-K partial products live in this file-scope
// device array (one copy per device, held until the process ends).  Host code never passes its name
// to a kernel: it takes the array's per-device address from hipGetSymbolAddress, once per device.
namespace {
constexpr int64_t kScratchDoubles = 4 * 1024 * 1024;  // 32 MiB
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) double scratch[kScratchDoubles];

static int scratch_address(void** out) {
  int dev = -1;
  int st = yh::hip_status(hipGetDevice(&dev), "current device");
  if (st) return st;
  YH_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    st = yh::hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
    if (st) return st;
    YH_CHECK_ARG(p != nullptr, "scratch address of device %d", dev);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return YH_OK;
}

extern "C" int yastn_hip_grouped_gemm_noalloc(void* const* outs, const void* const* ins, int dtype,
                                              const int64_t* d_gemms, int64_t ngemms, const int64_t* d_terms,
                                              int64_t nterms, const int32_t* d_tiles, int64_t ntiles, int cfg,
                                              int part_buf, int64_t part_numel, const int64_t* d_red, int64_t nred,
                                              hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(part_buf >= 0 && part_buf < YH_MAX_BUFS, "partials buffer slot %d", part_buf);
  YH_CHECK_ARG(outs[part_buf] == nullptr, "output slot %d of the partials is taken", part_buf);
  YH_CHECK_ARG(part_numel >= 0, "partials size %lld", (long long)part_numel);
  const int64_t doubles = part_numel * (dtype == YH_C128 ? 2 : 1);
  YH_CHECK_ARG(doubles <= kScratchDoubles, "split-K partials of %lld bytes exceed the %lld-byte scratch",
               (long long)(doubles * 8), (long long)(kScratchDoubles * 8));
  void* part = nullptr;
  int st = scratch_address(&part);
  if (st) return st;
  void* o[YH_MAX_BUFS];
  for (int i = 0; i < YH_MAX_BUFS; ++i) o[i] = outs[i];
  o[part_buf] = part;
  st = yastn_hip_grouped_gemm(o, ins, dtype, d_gemms, ngemms, d_terms, nterms, d_tiles, ntiles, cfg, stream);
  if (st || nred == 0) return st;
  return yastn_hip_splitk_reduce(o, part, dtype, d_red, nred, stream);
}
