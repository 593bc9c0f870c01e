// Batched one-sided Jacobi SVD (Hestenes) and Hermitian eigendecomposition, FP64 / complex128.
//
// One workgroup per block.  The block is copied into a column-major working
// matrix W (p x q, p = max(m,n), q = min(m,n); W = A, or A^H when m < n) and a
// q x q rotation accumulator V = I.  Each sweep visits all column pairs in the
// round-robin (circle) order, so the q/2 pairs of a round are disjoint and the
// 16 waves work on different pairs concurrently; a pair is rotated when
// |w_i^H w_j| > tol * ||w_i|| ||w_j||.  Converged when a sweep rotates nothing.
// Then sigma_j = ||w_j||, U = W Sigma^-1 (or V for m < n), sorted descending.
//
// Hermitian eigh reuses the same iteration on B = A + mu I, mu = max row
// abs-sum (Gershgorin, so B is positive semidefinite): the right rotations V are
// the eigenvectors and lambda_j = sigma_j - mu, written ascending.
//
// Blocks whose W and V fit in YH_SVD_LDS_BYTES go through an LDS-resident instance of the
// kernel (all sweeps in one launch); larger blocks use the panel path below.
#include <float.h>

#include <algorithm>

#include "common.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// Zero elements (never written), read in place of elements outside a matrix so that loads need no
// per-lane branch (a load under such a branch gets a wait of its own, which serialises the loads); a
// thread reads element (its index mod kZeroSpan), so a wave's fallback loads spread over cache lines.
constexpr int kZeroSpan = 256;
__device__ double2 kZeroElt[kZeroSpan];

// 16 waves per block: the q/2 disjoint pairs of a round are spread over more waves
constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;

template <bool CPLX>
struct Num;
template <>
struct Num<false> {
  typedef double T;
  static __device__ inline double re(double x) { return x; }
  static __device__ inline double im(double) { return 0.0; }
  static __device__ inline double make(double r, double) { return r; }
  static __device__ inline double conj(double x) { return x; }
  static __device__ inline double abs2(double x) { return x * x; }
};
template <>
struct Num<true> {
  typedef double2 T;
  static __device__ inline double re(double2 x) { return x.x; }
  static __device__ inline double im(double2 x) { return x.y; }
  static __device__ inline double2 make(double r, double i) { return make_double2(r, i); }
  static __device__ inline double2 conj(double2 x) { return make_double2(x.x, -x.y); }
  static __device__ inline double abs2(double2 x) { return x.x * x.x + x.y * x.y; }
};

// Wave-wide reductions: DPP moves within each row of 16 lanes (quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror: after the four steps every lane holds its row's value),
// then the four row values are read into scalar registers, so the result is wave-uniform.
template <int CTRL>
__device__ __attribute__((always_inline)) inline double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __attribute__((always_inline)) inline double read_lane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

__device__ __attribute__((always_inline)) inline double wave_sum(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

__device__ __attribute__((always_inline)) inline double wave_max(double v) {
  v = fmax(v, dpp_move<0xB1>(v));
  v = fmax(v, dpp_move<0x4E>(v));
  v = fmax(v, dpp_move<0x141>(v));
  v = fmax(v, dpp_move<0x140>(v));
  return fmax(fmax(read_lane(v, 0), read_lane(v, 16)), fmax(read_lane(v, 32), read_lane(v, 48)));
}

// circle-method pairing: qe (even) players, round r in [0, qe-1), pair k in [0, qe/2)
__device__ inline void pair_of(int qe, int r, int k, int& i, int& j) {
  const int n1 = qe - 1;
  if (k == 0) {
    i = r;
    j = n1;
  } else {
    i = (r + k) % n1;
    j = (r - k + n1) % n1;
  }
  if (i > j) {
    const int t = i;
    i = j;
    j = t;
  }
}

// 1/x and 1/sqrt(x) from the hardware approximations (v_rcp_f64, v_rsq_f64) refined by two
// Newton-Raphson steps each; the rotation parameters only need c^2 + s^2 = 1 to rounding accuracy.
__device__ __attribute__((always_inline)) inline double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

__device__ __attribute__((always_inline)) inline double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// Hestenes rotation of a column pair from a = ||w_i||^2, b = ||w_j||^2, g = w_i^H w_j = gr + i gi:
// returns false (no rotation) when |g| <= tol sqrt(a b) or a column is at the noise floor; else
// c, s, and the phase e^{i phi} = g / |g| = (er, ei).
__device__ __attribute__((always_inline)) inline bool rotation(double a, double b, double gr, double gi, double tol, double floor2, double& c,
                                double& s, double& er, double& ei) {
  const double g2 = gr * gr + gi * gi;
  // columns at the noise level (||w||^2 <= (eps ||A||_F)^2) carry no information beyond absolute
  // accuracy eps ||A||; rotating them against rounding noise would never converge
  if (!(g2 > tol * tol * a * b) || a <= floor2 || b <= floor2) return false;
  const double rg = rsq_nr(g2);
  er = gr * rg;
  ei = gi * rg;
  const double zeta = 0.5 * (b - a) * rg;  // (b - a) / (2 |g|)
  const double az = fabs(zeta);
  double t;
  if (az < 1e150) {
    const double x = 1.0 + zeta * zeta;
    t = rcp_nr(az + x * rsq_nr(x));
  } else {
    t = 0.5 / az;
  }
  if (zeta < 0) t = -t;
  c = rsq_nr(1.0 + t * t);
  s = c * t;
  return true;
}

// (u, v) <- (c u - s e^{-i phi} v, s e^{i phi} u + c v); for real data e^{i phi} = +-1 is folded into s.
template <bool CPLX>
__device__ __attribute__((always_inline)) inline void apply_rot(typename Num<CPLX>::T& u, typename Num<CPLX>::T& v, double c, double s, double er,
                                 double ei) {
  typedef Num<CPLX> N;
  if constexpr (CPLX) {
    const double vr = er * N::re(v) + ei * N::im(v), vi = er * N::im(v) - ei * N::re(v);
    const double ur = er * N::re(u) - ei * N::im(u), ui = er * N::im(u) + ei * N::re(u);
    const typename N::T u2 = N::make(c * N::re(u) - s * vr, c * N::im(u) - s * vi);
    v = N::make(s * ur + c * N::re(v), s * ui + c * N::im(v));
    u = u2;
  } else {
    const double se = s * er;
    const double u2 = c * u - se * v;
    v = se * u + c * v;
    u = u2;
  }
}

// One Hestenes rotation of the column pair (w_i, w_j) (length p) and of the matching accumulator
// columns (v_i, v_j) (length q), by one wave.  Returns true when the pair was rotated, i.e. when
// |w_i^H w_j| > tol ||w_i|| ||w_j|| and both squared norms exceed floor2.  J = [[c, s e^{i phi}], [-s e^{-i phi}, c]], e^{i phi} = g / |g|.
template <bool CPLX>
__device__ __attribute__((always_inline)) inline bool jacobi_pair(typename Num<CPLX>::T* wi, typename Num<CPLX>::T* wj, int64_t p,
                                   typename Num<CPLX>::T* vi_, typename Num<CPLX>::T* vj_, int64_t q, double tol,
                                   double floor2, int lane) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
  for (int64_t x = lane; x < p; x += 64) {
    const T u = wi[x], v = wj[x];
    a += N::abs2(u);
    b += N::abs2(v);
    gr += N::re(u) * N::re(v) + N::im(u) * N::im(v);  // conj(u) v
    gi += N::re(u) * N::im(v) - N::im(u) * N::re(v);
  }
  a = wave_sum(a);
  b = wave_sum(b);
  gr = wave_sum(gr);
  if (CPLX) gi = wave_sum(gi);
  double c, s, er, ei;
  if (!rotation(a, b, gr, gi, tol, floor2, c, s, er, ei)) return false;
  // w_i' = c w_i - s e^{-i phi} w_j ; w_j' = s e^{i phi} w_i + c w_j
  for (int64_t x = lane; x < p; x += 64) {
    T u = wi[x], v = wj[x];
    apply_rot<CPLX>(u, v, c, s, er, ei);
    wi[x] = u;
    wj[x] = v;
  }
  for (int64_t x = lane; x < q; x += 64) {
    T u = vi_[x], v = vj_[x];
    apply_rot<CPLX>(u, v, c, s, er, ei);
    vi_[x] = u;
    vj_[x] = v;
  }
  return true;
}

// The same rotation with the two columns held in registers between the inner products and the
// update (one read and one write per element instead of two reads and one write); p <= 64 * NR.
template <bool CPLX, int NR>
__device__ __attribute__((always_inline)) inline bool jacobi_pair_reg(typename Num<CPLX>::T* wi, typename Num<CPLX>::T* wj, int64_t p,
                                       typename Num<CPLX>::T* vi_, typename Num<CPLX>::T* vj_, int64_t q,
                                       double tol, double floor2, int lane) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  T xi[NR], xj[NR];
  double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int64_t x = lane + 64 * r;
    if (x < p) {
      xi[r] = wi[x];
      xj[r] = wj[x];
    } else {
      xi[r] = N::make(0.0, 0.0);
      xj[r] = N::make(0.0, 0.0);
    }
    const T u = xi[r], v = xj[r];
    a += N::abs2(u);
    b += N::abs2(v);
    gr += N::re(u) * N::re(v) + N::im(u) * N::im(v);
    gi += N::re(u) * N::im(v) - N::im(u) * N::re(v);
  }
  a = wave_sum(a);
  b = wave_sum(b);
  gr = wave_sum(gr);
  if (CPLX) gi = wave_sum(gi);
  double c, s, er, ei;
  if (!rotation(a, b, gr, gi, tol, floor2, c, s, er, ei)) return false;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int64_t x = lane + 64 * r;
    if (x < p) {
      T u = xi[r], v = xj[r];
      apply_rot<CPLX>(u, v, c, s, er, ei);
      wi[x] = u;
      wj[x] = v;
    }
  }
  for (int64_t x = lane; x < q; x += 64) {
    T u = vi_[x], v = vj_[x];
    apply_rot<CPLX>(u, v, c, s, er, ei);
    vi_[x] = u;
    vj_[x] = v;
  }
  return true;
}

// Rotation of one pair with the register-resident variant when the columns are short enough.
template <bool CPLX>
__device__ __attribute__((always_inline)) inline bool rotate_pair(typename Num<CPLX>::T* wi, typename Num<CPLX>::T* wj, int64_t p,
                                   typename Num<CPLX>::T* vi, typename Num<CPLX>::T* vj, int64_t q, double tol,
                                   double floor2, int lane) {
  if (p <= 64) return jacobi_pair_reg<CPLX, 1>(wi, wj, p, vi, vj, q, tol, floor2, lane);
  if (p <= 128) return jacobi_pair_reg<CPLX, 2>(wi, wj, p, vi, vj, q, tol, floor2, lane);
  if (p <= 256) return jacobi_pair_reg<CPLX, 4>(wi, wj, p, vi, vj, q, tol, floor2, lane);
  if (p <= 512) return jacobi_pair_reg<CPLX, 8>(wi, wj, p, vi, vj, q, tol, floor2, lane);
  return jacobi_pair<CPLX>(wi, wj, p, vi, vj, q, tol, floor2, lane);
}

struct BlockDesc {
  int64_t a_off, m, n, u_off, s_off, v_off, work_off;
};

__device__ inline BlockDesc load_desc(const int64_t* w) {
  BlockDesc d;
  d.a_off = w[0];
  d.m = w[1];
  d.n = w[2];
  d.u_off = w[3];
  d.s_off = w[4];
  d.v_off = w[5];
  d.work_off = w[6];
  return d;
}

// Column norms -> singular values (descending, ties by index), vectors written by whole waves.
// sigma_j is staged in the block's S slice; the ranks are kept in LDS.
// MODE 2 is the QR-preconditioned SVD: W = X (q x q, p == q here), V = V_J, and pf is the long
// dimension of the block; the q-side vectors x_j / sigma_j go to Vh (m >= n) or U (m < n) and the
// p-side vectors start as [v_j; 0] in the other output, to be multiplied by Q afterwards.
template <bool CPLX, int MODE>
__device__ void finalize_block(const BlockDesc& d, const typename Num<CPLX>::T* W, const typename Num<CPLX>::T* V,
                               int64_t p, int64_t q, bool tr, double mu, typename Num<CPLX>::T* __restrict__ U,
                               double* __restrict__ S, typename Num<CPLX>::T* __restrict__ Vh, int64_t pf = 0) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  __shared__ int s_rank_of[YH_SVD_MAX_Q];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // ---- finalize: sigma_j = ||w_j|| staged in the block's S slice, ranks by counting (descending,
  // ties by index), then vectors written by whole waves, then S in sorted order.
  for (int64_t j = wave; j < q; j += kWaves) {
    const T* wj = W + j * p;
    double a = 0.0;
    for (int64_t x = lane; x < p; x += 64) a += N::abs2(wj[x]);
    a = wave_sum(a);
    if (lane == 0) S[d.s_off + j] = sqrt(a);
  }
  __syncthreads();
  constexpr int kMaxCols = YH_SVD_MAX_Q / kThreads;
  int rk[kMaxCols];
  double sg[kMaxCols];
#pragma unroll
  for (int z = 0; z < kMaxCols; ++z) {
    const int64_t j = tid + (int64_t)z * kThreads;
    rk[z] = -1;
    sg[z] = 0.0;
    if (j < q) {
      const double sj = S[d.s_off + j];
      int cnt = 0;
      for (int64_t i = 0; i < q; ++i) {
        const double si = S[d.s_off + i];
        cnt += (si > sj) || (si == sj && i < j);
      }
      rk[z] = cnt;
      sg[z] = sj;
      s_rank_of[j] = cnt;
    }
  }
  __syncthreads();
  for (int64_t j = wave; j < q; j += kWaves) {
    const int64_t rkj = s_rank_of[j];
    const double sj = S[d.s_off + j];
    const T* wj = W + j * p;
    const T* vj = V + j * q;
    if (MODE == 1) {
      const int64_t pos = q - 1 - rkj;
      for (int64_t x = lane; x < q; x += 64) U[d.u_off + x * q + pos] = vj[x];
    } else if (MODE == 2) {
      const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
      if (!tr) {  // W0 = A = (Q V_J) Sigma U_J^H: Vh (q x q) row rkj = u_j^H, U (pf x q) column rkj = [v_j; 0]
        for (int64_t x = lane; x < q; x += 64) {
          const T w = wj[x];
          Vh[d.v_off + rkj * q + x] = N::make(N::re(w) * inv, -N::im(w) * inv);
        }
        for (int64_t x = lane; x < pf; x += 64) U[d.u_off + x * q + rkj] = x < q ? vj[x] : N::make(0.0, 0.0);
      } else {  // W0 = A^H: U (q x q) column rkj = u_j, Vh (q x pf) row rkj = [v_j; 0]^H
        for (int64_t x = lane; x < q; x += 64) {
          const T w = wj[x];
          U[d.u_off + x * q + rkj] = N::make(N::re(w) * inv, N::im(w) * inv);
        }
        for (int64_t x = lane; x < pf; x += 64) Vh[d.v_off + rkj * pf + x] = x < q ? N::conj(vj[x]) : N::make(0.0, 0.0);
      }
    } else if (!tr) {
      // A V = W: U (m x k) column rkj = w_j / sigma_j, Vh (k x n) row rkj = v_j^H
      const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
      for (int64_t x = lane; x < p; x += 64) {
        const T w = wj[x];
        U[d.u_off + x * q + rkj] = N::make(N::re(w) * inv, N::im(w) * inv);
      }
      for (int64_t x = lane; x < q; x += 64) Vh[d.v_off + rkj * q + x] = N::conj(vj[x]);
    } else {
      // A^H V = W, so A = V Sigma (W Sigma^-1)^H: U (m x m) column rkj = v_j, Vh (m x n) row rkj = w_j^H / sigma_j
      const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
      for (int64_t x = lane; x < q; x += 64) U[d.u_off + x * q + rkj] = vj[x];
      for (int64_t x = lane; x < p; x += 64) {
        const T w = wj[x];
        Vh[d.v_off + rkj * p + x] = N::make(N::re(w) * inv, -N::im(w) * inv);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int z = 0; z < kMaxCols; ++z) {
    if (rk[z] < 0) continue;
    if (MODE != 1)
      S[d.s_off + rk[z]] = sg[z];
    else
      S[d.s_off + (q - 1 - rk[z])] = sg[z] - mu;
  }
}

// MODE 0: SVD (U, S, Vh).  MODE 1: Hermitian eigh (S ascending, U).
template <bool CPLX, bool IN_LDS, int MODE>
__global__ void __launch_bounds__(kThreads)
    jacobi_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ order,
                  const typename Num<CPLX>::T* __restrict__ A, typename Num<CPLX>::T* __restrict__ U,
                  double* __restrict__ S, typename Num<CPLX>::T* __restrict__ Vh,
                  typename Num<CPLX>::T* __restrict__ work, int32_t* __restrict__ info, int max_sweeps) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  extern __shared__ double smem_raw[];
  __shared__ int s_rotated;
  __shared__ double s_red[kWaves];

  const int bidx = order[blockIdx.x];
  const BlockDesc d = load_desc(blocks + (int64_t)bidx * YH_SVD_WORDS);
  const bool tr = (MODE == 0) && (d.m < d.n);
  const int64_t p = tr ? d.n : d.m;
  const int64_t q = tr ? d.m : d.n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  T* W = IN_LDS ? reinterpret_cast<T*>(smem_raw) : work + d.work_off;
  T* V = W + p * q;

  // ---- load W (column-major p x q) and V = I
  for (int64_t e = tid; e < p * q; e += kThreads) {
    const int64_t c = e / p, r = e - c * p;
    T v = tr ? N::conj(A[d.a_off + c * d.n + r]) : A[d.a_off + r * d.n + c];
    W[e] = v;
  }
  for (int64_t e = tid; e < q * q; e += kThreads) {
    const int64_t c = e / q, r = e - c * q;
    V[e] = N::make(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();

  double mu = 0.0;
  if (MODE == 1) {
    // Gershgorin bound of the Hermitian block: max_r sum_c |a_rc|
    double local = 0.0;
    for (int64_t r = tid; r < p; r += kThreads) {
      double acc = 0.0;
      for (int64_t c = 0; c < q; ++c) acc += sqrt(N::abs2(W[r + c * p]));
      local = fmax(local, acc);
    }
    local = wave_max(local);
    if (lane == 0) s_red[wave] = local;
    __syncthreads();
    mu = s_red[0];
    for (int w = 1; w < kWaves; ++w) mu = fmax(mu, s_red[w]);
    for (int64_t r = tid; r < p; r += kThreads) {
      T x = W[r + r * p];
      W[r + r * p] = N::make(N::re(x) + mu, N::im(x));
    }
    __syncthreads();
  }

  const double tol = DBL_EPSILON * sqrt((double)(p > 1 ? p : 1));
  double floor2;
  {
    double local = 0.0;
    for (int64_t e = tid; e < p * q; e += kThreads) local += N::abs2(W[e]);
    local = wave_sum(local);
    __syncthreads();
    if (lane == 0) s_red[wave] = local;
    __syncthreads();
    double fro2 = 0.0;
    for (int w = 0; w < kWaves; ++w) fro2 += s_red[w];
    floor2 = DBL_EPSILON * DBL_EPSILON * fro2;
  }
  const int qe = (int)(q + (q & 1));
  int sweeps = 0;
  bool converged = (q <= 1);
  while (!converged && sweeps < max_sweeps) {
    if (tid == 0) s_rotated = 0;
    __syncthreads();
    for (int r = 0; r < qe - 1; ++r) {
      for (int k = wave; k < qe / 2; k += kWaves) {
        int i, j;
        pair_of(qe, r, k, i, j);
        if (j >= q) continue;
        if (!rotate_pair<CPLX>(W + (int64_t)i * p, W + (int64_t)j * p, p, V + (int64_t)i * q, V + (int64_t)j * q, q,
                               tol, floor2, lane))
          continue;
        if (lane == 0) s_rotated = 1;
      }
      __syncthreads();
    }
    ++sweeps;
    converged = (s_rotated == 0);
    __syncthreads();
  }

  finalize_block<CPLX, MODE>(d, W, V, p, q, tr, mu, U, S, Vh);
  if (tid == 0) info[bidx] = converged ? sweeps : -1;
}

// ---------------------------------------------------------------------------------------------
// Panel (block) Jacobi for blocks whose working matrices do not fit one CU.  The q columns are cut
// into panels of width b; an outer sweep visits all panel pairs in the circle order, one launch per
// round with one workgroup per (block, panel pair).  A workgroup stages the p x (2b) panel pair in
// LDS, applies one inner sweep of Hestenes rotations to it (accumulating them in a 2b x 2b Q), writes
// it back and updates the matching V columns, V <- V Q.  Pairs of different panels of the same
// block in one round are disjoint, so all of them run concurrently.
constexpr int kPanelThreads = 1024;

template <bool CPLX, int MODE>
__global__ void __launch_bounds__(kPanelThreads)
    panel_init_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ ids,
                      const typename Num<CPLX>::T* __restrict__ A, typename Num<CPLX>::T* __restrict__ work,
                      double* __restrict__ mu_out, double* __restrict__ floor_out) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  __shared__ double s_red[kPanelThreads / 64];
  const int bidx = ids[blockIdx.x];
  const BlockDesc d = load_desc(blocks + (int64_t)bidx * YH_SVD_WORDS);
  const bool tr = (MODE == 0) && (d.m < d.n);
  const int64_t p = tr ? d.n : d.m, q = tr ? d.m : d.n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T* W = work + d.work_off;
  T* V = W + p * q;
  for (int64_t e = tid; e < p * q; e += kPanelThreads) {
    const int64_t c = e / p, r = e - c * p;
    W[e] = tr ? N::conj(A[d.a_off + c * d.n + r]) : A[d.a_off + r * d.n + c];
  }
  for (int64_t e = tid; e < q * q; e += kPanelThreads) {
    const int64_t c = e / q, r = e - c * q;
    V[e] = N::make(r == c ? 1.0 : 0.0, 0.0);
  }
  double mu = 0.0;
  if (MODE == 1) {
    __syncthreads();
    double local = 0.0;
    for (int64_t r = tid; r < p; r += kPanelThreads) {
      double acc = 0.0;
      for (int64_t c = 0; c < q; ++c) acc += sqrt(N::abs2(W[r + c * p]));
      local = fmax(local, acc);
    }
    local = wave_max(local);
    if (lane == 0) s_red[wave] = local;
    __syncthreads();
    mu = s_red[0];
    for (int w = 1; w < kPanelThreads / 64; ++w) mu = fmax(mu, s_red[w]);
    for (int64_t r = tid; r < p; r += kPanelThreads) {
      const T x = W[r + r * p];
      W[r + r * p] = N::make(N::re(x) + mu, N::im(x));
    }
  }
  __syncthreads();
  double local = 0.0;
  for (int64_t e = tid; e < p * q; e += kPanelThreads) local += N::abs2(W[e]);
  local = wave_sum(local);
  __syncthreads();
  if (lane == 0) s_red[wave] = local;
  __syncthreads();
  if (tid == 0) {
    double fro2 = 0.0;
    for (int w = 0; w < kPanelThreads / 64; ++w) fro2 += s_red[w];
    mu_out[bidx] = mu;
    floor_out[bidx] = DBL_EPSILON * DBL_EPSILON * fro2;
  }
}

// One 16-row stripe of X (rows x nc, column c at col(c)) times Q (nc x nc, column c at Q + c nq) on
// the FP64 matrix cores, in two steps: stripe_load reads the stripe into registers (lane l holds
// A(l % 16, 4 s + l / 16), s = 0 .. nc / 4 - 1), stripe_mul_store multiplies (up to 4 output tiles of
// 16 columns, computed as Q^T X^T so that item v of a tile is D(l % 16, l / 16 + 4 v)) and writes the
// stripe back, so an update in
// place never overwrites what the wave has not read, and the loads can be issued before Q is ready.
// nc <= 64.
// KS k-steps of 4: nc <= 4 KS.
template <bool CPLX, int KS, typename COL>
__device__ __attribute__((always_inline)) inline void stripe_load(typename Num<CPLX>::T (&a)[KS], COL col, int rows,
                                                                  int st, int nc, int lane) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  const int ra = st * 16 + (lane & 15);
#pragma unroll
  for (int sidx = 0; sidx < KS; ++sidx) {
    const int k = 4 * sidx + (lane >> 4);
    a[sidx] = *(((k < nc) & (ra < rows)) ? col(k) + ra : (const T*)kZeroElt + lane);
  }
}

template <bool CPLX, int KS, typename COL>
__device__ __attribute__((always_inline)) inline void stripe_mul_store(const typename Num<CPLX>::T (&a)[KS], COL col,
                                                                       int rows, int st,
                                                                       const typename Num<CPLX>::T* Q, int nq,
                                                                       int nc, int lane) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  const int nk4 = (nc + 3) / 4, tcn = (nc + 15) / 16;
#pragma unroll
  for (int tj = 0; tj < KS / 4; ++tj) {
    if (tj < tcn) {
      const int cb = tj * 16 + (lane & 15);
      d4 accR = {0.0, 0.0, 0.0, 0.0}, accI = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int sidx = 0; sidx < KS; ++sidx) {
        if (sidx < nk4) {
          const int k = 4 * sidx + (lane >> 4);
          const T bq = (cb < nc && k < nc) ? Q[cb * nq + k] : N::make(0.0, 0.0);
          // transposed product (Q^T X^T): the result tile is D^T, so a lane's items run down a column
          // of X and each store covers 16 consecutive rows of 4 columns
          accR = __builtin_amdgcn_mfma_f64_16x16x4f64(N::re(bq), N::re(a[sidx]), accR, 0, 0, 0);
          if constexpr (CPLX) {
            accR = __builtin_amdgcn_mfma_f64_16x16x4f64(-N::im(bq), N::im(a[sidx]), accR, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f64_16x16x4f64(N::im(bq), N::re(a[sidx]), accI, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f64_16x16x4f64(N::re(bq), N::im(a[sidx]), accI, 0, 0, 0);
          }
        }
      }
      const int row = st * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int c = tj * 16 + (lane >> 4) + 4 * v;  // item v: D^T(l / 16 + 4 v, l % 16) = D(row, c)
        if (row < rows && c < nc) col(c)[row] = N::make(accR[v], accI[v]);
      }
    }
  }
}

// A round of the panel Jacobi runs two kernels over the (block, panel i, panel j) jobs of the round:
//  1. panel_gram_kernel: the Gram matrix G = Wp^H Wp (nc x nc) of the pair's nc = |I| + |J| columns
//     Wp, as partial sums over chunks of kGramRows rows, one workgroup per (chunk, job);
//  2. panel_inner_kernel: per job, `slices` workgroups each sum the partials and run the same inner
//     sweep of rotations on G (the same quantities and criterion as Hestenes rotations of the
//     columns; identical arithmetic, so identical results), accumulating the rotation Q (nc x nc), and
//     then update their slice of the pair's rows of W and V by Q in place.
// The Gram and update work, products over all p (and q) rows, is spread over many CUs; the inner
// sweep is repeated per slice instead of being followed by another launch.  A pair that did not
// rotate writes nothing.
constexpr int kGramRows = 64;      // rows of W per workgroup of kernel 1
constexpr int kChunkRows = 64;     // unit of the row-chunk count maxch_p passed by the host
constexpr int kChunkThreads = 256;  // 4 waves
constexpr int kInnerThreads = 1024;  // inner sweep: one thread per 2 x 2 block of G or (pair, row) of Q
// Iteration state (int32 words, zero before sweep 0): 1 plan too small for a panel pair, 3 result (sweeps
// used, -1 not converged, -3 plan too small), kStateHead + s: some pair rotated in sweep s.
constexpr int kStateHead = 4;

// The previous sweep rotated nothing (the iteration has converged): a round of this sweep has nothing
// to do.  Launches of later sweeps can thus be issued without a host read in between.
__device__ inline bool sweep_done(const int32_t* state, int sweep) {
  return state != nullptr && sweep > 0 &&
         __hip_atomic_load(&state[kStateHead + sweep - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
}

struct PanelPair {
  int64_t p, q, ci, cj, c0i, c0j;
  int nc;
  bool intra;
  __device__ int64_t col(int64_t l) const { return l < ci ? c0i + l : c0j + (l - ci); }
};

__device__ inline PanelPair panel_pair(const BlockDesc& d, const int32_t* jb, int b) {
  PanelPair pp;
  const bool tr = d.m < d.n;
  pp.p = tr ? d.n : d.m;
  pp.q = tr ? d.m : d.n;
  pp.intra = jb[1] == jb[2];  // one panel against itself, else two panels (cross pairs only)
  pp.c0i = (int64_t)jb[1] * b;
  pp.c0j = (int64_t)jb[2] * b;
  pp.ci = min((int64_t)b, pp.q - pp.c0i);
  pp.cj = pp.intra ? 0 : min((int64_t)b, pp.q - pp.c0j);
  pp.nc = (int)(pp.ci + pp.cj);
  return pp;
}

// Upper-triangle 16 x 16 tile tl of a T16 x T16 tile grid -> (ti, tj), ti <= tj.
__device__ inline void tile_of(int tl, int T16, int& ti, int& tj) {
  int rem = tl;
  ti = 0;
  while (rem >= T16 - ti) {
    rem -= T16 - ti;
    ++ti;
  }
  tj = ti + rem;
}

// Partial Gram tiles of one (row chunk, job): the chunk's nc columns are staged in LDS (odd column
// stride, so the 16 columns of a tile fall in different banks); each wave forms upper-triangle tiles
// on the matrix cores (lane l supplies A(l % 16, k) = conj(Wp[k, ti 16 + l % 16]) and B(k, l % 16) =
// Wp[k, tj 16 + l % 16], k = k0 + l / 16) and stores them in the MFMA result order (item v at v 64 + l).
// NST = 2 b kGramRows / kChunkThreads staged elements per thread (8 for b <= 16, 16 for b <= 32).
template <bool CPLX, int NST>
__global__ void __launch_bounds__(kChunkThreads)
    panel_gram_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ jobs,
                      const typename Num<CPLX>::T* __restrict__ work, double* __restrict__ part, int b, int maxch,
                      const int32_t* __restrict__ state, int sweep) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  extern __shared__ double smem_raw[];
  constexpr int ld = kGramRows + 1;
  if (sweep_done(state, sweep)) return;
  const int32_t* jb = jobs + 3 * (int64_t)blockIdx.y;
  const BlockDesc d = load_desc(blocks + (int64_t)jb[0] * YH_SVD_WORDS);
  const PanelPair pp = panel_pair(d, jb, b);
  const int64_t r0 = (int64_t)blockIdx.x * kGramRows;
  if (r0 >= pp.p) return;  // uniform: beyond this block's rows
  const int rows = (int)min((int64_t)kGramRows, pp.p - r0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const T* W = work + d.work_off;
  T* S = reinterpret_cast<T*>(smem_raw);
  {
    // all NST loads of a thread are issued before its first LDS store
    // (an element outside the chunk reads W[0] and is stored as zero)
    T v[NST];
#pragma unroll
    for (int i = 0; i < NST; ++i) {
      const int e = tid + i * kChunkThreads;
      const int l = e / kGramRows, r = e - l * kGramRows;
      const bool ok = (l < pp.nc) & (r < rows);
      v[i] = W[ok ? pp.col(l) * pp.p + r0 + r : 0];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the stores (no interleaving)
#pragma unroll
    for (int i = 0; i < NST; ++i) {
      const int e = tid + i * kChunkThreads;
      const int l = e / kGramRows, r = e - l * kGramRows;
      if (l < pp.nc) S[l * ld + r] = r < rows ? v[i] : N::make(0.0, 0.0);
    }
  }
  __syncthreads();
  const int T16 = (pp.nc + 15) / 16, ntile = T16 * (T16 + 1) / 2;
  const int T16m = (2 * b + 15) / 16, ntmax = T16m * (T16m + 1) / 2;
  for (int tl = wave; tl < ntile; tl += kChunkThreads / 64) {
    int ti, tj;
    tile_of(tl, T16, ti, tj);
    const int ca = ti * 16 + (lane & 15), cb = tj * 16 + (lane & 15);
    d4 accR = {0.0, 0.0, 0.0, 0.0}, accI = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < kGramRows; k0 += 4) {
      const int k = k0 + (lane >> 4);
      const T a = ca < pp.nc ? S[ca * ld + k] : N::make(0.0, 0.0);
      const T bb = cb < pp.nc ? S[cb * ld + k] : N::make(0.0, 0.0);
      accR = __builtin_amdgcn_mfma_f64_16x16x4f64(N::re(a), N::re(bb), accR, 0, 0, 0);
      if constexpr (CPLX) {  // conj(a) b
        accR = __builtin_amdgcn_mfma_f64_16x16x4f64(N::im(a), N::im(bb), accR, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f64_16x16x4f64(N::re(a), N::im(bb), accI, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f64_16x16x4f64(-N::im(a), N::re(bb), accI, 0, 0, 0);
      }
    }
    double* dst = part + (((int64_t)blockIdx.y * maxch + blockIdx.x) * ntmax + tl) * 256 * (CPLX ? 2 : 1);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      dst[v * 64 + lane] = accR[v];
      if constexpr (CPLX) dst[256 + v * 64 + lane] = accI[v];
    }
  }
}

// Inner sweep of one job on its Gram matrix (LDS: G and the rotation accumulator Q, nc x (2b + 1)
// each), then the update of slice blockIdx.y of the pair's rows of W and V.  Sets the sweep's rotation
// word of the state if anything rotated, and word 1 if lds_cap is too small for the pair.  (A variant
// with one barrier per sub-round, every thread computing its slots' rotations itself from a
// double-buffered G, ran the D=6 chi=64 CTMRG step slower: 96 against 91 ms.)
template <bool CPLX, int KS>
__global__ void __launch_bounds__(kInnerThreads)
    panel_inner_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ jobs,
                       typename Num<CPLX>::T* __restrict__ work, const double* __restrict__ floors,
                       const double* __restrict__ part, int32_t* __restrict__ state, int sweep, int b,
                       int maxch, int64_t lds_cap) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  constexpr int kMaxPairs = 32;  // nc <= 64
  extern __shared__ double smem_raw[];
  __shared__ int s_rot;
  __shared__ int4 s_ij[kMaxPairs];
  __shared__ double2 s_cs[kMaxPairs];
  __shared__ double s_si[kMaxPairs];
  if (sweep_done(state, sweep)) return;
  const int32_t* jb = jobs + 3 * (int64_t)blockIdx.x;
  const BlockDesc d = load_desc(blocks + (int64_t)jb[0] * YH_SVD_WORDS);
  const PanelPair pp = panel_pair(d, jb, b);
  const int64_t p = pp.p, ci = pp.ci, cj = pp.cj;
  const int nc = pp.nc;
  const bool intra = pp.intra;
  const int tid = threadIdx.x;
  const int nq = 2 * b + 1;  // leading dimension of G and Q (odd: column reads spread over banks)
  T* G = reinterpret_cast<T*>(smem_raw);
  T* Q = G + (int64_t)(2 * b) * nq;
  const int64_t need = (int64_t)2 * (2 * b) * nq * (int64_t)sizeof(T);
  if (need > lds_cap || nc > 2 * kMaxPairs || nc > 4 * KS) {
    if (tid == 0 && blockIdx.y == 0) state[1] = 1;  // uniform: plan too small for this panel pair
    return;
  }
  if (tid == 0) s_rot = 0;
  // ---- this workgroup's share of the update: the 16-row stripes of W (p rows) then V (q rows), cut
  // into gridDim.y equal slices, one stripe per wave at a time.  Each wave's first stripe is read now,
  // so that its loads overlap the inner sweep (nothing writes W or V of this pair before the update).
  const int lane = tid & 63, wave = tid >> 6;
  T* const Wd = work + d.work_off;
  const int sw = (int)((p + 15) / 16), sv = (int)((pp.q + 15) / 16);
  const int per = (sw + sv + (int)gridDim.y - 1) / (int)gridDim.y;
  const int s0 = (int)blockIdx.y * per, s1 = min(sw + sv, s0 + per);
  auto stripe_col = [&](int st) {  // column c of stripe st's matrix (W, or V after sw stripes)
    const bool isv = st >= sw;
    T* X = isv ? Wd + p * pp.q : Wd;
    const int64_t ldx = isv ? pp.q : p;
    return [=](int c) { return X + pp.col(c) * ldx; };
  };
  auto stripe_rows = [&](int st) { return st >= sw ? (int)pp.q : (int)p; };
  auto stripe_idx = [&](int st) { return st >= sw ? st - sw : st; };
  T a[KS];  // KS k-steps of 4 cover the pair's nc <= 4 KS columns
  const int st0 = s0 + wave;
  if (st0 < s1) stripe_load<CPLX, KS>(a, stripe_col(st0), stripe_rows(st0), stripe_idx(st0), nc, lane);
  // ---- G: sum of the chunks' partial tiles, mirrored to the lower triangle
  {
    const int T16 = (nc + 15) / 16, ntile = T16 * (T16 + 1) / 2;
    const int T16m = (2 * b + 15) / 16, ntmax = T16m * (T16m + 1) / 2;
    const int nch = (int)((p + kGramRows - 1) / kGramRows);
    for (int e = tid; e < ntile * 256; e += kInnerThreads) {
      const int tl = e >> 8, idx = e & 255;
      const int v = idx >> 6, ln = idx & 63;
      int ti, tj;
      tile_of(tl, T16, ti, tj);
      const int i = ti * 16 + (ln >> 4) + 4 * v, j = tj * 16 + (ln & 15);  // D(l / 16 + 4 v, l % 16)
      double gr = 0.0, gi = 0.0;
      // chunks in groups of kSumGroup with all loads of a group in flight; summed in chunk order
      constexpr int kSumGroup = 8;
      const double* zero = (const double*)kZeroElt + (idx & (kZeroSpan - 1));
      for (int c0 = 0; c0 < nch; c0 += kSumGroup) {
        double vr[kSumGroup], vi[kSumGroup];
#pragma unroll
        for (int g = 0; g < kSumGroup; ++g) {
          const int ch = c0 + g;
          const double* src = part + (((int64_t)blockIdx.x * maxch + ch) * ntmax + tl) * 256 * (CPLX ? 2 : 1);
          vr[g] = *(ch < nch ? src + idx : zero);
          if constexpr (CPLX) vi[g] = *(ch < nch ? src + 256 + idx : zero);
        }
#pragma unroll
        for (int g = 0; g < kSumGroup; ++g) {
          gr += vr[g];  // + 0.0 past the last chunk leaves the sum unchanged (it is never -0.0)
          if constexpr (CPLX) gi += vi[g];
        }
      }
      if (i < nc && j < nc) {
        G[j * nq + i] = N::make(gr, gi);  // G(i, j)
        if (ti != tj) G[i * nq + j] = N::make(gr, -gi);  // G(j, i) of the mirrored tile
      }
    }
  }
  for (int e = tid; e < nc * nq; e += kInnerThreads) {
    const int l = e / nq, k = e - l * nq;
    Q[e] = N::make(l == k ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  // ---- inner sweep on G.  Per sub-round: the rotation of each disjoint pair from the current G, then
  // G <- J^H G J in one pass over 2 x 2 blocks (rows of slot p, columns of slot p2: each output block
  // depends only on the same input block and the two rotations) and Q <- Q J, all threads.  A slot is
  // a pair or, where the ordering leaves a column idle, a single column (j >= nc).
  //  - one panel (intra): all pairs of its columns, circle order (nc - 1 sub-rounds of nc / 2 pairs);
  //  - two panels: only the cross pairs (i in I, j in J), max(ci, cj) sub-rounds, sub-round s pairing
  //    column k of the smaller panel with column (k + s) mod (size) of the larger one.
  // Every pair of a block's columns is thus rotated once per sweep (panels against themselves in the
  // first round of a sweep, then the circle order of panel pairs).
  // Slot k's rotation is published as s_ij[k] = (i, j, rotated) and s_cs[k] = (c, s Re e^{i phi}) (+ s_si[k] =
  // s Im e^{i phi}); each thread's items (2 x 2 blocks of G or (slot, row) pieces of Q) are decoded once.
  {
    const double tol = DBL_EPSILON * sqrt((double)(p > 1 ? p : 1));
    const double floor2 = floors[jb[0]];
    const int ne = nc + (nc & 1), n1 = ne - 1;
    const int ci_ = (int)ci, cj_ = (int)cj;
    const int npr = intra ? ne / 2 : (ci_ > cj_ ? ci_ : cj_);  // slots per sub-round
    const int nsub = intra ? ne - 1 : npr;
    const int ng = npr * npr, nitems = ng + npr * nc;
    constexpr int kItems = (3 * kMaxPairs * kMaxPairs + kInnerThreads - 1) / kInnerThreads;
    int ia[kItems], ib[kItems];  // G item: (slot pa, slot pb); Q item: (-1 - row, slot pb); none: ib = -1
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      const int e = tid + k * kInnerThreads;
      if (e < ng) {
        ia[k] = e / npr;
        ib[k] = e - ia[k] * npr;
      } else if (e < nitems) {
        const int ee = e - ng;
        ib[k] = ee / nc;
        ia[k] = -1 - (ee - ib[k] * nc);
      } else {
        ia[k] = 0;
        ib[k] = -1;
      }
    }
    // slot state of the parameter lanes (tid < npr), advanced by one per sub-round
    int sa = 0, sb = 0;
    if (intra) {
      sa = tid % (n1 > 0 ? n1 : 1);             // (r + k) mod n1 at r = 0
      sb = (n1 - tid % (n1 > 0 ? n1 : 1)) % (n1 > 0 ? n1 : 1);  // (r - k + n1) mod n1 at r = 0
    } else if (ci_ <= cj_) {
      sa = tid % cj_;  // (k + r) mod cj
    } else {
      sa = tid % ci_;  // (k + r) mod ci
    }
    for (int r = 0; r < nsub; ++r) {
      if (tid < npr) {
        int i, j;
        if (intra) {
          if (tid == 0) {
            i = r;
            j = n1;
          } else {
            i = min(sa, sb);
            j = max(sa, sb);
          }
          sa = sa + 1 == n1 ? 0 : sa + 1;
          sb = sb + 1 == n1 ? 0 : sb + 1;
        } else if (ci_ <= cj_) {
          i = tid < ci_ ? tid : ci_ + sa;
          j = tid < ci_ ? ci_ + sa : nc;  // tid >= ci: idle column of panel J
          sa = sa + 1 == cj_ ? 0 : sa + 1;
        } else {
          i = sa;
          j = tid < cj_ ? ci_ + tid : nc;  // tid >= cj: idle column of panel I
          sa = sa + 1 == ci_ ? 0 : sa + 1;
        }
        int go = 0;
        double c = 1.0, sn = 0.0, er = 1.0, ei = 0.0;
        if (j < nc) {
          const T gij = G[j * nq + i];
          go = rotation(N::re(G[i * nq + i]), N::re(G[j * nq + j]), N::re(gij), N::im(gij), tol, floor2, c, sn, er,
                        ei)
                   ? 1
                   : 0;
        }
        s_ij[tid] = make_int4(i, j, go, 0);
        s_cs[tid] = make_double2(c, sn * er);
        if constexpr (CPLX) s_si[tid] = sn * ei;
        if (go) s_rot = 1;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kItems; ++k) {
        if (ib[k] < 0) continue;
        const int pb = ib[k];
        const int4 qb = s_ij[pb];
        if (ia[k] >= 0) {
          const int4 qa = s_ij[ia[k]];
          if (!qa.z && !qb.z) continue;
          const int i = qa.x, j = qa.y, i2 = qb.x, j2 = qb.y;
          const bool hj = j < nc, hj2 = j2 < nc;
          T b00 = G[i2 * nq + i];
          T b01 = hj2 ? G[j2 * nq + i] : N::make(0.0, 0.0);
          T b10 = hj ? G[i2 * nq + j] : N::make(0.0, 0.0);
          T b11 = (hj && hj2) ? G[j2 * nq + j] : N::make(0.0, 0.0);
          if (qa.z) {  // rows: J^H
            const double2 cs = s_cs[ia[k]];
            const double si = CPLX ? s_si[ia[k]] : 0.0;
            apply_rot<CPLX>(b00, b10, cs.x, 1.0, cs.y, -si);
            apply_rot<CPLX>(b01, b11, cs.x, 1.0, cs.y, -si);
          }
          if (qb.z) {  // columns: J
            const double2 cs = s_cs[pb];
            const double si = CPLX ? s_si[pb] : 0.0;
            apply_rot<CPLX>(b00, b01, cs.x, 1.0, cs.y, si);
            apply_rot<CPLX>(b10, b11, cs.x, 1.0, cs.y, si);
          }
          G[i2 * nq + i] = b00;
          if (hj2) G[j2 * nq + i] = b01;
          if (hj) G[i2 * nq + j] = b10;
          if (hj && hj2) G[j2 * nq + j] = b11;
        } else {
          if (!qb.z) continue;
          const int row = -1 - ia[k];
          const double2 cs = s_cs[pb];
          const double si = CPLX ? s_si[pb] : 0.0;
          T u = Q[qb.x * nq + row], v = Q[qb.y * nq + row];
          apply_rot<CPLX>(u, v, cs.x, 1.0, cs.y, si);
          Q[qb.x * nq + row] = u;
          Q[qb.y * nq + row] = v;
        }
      }
      __syncthreads();
    }
  }
  if (s_rot && tid == 0 && blockIdx.y == 0) state[kStateHead + sweep] = 1;
  if (!s_rot) return;  // uniform: nothing rotated, W and V of the pair are unchanged
  // ---- the pair's columns of this workgroup's stripes <- (.) Q in place (stripes are disjoint rows,
  // and a wave reads its whole stripe before writing it)
  for (int st = st0; st < s1; st += kInnerThreads / 64) {
    if (st != st0) stripe_load<CPLX, KS>(a, stripe_col(st), stripe_rows(st), stripe_idx(st), nc, lane);
    stripe_mul_store<CPLX, KS>(a, stripe_col(st), stripe_rows(st), stripe_idx(st), Q, nq, nc, lane);
  }
}

// Sweep count from the per-sweep rotation flags: the first sweep that rotated nothing ends the
// iteration (result = its number + 1); none among the `launched` sweeps: -1; plan too small: -3.
__global__ void panel_result_kernel(int32_t* __restrict__ state, int launched) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int res = -1;
  for (int s = 0; s < launched; ++s)
    if (state[kStateHead + s] == 0) {
      res = s + 1;
      break;
    }
  state[3] = state[1] ? -3 : res;
}

template <bool CPLX, int MODE>
__global__ void __launch_bounds__(kPanelThreads)
    panel_finalize_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ ids,
                          const typename Num<CPLX>::T* __restrict__ work, const double* __restrict__ mu,
                          typename Num<CPLX>::T* __restrict__ U, double* __restrict__ S,
                          typename Num<CPLX>::T* __restrict__ Vh, int32_t* __restrict__ info, int sweeps,
                          const int32_t* __restrict__ d_sweeps) {
  const int bidx = ids[blockIdx.x];
  const BlockDesc d = load_desc(blocks + (int64_t)bidx * YH_SVD_WORDS);
  const bool tr = (MODE == 0) && (d.m < d.n);
  const int64_t p = tr ? d.n : d.m, q = tr ? d.m : d.n;
  const typename Num<CPLX>::T* W = work + d.work_off;
  finalize_block<CPLX, MODE>(d, W, W + p * q, p, q, tr, mu[bidx], U, S, Vh);
  if (threadIdx.x == 0) info[bidx] = d_sweeps ? *d_sweeps : sweeps;
}

// Block-wide sum over kThreads threads (kWaves partials in s_red); every thread receives the total.
__device__ inline double block_sum(double v, double* s_red, int lane, int wave) {
  v = wave_sum(v);
  __syncthreads();  // earlier readers of s_red are done
  if (lane == 0) s_red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kWaves; ++w) t += s_red[w];
  return t;
}

// Orthonormal completion of the singular vectors that belong to noise-level singular values.
// Rotations skip columns whose squared norm is at the noise floor eps^2 ||A||_F^2, so for such a
// column w_j / sigma_j is not orthogonal to the others.  The W-side vectors (columns of U when
// m >= n, rows of Vh when m < n) of the trailing singular values with sigma^2 <= 4 eps^2 sum(sigma^2)
// are replaced, in order, by vectors orthogonalised against every vector before them (classical
// Gram-Schmidt, two passes): first the scaled column itself, then unit vectors e_r if that leaves
// less than 1e-6 of the start.  An all-zero block thus gets the leading unit vectors, as LAPACK's
// SVD returns for a zero matrix.  The accumulated side (V) is orthonormal by construction.
//
// Two kernels share the work.  complete_project_kernel removes, for every noise column j >= c0 at
// once (grid.y workgroups per block), its components along the c0 kept vectors (two passes, after
// scaling the column to unit norm).  complete_kernel then runs the sequence above, whose first
// attempt for column c only has to project against the noise columns c0 ... c-1 (its acceptance
// test, against the start's norm, is unchanged: the start was scaled to norm 1); the unit-vector
// attempts project against every earlier vector as before.

// The kept count c0 of a block (S descending) and the vector layout of its normalised-column side.
struct CompleteSide {
  int64_t k, p, cs, xs, c0;
  bool side_vh;
};

template <bool CPLX>
__device__ inline CompleteSide complete_side(const BlockDesc& d, const double* __restrict__ S, int swap,
                                             double* s_red, int tid, int lane, int wave) {
  CompleteSide c;
  c.k = d.m < d.n ? d.m : d.n;
  // columns of U (length m) or rows of Vh (length n); without preconditioning it is U when m >= n,
  // with it (swap) the other one
  c.side_vh = (d.m < d.n) != (swap != 0);
  c.p = c.side_vh ? d.n : d.m;
  c.cs = c.side_vh ? d.n : 1;  // vector v, element x, at base[v * cs + x * xs]
  c.xs = c.side_vh ? 1 : c.k;
  double f = 0.0;
  for (int64_t j = tid; j < c.k; j += kThreads) f += S[d.s_off + j] * S[d.s_off + j];
  const double thr2 = 4.0 * DBL_EPSILON * DBL_EPSILON * block_sum(f, s_red, lane, wave);
  double cnt = 0.0;
  for (int64_t j = tid; j < c.k; j += kThreads) cnt += (S[d.s_off + j] * S[d.s_off + j] > thr2) ? 1.0 : 0.0;
  c.c0 = (int64_t)block_sum(cnt, s_red, lane, wave);
  return c;
}

// v_c -= sum_{l in [lb, c)} <v_l, v_c> v_l, twice (classical Gram-Schmidt, two passes), by the whole
// workgroup.  s_cr / s_ci hold the coefficients (indexed by l), s_pr / s_pim the partial sums.
template <bool CPLX>
__device__ inline void cgs2(typename Num<CPLX>::T* __restrict__ base, typename Num<CPLX>::T* __restrict__ vc,
                            int64_t lb, int64_t c, int64_t p, int64_t cs, int64_t xs, double* s_cr, double* s_ci,
                            double* s_pr, double* s_pim, int tid, int lane, int wave) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  for (int pass = 0; pass < 2; ++pass) {
    // h_l = <v_l, v_c>: a wave takes four l at a time, so their loads are in flight together
    for (int64_t l0 = lb + 4 * wave; l0 < c; l0 += 4 * kWaves) {
      double hr[4] = {0.0, 0.0, 0.0, 0.0}, hi[4] = {0.0, 0.0, 0.0, 0.0};
      for (int64_t x = lane; x < p; x += 64) {
        const T b = vc[x * xs];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (l0 + u < c) {
            const T a = base[(l0 + u) * cs + x * xs];
            hr[u] += N::re(a) * N::re(b) + N::im(a) * N::im(b);  // conj(a) b
            hi[u] += N::re(a) * N::im(b) - N::im(a) * N::re(b);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double r = wave_sum(hr[u]), i = wave_sum(hi[u]);
        if (lane == 0 && l0 + u < c) {
          s_cr[l0 + u] = r;
          s_ci[l0 + u] = i;
        }
      }
    }
    __syncthreads();
    // v_c -= sum_l h_l v_l.  With ng = kThreads / p >= 2 groups of p threads, thread (g, x) sums the
    // terms l = lb + g mod ng and the partials meet in LDS; otherwise one thread per x does all l.
    const int64_t ng = p * 2 <= kThreads ? kThreads / p : 1;
    for (int64_t e = tid; e < ng * p; e += kThreads) {
      const int64_t g = e / p, x = e - g * p;
      double ar = 0.0, ai = 0.0;
#pragma unroll 8
      for (int64_t l = lb + g; l < c; l += ng) {
        const T a = base[l * cs + x * xs];
        ar += s_cr[l] * N::re(a) - s_ci[l] * N::im(a);
        ai += s_cr[l] * N::im(a) + s_ci[l] * N::re(a);
      }
      if (ng > 1) {
        s_pr[e] = ar;
        if constexpr (CPLX) s_pim[e] = ai;
      } else {
        const T b = vc[x * xs];
        vc[x * xs] = N::make(N::re(b) - ar, N::im(b) - ai);
      }
    }
    if (ng > 1) {
      __syncthreads();
      for (int64_t x = tid; x < p; x += kThreads) {
        double ar = 0.0, ai = 0.0;
        for (int64_t g = 0; g < ng; ++g) {
          ar += s_pr[g * p + x];
          if constexpr (CPLX) ai += s_pim[g * p + x];
        }
        const T b = vc[x * xs];
        vc[x * xs] = N::make(N::re(b) - ar, N::im(b) - ai);
      }
    }
    __syncthreads();
  }
}

template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    complete_project_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ ids,
                            typename Num<CPLX>::T* __restrict__ U, const double* __restrict__ S,
                            typename Num<CPLX>::T* __restrict__ Vh, int swap) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  __shared__ double s_red[kWaves];
  __shared__ double s_cr[YH_SVD_MAX_Q], s_ci[YH_SVD_MAX_Q];
  __shared__ double s_pr[kThreads], s_pim[CPLX ? kThreads : 1];
  const BlockDesc d = load_desc(blocks + (int64_t)ids[blockIdx.x] * YH_SVD_WORDS);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const CompleteSide sd = complete_side<CPLX>(d, S, swap, s_red, tid, lane, wave);
  if (sd.c0 == 0 || sd.c0 == sd.k) return;  // uniform per workgroup
  T* base = sd.side_vh ? Vh + d.v_off : U + d.u_off;
  for (int64_t c = sd.c0 + blockIdx.y; c < sd.k; c += gridDim.y) {
    T* vc = base + c * sd.cs;
    double n0 = 0.0;
    for (int64_t x = tid; x < sd.p; x += kThreads) n0 += N::abs2(vc[x * sd.xs]);
    n0 = block_sum(n0, s_red, lane, wave);
    if (!(n0 > 0.0)) continue;  // uniform: a zero start stays zero and falls back to unit vectors
    const double inv = 1.0 / sqrt(n0);
    for (int64_t x = tid; x < sd.p; x += kThreads) {
      const T b = vc[x * sd.xs];
      vc[x * sd.xs] = N::make(N::re(b) * inv, N::im(b) * inv);
    }
    __syncthreads();
    cgs2<CPLX>(base, vc, 0, sd.c0, sd.p, sd.cs, sd.xs, s_cr, s_ci, s_pr, s_pim, tid, lane, wave);
  }
}

template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    complete_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ ids,
                    typename Num<CPLX>::T* __restrict__ U, const double* __restrict__ S,
                    typename Num<CPLX>::T* __restrict__ Vh, int swap) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  __shared__ double s_red[kWaves];
  __shared__ double s_cr[YH_SVD_MAX_Q], s_ci[YH_SVD_MAX_Q];
  __shared__ double s_pr[kThreads], s_pim[CPLX ? kThreads : 1];  // partial sums of the update
  const int bidx = ids[blockIdx.x];
  const BlockDesc d = load_desc(blocks + (int64_t)bidx * YH_SVD_WORDS);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const CompleteSide sd = complete_side<CPLX>(d, S, swap, s_red, tid, lane, wave);
  const int64_t k = sd.k, p = sd.p, cs = sd.cs, xs = sd.xs, c0 = sd.c0;
  T* base = sd.side_vh ? Vh + d.v_off : U + d.u_off;
  if (c0 == 0) {
    // an all-noise (e.g. all-zero) block: the procedure below would accept e_c for every c, so write
    // the leading unit vectors directly instead of k sequential orthogonalisations
    for (int64_t e = tid; e < k * p; e += kThreads) {
      const int64_t c = e / p, x = e - c * p;
      base[c * cs + x * xs] = N::make(x == c ? 1.0 : 0.0, 0.0);
    }
    return;
  }
  for (int64_t c = c0; c < k; ++c) {
    T* vc = base + c * cs;
    for (int64_t t = 0; t <= p; ++t) {
      if (t > 0) {
        const int64_t r = (c + t - 1) % p;
        for (int64_t x = tid; x < p; x += kThreads) vc[x * xs] = N::make(x == r ? 1.0 : 0.0, 0.0);
      }
      double n0 = 0.0;
      for (int64_t x = tid; x < p; x += kThreads) n0 += N::abs2(vc[x * xs]);
      n0 = block_sum(n0, s_red, lane, wave);
      if (!(n0 > 0.0)) continue;
      if (t == 0) {
        // the start: scaled to norm 1 and free of the kept vectors' components (complete_project_kernel)
        n0 = 1.0;
        cgs2<CPLX>(base, vc, c0, c, p, cs, xs, s_cr, s_ci, s_pr, s_pim, tid, lane, wave);
      } else {
        cgs2<CPLX>(base, vc, 0, c, p, cs, xs, s_cr, s_ci, s_pr, s_pim, tid, lane, wave);
      }
      double n1 = 0.0;
      for (int64_t x = tid; x < p; x += kThreads) n1 += N::abs2(vc[x * xs]);
      n1 = block_sum(n1, s_red, lane, wave);
      if (n1 > 1e-12 * n0) {
        const double inv = 1.0 / sqrt(n1);
        for (int64_t x = tid; x < p; x += kThreads) {
          const T b = vc[x * xs];
          vc[x * xs] = N::make(N::re(b) * inv, N::im(b) * inv);
        }
        __syncthreads();
        break;
      }
    }
  }
}

template <bool CPLX>
int launch_complete(const int64_t* d_blocks, const int32_t* d_ids, int64_t nb, void* U, const void* S, void* Vh,
                    hipStream_t stream, int swap = 0) {
  typedef typename Num<CPLX>::T T;
  if (nb == 0) return YH_OK;
  // workgroups per block of the projection stage, each taking every kCompleteSlices-th noise column
  constexpr unsigned kCompleteSlices = 32;
  hipLaunchKernelGGL(complete_project_kernel<CPLX>, dim3((unsigned)nb, kCompleteSlices), dim3(kThreads), 0, stream,
                     d_blocks, d_ids, (T*)U, (const double*)S, (T*)Vh, swap);
  hipLaunchKernelGGL(complete_kernel<CPLX>, dim3((unsigned)nb), dim3(kThreads), 0, stream, d_blocks, d_ids, (T*)U,
                     (const double*)S, (T*)Vh, swap);
  return yh::hip_status(hipGetLastError(), "svd completion launch");
}

template <bool CPLX, int MODE>
int launch_jacobi(const int64_t* d_blocks, const int32_t* d_order, int64_t n_lds, int64_t lds_bytes,
                  int64_t nblocks, const void* A, void* U, void* S, void* Vh, void* work, int32_t* info,
                  int max_sweeps, hipStream_t stream) {
  typedef typename Num<CPLX>::T T;
  if (n_lds > 0) {
    if (lds_bytes > 65536) {
      const int st = yh::hip_status(hipFuncSetAttribute((const void*)jacobi_kernel<CPLX, true, MODE>,
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes),
                                    "jacobi LDS size");
      if (st) return st;
    }
    hipLaunchKernelGGL((jacobi_kernel<CPLX, true, MODE>), dim3((unsigned)n_lds), dim3(kThreads), (size_t)lds_bytes,
                       stream, d_blocks, d_order, (const T*)A, (T*)U, (double*)S, (T*)Vh, (T*)work, info,
                       max_sweeps);
    const int st = yh::hip_status(hipGetLastError(), "jacobi (LDS) launch");
    if (st) return st;
  }
  if (nblocks > n_lds) {
    hipLaunchKernelGGL((jacobi_kernel<CPLX, false, MODE>), dim3((unsigned)(nblocks - n_lds)), dim3(kThreads), 0,
                       stream, d_blocks, d_order + n_lds, (const T*)A, (T*)U, (double*)S, (T*)Vh, (T*)work, info,
                       max_sweeps);
    const int st = yh::hip_status(hipGetLastError(), "jacobi (global) launch");
    if (st) return st;
  }
  if (MODE == 0) return launch_complete<CPLX>(d_blocks, d_order, nblocks, U, S, Vh, stream);
  return YH_OK;
}


// ---------------------------------------------------------------------------------------------
// QR preconditioning of the panel path for SVD blocks (Drmac & Veselic, "New fast and accurate
// Jacobi SVD algorithm I", SIAM J. Matrix Anal. Appl. 29 (2008) 1322-1342): with W0 = Q R
// (Householder QR of the p x q working matrix), the one-sided iteration runs on X = R^H (q x q),
// which needs about half the sweeps; X V_J = U_J Sigma gives W0 = (Q V_J) Sigma U_J^H.
//
// QR descriptor, YH_QR_WORDS int64 per block: 0 p, 1 q, 2 w_off (W0, p x q column-major; R in its
// upper triangle afterwards), 3 vall_off (Householder vectors, p x q column-major, unit diagonal;
// rows above the diagonal are never read), 4 tm_off (-T of each panel, YH_QR_NB x YH_QR_NB
// column-major, panel after panel), 5 x_off (X, q x q column-major, then V_J, q x q), 6 y_off
// (two YH_QR_NB x q row-major scratch matrices for the GEMM steps), 7 unused.
struct QrDesc {
  int64_t p, q, w_off, vall_off, tm_off, x_off, y_off, k;
};

__device__ inline QrDesc load_qr(const int64_t* w) {
  QrDesc d;
  d.p = w[0];
  d.q = w[1];
  d.w_off = w[2];
  d.vall_off = w[3];
  d.tm_off = w[4];
  d.x_off = w[5];
  d.y_off = w[6];
  d.k = w[7] > 0 ? w[7] : w[1];  // reflectors: q unless given (min(p, q) for a wide block)
  return d;
}

// Householder reflector of a column (LAPACK larfg): alpha = x[j], xn2 = |x[j+1:]|^2; H = I - tau v v^H
// with H^H x = beta e_1 (beta real), v = [1; s x[j+1:]], s = scr + i sci = 1 / (alpha - beta).
template <bool CPLX>
__device__ __attribute__((always_inline)) inline void house_coeffs(typename Num<CPLX>::T alpha, double xn2,
                                                                   double& beta, double& taur, double& taui,
                                                                   double& scr, double& sci) {
  typedef Num<CPLX> N;
  taur = taui = scr = sci = 0.0;
  if (xn2 == 0.0 && N::im(alpha) == 0.0) {
    beta = N::re(alpha);  // H = I
    return;
  }
  const double nrm = sqrt(N::abs2(alpha) + xn2);
  beta = N::re(alpha) >= 0.0 ? -nrm : nrm;
  const double rb = 1.0 / beta;
  taur = (beta - N::re(alpha)) * rb;
  taui = -N::im(alpha) * rb;
  const double dr = N::re(alpha) - beta, di = N::im(alpha);
  const double rden = 1.0 / (dr * dr + di * di);
  scr = dr * rden;
  sci = -di * rden;
}

// Column j - 1 of the panel becomes [R(0:j-1, j-1); beta; s x[j:]] (its reflector scaled in place).
template <bool CPLX>
__device__ __attribute__((always_inline)) inline void scale_column(typename Num<CPLX>::T* xp, int64_t j, int64_t L,
                                                                   double scr, double sci, double beta, int tid) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  for (int64_t r = j + tid; r < L; r += kThreads) {
    const T v = xp[r];
    xp[r] = N::make(N::re(v) * scr - N::im(v) * sci, N::re(v) * sci + N::im(v) * scr);
  }
  if (tid == 0) xp[j - 1] = N::make(beta, 0.0);
}

// The column loop of the panel factorisation, one workgroup barrier per column.  Column j's reflector
// is applied to the later columns with v = [1; s x[j+1:]] formed on the fly from the unscaled column
// (the same products as scaling first); the column itself is scaled during the next column's update,
// which does not read it.  The wave that updates column j + 1 also sums the squares of its new tail,
// the next reflector's norm, so that no block-wide reduction sits in the loop.
// qr_columns_mem: every column update reads and writes the panel (LDS or, in place, global memory),
// one wave per later column.
template <bool CPLX>
__device__ __attribute__((always_inline)) inline void qr_columns_mem(typename Num<CPLX>::T* P, int64_t ld,
                                                                     int64_t L, int nb, double xn2, double* s_tr,
                                                                     double* s_ti, double* s_nrm, int tid, int lane,
                                                                     int wave) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  double pscr = 0.0, psci = 0.0, pbeta = 0.0;  // the previous column's scale factor and beta
  for (int j = 0; j < nb; ++j) {
    T* x = P + (int64_t)j * ld;
    if (j > 0) xn2 = s_nrm[j & 1];
    const T alpha = x[j];
    double beta, taur, taui, scr, sci;
    house_coeffs<CPLX>(alpha, xn2, beta, taur, taui, scr, sci);
    if (j > 0) scale_column<CPLX>(P + (int64_t)(j - 1) * ld, j, L, pscr, psci, pbeta, tid);
    if (tid == 0) {
      s_tr[j] = taur;
      s_ti[j] = taui;
    }
    // H_j^H = I - conj(tau) v v^H on the later panel columns, one wave per column
    for (int c = j + 1 + wave; c < nb; c += kWaves) {
      T* y = P + (int64_t)c * ld;
      double wr = 0.0, wi = 0.0;
      for (int64_t r = j + 1 + lane; r < L; r += 64) {
        const T xr = x[r], u = y[r];
        const double vr = N::re(xr) * scr - N::im(xr) * sci, vi = N::re(xr) * sci + N::im(xr) * scr;
        wr += vr * N::re(u) + vi * N::im(u);
        wi += vr * N::im(u) - vi * N::re(u);
      }
      wr = wave_sum(wr) + N::re(y[j]);
      if (CPLX) wi = wave_sum(wi);
      wi += N::im(y[j]);
      const double cr = taur * wr + taui * wi, ci = taur * wi - taui * wr;  // conj(tau) w
      double loc = 0.0;
      for (int64_t r = j + 1 + lane; r < L; r += 64) {
        const T xr = x[r], u = y[r];
        const double vr = N::re(xr) * scr - N::im(xr) * sci, vi = N::re(xr) * sci + N::im(xr) * scr;
        const T nu = N::make(N::re(u) - (vr * cr - vi * ci), N::im(u) - (vr * ci + vi * cr));
        y[r] = nu;
        if (r > j + 1) loc += N::abs2(nu);
      }
      if (lane == 0) y[j] = N::make(N::re(y[j]) - cr, N::im(y[j]) - ci);
      if (c == j + 1) {
        loc = wave_sum(loc);
        if (lane == 0) s_nrm[(j + 1) & 1] = loc;
      }
    }
    pscr = scr;
    psci = sci;
    pbeta = beta;
    __syncthreads();  // updated columns and s_nrm visible; every thread has read alpha and s_nrm
  }
  scale_column<CPLX>(P + (int64_t)(nb - 1) * ld, nb, L, pscr, psci, pbeta, tid);
}

// Rows per lane of the register instance for the largest LDS-resident panels.
template <bool CPLX>
constexpr int kRegRows = CPLX ? 5 : 9;

// qr_columns_reg (panel in LDS, L <= 64 RPL): column c stays in the registers of wave c % kWaves
// (slot c / kWaves; lane l holds rows l + 64 u) until it becomes the reflector; its owner writes it
// back to the LDS panel one iteration earlier, so every column is in the LDS when it is read as a
// reflector and the LDS traffic of a column step is one read of the reflector per wave.
template <bool CPLX, int RPL>
__device__ __attribute__((always_inline)) inline void qr_columns_reg(typename Num<CPLX>::T* P, int64_t ld,
                                                                     int64_t L, int nb, double xn2, double* s_tr,
                                                                     double* s_ti, double* s_nrm, int tid, int lane,
                                                                     int wave) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  constexpr int NS = YH_QR_NB / kWaves;  // column slots per wave
  const T one = N::make(1.0, 0.0), zero = N::make(0.0, 0.0);
  T Y[NS][RPL];
#pragma unroll
  for (int sl = 0; sl < NS; ++sl) {
    const int c = wave + kWaves * sl;
    const int64_t cc = c < nb ? c : nb - 1;
#pragma unroll
    for (int u = 0; u < RPL; ++u) {
      const int64_t r = lane + 64 * u;
      const T v = P[(r < L ? r : L - 1) + cc * ld];
      Y[sl][u] = ((c < nb) & (r < L)) ? v : zero;
    }
  }
  double pscr = 0.0, psci = 0.0, pbeta = 0.0;
  for (int j = 0; j < nb; ++j) {
    const T* x = P + (int64_t)j * ld;
    if (j > 0) xn2 = s_nrm[j & 1];
    const T alpha = x[j];
    double beta, taur, taui, scr, sci;
    house_coeffs<CPLX>(alpha, xn2, beta, taur, taui, scr, sci);
    if (j > 0) scale_column<CPLX>(P + (int64_t)(j - 1) * ld, j, L, pscr, psci, pbeta, tid);
    if (tid == 0) {
      s_tr[j] = taur;
      s_ti[j] = taui;
    }
    if (wave + kWaves * (NS - 1) > j) {  // wave-uniform: this wave owns a later column
      T v[RPL];  // v = [0 above j; 1; s x[j+1:]]
#pragma unroll
      for (int u = 0; u < RPL; ++u) {
        const int64_t r = lane + 64 * u;
        const T xr = x[r < L ? r : L - 1];
        const T sx = N::make(N::re(xr) * scr - N::im(xr) * sci, N::re(xr) * sci + N::im(xr) * scr);
        v[u] = ((r > j) & (r < L)) ? sx : (r == j ? one : zero);
      }
#pragma unroll
      for (int sl = 0; sl < NS; ++sl) {
        const int c = wave + kWaves * sl;
        if (c > j && c < nb) {  // wave-uniform
          double wr = 0.0, wi = 0.0;  // w = v^H y
#pragma unroll
          for (int u = 0; u < RPL; ++u) {
            wr += N::re(v[u]) * N::re(Y[sl][u]) + N::im(v[u]) * N::im(Y[sl][u]);
            if constexpr (CPLX) wi += N::re(v[u]) * N::im(Y[sl][u]) - N::im(v[u]) * N::re(Y[sl][u]);
          }
          wr = wave_sum(wr);
          if constexpr (CPLX) wi = wave_sum(wi);
          const double cr = taur * wr + taui * wi, ci = taur * wi - taui * wr;  // conj(tau) w
          double loc = 0.0;
#pragma unroll
          for (int u = 0; u < RPL; ++u) {
            const int64_t r = lane + 64 * u;
            Y[sl][u] = N::make(N::re(Y[sl][u]) - (N::re(v[u]) * cr - N::im(v[u]) * ci),
                               N::im(Y[sl][u]) - (N::re(v[u]) * ci + N::im(v[u]) * cr));
            loc += r > j + 1 ? N::abs2(Y[sl][u]) : 0.0;  // rows past L hold zeros
          }
          if (c == j + 1) {  // the next reflector: its tail norm, and the column itself into the LDS
            loc = wave_sum(loc);
            if (lane == 0) s_nrm[(j + 1) & 1] = loc;
            T* y = P + (int64_t)c * ld;
#pragma unroll
            for (int u = 0; u < RPL; ++u) {
              const int64_t r = lane + 64 * u;
              if (r < L) y[r] = Y[sl][u];
            }
          }
        }
      }
    }
    pscr = scr;
    psci = sci;
    pbeta = beta;
    __syncthreads();  // the next reflector and s_nrm visible; every thread has read alpha and s_nrm
  }
  scale_column<CPLX>(P + (int64_t)(nb - 1) * ld, nb, L, pscr, psci, pbeta, tid);
}

constexpr int64_t kQrMaxRows = (int64_t)1 << 27;  // rows x YH_QR_NB panel elements < 2^32 (32-bit indices)

// One panel of the blocked Householder QR (LAPACK geqrf / larfg / larft conventions): the
// (p - k0) x nb panel at column k0 = step * YH_QR_NB is factored column by column,
// H_j = I - tau_j v_j v_j^H with H_j^H x = beta e_1 (beta real), applying H_j^H to the panel's
// later columns.  The vectors go to Vall, the R part back to W0, and -T, with H_1 ... H_nb =
// I - V T V^H, to Tm.  The columns right of the panel are updated by grouped GEMMs on the host's plan.
// The panel is staged in LDS (INPLACE false) or, when it does not fit, factored in place in W0
// (leading dimension p; the workgroup's barriers order its global reads and writes, and the
// panel of one block, a few hundred KB, stays in L2).
template <bool CPLX, bool INPLACE>
__global__ void __launch_bounds__(kThreads)
    qr_panel_kernel(const int64_t* __restrict__ qdesc, const int32_t* __restrict__ ids,
                    typename Num<CPLX>::T* __restrict__ work, int step) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  constexpr int NB = YH_QR_NB;
  extern __shared__ double smem_raw[];
  __shared__ double s_red[kWaves];
  __shared__ double s_tr[NB], s_ti[NB];          // tau_j
  __shared__ double s_Zr[NB][NB], s_Zi[NB][NB];  // Z = V^H V (upper triangle), [row][col]
  const QrDesc d = load_qr(qdesc + (int64_t)ids[blockIdx.x] * YH_QR_WORDS);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t k0 = (int64_t)step * NB;
  const int nb = (int)((d.k - k0) < NB ? (d.k - k0) : NB);
  const int64_t L = d.p - k0;
  const int64_t p = d.p;
  T* W0 = work + d.w_off;
  T* Vall = work + d.vall_off;
  // the L x nb panel, column-major with leading dimension ld
  const int64_t ld = INPLACE ? p : L;
  T* P = INPLACE ? W0 + k0 + k0 * p : reinterpret_cast<T*>(smem_raw);
  if (!INPLACE) {  // (the panel fits the LDS, so 32-bit element indices)
    const uint32_t Lu = (uint32_t)L, n = (uint32_t)(L * nb);
    for (uint32_t e = tid; e < n; e += kThreads) {
      const uint32_t c = e / Lu, r = e - c * Lu;
      P[e] = W0[(k0 + r) + (k0 + c) * p];
    }
    __syncthreads();
  }
  double xn2 = 0.0;
  {
    double loc = 0.0;
    for (int64_t r = 1 + tid; r < L; r += kThreads) loc += N::abs2(P[r]);
    xn2 = block_sum(loc, s_red, lane, wave);
  }
  __shared__ double s_nrm[2];  // by column parity: column j's norm is read while j + 1's is written
  bool done = false;
  if constexpr (!INPLACE) {  // the later columns in registers (rows per lane: the smallest instance)
    if (L <= 64 * 2) {
      qr_columns_reg<CPLX, 2>(P, ld, L, nb, xn2, s_tr, s_ti, s_nrm, tid, lane, wave);
      done = true;
    } else if (!CPLX && L <= 64 * 4) {
      qr_columns_reg<CPLX, 4>(P, ld, L, nb, xn2, s_tr, s_ti, s_nrm, tid, lane, wave);
      done = true;
    } else if (L <= 64 * kRegRows<CPLX>) {
      qr_columns_reg<CPLX, kRegRows<CPLX>>(P, ld, L, nb, xn2, s_tr, s_ti, s_nrm, tid, lane, wave);
      done = true;
    }
  }
  if (!done) qr_columns_mem<CPLX>(P, ld, L, nb, xn2, s_tr, s_ti, s_nrm, tid, lane, wave);
  __syncthreads();
  // T from T^-1 = diag(1 / tau) + striu(V^H V) (Joffrain, Low, Quintana-Orti, van de Geijn, Van Zee,
  // "Accumulating Householder transformations, revisited", ACM TOMS 32 (2006)): column j of T solves a
  // triangular system, T(j, j) = tau_j and T(i, j) = -tau_i sum_{i < l <= j} Z(i, l) T(l, j), Z = V^H V
  // (tau_i = 0 gives a zero row and column, as the larft recurrence does).
  // Z's upper 16 x 16 tiles (00, 01, 11) are formed on the matrix cores by waves 0-11 (tile wave / 4,
  // the rows cut into four parts whose products are summed in part order); then each wave solves two
  // columns of T with no workgroup barrier, lanes 0-31 column wave and lanes 32-63 column wave + 16,
  // lane l holding T(l, column).
  {
    const int tile = wave >> 2, part = wave & 3;
    const int ti = tile == 2 ? 1 : 0, tj = tile == 0 ? 0 : 1;
    const bool active = tile < 3 && 16 * tj < nb;  // wave-uniform
    d4 accR = {0.0, 0.0, 0.0, 0.0}, accI = {0.0, 0.0, 0.0, 0.0};
    if (active) {
      const int64_t rbeg = 16 * ti;  // V(r, c) = 0 above the diagonal: rows below 16 ti give nothing
      const int64_t ns = L > rbeg ? (L - rbeg + 3) / 4 : 0;
      const int64_t q4 = (ns + 3) / 4;
      const int64_t sb = part * q4, se = ns < sb + q4 ? ns : sb + q4;
      const int ca = 16 * ti + (lane & 15), cb = 16 * tj + (lane & 15);
      const int64_t cac = ca < nb ? ca : nb - 1, cbc = cb < nb ? cb : nb - 1;
      const T one = N::make(1.0, 0.0), zero = N::make(0.0, 0.0);
      for (int64_t st = sb; st < se; ++st) {
        const int64_t r = rbeg + 4 * st + (lane >> 4);
        const int64_t rc = r < L ? r : L - 1;
        const T va = P[rc + cac * ld], vb = P[rc + cbc * ld];
        // V(r, c): 0 above the diagonal, 1 on it, the stored vector below
        const T a = ((ca < nb) & (r < L) & (r >= ca)) ? (r == ca ? one : va) : zero;
        const T bb = ((cb < nb) & (r < L) & (r >= cb)) ? (r == cb ? one : vb) : zero;
        accR = __builtin_amdgcn_mfma_f64_16x16x4f64(N::re(a), N::re(bb), accR, 0, 0, 0);
        if constexpr (CPLX) {  // conj(a) b
          accR = __builtin_amdgcn_mfma_f64_16x16x4f64(N::im(a), N::im(bb), accR, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f64_16x16x4f64(N::re(a), N::im(bb), accI, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f64_16x16x4f64(-N::im(a), N::re(bb), accI, 0, 0, 0);
        }
      }
    }
    // item v of lane l is Z(16 ti + l / 16 + 4 v, 16 tj + l % 16)
    const int zr0 = 16 * ti + (lane >> 4), zc = 16 * tj + (lane & 15);
    for (int pp = 0; pp < 4; ++pp) {
      if (active && part == pp) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int zrow = zr0 + 4 * v;
          s_Zr[zrow][zc] = pp == 0 ? accR[v] : s_Zr[zrow][zc] + accR[v];
          if constexpr (CPLX) s_Zi[zrow][zc] = pp == 0 ? accI[v] : s_Zi[zrow][zc] + accI[v];
        }
      }
      __syncthreads();
    }
  }
  T* Tm = work + d.tm_off + (int64_t)step * NB * NB;
  {
    const int h = lane >> 5, l = lane & 31, jc = wave + 16 * h;
    const bool col_ok = jc < nb;
    double xr = 0.0, xi = 0.0;
    if (col_ok && l == jc) {
      xr = s_tr[jc];
      xi = s_ti[jc];
    }
    for (int i = wave + 15; i >= 0; --i) {  // wave-uniform trip count; half 0 needs only i < wave
      const double zr = s_Zr[i][l], zi = CPLX ? s_Zi[i][l] : 0.0;
      // Z(i, l) T(l, jc) over i < l <= jc (T(l, jc) is zero for l > jc; Z below the diagonal is unset)
      const bool use = (l > i) & (l <= jc) & col_ok;
      double cr = use ? zr * xr - zi * xi : 0.0;
      double cim = use ? zr * xi + zi * xr : 0.0;
      cr += dpp_move<0xB1>(cr);
      cr += dpp_move<0x4E>(cr);
      cr += dpp_move<0x141>(cr);
      cr += dpp_move<0x140>(cr);
      const double sr = h ? read_lane(cr, 32) + read_lane(cr, 48) : read_lane(cr, 0) + read_lane(cr, 16);
      double si = 0.0;
      if constexpr (CPLX) {
        cim += dpp_move<0xB1>(cim);
        cim += dpp_move<0x4E>(cim);
        cim += dpp_move<0x141>(cim);
        cim += dpp_move<0x140>(cim);
        si = h ? read_lane(cim, 32) + read_lane(cim, 48) : read_lane(cim, 0) + read_lane(cim, 16);
      }
      if ((l == i) & (i < jc) & col_ok) {
        const double tr_ = s_tr[i], ti_ = s_ti[i];
        xr = -(tr_ * sr - ti_ * si);
        xi = -(tr_ * si + ti_ * sr);
      }
    }
    Tm[l + jc * NB] = (col_ok && l <= jc) ? N::make(-xr, -xi) : N::make(0.0, 0.0);
  }
  for (int64_t e = tid; e < L * nb; e += kThreads) {
    const int64_t c = (int64_t)((uint32_t)e / (uint32_t)L), r = e - c * L;  // L nb < 2^32 (q <= YH_SVD_MAX_Q)
    const T v = P[r + c * ld];  // in place: the same element as the W0 store below
    W0[(k0 + r) + (k0 + c) * p] = r <= c ? v : N::make(0.0, 0.0);
    Vall[(k0 + r) + (k0 + c) * p] = r < c ? N::make(0.0, 0.0) : (r == c ? N::make(1.0, 0.0) : v);
  }
}

// X = R^H (R is the q x q upper triangle of W0) and V_J = I.
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    precond_init_kernel(const int64_t* __restrict__ qdesc, const int32_t* __restrict__ ids,
                        typename Num<CPLX>::T* __restrict__ work) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  const QrDesc d = load_qr(qdesc + (int64_t)ids[blockIdx.x] * YH_QR_WORDS);
  const T* W0 = work + d.w_off;
  T* X = work + d.x_off;
  T* VJ = X + d.q * d.q;
  for (int64_t e = threadIdx.x; e < d.q * d.q; e += kThreads) {
    const int64_t c = e / d.q, r = e - c * d.q;
    X[e] = r >= c ? N::conj(W0[c + r * d.p]) : N::make(0.0, 0.0);
    VJ[e] = N::make(r == c ? 1.0 : 0.0, 0.0);
  }
}

// Finalize of the preconditioned blocks: sigma, ranks, q-side vectors and the [V_J; 0] start of
// the p-side vectors (jblocks: the Jacobi-stage descriptors, work_off = x_off, m = n = q).
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    precond_finalize_kernel(const int64_t* __restrict__ blocks, const int64_t* __restrict__ qdesc,
                            const int32_t* __restrict__ ids, const typename Num<CPLX>::T* __restrict__ work,
                            typename Num<CPLX>::T* __restrict__ U, double* __restrict__ S,
                            typename Num<CPLX>::T* __restrict__ Vh, int32_t* __restrict__ info, int sweeps,
                            const int32_t* __restrict__ d_sweeps) {
  const int bidx = ids[blockIdx.x];
  const BlockDesc d = load_desc(blocks + (int64_t)bidx * YH_SVD_WORDS);
  const QrDesc qd = load_qr(qdesc + (int64_t)bidx * YH_QR_WORDS);
  const bool tr = d.m < d.n;
  const typename Num<CPLX>::T* X = work + qd.x_off;
  finalize_block<CPLX, 2>(d, X, X + qd.q * qd.q, qd.q, qd.q, tr, 0.0, U, S, Vh, qd.p);
  if (threadIdx.x == 0) info[bidx] = d_sweeps ? *d_sweeps : sweeps;
}

// ---- batched QR (the backend's qr): A = Q R per block, economic, diag(R) real and >= 0.
// Block words (YH_QROP_WORDS): 0 a_off, 1 q_off, 2 r_off (element offsets of A (m x n), Q (m x k) and
// R (k x n), all row-major); the QR descriptor as for the preconditioning, with x_off holding E (m x k
// column-major: [I_k; 0], then Q before the sign fix) and word 7 = k = min(m, n).
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    qr_op_init_kernel(const int64_t* __restrict__ qdesc, const int64_t* __restrict__ qblocks,
                      const typename Num<CPLX>::T* __restrict__ A, typename Num<CPLX>::T* __restrict__ work) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  const QrDesc d = load_qr(qdesc + (int64_t)blockIdx.x * YH_QR_WORDS);
  const int64_t a_off = qblocks[(int64_t)blockIdx.x * YH_QROP_WORDS];
  const int64_t m = d.p, n = d.q, k = d.k;
  T* W0 = work + d.w_off;
  T* E = work + d.x_off;
  for (int64_t e = threadIdx.x; e < m * n; e += kThreads) {  // row-major read, column-major write
    const int64_t r = e / n, c = e - r * n;
    W0[r + c * m] = A[a_off + e];
  }
  for (int64_t e = threadIdx.x; e < m * k; e += kThreads) {
    const int64_t c = e / m, r = e - c * m;
    E[e] = N::make(r == c ? 1.0 : 0.0, 0.0);
  }
}

// sR_j = sign(Re R_jj) (+1 for 0): Q(:, j) sR_j and sR_j R(j, :), the reference's positive diagonal.
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    qr_op_finalize_kernel(const int64_t* __restrict__ qdesc, const int64_t* __restrict__ qblocks,
                          const typename Num<CPLX>::T* __restrict__ work, typename Num<CPLX>::T* __restrict__ Q,
                          typename Num<CPLX>::T* __restrict__ R) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  const QrDesc d = load_qr(qdesc + (int64_t)blockIdx.x * YH_QR_WORDS);
  const int64_t* bw = qblocks + (int64_t)blockIdx.x * YH_QROP_WORDS;
  const int64_t q_off = bw[1], r_off = bw[2];
  const int64_t m = d.p, n = d.q, k = d.k;
  const T* W0 = work + d.w_off;
  const T* E = work + d.x_off;
  auto sgn = [&](int64_t j) { return N::re(W0[j + j * m]) < 0.0 ? -1.0 : 1.0; };
  for (int64_t e = threadIdx.x; e < k * n; e += kThreads) {
    const int64_t j = e / n, c = e - j * n;
    const double sj = sgn(j);
    const T w = W0[j + c * m];
    R[r_off + e] = c >= j ? N::make(sj * N::re(w), sj * N::im(w)) : N::make(0.0, 0.0);
  }
  for (int64_t e = threadIdx.x; e < m * k; e += kThreads) {
    const int64_t r = e / k, j = e - r * k;
    const double sj = sgn(j);
    const T v = E[r + j * m];
    Q[q_off + e] = N::make(sj * N::re(v), sj * N::im(v));
  }
}

// QR adjoint (per block, m x n, k = min(m, n), R1 = R(:, :k)); with M1 = R1 gR1^H and M2 = gQ'^H Q
// (gQ' = gQ + A(:, k:) gR(:, k:)^H for m < n), the core is C = copyltu(M1 - M2): M(i, j) below the
// diagonal, conj(M(j, i)) above it and Re M(i, i) on it (the real part only, because the gauge of the
// factorisation fixes diag(R) real: the imaginary diagonal of Q^H dQ is not free).  Words (YH_QRADJ_WORDS): 0 m, 1 n, 2 k, 3 m_off (M1, M2 and C,
// k x k row-major each, at the same offset of their buffers), 4 b_off (B = gQ' + Q C, m x k row-major),
// 5 ga_off (the block of gA, m x n row-major), 6 r_off (R, k x n row-major), 7 unused.
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    qr_adj_core_kernel(const int64_t* __restrict__ words, const typename Num<CPLX>::T* __restrict__ M1,
                       const typename Num<CPLX>::T* __restrict__ M2, typename Num<CPLX>::T* __restrict__ C) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  const int64_t* w = words + (int64_t)blockIdx.x * YH_QRADJ_WORDS;
  const int64_t k = w[2], off = w[3];
  for (int64_t e = threadIdx.x; e < k * k; e += kThreads) {
    const int64_t i = e / k, j = e - i * k;
    const int64_t src = i >= j ? e : j * k + i;
    const T a = M1[off + src], b = M2[off + src];
    const T d = N::make(N::re(a) - N::re(b), N::im(a) - N::im(b));
    C[off + e] = i > j ? d : (i == j ? N::make(N::re(d), 0.0) : N::conj(d));
  }
}

// gA(r, :k) = B(r, :) R1^{-H}: back substitution of x R1^H = b for one row per single-wave workgroup
// (grid: rows x blocks); the row lives in LDS, x_j = (b_j - sum_{l > j} x_l conj(R1(j, l))) / R1(j, j)
// (the diagonal of R is real), applied right-looking: once x_j is known, x_i -= x_j conj(R1(i, j)), i < j.
template <bool CPLX>
__global__ void __launch_bounds__(64)
    qr_adj_trsm_kernel(const int64_t* __restrict__ words, const typename Num<CPLX>::T* __restrict__ B,
                       const typename Num<CPLX>::T* __restrict__ R, typename Num<CPLX>::T* __restrict__ gA) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  extern __shared__ double smem_raw[];
  const int64_t* w = words + (int64_t)blockIdx.y * YH_QRADJ_WORDS;
  const int64_t m = w[0], n = w[1], k = w[2], b_off = w[4], ga_off = w[5], r_off = w[6];
  const int64_t r = blockIdx.x;
  if (r >= m) return;  // uniform: this block has fewer rows
  const int lane = threadIdx.x;
  T* x = reinterpret_cast<T*>(smem_raw);
  for (int64_t i = lane; i < k; i += 64) x[i] = B[b_off + r * k + i];
  __syncthreads();
  for (int64_t j = k - 1; j >= 0; --j) {
    const double rjj = N::re(R[r_off + j * n + j]);
    const T xj = N::make(N::re(x[j]) / rjj, N::im(x[j]) / rjj);
    __syncthreads();  // every lane has read x[j]
    if (lane == 0) x[j] = xj;
    for (int64_t i = lane; i < j; i += 64) {
      const T rij = R[r_off + i * n + j];  // x_i -= x_j conj(R1(i, j))
      const double cr = N::re(xj) * N::re(rij) + N::im(xj) * N::im(rij);
      const double ci = N::im(xj) * N::re(rij) - N::re(xj) * N::im(rij);
      x[i] = N::make(N::re(x[i]) - cr, N::im(x[i]) - ci);
    }
    __syncthreads();
  }
  for (int64_t i = lane; i < k; i += 64) gA[ga_off + r * n + i] = x[i];
}

}  // namespace

extern "C" int yastn_hip_check_svd_plan(const int64_t* h_blocks, int64_t nblocks, const int32_t* h_order,
                                        int64_t n_lds, int64_t lds_bytes, int mode, int dtype, int64_t a_numel,
                                        int64_t u_numel, int64_t s_numel, int64_t v_numel, int64_t work_numel) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(mode == 0 || mode == 1, "mode %d", mode);
  YH_CHECK_ARG(nblocks >= 0 && n_lds >= 0 && n_lds <= nblocks, "block counts");
  YH_CHECK_ARG(lds_bytes >= 0 && lds_bytes <= YH_SVD_LDS_BYTES, "lds bytes %lld", (long long)lds_bytes);
  const int64_t isz = yh::itemsize_of(dtype);
  char* seen = (char*)calloc(nblocks > 0 ? nblocks : 1, 1);
  YH_CHECK_ARG(seen != nullptr, "host allocation");
  int st = YH_OK;
  for (int64_t o = 0; o < nblocks && st == YH_OK; ++o) {
    const int32_t b = h_order[o];
    if (b < 0 || b >= nblocks || seen[b]) {
      yh::set_error("order[%lld] = %d invalid or repeated", (long long)o, b);
      st = YH_ERR_ARG;
      break;
    }
    seen[b] = 1;
    const int64_t* w = h_blocks + (int64_t)b * YH_SVD_WORDS;
    const int64_t m = w[1], n = w[2];
    const bool tr = (mode == 0) && (m < n);
    const int64_t p = tr ? n : m, q = tr ? m : n;
    if (m < 1 || n < 1 || q > YH_SVD_MAX_Q || (mode == 1 && m != n)) {
      yh::set_error("block %d: shape %lld x %lld", b, (long long)m, (long long)n);
      st = YH_ERR_ARG;
      break;
    }
    const bool ok_a = w[0] >= 0 && w[0] + m * n <= a_numel;
    const bool ok_u = w[3] >= 0 && w[3] + (mode == 0 ? m * q : n * n) <= u_numel;
    const bool ok_s = w[4] >= 0 && w[4] + q <= s_numel;
    const bool ok_v = mode == 1 || (w[5] >= 0 && w[5] + q * n <= v_numel);
    const int64_t need = (p * q + q * q);
    const bool ok_w = o < n_lds ? need * isz <= lds_bytes : (w[6] >= 0 && w[6] + need <= work_numel);
    if (!(ok_a && ok_u && ok_s && ok_v && ok_w)) {
      yh::set_error("block %d: offsets out of bounds (a%d u%d s%d v%d work%d)", b, ok_a, ok_u, ok_s, ok_v, ok_w);
      st = YH_ERR_ARG;
    }
  }
  free(seen);
  return st;
}

extern "C" int yastn_hip_svd(void* U, void* S, void* Vh, const void* A, int dtype, const int64_t* d_blocks,
                             int64_t nblocks, const int32_t* d_order, int64_t n_lds, int64_t lds_bytes, void* work,
                             int32_t* d_info, int max_sweeps, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(lds_bytes >= 0 && lds_bytes <= YH_SVD_LDS_BYTES, "lds bytes");
  if (nblocks == 0) return YH_OK;
  return dtype == YH_F64 ? launch_jacobi<false, 0>(d_blocks, d_order, n_lds, lds_bytes, nblocks, A, U, S, Vh, work,
                                                   d_info, max_sweeps, stream)
                         : launch_jacobi<true, 0>(d_blocks, d_order, n_lds, lds_bytes, nblocks, A, U, S, Vh, work,
                                                  d_info, max_sweeps, stream);
}

extern "C" int yastn_hip_eigh(void* S, void* U, const void* A, int dtype, const int64_t* d_blocks, int64_t nblocks,
                              const int32_t* d_order, int64_t n_lds, int64_t lds_bytes, void* work, int32_t* d_info,
                              int max_sweeps, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(lds_bytes >= 0 && lds_bytes <= YH_SVD_LDS_BYTES, "lds bytes");
  if (nblocks == 0) return YH_OK;
  return dtype == YH_F64 ? launch_jacobi<false, 1>(d_blocks, d_order, n_lds, lds_bytes, nblocks, A, U, S, nullptr,
                                                   work, d_info, max_sweeps, stream)
                         : launch_jacobi<true, 1>(d_blocks, d_order, n_lds, lds_bytes, nblocks, A, U, S, nullptr,
                                                  work, d_info, max_sweeps, stream);
}

extern "C" int yastn_hip_device_info(int* lds_per_block, int* num_cu) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(num_cu, hipDeviceAttributeMultiprocessorCount, dev);
  return yh::hip_status(e, "device info");
}

extern "C" int yastn_hip_svd_panel_init(const void* A, void* work, double* mu, double* floors, int dtype, int mode,
                                        const int64_t* d_blocks, const int32_t* d_ids, int64_t nb,
                                        hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(mode == 0 || mode == 1, "mode %d", mode);
  if (nb == 0) return YH_OK;
  const dim3 g((unsigned)nb), t(kPanelThreads);
  if (dtype == YH_F64) {
    if (mode == 0)
      hipLaunchKernelGGL((panel_init_kernel<false, 0>), g, t, 0, stream, d_blocks, d_ids, (const double*)A,
                         (double*)work, mu, floors);
    else
      hipLaunchKernelGGL((panel_init_kernel<false, 1>), g, t, 0, stream, d_blocks, d_ids, (const double*)A,
                         (double*)work, mu, floors);
  } else {
    if (mode == 0)
      hipLaunchKernelGGL((panel_init_kernel<true, 0>), g, t, 0, stream, d_blocks, d_ids, (const double2*)A,
                         (double2*)work, mu, floors);
    else
      hipLaunchKernelGGL((panel_init_kernel<true, 1>), g, t, 0, stream, d_blocks, d_ids, (const double2*)A,
                         (double2*)work, mu, floors);
  }
  return yh::hip_status(hipGetLastError(), "svd panel init launch");
}

extern "C" int yastn_hip_svd_panel_sweep(void* work, const double* floors, int dtype, const int64_t* d_blocks,
                                         const int32_t* d_jobs, const int64_t* h_rounds, int64_t nrounds, int b,
                                         int maxch_p, int slices, void* part, int64_t part_numel, int sweep,
                                         int max_sweeps, int32_t* d_state, int64_t state_numel, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(b >= 1 && b <= 32, "panel width %d", b);
  YH_CHECK_ARG(maxch_p >= 1 && maxch_p <= 65535, "chunk count %d", maxch_p);
  YH_CHECK_ARG(slices >= 1 && slices <= 64, "update slices %d", slices);
  YH_CHECK_ARG(max_sweeps >= 1 && max_sweeps <= 1000 && sweep >= 0 && sweep < max_sweeps, "sweep %d of %d", sweep,
               max_sweeps);
  YH_CHECK_ARG(d_state != nullptr && state_numel >= kStateHead + max_sweeps, "state buffer of %lld words",
               (long long)state_numel);
  const bool cplx = dtype == YH_C128;
  const int T16m = (2 * b + 15) / 16, ntmax = T16m * (T16m + 1) / 2;
  const int64_t isz = cplx ? 16 : 8;
  int64_t maxn = 0, end = 0;
  for (int64_t r = 0; r < nrounds; ++r) {
    YH_CHECK_ARG(h_rounds[2 * r] >= end && h_rounds[2 * r + 1] >= 0, "round %lld", (long long)r);
    end = h_rounds[2 * r] + h_rounds[2 * r + 1];
    maxn = std::max(maxn, h_rounds[2 * r + 1]);
  }
  YH_CHECK_ARG(maxn <= 65535, "jobs per round %lld", (long long)maxn);
  const int gch = (int)((maxch_p * (int64_t)kChunkRows + kGramRows - 1) / kGramRows);  // Gram chunks
  YH_CHECK_ARG(part_numel >= maxn * gch * ntmax * 256 * (cplx ? 2 : 1), "partial Gram buffer too small");
  const int64_t lds_inner = 2 * (2 * b) * (2 * b + 1) * isz;
  const int64_t lds_gram = (2 * b) * (kGramRows + 1) * isz;
  // the inner kernel's stripe registers cover 4 KS columns: 32 (KS = 8) for b <= 16, else 64
  const bool wide = 2 * b > 32;
  const void* inner = cplx ? (wide ? (const void*)panel_inner_kernel<true, 16> : (const void*)panel_inner_kernel<true, 8>)
                           : (wide ? (const void*)panel_inner_kernel<false, 16> : (const void*)panel_inner_kernel<false, 8>);
  int st = yh::hip_status(hipFuncSetAttribute(inner, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_inner),
                          "svd panel inner LDS size");
  // the Gram kernel stages 2 b x kGramRows elements, NST per thread: 8 for b <= 16, else 16
  const void* gram = cplx ? (wide ? (const void*)panel_gram_kernel<true, 16> : (const void*)panel_gram_kernel<true, 8>)
                          : (wide ? (const void*)panel_gram_kernel<false, 16> : (const void*)panel_gram_kernel<false, 8>);
  if (!st)
    st = yh::hip_status(hipFuncSetAttribute(gram, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_gram),
                        "svd panel gram LDS size");
  if (st) return st;
  for (int64_t r = 0; r < nrounds; ++r) {
    const int64_t off = h_rounds[2 * r], n = h_rounds[2 * r + 1];
    if (n == 0) continue;
    const int32_t* jobs = d_jobs + 3 * off;
    const dim3 gg((unsigned)gch, (unsigned)n), gi((unsigned)n, (unsigned)slices);
    if (cplx) {
      if (wide)
        hipLaunchKernelGGL((panel_gram_kernel<true, 16>), gg, dim3(kChunkThreads), (size_t)lds_gram, stream, d_blocks,
                           jobs, (const double2*)work, (double*)part, b, gch, d_state, sweep);
      else
        hipLaunchKernelGGL((panel_gram_kernel<true, 8>), gg, dim3(kChunkThreads), (size_t)lds_gram, stream, d_blocks,
                           jobs, (const double2*)work, (double*)part, b, gch, d_state, sweep);
      if (wide)
        hipLaunchKernelGGL((panel_inner_kernel<true, 16>), gi, dim3(kInnerThreads), (size_t)lds_inner, stream,
                           d_blocks, jobs, (double2*)work, floors, (const double*)part, d_state, sweep, b, gch,
                           lds_inner);
      else
        hipLaunchKernelGGL((panel_inner_kernel<true, 8>), gi, dim3(kInnerThreads), (size_t)lds_inner, stream,
                           d_blocks, jobs, (double2*)work, floors, (const double*)part, d_state, sweep, b, gch,
                           lds_inner);
    } else {
      if (wide)
        hipLaunchKernelGGL((panel_gram_kernel<false, 16>), gg, dim3(kChunkThreads), (size_t)lds_gram, stream, d_blocks,
                           jobs, (const double*)work, (double*)part, b, gch, d_state, sweep);
      else
        hipLaunchKernelGGL((panel_gram_kernel<false, 8>), gg, dim3(kChunkThreads), (size_t)lds_gram, stream, d_blocks,
                           jobs, (const double*)work, (double*)part, b, gch, d_state, sweep);
      if (wide)
        hipLaunchKernelGGL((panel_inner_kernel<false, 16>), gi, dim3(kInnerThreads), (size_t)lds_inner, stream,
                           d_blocks, jobs, (double*)work, floors, (const double*)part, d_state, sweep, b, gch,
                           lds_inner);
      else
        hipLaunchKernelGGL((panel_inner_kernel<false, 8>), gi, dim3(kInnerThreads), (size_t)lds_inner, stream,
                           d_blocks, jobs, (double*)work, floors, (const double*)part, d_state, sweep, b, gch,
                           lds_inner);
    }
    st = yh::hip_status(hipGetLastError(), "svd panel round launch");
    if (st) return st;
  }
  return YH_OK;
}

extern "C" int yastn_hip_svd_panel_result(int32_t* d_state, int64_t state_numel, int launched, hipStream_t stream) {
  YH_CHECK_ARG(launched >= 0 && state_numel >= kStateHead + launched, "%d sweeps, state of %lld words", launched,
               (long long)state_numel);
  hipLaunchKernelGGL(panel_result_kernel, dim3(1), dim3(64), 0, stream, d_state, launched);
  return yh::hip_status(hipGetLastError(), "svd panel result launch");
}

extern "C" int yastn_hip_check_panel_jobs(const int64_t* h_blocks, int64_t nblocks, const int32_t* h_jobs,
                                          int64_t njobs, int b, int mode, int maxch_p) {
  YH_CHECK_ARG(b >= 1 && b <= 32, "panel width %d", b);
  YH_CHECK_ARG(mode == 0 || mode == 1, "mode %d", mode);
  for (int64_t j = 0; j < njobs; ++j) {
    const int32_t blk = h_jobs[3 * j], pi = h_jobs[3 * j + 1], pj = h_jobs[3 * j + 2];
    YH_CHECK_ARG(blk >= 0 && blk < nblocks, "job %lld: block %d", (long long)j, blk);
    const int64_t m = h_blocks[blk * YH_SVD_WORDS + 1], n = h_blocks[blk * YH_SVD_WORDS + 2];
    const int64_t p = m < n ? n : m, q = m < n ? m : n;
    const int64_t npan = (q + b - 1) / b;
    YH_CHECK_ARG(m >= 1 && n >= 1 && pi >= 0 && pi <= pj && pj < npan, "job %lld: panels %d, %d of %lld",
                 (long long)j, pi, pj, (long long)npan);
    YH_CHECK_ARG((p + kChunkRows - 1) / kChunkRows <= maxch_p, "job %lld: %lld rows beyond the chunk grid",
                 (long long)j, (long long)p);
  }
  return YH_OK;
}

extern "C" int yastn_hip_svd_panel_finalize(void* U, void* S, void* Vh, const void* work, const double* mu,
                                            int dtype, int mode, const int64_t* d_blocks, const int32_t* d_ids,
                                            int64_t nb, int32_t* d_info, int sweeps, const int32_t* d_sweeps,
                                            hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(mode == 0 || mode == 1, "mode %d", mode);
  if (nb == 0) return YH_OK;
  const dim3 g((unsigned)nb), t(kPanelThreads);
  if (dtype == YH_F64) {
    if (mode == 0)
      hipLaunchKernelGGL((panel_finalize_kernel<false, 0>), g, t, 0, stream, d_blocks, d_ids, (const double*)work,
                         mu, (double*)U, (double*)S, (double*)Vh, d_info, sweeps, d_sweeps);
    else
      hipLaunchKernelGGL((panel_finalize_kernel<false, 1>), g, t, 0, stream, d_blocks, d_ids, (const double*)work,
                         mu, (double*)U, (double*)S, (double*)Vh, d_info, sweeps, d_sweeps);
  } else {
    if (mode == 0)
      hipLaunchKernelGGL((panel_finalize_kernel<true, 0>), g, t, 0, stream, d_blocks, d_ids, (const double2*)work,
                         mu, (double2*)U, (double*)S, (double2*)Vh, d_info, sweeps, d_sweeps);
    else
      hipLaunchKernelGGL((panel_finalize_kernel<true, 1>), g, t, 0, stream, d_blocks, d_ids, (const double2*)work,
                         mu, (double2*)U, (double*)S, (double2*)Vh, d_info, sweeps, d_sweeps);
  }
  const int st = yh::hip_status(hipGetLastError(), "svd panel finalize launch");
  if (st || mode == 1) return st;
  return dtype == YH_F64 ? launch_complete<false>(d_blocks, d_ids, nb, U, S, Vh, stream)
                         : launch_complete<true>(d_blocks, d_ids, nb, U, S, Vh, stream);
}

extern "C" int yastn_hip_check_qr_plan(const int64_t* h_qdesc, const int64_t* h_blocks, const int32_t* h_ids,
                                       int64_t n, int dtype, int64_t work_numel, int64_t lds_bytes) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  const int64_t isz = yh::itemsize_of(dtype);
  for (int64_t t = 0; t < n; ++t) {
    const int32_t b = h_ids[t];
    const int64_t* w = h_qdesc + (int64_t)b * YH_QR_WORDS;
    const int64_t* sw = h_blocks + (int64_t)b * YH_SVD_WORDS;
    const int64_t p = w[0], q = w[1];
    const int64_t sm = sw[1], sn = sw[2];
    const int64_t np = (q + YH_QR_NB - 1) / YH_QR_NB;
    YH_CHECK_ARG(b >= 0 && q >= 1 && p >= q && q <= YH_SVD_MAX_Q, "qr block %d: %lld x %lld", b, (long long)p,
                 (long long)q);
    YH_CHECK_ARG((sm >= sn ? sm : sn) == p && (sm < sn ? sm : sn) == q, "qr block %d: shape mismatch", b);
    YH_CHECK_ARG(p <= kQrMaxRows, "qr block %d: %lld rows (at most %lld)", b, (long long)p, (long long)kQrMaxRows);
    YH_CHECK_ARG(lds_bytes == 0 || p * YH_QR_NB * isz <= lds_bytes, "qr block %d: panel of %lld rows exceeds LDS",
                 b, (long long)p);
    const bool ok = w[2] >= 0 && w[2] + p * q <= work_numel && w[3] >= 0 && w[3] + p * q <= work_numel &&
                    w[4] >= 0 && w[4] + np * YH_QR_NB * YH_QR_NB <= work_numel && w[5] >= 0 &&
                    w[5] + 2 * q * q <= work_numel && w[6] >= 0 && w[6] + 2 * YH_QR_NB * q <= work_numel;
    YH_CHECK_ARG(ok, "qr block %d: workspace offsets out of bounds", b);
  }
  return YH_OK;
}

extern "C" int yastn_hip_svd_qr_panel(void* work, int dtype, const int64_t* d_qdesc, const int32_t* d_ids, int64_t n,
                                      int step, int64_t lds_bytes, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(step >= 0 && lds_bytes >= 0 && lds_bytes <= 160 * 1024, "qr panel step %d / lds %lld", step,
               (long long)lds_bytes);
  if (n == 0) return YH_OK;
  const dim3 g((unsigned)n), t(kThreads);
  if (lds_bytes == 0) {  // panels factored in place in the workspace
    if (dtype == YH_F64)
      hipLaunchKernelGGL((qr_panel_kernel<false, true>), g, t, 0, stream, d_qdesc, d_ids, (double*)work, step);
    else
      hipLaunchKernelGGL((qr_panel_kernel<true, true>), g, t, 0, stream, d_qdesc, d_ids, (double2*)work, step);
  } else if (dtype == YH_F64) {
    const int st = yh::hip_status(hipFuncSetAttribute((const void*)qr_panel_kernel<false, false>,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes),
                                  "svd qr panel LDS size");
    if (st) return st;
    hipLaunchKernelGGL((qr_panel_kernel<false, false>), g, t, (size_t)lds_bytes, stream, d_qdesc, d_ids,
                       (double*)work, step);
  } else {
    const int st = yh::hip_status(hipFuncSetAttribute((const void*)qr_panel_kernel<true, false>,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes),
                                  "svd qr panel LDS size");
    if (st) return st;
    hipLaunchKernelGGL((qr_panel_kernel<true, false>), g, t, (size_t)lds_bytes, stream, d_qdesc, d_ids,
                       (double2*)work, step);
  }
  return yh::hip_status(hipGetLastError(), "svd qr panel launch");
}

extern "C" int yastn_hip_svd_precond_init(void* work, int dtype, const int64_t* d_qdesc, const int32_t* d_ids,
                                          int64_t n, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  if (n == 0) return YH_OK;
  if (dtype == YH_F64)
    hipLaunchKernelGGL(precond_init_kernel<false>, dim3((unsigned)n), dim3(kThreads), 0, stream, d_qdesc, d_ids,
                       (double*)work);
  else
    hipLaunchKernelGGL(precond_init_kernel<true>, dim3((unsigned)n), dim3(kThreads), 0, stream, d_qdesc, d_ids,
                       (double2*)work);
  return yh::hip_status(hipGetLastError(), "svd precond init launch");
}

extern "C" int yastn_hip_svd_precond_finalize(void* U, void* S, void* Vh, const void* work, int dtype,
                                              const int64_t* d_blocks, const int64_t* d_qdesc, const int32_t* d_ids,
                                              int64_t n, int32_t* d_info, int sweeps, const int32_t* d_sweeps,
                                              hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  if (n == 0) return YH_OK;
  const dim3 g((unsigned)n), t(kThreads);
  if (dtype == YH_F64)
    hipLaunchKernelGGL(precond_finalize_kernel<false>, g, t, 0, stream, d_blocks, d_qdesc, d_ids,
                       (const double*)work, (double*)U, (double*)S, (double*)Vh, d_info, sweeps, d_sweeps);
  else
    hipLaunchKernelGGL(precond_finalize_kernel<true>, g, t, 0, stream, d_blocks, d_qdesc, d_ids,
                       (const double2*)work, (double2*)U, (double*)S, (double2*)Vh, d_info, sweeps, d_sweeps);
  const int st = yh::hip_status(hipGetLastError(), "svd precond finalize launch");
  if (st) return st;
  return dtype == YH_F64 ? launch_complete<false>(d_blocks, d_ids, n, U, S, Vh, stream, 1)
                         : launch_complete<true>(d_blocks, d_ids, n, U, S, Vh, stream, 1);
}

extern "C" int yastn_hip_check_qr_op_plan(const int64_t* h_qdesc, const int64_t* h_qblocks, int64_t nblocks,
                                          int dtype, int64_t a_numel, int64_t q_numel, int64_t r_numel,
                                          int64_t work_numel, int64_t lds_bytes) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  const int64_t isz = yh::itemsize_of(dtype);
  for (int64_t b = 0; b < nblocks; ++b) {
    const int64_t* w = h_qdesc + b * YH_QR_WORDS;
    const int64_t* bw = h_qblocks + b * YH_QROP_WORDS;
    const int64_t m = w[0], n = w[1], k = w[7];
    YH_CHECK_ARG(m >= 1 && n >= 1 && k == (m < n ? m : n), "qr block %lld: %lld x %lld, k %lld", (long long)b,
                 (long long)m, (long long)n, (long long)k);
    YH_CHECK_ARG(lds_bytes == 0 || m * YH_QR_NB * isz <= lds_bytes, "qr block %lld: panel of %lld rows exceeds LDS",
                 (long long)b, (long long)m);
    YH_CHECK_ARG(m <= kQrMaxRows, "qr block %lld: %lld rows (at most %lld)", (long long)b, (long long)m,
                 (long long)kQrMaxRows);
    const int64_t np = (k + YH_QR_NB - 1) / YH_QR_NB;
    const bool ok = w[2] >= 0 && w[2] + m * n <= work_numel && w[3] >= 0 && w[3] + m * k <= work_numel &&
                    w[4] >= 0 && w[4] + np * YH_QR_NB * YH_QR_NB <= work_numel && w[5] >= 0 &&
                    w[5] + m * k <= work_numel && w[6] >= 0 && w[6] + 2 * YH_QR_NB * n <= work_numel &&
                    bw[0] >= 0 && bw[0] + m * n <= a_numel && bw[1] >= 0 && bw[1] + m * k <= q_numel &&
                    bw[2] >= 0 && bw[2] + k * n <= r_numel;
    YH_CHECK_ARG(ok, "qr block %lld: offsets out of bounds", (long long)b);
  }
  return YH_OK;
}

extern "C" int yastn_hip_qr_init(const void* A, void* work, int dtype, const int64_t* d_qdesc,
                                 const int64_t* d_qblocks, int64_t nblocks, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  if (nblocks == 0) return YH_OK;
  if (dtype == YH_F64)
    hipLaunchKernelGGL(qr_op_init_kernel<false>, dim3((unsigned)nblocks), dim3(kThreads), 0, stream, d_qdesc,
                       d_qblocks, (const double*)A, (double*)work);
  else
    hipLaunchKernelGGL(qr_op_init_kernel<true>, dim3((unsigned)nblocks), dim3(kThreads), 0, stream, d_qdesc,
                       d_qblocks, (const double2*)A, (double2*)work);
  return yh::hip_status(hipGetLastError(), "qr init launch");
}

extern "C" int yastn_hip_qr_finalize(void* Q, void* R, const void* work, int dtype, const int64_t* d_qdesc,
                                     const int64_t* d_qblocks, int64_t nblocks, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  if (nblocks == 0) return YH_OK;
  if (dtype == YH_F64)
    hipLaunchKernelGGL(qr_op_finalize_kernel<false>, dim3((unsigned)nblocks), dim3(kThreads), 0, stream, d_qdesc,
                       d_qblocks, (const double*)work, (double*)Q, (double*)R);
  else
    hipLaunchKernelGGL(qr_op_finalize_kernel<true>, dim3((unsigned)nblocks), dim3(kThreads), 0, stream, d_qdesc,
                       d_qblocks, (const double2*)work, (double2*)Q, (double2*)R);
  return yh::hip_status(hipGetLastError(), "qr finalize launch");
}

extern "C" int yastn_hip_check_qr_adjoint_plan(const int64_t* h_words, int64_t nblocks, int dtype, int64_t m_numel,
                                               int64_t b_numel, int64_t ga_numel, int64_t r_numel) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  const int64_t isz = yh::itemsize_of(dtype);
  for (int64_t b = 0; b < nblocks; ++b) {
    const int64_t* w = h_words + b * YH_QRADJ_WORDS;
    const int64_t m = w[0], n = w[1], k = w[2];
    YH_CHECK_ARG(m >= 1 && n >= 1 && k == (m < n ? m : n) && k * isz <= 65536 && m <= 65535,
                 "qr adjoint block %lld: %lld x %lld", (long long)b, (long long)m, (long long)n);
    const bool ok = w[3] >= 0 && w[3] + k * k <= m_numel && w[4] >= 0 && w[4] + m * k <= b_numel && w[5] >= 0 &&
                    w[5] + m * n <= ga_numel && w[6] >= 0 && w[6] + k * n <= r_numel;
    YH_CHECK_ARG(ok, "qr adjoint block %lld: offsets out of bounds", (long long)b);
  }
  return YH_OK;
}

extern "C" int yastn_hip_qr_adjoint_core(void* C, const void* M1, const void* M2, int dtype, const int64_t* d_words,
                                         int64_t nblocks, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  if (nblocks == 0) return YH_OK;
  if (dtype == YH_F64)
    hipLaunchKernelGGL(qr_adj_core_kernel<false>, dim3((unsigned)nblocks), dim3(kThreads), 0, stream, d_words,
                       (const double*)M1, (const double*)M2, (double*)C);
  else
    hipLaunchKernelGGL(qr_adj_core_kernel<true>, dim3((unsigned)nblocks), dim3(kThreads), 0, stream, d_words,
                       (const double2*)M1, (const double2*)M2, (double2*)C);
  return yh::hip_status(hipGetLastError(), "qr adjoint core launch");
}

extern "C" int yastn_hip_qr_adjoint_trsm(void* gA, const void* B, const void* R, int dtype, const int64_t* d_words,
                                         int64_t nblocks, int64_t max_rows, int64_t max_k, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(max_rows >= 0 && max_rows <= 65535 && nblocks <= 65535, "qr adjoint grid");
  if (nblocks == 0 || max_rows == 0) return YH_OK;
  const int64_t lds = max_k * yh::itemsize_of(dtype);
  YH_CHECK_ARG(lds >= 0 && lds <= 65536, "qr adjoint row of %lld exceeds LDS", (long long)max_k);
  const dim3 g((unsigned)max_rows, (unsigned)nblocks);
  if (dtype == YH_F64)
    hipLaunchKernelGGL(qr_adj_trsm_kernel<false>, g, dim3(64), (size_t)lds, stream, d_words, (const double*)B,
                       (const double*)R, (double*)gA);
  else
    hipLaunchKernelGGL(qr_adj_trsm_kernel<true>, g, dim3(64), (size_t)lds, stream, d_words, (const double2*)B,
                       (const double2*)R, (double2*)gA);
  return yh::hip_status(hipGetLastError(), "qr adjoint trsm launch");
}
