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
// Blocks whose W and V fit in 64 KiB go through an LDS-resident instance of the
// kernel; larger blocks use the caller's workspace in global memory.
#include <float.h>

#include "common.h"

namespace {

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

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
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

// One Hestenes rotation of the column pair (w_i, w_j) (length p) and of the matching accumulator
// columns (v_i, v_j) (length q), by one wave.  Returns true when the pair was rotated, i.e. when
// |w_i^H w_j| > tol ||w_i|| ||w_j|| and both squared norms exceed floor2.  J = [[c, s e^{i phi}], [-s e^{-i phi}, c]], e^{i phi} = g / |g|.
template <bool CPLX>
__device__ inline bool jacobi_pair(typename Num<CPLX>::T* wi, typename Num<CPLX>::T* wj, int64_t p,
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
  gi = wave_sum(gi);
  const double g = sqrt(gr * gr + gi * gi);
  // columns at the noise level (||w||^2 <= (eps ||A||_F)^2) carry no information beyond absolute
  // accuracy eps ||A||; rotating them against rounding noise would never converge
  if (!(g > tol * sqrt(a * b)) || a <= floor2 || b <= floor2) return false;
  const double zeta = (b - a) / (2.0 * g);
  const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + t * t);
  const double s = c * t;
  const double er = gr / g, ei = gi / g;
  // w_i' = c w_i - s e^{-i phi} w_j ; w_j' = s e^{i phi} w_i + c w_j
  for (int64_t x = lane; x < p; x += 64) {
    const T u = wi[x], v = wj[x];
    const double vr = er * N::re(v) + ei * N::im(v), vi = er * N::im(v) - ei * N::re(v);
    const double ur = er * N::re(u) - ei * N::im(u), ui = er * N::im(u) + ei * N::re(u);
    wi[x] = N::make(c * N::re(u) - s * vr, c * N::im(u) - s * vi);
    wj[x] = N::make(s * ur + c * N::re(v), s * ui + c * N::im(v));
  }
  for (int64_t x = lane; x < q; x += 64) {
    const T u = vi_[x], v = vj_[x];
    const double vr = er * N::re(v) + ei * N::im(v), vv = er * N::im(v) - ei * N::re(v);
    const double ur = er * N::re(u) - ei * N::im(u), ui = er * N::im(u) + ei * N::re(u);
    vi_[x] = N::make(c * N::re(u) - s * vr, c * N::im(u) - s * vv);
    vj_[x] = N::make(s * ur + c * N::re(v), s * ui + c * N::im(v));
  }
  return true;
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
template <bool CPLX, int MODE>
__device__ void finalize_block(const BlockDesc& d, const typename Num<CPLX>::T* W, const typename Num<CPLX>::T* V,
                               int64_t p, int64_t q, bool tr, double mu, typename Num<CPLX>::T* __restrict__ U,
                               double* __restrict__ S, typename Num<CPLX>::T* __restrict__ Vh) {
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
    if (MODE == 0)
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
        if (!jacobi_pair<CPLX>(W + (int64_t)i * p, W + (int64_t)j * p, p, V + (int64_t)i * q, V + (int64_t)j * q, q,
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

template <bool CPLX>
__global__ void __launch_bounds__(kPanelThreads)
    panel_round_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ jobs,
                       typename Num<CPLX>::T* __restrict__ work, const double* __restrict__ floors,
                       int32_t* __restrict__ flags, int b) {
  typedef Num<CPLX> N;
  typedef typename N::T T;
  extern __shared__ double smem_raw[];
  __shared__ int s_rot;
  const int32_t* jb = jobs + 3 * (int64_t)blockIdx.x;
  const BlockDesc d = load_desc(blocks + (int64_t)jb[0] * YH_SVD_WORDS);
  const bool tr = d.m < d.n;
  const int64_t p = tr ? d.n : d.m, q = tr ? d.m : d.n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t c0i = (int64_t)jb[1] * b, c0j = (int64_t)jb[2] * b;
  const int64_t ci = min((int64_t)b, q - c0i), cj = min((int64_t)b, q - c0j);
  const int nc = (int)(ci + cj);
  const int nq = 2 * b;  // leading dimension of Q
  T* W = work + d.work_off;
  T* V = W + p * q;
  T* Wp = reinterpret_cast<T*>(smem_raw);
  T* Q = Wp + (int64_t)nq * p;
  auto gcol = [&](int64_t l) { return l < ci ? c0i + l : c0j + (l - ci); };
  for (int64_t e = tid; e < (int64_t)nc * p; e += kPanelThreads) {
    const int64_t l = e / p, r = e - l * p;
    Wp[e] = W[gcol(l) * p + r];
  }
  for (int e = tid; e < nc * nq; e += kPanelThreads) {
    const int l = e / nq, k = e - l * nq;
    Q[e] = N::make(l == k ? 1.0 : 0.0, 0.0);
  }
  if (tid == 0) s_rot = 0;
  __syncthreads();
  const double tol = DBL_EPSILON * sqrt((double)(p > 1 ? p : 1));
  const double floor2 = floors[jb[0]];
  const int ne = nc + (nc & 1);
  for (int r = 0; r < ne - 1; ++r) {
    for (int k = wave; k < ne / 2; k += kPanelThreads / 64) {
      int i, j;
      pair_of(ne, r, k, i, j);
      if (j >= nc) continue;
      if (jacobi_pair<CPLX>(Wp + (int64_t)i * p, Wp + (int64_t)j * p, p, Q + i * nq, Q + j * nq, nc, tol, floor2,
                            lane) &&
          lane == 0)
        s_rot = 1;
    }
    __syncthreads();
  }
  for (int64_t e = tid; e < (int64_t)nc * p; e += kPanelThreads) {
    const int64_t l = e / p, r = e - l * p;
    W[gcol(l) * p + r] = Wp[e];
  }
  __syncthreads();
  // V panel columns (q x nc) staged where the W panel was (q <= p), then V <- V Q
  T* Vp = Wp;
  for (int64_t e = tid; e < (int64_t)nc * q; e += kPanelThreads) {
    const int64_t l = e / q, r = e - l * q;
    Vp[e] = V[gcol(l) * q + r];
  }
  __syncthreads();
  for (int64_t e = tid; e < (int64_t)nc * q; e += kPanelThreads) {
    const int64_t c = e / q, r = e - c * q;
    double accr = 0.0, acci = 0.0;
    for (int k = 0; k < nc; ++k) {
      const T v = Vp[(int64_t)k * q + r], w = Q[c * nq + k];
      accr += N::re(v) * N::re(w) - N::im(v) * N::im(w);
      acci += N::re(v) * N::im(w) + N::im(v) * N::re(w);
    }
    V[gcol(c) * q + r] = N::make(accr, acci);
  }
  if (tid == 0) flags[blockIdx.x] = s_rot;
}

template <bool CPLX, int MODE>
__global__ void __launch_bounds__(kPanelThreads)
    panel_finalize_kernel(const int64_t* __restrict__ blocks, const int32_t* __restrict__ ids,
                          const typename Num<CPLX>::T* __restrict__ work, const double* __restrict__ mu,
                          typename Num<CPLX>::T* __restrict__ U, double* __restrict__ S,
                          typename Num<CPLX>::T* __restrict__ Vh, int32_t* __restrict__ info, int sweeps) {
  const int bidx = ids[blockIdx.x];
  const BlockDesc d = load_desc(blocks + (int64_t)bidx * YH_SVD_WORDS);
  const bool tr = (MODE == 0) && (d.m < d.n);
  const int64_t p = tr ? d.n : d.m, q = tr ? d.m : d.n;
  const typename Num<CPLX>::T* W = work + d.work_off;
  finalize_block<CPLX, MODE>(d, W, W + p * q, p, q, tr, mu[bidx], U, S, Vh);
  if (threadIdx.x == 0) info[bidx] = sweeps;
}

template <bool CPLX, int MODE>
int launch_jacobi(const int64_t* d_blocks, const int32_t* d_order, int64_t n_lds, int64_t lds_bytes,
                  int64_t nblocks, const void* A, void* U, void* S, void* Vh, void* work, int32_t* info,
                  int max_sweeps, hipStream_t stream) {
  typedef typename Num<CPLX>::T T;
  if (n_lds > 0) {
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
  return YH_OK;
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

extern "C" int yastn_hip_svd_panel_round(void* work, const double* floors, int dtype, const int64_t* d_blocks,
                                         const int32_t* d_jobs, int64_t njobs, int b, int64_t lds_bytes,
                                         int32_t* d_flags, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(b >= 1 && b <= 64 && lds_bytes > 0, "panel width %d / lds %lld", b, (long long)lds_bytes);
  if (njobs == 0) return YH_OK;
  const dim3 g((unsigned)njobs), t(kPanelThreads);
  if (dtype == YH_F64) {
    hipFuncSetAttribute((const void*)panel_round_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds_bytes);
    hipLaunchKernelGGL(panel_round_kernel<false>, g, t, (size_t)lds_bytes, stream, d_blocks, d_jobs,
                       (double*)work, floors, d_flags, b);
  } else {
    hipFuncSetAttribute((const void*)panel_round_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds_bytes);
    hipLaunchKernelGGL(panel_round_kernel<true>, g, t, (size_t)lds_bytes, stream, d_blocks, d_jobs,
                       (double2*)work, floors, d_flags, b);
  }
  return yh::hip_status(hipGetLastError(), "svd panel round launch");
}

extern "C" int yastn_hip_svd_panel_finalize(void* U, void* S, void* Vh, const void* work, const double* mu,
                                            int dtype, int mode, const int64_t* d_blocks, const int32_t* d_ids,
                                            int64_t nb, int32_t* d_info, int sweeps, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(mode == 0 || mode == 1, "mode %d", mode);
  if (nb == 0) return YH_OK;
  const dim3 g((unsigned)nb), t(kPanelThreads);
  if (dtype == YH_F64) {
    if (mode == 0)
      hipLaunchKernelGGL((panel_finalize_kernel<false, 0>), g, t, 0, stream, d_blocks, d_ids, (const double*)work,
                         mu, (double*)U, (double*)S, (double*)Vh, d_info, sweeps);
    else
      hipLaunchKernelGGL((panel_finalize_kernel<false, 1>), g, t, 0, stream, d_blocks, d_ids, (const double*)work,
                         mu, (double*)U, (double*)S, (double*)Vh, d_info, sweeps);
  } else {
    if (mode == 0)
      hipLaunchKernelGGL((panel_finalize_kernel<true, 0>), g, t, 0, stream, d_blocks, d_ids, (const double2*)work,
                         mu, (double2*)U, (double*)S, (double2*)Vh, d_info, sweeps);
    else
      hipLaunchKernelGGL((panel_finalize_kernel<true, 1>), g, t, 0, stream, d_blocks, d_ids, (const double2*)work,
                         mu, (double2*)U, (double*)S, (double2*)Vh, d_info, sweeps);
  }
  return yh::hip_status(hipGetLastError(), "svd panel finalize launch");
}
