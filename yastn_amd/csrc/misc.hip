// Small per-block kernels around the decompositions:
//  - fix_svd_signs: phase of the largest (2^40-quantised) entry of each U column
//  - the k x k core of the SVD adjoint (Lorentzian-broadened F, G terms)
//  - the n x n core of the Hermitian eigen adjoint
// The surrounding products (U^H gU, Vh gVh^H, U J Vh, ...) run through the grouped GEMM.
#include <algorithm>

#include "common.h"

namespace {

constexpr int kThreads = 256;

template <bool CPLX>
struct Z;
template <>
struct Z<false> {
  typedef double T;
  static __device__ inline double re(double x) { return x; }
  static __device__ inline double im(double) { return 0.0; }
  static __device__ inline double mk(double r, double) { return r; }
  static __device__ inline double absv(double x) { return fabs(x); }
};
template <>
struct Z<true> {
  typedef double2 T;
  static __device__ inline double re(double2 x) { return x.x; }
  static __device__ inline double im(double2 x) { return x.y; }
  static __device__ inline double2 mk(double r, double i) { return make_double2(r, i); }
  static __device__ inline double absv(double2 x) { return hypot(x.x, x.y); }
};

// Work split: kSplit workgroups per block (grid = nblocks * kSplit), so that one large block does not
// serialise on a single CU; a workgroup with nothing to do exits at once.
constexpr int kSplit = 16;

// fix_svd_signs.  Blocks given as SVD plan words (0 a_off, 1 m, 2 n, 3 u_off, 4 s_off, 5 v_off);
// U block m x k, Vh block k x n.  A workgroup takes 64-column chunks of U: lane = column (coalesced
// row reads), wave = stripe of rows; per column the argmax of int64(|U[r, j]| * 2^40) with ties to the
// smallest row, then the phase is applied to U[:, j] and Vh[j, :].
// With piv != nullptr only the pivot positions are written: piv[s_off + j] = flat index of the pivot
// of column j in U (the caller applies the phases with differentiable tensor ops).
template <bool CPLX>
__global__ void __launch_bounds__(kThreads) fix_signs_kernel(typename Z<CPLX>::T* __restrict__ U,
                                                             typename Z<CPLX>::T* __restrict__ Vh,
                                                             const int64_t* __restrict__ blocks,
                                                             int64_t* __restrict__ piv) {
  typedef Z<CPLX> N;
  typedef typename N::T T;
  constexpr int kW = kThreads / 64;
  __shared__ int64_t s_best[kW][64];
  __shared__ int64_t s_row[kW][64];
  __shared__ double s_ph[2][64];
  const int64_t* w = blocks + (int64_t)(blockIdx.x / kSplit) * YH_SVD_WORDS;
  const int split = blockIdx.x % kSplit;
  const int64_t m = w[1], n = w[2], u_off = w[3], s_off = w[4], v_off = w[5];
  const int64_t k = m < n ? m : n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t c0 = (int64_t)split * 64; c0 < k; c0 += (int64_t)kSplit * 64) {
    const int64_t j = c0 + lane;
    int64_t best = -1, brow = 0;
    if (j < k) {
      for (int64_t r = wave; r < m; r += kW) {
        const int64_t qv = (int64_t)(N::absv(U[u_off + r * k + j]) * 1099511627776.0);
        if (qv > best) {
          best = qv;
          brow = r;
        }
      }
    }
    s_best[wave][lane] = best;
    s_row[wave][lane] = brow;
    __syncthreads();
    if (wave == 0) {
      for (int x = 1; x < kW; ++x) {
        const int64_t ob = s_best[x][lane], orow = s_row[x][lane];
        if (ob > best || (ob == best && orow < brow)) {
          best = ob;
          brow = orow;
        }
      }
      if (j < k) {
        if (piv) {
          piv[s_off + j] = u_off + brow * k + j;
        } else {
          const T pv = U[u_off + brow * k + j];
          const double a = N::absv(pv);
          s_ph[0][lane] = N::re(pv) / a;
          s_ph[1][lane] = N::im(pv) / a;
        }
      }
    }
    __syncthreads();
    if (!piv) {
      const int64_t nc = (k - c0) < 64 ? (k - c0) : 64;
      if (lane < nc) {
        const double pr = s_ph[0][lane], pi = s_ph[1][lane];
        for (int64_t r = wave; r < m; r += kW) {  // u * conj(phase)
          const T u = U[u_off + r * k + j];
          U[u_off + r * k + j] = N::mk(N::re(u) * pr + N::im(u) * pi, N::im(u) * pr - N::re(u) * pi);
        }
      }
      for (int64_t l = wave; l < nc; l += kW) {  // Vh row c0 + l times phase
        const double pr = s_ph[0][l], pi = s_ph[1][l];
        T* row = Vh + v_off + (c0 + l) * n;
        for (int64_t c = lane; c < n; c += 64) {
          const T v = row[c];
          row[c] = N::mk(N::re(v) * pr - N::im(v) * pi, N::im(v) * pr + N::re(v) * pi);
        }
      }
    }
    __syncthreads();
  }
}

__device__ inline double lor(double x, double eps) { return x / (x * x + eps); }

// SVD adjoint core.  desc words: 0 m, 1 n, 2 k, 3 u_off, 4 s_off, 5 v_off, 6 kk_off, 7 t_off.
// J = 1/2 (F+G) o (X - X^H) + 1/2 (F-G) o (Y - Y^H) + diag(gS) + i diag(Im X_jj sinv_j)   [complex]
//     - [m>k] X diag(sinv) - [n>k] diag(sinv) Y^H
// T = [m>k] gU diag(sinv) | [n>k] diag(sinv) gVh | 0
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    svd_adj_core_kernel(typename Z<CPLX>::T* __restrict__ J, typename Z<CPLX>::T* __restrict__ T_out,
                        const typename Z<CPLX>::T* __restrict__ X, const typename Z<CPLX>::T* __restrict__ Y,
                        const double* __restrict__ S, const double* __restrict__ gS,
                        const typename Z<CPLX>::T* __restrict__ gU, const typename Z<CPLX>::T* __restrict__ gVh,
                        const int64_t* __restrict__ descs, double eps, int has_gS, int has_gU, int has_gV) {
  typedef Z<CPLX> N;
  typedef typename N::T T;
  const int64_t* w = descs + (int64_t)(blockIdx.x / kSplit) * 8;
  const int64_t m = w[0], n = w[1], k = w[2], u_off = w[3], s_off = w[4], v_off = w[5], kk = w[6], t_off = w[7];
  const double* s = S + s_off;
  const int64_t first = (int64_t)(blockIdx.x % kSplit) * kThreads + threadIdx.x, step = (int64_t)kSplit * kThreads;
  for (int64_t e = first; e < k * k; e += step) {
    const int64_t i = e / k, j = e - i * k;
    double jr = 0.0, ji = 0.0;
    const double sinv_j = lor(s[j], eps), sinv_i = lor(s[i], eps);
    if (i != j) {
      const double f = lor(s[j] - s[i], eps), g = lor(s[j] + s[i], eps);
      if (has_gU) {
        const T xij = X[kk + i * k + j], xji = X[kk + j * k + i];
        // X - X^H at (i, j) = x_ij - conj(x_ji)
        const double dr = N::re(xij) - N::re(xji), di = N::im(xij) + N::im(xji);
        jr += 0.5 * (f + g) * dr;
        ji += 0.5 * (f + g) * di;
      }
      if (has_gV) {
        const T yij = Y[kk + i * k + j], yji = Y[kk + j * k + i];
        const double dr = N::re(yij) - N::re(yji), di = N::im(yij) + N::im(yji);
        jr += 0.5 * (f - g) * dr;
        ji += 0.5 * (f - g) * di;
      }
    } else {
      if (has_gS) jr += gS[s_off + i];
      if (CPLX && has_gU) ji += N::im(X[kk + i * k + i]) * sinv_i;
    }
    if (has_gU && m > k) {  // - X diag(sinv): (i, j) -> x_ij sinv_j
      const T xij = X[kk + i * k + j];
      jr -= N::re(xij) * sinv_j;
      ji -= N::im(xij) * sinv_j;
    }
    if (has_gV && n > k) {  // - diag(sinv) Y^H: (i, j) -> sinv_i conj(y_ji)
      const T yji = Y[kk + j * k + i];
      jr -= sinv_i * N::re(yji);
      ji += sinv_i * N::im(yji);
    }
    J[kk + i * k + j] = N::mk(jr, ji);
  }
  if (m >= n) {  // T is m x k
    for (int64_t e = first; e < m * k; e += step) {
      const int64_t jcol = e % k;
      T v = N::mk(0.0, 0.0);
      if (has_gU && m > k) {
        const T g = gU[u_off + e];
        const double sv = lor(s[jcol], eps);
        v = N::mk(N::re(g) * sv, N::im(g) * sv);
      }
      T_out[t_off + e] = v;
    }
  } else {  // T is k x n
    for (int64_t e = first; e < k * n; e += step) {
      const int64_t irow = e / n;
      T v = N::mk(0.0, 0.0);
      if (has_gV) {
        const T g = gVh[v_off + e];
        const double sv = lor(s[irow], eps);
        v = N::mk(N::re(g) * sv, N::im(g) * sv);
      }
      T_out[t_off + e] = v;
    }
  }
}

// eigh adjoint core.  desc words: 0 n, 1 s_off, 2 kk_off.
// Xo = (X - X^H)/2 / E, E_ij = w_j - w_i (E_ii = 1); diag(Xo) = gS when given.
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    eigh_adj_core_kernel(typename Z<CPLX>::T* __restrict__ Xo, const typename Z<CPLX>::T* __restrict__ X,
                         const double* __restrict__ S, const double* __restrict__ gS,
                         const int64_t* __restrict__ descs, int has_gS, int has_gU) {
  typedef Z<CPLX> N;
  const int64_t* w = descs + (int64_t)(blockIdx.x / kSplit) * 8;
  const int64_t n = w[0], s_off = w[1], kk = w[2];
  const int64_t first = (int64_t)(blockIdx.x % kSplit) * kThreads + threadIdx.x, step = (int64_t)kSplit * kThreads;
  for (int64_t e = first; e < n * n; e += step) {
    const int64_t i = e / n, j = e - i * n;
    double r = 0.0, im = 0.0;
    if (has_gU) {
      const typename N::T xij = X[kk + i * n + j], xji = X[kk + j * n + i];
      r = 0.5 * (N::re(xij) - N::re(xji));
      im = 0.5 * (N::im(xij) + N::im(xji));
      const double E = (i == j) ? 1.0 : (S[s_off + j] - S[s_off + i]);
      r /= E;
      im /= E;
    }
    if (i == j && has_gS) {
      r = gS[s_off + i];
      im = 0.0;
    }
    Xo[kk + e] = N::mk(r, im);
  }
}

// dot_diag: every output block is a dense block B scaled along one axis by a diagonal block A
// (a vector).  Job words (YH_DIAG_WORDS int64): 0 out_off, 1 a_off, 2 b_off, 3 n (block elements),
// 4 inner (product of the dimensions after the axis), 5 D (dimension of the axis).  Element e of the
// block takes A[a_off + (e / inner) % D].  grid = (chunks, jobs); a workgroup strides over its block.
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    dot_diag_kernel(typename Z<CPLX>::T* __restrict__ out, const typename Z<CPLX>::T* __restrict__ A,
                    const typename Z<CPLX>::T* __restrict__ B, const int64_t* __restrict__ jobs, int conj_a) {
  typedef Z<CPLX> N;
  const int64_t* w = jobs + (int64_t)blockIdx.y * YH_DIAG_WORDS;
  const int64_t out_off = w[0], a_off = w[1], b_off = w[2], n = w[3], inner = w[4], D = w[5];
  const double sa = conj_a ? -1.0 : 1.0;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kThreads) {
    const typename N::T a = A[a_off + (e / inner) % D], b = B[b_off + e];
    const double ar = N::re(a), ai = sa * N::im(a), br = N::re(b), bi = N::im(b);
    out[out_off + e] = N::mk(ar * br - ai * bi, ar * bi + ai * br);
  }
}

__device__ inline double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Adjoint of dot_diag with respect to the diagonal: gA[a_off + i] = sum over the jobs of one group
// (all jobs reading the same A block, in list order) of sum_e g[out_off + e] conj(B[b_off + e]) over the
// elements e with (e / inner) % D == i.  Group words: 0 a_off, 1 D, 2 first index into job_ids,
// 3 count.  One wave per (group, i): 4 waves per workgroup, grid = (ceil(max D / 4), groups).  Lanes
// take the block's elements of index i in order and the wave sum is a fixed butterfly, so the result is
// deterministic.
template <bool CPLX>
__global__ void __launch_bounds__(kThreads)
    dot_diag_adj_a_kernel(typename Z<CPLX>::T* __restrict__ gA, const typename Z<CPLX>::T* __restrict__ g,
                          const typename Z<CPLX>::T* __restrict__ B, const int64_t* __restrict__ jobs,
                          const int64_t* __restrict__ groups, const int64_t* __restrict__ job_ids) {
  typedef Z<CPLX> N;
  const int64_t* gw = groups + (int64_t)blockIdx.y * 4;
  const int64_t a_off = gw[0], D = gw[1], first = gw[2], count = gw[3];
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (i >= D) return;  // wave-uniform
  double accr = 0.0, acci = 0.0;
  for (int64_t q = 0; q < count; ++q) {
    const int64_t* w = jobs + job_ids[first + q] * YH_DIAG_WORDS;
    const int64_t out_off = w[0], b_off = w[2], n = w[3], inner = w[4];
    const int64_t per = n / D;  // elements of index i in this block
    for (int64_t k = lane; k < per; k += 64) {
      const int64_t e = (k / inner) * D * inner + i * inner + k % inner;
      const typename N::T x = g[out_off + e], b = B[b_off + e];
      accr += N::re(x) * N::re(b) + N::im(x) * N::im(b);
      acci += N::im(x) * N::re(b) - N::re(x) * N::im(b);
    }
  }
  accr = wave_sum64(accr);
  acci = wave_sum64(acci);
  if (lane == 0) gA[a_off + i] = N::mk(accr, acci);
}

}  // namespace

extern "C" int yastn_hip_check_dot_diag_plan(const int64_t* h_jobs, int64_t njobs, const int64_t* h_groups,
                                             int64_t ngroups, const int64_t* h_job_ids, int64_t nids,
                                             int64_t out_numel, int64_t a_numel, int64_t b_numel) {
  for (int64_t j = 0; j < njobs; ++j) {
    const int64_t* w = h_jobs + j * YH_DIAG_WORDS;
    const int64_t out_off = w[0], a_off = w[1], b_off = w[2], n = w[3], inner = w[4], D = w[5];
    YH_CHECK_ARG(n >= 1 && inner >= 1 && D >= 1 && n % (inner * D) == 0, "dot_diag job %lld: shape", (long long)j);
    YH_CHECK_ARG(out_off >= 0 && out_off + n <= out_numel, "dot_diag job %lld: output range", (long long)j);
    YH_CHECK_ARG(b_off >= 0 && b_off + n <= b_numel, "dot_diag job %lld: dense operand range", (long long)j);
    YH_CHECK_ARG(a_off >= 0 && a_off + D <= a_numel, "dot_diag job %lld: diagonal range", (long long)j);
  }
  for (int64_t gi = 0; gi < ngroups; ++gi) {
    const int64_t* gw = h_groups + gi * 4;
    YH_CHECK_ARG(gw[2] >= 0 && gw[3] >= 1 && gw[2] + gw[3] <= nids, "dot_diag group %lld: job list", (long long)gi);
    YH_CHECK_ARG(gw[0] >= 0 && gw[1] >= 1 && gw[0] + gw[1] <= a_numel, "dot_diag group %lld: range", (long long)gi);
    for (int64_t q = gw[2]; q < gw[2] + gw[3]; ++q) {
      YH_CHECK_ARG(h_job_ids[q] >= 0 && h_job_ids[q] < njobs, "dot_diag group %lld: job id", (long long)gi);
      const int64_t* w = h_jobs + h_job_ids[q] * YH_DIAG_WORDS;
      YH_CHECK_ARG(w[1] == gw[0] && w[5] == gw[1], "dot_diag group %lld: job of another block", (long long)gi);
    }
  }
  return YH_OK;
}

extern "C" int yastn_hip_dot_diag(void* out, const void* A, const void* B, int dtype, const int64_t* d_jobs,
                                  int64_t njobs, int64_t max_n, int conj_a, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(njobs <= 65535, "dot_diag: %lld jobs", (long long)njobs);
  if (njobs == 0 || max_n <= 0) return YH_OK;
  const dim3 g((unsigned)std::min<int64_t>(yh::cdiv(max_n, kThreads), 1024), (unsigned)njobs);
  if (dtype == YH_F64)
    hipLaunchKernelGGL(dot_diag_kernel<false>, g, dim3(kThreads), 0, stream, (double*)out, (const double*)A,
                       (const double*)B, d_jobs, conj_a);
  else
    hipLaunchKernelGGL(dot_diag_kernel<true>, g, dim3(kThreads), 0, stream, (double2*)out, (const double2*)A,
                       (const double2*)B, d_jobs, conj_a);
  return yh::hip_status(hipGetLastError(), "dot_diag launch");
}

extern "C" int yastn_hip_dot_diag_adjoint_a(void* gA, const void* g, const void* B, int dtype, const int64_t* d_jobs,
                                            const int64_t* d_groups, int64_t ngroups, const int64_t* d_job_ids,
                                            int64_t max_D, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(ngroups <= 65535, "dot_diag adjoint: %lld groups", (long long)ngroups);
  if (ngroups == 0 || max_D <= 0) return YH_OK;
  const dim3 grid((unsigned)yh::cdiv(max_D, kThreads / 64), (unsigned)ngroups);
  if (dtype == YH_F64)
    hipLaunchKernelGGL(dot_diag_adj_a_kernel<false>, grid, dim3(kThreads), 0, stream, (double*)gA, (const double*)g,
                       (const double*)B, d_jobs, d_groups, d_job_ids);
  else
    hipLaunchKernelGGL(dot_diag_adj_a_kernel<true>, grid, dim3(kThreads), 0, stream, (double2*)gA,
                       (const double2*)g, (const double2*)B, d_jobs, d_groups, d_job_ids);
  return yh::hip_status(hipGetLastError(), "dot_diag adjoint launch");
}

extern "C" int yastn_hip_fix_svd_signs(void* U, void* Vh, int dtype, const int64_t* d_blocks, int64_t nblocks,
                                       hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  if (nblocks == 0) return YH_OK;
  if (dtype == YH_F64)
    hipLaunchKernelGGL(fix_signs_kernel<false>, dim3((unsigned)(nblocks * kSplit)), dim3(kThreads), 0, stream, (double*)U,
                       (double*)Vh, d_blocks, (int64_t*)nullptr);
  else
    hipLaunchKernelGGL(fix_signs_kernel<true>, dim3((unsigned)(nblocks * kSplit)), dim3(kThreads), 0, stream, (double2*)U,
                       (double2*)Vh, d_blocks, (int64_t*)nullptr);
  return yh::hip_status(hipGetLastError(), "fix_svd_signs launch");
}

extern "C" int yastn_hip_svd_sign_pivots(const void* U, int64_t* piv, int dtype, const int64_t* d_blocks,
                                         int64_t nblocks, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(piv != nullptr, "null pivot buffer");
  if (nblocks == 0) return YH_OK;
  if (dtype == YH_F64)
    hipLaunchKernelGGL(fix_signs_kernel<false>, dim3((unsigned)(nblocks * kSplit)), dim3(kThreads), 0, stream,
                       (double*)const_cast<void*>(U), (double*)nullptr, d_blocks, piv);
  else
    hipLaunchKernelGGL(fix_signs_kernel<true>, dim3((unsigned)(nblocks * kSplit)), dim3(kThreads), 0, stream,
                       (double2*)const_cast<void*>(U), (double2*)nullptr, d_blocks, piv);
  return yh::hip_status(hipGetLastError(), "svd_sign_pivots launch");
}

extern "C" int yastn_hip_svd_adjoint_core(void* J, void* T, const void* X, const void* Y, const void* S,
                                          const void* gS, const void* gU, const void* gVh, int dtype,
                                          const int64_t* d_descs, int64_t nblocks, double eps, int has_gS,
                                          int has_gU, int has_gV, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  if (nblocks == 0) return YH_OK;
  if (dtype == YH_F64)
    hipLaunchKernelGGL(svd_adj_core_kernel<false>, dim3((unsigned)(nblocks * kSplit)), dim3(kThreads), 0, stream, (double*)J,
                       (double*)T, (const double*)X, (const double*)Y, (const double*)S, (const double*)gS,
                       (const double*)gU, (const double*)gVh, d_descs, eps, has_gS, has_gU, has_gV);
  else
    hipLaunchKernelGGL(svd_adj_core_kernel<true>, dim3((unsigned)(nblocks * kSplit)), dim3(kThreads), 0, stream, (double2*)J,
                       (double2*)T, (const double2*)X, (const double2*)Y, (const double*)S, (const double*)gS,
                       (const double2*)gU, (const double2*)gVh, d_descs, eps, has_gS, has_gU, has_gV);
  return yh::hip_status(hipGetLastError(), "svd_adjoint_core launch");
}

extern "C" int yastn_hip_eigh_adjoint_core(void* Xo, const void* X, const void* S, const void* gS, int dtype,
                                           const int64_t* d_descs, int64_t nblocks, int has_gS, int has_gU,
                                           hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  if (nblocks == 0) return YH_OK;
  if (dtype == YH_F64)
    hipLaunchKernelGGL(eigh_adj_core_kernel<false>, dim3((unsigned)(nblocks * kSplit)), dim3(kThreads), 0, stream,
                       (double*)Xo, (const double*)X, (const double*)S, (const double*)gS, d_descs, has_gS, has_gU);
  else
    hipLaunchKernelGGL(eigh_adj_core_kernel<true>, dim3((unsigned)(nblocks * kSplit)), dim3(kThreads), 0, stream,
                       (double2*)Xo, (const double2*)X, (const double*)S, (const double*)gS, d_descs, has_gS,
                       has_gU);
  return yh::hip_status(hipGetLastError(), "eigh_adjoint_core launch");
}
