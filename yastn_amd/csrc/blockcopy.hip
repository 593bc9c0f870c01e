// Box-gather engine: transpose_and_merge, unmerge and their adjoints (+ masks).
//
// A plan is a list of jobs; each job copies one box (rank <= YH_MAXR, arbitrary
// source strides, destination-contiguous last dimension) or zero-fills one box.
// All jobs of one backend call go out in a single launch; workgroups pick their
// tile through tile_job[], so the grid is the plan's tile count.
//
// Two tile kinds:
//  kind 0  YH_GATHER_TILE consecutive destination elements, decoded per element;
//          used when the source stride of the last dim is 1 (both sides coalesced)
//          or when the transposed dims are too small for a 32x32 tile.
//  kind 1  32x32 tile over (tdim, last dim) staged through LDS, so that the read is
//          coalesced along the source-contiguous dim and the write along the
//          destination-contiguous dim.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "common.h"

namespace yh {

static thread_local char g_err[512];

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_status(hipError_t e, const char* where) {
  if (e == hipSuccess) return YH_OK;
  set_error("%s: %s", where, hipGetErrorString(e));
  return YH_ERR_HIP;
}

}  // namespace yh

extern "C" const char* yastn_hip_last_error(void) { return yh::g_err; }
extern "C" int yastn_hip_abi_version(void) { return YH_ABI_VERSION; }

namespace {

constexpr int kThreads = 256;
constexpr int kTE = YH_TRANSPOSE_TILE;
constexpr int kShape = 6, kSrcStride = 6 + YH_MAXR, kDstStride = 6 + 2 * YH_MAXR;

struct JobView {
  int64_t src_off, dst_off;
  int rank, kind, tdim;
  int64_t tile_begin;
  int64_t shape[YH_MAXR], ss[YH_MAXR], ds[YH_MAXR];
};

__device__ inline void load_job(const int64_t* __restrict__ w, JobView& j) {
  j.src_off = w[0];
  j.dst_off = w[1];
  j.rank = (int)w[2];
  j.kind = (int)w[3];
  j.tdim = (int)w[4];
  j.tile_begin = w[5];
#pragma unroll
  for (int d = 0; d < YH_MAXR; ++d) {
    j.shape[d] = w[kShape + d];
    j.ss[d] = w[kSrcStride + d];
    j.ds[d] = w[kDstStride + d];
  }
}

template <typename T>
__device__ inline T zero_of();
template <>
__device__ inline double zero_of<double>() { return 0.0; }
template <>
__device__ inline double2 zero_of<double2>() { return make_double2(0.0, 0.0); }

template <typename T>
__global__ void __launch_bounds__(kThreads) gather_kernel(T* __restrict__ dst, const T* __restrict__ src,
                                                          const int64_t* __restrict__ jobs,
                                                          const int32_t* __restrict__ tile_job) {
  __shared__ T tile[kTE][kTE + 1];
  const int64_t t = blockIdx.x;
  const int jid = tile_job[t];
  JobView j;
  load_job(jobs + (int64_t)jid * YH_JOB_WORDS, j);
  const int64_t local = t - j.tile_begin;
  const int r = j.rank;

  if (j.kind == 0) {
    int64_t numel = 1;
    for (int d = 0; d < r; ++d) numel *= j.shape[d];
    const int64_t e0 = local * YH_GATHER_TILE;
    const int64_t e1 = min(numel, e0 + (int64_t)YH_GATHER_TILE);
    const int64_t L = j.shape[r - 1];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kThreads) {
      // decode the destination-order multi-index of e
      int64_t rem = e / L;
      const int64_t c = e - rem * L;
      int64_t so = c * j.ss[r - 1];
      int64_t dof = c;  // dst stride of the last dim is 1
      for (int d = r - 2; d >= 0; --d) {
        const int64_t q = rem / j.shape[d];
        const int64_t i = rem - q * j.shape[d];
        so += i * j.ss[d];
        dof += i * j.ds[d];
        rem = q;
      }
      dst[j.dst_off + dof] = (j.src_off >= 0) ? src[j.src_off + so] : zero_of<T>();
    }
    return;
  }

  // kind 1: 32x32 transpose tile over dims (k = tdim, jd = r-1)
  const int k = j.tdim;
  const int jd = r - 1;
  const int64_t nk = yh::cdiv(j.shape[k], kTE);
  const int64_t nj = yh::cdiv(j.shape[jd], kTE);
  const int64_t per_outer = nk * nj;
  int64_t outer = local / per_outer;
  const int64_t in_outer = local - outer * per_outer;
  const int64_t kt = in_outer / nj;
  const int64_t jt = in_outer - kt * nj;
  int64_t sbase = j.src_off, dbase = j.dst_off;
  for (int d = r - 2; d >= 0; --d) {
    if (d == k) continue;
    const int64_t q = outer / j.shape[d];
    const int64_t i = outer - q * j.shape[d];
    sbase += i * j.ss[d];
    dbase += i * j.ds[d];
    outer = q;
  }
  const int64_t k0 = kt * kTE, j0 = jt * kTE;
  const int64_t kk_lim = min((int64_t)kTE, j.shape[k] - k0);
  const int64_t jj_lim = min((int64_t)kTE, j.shape[jd] - j0);
  const int lane32 = threadIdx.x & (kTE - 1);
  const int row8 = threadIdx.x >> 5;  // 0..7
  // read: consecutive lanes walk the source-contiguous dim k
#pragma unroll
  for (int p = 0; p < kTE / 8; ++p) {
    const int jj = row8 + 8 * p;
    if (lane32 < kk_lim && jj < jj_lim)
      tile[jj][lane32] = src[sbase + (k0 + lane32) /* ss[k]==1 */ + (j0 + jj) * j.ss[jd]];
  }
  __syncthreads();
  // write: consecutive lanes walk the destination-contiguous dim
#pragma unroll
  for (int p = 0; p < kTE / 8; ++p) {
    const int kk = row8 + 8 * p;
    if (lane32 < jj_lim && kk < kk_lim) dst[dbase + (j0 + lane32) + (k0 + kk) * j.ds[k]] = tile[lane32][kk];
  }
}

}  // namespace

extern "C" int yastn_hip_check_gather_plan(const int64_t* h_jobs, int64_t njobs, const int32_t* h_tile_job,
                                           int64_t ntiles, int64_t src_numel, int64_t dst_numel) {
  YH_CHECK_ARG(njobs >= 0 && ntiles >= 0, "negative plan length");
  YH_CHECK_ARG(njobs == 0 || (h_jobs && h_tile_job), "null plan");
  int64_t expect_tile = 0;
  for (int64_t q = 0; q < njobs; ++q) {
    const int64_t* w = h_jobs + q * YH_JOB_WORDS;
    const int64_t src_off = w[0], dst_off = w[1];
    const int rank = (int)w[2], kind = (int)w[3], tdim = (int)w[4];
    YH_CHECK_ARG(rank >= 1 && rank <= YH_MAXR, "job %lld: rank %d", (long long)q, rank);
    YH_CHECK_ARG(kind == 0 || kind == 1, "job %lld: kind %d", (long long)q, kind);
    YH_CHECK_ARG(w[5] == expect_tile, "job %lld: tile_begin %lld != %lld", (long long)q, (long long)w[5],
                 (long long)expect_tile);
    YH_CHECK_ARG(w[kDstStride + rank - 1] == 1, "job %lld: destination last stride must be 1", (long long)q);
    int64_t numel = 1, smax = 0, dmax = 0;
    for (int d = 0; d < rank; ++d) {
      const int64_t n = w[kShape + d];
      YH_CHECK_ARG(n >= 1, "job %lld: shape[%d]=%lld", (long long)q, d, (long long)n);
      YH_CHECK_ARG(w[kSrcStride + d] >= 0 && w[kDstStride + d] >= 0, "job %lld: negative stride", (long long)q);
      numel *= n;
      smax += (n - 1) * w[kSrcStride + d];
      dmax += (n - 1) * w[kDstStride + d];
    }
    YH_CHECK_ARG(dst_off >= 0 && dst_off + dmax < dst_numel, "job %lld: destination out of bounds", (long long)q);
    YH_CHECK_ARG(src_off == -1 || (src_off >= 0 && src_off + smax < src_numel),
                 "job %lld: source out of bounds", (long long)q);
    int64_t tiles;
    if (kind == 0) {
      tiles = yh::cdiv(numel, YH_GATHER_TILE);
    } else {
      YH_CHECK_ARG(src_off >= 0 && tdim >= 0 && tdim < rank - 1 && w[kSrcStride + tdim] == 1,
                   "job %lld: bad transpose dim", (long long)q);
      tiles = yh::cdiv(w[kShape + tdim], YH_TRANSPOSE_TILE) * yh::cdiv(w[kShape + rank - 1], YH_TRANSPOSE_TILE) *
              (numel / (w[kShape + tdim] * w[kShape + rank - 1]));
    }
    for (int64_t t = expect_tile; t < expect_tile + tiles; ++t) {
      YH_CHECK_ARG(t < ntiles && h_tile_job[t] == q, "tile %lld does not belong to job %lld", (long long)t,
                   (long long)q);
    }
    expect_tile += tiles;
  }
  YH_CHECK_ARG(expect_tile == ntiles, "plan has %lld tiles, jobs cover %lld", (long long)ntiles,
               (long long)expect_tile);
  return YH_OK;
}

extern "C" int yastn_hip_block_gather(void* dst, int64_t dst_numel, const void* src, int64_t src_numel, int dtype,
                                      const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job,
                                      int64_t ntiles, hipStream_t stream) {
  (void)dst_numel;
  (void)src_numel;
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(ntiles >= 0 && ntiles < (int64_t)INT32_MAX, "tile count %lld", (long long)ntiles);
  if (ntiles == 0) return YH_OK;
  YH_CHECK_ARG(dst && d_jobs && d_tile_job && njobs > 0, "null pointer");
  const dim3 grid((unsigned)ntiles), block(kThreads);
  if (dtype == YH_F64)
    hipLaunchKernelGGL(gather_kernel<double>, grid, block, 0, stream, (double*)dst, (const double*)src, d_jobs,
                       d_tile_job);
  else
    hipLaunchKernelGGL(gather_kernel<double2>, grid, block, 0, stream, (double2*)dst, (const double2*)src,
                       d_jobs, d_tile_job);
  return yh::hip_status(hipGetLastError(), "block_gather launch");
}

extern "C" int yastn_hip_transpose_and_merge(void* newdata, int64_t Dsize, const void* data, int64_t data_numel,
                                             int dtype, const int64_t* d_jobs, int64_t njobs,
                                             const int32_t* d_tile_job, int64_t ntiles, hipStream_t stream) {
  return yastn_hip_block_gather(newdata, Dsize, data, data_numel, dtype, d_jobs, njobs, d_tile_job, ntiles,
                                stream);
}

extern "C" int yastn_hip_transpose_and_merge_adjoint(void* grad_src, int64_t src_numel, const void* grad_new,
                                                     int64_t Dsize, int dtype, const int64_t* d_jobs,
                                                     int64_t njobs, const int32_t* d_tile_job, int64_t ntiles,
                                                     hipStream_t stream) {
  return yastn_hip_block_gather(grad_src, src_numel, grad_new, Dsize, dtype, d_jobs, njobs, d_tile_job, ntiles,
                                stream);
}

extern "C" int yastn_hip_unmerge(void* newdata, const void* data, int64_t numel, int dtype, const int64_t* d_jobs,
                                 int64_t njobs, const int32_t* d_tile_job, int64_t ntiles, hipStream_t stream) {
  return yastn_hip_block_gather(newdata, numel, data, numel, dtype, d_jobs, njobs, d_tile_job, ntiles, stream);
}

extern "C" int yastn_hip_unmerge_adjoint(void* grad_src, const void* grad_new, int64_t numel, int dtype,
                                         const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job,
                                         int64_t ntiles, hipStream_t stream) {
  return yastn_hip_block_gather(grad_src, numel, grad_new, numel, dtype, d_jobs, njobs, d_tile_job, ntiles,
                                stream);
}

extern "C" int yastn_hip_apply_mask(void* newdata, int64_t Dsize, const void* data, int64_t data_numel, int dtype,
                                    const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job, int64_t ntiles,
                                    hipStream_t stream) {
  return yastn_hip_block_gather(newdata, Dsize, data, data_numel, dtype, d_jobs, njobs, d_tile_job, ntiles, stream);
}

extern "C" int yastn_hip_embed_mask(void* newdata, int64_t Dsize, const void* data, int64_t data_numel, int dtype,
                                    const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job, int64_t ntiles,
                                    hipStream_t stream) {
  return yastn_hip_block_gather(newdata, Dsize, data, data_numel, dtype, d_jobs, njobs, d_tile_job, ntiles, stream);
}

extern "C" int yastn_hip_transpose(void* newdata, const void* data, int64_t numel, int dtype, const int64_t* d_jobs,
                                   int64_t njobs, const int32_t* d_tile_job, int64_t ntiles, hipStream_t stream) {
  return yastn_hip_block_gather(newdata, numel, data, numel, dtype, d_jobs, njobs, d_tile_job, ntiles, stream);
}
