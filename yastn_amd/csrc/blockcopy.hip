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
constexpr int kBuf = YH_JOB_BUFFER_WORD;  // 0: (dst0, src0), 1: (dst1, src1)
constexpr int kRun = YH_GATHER_TILE / kThreads;  // elements per thread in a kind-0 tile

template <typename T>
__device__ inline T zero_of();
template <>
__device__ inline double zero_of<double>() { return 0.0; }
template <>
__device__ inline double2 zero_of<double2>() { return make_double2(0.0, 0.0); }

// The job words are read in place (wave-uniform, so they live in scalar registers), with the loops
// over dimensions unrolled to YH_MAXR so that no per-thread array is indexed at run time; the
// multi-index arithmetic is 32-bit (a box holds fewer than 2^31 elements, checked on the host), the
// offsets 64-bit.  Every load of a thread is issued before the first wait: loads sit outside branches
// (an index past the end of the box reads the box's first element, whose value is then not stored),
// since a load under a per-lane branch gets its own wait.
// Kind-0 tile of a rank-R box (R >= 2): the multi-index of the thread's first element e (divisions
// once per thread), then each step of kThreads elements adds the uniform digits of kThreads in the box's
// mixed radix, with one carry per dimension.
template <typename T, int R>
__device__ __forceinline__ void gather_box(T* __restrict__ dst, const T* __restrict__ src, const int64_t* __restrict__ w,
                                           uint32_t e, uint32_t e1, int64_t src_off, int64_t dst_off) {
  uint32_t n[R], idx[R], dig[R];
  int64_t ss[R], ds[R];
#pragma unroll
  for (int d = 0; d < R; ++d) {
    n[d] = (uint32_t)w[kShape + d];
    ss[d] = w[kSrcStride + d];
    ds[d] = w[kDstStride + d];
  }
  uint32_t rem = e, dr = kThreads;
#pragma unroll
  for (int d = R - 1; d >= 0; --d) {
    const uint32_t q = rem / n[d], qd = dr / n[d];
    idx[d] = rem - q * n[d];
    dig[d] = dr - qd * n[d];
    rem = q;
    dr = qd;
  }
  T v[kRun];
  int64_t dofs[kRun];
#pragma unroll
  for (int i = 0; i < kRun; ++i) {
    if (i > 0) {
      uint32_t carry = 0;
#pragma unroll
      for (int d = R - 1; d >= 0; --d) {
        const uint32_t x = idx[d] + dig[d] + carry;
        carry = x >= n[d];
        idx[d] = carry ? x - n[d] : x;
      }
    }
    int64_t so = 0, dof = 0;
#pragma unroll
    for (int d = 0; d < R; ++d) {
      so += (int64_t)idx[d] * ss[d];
      dof += (int64_t)idx[d] * ds[d];
    }
    if (src_off >= 0) v[i] = src[src_off + (e + i * kThreads < e1 ? so : 0)];
    dofs[i] = dof;
  }
#pragma unroll
  for (int i = 0; i < kRun; ++i)
    if (e + i * kThreads < e1) dst[dst_off + dofs[i]] = (src_off >= 0) ? v[i] : zero_of<T>();
}

template <typename T>
__global__ void __launch_bounds__(kThreads) gather_kernel(T* __restrict__ dst0, const T* __restrict__ src0,
                                                          T* __restrict__ dst1, const T* __restrict__ src1,
                                                          const int64_t* __restrict__ jobs,
                                                          const int32_t* __restrict__ tile_job) {
  __shared__ T tile[kTE][kTE + 1];
  const int64_t t = blockIdx.x;
  const int jid = tile_job[t];
  const int64_t* __restrict__ w = jobs + (int64_t)jid * YH_JOB_WORDS;
  const bool second = w[kBuf] != 0;  // wave-uniform
  T* __restrict__ dst = second ? dst1 : dst0;
  const T* __restrict__ src = second ? src1 : src0;
  const int64_t src_off = w[0], dst_off = w[1];
  const int r = (int)w[2], kind = (int)w[3], tdim = (int)w[4];
  const int64_t local = t - w[5];

  if (kind == 0) {
    // Lanes take consecutive destination elements (coalesced); thread element i is e + i kThreads.
    uint32_t numel = 1;
#pragma unroll
    for (int d = 0; d < YH_MAXR; ++d)
      if (d < r) numel *= (uint32_t)w[kShape + d];
    const uint32_t base = (uint32_t)local * YH_GATHER_TILE;
    const uint32_t e1 = min(numel, base + (uint32_t)YH_GATHER_TILE);
    const uint32_t e = base + threadIdx.x;
    T v[kRun];
    if (r == 1) {  // one contiguous destination run: no index arithmetic
      const int64_t ss = w[kSrcStride];
      if (src_off >= 0) {
#pragma unroll
        for (int i = 0; i < kRun; ++i) {
          const uint32_t ei = e + i * kThreads;
          v[i] = src[src_off + (int64_t)(ei < e1 ? ei : 0) * ss];
        }
      }
#pragma unroll
      for (int i = 0; i < kRun; ++i) {
        const uint32_t ei = e + i * kThreads;
        if (ei < e1) dst[dst_off + ei] = (src_off >= 0) ? v[i] : zero_of<T>();
      }
      return;
    }
    switch (r) {
      case 2: gather_box<T, 2>(dst, src, w, e, e1, src_off, dst_off); break;
      case 3: gather_box<T, 3>(dst, src, w, e, e1, src_off, dst_off); break;
      case 4: gather_box<T, 4>(dst, src, w, e, e1, src_off, dst_off); break;
      case 5: gather_box<T, 5>(dst, src, w, e, e1, src_off, dst_off); break;
      case 6: gather_box<T, 6>(dst, src, w, e, e1, src_off, dst_off); break;
      case 7: gather_box<T, 7>(dst, src, w, e, e1, src_off, dst_off); break;
      default: gather_box<T, YH_MAXR>(dst, src, w, e, e1, src_off, dst_off); break;
    }
    return;
  }

  // kind 1: 32x32 transpose tile over dims (k = tdim, jd = r-1)
  const int k = tdim;
  const int jd = r - 1;
  const int64_t shk = w[kShape + k], shj = w[kShape + jd];
  const uint32_t nk = (uint32_t)yh::cdiv(shk, kTE);
  const uint32_t nj = (uint32_t)yh::cdiv(shj, kTE);
  const uint32_t per_outer = nk * nj;
  uint32_t outer = (uint32_t)local / per_outer;
  const uint32_t in_outer = (uint32_t)local - outer * per_outer;
  const uint32_t kt = in_outer / nj;
  const uint32_t jt = in_outer - kt * nj;
  int64_t sbase = src_off, dbase = dst_off;
#pragma unroll
  for (int d = YH_MAXR - 2; d >= 0; --d) {
    if (d <= r - 2 && d != k) {
      const uint32_t n = (uint32_t)w[kShape + d];
      const uint32_t q = outer / n;
      const uint32_t i = outer - q * n;
      sbase += (int64_t)i * w[kSrcStride + d];
      dbase += (int64_t)i * w[kDstStride + d];
      outer = q;
    }
  }
  const int64_t k0 = (int64_t)kt * kTE, j0 = (int64_t)jt * kTE;
  const int kk_lim = (int)min((int64_t)kTE, shk - k0);
  const int jj_lim = (int)min((int64_t)kTE, shj - j0);
  const int64_t ssj = w[kSrcStride + jd], dsk = w[kDstStride + k];
  const int lane32 = threadIdx.x & (kTE - 1);
  const int row8 = threadIdx.x >> 5;  // 0..7
  // read: consecutive lanes walk the source-contiguous dim k; cells outside the box are filled from
  // the tile's last in-range element and never written out
  const int kr = min(lane32, kk_lim - 1);
#pragma unroll
  for (int p = 0; p < kTE / 8; ++p) {
    const int jj = min(row8 + 8 * p, jj_lim - 1);
    tile[row8 + 8 * p][lane32] = src[sbase + (k0 + kr) /* ss[k]==1 */ + (j0 + jj) * ssj];
  }
  __syncthreads();
  // write: consecutive lanes walk the destination-contiguous dim
#pragma unroll
  for (int p = 0; p < kTE / 8; ++p) {
    const int kk = row8 + 8 * p;
    if (lane32 < jj_lim && kk < kk_lim) dst[dbase + (j0 + lane32) + (k0 + kk) * dsk] = tile[lane32][kk];
  }
}

}  // namespace

namespace {

// Bounds check of a plan over nbuf (source, destination) pairs; a job's buffer word selects the pair.
int check_plan(const int64_t* h_jobs, int64_t njobs, const int32_t* h_tile_job, int64_t ntiles,
               const int64_t* src_numels, const int64_t* dst_numels, int nbuf) {
  YH_CHECK_ARG(njobs >= 0 && ntiles >= 0, "negative plan length");
  YH_CHECK_ARG(njobs == 0 || (h_jobs && h_tile_job), "null plan");
  int64_t expect_tile = 0;
  for (int64_t q = 0; q < njobs; ++q) {
    const int64_t* w = h_jobs + q * YH_JOB_WORDS;
    const int64_t src_off = w[0], dst_off = w[1];
    const int rank = (int)w[2], kind = (int)w[3], tdim = (int)w[4];
    YH_CHECK_ARG(w[kBuf] >= 0 && w[kBuf] < nbuf, "job %lld: buffer %lld of %d", (long long)q, (long long)w[kBuf],
                 nbuf);
    const int64_t src_numel = src_numels[w[kBuf]], dst_numel = dst_numels[w[kBuf]];
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
    YH_CHECK_ARG(numel < ((int64_t)1 << 31), "job %lld: box of %lld elements", (long long)q, (long long)numel);
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

}  // namespace

extern "C" int yastn_hip_check_gather_plan(const int64_t* h_jobs, int64_t njobs, const int32_t* h_tile_job,
                                           int64_t ntiles, int64_t src_numel, int64_t dst_numel) {
  return check_plan(h_jobs, njobs, h_tile_job, ntiles, &src_numel, &dst_numel, 1);
}

extern "C" int yastn_hip_check_gather_plan2(const int64_t* h_jobs, int64_t njobs, const int32_t* h_tile_job,
                                            int64_t ntiles, int64_t src_numel0, int64_t dst_numel0,
                                            int64_t src_numel1, int64_t dst_numel1) {
  const int64_t sn[2] = {src_numel0, src_numel1}, dn[2] = {dst_numel0, dst_numel1};
  return check_plan(h_jobs, njobs, h_tile_job, ntiles, sn, dn, 2);
}

namespace {

int gather_launch(void* dst0, const void* src0, void* dst1, const void* src1, int dtype, const int64_t* d_jobs,
                  int64_t njobs, const int32_t* d_tile_job, int64_t ntiles, hipStream_t stream) {
  YH_CHECK_ARG(dtype == YH_F64 || dtype == YH_C128, "dtype %d", dtype);
  YH_CHECK_ARG(ntiles >= 0 && ntiles < (int64_t)INT32_MAX, "tile count %lld", (long long)ntiles);
  if (ntiles == 0) return YH_OK;
  YH_CHECK_ARG(dst0 && d_jobs && d_tile_job && njobs > 0, "null pointer");
  const dim3 grid((unsigned)ntiles), block(kThreads);
  if (dtype == YH_F64)
    hipLaunchKernelGGL(gather_kernel<double>, grid, block, 0, stream, (double*)dst0, (const double*)src0,
                       (double*)dst1, (const double*)src1, d_jobs, d_tile_job);
  else
    hipLaunchKernelGGL(gather_kernel<double2>, grid, block, 0, stream, (double2*)dst0, (const double2*)src0,
                       (double2*)dst1, (const double2*)src1, d_jobs, d_tile_job);
  return yh::hip_status(hipGetLastError(), "block_gather launch");
}

}  // namespace

extern "C" int yastn_hip_block_gather(void* dst, int64_t dst_numel, const void* src, int64_t src_numel, int dtype,
                                      const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job,
                                      int64_t ntiles, hipStream_t stream) {
  (void)dst_numel;
  (void)src_numel;
  return gather_launch(dst, src, dst, src, dtype, d_jobs, njobs, d_tile_job, ntiles, stream);
}

extern "C" int yastn_hip_block_gather2(void* dst0, int64_t dst_numel0, const void* src0, int64_t src_numel0,
                                       void* dst1, int64_t dst_numel1, const void* src1, int64_t src_numel1,
                                       int dtype, const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job,
                                       int64_t ntiles, hipStream_t stream) {
  (void)dst_numel0;
  (void)src_numel0;
  (void)dst_numel1;
  (void)src_numel1;
  YH_CHECK_ARG(ntiles == 0 || dst1, "null pointer");
  return gather_launch(dst0, src0, dst1, src1, dtype, d_jobs, njobs, d_tile_job, ntiles, stream);
}

extern "C" int yastn_hip_transpose_and_merge_pair(void* newdata0, int64_t Dsize0, const void* data0,
                                                  int64_t data_numel0, void* newdata1, int64_t Dsize1,
                                                  const void* data1, int64_t data_numel1, int dtype,
                                                  const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job,
                                                  int64_t ntiles, hipStream_t stream) {
  return yastn_hip_block_gather2(newdata0, Dsize0, data0, data_numel0, newdata1, Dsize1, data1, data_numel1, dtype,
                                 d_jobs, njobs, d_tile_job, ntiles, stream);
}

extern "C" int yastn_hip_transpose_and_merge_adjoint_pair(void* grad_src0, int64_t src_numel0, const void* grad_new0,
                                                          int64_t Dsize0, void* grad_src1, int64_t src_numel1,
                                                          const void* grad_new1, int64_t Dsize1, int dtype,
                                                          const int64_t* d_jobs, int64_t njobs,
                                                          const int32_t* d_tile_job, int64_t ntiles,
                                                          hipStream_t stream) {
  return yastn_hip_block_gather2(grad_src0, src_numel0, grad_new0, Dsize0, grad_src1, src_numel1, grad_new1, Dsize1,
                                 dtype, d_jobs, njobs, d_tile_job, ntiles, stream);
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
