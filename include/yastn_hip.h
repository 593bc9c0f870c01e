/*
 * yastn_hip -- C ABI of the MI355X (gfx950) block-sparse backend hot path.
 *
 * Every entry point takes raw device pointers, element counts, a device-resident
 * work list ("plan") built by the host from the reference planner's meta tuples,
 * and a hipStream_t.  No torch types cross this boundary.  Entry points return
 * YH_OK (0) or a negative status; yastn_hip_last_error() returns the message.
 * They never allocate or free device memory: every output and scratch buffer is
 * passed in by the caller.  All launches are asynchronous on `stream`.
 *
 * Data model (reference yastn/tensor/_auxiliary.py:26-38): a tensor's data is one
 * flat 1-D array of float64 or complex128 holding C-ordered blocks; meta slices
 * are element offsets into it.
 *
 * The reference's "FFI" for this path is the duck-typed backend module
 * (yastn/backend/backend_torch.py:32-49); each entry point below names the
 * backend function it replaces.  yastn_amd/backend_hip.py is the binding.
 */
#ifndef YASTN_HIP_H
#define YASTN_HIP_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status / dtypes ---------------------------------------------------- */
#define YH_OK 0
#define YH_ERR_ARG -1      /* invalid argument or plan out of bounds          */
#define YH_ERR_HIP -2      /* HIP runtime error                               */
#define YH_ERR_NUMERIC -3  /* numerical failure flagged by a kernel           */

#define YH_F64 0
#define YH_C128 1

#define YH_ABI_VERSION 13

/* ---- box-gather plan (one job per copied or zero-filled box) -------------
 * int64 words per job (YH_JOB_WORDS):
 *   0 src_off  element offset in the source, -1 = zero fill
 *   1 dst_off  element offset in the destination
 *   2 rank     1..YH_MAXR (after merging dimensions contiguous in both)
 *   3 kind     0 = element copy along the destination-contiguous dim,
 *              1 = 32x32 LDS-staged transpose of (tdim, rank-1)
 *   4 tdim     kind 1: dimension whose source stride is 1
 *   5 tile_begin  index of the job's first tile
 *   6 ..  6+YH_MAXR-1    shape[YH_MAXR] (destination C order; the last dim has dst stride 1)
 *   6+YH_MAXR ..         src_stride[YH_MAXR]
 *   6+2*YH_MAXR ..       dst_stride[YH_MAXR]
 *   YH_JOB_BUFFER_WORD   buffer pair of the job: 0, or 1 for the second (destination, source) pair of
 *                        the paired entry points (0 everywhere else)
 *   (then padding to YH_JOB_WORDS)
 * tile_job[t] (int32) is the job owning tile t.
 */
#define YH_MAXR 8
#define YH_JOB_WORDS 32
#define YH_JOB_BUFFER_WORD 30
#define YH_GATHER_TILE 2048 /* elements per kind-0 tile */
#define YH_TRANSPOSE_TILE 32 /* kind-1 tile edge */

/* Host-side bounds check of a gather plan against buffer sizes (elements). */
int yastn_hip_check_gather_plan(const int64_t* h_jobs, int64_t njobs, const int32_t* h_tile_job,
                                int64_t ntiles, int64_t src_numel, int64_t dst_numel);

/* The same check for a plan over two (source, destination) pairs (buffer word 0 or 1). */
int yastn_hip_check_gather_plan2(const int64_t* h_jobs, int64_t njobs, const int32_t* h_tile_job,
                                 int64_t ntiles, int64_t src_numel0, int64_t dst_numel0, int64_t src_numel1,
                                 int64_t dst_numel1);

/* Generic engine behind the single-pair entry points that follow. */
int yastn_hip_block_gather(void* dst, int64_t dst_numel, const void* src, int64_t src_numel, int dtype,
                           const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job, int64_t ntiles,
                           hipStream_t stream);

/* A1: replaces backend.transpose_and_merge(data, order, meta_new, meta_mrg, Dsize)
 *     (reference yastn/backend/backend_np.py:653-661, backend_torch.py:588-589,
 *      _backend_torch_backwards.py:338-364).  Uncovered destination boxes are
 *     zero-fill jobs of the same plan, so every destination element is written once. */
int yastn_hip_transpose_and_merge(void* newdata, int64_t Dsize, const void* data, int64_t data_numel, int dtype,
                                  const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job,
                                  int64_t ntiles, hipStream_t stream);

/* A6: replaces kernel_transpose_and_merge.backward (_backend_torch_backwards.py:376-392):
 *     gather under the inverse permutation. */
int yastn_hip_transpose_and_merge_adjoint(void* grad_src, int64_t src_numel, const void* grad_new,
                                          int64_t Dsize, int dtype, const int64_t* d_jobs, int64_t njobs,
                                          const int32_t* d_tile_job, int64_t ntiles, hipStream_t stream);

/* Generic engine of the paired entry points: one launch over the jobs of two gathers, each job writing
 * dst0 from src0 or dst1 from src1 by its buffer word. */
int yastn_hip_block_gather2(void* dst0, int64_t dst_numel0, const void* src0, int64_t src_numel0, void* dst1,
                            int64_t dst_numel1, const void* src1, int64_t src_numel1, int dtype,
                            const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job, int64_t ntiles,
                            hipStream_t stream);

/* A1 for both operands of one tensordot in one launch: replaces the pair of
 *     _transpose_and_merge calls of _tensordot_fc (reference yastn/tensor/_contractions.py:169,174) or of
 *     _merge_to_matrix in _tensordot_f2m (_contractions.py:145-146), when both operands need a merge and
 *     share a dtype.  Buffer 0 is operand a, buffer 1 operand b; outputs equal two
 *     yastn_hip_transpose_and_merge calls bit for bit. */
int yastn_hip_transpose_and_merge_pair(void* newdata0, int64_t Dsize0, const void* data0, int64_t data_numel0,
                                       void* newdata1, int64_t Dsize1, const void* data1, int64_t data_numel1,
                                       int dtype, const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job,
                                       int64_t ntiles, hipStream_t stream);

/* A6 of the pair above (two kernel_transpose_and_merge.backward gathers,
 *     _backend_torch_backwards.py:376-392, in one launch). */
int yastn_hip_transpose_and_merge_adjoint_pair(void* grad_src0, int64_t src_numel0, const void* grad_new0,
                                               int64_t Dsize0, void* grad_src1, int64_t src_numel1,
                                               const void* grad_new1, int64_t Dsize1, int dtype,
                                               const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job,
                                               int64_t ntiles, hipStream_t stream);

/* A2: replaces backend.unmerge(data, meta) (backend_np.py:664-669,
 *     _backend_torch_backwards.py:395-408). */
int yastn_hip_unmerge(void* newdata, const void* data, int64_t numel, int dtype, const int64_t* d_jobs,
                      int64_t njobs, const int32_t* d_tile_job, int64_t ntiles, hipStream_t stream);

/* A8: replaces kernel_unmerge.backward (_backend_torch_backwards.py:417-426). */
int yastn_hip_unmerge_adjoint(void* grad_src, const void* grad_new, int64_t numel, int dtype,
                              const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job, int64_t ntiles,
                              hipStream_t stream);

/* ---- grouped GEMM plan ---------------------------------------------------
 * C_g (=|+=) sum_t op(A_t) op(B_t), FP64 MFMA (v_mfma_f64_16x16x4f64); complex128 as
 * four real products.  Operands are read through (row, col) strides so that
 * transposes and conjugate transposes need no copies.
 * gemm (int64 x YH_GEMM_WORDS): 0 M, 1 N, 2 c_off, 3 c_rs, 4 c_cs, 5 term_begin,
 *                               6 term_count, 7 flags (bit0 accumulate, bits8.. out buffer id)
 * term (int64 x YH_TERM_WORDS): 0 a_off, 1 a_rs, 2 a_cs, 3 b_off, 4 b_rs, 5 b_cs, 6 K,
 *                               7 flags (bit0 conj A, bit1 conj B, bits8.. A buffer id, bits16.. B buffer id)
 * tile (int32 x 3): gemm index, first row, first column.
 */
#define YH_GEMM_WORDS 8
#define YH_TERM_WORDS 8
#define YH_MAX_BUFS 4

/* tile configurations selectable by the autotuner */
#define YH_TILE_32x32 0
#define YH_TILE_64x64 1
#define YH_TILE_128x64 2
#define YH_TILE_16x64 3   /* skinny M (e.g. MPO contractions) */
#define YH_TILE_64x16 4   /* skinny N */
#define YH_TILE_128x128 5 /* large sectors */
#define YH_TILE_SMALLM 6  /* M <= 16: streaming kernel, one thread per column (16 x 256 tiles) */
#define YH_TILE_RD32 7    /* 32 x 32, operands read straight into registers, the 4 waves split K (long K) */

int yastn_hip_check_gemm_plan(const int64_t* h_gemms, int64_t ngemms, const int64_t* h_terms, int64_t nterms,
                              const int32_t* h_tiles, int64_t ntiles, int tile_config,
                              const int64_t* in_numel, int nin, const int64_t* out_numel, int nout);

int yastn_hip_grouped_gemm(void* const* outs, const void* const* ins, int dtype, const int64_t* d_gemms,
                           int64_t ngemms, const int64_t* d_terms, int64_t nterms, const int32_t* d_tiles,
                           int64_t ntiles, int tile_config, hipStream_t stream);

/* A3: replaces backend.dot(Adata, Bdata, meta_dot, Dsize) (backend_np.py:580-587,
 *     _backend_torch_backwards.py:98-109): all sector GEMMs of one tensordot in one launch. */
int yastn_hip_dot(void* C, const void* A, const void* B, int dtype, const int64_t* d_gemms, int64_t ngemms,
                  const int64_t* d_terms, int64_t nterms, const int32_t* d_tiles, int64_t ntiles, int tile_config,
                  hipStream_t stream);

/* A7: replaces kernel_dot.backward (_backend_torch_backwards.py:117-138): gA = sum gC B^H and
 *     gB = sum A^H gC, accumulated per operand block without atomics, one launch. */
int yastn_hip_dot_adjoint(void* gA, void* gB, const void* gC, const void* A, const void* B, int dtype,
                          const int64_t* d_gemms, int64_t ngemms, const int64_t* d_terms, int64_t nterms,
                          const int32_t* d_tiles, int64_t ntiles, int tile_config, hipStream_t stream);

/* ---- batched one-sided Jacobi SVD / Hermitian Jacobi eigh -----------------
 * block (int64 x YH_SVD_WORDS): 0 a_off, 1 m, 2 n, 3 u_off, 4 s_off, 5 v_off, 6 work_off, 7 unused
 * (eigh: m == n, v_off unused).  order[] lists block indices: the first n_lds blocks run with their
 * working matrices in LDS (lds_bytes per workgroup, <= YH_SVD_LDS_BYTES), the rest in `work`
 * (global memory), each needing (p*q + q*q) elements at work_off, p = max(m,n), q = min(m,n).
 * info[b] = sweeps used, or -1 if not converged within max_sweeps.  q <= YH_SVD_MAX_Q.
 */
#define YH_SVD_WORDS 8
#define YH_SVD_MAX_Q 2048
#define YH_SVD_LDS_BYTES 131072

int yastn_hip_check_svd_plan(const int64_t* h_blocks, int64_t nblocks, const int32_t* h_order, int64_t n_lds,
                             int64_t lds_bytes, int mode, int dtype, int64_t a_numel, int64_t u_numel,
                             int64_t s_numel, int64_t v_numel, int64_t work_numel);

/* A4: replaces backend.svd(data, meta, sizes) (backend_np.py:330-340,
 *     _backend_torch_backwards.py:26-39): thin SVD of every block, S descending. */
int yastn_hip_svd(void* U, void* S, void* Vh, const void* A, int dtype, const int64_t* d_blocks, int64_t nblocks,
                  const int32_t* d_order, int64_t n_lds, int64_t lds_bytes, void* work, int32_t* d_info,
                  int max_sweeps, hipStream_t stream);

/* A5: replaces backend.eigh(data, meta, sizes) (backend_np.py:461-474, backend_torch.py:369-384):
 *     Hermitian eigendecomposition of every block, eigenvalues ascending. */
int yastn_hip_eigh(void* S, void* U, const void* A, int dtype, const int64_t* d_blocks, int64_t nblocks,
                   const int32_t* d_order, int64_t n_lds, int64_t lds_bytes, void* work, int32_t* d_info,
                   int max_sweeps, hipStream_t stream);

/* Large blocks of A4 / A5 (working matrices beyond one CU): panel Jacobi.  init copies A (or A^H,
 * or A + mu I with the Gershgorin mu for eigh) into `work` and sets V = I.  A sweep visits the
 * (block, panel i, panel j) int32 jobs of its rounds (h_rounds: nrounds (offset, count) pairs into
 * d_jobs); per round, two launches: partial Gram matrices of each job's 2b columns over chunks of 64
 * rows (part), then per job `slices` workgroups that each run the job's inner sweep of rotations on
 * its Gram matrix and update their share of the pair's rows of W and V in place.  d_state (int32,
 * state_numel >= 4 + max_sweeps, ZERO before sweep 0): word 4 + sweep is set when a pair rotated in
 * that sweep, word 1 when the LDS cannot hold a pair; a round of sweep s > 0 whose previous sweep
 * rotated nothing exits at once, so sweeps can be issued without a host read in between (HIP graph
 * capture).  yastn_hip_svd_panel_result turns the words into the sweep count (word 3: -3 for an LDS
 * plan too small); finalize writes U, S, Vh (or S, U) like yastn_hip_svd.  Columns whose squared norm
 * is below (eps ||A||_F)^2 (floors[b], set by init) are not rotated, here and in yastn_hip_svd: they are
 * at the rounding-noise level of the block.  maxch_p = ceil(max p / 64); part holds >= (max jobs per
 * round) maxch_p T (T + 1) / 2 256 (x2 complex) doubles, T = ceil(2b / 16).  b <= 32.
 * yastn_hip_check_panel_jobs validates a job list against the blocks on the host. */
int yastn_hip_device_info(int* lds_per_block, int* num_cu);
int yastn_hip_svd_panel_init(const void* A, void* work, double* mu, double* floors, int dtype, int mode,
                             const int64_t* d_blocks, const int32_t* d_ids, int64_t nb, hipStream_t stream);
int yastn_hip_svd_panel_sweep(void* work, const double* floors, int dtype, const int64_t* d_blocks,
                              const int32_t* d_jobs, const int64_t* h_rounds, int64_t nrounds, int b, int maxch_p,
                              int slices, void* part, int64_t part_numel, int sweep, int max_sweeps,
                              int32_t* d_state, int64_t state_numel, hipStream_t stream);
int yastn_hip_check_panel_jobs(const int64_t* h_blocks, int64_t nblocks, const int32_t* h_jobs, int64_t njobs, int b,
                               int mode, int maxch_p);
/* Sweep count of the panel rounds: word 3 of the state <- the first of the `launched` sweeps that rotated
 * nothing, + 1; -1 if every launched sweep rotated; -3 if a round found an LDS plan too small for a pair. */
int yastn_hip_svd_panel_result(int32_t* d_state, int64_t state_numel, int launched, hipStream_t stream);
/* d_sweeps (nullable): a device word holding the sweep count written to info (then `sweeps` is
 * ignored), e.g. word 3 of the panel sweep state. */
int yastn_hip_svd_panel_finalize(void* U, void* S, void* Vh, const void* work, const double* mu, int dtype, int mode,
                                 const int64_t* d_blocks, const int32_t* d_ids, int64_t nb, int32_t* d_info,
                                 int sweeps, const int32_t* d_sweeps, hipStream_t stream);

/* Split-K for GEMMs with few output tiles and a long K (e.g. a full contraction to a scalar): the
 * host plan cuts K into S chunks whose GEMMs write partial results (row-major M x N each) to a
 * scratch buffer, then this reduction sums them in a fixed order, so results stay deterministic.
 * Record words: 0 M, 1 N, 2 c_off, 3 c_rs, 4 c_cs, 5 p_off, 6 S, 7 flags (bit 0 accumulate,
 * bits 8-15 output buffer).  Part of A3 dot (backend_np.py:580-587). */
#define YH_SPLITK_WORDS 8
int yastn_hip_check_splitk_plan(const int64_t* h_recs, int64_t nrec, const int64_t* out_numel, int nout,
                                int64_t part_numel);
int yastn_hip_splitk_reduce(void* const* outs, const void* partials, int dtype, const int64_t* d_recs, int64_t nrec,
                            hipStream_t stream);

/* No counterpart in the reference: yastn_hip_grouped_gemm followed by yastn_hip_splitk_reduce for a plan
 * with split-K GEMMs, without any allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  The partial products go to a file-scope device scratch array of
 * YH_NOALLOC_SCRATCH_BYTES (one per device, kept until the process ends; its address is taken with
 * hipGetSymbolAddress on the first call on a device); part_numel elements (complex: 2 doubles each)
 * beyond it are rejected.  outs[part_buf] must be NULL (the slot the plan's GEMMs write partials to).
 * Every call on a device shares that scratch: calls that overlap on one device race. */
#define YH_NOALLOC_SCRATCH_BYTES 33554432
int yastn_hip_grouped_gemm_noalloc(void* const* outs, const void* const* ins, int dtype, const int64_t* d_gemms,
                                   int64_t ngemms, const int64_t* d_terms, int64_t nterms, const int32_t* d_tiles,
                                   int64_t ntiles, int tile_config, int part_buf, int64_t part_numel,
                                   const int64_t* d_red, int64_t nred, hipStream_t stream);

/* QR preconditioning of the panel path (SVD blocks only; no reference counterpart: the reference
 * calls LAPACK gesdd / rocSOLVER gesvd per block, backend_np.py:281-286, torch_svd_gesdd.py:15-21).
 * W0 = Q R by blocked Householder QR, one-sided Jacobi on X = R^H (q x q), W0 = (Q V_J) Sigma U_J^H.
 * QR descriptor, YH_QR_WORDS int64 per block (indexed like the SVD block words):
 *   0 p, 1 q, 2 w_off (W0, p x q col-major), 3 vall_off (Householder vectors, p x q col-major),
 *   4 tm_off (-T per panel, YH_QR_NB^2 col-major each), 5 x_off (X then V_J, q x q each),
 *   6 y_off (2 x YH_QR_NB x q scratch), 7 number of reflectors k (0: q).
 * The trailing-column updates and the final multiplication by Q run through yastn_hip_grouped_gemm
 * with plans built by the host. */
#define YH_QR_WORDS 8
#define YH_QR_NB 32
int yastn_hip_check_qr_plan(const int64_t* h_qdesc, const int64_t* h_blocks, const int32_t* h_ids, int64_t n,
                            int dtype, int64_t work_numel, int64_t lds_bytes);
/* Factor panel `step` (columns step*YH_QR_NB ...) of every listed block: staged in LDS when
 * lds_bytes >= p*YH_QR_NB*itemsize for every block, in place in the workspace when lds_bytes == 0. */
int yastn_hip_svd_qr_panel(void* work, int dtype, const int64_t* d_qdesc, const int32_t* d_ids, int64_t n, int step,
                           int64_t lds_bytes, hipStream_t stream);
/* X = R^H, V_J = I. */
int yastn_hip_svd_precond_init(void* work, int dtype, const int64_t* d_qdesc, const int32_t* d_ids, int64_t n,
                               hipStream_t stream);
/* sigma, sorted; q-side vectors x_j / sigma_j (completed where sigma is at the noise level); p-side
 * vectors initialised to [V_J; 0] (U for m >= n, conjugated rows of Vh for m < n), to be multiplied
 * by Q.  d_blocks are the block's SVD words (original m, n and output offsets). */
int yastn_hip_svd_precond_finalize(void* U, void* S, void* Vh, const void* work, int dtype, const int64_t* d_blocks,
                                   const int64_t* d_qdesc, const int32_t* d_ids, int64_t n, int32_t* d_info,
                                   int sweeps, const int32_t* d_sweeps, hipStream_t stream);

/* Batched QR: replaces backend.qr(data, meta, sizes) (backend_np.py:511-520, backend_torch.py:475-484):
 * per block A (m x n) = Q (m x k) R (k x n), k = min(m, n), economic, with the sign of each R row and Q
 * column chosen so that diag(R) is real and >= 0.  Blocked Householder QR on the same panel kernel and
 * grouped-GEMM steps as the preconditioning (QR descriptor as above, with 0 p = m, 1 q = n, 5 x_off:
 * E (m x k col-major, [I_k; 0], then Q = H_1 ... H_k E), 7 k); block words (YH_QROP_WORDS int64):
 * 0 a_off, 1 q_off, 2 r_off, element offsets of the row-major blocks. */
#define YH_QROP_WORDS 3
int yastn_hip_check_qr_op_plan(const int64_t* h_qdesc, const int64_t* h_qblocks, int64_t nblocks, int dtype,
                               int64_t a_numel, int64_t q_numel, int64_t r_numel, int64_t work_numel,
                               int64_t lds_bytes);
/* W0 = A (column-major), E = [I_k; 0]. */
int yastn_hip_qr_init(const void* A, void* work, int dtype, const int64_t* d_qdesc, const int64_t* d_qblocks,
                      int64_t nblocks, hipStream_t stream);
/* Q, R from E and the upper triangle of W0, signs fixed. */
int yastn_hip_qr_finalize(void* Q, void* R, const void* work, int dtype, const int64_t* d_qdesc,
                          const int64_t* d_qblocks, int64_t nblocks, hipStream_t stream);

/* Adjoint of the batched QR (the reference differentiates torch.linalg.qr through torch's autograd,
 * backend_torch.py:475-484).  Per block (m x n, k = min(m, n), R1 = R(:, :k)):
 *   gQ' = gQ (+ A(:, k:) gR(:, k:)^H for m < n),  M1 = R1 gR1^H,  M2 = gQ'^H Q,
 *   C = copyltu(M1 - M2) (lower part, its conjugate transpose above, real part on the diagonal),
 *   B = gQ' + Q C,  gA(:, :k) = B R1^{-H}  (+ gA(:, k:) = Q gR(:, k:) for m < n);
 * the products run as grouped GEMMs of the host's plan, these two entry points do the rest.
 * Words (YH_QRADJ_WORDS int64 per block): 0 m, 1 n, 2 k, 3 m_off (M1, M2, C: k x k row-major), 4 b_off
 * (B, m x k row-major), 5 ga_off (block of gA, m x n row-major), 6 r_off (R, k x n row-major), 7 unused.
 * trsm: one 64-thread workgroup per (row, block), k <= 65536 / itemsize. */
#define YH_QRADJ_WORDS 8
int yastn_hip_check_qr_adjoint_plan(const int64_t* h_words, int64_t nblocks, int dtype, int64_t m_numel,
                                    int64_t b_numel, int64_t ga_numel, int64_t r_numel);
int yastn_hip_qr_adjoint_core(void* C, const void* M1, const void* M2, int dtype, const int64_t* d_words,
                              int64_t nblocks, hipStream_t stream);
int yastn_hip_qr_adjoint_trsm(void* gA, const void* B, const void* R, int dtype, const int64_t* d_words,
                              int64_t nblocks, int64_t max_rows, int64_t max_k, hipStream_t stream);

/* A11: replaces backend.fix_svd_signs(Udata, Vdata, meta) (backend_np.py:447-458,
 *      backend_torch.py:353-366): in place on U and Vh; blocks are SVD plan words. */
int yastn_hip_fix_svd_signs(void* U, void* Vh, int dtype, const int64_t* d_blocks, int64_t nblocks,
                            hipStream_t stream);

/* A11 (differentiable form): only the pivot of each U column, piv[s_off + j] = flat index in U of
 *      the first entry of largest 2^40-quantised modulus (backend_torch.py:353-366 takes the phase of
 *      that entry with differentiable tensor ops; the binding does the same with the pivots). */
int yastn_hip_svd_sign_pivots(const void* U, int64_t* piv, int dtype, const int64_t* d_blocks, int64_t nblocks,
                              hipStream_t stream);

/* A9 core: the k x k middle factor of kernel_svd.backward / SVDGESDD.backward
 *     (_backend_torch_backwards.py:53-66, linalg/torch_svd_gesdd.py:296-428) for every block:
 *       J = 1/2 (F+G) o (X - X^H) + 1/2 (F-G) o (Y - Y^H) + diag(gS) + i diag(Im X_jj / s_j)
 *           - [m>k] X diag(1/s) - [n>k] diag(1/s) Y^H
 *     with X = U^H gU, Y = Vh gVh^H, F_ij = l(s_j - s_i), G_ij = l(s_j + s_i), l(x) = x/(x^2+eps),
 *     1/s -> l(s); and T = [m>k] gU diag(1/s) or [n>k] diag(1/s) gVh (else 0).  The caller then
 *     forms gA = (T + U J) Vh (m >= n) or U (T + J Vh) (m < n) with grouped GEMMs.
 *     desc words (x8): 0 m, 1 n, 2 k, 3 u_off, 4 s_off, 5 v_off, 6 offset in X/Y/J, 7 offset in T. */
int yastn_hip_svd_adjoint_core(void* J, void* T, const void* X, const void* Y, const void* S, const void* gS,
                               const void* gU, const void* gVh, int dtype, const int64_t* d_descs,
                               int64_t nblocks, double eps, int has_gS, int has_gU, int has_gV,
                               hipStream_t stream);

/* A10 core: the n x n middle factor of the Hermitian eigen adjoint (torch.linalg.eigh backward,
 *     used by backend_torch.py:378): Xo = skew(X) / E, E_ij = w_j - w_i (E_ii = 1), diag(Xo) = gS
 *     when given; X = U^H gU.  gA = U Xo U^H follows with grouped GEMMs.
 *     desc words (x8): 0 n, 1 s_off, 2 offset in X/Xo. */
int yastn_hip_eigh_adjoint_core(void* Xo, const void* X, const void* S, const void* gS, int dtype,
                                const int64_t* d_descs, int64_t nblocks, int has_gS, int has_gU,
                                hipStream_t stream);

/* A12: replaces backend.apply_mask(Adata, mask, meta, Dsize, axis, ndim) and embed_mask
 *      (backend_np.py:626-643, _backend_torch_backwards.py:251-310): the kept index list of each
 *      block is split into runs of consecutive indices, one box-gather job per run; embed_mask adds
 *      zero-fill jobs for the complement.  Also the adjoints (apply <-> embed). */
int yastn_hip_apply_mask(void* newdata, int64_t Dsize, const void* data, int64_t data_numel, int dtype,
                         const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job, int64_t ntiles,
                         hipStream_t stream);
int yastn_hip_embed_mask(void* newdata, int64_t Dsize, const void* data, int64_t data_numel, int dtype,
                         const int64_t* d_jobs, int64_t njobs, const int32_t* d_tile_job, int64_t ntiles,
                         hipStream_t stream);

/* Backend dot_diag(Adata, Bdata, meta, Dsize, axis, a_ndim) (backend_np.py:604-611,
 *      backend_torch.py:557-564; caller _contractions.py:496): every output block is a dense block of B
 *      scaled along `axis` by a diagonal block of A, all blocks in one launch.  Job words
 *      (YH_DIAG_WORDS int64): 0 out_off, 1 a_off, 2 b_off, 3 n (block elements), 4 inner (product of the
 *      block dimensions after `axis`), 5 D (dimension of `axis`); max_n = largest n.  conj_a = 1 uses
 *      conj(A), which is the adjoint with respect to B (gB = g * conj(A), same jobs with B := g). */
#define YH_DIAG_WORDS 8
int yastn_hip_check_dot_diag_plan(const int64_t* h_jobs, int64_t njobs, const int64_t* h_groups,
                                  int64_t ngroups, const int64_t* h_job_ids, int64_t nids, int64_t out_numel,
                                  int64_t a_numel, int64_t b_numel);
int yastn_hip_dot_diag(void* out, const void* A, const void* B, int dtype, const int64_t* d_jobs, int64_t njobs,
                       int64_t max_n, int conj_a, hipStream_t stream);

/* Adjoint of dot_diag with respect to the diagonal (the reference differentiates the torch ops of
 *      backend_torch.py:557-564): gA[a_off + i] = sum of g conj(B) over the elements of index i along
 *      `axis`, summed over every job that reads the same A block (group words x4: 0 a_off, 1 D,
 *      2 first index into job_ids, 3 count), in a fixed order (no atomics). */
int yastn_hip_dot_diag_adjoint_a(void* gA, const void* g, const void* B, int dtype, const int64_t* d_jobs,
                                 const int64_t* d_groups, int64_t ngroups, const int64_t* d_job_ids, int64_t max_D,
                                 hipStream_t stream);

/* Backend transpose(data, axes, meta_transpose) (backend_np.py:646-650): per-block permutation. */
int yastn_hip_transpose(void* newdata, const void* data, int64_t numel, int dtype, const int64_t* d_jobs,
                        int64_t njobs, const int32_t* d_tile_job, int64_t ntiles, hipStream_t stream);

const char* yastn_hip_last_error(void);
int yastn_hip_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* YASTN_HIP_H */
