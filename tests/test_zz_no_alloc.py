"""yastn_hip_grouped_gemm_noalloc (split-K partial products in the library's static device scratch, no
allocation) against the allocating path (yastn_hip_grouped_gemm with a torch-allocated partials buffer, then
yastn_hip_splitk_reduce): bitwise equal, on two calls in a row (the second with a smaller input), real and
complex; a plan whose partials exceed the scratch is rejected with an error."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')


def _case(rng, shapes, cplx):
    A, Bm, meta, ao, bo, co = [], [], [], 0, 0, 0
    for M, N, K in shapes:
        A.append(rng.standard_normal(M * K) + (1j * rng.standard_normal(M * K) if cplx else 0))
        Bm.append(rng.standard_normal(K * N) + (1j * rng.standard_normal(K * N) if cplx else 0))
        meta.append(((co, co + M * N), (M, N), (ao, ao + M * K), (M, K), (bo, bo + K * N), (K, N)))
        ao, bo, co = ao + M * K, bo + K * N, co + M * N
    return (torch.as_tensor(np.concatenate(A), device=DEV), torch.as_tensor(np.concatenate(Bm), device=DEV),
            tuple(meta), co)


def _noalloc(plan, out, A, B):
    from yastn_amd import _lib as L
    from yastn_amd import _functions as F
    outs = [out] + [None] * (L.YH_MAX_BUFS - 1)
    ins = [A, B] + [None] * (L.YH_MAX_BUFS - 2)
    po = (ctypes.c_void_p * L.YH_MAX_BUFS)(*[F.ptr(t) for t in outs])
    pi = (ctypes.c_void_p * L.YH_MAX_BUFS)(*[F.ptr(t) for t in ins])
    L.call('yastn_hip_grouped_gemm_noalloc', po, pi, F.dtype_code(out), *plan.args(), plan.part_buf,
           plan.part_numel, plan.d_red.data_ptr(), len(plan.h_red), F.stream_of(out))


@pytest.mark.parametrize('cplx', [False, True])
def test_noalloc_split_k_equals_allocating_path(cplx):
    from yastn_amd import _functions as F
    from yastn_amd import plans as P
    rng = np.random.default_rng(70 + cplx)
    for shapes in ([(1, 1, 200000), (3, 5, 20000), (40, 40, 300), (8, 8, 50000)],
                   [(1, 1, 30000), (2, 3, 9000)]):  # the second call is smaller
        A, B, meta, Dsize = _case(rng, shapes, cplx)
        plan = P.dot_plan(meta, A.numel(), B.numel(), Dsize, DEV)
        assert plan.part_numel > 0 and plan.part_buf == 1
        print('noalloc split-k', cplx, shapes, plan.part_numel, flush=True)
        ref = torch.zeros(Dsize, dtype=A.dtype, device=DEV)
        F.launch_plan(plan, (ref,), (A, B), F.dtype_code(ref), F.stream_of(ref))
        got = torch.zeros(Dsize, dtype=A.dtype, device=DEV)
        _noalloc(plan, got, A, B)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)


def test_noalloc_rejects_oversized_partials():
    from yastn_amd import _lib as L
    from yastn_amd import _functions as F
    from yastn_amd import plans as P
    rng = np.random.default_rng(72)
    A, B, meta, Dsize = _case(rng, [(1, 1, 200000)], False)
    plan = P.dot_plan(meta, A.numel(), B.numel(), Dsize, DEV)
    out = torch.zeros(Dsize, dtype=A.dtype, device=DEV)
    outs = (ctypes.c_void_p * L.YH_MAX_BUFS)(F.ptr(out), None, None, None)
    ins = (ctypes.c_void_p * L.YH_MAX_BUFS)(F.ptr(A), F.ptr(B), None, None)
    with pytest.raises(L.HipBackendError, match='scratch'):
        L.call('yastn_hip_grouped_gemm_noalloc', outs, ins, F.dtype_code(out), *plan.args(), plan.part_buf,
               L.YH_NOALLOC_SCRATCH_BYTES // 8 + 1, plan.d_red.data_ptr(), len(plan.h_red), F.stream_of(out))
