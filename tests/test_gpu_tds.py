"""GPU parity of the no-fusion tensordot kernel (backend.transpose_dot_sum) and its adjoint.

Tolerances: output vs the oracle <= 1e-12 relative Frobenius; gradients vs torch autograd of the
reference torch kernel (_backend_torch_backwards.py:141-226, restated below with torch ops on the CPU)
<= 1e-12 relative.
"""
import numpy as np
import pytest
import torch

import tds_case
from oracle import np_backend as O

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')


def _torch_ref(A, B, meta_dot, Ar, Br, Ao, Bo, Dsize):
    dtype = torch.promote_types(A.dtype, B.dtype)
    A, B = A.to(dtype), B.to(dtype)
    At = [A[sl[0]:sl[1]].view(Di).permute(Ao).reshape(Dl, Dr) for sl, Di, Dl, Dr in Ar]
    Bt = [B[sl[0]:sl[1]].view(Di).permute(Bo).reshape(Dl, Dr) for sl, Di, Dl, Dr in Br]
    parts = []
    for sl, Dslc, tab in meta_dot:
        acc = 0
        for ta, tb in tab:
            acc = acc + At[ta] @ Bt[tb]
        parts.append(acc.reshape(-1))
    return torch.cat(parts)


@pytest.mark.parametrize('ca,cb', [(False, False), (True, False), (True, True)])
def test_transpose_dot_sum(ca, cb):
    from yastn_amd import backend_hip as B
    rng = np.random.default_rng(41)
    A, Bd, meta_dot, Ar, Br, Ao, Bo, Dsize = tds_case.make(rng, ca, cb)
    print('transpose_dot_sum complex', ca, cb, flush=True)
    C = B.transpose_dot_sum(torch.as_tensor(A, device=DEV), torch.as_tensor(Bd, device=DEV), meta_dot, Ar, Br,
                            Ao, Bo, Dsize)
    ref = O.transpose_dot_sum(A, Bd, meta_dot, Ar, Br, Ao, Bo, Dsize)
    got = C.cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    # adjoint
    w = rng.standard_normal(Dsize) + (1j * rng.standard_normal(Dsize) if (ca or cb) else 0)
    ta = torch.as_tensor(A, device=DEV).requires_grad_(True)
    tb = torch.as_tensor(Bd, device=DEV).requires_grad_(True)
    out = B.transpose_dot_sum(ta, tb, meta_dot, Ar, Br, Ao, Bo, Dsize)
    (out * torch.as_tensor(w, device=DEV)).real.sum().backward()
    ca_ = torch.as_tensor(A).requires_grad_(True)
    cb_ = torch.as_tensor(Bd).requires_grad_(True)
    (_torch_ref(ca_, cb_, meta_dot, Ar, Br, Ao, Bo, Dsize) * torch.as_tensor(w)).real.sum().backward()
    for g, gr in ((ta.grad, ca_.grad), (tb.grad, cb_.grad)):
        g, gr = g.cpu().numpy(), gr.numpy()
        assert np.linalg.norm(g - gr) <= 1e-12 * np.linalg.norm(gr)


@pytest.mark.parametrize('ca,cb', [(False, False), (True, True)])
def test_transpose_dot_sum_every_tile_config(ca, cb):
    """The multi-term grouped GEMM of transpose_dot_sum (outputs summing several products) on every tile
    shape, the register-direct one included, against the oracle (<= 1e-12)."""
    from yastn_amd import backend_hip as B, plans as P
    rng = np.random.default_rng(43)
    A, Bd, meta_dot, Ar, Br, Ao, Bo, Dsize = tds_case.make(rng, ca, cb)
    ref = O.transpose_dot_sum(A, Bd, meta_dot, Ar, Br, Ao, Bo, Dsize)
    tA, tB = torch.as_tensor(A, device=DEV), torch.as_tensor(Bd, device=DEV)
    B.transpose_dot_sum(tA, tB, meta_dot, Ar, Br, Ao, Bo, Dsize)  # builds and tunes the plan
    plan = P.tds_plan(meta_dot, Ar, Br, Ao, Bo, A.size, Bd.size, Dsize, DEV).dot
    assert max(int(r[6]) for r in plan.h_gemms) > 1  # some outputs sum several products
    keep = plan.cfg
    for cfg in P.L.TILE_DIMS:
        if not plan.applicable(cfg):
            continue
        print('tile config', cfg, flush=True)
        plan.set_cfg(cfg)
        got = B.transpose_dot_sum(tA, tB, meta_dot, Ar, Br, Ao, Bo, Dsize).cpu().numpy()
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref), cfg
    plan.set_cfg(keep)


def test_transpose_dot_sum_reference_metas():
    """The reference planner's own no-fusion calls (tests/golden nofusion_u1: list-typed meta_dot with
    [a block, b block] product lists, real, mixed real x complex and complex): output vs the reference numpy
    backend's recorded result <= 1e-12, gradients vs torch autograd of the reference kernel <= 1e-12."""
    import golden_io as G
    from yastn_amd import backend_hip as B
    rng = np.random.default_rng(43)
    calls = G.calls('nofusion_u1', 'transpose_dot_sum')
    assert len(calls) == 5
    for fn, args, kw, outs in calls:
        A, Bd, meta_dot, Ar, Br, Ao, Bo, Dsize = args
        assert isinstance(meta_dot, list)
        print('transpose_dot_sum reference meta', len(meta_dot), 'blocks', flush=True)
        ta = torch.as_tensor(A, device=DEV).requires_grad_(True)
        tb = torch.as_tensor(Bd, device=DEV).requires_grad_(True)
        out = B.transpose_dot_sum(ta, tb, meta_dot, Ar, Br, Ao, Bo, Dsize)
        got = out.detach().cpu().numpy()
        assert got.dtype == outs[0].dtype
        assert np.linalg.norm(got - outs[0]) <= 1e-12 * np.linalg.norm(outs[0])
        cplx = np.iscomplexobj(got)
        w = rng.standard_normal(Dsize) + (1j * rng.standard_normal(Dsize) if cplx else 0)
        (out * torch.as_tensor(w, device=DEV)).real.sum().backward()
        ca_ = torch.as_tensor(A).requires_grad_(True)
        cb_ = torch.as_tensor(Bd).requires_grad_(True)
        (_torch_ref(ca_, cb_, meta_dot, Ar, Br, Ao, Bo, Dsize) * torch.as_tensor(w)).real.sum().backward()
        for g, gr in ((ta.grad, ca_.grad), (tb.grad, cb_.grad)):
            g, gr = g.cpu().numpy(), gr.numpy()
            assert np.linalg.norm(g - gr) <= 1e-12 * np.linalg.norm(gr)
