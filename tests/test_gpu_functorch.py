"""The hot functions are torch.autograd.Functions in forward + setup_context style, as the reference's
fixed-point AD (torch.func.vjp, fixed_pt.py:853) and checkpointed recomputation need: torch.func.vjp and
torch.utils.checkpoint through the HIP backend give the plain autograd gradient (to rounding)."""
import numpy as np
import pytest
import torch
from torch.utils.checkpoint import checkpoint

import golden_io as G

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')


def _pipeline(B):
    """dot of the recorded operands, then the recorded SVD's shapes on a matching buffer: scalar loss."""
    _, (A, Bd, meta_dot, Dsize), _, _ = G.calls('mps_u1_fuse_contracted', 'dot')[0]
    _, (S_in, meta_svd, sizes), _, _ = G.calls('mps_u1_fuse_contracted', 'svd')[0]
    rng = np.random.default_rng(51)
    wC = torch.as_tensor(rng.standard_normal(Dsize), device=DEV)
    wU = torch.as_tensor(rng.standard_normal(sizes[0]), device=DEV)
    Bt = torch.as_tensor(Bd, device=DEV)
    S0 = torch.as_tensor(S_in, device=DEV)

    def f(a, s):
        C = B.dot(a, Bt, meta_dot, Dsize)
        U, S, V = B.svd(s, meta_svd, sizes)
        return (C * wC).sum() + S.sum() + (U * wU).sum() ** 2
    return f, torch.as_tensor(A, device=DEV), S0


def test_vjp_and_checkpoint_match_autograd():
    from yastn_amd import backend_hip as B
    f, a0, s0 = _pipeline(B)
    a, s = a0.clone().requires_grad_(True), s0.clone().requires_grad_(True)
    f(a, s).backward()
    ga, gs = a.grad.clone(), s.grad.clone()
    print('functorch vjp', flush=True)
    _, vjp_fn = torch.func.vjp(f, a0, s0)
    va, vs = vjp_fn(torch.ones((), dtype=a0.dtype, device=DEV))
    assert torch.linalg.norm(va - ga) <= 1e-12 * torch.linalg.norm(ga)
    assert torch.linalg.norm(vs - gs) <= 1e-12 * torch.linalg.norm(gs)
    print('checkpoint', flush=True)
    a2, s2 = a0.clone().requires_grad_(True), s0.clone().requires_grad_(True)
    checkpoint(f, a2, s2, use_reentrant=False).backward()
    assert torch.linalg.norm(a2.grad - ga) <= 1e-12 * torch.linalg.norm(ga)
    assert torch.linalg.norm(s2.grad - gs) <= 1e-12 * torch.linalg.norm(gs)
    print('reentrant checkpoint', flush=True)
    a3, s3 = a0.clone().requires_grad_(True), s0.clone().requires_grad_(True)
    checkpoint(f, a3, s3, use_reentrant=True).backward()
    assert torch.linalg.norm(a3.grad - ga) <= 1e-12 * torch.linalg.norm(ga)
    assert torch.linalg.norm(s3.grad - gs) <= 1e-12 * torch.linalg.norm(gs)


@pytest.mark.parametrize('D,chi,dtype', [(3, 24, 'float64'), (2, 16, 'complex128')])
def test_vjp_of_whole_ctm_step(D, chi, dtype):
    """torch.func.vjp of one whole CTM step f(C, a) through the HIP backend (merges, grouped GEMMs, projector
    SVDs with fix_signs, dot_diag, masks, unmerges), as the reference's fixed-point AD takes it
    (fixed_pt.py:853), equals torch.autograd.grad of the same step: <= 1e-11 relative."""
    from yastn_amd import backend_hip
    from yastn_amd import blocksparse as bs
    from yastn_amd import ctmrg as C
    cfg = bs.make_config(backend_hip, 'U1', device='cuda:0')
    A, env, _ = C.setup(cfg, D=D, chi=chi, dtype=dtype, seed=1)
    opts = C.static_svd_opts(chi)
    tmpl = C.ctm_step(env, C.double_layer(A), chi, opts)
    f = C.ctm_step_flat(A, tmpl, chi, opts)
    C0 = C.env_flat(tmpl).detach()
    print('vjp of a CTM step', D, chi, dtype, flush=True)
    out, vjp = torch.func.vjp(f, C0, A.data)
    rng = np.random.default_rng(4)
    v = torch.as_tensor(rng.standard_normal(out.numel()) + (1j * rng.standard_normal(out.numel())
                        if out.is_complex() else 0), device=DEV)
    gC, gA = vjp(v)
    c = C0.clone().requires_grad_(True)
    x = A.data.clone().requires_grad_(True)
    rC, rA = torch.autograd.grad(f(c, x), (c, x), grad_outputs=v)
    for g, r in ((gC, rC), (gA, rA)):
        assert float((g - r).norm()) <= 1e-11 * float(r.norm())
