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
