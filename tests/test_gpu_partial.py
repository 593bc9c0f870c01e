"""Partial decompositions of the backend module on the GPU.  svd_lowrank, svds_scipy and eigh_lowrank (the
reference: ARPACK / scipy eigsh per block on the host, backend_torch.py:308-350, 424-449) come from one
batched Jacobi SVD / eigh of all blocks, so they are exact: the k leading singular triplets (eigenpairs in
the ``which`` order) against numpy's full decompositions, to 1e-10 of the block's scale.  svd_randomized
(the reference: torch.svd_lowrank per block) is a batched randomized range finder with power iterations,
checked on blocks with decaying spectra against the bound of the method."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')


@pytest.fixture(scope='module')
def B():
    from yastn_amd import backend_hip
    return backend_hip


def _svd_meta(shapes):
    meta, off, ou, os_, ov = [], 0, 0, 0, 0
    for m, n, k in shapes:
        meta.append(((off, off + m * n), (m, n), (ou, ou + m * k), (m, k), (os_, os_ + k), (ov, ov + k * n), (k, n)))
        off, ou, os_, ov = off + m * n, ou + m * k, os_ + k, ov + k * n
    return tuple(meta), off, (ou, os_, ov)


@pytest.mark.parametrize('cplx', [False, True])
@pytest.mark.parametrize('policy', ['svd_lowrank', 'svds_scipy'])
def test_svd_partial(B, policy, cplx):
    rng = np.random.default_rng(71)
    shapes = [(30, 20, 5), (64, 64, 10), (200, 120, 7), (15, 40, 15), (1, 3, 1)]
    meta, na, sizes = _svd_meta(shapes)
    a = rng.standard_normal(na) + (1j * rng.standard_normal(na) if cplx else 0)
    print('partial svd', policy, 'complex', cplx, flush=True)
    U, S, V = getattr(B, policy)(torch.as_tensor(a, device=DEV), meta, sizes)
    U, S, V = U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy()
    for sl, D, slU, DU, slS, slV, DV in meta:
        blk = a[sl[0]:sl[1]].reshape(D)
        k = DU[1]
        u, s, vh = np.linalg.svd(blk, full_matrices=False)
        scale = s[0]
        uu, ss, vv = U[slU[0]:slU[1]].reshape(DU), S[slS[0]:slS[1]], V[slV[0]:slV[1]].reshape(DV)
        assert np.max(np.abs(ss - s[:k])) <= 1e-10 * scale, D
        assert np.linalg.norm(uu.conj().T @ uu - np.eye(k)) <= 1e-10 * k, D
        assert np.linalg.norm(vv @ vv.conj().T - np.eye(k)) <= 1e-10 * k, D
        # the best rank-k approximation (unique for the gaps of random blocks)
        assert np.linalg.norm((uu * ss) @ vv - (u[:, :k] * s[:k]) @ vh[:k]) <= 1e-10 * scale * k, D


@pytest.mark.parametrize('which', ['LM', 'SM', 'LR', 'SR'])
def test_eigh_lowrank(B, which):
    rng = np.random.default_rng(72)
    shapes = [(12, 4), (40, 9), (100, 5), (3, 3)]
    meta, off, ou, os_ = [], 0, 0, 0
    blocks = []
    for n, k in shapes:
        x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        h = x + x.conj().T
        blocks.append(h)
        meta.append(((off, off + n * n), (n, n), (ou, ou + n * k), (n, k), (os_, os_ + k)))
        off, ou, os_ = off + n * n, ou + n * k, os_ + k
    a = np.concatenate([h.ravel() for h in blocks])
    print('eigh_lowrank', which, flush=True)
    S, U = B.eigh_lowrank(torch.as_tensor(a, device=DEV), tuple(meta), (os_, ou), which=which)
    S, U = S.cpu().numpy(), U.cpu().numpy()
    key = {'LM': lambda v: -np.abs(v), 'SM': np.abs, 'LR': lambda v: -v, 'SR': lambda v: v}[which]
    for h, (sl, D, slU, DU, slS) in zip(blocks, meta):
        k = DU[1]
        w = np.linalg.eigvalsh(h)
        ref = w[np.argsort(key(w), kind='stable')][:k]
        s, u = S[slS[0]:slS[1]], U[slU[0]:slU[1]].reshape(DU)
        scale = np.abs(w).max()
        assert np.max(np.abs(s - ref)) <= 1e-10 * scale, (D, which)
        assert np.linalg.norm(h @ u - u * s) <= 1e-10 * scale * k, (D, which)
        assert np.linalg.norm(u.conj().T @ u - np.eye(k)) <= 1e-10 * k, (D, which)


@pytest.mark.parametrize('cplx', [False, True])
def test_svd_randomized_batched(B, cplx):
    """Blocks A = W diag(0.6^i) X^H: with q = 2k and three power iterations the range finder's error factor is
    (sigma_q+1 / sigma_k)^7 ~ 1e-11, so the k leading singular values and the rank-k approximation agree
    with numpy to 1e-8 of the scale; d(sum S)/dA = sum_i u_i v_i^H to 1e-7; two calls give identical
    factors (fixed random seed)."""
    rng = np.random.default_rng(73)
    shapes = [(200, 120, 6), (64, 64, 8), (30, 20, 5), (15, 40, 7)]
    meta, na, sizes = _svd_meta(shapes)
    parts = []
    for m, n, k in shapes:
        r = min(m, n)
        w = np.linalg.qr(rng.standard_normal((m, r)) + (1j * rng.standard_normal((m, r)) if cplx else 0))[0]
        x = np.linalg.qr(rng.standard_normal((n, r)) + (1j * rng.standard_normal((n, r)) if cplx else 0))[0]
        parts.append(((w * 0.6 ** np.arange(r)) @ x.conj().T).ravel())
    a = np.concatenate(parts)
    ta = torch.as_tensor(a, device=DEV).requires_grad_(True)
    print('svd_randomized complex', cplx, flush=True)
    U, S, V = B.svd_randomized(ta, meta, sizes)
    S.sum().backward()
    U2, S2, V2 = B.svd_randomized(ta.detach(), meta, sizes)
    assert torch.equal(U.detach(), U2) and torch.equal(S.detach(), S2) and torch.equal(V.detach(), V2)
    U, S, V, g = U.detach().cpu().numpy(), S.detach().cpu().numpy(), V.detach().cpu().numpy(), ta.grad.cpu().numpy()
    for sl, D, slU, DU, slS, slV, DV in meta:
        blk = a[sl[0]:sl[1]].reshape(D)
        k = DU[1]
        u, s, vh = np.linalg.svd(blk, full_matrices=False)
        uu, ss, vv = U[slU[0]:slU[1]].reshape(DU), S[slS[0]:slS[1]], V[slV[0]:slV[1]].reshape(DV)
        assert np.max(np.abs(ss - s[:k])) <= 1e-8 * s[0], D
        assert np.linalg.norm(uu.conj().T @ uu - np.eye(k)) <= 1e-10 * k, D
        assert np.linalg.norm((uu * ss) @ vv - (u[:, :k] * s[:k]) @ vh[:k]) <= 1e-8 * s[0] * k, D
        gref = u[:, :k] @ vh[:k]
        assert np.linalg.norm(g[sl[0]:sl[1]].reshape(D) - gref) <= 1e-7 * k, D
