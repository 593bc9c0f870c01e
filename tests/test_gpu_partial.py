"""Partial decompositions of the backend module on the GPU (svd_lowrank, svds_scipy, svd_randomized,
eigh_lowrank): the reference computes them per block on the host (ARPACK / scipy eigsh,
backend_torch.py:308-350, 424-449) or with torch.svd_lowrank; here they come from one batched Jacobi
SVD / eigh of all blocks, so they are exact: the k leading singular triplets (eigenpairs in the
``which`` order) against numpy's full decompositions, to 1e-10 of the block's scale."""
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
@pytest.mark.parametrize('policy', ['svd_lowrank', 'svds_scipy', 'svd_randomized'])
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
