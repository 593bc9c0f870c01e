"""GPU parity of the batched Householder QR (backend.qr) and of its adjoint.

Tolerances (stated per result):
  Q, R vs the oracle (positive-diagonal QR, unique for full-rank blocks)   <= 1e-12 relative
  Q R = A, Q^H Q = I                                                       <= 1e-12
  gradient vs torch autograd through the reference torch backend's qr
  (torch.linalg.qr + sign fix, backend_torch.py:475-484) on the CPU        <= 1e-9 relative
  gradient vs central finite differences (h = 1e-6)                        <= 1e-7 relative
"""
import numpy as np
import pytest
import torch

from oracle import np_backend as O

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')
SHAPES = [(1, 1), (5, 3), (3, 5), (40, 40), (100, 37), (37, 100), (300, 120), (120, 300), (440, 440), (64, 17)]


def _meta(shapes):
    meta, off, oq, orr = [], 0, 0, 0
    for m, n in shapes:
        k = min(m, n)
        meta.append(((off, off + m * n), (m, n), (oq, oq + m * k), (m, k), (orr, orr + k * n), (k, n)))
        off, oq, orr = off + m * n, oq + m * k, orr + k * n
    return tuple(meta), off, (oq, orr)


def _data(rng, n, cplx):
    return rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)


@pytest.fixture(scope='module')
def B():
    from yastn_amd import backend_hip
    return backend_hip


# blocks whose Householder panel exceeds LDS: the whole call factors its panels in place
LONG = [(1300, 40), (700, 90), (90, 700)]


@pytest.mark.parametrize('cplx', [False, True])
@pytest.mark.parametrize('long_blocks', [False, True])
def test_qr_matches_oracle(B, cplx, long_blocks):
    rng = np.random.default_rng(31)
    meta, na, sizes = _meta(SHAPES + (LONG if long_blocks else []))
    a = _data(rng, na, cplx)
    print('qr batched blocks', len(meta), 'complex', cplx, 'long', long_blocks, flush=True)
    Q, R = B.qr(torch.as_tensor(a, device=DEV), meta, sizes)
    Q, R = Q.cpu().numpy(), R.cpu().numpy()
    Qo, Ro = O.qr(a, meta, sizes)
    for sl, D, slQ, DQ, slR, DR in meta:
        q, r = Q[slQ[0]:slQ[1]].reshape(DQ), R[slR[0]:slR[1]].reshape(DR)
        blk = a[sl[0]:sl[1]].reshape(D)
        nrm = np.linalg.norm(blk)
        assert np.linalg.norm(q @ r - blk) <= 1e-12 * nrm, D
        assert np.linalg.norm(q.conj().T @ q - np.eye(DQ[1])) <= 1e-12 * DQ[1], D
        assert np.all(np.real(np.diag(r)) >= 0) and np.allclose(np.tril(r, -1), 0), D
        qo, ro = Qo[slQ[0]:slQ[1]].reshape(DQ), Ro[slR[0]:slR[1]].reshape(DR)
        assert np.linalg.norm(q - qo) <= 1e-12 * np.sqrt(DQ[1]), D
        assert np.linalg.norm(r - ro) <= 1e-12 * nrm, D


def _torch_ref_qr(a, meta, sizes):
    """The reference torch backend's qr (backend_torch.py:475-484), differentiated by torch on the CPU."""
    Q = torch.zeros(sizes[0], dtype=a.dtype)
    R = torch.zeros(sizes[1], dtype=a.dtype)
    for sl, D, slQ, DQ, slR, DR in meta:
        q, r = torch.linalg.qr(a[sl[0]:sl[1]].view(D))
        s = torch.sign(torch.real(r.diag()))
        s = torch.where(s == 0, torch.ones_like(s), s)
        Q[slQ[0]:slQ[1]].view(DQ)[:] = q * s
        R[slR[0]:slR[1]].view(DR)[:] = s.reshape(-1, 1) * r
    return Q, R


@pytest.mark.parametrize('cplx', [False, True])
def test_qr_adjoint(B, cplx):
    rng = np.random.default_rng(32)
    shapes = [(6, 4), (4, 6), (30, 30), (50, 20), (20, 50), (130, 70)]
    meta, na, sizes = _meta(shapes)
    a = _data(rng, na, cplx)
    wq, wr = _data(rng, sizes[0], cplx), _data(rng, sizes[1], cplx)

    def loss(Q, R, wq_, wr_):
        return (Q * wq_).real.sum() + (R * wr_).real.sum()

    ta = torch.as_tensor(a, device=DEV).requires_grad_(True)
    print('qr adjoint blocks', len(meta), 'complex', cplx, flush=True)
    Q, R = B.qr(ta, meta, sizes)
    loss(Q, R, torch.as_tensor(wq, device=DEV), torch.as_tensor(wr, device=DEV)).backward()
    g = ta.grad.cpu().numpy()
    # torch's own adjoint of the reference formula, on the CPU
    tc = torch.as_tensor(a).requires_grad_(True)
    Qc, Rc = _torch_ref_qr(tc, meta, sizes)
    loss(Qc, Rc, torch.as_tensor(wq), torch.as_tensor(wr)).backward()
    gc = tc.grad.numpy()
    assert np.linalg.norm(g - gc) <= 1e-9 * np.linalg.norm(gc)
    # central finite differences along random directions (the oracle's forward)
    h = 1e-6

    def f(x):
        Qo, Ro = O.qr(x, meta, sizes)
        return np.sum(np.real(Qo * wq)) + np.sum(np.real(Ro * wr))
    for _ in range(4):
        d = _data(rng, na, cplx)
        fd = (f(a + h * d) - f(a - h * d)) / (2 * h)
        an = np.real(np.vdot(g, d)) if cplx else float(g @ d)
        assert abs(fd - an) <= 1e-7 * max(1.0, abs(fd)), (fd, an)
