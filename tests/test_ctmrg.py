"""CTMRG harness on the CPU: numpy oracle forward vs the plain torch path, and finite differences."""
import numpy as np
import pytest
import torch

from oracle import np_module as NB
from yastn_amd import blocksparse as bs
from yastn_amd import ctmrg as C


def _plain():
    from yastn_amd import backend_plain
    return backend_plain


@pytest.mark.parametrize('D,chi', [(2, 8), (3, 12)])
def test_forward_oracle_vs_plain_cpu(D, chi):
    cn = bs.make_config(NB, 'U1')
    A, env, sz = C.setup(cn, D=D, chi=chi, seed=1)
    v_np = C.loss_fn(A, env, sz, chi)
    cp = bs.make_config(_plain(), 'U1', device='cpu')
    A, env, sz = C.setup(cp, D=D, chi=chi, seed=1)
    v_t = C.loss_fn(A, env, sz, chi)
    assert abs(float(v_np) - float(v_t)) <= 1e-10 * max(1.0, abs(float(v_np)))


def test_gradient_central_fd_plain_cpu():
    """Gradient of one CTMRG step + <Sz> w.r.t. the site tensor vs central finite differences
    (h = 1e-6, 4 random directions), with the SVD adjoint unbroadened (eps 1e-30): the reference's
    fixed eps = 1e-12 biases the gradient when kept singular values are ~1e-6."""
    B = _plain()
    B.set_ad_broadening(1e-30)
    try:
        _fd_check(B)
    finally:
        B.set_ad_broadening(1e-12)


def test_broadening_biases_small_spectra():
    """With the reference's eps = 1e-12 the same gradient differs from finite differences by percent
    (kept singular values near 1e-6 in the vertical move): the quirk is reproduced, not fixed."""
    B = _plain()
    cp = bs.make_config(B, 'U1', device='cpu')
    A, env, sz = C.setup(cp, D=2, chi=8, seed=2)
    loss, g = C.fwd_bwd(A, env, sz, 8)
    d = torch.as_tensor(np.random.default_rng(0).standard_normal(A.data.numel()))
    with torch.no_grad():
        fd = float((C.loss_fn(A._replace(data=A.data + 1e-6 * d), env, sz, 8) -
                    C.loss_fn(A._replace(data=A.data - 1e-6 * d), env, sz, 8)) / 2e-6)
    an = float((g * d).sum())
    assert 1e-3 < abs(fd - an) / abs(fd) < 0.2


def central_fd(f, h=1e-6):
    """Five-point central difference with step h (fourth order)."""
    return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)


def _fd_check(B):
    cp = bs.make_config(B, 'U1', device='cpu')
    A, env, sz = C.setup(cp, D=2, chi=8, seed=2)
    loss, g = C.fwd_bwd(A, env, sz, 8)
    rng = np.random.default_rng(0)
    for _ in range(4):
        d = torch.as_tensor(rng.standard_normal(A.data.numel()))

        def f(h):
            with torch.no_grad():
                return float(C.loss_fn(A._replace(data=A.data + h * d), env, sz, 8))
        fd = central_fd(f)
        an = float((g * d).sum())
        # relative to the gradient scale |g| |d| (directions with cancelling projections included)
        assert abs(fd - an) <= 1e-7 * float(g.norm() * d.norm()), (fd, an)
