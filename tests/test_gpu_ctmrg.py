"""CTMRG step on the GPU: batched HIP backend vs the plain per-block path and finite differences."""
import numpy as np
import pytest
import torch

from yastn_amd import blocksparse as bs
from yastn_amd import ctmrg as C

pytestmark = pytest.mark.gpu


def _backends():
    from yastn_amd import backend_hip, backend_plain
    return backend_hip, backend_plain


@pytest.mark.parametrize('D,chi,dtype', [(3, 24, 'float64'), (2, 16, 'complex128')])
def test_hip_vs_plain_loss_and_gradient(D, chi, dtype):
    BH, BP = _backends()
    print('ctmrg', D, chi, dtype, flush=True)
    res = []
    for B in (BH, BP):
        cfg = bs.make_config(B, 'U1', device='cuda:0')
        A, env, sz = C.setup(cfg, D=D, chi=chi, dtype=dtype, seed=3)
        res.append(C.fwd_bwd(A, env, sz, chi))
    (lh, gh), (lp, gp) = res
    assert abs(float(lh) - float(lp)) <= 1e-10 * max(1.0, abs(float(lp)))
    rel = float((gh - gp).norm() / gp.norm())
    assert rel <= 1e-8, rel


def test_hip_gradient_central_fd():
    """20 random directions, h = 1e-6 (five-point central stencil), unbroadened SVD adjoint."""
    BH, _ = _backends()
    BH.set_ad_broadening(1e-30)
    try:
        cfg = bs.make_config(BH, 'U1', device='cuda:0')
        A, env, sz = C.setup(cfg, D=2, chi=8, seed=2)
        loss, g = C.fwd_bwd(A, env, sz, 8)
        rng = np.random.default_rng(0)
        worst = 0.0
        for _ in range(20):
            d = torch.as_tensor(rng.standard_normal(A.data.numel()), device=g.device)

            def f(h):
                with torch.no_grad():
                    return float(C.loss_fn(A._replace(data=A.data + h * d), env, sz, 8))
            h = 1e-6
            fd = (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)
            an = float((g * d).sum())
            worst = max(worst, abs(fd - an) / float(g.norm() * d.norm()))
        assert worst <= 1e-7, worst
    finally:
        BH.set_ad_broadening(1e-12)


def test_hip_unrolled_energy_checkpointed_vs_plain():
    """Three CTMRG steps + nearest-neighbour Heisenberg energy, complex128, checkpoint segments of two steps
    on the HIP backend, against the plain per-block path without checkpointing."""
    BH, BP = _backends()
    print('unrolled ctmrg D=3 chi=12 complex128', flush=True)
    res = []
    for B, seg in ((BH, 2), (BP, 0)):
        cfg = bs.make_config(B, 'U1', device='cuda:0')
        A, env, _ = C.setup(cfg, D=3, chi=12, dtype='complex128', seed=4)
        res.append(C.energy_fwd_bwd(A, env, C.heisenberg_bond(cfg), 12, 3, segment=seg))
    (lh, gh), (lp, gp) = res
    assert abs(float(lh) - float(lp)) <= 1e-10 * max(1.0, abs(float(lp)))
    rel = float((gh - gp).norm() / gp.norm())
    assert rel <= 1e-8, rel


def test_hip_move_batched_projector_svds():
    """The two projector SVDs of each move in one batched HIP call (blocksparse.svd_many) against the
    plain path's one-after-the-other SVDs: loss and gradient."""
    BH, BP = _backends()
    print('ctmrg move-batched projector svds D=3 chi=24', flush=True)
    res = []
    for B, batch in ((BH, True), (BP, False)):
        cfg = bs.make_config(B, 'U1', device='cuda:0')
        A, env, sz = C.setup(cfg, D=3, chi=24, seed=3)
        res.append(C.fwd_bwd(A, env, sz, 24, batch=batch))
    (lh, gh), (lp, gp) = res
    assert abs(float(lh) - float(lp)) <= 1e-10 * max(1.0, abs(float(lp)))
    rel = float((gh - gp).norm() / gp.norm())
    assert rel <= 1e-8, rel


@pytest.mark.parametrize('D,chi,dtype', [(3, 24, 'float64'), (2, 16, 'complex128')])
def test_hip_qr_projectors_vs_plain(D, chi, dtype):
    """use_qr=True (R factors of the half-systems through the batched HIP QR and its adjoint) against the
    plain path's per-block torch QR, autograd-differentiated: loss and gradient of one horizontal move +
    <Sz>.  One move: after it the truncated environment makes the next move's half-systems numerically
    rank-deficient (|R_jj| ~ 1e-21), where the QR adjoint, the reference's included, divides by ~0 and
    the gradient is rounding noise or not finite."""
    BH, BP = _backends()
    print('ctmrg use_qr', D, chi, dtype, flush=True)
    res = []
    for B in (BH, BP):
        cfg = bs.make_config(B, 'U1', device='cuda:0')
        A, env, sz = C.setup(cfg, D=D, chi=chi, dtype=dtype, seed=3)
        x = A.data.detach().clone().requires_grad_(True)
        A2 = A._replace(data=x)
        val = C.cluster_expectation(C.move_h(env, C.double_layer(A2), chi, {'use_qr': True}), A2, sz)
        loss = val.real if val.is_complex() else val
        loss.backward()
        res.append((loss.detach(), x.grad))
    (lh, gh), (lp, gp) = res
    assert abs(float(lh) - float(lp)) <= 1e-10 * max(1.0, abs(float(lp)))
    rel = float((gh - gp).norm() / gp.norm())
    assert rel <= 1e-8, rel


def test_rvb_kagome_qr_projectors_on_gpu():
    """The RVB state, 2 sweeps from 'dl', use_qr=True through the HIP backend: the loss of the reference
    CTMRG with use_qr=True (<= 1e-12)."""
    import rvb_case as R
    from yastn_amd import backend_hip
    sweeps, init, loss_qr, _ = R.runs_qr()[1]
    print('rvb kagome use_qr', sweeps, 'sweeps from', init, flush=True)
    cfg = bs.make_config(backend_hip, 'U1', device='cuda:0')
    with torch.no_grad():
        loss = C.rdm_loss(R.state(cfg), R.CHI, sweeps, init, dict(R.svd_opts(), use_qr=True), block=R.BLOCK)
    assert abs(float(loss) - loss_qr) <= 1e-12


@pytest.mark.parametrize('k', [0, 2, 3])
def test_rvb_kagome_reference_ctmrg_on_gpu(k):
    """The Kagome RVB D=3 U(1) complex state of the reference's tests/peps/test_ad.py through the HIP backend
    (chi=64, multiplet truncation, use_qr=False, fix_signs=False): 1 sweep from 'eye' and 2 sweeps from 'dl'
    against the reference CTMRG's loss (<= 1e-12) and gradient (<= 1e-9 relative); 20 sweeps from 'dl'
    against the reference test's grad_expected (rtol 1e-3, atol 1e-5) and the reference CTMRG's gradient
    (<= 1e-8 relative)."""
    import rvb_case as R
    from yastn_amd import backend_hip
    sweeps, init, loss_ref, g_ref = R.runs()[k]
    print('rvb kagome', sweeps, 'sweeps from', init, flush=True)
    cfg = bs.make_config(backend_hip, 'U1', device='cuda:0')
    A = R.state(cfg)
    x = A.data.detach().clone().requires_grad_(True)
    loss = C.rdm_loss(A._replace(data=x), R.CHI, sweeps, init, R.svd_opts(), block=R.BLOCK)
    loss.backward()
    g = x.grad.cpu().numpy()
    if sweeps == 20:
        assert np.allclose(g, R.grad_expected(), rtol=1e-3, atol=1e-5)
        assert abs(float(loss) - loss_ref) <= 1e-10
        assert np.abs(g - g_ref).max() <= 1e-8 * np.abs(g_ref).max()
    else:
        assert abs(float(loss) - loss_ref) <= 1e-12
        assert np.abs(g - g_ref).max() <= 1e-9 * np.abs(g_ref).max()


@pytest.mark.parametrize('batch', [False, True])
def test_captured_step_replays_equal_eager(batch):
    """D=6, chi=64 CTMRG step + <Sz>, forward and backward, with static-structure projectors, captured as
    one HIP graph (also with the two projector SVDs of each move in one call): every replay (with eager work
    and synchronisations in between) gives the eager loss and gradient bitwise; the Jacobi convergence
    flags and the structure check are read after each replay."""
    from yastn_amd import backend_hip
    cfg = bs.make_config(backend_hip, 'U1', device='cuda:0')
    A, env, sz = C.setup(cfg, D=6, chi=64, seed=0)
    print('capturing D=6 chi=64 fwd+bwd, move-batched', batch, flush=True)
    cs, (l_eager, g_eager) = C.captured_fwd_bwd(A, env, sz, 64, batch=batch)
    for k in range(3):
        loss, grad = cs.replay()
        other = cs.validate()
        print('replay', k, float(loss), 'global-rule differences', other, flush=True)
        assert torch.equal(loss, l_eager) and torch.equal(grad, g_eager)
        x = torch.randn(1000, device='cuda:0')  # eager work between replays
        (x @ x).item()
