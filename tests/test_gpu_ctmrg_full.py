"""CTMRG at the full workload sizes on the GPU.

* D=6, chi=64 U(1) float64 'hv' step, forward + backward: the HIP backend against the plain per-block
  path (which follows the reference torch backend, yastn/backend/backend_torch.py), and against central
  finite differences on 20 seeded random directions, judged per direction.
* D=8, chi=128 U(1) complex128 CTMRG + nearest-neighbour Heisenberg energy: checkpoint segments of one
  step (the reference's checkpoint_move, yastn/tn/fpeps/envs/_env_ctm.py:556-589) against the unrolled
  graph, and the HIP backend against the plain path for one step.
"""
import numpy as np
import pytest
import torch

from yastn_amd import blocksparse as bs
from yastn_amd import ctmrg as C
from yastn_amd import _functions as F

pytestmark = pytest.mark.gpu


def _backends():
    from yastn_amd import backend_hip, backend_plain
    return backend_hip, backend_plain


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def test_d6_chi64_hip_vs_plain(monkeypatch):
    """Loss <= 1e-10 and gradient <= 1e-8 (relative) against the plain path; the projector SVDs run the
    QR-preconditioned panel Jacobi (blocks up to 440 x 440), not only the LDS-resident kernel."""
    BH, BP = _backends()
    seen = []
    orig = F._run_jacobi

    def spy(mode, plan, *a, **k):
        seen.append((plan.n_big, getattr(plan, 'n_pre', 0), plan.nblocks))
        return orig(mode, plan, *a, **k)
    monkeypatch.setattr(F, '_run_jacobi', spy)
    res = []
    for B in (BH, BP):
        print('ctmrg D=6 chi=64 fwd+bwd', B.__name__, flush=True)
        cfg = bs.make_config(B, 'U1', device='cuda:0')
        A, env, sz = C.setup(cfg, D=6, chi=64, seed=0)
        res.append(C.fwd_bwd(A, env, sz, 64))
    (lh, gh), (lp, gp) = res
    assert abs(float(lh) - float(lp)) <= 1e-10 * abs(float(lp))
    assert _rel(gh, gp) <= 1e-8, _rel(gh, gp)
    assert any(nb > 0 and npre > 0 for nb, npre, _ in seen), seen


def test_d6_chi64_gradient_central_fd_per_direction():
    """20 seeded random directions, h = 1e-6, five-point central stencil, unbroadened SVD adjoint
    (broadening 1e-30); each direction judged on its own: |fd - an| / |an| <= 1e-7."""
    BH, _ = _backends()
    BH.set_ad_broadening(1e-30)
    try:
        cfg = bs.make_config(BH, 'U1', device='cuda:0')
        A, env, sz = C.setup(cfg, D=6, chi=64, seed=0)
        loss, g = C.fwd_bwd(A, env, sz, 64)
        # directions from another seed: the site tensor's own seed gives d ~ A, along which the loss
        # (a normalised expectation value) is constant
        rng = np.random.default_rng(12345)
        errs = []
        for k in range(20):
            d = torch.as_tensor(rng.standard_normal(A.data.numel()), device=g.device)

            def f(h):
                with torch.no_grad():
                    return float(C.loss_fn(A._replace(data=A.data + h * d), env, sz, 64))
            h = 1e-6
            fd = (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)
            an = float((g * d).sum())
            errs.append(abs(fd - an) / abs(an))
            print(f'direction {k}: an {an:.6e} rel {errs[-1]:.2e}', flush=True)
        assert max(errs) <= 1e-7, errs
    finally:
        BH.set_ad_broadening(1e-12)


def test_d8_chi128_complex_checkpoint_segments_vs_unrolled():
    """Two CTMRG steps + energy, complex128, D=8, chi=128: checkpoint segments of one step (each step's
    forward recomputed in the backward pass) against the unrolled graph; the recomputation is the same
    deterministic kernels on the same inputs, so energy and gradient agree to rounding (asserted <= 1e-12)."""
    BH, _ = _backends()
    cfg = bs.make_config(BH, 'U1', device='cuda:0')
    A, env, _ = C.setup(cfg, D=8, chi=128, dtype='complex128', seed=0)
    H = C.heisenberg_bond(cfg)
    out = []
    for seg in (1, 0):
        print('ctmrg D=8 chi=128 c128, 2 steps + energy, segment', seg, flush=True)
        out.append(C.energy_fwd_bwd(A, env, H, 128, 2, segment=seg))
    (l1, g1), (l0, g0) = out
    print('bitwise equal gradients:', bool(torch.equal(g1, g0)), flush=True)
    assert abs(float(l1) - float(l0)) <= 1e-12 * abs(float(l0))
    assert _rel(g1, g0) <= 1e-12, _rel(g1, g0)


def test_d8_chi128_complex_one_step_hip_vs_plain():
    """One CTMRG step + energy at D=8, chi=128, complex128 (projector blocks up to 1134 x 1134): the HIP
    backend against the plain per-block path, energy <= 1e-10 and gradient <= 1e-8 (relative)."""
    BH, BP = _backends()
    res = []
    for B in (BH, BP):
        print('ctmrg D=8 chi=128 c128, 1 step + energy', B.__name__, flush=True)
        cfg = bs.make_config(B, 'U1', device='cuda:0')
        A, env, _ = C.setup(cfg, D=8, chi=128, dtype='complex128', seed=0)
        res.append(C.energy_fwd_bwd(A, env, C.heisenberg_bond(cfg), 128, 1, segment=1))
    (lh, gh), (lp, gp) = res
    assert abs(float(lh) - float(lp)) <= 1e-10 * abs(float(lp))
    assert _rel(gh, gp) <= 1e-8, _rel(gh, gp)
