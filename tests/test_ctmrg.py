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


# ------------------------------------------------------------------ unrolled CTMRG + bond energy

def _trivial_env_plaquette(cfg, dtype, seed):
    """Site tensor with virtual charges {-1, 0, 1} (dimension 1 each) and an environment of chi = 1, charge 0,
    all entries 1: the double-layer edges then trace the outer virtual legs of the 2x2 plaquette (ket = bra),
    so the cluster's reduced density matrices are those of the plaquette state with open outer legs."""
    L = lambda s, t, D: bs.leg(cfg, s, t, D)
    vt, vD = (-1, 0, 1), (1, 1, 1)
    legs = [L(-1, vt, vD), L(1, vt, vD), L(1, vt, vD), L(-1, vt, vD), L(1, (0, 1), (1, 1))]
    st, _ = bs._fill_struct(cfg, legs, 0)
    A = bs.from_flat(cfg, legs, 0, bs.seeded_normal(st.size, dtype, seed))
    a = C.double_layer(A)
    aL = [a.get_legs(i) for i in range(4)]
    one = lambda s: L(s, (0,), (1,))
    DL = lambda i, s: bs.Leg(s, aL[i].t, aL[i].D)
    specs = {'tl': [one(-1), one(1)], 'tr': [one(-1), one(1)], 'br': [one(-1), one(1)], 'bl': [one(-1), one(1)],
             't': [one(-1), DL(0, 1), one(1)], 'r': [one(-1), DL(3, 1), one(1)],
             'b': [one(-1), DL(2, -1), one(1)], 'l': [one(-1), DL(1, -1), one(1)]}
    fused_of = {'t': (0, True), 'r': (3, True), 'b': (2, False), 'l': (1, False)}
    env = {}
    for k in C.Env.NAMES:
        st2, _ = bs._fill_struct(cfg, specs[k], 0)
        x = bs.from_flat(cfg, specs[k], 0, np.ones(st2.size))
        if k in fused_of:
            ia, flip = fused_of[k]
            x = x._replace(hfs=(x.hfs[0], a.hfs[ia].conj() if flip else a.hfs[ia], x.hfs[2]))
        env[k] = x
    return A, C.Env(**env)


def _dense_plaquette_energy(Ad):
    """<S.S> on the horizontal (s00, s01) plus the vertical (s10, s00) bond of the 2x2 plaquette state
    psi = s00 s01 s10 s11 (inner bonds contracted, outer legs open), by dense linear algebra."""
    psi = np.einsum('abcdp,edfgq,chijr,fjkls->pqrsabeghikl', Ad, Ad, Ad, Ad).reshape(2, 2, 2, 2, -1)
    Sz = np.diag([-0.5, 0.5])
    Sp = np.array([[0., 0.], [1., 0.]])
    H = (np.kron(Sz, Sz) + 0.5 * (np.kron(Sp, Sp.T) + np.kron(Sp.T, Sp))).reshape(2, 2, 2, 2)
    rho_h = np.einsum('abcdo,efcdo->abef', psi, psi.conj())   # (i0, i1, o0, o1)
    rho_v = np.einsum('abcdo,ebgdo->cage', psi, psi.conj())   # (i2, i0, o2, o0)
    e = lambda r: np.einsum('abef,efab->', r, H) / np.einsum('abab->', r)
    return e(rho_h) + e(rho_v)


@pytest.mark.parametrize('dtype', ['float64', 'complex128'])
def test_bond_energy_matches_dense_plaquette(dtype):
    """The bond energy contraction (open double layer, two-site density matrices of the 2x2 cluster, the
    U(1) S.S tensor) against an independent dense computation on a plaquette."""
    for B in (NB, _plain()):
        cfg = bs.make_config(B, 'U1')
        A, env = _trivial_env_plaquette(cfg, dtype, seed=5)
        e = complex(C.bond_energy(env, A, C.heisenberg_bond(cfg)))
        ref = _dense_plaquette_energy(A.to_dense())
        assert abs(e - ref) <= 1e-12 * max(1.0, abs(ref)), (e, ref)
        assert abs(ref.imag) <= 1e-12


def test_open_double_layer_traces_to_double_layer():
    cfg = bs.make_config(NB, 'U1')
    A, env, sz = C.setup(cfg, D=3, chi=12, seed=1)
    ao = C.double_layer_open(A)
    pt = (0, 1)
    one = bs.from_flat(cfg, [bs.leg(cfg, 1, pt, (1, 1)), bs.leg(cfg, -1, pt, (1, 1))], 0, np.array([1.0, 1.0]))
    for op, ref in ((one, C.double_layer(A)), (sz, C.double_layer(A, sz))):
        got = bs.tensordot(ao, op, axes=((4, 5), (1, 0)))
        assert got.struct.t == ref.struct.t and got.struct.D == ref.struct.D
        assert np.allclose(got.to_dense(), ref.to_dense(), rtol=0, atol=1e-14)


@pytest.mark.parametrize('D,chi,dtype', [(2, 8, 'float64'), (3, 12, 'complex128')])
def test_unrolled_energy_oracle_vs_plain_and_checkpoint_cpu(D, chi, dtype):
    """Three CTMRG steps + bond energy: numpy oracle vs the plain torch path; the gradient with checkpointed
    segments of two steps equals the gradient without checkpointing."""
    cn = bs.make_config(NB, 'U1')
    A, env, _ = C.setup(cn, D=D, chi=chi, dtype=dtype, seed=1)
    v_np = float(np.real(C.energy_loss_fn(A, env, C.heisenberg_bond(cn), chi, 3)))
    cp = bs.make_config(_plain(), 'U1', device='cpu')
    A, env, _ = C.setup(cp, D=D, chi=chi, dtype=dtype, seed=1)
    H = C.heisenberg_bond(cp)
    l0, g0 = C.energy_fwd_bwd(A, env, H, chi, 3, segment=0)
    l2, g2 = C.energy_fwd_bwd(A, env, H, chi, 3, segment=2)
    assert abs(float(l0) - v_np) <= 1e-10 * max(1.0, abs(v_np))
    assert float(l0) == float(l2)
    assert float((g0 - g2).norm()) <= 1e-12 * float(g0.norm())


@pytest.mark.parametrize('dtype', ['float64', 'complex128'])
def test_batched_projector_svds_equal_sequential_cpu(dtype):
    """The two projector SVDs of a move in one batched call (blocksparse.svd_many) give the same loss and
    gradient as the reference's one-after-the-other calls."""
    B = _plain()
    cp = bs.make_config(B, 'U1', device='cpu')
    A, env, sz = C.setup(cp, D=3, chi=12, dtype=dtype, seed=4)
    l0, g0 = C.fwd_bwd(A, env, sz, 12)
    l1, g1 = C.fwd_bwd(A, env, sz, 12, batch=True)
    assert abs(float(l0) - float(l1)) <= 1e-13 * max(1.0, abs(float(l0)))
    assert float((g0 - g1).norm()) <= 1e-12 * float(g0.norm())
    cn = bs.make_config(NB, 'U1')
    A, env, sz = C.setup(cn, D=3, chi=12, dtype=dtype, seed=4)
    v0, v1 = C.loss_fn(A, env, sz, 12), C.loss_fn(A, env, sz, 12, batch=True)
    assert abs(complex(v0) - complex(v1)) <= 1e-13 * max(1.0, abs(complex(v0)))


@pytest.mark.parametrize('k', [0, 1, 2])
def test_rvb_kagome_matches_reference_ctmrg(k):
    """The reference's own CTMRG (yastn.tn.fpeps EnvCTM.ctmrg_, torch backend on the CPU; golden vectors) on
    the Kagome RVB D=3 U(1) complex state of tests/peps/test_ad.py, chi=64, multiplet truncation, use_qr=False:
    1 sweep from the 'eye' and from the 'dl' environment, 2 sweeps from 'dl'.  The restated harness (2x2
    corners, proj_corners, _update_env_, expand_outward_, rdm1x1) on the plain torch path gives the same
    loss (<= 1e-12) and gradient (<= 1e-9 relative): a wrong projector leg, a wrong environment update or a
    wrong truncation would not."""
    import rvb_case as R
    sweeps, init, loss_ref, g_ref = R.runs()[k]
    cfg = bs.make_config(_plain(), 'U1', device='cpu')
    A = R.state(cfg)
    x = A.data.detach().clone().requires_grad_(True)
    loss = C.rdm_loss(A._replace(data=x), R.CHI, sweeps, init, R.svd_opts(), block=R.BLOCK)
    loss.backward()
    assert abs(float(loss) - loss_ref) <= 1e-12
    g = x.grad.numpy()
    assert np.abs(g - g_ref).max() <= 1e-9 * np.abs(g_ref).max()


@pytest.mark.parametrize('k', [0, 1])
def test_rvb_kagome_qr_projectors_match_reference(k):
    """use_qr=True (the reference's default, _env_ctm.py:1095-1096 ...): the projectors come from the R
    factors of the half-systems.  The loss equals the reference CTMRG's with use_qr=True (golden) and with
    use_qr=False to 1e-12, and the gradient is non-finite exactly when the reference's is (this state's
    half-systems are rank-deficient)."""
    import rvb_case as R
    sweeps, init, loss_qr, finite = R.runs_qr()[k]
    _, _, loss_ref, _ = R.runs()[k + 1]
    cfg = bs.make_config(_plain(), 'U1', device='cpu')
    A = R.state(cfg)
    x = A.data.detach().clone().requires_grad_(True)
    opts = dict(R.svd_opts(), use_qr=True)
    loss = C.rdm_loss(A._replace(data=x), R.CHI, sweeps, init, opts, block=R.BLOCK)
    assert abs(float(loss) - loss_qr) <= 1e-12 and abs(float(loss) - loss_ref) <= 1e-12
    loss.backward()
    assert bool(np.isfinite(x.grad.numpy()).all()) == finite


def test_qr_projectors_same_gradient_and_fd_cpu():
    """On a random state (full-rank half-systems) use_qr=True and use_qr=False are the same function: equal
    loss and gradient, and the use_qr gradient matches central finite differences (unbroadened adjoint)."""
    B = _plain()
    B.set_ad_broadening(1e-30)
    try:
        cfg = bs.make_config(B, 'U1', device='cpu')
        A, env, sz = C.setup(cfg, D=2, chi=8, seed=2)
        out = []
        for q in (False, True):
            x = A.data.detach().clone().requires_grad_(True)
            loss = C.loss_fn(A._replace(data=x), env, sz, 8, {'use_qr': q})
            loss.backward()
            out.append((float(loss), x.grad))
        (l0, g0), (l1, g1) = out
        assert abs(l0 - l1) <= 1e-12 * max(1.0, abs(l0))
        assert float((g0 - g1).norm()) <= 1e-8 * float(g0.norm())
        rng = np.random.default_rng(5)
        for _ in range(3):
            d = torch.as_tensor(rng.standard_normal(A.data.numel()))
            f = lambda h: float(C.loss_fn(A._replace(data=A.data + h * d), env, sz, 8, {'use_qr': True}))
            h = 1e-6
            fd = (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)
            an = float((g1 * d).sum())
            assert abs(fd - an) <= 1e-6 * abs(an)
    finally:
        B.set_ad_broadening(1e-12)


def test_rvb_kagome_20_sweeps_known_answer():
    """tests/peps/test_ad.py:264-280 (test_Kagome_RVB_D3_U1_sym_vs_pepstorch): 20 sweeps from 'dl', the
    gradient equals grad_expected (rtol 1e-3, atol 1e-5, as the reference asserts); it also equals the
    reference CTMRG's own gradient to 1e-8 (relative)."""
    import rvb_case as R
    sweeps, init, loss_ref, g_ref = R.runs()[3]
    assert (sweeps, init) == (20, 'dl')
    cfg = bs.make_config(_plain(), 'U1', device='cpu')
    A = R.state(cfg)
    x = A.data.detach().clone().requires_grad_(True)
    loss = C.rdm_loss(A._replace(data=x), R.CHI, sweeps, init, R.svd_opts(), block=R.BLOCK)
    loss.backward()
    g = x.grad.numpy()
    assert np.allclose(g, R.grad_expected(), rtol=1e-3, atol=1e-5)
    assert abs(float(loss) - loss_ref) <= 1e-10
    assert np.abs(g - g_ref).max() <= 1e-8 * np.abs(g_ref).max()


def test_static_truncation_mask_equals_eager_rule():
    """truncation_mask_static(S, D_block) keeps the same entries as the reference rule
    truncation_mask(S, tol=0, D_block=D_block, D_total=sum(D_block)) when the kept values are positive, and
    its device counters flag a zero kept value and a dropped value above the smallest kept one."""
    cfg = bs.make_config(_plain(), 'U1', device='cpu')
    L = lambda s, t, D: bs.leg(cfg, s, t, D)
    a = bs.rand(cfg, [L(1, (-1, 0, 1), (4, 6, 4)), L(-1, (-1, 0, 1), (5, 7, 5))], n=0, seed=2)
    U, S, V = bs.svd(a, axes=(0, 1))
    D_block = {(-1,): 2, (0,): 3, (1,): 1}
    want = bs.truncation_mask(S, tol=0, D_block=D_block, D_total=6)
    check = torch.zeros(2, dtype=torch.int64)
    got = bs.truncation_mask_static(S, D_block, check)
    assert torch.equal(got.data, want.data)
    assert {k: v.tolist() for k, v in got.host_index.items()} == {(-1,): [0, 1], (0,): [0, 1, 2], (1,): [0]}
    _, s1, _ = bs.apply_mask(got, U, S, V, axes=(-1, 0, 0))
    _, s2, _ = bs.apply_mask(want, U, S, V, axes=(-1, 0, 0))
    assert torch.equal(s1.data, s2.data) and s1.struct == s2.struct
    # global top-6 here differs from the fixed 2/3/1 split unless the spectrum happens to agree
    top = np.sort(S.data.numpy())[::-1][:6]
    kept = np.sort(S.data.numpy()[want.data.numpy()])[::-1]
    assert check.tolist() == [0, int(not np.array_equal(top, kept))]
    S0 = S._replace(data=torch.where(torch.arange(S.size) == 0, torch.zeros(()), S.data))
    check.zero_()
    bs.truncation_mask_static(S0, D_block, check)
    assert check[0] == 1


def test_dict_serialization_and_split_combine():
    """Tensor.to_dict / from_dict at levels 0 and 2, and split_data_and_meta / combine_data_and_meta (the
    reference's _split_combine_dict.py semantics: data arrays in sorted-key order, replaced by positions)."""
    cfg = bs.make_config(_plain(), 'U1', device='cpu')
    A, env, sz = C.setup(cfg, D=2, chi=8, seed=1)
    a = C.double_layer(A)
    for level in (0, 2):
        d = a.to_dict(level)
        b = bs.Tensor.from_dict(d, cfg)
        assert b.struct == a.struct and b.slices == a.slices and b.hfs == a.hfs and b.trans == a.trans
        assert torch.equal(torch.as_tensor(b.data), a.data)
    d = env.to_dict(level=0)
    data, meta = bs.split_data_and_meta(d)
    assert len(data) == 8 and meta['env']['b']['data'] == 0 and meta['env']['tl']['data'] == 6  # sorted names
    back = C.Env.from_dict(bs.combine_data_and_meta(data, meta), cfg)
    for x, y in zip(back.tensors(), env.tensors()):
        assert x.struct == y.struct and torch.equal(x.data, y.data)
    one, m1 = bs.split_data_and_meta(A.to_dict(0), squeeze=True)
    assert one is A.data and bs.combine_data_and_meta(one, m1)['data'] is A.data


@pytest.mark.parametrize('mode', ['nonreentrant', 'reentrant'])
def test_checkpoint_move_equals_plain(mode):
    """Per-move checkpointing through split_data_and_meta (the reference's checkpoint_move) gives the same
    loss and gradient as the unrolled step (<= 1e-12)."""
    cfg = bs.make_config(_plain(), 'U1', device='cpu')
    A, env, sz = C.setup(cfg, D=2, chi=8, seed=1)
    out = []
    for cm in (False, mode):
        x = A.data.detach().clone().requires_grad_(True)
        Ax = A._replace(data=x)
        e2 = C.ctm_step(env, C.double_layer(Ax), 8, checkpoint_move=cm, A=Ax)
        loss = C.cluster_expectation(e2, Ax, sz)
        loss.backward()
        out.append((float(loss), x.grad))
    assert abs(out[0][0] - out[1][0]) <= 1e-12
    assert float((out[0][1] - out[1][1]).norm()) <= 1e-12 * float(out[0][1].norm())


def test_functorch_vjp_of_ctm_step_plain_cpu():
    """torch.func.vjp of one whole CTM step f(C, a) (environment and site buffers in, environment buffer
    out, static-structure truncation), as the reference's fixed-point AD takes it (fixed_pt.py:853), equals
    torch.autograd.grad of the same step (plain path on the CPU; the HIP path is checked on the GPU)."""
    cfg = bs.make_config(_plain(), 'U1', device='cpu')
    A, env, sz = C.setup(cfg, D=2, chi=8, seed=1)
    opts = C.static_svd_opts(8)
    tmpl = C.ctm_step(env, C.double_layer(A), 8, opts)
    f = C.ctm_step_flat(A, tmpl, 8, opts)
    C0 = C.env_flat(tmpl).detach()
    out, vjp = torch.func.vjp(f, C0, A.data)
    v = torch.as_tensor(np.random.default_rng(3).standard_normal(out.numel()))
    gC, gA = vjp(v)
    c = C0.clone().requires_grad_(True)
    x = A.data.clone().requires_grad_(True)
    rC, rA = torch.autograd.grad(f(c, x), (c, x), grad_outputs=v)
    assert torch.allclose(gC, rC, rtol=1e-12, atol=1e-13) and torch.allclose(gA, rA, rtol=1e-12, atol=1e-13)
