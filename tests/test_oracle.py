"""Pin the CPU oracle (oracle/np_backend.py) before trusting it.

Checks: the reference numpy backend's golden vectors, the reference's own C merge worker
(oracle/_ref, when built), dense closed forms, round trips, adjoint identities, finite
differences, the reference torch backend's recorded gradients and the known-answer tests of
the reference test-suite.
"""
import ctypes
import os

import numpy as np
import pytest

import golden_io as G
from oracle import np_backend as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TM_WORKER = os.path.join(ROOT, 'oracle', '_ref', 'libtm_worker.so')


def _cmp_exact(a, b):
    assert a.dtype == b.dtype and a.shape == b.shape
    assert np.array_equal(a, b)


@pytest.mark.parametrize('scenario', G.SCENARIOS)
def test_oracle_copies_match_golden(scenario):
    n = 0
    for fn, args, kw, outs in G.calls(scenario):
        if fn == 'transpose_and_merge':
            _cmp_exact(O.transpose_and_merge(*args), outs[0]); n += 1
        elif fn == 'unmerge':
            _cmp_exact(O.unmerge(*args), outs[0]); n += 1
        elif fn == 'transpose':
            _cmp_exact(O.transpose(*args), outs[0]); n += 1
        elif fn == 'apply_mask':
            _cmp_exact(O.apply_mask(*args), outs[0]); n += 1
    assert n > 0 or scenario == 'fix_signs_kat'


@pytest.mark.parametrize('scenario', G.SCENARIOS)
def test_oracle_dot_matches_golden(scenario):
    for fn, args, kw, outs in G.calls(scenario, 'dot'):
        got = O.dot(*args)
        ref = outs[0]
        assert got.dtype == ref.dtype
        assert np.linalg.norm(got - ref) <= 1e-13 * max(1.0, np.linalg.norm(ref))


@pytest.mark.parametrize('scenario', G.SCENARIOS)
def test_oracle_svd_matches_golden(scenario):
    for fn, args, kw, outs in G.calls(scenario, 'svd'):
        data, meta, sizes = args[:3]
        U, S, V = O.svd(data, meta, sizes)
        np.testing.assert_allclose(S, outs[1], rtol=1e-12, atol=1e-13 * max(1.0, outs[1].max(initial=0)))
        for sl, D, slU, DU, slS, slV, DV in meta:
            a = data[sl[0]:sl[1]].reshape(D)
            u = U[slU[0]:slU[1]].reshape(DU)
            v = V[slV[0]:slV[1]].reshape(DV)
            s = S[slS[0]:slS[1]]
            assert np.linalg.norm(u * s @ v - a) <= 1e-12 * max(1.0, np.linalg.norm(a))
            # gauge-invariant: the recorded factors reconstruct the same block
            ur = outs[0][slU[0]:slU[1]].reshape(DU)
            vr = outs[2][slV[0]:slV[1]].reshape(DV)
            assert np.linalg.norm(ur * outs[1][slS[0]:slS[1]] @ vr - a) <= 1e-12 * max(1.0, np.linalg.norm(a))


def test_oracle_eigh_matches_golden():
    for fn, args, kw, outs in G.calls('eigh_u1_c128', 'eigh'):
        data, meta, sizes = args
        S, U = O.eigh(data, meta, sizes)
        np.testing.assert_allclose(S, outs[0], rtol=1e-12, atol=1e-12)
        for sl, D, slU, DU, slS in meta:
            a = data[sl[0]:sl[1]].reshape(D)
            u = U[slU[0]:slU[1]].reshape(DU)
            w = S[slS[0]:slS[1]]
            assert np.linalg.norm(u * w @ u.conj().T - a) < 1e-12 * np.linalg.norm(a)
            assert np.all(np.diff(w) >= 0)


def test_oracle_fix_signs_kat():
    """Reference tests/tensor/test_svd.py:159-186: the fixed U has real positive largest entries."""
    calls = G.calls('fix_signs_kat')
    fn, args, kw, outs = [c for c in calls if c[0] == 'fix_svd_signs'][0]
    U, V = O.fix_svd_signs(*args)
    np.testing.assert_allclose(U, outs[0], atol=1e-14)
    np.testing.assert_allclose(V, outs[1], atol=1e-14)
    is2 = 0.5 ** 0.5
    expect = np.array([[is2, 1j / 2], [-is2, 1j / 2], [0, is2]])
    t = G.tensors('fix_signs_kat')['U']
    blk0 = t['data'][t['slices'][0][0]:t['slices'][0][1]].reshape(t['D'][0])
    np.testing.assert_allclose(blk0, expect, atol=1e-12)


def test_oracle_fix_signs_golden():
    for fn, args, kw, outs in G.calls('u1xu1_c128', 'fix_svd_signs'):
        U, V = O.fix_svd_signs(*args)
        np.testing.assert_allclose(U, outs[0], atol=1e-14)
        np.testing.assert_allclose(V, outs[1], atol=1e-14)


def _tm_worker():
    if not os.path.exists(TM_WORKER):
        pytest.skip('reference C worker not built (oracle/Makefile needs the reference tree)')
    lib = ctypes.CDLL(TM_WORKER)
    f = lib.tm_worker
    ip = ctypes.POINTER(ctypes.c_int)
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ip, ip, ctypes.c_void_p, ctypes.c_int, ip, ip, ip,
                  ctypes.c_int, ctypes.c_int]
    f.restype = ctypes.c_int
    return f


def _ints(x):
    arr = (ctypes.c_int * max(1, len(x)))(*[int(v) for v in x])
    return arr


@pytest.mark.parametrize('scenario', G.SCENARIOS)
def test_oracle_merge_vs_reference_c_worker(scenario):
    """The numpy restatement of transpose_and_merge agrees bit for bit with the reference's own C
    merge worker (experimental/tm_worker.c, built from the reference tree into oracle/_ref)."""
    f = _tm_worker()
    for fn, args, kw, outs in G.calls(scenario, 'transpose_and_merge'):
        data, order, meta_new, meta_mrg, Dsize = args
        data = np.ascontiguousarray(data)
        out = np.zeros(Dsize, dtype=data.dtype)
        for (tn, Dn, sln), rows in O._groups(meta_new, meta_mrg):
            for _, slo, Do, Dslc, Drsh in rows:
                st = f(out.ctypes.data, sln[0], len(Dn), _ints(Dn), _ints([s[0] for s in Dslc]),
                       data.ctypes.data, slo[0], _ints(Do), _ints(order), _ints(Drsh), len(Do), data.itemsize)
                assert st == 0
        _cmp_exact(out, O.transpose_and_merge(*args))


def _densify(t):
    """Dense array of a golden tensor from its block charges (legs sorted by charge)."""
    s = t['s']
    nsym = len(t['n'])
    legs = []
    for ax in range(len(s)):
        td = {}
        for tt, DD in zip(t['t'], t['D']):
            td[tuple(tt[ax * nsym:(ax + 1) * nsym])] = DD[ax]
        keys = sorted(td)
        offs, acc = {}, 0
        for k in keys:
            offs[k] = acc
            acc += td[k]
        legs.append((offs, acc))
    out = np.zeros(tuple(l[1] for l in legs), dtype=t['data'].dtype)
    for tt, DD, sl in zip(t['t'], t['D'], t['slices']):
        idx = tuple(slice(legs[ax][0][tuple(tt[ax * nsym:(ax + 1) * nsym])],
                          legs[ax][0][tuple(tt[ax * nsym:(ax + 1) * nsym])] + DD[ax]) for ax in range(len(s)))
        out[idx] = t['data'][sl[0]:sl[1]].reshape(DD)
    return out, legs


@pytest.mark.parametrize('scenario', ['mps_u1_fuse_contracted', 'mps_u1_fuse_to_matrix', 'u1xu1_c128'])
def test_golden_transfer_matrix_closed_form(scenario):
    """tensordot(A, A*, ((0,1),(0,1))) equals np.tensordot of the densified tensors (both policies)."""
    ts = G.tensors(scenario)
    A, _ = _densify(ts['A'])
    T, _ = _densify(ts['T'])
    ref = np.tensordot(A, A.conj(), axes=((0, 1), (0, 1)))
    assert T.shape == ref.shape
    assert np.linalg.norm(T - ref) <= 1e-12 * np.linalg.norm(ref)


def test_merge_unmerge_round_trip_and_adjoint():
    rng = np.random.default_rng(0)
    for scenario in ('fuse_u1_fuse_contracted', 'mps_u1_fuse_to_matrix', 'u1xu1_c128'):
        for fn, args, kw, outs in G.calls(scenario, 'transpose_and_merge'):
            data, order, meta_new, meta_mrg, Dsize = args
            x = rng.standard_normal(data.size)
            y = rng.standard_normal(Dsize)
            mx = O.transpose_and_merge(x, order, meta_new, meta_mrg, Dsize)
            aty = O.transpose_and_merge_adjoint(y, order, meta_new, meta_mrg, data.size)
            assert abs(mx @ y - x @ aty) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y)
        for fn, args, kw, outs in G.calls(scenario, 'unmerge'):
            data, meta = args
            x = rng.standard_normal(data.size)
            y = rng.standard_normal(data.size)
            ux = O.unmerge(x, meta)
            aty = O.unmerge_adjoint(y, meta, data.size)
            assert abs(ux @ y - x @ aty) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y)
            # unmerge is a permutation of the elements: its adjoint is its inverse
            np.testing.assert_array_equal(O.unmerge_adjoint(ux, meta, data.size), x)


def _num_grad(f, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = h
        g.flat[i] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def test_oracle_dot_adjoint_fd():
    fn, args, kw, outs = G.calls('mps_u1_fuse_contracted', 'dot')[0]
    A, B, meta, Dsize = args
    rng = np.random.default_rng(1)
    W = rng.standard_normal(Dsize)
    gA, gB = O.dot_adjoint(W, A, B, meta)
    fA = _num_grad(lambda a: O.dot(a, B, meta, Dsize) @ W, A)
    fB = _num_grad(lambda b: O.dot(A, b, meta, Dsize) @ W, B)
    np.testing.assert_allclose(gA, fA, rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(gB, fB, rtol=1e-7, atol=1e-7)


def test_oracle_svd_adjoint_vs_reference_torch_gradients():
    js, arrays = G.load('torch_grads')
    for bi, blk in enumerate(js['svd_blocks']):
        a = arrays[f'svdb{bi}_a']
        g_ref = arrays[f'svdb{bi}_g']
        u, s, vh = np.linalg.svd(a, full_matrices=False)
        if blk['loss'] == 'linear':
            # reference factors, so that the gauge of gU/gV matches
            u, s, vh = arrays[f'svdb{bi}_u'], arrays[f'svdb{bi}_s'], arrays[f'svdb{bi}_vh']
            g = O.svd_block_adjoint(u, s, vh, arrays[f'svdb{bi}_gu'], arrays[f'svdb{bi}_gs'], arrays[f'svdb{bi}_gv'])
        else:
            g = _complex_svd_grad_fd_free(a, arrays[f'svdb{bi}_w'])
        np.testing.assert_allclose(g, g_ref, rtol=1e-8, atol=1e-9)


def _complex_svd_grad_fd_free(a, w):
    """Adjoint of loss(a) = Re tr(U S Vh W) + Im tr(U Vh W) + sum s^3 via the oracle's svd adjoint.

    For a real loss L(z), torch's gradient is conj(dL/dz) (= 2 dL/d conj z / 2 per its convention,
    reported as dL/dRe + i dL/dIm).  With M = W:  Re tr(U S Vh W) gives gU = conj(S Vh W)^T, gVh =
    conj(W U S)^T, gS = Re diag(Vh W U); Im tr(U Vh W) = Re tr(-i U Vh W) gives gU = conj(-i Vh W)^T, gVh =
    conj(-i W U)^T.
    """
    u, s, vh = np.linalg.svd(a, full_matrices=False)
    gu = np.conj((s[:, None] * vh) @ w).T + np.conj(-1j * (vh @ w)).T
    gvh = np.conj(w @ (u * s[None, :])).T + np.conj(-1j * (w @ u)).T
    gs = np.real(np.diag(vh @ w @ u)) + 3 * s ** 2
    return O.svd_block_adjoint(u, s, vh, gu, gs, gvh)


def test_oracle_svd_adjoint_fd_real():
    rng = np.random.default_rng(3)
    for m, n in ((5, 5), (6, 3), (3, 7)):
        a = rng.standard_normal((m, n))
        k = min(m, n)
        cu, cs, cv = rng.standard_normal((m, k)), rng.standard_normal(k), rng.standard_normal((k, n))

        def loss(x):
            u, s, vh = np.linalg.svd(x, full_matrices=False)
            # fix the sign gauge so that the loss is a smooth function of x
            sg = np.sign(u[np.argmax(np.abs(u), axis=0), np.arange(k)])
            u, vh = u * sg, vh * sg[:, None]
            return np.sum(u * cu) + s @ cs + np.sum(vh * cv)
        u, s, vh = np.linalg.svd(a, full_matrices=False)
        sg = np.sign(u[np.argmax(np.abs(u), axis=0), np.arange(k)])
        u, vh = u * sg, vh * sg[:, None]
        g = O.svd_block_adjoint(u, s, vh, cu, cs, cv)
        np.testing.assert_allclose(g, _num_grad(loss, a), rtol=1e-6, atol=1e-6)


def test_oracle_eigh_adjoint_fd():
    rng = np.random.default_rng(4)
    n = 6
    x = rng.standard_normal((n, n))
    a = x + x.T
    cw, cu = rng.standard_normal(n), rng.standard_normal((n, n))

    def loss(y):
        y = (y + y.T) / 2
        w, u = np.linalg.eigh(y)
        sg = np.sign(u[np.argmax(np.abs(u), axis=0), np.arange(n)])
        return w @ cw + np.sum(u * sg * cu)
    w, u = np.linalg.eigh(a)
    sg = np.sign(u[np.argmax(np.abs(u), axis=0), np.arange(n)])
    g = O.eigh_block_adjoint(w, u * sg, cw, cu)
    # the loss symmetrises its argument, so compare the symmetric part
    fd = _num_grad(loss, a)
    np.testing.assert_allclose((g + g.T) / 2, (fd + fd.T) / 2, rtol=1e-6, atol=1e-6)


def test_mask_adjoint_and_embed_mask():
    """apply_mask's adjoint (zero-filled scatter) against the adjoint identity, and embed_mask against
    its closed form: embedding the selected entries back reproduces the masked tensor with zeros in the
    dropped rows, and apply_mask of the embedding is the identity."""
    rng = np.random.default_rng(3)
    n = 0
    for fn, args, kw, outs in G.calls('truncation_u1', 'apply_mask'):
        A, mask, meta, Dsize, axis, ndim = args
        x = rng.standard_normal(A.size)
        y = rng.standard_normal(Dsize)
        mx = O.apply_mask(x, mask, meta, Dsize, axis, ndim)
        aty = O.apply_mask_adjoint(y, mask, meta, A.size, axis, ndim)
        assert abs(mx @ y - x @ aty) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y)
        # embed_mask maps the small blocks (sln side) into zero-filled large ones (sla side)
        emb_meta = tuple((sla, Da, sln, Dn, tm) for sln, Dn, sla, Da, tm in meta)
        e = O.embed_mask(y, mask, emb_meta, A.size, axis, ndim)
        np.testing.assert_array_equal(e, O.apply_mask_adjoint(y, mask, meta, A.size, axis, ndim))
        np.testing.assert_array_equal(O.apply_mask(e, mask, meta, Dsize, axis, ndim), y)
        kept = np.zeros(A.size, dtype=bool)
        kept[O.apply_mask_adjoint(np.arange(1, Dsize + 1, dtype=float), mask, meta, A.size, axis, ndim) > 0] = True
        assert kept.sum() == Dsize and np.all(e[~kept] == 0)
        n += 1
    assert n > 0


def test_oracle_qr_invariants_and_gram_schmidt():
    """Oracle qr: A = Q R, Q^H Q = I, R upper triangular with a real non-negative diagonal, and, for
    full-rank blocks, equal to modified Gram-Schmidt on the columns (the positive-diagonal QR is unique)."""
    rng = np.random.default_rng(21)
    shapes = [(5, 3), (3, 5), (4, 4), (7, 2), (1, 3)]
    for cplx in (False, True):
        meta, off, oq, orr = [], 0, 0, 0
        blocks = []
        for m, n in shapes:
            k = min(m, n)
            a = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
            blocks.append(a)
            meta.append(((off, off + m * n), (m, n), (oq, oq + m * k), (m, k), (orr, orr + k * n), (k, n)))
            off, oq, orr = off + m * n, oq + m * k, orr + k * n
        data = np.concatenate([b.ravel() for b in blocks])
        Q, R = O.qr(data, tuple(meta), (oq, orr))
        for (sl, D, slQ, DQ, slR, DR), a in zip(meta, blocks):
            q = Q[slQ[0]:slQ[1]].reshape(DQ)
            r = R[slR[0]:slR[1]].reshape(DR)
            k = DQ[1]
            assert np.linalg.norm(q @ r - a) <= 1e-13 * np.linalg.norm(a)
            assert np.linalg.norm(q.conj().T @ q - np.eye(k)) <= 1e-13
            assert np.allclose(np.tril(r, -1), 0) and np.all(np.real(np.diag(r)) >= 0)
            assert np.allclose(np.imag(np.diag(r)), 0)
            # modified Gram-Schmidt of the first k columns; the rest of R is Q^H A
            g = np.zeros((D[0], k), dtype=a.dtype)
            for j in range(k):
                v = a[:, j].copy()
                for i in range(j):
                    v = v - g[:, i] * (g[:, i].conj() @ v)
                g[:, j] = v / np.linalg.norm(v)
            assert np.linalg.norm(g - q) <= 1e-12
            assert np.linalg.norm(g.conj().T @ a - r) <= 1e-12 * np.linalg.norm(a)


def test_oracle_transpose_dot_sum_closed_form():
    """Oracle transpose_dot_sum against np.tensordot of each block pair over the contracted legs."""
    import tds_case
    rng = np.random.default_rng(23)
    for ca, cb in ((False, False), (True, False), (True, True)):
        A, B, meta_dot, Ar, Br, Ao, Bo, Dsize = tds_case.make(rng, ca, cb)
        C = O.transpose_dot_sum(A, B, meta_dot, Ar, Br, Ao, Bo, Dsize)
        for sl, Dslc, tab in meta_dot:
            ref = 0
            for ta, tb in tab:
                a = A[Ar[ta][0][0]:Ar[ta][0][1]].reshape(Ar[ta][1])  # (c0, m, c2)
                b = B[Br[tb][0][0]:Br[tb][0][1]].reshape(Br[tb][1])  # (c2, c0, n)
                ref = ref + np.tensordot(a, b, axes=((0, 2), (1, 0)))
            assert np.linalg.norm(C[sl[0]:sl[1]].reshape(Dslc) - ref) <= 1e-13 * max(1.0, np.linalg.norm(ref))


def test_oracle_dot_diag_embed_and_nofusion_match_golden():
    """dot_diag and embed_mask (index and slice masks) bit-exact against the reference numpy backend's
    recorded calls; transpose_dot_sum against its no-fusion calls (list-typed metas) to 1e-13."""
    n = {'dot_diag': 0, 'embed_mask': 0, 'transpose_dot_sum': 0}
    for scenario in ('diag_embed_u1', 'nofusion_u1'):
        for fn, args, kw, outs in G.calls(scenario):
            if fn == 'dot_diag':
                _cmp_exact(O.dot_diag(*args), outs[0])
            elif fn == 'embed_mask':
                _cmp_exact(O.embed_mask(*args), outs[0])
            elif fn == 'transpose_dot_sum':
                got = O.transpose_dot_sum(*args)
                assert got.dtype == outs[0].dtype
                assert np.linalg.norm(got - outs[0]) <= 1e-13 * np.linalg.norm(outs[0])
            else:
                continue
            n[fn] += 1
    assert n == {'dot_diag': 5, 'embed_mask': 2, 'transpose_dot_sum': 5}, n


def test_oracle_dot_diag_adjoint_vs_torch_autograd():
    """The oracle's dot_diag adjoint equals torch's autograd of the reference torch backend's dot_diag
    (backend_torch.py:557-564, restated with torch ops on the CPU) on the recorded calls, real and complex."""
    import torch
    rng = np.random.default_rng(31)
    for fn, args, kw, outs in G.calls('diag_embed_u1', 'dot_diag'):
        A, B, meta, Dsize, axis, nd = args
        cplx = np.iscomplexobj(A) or np.iscomplexobj(B)
        g = rng.standard_normal(Dsize) + (1j * rng.standard_normal(Dsize) if cplx else 0)
        gA, gB = O.dot_diag_adjoint(g, A, B, meta, axis, nd)
        ta = torch.as_tensor(A).requires_grad_(True)
        tb = torch.as_tensor(B).requires_grad_(True)
        dim = [1] * nd
        dim[axis] = -1
        out = torch.cat([(ta[sla[0]:sla[1]].reshape(dim) * tb[slb[0]:slb[1]].reshape(Db)).reshape(-1)
                         for sln, slb, Db, sla in meta])
        (out * torch.as_tensor(np.conj(g))).real.sum().backward()
        np.testing.assert_allclose(gA, ta.grad.numpy(), rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(gB, tb.grad.numpy(), rtol=1e-13, atol=1e-14)


def test_oracle_eigh_adjoint_fd_complex():
    """The oracle's eigh adjoint on a complex Hermitian block (torch's gradient convention, gA = dL/dRe A +
    i dL/dIm A) against central finite differences along random Hermitian directions, for a loss of the
    eigenvalues and of a phase-invariant projector u0 u0^H."""
    rng = np.random.default_rng(6)
    n = 6
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = x + x.conj().T
    cw = rng.standard_normal(n)
    Y = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))

    def loss(y):
        w, u = np.linalg.eigh(y)
        return w @ cw + np.real(u[:, 0].conj() @ Y @ u[:, 0])
    w, u = np.linalg.eigh(a)
    gu = np.zeros((n, n), dtype=complex)
    gu[:, 0] = (Y + Y.conj().T) @ u[:, 0]
    g = O.eigh_block_adjoint(w, u, cw, gu)
    for _ in range(4):
        d = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        d = d + d.conj().T
        h = 1e-6
        fd = (loss(a + h * d) - loss(a - h * d)) / (2 * h)
        an = np.real(np.sum(np.conj(g) * d))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)
