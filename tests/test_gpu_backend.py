"""GPU parity of the HIP backend against the reference numpy backend's golden vectors and the oracle.

Tolerances (stated per result):
  copies (transpose_and_merge, unmerge, apply_mask, adjoints)   bit-exact
  dot / dot adjoint                                             <= 1e-12 relative Frobenius
  svd singular values                                           <= 1e-10 relative to the block maximum
  svd factors                                                   ||U S Vh - A|| <= 1e-12 ||A||, orthonormal to 1e-12
  eigh eigenvalues                                              <= 1e-12 ||A||
  svd adjoint vs the oracle's per-block formula                  <= 1e-9 relative
"""
import numpy as np
import pytest
import torch

import golden_io as G
from oracle import np_backend as O

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')


@pytest.fixture(scope='module')
def B():
    from yastn_amd import backend_hip
    return backend_hip


def _t(x):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV)


def _n(x):
    return x.detach().cpu().resolve_conj().numpy()


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize('scenario', G.SCENARIOS)
def test_copies_bit_exact(B, scenario):
    print('scenario', scenario, flush=True)
    for fn, args, kw, outs in G.calls(scenario):
        if fn == 'transpose_and_merge':
            data, order, meta_new, meta_mrg, Dsize = args
            got = B.transpose_and_merge(_t(data), order, meta_new, meta_mrg, Dsize)
        elif fn == 'unmerge':
            got = B.unmerge(_t(args[0]), args[1])
        elif fn == 'apply_mask':
            A, mask, meta, Dsize, axis, ndim = args
            got = B.apply_mask(_t(A), {k: torch.as_tensor(v) for k, v in mask.items()}, meta, Dsize, axis, ndim)
        else:
            continue
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_n(got), outs[0])


def _random_box_jobs(rng, ranks, max_dim):
    """Permuted copies of C-ordered source blocks of the given ranks (sides 2 ... max_dim up to rank 4, 2 ... 5
    above) into consecutive C-ordered destination boxes, plus zero-fill boxes between them:
    (jobs, src_numel, dst_numel)."""
    jobs, so, do = [], 0, 0
    for r in ranks:
        shape = tuple(int(x) for x in rng.integers(2, (max_dim if r <= 4 else 5) + 1, size=r))
        perm = rng.permutation(r)
        cs = np.cumprod((1,) + shape[::-1])[:-1][::-1]  # C strides of the source block
        dshape = tuple(shape[k] for k in perm)
        dcs = np.cumprod((1,) + dshape[::-1])[:-1][::-1]
        jobs.append((so, do, dshape, tuple(int(cs[k]) for k in perm), tuple(int(x) for x in dcs)))
        n = int(np.prod(shape))
        so, do = so + n, do + n
        gap = int(rng.integers(0, 3000))
        if gap:
            jobs.append((-1, do, (gap,), (0,), (1,)))
            do += gap
    return jobs, so, do


def _np_gather(jobs, src, dst_numel):
    out = np.full(dst_numel, np.nan, dtype=src.dtype)
    for src_off, dst_off, shape, ss, ds in jobs:
        idx = np.indices(shape).reshape(len(shape), -1)
        d = dst_off + (np.array(ds)[:, None] * idx).sum(0)
        out[d] = 0 if src_off < 0 else src[src_off + (np.array(ss)[:, None] * idx).sum(0)]
    return out


@pytest.mark.parametrize('dtype', [np.float64, np.complex128])
def test_box_gather_every_rank(dtype):
    """The box-gather engine on permuted boxes of rank 1 ... 8 (each rank its own specialised path; boxes
    span several 2048-element tiles, so the per-thread index stepping crosses tile and dimension
    boundaries), the 32 x 32 transpose tiles, zero fills, and the paired launch: bit-exact against numpy."""
    from yastn_amd import plans as P, _functions as F
    rng = np.random.default_rng(7)
    for trial in range(3):
        jobs, sn, dn = _random_box_jobs(rng, [1, 2, 3, 4, 5, 6, 7, 8, 3, 2], 5 if trial < 2 else 9)
        src = rng.standard_normal(sn)
        if dtype == np.complex128:
            src = src + 1j * rng.standard_normal(sn)
        plan = P.GatherPlan(jobs, sn, dn, DEV)
        print('trial', trial, 'jobs', plan.njobs, 'tiles', plan.ntiles, 'kinds', sorted(set(plan.h_jobs[:, 3])), flush=True)
        out = torch.empty(dn, dtype=_t(src).dtype, device=DEV)
        F.run_gather('yastn_hip_transpose_and_merge', plan, out, _t(src))
        np.testing.assert_array_equal(_n(out), _np_gather(jobs, src, dn))
        # the same two plans in one paired launch
        jobs2, sn2, dn2 = _random_box_jobs(rng, [2, 5, 3], 6)
        src2 = rng.standard_normal(sn2).astype(src.dtype)
        plan2 = P.GatherPlan(jobs2, sn2, dn2, DEV)
        pp = P.pair_plan(plan, plan2, DEV)
        o1 = torch.empty(dn, dtype=out.dtype, device=DEV)
        o2 = torch.empty(dn2, dtype=out.dtype, device=DEV)
        from yastn_amd import _lib as L
        s1, s2 = _t(src), _t(src2)  # held: a raw pointer does not keep its tensor alive
        L.call('yastn_hip_transpose_and_merge_pair', o1.data_ptr(), dn, s1.data_ptr(), sn, o2.data_ptr(), dn2,
               s2.data_ptr(), sn2, F.dtype_code(o1), *pp.args(), F.stream_of(o1))
        np.testing.assert_array_equal(_n(o1), _np_gather(jobs, src, dn))
        np.testing.assert_array_equal(_n(o2), _np_gather(jobs2, src2, dn2))


@pytest.mark.parametrize('scenario', ['mps_u1_fuse_contracted', 'u1xu1_c128', 'fuse_u1_fuse_to_matrix'])
def test_merge_pair_equals_two_merges(B, scenario):
    """transpose_and_merge_pair (both operands of a tensordot in one launch) against the golden merges
    bit for bit, forward, and its adjoint against the single-merge adjoints bit for bit."""
    merges = [args for fn, args, kw, outs in G.calls(scenario) if fn == 'transpose_and_merge']
    refs = [outs[0] for fn, args, kw, outs in G.calls(scenario) if fn == 'transpose_and_merge']
    pairs = [(i, j) for i in range(len(merges)) for j in range(i + 1, min(len(merges), i + 3))
             if merges[i][0].dtype == merges[j][0].dtype]
    assert pairs
    for i, j in pairs:
        print('pair', i, j, flush=True)
        a, b = merges[i], merges[j]
        da = _t(a[0]).requires_grad_(True)
        db = _t(b[0]).requires_grad_(True)
        oa, ob = B.transpose_and_merge_pair(da, *a[1:], db, *b[1:])
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_n(oa.detach()), refs[i])
        np.testing.assert_array_equal(_n(ob.detach()), refs[j])
        ga = torch.randn(oa.shape, dtype=torch.float64, device=DEV).to(oa.dtype)
        gb = torch.randn(ob.shape, dtype=torch.float64, device=DEV).to(ob.dtype)
        gpa, gpb = torch.autograd.grad((oa, ob), (da, db), (ga, gb))
        sa = _t(a[0]).requires_grad_(True)
        sb = _t(b[0]).requires_grad_(True)
        (gsa,) = torch.autograd.grad(B.transpose_and_merge(sa, *a[1:]), sa, ga)
        (gsb,) = torch.autograd.grad(B.transpose_and_merge(sb, *b[1:]), sb, gb)
        np.testing.assert_array_equal(_n(gpa), _n(gsa))
        np.testing.assert_array_equal(_n(gpb), _n(gsb))


@pytest.mark.parametrize('scenario', G.SCENARIOS)
def test_dot(B, scenario):
    for fn, args, kw, outs in G.calls(scenario, 'dot'):
        A, Bd, meta, Dsize = args
        got = _n(B.dot(_t(A), _t(Bd), meta, Dsize))
        assert got.dtype == outs[0].dtype
        assert _rel(got, outs[0]) <= 1e-12, _rel(got, outs[0])


def test_dot_each_tile_config(B):
    from yastn_amd import plans as P
    for scenario in ('mps_u1_fuse_contracted', 'u1xu1_c128', 'fuse_u1_fuse_to_matrix'):
        for fn, args, kw, outs in G.calls(scenario, 'dot'):
            A, Bd, meta, Dsize = args
            dt = np.promote_types(A.dtype, Bd.dtype)
            tA, tB = _t(A.astype(dt)), _t(Bd.astype(dt))
            plan = P.dot_plan(meta, A.size, Bd.size, Dsize, DEV)
            keep = plan.cfg
            for cfg in P.L.TILE_DIMS:
                if not plan.applicable(cfg):
                    continue
                plan.set_cfg(cfg)
                got = _n(B.dot(tA, tB, meta, Dsize))
                assert _rel(got, outs[0]) <= 1e-12, (scenario, cfg)
            plan.set_cfg(keep)


def _random_dot_meta(shapes):
    """meta_dot of one GEMM per block, blocks laid out one after the other in A, B and C."""
    meta, oa, ob, oc = [], 0, 0, 0
    for M, K, N in shapes:
        meta.append(((oc, oc + M * N), (M, N), (oa, oa + M * K), (M, K), (ob, ob + K * N), (K, N)))
        oa, ob, oc = oa + M * K, ob + K * N, oc + M * N
    return tuple(meta), oa, ob, oc


@pytest.mark.parametrize('dtype', [np.float64, np.complex128])
def test_dot_and_adjoint_every_tile_config_ragged(B, dtype):
    """Every tile shape on ragged sector GEMMs: M, N off every tile multiple, K of 1, below, at and past
    the LDS slices (tails), the Heff2 sector sizes
    (99 x 473 x 99, 106 x 106 x 318): the dot and both adjoint products (conjugated, transposed operand
    strides) against numpy, <= 1e-12 relative."""
    from yastn_amd import plans as P
    shapes = [(1, 1, 1), (3, 17, 5), (16, 16, 16), (17, 33, 15), (33, 47, 65), (64, 64, 64), (65, 129, 63),
              (99, 473, 99), (106, 106, 318), (2, 500, 3), (130, 5, 70)]
    meta, na, nb, nc = _random_dot_meta(shapes)
    rng = np.random.default_rng(11)
    rnd = lambda n: rng.standard_normal(n) + (1j * rng.standard_normal(n) if dtype == np.complex128 else 0)
    A, Bd, gC = rnd(na).astype(dtype), rnd(nb).astype(dtype), rnd(nc).astype(dtype)
    want = O.dot(A, Bd, meta, nc)
    wA, wB = O.dot_adjoint(gC, A, Bd, meta)
    plan = P.dot_plan(meta, na, nb, nc, DEV)
    adj = P.dot_adjoint_plan(meta, na, nb, nc, DEV)
    keep = plan.cfg, adj.cfg
    plan.tuned = adj.tuned = True
    try:
        for cfg in P.L.TILE_DIMS:
            if not (plan.applicable(cfg) and adj.applicable(cfg)):
                continue
            print('tile config', cfg, np.dtype(dtype).name, flush=True)
            plan.set_cfg(cfg)
            adj.set_cfg(cfg)
            tA, tB = _t(A).requires_grad_(True), _t(Bd).requires_grad_(True)
            C = B.dot(tA, tB, meta, nc)
            assert _rel(_n(C), want) <= 1e-12, cfg
            C.backward(_t(gC))
            assert _rel(_n(tA.grad), wA) <= 1e-12, cfg
            assert _rel(_n(tB.grad), wB) <= 1e-12, cfg
    finally:
        plan.set_cfg(keep[0])
        adj.set_cfg(keep[1])


@pytest.mark.parametrize('scenario', ['mps_u1_fuse_contracted', 'u1xu1_c128', 'fuse_u1_fuse_contracted'])
def test_dot_adjoint(B, scenario):
    rng = np.random.default_rng(7)
    for fn, args, kw, outs in G.calls(scenario, 'dot'):
        A, Bd, meta, Dsize = args
        dt = np.promote_types(A.dtype, Bd.dtype)
        tA = _t(A.astype(dt)).requires_grad_(True)
        tB = _t(Bd.astype(dt)).requires_grad_(True)
        gC = rng.standard_normal(Dsize) + (1j * rng.standard_normal(Dsize) if dt == np.complex128 else 0)
        C = B.dot(tA, tB, meta, Dsize)
        C.backward(_t(gC.astype(dt)))
        rA, rB = O.dot_adjoint(gC.astype(dt), A.astype(dt), Bd.astype(dt), meta)
        assert _rel(_n(tA.grad), rA) <= 1e-12
        assert _rel(_n(tB.grad), rB) <= 1e-12


@pytest.mark.parametrize('scenario', ['mps_u1_fuse_contracted', 'u1xu1_c128', 'fuse_u1_fuse_to_matrix'])
def test_merge_and_unmerge_adjoints(B, scenario):
    rng = np.random.default_rng(8)
    for fn, args, kw, outs in G.calls(scenario):
        if fn == 'transpose_and_merge':
            data, order, meta_new, meta_mrg, Dsize = args
            x = _t(data).requires_grad_(True)
            y = B.transpose_and_merge(x, order, meta_new, meta_mrg, Dsize)
            g = rng.standard_normal(Dsize).astype(data.dtype)
            y.backward(_t(g))
            ref = O.transpose_and_merge_adjoint(g, order, meta_new, meta_mrg, data.size)
            np.testing.assert_array_equal(_n(x.grad), ref)
        elif fn == 'unmerge':
            data, meta = args
            x = _t(data).requires_grad_(True)
            y = B.unmerge(x, meta)
            g = rng.standard_normal(data.size).astype(data.dtype)
            y.backward(_t(g))
            np.testing.assert_array_equal(_n(x.grad), O.unmerge_adjoint(g, meta, data.size))


def _check_svd(data, meta, U, S, V, S_ref=None):
    for sl, D, slU, DU, slS, slV, DV in meta:
        a = data[sl[0]:sl[1]].reshape(D)
        u = U[slU[0]:slU[1]].reshape(DU)
        v = V[slV[0]:slV[1]].reshape(DV)
        s = S[slS[0]:slS[1]]
        k = len(s)
        assert np.all(np.diff(s) <= 0), 'descending'
        assert np.linalg.norm(u * s @ v - a) <= 1e-12 * max(np.linalg.norm(a), 1e-300)
        assert np.linalg.norm(u.conj().T @ u - np.eye(k)) <= 1e-12 * max(1, k)
        assert np.linalg.norm(v @ v.conj().T - np.eye(k)) <= 1e-12 * max(1, k)
        if S_ref is not None:
            sr = S_ref[slS[0]:slS[1]]
            assert np.max(np.abs(s - sr)) <= 1e-10 * max(sr.max(initial=0), 1e-300)


@pytest.mark.parametrize('scenario', ['mps_u1_fuse_contracted', 'mps_u1_fuse_to_matrix', 'u1xu1_c128',
                                      'truncation_u1', 'fix_signs_kat'])
def test_svd(B, scenario):
    for fn, args, kw, outs in G.calls(scenario, 'svd'):
        data, meta, sizes = args[:3]
        print('svd', scenario, [m[1] for m in meta], flush=True)
        U, S, V = B.svd(_t(data), meta, sizes)
        _check_svd(data, meta, _n(U), _n(S), _n(V), outs[1])


def test_fix_svd_signs(B):
    for scenario in ('fix_signs_kat', 'u1xu1_c128'):
        for fn, args, kw, outs in G.calls(scenario, 'fix_svd_signs'):
            Ud, Vd, meta = args
            U2, V2 = B.fix_svd_signs(_t(Ud), _t(Vd), meta)
            np.testing.assert_allclose(_n(U2), outs[0], atol=1e-14)
            np.testing.assert_allclose(_n(V2), outs[1], atol=1e-14)
            # differentiable route gives the same values
            tU = _t(Ud).requires_grad_(True)
            U3, V3 = B.fix_svd_signs(tU, _t(Vd), meta)
            np.testing.assert_allclose(_n(U3), outs[0], atol=1e-14)
            np.testing.assert_allclose(_n(V3), outs[1], atol=1e-14)
            # its adjoint against torch autograd of the per-block torch formulation (same pivots)
            from yastn_amd import backend_plain as BPl
            gu = torch.randn(U3.shape, dtype=torch.float64, device=DEV).to(U3.dtype)
            gv = torch.randn(V3.shape, dtype=torch.float64, device=DEV).to(V3.dtype)
            tU, tV = _t(Ud).requires_grad_(True), _t(Vd).requires_grad_(True)
            hU, hV = torch.autograd.grad(B.fix_svd_signs(tU, tV, meta), (tU, tV), (gu, gv))
            pU, pV = _t(Ud).requires_grad_(True), _t(Vd).requires_grad_(True)
            rU, rV = torch.autograd.grad(BPl.fix_svd_signs(pU, pV, meta), (pU, pV), (gu, gv))
            assert _rel(_n(hU), _n(rU)) <= 1e-13 and _rel(_n(hV), _n(rV)) <= 1e-13
            # deterministic: the same adjoint twice, bit for bit
            tU2, tV2 = _t(Ud).requires_grad_(True), _t(Vd).requires_grad_(True)
            hU2, hV2 = torch.autograd.grad(B.fix_svd_signs(tU2, tV2, meta), (tU2, tV2), (gu, gv))
            np.testing.assert_array_equal(_n(hU), _n(hU2))
            np.testing.assert_array_equal(_n(hV), _n(hV2))


def test_eigh(B):
    for fn, args, kw, outs in G.calls('eigh_u1_c128', 'eigh'):
        data, meta, sizes = args
        S, U = B.eigh(_t(data), meta, sizes)
        S, U = _n(S), _n(U)
        for sl, D, slU, DU, slS in meta:
            a = data[sl[0]:sl[1]].reshape(D)
            u = U[slU[0]:slU[1]].reshape(DU)
            w = S[slS[0]:slS[1]]
            assert np.all(np.diff(w) >= 0)
            assert np.max(np.abs(w - outs[0][slS[0]:slS[1]])) <= 1e-12 * np.linalg.norm(a)
            assert np.linalg.norm(u * w @ u.conj().T - a) <= 1e-12 * np.linalg.norm(a)


def test_svd_random_shapes(B):
    """Single blocks of assorted shapes (LDS and workspace instances; the workspace instance is QR
    preconditioned, with the Householder panel in LDS, or in place for 1300 rows, and 700 complex
    rows), real and complex."""
    rng = np.random.default_rng(9)
    for (m, n) in ((1, 1), (1, 7), (7, 1), (2, 2), (17, 5), (5, 17), (64, 64), (90, 40), (33, 130), (150, 150),
                   (300, 170), (170, 300), (440, 440), (1300, 40), (40, 1300), (700, 300)):
        for cplx in (False, True):
            a = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
            k = min(m, n)
            meta = (((0, m * n), (m, n), (0, m * k), (m, k), (0, k), (0, k * n), (k, n)),)
            print('svd shape', m, n, cplx, flush=True)
            U, S, V = B.svd(_t(a.ravel()), meta, (m * k, k, k * n))
            s_ref = np.linalg.svd(a, compute_uv=False)
            _check_svd(a.ravel(), meta, _n(U), _n(S), _n(V), s_ref)


def _sector_sizes(chi, d2):
    """Sector sizes of the fused chi x D^2 leg: charges q1 + q2 with dims chi(q1) D^2(q2)."""
    dim = lambda d, q: d.get(abs(q), 0)
    out = [sum(dim(chi, q1) * dim(d2, q - q1) for q1 in range(-8, 9)) for q in range(-16, 17)]
    return [n for n in out if n > 0]


@pytest.mark.parametrize('case', ['d6_chi64_real', 'd8_chi128_complex'])
def test_svd_ctmrg_projector_sizes(B, case):
    """One batched call over all projector blocks of a CTMRG workload (13 real blocks up to 440, or 21
    complex blocks up to 1134): many panel pairs per round with several workgroups each, next to the
    LDS-kernel launch on the side stream, the setting in which the workgroups of a pair start at
    different times."""
    if case == 'd6_chi64_real':
        sizes, cplx = _sector_sizes({0: 14, 1: 13, 2: 8, 3: 3, 4: 1}, {0: 12, 1: 8, 2: 4}), False
    else:
        sizes, cplx = _sector_sizes({0: 22, 1: 21, 2: 15, 3: 9, 4: 5, 5: 2, 6: 1}, {0: 14, 1: 12, 2: 8, 3: 4, 4: 1}), True
    rng = np.random.default_rng(12)
    meta, off = [], 0
    for n in sizes:
        meta.append(((off, off + n * n), (n, n), (off, off + n * n), (n, n), (0, 0), (off, off + n * n), (n, n)))
        off += n * n
    so = 0
    for i, n in enumerate(sizes):
        sl, D, slU, DU, _, slV, DV = meta[i]
        meta[i] = (sl, D, slU, DU, (so, so + n), slV, DV)
        so += n
    meta = tuple(meta)
    a = rng.standard_normal(off) + (1j * rng.standard_normal(off) if cplx else 0)
    print('svd projector sizes', case, sizes, flush=True)
    U, S, V = B.svd(_t(a), meta, (off, so, off))
    s_ref = np.concatenate([np.linalg.svd(a[m[0][0]:m[0][1]].reshape(m[1]), compute_uv=False) for m in meta])
    _check_svd(a, meta, _n(U), _n(S), _n(V), s_ref)


def test_svd_graded_spectrum(B):
    """Singular values spanning 12 decades keep their relative accuracy (one-sided Jacobi)."""
    rng = np.random.default_rng(10)
    n = 40
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -12, n)
    a = (q1 * s) @ q2.T
    meta = (((0, n * n), (n, n), (0, n * n), (n, n), (0, n), (0, n * n), (n, n)),)
    U, S, V = B.svd(_t(a.ravel()), meta, (n * n, n, n * n))
    np.testing.assert_allclose(_n(S), s, rtol=1e-8, atol=1e-13)


@pytest.mark.parametrize('scenario', ['mps_u1_fuse_contracted', 'u1xu1_c128', 'truncation_u1'])
def test_svd_adjoint_vs_oracle(B, scenario):
    rng = np.random.default_rng(11)
    for fn, args, kw, outs in G.calls(scenario, 'svd')[:2]:
        data, meta, sizes = args[:3]
        x = _t(data).requires_grad_(True)
        U, S, V = B.svd(x, meta, sizes)
        cplx = U.is_complex()
        mk = (lambda n: rng.standard_normal(n) + 1j * rng.standard_normal(n)) if cplx else rng.standard_normal
        gU, gS, gV = mk(U.numel()), rng.standard_normal(S.numel()), mk(V.numel())
        torch.autograd.backward([U, S, V], [_t(gU.astype(data.dtype)), _t(gS), _t(gV.astype(data.dtype))])
        ref = O.svd_adjoint(gU.astype(data.dtype), gS, gV.astype(data.dtype), _n(U), _n(S), _n(V), meta, data.size)
        assert _rel(_n(x.grad), ref) <= 1e-9, _rel(_n(x.grad), ref)


def test_svd_adjoint_vs_reference_torch_blocks(B):
    """Per-block gradients recorded from the reference torch backend (kernel_svd / SVDGESDD.backward)."""
    js, arrays = G.load('torch_grads')
    for bi, blk in enumerate(js['svd_blocks']):
        a = arrays[f'svdb{bi}_a']
        m, n = a.shape
        k = min(m, n)
        meta = (((0, m * n), (m, n), (0, m * k), (m, k), (0, k), (0, k * n), (k, n)),)
        x = _t(a.ravel()).requires_grad_(True)
        U, S, V = B.svd(x, meta, (m * k, k, k * n))
        if blk['loss'] == 'linear':
            # the reference's gradient uses its own singular-vector gauge: align ours to it first
            ur, vr = arrays[f'svdb{bi}_u'], arrays[f'svdb{bi}_vh']
            sg = np.sign(np.sum(_n(U).reshape(m, k) * ur, axis=0))
            Uf = U.reshape(m, k) * _t(sg)
            Vf = V.reshape(k, n) * _t(sg[:, None])
            loss = (Uf * _t(arrays[f'svdb{bi}_gu'])).sum() + (S * _t(arrays[f'svdb{bi}_gs'])).sum() + \
                (Vf * _t(arrays[f'svdb{bi}_gv'])).sum()
        else:
            w = _t(arrays[f'svdb{bi}_w'])
            uu, vv = U.reshape(m, k), V.reshape(k, n)
            loss = ((uu * S[None, :]) @ vv @ w).trace().real + ((uu @ vv) @ w).trace().imag + (S ** 3).sum()
        loss.backward()
        np.testing.assert_allclose(_n(x.grad).reshape(m, n), arrays[f'svdb{bi}_g'], rtol=1e-8, atol=1e-9)


def test_eigh_adjoint_fd(B):
    rng = np.random.default_rng(12)
    n = 7
    x = rng.standard_normal((n, n))
    a = (x + x.T) / 2
    meta = (((0, n * n), (n, n), (0, n * n), (n, n), (0, n)),)
    cw = rng.standard_normal(n)
    ta = _t(a.ravel()).requires_grad_(True)
    S, U = B.eigh(ta, meta, (n, n * n))
    u = U.reshape(n, n)
    # gauge-invariant loss: eigenvalues and projectors
    loss = (S * _t(cw)).sum() + ((u[:, 0:1] @ u[:, 0:1].T) * _t(x)).sum()
    loss.backward()
    w, v = np.linalg.eigh(a)
    gw = cw
    # projector derivative through the oracle formula: dL/du0 = (X + X^T) u0
    gu = np.zeros((n, n))
    gu[:, 0] = (x + x.T) @ v[:, 0]
    ref = O.eigh_block_adjoint(w, v, gw, gu)
    np.testing.assert_allclose(_n(ta.grad).reshape(n, n), ref, rtol=1e-9, atol=1e-10)


def test_heff2_graph_replay_matches_eager(B):
    """A captured Heff2 step replays to the eager result, with eager work between replays."""
    from yastn_amd import blocksparse as bs
    from yastn_amd import workloads as W
    cfg = bs.make_config(B, 'U1', device='cuda:0')
    AA, FL, FR, Wm = W.dmrg_heff2_setup(cfg, chi=96, seed=5)
    with torch.no_grad():
        ref = W.heff2(AA, FL, FR, Wm).data.clone()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(2):
                W.heff2(AA, FL, FR, Wm)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            out = W.heff2(AA, FL, FR, Wm)
        for k in range(4):
            out.data.zero_()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out.data, ref), k
            W.heff2(AA, FL, FR, Wm)  # eager work between replays
            torch.cuda.synchronize()


def test_heff2_vs_oracle_full_size(B):
    """Heff2 at chi=1024 on the GPU vs the numpy oracle through the same restated planner."""
    from oracle import np_module as NB
    from yastn_amd import blocksparse as bs
    from yastn_amd import workloads as W
    cg = bs.make_config(B, 'U1', device='cuda:0')
    cn = bs.make_config(NB, 'U1')
    g = W.heff2(*W.dmrg_heff2_setup(cg, chi=1024, seed=0))
    n = W.heff2(*W.dmrg_heff2_setup(cn, chi=1024, seed=0))
    assert g.struct == n.struct
    assert _rel(_n(g.data), n.data) <= 1e-12
    U, S, V = W.two_site_truncation(g, 1024)
    Un, Sn, Vn = W.two_site_truncation(n, 1024)
    assert S.struct.t == Sn.struct.t and S.struct.D == Sn.struct.D
    s, sn = _n(S.data), Sn.data
    assert np.max(np.abs(s - sn)) <= 1e-10 * sn.max()


def test_u1xu1_transfer_vs_oracle(B):
    from oracle import np_module as NB
    from yastn_amd import blocksparse as bs
    from yastn_amd import workloads as W
    cg = bs.make_config(B, 'U1xU1', device='cuda:0')
    cn = bs.make_config(NB, 'U1xU1')
    g = W.u1xu1_transfer(W.u1xu1_transfer_setup(cg, seed=0))
    n = W.u1xu1_transfer(W.u1xu1_transfer_setup(cn, seed=0))
    assert g.struct == n.struct
    assert _rel(_n(g.data), n.data) <= 1e-12


def test_svd_panel_path_graded_and_batched(B):
    """Blocks beyond the LDS kernel (panel Jacobi): a 12-decade graded spectrum and several blocks of
    different sizes in one call."""
    rng = np.random.default_rng(13)
    n = 260
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -12, n)
    a = (q1 * s) @ q2.T
    meta = (((0, n * n), (n, n), (0, n * n), (n, n), (0, n), (0, n * n), (n, n)),)
    U, S, V = B.svd(_t(a.ravel()), meta, (n * n, n, n * n))
    np.testing.assert_allclose(_n(S), s, rtol=1e-8, atol=1e-13)
    # batch of mixed blocks (LDS + panel) in one call
    shapes = [(120, 90), (30, 30), (210, 240), (5, 9), (180, 180)]
    data, meta, uo, so, vo, off = [], [], 0, 0, 0, 0
    for m, k2 in shapes:
        k = min(m, k2)
        x = rng.standard_normal(m * k2)
        data.append(x)
        meta.append(((off, off + m * k2), (m, k2), (uo, uo + m * k), (m, k), (so, so + k), (vo, vo + k * k2), (k, k2)))
        off, uo, so, vo = off + m * k2, uo + m * k, so + k, vo + k * k2
    data = np.concatenate(data)
    meta = tuple(meta)
    U, S, V = B.svd(_t(data), meta, (uo, so, vo))
    ref = np.concatenate([np.linalg.svd(data[m[0][0]:m[0][1]].reshape(m[1]), compute_uv=False) for m in meta])
    _check_svd(data, meta, _n(U), _n(S), _n(V), ref)


def test_svd_rank_deficient_and_zero_blocks(B):
    """Blocks of rank r < min(m, n) and all-zero blocks, LDS and panel sizes, both orientations: the
    singular vectors of the (numerically) zero singular values are still orthonormal, as LAPACK's are."""
    rng = np.random.default_rng(15)
    for cplx in (False, True):
        cases = [(7, 15, 3), (15, 7, 2), (40, 40, 10), (12, 12, 0), (9, 4, 0), (200, 230, 60), (230, 200, 1),
                 (180, 180, 0)]
        data, meta, uo, so, vo, off = [], [], 0, 0, 0, 0
        for m, n, r in cases:
            k = min(m, n)
            x = rng.standard_normal((m, r)) + (1j * rng.standard_normal((m, r)) if cplx else 0)
            y = rng.standard_normal((r, n)) + (1j * rng.standard_normal((r, n)) if cplx else 0)
            data.append((x @ y).ravel())
            meta.append(((off, off + m * n), (m, n), (uo, uo + m * k), (m, k), (so, so + k), (vo, vo + k * n), (k, n)))
            off, uo, so, vo = off + m * n, uo + m * k, so + k, vo + k * n
        data = np.concatenate(data)
        meta = tuple(meta)
        print('svd rank deficient', cplx, cases, flush=True)
        U, S, V = B.svd(_t(data), meta, (uo, so, vo))
        ref = np.concatenate([np.linalg.svd(data[m[0][0]:m[0][1]].reshape(m[1]), compute_uv=False) for m in meta])
        _check_svd(data, meta, _n(U), _n(S), _n(V), ref)


def test_eigh_panel_path(B):
    rng = np.random.default_rng(14)
    for n, cplx in ((200, False), (130, True)):
        x = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
        a = (x + x.conj().T) / 2
        meta = (((0, n * n), (n, n), (0, n * n), (n, n), (0, n)),)
        S, U = B.eigh(_t(a.ravel()), meta, (n, n * n))
        w, u = _n(S), _n(U).reshape(n, n)
        np.testing.assert_allclose(w, np.linalg.eigvalsh(a), atol=1e-11 * np.abs(w).max())
        assert np.linalg.norm(u * w @ u.conj().T - a) <= 1e-11 * np.linalg.norm(a)


def test_dot_split_k_long_contractions(B):
    """Split-K GEMMs (small outputs, long K) against numpy, real and complex, and through autograd."""
    rng = np.random.default_rng(16)
    for cplx in (False, True):
        shapes = [(1, 1, 200000), (3, 5, 20000), (40, 40, 300)]
        A, Bm, meta, ao, bo, co = [], [], [], 0, 0, 0
        for M, N, K in shapes:
            A.append(rng.standard_normal(M * K) + (1j * rng.standard_normal(M * K) if cplx else 0))
            Bm.append(rng.standard_normal(K * N) + (1j * rng.standard_normal(K * N) if cplx else 0))
            meta.append(((co, co + M * N), (M, N), (ao, ao + M * K), (M, K), (bo, bo + K * N), (K, N)))
            ao, bo, co = ao + M * K, bo + K * N, co + M * N
        A, Bm, meta = np.concatenate(A), np.concatenate(Bm), tuple(meta)
        print('split-k dot', cplx, shapes, flush=True)
        got = _n(B.dot(_t(A), _t(Bm), meta, co))
        ref = O.dot(A, Bm, meta, co)
        assert _rel(got, ref) <= 1e-12


@pytest.mark.parametrize('n', [40, 300])
def test_svd_nonconvergence_raises(B, n):
    """A block that needs more sweeps than the cap raises (the reference raises LAPACK's error,
    backend_np.py:397-416).  The cap is lowered from the host to one sweep, so every kernel still exits
    normally: the LDS kernel (n = 40) stops at its sweep bound, the panel path (n = 300) after one host
    sweep call; the finalize kernels mark the blocks and the host check raises."""
    from yastn_amd import _functions as F, _lib as L
    rng = np.random.default_rng(13)
    a = rng.standard_normal(n * n)
    meta = (((0, n * n), (n, n), (0, n * n), (n, n), (0, n), (0, n * n), (n, n)),)
    print('svd non-convergence', n, flush=True)
    cap = F.MAX_SWEEPS
    F.MAX_SWEEPS = 1
    try:
        with pytest.raises(L.HipBackendError, match='did not converge'):
            B.svd(_t(a), meta, (n * n, n, n * n))
    finally:
        F.MAX_SWEEPS = cap
    U, S, V = B.svd(_t(a), meta, (n * n, n, n * n))  # and the normal cap converges
    _check_svd(a, meta, _n(U), _n(S), _n(V), np.linalg.svd(a.reshape(n, n), compute_uv=False))


def test_dot_diag_embed_mask_transpose_golden(B):
    """dot_diag, embed_mask (index-list and slice masks) and transpose against the reference numpy
    backend's recorded calls: bit-exact (the diagonal is real in every recorded dot_diag, so each
    element is one rounded product, as in numpy)."""
    n = {'dot_diag': 0, 'embed_mask': 0, 'transpose': 0}
    for scenario in G.SCENARIOS + ('diag_embed_u1',):
        for fn, args, kw, outs in G.calls(scenario):
            if fn == 'dot_diag':
                A, Bd, meta, Dsize, axis, nd = args
                assert not np.iscomplexobj(A)
                got = B.dot_diag(_t(A), _t(Bd), meta, Dsize, axis, nd)
            elif fn == 'embed_mask':
                A, mask, meta, Dsize, axis, nd = args
                m = {k: v if isinstance(v, slice) else torch.as_tensor(v) for k, v in mask.items()}
                got = B.embed_mask(_t(A), m, meta, Dsize, axis, nd)
            elif fn == 'transpose':
                got = B.transpose(_t(args[0]), args[1], args[2])
            else:
                continue
            print(scenario, fn, flush=True)
            torch.cuda.synchronize()
            res = _n(got)
            assert res.dtype == outs[0].dtype
            np.testing.assert_array_equal(res, outs[0])
            n[fn] += 1
    assert n == {'dot_diag': 5, 'embed_mask': 2, 'transpose': 2}, n


def test_dot_diag_adjoint_vs_oracle(B):
    """Gradients of dot_diag with respect to both operands (one launch each) against the oracle's adjoint,
    which is pinned to torch's autograd of the reference torch backend's dot_diag: <= 1e-14 relative."""
    rng = np.random.default_rng(32)
    for fn, args, kw, outs in G.calls('diag_embed_u1', 'dot_diag'):
        A, Bd, meta, Dsize, axis, nd = args
        cplx = np.iscomplexobj(A) or np.iscomplexobj(Bd)
        g = rng.standard_normal(Dsize) + (1j * rng.standard_normal(Dsize) if cplx else 0)
        gA, gB = O.dot_diag_adjoint(g, A, Bd, meta, axis, nd)
        ta, tb = _t(A).requires_grad_(True), _t(Bd).requires_grad_(True)
        out = B.dot_diag(ta, tb, meta, Dsize, axis, nd)
        (out * _t(np.conj(g))).real.sum().backward()
        assert _rel(_n(ta.grad), gA) <= 1e-14 and _rel(_n(tb.grad), gB) <= 1e-14


@pytest.mark.parametrize('n,cplx', [(9, True), (120, True), (150, False)])
def test_eigh_adjoint_vs_oracle_complex_and_panel(B, n, cplx):
    """eigh adjoint (HIP grouped GEMMs + core kernel) against the oracle's formula, which is pinned by central
    finite differences, real and complex: an LDS-resident block (n=9) and panel-path blocks (n=120, 150)."""
    rng = np.random.default_rng(20 + n)
    x = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    a = (x + x.conj().T) / 2
    meta = (((0, n * n), (n, n), (0, n * n), (n, n), (0, n)),)
    cw = rng.standard_normal(n)
    Y = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    ta = _t(a.ravel()).requires_grad_(True)
    print('eigh adjoint', n, cplx, flush=True)
    S, U = B.eigh(ta, meta, (n, n * n))
    u = U.reshape(n, n)
    loss = (S * _t(cw)).sum() + (u[:, 0].conj() @ _t(Y) @ u[:, 0]).real
    loss.backward()
    w, v = np.linalg.eigh(a)
    gu = np.zeros((n, n), dtype=a.dtype)
    gu[:, 0] = (Y + Y.conj().T) @ v[:, 0]
    ref = O.eigh_block_adjoint(w, v, cw, gu)
    got = _n(ta.grad).reshape(n, n)
    assert _rel(got, ref) <= 1e-9, _rel(got, ref)


@pytest.mark.parametrize('case', ['d6_chi64_real'])
def test_svd_adjoint_vs_oracle_panel_sizes(B, case):
    """SVD adjoint against the oracle's per-block formula (the reference's SVDGESDD.backward restated, pinned by
    the reference torch gradients and finite differences) on the CTMRG projector block sizes, where the panel
    path with QR preconditioning runs: <= 1e-9 relative."""
    rng = np.random.default_rng(33)
    shapes = [(440, 440), (300, 260), (180, 210), (64, 64), (12, 7)]
    meta, off, uo, so, vo = [], 0, 0, 0, 0
    for m, n in shapes:
        k = min(m, n)
        meta.append(((off, off + m * n), (m, n), (uo, uo + m * k), (m, k), (so, so + k), (vo, vo + k * n), (k, n)))
        off, uo, so, vo = off + m * n, uo + m * k, so + k, vo + k * n
    meta, sizes = tuple(meta), (uo, so, vo)
    a = rng.standard_normal(off)
    ta = _t(a).requires_grad_(True)
    U, S, V = B.svd(ta, meta, sizes)
    gU, gS, gV = (rng.standard_normal(x) for x in sizes)
    print('svd adjoint panel sizes', shapes, flush=True)
    torch.autograd.backward([U, S, V], [_t(gU), _t(gS), _t(gV)])
    ref = O.svd_adjoint(gU, gS, gV, _n(U), _n(S), _n(V), meta, off)
    assert _rel(_n(ta.grad), ref) <= 1e-9, _rel(_n(ta.grad), ref)


def test_decompositions_reject_nonfinite_input(B):
    """As the reference (scipy's check_finite, backend_np.py:281-286), an SVD or eigh input with a NaN or an
    infinity raises ValueError."""
    n = 5
    meta = (((0, n * n), (n, n), (0, n * n), (n, n), (0, n), (0, n * n), (n, n)),)
    emeta = (((0, n * n), (n, n), (0, n * n), (n, n), (0, n)),)
    for bad in (np.nan, np.inf):
        a = np.random.default_rng(1).standard_normal(n * n)
        a[7] = bad
        with pytest.raises(ValueError, match='infs or NaNs'):
            B.svd(_t(a), meta, (n * n, n, n * n))
        with pytest.raises(ValueError, match='infs or NaNs'):
            B.eigh(_t(a), emeta, (n, n * n))


def test_svd_and_ctm_step_bitwise_deterministic(B):
    """The Jacobi SVD (LDS and panel paths, one-launch sweeps) and a whole CTMRG fwd+bwd step run twice give
    bitwise identical results (checkpoint recomputation relies on it)."""
    from yastn_amd import blocksparse as bs
    from yastn_amd import ctmrg as C
    rng = np.random.default_rng(9)
    shapes = [(300, 280), (40, 30), (150, 150)]
    meta, off, uo, so, vo = [], 0, 0, 0, 0
    for m, n in shapes:
        k = min(m, n)
        meta.append(((off, off + m * n), (m, n), (uo, uo + m * k), (m, k), (so, so + k), (vo, vo + k * n), (k, n)))
        off, uo, so, vo = off + m * n, uo + m * k, so + k, vo + k * n
    a = _t(rng.standard_normal(off))
    r1 = B.svd(a, tuple(meta), (uo, so, vo))
    r2 = B.svd(a, tuple(meta), (uo, so, vo))
    assert all(torch.equal(x, y) for x, y in zip(r1, r2))
    cfg = bs.make_config(B, 'U1', device='cuda:0')
    A, env, sz = C.setup(cfg, D=6, chi=64, seed=0)
    print('ctm step twice', flush=True)
    l1, g1 = C.fwd_bwd(A, env, sz, 64)
    l2, g2 = C.fwd_bwd(A, env, sz, 64)
    assert torch.equal(l1, l2) and torch.equal(g1, g2)


def test_eigh_dense_meta_none(B):
    """eigh(data, meta=None) on a dense Hermitian matrix, the form the reference's Krylov solvers call
    (krylov/_krylov.py:230): ascending eigenvalues vs numpy, U diag(w) U^H = A."""
    rng = np.random.default_rng(11)
    for n, cplx in ((12, False), (30, True)):
        x = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
        a = (x + x.conj().T) / 2
        S, U = B.eigh(_t(a))
        w, u = _n(S), _n(U)
        np.testing.assert_allclose(w, np.linalg.eigvalsh(a), rtol=0, atol=1e-12 * np.abs(w).max())
        np.testing.assert_allclose(u @ np.diag(w) @ u.conj().T, a, atol=1e-12 * np.abs(a).max())
