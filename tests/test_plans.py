"""Work lists (yastn_amd/plans.py) checked on the CPU by interpreting them with numpy.

Every plan is also validated by the library's own host-side checker while it is built.
"""
import numpy as np
import pytest
import torch

import golden_io as G
import plan_emulator as E
from oracle import np_backend as O
from yastn_amd import plans as P
from yastn_amd import _functions as F

CPU = torch.device('cpu')


@pytest.mark.parametrize('scenario', G.SCENARIOS)
def test_gather_plans_reproduce_golden(scenario):
    for fn, args, kw, outs in G.calls(scenario):
        if fn == 'transpose_and_merge':
            data, order, meta_new, meta_mrg, Dsize = args
            plan = P.merge_plan(order, meta_new, meta_mrg, Dsize, data.size, CPU)
            dst = np.full(Dsize, np.nan, dtype=data.dtype)
            np.testing.assert_array_equal(E.gather(plan.h_jobs, data, dst), outs[0])
            # adjoint plan = oracle adjoint
            g = np.random.default_rng(0).standard_normal(Dsize).astype(data.dtype)
            pa = P.merge_adjoint_plan(order, meta_new, meta_mrg, Dsize, data.size, CPU)
            got = E.gather(pa.h_jobs, g, np.full(data.size, np.nan, dtype=data.dtype))
            np.testing.assert_array_equal(got, O.transpose_and_merge_adjoint(g, order, meta_new, meta_mrg, data.size))
        elif fn == 'unmerge':
            data, meta = args
            plan = P.unmerge_plan(meta, data.size, CPU)
            np.testing.assert_array_equal(E.gather(plan.h_jobs, data, np.full(data.size, np.nan, data.dtype)), outs[0])
            pa = P.unmerge_adjoint_plan(meta, data.size, CPU)
            back = E.gather(pa.h_jobs, outs[0], np.full(data.size, np.nan, data.dtype))
            np.testing.assert_array_equal(back, data)
        elif fn == 'apply_mask':
            A, mask, meta, Dsize, axis, ndim = args
            mh = F._mask_host(mask)
            plan = P.mask_plan(mh, meta, Dsize, A.size, axis, CPU, embed=False, mask_key=F._mask_key(mh))
            got = E.gather(plan.h_jobs, A, np.full(Dsize, np.nan, A.dtype))
            np.testing.assert_array_equal(got, outs[0])
            sw = F._swapped_mask_meta(meta)
            pe = P.mask_plan(mh, sw, A.size, Dsize, axis, CPU, embed=True, mask_key=F._mask_key(mh))
            emb = E.gather(pe.h_jobs, outs[0], np.full(A.size, np.nan, A.dtype))
            np.testing.assert_array_equal(emb, O.apply_mask_adjoint(outs[0], mask, meta, A.size, axis, ndim))


@pytest.mark.parametrize('scenario', G.SCENARIOS)
def test_dot_plans_reproduce_golden(scenario):
    for fn, args, kw, outs in G.calls(scenario, 'dot'):
        A, B, meta, Dsize = args
        dt = np.promote_types(A.dtype, B.dtype)
        A, B = A.astype(dt), B.astype(dt)
        plan = P.dot_plan(meta, A.size, B.size, Dsize, CPU)
        C = np.full(Dsize, np.nan, dtype=dt)
        E.gemm(plan.h_gemms, plan.h_terms, [C], [A, B])
        assert np.linalg.norm(C - outs[0]) <= 1e-13 * np.linalg.norm(outs[0])
        for cfg in P.L.TILE_DIMS:
            if plan.applicable(cfg):
                plan.tiles_for(cfg)  # host-checked for every tile shape
            else:  # the checker refuses the small-M kernel for a plan with M > 16
                with pytest.raises(RuntimeError):
                    plan.tiles_for(cfg)
        # adjoint
        rng = np.random.default_rng(2)
        gC = rng.standard_normal(Dsize).astype(dt)
        pa = P.dot_adjoint_plan(meta, A.size, B.size, Dsize, CPU)
        gA = np.zeros(A.size, dt) if pa.zero_a else np.full(A.size, np.nan, dt)
        gB = np.zeros(B.size, dt) if pa.zero_b else np.full(B.size, np.nan, dt)
        E.gemm(pa.h_gemms, pa.h_terms, [gA, gB], [A, B, gC])
        rA, rB = O.dot_adjoint(gC, A, B, meta)
        assert np.linalg.norm(gA - rA) <= 1e-13 * np.linalg.norm(rA)
        assert np.linalg.norm(gB - rB) <= 1e-13 * np.linalg.norm(rB)


@pytest.mark.parametrize('scenario', ['mps_u1_fuse_contracted', 'u1xu1_c128', 'truncation_u1'])
def test_svd_adjoint_plan_algebra(scenario):
    """The composed adjoint (GEMM plans + core formula) equals the oracle's per-block adjoint."""
    rng = np.random.default_rng(5)
    for fn, args, kw, outs in G.calls(scenario, 'svd')[:2]:
        data, meta, sizes = args[:3]
        U, S, V = O.svd(data, meta, sizes)
        dt = U.dtype
        cplx = np.iscomplexobj(U)
        mk = (lambda n: rng.standard_normal(n) + 1j * rng.standard_normal(n)) if cplx else rng.standard_normal
        gU, gS, gV = mk(U.size).astype(dt), rng.standard_normal(S.size), mk(V.size).astype(dt)
        ref = O.svd_adjoint(gU, gS, gV, U, S, V, meta, data.size)
        ap = F._SvdAdjointPlans(meta, (U.size, S.size, V.size, data.size), CPU)
        X = np.zeros(max(ap.kk, 1), dt); Y = np.zeros(max(ap.kk, 1), dt)
        J = np.zeros(max(ap.kk, 1), dt); T = np.zeros(max(ap.tt, 1), dt)
        E.gemm(ap.p1.h_gemms, ap.p1.h_terms, [X, Y], [U, gU, V, gV])
        E.svd_adjoint_core(ap.h_descs, J, T, X, Y, S, gS, gU, gV, 1e-12, 1, 1, 1)
        E.gemm(ap.p2.h_gemms, ap.p2.h_terms, [T], [U, J, V])
        gA = np.zeros(data.size, dt)
        E.gemm(ap.p3.h_gemms, ap.p3.h_terms, [gA], [T, V, U])
        assert np.linalg.norm(gA - ref) <= 1e-10 * np.linalg.norm(ref)


def test_svd_plan_split_and_bounds():
    for scenario in ('mps_u1_fuse_contracted', 'u1xu1_c128'):
        for fn, args, kw, outs in G.calls(scenario, 'svd'):
            data, meta, sizes = args[:3]
            plan = P.svd_plan(meta, data.itemsize, data.size, sizes, CPU)
            assert plan.nblocks == len(meta)
            assert sorted(plan.h_order.tolist()) == list(range(len(meta)))


def test_merge_groups_and_coalesce():
    assert P._merge_groups((6, 5), (2, 3, 5)) == [[0, 1], [2]]
    assert P._merge_groups((2, 5), (2, 1, 5)) == [[0], [1, 2]]
    assert P._merge_groups((1, 4), (4,)) == [[], [0]]
    with pytest.raises(ValueError):
        P._merge_groups((4, 5), (2, 3, 5))
    # contiguous box collapses to one dim
    assert P._coalesce((3, 4, 5), (20, 5, 1), (20, 5, 1)) == ((60,), (1,), (1,))
    # transposed inner dims stay separate
    sh, ss, ds = P._coalesce((4, 5), (1, 4), (5, 1))
    assert sh == (4, 5) and ss == (1, 4) and ds == (5, 1)


def test_uncovered_boxes():
    boxes = [((0, 2), (0, 3)), ((2, 5), (3, 7))]
    un = P._uncovered_boxes((5, 7), boxes)
    cover = np.zeros((5, 7), int)
    for b in boxes + un:
        cover[b[0][0]:b[0][1], b[1][0]:b[1][1]] += 1
    assert np.all(cover == 1)


def test_tile_choice_prefers_small_tiles_for_small_sectors():
    tiny = P.dot_plan(tuple(((0, 64), (8, 8), (0, 64), (8, 8), (0, 64), (8, 8)) for _ in range(1)), 64, 64, 64, CPU)
    assert tiny.cfg == P.L.YH_TILE_SMALLM  # a few rows: the streaming kernel
    small = P.dot_plan(tuple(((0, 576), (24, 24), (0, 576), (24, 24), (0, 576), (24, 24)) for _ in range(1)),
                       576, 576, 576, CPU)
    assert small.cfg == P.L.YH_TILE_32x32


def test_autotune_near_ties_go_to_less_operand_traffic():
    """Operand traffic of a tile shape (A's row panel per column tile, B's column panel per row tile):
    for six 99 x 473 x 99 sector GEMMs larger tiles re-read less, and a 32 x 32 launch timed within 5% of a
    16 x 64 one wins over it (74.2 against 46.6 MB measured by the HBM counters of those two launches),
    while a clearly faster shape still wins."""
    L = P.L
    meta = tuple(((i * 9801, (i + 1) * 9801), (99, 99), (i * 46827, (i + 1) * 46827), (99, 473),
                  (i * 46827, (i + 1) * 46827), (473, 99)) for i in range(6))
    plan = P.dot_plan(meta, 6 * 46827, 6 * 46827, 6 * 9801, CPU)
    t = {c: plan.operand_traffic(c) for c in (L.YH_TILE_16x64, L.YH_TILE_32x32, L.YH_TILE_64x64, L.YH_TILE_128x64)}
    assert t[L.YH_TILE_128x64] < t[L.YH_TILE_64x64] < t[L.YH_TILE_32x32] < t[L.YH_TILE_16x64]
    assert t[L.YH_TILE_32x32] == 6 * (99 * 473 * 4 + 473 * 99 * 4 + 99 * 99)
    assert plan.pick_tile({L.YH_TILE_16x64: 21.0, L.YH_TILE_32x32: 21.9, L.YH_TILE_64x64: 36.0}) == L.YH_TILE_32x32
    assert plan.pick_tile({L.YH_TILE_16x64: 20.0, L.YH_TILE_32x32: 21.9}) == L.YH_TILE_16x64


@pytest.mark.parametrize('cplx', [False, True])
def test_qr_preconditioned_svd_plan_algebra(cplx):
    """The QR-preconditioned panel path, executed with numpy interpreters of its kernels and of its
    grouped-GEMM plans (the Jacobi stage replaced by numpy's SVD of X = R^H), reconstructs every
    block with orthonormal factors, for m >= n and m < n blocks of several panels."""
    rng = np.random.default_rng(5)
    shapes = [(100, 80), (70, 120), (130, 130), (90, 37)]
    dt = np.complex128 if cplx else np.float64
    data, meta, uo, so, vo, off = [], [], 0, 0, 0, 0
    for m, n in shapes:
        k = min(m, n)
        x = rng.standard_normal(m * n) + (1j * rng.standard_normal(m * n) if cplx else 0)
        data.append(x.astype(dt))
        meta.append(((off, off + m * n), (m, n), (uo, uo + m * k), (m, k), (so, so + k), (vo, vo + k * n), (k, n)))
        off, uo, so, vo = off + m * n, uo + m * k, so + k, vo + k * n
    data = np.concatenate(data)
    plan = P.SvdPlan([(sl[0], D[0], D[1], slU[0], slS[0], slV[0]) for sl, D, slU, DU, slS, slV, DV in meta],
                     data.itemsize, 0, data.size, uo, so, vo, CPU)
    pre = plan.d_pre.numpy().tolist()
    assert plan.n_pre >= 3
    work = np.zeros(plan.work_numel, dtype=dt)
    U = np.zeros(uo, dtype=dt)
    S = np.zeros(so)
    Vh = np.zeros(vo, dtype=dt)
    for b in pre:  # panel_init: W0 = A or A^H, column-major
        a_off, m, n = (int(x) for x in plan.h_blocks[b, :3])
        A = data[a_off:a_off + m * n].reshape(m, n)
        W0 = A if m >= n else A.conj().T
        p, q = W0.shape
        E._colmajor(work, int(plan.h_qdesc[b, 2]), p, q)[:] = W0
    for step, (d_ids, n_act, trail) in enumerate(plan.qr_steps):
        for b in d_ids.numpy().tolist():
            E.qr_panel(work, plan.h_qdesc[b], step)
        if trail is not None:
            for gp in trail:
                E.gemm(gp.h_gemms, gp.h_terms, [work], [work])
    for b in pre:
        p, q, w_off, vall_off, tm_off, x_off, y_off, _ = (int(x) for x in plan.h_qdesc[b])
        R = np.triu(E._colmajor(work, w_off, p, q)[:q, :])
        a_off, m, n, u_off, s_off, v_off = (int(x) for x in plan.h_blocks[b, :6])
        A = data[a_off:a_off + m * n].reshape(m, n)
        W0 = A if m >= n else A.conj().T
        # R is the R factor of W0 (up to the Householder signs): R^H R = W0^H W0
        np.testing.assert_allclose(R.conj().T @ R, W0.conj().T @ W0, atol=1e-10 * np.linalg.norm(W0) ** 2)
        X = R.conj().T
        Uj, sj, Vjh = np.linalg.svd(X)
        Vj = Vjh.conj().T
        S[s_off:s_off + q] = sj
        if m >= n:  # Vh (q x q) rows u_j^H, U (p x q) columns [v_j; 0]
            Vh[v_off:v_off + q * q] = Uj.conj().T.ravel()
            Ub = np.zeros((p, q), dtype=dt)
            Ub[:q] = Vj
            U[u_off:u_off + p * q] = Ub.ravel()
        else:  # U (q x q) columns u_j, Vh (q x p) rows conj([v_j; 0])
            U[u_off:u_off + q * q] = Uj.ravel()
            Vb = np.zeros((q, p), dtype=dt)
            Vb[:, :q] = Vj.conj().T
            Vh[v_off:v_off + q * p] = Vb.ravel()
    for h1, h2, h3 in plan.q_apply:
        E.gemm(h1.h_gemms, h1.h_terms, [work], [work, U, Vh])
        E.gemm(h2.h_gemms, h2.h_terms, [work], [work])
        E.gemm(h3.h_gemms, h3.h_terms, [work, U, Vh], [work])
    for b in pre:  # the blocks of the LDS kernel are not part of this path
        sl, D, slU, DU, slS, slV, DV = meta[b]
        a = data[sl[0]:sl[1]].reshape(D)
        u = U[slU[0]:slU[1]].reshape(DU)
        v = Vh[slV[0]:slV[1]].reshape(DV)
        s = S[slS[0]:slS[1]]
        k = len(s)
        assert np.linalg.norm(u * s @ v - a) <= 1e-12 * np.linalg.norm(a)
        assert np.linalg.norm(u.conj().T @ u - np.eye(k)) <= 1e-12 * k
        assert np.linalg.norm(v @ v.conj().T - np.eye(k)) <= 1e-12 * k


def test_split_k_plan_reproduces_long_contractions():
    """GEMMs with small outputs and long K (a full contraction to a scalar) are cut into K chunks whose
    partial products are summed in a fixed order; the interpreted plan equals the direct product."""
    rng = np.random.default_rng(6)
    shapes = [(1, 1, 50000), (3, 5, 20000), (40, 40, 300)]
    A, B, meta, ao, bo, co = [], [], [], 0, 0, 0
    for M, N, K in shapes:
        A.append(rng.standard_normal(M * K))
        B.append(rng.standard_normal(K * N))
        meta.append(((co, co + M * N), (M, N), (ao, ao + M * K), (M, K), (bo, bo + K * N), (K, N)))
        ao, bo, co = ao + M * K, bo + K * N, co + M * N
    A, B = np.concatenate(A), np.concatenate(B)
    plan = P.dot_plan(tuple(meta), A.size, B.size, co, CPU)
    assert len(plan.h_red) == 2 and plan.part_numel > 0
    out = np.zeros(co)
    part = np.zeros(plan.part_numel)
    outs = [out, None, None, None]
    outs[plan.part_buf] = part
    E.gemm(plan.h_gemms, plan.h_terms, outs, [A, B])
    E.splitk_reduce(plan.h_red, outs, part)
    for (slc, Dc, sla, Da, slb, Db) in meta:
        ref = A[sla[0]:sla[1]].reshape(Da) @ B[slb[0]:slb[1]].reshape(Db)
        np.testing.assert_allclose(out[slc[0]:slc[1]].reshape(Dc), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_plan_cache_bounded_for_fresh_equal_metas():
    """Metas rebuilt on every call (equal content, new objects) hit the content map and keep both maps
    bounded; the LRU evicts one entry at a time instead of clearing the cache."""
    cache = P._Cache(maxsize=8)
    builds = []

    def meta():
        return tuple((i, (i, i + 1)) for i in range(5))

    first = None
    for _ in range(100):
        m = meta()
        v = cache.get(('k', id(m), 3), (m,), lambda: builds.append(1) or object())
        first = first or v
        assert v is first
    assert len(builds) == 1
    assert len(cache.d) <= 8 and len(cache.by_content) <= 8
    # distinct structures beyond the bound: the oldest go, the recent ones stay
    keep = []
    for i in range(20):
        m = (i,)
        keep.append(m)
        cache.get(('k', id(m)), (m,), lambda i=i: ('plan', i))
    assert len(cache.d) == 8 and len(cache.by_content) <= 8
    assert cache.get(('k', id(keep[-1])), (keep[-1],), lambda: 'rebuilt') == ('plan', 19)


def test_partial_svd_meta_cache_is_content_keyed():
    from yastn_amd import backend_hip as B
    m1 = (((0, 12), (3, 4), (0, 6), (3, 2), (0, 2), (0, 8), (2, 4)),)
    m2 = tuple(tuple(x) for x in m1)  # equal content, new object
    a = B._full_meta_of_partial(m1, CPU)
    n0 = len(B._PARTIAL.by_content)
    b = B._full_meta_of_partial(m2, CPU)
    assert a is b and len(B._PARTIAL.by_content) == n0
    full, sizes, (iu, is_, iv) = a
    assert sizes == (9, 3, 12) and iu.device == CPU
    np.testing.assert_array_equal(is_.numpy(), [0, 1])
    np.testing.assert_array_equal(iu.numpy(), [0, 1, 3, 4, 6, 7])


class _Twice(F._Fn):
    @staticmethod
    def forward(x, scale=2.0):
        return x * scale

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.scale = inputs[1] if len(inputs) > 1 else 2.0

    @staticmethod
    def backward(ctx, g):
        return g * ctx.scale, None


@pytest.mark.parametrize('fast', [True, False])
def test_fast_apply_matches_public_apply(fast, monkeypatch):
    """_Fn.apply's fast path (torch internals) and torch's public Function.apply give the same outputs
    and gradients, with and without defaulted arguments."""
    monkeypatch.setattr(F, 'FAST_APPLY', fast and F._UNWRAP_DEAD is not None)
    x = torch.arange(5.0, dtype=torch.float64, requires_grad=True)
    y = _Twice.apply(x, 3.0)
    y.sum().backward()
    np.testing.assert_array_equal(y.detach().numpy(), 3.0 * np.arange(5.0))
    np.testing.assert_array_equal(x.grad.numpy(), np.full(5, 3.0))
    x.grad = None
    z = _Twice.apply(x)
    (z * z).sum().backward()
    np.testing.assert_array_equal(x.grad.numpy(), 8.0 * np.arange(5.0))
    gx, = torch.func.vjp(lambda t: _Twice.apply(t, 3.0), x.detach())[1](torch.ones(5, dtype=torch.float64))
    np.testing.assert_array_equal(gx.numpy(), np.full(5, 3.0))


def test_panel_round_schedule_emulated():
    """The schedule of the panel rounds (yastn_hip_svd_panel_sweep's inner kernel), run in numpy on a
    real SvdPlan's job and round lists: on the QR-preconditioned blocks (X = R^H) the items of every
    sub-round of every pair cover every entry of its Gram matrix, each column pair of a block is rotated
    once per sweep, the iteration converges, and the column norms are the singular values."""
    import plan_emulator as E
    rng = np.random.default_rng(5)
    shapes = [(200, 150), (130, 130), (300, 100), (90, 140), (17, 12)]
    blocks, a_off, u_off, s_off, v_off, mats = [], 0, 0, 0, 0, []
    for m, n in shapes:
        k = min(m, n)
        blocks.append((a_off, m, n, u_off, s_off, v_off))
        mats.append(rng.standard_normal((m, n)) * np.logspace(0, -8, n)[None, :])
        a_off, u_off, s_off, v_off = a_off + m * n, u_off + m * k, s_off + k, v_off + k * n
    plan = P.SvdPlan(blocks, 8, 0, a_off, u_off, s_off, v_off, torch.device('cpu'))
    assert plan.n_big >= 3 and plan.n_pre == plan.n_big and plan.panel_width <= 16
    work = np.zeros(plan.work_numel)
    floors = np.zeros(len(blocks))
    big = set(int(x) for x in plan.h_big)
    for bi, ((m, n), A) in enumerate(zip(shapes, mats)):
        if bi not in big:
            continue
        W0 = A.conj().T if m < n else A
        R = np.linalg.qr(W0, mode='r')
        X = R.conj().T
        q = X.shape[1]
        x_off = int(plan.h_jblocks[bi][6])
        work[x_off:x_off + q * q] = X.T.ravel()          # column-major X
        work[x_off + q * q:x_off + 2 * q * q] = np.eye(q).ravel()
        floors[bi] = np.finfo(float).eps ** 2 * np.sum(X ** 2)
    sweeps = E.panel_sweeps(plan.h_jobs, plan.h_rounds, plan.h_jblocks, work, floors, plan.panel_width, 60, None)
    assert 0 < sweeps < 30, sweeps
    for bi, ((m, n), A) in enumerate(zip(shapes, mats)):
        if bi not in big:
            continue
        q = min(m, n)
        x_off = int(plan.h_jblocks[bi][6])
        W = work[x_off:x_off + q * q].reshape(q, q).T
        s = np.sort(np.linalg.norm(W, axis=0))[::-1]
        s_ref = np.linalg.svd(A, compute_uv=False)
        assert np.max(np.abs(s - s_ref)) <= 1e-12 * s_ref[0]
