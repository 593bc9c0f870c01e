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
            plan.tiles_for(cfg)  # host-checked for every tile shape
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
    small = P.dot_plan(tuple(((0, 64), (8, 8), (0, 64), (8, 8), (0, 64), (8, 8)) for _ in range(1)), 64, 64, 64, CPU)
    assert small.cfg == P.L.YH_TILE_32x32
