"""Workload construction and the restated tensor layer on the numpy oracle backend (CPU)."""
import numpy as np

from oracle import np_module as NB
from yastn_amd import blocksparse as bs
from yastn_amd import workloads as W


def test_split_dims_matches_stated_sectors():
    assert W.split_dims(1024, list(range(-12, 13)), lambda q: np.exp(-q * q / 32)) == \
        (1, 2, 4, 8, 14, 22, 33, 47, 62, 77, 90, 99, 106, 99, 90, 77, 62, 47, 33, 22, 14, 8, 4, 2, 1)
    assert W.split_dims(64, list(range(-4, 5)), lambda q: np.exp(-q * q / 6)) == (1, 3, 8, 13, 14, 13, 8, 3, 1)
    tv, Dv = W.u1xu1_sectors()
    assert len(tv) == 150 and sum(Dv) == 2048


def test_mps_transfer_is_the_golden_tensor():
    import golden_io as G
    cfg = bs.make_config(NB, 'U1')
    A = W.mps_transfer(cfg, seed=0)
    t = G.tensors('mps_u1_fuse_contracted')['A']
    assert A.struct.t == t['t'] and A.struct.D == t['D']
    np.testing.assert_array_equal(A.data, t['data'])


def test_heff2_dense_closed_form():
    cfg = bs.make_config(NB, 'U1')
    AA, FL, FR, Wm = W.dmrg_heff2_setup(cfg, chi=48, seed=3)
    out = W.heff2(AA, FL, FR, Wm)
    a, fl, fr, w = (x.to_dense() for x in (AA, FL, FR, Wm))
    t = np.tensordot(a, fr, axes=(3, 0))                      # l p1 p2 w r'
    t = np.tensordot(w, t, axes=((2, 3), (3, 2)))             # wl po l p1 r'
    t = np.tensordot(w, t, axes=((2, 3), (0, 3)))             # wl1 po1 po2 l r'
    t = np.tensordot(fl, t, axes=((0, 1), (3, 0)))            # l' po1 po2 r'
    got = out.to_dense()
    assert got.shape == t.shape
    assert np.linalg.norm(got - t) <= 1e-12 * np.linalg.norm(t)


def test_xxz_mpo_is_hermitian_heisenberg_bond():
    """Two MPO sites contracted with open boundary vectors give the two-site XXZ Hamiltonian."""
    cfg = bs.make_config(NB, 'U1')
    w = W.xxz_mpo(cfg, delta=1.0).to_dense()          # (wl, po, wr, pi), virtual basis ordered (-2 | 0,0,0 | +2)
    vbasis = [1, 0, 3, 4, 2]                           # dense index -> original MPO index
    start = np.zeros(5); start[vbasis.index(4)] = 1
    end = np.zeros(5); end[vbasis.index(0)] = 1
    h = np.einsum('a,aibj,bkcl,c->ikjl', start, w, w, end).reshape(4, 4)
    sz = np.diag([-0.5, 0.5]); sp = np.array([[0, 0], [1, 0.]]); sm = sp.T
    ref = np.kron(sz, sz) + 0.5 * (np.kron(sp, sm) + np.kron(sm, sp))
    np.testing.assert_allclose(h, ref, atol=1e-14)


def test_u1xu1_transfer_small_dense():
    cfg = bs.make_config(NB, 'U1xU1')
    A = W.u1xu1_transfer_setup(cfg, chi=60, nsect=12, seed=1)
    T = W.u1xu1_transfer(A)
    a = A.to_dense()
    ref = np.tensordot(a, a.conj(), axes=((0, 1), (0, 1)))
    assert np.linalg.norm(T.to_dense() - ref) <= 1e-12 * np.linalg.norm(ref)
