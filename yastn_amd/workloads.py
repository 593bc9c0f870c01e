"""Synthetic workloads of the block-sparse hot path, built with the restated tensor layer.

Sector dimensions "D(q) proportional to w(q), normalised to chi" use D(q) = max(1, round(chi w / sum w))
with the rounding residual added to the charge-0 sector (keeps q <-> -q symmetry).  Block data are
N(0,1) from ``np.random.default_rng(seed)`` over the flat buffer in block order (complex: independent
real and imaginary parts).
"""
import math

import numpy as np

from . import blocksparse as bs


def split_dims(chi, charges, weight):
    w = np.array([weight(q) for q in charges], dtype=np.float64)
    D = np.maximum(1, np.round(chi * w / w.sum())).astype(np.int64)
    zero = charges.index(0) if 0 in charges else int(np.argmax(w))
    D[zero] += chi - int(D.sum())
    return tuple(int(x) for x in D)


# ------------------------------------------------------------ MPS transfer contraction (U1, numpy-sized)

def mps_transfer(config, seed=0):
    """A(l, p, r), s=(+1,+1,-1); virtual charges -4..4 with D(q)=round(16 exp(-q^2/8)); physical {0,1}."""
    tq = tuple(range(-4, 5))
    Dq = tuple(int(round(16 * math.exp(-q * q / 8))) for q in tq)
    legs = [bs.leg(config, 1, tq, Dq), bs.leg(config, 1, (0, 1), (1, 1)), bs.leg(config, -1, tq, Dq)]
    return bs.rand(config, legs, n=0, dtype='float64', seed=seed)


def mps_transfer_step(A):
    T = bs.tensordot(A, A.conj(), axes=((0, 1), (0, 1)))
    AA = A @ A
    U, S, V = bs.svd_with_truncation(AA, axes=((0, 1), (2, 3)), D_total=64)
    return T, U, S, V


# ------------------------------------------------------------ DMRG two-site Heff (U1 spin-1/2, XXZ MPO)

def xxz_mpo(config, delta=1.0):
    """XXZ MPO tensor W(wl, p_out, wr, p_in), s=(+1,+1,-1,-1), bond 5, charges 2*Sz.

    Lower-triangular form: W = [[I,0,0,0,0],[S+,0,0,0,0],[S-,0,0,0,0],[Sz,0,0,0,0],[0,S-/2,S+/2,delta Sz,I]].
    Virtual index -> charge: 0 (done) 0, 1: -2, 2: +2, 3 (Sz): 0, 4 (start) 0.
    """
    vq = {0: 0, 1: -2, 2: 2, 3: 0, 4: 0}
    # operators in the basis (charge -1, charge +1) = (down, up), as (p_out, p_in) matrices
    I = np.eye(2)
    Sz = np.diag([-0.5, 0.5])
    Sp = np.array([[0, 0], [1, 0]], dtype=float)  # up <- down
    Sm = Sp.T
    W = np.zeros((5, 2, 5, 2))
    for (a, b, op) in ((0, 0, I), (1, 0, Sp), (2, 0, Sm), (3, 0, Sz), (4, 1, Sm / 2), (4, 2, Sp / 2),
                       (4, 3, delta * Sz), (4, 4, I)):
        W[a, :, b, :] = op
    # order the virtual basis by charge: charge -2: [1], 0: [0, 3, 4], +2: [2]
    order = {-2: [1], 0: [0, 3, 4], 2: [2]}
    vleg_t, vleg_D = (-2, 0, 2), (1, 3, 1)
    pq = (-1, 1)
    legs = [bs.leg(config, 1, vleg_t, vleg_D), bs.leg(config, 1, pq, (1, 1)),
            bs.leg(config, -1, vleg_t, vleg_D), bs.leg(config, -1, pq, (1, 1))]
    T = bs.zeros(config, legs, n=0, dtype='float64')
    flat = np.zeros(T.size)
    for (t, D, sl) in zip(T.struct.t, T.struct.D, T.slices):
        ql, po, qr, pi = t
        blk = W[np.ix_(order[ql], [pq.index(po)], order[qr], [pq.index(pi)])]
        flat[sl.slcs[0][0]:sl.slcs[0][1]] = blk.reshape(-1)
    return bs.from_flat(config, legs, 0, flat)


def dmrg_heff2_setup(config, chi=1024, seed=0):
    """AA(l, p1, p2, r), F_left(l, w, l'), F_right(r, w, r') and the XXZ MPO at bond dimension chi."""
    tq = tuple(range(-12, 13))
    Dq = split_dims(chi, list(tq), lambda q: math.exp(-q * q / 32))
    pq = (-1, 1)
    wt, wD = (-2, 0, 2), (1, 3, 1)
    AA = bs.rand(config, [bs.leg(config, 1, tq, Dq), bs.leg(config, 1, pq, (1, 1)), bs.leg(config, 1, pq, (1, 1)),
                          bs.leg(config, -1, tq, Dq)], n=0, seed=seed)
    FR = bs.rand(config, [bs.leg(config, 1, tq, Dq), bs.leg(config, 1, wt, wD), bs.leg(config, -1, tq, Dq)], n=0,
                 seed=seed + 1)
    FL = bs.rand(config, [bs.leg(config, -1, tq, Dq), bs.leg(config, -1, wt, wD), bs.leg(config, 1, tq, Dq)], n=0,
                 seed=seed + 2)
    W = xxz_mpo(config)
    return AA, FL, FR, W


def heff2(AA, FL, FR, W):
    """Env3.Heff2 contraction order (reference tn/mps/_env.py:512-518): 4 tensordots."""
    tmp = AA @ FR
    tmp = bs.tensordot(W, tmp, axes=((2, 3), (3, 2)))
    tmp = bs.tensordot(W, tmp, axes=((2, 3), (0, 3)))
    return bs.tensordot(FL, tmp, axes=((0, 1), (3, 0)))


def two_site_truncation(AA, chi=1024):
    """post_2site truncation (reference tn/mps/_mps_obc.py:521-557): SVD of the fused two-site tensor."""
    return bs.svd_with_truncation(AA, axes=((0, 1), (2, 3)), D_total=chi)


# ------------------------------------------------------------ U1 x U1 uneven sectors (complex)

def u1xu1_sectors(chi=2048, nsect=150, sigma=3.0):
    cands = [(a, b) for a in range(-12, 13) for b in range(-12, 13)]
    w = {c: math.exp(-(c[0] ** 2 + c[1] ** 2) / (2 * sigma ** 2)) for c in cands}
    keep = sorted(sorted(cands, key=lambda c: (-w[c], c))[:nsect])
    D = split_dims(chi, keep, lambda c: w[c])
    return tuple(keep), D


def u1xu1_transfer_setup(config, chi=2048, nsect=150, sigma=3.0, seed=0):
    tv, Dv = u1xu1_sectors(chi, nsect, sigma)
    tp = ((0, 0), (1, 0), (0, 1), (1, 1))
    legs = [bs.leg(config, 1, tv, Dv), bs.leg(config, 1, tp, (1, 1, 1, 1)), bs.leg(config, -1, tv, Dv)]
    struct, _ = bs._fill_struct(config, legs, (0, 0))
    return bs.from_flat(config, legs, (0, 0), bs.seeded_normal(struct.size, 'complex128', seed))


def u1xu1_transfer(A):
    return bs.tensordot(A, A.conj(), axes=((0, 1), (0, 1)))
