"""The Kagome RVB D=3 U(1) iPEPS of the reference's tests/peps/test_ad.py:37-180 and its CTMRG settings
(test infrastructure).  The state and gradients are golden vectors (tests/golden/rvb_kagome.*): the state
and grad_expected as the reference test holds them, and loss + gradient of the reference torch backend
(CPU) after 1 sweep from the 'eye' and 'dl' environments, 2 sweeps and 20 sweeps from 'dl'."""
import numpy as np

import golden_io as G
from yastn_amd import blocksparse as bs

CHI = 64
BLOCK = ((-1,), (-1,))  # the loss is Re Tr rho[(-1, -1)] (test_ad.py:165)


def state(config):
    """A(t, l, b, r, p): the stored blocks (physical leg first) with the lazy transpose (1, 2, 3, 4, 0)."""
    t = G.tensors('rvb_kagome')['A']
    A = bs.from_blocks(config, t['s'], t['n'], t['t'], t['D'], np.array(t['data']))
    return A._replace(trans=tuple(t['trans']))


def svd_opts():
    """opts_svd {'D_total': 64, 'fix_signs': False} with truncation_mask_multiplets(keep_multiplets=True,
    D_total=64, tol=1e-8, tol_block=0, eps_multiplet=1e-8) (test_ad.py:153-163)."""
    mask_f = lambda S: bs.truncation_mask_multiplets(S, keep_multiplets=True, D_total=CHI, tol=1.0e-8,
                                                     tol_block=0.0, eps_multiplet=1.0e-8)
    return {'fix_signs': False, 'mask_f': mask_f}


def runs():
    js, ar = G.load('rvb_kagome')
    return [(r['sweeps'], r['init'], r['loss'], ar['grad_' + r['tag']]) for r in js['runs']]


def grad_expected():
    return G.load('rvb_kagome')[1]['grad_expected']


def runs_qr():
    """The reference CTMRG with use_qr=True (its default) on the same state: (sweeps, init, loss, whether the
    reference torch gradient is finite).  The half-systems of this state are rank-deficient, so the QR
    adjoint divides by zero there and the reference's gradient is not finite."""
    js, _ = G.load('rvb_kagome_qr')
    return [(r['sweeps'], r['init'], r['loss'], r['finite']) for r in js['runs']]
