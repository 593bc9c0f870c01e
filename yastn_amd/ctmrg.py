"""Differentiable CTMRG step for a U(1) iPEPS on a 1x1 unit cell (the harness of the CTMRG workloads).

Structure restated from the reference (yastn/tn/fpeps/envs/_env_ctm.py, _env_contractions.py):
  * environment tensors per direction with the reference's leg conventions: corners tl (b, r),
    tr (l, b), br (t, l), bl (r, t); edges t (l, D^2, r), l (b, D^2, t), b (r, D^2, l), r (t, D^2, b);
  * 2x2 enlarged corners (corner2x2_*, _env_contractions.py:429-475) on the double-layer site;
  * projectors from proj_corners (_env_ctm.py:1230-1258): rr = r0 r1^T, svd_with_truncation with
    fix_signs=True, p0 = r1 (S^-1/2 V)^*, p1 = r0 (U S^-1/2)^*; r0, r1 the half-systems (use_qr=False,
    r_t = cor_tt, ...) or their R factors (use_qr=True, the reference's default), as in
    update_extended_2x2_projectors_ (_env_ctm.py:1050-1183);
  * environment renormalisation of _update_env_ (_env_ctm.py:692-775) with norm(p='inf') rescaling;
  * an 'hv' step = horizontal move (projectors for the left and right bonds, then l, tl, bl, r, tr, br)
    followed by the vertical move (t, tl, tr, b, bl, br).
The double-layer site is formed explicitly (ket x bra contracted over the physical leg, legs fused
pairwise); the reference contracts ket and bra lazily.  The observable is <O> on the 2x2 cluster:
Z(O)/Z(1) with Z(X) = Tr(cor_tl(X) cor_tr cor_br cor_bl).
"""
import math
import os
import time

import numpy as np

from . import blocksparse as bs
from .workloads import split_dims


class Env:
    """The eight environment tensors of a 1x1 unit cell."""
    NAMES = ('tl', 't', 'tr', 'r', 'br', 'b', 'bl', 'l')

    def __init__(self, **kw):
        for k in self.NAMES:
            setattr(self, k, kw[k])

    def tensors(self):
        return [getattr(self, k) for k in self.NAMES]

    def to_dict(self, level=2):
        """{'type', 'env': {name: Tensor.to_dict}} (the environment part of the reference's EnvCTM.to_dict)."""
        return {'type': 'EnvCTM1x1', 'dict_ver': 1, 'env': {k: getattr(self, k).to_dict(level) for k in self.NAMES}}

    @classmethod
    def from_dict(cls, d, config=None):
        return cls(**{k: bs.Tensor.from_dict(v, config) for k, v in d['env'].items()})


def double_layer(A, op=None):
    """a(t, l, b, r) = sum_p A(t, l, b, r, p) [op] A*(t, l, b, r, p), legs fused (ket, bra) pairwise."""
    Ao = A if op is None else bs.tensordot(A, op, axes=(4, 1))
    x = bs.tensordot(Ao, A.conj(), axes=(4, 4))  # (t, l, b, r, t*, l*, b*, r*)
    return x.fuse_legs(axes=((0, 4), (1, 5), (2, 6), (3, 7)))


def _extra(a):
    """Legs of the double-layer site beyond (t, l, b, r): the open physical legs of an energy contraction."""
    return tuple(range(4, a.ndim))


def corner_tl(env, a):
    c = env.l @ env.tl @ env.t
    c = bs.tensordot(c, a, axes=((2, 1), (0, 1)))
    return c.fuse_legs(axes=((0, 2), (1, 3)) + _extra(a))


def corner_tr(env, a):
    c = env.t @ env.tr @ env.r
    c = bs.tensordot(c, a, axes=((1, 2), (0, 3)))
    return c.fuse_legs(axes=((0, 2), (1, 3)) + _extra(a))


def corner_br(env, a):
    c = env.r @ env.br @ env.b
    c = bs.tensordot(c, a, axes=((2, 1), (2, 3)))
    return c.fuse_legs(axes=((0, 2), (1, 3)) + _extra(a))


def corner_bl(env, a):
    c = env.b @ env.bl @ env.l
    c = bs.tensordot(c, a, axes=((2, 1), (1, 2)))
    return c.fuse_legs(axes=((0, 3), (1, 2)) + _extra(a))


def proj_corners(r0, r1, chi, svd_opts=None):
    """reference _env_ctm.py:1230-1258."""
    rr = bs.tensordot(r0, r1, axes=(1, 1))
    opts = dict(D_total=chi, fix_signs=True)
    opts.update(svd_opts or {})
    u, s, v = bs.svd_with_truncation(rr, axes=(0, 1), sU=r0.s[1], **opts)
    rs = s.rsqrt()
    p0 = bs.tensordot(r1, (rs @ v).conj(), axes=(0, 1)).unfuse_legs(axes=0)
    p1 = bs.tensordot(r0, (u @ rs).conj(), axes=(0, 0)).unfuse_legs(axes=0)
    return p0, p1


def proj_corners_many(pairs, chi, svd_opts=None):
    """proj_corners for several (r0, r1) pairs, with all their projector SVDs in one batched backend call
    (blocksparse.svd_many); truncation and the projectors per pair as in proj_corners."""
    opts = dict(D_total=chi)
    opts.update(svd_opts or {})
    rrs = [bs.tensordot(r0, r1, axes=(1, 1)) for r0, r1 in pairs]
    usv = bs.svd_many([(rr, (0, 1), r0.s[1], True) for rr, (r0, _) in zip(rrs, pairs)], fix_signs=True)
    out = []
    for (r0, r1), (u, s, v) in zip(pairs, usv):
        mask = bs.truncation_mask(s, **opts)
        u, s, v = bs.apply_mask(mask, u, s, v, axes=(-1, 0, 0))
        rs = s.rsqrt()
        p0 = bs.tensordot(r1, (rs @ v).conj(), axes=(0, 1)).unfuse_legs(axes=0)
        p1 = bs.tensordot(r0, (u @ rs).conj(), axes=(0, 0)).unfuse_legs(axes=0)
        out.append((p0, p1))
    return out


def _normalize(t):
    return t / t.config.backend.norm(t.data, 'inf')


def _r_pairs(h1, h2, svd_opts):
    """The (r0, r1) pairs of the two projector SVDs between the half-systems h1 and h2 (reference
    _env_ctm.py:1095-1096, 1122-1123, 1153-1154, 1180-1181).  With ``use_qr`` (a key of svd_opts, default
    False) r0 and r1 are the R factors of the half-systems, otherwise the half-systems themselves; both give
    the same projectors up to the signs fix_signs picks.  Returns the pairs and svd_opts without the key."""
    opts = dict(svd_opts or {})
    if opts.pop('use_qr', False):
        return [(bs.qr(h1, axes=(0, 1))[1], bs.qr(h2, axes=(1, 0))[1]),
                (bs.qr(h1, axes=(1, 0))[1], bs.qr(h2, axes=(0, 1))[1])], opts
    return [(h1, h2.transpose((1, 0))), (h1.transpose((1, 0)), h2)], opts


def move_h(env, a, chi, svd_opts=None, batch=False):
    ctl, ctr, cbr, cbl = corner_tl(env, a), corner_tr(env, a), corner_br(env, a), corner_bl(env, a)
    cor_tt = ctl @ ctr
    cor_bb = cbr @ cbl
    pairs, svd_opts = _r_pairs(cor_tt, cor_bb, svd_opts)   # 'rh', 'lh'
    if batch:
        (hrb, hrt), (hlb, hlt) = proj_corners_many(pairs, chi, svd_opts)
    else:
        hrb, hrt = proj_corners(*pairs[0], chi, svd_opts)
        hlb, hlt = proj_corners(*pairs[1], chi, svd_opts)
    new = {}
    tmp = env.l @ hlt
    tmp = bs.tensordot(a, tmp, axes=((0, 1), (2, 1)))
    tmp = bs.tensordot(hlb, tmp, axes=((0, 1), (2, 0)))
    new['l'] = _normalize(tmp)
    new['tl'] = _normalize(bs.tensordot(hlb, env.tl @ env.t, axes=((0, 1), (0, 1))))
    new['bl'] = _normalize(bs.tensordot(env.b, env.bl @ hlt, axes=((2, 1), (0, 1))))
    tmp = env.r @ hrb
    tmp = bs.tensordot(a, tmp, axes=((2, 3), (2, 1)))
    tmp = bs.tensordot(hrt, tmp, axes=((0, 1), (2, 0)))
    new['r'] = _normalize(tmp)
    new['tr'] = _normalize(bs.tensordot(env.t, env.tr @ hrb, axes=((2, 1), (0, 1))))
    new['br'] = _normalize(bs.tensordot(hrt, env.br @ env.b, axes=((0, 1), (0, 1))))
    kw = {k: getattr(env, k) for k in Env.NAMES}
    kw.update(new)
    return Env(**kw)


def move_v(env, a, chi, svd_opts=None, batch=False):
    ctl, ctr, cbr, cbl = corner_tl(env, a), corner_tr(env, a), corner_br(env, a), corner_bl(env, a)
    cor_ll = cbl @ ctl
    cor_rr = ctr @ cbr
    pairs, svd_opts = _r_pairs(cor_ll, cor_rr, svd_opts)   # 'tv', 'bv'
    if batch:
        (vtr, vtl), (vbr, vbl) = proj_corners_many(pairs, chi, svd_opts)
    else:
        vtr, vtl = proj_corners(*pairs[0], chi, svd_opts)
        vbr, vbl = proj_corners(*pairs[1], chi, svd_opts)
    new = {}
    tmp = bs.tensordot(vtl, env.t, axes=(0, 0))
    tmp = bs.tensordot(tmp, a, axes=((2, 0), (0, 1)))
    tmp = bs.tensordot(tmp, vtr, axes=((1, 3), (0, 1)))
    new['t'] = _normalize(tmp)
    new['tl'] = _normalize(bs.tensordot(env.l, env.tl @ vtr, axes=((2, 1), (0, 1))))
    new['tr'] = _normalize(bs.tensordot(vtl, env.tr @ env.r, axes=((0, 1), (0, 1))))
    tmp = bs.tensordot(vbr, env.b, axes=(0, 0))
    tmp = bs.tensordot(tmp, a, axes=((2, 0), (2, 3)))
    tmp = bs.tensordot(tmp, vbl, axes=((1, 3), (0, 1)))
    new['b'] = _normalize(tmp)
    new['bl'] = _normalize(bs.tensordot(vbr, env.bl @ env.l, axes=((0, 1), (0, 1))))
    new['br'] = _normalize(bs.tensordot(env.r, env.br @ vbl, axes=((2, 1), (0, 1))))
    kw = {k: getattr(env, k) for k in Env.NAMES}
    kw.update(new)
    return Env(**kw)


def ctm_step(env, a, chi, svd_opts=None, batch=False, checkpoint_move=False, A=None):
    """One 'hv' CTMRG step: 4 projector SVDs (two per move; with ``batch`` the two of a move go out as one
    batched SVD call, the reference computes them one after the other).

    checkpoint_move ('reentrant' / 'nonreentrant', needs the site tensor A): each move is a checkpoint region
    whose inputs and outputs cross it as split_data_and_meta of the environment and site dictionaries, as the
    reference's EnvCTM.update_ does (_env_ctm.py:566-589); the move is recomputed in the backward pass."""
    if not checkpoint_move:
        env = move_h(env, a, chi, svd_opts, batch)
        return move_v(env, a, chi, svd_opts, batch)
    import torch.utils.checkpoint as tc
    cfg = A.config

    def f_update(move, meta, *inputs):
        d = bs.combine_data_and_meta(inputs, meta)
        e = Env.from_dict(d, cfg)
        aa = double_layer(bs.Tensor.from_dict(d['psi'], cfg))
        e = (move_h if move == 'h' else move_v)(e, aa, chi, svd_opts, batch)
        data, meta_out = bs.split_data_and_meta(e.to_dict(level=0))
        return meta_out, *data

    for move in 'hv':
        d = env.to_dict(level=0)
        d['psi'] = A.to_dict(level=0)
        inputs, meta = bs.split_data_and_meta(d)
        meta_out, *data = tc.checkpoint(f_update, move, meta, *inputs,
                                        use_reentrant=checkpoint_move == 'reentrant')
        env = Env.from_dict(bs.combine_data_and_meta(data, meta_out), cfg)
    return env


def cluster_expectation(env, A, op):
    """<op> on the 2x2 cluster: Z(op) / Z(1), Z(X) = Tr(cor_tl(X) cor_tr cor_br cor_bl)."""
    a = double_layer(A)
    aO = double_layer(A, op)
    ctr, cbr, cbl = corner_tr(env, a), corner_br(env, a), corner_bl(env, a)
    bottom = cbr @ cbl

    def Z(corner):
        top = corner @ ctr
        return bs.tensordot(top, bottom, axes=((0, 1), (1, 0))).to_number()
    return Z(corner_tl(env, aO)) / Z(corner_tl(env, a))


def double_layer_open(A):
    """a(t, l, b, r, p, p*) = A(t, l, b, r, p) A*(t, l, b, r, p*), legs fused (ket, bra) pairwise and the
    physical legs left open.  The outer product is a contraction over a trailing dimension-1, charge-0
    leg, which leaves the block buffer unchanged."""
    cfg = A.config
    A = A.consume_transpose()
    z = cfg.sym.zero()
    Ax = bs.from_blocks(cfg, A.struct.s + (1,), A.n, tuple(t + z for t in A.struct.t),
                        tuple(D + (1,) for D in A.struct.D), A.data)._replace(hfs=A.hfs + (bs.Fusion(s=(1,)),))
    x = bs.tensordot(Ax, Ax.conj(), axes=(5, 5))  # (t, l, b, r, p, t*, l*, b*, r*, p*)
    return x.fuse_legs(axes=((0, 5), (1, 6), (2, 7), (3, 8), 4, 9))


def heisenberg_bond(config):
    """S.S on two spin-1/2 sites as a U(1)-symmetric 4-leg tensor H(o1, o2, i1, i2), s = (+1, +1, -1, -1),
    physical charges 0 (Sz = -1/2) and 1 (Sz = +1/2): <o1 o2|Sz Sz + (S+S- + S-S+)/2|i1 i2>."""
    pt, pD = (0, 1), (1, 1)
    legs = [bs.leg(config, s, pt, pD) for s in (1, 1, -1, -1)]
    st, _ = bs._fill_struct(config, legs, 0)
    sz = {0: -0.5, 1: 0.5}
    vals = []
    for o1, o2, i1, i2 in st.t:
        if (o1, o2) == (i1, i2):
            vals.append(sz[o1] * sz[o2])
        else:  # (o1, o2) = (i2, i1) with o1 != o2: one spin flip on each site
            vals.append(0.5)
    return bs.from_flat(config, legs, 0, np.array(vals))


def _bond_identity(config):
    pt, pD = (0, 1), (1, 1)
    legs = [bs.leg(config, s, pt, pD) for s in (1, 1, -1, -1)]
    st, _ = bs._fill_struct(config, legs, 0)
    return bs.from_flat(config, legs, 0, np.array([1.0 if t[:2] == t[2:] else 0.0 for t in st.t]))


def bond_energy(env, A, H):
    """Energy per site of a nearest-neighbour bond operator on a 1x1 unit cell: <H> on the horizontal bond
    (tl, tr) plus <H> on the vertical bond (bl, tl) of the 2x2 cluster, each as Tr(rho H) / Tr(rho) with
    rho the two-site reduced density matrix (the bond expectation values of the reference's measure_nn,
    _env_ctm_measure.py:79-175, in the 2x2-cluster form)."""
    a, ao = double_layer(A), double_layer_open(A)
    I = _bond_identity(A.config)
    ctl, ctr, cbr, cbl = corner_tl(env, ao), corner_tr(env, ao), corner_br(env, a), corner_bl(env, ao)

    def expect(rho):  # rho(p1, p1*, p2, p2*)
        num = bs.tensordot(rho, H, axes=((0, 1, 2, 3), (2, 0, 3, 1))).to_number()
        den = bs.tensordot(rho, I, axes=((0, 1, 2, 3), (2, 0, 3, 1))).to_number()
        return num / den
    # horizontal: Tr(ctl ctr cbr cbl) with open legs on tl and tr
    top = bs.tensordot(ctl, ctr, axes=(1, 0))                 # (x, p1, p1*, y, p2, p2*)
    bottom = cbr @ corner_bl(env, a)                          # (y, x)
    e_h = expect(bs.tensordot(top, bottom, axes=((0, 3), (1, 0))))
    # vertical: open legs on bl and tl
    left = bs.tensordot(cbl, corner_tl(env, ao), axes=(1, 0))  # (x, pb, pb*, y, pt, pt*)
    right = corner_tr(env, a) @ cbr                           # (y, x)
    e_v = expect(bs.tensordot(left, right, axes=((0, 3), (1, 0))))
    return e_h + e_v


def ctm_iterations(A, env, chi, n_iter, segment=0, svd_opts=None, on_step=None, batch=False, a=None):
    """``n_iter`` 'hv' steps from ``env``.  With ``segment`` > 0 and gradients enabled, each run of
    ``segment`` steps is a non-reentrant checkpoint region (the reference's checkpoint_move,
    _env_ctm.py:556-589): its environments are recomputed in the backward pass instead of kept."""
    import torch
    from torch.utils.checkpoint import checkpoint
    done = 0
    while done < n_iter:
        k = min(segment, n_iter - done) if segment else n_iter - done
        if segment and torch.is_grad_enabled():
            held = []

            def run(adata, *edata, k=k, env0=env):
                e = Env(**{n: getattr(env0, n)._replace(data=d) for n, d in zip(Env.NAMES, edata)})
                aa = double_layer(A._replace(data=adata))
                for _ in range(k):
                    e = ctm_step(e, aa, chi, svd_opts, batch)
                    if on_step is not None:
                        on_step()
                held[:] = [e]
                return tuple(t.data for t in e.tensors())
            out = checkpoint(run, A.data, *(t.data for t in env.tensors()), use_reentrant=False)
            env = Env(**{n: getattr(held[0], n)._replace(data=d) for n, d in zip(Env.NAMES, out)})
        else:
            a = double_layer(A) if a is None else a
            for _ in range(k):
                env = ctm_step(env, a, chi, svd_opts, batch)
                if on_step is not None:
                    on_step()
        done += k
    return env


def energy_loss_fn(A, env, H, chi, n_iter, segment=0, svd_opts=None, on_step=None, batch=False):
    """``n_iter`` CTMRG steps followed by the nearest-neighbour bond energy per site (real part)."""
    env2 = ctm_iterations(A, env, chi, n_iter, segment, svd_opts, on_step, batch)
    val = bond_energy(env2, A, H)
    return val.real if hasattr(val, 'is_complex') and val.is_complex() else val


def energy_fwd_bwd(A, env, H, chi, n_iter, segment=0, svd_opts=None, on_step=None, batch=False):
    """Unrolled CTMRG + energy, forward and backward w.r.t. the site tensor's data; returns (loss, grad)."""
    data = A.data.detach().clone().requires_grad_(True)
    loss = energy_loss_fn(A._replace(data=data), env, H, chi, n_iter, segment, svd_opts, on_step, batch)
    loss.backward()
    return loss.detach(), data.grad


# ------------------------------------------------------------------ initial environments, rdm1x1

def _pad_legs(x, s_left, s_right):
    """A rank-1 tensor x(k) as the rank-3 tensor (1, k, 1) with charge-zero legs of dimension 1 and
    signatures s_left, s_right (the buffer is unchanged)."""
    cfg = x.config
    z = cfg.sym.zero()
    t = tuple(z + tk + z for tk in x.struct.t)
    D = tuple((1, Dk[0], 1) for Dk in x.struct.D)
    y = bs.from_blocks(cfg, (s_left, x.struct.s[0], s_right), x.struct.n, t, D, x.data)
    return y._replace(hfs=(bs.Fusion(s=(s_left,)), x.hfs[0], bs.Fusion(s=(s_right,))))


def _identity(config, leg):
    """The identity matrix on leg.conj() x leg (blocks (q, q), unit diagonal)."""
    blocks = [((q, q), np.eye(d).ravel()) for q, d in zip(leg.t, leg.D)]
    t = tuple(tuple(q1) + tuple(q2) for (q1, q2), _ in blocks)
    D = tuple((d, d) for d in leg.D)
    data = np.concatenate([b for _, b in blocks]) if blocks else np.zeros(0)
    return bs.from_blocks(config, (-leg.s, leg.s), config.sym.zero(), t, D, data)


def env_eye(A, corner_s=(1, -1)):
    """The 'eye' initial environment of the reference (EnvCTM.reset_(init='eye'), _env_ctm.py:284-320):
    corners are the 1 x 1 identity of charge zero with signatures corner_s; each edge is
    identity_boundary (_env_contractions.py:46-50) of the double-layer leg it faces, i.e. the ket leg traced
    against the bra leg, between two dimension-1 legs of signatures (-corner_s[1], -corner_s[0])."""
    cfg = A.config
    z = cfg.sym.zero()
    one = lambda s: bs.leg(cfg, s, (z,), (1,))
    corner = bs.from_flat(cfg, [one(corner_s[0]), one(corner_s[1])], z, np.ones(1))
    edges = {}
    for k, i in (('t', 0), ('l', 1), ('b', 2), ('r', 3)):
        # the double layer fuses (ket leg i, bra leg i); the edge meets it with the conjugate fused leg
        leg = A.get_legs(i)
        bnd = _identity(cfg, leg).fuse_legs(axes=((0, 1),))
        edges[k] = _pad_legs(bnd, -corner_s[1], -corner_s[0])
    return Env(tl=corner, tr=corner, br=corner, bl=corner, **edges)


def env_expand(env, a):
    """One layer of double-layer sites added to every environment tensor without truncation (the reference's
    expand_outward_, _env_ctm.py:322-467, whose identity projectors amount to fusing each edge's outer leg with
    the site leg beside it); every tensor is rescaled by its largest modulus."""
    new = {'tl': corner_tl(env, a), 'tr': corner_tr(env, a), 'br': corner_br(env, a), 'bl': corner_bl(env, a)}
    x = bs.tensordot(env.l, a, axes=(1, 1))                       # (b, t, T, B, R)
    new['l'] = x.transpose((0, 3, 4, 1, 2)).fuse_legs(axes=((0, 1), 2, (3, 4)))
    x = bs.tensordot(env.r, a, axes=(1, 3))                       # (t, b, T, L, B)
    new['r'] = x.transpose((0, 2, 3, 1, 4)).fuse_legs(axes=((0, 1), 2, (3, 4)))
    x = bs.tensordot(env.t, a, axes=(1, 0))                       # (l, r, L, B, R)
    new['t'] = x.transpose((0, 2, 3, 1, 4)).fuse_legs(axes=((0, 1), 2, (3, 4)))
    x = bs.tensordot(env.b, a, axes=(1, 2))                       # (r, l, T, L, R)
    new['b'] = x.transpose((0, 4, 2, 1, 3)).fuse_legs(axes=((0, 1), 2, (3, 4)))
    return Env(**{k: _normalize(v) for k, v in new.items()})


def env_dl(A, a=None, corner_s=(1, -1)):
    """The 'dl' initial environment (EnvCTM.reset_(init='dl'), _env_ctm.py:278-285): 'eye', expanded by one
    layer of the double-layer site, each tensor rescaled by its largest modulus."""
    a = double_layer(A) if a is None else a
    return env_expand(env_eye(A, corner_s), a)


def rdm1x1(env, A):
    """One-site reduced density matrix rho(p, p*) / Tr(rho) of a 1x1 cell (the reference's rdm1x1,
    tn/fpeps/envs/rdm.py:211-243): the eight environment tensors around the double-layer site with open
    physical legs (ket p, bra p*).  Returns (rho normalised, Tr(rho))."""
    ao = double_layer_open(A)
    vect = (env.l @ env.tl) @ (env.t @ env.tr)                    # (b_l, L, T, b_tr)
    vecb = (env.r @ env.br) @ (env.b @ env.bl)                    # (t_r, R, B, t_bl)
    tmp = bs.tensordot(vect, ao, axes=((1, 2), (1, 0)))           # (b_l, b_tr, B, R, p, p*)
    rho = bs.tensordot(vecb, tmp, axes=((0, 1, 2, 3), (1, 3, 2, 0)))  # (p, p*)
    norm = trace(rho)
    return rho / norm, norm


def trace(x, block=None):
    """Trace of a rank-2 tensor (of one diagonal block when ``block`` names its charges (q, q))."""
    cfg = x.config
    be = cfg.backend
    x = x.consume_transpose()
    if block is not None:  # ((q,), (q,)) or (q, q)
        block = tuple(c for q in block for c in (q if isinstance(q, tuple) else (q,)))
    parts = []
    for t, D, sl in zip(x.struct.t, x.struct.D, x.slices):
        if block is not None and tuple(t) != block:
            continue
        if D[0] != D[1]:
            raise bs.BlockSparseError('trace of a non-square block')
        lo, hi = sl.slcs[0]
        parts.append(x._data[lo:hi].reshape(D).diagonal().sum())
    if not parts:
        return be.zeros((), dtype=x.yastn_dtype, device=cfg.device)
    out = parts[0]
    for p in parts[1:]:
        out = out + p
    return out


def rdm_loss(A, chi, sweeps, init='dl', svd_opts=None, segment=0, block=None):
    """Re Tr_block rho after ``sweeps`` 'hv' CTMRG steps from the ``init`` environment (the cost function of the
    reference's tests/peps/test_ad.py:147-165: use_qr=False, projectors truncated by ``svd_opts``)."""
    a = double_layer(A)
    env = env_dl(A, a) if init == 'dl' else env_eye(A)
    env = ctm_iterations(A, env, chi, sweeps, segment=segment, svd_opts=svd_opts, a=a)
    rho, _ = rdm1x1(env, A)
    val = trace(rho, block)
    return val.real if hasattr(val, 'is_complex') and val.is_complex() else val


# ------------------------------------------------------------------ CTM step as a function of flat buffers

def env_flat(env):
    """The eight environment buffers concatenated (the 1-D form the reference's FixedPoint differentiates,
    fixed_pt.py _concat_data)."""
    import torch
    return torch.cat([t.data.reshape(-1) for t in env.tensors()])


def env_from_flat(template, flat):
    """An environment with template's structures and the buffers cut from ``flat``."""
    out, off = {}, 0
    for k in Env.NAMES:
        t = getattr(template, k)
        out[k] = t._replace(data=flat[off:off + t.size])
        off += t.size
    return Env(**out)


def ctm_step_flat(A, template, chi, svd_opts=None):
    """f(C, a) = one CTM step of the environment buffer C for the site buffer a, flattened: the map whose
    vector-Jacobian products the reference's fixed-point AD takes with torch.func.vjp (fixed_pt.py:853).
    The output structure must stay that of ``template`` (static-structure truncation)."""
    def f(C, adata):
        e = ctm_step(env_from_flat(template, C), double_layer(A._replace(data=adata)), chi, svd_opts)
        return env_flat(e)
    return f


# ------------------------------------------------------------------ workload setup

def site_legs(config, D):
    """Virtual leg with charges {-1:2, 0:2, 1:2} for D=6, {-2:1,-1:2,0:2,1:2,2:1} for D=8."""
    if D == 6:
        vt, vD = (-1, 0, 1), (2, 2, 2)
    elif D == 8:
        vt, vD = (-2, -1, 0, 1, 2), (1, 2, 2, 2, 1)
    elif D == 2:
        vt, vD = (0, 1), (1, 1)
    elif D == 3:
        vt, vD = (-1, 0, 1), (1, 1, 1)
    else:
        raise ValueError(D)
    return vt, vD


def env_charges(chi):
    if chi == 64:
        qs, w = list(range(-4, 5)), (lambda q: math.exp(-q * q / 6))
    elif chi == 128:
        qs, w = list(range(-6, 7)), (lambda q: math.exp(-q * q / 10))
    else:  # small test sizes
        r = 1 if chi < 16 else 2
        qs, w = list(range(-r, r + 1)), (lambda q: math.exp(-q * q / 4))
    D = split_dims(chi, qs, w)
    if min(D) < 1:
        raise ValueError(f"chi={chi} too small for charges {qs}")
    return tuple(qs), D


def setup(config, D=6, chi=64, dtype='float64', seed=0):
    """Site tensor A(t, l, b, r, p), s=(-1, 1, 1, -1, 1), N(0,1) / sqrt(numel); environment N(0,1)."""
    vt, vD = site_legs(config, D)
    pt, pD = (0, 1), (1, 1)
    L = lambda s, t, Dd: bs.leg(config, s, t, Dd)
    legs = [L(-1, vt, vD), L(1, vt, vD), L(1, vt, vD), L(-1, vt, vD), L(1, pt, pD)]
    st, _ = bs._fill_struct(config, legs, 0)
    A = bs.from_flat(config, legs, 0, bs.seeded_normal(st.size, dtype, seed) / math.sqrt(st.size))
    a = double_layer(A)
    aL = [a.get_legs(i) for i in range(4)]  # (T, L, B, R) double-layer legs
    et, eD = env_charges(chi)
    E = lambda s: bs.leg(config, s, et, eD)
    DL = lambda i, s: bs.Leg(s, aL[i].t, aL[i].D)
    specs = {
        'tl': [E(-1), E(1)], 'tr': [E(-1), E(1)], 'br': [E(-1), E(1)], 'bl': [E(-1), E(1)],
        't': [E(-1), DL(0, 1), E(1)], 'r': [E(-1), DL(3, 1), E(1)],
        'b': [E(-1), DL(2, -1), E(1)], 'l': [E(-1), DL(1, -1), E(1)],
    }
    # the D^2 leg of each edge carries the fusion structure of the double-layer leg it meets
    fused_of = {'t': (0, True), 'r': (3, True), 'b': (2, False), 'l': (1, False)}
    env = {}
    for i, k in enumerate(Env.NAMES):
        st, _ = bs._fill_struct(config, specs[k], 0)
        x = bs.from_flat(config, specs[k], 0, bs.seeded_normal(st.size, dtype, seed + 10 + i))
        if k in fused_of:
            ia, flip = fused_of[k]
            hf = a.hfs[ia].conj() if flip else a.hfs[ia]
            x = x._replace(hfs=(x.hfs[0], hf, x.hfs[2]))
        env[k] = x
    sz = bs.from_flat(config, [L(1, pt, pD), L(-1, pt, pD)], 0, np.array([-0.5, 0.5]))
    return A, Env(**env), sz


def loss_fn(A, env, sz, chi, svd_opts=None, batch=False):
    """Forward of one CTMRG 'hv' step followed by the 2x2-cluster <Sz>."""
    a = double_layer(A)
    env2 = ctm_step(env, a, chi, svd_opts, batch)
    val = cluster_expectation(env2, A, sz)
    return val.real if hasattr(val, 'is_complex') and val.is_complex() else val


def fwd_bwd(A, env, sz, chi, svd_opts=None, batch=False):
    """One CTMRG step forward + backward w.r.t. the site tensor's data; returns (loss, grad)."""
    import torch
    data = A.data.detach().clone().requires_grad_(True)
    A2 = A._replace(data=data)
    loss = loss_fn(A2, env, sz, chi, svd_opts, batch)
    loss.backward()
    return loss.detach(), data.grad


def env_split(chi):
    """Per-charge dimensions of the environment leg at bond dimension chi (the setup's split)."""
    et, eD = env_charges(chi)
    return {(t,): D for t, D in zip(et, eD)}


def static_svd_opts(chi, check=None):
    """Projector truncation with a fixed structure: every projector keeps env_split(chi)[t] values of sector
    t (blocksparse.truncation_mask_static, passed as the reference's custom mask function mask_f), so the
    environment keeps its sectors from step to step and a step has no host synchronisation."""
    D_block = env_split(chi)
    return {'mask_f': lambda S: bs.truncation_mask_static(S, D_block, check)}


class CapturedStep:
    """A CTMRG step and its loss, forward and backward, captured as one HIP graph (torch.cuda.CUDAGraph):
    static-structure truncation, the Jacobi convergence flags and the structure check read after a replay
    (one synchronisation), all plans built and tuned by ``warmup`` eager runs first.

    ``run(data)`` is the step (it must call backward on the loss and return it); the graph's inputs are the
    tensors it closes over, ``data`` (requires grad) among them.  ``replay()`` returns the loss and gradient
    tensors of the graph (overwritten by the next replay)."""

    def __init__(self, run, data, check, warmup=2):
        import torch
        from . import _functions as F
        self.run, self.data, self.check = run, data, check
        for _ in range(warmup):
            self.data.grad = None
            self.run(self.data)
        torch.cuda.synchronize()
        self.data.grad = None
        self.graph = torch.cuda.CUDAGraph()
        self.checks = F.deferred_checks()
        with self.checks:
            with torch.cuda.graph(self.graph):
                self.loss = self.run(self.data)
        self.grad = self.data.grad

    def replay(self):
        self.check.zero_()
        self.graph.replay()
        return self.loss, self.grad

    def validate(self):
        """Raise if a Jacobi call of the last replay did not converge or the fixed structure was invalid;
        return the number of projector truncations where the global rule would have kept another set."""
        import torch
        torch.cuda.synchronize()
        self.checks.check()
        zero_kept, other_set = self.check.tolist()
        if zero_kept:
            raise RuntimeError(f"static truncation kept a zero singular value in {zero_kept} projector(s)")
        return other_set


def captured_fwd_bwd(A, env, sz, chi, warmup=2, batch=False):
    """One CTMRG step + <Sz> with static-structure projectors (fwd + bwd) as a CapturedStep; also returns the
    eager (loss, grad) of the same step for comparison.  ``batch``: the two projector SVDs of each move in one
    batched backend call."""
    import torch
    check = torch.zeros(2, dtype=torch.int64, device=A.data.device)
    opts = static_svd_opts(chi, check)

    def run(data):
        loss = loss_fn(A._replace(data=data), env, sz, chi, opts, batch)
        loss.backward()
        return loss
    data = A.data.detach().clone().requires_grad_(True)
    cs = CapturedStep(run, data, check, warmup)
    d2 = A.data.detach().clone().requires_grad_(True)
    l2 = run(d2)
    return cs, (l2.detach(), d2.grad)


def host_info():
    """Host threads and BLAS of the CPU baselines (torch and numpy on the host cores)."""
    import torch
    info = {'torch_threads': torch.get_num_threads(), 'os_cpu_count': os.cpu_count(),
            'OMP_NUM_THREADS': os.environ.get('OMP_NUM_THREADS'),
            'OPENBLAS_NUM_THREADS': os.environ.get('OPENBLAS_NUM_THREADS')}
    try:
        from threadpoolctl import threadpool_info
        info['blas'] = [{k: x.get(k) for k in ('internal_api', 'version', 'num_threads', 'prefix')}
                        for x in threadpool_info() if x.get('user_api') == 'blas']
    except Exception:  # threadpoolctl absent
        pass
    try:
        info['torch_blas'] = torch.__config__.show().split('BLAS_INFO=')[1].split(',')[0]
    except Exception:
        pass
    return info


def bench(steps=3, D=6, chi=64, plain=True, cpu=True, host_runs=10, host_warmup=3):
    """Seconds per CTMRG fwd+bwd step (D=6, chi=64, U(1), float64): batched HIP backend (eager, the
    reference's data-dependent truncation) and the same step with static-structure truncation replayed as
    one HIP graph, vs the plain per-block GPU path vs the plain path on the host cores (torch CPU, median of
    host_runs runs after host_warmup warm-up steps)."""
    import torch
    from . import backend_hip, backend_plain
    from . import _functions as F
    out = {'workload': f'CTMRG hv step fwd+bwd, U(1) iPEPS 1x1, D={D}, chi={chi}, float64; loss = 2x2 cluster <Sz>'}

    def run(backend, device, n, batch=False, median=False, warmup=1):
        cfg = bs.make_config(backend, 'U1', device=device)
        A, env, sz = setup(cfg, D=D, chi=chi, seed=0)
        for _ in range(warmup):  # warm-up (plans)
            fwd_bwd(A, env, sz, chi, batch=batch)
        sync = torch.cuda.synchronize if device != 'cpu' else (lambda: None)
        sync()
        ts = []
        for _ in range(n if median else 1):
            t0 = time.perf_counter()
            for _ in range(1 if median else n):
                loss, g = fwd_bwd(A, env, sz, chi, batch=batch)
            sync()
            ts.append((time.perf_counter() - t0) / (1 if median else n))
        return float(np.median(ts)), float(loss), g
    tb, lb, gb = run(backend_hip, 'cuda:0', steps)
    out['s_per_step_batched'] = tb
    out['loss'] = lb
    # the two projector SVDs of each move in one batched call (the reference calls them one by one)
    tm, lm, gm = run(backend_hip, 'cuda:0', steps, batch=True)
    out['move_batched_svds'] = {'s_per_step': tm, 'loss_diff': abs(lm - lb),
                                'grad_rel_diff': float((gm - gb).norm() / gb.norm())}
    # static-structure truncation, the whole fwd+bwd step captured as one HIP graph
    cfg = bs.make_config(backend_hip, 'U1', device='cuda:0')
    A, env, sz = setup(cfg, D=D, chi=chi, seed=0)
    cs, (le, ge) = captured_fwd_bwd(A, env, sz, chi)
    cs.replay()
    other = cs.validate()
    t0 = time.perf_counter()
    for _ in range(steps):
        cs.replay()
    torch.cuda.synchronize()
    tg = (time.perf_counter() - t0) / steps
    cs.validate()
    loss_g, grad_g = cs.replay()
    cs.validate()
    out['graph_static'] = {'s_per_step': tg, 'loss': float(loss_g),
                           'bitwise_equal_eager': bool(torch.equal(loss_g, le) and torch.equal(grad_g, ge)),
                           'truncation': 'static: per-sector kept counts = the environment split '
                                         f'{ {k[0]: v for k, v in env_split(chi).items()} }',
                           'projectors_where_global_rule_differs': other}
    loss_g, grad_g = loss_g.clone(), grad_g.clone()  # (the graph's outputs go with the graph)
    del cs
    # the same, with the two projector SVDs of each move in one batched call
    cs, (lm_e, gm_e) = captured_fwd_bwd(A, env, sz, chi, batch=True)
    cs.replay()
    cs.validate()
    t0 = time.perf_counter()
    for _ in range(steps):
        cs.replay()
    torch.cuda.synchronize()
    tgm = (time.perf_counter() - t0) / steps
    cs.validate()
    loss_gm, grad_gm = cs.replay()
    cs.validate()
    out['graph_static_move_batched'] = {
        's_per_step': tgm, 'bitwise_equal_eager': bool(torch.equal(loss_gm, lm_e) and torch.equal(grad_gm, gm_e)),
        'loss_diff_vs_graph_static': abs(float(loss_gm) - float(loss_g)),
        'grad_rel_diff_vs_graph_static': float((grad_gm - grad_g).norm() / grad_g.norm())}
    del cs
    if plain:
        tp, lp, gp = run(backend_plain, 'cuda:0', max(1, steps // 2))
        out['s_per_step_plain_gpu'] = tp
        out['speedup_vs_plain_gpu'] = tp / tb
        out['move_batched_svds']['speedup_vs_plain_gpu'] = tp / tm
        out['graph_static']['speedup_vs_plain_gpu'] = tp / tg
        out['graph_static_move_batched']['speedup_vs_plain_gpu'] = tp / tgm
        out['grad_rel_diff_vs_plain'] = float((gb - gp).norm() / gp.norm())
        out['plain_svd'] = ('torch.linalg.svd per block, full_matrices=False, preferred linalg library '
                            f'{torch.backends.cuda.preferred_linalg_library()}')
    if cpu:
        tc, lc, gc = run(backend_plain, 'cpu', host_runs, median=True, warmup=host_warmup)
        out['s_per_step_host_torch'] = tc
        out['host_runs'] = host_runs
        out['host_warmup'] = host_warmup
        out['host'] = host_info()
        out['speedup_vs_host'] = tc / tb
        out['move_batched_svds']['speedup_vs_host'] = tc / tm
        out['graph_static']['speedup_vs_host'] = tc / tg
        out['graph_static_move_batched']['speedup_vs_host'] = tc / tgm
    return out


def bench_unrolled(D=8, chi=128, dtype='complex128', n_iter=20, segment=4, repeats=1, plain_iter=0,
                   no_checkpoint_memory=False):
    """Seconds per step of ``n_iter`` CTMRG 'hv' steps + the nearest-neighbour Heisenberg energy, forward and
    backward with checkpoint segments of ``segment`` steps, on the batched HIP backend, with the peak device
    memory of the step.  With ``plain_iter`` > 0 the batched and the plain per-block GPU paths also run a
    ``plain_iter``-step version of the same step (the plain path is too slow for the full unrolling)."""
    import torch
    from . import backend_hip, backend_plain
    out = {'workload': f'{n_iter} CTMRG hv steps + nearest-neighbour Heisenberg energy per site, fwd+bwd, U(1) '
                       f'iPEPS 1x1, D={D}, chi={chi}, {dtype}; checkpoint segments of {segment} steps'}
    count = [0]

    def tick():
        count[0] += 1
        if count[0] % 4 == 0:
            print(f'  ctm steps run: {count[0]}', flush=True)

    def run(backend, n, nit, seg, batch=False):
        cfg = bs.make_config(backend, 'U1', device='cuda:0')
        A, env, _ = setup(cfg, D=D, chi=chi, dtype=dtype, seed=0)
        H = heisenberg_bond(cfg)
        best, peak = float('inf'), 0
        for i in range(n + 1):  # the first run builds and tunes the plans
            torch.cuda.synchronize()
            torch.cuda.reset_peak_memory_stats()
            base = torch.cuda.memory_allocated()
            t0 = time.perf_counter()
            loss, g = energy_fwd_bwd(A, env, H, chi, nit, seg, on_step=tick, batch=batch)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            print(f'  {backend.__name__.split(".")[-1]} {nit} steps segment {seg}: {dt:.3f} s', flush=True)
            if i > 0 or n == 0:
                best = min(best, dt)
                peak = torch.cuda.max_memory_allocated() - base
        return best, float(loss), g, peak
    tb, lb, gb, pk = run(backend_hip, repeats, n_iter, segment)
    out.update(s_per_step_batched=tb, energy=lb, peak_hbm_gb=pk / 1e9)
    if no_checkpoint_memory:
        t0, l0, g0, pk0 = run(backend_hip, 0, n_iter, 0)
        out['no_checkpoint'] = {'s_per_step': t0, 'peak_hbm_gb': pk0 / 1e9,
                                'grad_rel_diff_vs_checkpointed': float((g0 - gb).norm() / gb.norm())}
    if plain_iter:
        seg = min(segment, plain_iter)
        ts, ls, gs, _ = run(backend_hip, repeats, plain_iter, seg)
        tp, lp, gp, _ = run(backend_plain, 0, plain_iter, seg)
        tm, lm, gm, _ = run(backend_hip, repeats, plain_iter, seg, batch=True)
        out['short'] = {'ctm_steps': plain_iter, 'segment': seg, 's_batched': ts, 's_plain_gpu': tp,
                        'speedup_vs_plain_gpu': tp / ts, 'energy_diff': abs(ls - lp),
                        'grad_rel_diff_vs_plain': float((gs - gp).norm() / gp.norm()),
                        'move_batched_svds': {'s': tm, 'speedup_vs_plain_gpu': tp / tm,
                                              'energy_diff': abs(lm - lp),
                                              'grad_rel_diff_vs_plain': float((gm - gp).norm() / gp.norm())}}
    return out


def _main():
    """python -m yastn_amd.ctmrg [--steps N] [--graph] [--batch]: N fwd+bwd steps of the D=6, chi=64 CTMRG
    workload on the HIP backend (eager, or replays of the captured static-structure step), for profiling."""
    import argparse
    import torch
    from . import backend_hip
    ap = argparse.ArgumentParser()
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--graph', action='store_true')
    ap.add_argument('--batch', action='store_true', help='move-batched projector SVDs (eager only)')
    ap.add_argument('--D', type=int, default=6)
    ap.add_argument('--chi', type=int, default=64)
    args = ap.parse_args()
    cfg = bs.make_config(backend_hip, 'U1', device='cuda:0')
    A, env, sz = setup(cfg, D=args.D, chi=args.chi, seed=0)
    if args.graph:
        cs, _ = captured_fwd_bwd(A, env, sz, args.chi)
        run = lambda: cs.replay()
    else:
        fwd_bwd(A, env, sz, args.chi, batch=args.batch)
        run = lambda: fwd_bwd(A, env, sz, args.chi, batch=args.batch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    torch.cuda.synchronize()
    print(f'{args.steps} steps, {(time.perf_counter() - t0) / args.steps * 1e3:.2f} ms/step '
          f'({"graph" if args.graph else "eager"})', flush=True)


if __name__ == '__main__':
    _main()
