"""Differentiable CTMRG step for a U(1) iPEPS on a 1x1 unit cell (the harness of the CTMRG workloads).

Structure restated from the reference (yastn/tn/fpeps/envs/_env_ctm.py, _env_contractions.py):
  * environment tensors per direction with the reference's leg conventions: corners tl (b, r),
    tr (l, b), br (t, l), bl (r, t); edges t (l, D^2, r), l (b, D^2, t), b (r, D^2, l), r (t, D^2, b);
  * 2x2 enlarged corners (corner2x2_*, _env_contractions.py:429-475) on the double-layer site;
  * projectors from proj_corners (_env_ctm.py:1230-1258): rr = r0 r1^T, svd_with_truncation with
    fix_signs=True, p0 = r1 (S^-1/2 V)^*, p1 = r0 (U S^-1/2)^*; use_qr=False (r_t = cor_tt, ...) as in
    update_extended_2x2_projectors_ (_env_ctm.py:1050-1183);
  * environment renormalisation of _update_env_ (_env_ctm.py:692-775) with norm(p='inf') rescaling;
  * an 'hv' step = horizontal move (projectors for the left and right bonds, then l, tl, bl, r, tr, br)
    followed by the vertical move (t, tl, tr, b, bl, br).
The double-layer site is formed explicitly (ket x bra contracted over the physical leg, legs fused
pairwise); the reference contracts ket and bra lazily.  The observable is <O> on the 2x2 cluster:
Z(O)/Z(1) with Z(X) = Tr(cor_tl(X) cor_tr cor_br cor_bl).
"""
import math
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


def double_layer(A, op=None):
    """a(t, l, b, r) = sum_p A(t, l, b, r, p) [op] A*(t, l, b, r, p), legs fused (ket, bra) pairwise."""
    Ao = A if op is None else bs.tensordot(A, op, axes=(4, 1))
    x = bs.tensordot(Ao, A.conj(), axes=(4, 4))  # (t, l, b, r, t*, l*, b*, r*)
    return x.fuse_legs(axes=((0, 4), (1, 5), (2, 6), (3, 7)))


def corner_tl(env, a):
    c = env.l @ env.tl @ env.t
    c = bs.tensordot(c, a, axes=((2, 1), (0, 1)))
    return c.fuse_legs(axes=((0, 2), (1, 3)))


def corner_tr(env, a):
    c = env.t @ env.tr @ env.r
    c = bs.tensordot(c, a, axes=((1, 2), (0, 3)))
    return c.fuse_legs(axes=((0, 2), (1, 3)))


def corner_br(env, a):
    c = env.r @ env.br @ env.b
    c = bs.tensordot(c, a, axes=((2, 1), (2, 3)))
    return c.fuse_legs(axes=((0, 2), (1, 3)))


def corner_bl(env, a):
    c = env.b @ env.bl @ env.l
    c = bs.tensordot(c, a, axes=((2, 1), (1, 2)))
    return c.fuse_legs(axes=((0, 3), (1, 2)))


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


def _normalize(t):
    return t / t.config.backend.norm(t.data, 'inf')


def move_h(env, a, chi, svd_opts=None):
    ctl, ctr, cbr, cbl = corner_tl(env, a), corner_tr(env, a), corner_br(env, a), corner_bl(env, a)
    cor_tt = ctl @ ctr
    cor_bb = cbr @ cbl
    hrb, hrt = proj_corners(cor_tt, cor_bb.transpose((1, 0)), chi, svd_opts)   # 'rh'
    hlb, hlt = proj_corners(cor_tt.transpose((1, 0)), cor_bb, chi, svd_opts)   # 'lh'
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


def move_v(env, a, chi, svd_opts=None):
    ctl, ctr, cbr, cbl = corner_tl(env, a), corner_tr(env, a), corner_br(env, a), corner_bl(env, a)
    cor_ll = cbl @ ctl
    cor_rr = ctr @ cbr
    vtr, vtl = proj_corners(cor_ll, cor_rr.transpose((1, 0)), chi, svd_opts)   # 'tv'
    vbr, vbl = proj_corners(cor_ll.transpose((1, 0)), cor_rr, chi, svd_opts)   # 'bv'
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


def ctm_step(env, a, chi, svd_opts=None):
    """One 'hv' CTMRG step: 4 projector SVDs (two per move)."""
    env = move_h(env, a, chi, svd_opts)
    return move_v(env, a, chi, svd_opts)


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


def loss_fn(A, env, sz, chi, svd_opts=None):
    """Forward of one CTMRG 'hv' step followed by the 2x2-cluster <Sz>."""
    a = double_layer(A)
    env2 = ctm_step(env, a, chi, svd_opts)
    val = cluster_expectation(env2, A, sz)
    return val.real if hasattr(val, 'is_complex') and val.is_complex() else val


def fwd_bwd(A, env, sz, chi, svd_opts=None):
    """One CTMRG step forward + backward w.r.t. the site tensor's data; returns (loss, grad)."""
    import torch
    data = A.data.detach().clone().requires_grad_(True)
    A2 = A._replace(data=data)
    loss = loss_fn(A2, env, sz, chi, svd_opts)
    loss.backward()
    return loss.detach(), data.grad


def bench(steps=3, D=6, chi=64, plain=True, cpu=True):
    """Seconds per CTMRG fwd+bwd step (D=6, chi=64, U(1), float64): batched HIP backend vs the plain
    per-block GPU path vs the plain path on the host cores (torch CPU)."""
    import torch
    from . import backend_hip, backend_plain
    out = {'workload': f'CTMRG hv step fwd+bwd, U(1) iPEPS 1x1, D={D}, chi={chi}, float64; loss = 2x2 cluster <Sz>'}

    def run(backend, device, n):
        cfg = bs.make_config(backend, 'U1', device=device)
        A, env, sz = setup(cfg, D=D, chi=chi, seed=0)
        fwd_bwd(A, env, sz, chi)  # warm-up (plans)
        sync = torch.cuda.synchronize if device != 'cpu' else (lambda: None)
        sync()
        t0 = time.perf_counter()
        for _ in range(n):
            loss, g = fwd_bwd(A, env, sz, chi)
        sync()
        return (time.perf_counter() - t0) / n, float(loss), g
    tb, lb, gb = run(backend_hip, 'cuda:0', steps)
    out['s_per_step_batched'] = tb
    out['loss'] = lb
    if plain:
        tp, lp, gp = run(backend_plain, 'cuda:0', max(1, steps // 2))
        out['s_per_step_plain_gpu'] = tp
        out['speedup_vs_plain_gpu'] = tp / tb
        out['grad_rel_diff_vs_plain'] = float((gb - gp).norm() / gp.norm())
    if cpu:
        import os
        tc, lc, gc = run(backend_plain, 'cpu', 1)
        out['s_per_step_host_torch'] = tc
        out['host_threads'] = torch.get_num_threads()
        out['speedup_vs_host'] = tc / tb
    return out
