"""Write golden vectors for the block-sparse backend hot path.

Run in a container where the reference (yastn) source tree is present:

    python tests/golden/make_golden.py /root/reference

It imports the reference, swaps its numpy backend for a recording proxy, runs a
few small scenarios through the reference Tensor layer, and writes for every
hot-path backend call (transpose_and_merge, unmerge, dot, svd, eigh,
fix_svd_signs, apply_mask, transpose) its meta arguments (JSON) and its input
and output arrays (NPZ).  It also records input/output structure of the
tensors of each scenario (so a restated planner can be checked for identical
block order and offsets), and gradients computed by the reference torch
backend on the CPU (so hand-written adjoints can be checked against it).

Only the vectors travel; the reference itself is never copied.
The metas pin what nothing else pins: the reference planner's block order,
offsets and fused-leg decompositions (yastn/tensor/_merging.py,
_contractions.py, linalg.py), and the reference torch backend's adjoints.
"""
import json
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# hot-path backend functions whose calls are recorded
RECORDED = ('transpose_and_merge', 'unmerge', 'dot', 'svd', 'eigh', 'fix_svd_signs',
            'apply_mask', 'transpose', 'embed_mask', 'transpose_dot_sum', 'dot_diag', 'qr')


def _stub_opt_einsum():
    """yastn/__init__.py imports opt_einsum (absent here); it is only used in path optimisation."""
    oe = types.ModuleType('opt_einsum')
    oec = types.ModuleType('opt_einsum.contract')

    class PathInfo:  # placeholder type, never instantiated
        pass
    oec.PathInfo = PathInfo
    oe.contract = oec
    sys.modules.setdefault('opt_einsum', oe)
    sys.modules.setdefault('opt_einsum.contract', oec)


def _jsonable(x):
    if isinstance(x, (tuple, list)):
        return [_jsonable(y) for y in x]
    if isinstance(x, (np.integer,)):
        return int(x)
    if isinstance(x, (np.floating,)):
        return float(x)
    if isinstance(x, float) and x == float('inf'):
        return 'inf'
    return x


class Recorder:
    def __init__(self):
        self.calls = []
        self.arrays = {}
        self.on = True

    def add_array(self, key, arr):
        self.arrays[key] = np.array(arr)
        return key


def make_proxy(base, rec):
    """A module object forwarding to ``base`` and recording the hot-path calls."""
    proxy = types.ModuleType('recording_backend')
    for name in dir(base):
        if not name.startswith('__'):
            setattr(proxy, name, getattr(base, name))

    def wrap(name):
        fn = getattr(base, name)

        def wrapped(*args, **kwargs):
            out = fn(*args, **kwargs)
            if rec.on:
                i = len(rec.calls)
                entry = {'fn': name, 'args': [], 'kwargs': {}, 'out': []}
                for j, a in enumerate(args):
                    if isinstance(a, np.ndarray):
                        entry['args'].append({'array': rec.add_array(f'c{i}_a{j}', a)})
                    elif isinstance(a, dict):  # mask dict: charge -> index array
                        entry['args'].append({'mask': [[_jsonable(k), {'slice': [v.start, v.stop, v.step]}
                                                        if isinstance(v, slice) else _jsonable(np.asarray(v).tolist())]
                                                       for k, v in a.items()]})
                    else:
                        entry['args'].append({'value': _jsonable(a)})
                for k, v in kwargs.items():
                    entry['kwargs'][k] = _jsonable(v)
                outs = out if isinstance(out, tuple) else (out,)
                for j, o in enumerate(outs):
                    entry['out'].append(rec.add_array(f'c{i}_o{j}', o))
                rec.calls.append(entry)
            return out
        return wrapped

    for name in RECORDED:
        setattr(proxy, name, wrap(name))
    proxy.BACKEND_ID = base.BACKEND_ID
    return proxy


def struct_dict(t):
    return {'s': list(t.struct.s), 'n': list(t.struct.n), 'diag': bool(t.struct.diag),
            't': _jsonable(t.struct.t), 'D': _jsonable(t.struct.D), 'size': int(t.struct.size),
            'slices': [list(sl.slcs[0]) for sl in t.slices], 'trans': list(t.trans),
            'hfs': [{'tree': list(h.tree), 'op': h.op, 's': list(h.s), 't': _jsonable(h.t), 'D': _jsonable(h.D)}
                    for h in t.hfs],
            'mfs': _jsonable(t.mfs)}


def seeded(n, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == 'complex128':
        re = rng.standard_normal(n)
        im = rng.standard_normal(n)
        return re + 1j * im
    return rng.standard_normal(n)


class Scenario:
    def __init__(self, name, rec):
        self.name = name
        self.rec = rec
        self.tensors = {}

    def keep(self, label, t):
        key = self.rec.add_array(f't_{label}', t.data)
        d = struct_dict(t)
        d['data'] = key
        d['dtype'] = t.yastn_dtype
        self.tensors[label] = d

    def dump(self, extra=None):
        out = {'scenario': self.name, 'calls': self.rec.calls, 'tensors': self.tensors}
        if extra:
            out.update(extra)
        with open(os.path.join(HERE, f'{self.name}.json'), 'w') as f:
            json.dump(out, f, separators=(',', ':'))
        np.savez_compressed(os.path.join(HERE, f'{self.name}.npz'), **self.rec.arrays)
        print(self.name, len(self.rec.calls), 'calls', sum(a.nbytes for a in self.rec.arrays.values()), 'bytes')


def main(ref_root):
    _stub_opt_einsum()
    sys.path.insert(0, ref_root)
    import yastn
    from yastn.backend import backend_np

    def cfg(sym, rec, policy='fuse_contracted'):
        return yastn.make_config(sym=sym, backend=make_proxy(backend_np, rec), tensordot_policy=policy)

    def tensor(config, legs, n, dtype, seed):
        a = yastn.zeros(config=config, legs=legs, n=n, dtype=dtype)
        a._data = seeded(a.size, dtype, seed).astype(dtype)
        return a

    # ---------------- MPS transfer contraction + two-site truncated SVD (U1) -------------
    # A(l, p, r) with s=(+1,+1,-1); virtual charges -4..4, D(q)=round(16 exp(-q^2/8)); physical {0,1}.
    tq = tuple(range(-4, 5))
    Dq = tuple(int(round(16 * np.exp(-q * q / 8))) for q in tq)
    assert Dq == (2, 5, 10, 14, 16, 14, 10, 5, 2)
    for policy in ('fuse_contracted', 'fuse_to_matrix'):
        rec = Recorder()
        sc = Scenario(f'mps_u1_{policy}', rec)
        c = cfg('U1', rec, policy)
        legs = [yastn.Leg(c, s=1, t=tq, D=Dq), yastn.Leg(c, s=1, t=(0, 1), D=(1, 1)), yastn.Leg(c, s=-1, t=tq, D=Dq)]
        A = tensor(c, legs, 0, 'float64', 0)
        sc.keep('A', A)
        T = yastn.tensordot(A, A.conj(), axes=((0, 1), (0, 1)))
        sc.keep('T', T)
        AA = A @ A
        sc.keep('AA', AA)
        U, S, V = yastn.linalg.svd_with_truncation(AA, axes=((0, 1), (2, 3)), D_total=64)
        sc.keep('U', U); sc.keep('S', S); sc.keep('V', V)
        U2, S2, V2 = yastn.linalg.svd_with_truncation(AA, axes=((0, 1), (2, 3)), D_total=64, fix_signs=True)
        sc.keep('U2', U2); sc.keep('S2', S2); sc.keep('V2', V2)
        sc.dump()

    # ---------------- U1xU1 complex, uneven sectors (small) ---------------------------------
    rec = Recorder()
    sc = Scenario('u1xu1_c128', rec)
    c = cfg('U1xU1', rec)
    tv, Dv = [], []
    for nu in range(-2, 3):
        for nd in range(-2, 3):
            w = np.exp(-(nu * nu + nd * nd) / (2 * 1.5 ** 2))
            d = int(round(9 * w))
            if d > 0:
                tv.append((nu, nd)); Dv.append(d)
    tp = ((0, 0), (1, 0), (0, 1), (1, 1))
    legs = [yastn.Leg(c, s=1, t=tv, D=Dv), yastn.Leg(c, s=1, t=tp, D=(1, 1, 1, 1)), yastn.Leg(c, s=-1, t=tv, D=Dv)]
    A = tensor(c, legs, (0, 0), 'complex128', 3)
    sc.keep('A', A)
    T = yastn.tensordot(A, A.conj(), axes=((0, 1), (0, 1)))
    sc.keep('T', T)
    B = tensor(c, legs, (0, 0), 'float64', 4)   # mixed dtype: f64 x c128 promotes
    sc.keep('B', B)
    TB = yastn.tensordot(B, A, axes=((2,), (0,)))
    sc.keep('TB', TB)
    U, S, V = yastn.linalg.svd_with_truncation(TB, axes=((0, 1), (2, 3)), D_total=40, fix_signs=True)
    sc.keep('U', U); sc.keep('S', S); sc.keep('V', V)
    sc.dump({'legs': {'tv': tv, 'Dv': Dv}})

    # ---------------- fusion / unfusion / transposed operands (U1, rank 4) -------------------
    for policy in ('fuse_contracted', 'fuse_to_matrix'):
        rec = Recorder()
        sc = Scenario(f'fuse_u1_{policy}', rec)
        c = cfg('U1', rec, policy)
        la = [yastn.Leg(c, s=1, t=(-1, 0, 1), D=(2, 3, 2)), yastn.Leg(c, s=1, t=(0, 1), D=(2, 3)),
              yastn.Leg(c, s=-1, t=(-1, 0, 1, 2), D=(1, 2, 3, 1)), yastn.Leg(c, s=-1, t=(0, 1), D=(3, 1))]
        a = tensor(c, la, 1, 'float64', 5)
        sc.keep('a', a)
        f = a.fuse_legs(axes=(0, (1, 2), 3), mode='hard')
        sc.keep('f', f)
        g = f.fuse_legs(axes=((0, 2), 1), mode='hard')
        sc.keep('g', g)
        u = g.unfuse_legs(axes=0)
        sc.keep('u', u)
        uu = u.unfuse_legs(axes=2)
        sc.keep('uu', uu)
        lb = [la[2].conj(), la[0].conj(), yastn.Leg(c, s=1, t=(0, 1, 2), D=(2, 1, 2))]
        b = tensor(c, lb, 0, 'float64', 6)
        sc.keep('b', b)
        at = a.transpose(axes=(3, 2, 1, 0))
        r1 = yastn.tensordot(at, b, axes=((1, 3), (0, 1)))
        sc.keep('r1', r1)
        r2 = yastn.tensordot(b, a, axes=((0, 1), (2, 0)))
        sc.keep('r2', r2)
        r3 = yastn.tensordot(f, f.conj(), axes=((0, 2), (0, 2)))
        sc.keep('r3', r3)
        ct = r2.transpose(axes=(2, 0, 1))
        ct = ct.consume_transpose() if hasattr(ct, 'consume_transpose') else ct
        sc.keep('ct', ct)
        sc.dump()

    # ---------------- truncation known answers (reference tests/tensor/test_svd.py:189-266) ----
    rec = Recorder()
    sc = Scenario('truncation_u1', rec)
    c = cfg('U1', rec)
    legs = [yastn.Leg(c, s=1, t=(0, 1), D=(5, 6)), yastn.Leg(c, s=1, t=(-1, 0), D=(5, 6)),
            yastn.Leg(c, s=-1, t=(-1, 0, 1), D=(2, 3, 4)), yastn.Leg(c, s=-1, t=(-1, 0, 1), D=(2, 3, 4))]
    a = tensor(c, legs, 1, 'float64', 7)
    U, S, V = yastn.linalg.svd(a, axes=((0, 1), (2, 3)), sU=-1)
    S.set_block(ts=(-2, -2), Ds=4, val=[2 ** (-ii - 6) for ii in range(4)])
    S.set_block(ts=(-1, -1), Ds=12, val=[2 ** (-ii - 2) for ii in range(12)])
    S.set_block(ts=(0, 0), Ds=25, val=[2 ** (-ii - 1) for ii in range(25)])
    a = yastn.ncon([U, S, V], [(-1, -2, 1), (1, 2), (2, -3, -4)])
    sc.keep('a', a)
    rec.calls.clear(); rec.arrays = {k: v for k, v in rec.arrays.items() if k.startswith('t_')}
    cases = [
        ({'tol': 0.01, 'D_block': 100, 'D_total': 12}, -1, True),
        ({'tol': 0.01, 'D_block': 100, 'D_total': 12}, 1, True),
        ({'tol': 0.01, 'D_block': 100, 'D_total': 12}, -1, False),
        ({'tol': 0.02, 'D_block': 5, 'D_total': 100}, 1, True),
        ({'tol': 0, 'tol_block': 0.2, 'D_total': 100}, 1, True),
        ({'D_total': 10, 'truncate_multiplets': True}, -1, True),
        ({'D_total': 0}, -1, False),
    ]
    results = []
    for k, (opts, sU, nU) in enumerate(cases):
        U1, S1, V1 = yastn.linalg.svd_with_truncation(a, axes=((0, 1), (2, 3)), sU=sU, nU=nU, **opts)
        sc.keep(f'S_{k}', S1); sc.keep(f'U_{k}', U1); sc.keep(f'V_{k}', V1)
        results.append({'opts': _jsonable(opts), 'sU': sU, 'nU': nU, 'shape': list(S1.get_shape()),
                        'charges': _jsonable(S1.get_blocks_charge())})
    # per-charge dict options are not JSON keys; record separately
    opts = {'D_block': {(0,): 2, (-1,): 3}, 'tol_block': {(0,): 0.6, (-1,): 0.1}}
    _, S1, _ = yastn.linalg.svd_with_truncation(a, axes=((0, 1), (2, 3)), sU=-1, **opts)
    results.append({'opts': 'dict_D_block_tol_block', 'sU': -1, 'nU': True, 'shape': list(S1.get_shape()),
                    'charges': _jsonable(S1.get_blocks_charge())})
    sc.dump({'cases': results})

    # ---------------- eigh (hermitian, U1, complex) ------------------------------------------
    rec = Recorder()
    sc = Scenario('eigh_u1_c128', rec)
    c = cfg('U1', rec)
    l4 = [yastn.Leg(c, s=1, t=(-1, 0, 1), D=(2, 4, 3)), yastn.Leg(c, s=1, t=(0, 1), D=(2, 2)),
          yastn.Leg(c, s=-1, t=(-1, 0, 1), D=(2, 4, 3)), yastn.Leg(c, s=-1, t=(0, 1), D=(2, 2))]
    m = tensor(c, l4, 0, 'complex128', 9)
    rec.on = False
    mh = m + m.transpose(axes=(2, 3, 0, 1)).conj()
    mh = mh.consume_transpose() if hasattr(mh, 'consume_transpose') else mh
    rec.on = True
    sc.keep('mh', mh)
    Sd, Ud = yastn.linalg.eigh(mh, axes=((0, 1), (2, 3)))
    sc.keep('S', Sd); sc.keep('U', Ud)
    sc.dump()

    # ---------------- fix_svd_signs known answer (reference tests/tensor/test_svd.py:159-186) -
    rec = Recorder()
    sc = Scenario('fix_signs_kat', rec)
    c = cfg('U1', rec)
    U = yastn.Tensor(config=c, s=(-1, 1), dtype='complex128')
    S = yastn.Tensor(config=c, s=(-1, 1), isdiag=True)
    V = yastn.Tensor(config=c, s=(-1, 1), dtype='complex128')
    is2 = 0.5 ** 0.5
    U.set_block(ts=(0, 0), Ds=(3, 2), val=[[-is2, -1 / 2], [is2, -1 / 2], [0, is2 * 1j]])
    U.set_block(ts=(1, 1), Ds=(2, 2), val=[[0, 1j], [1, 0]])
    S.set_block(ts=(0, 0), Ds=(2, 2), val=[0.5, 0.2])
    S.set_block(ts=(1, 1), Ds=(2, 2), val=[0.3, 0.2])
    V.set_block(ts=(0, 0), Ds=(2, 2), val=[[0, 1], [1, 0]])
    V.set_block(ts=(1, 1), Ds=(2, 3), val=[[1, 0, 0], [0, 0.8j, 0.6]])
    rec.on = False
    USV = U @ S @ V
    rec.on = True
    sc.keep('USV', USV)
    nU, nS, nV = yastn.linalg.svd(USV, axes=(0, 1), fix_signs=True)
    sc.keep('U', nU); sc.keep('S', nS); sc.keep('V', nV)
    sc.dump()

    # ---------------- reference torch backend gradients (CPU) ---------------------------------
    import torch
    from yastn.backend import backend_torch
    torch.set_default_dtype(torch.float64)
    recg = Recorder()
    sc = Scenario('torch_grads', recg)
    grads = {}
    ct = yastn.make_config(sym='U1', backend=make_proxy(backend_torch, recg))
    recg.on = False

    def tt(config, legs, n, dtype, seed):
        a = yastn.zeros(config=config, legs=legs, n=n, dtype=dtype)
        a._data = torch.as_tensor(seeded(a.size, dtype, seed))
        return a

    legs = [yastn.Leg(ct, s=1, t=tq, D=Dq), yastn.Leg(ct, s=1, t=(0, 1), D=(1, 1)), yastn.Leg(ct, s=-1, t=tq, D=Dq)]
    for dtype, seed in (('float64', 10), ('complex128', 11)):
        A = tt(ct, legs, 0, dtype, seed)
        A.requires_grad_(True)
        # svd adjoint through a merge, an svd with truncation, unmerges and contractions
        AA = A @ A
        U, S, V = yastn.linalg.svd_with_truncation(AA, axes=((0, 1), (2, 3)), D_total=40)
        X = yastn.tensordot(U, V, axes=(2, 0))
        Z = yastn.tensordot(X, AA.conj(), axes=((0, 1, 2, 3), (0, 1, 2, 3)))
        l = Z.to_number() + S.norm() ** 2
        l = l.real if torch.is_tensor(l) and l.is_complex() else l
        l.backward()
        grads[dtype] = {'loss': float(l.detach())}
        sc.rec.add_array(f'A_{dtype}', A.data.detach().numpy())
        sc.rec.add_array(f'gA_{dtype}', A.data.grad.numpy() if A.data.grad is not None else A.grad().data.numpy())
        sc.keep(f'A_{dtype}', A.detach())
    # isolated per-block svd adjoint (through the reference kernel_svd with a one-block meta)
    rng = np.random.default_rng(12)
    blocks = []
    for bi, (m, n, dtype) in enumerate(((6, 6, 'float64'), (7, 4, 'float64'), (3, 8, 'float64'),
                                        (5, 5, 'complex128'), (6, 3, 'complex128'), (2, 7, 'complex128'))):
        cplx = dtype == 'complex128'
        a = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
        k = min(m, n)
        meta = (((0, m * n), (m, n), (0, m * k), (m, k), (0, k), (0, k * n), (k, n)),)
        ta = torch.as_tensor(a.ravel()).requires_grad_(True)
        u, s, vh = backend_torch.svd(ta, meta, (m * k, k, k * n))
        if cplx:
            # a loss that is invariant under the column phases of u (paired with rows of vh)
            w = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
            tw = torch.as_tensor(w)
            uu, vv = u.view(m, k), vh.view(k, n)
            lf = ((uu * s.unsqueeze(0)) @ vv @ tw).trace().real + ((uu @ vv) @ tw).trace().imag + (s ** 3).sum()
            lf.backward()
            blocks.append({'m': m, 'n': n, 'dtype': dtype, 'loss': 'uv_w'})
            recg.add_array(f'svdb{bi}_a', a); recg.add_array(f'svdb{bi}_w', w)
            recg.add_array(f'svdb{bi}_g', ta.grad.numpy().reshape(m, n))
        else:
            gu = rng.standard_normal((m, k)); gs = rng.standard_normal(k); gv = rng.standard_normal((k, n))
            torch.autograd.backward([u, s, vh], [torch.as_tensor(gu.ravel()), torch.as_tensor(gs),
                                                 torch.as_tensor(gv.ravel())])
            blocks.append({'m': m, 'n': n, 'dtype': dtype, 'loss': 'linear'})
            recg.add_array(f'svdb{bi}_a', a); recg.add_array(f'svdb{bi}_gu', gu)
            recg.add_array(f'svdb{bi}_gs', gs); recg.add_array(f'svdb{bi}_gv', gv)
            recg.add_array(f'svdb{bi}_u', u.detach().numpy().reshape(m, k))
            recg.add_array(f'svdb{bi}_s', s.detach().numpy())
            recg.add_array(f'svdb{bi}_vh', vh.detach().numpy().reshape(k, n))
            recg.add_array(f'svdb{bi}_g', ta.grad.numpy().reshape(m, n))
    sc.dump({'grads': grads, 'svd_blocks': blocks})

    # ---------------- the backend attribute surface used by the reference Tensor layer -----------
    import re
    names = set()
    for root, _, files in os.walk(os.path.join(ref_root, 'yastn')):
        for fn in files:
            if fn.endswith('.py'):
                with open(os.path.join(root, fn)) as f:
                    names.update(re.findall(r'backend\.(\w+)', f.read()))
    with open(os.path.join(HERE, 'backend_surface.json'), 'w') as f:
        json.dump(sorted(names), f, indent=0)
    print('backend surface', len(names), 'names')


def more(ref_root, only=None):
    """Scenarios added later (each writes its own files; the ones above are left as they are):
    the no-fusion tensordot policy (transpose_dot_sum with the planner's list-typed metas), dot_diag and
    embed_mask calls, truncation_mask_multiplets decisions, and the Kagome RVB state of the reference's
    tests/peps/test_ad.py with CTMRG losses and gradients of the reference torch backend on the CPU."""
    _stub_opt_einsum()
    if ref_root not in sys.path:
        sys.path.insert(0, ref_root)
    import yastn
    from yastn.backend import backend_np

    def want(name):
        return only is None or name in only

    def cfg(sym, rec, policy='fuse_contracted'):
        return yastn.make_config(sym=sym, backend=make_proxy(backend_np, rec), tensordot_policy=policy)

    def tensor(config, legs, n, dtype, seed):
        a = yastn.zeros(config=config, legs=legs, n=n, dtype=dtype)
        a._data = seeded(a.size, dtype, seed).astype(dtype)
        return a

    tq = tuple(range(-4, 5))
    Dq = (2, 5, 10, 14, 16, 14, 10, 5, 2)

    def rank4_legs(c):
        la = [yastn.Leg(c, s=1, t=(-1, 0, 1), D=(2, 3, 2)), yastn.Leg(c, s=1, t=(0, 1), D=(2, 3)),
              yastn.Leg(c, s=-1, t=(-1, 0, 1, 2), D=(1, 2, 3, 1)), yastn.Leg(c, s=-1, t=(0, 1), D=(3, 1))]
        lb = [la[2].conj(), la[0].conj(), yastn.Leg(c, s=1, t=(0, 1, 2), D=(2, 1, 2))]
        return la, lb

    # ---------------- no-fusion tensordot: transpose_dot_sum with list-typed metas ---------------
    if want('nofusion_u1'):
        rec = Recorder()
        sc = Scenario('nofusion_u1', rec)
        c = cfg('U1', rec, 'no_fusion')
        legs = [yastn.Leg(c, s=1, t=tq, D=Dq), yastn.Leg(c, s=1, t=(0, 1), D=(1, 1)),
                yastn.Leg(c, s=-1, t=tq, D=Dq)]
        A = tensor(c, legs, 0, 'float64', 0)
        sc.keep('A', A)
        sc.keep('T', yastn.tensordot(A, A.conj(), axes=((0, 1), (0, 1))))
        la, lb = rank4_legs(c)
        a = tensor(c, la, 1, 'float64', 5)
        b = tensor(c, lb, 0, 'float64', 6)
        ac = tensor(c, la, 1, 'complex128', 13)
        sc.keep('a', a); sc.keep('b', b); sc.keep('ac', ac)
        sc.keep('r1', yastn.tensordot(a.transpose(axes=(3, 2, 1, 0)), b, axes=((1, 3), (0, 1))))
        sc.keep('r2', yastn.tensordot(b, a, axes=((0, 1), (2, 0))))
        sc.keep('r4', yastn.tensordot(b, ac, axes=((0, 1), (2, 0))))
        sc.keep('r5', yastn.tensordot(ac, ac.conj(), axes=((0, 1), (0, 1))))
        sc.dump()

    # ---------------- dot_diag and embed_mask calls ----------------------------------------------
    if want('diag_embed_u1'):
        rec = Recorder()
        sc = Scenario('diag_embed_u1', rec)
        c = cfg('U1', rec)
        la, _ = rank4_legs(c)
        rec.on = False
        a = tensor(c, la, 1, 'float64', 7)
        U, S, V = yastn.linalg.svd(a, axes=((0, 1), (2, 3)), sU=-1)
        ac = tensor(c, la, 1, 'complex128', 8)
        Uc, Sc, Vc = yastn.linalg.svd(ac, axes=((0, 1), (2, 3)), sU=1)
        rec.on = True
        for lab, t in (('U', U), ('S', S), ('V', V), ('Uc', Uc), ('Sc', Sc), ('Vc', Vc)):
            sc.keep(lab, t)
        sc.keep('US', U @ S)
        sc.keep('SV', S @ V)
        sc.keep('rsV', S.rsqrt() @ V)
        sc.keep('UcSc', Uc @ Sc)
        sc.keep('ScVc', Sc @ Vc)
        # embed_mask: flip_charges (mask {0: slice(None)}) and addition of tensors whose hard-fused legs
        # differ (_embed_tensor)
        sc.keep('a', a)
        sc.keep('flip', a.flip_charges(axes=1))
        l0 = [yastn.Leg(c, s=1, t=(-1, 0, 1), D=(2, 3, 2)), yastn.Leg(c, s=1, t=(0, 1), D=(2, 3)),
              yastn.Leg(c, s=-1, t=(0, 1), D=(2, 2))]
        l1 = [yastn.Leg(c, s=1, t=(0, 1), D=(3, 2)), yastn.Leg(c, s=1, t=(0, 1), D=(2, 3)),
              yastn.Leg(c, s=-1, t=(0, 1), D=(2, 2))]
        x = tensor(c, l0, 0, 'float64', 14).fuse_legs(axes=((0, 1), 2), mode='hard')
        y = tensor(c, l1, 0, 'float64', 15).fuse_legs(axes=((0, 1), 2), mode='hard')
        sc.keep('x', x); sc.keep('y', y)
        sc.keep('xy', x + y)
        sc.dump()

    # ---------------- truncation_mask_multiplets (yastn/tensor/linalg.py:539-634) ------------------
    if want('multiplets_u1'):
        rec = Recorder()
        sc = Scenario('multiplets_u1', rec)
        c = cfg('U1', rec)
        rec.on = False
        legs = [yastn.Leg(c, s=1, t=(-1, 0, 1), D=(2, 3, 2)), yastn.Leg(c, s=1, t=(-1, 0, 1), D=(3, 4, 3)),
                yastn.Leg(c, s=-1, t=(-1, 0, 1), D=(4, 5, 4)), yastn.Leg(c, s=-1, t=(-1, 0, 1), D=(5, 6, 5))]
        a = tensor(c, legs, 0, 'float64', 16)
        _, S, _ = yastn.linalg.svd(a, axes=((0, 1), (2, 3)))
        # the spectrum of the reference's test_svd_multiplets (tests/tensor/test_svd.py:344-388)
        v00 = [1, 1, 0.1001, 0.1000, 0.1000, 0.0999, 0.001001, 0.001000] + [0] * 16
        v11 = [1, 1, 0.1001, 0.1000, 0.0999, 0.001000, 0.000999] + [0] * 10
        v22 = [1, 1, 0.1001, 0.1000, 0.001000, 0]
        S.set_block(ts=(0, 0), Ds=24, val=v00)
        S.set_block(ts=(1, 1), Ds=17, val=v11)
        S.set_block(ts=(-1, -1), Ds=17, val=v11)
        S.set_block(ts=(2, 2), Ds=6, val=v22)
        S.set_block(ts=(-2, -2), Ds=6, val=v22)
        # a second spectrum: exact multiplets of sizes 1-3 spread over the blocks, and a random tail
        rng = np.random.default_rng(17)
        S2 = S.copy()
        vals = {}
        levels = np.sort(rng.uniform(0.01, 1, 12))[::-1]
        for t in S2.struct.t:
            D = S2[t].shape[0] if hasattr(S2[t], 'shape') else len(S2[t])
            v = list(levels[rng.integers(0, 12, size=D)])
            v = sorted(v, reverse=True)
            vals[t] = v
            S2.set_block(ts=t, Ds=D, val=v)
        cases = [(0, {'tol': 1e-4, 'D_total': 30, 'eps_multiplet': 1e-3}),
                 (0, {'tol': 1e-4, 'D_total': 20, 'eps_multiplet': 1e-3}),
                 (0, {'tol': 1e-4, 'D_total': 12, 'eps_multiplet': 1e-3}),
                 (0, {'tol': 0.05, 'D_total': 64, 'eps_multiplet': 1e-3}),
                 (0, {'tol': 1e-4, 'D_total': 30, 'eps_multiplet': 1e-2}),
                 (0, {'tol': 0.0, 'D_total': 1000, 'eps_multiplet': 1e-8}),
                 (0, {'tol': 1e-8, 'D_total': 25, 'eps_multiplet': 1e-8, 'hermitian': True}),
                 (1, {'tol': 1e-8, 'D_total': 17, 'eps_multiplet': 1e-8}),
                 (1, {'tol': 1e-8, 'D_total': 40, 'eps_multiplet': 1e-8}),
                 (1, {'tol': 0.3, 'D_total': 40, 'eps_multiplet': 1e-8}),
                 (1, {'tol': 1e-8, 'D_total': 9, 'eps_multiplet': 1e-8, 'hermitian': True})]
        sc.keep('S0', S)
        sc.keep('S1', S2)
        res = []
        for k, (which, opts) in enumerate(cases):
            m = yastn.linalg.truncation_mask_multiplets((S, S2)[which], **opts)
            sc.keep(f'mask_{k}', m)
            res.append({'S': which, 'opts': _jsonable(opts)})
        sc.dump({'cases': res})

    # ---------------- qr (yastn/tensor/linalg.py:738-830): metas for both signatures and charges -------
    if want('qr_u1'):
        rec = Recorder()
        sc = Scenario('qr_u1', rec)
        c = cfg('U1', rec)
        la, _ = rank4_legs(c)
        a = tensor(c, la, 1, 'float64', 18)
        ac = tensor(c, la, 0, 'complex128', 19)
        sc.keep('a', a); sc.keep('ac', ac)
        for k, (t, axes, sQ, Qaxis, Raxis) in enumerate(((a, ((0, 1), (2, 3)), 1, -1, 0),
                                                         (a, ((2, 0), (3, 1)), -1, 0, -1),
                                                         (ac, ((0, 3), (1, 2)), 1, 1, 1),
                                                         (ac, ((1, 2, 3), 0), -1, -1, 0))):
            Q, R = yastn.linalg.qr(t, axes=axes, sQ=sQ, Qaxis=Qaxis, Raxis=Raxis)
            sc.keep(f'Q{k}', Q); sc.keep(f'R{k}', R)
        sc.dump({'cases': [{'a': 'a', 'axes': [[0, 1], [2, 3]], 'sQ': 1, 'Qaxis': -1, 'Raxis': 0},
                           {'a': 'a', 'axes': [[2, 0], [3, 1]], 'sQ': -1, 'Qaxis': 0, 'Raxis': -1},
                           {'a': 'ac', 'axes': [[0, 3], [1, 2]], 'sQ': 1, 'Qaxis': 1, 'Raxis': 1},
                           {'a': 'ac', 'axes': [[1, 2, 3], [0]], 'sQ': -1, 'Qaxis': -1, 'Raxis': 0}]})

    # ---------------- Kagome RVB D=3 U(1) state and CTMRG gradients (tests/peps/test_ad.py) -------
    if want('rvb_kagome'):
        import torch
        sys.path.insert(0, os.path.join(ref_root, 'tests', 'peps'))
        import test_ad as T
        torch.set_default_dtype(torch.float64)
        kw = {'backend': 'torch', 'default_device': 'cpu'}
        A, Ag, cost = T.prepare_RVB((dict(kw), torch, None))
        rec = Recorder()
        sc = Scenario('rvb_kagome', rec)
        sc.keep('A', A.detach())
        rec.add_array('grad_expected', Ag.data.detach().numpy())
        runs = []
        for sweeps, init in ((1, 'eye'), (1, 'dl'), (2, 'dl'), (20, 'dl')):
            x = A.data.detach().clone().requires_grad_(True)
            loss = cost(x, slice(0, x.numel()), sweeps, ctm_init=init, fix_signs=False,
                        truncate_multiplets_mode='truncate', checkpoint_move=False)
            loss.backward()
            tag = f's{sweeps}_{init}'
            rec.add_array(f'grad_{tag}', x.grad.detach().numpy())
            runs.append({'sweeps': sweeps, 'init': init, 'loss': float(loss.detach()), 'tag': tag})
            print('rvb', tag, float(loss.detach()), flush=True)
        sc.dump({'runs': runs, 'leg_order': 't,l,b,r,p (after transpose (1, 2, 3, 4, 0))'})

    # ---------------- the same state with QR-reduced projectors (use_qr=True, the reference default) ---
    if want('rvb_kagome_qr'):
        import torch
        import yastn.tn.fpeps as fpeps
        from yastn.tn.fpeps.envs.rdm import rdm1x1
        sys.path.insert(0, os.path.join(ref_root, 'tests', 'peps'))
        import test_ad as T
        torch.set_default_dtype(torch.float64)
        kw = {'backend': 'torch', 'default_device': 'cpu'}
        A, _, _ = T.prepare_RVB((dict(kw), torch, None))

        def truncation_f(S):
            return yastn.linalg.truncation_mask_multiplets(S, keep_multiplets=True, D_total=64, tol=1.0e-8,
                                                           tol_block=0.0, eps_multiplet=1.0e-8)

        def cost(x, sweeps, init):
            A0 = A.clone()
            A0._data = x
            g = fpeps.SquareLattice(dims=(1, 1), boundary='infinite')
            psi = fpeps.Peps(g, tensors=dict(zip(g.sites(), [A0])))
            env = fpeps.EnvCTM(psi, init=init)
            env.ctmrg_(opts_svd={'D_total': 64, 'fix_signs': False}, max_sweeps=sweeps, truncation_f=truncation_f,
                       use_qr=True, checkpoint_move=False)
            r1x1, _ = rdm1x1((0, 0), psi, env)
            return r1x1[(-1, -1)].trace().real

        rec = Recorder()
        sc = Scenario('rvb_kagome_qr', rec)
        runs = []
        for sweeps, init in ((1, 'dl'), (2, 'dl')):
            x = A.data.detach().clone().requires_grad_(True)
            loss = cost(x, sweeps, init)
            loss.backward()
            tag = f's{sweeps}_{init}'
            rec.add_array(f'grad_{tag}', x.grad.detach().numpy())
            runs.append({'sweeps': sweeps, 'init': init, 'loss': float(loss.detach()), 'tag': tag,
                         'finite': bool(np.isfinite(x.grad.detach().numpy()).all())})
            print('rvb qr', tag, runs[-1], flush=True)
        sc.dump({'runs': runs})


if __name__ == '__main__':
    args = [a for a in sys.argv[1:] if not a.startswith('--')]
    root = args[0] if args else '/root/reference'
    if '--more' in sys.argv:
        more(root, set(args[1:]) or None)
    else:
        main(root)
        more(root)
