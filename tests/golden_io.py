"""Load the golden vectors written by tests/golden/make_golden.py (JSON meta + NPZ arrays)."""
import functools
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
SCENARIOS = ('mps_u1_fuse_contracted', 'mps_u1_fuse_to_matrix', 'u1xu1_c128', 'fuse_u1_fuse_contracted',
             'fuse_u1_fuse_to_matrix', 'truncation_u1', 'eigh_u1_c128', 'fix_signs_kat')


def tup(x):
    """JSON lists back to the nested tuples the reference planner emits."""
    if isinstance(x, list):
        return tuple(tup(y) for y in x)
    return x


@functools.lru_cache(maxsize=None)
def load(name):
    with open(os.path.join(GOLDEN, f'{name}.json')) as f:
        js = json.load(f)
    arrays = dict(np.load(os.path.join(GOLDEN, f'{name}.npz')))
    return js, arrays


def calls(name, fn=None):
    """Decoded recorded backend calls: list of (fn, args, kwargs, outputs)."""
    js, arrays = load(name)
    out = []
    for c in js['calls']:
        if fn is not None and c['fn'] != fn:
            continue
        args = []
        for a in c['args']:
            if 'array' in a:
                args.append(arrays[a['array']])
            elif 'mask' in a:
                args.append({tup(k): slice(*v['slice']) if isinstance(v, dict) else np.asarray(v, dtype=np.int64)
                             for k, v in a['mask']})
            else:
                args.append(tup(a['value']))
        if c['fn'] == 'transpose_dot_sum':
            # the no-fusion planner's meta_dot is a list of (slice, (Dl, Dr), [[ia, ib], ...])
            # (yastn/tensor/_contractions.py:444-446); restore that typing
            args[2] = [(sl, D, [list(p) for p in grp]) for sl, D, grp in args[2]]
        outs = [arrays[k] for k in c['out']]
        out.append((c['fn'], args, c['kwargs'], outs))
    return out


def tensors(name):
    js, arrays = load(name)
    res = {}
    for label, d in js['tensors'].items():
        d = dict(d)
        d['t'] = tup(d['t'])
        d['D'] = tup(d['D'])
        d['data'] = arrays[d['data']]
        res[label] = d
    return res
