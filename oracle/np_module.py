"""numpy backend module (oracle) for the restated tensor layer (yastn_amd.blocksparse): CPU checks and
the CPU baseline leg of bench.py.

Its hot-path functions are the oracle's (oracle/np_backend.py); the rest are numpy one-liners.
Test infrastructure only (oracle package).
"""
import numpy as np

from oracle import np_backend as O

BACKEND_ID = 'np_test'
DTYPE = {'float64': np.float64, 'complex128': np.complex128, 'bool': bool}

transpose_and_merge = O.transpose_and_merge
unmerge = O.unmerge
dot = O.dot
svd = lambda data, meta, sizes, **kw: O.svd(data, meta, sizes)
fix_svd_signs = O.fix_svd_signs
eigh = lambda data, meta, sizes: O.eigh(data, meta, sizes)
apply_mask = O.apply_mask
embed_mask = O.embed_mask
transpose = O.transpose
transpose_dot_sum = O.transpose_dot_sum
dot_diag = O.dot_diag
qr = O.qr


def zeros(D, dtype='float64', device='cpu'):
    return np.zeros(D, dtype=DTYPE[dtype])


def to_tensor(val, Ds=None, dtype='float64', device='cpu'):
    x = np.array(val, dtype=DTYPE[dtype])
    return x if Ds is None else x.reshape(Ds)


def to_numpy(x):
    return np.array(x)


def get_size(x):
    return x.size


def get_yastn_dtype(x):
    return x.dtype.name


def conj(x):
    return x.conj()


def detach(x):
    return x


def sum_elements(x):
    return x.sum().reshape(1)


def max_abs(x):
    return np.abs(x).max() if x.size else 0


def argsort(x):
    return np.argsort(x)


def to_mask(x):
    return x.nonzero()[0]


def rsqrt(x, cutoff=0):
    out = np.zeros_like(x)
    k = np.abs(x) > cutoff
    out[k] = 1 / np.sqrt(x[k])
    return out


def sqrt(x):
    return np.sqrt(x)


def norm(x, p):
    """backend_np.py:154-158: 'fro' is the Frobenius norm, 'inf' the largest modulus."""
    if p == 'fro':
        return np.linalg.norm(x)
    return np.max(np.abs(x)) if x.size > 0 else np.float64(0.)


def vdot(a, b, meta):
    return sum(np.dot(a[x[0]:x[1]], b[y[0]:y[1]]) for x, y in meta)


def maximum(x, y):
    return np.maximum(x, y)


def absolute(x):
    return np.abs(x)
