"""The C-ABI library loads without a GPU and exports every entry point include/yastn_hip.h declares."""
import ctypes
import os
import re

import numpy as np

from yastn_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, 'include', 'yastn_hip.h')) as f:
        text = f.read()
    return sorted(set(re.findall(r'\b(yastn_hip_\w+)\s*\(', text)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.exported_symbols())


def test_abi_version_and_constants():
    assert L.lib().yastn_hip_abi_version() == L.YH_ABI_VERSION
    with open(os.path.join(ROOT, 'include', 'yastn_hip.h')) as f:
        text = f.read()
    for name in ('YH_MAXR', 'YH_JOB_WORDS', 'YH_GATHER_TILE', 'YH_TRANSPOSE_TILE', 'YH_GEMM_WORDS',
                 'YH_TERM_WORDS', 'YH_SVD_WORDS', 'YH_SVD_MAX_Q', 'YH_SVD_LDS_BYTES', 'YH_ABI_VERSION'):
        m = re.search(rf'#define {name} (\d+)', text)
        assert m and int(m.group(1)) == getattr(L, name), name


def test_plan_checker_rejects_out_of_bounds():
    job = np.zeros((1, L.YH_JOB_WORDS), dtype=np.int64)
    job[0, :6] = (0, 0, 1, 0, 0, 0)
    M = L.YH_MAXR
    job[0, 6:6 + M] = 1
    job[0, 6] = 10
    job[0, 6 + M] = 1
    job[0, 6 + 2 * M] = 1
    tiles = np.zeros(1, dtype=np.int32)
    lib = L.lib()
    ok = lib.yastn_hip_check_gather_plan(job.ctypes.data, 1, tiles.ctypes.data, 1, 10, 10)
    assert ok == L.YH_OK
    bad = lib.yastn_hip_check_gather_plan(job.ctypes.data, 1, tiles.ctypes.data, 1, 9, 10)
    assert bad == L.YH_ERR_ARG
    assert b'out of bounds' in lib.yastn_hip_last_error()


def test_backend_module_surface():
    """backend_hip exports every name the reference Tensor layer looks up on config.backend."""
    import json
    from yastn_amd import backend_hip
    with open(os.path.join(ROOT, 'tests', 'golden', 'backend_surface.json')) as f:
        names = json.load(f)
    not_backend = {'backend_np', 'backend_torch', 'kernel_tensordot_bs'}  # module refs / torch_cpp only
    missing = [n for n in names if n not in not_backend and not hasattr(backend_hip, n)]
    assert not missing, missing
    assert 'torch' in backend_hip.BACKEND_ID and backend_hip.BACKEND_ID != 'torch_cpp'
