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
    for name in ('YH_MAXR', 'YH_JOB_WORDS', 'YH_JOB_BUFFER_WORD', 'YH_GATHER_TILE', 'YH_TRANSPOSE_TILE', 'YH_GEMM_WORDS',
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


def test_pair_plan_buffer_word_checked():
    """A job's buffer word selects the (source, destination) sizes it is checked against; the
    single-pair checker accepts buffer 0 only, the paired one 0 and 1."""
    job = np.zeros((1, L.YH_JOB_WORDS), dtype=np.int64)
    job[0, :6] = (0, 0, 1, 0, 0, 0)
    M = L.YH_MAXR
    job[0, 6:6 + M] = 1
    job[0, 6] = 10
    job[0, 6 + M] = 1
    job[0, 6 + 2 * M] = 1
    tiles = np.zeros(1, dtype=np.int32)
    lib = L.lib()
    job[0, L.YH_JOB_BUFFER_WORD] = 1
    assert lib.yastn_hip_check_gather_plan(job.ctypes.data, 1, tiles.ctypes.data, 1, 10, 10) == L.YH_ERR_ARG
    assert lib.yastn_hip_check_gather_plan2(job.ctypes.data, 1, tiles.ctypes.data, 1, 5, 5, 10, 10) == L.YH_OK
    assert lib.yastn_hip_check_gather_plan2(job.ctypes.data, 1, tiles.ctypes.data, 1, 10, 10, 9, 10) == L.YH_ERR_ARG
    job[0, L.YH_JOB_BUFFER_WORD] = 2
    assert lib.yastn_hip_check_gather_plan2(job.ctypes.data, 1, tiles.ctypes.data, 1, 10, 10, 10, 10) == L.YH_ERR_ARG


def test_pair_plan_concatenates_jobs():
    """GatherPairPlan: the second plan's jobs follow with buffer word 1 and shifted tiles."""
    import torch
    from yastn_amd import plans as P
    dev = torch.device('cpu')
    pa = P.GatherPlan([(0, 0, (3, 4), (4, 1), (4, 1)), (-1, 12, (5,), (0,), (1,))], 12, 17, dev)
    pb = P.GatherPlan([(0, 0, (2, 3), (1, 2), (3, 1))], 6, 6, dev)
    pp = P.GatherPairPlan(pa, pb, dev)
    assert pp.njobs == pa.njobs + pb.njobs and pp.ntiles == pa.ntiles + pb.ntiles
    assert (pp.h_jobs[:pa.njobs, L.YH_JOB_BUFFER_WORD] == 0).all()
    assert (pp.h_jobs[pa.njobs:, L.YH_JOB_BUFFER_WORD] == 1).all()
    assert (pp.h_jobs[pa.njobs:, 5] == pb.h_jobs[:, 5] + pa.ntiles).all()
    assert (pp.h_tile_job[pa.ntiles:] == pb.h_tile_job + pa.njobs).all()
    assert P.pair_plan(pa, pb, dev) is P.pair_plan(pa, pb, dev)


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
