"""ctypes binding of the C-ABI library ``libyastn_hip.so`` (declared in include/yastn_hip.h).

The library must exist (built by ``__graft_entry__.build()`` / ``python -m yastn_amd.build``);
there is no fallback: a missing library or a non-zero status raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get('YASTN_HIP_LIB', os.path.join(HERE, 'libyastn_hip.so'))

# constants mirrored from include/yastn_hip.h
YH_OK, YH_ERR_ARG, YH_ERR_HIP, YH_ERR_NUMERIC = 0, -1, -2, -3
YH_F64, YH_C128 = 0, 1
YH_ABI_VERSION = 13
YH_MAXR = 8
YH_JOB_WORDS = 32
YH_JOB_BUFFER_WORD = 30
YH_GATHER_TILE = 2048
YH_TRANSPOSE_TILE = 32
YH_GEMM_WORDS = 8
YH_TERM_WORDS = 8
YH_MAX_BUFS = 4
YH_TILE_32x32, YH_TILE_64x64, YH_TILE_128x64, YH_TILE_16x64, YH_TILE_64x16, YH_TILE_128x128 = 0, 1, 2, 3, 4, 5
YH_TILE_SMALLM = 6
YH_TILE_RD32 = 7
TILE_DIMS = {YH_TILE_32x32: (32, 32), YH_TILE_64x64: (64, 64), YH_TILE_128x64: (128, 64),
             YH_TILE_16x64: (16, 64), YH_TILE_64x16: (64, 16), YH_TILE_128x128: (128, 128),
             YH_TILE_SMALLM: (16, 256), YH_TILE_RD32: (32, 32)}
YH_SVD_WORDS = 8
YH_SVD_MAX_Q = 2048
YH_QR_WORDS = 8
YH_SPLITK_WORDS = 8
YH_QR_NB = 32
YH_QROP_WORDS = 3
YH_QRADJ_WORDS = 8
YH_SVD_LDS_BYTES = 131072
YH_DIAG_WORDS = 8
YH_NOALLOC_SCRATCH_BYTES = 33554432

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_dbl = ctypes.c_double

# name -> argtypes (all return int)
_SIGS = {
    'yastn_hip_abi_version': [],
    'yastn_hip_check_gather_plan': [_p, _i64, _p, _i64, _i64, _i64],
    'yastn_hip_check_gather_plan2': [_p, _i64, _p, _i64, _i64, _i64, _i64, _i64],
    'yastn_hip_block_gather': [_p, _i64, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_block_gather2': [_p, _i64, _p, _i64, _p, _i64, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_transpose_and_merge_pair': [_p, _i64, _p, _i64, _p, _i64, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_transpose_and_merge_adjoint_pair': [_p, _i64, _p, _i64, _p, _i64, _p, _i64, _i32, _p, _i64, _p,
                                                   _i64, _p],
    'yastn_hip_transpose_and_merge': [_p, _i64, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_transpose_and_merge_adjoint': [_p, _i64, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_unmerge': [_p, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_unmerge_adjoint': [_p, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_transpose': [_p, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_apply_mask': [_p, _i64, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_embed_mask': [_p, _i64, _p, _i64, _i32, _p, _i64, _p, _i64, _p],
    'yastn_hip_check_gemm_plan': [_p, _i64, _p, _i64, _p, _i64, _i32, _p, _i32, _p, _i32],
    'yastn_hip_grouped_gemm': [_p, _p, _i32, _p, _i64, _p, _i64, _p, _i64, _i32, _p],
    'yastn_hip_check_splitk_plan': [_p, _i64, _p, _i32, _i64],
    'yastn_hip_splitk_reduce': [_p, _p, _i32, _p, _i64, _p],
    'yastn_hip_grouped_gemm_noalloc': [_p, _p, _i32, _p, _i64, _p, _i64, _p, _i64, _i32, _i32, _i64, _p, _i64, _p],
    'yastn_hip_dot': [_p, _p, _p, _i32, _p, _i64, _p, _i64, _p, _i64, _i32, _p],
    'yastn_hip_dot_adjoint': [_p, _p, _p, _p, _p, _i32, _p, _i64, _p, _i64, _p, _i64, _i32, _p],
    'yastn_hip_check_svd_plan': [_p, _i64, _p, _i64, _i64, _i32, _i32, _i64, _i64, _i64, _i64, _i64],
    'yastn_hip_svd': [_p, _p, _p, _p, _i32, _p, _i64, _p, _i64, _i64, _p, _p, _i32, _p],
    'yastn_hip_eigh': [_p, _p, _p, _i32, _p, _i64, _p, _i64, _i64, _p, _p, _i32, _p],
    'yastn_hip_fix_svd_signs': [_p, _p, _i32, _p, _i64, _p],
    'yastn_hip_device_info': [_p, _p],
    'yastn_hip_svd_panel_init': [_p, _p, _p, _p, _i32, _i32, _p, _p, _i64, _p],
    'yastn_hip_svd_panel_sweep': [_p, _p, _i32, _p, _p, _p, _i64, _i32, _i32, _i32, _p, _i64, _i32, _i32, _p, _i64,
                                  _p],
    'yastn_hip_check_panel_jobs': [_p, _i64, _p, _i64, _i32, _i32, _i32],
    'yastn_hip_check_qr_op_plan': [_p, _p, _i64, _i32, _i64, _i64, _i64, _i64, _i64],
    'yastn_hip_qr_init': [_p, _p, _i32, _p, _p, _i64, _p],
    'yastn_hip_qr_finalize': [_p, _p, _p, _i32, _p, _p, _i64, _p],
    'yastn_hip_check_qr_adjoint_plan': [_p, _i64, _i32, _i64, _i64, _i64, _i64],
    'yastn_hip_qr_adjoint_core': [_p, _p, _p, _i32, _p, _i64, _p],
    'yastn_hip_qr_adjoint_trsm': [_p, _p, _p, _i32, _p, _i64, _i64, _i64, _p],
    'yastn_hip_svd_panel_finalize': [_p, _p, _p, _p, _p, _i32, _i32, _p, _p, _i64, _p, _i32, _p, _p],
    'yastn_hip_svd_panel_result': [_p, _i64, _i32, _p],
    'yastn_hip_check_qr_plan': [_p, _p, _p, _i64, _i32, _i64, _i64],
    'yastn_hip_svd_qr_panel': [_p, _i32, _p, _p, _i64, _i32, _i64, _p],
    'yastn_hip_svd_precond_init': [_p, _i32, _p, _p, _i64, _p],
    'yastn_hip_svd_precond_finalize': [_p, _p, _p, _p, _i32, _p, _p, _p, _i64, _p, _i32, _p, _p],
    'yastn_hip_svd_sign_pivots': [_p, _p, _i32, _p, _i64, _p],
    'yastn_hip_svd_adjoint_core': [_p, _p, _p, _p, _p, _p, _p, _p, _i32, _p, _i64, _dbl, _i32, _i32, _i32, _p],
    'yastn_hip_eigh_adjoint_core': [_p, _p, _p, _p, _i32, _p, _i64, _i32, _i32, _p],
    'yastn_hip_check_dot_diag_plan': [_p, _i64, _p, _i64, _p, _i64, _i64, _i64, _i64],
    'yastn_hip_dot_diag': [_p, _p, _p, _i32, _p, _i64, _i64, _i32, _p],
    'yastn_hip_dot_diag_adjoint_a': [_p, _p, _p, _i32, _p, _p, _i64, _p, _i64, _p],
}

_lib = None


class HipBackendError(RuntimeError):
    """Raised when the C-ABI library reports a non-zero status."""


def lib():
    """Load the library once; raise loudly if it is absent or of a different ABI version."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"yastn_amd: HIP library not built ({LIB_PATH}); run `python -m yastn_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, argt in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = argt
            f.restype = ctypes.c_int
        L.yastn_hip_last_error.argtypes = []
        L.yastn_hip_last_error.restype = ctypes.c_char_p
        ver = L.yastn_hip_abi_version()
        if ver != YH_ABI_VERSION:
            raise ImportError(f"yastn_amd: ABI version {ver} != {YH_ABI_VERSION}; rebuild the library")
        _lib = L
    return _lib


def exported_symbols():
    return tuple(_SIGS) + ('yastn_hip_last_error',)


def call(name, *args):
    """Call an entry point and raise HipBackendError on a non-zero status."""
    L = lib()
    st = getattr(L, name)(*args)
    if st != YH_OK:
        msg = L.yastn_hip_last_error().decode(errors='replace')
        raise HipBackendError(f"{name} failed (status {st}): {msg}")
    return st


_DEVICE_INFO = None


def _device_info():
    """(LDS bytes per workgroup, compute units) of the current device; (64 KiB, 256) without a GPU."""
    global _DEVICE_INFO
    if _DEVICE_INFO is None:
        try:
            import torch
            if not torch.cuda.is_available():
                return 65536, 256
        except ImportError:
            return 65536, 256
        lds, cu = ctypes.c_int(0), ctypes.c_int(0)
        call('yastn_hip_device_info', ctypes.byref(lds), ctypes.byref(cu))
        _DEVICE_INFO = (int(lds.value), int(cu.value))
    return _DEVICE_INFO


def device_lds_per_block():
    """LDS bytes one workgroup may allocate on the current device (64 KiB when no GPU is visible)."""
    return _device_info()[0]


def device_cu_count():
    """Compute units of the current device (256 when no GPU is visible)."""
    return _device_info()[1]
