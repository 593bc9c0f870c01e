"""Build the C-ABI HIP library ``libyastn_hip.so`` in-tree for gfx950 (hipcc cross-compiles; no GPU needed)."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, 'libyastn_hip.so')
SOURCES = ('blockcopy.hip', 'gemm.hip', 'jacobi.hip', 'misc.hip')
ARCH = os.environ.get('YASTN_HIP_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['-O3', '-fPIC', '-std=c++17', f'--offload-arch={ARCH}',
         # packed FP32 instructions are kept out of every kernel
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I', os.path.join(ROOT, 'include'), '-Wall']


def _compile(src, objdir):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC, *FLAGS, '-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir), SOURCES))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', *objs, '-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
