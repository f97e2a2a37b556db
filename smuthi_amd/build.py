"""Build libsmh.so (the HIP/CDNA4 kernels and the C ABI) in-tree for gfx950.

Usage: python -m smuthi_amd.build   (or smuthi_amd.build.build())
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libsmh.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950',
         # no packed FP32 instructions anywhere in the device code
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I', os.path.join(HERE, '..', 'include')]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src).replace('.hip', '.o'))
    cmd = [HIPCC, '-c'] + FLAGS + ['-o', obj, src]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, res.stderr[-4000:]))
    return obj


def build(verbose=True):
    """Compile every source under csrc/ and link libsmh.so next to this file."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))
    workers = min(len(srcs), max(1, (os.cpu_count() or 2)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    # rocSOLVER / rocBLAS: the LU factorization of the explicit master matrix
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950', '-o', tmp] + objs + [
        '-L/opt/rocm/lib', '-lrocsolver', '-lrocblas', '-Wl,-rpath,/opt/rocm/lib']
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('link failed:\n%s' % res.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
