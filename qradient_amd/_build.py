"""Build libqradient_amd.so in-tree with hipcc for gfx950.

Every translation unit is compiled with the packed-FP32 target feature
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``), and the
library is written next to this file so it ships with the package.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libqradient_amd.so')
SOURCES = ['qg_kernels_c64.hip', 'qg_kernels_c64x.hip', 'qg_kernels_c128.hip', 'qg_finalize.hip', 'qg_fused.hip', 'qg_simple.hip',
           'qg_noalloc.hip', 'qg_stream.hip', 'qg_api.cpp', 'qg_server.cpp']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I' + os.path.join(HERE, '..', 'include')]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for {}:\n{}'.format(src, r.stderr))
    return obj


def build(verbose=False):
    """Compile all sources (in parallel) and link the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
