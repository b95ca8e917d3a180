"""Build libsnoo_hip.so (gfx950) in-tree with hipcc.

Usage: python snoo.py_amd/build_lib.py  (or __graft_entry__.build()).
Each translation unit is compiled to an object in _build/ and linked into
snoo.py_amd/libsnoo_hip.so next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(ROOT, '_build')
OUT = os.path.join(HERE, 'libsnoo_hip.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

SOURCES = ['sn_sources.hip', 'sn_pattern.hip', 'sn_assemble.hip', 'sn_cg.hip', 'sn_ams.hip',
           'sn_mlp.hip', 'sn_reftable.hip', 'sn_galerkin.hip', 'sn_vcg.hip']

FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950',
         # packed FP32 instructions are kept out of the generated code
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result', '-I' + os.path.join(ROOT, 'include')]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for {}:\n{}\n{}'.format(src, ' '.join(cmd), r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(len(srcs), 8)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n{}\n{}'.format(' '.join(cmd), r.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT)
    return OUT


if __name__ == '__main__':
    build()
    sys.exit(0)
