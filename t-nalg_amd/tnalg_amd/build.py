"""Build the gfx950 HIP library ``libtnalg_hip.so`` in-tree.

``python -m tnalg_amd.build`` (with ``t-nalg_amd`` on ``sys.path``) or
``__graft_entry__.build()`` compiles every ``csrc/*.hip`` with hipcc for
gfx950 and links ``tnalg_amd/_lib/libtnalg_hip.so``.  Packed-FP32 VALU forms
are disabled (``-packed-fp32-ops``) for every device compile.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
SRC_DIR = os.path.join(os.path.dirname(PKG_DIR), 'csrc')
LIB_DIR = os.path.join(PKG_DIR, '_lib')
LIB_PATH = os.path.join(LIB_DIR, 'libtnalg_hip.so')
OBJ_DIR = os.path.join(LIB_DIR, 'obj')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = 'gfx950'
DEVICE_FLAGS = ['-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
COMMON = ['--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-Wall',
          '-Wno-unused-function', '-Wno-unused-variable']


def sources():
    return sorted(os.path.join(SRC_DIR, f) for f in os.listdir(SRC_DIR) if f.endswith('.hip'))


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src)[:-4] + '.o')
    cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ['-c', src, '-o', obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = '\n'.join(l for l in (p.stdout + p.stderr).splitlines()
                    if 'packed-fp32-ops' not in l)
    if p.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, out))
    return obj, out


def build_library(verbose=False, jobs=8):
    """Compile all sources and link the shared library; returns its path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for obj, out in results:
            if out.strip():
                print(out)
    tmp = LIB_PATH + '.tmp'
    cmd = [HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + [o for o, _ in results]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('link failed:\n%s%s' % (p.stdout, p.stderr))
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    print(build_library(verbose='-v' in sys.argv))
