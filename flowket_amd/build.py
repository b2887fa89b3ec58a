"""Builds libflowket_hip.so in-tree with hipcc for gfx950.

Usage: python -m flowket_amd.build  (or flowket_amd.build.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libflowket_hip.so')
OBJ_DIR = os.path.join(HERE, '_build')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = 'gfx950'
# Packed FP32 is disabled for device code: hand-scheduled float code stays on
# the scalar-per-lane VALU path (the host pass warns that the feature is
# unknown to it, which is harmless).
COMMON = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wno-unused-result', '-I', os.path.join(HERE, '..', 'include')]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + '.o')
    cmd = [HIPCC] + COMMON + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{r.stderr}')
    return obj


def _stale(src, headers_mtime):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + '.o')
    if not os.path.exists(obj):
        return True
    return os.path.getmtime(obj) < max(os.path.getmtime(src), headers_mtime)


def build(verbose=False, changed_only=False):
    """Compiles every source (or, with changed_only, those whose object is
    older than the source or any header) and links the library."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    todo = srcs
    if changed_only:
        hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.h')]
        hdrs.append(os.path.join(HERE, '..', 'include', 'flowket_hip.h'))
        hm = max(os.path.getmtime(h) for h in hdrs)
        todo = [s for s in srcs if _stale(s, hm)]
    jobs = min(max(len(todo), 1), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(OBJ_DIR, os.path.basename(s) + '.o') for s in srcs]
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', f'--offload-arch={ARCH}', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True, changed_only='--changed' in sys.argv)
    sys.exit(0)
