'''Build libsearay_uppe.so in-tree with hipcc for gfx950.

Usage: python -m searay_amd.build   (or searay_amd.build.build_library()).
Objects go to searay_amd/_build/, the library to searay_amd/_lib/libsearay_uppe.so.
'''
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIBDIR = os.path.join(HERE, '_lib')
LIBNAME = 'libsearay_uppe.so'
SOURCES = ['uppe_api.hip', 'uppe_plain.hip', 'uppe_fused.hip']
ARCH = os.environ.get('SEARAY_HIP_ARCH', 'gfx950')

HIPCC_FLAGS = [
    '-std=c++17', '-O3', '-fPIC', '--offload-arch=' + ARCH,
    # packed FP32 instructions stay off: hipcc would otherwise form them from scalar float code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def _hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [_hipcc()] + HIPCC_FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, r.stdout, r.stderr))
    return obj


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + '.tmp'
    cmd = [_hipcc(), '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (r.stdout, r.stderr))
    os.replace(tmp, out)
    return out


if __name__ == '__main__':
    print(build_library(verbose='-v' in sys.argv))
