'''Build the FP64 CPU oracle (test infrastructure) into oracle/_build/liboracle_uppe.so.

gcc -O2 without -ffast-math and without FMA contraction flags, so every operation is a plain IEEE double
operation as in the reference's OpenCL C; OpenMP parallelises the per-work-item loops.
'''
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, '_build')
LIB = os.path.join(OUT_DIR, 'liboracle_uppe.so')


def build_oracle(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + '.tmp'
    cmd = ['gcc', '-O2', '-std=c11', '-fopenmp', '-fPIC', '-shared', '-ffp-contract=off', '-Wall',
           '-Wno-unknown-pragmas', '-o', tmp, os.path.join(HERE, 'uppe_oracle.c'), '-lm']
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('oracle build failed:\n%s\n%s' % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build_oracle(verbose='-v' in sys.argv))
