"""Builds libpercolate_hip.so in-tree for gfx950 (python -m pypercolate_amd.build)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, 'csrc', 'percolate_hip.hip')]
OUTPUT = os.path.join(HERE, 'libpercolate_hip.so')

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    '-Wall', '-I' + os.path.join(ROOT, 'include'),
    # no packed-FP32 instructions in any kernel
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
]


def hipcc():
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    cand = os.path.join(rocm, 'bin', 'hipcc')
    return cand if os.path.exists(cand) else 'hipcc'


def build(verbose=False):
    tmp = OUTPUT + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-o', tmp] + SOURCES
    if verbose:
        print(' '.join(cmd))
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          universal_newlines=True)
    out = '\n'.join(l for l in proc.stdout.splitlines()
                    if 'packed-fp32-ops' not in l)
    if proc.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + out)
    if out.strip() and verbose:
        print(out)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
