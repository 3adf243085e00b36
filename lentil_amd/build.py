"""Build the in-tree HIP library ``lentil_amd/libmftx.so`` for gfx950.

    python -m lentil_amd.build          # or __graft_entry__.build()

Compiles every HIP source under ``lentil_amd/csrc`` with hipcc in one
shared-library link.  Packed FP32 instructions are disabled for all device
code (``-packed-fp32-ops``); the host pass warns that it does not know the
feature, which is harmless.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libmftx.so')
ARCH = os.environ.get('MFTX_OFFLOAD_ARCH', 'gfx950')


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(('.hip', '.cpp')))


def command(out=LIB):
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    return [hipcc, f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC', '-shared',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-Wno-unused-result', '-o', out] + sources()


def build(verbose=False):
    tmp = LIB + '.tmp'
    cmd = command(tmp)
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = '\n'.join(l for l in res.stderr.splitlines() if 'packed-fp32-ops' not in l)
    if res.returncode != 0:
        raise RuntimeError(f'hipcc failed ({res.returncode}):\n{err}')
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
