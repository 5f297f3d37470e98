"""Build the gfx950 shared library ``nipet_amd/libnipet_hip.so`` with hipcc.

Used by ``__graft_entry__.build()``; can also be run as
``python -m nipet_amd._build``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libnipet_hip.so')
SOURCES = ['nipet_geom.hip', 'nipet_prj.hip', 'nipet_hist.hip', 'nipet_psf.hip',
           'nipet_rand.hip', 'nipet_norm.hip']

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    '-munsafe-fp-atomics',
    # no packed FP32 instructions (v_pk_*_f32)
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result',
]


def build(verbose=False):
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + '.tmp'
    cmd = [hipcc, *HIPCC_FLAGS, '-I', os.path.join(ROOT, 'include'), '-I', CSRC, *srcs,
           '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    build(verbose='-v' in sys.argv)
