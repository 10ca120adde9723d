"""Builds libqsx.so (all HIP sources of this folder) for gfx950 in place.

    python -m qspectra_amd.csrc.build        # or __graft_entry__.build()

One hipcc invocation compiles every .hip file into
qspectra_amd/libqsx.so.  Packed-FP32 instructions are disabled for all
device code (-packed-fp32-ops).
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB = os.path.join(PKG, 'libqsx.so')
ARCH = os.environ.get('QSX_OFFLOAD_ARCH', 'gfx950')


def hipcc():
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    path = os.path.join(rocm, 'bin', 'hipcc')
    return path if os.path.exists(path) else 'hipcc'


def build(verbose=False):
    sources = sorted(glob.glob(os.path.join(HERE, '*.hip')))
    tmp = LIB + '.tmp%d' % os.getpid()
    cmd = [hipcc(), '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-shared',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-Wno-unused-result', '-I', os.path.join(os.path.dirname(PKG), 'include'),
           '-o', tmp] + sources
    if verbose:
        print(' '.join(cmd))
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = '\n'.join(l for l in proc.stdout.splitlines()
                    if 'packed-fp32-ops' not in l)
    if proc.returncode != 0:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise RuntimeError('hipcc failed (%d):\n%s' % (proc.returncode, out))
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
