'''
Build the owmi HIP library (libowmi.so) for gfx950 in-tree with hipcc.

    python -m optweight_amd.build

Objects go to optweight_amd/_build/, the shared library to
optweight_amd/_lib/libowmi.so. Every HIP source is compiled with packed
FP32 instructions disabled.
'''
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIBDIR = os.path.join(HERE, '_lib')
LIB = os.path.join(LIBDIR, 'libowmi.so')

SOURCES = ['streaming.hip', 'sht_plan.hip', 'legendre.hip', 'ringfft.hip', 'fourier.hip']

HIPCC = os.environ.get('HIPCC', shutil.which('hipcc') or '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('OWMI_OFFLOAD_ARCH', 'gfx950')

CFLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wno-unused-result', '-I', os.path.join(ROOT, 'include'), '-I', CSRC]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + CFLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{res.stderr}')
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'link failed:\n{res.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB}')
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
