"""Build libtomohip.so (all HIP sources of the package) for gfx950 with hipcc.

Usage: ``python -m tomocupy_amd.build`` or ``tomocupy_amd.build.build()``.
Always recompiles; objects go to ``tomocupy_amd/_build/`` and the shared
library to ``tomocupy_amd/libtomohip.so``.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'libtomohip.so')
BUILD = os.path.join(HERE, '_build')
SOURCES = ['common.hip', 'fourierrec.hip', 'filter.hip', 'preproc.hip', 'stripe.hip',
           'outliers.hip', 'stripe_ti.hip', 'pad360.hip', 'binning.hip',
           'stripe_vo.hip', 'phase.hip', 'rotate.hip']
ARCH = os.environ.get('TOMOHIP_ARCH', 'gfx950')

FLAGS = [
    '--offload-arch=' + ARCH, '-O3', '-fPIC', '-std=c++17',
    '-munsafe-fp-atomics',
    # no packed FP32 (v_pk_*_f32) code generation
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    # matrix-core accumulators in ordinary VGPRs: with the accumulation
    # registers (AGPRs) the allocator rotated the gridding accumulators through
    # ~50 copy instructions per iteration
    '-mllvm', '-amdgpu-mfma-vgpr-form',
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable', '-Werror=return-type',
]


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found (set HIPCC)')


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [hipcc()] + FLAGS + ['-I', CSRC, '-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, ' '.join(cmd), r.stderr))
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for (_, err), src in zip(results, SOURCES):
            if err.strip():
                print('[%s]\n%s' % (src, err), file=sys.stderr)
    tmp = OUT + '.tmp'
    cmd = [hipcc(), '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (' '.join(cmd), r.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
