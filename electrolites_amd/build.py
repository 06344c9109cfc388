"""Ahead-of-time build of ``libmi355df.so`` for gfx950 with hipcc.

Every ``csrc/*.hip`` file (hand-written and generated) is compiled to an
object and linked into one shared library next to this file.  Run as
``python -m electrolites_amd.build``.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libmi355df.so')
ARCH = 'gfx950'

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=' + ARCH, '-O3', '-fPIC', '-std=c++17',
         # packed-FP32 instructions stay off (compiler-generated v_pk_*_f32)
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         # MFMA accumulators in VGPRs, never AGPRs: a register-only loop of
         # v_mfma_f64_16x16x4 runs at 77.6 TF/s with VGPR accumulators and
         # 47 TF/s with AGPR ones on gfx950, and the K step-1 tile of 160
         # columns (AGPR accumulators by default) ran 65.5 -> 48.9 ms per pass
         '-mllvm', '-amdgpu-mfma-vgpr-form',
         '-Wno-unused-command-line-argument', '-Wno-unknown-warning-option',
         '-I' + os.path.join(HERE, '..', 'include')]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith('.hip'))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'hipcc failed on {src}:\n{r.stderr}')
    return obj


def build(jobs=None, verbose=True):
    """Compile every source and link the library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB} from {len(srcs)} sources')
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
