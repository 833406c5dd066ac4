"""Build the wavebit_amd HIP engine for gfx950: ``python -m wavebit_amd.build``.

Compiles every ``csrc/*.hip`` with hipcc into ``wavebit_amd/libwavebit_hip.so`` (in-tree, so
it travels with the source tree). Packed-FP32 instruction selection is disabled for all device
code. Builds always run when asked for; there is no up-to-date check.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / 'csrc'
LIB = HERE / 'libwavebit_hip.so'
ARCH = os.environ.get('WAVEBIT_AMD_ARCH', 'gfx950')

HIPCC_FLAGS = ['-O3', '-fPIC', '-std=c++17', f'--offload-arch={ARCH}',
               '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
               '-Wno-unused-result']


def hipcc():
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not Path(path).exists():
        raise RuntimeError('hipcc not found: the wavebit_amd HIP engine needs ROCm to build')
    return path


def _compile(src, obj):
    cmd = [hipcc(), *HIPCC_FLAGS, '-c', str(src), '-o', str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f'hipcc failed on {src.name}:\n{proc.stderr}')
    return obj


def build(verbose=True):
    """Compile all sources and link the shared library. Returns its path."""
    objdir = HERE / '_build'
    objdir.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob('*.hip'))
    jobs = min(len(sources), int(os.environ.get('MAX_JOBS', '8')))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(lambda s: _compile(s, objdir / (s.stem + '.o')), sources))
    tmp = LIB.with_suffix('.so.tmp')
    cmd = [hipcc(), '-shared', f'--offload-arch={ARCH}', '-fPIC', *map(str, objs), '-o', str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f'link failed:\n{proc.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB} from {len(sources)} sources', file=sys.stderr)
    return LIB


if __name__ == '__main__':
    build()
