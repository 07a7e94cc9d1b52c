"""Build recipe for the oopao_amd HIP library (gfx950).

Compiles every ``oopao_amd/csrc/*.hip`` / ``*.cpp`` with hipcc into the in-tree
shared library ``oopao_amd/liboopao_amd.so``. Run ``python -m oopao_amd.build``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')
OBJDIR = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'liboopao_amd.so')
ARCH = 'gfx950'
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

DEVICE_FLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-fPIC', '-std=c++17',
    # keep the compiler from forming packed-FP32 instructions
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith('.hip') or f.endswith('.cpp'):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + '.o')
    cmd = [HIPCC] + DEVICE_FLAGS + ['-I', CSRC, '-I', INCLUDE, '-c', src, '-o', obj]
    if src.endswith('.cpp'):
        cmd += ['-x', 'hip']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{r.stderr}')
    return obj


def build(verbose=False, jobs=None):
    """Compile all sources and link liboopao_amd.so (always rebuilds)."""
    os.makedirs(OBJDIR, exist_ok=True)
    from concurrent.futures import ThreadPoolExecutor
    srcs = sources()
    jobs = jobs or min(len(srcs), 8)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB}')
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
