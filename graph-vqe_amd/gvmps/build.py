"""Builds libgvmps.so in-tree for gfx950 (hipcc, no JIT cache).

    python -m gvmps.build          (from graph-vqe_amd/)

Host C++ (plan.cpp) is compiled with g++, HIP sources with hipcc
--offload-arch=gfx950; every HIP compile carries the packed-FP32 opt-out.
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = ROOT / 'csrc'
LIB = PKG / 'libgvmps.so'
HIP_SOURCES = ['kernels_prep.hip', 'kernels_energy.hip', 'kernels_real.hip', 'kernels_sample.hip', 'kernels_grad.hip',
               'api.hip',
               'kernels_gnn.hip', 'gnn_api.hip']
CPP_SOURCES = ['plan.cpp']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
HIP_FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-Wno-unused-result', f'-I{CSRC}']
CXX_FLAGS = ['-O2', '-std=c++17', '-fPIC', '-Wall', '-Wextra', f'-I{CSRC}']


def _run(cmd, verbose):
    if verbose:
        print(' '.join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(str(c) for c in cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose=False, jobs=8):
    objdir = ROOT / '_build'
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in CPP_SOURCES:
        obj = objdir / (src + '.o')
        objs.append(obj)
        procs.append(['g++', *CXX_FLAGS, '-c', CSRC / src, '-o', obj])
    for src in HIP_SOURCES:
        obj = objdir / (src + '.o')
        objs.append(obj)
        procs.append([HIPCC, *HIP_FLAGS, '-c', CSRC / src, '-o', obj])
    running = []
    for cmd in procs:
        if verbose:
            print(' '.join(str(c) for c in cmd), flush=True)
        running.append((cmd, subprocess.Popen([str(c) for c in cmd], stdout=subprocess.PIPE,
                                              stderr=subprocess.PIPE, text=True)))
        if len(running) >= jobs:
            _wait(running.pop(0))
    for r in running:
        _wait(r)
    tmp = LIB.with_suffix('.so.tmp')
    _run([HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950', *objs, '-o', tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def _wait(item):
    cmd, p = item
    out, err = p.communicate()
    if p.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(str(c) for c in cmd)}\n{out}\n{err}")


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
