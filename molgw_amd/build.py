"""Build libmgw.so (HIP, gfx950) in-tree.

Usage: ``python -m molgw_amd.build`` (also called by ``__graft_entry__.build``).
Compiles every source under ``molgw_amd/csrc`` with hipcc for gfx950 and
links them with rocSOLVER/rocBLAS into ``molgw_amd/libmgw.so``. Always
rebuilds (the call is the "does it build" check).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libmgw.so')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
HIPCC = os.path.join(ROCM, 'bin', 'hipcc')
ARCH = 'gfx950'

SOURCES = ['gemm_f64.hip', 'gemm_lowp.hip', 'chi0_split.hip', 'wb_split.hip', 'chol.hip', 'kernels.hip', 'api.hip',
           'ints.hip', 'ints_l0.hip', 'ints_l1.hip', 'ints_l2.hip', 'ints_l3.hip', 'ints_l3_lb2.hip',
           'ints_l3_lb3.hip', 'ints_l4.hip']

# compile time in seconds, measured here (to start the slow ones first)
SLOW = {'ints_l3_lb3.hip': 200, 'ints_l3_lb2.hip': 90, 'ints_l2.hip': 60, 'ints_l3.hip': 50, 'ints_l4.hip': 40,
        'api.hip': 40, 'chi0_split.hip': 30}

CFLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC', '-Wall', '-Wno-unused-result',
    # the integral kernels' runtime-bounded primitive loops are left rolled on purpose
    '-Wno-pass-failed',
    # packed FP32 instructions stay off: the compiler may not form them by itself
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    f'-I{os.path.join(HERE, "..", "include")}',
]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError('command failed: ' + ' '.join(cmd))
    noise = ("is not a recognized feature for this target",)
    for line in (proc.stdout + proc.stderr).splitlines():
        if line.strip() and not any(n in line for n in noise):
            sys.stderr.write(line + '\n')


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace('.hip', '.o'))
        objs.append(obj)
        cmds.append([HIPCC, *CFLAGS, '-c', os.path.join(CSRC, src), '-o', obj])
    # the integral kernels of l_a = 3 take minutes: start them first
    order = sorted(range(len(cmds)), key=lambda k: -SLOW.get(SOURCES[k], 0))
    cmds = [cmds[k] for k in order]
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(4, min(8, os.cpu_count() or 4)))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + '.tmp'
    _run([HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp, *objs,
          f'-L{os.path.join(ROCM, "lib")}', '-lrocsolver', '-lrocblas', '-lamdhip64',
          f'-Wl,-rpath,{os.path.join(ROCM, "lib")}'])
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    build_ints(verbose)
    build_fortran(verbose)
    return LIB


INTS_SRC = os.path.join(HERE, 'frontend', 'ints.cpp')
INTS_LIB = os.path.join(HERE, 'frontend', 'libmgw_ints.so')


def build_ints(verbose: bool = False) -> str:
    """The front end's native integral engine (host C++, OpenMP)."""
    _run(['g++', '-O3', '-fopenmp', '-std=c++17', '-shared', '-fPIC', '-Wall', INTS_SRC, '-o', INTS_LIB + '.tmp'])
    os.replace(INTS_LIB + '.tmp', INTS_LIB)
    if verbose:
        print('built', INTS_LIB)
    return INTS_LIB


FLANG = os.path.join(ROCM, 'lib', 'llvm', 'bin', 'amdflang')
FDIR = os.path.join(HERE, 'fortran')
FBUILD = os.path.join(FDIR, '_build')
FDRIVER = os.path.join(FBUILD, 'mgw_fdriver')


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the iso_c_binding module m_mgw and the Fortran driver that runs
    the one-shot flow through it (linked against libmgw.so)."""
    if not os.path.exists(FLANG):
        if verbose:
            print('amdflang not found: Fortran interface not built')
        return None
    os.makedirs(FBUILD, exist_ok=True)
    obj = os.path.join(FBUILD, 'm_mgw.o')
    _run([FLANG, '-O2', '-c', os.path.join(FDIR, 'm_mgw.f90'), '-module-dir', FBUILD, '-o', obj])
    _run([FLANG, '-O2', '-I', FBUILD, os.path.join(FDIR, 'mgw_fdriver.f90'), obj, f'-L{HERE}', '-lmgw',
          '-Wl,-rpath,$ORIGIN/../..', '-o', FDRIVER + '.tmp'])
    os.replace(FDRIVER + '.tmp', FDRIVER)
    if verbose:
        print('built', FDRIVER)
    return FDRIVER


if __name__ == '__main__':
    build(verbose=True)
