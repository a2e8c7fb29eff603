'''
In-tree build of the two native libraries of the package:

- `libtentorch_hip.so`: the HIP kernels and the C ABI of `include/tentorch_amd.h`, compiled by
  `hipcc --offload-arch=gfx950`;
- `_engine*.so`: the C++ host engine (categories, legs, trees, leg-move and contraction planning)
  with its Python binding, compiled by g++ and linked against the first.

Both land next to this file, so they travel with the source tree.  `python tentorch_amd/build.py`
rebuilds them (run it by path: importing the package loads the libraries being rebuilt).
'''

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(ROOT, '_build')
HIP_LIB = os.path.join(HERE, 'libtentorch_hip.so')
HIP_SOURCES = ['device.hip', 'ggemm.hip', 'mlp.hip', 'eigh.hip', 'split.hip']
ENGINE_SOURCES = ['category.cpp', 'tree.cpp', 'move.cpp', 'program.cpp', 'split.cpp', 'bindings.cpp']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
# The soname makes the engine record a bare library name, found through its $ORIGIN run path, so
# the two libraries load from wherever the tree lies.
# Kernel arguments are preloaded into scalar registers (gfx950 supports up to 16 dwords): a launch
# then starts its first memory access without waiting for a load of its arguments, which shortens
# every latency-bound launch (the eigensolver's steps most of all).
HIP_FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared', '-Wl,-soname,libtentorch_hip.so',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-mllvm', '-amdgpu-kernarg-preload-count=16']


def engine_path() -> str:
    return os.path.join(HERE, '_engine' + sysconfig.get_config_var('EXT_SUFFIX'))


def _run(command):
    result = subprocess.run(command, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if result.returncode != 0:
        raise RuntimeError('build step failed: ' + ' '.join(command) + '\n' + result.stdout)
    return result.stdout


def build_hip() -> str:
    '''Compile the device library for gfx950.'''
    os.makedirs(BUILD, exist_ok=True)
    out = HIP_LIB + '.tmp'
    _run([HIPCC, *HIP_FLAGS, '-o', out, *[os.path.join(CSRC, 'hip', name) for name in HIP_SOURCES]])
    os.replace(out, HIP_LIB)
    return HIP_LIB


def build_engine() -> str:
    '''Compile the host engine and its Python binding, linked to the device library.'''
    import pybind11
    os.makedirs(BUILD, exist_ok=True)
    includes = [f'-I{pybind11.get_include()}', f'-I{sysconfig.get_paths()["include"]}']
    flags = ['-std=c++17', '-O2', '-fPIC', '-fvisibility=hidden', '-Wall', '-Wno-unused-variable']

    def compile_one(name):
        obj = os.path.join(BUILD, name.replace('.cpp', '.o'))
        _run(['g++', *flags, *includes, '-c', os.path.join(CSRC, 'engine', name), '-o', obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(ENGINE_SOURCES))) as pool:
        objects = list(pool.map(compile_one, ENGINE_SOURCES))
    out = engine_path() + '.tmp'
    _run(['g++', '-shared', '-o', out, *objects, HIP_LIB, "-Wl,-rpath,$ORIGIN",
          '-Wl,-rpath,/opt/rocm/lib'])
    os.replace(out, engine_path())
    return engine_path()


def build() -> None:
    '''Build both libraries (the device library first: the engine links it).'''
    build_hip()
    build_engine()


if __name__ == '__main__':
    build()
    print('built', HIP_LIB, engine_path(), file=sys.stderr)
