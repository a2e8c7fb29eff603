'''
Builds the reference's engine (`tentorch/_ext`, the C++ behind `einsum`, `contract`, `svd`) from its
sources where they lie under /root/reference, with g++ directly (not the reference's JIT build), into
`oracle/_ref/symmetry_ext.so`.  TEST INFRASTRUCTURE ONLY: the CPU baseline and the oracle's checks
load it; the product never does.

The engine is a torch C++ extension: its blocks are ATen tensors, so it compiles against the torch
headers and links torch's libraries, with the compile flags torch uses for its CPU extensions.  The
sources are only read; objects and the library go to `oracle/_ref/` (ignored by git).

    python oracle/build_ref.py [--jobs N]
'''

import argparse
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, '_ref')
LIBRARY = os.path.join(OUT, 'symmetry_ext.so')
# the engine's translation units (`tentorch/_ext/build.py:13-29` lists the same)
UNITS = ['_ext/bindings.cpp', 'symmetry/cat/cat.cpp', 'symmetry/cat/group.cpp', 'symmetry/cat/abelian/abelian.cpp',
         'symmetry/cat/fibonacci/fibonacci.cpp', 'symmetry/cat/su2/su2.cpp', 'symmetry/cat/su3/su3.cpp',
         'symmetry/tree/tree.cpp', 'tensors/layout.cpp', 'tensors/regroup.cpp', 'tensors/cache.cpp',
         'tensors/place.cpp', 'tensors/tensor.cpp', 'linalg/graded/contract.cpp', 'linalg/graded/svd.cpp',
         'linalg/graded/compare.cpp', 'linalg/graded/einsum.cpp']


def flags():
    import torch
    from torch.utils.cpp_extension import include_paths
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    includes = include_paths() + [sysconfig.get_paths()['include']]
    # every unit is compiled with the engine's shared header first, as its own build does
    pch = os.path.join(REFERENCE, 'tentorch', '_ext', 'pch.h')
    return ['-DTORCH_EXTENSION_NAME=symmetry_ext', '-DTORCH_API_INCLUDE_EXTENSION_H', f'-D_GLIBCXX_USE_CXX11_ABI={abi}',
            *[f'-isystem{p}' for p in includes], '-fPIC', '-std=c++20', '-O3', '-w', '-include', pch]


def build(jobs: int = 8) -> str:
    if not os.path.isdir(REFERENCE):
        raise FileNotFoundError(f'{REFERENCE} is not here: the reference engine builds only where its sources are')
    from torch.utils.cpp_extension import library_paths
    os.makedirs(os.path.join(OUT, 'obj'), exist_ok=True)
    root = os.path.join(REFERENCE, 'tentorch')
    common = flags()

    def compile_one(unit):
        source = os.path.join(root, unit)
        obj = os.path.join(OUT, 'obj', unit.replace('/', '_').replace('.cpp', '.o'))
        if os.path.exists(obj) and os.path.getmtime(obj) > os.path.getmtime(source):
            return obj
        result = subprocess.run(['g++', *common, '-c', source, '-o', obj], capture_output=True, text=True)
        if result.returncode != 0:
            raise RuntimeError(f'g++ failed on {unit}:\n{result.stderr[-4000:]}')
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(compile_one, UNITS))
    libs = library_paths()
    command = ['g++', '-shared', '-o', LIBRARY + '.tmp', *objects, *[f'-L{p}' for p in libs],
               *[f'-Wl,-rpath,{p}' for p in libs], '-lc10', '-ltorch', '-ltorch_cpu', '-ltorch_python']
    result = subprocess.run(command, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f'link failed:\n{result.stderr[-4000:]}')
    os.replace(LIBRARY + '.tmp', LIBRARY)
    return LIBRARY


if __name__ == '__main__':
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument('--jobs', type=int, default=8)
    print(build(parser.parse_args().jobs))
    sys.exit(0)
