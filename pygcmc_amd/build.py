"""Build the native parts of pygcmc_amd in-tree.

* ``libpygcmc_hip.so`` -- HIP engine + C ABI (``include/pygcmc_hip.h``),
  hipcc for gfx950, plus host-only set-up code (g++).
* ``gpu.cpython-*.so`` -- CPython module ``pygcmc_amd.gpu`` exposing
  ``runGCMC(info, fragmentInfo, residueInfo, atomInfo, grid, ff, moveArray)``
  like pyGCMC's ``gcmc.gpu`` (``gcmc/cpp/gcmc.cpp:162-225``), linked against
  the library above.

Run ``python -m pygcmc_amd.build``.
"""

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libpygcmc_hip.so')
EXT_SUFFIX = sysconfig.get_config_var('EXT_SUFFIX') or '.so'
MODULE = os.path.join(HERE, 'gpu' + EXT_SUFFIX)
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('PYGCMC_ARCH', 'gfx950')

HIP_SOURCES = ['engine.hip', 'grid.hip', 'pme.hip']
# host-only C++ (system set-up), compiled with the host compiler
HOST_SOURCES = ['setup_host.cpp']

HIP_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
    # bit-exact FP32 arithmetic: no contraction, IEEE division and sqrt
    '-ffp-contract=off', '-fhip-fp32-correctly-rounded-divide-sqrt',
    # no packed FP32 instructions
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable', '-Wno-unused-value',
]


def _run(cmd):
    print(' '.join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_library():
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(HERE, '_build', os.path.splitext(src)[0] + '.o')
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        _run([HIPCC, *HIP_FLAGS, '-I', os.path.join(ROOT, 'include'), '-c',
              os.path.join(CSRC, src), '-o', obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(HERE, '_build', os.path.splitext(src)[0] + '.o')
        _run(['g++', '-O2', '-std=c++17', '-fPIC', '-Wall', '-I', os.path.join(ROOT, 'include'), '-c',
              os.path.join(CSRC, src), '-o', obj])
        objs.append(obj)
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), 'lib')
    _run([HIPCC, '-shared', f'--offload-arch={ARCH}', *objs, '-L', rocm_lib, '-lhipfft',
          f'-Wl,-rpath,{rocm_lib}', '-o', LIB + '.tmp'])
    os.replace(LIB + '.tmp', LIB)


def build_module():
    inc = sysconfig.get_paths()['include']
    _run(['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-Wall',
          '-I', inc, '-I', os.path.join(ROOT, 'include'),
          os.path.join(CSRC, 'gpu_module.cpp'), '-o', MODULE + '.tmp',
          '-L', HERE, '-lpygcmc_hip', '-Wl,-rpath,$ORIGIN'])
    os.replace(MODULE + '.tmp', MODULE)


def build():
    build_library()
    build_module()


if __name__ == '__main__':
    build()
    sys.exit(0)
