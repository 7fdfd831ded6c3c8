"""Build of the HIP library for gfx950.

Replaces the reference's ``libgen`` (``pyjac/libgen/libgen.py:24-283``, which
compiles the generated C/OpenCL with ``g++ -O3 -fopenmp``): here one
mechanism-independent shared library, ``libpyjac_hip.so``, is compiled once
with ``hipcc --offload-arch=gfx950`` and every mechanism is handed to it as
tables at run time.
"""

import os
import shutil
import subprocess

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(ROOT, 'csrc')
INCLUDE = os.path.join(os.path.dirname(ROOT), 'include')
LIB_NAME = 'libpyjac_hip.so'
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)

SOURCES = ['pyjac_kernels.hip', 'pyjac_wave.hip', 'pyjac_block.hip', 'pyjac_fd.hip',
           'pyjac_checks.hip', 'pyjac_noalloc.hip', 'pyjac_integrator.hip',
           'surrogate_kernels.hip', 'runtime.cpp']
HEADERS = ['device_mech.h', 'kernels.h']

# every HIP translation unit is compiled with packed FP32 disabled
HIPCC_FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
               '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def hipcc():
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(path):
        raise RuntimeError('hipcc not found')
    return path


def build_library(out_path=LIB_PATH, extra_flags=(), verbose=False, jobs=8):
    """Compile ``libpyjac_hip.so`` into ``out_path`` (always recompiles):
    every source to an object in parallel, then one link."""
    import concurrent.futures
    import tempfile
    for s in SOURCES:
        if not os.path.exists(os.path.join(CSRC, s)):
            raise RuntimeError('missing source ' + s)
    with tempfile.TemporaryDirectory(prefix='pyjac_build_') as tmpdir:
        def compile_one(src):
            obj = os.path.join(tmpdir, os.path.basename(src) + '.o')
            cmd = [hipcc()] + HIPCC_FLAGS + list(extra_flags) + \
                ['-I', INCLUDE, '-c', os.path.join(CSRC, src), '-o', obj]
            res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                 text=True)
            return obj, res
        with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
            results = list(ex.map(compile_one, SOURCES))
        for obj, res in results:
            if res.returncode != 0:
                raise RuntimeError('hipcc failed:\n' + res.stdout)
            if verbose:
                print(res.stdout)
        tmp = out_path + '.tmp'
        cmd = [hipcc(), '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + \
            [obj for obj, _ in results]
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if res.returncode != 0:
            raise RuntimeError('link failed:\n' + res.stdout)
    os.replace(tmp, out_path)
    if os.path.basename(out_path) == LIB_NAME:
        build_main(os.path.join(os.path.dirname(out_path), MAIN_NAME), os.path.dirname(out_path))
    return out_path


MAIN_NAME = 'pyjac_hip_main'


def build_main(out_path, lib_dir):
    """Compile the executable counterpart of the reference's generated main
    (csrc/pyjac_main.cpp, kernel_utils/common/kernel.cpp.in:645-748) against
    libpyjac_hip.so in lib_dir (found next to the executable at run time)."""
    cmd = [shutil.which('g++') or 'g++', '-O2', '-std=c++17', '-I', INCLUDE,
           os.path.join(CSRC, 'pyjac_main.cpp'), '-L', lib_dir, '-lpyjac_hip',
           '-Wl,-rpath,$ORIGIN:/opt/rocm/lib', '-o', out_path]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError('g++ failed:\n' + res.stdout)
    return out_path


def generate_library(lang='hip', source_dir=None, out_dir=None, **kwargs):
    """Counterpart of the reference's ``libgen.generate_library``
    (libgen.py:283): builds the HIP library and returns its path.  With
    ``out_dir`` the library is also copied there."""
    if lang != 'hip':
        raise NotImplementedError('this build provides the hip target only')
    path = build_library()
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        dst = os.path.join(out_dir, LIB_NAME)
        shutil.copy2(path, dst)
        return dst
    return path
