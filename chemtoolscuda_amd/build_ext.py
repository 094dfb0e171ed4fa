"""Builds the native parts in-tree with hipcc / g++ (no CMake, no JIT cache).

Outputs (next to this file):
  * ``libchemtools_amd.so``  -- HIP kernels + C ABI (include/chemtools_amd.h), gfx950
  * ``_chemtools_amd<ext>``  -- pybind11 module linking the library via rpath $ORIGIN

Run:  python chemtoolscuda_amd/build_ext.py
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libchemtools_amd.so")
ARCH = os.environ.get("CT_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["ao_kernels.hip", "contract_f64.hip", "contract_mixed.hip", "finalize.hip", "esp.hip", "adjoint.hip",
               "adjoint_mixed.hip", "screen.hip", "noalloc.hip", "peak.hip"]
HOST_SOURCES = ["mol.cpp", "capi.cpp", "fchk.cpp", "wavefunction.cpp", "esp_host.cpp", "screen_host.cpp",
                "host_copy.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-I", INCLUDE, "-I", CSRC]


def ext_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    srcs = HIP_SOURCES + HOST_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", LIB + ".tmp"])
    os.replace(LIB + ".tmp", LIB)
    import pybind11

    pyinc = sysconfig.get_paths()["include"]
    mod = os.path.join(HERE, "_chemtools_amd" + ext_suffix())
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", pybind11.get_include(), "-I", pyinc, "-I", INCLUDE,
          os.path.join(CSRC, "pymolecule.cpp"), "-L", HERE, "-lchemtools_amd", "-Wl,-rpath,$ORIGIN",
          "-o", mod + ".tmp"])
    os.replace(mod + ".tmp", mod)
    if verbose:
        print("built", LIB, "and", mod)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
