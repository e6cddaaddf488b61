"""In-tree build of the native libraries (no JIT cache: the .so files travel with the tree).

* ``libeqhip.so``  — HIP kernels + C ABI (include/eqhip.h), compiled for gfx950 with hipcc.
* ``_engine*.so``  — pybind11 module of the C++ host (registry / measures / price), linked
  against libeqhip.so through an rpath of ``$ORIGIN``.
"""
from __future__ import annotations

import os
import subprocess
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for every HIP translation unit of this library.
# Matrix-core results go straight to VGPRs (no accumulator-register reads in the network kernels).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form", "-Wall", "-Wno-unused-result"]

EQHIP_SOURCES = ["eqhip_kernels.hip", "eqhip_api.hip", "eqhip_berm_kernels.hip", "eqhip_berm.hip"]
HOST_SOURCES = ["host/params.cpp", "host/market.cpp", "host/model_product.cpp", "host/contexts.cpp", "host/greeks.cpp",
                "host/measures.cpp", "host/price.cpp"]
PY_SOURCES = ["py/engine_py.cpp"]

LIBEQHIP = os.path.join(PKG, "libeqhip.so")


def _run(cmd):
    proc = subprocess.run(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_eqhip() -> str:
    srcs = [os.path.join(CSRC, s) for s in EQHIP_SOURCES]
    tmp = LIBEQHIP + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{CSRC}", *srcs, "-o", tmp])
    os.replace(tmp, LIBEQHIP)  # atomic: a concurrent reader never sees a partial library
    return LIBEQHIP


def engine_module_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "_engine" + suffix)


def build_engine() -> str:
    """C++ host mirror of the reference registry + its pybind11 module."""
    import pybind11

    out = engine_module_path()
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES + PY_SOURCES]
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-fvisibility=hidden",
          f"-I{INCLUDE}", f"-I{CSRC}", f"-I{py_inc}", f"-I{pybind11.get_include()}", *srcs,
          f"-L{PKG}", "-leqhip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all() -> None:
    build_eqhip()
    if all(os.path.exists(os.path.join(CSRC, s)) for s in HOST_SOURCES + PY_SOURCES):
        build_engine()


if __name__ == "__main__":
    build_all()
