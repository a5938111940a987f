"""Build recipe of the native parts, compiled in-tree for gfx950.

* ``libplainmp_amd.so``: the gfx950 kernels and the C ABI
  (include/plainmp_amd.h), compiled by hipcc.
* ``_native*.so``: the host-side mirror of plainmp's classes with its
  Python bindings, linked against ``libplainmp_amd.so``.

``python plainmp_amd/build.py`` rebuilds both (run as a file: importing the
package needs the extension).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("PMP_OFFLOAD_ARCH", "gfx950")
LIB = os.path.join(HERE, "libplainmp_amd.so")
EXT = os.path.join(HERE, "_native" + sysconfig.get_config_var("EXT_SUFFIX"))

# Packed FP32 instructions stay disabled: hipcc forms them from plain float
# code, and the device code here is written for scalar FP32 issue.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")


def build_library(verbose: bool = False) -> str:
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    srcs = [os.path.join(CSRC, "kernels.hip"), os.path.join(CSRC, "cloud.hip"),
            os.path.join(CSRC, "cloud_build.hip"), os.path.join(CSRC, "capi.cpp")]
    tmp = LIB + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-shared", *srcs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_extension(verbose: bool = False) -> str:
    import pybind11

    cxx = os.environ.get("CXX", "g++")
    inc = [
        f"-I{sysconfig.get_paths()['include']}",
        f"-I{pybind11.get_include()}",
        f"-I{os.path.join(ROCM, 'include')}",
    ]
    tmp = EXT + ".tmp"
    _run(
        [
            cxx,
            "-O2",
            "-std=c++17",
            "-fPIC",
            "-shared",
            "-D__HIP_PLATFORM_AMD__",
            *inc,
            os.path.join(CSRC, "host", "bindings.cpp"),
            "-o",
            tmp,
            f"-L{HERE}",
            "-lplainmp_amd",
            f"-L{os.path.join(ROCM, 'lib')}",
            "-lamdhip64",
            "-Wl,-rpath,$ORIGIN",
            f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}",
        ],
        verbose,
    )
    os.replace(tmp, EXT)
    return EXT


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_extension(verbose)


if __name__ == "__main__":
    build(verbose=True)
