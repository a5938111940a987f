"""Build recipe of the native parts, compiled in-tree for gfx950.

* ``libplainmp_amd.so``: the gfx950 kernels and the C ABI
  (include/plainmp_amd.h), compiled by hipcc.
* ``_native*.so``: the host-side mirror of plainmp's classes with its
  Python bindings, linked against ``libplainmp_amd.so``.

``python plainmp_amd/build.py`` rebuilds both (run as a file: importing the
package needs the extension); ``--incremental`` recompiles only the sources
that changed since their objects were built.
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


def _stale(obj, src):
    """An object is rebuilt when it is older than its source or any header."""
    if not os.path.exists(obj):
        return True
    heads = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".h", ".hpp"))]
    heads += [os.path.join(HERE, "..", "include", f) for f in os.listdir(os.path.join(HERE, "..", "include"))]
    return os.path.getmtime(obj) < max(os.path.getmtime(f) for f in [src] + heads)


def build_library(verbose: bool = False, force: bool = True) -> str:
    """Each source is compiled to an object in parallel, then linked
    (``force=False`` keeps objects newer than their sources and headers)."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    srcs = ["kernels.hip", "cloud.hip", "cloud_build.hip", "noalloc.hip", "ik.hip", "capi.cpp"]
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, s + ".o") for s in srcs]
    jobs = [[hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(srcs, objs)
            if force or _stale(o, os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=max(1, len(jobs))) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-shared", *objs, "-o", tmp], verbose)
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


def build(verbose: bool = False, force: bool = True) -> None:
    build_library(verbose, force)
    build_extension(verbose)


if __name__ == "__main__":
    build(verbose=True, force="--incremental" not in sys.argv)
