"""Build helpers: compile libhipfem.so for gfx950 in-tree.

The library is built with hipcc directly (no JIT cache), so the shared
object under oofem_amd/lib/ travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libhipfem.so")
HOST_LIB = os.path.join(LIBDIR, "libhipfem_host.so")
SOURCES = ["hipfem_csr.hip", "hipfem_pcg.hip", "hipfem_model.hip", "hipfem_mlp.hip", "hipfem_batch.hip"]
HOST_SOURCES = ["classfactory.C", "hipcsr.C", "hippcg.C", "structmodel.C", "host_capi.C"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off (compiler-generated forms only are avoided)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_device_lib(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), *HIPCC_FLAGS, "-shared", "-o", tmp, *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_host_lib(verbose=False):
    """C++ host mirror of the reference's plug-in classes, linked against
    libhipfem.so (plain g++: no device code)."""
    hdir = os.path.join(HERE, "host")
    srcs = [os.path.join(hdir, s) for s in HOST_SOURCES if os.path.exists(os.path.join(hdir, s))]
    if not srcs:
        return None
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-o", tmp, *srcs, "-L", LIBDIR, "-lhipfem", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def build_all(verbose=False):
    build_device_lib(verbose)
    build_host_lib(verbose)
