"""In-tree build of the native parts of s_spi_spring_amd.

  lib/libspring_hip.so   HIP kernels + C ABI (include/spring_hip.h), gfx950 only;
                         needs the HIP runtime by the soname `libamdhip64.so`
                         (RUNPATH /opt/rocm/lib), see _runtime_stub

Run ``python -m s_spi_spring_amd.build`` or call ``build_all()``. The build runs
without a GPU (hipcc cross-compiles for gfx950).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 VALU forms are disabled for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


ROCM_LIB = os.environ.get("ROCM_LIB", "/opt/rocm/lib")


def _runtime_stub(obj: str, builddir: str, verbose: bool) -> None:
    """Link-time stand-in for the HIP runtime with soname `libamdhip64.so`.

    PyTorch-ROCm wheels ship their own HIP runtime under that unversioned soname,
    while /opt/rocm's carries `libamdhip64.so.7`. Two HIP (and HSA) runtimes in one
    process do not share the device: the second one to initialise finds no GPU.
    Linking against a stub with the unversioned soname makes the dynamic loader
    bind libspring_hip.so to a runtime the process has already loaded under that
    name (torch's, when torch was imported first), and otherwise load
    /opt/rocm/lib/libamdhip64.so through the RUNPATH. The stub defines the HIP
    entry points the object file references and is deleted after the link."""
    r = _run(["nm", "-u", "--format=posix", obj])
    syms = sorted({ln.split()[0] for ln in r.stdout.splitlines()
                   if ln.split() and ln.split()[0].startswith(("hip", "__hip"))})
    src = os.path.join(builddir, "hip_stub.c")
    with open(src, "w") as f:
        f.write("".join(f"void {s}(void) {{}}\n" for s in syms))
    cmd = ["gcc", "-shared", "-fPIC", "-Wl,-soname,libamdhip64.so", "-o",
           os.path.join(builddir, "libamdhip64.so"), src]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)


def build_hip(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    builddir = os.path.join(LIBDIR, "_build")
    os.makedirs(builddir, exist_ok=True)
    out = os.path.join(LIBDIR, "libspring_hip.so")
    obj = os.path.join(builddir, "spr_backend.o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", "-I", INCLUDE, "-I", CSRC, "-o", obj, os.path.join(CSRC, "spr_backend.hip")]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    _runtime_stub(obj, builddir, verbose)
    tmp = out + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-shared", "-fPIC", "-o", tmp, obj, "-L", builddir, "-lamdhip64",
           "-Wl,-z,defs", "-Wl,--enable-new-dtags", "-Wl,-rpath," + ROCM_LIB]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.remove(os.path.join(builddir, "libamdhip64.so"))
    os.replace(tmp, out)
    return out


def build_core(verbose: bool = False) -> str:
    """pybind11 module spring_core (host C++: Solver + bindings), linked against
    lib/libspring_hip.so. Built with g++ (no device code)."""
    import sysconfig

    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "spring_core" + suffix)
    tmp = out + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           "-I", INCLUDE, "-I", CSRC, "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           os.path.join(CSRC, "solver.cpp"), os.path.join(CSRC, "pattern_prep.cpp"), os.path.join(CSRC, "spring_core.cpp"),
           "-L", LIBDIR, "-lspring_hip", "-Wl,-rpath,$ORIGIN/lib", "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all(verbose: bool = False):
    return [build_hip(verbose), build_core(verbose)]


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print("built", p)
