"""Build the gfx950 shared library ``hydroexa_amd/lib/libhydroexa.so`` with hipcc.

The library is built in-tree so the same file is loaded by tests, smoke and
the benchmark.  Parity FMA policy: ``-ffp-contract=off`` (the FP64 kernels
then round exactly like the CPU restatement).  Packed FP32 instructions are
disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libhydroexa.so")
# (source, FP contraction policy): the exact build must not contract; the
# fast build contracts within an expression only, so that every inlined copy
# of a face evaluates it identically (a y-invariant flow keeps hv = 0 exactly).
SOURCES = [(os.path.join(CSRC, "hx_kernels.hip"), "off"),
           (os.path.join(CSRC, "hx_fast.hip"), "on"),
           (os.path.join(CSRC, "hx_adjoint.hip"), "fast")]
HOST_SOURCES = [os.path.join(CSRC, "hx_regrid.cpp")]
ARCH = os.environ.get("HYDROEXA_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def flags() -> list:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-command-line-argument",
        f"-I{os.path.join(ROOT, 'include')}",
    ]


HDF5_ROOT = os.environ.get("HYDROEXA_HDF5", "/opt/conda")
IO_LIB = os.path.join(LIBDIR, "libhydroexa_io.so")


def build_io(verbose: bool = False) -> str:
    """Host-only HDF5 input library (libhydroexa_io.so) with g++ against the
    serial HDF5 C library.  Returns the path, or "" when HDF5 is absent."""
    inc = os.path.join(HDF5_ROOT, "include")
    lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        sys.stderr.write(f"HDF5 not found under {HDF5_ROOT}: libhydroexa_io.so not built\n")
        return ""
    os.makedirs(LIBDIR, exist_ok=True)
    out = IO_LIB
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-static-libstdc++",
           "-static-libgcc",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{inc}",
           os.path.join(CSRC, "hx_io.cpp"), "-o", tmp,
           f"-L{lib}", f"-Wl,-rpath,{lib}", "-lhdf5"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building libhydroexa_io.so failed")
    os.replace(tmp, out)
    return out


CXX_CHECK_SRC = os.path.join(ROOT, "tests", "cxx", "binding_check.cpp")
CXX_CHECK = os.path.join(LIBDIR, "hx_binding_check")


def build_cxx_check(verbose: bool = False) -> str:
    """The C++ caller of the C-ABI (tests/cxx/binding_check.cpp), compiled with
    the host compiler against include/hydroexa.h and the HIP runtime API and
    linked to libhydroexa.so, as a HydroEXA solver translation unit would be."""
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(hipcc())))
    tmp = CXX_CHECK + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(rocm, 'include')}",
           CXX_CHECK_SRC, "-o", tmp, f"-L{LIBDIR}", "-lhydroexa",
           f"-L{os.path.join(rocm, 'lib')}", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{os.path.join(rocm, 'lib')}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building the C++ binding check failed")
    os.replace(tmp, CXX_CHECK)
    return CXX_CHECK


def build(verbose: bool = False) -> str:
    """Compile every HIP source (the three translation units in parallel) and
    link the in-tree shared library, then the HDF5 input library.  Always
    rebuilds."""
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    for src, contract in SOURCES:
        obj = os.path.join(LIBDIR, os.path.basename(src) + ".o")
        cmd = [hipcc()] + flags() + [f"-ffp-contract={contract}", "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        log = tempfile.TemporaryFile(mode="w+")      # a pipe could fill and stall hipcc
        jobs.append((src, obj, log, subprocess.Popen(cmd, stdout=log, stderr=subprocess.STDOUT)))
    failed = []
    for src, obj, log, proc in jobs:
        proc.wait()
        if proc.returncode != 0:
            log.seek(0)
            sys.stderr.write(log.read())
            failed.append(os.path.basename(src))
        log.close()
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    # host-only bookkeeping (grid generation) in plain C++
    host_objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(LIBDIR, os.path.basename(src) + ".o")
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
               "-c", src, "-o", obj]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"g++ failed on {os.path.basename(src)}")
        host_objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC"] + [o for _, o, _, _ in jobs] + \
          host_objs + ["-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking libhydroexa.so failed")
    os.replace(tmp, LIB)
    build_io(verbose)
    build_cxx_check(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
