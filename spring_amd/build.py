"""Build libspring_gpu.so (gfx950) in-tree with hipcc.

Every HIP source under spring_amd/csrc is compiled for gfx950 only, with packed FP32 VALU
instructions disabled, and linked into spring_amd/libspring_gpu.so.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspring_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
]


def build(verbose: bool = False, jobs: int = 8) -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [HIPCC, *FLAGS, "-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


def _drain(procs):
    errs = []
    for s, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        if p.returncode != 0:
            errs.append(f"{s}:\n{txt}")
        elif "warning" in txt and os.environ.get("SPR_BUILD_VERBOSE"):
            sys.stderr.write(txt)
    procs.clear()
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))


H5_SRC = os.path.join(HERE, "h5", "spr_h5.cpp")
H5_LIB = os.path.join(HERE, "libspring_h5.so")
H5_ROOT = os.environ.get("SPR_HDF5_ROOT", "/opt/conda")


def build_h5(verbose: bool = False) -> str | None:
    """libspring_h5.so: the HDF5 file boundary (host code, g++ over the HDF5 C library).
    Returns None when no HDF5 installation is found (the GPU library does not depend on it)."""
    inc, lib = os.path.join(H5_ROOT, "include"), os.path.join(H5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print(f"no hdf5.h under {inc}: libspring_h5.so not built", flush=True)
        return None
    tmp = H5_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{inc}", H5_SRC, "-o", tmp,
           f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("libspring_h5.so build failed:\n" + r.stdout)
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_h5(verbose=True))
