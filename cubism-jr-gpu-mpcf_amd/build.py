"""Build libmpcf_hip.so (gfx950) in-tree with hipcc.

    python cubism-jr-gpu-mpcf_amd/build.py        # or __graft_entry__.build()

Objects:
  parity.o  mpcf_kernels.hip, MPCF_FAST=0, -ffp-contract=off  (bit-exact scheme)
  fast.o    mpcf_kernels.hip, MPCF_FAST=1, -ffp-contract=on (a*b+c inside one
            expression becomes an FMA; the same in every copy of the code, so
            results do not depend on which unrolled or peeled step, chunk or
            batch computes a cell)
  fastp.o   mpcf_kernels.hip, fast + MPCF_P64=1 (binary64 pressure recovery and reductions)
  fastpe.o  mpcf_kernels.hip, fastp + MPCF_EPAIR=1 (E stored as a binary32 hi/lo pair)
  f64.o     mpcf_kernels.hip, MPCF_F64=1, -ffp-contract=off (the scheme in binary64)
  reduce.o  mpcf_reduce.hip  (binary64 block comparisons and integrals)
  wavelet.o mpcf_wavelet.hip (wavelet-compressed dumps; linked with zlib)
  noalloc.o mpcf_noalloc.hip (the allocation-free speed-of-sound / dt entry point)
  capi.o    the C-ABI (include/mpcf_hip.h)
  ic.o      initial conditions;  sic.o  the shock-induced cloud input and IC
  sim_ic.o  the Sod, 2D shock-bubble and static-index drivers' ICs (-ffp-contract=off)
  io.o      dumps (raw + XDMF) and checkpoints (save.info + save.data)
  flops.o   host instantiations of the scheme
and the C++ driver mpcf_run (csrc/mpcf_run.cpp over csrc/mpcf.hpp).
All HIP code is compiled with packed-FP32 instructions disabled
(-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import concurrent.futures as cf

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmpcf_hip.so")

ARCH = os.environ.get("MPCF_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def common_flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        "-Wno-unused-command-line-argument",
    ]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    units = [
        ("parity.o", "mpcf_kernels.hip", ["-DMPCF_FAST=0", "-ffp-contract=off"]),
        ("fast.o", "mpcf_kernels.hip", ["-DMPCF_FAST=1", "-ffp-contract=on"]),
        ("fastp.o", "mpcf_kernels.hip", ["-DMPCF_FAST=1", "-DMPCF_P64=1", "-ffp-contract=on"]),
        ("fastpe.o", "mpcf_kernels.hip", ["-DMPCF_FAST=1", "-DMPCF_P64=1", "-DMPCF_EPAIR=1", "-ffp-contract=on"]),
        ("f64.o", "mpcf_kernels.hip", ["-DMPCF_FAST=0", "-DMPCF_F64=1", "-ffp-contract=off"]),
        ("reduce.o", "mpcf_reduce.hip", ["-ffp-contract=off"]),
        ("wavelet.o", "mpcf_wavelet.hip", ["-ffp-contract=off"]),
        ("noalloc.o", "mpcf_noalloc.hip", []),
        ("capi.o", "mpcf_capi.cpp", ["-x", "hip"]),
        ("serve.o", "mpcf_serve.cpp", ["-x", "hip"]),
        ("ic.o", "ic.cpp", ["-x", "hip"]),
        ("sic.o", "sic.cpp", ["-x", "hip"]),
        ("sim_ic.o", "sim_ic.cpp", ["-x", "hip", "-ffp-contract=off"]),
        ("io.o", "io.cpp", ["-x", "hip"]),
        ("flops.o", "flops.cpp", ["-x", "hip", "-ffp-contract=off"]),
    ]
    cmds = []
    for obj, src, extra in units:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        cmds.append([cc, *common_flags(), *extra, "-c", path, "-o", os.path.join(BUILD, obj)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for r, cmd in zip(ex.map(_run, cmds), cmds):
            if verbose:
                print(" ".join(cmd))
    objs = [c[-1] for c in cmds]
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"])
    os.replace(tmp, LIB)
    # the C++ driver over the C-ABI (mpcf.hpp), next to the library
    exe = os.path.join(PKG, "mpcf_run")
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          os.path.join(CSRC, "mpcf_run.cpp"), "-o", exe + ".tmp", "-L", PKG, "-lmpcf_hip",
          "-L/opt/rocm/lib", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
