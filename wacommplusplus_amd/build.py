"""Builds libwacomm_hip.so in-tree for gfx950 with hipcc.

    python -m wacommplusplus_amd.build        (or __graft_entry__.build())

Every HIP source is compiled with contraction off by default (the parity
kernels keep the reference's rounding points; the mixed-precision source turns
contraction on with a pragma), IEEE denormals, and packed-FP32 instructions
disabled.
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
LIB = os.path.join(PKG, "libwacomm_hip.so")
SOURCES = ["wc_api.hip", "wc_move_parity.hip", "wc_move_fast.hip", "wc_grid.hip", "wc_driver.hip",
           "wc_roms.hip", "wc_unet.hip", "wc_unet_api.hip",
           "wc_netcdf.hip", "wc_noalloc.hip"]
ARCH = os.environ.get("WC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libwacomm_hip.so)")


FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument", "-Wno-unused-value",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    cc = hipcc()
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    with tempfile.TemporaryDirectory(prefix="wc_build_") as tmp:
        objs = []
        for src in SOURCES:
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            cmd = [cc, *FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
            objs.append(obj)
        tmp_lib = LIB + ".tmp"
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp_lib, *objs]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp_lib, LIB)
    build_cpp_tests(verbose)
    return LIB


CPP_TESTS = os.path.join(ROOT, "tests", "cpp")


def build_cpp_tests(verbose: bool = False) -> None:
    """The C++ host-layer tests (tests/cpp/*.cpp), built with g++ against
    include/wacomm_hip.hpp and the library, into tests/cpp/_bin/."""
    out = os.path.join(CPP_TESTS, "_bin")
    os.makedirs(out, exist_ok=True)
    for name in sorted(os.listdir(CPP_TESTS)):
        if not name.endswith(".cpp"):
            continue
        exe = os.path.join(out, name[:-4])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(CPP_TESTS, name), "-L", PKG, "-lwacomm_hip",
               "-Wl,-rpath,$ORIGIN/../../../wacommplusplus_amd", "-o", exe + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
