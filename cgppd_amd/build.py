"""In-tree build of the native libraries.

libcgppd_hip.so  -- gfx950 kernels + C ABI (include/cgppd_hip.h), hipcc
libcgppd_host.so -- C++ host side mirroring the reference's Replica /
                    Simulation / loader API on top of the C ABI

Run ``python -m cgppd_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

HIP_LIB = PKG / "libcgppd_hip.so"
HOST_LIB = PKG / "libcgppd_host.so"

ARCH = os.environ.get("CGPPD_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {cmd[0]}")
    return res


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build_hip(verbose=False):
    tmp = HIP_LIB.with_suffix(".so.tmp")
    _run([hipcc(), *HIP_FLAGS, "-o", tmp, CSRC / "cgppd_hip.hip", CSRC / "topology.cpp"], verbose)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


CLI = PKG / "cgppd"


def build_host(verbose=False):
    """C++ host library (Replica/Simulation API) and the cgppd command line."""
    srcs = [CSRC / "host" / "cgppd_host.cpp", CSRC / "host" / "host_capi.cpp"]
    common = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{CSRC / 'host'}"]
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run([*common, "-shared", "-o", tmp, *srcs, f"-L{PKG}", "-lcgppd_hip", "-Wl,-rpath,$ORIGIN", "-lpthread"], verbose)
    os.replace(tmp, HOST_LIB)
    tmp = CLI.with_suffix(".tmp")
    _run([*common, "-o", tmp, CSRC / "host" / "main.cpp", f"-L{PKG}", "-lcgppd_host", "-lcgppd_hip",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, CLI)
    return HOST_LIB


def build_oracle(verbose=False):
    """The CPU oracle (test infrastructure) and, where the reference tree is
    present, oracle/_ref."""
    _run(["make", "-s", "-C", ROOT / "oracle", "oracle"], verbose)
    if Path("/root/reference/src/Potential.cpp").exists():
        _run(["make", "-s", "-C", ROOT / "oracle", "ref", "ref_flex"], verbose)


def build_all(verbose=False):
    build_hip(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
