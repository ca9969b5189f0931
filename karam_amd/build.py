"""Build libkaram_nr.so in-tree for gfx950 with hipcc.

    python -m karam_amd.build [--verbose]

The library is one HIP translation unit (kernels + host driver + C ABI)
plus the host-only topology builder and SPICE parser.  Packed-FP32 instruction selection is
disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libkaram_nr.so")
SOURCES = [os.path.join(PKG, "csrc", f) for f in ("nr_device.hip", "topology.cpp", "spice.cpp")]
DEPENDS = SOURCES + [os.path.join(PKG, "csrc", f) for f in ("nr_host.inc", "seeder.inc", "topology.h", "spice.h")] + \
    [os.path.join(ROOT, "include", "karam_nr.h")]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def command(out=LIB, arch="gfx950"):
    return [hipcc(), f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared", "-DKARAM_NR_BUILD",
            "-fvisibility=hidden", "-I" + os.path.join(ROOT, "include"),
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *SOURCES, "-o", out]


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libkaram_nr.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
