"""Builds libb2g.so (the gfx950 HIP library) in-tree with hipcc.

    python -m pypybox2d_amd.build

Device code is compiled with -ffp-contract=off (no FMA contraction, so the
FP64 pipeline follows the reference's operation order) and with packed FP32
instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libb2g.so")
SOURCES = [os.path.join(CSRC, "b2g_engine.hip")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("b2g_math.hpp", "b2g_collide.hpp", "b2g_raycast.hpp", "b2g_toi.hpp", "b2g_joints.hpp")] + [
    os.path.join(REPO, "include", f) for f in ("b2g.h", "b2g_layout.h")]

ARCH = os.environ.get("B2G_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def command(out: str = LIB):
    return [hipcc(), "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-fvisibility=hidden",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(REPO, "include"), "-o", out] + SOURCES


def build(force: bool = True, verbose: bool = False) -> str:
    if not force and os.path.exists(LIB):
        lib_t = os.path.getmtime(LIB)
        if all(os.path.getmtime(p) <= lib_t for p in DEPS):
            return LIB
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libb2g.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True, verbose=True))
