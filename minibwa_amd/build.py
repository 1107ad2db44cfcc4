"""Compile libmbg.so (HIP, gfx950) and libmbsim.so (C) in-tree.

    python -m minibwa_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIP_SRC = ["csrc/mbg_api.hip", "csrc/mbg_fm.hip", "csrc/mbg_dp.hip", "csrc/mbg_pipe.hip", "csrc/mbg_scan.hip",
           "csrc/mbg_ll.hip", "csrc/mbg_rescue.hip",
           "csrc/mbg_extra.hip", "csrc/mbg_align.hip", "csrc/mbg_anchor.hip", "csrc/mbg_front.hip", "csrc/mbg_io.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wno-unused-value",
             # packed fp32 VALU ops stay off in every HIP object of this library
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # the few floating-point results (mate rescue's dp_max2, mb_update_extra's log-gap
             # scores) must round like the host code they reproduce: no fused multiply-add
             "-ffp-contract=off"]


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return r


def build_mbg() -> str:
    out = os.path.join(HERE, "libmbg.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *HIP_SRC, "-lz", "-lpthread"], HERE)
    os.replace(tmp, out)
    return out


def build_sim() -> str:
    out = os.path.join(HERE, "libmbsim.so")
    _run(["gcc", "-O2", "-Wall", "-fPIC", "-shared", "-o", out + ".tmp", "sim/mbsim.c", "-lm"], HERE)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_mbg()
    build_sim()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(HERE, "libmbg.so"), os.path.join(HERE, "libmbsim.so"))
