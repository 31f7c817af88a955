"""Build recipes: the HIP library (product) and the CPU checkers.

    python -m magic_master_amd.build          # HIP library + oracle (+ reference if mounted)

The HIP library is compiled with hipcc for gfx950 only, in-tree, into
magic_master_amd/libmagic_hip.so.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIP_SOURCES = ["sweep.hip", "bc_src.hip", "capi.hip"]
HIP_LIB = os.path.join(HERE, "libmagic_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             # no packed FP32 instructions (v_pk_*_f32) anywhere in this library
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r.stdout


def build_hip() -> str:
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp] + HIP_SOURCES, cwd=CSRC)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "libmagic_oracle.so")


def build_reference(ref_root: str = "/root/reference"):
    """Compile the reference Fortran (checker only) when its sources exist."""
    if not os.path.isdir(ref_root):
        return None
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={ref_root}"])
    return os.path.join(ROOT, "oracle", "_ref", "libmagic_ref.so")


def build_all():
    out = [build_hip(), build_oracle()]
    ref = build_reference()
    if ref:
        out.append(ref)
    return out


if __name__ == "__main__":
    for p in build_all():
        print(p)
