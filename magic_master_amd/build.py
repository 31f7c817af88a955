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
HIP_SOURCES = ["sweep.hip", "sweep_tile.hip", "sweep_tile32.hip", "bc_src.hip", "bc_src32.hip", "capi.hip",
               "batch.hip", "sums.hip", "sums_noalloc.hip"]
# the tiled sweeps are compiled once per tile shape (TILE_NSHAPE in magic_dev.h)
PER_SHAPE = ("sweep_tile.hip", "sweep_tile32.hip")
TILE_NSHAPE = 4


def hip_units():
    """(source, object name, extra flags) of every compile of the library."""
    units = []
    for src in HIP_SOURCES:
        stem = os.path.splitext(src)[0]
        if src in PER_SHAPE:
            units += [(src, f"{stem}_s{k}.o", [f"-DMAGIC_TILE_SHAPE={k}"]) for k in range(TILE_NSHAPE)]
        else:
            units.append((src, stem + ".o", []))
    return units
HIP_LIB = os.path.join(HERE, "libmagic_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             # the FP64 step rounds every product and sum as the reference's
             # Fortran does (no contraction into fused multiply-adds): at rest
             # the momenta are differences of O(rho*g*dt) terms, so a single
             # fused rounding moves them by ~1e-9 of their own size
             "-ffp-contract=off",
             # no packed FP32 instructions (v_pk_*_f32) anywhere in this library
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r.stdout


def build_hip() -> str:
    """Each source compiles to its own object (in parallel), then one link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    units = hip_units()
    objs = [os.path.join(objdir, obj) for _, obj, _ in units]
    jobs = max(1, min(len(units), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, [HIPCC] + compile_flags + extra + ["-c", src, "-o", obj], CSRC)
                  for (src, _, extra), obj in zip(units, objs)]:
            f.result()
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


H5_LIB = os.path.join(HERE, "libmagic_h5.so")
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")


def build_h5() -> str:
    """Host-only HDF5 frame writer/reader (include/magic_h5.h) on the HDF5 C
    library of the image."""
    tmp = H5_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", tmp, os.path.join(CSRC, "h5out.cpp"),
          "-I" + os.path.join(HDF5_ROOT, "include"), "-L" + os.path.join(HDF5_ROOT, "lib"), "-lhdf5",
          "-Wl,-rpath," + os.path.join(HDF5_ROOT, "lib")])
    os.replace(tmp, H5_LIB)
    return H5_LIB


FC = os.environ.get("FC", "/opt/rocm/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran")
BIN = os.path.join(HERE, "bin")


def build_fortran() -> str:
    """Fortran host: iso_c_binding module, drop-in claw3/step3, driver program
    (claw3ez_hip) linked against libmagic_hip.so."""
    os.makedirs(BIN, exist_ok=True)
    mod = os.path.join(BIN, "mod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in ("magic_hip.f90", "claw3_hip.f90", "claw3ez_hip.f90", "out3_h5.f90"):
        obj = os.path.join(mod, src.replace(".f90", ".o"))
        # no contraction: the host's background arithmetic rounds as the
        # reference's setaux.f does
        _run([FC, "-c", "-O2", "-fPIC", "-ffp-contract=off", "-module-dir", mod, "-I", mod,
              os.path.join(FORTRAN, src), "-o", obj])
        objs.append(obj)
    exe = os.path.join(BIN, "claw3ez_hip")
    _run([FC, "-o", exe] + objs + ["-L" + HERE, "-lmagic_hip", "-lmagic_h5", "-Wl,-rpath,$ORIGIN/.."])
    # the drop-in operators as a library a reference-style caller links against
    lib = os.path.join(BIN, "libmagic_f77.so")
    _run([FC, "-shared", "-o", lib, objs[0], objs[1], "-L" + HERE, "-lmagic_hip",
          "-Wl,-rpath,$ORIGIN/.."])
    # callers written for the reference: one that calls claw3, one that keeps
    # its own step control and calls step3, bc3, b4step3 and src3
    for prog in ("dropin_check", "step3_loop_check"):
        chk = os.path.join(mod, prog + ".o")
        _run([FC, "-c", "-O2", "-module-dir", mod, "-I", mod, os.path.join(FORTRAN, prog + ".f90"), "-o", chk])
        _run([FC, "-o", os.path.join(BIN, prog), chk, "-L" + BIN, "-lmagic_f77",
              "-L" + HERE, "-lmagic_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."])
    return exe


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "libmagic_oracle.so")


def build_reference(ref_root: str = "/root/reference"):
    """Compile the reference Fortran (checker only) when its sources exist."""
    if not os.path.isdir(ref_root):
        return None
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={ref_root}"])
    # the timing build of the same sources (the CPU baseline's reference kind)
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={ref_root}", "speed"])
    return os.path.join(ROOT, "oracle", "_ref", "libmagic_ref.so")


def build_all():
    out = [build_hip(), build_h5(), build_fortran(), build_oracle()]
    ref = build_reference()
    if ref:
        out.append(ref)
    return out


if __name__ == "__main__":
    for p in build_all():
        print(p)
