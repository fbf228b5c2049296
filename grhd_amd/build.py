"""Build the HIP library (libgrhd_hip.so) in-tree for gfx950.

``python -m grhd_amd.build`` or ``grhd_amd.build.build()``.  The library is a
plain C-ABI shared object (include/grhd_c.h); no PyTorch extension machinery.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libgrhd_hip.so"
KINDS = ["kslog", "ks", "ef", "mink"]
SOURCES = ([f"grhd_stage_f64_{k}.hip" for k in KINDS] + [f"grhd_stage_f32_{k}.hip" for k in KINDS] +
           ["grhd_batch_f64_a.hip", "grhd_batch_f64_b.hip", "grhd_batch_mixed_a.hip", "grhd_batch_mixed_b.hip",
            "grhd_kernels.hip", "grhd_stage_dispatch.hip", "grhd_initial.hip", "grhd_capi.hip", "grhd_batch.hip"])
# The mixed-precision stage is held to an error bound against the FP64 path, so
# it may contract multiply-adds; every other unit keeps the reference's
# rounding sequence.  Contraction is "on" (within one source expression), not
# "fast": the fused operations are then fixed by the source alone, so the
# single-problem and the batched kernels -- different code around the same
# device functions -- round identically.
# FP32 division and square root use the hardware reciprocal and square-root
# instructions (1 ulp) instead of the correctly rounded sequences; FP64
# arithmetic in these units (metric, sources, con2prim, RK) is unaffected.
MIXED_FLAGS = ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]
EXTRA = {"grhd_batch_mixed_a.hip": MIXED_FLAGS, "grhd_batch_mixed_b.hip": MIXED_FLAGS,
         **{f"grhd_stage_f32_{k}.hip": MIXED_FLAGS for k in KINDS}}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GRHD_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # The FP64 path keeps the reference's rounding sequence (no fused
    # multiply-add contraction) so that one RK3 step agrees with the reference
    # to round-off of the transcendental functions only.
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    # validity flags of the flat HLLC solver are combined with & on purpose
    "-Wno-bitwise-instead-of-logical",
]


def build(verbose: bool = True, jobs: int = 8) -> Path:
    objs = []
    procs = []
    for src in SOURCES:
        obj = CSRC / (Path(src).stem + ".o")
        cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), obj))
        objs.append(obj)
        if len(procs) >= jobs:
            p, _ = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError("hipcc failed")
    for p, _ in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        o.unlink(missing_ok=True)
    return LIB


FORTRAN = PKG / "fortran"
DRIVER = PKG / "bin" / "grhd2_hip"
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def build_driver(verbose: bool = True) -> Path:
    """Fortran driver (grhd_amd/bin/grhd2_hip) linked against libgrhd_hip.so."""
    out = DRIVER.parent
    obj = out / "_obj"
    obj.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in ["grhd_c.f90", "grhd_par.f90", "grhd_ckpt.f90", "grhd_vtk.f90", "grhd2_hip.f90"]:
        o = obj / (Path(src).stem + ".o")
        cmd = [FC, "-O2", f"-J{obj}", "-c", str(FORTRAN / src), "-o", str(o)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(o)
    cmd = [FC, "-O2", "-o", str(DRIVER), *map(str, objs), f"-L{PKG}", "-lgrhd_hip",
           "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return DRIVER


PROBE = PKG / "bin" / "valu_rate"


def build_probe(verbose: bool = True) -> Path:
    """VALU issue-rate probe (grhd_amd/bin/valu_rate), run by `bench.py --valu-rate`."""
    PROBE.parent.mkdir(parents=True, exist_ok=True)
    cmd = [HIPCC, *FLAGS, "-o", str(PROBE), str(CSRC / "valu_rate.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return PROBE


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
    build_driver(verbose="-q" not in sys.argv)
    build_probe(verbose="-q" not in sys.argv)
