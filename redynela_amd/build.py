"""Builds libdnlgpu.so (HIP for gfx950) in-tree with hipcc.

Usage: python -m redynela_amd.build  (or call build()).  Every call
recompiles.  The FP64 element translation unit is compiled with
-ffp-contract=off so that its arithmetic is the reference's, operation for
operation; every HIP unit disables packed FP32 instructions.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdnlgpu.so")
BUILD = os.path.join(HERE, "_build")

ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
UNITS = [
    ("dnlg_f64.hip", ["-ffp-contract=off"]),
    ("dnlg_f32.hip", ["-ffast-math", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
    ("dnlg_batch.hip", ["-ffp-contract=off"]),
    ("dnlg_pulse.hip", ["-ffp-contract=off"]),
    ("dnlg_history.hip", ["-ffp-contract=off"]),
    ("dnlg_batch_f32.hip", ["-ffp-contract=fast-honor-pragmas", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
    ("dnlg_api.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("dnlg_mesh.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("dnlg_batch_api.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("dnlg_adjoint.hip", ["-ffp-contract=fast-honor-pragmas"]),
    ("dnlg_adjoint_api.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("dnlg_noalloc.hip", ["-ffp-contract=off"]),
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
