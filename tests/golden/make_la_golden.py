"""Writes tests/golden/la_golden.npz from the reference's own linear algebra.

Runs oracle/_ref/ref_la_harness (built by `make -C oracle ref` from the
reference sources, in the container that holds /root/reference) and packs its
raw float64 outputs into one compressed fixture.  The fixture holds inputs and
the reference's outputs for Tensor2::polarDecomposeLnU, Matrix::inverse,
Matrix::operator* and the Tensor2 helpers of the stress update.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_la_harness")

SHAPES = {
    "polar_F": (-1, 3, 3), "polar_lnU": (-1, 3, 3), "polar_R": (-1, 3, 3),
    "inv_J": (-1, 3, 3), "inv_Jinv": (-1, 3, 3),
    "gemm_dN0": (-1, 8, 3), "gemm_dN": (-1, 8, 3),
    "gemm2_A": (-1, 4, 2), "gemm2_B": (-1, 2, 2), "gemm2_C": (-1, 4, 2),
    "t2_S": (-1, 3, 3), "t2_R": (-1, 3, 3), "t2_E": (-1, 3, 3), "t2_RSRt": (-1, 3, 3),
    "t2_devE": (-1, 3, 3), "t2_J2": (-1,), "t2_norme": (-1,),
}


def main():
    if not os.path.exists(HARNESS):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    with tempfile.TemporaryDirectory() as d:
        subprocess.check_call([HARNESS, d])
        arrs = {k: np.fromfile(os.path.join(d, k + ".bin"), dtype="<f8").reshape(s)
                for k, s in SHAPES.items()}
    out = os.path.join(HERE, "la_golden.npz")
    np.savez_compressed(out, **arrs)
    print("wrote", out, {k: v.shape for k, v in arrs.items()})


if __name__ == "__main__":
    sys.exit(main())
