"""The oracle's linear algebra against the reference's own doubles.

tests/golden/la_golden.npz holds inputs and outputs of the reference's
Tensor2::polarDecomposeLnU (LA/Tensor2.C:1154-1259), Matrix::inverse
(LA/Matrix.C:1653-1723, LAPACK dgetrf/dgetri), Matrix::operator*
(LA/Matrix.C:644-671, BLAS dgemm) and Tensor2 helpers, produced by
tests/golden/make_la_golden.py from the reference sources.  With libm
math (mathmode 0, the reference's own log) the oracle must reproduce them
bit for bit.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "la_golden.npz"))


def test_polar_bit_exact_vs_reference():
    F = G["polar_F"]
    bad = 0
    for k in range(F.shape[0]):
        st, U, R = O.polar(F[k], mathmode=0)
        assert st == 0
        if not (np.array_equal(U, G["polar_lnU"][k]) and np.array_equal(R, G["polar_R"][k])):
            bad += 1
    assert bad == 0, "%d of %d polar decompositions differ from the reference" % (bad, F.shape[0])


def test_polar_det_mode_close_to_reference():
    # det mode changes only the log (deterministic own implementation): lnU moves by ulps
    F = G["polar_F"]
    for k in range(0, F.shape[0], 7):
        st, U, R = O.polar(F[k], mathmode=1)
        assert st == 0
        assert np.array_equal(R, G["polar_R"][k])
        assert np.max(np.abs(U - G["polar_lnU"][k])) <= 4e-16 * max(1.0, np.max(np.abs(U)))


def test_inverse_bit_exact_vs_reference():
    J = G["inv_J"]
    L = O.lib()
    for k in range(J.shape[0]):
        inv = np.zeros(9)
        assert L.dnlo_inv3(np.ascontiguousarray(J[k].reshape(9)), inv) == 0
        assert np.array_equal(inv.reshape(3, 3), G["inv_Jinv"][k]), k


def test_gemm_bit_exact_vs_reference():
    L = O.lib()
    for k in range(G["inv_J"].shape[0]):
        out = np.zeros(24)
        L.dnlo_gemm(8, 3, 3, np.ascontiguousarray(G["gemm_dN0"][k].reshape(-1)),
                    np.ascontiguousarray(G["inv_Jinv"][k].reshape(-1)), out)
        assert np.array_equal(out.reshape(8, 3), G["gemm_dN"][k]), k
    for k in range(G["gemm2_A"].shape[0]):
        out = np.zeros(8)
        L.dnlo_gemm(4, 2, 2, np.ascontiguousarray(G["gemm2_A"][k].reshape(-1)),
                    np.ascontiguousarray(G["gemm2_B"][k].reshape(-1)), out)
        assert np.array_equal(out.reshape(4, 2), G["gemm2_C"][k]), k


def test_tensor_helpers_bit_exact_vs_reference():
    L = O.lib()
    for k in range(G["t2_S"].shape[0]):
        S = np.ascontiguousarray(G["t2_S"][k].reshape(9))
        R = np.ascontiguousarray(G["t2_R"][k].reshape(9))
        E = np.ascontiguousarray(G["t2_E"][k].reshape(9))
        out = np.zeros(9)
        L.dnlo_t2_rfi(R, S, out)
        assert np.array_equal(out.reshape(3, 3), G["t2_RSRt"][k])
        L.dnlo_t2_deviator(E, out)
        assert np.array_equal(out.reshape(3, 3), G["t2_devE"][k])
        assert L.dnlo_t2_J2(S) == G["t2_J2"][k]
        assert L.dnlo_t2_norme(S) == G["t2_norme"][k]


def test_polar_against_eigendecomposition():
    # independent check of the mathematics: lnU = 1/2 ln(F^T F), R = F (F^T F)^-1/2
    rng = np.random.default_rng(7)
    for sc in (1e-4, 1e-2, 0.2):
        for _ in range(50):
            F = np.eye(3) + sc * rng.uniform(-1, 1, (3, 3))
            st, U, R = O.polar(F, mathmode=1)
            assert st == 0
            w, V = np.linalg.eigh(F.T @ F)
            U_ref = V @ np.diag(0.5 * np.log(w)) @ V.T
            R_ref = F @ V @ np.diag(w ** -0.5) @ V.T
            # the reference stops the Jacobi sweep at |C_pq| < 1e-9
            assert np.max(np.abs(U - U_ref)) < 2e-9
            assert np.max(np.abs(R - R_ref)) < 2e-9
            assert np.max(np.abs(R @ R.T - np.eye(3))) < 1e-8


def test_polar_pure_rotation_gives_zero_strain():
    c, s = np.cos(0.3), np.sin(0.3)
    Q = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    st, U, R = O.polar(Q, mathmode=1)
    assert st == 0
    assert np.max(np.abs(U)) < 1e-15
    assert np.max(np.abs(R - Q)) < 1e-15


def test_inverse_matches_numpy():
    rng = np.random.default_rng(3)
    L = O.lib()
    for _ in range(200):
        J = rng.uniform(-1, 1, (3, 3)) + 2 * np.eye(3)
        inv = np.zeros(9)
        assert L.dnlo_inv3(np.ascontiguousarray(J.reshape(9)), inv) == 0
        assert np.allclose(inv.reshape(3, 3) @ J, np.eye(3), atol=1e-13)
