"""The oracle's element and solver arithmetic pinned by what the mathematics fixes.

The reference ships no tests or golden data (its own test strategy is manual
runs of sample decks).  Each check below is a closed form, an invariant or a
special case that a plausible transcription mistake (a dropped term, a wrong
sign, a transposed index) would fail:
  * lumped mass of a box / a revolved rectangle = rho V      (Element::computeMassMatrix)
  * uniaxial Hencky strain: Sig_xx = (K + 4G/3) ln(1+e), nodal forces = -int grad(N) . Sig
  * linear-hardening radial return lands on the yield surface (FE/Element.C:283-348)
  * rigid rotation: no strain increment, stress rotated by R S R^T (FE/Element.C:383-389)
  * patch test: a linear increment field gives identical points on a distorted hexahedron
  * axisymmetric radial expansion: hoop strain u_r / r (FE/ElementAx.C:110-136)
  * Courant lengths of a rectangle (FE/ElQua4nAx.C:131-143) and of a cube (hex8 V/A_max)
  * Chung-Hulbert parameters (FE/ExplicitSolverCH.C:336-343)
  * sum of nodal internal forces = 0 (no body force)
"""
import numpy as np
import pytest

import oracle_lib as O

EL = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0, **{"yield": 0})
PL = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)
SG = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
              dtype=float)


def cube(h=1e-3, perturb=0.0, seed=0):
    x = (SG + 1.0) * 0.5 * h
    if perturb:
        x = x + np.random.default_rng(seed).uniform(-perturb, perturb, x.shape) * h
    return dict(dim=3, x=x, conn=np.arange(8, dtype=np.int32).reshape(1, 8), bc=np.zeros(8, np.uint8),
                v0=np.zeros((8, 3)))


def quad(a, b, r0=0.0, z0=0.0):
    x = np.array([[r0, z0, 0], [r0 + a, z0, 0], [r0 + a, z0 + b, 0], [r0, z0 + b, 0]], dtype=float)
    return dict(dim=2, x=x, conn=np.arange(4, dtype=np.int32).reshape(1, 4), bc=np.zeros(4, np.uint8),
                v0=np.zeros((4, 3)))


def constants(mat):
    E, nu = mat["E"], mat["nu"]
    return E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))


def test_shape_tables_are_the_trilinear_closed_forms():
    N = np.zeros((8, 8))
    dN = np.zeros((8, 8, 3))
    O.lib().dnlo_shape_tables(3, N.reshape(-1), dN.reshape(-1))
    g = SG / np.sqrt(3.0)
    for p in range(8):
        for k in range(8):
            f = (1 + SG[k] * g[p]) / 2.0
            assert abs(N[p, k] - f.prod()) < 1e-16
            for d in range(3):
                dd = SG[k, d] / 2.0 * np.prod([f[e] for e in range(3) if e != d])
                assert abs(dN[p, k, d] - dd) < 1e-16
    assert np.allclose(N.sum(axis=1), 1.0, atol=1e-15)
    N2 = np.zeros((4, 4))
    dN2 = np.zeros((4, 4, 2))
    O.lib().dnlo_shape_tables(2, N2.reshape(-1), dN2.reshape(-1))
    for p in range(4):
        for k in range(4):
            f = (1 + SG[k, :2] * g[p, :2]) / 2.0
            assert abs(N2[p, k] - f.prod()) < 1e-16
            assert abs(dN2[p, k, 0] - SG[k, 0] / 2 * f[1]) < 1e-16
            assert abs(dN2[p, k, 1] - SG[k, 1] / 2 * f[0]) < 1e-16


def test_lumped_mass_box_and_revolved_rectangle():
    m = O.Model(cube(2e-3, perturb=0.0), EL, {})
    assert m.begin(1.0) == 1
    M = m.get(O.F_MASS)
    V = 8e-9
    assert abs(M.sum() - 8930 * V) < 1e-14 * 8930 * V
    assert np.allclose(M, 8930 * V / 8, rtol=1e-14)
    m = O.Model(quad(1e-3, 2e-3, r0=1e-3), EL, {})
    m.begin(1.0)
    M = m.get(O.F_MASS)
    V = np.pi * ((2e-3) ** 2 - (1e-3) ** 2) * 2e-3
    assert abs(M.sum() - 8930 * V) < 1e-13 * 8930 * V


@pytest.mark.parametrize("eps", [1e-4, -3e-3])
def test_uniaxial_hencky_elastic(eps):
    h = 1e-3
    mesh = cube(h)
    m = O.Model(mesh, EL, {})
    du = np.zeros((8, 3))
    du[:, 0] = eps * mesh["x"][:, 0]
    m.set_du(du)
    assert m.compute_internal_forces() == 0
    K, G = constants(EL)
    le = np.log1p(eps)
    S = m.get(O.F_S).reshape(8, 9)
    p = m.get(O.F_P)
    Sig = m.get(O.F_SIG).reshape(8, 9)
    # ln U is taken of C = F^T F = (1+e)^2, whose rounding is 1e-16 absolute: ~1e-16/|e| relative
    tol = 20e-16 / abs(eps)
    assert np.allclose(S[:, 0], 2 * G * 2 / 3 * le, rtol=tol)
    assert np.allclose(S[:, 4], -2 * G / 3 * le, rtol=tol)
    assert np.allclose(p, K * le, rtol=tol)
    assert np.allclose(Sig[:, 0], (K + 4 * G / 3) * le, rtol=tol)
    assert np.max(np.abs(S[:, [1, 2, 3, 5, 6, 7]])) < 1e-9 * abs(S[0, 0])
    # nodal force of a uniform stress in a box: F_I = -int dN_I/dx Sig dV = -sign_x Sig_xx h^2/4
    F = m.get(O.F_FINT).reshape(8, 3)
    sxx, syy = Sig[0, 0], Sig[0, 4]
    assert np.allclose(F[:, 0], -SG[:, 0] * sxx * h * h / 4, rtol=1e-13)
    assert np.allclose(F[:, 1], -SG[:, 1] * syy * h * h / 4, rtol=1e-13)


def test_radial_return_lands_on_yield_surface():
    mesh = cube(1e-3)
    m = O.Model(mesh, PL, {})
    du = np.zeros((8, 3))
    du[:, 0] = 2e-2 * mesh["x"][:, 0]
    du[:, 1] = -1e-2 * mesh["x"][:, 1]
    m.set_du(du)
    assert m.compute_internal_forces() == 0
    S = m.get(O.F_S).reshape(8, 3, 3)
    evp = m.get(O.F_EVP)
    K, G = constants(PL)
    # trial deviator from the Hencky strain
    le = np.array([np.log1p(2e-2), np.log1p(-1e-2), 0.0])
    dev = le - le.mean()
    Str = 2 * G * dev
    ntr = np.linalg.norm(Str)
    sy0 = 400e6 + 100e6 * 1e-10
    gamma = (ntr - np.sqrt(2 / 3) * sy0) / (2 * G * (1 + 100e6 / (3 * G)))
    assert np.allclose(evp, 1e-10 + np.sqrt(2 / 3) * gamma, rtol=1e-9)
    for q in range(8):
        eq = np.sqrt(1.5) * np.linalg.norm(S[q])
        assert abs(eq - (400e6 + 100e6 * evp[q])) < 1e-8 * eq
        # radial: the returned deviator is parallel to the trial one
        assert np.allclose(np.diag(S[q]) / np.linalg.norm(S[q]), Str / ntr, atol=1e-9)


def test_rigid_rotation_rotates_stress():
    mesh = cube(1e-3)
    m = O.Model(mesh, EL, {})
    du = np.zeros((8, 3))
    du[:, 0] = 1e-3 * mesh["x"][:, 0]
    du[:, 2] = -2e-3 * mesh["x"][:, 2] + 5e-4 * mesh["x"][:, 1]
    m.set_du(du)
    m.compute_internal_forces()
    S1 = m.get(O.F_S).reshape(8, 3, 3)
    evp1 = m.get(O.F_EVP)
    th = 0.2
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    du = (mesh["x"] @ (Q - np.eye(3)).T)
    m.set_du(du)
    m.compute_internal_forces()
    S2 = m.get(O.F_S).reshape(8, 3, 3)
    for q in range(8):
        assert np.allclose(S2[q], Q @ S1[q] @ Q.T, atol=1e-9 * np.abs(S1[q]).max())
    assert np.array_equal(m.get(O.F_EVP), evp1)


def test_patch_test_distorted_hexahedron():
    mesh = cube(1e-3, perturb=0.12, seed=4)
    m = O.Model(mesh, PL, {})
    H = np.array([[3e-3, 1e-3, -2e-3], [0.5e-3, -4e-3, 1e-3], [2e-3, 0, 1e-3]])
    m.set_du(mesh["x"] @ H.T)
    m.compute_internal_forces()
    S = m.get(O.F_S).reshape(8, 9)
    assert np.max(np.abs(S - S[0])) < 1e-10 * np.abs(S[0]).max()
    F = m.get(O.F_FINT).reshape(8, 3)
    assert np.max(np.abs(F.sum(axis=0))) < 1e-12 * np.abs(F).max()


def test_axisymmetric_radial_expansion():
    mesh = quad(1e-3, 1e-3, r0=2e-3)
    m = O.Model(mesh, EL, {})
    eps = 1e-3
    du = np.zeros((4, 3))
    du[:, 0] = eps * mesh["x"][:, 0]
    m.set_du(du)
    m.compute_internal_forces()
    K, G = constants(EL)
    le = np.log1p(eps)
    S = m.get(O.F_S).reshape(4, 9)
    assert np.allclose(S[:, 0], 2 * G * le / 3, rtol=1e-12)   # (2 le - le*2/3*... ) deviator: le - 2le/3
    assert np.allclose(S[:, 8], 2 * G * le / 3, rtol=1e-12)
    assert np.allclose(S[:, 4], -2 * G * 2 * le / 3, rtol=1e-12)
    assert np.allclose(m.get(O.F_P), K * 2 * le, rtol=1e-12)


def test_courant_lengths():
    a, b = 1e-3, 3e-3
    m = O.Model(quad(a, b, r0=1e-3), EL, {})
    assert abs(O.lib().dnlo_element_length(m.h, 0) - a * b / np.hypot(a, b)) < 1e-18
    m = O.Model(cube(2e-3), EL, {})
    assert abs(O.lib().dnlo_element_length(m.h, 0) - 2e-3) < 1e-18
    # cube stretched to 2h along x: V = 2h^3, A_max = 2h^2 -> h
    mesh = cube(1e-3)
    mesh["x"][:, 0] *= 2
    m = O.Model(mesh, EL, {})
    assert abs(O.lib().dnlo_element_length(m.h, 0) - 1e-3) < 1e-18


@pytest.mark.parametrize("rb", [1.0, 0.8182, 0.5, 0.0])
def test_chung_hulbert_parameters(rb):
    m = O.Model(cube(), EL, dict(rho_b=rb))
    t = m.time()
    aM = (2 * rb - 1) / (1 + rb)
    assert abs(t["alphaM"] - aM) < 1e-15
    assert abs(t["beta"] - (5 - 3 * rb) / ((1 + rb) ** 2 * (2 - rb))) < 1e-15
    assert abs(t["gamma"] - (1.5 - aM)) < 1e-15
    om = np.sqrt(12 * (1 + rb) ** 3 * (2 - rb) / (10 + 15 * rb - rb ** 2 + rb ** 3 - rb ** 4))
    assert abs(t["omegaS"] - om) < 1e-15
    if rb == 1.0:
        assert (t["alphaM"], t["beta"], t["gamma"], t["omegaS"]) == (0.5, 0.5, 1.0, 2.0)


def test_internal_forces_sum_to_zero_3d():
    import redynela_amd as R
    mesh = R.mesh.ogrid_bar(6, 3, 3, 3.2e-3, 2.0e-3, 0.4, 0.0)
    m = O.Model(mesh, PL, {})
    rng = np.random.default_rng(5)
    m.set_du(rng.uniform(-1, 1, mesh["x"].shape) * 3e-6)
    m.compute_internal_forces()
    F = m.get(O.F_FINT).reshape(-1, 3)
    assert np.max(np.abs(F.sum(axis=0))) < 1e-11 * np.abs(F).max()


def _jc_model(C=0.025, tq=0.9, n=0.31):
    mat = dict(E=115e9, nu=0.3, rho=8930.0, A=90e6, B=292e6, n=n,
               jc=dict(C=C, epsd0=1.0, m=1.09, T0=293.0, Tm=1356.0, cp=383.0, tq=tq))
    return O.Model(cube(1e-3), mat, dict(rho_b=1.0)), mat


def test_johnson_cook_extension_lands_on_its_yield_surface():
    """One uniaxial-strain increment at a known rate: the returned state satisfies
    sqrt(3/2)|S| = (A + B evp^n)(1 + C ln(depsp/dt/epsd0)) theta(T0) and the point
    heats by chi sigma_y depsp / (rho cp)."""
    m, mat = _jc_model()
    j = mat["jc"]
    L = O.lib()
    assert m.begin(1.0) == 1                # sets a Courant time step
    dt = m.time()["dt"]
    du = np.zeros((8, 3))
    du[:, 0] = 1.5e-2 * cube(1e-3)["x"][:, 0]
    m.set_du(du)
    assert m.compute_internal_forces() == 0
    S = m.get(O.F_S).reshape(8, 3, 3)
    evp = m.get(O.F_EVP)
    T = m.get(O.F_TEMP)
    for q in range(8):
        dep = evp[q] - 1e-10
        assert dep > 0
        rate = 1.0 + j["C"] * np.log(dep / dt / j["epsd0"])
        sy = (mat["A"] + mat["B"] * evp[q] ** mat["n"]) * rate  # theta(T0) = 1
        eq = np.sqrt(1.5) * np.linalg.norm(S[q])
        assert abs(eq - sy) < 1e-8 * sy
        assert abs((T[q] - 293.0) - j["tq"] * sy * dep / (8930.0 * 383.0)) < 1e-9 * (T[q] - 293.0)


def test_johnson_cook_extension_without_rate_and_heat_is_bitwise_the_reference_law():
    a, _ = _jc_model(C=0.0, tq=0.0)
    mat = dict(E=115e9, nu=0.3, rho=8930.0, A=90e6, B=292e6, n=0.31)
    b = O.Model(cube(1e-3), mat, dict(rho_b=1.0))
    du = np.zeros((8, 3))
    du[:, 0] = 1.5e-2 * cube(1e-3)["x"][:, 0]
    for m in (a, b):
        m.set_du(du)
        m.compute_internal_forces()
    assert np.array_equal(a.get(O.F_S), b.get(O.F_S))
    assert np.array_equal(a.get(O.F_EVP), b.get(O.F_EVP))
