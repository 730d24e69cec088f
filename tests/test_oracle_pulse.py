"""The oracle's Pulse time step (Domain::getMaximumFrequency, FE/Domain.C:816-946)
pinned by what the mathematics fixes, on the CPU.

* Material::getHookeMatrix (FE/Material.C:270-331) against the Lame form
  lambda + 2 mu, lambda, mu.
* Element stiffness matrices (FE/Element3D.C:242-333, FE/Element2D.C:129-222):
  symmetry, the rigid-body null space, and the strain energy of a uniform
  strain field, u^T K u = V eps^T C eps, which 2 x 2 (x 2) Gauss integrates
  exactly; V from closed forms (frustum volume, Pappus' theorem).  A dropped
  or misplaced term of B, C or the hoop row breaks one of them.
* The power iteration: repeated calls continue from the last iterate, so the
  value converges to an eigenvalue of the assembled M^-1 K (numpy), here the
  largest; the second call on an unchanged mesh stops after one iteration
  (fmax0 carried over).
"""
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
STEEL = dict(E=206e9, nu=0.3, rho=7830.0, A=350e6, B=275e6, n=0.36)


def lame(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return lam, mu


def one_element(dim, x):
    n = x.shape[0]
    conn = np.arange(n, dtype=np.int32)[None, :]
    mesh = dict(dim=dim, x=x, conn=conn, bc=np.zeros(n, np.uint8), v0=np.zeros((n, 3)))
    return O.Model(mesh, STEEL, dict(ts_method="Pulse"))


def test_hooke_matrices():
    lam, mu = lame(STEEL["E"], STEEL["nu"])
    C3 = one_element(3, np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1],
                                  [0, 1, 1]], float)).hooke()
    want = np.zeros((6, 6))
    want[:3, :3] = lam
    want[np.arange(3), np.arange(3)] = lam + 2 * mu
    want[np.arange(3, 6), np.arange(3, 6)] = mu
    np.testing.assert_allclose(C3, want, rtol=1e-14)
    C2 = one_element(2, np.array([[1, 0, 0], [2, 0, 0], [2, 1, 0], [1, 1, 0]], float)).hooke()
    want = np.full((4, 4), lam)
    want[2, :] = want[:, 2] = 0.0  # (rr, zz, rz, theta-theta)
    want[[0, 1, 3], [0, 1, 3]] = lam + 2 * mu
    want[2, 2] = mu
    np.testing.assert_allclose(C2, want, rtol=1e-14)


def frustum(h=0.7e-3, a=1.0e-3, b=0.6e-3):
    """square frustum: base a x a at z=0, top b x b centred above it (planar faces)"""
    o = (a - b) / 2
    x = np.array([[0, 0, 0], [a, 0, 0], [a, a, 0], [0, a, 0],
                  [o, o, h], [o + b, o, h], [o + b, o + b, h], [o, o + b, h]], float)
    return x, h / 3 * (a * a + b * b + a * b)


def voigt3(G):
    return np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])


def test_hex_stiffness_physics():
    rng = np.random.default_rng(3)
    x, V = frustum()
    # a rigid motion of the frustum changes nothing of the physics checks
    th = 0.3
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    x = x @ Rz.T + np.array([1e-3, -2e-3, 5e-4])
    o = one_element(3, x)
    K = o.element_stiffness(0)
    C = o.hooke()
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-13 * scale
    # rigid translations and infinitesimal rotations
    for d in range(3):
        u = np.zeros((8, 3))
        u[:, d] = 1.0
        assert np.abs(K @ u.reshape(-1)).max() <= 1e-12 * scale
    for W in (np.array([[0, -1, 0], [1, 0, 0], [0, 0, 0]]), np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0]]),
              np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0]])):
        u = x @ W.T
        assert np.abs(K @ u.reshape(-1)).max() <= 1e-12 * scale * np.abs(u).max()
    # uniform strain: u = G x, energy = V eps^T C eps (engineering shears)
    for _ in range(3):
        G = rng.normal(size=(3, 3)) * 1e-3
        u = (x @ G.T).reshape(-1)
        ev = voigt3(G)
        np.testing.assert_allclose(u @ K @ u, V * ev @ C @ ev, rtol=1e-12)
    w = np.linalg.eigvalsh((K + K.T) / 2)
    assert np.sum(np.abs(w) < 1e-10 * w.max()) == 6 and np.all(w > -1e-10 * w.max())


def test_axi_stiffness_physics():
    rng = np.random.default_rng(5)
    # a general quadrilateral section away from the axis
    x = np.array([[1.0e-3, 0.1e-3, 0], [2.1e-3, 0.0, 0], [2.0e-3, 1.2e-3, 0], [1.2e-3, 0.9e-3, 0]])
    o = one_element(2, x)
    K = o.element_stiffness(0)
    C = o.hooke()
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-13 * scale
    uz = np.zeros((4, 2))
    uz[:, 1] = 1.0
    assert np.abs(K @ uz.reshape(-1)).max() <= 1e-12 * scale
    ur = np.zeros((4, 2))
    ur[:, 0] = 1.0  # a radial shift strains the hoop direction
    assert np.abs(K @ ur.reshape(-1)).max() > 1e-3 * scale
    # Pappus: V = 2 pi r_centroid A
    r, z = x[:, 0], x[:, 1]
    cr = np.roll(r, -1)
    cz = np.roll(z, -1)
    cross = r * cz - cr * z
    A = cross.sum() / 2
    rc = ((r + cr) * cross).sum() / (6 * A)
    V = 2 * np.pi * rc * A
    for _ in range(3):
        a, c, d = rng.normal(size=3) * 1e-3
        # u_r = a r, u_z = c r + d z: eps_rr = a, eps_zz = d, gamma_rz = c, eps_tt = a
        u = np.stack([a * r, c * r + d * z], axis=1).reshape(-1)
        ev = np.array([a, d, c, a])
        np.testing.assert_allclose(u @ K @ u, V * ev @ C @ ev, rtol=1e-12)


def assembled(o, mesh):
    dim, nen = mesh["dim"], mesh["conn"].shape[1]
    nn = mesh["x"].shape[0]
    K = np.zeros((nn * dim, nn * dim))
    for e in range(mesh["conn"].shape[0]):
        idx = (mesh["conn"][e][:, None] * dim + np.arange(dim)[None, :]).reshape(-1)
        K[np.ix_(idx, idx)] += o.element_stiffness(e)
    m = np.repeat(o.get(O.F_MASS), dim)
    return np.sort(np.real(np.linalg.eigvals(K / m[:, None])))[::-1]


@pytest.mark.parametrize("name", ["hex_block", "taylor_axi_small"])
def test_power_iteration_converges_to_the_largest_eigenvalue(name):
    from redynela_amd import deck
    d = deck.load(os.path.join(HERE, "decks", name + ".dyn"))
    mesh, _, _ = d.model()
    o = O.from_deck(d)
    f1, it1 = o.max_frequency()
    f2, it2 = o.max_frequency()
    assert it1 > 3 and it2 == 1  # fmax0 and the iterate carry over
    assert abs(f2 - f1) / f1 < 1e-4
    lam = assembled(o, mesh)
    for _ in range(3000):
        f, _ = o.max_frequency()
    assert abs(f * f - lam[0]) / lam[0] < 1e-8
    # the first call stops within 1 % of it (the reference's 1e-4 test on successive fmax)
    assert abs(f1 * f1 - lam[0]) / lam[0] < 0.02


def hex_grid(n, nz, h, hz):
    """n x n x nz regular hex8 grid (ElHex8n3D node order), free"""
    idx, x, conn = {}, [], []
    for k in range(nz + 1):
        for j in range(n + 1):
            for i in range(n + 1):
                idx[(i, j, k)] = len(x)
                x.append((i * h, j * h, k * hz))
    for k in range(nz):
        for j in range(n):
            for i in range(n):
                c = [(i, j, k), (i + 1, j, k), (i + 1, j + 1, k), (i, j + 1, k),
                     (i, j, k + 1), (i + 1, j, k + 1), (i + 1, j + 1, k + 1), (i, j + 1, k + 1)]
                conn.append([idx[t] for t in c])
    x = np.array(x)
    return dict(dim=3, x=x, conn=np.array(conn, np.int32), bc=np.zeros(len(x), np.uint8), v0=np.zeros_like(x))


@pytest.mark.parametrize("n, nz, hz, lo, hi", [(1, 1, 1.0e-3, 1.30, 1.40), (4, 4, 1.0e-3, 1.02, 1.06),
                                               (4, 6, 0.4e-3, 0.99, 1.01), (4, 6, 0.5e-3, 0.99, 1.01)])
def test_hex8_courant_length_against_the_top_frequency(n, nz, hz, lo, hi):
    """the build-defined hex8 Courant length V / A_max against the true top
    frequency of the assembled mesh (eigenvalues of M^-1 K with lumped mass,
    numpy): regular meshes of cubes and of flat plates stay inside the 0.9
    safety factor of the time step; one lone cube does not"""
    mesh = hex_grid(n, nz, 1e-3, hz)
    o = O.Model(mesh, STEEL, dict(ts_method="Pulse"))
    o.compute_internal_matrices()
    f_courant = 2.0 / o.courant_crit()  # f_max = 2 / min(l / c), FE/Solver.C:220-267
    o.max_frequency()  # lumped mass
    ratio = np.sqrt(assembled(o, mesh)[0]) / f_courant
    assert lo < ratio < hi, ratio
    assert (ratio * 0.9 < 1.0) == (n > 1)
