"""Whole-run properties of the oracle's ExplicitSolverCH restatement.

  * the Chung-Hulbert update with rho_b = 1 is the central-difference scheme
    (FE/ExplicitSolverCH.C:387-545): a_{n+1} = 2 M^-1 f - a_n,
    v_{n+1} = v_n + dt a_{n+1}, du_{n+1} = dt v_n + dt^2/2 a_{n+1}
  * the time-step control of FE/Solver.C:220-267: a Courant recompute every
    10 increments and linear extrapolation in between
  * the 3D O-grid bar keeps the fourfold symmetry of its mesh
  * runs are bit-reproducible; det and libm math agree to rounding
  * the axisymmetric Taylor bar ends where the classic copper Taylor test
    does (final length and mushroom radius of the 10x50 model)
"""
import numpy as np

import oracle_lib as O

CU = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)


def test_central_difference_identity_and_timestep_control():
    import redynela_amd as R
    mesh = R.mesh.taylor_axi(6, 30, 3.2e-3, 32.4e-3, 227.0)
    m = O.Model(mesh, CU, dict(rho_b=1.0))
    assert m.begin(80e-6) == 1
    dts, ts = [], []
    for k in range(60):
        v0, a0, t0 = m.get(O.F_V), m.get(O.F_A), m.time()
        ts.append(t0["t"])
        x0 = m.get(O.F_X)
        assert m.step() == 1
        dt = t0["dt"]  # the step just taken used the dt set after the previous increment
        dts.append(dt)
        v1, a1, du1 = m.get(O.F_V), m.get(O.F_A), m.get(O.F_DU)
        f = m.get(O.F_FINT).reshape(-1, 2)
        M = m.get(O.F_MASS)
        acc = np.zeros_like(v0).reshape(-1, 3)
        acc[:, :2] = f / M[:, None]
        free = mesh["bc"] == 0
        sel = np.repeat(free, 3)
        a_exp = (2 * acc.reshape(-1) - a0)
        assert np.allclose(a1[sel], a_exp[sel], rtol=1e-12, atol=1e-9 * np.abs(a_exp).max())
        assert np.allclose(v1[sel], (v0 + dt * a1)[sel], rtol=1e-13, atol=1e-12)
        assert np.allclose(du1[sel], (dt * v0 + 0.5 * dt * dt * a1)[sel], rtol=1e-12, atol=1e-20)
        assert np.allclose(m.get(O.F_X), x0 + du1, rtol=0, atol=1e-18)
        # restrained components stay at rest
        assert np.all(du1.reshape(-1, 3)[mesh["bc"] & 2 > 0, 1] == 0.0)
        assert np.all(v1.reshape(-1, 3)[mesh["bc"] & 1 > 0, 0] == 0.0)
    # time-step control: recompute at increments 10, 20, ... ; in between, linear extrapolation
    # dts[i] is the step size of increment i+1, ts[i] the time before it
    assert all(d == dts[0] for d in dts[:10])
    for i in range(11, 60):
        mm, r = divmod(i, 10)
        if r == 0:
            continue
        lo, hi = 10 * (mm - 1), 10 * mm
        lam = (ts[i] - ts[lo]) / (ts[hi] - ts[lo])
        assert dts[i] == (lam * dts[hi] + (1.0 - lam) * dts[lo])  # LINEARdbl, BT/Math.h:91


def test_oracle_runs_are_bit_reproducible():
    import redynela_amd as R
    mesh = R.mesh.ogrid_bar(6, 3, 4, 3.2e-3, 2.6e-3, 0.4, 250.0)
    outs = []
    for _ in range(2):
        m = O.Model(mesh, CU, dict(rho_b=1.0))
        m.begin(1.0)
        for _ in range(80):
            m.step()
        outs.append((m.get(O.F_X), m.get(O.F_S), m.get(O.F_EVP)))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_ogrid_fourfold_symmetry():
    import redynela_amd as R
    mesh = R.mesh.ogrid_bar(6, 3, 4, 3.2e-3, 2.6e-3, 0.4, 250.0)
    x0 = mesh["x"]
    m = O.Model(mesh, CU, dict(rho_b=1.0))
    m.begin(1.0)
    for _ in range(120):
        m.step()
    x = m.get(O.F_X).reshape(-1, 3)
    assert np.max(m.get(O.F_EVP)) > 1e-3
    # node images under a 90 degree rotation about z
    rot = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    key = {tuple(np.round(p * 1e9).astype(np.int64)): i for i, p in enumerate(x0)}
    img = np.array([key[tuple(np.round((rot @ p) * 1e9).astype(np.int64))] for p in x0])
    scale = np.abs(x - x0).max()
    assert np.max(np.abs(x[img] - x @ rot.T)) < 1e-9 * scale


def test_det_and_libm_math_agree_on_taylor_axi_10x50():
    import redynela_amd as R
    mesh = R.mesh.taylor_axi_10x50()
    res = []
    for mode in (0, 1):
        m = O.Model(mesh, CU, dict(rho_b=1.0), mathmode=mode)
        n = m.solve(20e-6)
        res.append((n, m.get(O.F_X), m.get(O.F_EVP)))
    assert res[0][0] == res[1][0]
    assert np.max(np.abs(res[0][1] - res[1][1])) <= 1e-12 * np.abs(res[0][1]).max()
    assert np.max(np.abs(res[0][2] - res[1][2])) <= 1e-9 * res[0][2].max()


def test_taylor_axi_10x50_final_shape_and_mass():
    import redynela_amd as R
    mesh = R.mesh.taylor_axi_10x50()
    m = O.Model(mesh, CU, dict(rho_b=1.0), mathmode=1, ke=True)
    m.begin(80e-6)
    M = m.get(O.F_MASS)
    V = np.pi * 3.2e-3 ** 2 * 32.4e-3
    assert abs(M.sum() - 8930 * V) < 1e-12 * 8930 * V
    ke0 = 0.5 * M.sum() * 227.0 ** 2
    while m.step() == 1:
        pass
    t = m.time()
    assert t["t"] == 80e-6
    x = m.get(O.F_X).reshape(-1, 3)
    # classic copper Taylor test (227 m/s, 32.4 x 6.4 mm): ~21.4 mm long, ~7 mm mushroom radius
    assert 0.0205 < x[:, 1].max() < 0.0222
    assert 0.0065 < x[:11, 0].max() < 0.0076
    assert 2.5 < m.get(O.F_EVP).max() < 3.8
    assert 0 <= t["ke"] < 0.05 * ke0
