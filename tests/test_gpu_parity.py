"""GPU FP64 path against the oracle, through the C ABI.

The oracle (oracle/liboracle.so, det math mode: the same deterministic log/pow
as the GPU) restates the reference's arithmetic operation for operation, so
the bar here is bit equality: per-step nodal internal forces, coordinates,
equivalent plastic strain, stresses and time scalars must be identical
doubles.  (Tolerances of 1e-12 per step on Fint and 1e-9 on the final
state are implied.)
"""
import hashlib
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

COPPER = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)
JC = dict(E=115e9, nu=0.3, rho=8930.0, A=90e6, B=292e6, n=0.31)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(mesh, mat, rho_b=1.0, track=False):
    import redynela_amd as R
    m = R.Material(mat["E"], mat["nu"], mat["rho"], R.IsoHardElastoplastic(mat["A"], mat["B"], mat["n"]))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(rho_b)
    g = R.Domain(mesh, m, s, track_strain=track)
    o = O.Model(mesh, mat, dict(rho_b=rho_b), mathmode=1)
    return g, o


def same(a, b):
    return a.shape == b.shape and np.array_equal(a, b)


def tens(g, field):
    """GPU point tensors are component-major [9][n_points]; the oracle's are [n_points][9]."""
    return g.field(field).reshape(9, -1).T.reshape(-1)


def rel(a, b):
    d = np.max(np.abs(a - b))
    n = max(np.max(np.abs(b)), 1e-300)
    return d / n


def compare_state(g, o, what):
    from redynela_amd import _lib as L
    xg, xo = g.field(L.F_COORDS), o.get(O.F_X)
    eg, eo = g.field(L.F_EVP), o.get(O.F_EVP)
    fg, fo = g.field(L.F_FINT), o.get(O.F_FINT)
    assert same(fg, fo), "%s: Fint differs, rel %.3e" % (what, rel(fg, fo))
    assert same(xg, xo), "%s: coords differ, rel %.3e" % (what, rel(xg, xo))
    assert same(eg, eo), "%s: evp differs, rel %.3e" % (what, rel(eg, eo))


def test_axi_10x50_per_step_and_final():
    import redynela_amd as R
    say("case: axi 10x50 FP64, rho_b=1, per-step bit parity for 300 steps then to 80 us")
    mesh = R.mesh.taylor_axi_10x50()
    g, o = make_pair(mesh, COPPER)
    assert g.begin(80e-6) and o.begin(80e-6) == 1
    compare_state(g, o, "preamble")
    from redynela_amd import _lib as L
    assert same(g.field(L.F_MASS), o.get(O.F_MASS))  # Node::mass
    for k in range(300):
        rg = g.step(1)
        ro = o.step()
        compare_state(g, o, "step %d" % (k + 1))
        tg, to = g.time(), o.time()
        assert tg["t"] == to["t"] and tg["dt"] == to["dt"] and tg["increment"] == to["increment"]
        assert rg == (ro == 1)
    say("  per-step ok; running to the end")
    while g.step(64):
        pass
    while o.step() == 1:
        pass
    compare_state(g, o, "final")
    tg, to = g.time(), o.time()
    assert tg["t"] == to["t"] == 80e-6 and tg["increment"] == to["increment"]
    from redynela_amd import _lib as L
    assert same(tens(g, L.F_S), o.get(O.F_S))
    assert same(g.field(L.F_V), o.get(O.F_V))
    say("  final ok, increments", tg["increment"])


def test_axi_10x50_generalized_alpha_rho_08182():
    import redynela_amd as R
    say("case: axi 10x50, rho_b=0.8182 (reference default), 400 steps")
    mesh = R.mesh.taylor_axi_10x50()
    g, o = make_pair(mesh, COPPER, rho_b=0.8182)
    g.begin(80e-6)
    o.begin(80e-6)
    g.step(400)
    for _ in range(400):
        o.step()
    compare_state(g, o, "400 steps")
    assert g.time()["dt"] == o.time()["dt"]


def test_hex8_ogrid_per_step():
    import redynela_amd as R
    say("case: hex8 O-grid 8x4 x 6 layers, 300 m/s, per-step bit parity for 200 steps")
    mesh = R.mesh.ogrid_bar(8, 4, 6, 3.2e-3, 3.9e-3, 0.4, 300.0)
    g, o = make_pair(mesh, COPPER)
    g.begin(1.0)
    o.begin(1.0)
    compare_state(g, o, "preamble")
    for k in range(200):
        g.step(1)
        o.step()
        compare_state(g, o, "step %d" % (k + 1))
    from redynela_amd import _lib as L
    assert same(tens(g, L.F_S), o.get(O.F_S))
    assert np.max(o.get(O.F_EVP)) > 1e-3  # plastic flow happened


def test_hex8_johnson_cook_power_law():
    import redynela_amd as R
    say("case: hex8 O-grid 6x3 x 4 layers, A+B evp^0.31, 150 steps")
    mesh = R.mesh.ogrid_bar(6, 3, 4, 3.2e-3, 2.6e-3, 0.4, 250.0)
    g, o = make_pair(mesh, JC)
    g.begin(1.0)
    o.begin(1.0)
    g.step(150)
    for _ in range(150):
        o.step()
    compare_state(g, o, "JC 150 steps")
    assert np.max(o.get(O.F_EVP)) > 1e-3


def test_compute_internal_forces_random_increment():
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: Domain::computeInternalForces with a random increment, hex8 and axi")
    rng = np.random.default_rng(11)
    for mesh in (R.mesh.ogrid_bar(6, 3, 3, 3.2e-3, 2.0e-3, 0.4, 0.0), R.mesh.taylor_axi(6, 12)):
        g, o = make_pair(mesh, COPPER, track=True)
        du = rng.uniform(-1, 1, mesh["x"].shape) * 2e-6
        if mesh["dim"] == 2:
            du[:, 2] = 0.0
        g.set_field(L.F_DU, du)
        o.set_du(du)
        for rep in range(3):
            fg = g.computeInternalForces()
            assert o.compute_internal_forces() == 0
            assert same(fg.reshape(-1), o.get(O.F_FINT)), "rep %d" % rep
        assert same(tens(g, L.F_S), o.get(O.F_S))
        assert same(tens(g, L.F_EPS), o.get(O.F_EPS))
        assert same(tens(g, L.F_EPSPLAS), o.get(O.F_EPSPLAS))
        assert same(tens(g, L.F_DEPS), o.get(O.F_DEPS))
        assert same(g.field(L.F_P), o.get(O.F_P)[::g.nip])


def test_bit_reproducible_runs():
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: three identical hex8 runs hash equal")
    mesh = R.mesh.ogrid_bar(10, 5, 8, 3.2e-3, 5.2e-3, 0.4, 250.0)
    hashes = []
    for rep in range(3):
        g, _ = make_pair(mesh, COPPER)
        g.solve(2e-6)
        h = hashlib.sha256()
        for f in (L.F_COORDS, L.F_V, L.F_S, L.F_EVP, L.F_FINT):
            h.update(g.field(f).tobytes())
        hashes.append(h.hexdigest())
        g.close()
    assert hashes[0] == hashes[1] == hashes[2]


def test_solve_matches_oracle_solve():
    import redynela_amd as R
    say("case: Domain.solve(upTime) == oracle solve, axi 8x40 at 150 m/s to 30 us")
    mesh = R.mesh.taylor_axi(8, 40, 3.2e-3, 32.4e-3, 150.0)
    g, o = make_pair(mesh, COPPER)
    ng = g.solve(30e-6)
    no = o.solve(30e-6)
    assert ng == no
    compare_state(g, o, "solve")


def test_error_negative_jacobian_reported():
    import redynela_amd as R
    say("case: inverted axisymmetric element raises NEG_DETJ")
    mesh = R.mesh.taylor_axi(2, 2)
    mesh["conn"] = np.ascontiguousarray(mesh["conn"][:, [0, 3, 2, 1]])  # clockwise: detJ < 0
    m = R.Material(COPPER["E"], COPPER["nu"], COPPER["rho"], R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    g = R.Domain(mesh, m, s)
    with pytest.raises(R.DnlgError, match="NEG_DETJ"):
        g.begin(1e-6)


def test_element_error_mid_run_stops_like_the_oracle():
    """a negative Jacobian in mid-run: nothing runs after the element pass
    that finds it, so time, coordinates and velocities equal the oracle's,
    whether the steps were issued one at a time or as one batch of 50.  The
    reference finds a 2D negative Jacobian in computeInternalMatrices at the
    end of the previous increment (FE/Element2D.C, after endStep); the GPU
    finds it in the fused element pass, after the next initStep, so its
    increment count is one higher"""
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: axi bar at 20 km/s inverts in its first increments; GPU state after the error vs the oracle")
    mesh = R.mesh.taylor_axi(4, 20, 3.2e-3, 32.4e-3, 2.0e4)
    mat = dict(COPPER, A=400e6, B=100e6, n=1.0)

    def gpu():
        m = R.Material(mat["E"], mat["nu"], mat["rho"], R.IsoHardElastoplastic(mat["A"], mat["B"], mat["n"]))
        s = R.ExplicitSolverCH()
        s.setTimeStepMethod("Courant")
        s.setDissipation(1.0)
        g = R.Domain(mesh, m, s)
        g.begin(1.0)
        return g

    o = O.Model(mesh, mat, dict(rho_b=1.0), mathmode=1)
    assert o.begin(1.0) == 1
    k = 0
    while True:
        r = o.step()
        k += 1
        if r < 0:
            break
        assert r == 1 and k < 40
    assert r == -4  # the oracle's negative-Jacobian code (FE/Element2D.C:101-109)
    g1 = gpu()
    with pytest.raises(R.DnlgError, match="NEG_DETJ"):
        for _ in range(40):
            g1.step(1)
    g2 = gpu()
    with pytest.raises(R.DnlgError, match="NEG_DETJ"):
        g2.step(50)
    to = o.time()
    for g in (g1, g2):
        tg = g.time()
        assert tg["t"] == to["t"] and tg["increment"] == to["increment"] + 1, (tg, to)
        assert np.array_equal(g.field(L.F_COORDS), o.get(O.F_X))
        assert np.array_equal(g.field(L.F_V), o.get(O.F_V))
    say("  error at increment %d, t = %.6e" % (to["increment"], to["t"]))
    g1.close()
    g2.close()
