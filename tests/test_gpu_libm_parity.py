"""GPU FP64 path against the oracle in libm mode: glibc log / pow, as the
reference calls them (LA/Tensor2.C:1185-1194, FE/IsoHardElastoplastic.C:175-183).

The GPU uses a deterministic log / pow (the reference's glibc cannot run on a
GPU), so it is not bit-equal to a libm run.  Two checks at the tolerances of
the element pass and the whole run:

* per step (local): at every increment the GPU computes Domain::Fint from the
  libm oracle's own state of that increment -- coordinates, Current
  displacement increment (through dnlg_compute_internal_forces_from, the
  per-call boundary) and point state S, evp, p -- and must agree with the
  oracle's Fint to 1e-12 normwise (max |dF| / max |F|);
* trajectory: the GPU runs the whole loop itself; final displacements and
  equivalent plastic strains must agree with the libm run to 1e-9 normwise.
  The per-step Fint of the two trajectories is reported too: last-bit log
  differences seed a round-off drift of the state (1e-14 of the displacement),
  which the cancellation in the nodal sums shows as ~1e-11 of max |Fint|; it
  is bounded here at 1e-10.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

COPPER = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)
POWER = dict(E=115e9, nu=0.3, rho=8930.0, A=90e6, B=292e6, n=0.31)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_domain(mesh, mat):
    import redynela_amd as R
    m = R.Material(mat["E"], mat["nu"], mat["rho"], R.IsoHardElastoplastic(mat["A"], mat["B"], mat["n"]))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    return R.Domain(mesh, m, s)


def nrel(a, b):
    n = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / n if n > 0 else float(np.max(np.abs(a - b)))


def local_run(mesh, mat, up_time, max_inc, what):
    """GPU Fint from the libm oracle's state at every increment; returns (increments, worst)."""
    from redynela_amd import _lib as L
    g = gpu_domain(mesh, mat)
    o = O.Model(mesh, mat, dict(rho_b=1.0), mathmode=0)
    nip = 8 if mesh["dim"] == 3 else 4

    def push_points():
        S = o.get(O.F_S).reshape(-1, 9)
        g.set_field(L.F_S, np.ascontiguousarray(S.T))
        g.set_field(L.F_EVP, o.get(O.F_EVP))
        g.set_field(L.F_P, np.ascontiguousarray(o.get(O.F_P)[::nip]))

    assert o.begin_pre(up_time) == 1
    push_points()
    fg = g.computeInternalForces(o.get(O.F_X).reshape(-1, 3), o.get(O.F_DU).reshape(-1, 3)).reshape(-1)
    o.compute_internal_forces()
    worst = nrel(fg, o.get(O.F_FINT))
    o.begin_post()
    k = 0
    while k < max_inc:
        o.step_pre()
        push_points()
        fg = g.computeInternalForces(o.get(O.F_X).reshape(-1, 3), o.get(O.F_DU).reshape(-1, 3)).reshape(-1)
        o.compute_internal_forces()
        r = nrel(fg, o.get(O.F_FINT))
        assert r <= 1e-12, "%s: increment %d, GPU Fint from the libm state differs by %.3e" % (what, k + 1, r)
        worst = max(worst, r)
        k += 1
        if o.step_post() != 1:
            break
    g.close()
    say("%s: %d increments, worst per-step |dFint|/|Fint| = %.3e (bar 1e-12)" % (what, k, worst))
    return k, worst


def trajectory_run(mesh, mat, up_time, max_inc, what, drift_bound=1e-10):
    from redynela_amd import _lib as L
    g = gpu_domain(mesh, mat)
    o = O.Model(mesh, mat, dict(rho_b=1.0), mathmode=0)
    assert g.begin(up_time) and o.begin(up_time) == 1
    drift = nrel(g.field(L.F_FINT), o.get(O.F_FINT))
    k = 0
    while k < max_inc:
        rg = g.step(1)
        ro = o.step()
        k += 1
        assert rg == (ro == 1), "%s: loop flags differ at increment %d" % (what, k)
        drift = max(drift, nrel(g.field(L.F_FINT), o.get(O.F_FINT)))
        if not rg:
            break
    x0 = np.asarray(mesh["x"]).reshape(-1)
    du = nrel(g.field(L.F_COORDS) - x0, o.get(O.F_X) - x0)
    de = nrel(g.field(L.F_EVP), o.get(O.F_EVP))
    tg, to = g.time(), o.time()
    g.close()
    say("%s: %d increments, final displacement %.3e, final evp %.3e (bar 1e-9); trajectory Fint drift %.3e; "
        "t %.9e / %.9e" % (what, k, du, de, drift, tg["t"], to["t"]))
    assert tg["increment"] == to["increment"]
    assert du <= 1e-9 and de <= 1e-9
    assert drift <= drift_bound
    return du, de, drift


def test_axi_10x50_local_every_increment_to_80us():
    import redynela_amd as R
    say("case: axi 10x50, copper, local per-step parity against libm over the whole 80 us run")
    k, worst = local_run(R.mesh.taylor_axi_10x50(), COPPER, 80e-6, 10 ** 9, "axi 10x50 local")
    assert k == 9005


def test_axi_10x50_trajectory_to_80us():
    import redynela_amd as R
    say("case: axi 10x50, copper, GPU trajectory against the libm run to 80 us")
    trajectory_run(R.mesh.taylor_axi_10x50(), COPPER, 80e-6, 10 ** 9, "axi 10x50 trajectory")


def test_axi_20x100_power_law_local_4000_increments():
    import redynela_amd as R
    say("case: axi 20x100, 200 m/s, A=90 MPa B=292 MPa n=0.31, local per-step parity against libm, 4000 increments")
    mesh = R.mesh.taylor_axi(20, 100, 3.2e-3, 32.4e-3, 200.0)
    k, worst = local_run(mesh, POWER, 1e300, 4000, "axi 20x100 n=0.31 local")
    assert k == 4000


def test_axi_20x100_power_law_trajectory_4000_increments():
    import redynela_amd as R
    say("case: axi 20x100, 200 m/s, n=0.31, GPU trajectory against the libm run, 4000 increments")
    mesh = R.mesh.taylor_axi(20, 100, 3.2e-3, 32.4e-3, 200.0)
    trajectory_run(mesh, POWER, 1e300, 4000, "axi 20x100 n=0.31 trajectory")


def test_ogrid_1M_hex8_local_and_trajectory():
    import redynela_amd as R
    O.lib().dnlo_set_threads(16)
    say("case: hex8 O-grid 1M elements (2000/section x 500 layers), libm parity, preamble + 5 increments")
    mesh = R.mesh.ogrid_taylor_bar(500)
    local_run(mesh, COPPER, 1e300, 5, "hex8 1M local")
    trajectory_run(mesh, COPPER, 1e300, 5, "hex8 1M trajectory", drift_bound=1e-12)
