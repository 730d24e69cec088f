"""Pulse time step on the GPU (Domain::getMaximumFrequency, FE/Domain.C:816-946)
against the oracle, bit for bit.

Each recomputing increment forms every element's stiffness matrix and runs the
power iteration; the frequency sets the time step, so any difference in the
matrices, the assembly order, the mass division or the 1/fmax scaling would
show in dt and, through it, in every coordinate.  Checked per step: time
scalars, coordinates, the number of power iterations of the last call and
fmax0.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
COPPER = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(mesh, mat, rho_b=0.8182, form="fp64"):
    import redynela_amd as R
    m = R.Material(mat["E"], mat["nu"], mat["rho"], R.IsoHardElastoplastic(mat["A"], mat["B"], mat["n"]))
    s = R.ExplicitSolverCH()  # Pulse, the reference's default for this solver
    s.setDissipation(rho_b)
    s.setPrecision(form)
    g = R.Domain(mesh, m, s)
    o = O.Model(mesh, mat, dict(rho_b=rho_b, ts_method="Pulse"), mathmode=1)
    return g, o


def same(a, b):
    return a.shape == b.shape and np.array_equal(a, b)


def check(g, o, what):
    from redynela_amd import _lib as L
    tg, to = g.time(), o.time()
    for k in ("t", "dt", "increment", "computed_dt"):
        assert tg[k] == to[k], "%s: %s %r vs %r" % (what, k, tg[k], to[k])
    assert same(g.field(L.F_COORDS), o.get(O.F_X)), what
    pi = g.pulse_info()
    assert pi["iterations"] == O.lib().dnlo_pulse_iterations(o.h), what
    assert pi["fmax0"] == O.lib().dnlo_pulse_fmax0(o.h), what


@pytest.mark.parametrize("kind", ["hex8", "axi"])
def test_pulse_per_step(kind):
    import redynela_amd as R
    say("case: Pulse time step per step, %s" % kind)
    if kind == "hex8":
        mesh = R.mesh.ogrid_bar(6, 3, 6, 3.2e-3, 4.0e-3, 0.4, 250.0)
    else:
        mesh = R.mesh.taylor_axi(6, 14, 3.2e-3, 9.0e-3, 250.0)
    g, o = pair(mesh, COPPER)
    assert g.begin(1.0)
    assert o.begin(1.0) == 1
    check(g, o, "preamble")
    assert g.pulse_info()["iterations"] > 3
    for k in range(45):  # four recomputations (every 10 increments)
        g.step(1)
        assert o.step() >= 0
        check(g, o, "step %d" % (k + 1))
    assert g.pulse_info()["calls"] == 5
    g.close()


def test_pulse_async_steps_and_solve():
    """steps issued without waiting still run the power iteration in order;
    solve() over several save intervals matches the oracle's solve"""
    import redynela_amd as R
    say("case: Pulse with step_async and solve")
    mesh = R.mesh.ogrid_bar(5, 3, 5, 3.2e-3, 3.5e-3, 0.4, 200.0)
    g, o = pair(mesh, COPPER)
    g.begin(1.0)
    o.begin(1.0)
    g.step_async(33)
    g.synchronize()
    for _ in range(33):
        o.step()
    check(g, o, "async 33")
    g.close()
    g, o = pair(mesh, COPPER, rho_b=1.0)
    for t in (1e-6, 2e-6, 3e-6):
        g.solve(t)
        assert o.solve(t) >= 0
        check(g, o, "solve to %g" % t)
    assert np.max(g.evp) > 0
    g.close()


def test_pulse_deck_result_files(tmp_path):
    """the hex deck names no method: Pulse, as the reference; result files byte-equal"""
    from redynela_amd import deck
    from redynela_amd import results as RF
    say("case: hex deck with the default Pulse method, result files against the oracle")
    d = deck.load(os.path.join(HERE, "decks", "hex_block.dyn"))
    assert d.solver.method == "Pulse"
    _, mat, _ = d.model()
    dom = d.run(results=str(tmp_path / "gpu" / "hex"))
    assert dom.pulse_info()["calls"] >= 2
    o = O.from_deck(d)
    paths = []

    def save(t):
        _, nodes, ips = O.result_blocks(o, mat.density, mat.init_temp)
        paths.append(str(tmp_path / ("hex.%03d.bin" % (len(paths) + 1))))
        RF.write(paths[-1], t, nodes, ips)

    def solve(t):
        assert o.solve(t) >= 0
        return o.time()["t"]

    deck.structure_solve(d.solver.stop, d.save, solve, save)
    gpu = sorted(f for f in os.listdir(tmp_path / "gpu") if f.endswith(".bin"))
    assert gpu == [os.path.basename(p) for p in paths]
    for gname, p in zip(gpu, paths):
        assert open(tmp_path / "gpu" / gname, "rb").read() == open(p, "rb").read(), gname
    dom.close()


def test_pulse_mixed_form_and_switch():
    """mixed form with Pulse runs close to FP64; switching Courant -> Pulse mid-run follows the oracle"""
    import redynela_amd as R
    say("case: Pulse in the mixed form; method switch")
    mesh = R.mesh.ogrid_bar(5, 3, 5, 3.2e-3, 3.5e-3, 0.4, 200.0)
    a, _ = pair(mesh, COPPER)
    b, _ = pair(mesh, COPPER, form="mixed")
    a.solve(2e-6)
    b.solve(2e-6)
    xa, xb = a.coords, b.coords
    assert np.max(np.abs(xa - xb)) / np.max(np.abs(xa)) < 1e-5
    a.close()
    b.close()
    g, o = pair(mesh, COPPER)
    g.solver.setTimeStepMethod("Courant")
    O.lib().dnlo_set_ts_method(o.h, 0)
    g.begin(1.0)
    o.begin(1.0)
    for _ in range(12):
        g.step(1)
        o.step()
    g.solver.setTimeStepMethod("Pulse")
    O.lib().dnlo_set_ts_method(o.h, 1)
    for k in range(25):
        g.step(1)
        o.step()
        check(g, o, "after switch, step %d" % (k + 1))
    g.close()


def test_pulse_rejected_by_adjoint():
    import redynela_amd as R
    from redynela_amd._lib import DnlgError
    say("case: the adjoint refuses the Pulse method")
    mesh = R.mesh.taylor_axi(4, 8, 3.2e-3, 6.0e-3, 200.0)
    g, _ = pair(mesh, COPPER)
    with pytest.raises(DnlgError, match="Courant time step"):
        R._lib.check(R._lib.lib().dnlg_adjoint_forward(g.h, 10, 5), g.h)
    g.close()
