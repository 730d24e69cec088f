"""The per-call boundary and the C++ facade on the GPU.

tests/capi/reference_loop (built from tests/capi/reference_loop.cpp against
include/dnlg_domain.hpp) runs the reference's ExplicitSolverCH loop on the
host in the oracle and calls Domain::computeInternalForces() on the GPU
through the facade, from the host's coordinates and Current displacement
increment, then the host corrector; it requires every increment's Fint,
coordinates and time to be the oracle's own doubles.  The Python checks here
drive dnlg_compute_internal_forces_from and dnlg_courant_time_step directly.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "capi", "reference_loop")
COPPER = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", ["axi", "axi-rho", "hex"])
def test_reference_loop_with_gpu_forces(case):
    assert os.path.exists(PROG), "tests/capi/reference_loop is not built (run __graft_entry__.build())"
    say("case: compiled host loop with GPU forces through the C++ facade:", case)
    r = subprocess.run([PROG, case], capture_output=True, text=True, timeout=240)
    say(r.stdout.strip())
    if r.stderr.strip():
        say(r.stderr.strip()[-2000:])
    assert r.returncode == 0, r.stdout[-2000:]
    assert "bit-equal to the oracle" in r.stdout and "MISMATCH" not in r.stdout


def test_state_in_forces_and_courant_from_python():
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: dnlg_compute_internal_forces_from with random increments, dnlg_courant_time_step, axi 6x20")
    mesh = R.mesh.taylor_axi(6, 20, 3.2e-3, 32.4e-3, 227.0)
    m = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    g = R.Domain(mesh, m, s)
    o = O.Model(mesh, COPPER, dict(rho_b=1.0), mathmode=1)
    rng = np.random.default_rng(7)
    x0 = np.asarray(mesh["x"], dtype=np.float64)
    for k in range(5):
        x = x0 + rng.uniform(-2e-6, 2e-6, x0.shape) * (k > 0)
        x[:, 2] = 0.0
        du = rng.uniform(-3e-6, 3e-6, x0.shape)
        du[:, 2] = 0.0
        o.set(O.F_X, x)
        o.set_du(du)
        o.compute_internal_matrices()
        assert o.compute_internal_forces() == 0
        fg = g.computeInternalForces(x, du).reshape(-1)
        assert np.array_equal(fg, o.get(O.F_FINT)), "call %d" % k
        assert g.computeCourantTimeStep() == o.courant_crit(), "call %d" % k
    # NULL coordinates keep the device copy
    fg2 = g.computeInternalForces(None, du).reshape(-1)
    o.compute_internal_forces()
    assert np.array_equal(fg2, o.get(O.F_FINT))
    assert np.array_equal(g.field(L.F_COORDS), x.reshape(-1))
    g.close()
