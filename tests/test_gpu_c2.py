"""GPU at the precision workload's full size: the hex8 O-grid bar of 4,000
elements per section (40 x 40 core + 4 rings of 40 x 15) x 1,000 layers
(4,000,000 elements, 4,085,081 nodes), 250 m/s, Johnson-Cook with the rate
term and adiabatic heating (A = 90 MPa, B = 292 MPa, n = 0.31, C = 0.025,
epsdot0 = 1/s, m = 1.09, T0 = 293 K, Tm = 1356 K, cp = 383 J/kg/K,
Taylor-Quinney 0.9).

* FP64 against the oracle (the same law, det math): the preamble and three
  increments, Fint, coordinates, evp and temperature bit for bit;
* the mixed form against FP64 over the same 1,000 increments: final length,
  mushroom radius, peak plastic strain and the displacement field within
  1e-4 relative.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

JCMAT = dict(E=115e9, nu=0.3, rho=8930.0, A=90e6, B=292e6, n=0.31,
             jc=dict(C=0.025, epsd0=1.0, m=1.09, T0=293.0, Tm=1356.0, cp=383.0, tq=0.9))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def c2_mesh():
    import redynela_amd as R
    return R.mesh.ogrid_bar(40, 15, 1000, 3.2e-3, 32.4e-3, 0.4, 250.0)


def gpu_domain(mesh, form="fp64"):
    import redynela_amd as R
    j = JCMAT["jc"]
    h = R.IsoHardJohnsonCook(JCMAT["A"], JCMAT["B"], JCMAT["n"], C=j["C"], m=j["m"], depsp0=j["epsd0"], Tm=j["Tm"],
                             T0=j["T0"])
    h.enableRateAndThermal(j["cp"], j["tq"])
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    s.setPrecision(form)
    return R.Domain(mesh, R.Material(JCMAT["E"], JCMAT["nu"], JCMAT["rho"], h), s)


def shape(mesh, x, evp):
    base = mesh["x"][:, 2] == 0.0
    return x[:, 2].max(), np.hypot(x[base, 0], x[base, 1]).max(), evp.max()


def test_c2_fp64_full_size_bit_parity():
    from redynela_amd import _lib as L
    say("case: C2 hex8 O-grid 4,000/section x 1,000 layers, JC rate + heating, FP64 preamble + 3 increments vs oracle")
    mesh = c2_mesh()
    assert mesh["conn"].shape[0] == 4000000
    O.lib().dnlo_set_threads(16)
    o = O.Model(mesh, JCMAT, dict(rho_b=1.0), mathmode=1)
    g = gpu_domain(mesh)
    g.set_option(L.OPT_STORE_FINT, 1)
    g.begin(1.0)
    o.begin(1.0)
    for k in range(4):
        if k:
            g.step(1)
            o.step()
        for fg, fo, name in ((L.F_FINT, O.F_FINT, "Fint"), (L.F_COORDS, O.F_X, "coords"), (L.F_EVP, O.F_EVP, "evp"),
                             (L.F_TEMP, O.F_TEMP, "temperature")):
            a, b = g.field(fg), o.get(fo)
            assert np.array_equal(a, b), "increment %d: %s differs (max %.3e)" % (k, name, np.max(np.abs(a - b)))
        say("  increment %d: Fint, coords, evp, temperature identical" % k)
    tg, to = g.time(), o.time()
    assert tg["dt"] == to["dt"] and tg["t"] == to["t"]
    assert g.evp.max() > 1e-3  # the impact face is plastic
    g.close()


def test_c2_mixed_form_against_fp64_1000_increments():
    say("case: C2, mixed form against FP64 over 1,000 increments")
    mesh = c2_mesh()
    res = {}
    for form in ("fp64", "mixed"):
        d = gpu_domain(mesh, form)
        d.begin(1.0)
        d.step(1000)
        tm = d.time()
        res[form] = (d.coords, d.evp, tm["t"])
        say("  %s: t = %.6e s, length %.9e, radius %.9e, max evp %.6e" % ((form, tm["t"]) + shape(mesh, d.coords, d.evp)))
        d.close()
    (x0, e0, t0), (x1, e1, t1) = res["fp64"], res["mixed"]
    a, b = shape(mesh, x0, e0), shape(mesh, x1, e1)
    err = [abs(q - p) / abs(p) for p, q in zip(a, b)]
    u0, u1 = x0 - mesh["x"], x1 - mesh["x"]
    du = np.abs(u1 - u0).max() / np.abs(u0).max()
    say("  relative error: length %.3e, radius %.3e, max evp %.3e, displacement field %.3e, time %.3e"
        % (err[0], err[1], err[2], du, abs(t1 - t0) / t0))
    assert max(err) <= 1e-4 and du <= 1e-4
