"""GPU: the Johnson-Cook rate/thermal extension and the mixed-precision form.

* hardening = 2 (rate and adiabatic-heating factors documented in
  FE/IsoHardJohnsonCook.h:51-68, not evaluated by the reference) in the FP64
  form is bit-identical to the oracle's restatement of the same law, and with
  C = 0 and no heating it is bit-identical to the reference law;
* the mixed form (FP32 element pass, FP64 nodal state) tracks the FP64 form:
  final length, mushroom radius and peak plastic strain within the stated
  tolerances.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

JCMAT = dict(E=115e9, nu=0.3, rho=8930.0, A=90e6, B=292e6, n=0.31,
             jc=dict(C=0.025, epsd0=1.0, m=1.09, T0=293.0, Tm=1356.0, cp=383.0, tq=0.9))
CU = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_domain(mesh, mat, form="fp64", rho_b=1.0):
    import redynela_amd as R
    if "jc" in mat:
        j = mat["jc"]
        h = R.IsoHardJohnsonCook(mat["A"], mat["B"], mat["n"], C=j["C"], m=j["m"], depsp0=j["epsd0"], Tm=j["Tm"],
                                 T0=j["T0"])
        h.enableRateAndThermal(j["cp"], j["tq"])
    else:
        h = R.IsoHardElastoplastic(mat["A"], mat["B"], mat["n"])
    m = R.Material(mat["E"], mat["nu"], mat["rho"], h)
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(rho_b)
    s.setPrecision(form)
    return R.Domain(mesh, m, s)


def test_johnson_cook_rate_thermal_fp64_bit_parity():
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: hex8 O-grid 6x3 x 4 layers, JC rate + adiabatic heating, FP64, 200 steps")
    mesh = R.mesh.ogrid_bar(6, 3, 4, 3.2e-3, 2.6e-3, 0.4, 250.0)
    g = gpu_domain(mesh, JCMAT)
    o = O.Model(mesh, JCMAT, dict(rho_b=1.0), mathmode=1)
    g.begin(1.0)
    o.begin(1.0)
    for k in range(200):
        g.step(1)
        o.step()
        if k % 20 == 19:
            assert np.array_equal(g.field(L.F_FINT), o.get(O.F_FINT)), k
    assert np.array_equal(g.field(L.F_COORDS), o.get(O.F_X))
    assert np.array_equal(g.field(L.F_EVP), o.get(O.F_EVP))
    T = g.field(L.F_TEMP)
    assert np.array_equal(T, o.get(O.F_TEMP))
    assert T.max() > 293.0 + 1.0  # adiabatic heating happened


def test_johnson_cook_without_rate_and_heat_is_the_reference_law():
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: JC with C = 0 and no heating == A + B evp^n, bitwise")
    mesh = R.mesh.ogrid_bar(6, 3, 4, 3.2e-3, 2.6e-3, 0.4, 250.0)
    plain = dict(JCMAT)
    plain.pop("jc")
    j0 = dict(JCMAT)
    j0["jc"] = dict(JCMAT["jc"], C=0.0, tq=0.0)
    a = gpu_domain(mesh, plain)
    b = gpu_domain(mesh, j0)
    for d in (a, b):
        d.begin(1.0)
        d.step(120)
    for f in (L.F_COORDS, L.F_EVP, L.F_S):
        assert np.array_equal(a.field(f), b.field(f))


@pytest.mark.parametrize("case", ["axi", "hex8"])
def test_mixed_form_tracks_fp64(case):
    import redynela_amd as R
    from redynela_amd import _lib as L
    if case == "axi":
        say("case: mixed form vs FP64, axi 10x50 copper to 80 us")
        mesh, mat, t_end, nr = R.mesh.taylor_axi_10x50(), CU, 80e-6, 11
    else:
        say("case: mixed form vs FP64, hex8 O-grid 8x4 x 40 layers JC rate+thermal to 20 us")
        mesh, mat, t_end = R.mesh.ogrid_bar(8, 4, 40, 3.2e-3, 32.4e-3, 0.4, 250.0), JCMAT, 20e-6
    out = {}
    for form in ("fp64", "mixed"):
        d = gpu_domain(mesh, mat, form)
        n = d.solve(t_end)
        x = d.coords
        if case == "axi":
            length, radius = x[:, 1].max(), x[:nr, 0].max()
        else:
            base = mesh["x"][:, 2] == 0.0
            length, radius = x[:, 2].max(), np.hypot(x[base, 0], x[base, 1]).max()
        out[form] = (length, radius, d.evp.max(), n)
        say("  %s: steps %d length %.9e radius %.9e max evp %.6e" % (form, n, length, radius, d.evp.max()))
    (l0, r0, e0, _), (l1, r1, e1, _) = out["fp64"], out["mixed"]
    assert abs(l1 - l0) / l0 < 1e-4
    assert abs(r1 - r0) / r0 < 1e-4
    assert abs(e1 - e0) / e0 < 2e-3
