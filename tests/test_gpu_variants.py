"""Every instance of the FP64 element kernels against the oracle, bit for bit.

The hex8 and the axisymmetric quad4 kernels are each compiled in eight
instances: with or without a constant BoundarySpeed on the top face
(FE/Boundary.C:478-524), with or without the strain diagnostics Eps, EpsPlas
and dEps (FE/Element.C:362-391), and with the reference's hardening
A + B evp^n (n = 0.31) or the Johnson-Cook rate and temperature extension.
Each of the 16 runs 40 increments, and every 10 increments Fint, coordinates,
velocities, evp, S and p -- and Eps, EpsPlas, dEps and the temperature where
the instance keeps them -- must equal the oracle's doubles.
"""
import itertools
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

MAT = dict(E=115e9, nu=0.3, rho=8930.0, A=90e6, B=292e6, n=0.31)
JCX = dict(C=0.025, epsd0=1.0, m=1.09, T0=293.0, Tm=1356.0, cp=383.0, tq=0.9)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tens(g, field):
    """GPU point tensors are component-major [9][n_points]; the oracle's are [n_points][9]."""
    return g.field(field).reshape(9, -1).T.reshape(-1)


def mesh_of(dim, speed):
    import redynela_amd as R
    if dim == 3:
        m = R.mesh.ogrid_bar(4, 2, 6, 3.2e-3, 3.9e-3, 0.4, 250.0)
        axis = 2
    else:
        m = R.mesh.taylor_axi(4, 16, 3.2e-3, 5.2e-3, 250.0)
        axis = 1
    if speed:
        # the top face driven at the impact speed along the axis
        top = m["x"][:, axis] == m["x"][:, axis].max()
        mask = np.zeros(m["x"].shape[0], dtype=np.uint8)
        mask[top] = 1 << axis
        sp = np.zeros_like(m["x"])
        sp[top, axis] = -250.0
        m = dict(m, speed_mask=mask, speed=sp)
    return m


@pytest.mark.parametrize("dim,speed,track,jc", list(itertools.product((3, 2), (False, True), (False, True),
                                                                        (False, True))))
def test_kernel_instance_bit_parity(dim, speed, track, jc):
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: %s, speed BC %s, strain diagnostics %s, %s" % ("hex8" if dim == 3 else "axi quad4", speed, track,
                                                               "Johnson-Cook rate + heating" if jc else "A + B evp^n"))
    mesh = mesh_of(dim, speed)
    if jc:
        h = R.IsoHardJohnsonCook(MAT["A"], MAT["B"], MAT["n"], C=JCX["C"], m=JCX["m"], depsp0=JCX["epsd0"],
                                 Tm=JCX["Tm"], T0=JCX["T0"])
        h.enableRateAndThermal(JCX["cp"], JCX["tq"])
        omat = dict(MAT, jc=JCX)
    else:
        h = R.IsoHardElastoplastic(MAT["A"], MAT["B"], MAT["n"])
        omat = MAT
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(0.8182)
    g = R.Domain(mesh, R.Material(MAT["E"], MAT["nu"], MAT["rho"], h), s, speed_mask=mesh.get("speed_mask"),
                 speed=mesh.get("speed"), track_strain=track)
    g.set_option(L.OPT_STORE_FINT, 1)
    o = O.Model(mesh, omat, dict(rho_b=0.8182), mathmode=1)
    g.begin(1.0)
    o.begin(1.0)
    pairs = [(L.F_FINT, O.F_FINT, None), (L.F_COORDS, O.F_X, None), (L.F_V, O.F_V, None), (L.F_EVP, O.F_EVP, None)]
    for k in range(4):
        g.step(10)
        for _ in range(10):
            o.step()
        for fg, fo, _ in pairs:
            a, b = g.field(fg), o.get(fo)
            assert np.array_equal(a, b), "increment %d: field %d differs (max %.3e)" % (10 * (k + 1), fg,
                                                                                      np.max(np.abs(a - b)))
        assert np.array_equal(tens(g, L.F_S), o.get(O.F_S))
        assert np.array_equal(g.field(L.F_P), o.get(O.F_P)[::g.nip])
        if track:
            for fg, fo in ((L.F_EPS, O.F_EPS), (L.F_EPSPLAS, O.F_EPSPLAS), (L.F_DEPS, O.F_DEPS)):
                assert np.array_equal(tens(g, fg), o.get(fo)), "increment %d: strain field %d" % (10 * (k + 1), fg)
        if jc:
            assert np.array_equal(g.field(L.F_TEMP), o.get(O.F_TEMP))
        tg, to = g.time(), o.time()
        assert tg["t"] == to["t"] and tg["dt"] == to["dt"]
    assert g.evp.max() > 1e-3  # the runs are plastic
    if jc:
        assert g.field(L.F_TEMP).max() > JCX["T0"]
    g.close()
