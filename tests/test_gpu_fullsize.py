"""GPU at the benchmark's full size: the 1,000,000-element hex8 O-grid bar
(2,000 elements per section x 500 layers, 1,022,541 nodes), FP64, against the
oracle running the same mesh -- the preamble and three increments, bit for
bit, then on to increment 100 (the benchmark's warm-up; its 5 + 20 window
ends at 25) with Fint, coordinates and evp identical at 10, 25, 50 and 100 --
plus run-to-run identity and the zero sum of the internal forces."""
import hashlib
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

COPPER = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_domain(mesh):
    import redynela_amd as R
    m = R.Material(COPPER["E"], COPPER["nu"], COPPER["rho"], R.IsoHardElastoplastic(COPPER["A"], COPPER["B"],
                                                                                   COPPER["n"]))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    return R.Domain(mesh, m, s)


def test_hex8_1m_full_size_bit_parity_and_invariants():
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: 1M hex8 O-grid, FP64, preamble + 3 increments against the oracle")
    mesh = R.mesh.ogrid_taylor_bar(500)
    assert mesh["conn"].shape[0] == 1000000 and mesh["x"].shape[0] == 1022541
    O.lib().dnlo_set_threads(16)
    o = O.Model(mesh, COPPER, dict(rho_b=1.0), mathmode=1)
    g = gpu_domain(mesh)
    g.begin(1.0)
    o.begin(1.0)
    for k in range(4):
        if k:
            g.step(1)
            o.step()
        for fg, fo, name in ((L.F_FINT, O.F_FINT, "Fint"), (L.F_COORDS, O.F_X, "coords"), (L.F_EVP, O.F_EVP, "evp")):
            a, b = g.field(fg), o.get(fo)
            assert np.array_equal(a, b), "increment %d: %s differs (max %.3e)" % (k, name, np.max(np.abs(a - b)))
        say("  increment %d: Fint, coords, evp identical" % k)
    tg, to = g.time(), o.time()
    assert tg["dt"] == to["dt"] and tg["increment"] == to["increment"]
    # internal forces of a free body sum to zero up to round-off
    f = g.field(L.F_FINT).reshape(-1, 3)
    assert np.all(np.abs(f.sum(axis=0)) <= 1e-10 * np.abs(f).sum(axis=0))
    h1 = hashlib.sha1(g.coords.tobytes()).hexdigest()
    done = 3
    for stop in (10, 25, 50, 100):
        g.step(stop - done)
        for _ in range(stop - done):
            o.step()
        done = stop
        for fg, fo, name in ((L.F_FINT, O.F_FINT, "Fint"), (L.F_COORDS, O.F_X, "coords"), (L.F_EVP, O.F_EVP, "evp")):
            a, b = g.field(fg), o.get(fo)
            assert np.array_equal(a, b), "increment %d: %s differs (max %.3e)" % (stop, name, np.max(np.abs(a - b)))
        tg, to = g.time(), o.time()
        assert tg["t"] == to["t"] and tg["dt"] == to["dt"]
        say("  increment %d: Fint, coords, evp, t, dt identical (max evp %.4f)" % (stop, g.evp.max()))
    g.close()
    g2 = gpu_domain(mesh)
    g2.begin(1.0)
    g2.step(3)
    assert hashlib.sha1(g2.coords.tobytes()).hexdigest() == h1
    g2.close()
