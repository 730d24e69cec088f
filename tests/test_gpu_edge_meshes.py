"""Meshes at the edges of the launch shapes, GPU FP64 against the oracle bit
for bit at every increment: one hexahedron (8 of a workgroup's 256 lanes
valid), one axisymmetric quadrilateral, and a column of 33 hexahedra (one
element past a full 32-element workgroup, so the second workgroup holds a
single element).  Each is driven hard enough into its wall to flow
plastically, so the radial return and the Jacobi sweep both run.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_parity import COPPER, compare_state, make_pair, same, tens

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hex_column(n, h=1e-3, hz=0.4e-3, v=-150.0, shear=20.0):
    """n square plates (h x h x hz) stacked along z (ElHex8n3D node order:
    base face counter-clockwise, then the top face), base restrained along z,
    every node at v along z, with an x shear so that the sweep rotates.  The
    plates are thinner than wide, as the O-grid bar's layers are: the hex8
    Courant length V / A_max then bounds the element's highest frequency (for one
    lone cube it does not, and the central-difference step is unstable)"""
    x = []
    for k in range(n + 1):
        for (i, j) in ((0, 0), (1, 0), (1, 1), (0, 1)):
            x.append((i * h + 0.05 * h * k / max(n, 1) * j, j * h, k * hz))
    conn = [[4 * k + c for c in range(8)] for k in range(n)]
    x = np.array(x, dtype=np.float64)
    bc = np.where(x[:, 2] == 0.0, 4, 0).astype(np.uint8)
    v0 = np.zeros_like(x)
    v0[:, 2] = v
    v0[:, 0] = shear * x[:, 2] / (n * hz)
    return dict(dim=3, x=x, conn=np.array(conn, dtype=np.int32), bc=bc, v0=v0)


def quad_single(h=1e-3, v=-150.0):
    """one axisymmetric quadrilateral (r, z): the axis restrained in r, the base in z"""
    x = np.array([(0, 0, 0), (h, 0, 0), (h, h, 0), (0, h, 0)], dtype=np.float64)
    bc = np.zeros(4, np.uint8)
    bc[x[:, 0] == 0.0] |= 1
    bc[x[:, 1] == 0.0] |= 2
    v0 = np.zeros_like(x)
    v0[:, 1] = v
    return dict(dim=2, x=x, conn=np.array([[0, 1, 2, 3]], dtype=np.int32), bc=bc, v0=v0)


@pytest.mark.parametrize("name", ["hex8 x1", "hex8 x33", "quad4 axi x1"])
def test_edge_mesh_per_step(name):
    from redynela_amd import _lib as L
    mesh = {"hex8 x1": lambda: hex_column(1, v=-20.0, shear=5.0), "hex8 x33": lambda: hex_column(33),
            "quad4 axi x1": quad_single}[name]()
    say("case: %s (%d elements) driven into the wall, per-step bit parity for 150 increments"
        % (name, mesh["conn"].shape[0]))
    g, o = make_pair(mesh, COPPER)
    assert g.begin(1.0) and o.begin(1.0) == 1
    compare_state(g, o, "%s preamble" % name)
    for k in range(150):
        rg = g.step(1)
        ro = o.step()
        compare_state(g, o, "%s step %d" % (name, k + 1))
        tg, to = g.time(), o.time()
        assert tg["t"] == to["t"] and tg["dt"] == to["dt"] and rg == (ro == 1)
        if not rg:
            break
    assert same(tens(g, L.F_S), o.get(O.F_S))
    assert same(g.field(L.F_V), o.get(O.F_V))
    assert np.max(o.get(O.F_EVP)) > 1e-4, "no plastic flow: the radial return did not run"
    g.close()


@pytest.mark.parametrize("name", ["hex8 x1", "hex8 x33", "quad4 axi x1"])
def test_edge_mesh_mixed_form_tracks_fp64(name):
    """the mixed form on the same meshes (its hex8 pass keeps each lane's node
    rows in LDS; lanes past the last element return before touching it):
    after 150 increments the displacement field is within 1e-4 of FP64,
    the final-shape bar"""
    import redynela_amd as R
    from redynela_amd import _lib as L
    mesh = {"hex8 x1": lambda: hex_column(1, v=-20.0, shear=5.0), "hex8 x33": lambda: hex_column(33),
            "quad4 axi x1": quad_single}[name]()
    say("case: %s, mixed form against FP64 on the GPU, 150 increments" % name)
    out = {}
    for form in ("fp64", "mixed"):
        m = R.Material(COPPER["E"], COPPER["nu"], COPPER["rho"],
                       R.IsoHardElastoplastic(COPPER["A"], COPPER["B"], COPPER["n"]))
        s = R.ExplicitSolverCH()
        s.setTimeStepMethod("Courant")
        s.setDissipation(1.0)
        s.setPrecision(form)
        g = R.Domain(mesh, m, s)
        g.begin(1.0)
        g.step(150)
        out[form] = (g.field(L.F_COORDS).reshape(-1, 3), g.time()["increment"])
        g.close()
    x0 = np.asarray(mesh["x"], dtype=np.float64)
    d64, dmx = out["fp64"][0] - x0, out["mixed"][0] - x0
    err = np.max(np.abs(dmx - d64)) / np.max(np.abs(d64))
    say("  displacement error %.3e" % err)
    assert out["fp64"][1] == out["mixed"][1] and err < 1e-4


def test_single_cube_pulse_per_step():
    """one 1 mm cube under the Pulse time step, the reference's hex8 method:
    GPU against the oracle bit for bit per increment (time scalars,
    coordinates, iteration counts of the power iteration), and bounded,
    where the build-defined Courant length would be unstable on a lone cube"""
    from redynela_amd import _lib as L
    from test_gpu_pulse import check, pair
    mesh = hex_column(1, hz=1e-3, v=-10.0, shear=2.0)
    say("case: one 1 mm hex8 cube, Pulse time step, 120 increments, GPU against the oracle")
    g, o = pair(mesh, COPPER)
    assert g.begin(1.0) and o.begin(1.0) == 1
    check(g, o, "cube preamble")
    for k in range(120):
        g.step(1)
        assert o.step() >= 0
        check(g, o, "cube step %d" % (k + 1))
    assert np.max(np.abs(g.field(L.F_V))) < 50.0  # no blow-up
    g.close()
