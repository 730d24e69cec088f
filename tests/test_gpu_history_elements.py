"""HistoryFile element fields on the GPU (stress, strain, strainInc,
plasticStrain: FE/Node.C:350-418 with FE/Element.C:1180-1198) against the
oracle's restatement of Node::getNodalValue, value for value, at every
increment of a hex8 and an axisymmetric run."""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

CODES = {"stress": 13, "strain": 14, "strainInc": 15, "plasticStrain": 16}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", ["hex8", "axi"])
def test_element_fields_every_increment(case):
    import redynela_amd as R
    if case == "hex8":
        say("case: HistoryFile element fields, hex8 O-grid 6+3 x 4 layers, 250 m/s, 60 increments")
        mesh = R.mesh.ogrid_bar(6, 3, 4, 3.2e-3, 3.0e-3, 0.4, 250.0)
        nodes = [0, 5, int(np.argmax(mesh["x"][:, 2])), 130]
    else:
        say("case: HistoryFile element fields, axi 10x50, 60 increments")
        mesh = R.mesh.taylor_axi_10x50()
        nodes = [0, 5, 12, 300, 560]
    mat = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)
    m = R.Material(mat["E"], mat["nu"], mat["rho"], R.IsoHardElastoplastic(mat["A"], mat["B"], mat["n"]))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    g = R.Domain(mesh, m, s)  # strain tracking is turned on by the strain fields
    hf = R.HistoryFile()
    items = []
    for node in nodes:
        for field in ("stress", "strain", "strainInc", "plasticStrain"):
            for comp in (0, 11, 12, 22, 33, 21):
                hf.append([node], field, comp)
                items.append((node, field, comp))
    hf.setSaveTime(1e-12, 0.0, 1e-12)  # a row at every increment
    g.addHistoryFile(hf)
    o = O.Model(mesh, mat, dict(rho_b=1.0), mathmode=1)
    g.begin(1.0)
    o.begin(1.0)
    g.step(60)
    got = g.history_rows(hf)
    assert got.shape == (60, 1 + len(items)), got.shape
    for k in range(60):
        o.step()
        want = [o.time()["t"]] + [o.nodal_value(n, CODES[f], c) for n, f, c in items]
        assert np.array_equal(got[k], np.array(want)), (k, np.abs(got[k] - np.array(want)).max())
    assert np.abs(got[-1, 1:]).max() > 0.0
    g.close()


def test_element_fields_refused_where_they_cannot_be_kept():
    import redynela_amd as R
    from redynela_amd import DnlgError
    mesh = R.mesh.taylor_axi(2, 4)
    m = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    g = R.Domain(mesh, m, s)
    hf = R.HistoryFile()
    hf.append([0], "stress", 14)  # no (1, 4) entry
    hf.setSaveTime(1e-7)
    with pytest.raises(DnlgError):
        g.addHistoryFile(hf)
    g.begin(1.0)
    g.step(2)
    hf2 = R.HistoryFile()
    hf2.append([0], "strain", 11)  # strains need tracking from the first pass
    hf2.setSaveTime(1e-7)
    with pytest.raises(DnlgError):
        g.addHistoryFile(hf2)
    g.close()
