"""HistoryFile element fields in the oracle (Node::getNodalValue,
FE/Node.C:350-418, with Element's get_IntegrationValue_atNode,
FE/Element.C:1180-1198), pinned by what the extrapolation must do:

* integrationPointsToNode inverts the shape functions at the Gauss points:
  a field that is trilinear (hex8) or bilinear (quad4) in the reference
  coordinates, sampled at the points, comes back at each node as its value
  there -- a transposed or misplaced coefficient fails this;
* a node shared by several elements gets the mean of their extrapolations;
* component 0 is the von Mises equivalent of a stress and sqrt(2/3 e:e) of a
  strain; 10 i + j is entry (i, j).
"""
import numpy as np
import pytest

import oracle_lib as O

EL = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0, **{"yield": 0})
SG = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
              dtype=float)
STRESS, STRAIN, STRAIN_INC, PLASTIC = 13, 14, 15, 16


def hexes(n):
    """n unit hexahedra stacked along z, sharing faces"""
    x = []
    for k in range(n + 1):
        for (i, j) in ((0, 0), (1, 0), (1, 1), (0, 1)):
            x.append([i * 1e-3, j * 1e-3, k * 1e-3])
    conn = [[4 * k + a for a in (0, 1, 2, 3)] + [4 * (k + 1) + a for a in (0, 1, 2, 3)] for k in range(n)]
    x = np.array(x)
    return dict(dim=3, x=x, conn=np.array(conn, dtype=np.int32), bc=np.zeros(len(x), np.uint8),
                v0=np.zeros_like(x))


def quad():
    x = np.array([[1e-3, 0, 0], [2e-3, 0, 0], [2e-3, 1e-3, 0], [1e-3, 1e-3, 0]], dtype=float)
    return dict(dim=2, x=x, conn=np.arange(4, dtype=np.int32).reshape(1, 4), bc=np.zeros(4, np.uint8),
                v0=np.zeros((4, 3)))


def f3(p):  # trilinear in the reference coordinates
    x, y, z = p
    return 1.5 + 0.3 * x - 0.7 * y + 0.2 * z + 0.11 * x * y - 0.05 * y * z + 0.08 * x * z + 0.03 * x * y * z


M = np.array([[2.0, 0.3, -0.1], [0.4, -1.0, 0.25], [-0.2, 0.6, 0.5]])


@pytest.mark.parametrize("field", [STRESS, STRAIN, STRAIN_INC, PLASTIC])
def test_hex8_extrapolation_reproduces_a_trilinear_field(field):
    mesh = hexes(1)
    m = O.Model(mesh, EL, dict(rho_b=1.0), mathmode=0)
    g = SG / np.sqrt(3.0)
    T = np.array([f3(g[p]) * M for p in range(8)]).reshape(8, 9)
    fld = {STRESS: O.F_SIG, STRAIN: O.F_EPS, STRAIN_INC: O.F_DEPS, PLASTIC: O.F_EPSPLAS}[field]
    m.set(fld, T.reshape(-1))
    for node in range(8):
        want = f3(SG[node]) * M
        for comp in (11, 12, 13, 21, 22, 23, 31, 32, 33):
            got = m.nodal_value(node, field, comp)
            assert abs(got - want[comp // 10 - 1, comp % 10 - 1]) <= 1e-13, (node, comp, got)
        eq = m.nodal_value(node, field, 0)
        if field == STRESS:
            s = want
            vm = np.sqrt(((s[0, 0] - s[1, 1]) ** 2 + (s[0, 0] - s[2, 2]) ** 2 + (s[1, 1] - s[2, 2]) ** 2
                          + 6 * (s[0, 1] ** 2 + s[0, 2] ** 2 + s[1, 2] ** 2)) / 2)
            assert abs(eq - vm) <= 1e-13 * vm
        else:
            assert abs(eq - np.sqrt(2.0 / 3.0 * (want ** 2).sum())) <= 1e-13


def test_quad4_extrapolation_reproduces_a_bilinear_field():
    mesh = quad()
    m = O.Model(mesh, EL, dict(rho_b=1.0), mathmode=0)
    g = SG[:4, :2] / np.sqrt(3.0)
    f2 = lambda p: 0.9 - 0.4 * p[0] + 0.6 * p[1] + 0.25 * p[0] * p[1]
    T = np.array([f2(g[p]) * M for p in range(4)]).reshape(4, 9)
    m.set(O.F_SIG, T.reshape(-1))
    for node in range(4):
        for comp in (11, 22, 33, 12):
            want = f2(SG[node, :2]) * M[comp // 10 - 1, comp % 10 - 1]
            assert abs(m.nodal_value(node, STRESS, comp) - want) <= 1e-13


def test_shared_node_is_the_mean_of_its_elements():
    mesh = hexes(2)  # nodes 4..7 are shared by both elements
    m = O.Model(mesh, EL, dict(rho_b=1.0), mathmode=0)
    c = np.array([3.0e8] * 8 * 9).reshape(8, 9)
    T = np.concatenate([c, 2 * c]).reshape(-1)  # uniform in each element: 3e8, then 6e8
    m.set(O.F_SIG, T)
    assert abs(m.nodal_value(0, STRESS, 11) - 3.0e8) <= 1e-6
    assert abs(m.nodal_value(5, STRESS, 11) - 4.5e8) <= 1e-6
    assert abs(m.nodal_value(10, STRESS, 11) - 6.0e8) <= 1e-6
