"""The numpy restatement of the adjoint (tests/adjoint_ref.py) against central
differences of its own forward: pins the hand-derived reverse formulas (host
only, no GPU)."""
import numpy as np

import adjoint_ref as AR


def test_adjoint_formulas_match_central_differences():
    import redynela_amd as R
    mesh = R.mesh.taylor_axi(3, 12, 3.2e-3, 3.888e-3, 200.0)
    m = AR.Model(mesh)
    th0 = np.array([90e6, 292e6, 0.31])
    nst = 120  # covers Courant recomputations, LINEARdbl extrapolation and plastic flow
    xt, _ = m.run(th0, nst)
    th = th0 * np.array([1.1, 0.85, 1.15])
    xN, g = m.grad(th, nst, xt)
    for i in range(3):
        h = 1e-5
        tp, tm = th.copy(), th.copy()
        tp[i] *= 1 + h
        tm[i] *= 1 - h
        fd = (np.sum((m.run(tp, nst)[0] - xt) ** 2) - np.sum((m.run(tm, nst)[0] - xt) ** 2)) / (2 * h * th[i])
        assert abs(fd - g[i]) <= 1e-7 * abs(fd), (i, fd, g[i])


def test_restrained_nodes_and_zero_seed_give_zero_gradient():
    import redynela_amd as R
    mesh = R.mesh.taylor_axi(2, 6, 3.2e-3, 1.9e-3, 150.0)
    m = AR.Model(mesh)
    th = np.array([90e6, 292e6, 0.31])
    xN, _ = m.run(th, 30)
    _, g = m.grad(th, 30, xN)  # loss at its minimum: zero seed
    assert np.all(g == 0.0)


def test_thermal_adjoint_formulas_match_central_differences():
    """the Johnson-Cook temperature factor and adiabatic heating (no rate
    term): the point temperature's adjoint and the heating's dependence on
    (A, B, n) against central differences of the same forward"""
    import redynela_amd as R
    mesh = R.mesh.taylor_axi(3, 12, 3.2e-3, 3.888e-3, 250.0)
    th_m = dict(m=1.09, T0=293.0, Tm=1356.0, cp=383.0, tq=0.9)
    m = AR.Model(mesh, thermal=th_m)
    th0 = np.array([90e6, 292e6, 0.31])
    nst = 120
    xt, _ = m.run(th0, nst)
    # the heating is felt: the same run without it ends elsewhere
    assert np.max(np.abs(AR.Model(mesh).run(th0, nst)[0] - xt)) > 1e-9
    th = th0 * np.array([1.1, 0.85, 1.15])
    xN, g = m.grad(th, nst, xt)
    for i in range(3):
        h = 1e-5
        tp, tm = th.copy(), th.copy()
        tp[i] *= 1 + h
        tm[i] *= 1 - h
        fd = (np.sum((m.run(tp, nst)[0] - xt) ** 2) - np.sum((m.run(tm, nst)[0] - xt) ** 2)) / (2 * h * th[i])
        assert abs(fd - g[i]) <= 1e-7 * abs(fd), (i, fd, g[i])
