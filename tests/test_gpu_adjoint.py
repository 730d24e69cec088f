"""GPU: the hand-written adjoint (dnlg_adjoint_forward / _backward).

* its forward is the production stepping, bit for bit;
* its gradient equals the numpy restatement's adjoint (tests/adjoint_ref.py)
  on a small bar, with uneven checkpoint segments;
* on the parameter-identification workload (axi 20 x 100, 200 m/s, 4,000
  increments, checkpoints every 100) the gradient of the final-shape misfit
  w.r.t. (A, B, n) matches central differences of the oracle (the C
  restatement of the reference solver): within 1e-5 at relative step 1e-6
  once the solver's stops are tightened, and within its noise floor (2e-4)
  with the reference's own stops;
* the torch.autograd.Function returns the same gradient;
* with the Johnson-Cook temperature factor and adiabatic heating (the
  precision workload's m, T0, Tm, cp and Taylor-Quinney fraction, no rate
  term) the gradient equals the numpy restatement's and, on the 20 x 100
  workload, central differences of the oracle's same law within 1e-5 (stops
  tightened); the reference evaluates A + B evp^n only, so this part has no
  reference counterpart;
* hex8 (3D polar decomposition through the exact eigen-decomposition of
  F^T F, the hex8 Courant length V / A_max): on a small O-grid bar the
  gradient matches central differences of the oracle within 1e-5 (stops
  tightened), for the plain law and with the temperature factor, and equals
  the sweep that keeps every increment.
"""
import sys
import time

import numpy as np
import pytest

import adjoint_ref as AR
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_final(mesh, theta, nsteps, smooth=False, thermal=None, dims=2):
    """Final (r, z) of the oracle run.  smooth: test-only tighter stops of the
    radial return (1e-13 instead of 1e-8) and of the Jacobi polar iteration
    (1e-15 instead of 1e-9), which take the reference's iteration-count jumps
    and its shear dead zone |(F^T F)_01| < 1e-9 out of the finite differences."""
    A, B, n = theta
    sol = dict(rho_b=1.0)
    if smooth:
        sol["rr_prec"] = 1e-13
    mat = dict(E=115e9, nu=0.3, rho=8930.0, A=A, B=B, n=n)
    if thermal:
        mat["jc"] = dict(C=0.0, epsd0=1.0, m=thermal["m"], T0=thermal["T0"], Tm=thermal["Tm"], cp=thermal["cp"],
                         tq=thermal["tq"])
    m = O.Model(mesh, mat, sol, mathmode=1)
    O.lib().dnlo_set_polar_precision(1e-15 if smooth else 10.0e-10)
    try:
        m.begin(1e300)
        for _ in range(nsteps):
            m.step()
    finally:
        O.lib().dnlo_set_polar_precision(10.0e-10)
    return m.get(O.F_X).reshape(-1, 3)[:, :dims]


def central(mesh, th, i, h, nst, loss, smooth=False, thermal=None, dims=2):
    tp, tm = th.copy(), th.copy()
    tp[i] *= 1 + h
    tm[i] *= 1 - h
    return (loss(oracle_final(mesh, tp, nst, smooth, thermal, dims))
            - loss(oracle_final(mesh, tm, nst, smooth, thermal, dims))) / (2 * h * th[i])


def test_adjoint_forward_is_the_production_stepping():
    import redynela_amd as R
    from redynela_amd.adjoint import HardeningAdjoint
    say("case: axi 4x16, 150 increments, checkpoints every 40: adjoint forward == dnlg_step")
    mesh = R.mesh.taylor_axi(4, 16, 3.2e-3, 5.2e-3, 200.0)
    th = (90e6, 292e6, 0.31)
    p = HardeningAdjoint(mesh, nsteps=150, ckpt_every=40)
    x = p.forward(th)
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    d = R.Domain(mesh, R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(*th)), s)
    d.begin(1e300)
    d.step(150)
    assert np.array_equal(x, d.coords)
    # the backward leaves the post-forward state in place, and a second
    # backward over the same checkpoints gives the same gradient
    g1 = p.backward(np.ones((x.shape[0], 2)))
    assert np.array_equal(p.dom.coords, d.coords)
    assert p.dom.time()["increment"] == 150
    g2 = p.backward(np.ones((x.shape[0], 2)))
    assert np.array_equal(g1, g2)
    p.close()
    d.close()


def test_adjoint_equals_numpy_restatement():
    import redynela_amd as R
    from redynela_amd.adjoint import HardeningAdjoint
    say("case: axi 4x16, 230 increments, checkpoints every 60 (uneven), GPU vs numpy adjoint")
    mesh = R.mesh.taylor_axi(4, 16, 3.2e-3, 5.2e-3, 200.0)
    th0 = np.array([90e6, 292e6, 0.31])
    nst = 230
    m = AR.Model(mesh)
    xt, _ = m.run(th0, nst)
    th = th0 * np.array([1.12, 0.9, 1.1])
    xr, gr = m.grad(th, nst, xt)
    p = HardeningAdjoint(mesh, nsteps=nst, ckpt_every=60)
    x = p.forward(th)[:, :2]
    g = p.backward(2 * (x - xt))
    say("  numpy", gr, "\n  gpu  ", g, "\n  rel  ", np.abs(g - gr) / np.abs(gr))
    assert np.max(np.abs(x - xr)) < 1e-10 * np.max(np.abs(xr))  # Jacobi vs closed-form polar, round-off
    assert np.all(np.abs(g - gr) <= 1e-7 * np.abs(gr))
    p.close()


def test_adjoint_bar_20x100_matches_oracle_central_differences():
    from redynela_amd import adjoint as AD
    mesh = AD.bar_20x100()
    nst = AD.ADJ_STEPS
    say("case: axi 20x100, 200 m/s, %d increments, checkpoints every %d; target (A,B,n) = (90 MPa, 292 MPa, 0.31)"
        % (nst, AD.ADJ_CKPT))
    xt = oracle_final(mesh, AD.TARGET_THETA, nst)
    th = AD.perturbed_start()
    p = AD.HardeningAdjoint(mesh, nsteps=nst, ckpt_every=AD.ADJ_CKPT)
    t0 = time.perf_counter()
    x = p.forward(th)[:, :2]
    g = p.backward(2 * (x - xt))
    wall = time.perf_counter() - t0
    st = p.stats
    say("  theta", th, "\n  adjoint grad", g)
    say("  forward %.1f ms, backward %.1f ms (recompute %.1f ms), device memory %.1f MB (%d checkpoints, %.1f MB), "
        "wall %.2f s" % (st["forward_ms"], st["backward_ms"], st["recompute_ms"], st["device_bytes"] / 1e6,
                         st["n_checkpoints"], st["checkpoint_bytes"] / 1e6, wall))
    xo = oracle_final(mesh, th, nst)
    assert np.array_equal(x, xo)  # the forward is the oracle's, bit for bit
    # central differences of the reference solver: its map has jumps of
    # ~1e-9 (radial-return iteration count, Jacobi shear dead zone), so the
    # difference quotient has a noise floor; reported, bounded loosely
    loss = lambda xx: np.sum((xx - xt) ** 2)  # noqa: E731
    for i in range(3):
        row = [(h, central(mesh, th, i, h, nst, loss)) for h in (1e-4, 1e-5)]
        rel = [abs(fd - g[i]) / abs(fd) for _, fd in row]
        say("  reference FD, param %d: " % i + "  ".join("h=%.0e fd=%.9e rel=%.2e" % (h, fd, r)
                                                        for (h, fd), r in zip(row, rel)))
        assert min(rel) <= 2e-4
    # central differences of the smoothed solver at relative step 1e-6: <= 1e-5
    xts = oracle_final(mesh, AD.TARGET_THETA, nst, smooth=True)
    loss_s = lambda xx: np.sum((xx - xts) ** 2)  # noqa: E731
    for i in range(3):
        fd = central(mesh, th, i, 1e-6, nst, loss_s, smooth=True)
        rel = abs(fd - g[i]) / abs(fd)
        say("  smoothed FD, param %d: h=1e-06 fd=%.9e rel=%.2e" % (i, fd, rel))
        assert rel <= 1e-5
    p.close()


def test_torch_autograd_function():
    import torch
    import redynela_amd as R
    from redynela_amd import adjoint as AD
    say("case: torch.autograd through FinalShape, axi 4x16, 120 increments")
    mesh = R.mesh.taylor_axi(4, 16, 3.2e-3, 5.2e-3, 200.0)
    prob = AD.HardeningAdjoint(mesh, nsteps=120, ckpt_every=50)
    xt = torch.from_numpy(prob.forward(AD.TARGET_THETA)[:, :2].copy())
    theta = torch.tensor([100e6, 280e6, 0.3], dtype=torch.float64, requires_grad=True)
    loss = ((AD.FinalShape.apply(theta, prob) - xt) ** 2).sum()
    loss.backward()
    g = prob.backward(2 * (prob.dom.coords[:, :2] - xt.numpy()))
    assert torch.allclose(theta.grad, torch.from_numpy(g), rtol=1e-12, atol=0)
    prob.close()


def test_adjoint_checkpoint_every_increment_needs_no_recompute():
    """checkpoints after every increment: the sweep reads them directly, with
    no recomputation, and gives the gradient of the segment-recomputing sweep
    bit for bit"""
    from redynela_amd.adjoint import HardeningAdjoint
    import redynela_amd as R
    say("case: axi 4x16, 150 increments: gradient with checkpoints every increment == every 40")
    mesh = R.mesh.taylor_axi(4, 16, 3.2e-3, 5.2e-3, 200.0)
    th = (90e6, 292e6, 0.31)
    out = []
    for every in (40, 1):
        p = HardeningAdjoint(mesh, nsteps=150, ckpt_every=every)
        x = p.forward(th)
        g = p.backward(np.ones((x.shape[0], 2)))
        st = p.stats
        say("  every %d: backward %.2f ms, recompute %.2f ms, %d checkpoints" % (every, st["backward_ms"],
                                                                             st["recompute_ms"], st["n_checkpoints"]))
        out.append((x, g, st))
        p.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[1][2]["recompute_ms"] == 0.0 and out[1][2]["n_checkpoints"] == 150


def test_thermal_adjoint_equals_numpy_restatement():
    import redynela_amd as R
    from redynela_amd import adjoint as AD
    say("case: axi 4x16, 250 m/s, JC temperature factor + heating, 230 increments, checkpoints every 60, "
        "GPU vs numpy adjoint")
    mesh = R.mesh.taylor_axi(4, 16, 3.2e-3, 5.2e-3, 250.0)
    th0 = np.array([90e6, 292e6, 0.31])
    nst = 230
    m = AR.Model(mesh, thermal=AD.THERMAL)
    xt, _ = m.run(th0, nst)
    th = th0 * np.array([1.12, 0.9, 1.1])
    xr, gr = m.grad(th, nst, xt)
    p = AD.HardeningAdjoint(mesh, nsteps=nst, ckpt_every=60, thermal=AD.THERMAL)
    x = p.forward(th)[:, :2]
    g = p.backward(2 * (x - xt))
    say("  numpy", gr, "\n  gpu  ", g, "\n  rel  ", np.abs(g - gr) / np.abs(gr))
    assert np.max(np.abs(x - xr)) < 1e-10 * np.max(np.abs(xr))
    assert np.all(np.abs(g - gr) <= 1e-7 * np.abs(gr))
    # the temperature factor matters: the plain law's gradient is another one
    p2 = AD.HardeningAdjoint(mesh, nsteps=nst, ckpt_every=60)
    x2 = p2.forward(th)[:, :2]
    g2 = p2.backward(2 * (x2 - xt))
    assert np.max(np.abs(g2 - g) / np.abs(g)) > 1e-4
    p.close()
    p2.close()


def test_thermal_adjoint_bar_20x100_matches_oracle_central_differences():
    from redynela_amd import adjoint as AD
    mesh = AD.bar_20x100()
    nst = AD.ADJ_STEPS
    say("case: axi 20x100, 200 m/s, JC temperature factor + heating (m 1.09, T0 293 K, Tm 1356 K, cp 383, "
        "chi 0.9), %d increments, checkpoints every %d" % (nst, AD.ADJ_CKPT))
    T = AD.THERMAL
    th = AD.perturbed_start()
    p = AD.HardeningAdjoint(mesh, nsteps=nst, ckpt_every=AD.ADJ_CKPT, thermal=T)
    x = p.forward(th)[:, :2]
    xo = oracle_final(mesh, th, nst, thermal=T)
    assert np.array_equal(x, xo)  # the forward is the oracle's same law, bit for bit
    xts = oracle_final(mesh, AD.TARGET_THETA, nst, smooth=True, thermal=T)
    g = p.backward(2 * (x - xts))
    st = p.stats
    say("  theta", th, "\n  adjoint grad", g)
    say("  forward %.1f ms, backward %.1f ms, device memory %.1f MB" % (st["forward_ms"], st["backward_ms"],
                                                                     st["device_bytes"] / 1e6))
    loss_s = lambda xx: np.sum((xx - xts) ** 2)  # noqa: E731
    for i in range(3):
        fd = central(mesh, th, i, 1e-6, nst, loss_s, smooth=True, thermal=T)
        rel = abs(fd - g[i]) / abs(fd)
        say("  smoothed FD, param %d: h=1e-06 fd=%.9e rel=%.2e" % (i, fd, rel))
        assert rel <= 1e-5
    p.close()


def test_adjoint_refuses_the_rate_term():
    import redynela_amd as R
    from redynela_amd import _lib as L
    from redynela_amd._lib import DnlgError
    mesh = R.mesh.taylor_axi(2, 6, 3.2e-3, 1.9e-3, 150.0)
    h = R.IsoHardJohnsonCook(90e6, 292e6, 0.31, C=0.025, m=1.09, Tm=1356.0, T0=293.0)
    h.enableRateAndThermal(383.0, 0.9)
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    d = R.Domain(mesh, R.Material(115e9, 0.3, 8930.0, h), s)
    with pytest.raises(DnlgError):
        L.check(L.lib().dnlg_adjoint_forward(d.h, 10, 5), d.h)
    d.close()


def hex_bar():
    import redynela_amd as R
    # 48 elements per section (4 x 4 core, four 4 x 2 rings) x 10 layers, 250 m/s
    return R.mesh.ogrid_bar(4, 2, 10, 3.2e-3, 6.48e-3, 0.4, 250.0)


@pytest.mark.parametrize("law", ["plain", "thermal"])
def test_hex8_adjoint_matches_oracle_central_differences(law):
    from redynela_amd import adjoint as AD
    mesh = hex_bar()
    nst, every = 300, 50
    T = AD.THERMAL if law == "thermal" else None
    say("case: hex8 O-grid 48/section x 10 layers, 250 m/s, %s law, %d increments, checkpoints every %d"
        % (law, nst, every))
    th = AD.perturbed_start()
    p = AD.HardeningAdjoint(mesh, nsteps=nst, ckpt_every=every, thermal=T)
    x = p.forward(th)
    xo = oracle_final(mesh, th, nst, thermal=T, dims=3)
    assert np.array_equal(x, xo)  # the forward is the oracle's, bit for bit
    xts = oracle_final(mesh, AD.TARGET_THETA, nst, smooth=True, thermal=T, dims=3)
    g = p.backward(2 * (x - xts))
    st = p.stats
    say("  adjoint grad", g, "\n  forward %.1f ms, backward %.1f ms" % (st["forward_ms"], st["backward_ms"]))
    loss_s = lambda xx: np.sum((xx - xts) ** 2)  # noqa: E731
    for i in range(3):
        fd = central(mesh, th, i, 1e-6, nst, loss_s, smooth=True, thermal=T, dims=3)
        rel = abs(fd - g[i]) / abs(fd)
        say("  smoothed FD, param %d: h=1e-06 fd=%.9e rel=%.2e" % (i, fd, rel))
        assert rel <= 1e-5
    # the sweep over every increment's kept state gives the same gradient
    q = AD.HardeningAdjoint(mesh, nsteps=nst, ckpt_every=1, thermal=T)
    xq = q.forward(th)
    assert np.array_equal(q.backward(2 * (xq - xts)), g)
    p.close()
    q.close()


def test_adjoint_refuses_a_finite_end_time():
    """an end time would stop the loop before nsteps increments and leave
    no-op increments for the reverse sweep: refused before anything runs"""
    import redynela_amd as R
    from redynela_amd import _lib as L
    from redynela_amd._lib import DnlgError
    mesh = R.mesh.taylor_axi(2, 6, 3.2e-3, 1.9e-3, 150.0)
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    s.setTimes(0.0, 1e-7)
    d = R.Domain(mesh, R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(90e6, 292e6, 0.31)), s)
    with pytest.raises(DnlgError, match="end_time"):
        L.check(L.lib().dnlg_adjoint_forward(d.h, 50, 10), d.h)
    assert d.time()["increment"] == 0
    d.close()
