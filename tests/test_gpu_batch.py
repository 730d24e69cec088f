"""GPU: the batched mode (many bars, each with its own time step).

* FP64: sampled bars of the 512-bar workload, run together to 40 us by the
  stepped engine, are bit-identical to the oracle running each bar alone, and
  to the single-domain GPU path (also across a split solve); every engine
  (persistent, stepped in 64- or 256-lane blocks, without graphs, and the
  autotuned one, whose probes must leave the state as it was) gives the same
  bits;
* mixed form: every bar's final length and mushroom radius within 1e-4 of
  the FP64 batch, peak plastic strain within 2e-3;
* an exhausted increment budget is an error; a solve with nothing to do
  leaves the state alone.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

T_END = 40e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solver(form="fp64"):
    import redynela_amd as R
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    s.setPrecision(form)
    return s


def sample_bars(nr):
    """smallest, largest and a few in between"""
    idx = [int(np.argmin(nr)), int(np.argmax(nr))]
    for k in (3, 100, 257, 511):
        if k not in idx:
            idx.append(k)
    return idx


def shape(x, nr):
    """final length (top of the axis) and mushroom radius (outer base node)"""
    return x[:, 1].max(), x[: nr + 1, 0].max()


@pytest.fixture(scope="module")
def workload():
    from redynela_amd import batch
    meshes, mats = batch.taylor_batch(512)
    return meshes, mats


@pytest.fixture(scope="module")
def fp64_batch(workload):
    import redynela_amd as R
    meshes, mats = workload
    say("case: 512-bar batch, FP64, to 40 us")
    b = R.Batch(meshes, mats, solver("fp64"))
    w, ms = b.solve(T_END)
    say("  launch %.2f ms, per-bar wall p50 %.1f us p99 %.1f us" % (ms, np.percentile(w, 50), np.percentile(w, 99)))
    yield b
    b.close()


def test_batch_fp64_bit_identical_to_oracle(workload, fp64_batch):
    from redynela_amd import _lib as L
    meshes, mats = workload
    nrs = [int(np.sqrt(m["conn"].shape[0] / 5) + 0.5) for m in meshes]
    for i in sample_bars(np.array(nrs)):
        say("  bar %d: nr %d, %d elements" % (i, nrs[i], meshes[i]["conn"].shape[0]))
        o = O.Model(meshes[i], mats[i], dict(rho_b=1.0), mathmode=1)
        o.solve(T_END)
        to, tg = o.time(), fp64_batch.time(i)
        assert tg["increment"] == to["increment"] and tg["t"] == to["t"] and tg["dt"] == to["dt"], (i, tg, to)
        assert np.array_equal(fp64_batch.coords(i).reshape(-1), o.get(O.F_X)), i
        assert np.array_equal(fp64_batch.field(i, L.F_V), o.get(O.F_V)), i
        assert np.array_equal(fp64_batch.evp(i), o.get(O.F_EVP)), i
        assert np.array_equal(fp64_batch.field(i, L.F_P), o.get(O.F_P)[::4]), i  # the oracle keeps p per point
        S = fp64_batch.field(i, L.F_S).reshape(9, -1).T.reshape(-1)
        assert np.array_equal(S, o.get(O.F_S)), i
        assert fp64_batch.evp(i).max() > 0.05  # the bar did yield


def test_batch_bar_equals_single_domain_gpu_run_split_solve(workload):
    import redynela_amd as R
    meshes, mats = workload
    i = 1  # nr 18
    say("case: bar %d in a batch of 3 vs alone, solve to 15 us then 40 us" % i)
    b = R.Batch([meshes[0], meshes[i], meshes[2]], [mats[0], mats[i], mats[2]], solver())
    t = mats[i]
    d = R.Domain(meshes[i], R.Material(t["E"], t["nu"], t["rho"], R.IsoHardElastoplastic(t["A"], t["B"], t["n"])),
                 solver())
    for up in (15e-6, T_END):
        b.solve(up)
        d.solve(up)
        assert b.time(1)["increment"] == d.time()["increment"]
        assert np.array_equal(b.coords(1), d.coords)
        assert np.array_equal(b.evp(1), d.evp)
    b.close()
    d.close()


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_persistent_engine_equals_stepped(workload, form):
    import redynela_amd as R
    from redynela_amd import _lib as L
    meshes, mats = workload
    idx = list(range(0, 512, 21))
    say("case: %d bars, %s, every engine (persistent, stepped 64/256 lanes, plain, auto), to 40 us"
        % (len(idx), form))
    out = []
    for eng in ("persistent", "stepped", "stepped-wide", "stepped-plain", "auto"):
        b = R.Batch([meshes[i] for i in idx], [mats[i] for i in idx], solver(form))
        b.set_engine(eng)
        if eng == "auto":
            with pytest.raises(R.DnlgError, match="autotuner has not run"):
                b.tuning()
        # for 'auto' the probes run first, from a copy of the state that is
        # put back, so the solve still lands on the same bits
        w, ms = b.solve(T_END)
        say("  %s: %.2f ms" % (eng, ms))
        if eng == "auto":
            choice, probes = b.tuning()
            say("  auto: chose %s, probes %s" % (choice, probes))
            assert choice in ("stepped", "stepped-wide", "persistent")
            assert all(v > 0 for v in probes.values())
            assert probes[choice] == min(probes.values())
        out.append([(b.time(k)["increment"], b.coords(k), b.evp(k), b.field(k, L.F_S)) for k in range(len(idx))])
        b.close()
    for name, other in zip(("stepped", "stepped-wide", "stepped-plain", "auto"), out[1:]):
        for k, (p, q) in enumerate(zip(out[0], other)):
            assert p[0] == q[0], (name, k)
            for u, v in zip(p[1:], q[1:]):
                assert np.array_equal(u, v), (name, k)


def test_batch_mixed_tracks_fp64(workload, fp64_batch):
    import redynela_amd as R
    meshes, mats = workload
    say("case: 512-bar batch, mixed form, to 40 us")
    b = R.Batch(meshes, mats, solver("mixed"))
    w, ms = b.solve(T_END)
    say("  launch %.2f ms, per-bar wall p50 %.1f us p99 %.1f us" % (ms, np.percentile(w, 50), np.percentile(w, 99)))
    worst = [0.0, 0.0, 0.0]
    for i, m in enumerate(meshes):
        nr = int(np.sqrt(m["conn"].shape[0] / 5) + 0.5)
        l0, r0 = shape(fp64_batch.coords(i), nr)
        l1, r1 = shape(b.coords(i), nr)
        e0, e1 = fp64_batch.evp(i).max(), b.evp(i).max()
        worst = [max(worst[0], abs(l1 - l0) / l0), max(worst[1], abs(r1 - r0) / r0), max(worst[2], abs(e1 - e0) / e0)]
    say("  worst relative error: length %.2e radius %.2e max evp %.2e" % tuple(worst))
    assert worst[0] < 1e-4 and worst[1] < 1e-4 and worst[2] < 2e-3
    b.close()


def test_batch_budget_and_idle_solve(workload):
    import redynela_amd as R
    meshes, mats = workload
    say("case: batch of 2, max_steps 50 to 40 us (budget error), then a solve to t = 0")
    b = R.Batch(meshes[:2], mats[:2], solver())
    with pytest.raises(R.DnlgError, match="max_steps"):
        b.solve(T_END, max_steps=50)
    assert b.time(0)["increment"] == 50 and b.time(1)["increment"] == 50
    b.close()
    b = R.Batch(meshes[:2], mats[:2], solver())
    b.solve(0.0)
    assert b.time(0)["increment"] == 0 and np.array_equal(b.coords(0), meshes[0]["x"])
    b.close()


def test_batch_element_error_same_under_autotune():
    """a bar that inverts in its first increments: every engine stops it at the
    same point (the failing increment counted, no nodal update after it) and
    reports the same error; the autotuned solve, whose probes meet the error
    first and are undone, too; the other bar still runs to the end"""
    import redynela_amd as R
    say("case: batch of 2, one bar at 20 km/s (inverts at once), stepped / persistent / auto")
    good = R.mesh.taylor_axi(5, 25, 3.2e-3, 32.4e-3, 200.0)
    bad = R.mesh.taylor_axi(5, 25, 3.2e-3, 32.4e-3, 2.0e4)
    mat = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)
    out = []
    for eng in ("stepped", "persistent", "auto"):
        b = R.Batch([good, bad], [mat, mat], solver())
        b.set_engine(eng)
        with pytest.raises(R.DnlgError) as ei:
            b.solve(T_END)
        say("  %s: %s" % (eng, ei.value))
        out.append((str(ei.value), b.time(0)["increment"], b.time(1)["increment"], b.coords(0), b.coords(1)))
        b.close()
    assert "domain 1" in out[0][0] and out[0][2] < 48 < out[0][1]
    for other in out[1:]:
        assert other[0] == out[0][0]
        assert other[1:3] == out[0][1:3]
        assert np.array_equal(other[3], out[0][3]) and np.array_equal(other[4], out[0][4])
