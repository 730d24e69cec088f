"""Decks run on the GPU through the deck reader, with result files, against
the oracle running the same deck through the same Structure::solve loop.

The bar is byte equality of every `<base>.NNN.bin` the two runs write: the
GPU FP64 path is bit-identical to the oracle, so coordinates, nodal
velocities and accelerations, stresses, strains and the ro/T swap parity
all agree to the last bit.  The hex deck adds constant BoundarySpeed nodes.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_files(d, base):
    """the oracle through the same Structure::solve loop: result files and
    the .plot lines (ExplicitSolverCH::solve stepped one increment at a time)"""
    from redynela_amd import deck
    from redynela_amd import results as RF
    _, mat, _ = d.model()
    o = O.from_deck(d, ke=True)
    paths, rows = [], []

    def save(t):
        _, nodes, ips = O.result_blocks(o, mat.density, mat.init_temp)
        paths.append("%s.%03d.bin" % (base, len(paths) + 1))
        RF.write(paths[-1], t, nodes, ips)

    def solve(t):
        r = o.begin(t)
        assert r >= 0
        while r == 1:
            r = o.step()
            assert r >= 0
            rows.append(o.plot_row())
        return o.time()["t"]

    deck.structure_solve(d.solver.stop, d.save, solve, save)
    return paths, np.array(rows)


@pytest.mark.parametrize("name", ["taylor_axi_small", "hex_block"])
def test_deck_result_files_match_oracle(tmp_path, name):
    from redynela_amd import deck
    say("case: deck %s run with result files, GPU FP64 against the oracle" % name)
    d = deck.load(os.path.join(HERE, "decks", name + ".dyn"))
    seen = []
    dom = d.run(results=str(tmp_path / "gpu" / name), on_save=lambda t, dm: seen.append(t))
    gpu = sorted(f for f in os.listdir(tmp_path / "gpu") if f.endswith(".bin"))
    ref, rows = oracle_files(d, str(tmp_path / name))
    assert gpu == [os.path.basename(p) for p in ref] and len(gpu) == len(seen) >= 2
    for g, r in zip(gpu, ref):
        a = open(tmp_path / "gpu" / g, "rb").read()
        b = open(r, "rb").read()
        assert len(a) == len(b)
        if a != b:
            x, y = np.frombuffer(a, "<f8"), np.frombuffer(b, "<f8")
            k = np.nonzero(x.view(np.int64) != y.view(np.int64))[0]
            pytest.fail("%s: %d doubles differ, first at %d (%r vs %r)" % (g, k.size, k[0], x[k[0]], y[k[0]]))
    assert np.max(dom.evp) > 0.0
    dom.close()
    # the .plot history: one line per increment; time and time steps print
    # identically, the kinetic energy (a tree sum on the GPU, a node-order sum
    # in the oracle and the reference) to the printed 5 digits
    from redynela_amd import results as RF
    plot = str(tmp_path / "gpu" / (name + "_defaultDomain.plot"))
    lines = open(plot).read().splitlines()
    assert lines[0] == "#DynELA_plot : Time history of computation"
    assert lines[1] == "#plotted : Used_timeStep timeStep kinetic_energy"
    assert len(lines) - 2 == len(rows) > 10
    for ln, r in zip(lines[2:], rows):
        f = ln.split(" ")
        assert f[:3] == ["%8.4E" % r[0], "%8.4E" % r[1], "%8.4E" % r[2]], ln
        assert abs(float(f[3]) - r[3]) <= 1e-4 * abs(r[3]), ln
    assert np.all(RF.read_plot(plot)[:, 3] > 0)


def test_hex_deck_two_component_speed_matches_oracle(tmp_path):
    """the hex deck with its top face driven at (20, 0, -150) m/s: speed mask 5
    (x and z driven, y free, FE/Boundary.C:478-524); every result file and the
    final state equal the oracle's bit for bit"""
    from redynela_amd import deck
    say("case: hex deck, constant BoundarySpeed on x and z (mask 5), GPU FP64 against the oracle")
    text = open(os.path.join(HERE, "decks", "hex_block.dyn")).read()
    assert text.count("vt.set(0, 0, -speed);") == 1
    path = tmp_path / "hex_block_xz.dyn"
    path.write_text(text.replace("vt.set(0, 0, -speed);", "vt.set(20, 0, -speed);"))
    d = deck.load(str(path))
    m, _, _ = d.model()
    top = m["x"][:, 2] == 3e-3
    assert np.all(m["speed_mask"][top] == 5) and np.all(m["v0"][top] == [20.0, 0.0, -150.0])
    dom = d.run(results=str(tmp_path / "gpu" / "hex_block_xz"))
    gpu = sorted(f for f in os.listdir(tmp_path / "gpu") if f.endswith(".bin"))
    ref, _ = oracle_files(d, str(tmp_path / "hex_block_xz"))
    assert gpu == [os.path.basename(p) for p in ref] and len(gpu) >= 2
    for g, r in zip(gpu, ref):
        assert open(tmp_path / "gpu" / g, "rb").read() == open(r, "rb").read(), g
    x = dom.coords.reshape(-1, 3)
    # the driven x and z moved at their speeds; y of the top face is free
    assert np.all(x[top, 0] > m["x"][top, 0]) and np.all(x[top, 2] < m["x"][top, 2])
    dom.close()


def test_plot_rows_per_increment():
    """the history lines against the oracle's endStep values, per increment:
    time and time steps bit-equal, kinetic energy to 1e-13"""
    import redynela_amd as R
    from redynela_amd import _lib as L
    say("case: endStep history lines, hex8 Courant, 40 increments in chunks")
    mesh = R.mesh.ogrid_bar(6, 3, 4, 3.2e-3, 3.0e-3, 0.4, 250.0)
    m = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    g = R.Domain(mesh, m, s)
    g.set_option(L.OPT_PLOT_HISTORY, 1)
    o = O.Model(mesh, dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0), dict(), mathmode=1, ke=True)
    g.begin(1.0)
    o.begin(1.0)
    want = []
    for chunk in (1, 7, 12, 20):
        g.step_async(chunk)
        g.synchronize()
        for _ in range(chunk):
            o.step()
            want.append(o.plot_row())
    got = g.plot_rows()
    want = np.array(want)
    assert got.shape == (40, 4)
    assert np.array_equal(got[:, :3], want[:, :3])
    np.testing.assert_allclose(got[:, 3], want[:, 3], rtol=1e-13)
    assert g.plot_rows().shape == (0, 4)
    g.close()


def oracle_history(o, items, start, stop, every, density, T0, nsteps):
    """HistoryFile::save restated over oracle increments (FE/HistoryFile.C:114-141):
    the Current field before the swap (read before the step), coordinates after"""
    vec = {"speed": O.F_V, "speedInc": O.F_A, "displacementInc": O.F_DU, "displacement": O.F_D}
    rows, nxt = [], start
    for _ in range(nsteps):
        cur = {k: o.get(f).reshape(-1, 3).copy() for k, f in vec.items()}
        o.step()
        tm, inc = o.time()["t"], o.time()["increment"]
        if (stop > 0 and tm >= stop + every) or tm < nxt:
            continue
        nxt += every
        x = o.get(O.F_X).reshape(-1, 3)
        row = [tm]
        for node, field, comp in items:
            v = x[node] if field == "coords" else cur.get(field, None)
            if field in ("coords",) or field in vec:
                vv = v if field == "coords" else v[node]
                row.append(np.sqrt(vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]) if comp == 0 else vv[comp - 1])
            elif field == "density":
                row.append(density if (inc - 1) % 2 == 0 else 0.0)
            elif field == "mass":
                row.append(o.get(O.F_MASS)[node])
            else:
                raise AssertionError(field)
        rows.append(row)
    return np.array(rows)


def test_history_file_rows_per_increment():
    """HistoryFile series against the oracle, value for value"""
    import redynela_amd as R
    say("case: HistoryFile nodal series, hex8 Courant, 60 increments")
    mesh = R.mesh.ogrid_bar(6, 3, 4, 3.2e-3, 3.0e-3, 0.4, 250.0)
    m = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    g = R.Domain(mesh, m, s)
    hf = R.HistoryFile()
    top = int(np.argmax(mesh["x"][:, 2]))
    items = [(top, "speed", 0), (top, "speed", 3), (5, "speedInc", 1), (top, "displacement", 3),
             (7, "displacementInc", 2), (top, "coords", 0), (top, "coords", 3), (3, "density", 0), (3, "mass", 0)]
    for node, field, comp in items:
        hf.append([node], field, comp)
    hf.setSaveTime(5e-8, 1.5e-6, 1e-7)
    g.addHistoryFile(hf)
    o = O.Model(mesh, dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0), dict(), mathmode=1)
    g.begin(1.0)
    o.begin(1.0)
    g.step(25)
    g.step_async(35)
    g.synchronize()
    got = g.history_rows(hf)
    want = oracle_history(o, items, 5e-8, 1.5e-6, 1e-7, 8930.0, 0.0, 60)
    assert got.shape == want.shape and got.shape[0] >= 5, (got.shape, want.shape)
    assert np.array_equal(got, want)
    assert g.history_rows(hf).shape[0] == 0
    g.close()


def test_deck_history_file(tmp_path):
    """the hex deck's HistoryFile, written by the deck runner, against the oracle"""
    from redynela_amd import deck
    say("case: deck HistoryFile text against the oracle")
    d = deck.load(os.path.join(HERE, "decks", "hex_block.dyn"))
    mesh, mat, sol = d.model()
    dom = d.run(results=str(tmp_path / "hex_block"))
    text = open(tmp_path / "hex_block_history.plot").read().splitlines()
    assert text[0] == "#DynELA_plot history file"
    (hfo,) = d.history_files
    items = [(d._rank[n], f, c) for nums, f, c in hfo.items for n in nums]
    # the oracle through the same loop
    o = O.from_deck(d)
    rows = []
    start, stop, every = hfo.save
    # ExplicitSolverCH::solve to each save time, the series taken per increment
    nxt = start
    vec = {"speed": O.F_V, "speedInc": O.F_A, "displacementInc": O.F_DU, "displacement": O.F_D}

    def run_to(t):
        nonlocal nxt
        r = o.begin(t)
        while r == 1:
            cur = {k: o.get(f).reshape(-1, 3).copy() for k, f in vec.items()}
            r = o.step()
            tm, inc = o.time()["t"], o.time()["increment"]
            if (stop > 0 and tm >= stop + every) or tm < nxt:
                continue
            nxt += every
            x = o.get(O.F_X).reshape(-1, 3)
            row = [tm]
            for node, field, comp in items:
                vv = x[node] if field == "coords" else (cur[field][node] if field in vec else None)
                if vv is not None:
                    row.append(np.sqrt(vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]) if comp == 0 else vv[comp - 1])
                elif field == "density":
                    row.append(mat.density if (inc - 1) % 2 == 0 else 0.0)
                else:
                    row.append(o.get(O.F_MASS)[node])
            rows.append(row)
        return o.time()["t"]

    deck.structure_solve(sol.stop, d.save, run_to, lambda t: None)
    want = ["".join("%10.7E " % v for v in r) for r in rows]
    assert text[2:] == want and len(want) >= 5
    dom.close()


def test_deck_mixed_form_runs():
    """the same deck in the mixed form: final coordinates near FP64's"""
    from redynela_amd import deck
    say("case: deck taylor_axi_small in the mixed form")
    d = deck.load(os.path.join(HERE, "decks", "taylor_axi_small.dyn"))
    a = d.run(form="fp64")
    b = d.run(form="mixed")
    xa, xb = a.coords, b.coords
    err = np.max(np.abs(xa - xb)) / np.max(np.abs(xa))
    assert err < 1e-5, err
    a.close()
    b.close()


@pytest.mark.parametrize("parity", ["even", "odd"])
def test_deck_restart_matches_oracle_restart(tmp_path, parity):
    """DynELA_solve's restart (Sources/Applications/DynELA_solve/main.C:106-130):
    the axisymmetric deck is run, then restarted from its result file nfile;
    the files it writes from there on equal, byte for byte, those of the
    oracle restarted from the same file (state read back, mass from the
    restored coordinates, fresh time-step history), and they differ from the
    files of the uninterrupted run, as the reference's restart is no exact
    continuation.  The nodal ro and T columns follow Node::read (FE/Node.C:148):
    the file's values go into the Current field and the other field stays
    cleared, so after k increments from the restart they are the file's when
    k is even and 0 when it is odd -- stated here, not taken from the writer.
    The restart file is the first one written after an even (ro = density)
    or an odd (ro = 0) number of increments."""
    from redynela_amd import deck
    from redynela_amd import results as RF
    d = deck.load(os.path.join(HERE, "decks", "taylor_axi_small.dyn"))
    _, mat, sol = d.model()
    o0 = O.from_deck(d)
    nn0, nipt0 = o0.nn, o0.get(O.F_EVP).size
    want_ro = (lambda r: r != 0.0) if parity == "even" else (lambda r: r == 0.0)
    # the deck's save period, or another one whose save points fall at the parity wanted
    nfile = None
    for nsaves in (3, 4, 5, 7, 9):
        d.save = (0.0, sol.stop, sol.stop / nsaves)
        base = str(tmp_path / ("run%d" % nsaves) / "axi")
        d.run(results=base).close()
        full = sorted(f for f in os.listdir(os.path.dirname(base)) if f.endswith(".bin"))
        cands = [k for k in range(1, len(full))
                 if want_ro(RF.read("%s.%03d.bin" % (base, k), nn0, nipt0)["nodes"][0, 3])]
        if cands:
            nfile = cands[0]
            break
    assert nfile is not None, "no result file of %s parity before the last" % parity
    run_dir = os.path.dirname(base)
    say("case: axi deck restarted from result file %d (%s parity), GPU against the oracle" % (nfile, parity))
    # GPU restart from file nfile into a new base: files nfile + 1 .. are written
    rbase = str(tmp_path / "restart" / "axi")
    os.makedirs(tmp_path / "restart")
    import shutil
    shutil.copy("%s.%03d.bin" % (base, nfile), "%s.%03d.bin" % (rbase, nfile))
    dom = d.run(results=rbase, restart=nfile)
    dom.close()
    # the oracle, restarted from the same file
    o = O.from_deck(d)
    nn, nipt = o.nn, o.get(O.F_EVP).size
    rec = RF.read("%s.%03d.bin" % (base, nfile), nn, nipt)
    nodes, ips = rec["nodes"], rec["ips"]
    ro_file, T_file = nodes[:, 3].copy(), nodes[:, 7].copy()
    for f, c in ((O.F_X, 0), (O.F_V, 8), (O.F_A, 11), (O.F_DU, 17), (O.F_D, 20)):
        o.set(f, nodes[:, c:c + 3])
    o.set(O.F_P, ips[:, 0])
    o.set(O.F_EVP, ips[:, 1])
    o.set(O.F_S, ips[:, 2:11])
    o.set(O.F_EPS, ips[:, 20:29])
    o.set(O.F_EPSPLAS, ips[:, 38:47])
    o.set_time(rec["time"])
    paths = []

    def save(t):
        _, nb, ib = O.result_blocks(o, 0.0, 0.0)
        even = o.time()["increment"] % 2 == 0
        nb[:, 3] = ro_file if even else 0.0
        nb[:, 7] = T_file if even else 0.0
        paths.append(str(tmp_path / ("oracle.%03d.bin" % (nfile + 1 + len(paths)))))
        RF.write(paths[-1], t, nb, ib)

    def solve(t):
        assert o.solve(t) >= 0
        return o.time()["t"]

    save_times = d.save if d.save is not None else (0.0, sol.stop, sol.stop)
    deck.structure_solve(sol.stop, save_times, solve, save, resume=(rec["time"], (nfile + 1) * save_times[2]))
    gpu = sorted(f for f in os.listdir(tmp_path / "restart") if f.endswith(".bin"))[1:]
    assert len(gpu) == len(paths) == len(full) - nfile
    for g, r in zip(gpu, paths):
        assert open(tmp_path / "restart" / g, "rb").read() == open(r, "rb").read(), g
    # not the uninterrupted run's files (fresh mass and time-step history)
    assert open(tmp_path / "restart" / gpu[-1], "rb").read() != open(os.path.join(run_dir, full[-1]), "rb").read()
    say("  restart file %d of %d (save period stop/%d): ro column %s" % (nfile, len(full), nsaves,
        "density" if ro_file[0] != 0.0 else "0 (odd parity)"))
