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
    from redynela_amd import deck
    from redynela_amd import results as RF
    _, mat, _ = d.model()
    o = O.from_deck(d)
    paths = []

    def save(t):
        _, nodes, ips = O.result_blocks(o, mat.density, mat.init_temp)
        paths.append("%s.%03d.bin" % (base, len(paths) + 1))
        RF.write(paths[-1], t, nodes, ips)

    def solve(t):
        assert o.solve(t) >= 0
        return o.time()["t"]

    deck.structure_solve(d.solver.stop, d.save, solve, save)
    return paths


@pytest.mark.parametrize("name", ["taylor_axi_small", "hex_block"])
def test_deck_result_files_match_oracle(tmp_path, name):
    from redynela_amd import deck
    say("case: deck %s run with result files, GPU FP64 against the oracle" % name)
    d = deck.load(os.path.join(HERE, "decks", name + ".dyn"))
    seen = []
    dom = d.run(results=str(tmp_path / "gpu" / name), on_save=lambda t, dm: seen.append(t))
    gpu = sorted(os.listdir(tmp_path / "gpu"))
    ref = oracle_files(d, str(tmp_path / name))
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
