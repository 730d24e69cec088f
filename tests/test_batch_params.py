"""The batched Taylor workload's parameter draw (host only, no GPU): the C
library's dnlg_batch_taylor_params reproduces the committed 512-bar table,
which an independent pure-Python MT19937-64 wrote; that generator is pinned
by the C++ standard's known answer for std::mt19937_64."""
import csv
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_batch_params as P  # noqa: E402


def test_mt19937_64_known_answer():
    # [rand.predef]: the 10000th invocation of a default-constructed mt19937_64 (seed 5489)
    g = P.MT64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_library_draw_matches_committed_table():
    from redynela_amd import batch
    rows = list(csv.DictReader(open(os.path.join(HERE, "golden", "batch_taylor_params.csv"))))
    assert len(rows) == 512
    nr, v0, sy, hm = batch.taylor_params(512)
    assert np.array_equal(nr, [int(r["nr"]) for r in rows])
    assert np.array_equal(v0, [float(r["v0"]) for r in rows])
    assert np.array_equal(sy, [float(r["sigma_y0"]) for r in rows])
    assert np.array_equal(hm, [float(r["hardening_modulus"]) for r in rows])
    # the workload's shape: 125..2000 quads per bar, uneven
    ne = 5 * nr * nr
    assert ne.min() == 125 and ne.max() == 2000
    assert 150 <= v0.min() and v0.max() < 300 and 300e6 <= sy.min() and sy.max() < 500e6


def test_batch_create_rejects_mixed_meshes_without_touching_the_gpu():
    import ctypes as C
    from redynela_amd import _lib
    L = _lib.lib()
    h = C.c_void_p()
    x = np.zeros((8, 3))
    conn = np.array([[0, 1, 2, 3, 4, 5, 6, 7]], dtype=np.int32)
    m = (_lib.Mesh * 1)(_lib.Mesh(3, 8, 1, x.ctypes.data, conn.ctypes.data))
    t = (_lib.Material * 1)(_lib.Material(115e9, 0.3, 8930.0, 4e8, 1e8, 1.0, 1))
    sp = _lib.SolverParams(1.0, 0.9, 10, 1.0, 0)
    assert L.dnlg_batch_create(1, C.cast(m, C.c_void_p), C.cast(t, C.c_void_p), None, C.byref(sp), C.byref(h)) == 1
    assert b"axisymmetric" in L.dnlg_batch_last_error(None)
    assert L.dnlg_batch_create(0, None, None, None, C.byref(sp), C.byref(h)) == 1
