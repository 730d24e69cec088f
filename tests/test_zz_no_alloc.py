"""dnlg_courant_time_step_noalloc against dnlg_courant_time_step, bit for bit.

The no-alloc entry point takes device buffers from the caller (the test
allocates them with hipMalloc; the entry point itself allocates nothing) and keeps its
per-block minima in the library's file-scope scratch array.  Two calls in a
row, the second on a smaller mesh, must each equal the handle-based entry
point on the same mesh; a mesh beyond the scratch's capacity is refused on
the host before any launch.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def handle_crit(mesh):
    import redynela_amd as R
    m = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    s.setDissipation(1.0)
    g = R.Domain(mesh, m, s)
    c = g.computeCourantTimeStep()
    g.close()
    return c


def noalloc_crit(mesh):
    from hip_mem import DeviceArray
    from redynela_amd import _lib as L
    x = DeviceArray(np.asarray(mesh["x"], dtype=np.float64))
    conn = DeviceArray(np.asarray(mesh["conn"], dtype=np.int32))
    out = DeviceArray(np.full(1, -1.0))
    mat = L.Material(115e9, 0.3, 8930.0, 400e6, 100e6, 1.0, 1)
    rc = L.lib().dnlg_courant_time_step_noalloc(int(mesh["dim"]), int(mesh["conn"].shape[0]), x.ptr, conn.ptr,
                                                C.byref(mat), out.ptr, None)
    assert rc == 0, L.lib().dnlg_noalloc_last_error()
    crit = float(out.get()[0])
    for b in (x, conn, out):
        b.free()
    return crit


def test_noalloc_courant_equals_handle_entry_point_twice():
    import redynela_amd as R
    say("case: dnlg_courant_time_step_noalloc, hex8 O-grid 20+20 x 40 layers (80,000 elements), then axi 10x50")
    big = R.mesh.ogrid_bar(20, 20, 40, 3.2e-3, 32.4e-3, 0.4, 227.0)
    small = R.mesh.taylor_axi(10, 50, 3.2e-3, 32.4e-3, 227.0)
    for mesh, what in ((big, "hex8 80,000"), (small, "axi 500")):
        ref = handle_crit(mesh)
        got = noalloc_crit(mesh)
        say("%s: handle %.17g, no-alloc %.17g" % (what, ref, got))
        assert got == ref and ref > 0


def test_noalloc_refuses_more_elements_than_the_scratch_holds():
    from redynela_amd import _lib as L
    say("case: dnlg_courant_time_step_noalloc, n_elems above 16,777,216 (host check, nothing launched)")
    mat = L.Material(115e9, 0.3, 8930.0, 400e6, 100e6, 1.0, 1)
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    rc = L.lib().dnlg_courant_time_step_noalloc(3, 16777217, dummy, dummy, C.byref(mat), dummy, None)
    assert rc == 1
    assert b"16777216" in L.lib().dnlg_noalloc_last_error()
