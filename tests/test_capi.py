"""The C ABI library: it loads without a GPU, exports every function that
include/dnlg.h declares, and its host-only mesh generators are right."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "dnlg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dnlg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from redynela_amd import _lib
    L = _lib.lib()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTS)
    assert L.dnlg_abi_version() == 1


def test_create_rejects_bad_arguments_without_touching_the_gpu():
    from redynela_amd import _lib
    L = _lib.lib()
    h = C.c_void_p()
    assert L.dnlg_create(None, None, None, None, C.byref(h)) == 1
    assert b"bad mesh" in L.dnlg_last_error(None)
    x = np.zeros((4, 3))
    conn = np.array([[0, 1, 2, 7]], dtype=np.int32)  # node 7 does not exist
    mesh = _lib.Mesh(2, 4, 1, x.ctypes.data, conn.ctypes.data)
    mat = _lib.Material(115e9, 0.3, 8930.0, 4e8, 1e8, 1.0, 1)
    sp = _lib.SolverParams(1.0, 0.9, 10, 1.0, 0)
    assert L.dnlg_create(C.byref(mesh), C.byref(mat), None, C.byref(sp), C.byref(h)) == 1
    assert b"connectivity" in L.dnlg_last_error(None)
    sp = _lib.SolverParams(1.5, 0.9, 10, 1.0, 0)
    conn = np.array([[0, 1, 2, 3]], dtype=np.int32)
    mesh = _lib.Mesh(2, 4, 1, x.ctypes.data, conn.ctypes.data)
    assert L.dnlg_create(C.byref(mesh), C.byref(mat), None, C.byref(sp), C.byref(h)) == 1
    assert b"dissipation" in L.dnlg_last_error(None)


def test_set_dissipation_range_like_the_reference():
    import redynela_amd as R
    s = R.ExplicitSolverCH()
    s.setTimeStepMethod("Courant")
    with pytest.raises(R.DnlgError):
        s.setDissipation(1.2)
    s.setDissipation(1.0)
    assert s.rho_b == 1.0


def test_axi_mesh_matches_sample_deck_recipe():
    import redynela_amd as R
    m = R.mesh.taylor_axi(25, 250, 3.2e-3, 32.4e-3, 227.0)  # the sample deck's 25 x 250 bar
    assert m["x"].shape == (26 * 251, 3) and m["conn"].shape == (6250, 4)
    # deck numbering x1 = i + j(nl+1) + 1 ... (1-based)
    i, j = 3, 7
    e = j * 25 + i
    assert list(m["conn"][e] + 1) == [i + j * 26 + 1, i + j * 26 + 2, i + (j + 1) * 26 + 2, i + (j + 1) * 26 + 1]
    assert np.allclose(m["x"][i + j * 26], [i * (3.2e-3 / 25), j * (32.4e-3 / 250), 0])
    assert np.all(m["bc"][:26] & 2) and np.all(m["bc"][::26] & 5 == 5)
    assert np.all(m["v0"][:, 1] == -227.0)


def test_ogrid_mesh_volume_orientation_and_counts():
    import redynela_amd as R
    nn, ne = R.mesh.ogrid_sizes(20, 20, 500)
    assert (nn, ne) == (1022541, 1000000)
    m = R.mesh.ogrid_bar(20, 20, 2, 3.2e-3, 0.13e-3, 0.4, 227.0)
    mod = O.Model(m, dict(E=115e9, nu=0.3, rho=8930.0, A=4e8, B=1e8, n=1.0), {})
    mod.begin(1.0)
    M = mod.get(O.F_MASS)
    # the polygonal section approaches pi R^2 from below
    V = M.sum() / 8930.0
    Vc = np.pi * 3.2e-3 ** 2 * 0.13e-3
    assert 0.995 * Vc < V < Vc
    assert np.all(M > 0)
    assert np.all(m["bc"][:2041] == 4) and np.all(m["bc"][2041:] == 0)


def test_ctypes_structs_match_the_header_layout():
    """Compile a probe against include/dnlg.h and compare sizes/offsets with the ctypes mirror."""
    import subprocess
    import tempfile
    from redynela_amd import _lib
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "dnlg.h"
#define P(T, F) printf("%s.%s %zu\n", #T, #F, offsetof(T, F));
int main(void) {
  printf("dnlg_mesh %zu\ndnlg_material %zu\ndnlg_bc %zu\ndnlg_solver_params %zu\n", sizeof(dnlg_mesh),
         sizeof(dnlg_material), sizeof(dnlg_bc), sizeof(dnlg_solver_params));
  P(dnlg_material, hardening) P(dnlg_material, C) P(dnlg_material, taylor_quinney)
  P(dnlg_solver_params, end_time) P(dnlg_solver_params, form) P(dnlg_mesh, coords)
  return 0; }
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "p")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = dict(l.rsplit(" ", 1) for l in subprocess.check_output([exe]).decode().split("\n") if l)
    py = {"dnlg_mesh": C.sizeof(_lib.Mesh), "dnlg_material": C.sizeof(_lib.Material),
          "dnlg_bc": C.sizeof(_lib.BC), "dnlg_solver_params": C.sizeof(_lib.SolverParams),
          "dnlg_material.hardening": _lib.Material.hardening.offset, "dnlg_material.C": _lib.Material.C.offset,
          "dnlg_material.taylor_quinney": _lib.Material.taylor_quinney.offset,
          "dnlg_solver_params.end_time": _lib.SolverParams.end_time.offset,
          "dnlg_solver_params.form": _lib.SolverParams.form.offset, "dnlg_mesh.coords": _lib.Mesh.coords.offset}
    for k, v in py.items():
        assert int(out[k]) == v, (k, out[k], v)


def test_batch_create_rejects_bad_arguments_without_touching_the_gpu():
    """dnlg_batch_create validates every domain before any device call; the
    other batch entry points reject a null handle"""
    from redynela_amd import _lib
    L = _lib.lib()
    b = C.c_void_p()
    x = np.array([[0, 0, 0], [1e-3, 0, 0], [1e-3, 1e-3, 0], [0, 1e-3, 0]], dtype=np.float64)
    conn = np.array([[0, 1, 2, 3]], dtype=np.int32)
    good = _lib.Mesh(2, 4, 1, x.ctypes.data, conn.ctypes.data)
    mat = _lib.Material(115e9, 0.3, 8930.0, 4e8, 1e8, 1.0, 1)
    sp = _lib.SolverParams(1.0, 0.9, 10, 1.0, 0)

    def create(meshes, mats, solver, n=None):
        ma = (_lib.Mesh * len(meshes))(*meshes)
        mt = (_lib.Material * len(mats))(*mats)
        return L.dnlg_batch_create(len(meshes) if n is None else n, ma, mt, None, C.byref(solver), C.byref(b))

    assert create([good], [mat], sp, n=0) == 1
    assert b"bad arguments" in L.dnlg_batch_last_error(None)
    assert create([good], [mat], _lib.SolverParams(1.0, 0.9, 10, 1.0, 7)) == 1
    hexmesh = _lib.Mesh(3, 4, 1, x.ctypes.data, conn.ctypes.data)
    assert create([good, hexmesh], [mat, mat], sp) == 1
    assert b"axisymmetric quad4" in L.dnlg_batch_last_error(None)
    other_e = _lib.Material(200e9, 0.3, 8930.0, 4e8, 1e8, 1.0, 1)
    assert create([good, good], [mat, other_e], sp) == 1
    assert b"share E, nu and density" in L.dnlg_batch_last_error(None)
    bad_conn = np.array([[0, 1, 2, 4]], dtype=np.int32)
    outside = _lib.Mesh(2, 4, 1, x.ctypes.data, bad_conn.ctypes.data)
    assert create([good, outside], [mat, mat], sp) == 1
    assert b"outside its domain" in L.dnlg_batch_last_error(None)
    assert create([good], [mat], _lib.SolverParams(1.2, 0.9, 10, 1.0, 0)) == 1
    assert b"dissipation" in L.dnlg_batch_last_error(None)
    assert not b.value
    assert L.dnlg_batch_set_engine(None, 0) == 1
    assert L.dnlg_batch_get_tuning(None, None, None) == 1
    assert L.dnlg_batch_solve(None, 1e-6, 10, None, None) == 1


def test_compile_time_shape_tables_equal_the_oracle_tables():
    """The FP64 element kernels copy their shape-function tables from
    compile-time constants (redynela_amd/csrc/dnlg_tables.hpp), written as
    hexadecimal literals and expanded by constexpr code. tests/capi/shape_tables
    prints them; every entry must be the oracle's dN0 (and, for quad4, N) of the
    same Gauss point and node, bit for bit (the oracle's tables are pinned to
    the trilinear / bilinear closed forms in test_oracle_element.py)."""
    import subprocess
    prog = os.path.join(ROOT, "tests", "capi", "shape_tables")
    if not os.path.exists(prog):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "capi"), "shape_tables"])
    vals = [float.fromhex(t) for t in subprocess.check_output([prog], text=True).split()]
    assert len(vals) == 196 + 48
    hexv, quadv = vals[:196], vals[196:]
    L = O.lib()
    N8, dN8 = np.zeros((8, 8)), np.zeros((8, 8, 3))
    L.dnlo_shape_tables(3, N8.reshape(-1), dN8.reshape(-1))
    for ip in range(8):
        for k in range(8):
            for d in range(3):
                got = hexv[(ip * 8 + k) * 3 + (4 if ip >= 4 else 0) + d]
                assert got == dN8[ip, k, d] and np.signbit(got) == np.signbit(dN8[ip, k, d]), (ip, k, d)
    N4, dN4 = np.zeros((4, 4)), np.zeros((4, 4, 2))
    L.dnlo_shape_tables(2, N4.reshape(-1), dN4.reshape(-1))
    for ip in range(4):
        for k in range(4):
            assert quadv[(ip * 4 + k) * 3 + 0] == dN4[ip, k, 0]
            assert quadv[(ip * 4 + k) * 3 + 1] == dN4[ip, k, 1]
            assert quadv[(ip * 4 + k) * 3 + 2] == N4[ip, k]


def test_missing_library_fails_loudly():
    """without the built HIP library every entry point raises: there is no
    CPU path behind the Domain API (run in a child process, so that this
    process keeps its loaded library)"""
    import subprocess
    import sys
    code = ("import redynela_amd as R\n"
            "from redynela_amd import _lib as L\n"
            "L.LIB_PATH = '/nonexistent/libdnlgpu.so'\n"
            "m = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))\n"
            "s = R.ExplicitSolverCH()\n"
            "try:\n"
            "    R.Domain(R.mesh.taylor_axi(2, 4), m, s)\n"
            "except L.DnlgError as e:\n"
            "    print('raised:', e)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert "raised: libdnlgpu.so is not built" in r.stdout, (r.stdout, r.stderr)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; a request for more exits with a message,
    before anything touches a device"""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "--gpus 2 is not supported" in (r.stdout + r.stderr)
