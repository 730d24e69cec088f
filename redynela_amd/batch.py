"""Batched mode: many independent axisymmetric domains advanced by one
persistent GPU launch (one workgroup per domain, each with its own time step
and increment count).

The reference runs one ExplicitSolverCH::solve per Structure
(FE/ExplicitSolverCH.C:115-278, FE/Structure.C:809-897); Batch runs N of them
at once and domain i follows, bit for bit in the FP64 form, the single-domain
run of the same inputs.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import DnlgError
from .mesh import taylor_axi

SEED = 20261203  # the batched Taylor workload's parameter draw


def taylor_params(n, seed=SEED):
    """(nr, v0, sigma_y0, hardening modulus) of n bars, drawn as documented at
    dnlg_batch_taylor_params in include/dnlg.h."""
    nr = np.zeros(n, np.int32)
    v0, sy0, hm = np.zeros(n), np.zeros(n), np.zeros(n)
    _lib.check(_lib.lib().dnlg_batch_taylor_params(n, seed, nr.ctypes.data, v0.ctypes.data, sy0.ctypes.data,
                                                    hm.ctypes.data))
    return nr, v0, sy0, hm


def taylor_batch(n, seed=SEED, radius=3.2e-3, length=32.4e-3):
    """n copper Taylor bars of nr x 5 nr quads with drawn speed and linear
    hardening sigma_y = sigma_y0 + H evp (E, nu, rho of copper).  Returns
    (meshes, materials) as lists of mesh dicts and material dicts."""
    nr, v0, sy0, hm = taylor_params(n, seed)
    meshes, mats = [], []
    for i in range(n):
        meshes.append(taylor_axi(int(nr[i]), 5 * int(nr[i]), radius, length, float(v0[i])))
        mats.append(dict(E=115e9, nu=0.3, rho=8930.0, A=float(sy0[i]), B=float(hm[i]), n=1.0))
    return meshes, mats


class Batch:
    """meshes: list of axisymmetric mesh dicts (see mesh.taylor_axi);
    materials: list of dicts E, nu, rho, A, B, n (E, nu, rho common);
    solver: an ExplicitSolverCH (dissipation, safety, frequency, end time, precision)."""

    def __init__(self, meshes, materials, solver):
        L = _lib.lib()
        n = len(meshes)
        if n == 0 or len(materials) != n:
            raise DnlgError("Batch: one material per mesh, at least one domain")
        self.n = n
        self._keep = []
        cm = (_lib.Mesh * n)()
        ct = (_lib.Material * n)()
        cb = (_lib.BC * n)()
        self.sizes = []
        for i, (m, t) in enumerate(zip(meshes, materials)):
            x = np.ascontiguousarray(m["x"], dtype=np.float64)
            conn = np.ascontiguousarray(m["conn"], dtype=np.int32)
            bc = np.ascontiguousarray(m.get("bc", np.zeros(len(x), np.uint8)), dtype=np.uint8)
            v0 = np.ascontiguousarray(m.get("v0", np.zeros_like(x)), dtype=np.float64)
            self._keep += [x, conn, bc, v0]
            cm[i] = _lib.Mesh(int(m["dim"]), x.shape[0], conn.shape[0], x.ctypes.data, conn.ctypes.data)
            ct[i] = _lib.Material(t["E"], t["nu"], t["rho"], t["A"], t["B"], t["n"], 1, 0.0, 1.0, 1.0, 0.0, 1.0,
                                  0.0, 0.0)
            cb[i] = _lib.BC(bc.ctypes.data, v0.ctypes.data, None, None)
            self.sizes.append((x.shape[0], conn.shape[0]))
        self.form = solver.form
        sp = solver._c()
        h = C.c_void_p()
        rc = L.dnlg_batch_create(n, C.cast(cm, C.c_void_p), C.cast(ct, C.c_void_p), C.cast(cb, C.c_void_p),
                                 C.byref(sp), C.byref(h))
        if rc != 0:
            raise DnlgError("%s: %s" % (_lib.ERRORS.get(rc, rc), L.dnlg_batch_last_error(None).decode()))
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise DnlgError("%s: %s" % (_lib.ERRORS.get(rc, rc), _lib.lib().dnlg_batch_last_error(self.h).decode()))

    ENGINES = {"auto": _lib.BATCH_AUTO, "persistent": _lib.BATCH_PERSISTENT, "stepped": _lib.BATCH_STEPPED,
               "stepped-plain": _lib.BATCH_STEPPED_PLAIN, "stepped-wide": _lib.BATCH_STEPPED_WIDE}

    def set_engine(self, engine):
        """'auto' (the default: the fastest of the next three, probed at the
        first solve), 'persistent' (one workgroup per domain, one launch),
        'stepped' (one increment of every domain per launch group in 64-lane
        blocks, replayed from a captured graph), 'stepped-wide' (the same in
        256-lane blocks) or 'stepped-plain' (stepped without graphs)."""
        self._check(_lib.lib().dnlg_batch_set_engine(self.h, self.ENGINES[engine]))

    def tuning(self):
        """The 'auto' engine's choice after the first solve: (engine name,
        {candidate: probe device time in ms})."""
        e = C.c_int()
        ms = np.zeros(3)
        self._check(_lib.lib().dnlg_batch_get_tuning(self.h, C.byref(e), ms.ctypes.data))
        name = {v: k for k, v in self.ENGINES.items()}[e.value]
        return name, dict(zip(("stepped", "stepped-wide", "persistent"), ms.tolist()))

    def solve(self, upTime, max_steps=200000):
        """Every domain to upTime, at most max_steps increments each (a domain
        that needs more raises); returns (per-domain device wall time in us,
        launch time in ms)."""
        w = np.zeros(self.n)
        ms = C.c_double()
        self._check(_lib.lib().dnlg_batch_solve(self.h, float(upTime), int(max_steps), w.ctypes.data, C.byref(ms)))
        return w, ms.value

    def time(self, i):
        out = np.zeros(9)
        self._check(_lib.lib().dnlg_batch_get_time(self.h, i, out.ctypes.data))
        return dict(t=out[0], dt=out[1], increment=int(out[2]), computed_dt=out[3], ke=out[4],
                    alphaM=out[5], beta=out[6], gamma=out[7], omegaS=out[8])

    def field(self, i, f):
        nn, ne = self.sizes[i]
        mixed = self.form == _lib.FORM_MIXED
        size = {_lib.F_COORDS: nn * 3, _lib.F_V: nn * 3, _lib.F_A: nn * 3, _lib.F_DU: nn * 3, _lib.F_D: nn * 3,
                _lib.F_MASS: nn, _lib.F_P: ne, _lib.F_EVP: ne * 4,
                _lib.F_S: ne * 4 * (6 if mixed else 9)}[f]
        out = np.zeros(size)
        self._check(_lib.lib().dnlg_batch_get_field(self.h, i, f, out.ctypes.data, size))
        return out

    def coords(self, i):
        return self.field(i, _lib.F_COORDS).reshape(-1, 3)

    def evp(self, i):
        """Equivalent plastic strain, [ne*4] in element-major point order."""
        e = self.field(i, _lib.F_EVP)
        if self.form == _lib.FORM_MIXED:
            e = e.reshape(4, -1).T.reshape(-1)
        return e

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().dnlg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
