"""Host-side mirror of the reference's Domain / Material / ExplicitSolverCH
interface over the libdnlgpu C ABI.

Names, argument meaning and error behaviour follow the reference:
  Domain::computeInternalForces()  FE/Domain.C:130-738   -> Domain.computeInternalForces(), Domain.Fint
  ExplicitSolverCH::solve(upTime)  FE/ExplicitSolverCH.C:115-278 -> ExplicitSolverCH.solve(upTime)
  ExplicitSolverCH::setDissipation FE/ExplicitSolverCH.C:357-380 (raises on a value outside [0, 1])
  IsoHardElastoplastic / IsoHardJohnsonCook set* (FE/IsoHardElastoplastic.C, FE/IsoHardJohnsonCook.C)
A failing kernel raises DnlgError carrying the reference's fatalError text
(e.g. "Too many atempts in plastic procedure").
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import DnlgError


class IsoHardElastoplastic:
    """sigma_y = A + B evp^n (FE/IsoHardElastoplastic.C:171-184)."""

    def __init__(self, A=0.0, B=0.0, n=1.0):
        self.A, self.B, self.n = float(A), float(B), float(n)

    def setYieldStress(self, v):
        self.A = float(v)

    def setHardParameter(self, v):
        self.B = float(v)

    def setHardExponent(self, v):
        self.n = float(v)


class IsoHardJohnsonCook(IsoHardElastoplastic):
    """The reference's Johnson-Cook law evaluates A + B evp^n only
    (FE/IsoHardJohnsonCook.C:185-198); rate and temperature parameters are
    stored, as in FE/IsoHardJohnsonCook.h:79, and not used on this path."""

    def __init__(self, A=0.0, B=0.0, n=1.0, C=0.0, m=1.0, depsp0=1.0, Tm=0.0, T0=0.0):
        super().__init__(A, B, n)
        self.C, self.m, self.depsp0, self.Tm, self.T0 = C, m, depsp0, Tm, T0
        self.cp, self.taylor_quinney = 0.0, 0.0
        self.rate_and_thermal = False

    def defineLaw(self, A, B, C, n, m, epsp0, Tm, T0):
        """IsoHardJohnsonCook::defineLaw argument order (FE/IsoHardJohnsonCook.h)."""
        self.A, self.B, self.C, self.n, self.m, self.depsp0, self.Tm, self.T0 = A, B, C, n, m, epsp0, Tm, T0

    def enableRateAndThermal(self, cp, taylor_quinney=0.9):
        """Build extension: evaluate the rate and temperature factors of
        FE/IsoHardJohnsonCook.h:51-68 with adiabatic heating."""
        self.cp, self.taylor_quinney = float(cp), float(taylor_quinney)
        self.rate_and_thermal = True


class HistoryFile:
    """HistoryFile (FE/HistoryFile.C): nodal series written at endStep every
    save period.  append(nodes, field, comp) takes node indices (and, for
    the column names, their numbers), a Node::getNodalValue field name
    (FE/Node.C:326-342) and a component (0 the norm of a vector, 1..3 one
    component)."""

    def __init__(self, name=""):
        self.name = name
        self.fileName = ""
        self.start = self.every = 0.0
        self.stop = 0.0
        self.items = []  # (node index, field, comp, column name)
        self.domain, self.id = None, None

    def setName(self, name):
        self.fileName = name

    def setSaveTime(self, *a):
        """(period) or (start, stop, period), FE/HistoryFile.C:241-259"""
        if len(a) == 1:
            self.start = self.every = float(a[0])
            self.stop = 0.0
        else:
            self.start, self.stop, self.every = float(a[0]), float(a[1]), float(a[2])

    def append(self, nodes, field, comp=0, numbers=None):
        """HistoryFile::append(NodeSet*, String, Indice) (FE/HistoryFile.C:192-239)"""
        if field not in _lib.HF_FIELDS:
            raise DnlgError("getNodalValue: undefined field %s" % field)
        nodes = list(nodes)
        numbers = list(numbers) if numbers is not None else [i + 1 for i in nodes]
        base = field + ("_%d" % comp if comp != 0 else "")
        for i, num in zip(nodes, numbers):
            self.items.append((int(i), field, int(comp), base + "_n%d" % num if len(nodes) > 1 else base))

    @property
    def columns(self):
        return [it[3] for it in self.items]


class Material:
    def __init__(self, young=0.0, poisson=0.0, density=0.0, hardening=None):
        self.young, self.poisson, self.density = float(young), float(poisson), float(density)
        self.hardening = hardening

    def setYoung(self, v):
        self.young = float(v)

    def setPoisson(self, v):
        self.poisson = float(v)

    def setDensity(self, v):
        self.density = float(v)

    def setHardening(self, h):
        self.hardening = h

    def _c(self):
        h = self.hardening
        jc = isinstance(h, IsoHardJohnsonCook) and h.rate_and_thermal
        law = (2 if jc else 1) if h else 0
        return _lib.Material(self.young, self.poisson, self.density, h.A if h else 0.0, h.B if h else 0.0,
                             h.n if h else 1.0, law,
                             h.C if jc else 0.0, h.depsp0 if jc else 1.0, h.m if jc else 1.0,
                             h.T0 if jc else 0.0, h.Tm if jc else 1.0, h.cp if jc else 0.0,
                             h.taylor_quinney if jc else 0.0)


class ExplicitSolverCH:
    """Chung-Hulbert explicit solver parameters; default spectral radius 0.8182
    (FE/DefaultValues.h:25), safety factor 0.9, time step recomputed every 10
    increments (FE/DefaultValues.h:24) by the Pulse method, the reference's
    default for this solver (FE/ExplicitSolverCH.C:83)."""

    def __init__(self):
        self.rho_b = 0.8182
        self.safety = 0.9
        self.freq = 10
        self.start_time, self.end_time = 0.0, 1e300
        self.form = _lib.FORM_FP64
        self.method = "Pulse"
        self.domain = None

    def setPrecision(self, form):
        """'fp64' (the reference's arithmetic) or 'mixed' (FP32 element pass, FP64 nodal state)."""
        self.form = {"fp64": _lib.FORM_FP64, "mixed": _lib.FORM_MIXED}[form]

    def setDissipation(self, v):
        if v < 0.0 or v > 1.0:
            raise DnlgError("ExplicitSolverCH::setDissipation: dissipation value out of range [0.0, 1.0]")
        self.rho_b = float(v)

    def setTimes(self, start, stop):
        self.start_time, self.end_time = float(start), float(stop)

    def setTimeStepSafetyFactor(self, v):
        self.safety = float(v)

    def setComputeTimeStepFrequency(self, n):
        self.freq = int(n)

    def setTimeStepMethod(self, name):
        """'Courant' (characteristic length) or 'Pulse' (power iteration on
        M^-1 K), Solver::setTimeStepMethod (FE/Solver.C:312-327)."""
        if name not in ("Courant", "Pulse"):
            raise DnlgError("setTimeStepMethod: Unknown method %s" % name)
        self.method = name
        if self.domain is not None:
            self.domain.set_option(_lib.OPT_TIME_STEP_METHOD, _lib.TS_PULSE if name == "Pulse" else _lib.TS_COURANT)

    def solve(self, upTime):
        return self.domain.solve(upTime)

    def _c(self):
        return _lib.SolverParams(self.rho_b, self.safety, self.freq, self.end_time, self.form)


class Domain:
    """One domain on one GPU.  mesh: dict(dim, x, conn, bc, v0) (see mesh.py)."""

    def __init__(self, mesh, material, solver, speed_mask=None, speed=None, track_strain=False):
        L = _lib.lib()
        self.dim = int(mesh["dim"])
        self._x = np.ascontiguousarray(mesh["x"], dtype=np.float64)
        self._conn = np.ascontiguousarray(mesh["conn"], dtype=np.int32)
        self._bc = np.ascontiguousarray(mesh.get("bc", np.zeros(len(self._x), np.uint8)), dtype=np.uint8)
        self._v0 = np.ascontiguousarray(mesh.get("v0", np.zeros_like(self._x)), dtype=np.float64)
        self.n_nodes, self.n_elems = self._x.shape[0], self._conn.shape[0]
        self.nen = self._conn.shape[1]
        self.nip = self.nen
        m = _lib.Mesh(self.dim, self.n_nodes, self.n_elems, self._x.ctypes.data, self._conn.ctypes.data)
        self._sm = None if speed_mask is None else np.ascontiguousarray(speed_mask, dtype=np.uint8)
        self._sp = None if speed is None else np.ascontiguousarray(speed, dtype=np.float64)
        b = _lib.BC(self._bc.ctypes.data, self._v0.ctypes.data, _lib.ptr(self._sm), _lib.ptr(self._sp))
        self.material = material
        self.solver = solver
        solver.domain = self
        mat = material._c()
        sp = solver._c()
        h = C.c_void_p()
        rc = L.dnlg_create(C.byref(m), C.byref(mat), C.byref(b), C.byref(sp), C.byref(h))
        if rc != 0:
            raise DnlgError("%s: %s" % (_lib.ERRORS.get(rc, rc), L.dnlg_last_error(None).decode()))
        self.h = h
        if track_strain:
            self.set_option(_lib.OPT_TRACK_STRAIN, 1)
        if solver.method == "Pulse":
            self.set_option(_lib.OPT_TIME_STEP_METHOD, _lib.TS_PULSE)

    # -- reference interface -------------------------------------------------
    def computeInternalForces(self, coords=None, delta_disp=None):
        """Domain::computeInternalForces(): stress update + Fint assembly.

        With coords and/or delta_disp ([n_nodes, 3]) the caller's node
        coordinates and Current displacement increment replace the device
        copies first, as when the caller runs the time loop itself
        (FE/ExplicitSolverCH.C:471-481)."""
        out = np.zeros((self.n_nodes, self.dim))
        L = _lib.lib()
        if coords is None and delta_disp is None:
            _lib.check(L.dnlg_compute_internal_forces(self.h, out.ctypes.data), self.h)
            return out
        x = None if coords is None else np.ascontiguousarray(coords, dtype=np.float64).reshape(self.n_nodes, 3)
        u = None if delta_disp is None else np.ascontiguousarray(delta_disp, dtype=np.float64).reshape(self.n_nodes, 3)
        _lib.check(L.dnlg_compute_internal_forces_from(self.h, _lib.ptr(x), _lib.ptr(u), out.ctypes.data), self.h)
        return out

    def computeCourantTimeStep(self):
        """Domain::computeCourantTimeStep(): min over elements of l / c (FE/Domain.C:1613-1646)."""
        c = C.c_double()
        _lib.check(_lib.lib().dnlg_courant_time_step(self.h, C.byref(c)), self.h)
        return c.value

    @property
    def Fint(self):
        return self.field(_lib.F_FINT).reshape(self.n_nodes, self.dim)

    def solve(self, upTime):
        """ExplicitSolverCH::solve(upTime); returns the number of increments run."""
        n = C.c_longlong()
        _lib.check(_lib.lib().dnlg_solve(self.h, float(upTime), C.byref(n)), self.h)
        return n.value

    # -- stepping ------------------------------------------------------------
    def begin(self, upTime):
        r = C.c_int()
        _lib.check(_lib.lib().dnlg_begin(self.h, float(upTime), C.byref(r)), self.h)
        return bool(r.value)

    def step(self, n=1):
        r = C.c_int()
        _lib.check(_lib.lib().dnlg_step(self.h, int(n), C.byref(r)), self.h)
        return bool(r.value)

    def step_async(self, n):
        _lib.check(_lib.lib().dnlg_step_async(self.h, int(n)), self.h)

    def synchronize(self):
        _lib.check(_lib.lib().dnlg_synchronize(self.h), self.h)

    def set_option(self, opt, value):
        _lib.check(_lib.lib().dnlg_set_option(self.h, opt, int(value)), self.h)

    # -- state ---------------------------------------------------------------
    def field(self, f):
        L = _lib.lib()
        n = L.dnlg_field_size(self.h, f)
        if n <= 0:
            raise DnlgError("field %d not available" % f)
        out = np.zeros(n)
        _lib.check(L.dnlg_get_field(self.h, f, out.ctypes.data, n), self.h)
        return out

    def set_time(self, t):
        """Structure::readData's currentTime on a restart (before the first solve)."""
        _lib.check(_lib.lib().dnlg_set_time(self.h, float(t)), self.h)

    def set_field(self, f, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _lib.check(_lib.lib().dnlg_set_field(self.h, f, a.ctypes.data, a.size), self.h)

    @property
    def coords(self):
        return self.field(_lib.F_COORDS).reshape(-1, 3)

    @property
    def evp(self):
        return self.field(_lib.F_EVP)

    def time(self):
        out = np.zeros(9)
        _lib.check(_lib.lib().dnlg_get_time(self.h, out.ctypes.data), self.h)
        return dict(t=out[0], dt=out[1], increment=int(out[2]), computed_dt=out[3], ke=out[4],
                    alphaM=out[5], beta=out[6], gamma=out[7], omegaS=out[8])

    def addHistoryFile(self, hf, init_temp=0.0):
        """Structure::addHistoryFile for this domain (FE/Structure.C:601-613)."""
        n = len(hf.items)
        nodes = np.array([it[0] for it in hf.items], dtype=np.int64)
        fields = np.array([_lib.HF_FIELDS[it[1]][0] for it in hf.items], dtype=np.int32)
        comps = np.array([it[2] for it in hf.items], dtype=np.int32)
        i = C.c_int()
        _lib.check(_lib.lib().dnlg_history_file_add(self.h, hf.start, hf.stop, hf.every, n, nodes.ctypes.data,
                                                    fields.ctypes.data, comps.ctypes.data, float(init_temp),
                                                    C.byref(i)), self.h)
        hf.domain, hf.id = self, i.value
        return hf

    def history_rows(self, hf):
        """rows written since the last call: [n][1 + items] (time, values)."""
        L = _lib.lib()
        w = 1 + len(hf.items)
        n = L.dnlg_history_file_rows(self.h, hf.id, None, 0)
        if n < 0:
            _lib.check(-n, self.h)
        out = np.zeros((n, w))
        if n:
            got = L.dnlg_history_file_rows(self.h, hf.id, out.ctypes.data, n)
            out = out[:got]
        return out

    def plot_rows(self):
        """endStep history lines since the last call (DNLG_OPT_PLOT_HISTORY on):
        [n][4] time, time step, computed time step, kinetic energy."""
        L = _lib.lib()
        n = L.dnlg_plot_rows(self.h, None, 0)
        if n < 0:
            _lib.check(-n, self.h)
        out = np.zeros((n, 4))
        if n:
            got = L.dnlg_plot_rows(self.h, out.ctypes.data, n)
            out = out[:got]
        return out

    def pulse_info(self):
        """Pulse time step: calls, iterations of the last call, fmax0 and its convergence measure."""
        out = np.zeros(4)
        _lib.check(_lib.lib().dnlg_pulse_info(self.h, out.ctypes.data), self.h)
        return dict(calls=int(out[0]), iterations=int(out[1]), fmax0=out[2], convergence=out[3])

    def enable_kernel_timing(self, on=True):
        _lib.check(_lib.lib().dnlg_enable_kernel_timing(self.h, 1 if on else 0), self.h)

    def kernel_stats(self, which):
        n, ms = C.c_longlong(), C.c_double()
        _lib.check(_lib.lib().dnlg_kernel_stats(self.h, which, C.byref(n), C.byref(ms)), self.h)
        return n.value, ms.value

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().dnlg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
