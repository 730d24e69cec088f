"""Adjoint of the axisymmetric or hex8 explicit solve: gradients of a function of the
final nodal coordinates with respect to the hardening parameters (A, B, n) of
sigma_y = A + B evp^n -- or of the Johnson-Cook law with its temperature
factor and adiabatic heating but no rate term, sigma_y = (A + B evp^n)
(1 - T*^m) -- by the hand-written reverse sweep of libdnlgpu
(dnlg_adjoint_forward / dnlg_adjoint_backward) with checkpoint-recompute.

FinalShape is the torch.autograd.Function over it: final (r, z) of every
node of an axisymmetric mesh as a function of theta = (A, B, n).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import DnlgError
from .domain import Domain, ExplicitSolverCH, IsoHardElastoplastic, IsoHardJohnsonCook, Material
from .mesh import taylor_axi

# the parameter-identification workload: copper, Johnson-Cook hardening
# without rate or temperature terms, 20 x 100 quads, 200 m/s, 4,000 increments
TARGET_THETA = np.array([90e6, 292e6, 0.31])
ADJ_STEPS, ADJ_CKPT = 4000, 100
# the same law with the temperature factor and adiabatic heating of the
# precision workload (m = 1.09, T0 = 293 K, Tm = 1356 K, cp = 383 J/kg/K,
# Taylor-Quinney 0.9), still without the rate term
THERMAL = dict(m=1.09, T0=293.0, Tm=1356.0, cp=383.0, tq=0.9)


def bar_20x100():
    return taylor_axi(20, 100, 3.2e-3, 32.4e-3, 200.0)


def perturbed_start(seed=20261203):
    """(A, B, n) perturbed by a factor 1 + U(-0.2, 0.2) each (numpy PCG64 from seed)."""
    u = np.random.Generator(np.random.PCG64(seed)).uniform(-0.2, 0.2, 3)
    return TARGET_THETA * (1.0 + u)


class HardeningAdjoint:
    """One mesh and solver setting; forward(theta) runs nsteps increments from
    the initial state, backward(xbar) returns dL/dtheta for the adjoint xbar
    of the final coordinates.  thermal=dict(m, T0, Tm, cp, tq) selects the
    Johnson-Cook temperature factor with adiabatic heating (no rate term)."""

    def __init__(self, mesh, E=115e9, nu=0.3, rho=8930.0, nsteps=ADJ_STEPS, ckpt_every=ADJ_CKPT, rho_b=1.0,
                 thermal=None):
        self.mesh, self.E, self.nu, self.rho = mesh, E, nu, rho
        self.nsteps, self.ckpt_every, self.rho_b = int(nsteps), int(ckpt_every), rho_b
        self.thermal = dict(thermal) if thermal else None
        self.dom = None
        self.stats = None

    def forward(self, theta):
        A, B, n = (float(t) for t in theta)
        if self.dom is not None:
            self.dom.close()
        s = ExplicitSolverCH()
        s.setTimeStepMethod("Courant")
        s.setDissipation(self.rho_b)
        if self.thermal:
            t = self.thermal
            hard = IsoHardJohnsonCook(A, B, n, C=0.0, m=t["m"], Tm=t["Tm"], T0=t["T0"])
            hard.enableRateAndThermal(t["cp"], t["tq"])
        else:
            hard = IsoHardElastoplastic(A, B, n)
        self.dom = Domain(self.mesh, Material(self.E, self.nu, self.rho, hard), s)
        self.dom.set_option(_lib.OPT_STORE_FINT, 0)
        _lib.check(_lib.lib().dnlg_adjoint_forward(self.dom.h, self.nsteps, self.ckpt_every), self.dom.h)
        return self.dom.coords

    def backward(self, xbar):
        if self.dom is None:
            raise DnlgError("HardeningAdjoint.backward: call forward first")
        xb = np.zeros((self.dom.n_nodes, 3))
        xbar = np.asarray(xbar, dtype=np.float64)
        xb[:, : xbar.shape[1]] = xbar
        g = np.zeros(3)
        st = _lib.AdjointStats()
        _lib.check(_lib.lib().dnlg_adjoint_backward(self.dom.h, xb.ctypes.data, g.ctypes.data, C.byref(st)),
                   self.dom.h)
        self.stats = {f: getattr(st, f) for f, _ in _lib.AdjointStats._fields_}
        return g

    def close(self):
        if self.dom is not None:
            self.dom.close()
            self.dom = None


def _final_shape_class():
    import torch

    class FinalShape(torch.autograd.Function):
        """x_final[:, :2] = FinalShape.apply(theta, problem), theta = (A, B, n) float64."""

        @staticmethod
        def forward(ctx, theta, problem):
            x = problem.forward(theta.detach().cpu().numpy())
            ctx.problem = problem
            return torch.from_numpy(np.ascontiguousarray(x[:, :2])).to(theta.dtype)

        @staticmethod
        def backward(ctx, gx):
            g = ctx.problem.backward(gx.detach().cpu().double().numpy())
            return torch.from_numpy(g).to(gx.dtype), None

    return FinalShape


_FS = None


def __getattr__(name):
    global _FS
    if name == "FinalShape":
        if _FS is None:
            _FS = _final_shape_class()
        return _FS
    raise AttributeError(name)
