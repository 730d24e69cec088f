"""Mesh generators (thin wrappers of the C ABI generators, host only).

taylor_axi: the axisymmetric Taylor bar of the reference's sample deck
(Samples/Taylor/Taylor_Axi_Parametric.dyn:39-117).
ogrid_bar: 3D O-grid cylinder of hexahedra (build-defined; the reference has
no 3D mesher).
Each returns a dict with dim, x [nn,3], conn [ne,nen] int32, bc [nn] uint8
(restraint component mask) and v0 [nn,3].
"""
import ctypes as C

import numpy as np

from . import _lib


def taylor_axi(nr, nz, radius=3.2e-3, length=32.4e-3, v0=227.0):
    nn, ne = (nr + 1) * (nz + 1), nr * nz
    x = np.zeros((nn, 3))
    conn = np.zeros((ne, 4), dtype=np.int32)
    bc = np.zeros(nn, dtype=np.uint8)
    vel = np.zeros((nn, 3))
    _lib.check(_lib.lib().dnlg_mesh_axi_bar(nr, nz, radius, length, v0, x.ctypes.data, conn.ctypes.data,
                                             bc.ctypes.data, vel.ctypes.data))
    return dict(dim=2, x=x, conn=conn, bc=bc, v0=vel)


def ogrid_sizes(n, m, nz):
    nn, ne = C.c_longlong(), C.c_longlong()
    _lib.check(_lib.lib().dnlg_mesh_ogrid_sizes(n, m, nz, C.byref(nn), C.byref(ne)))
    return nn.value, ne.value


def ogrid_bar(n, m, nz, radius=3.2e-3, length=32.4e-3, s_frac=0.4, v0=227.0):
    nn, ne = ogrid_sizes(n, m, nz)
    x = np.zeros((nn, 3))
    conn = np.zeros((ne, 8), dtype=np.int32)
    bc = np.zeros(nn, dtype=np.uint8)
    vel = np.zeros((nn, 3))
    _lib.check(_lib.lib().dnlg_mesh_ogrid_bar(n, m, nz, radius, length, s_frac, v0, x.ctypes.data,
                                               conn.ctypes.data, bc.ctypes.data, vel.ctypes.data))
    return dict(dim=3, x=x, conn=conn, bc=bc, v0=vel)


# the Taylor-bar workloads of the tests and the benchmark
COPPER = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)


def taylor_axi_10x50():
    """Axisymmetric Taylor bar, 10 x 50 quads, 227 m/s (500 elements, 561 nodes)."""
    return taylor_axi(10, 50, 3.2e-3, 32.4e-3, 227.0)


def ogrid_taylor_bar(nz=500):
    """3D hex8 O-grid Taylor bar: 2,000 elements per section (20x20 core + 4 rings of 20x20) x nz layers."""
    return ogrid_bar(20, 20, nz, 3.2e-3, 32.4e-3, 0.4, 227.0)
