"""redynela_amd: MI355X (gfx950) explicit FEM core with the Domain / Material /
ExplicitSolverCH interface of pantale/ReDynELA.

The compute path is libdnlgpu.so (hand-written HIP kernels behind the C ABI of
include/dnlg.h); this package is its host-side mirror.
"""
from ._lib import DnlgError, lib  # noqa: F401
from .domain import (Domain, ExplicitSolverCH, HistoryFile, IsoHardElastoplastic,  # noqa: F401
                     IsoHardJohnsonCook, Material)
from . import mesh  # noqa: F401
from .batch import Batch  # noqa: F401

__all__ = ["Domain", "ExplicitSolverCH", "HistoryFile", "IsoHardElastoplastic", "IsoHardJohnsonCook", "Material",
           "mesh", "Batch", "DnlgError", "lib"]
