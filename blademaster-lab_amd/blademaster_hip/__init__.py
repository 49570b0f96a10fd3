"""BladeMaster TLMPM cutting substep on MI355X: HIP kernels behind a ctypes C-ABI.

Public surface mirrors the reference package blademaster_warp: Config, Solver,
ExternalToolPort, ToolState, Reaction, components, Metrics, points, stencil.
"""

from .config import Config
from .geometry import points, stencil, levelset
from .topology import components, compatible

__all__ = [
    "Config",
    "points",
    "stencil",
    "levelset",
    "components",
    "compatible",
    "Solver",
    "ExternalToolPort",
    "ToolState",
    "Reaction",
    "Metrics",
]


def __getattr__(name):
    # Device-side classes import torch and the HIP library lazily.
    if name == "Solver":
        from .solver import Solver

        return Solver
    if name in ("ExternalToolPort", "ToolState", "Reaction"):
        from . import coupling

        return getattr(coupling, name)
    if name == "Metrics":
        from .diagnostics import Metrics

        return Metrics
    raise AttributeError(name)
