"""graptor_amd: MI355X-native multi-channel TGV primal-dual tomographic
reconstruction (the hot path of Graptor), on hand-written gfx950 HIP
kernels behind a C ABI (include/graptor_amd.h).

Public surface, mirroring the reference's module-level and closure API:
  radon_struct, radon, radon_ad, radon_normest, make_function_feasible
  Reconstructionapproach2d, Reconstructionapproach3d
  TgvProblem (device-resident solver state)
"""
from ._lib import GraptorError, GraptorMemoryError, load as load_library  # noqa: F401

__all__ = [
    "radon_struct", "radon", "radon_ad", "radon_normest", "make_function_feasible",
    "RadonGeometry", "TgvProblem", "Reconstructionapproach2d", "Reconstructionapproach3d",
    "GraptorError", "GraptorMemoryError", "load_library",
]


def __getattr__(name):
    # torch-dependent pieces are imported lazily so that the C ABI can be
    # inspected without initialising torch
    if name in ("radon_struct", "radon", "radon_ad", "radon_normest",
                "make_function_feasible", "RadonGeometry"):
        from . import projector as _r
        return getattr(_r, name)
    if name == "TgvProblem":
        from .solver import TgvProblem
        return TgvProblem
    if name == "Reconstructionapproach2d":
        from .radon_tgv_primal_dual_2d import Reconstructionapproach2d
        return Reconstructionapproach2d
    if name == "Reconstructionapproach3d":
        from .radon_tgv_primal_dual_3d import Reconstructionapproach3d
        return Reconstructionapproach3d
    raise AttributeError(name)
