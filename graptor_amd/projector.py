"""Parallel-beam Radon operator pair over torch-ROCm tensors.

Module-level counterparts of the closures the reference defines inside its
solver entry points (radon_tgv_primal_dual_2d.py 2D:503-588,
radon_tgv_primal_dual_3d.py 3D:947-1030): ``radon_struct``, ``radon``,
``radon_ad``, ``radon_normest`` and ``make_function_feasible``.  Arrays keep
the reference's memory layout (pyopencl ``order='F'``, first index fastest):
an image stack is either a C-contiguous tensor shaped ``(S, Ny, Nx)`` or an
F-ordered view shaped ``(Nx, Ny, S)``; sinograms likewise ``(S, na, nd)`` /
``(nd, na, S)``.  All compute runs in libgraptor_amd on the GPU.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib


class RadonGeometry:
    """Device-side geometry: the float32 angle table plus the projector's
    per-angle constants (owns a ``gr_geometry`` handle)."""

    def __init__(self, shape, angles, n_detectors=None, detector_width=1.0,
                 detector_shift=0.0):
        lib = _lib.load()
        if np.isscalar(angles):
            angles = np.linspace(0, np.pi, int(angles) + 1)[:-1]
        self.angles = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(-1))
        self.shape = (int(shape[0]), int(shape[1]))
        nx, ny = self.shape
        if n_detectors is None:
            nd = int(math.ceil(math.hypot(nx, ny)))
        else:
            nd = int(n_detectors)
        self.nd = nd
        self.na = int(self.angles.size)
        ofs = np.zeros((self.na, 4), dtype=np.float32)
        _lib.check(lib.gr_radon_struct(self.angles.ctypes.data, self.na, nx, ny, nd,
                                       float(detector_width), float(detector_shift),
                                       ofs.ctypes.data), "radon_struct")
        self.ofs = ofs  # (na, 4): cos, sin, offset, 1/cos
        h = ctypes.c_void_p()
        _lib.check(lib.gr_geometry_create(ofs.ctypes.data, self.na, nx, ny, nd,
                                          ctypes.byref(h)), "geometry_create")
        self._h = h
        self.fuse_primal = False

    @property
    def handle(self):
        return self._h

    @property
    def sinogram_shape(self):
        return (self.nd, self.na)

    def set_tuning(self, fwd_band=0, adj_tile_y=0):
        _lib.check(_lib.load().gr_geometry_set_tuning(self._h, int(fwd_band), int(adj_tile_y)),
                   "set_tuning")

    def set_fusion(self, fuse_primal: bool):
        """Run the primal step inside the backprojection kernel (True) or as
        a separate kernel reading K^T lam from memory (False, default)."""
        _lib.check(_lib.load().gr_geometry_set_fusion(self._h, 1 if fuse_primal else 0),
                   "set_fusion")
        self.fuse_primal = bool(fuse_primal)

    def get_tuning(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().gr_geometry_get_tuning(self._h, ctypes.byref(a), ctypes.byref(b)),
                   "get_tuning")
        return a.value, b.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().gr_geometry_destroy(h)
            except Exception:
                pass
            self._h = None


def radon_struct(queue, shape, angles, n_detectors=None, detector_width=1.0,
                 detector_shift=0.0):
    """Counterpart of ``radon_struct`` (2D:503-535, 3D:947-979).

    ``queue`` is accepted for call-compatibility and ignored (the stream is
    the current torch stream at call time).  Returns the reference's triple
    ``(geometry, shape, (nd, na))`` where ``geometry`` plays the role of the
    device buffer ``ofs_buf``.
    """
    g = RadonGeometry(shape, angles, n_detectors, detector_width, detector_shift)
    return (g, tuple(shape), g.sinogram_shape)


def _stream_ptr(t: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _as_stack(t: torch.Tensor, last2, what):
    """Return (nslices, contiguous tensor view) for an image/sinogram stack in
    either C-order (S..., a, b) or the reference's F-order (b, a, S...)."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{what}: expected a CUDA tensor")
    if t.dtype != torch.float32:
        raise TypeError(f"{what}: expected float32, got {t.dtype}")
    a, b = last2
    if t.is_contiguous() and t.dim() >= 2 and tuple(t.shape[-2:]) == (a, b):
        ns = int(np.prod(t.shape[:-2])) if t.dim() > 2 else 1
        return ns, t
    rev = t.permute(*reversed(range(t.dim())))
    if rev.is_contiguous() and rev.dim() >= 2 and tuple(rev.shape[-2:]) == (a, b):
        ns = int(np.prod(rev.shape[:-2])) if rev.dim() > 2 else 1
        return ns, rev
    raise ValueError(f"{what}: shape {tuple(t.shape)} does not match ({a}, {b}) in "
                     "C-order or the reference's F-order, or is not contiguous")


def _geom(r_struct) -> RadonGeometry:
    g = r_struct[0] if isinstance(r_struct, tuple) else r_struct
    if not isinstance(g, RadonGeometry):
        raise TypeError("r_struct must come from radon_struct()")
    return g


def radon(sino, img, r_struct, wait_for=None):
    """Forward projection into ``sino`` (2D:544-550; kernel 2D:391-434).

    ``wait_for`` is accepted for call-compatibility; ordering follows the
    current torch stream.  Returns None (the reference returns an event).
    """
    g = _geom(r_struct)
    nx, ny = g.shape
    ns_i, imv = _as_stack(img, (ny, nx), "radon img")
    ns_s, sv = _as_stack(sino, (g.na, g.nd), "radon sino")
    if ns_i != ns_s:
        raise ValueError(f"radon: {ns_i} image slices vs {ns_s} sinogram slices")
    _lib.check(_lib.load().gr_radon(g.handle, sv.data_ptr(), imv.data_ptr(), ns_i,
                                    _stream_ptr(img)), "radon")


def radon_ad(img, sino, r_struct, wait_for=None):
    """Backprojection into ``img`` (2D:559-565; kernel 2D:436-462)."""
    g = _geom(r_struct)
    nx, ny = g.shape
    ns_i, imv = _as_stack(img, (ny, nx), "radon_ad img")
    ns_s, sv = _as_stack(sino, (g.na, g.nd), "radon_ad sino")
    if ns_i != ns_s:
        raise ValueError(f"radon_ad: {ns_i} image slices vs {ns_s} sinogram slices")
    _lib.check(_lib.load().gr_radon_ad(g.handle, imv.data_ptr(), sv.data_ptr(), ns_i,
                                       _stream_ptr(img)), "radon_ad")


def radon_normest(queue, r_struct, mode="reference", iters=None, x0=None, seed=None,
                  tol=1e-10, device=None):
    """Operator-norm estimate ||R|| (2D:576-588), computed in FP64 on the GPU.

    ``mode='reference'`` repeats the reference's 10 sum-normalised rounds
    from a standard-normal start image (seeded here with ``seed``; the
    reference's start is unseeded).  ``mode='power'`` runs an L2-normalised
    power iteration to relative tolerance ``tol`` (default 200 rounds).
    """
    g = _geom(r_struct)
    nx, ny = g.shape
    if iters is None:
        iters = 10 if mode == "reference" else 200
    if x0 is None and mode == "reference":
        rng = np.random.default_rng(seed)
        # the reference draws randn(Nx, Ny) and stores it as float32 (2D:577)
        x0 = rng.standard_normal((nx, ny)).astype(np.float32)
    x0p = None
    if x0 is not None:
        # the start image is given in the reference's (Nx, Ny) indexing
        x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64).reshape(nx, ny).T)
        x0p = x0.ctypes.data
    dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    out = ctypes.c_double()
    _lib.check(_lib.load().gr_normest(g.handle, x0p, int(iters), _lib.NORMEST_MODE[mode],
                                      float(tol), ctypes.byref(out), stream), "normest")
    return out.value


def make_function_feasible(f, r_struct):
    """Zero the bins of the device sinogram stack ``f`` that no pixel reaches
    (R 1 <= 0), in place (2D:476-486, 3D:860-870)."""
    g = _geom(r_struct)
    ns, fv = _as_stack(f, (g.na, g.nd), "make_function_feasible")
    _lib.check(_lib.load().gr_make_feasible(g.handle, fv.data_ptr(), ns, _stream_ptr(f)),
               "make_feasible")
    return f
