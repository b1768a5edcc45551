"""Primal-dual TGV/TV reconstruction driver over libgraptor_amd.

``TgvProblem`` is the device-resident counterpart of the state that
``tgv_radon`` (radon_tgv_primal_dual_2d.py 2D:710-851,
radon_tgv_primal_dual_3d.py 3D:1135-1292) allocates: it holds the data,
the operator norm, the derived step sizes and every primal/dual array as
torch-ROCm tensors, and advances the iteration through ``gr_tgv_iterate``
(one fused dual kernel, one forward-projection+prox kernel, one
backprojection kernel and one fused primal+extrapolation kernel per
iteration).  Everything numeric runs in the HIP library.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .projector import RadonGeometry, make_function_feasible, radon_normest

LSQR_2D = 0.5 * (18.0 + math.sqrt(33.0))   # 2D:754
LSQR_3D = 17.0                              # 3D:1181

# method strings of the reference's dispatch (2D:640-690, 3D:1079-1130)
METHODS = {
    "Uncorrelated 2D Reconstruction": ("uncorrelated", 2),
    "Frobenius 2D Reconstruction": ("frobenius", 2),
    "Nuclear 2D Reconstruction": ("nuclear", 2),
    "Uncorrelated 3D Reconstruction": ("uncorrelated", 3),
    "Frobenius 3D Reconstruction": ("frobenius", 3),
    "Nuclear 3D Reconstruction": ("nuclear", 3),
}


def resolve_device(ctx=None) -> torch.device:
    """Map the reference's ``ctx`` argument to a torch device: None, an int
    ordinal, a string or a torch.device.  Anything else (e.g. a pyopencl
    context) selects the current CUDA device."""
    if not torch.cuda.is_available():
        raise RuntimeError("graptor_amd needs a ROCm GPU; torch.cuda.is_available() is False")
    if isinstance(ctx, torch.device):
        return ctx
    if isinstance(ctx, int):
        return torch.device("cuda", ctx)
    if isinstance(ctx, str):
        return torch.device(ctx)
    return torch.device("cuda", torch.cuda.current_device())


@dataclass
class StepConstants:
    """Scalars derived as 2D:752-762 / 3D:1180-1189 derive them."""
    L: float
    sigma: float
    tau: float
    sigma_p: float
    norming: float
    alpha_inv0: float
    alpha_inv1: float
    mfac: np.ndarray

    @staticmethod
    def make(dim, L, alpha, mu, discrepancy):
        lsqr = LSQR_2D if dim == 2 else LSQR_3D
        sigma = 1.0 / math.sqrt(lsqr)
        tau = sigma
        at = np.asarray(alpha, dtype=np.float64).reshape(-1) / L ** 2
        mu = np.asarray(mu, dtype=np.float64).reshape(-1)
        mfac = mu if discrepancy == "KL" else mu / (sigma + mu)
        return StepConstants(L=float(L), sigma=sigma, tau=tau, sigma_p=sigma / L,
                             norming=1.0 / L, alpha_inv0=1.0 / at[0], alpha_inv1=1.0 / at[1],
                             mfac=mfac)


class TgvProblem:
    """One reconstruction problem resident on one GPU.

    f        : sinogram stack, array-like (C, Nz, na, nd) (C-order), float32.
    geometry : RadonGeometry of an nd x nd grid.
    alpha    : (alpha0, alpha1); mu: C channel weights.
    dim      : 2 (Nz must be 1) or 3.
    L        : operator norm to use; None estimates it (``normest_mode``).
    u0       : optional warm start (C, Nz, nd, nd); the reference starts at 0.
    fused    : allocate the second state buffers with which ``iterate`` runs
               the primal step and the next dual step as one kernel (same
               values; 112 instead of 176 algorithmic bytes per voxel-channel,
               ~10 more floats of memory per voxel-channel).  Off by default:
               the fused kernel re-evaluates the primal step at the six
               neighbours it needs and measured 1.47 ms against 0.70 ms for
               the two separate kernels on 256x256x64x4.  At most 8 channels.
    """

    def __init__(self, f, geometry: RadonGeometry, alpha, mu, dim=3,
                 coupling="frobenius", discrepancy="KL", regularisation="TGV",
                 stepsize=1.0, L=None, normest_mode="reference", normest_seed=None,
                 u0=None, feasible=True, device=None, fused=False):
        self.fused = bool(fused)
        if coupling not in _lib.COUPLING:
            raise ValueError(f"unknown coupling {coupling!r}")
        if discrepancy not in _lib.DISCREPANCY:
            raise ValueError(f"unknown discrepancy {discrepancy!r}")
        reg = regularisation.upper()
        if reg not in _lib.REGULARISATION:
            raise ValueError(f"unknown regularisation {regularisation!r}")
        self.device = resolve_device(device)
        self.geometry = geometry
        self.dim = int(dim)
        self.coupling, self.discrepancy, self.regularisation = coupling, discrepancy, reg
        self.stepsize = float(stepsize)
        ft = torch.as_tensor(np.asarray(f, dtype=np.float32) if not isinstance(f, torch.Tensor)
                             else f, dtype=torch.float32, device=self.device).contiguous()
        if ft.dim() != 4:
            raise ValueError("f must be (C, Nz, na, nd)")
        C, Nz, na, nd = ft.shape
        if (na, nd) != (geometry.na, geometry.nd):
            raise ValueError(f"f has (na, nd) = {(na, nd)}, geometry {(geometry.na, geometry.nd)}")
        if self.dim == 2 and Nz != 1:
            raise ValueError("2D problems have Nz == 1")
        self.C, self.Nz, self.na, self.nd = int(C), int(Nz), int(na), int(nd)
        self.nx, self.ny = geometry.shape
        self.alpha = tuple(float(a) for a in np.asarray(alpha, dtype=np.float64).reshape(-1))
        if len(self.alpha) != 2:
            raise ValueError("alpha must have two entries (alpha0, alpha1)")
        self.mu = np.asarray(mu, dtype=np.float64).reshape(-1)
        if self.mu.size != self.C:
            raise ValueError(f"mu has {self.mu.size} entries for {self.C} channels")
        with torch.cuda.device(self.device):
            if feasible:
                make_function_feasible(ft, geometry)
            self.f = ft
            if L is None:
                L = radon_normest(None, geometry, mode=normest_mode, seed=normest_seed,
                                  device=self.device)
            self.k = StepConstants.make(self.dim, float(L), self.alpha, self.mu, discrepancy)
            self._alloc(u0)

    # ------------------------------------------------------------------
    def _alloc(self, u0):
        d = 2 if self.dim == 2 else 3
        m = 3 if self.dim == 2 else 6
        img = (self.C, self.Nz, self.ny, self.nx)
        z = lambda *s: torch.zeros(s, dtype=torch.float32, device=self.device)
        if u0 is None:
            self.u = z(*img)
        else:
            self.u = torch.as_tensor(np.asarray(u0, np.float32) if not isinstance(u0, torch.Tensor)
                                     else u0, dtype=torch.float32,
                                     device=self.device).reshape(img).contiguous().clone()
        self.ubar = self.u.clone()
        tgv = self.regularisation == "TGV"
        self.p = z(d, *img) if tgv else None
        self.pbar = z(d, *img) if tgv else None
        self.v = z(d, *img)
        self.w = z(m, *img) if tgv else None
        self.lam = torch.zeros_like(self.f)
        # K^T lam: needed unless the geometry fuses the primal step into the
        # backprojection
        self.ktl = None if getattr(self.geometry, "fuse_primal", False) else z(*img)
        # second state buffers of the fused primal + next-dual kernel
        if self.fused and self.C > 8:
            raise ValueError("fused=True: the fused kernel holds at most 8 channels per voxel")
        alt = self.fused and self.ktl is not None
        self.u2 = z(*img) if alt else None
        self.v2 = z(d, *img) if alt else None
        self.w2 = z(m, *img) if alt and tgv else None
        self.ku = torch.zeros_like(self.f)
        self.mfac = torch.as_tensor(self.k.mfac.astype(np.float32), device=self.device)
        self.params = _lib.TgvParams(
            dim=self.dim, nx=self.nx, ny=self.ny, nz=self.Nz, nch=self.C,
            coupling=_lib.COUPLING[self.coupling],
            discrepancy=_lib.DISCREPANCY[self.discrepancy],
            regularisation=_lib.REGULARISATION[self.regularisation],
            stepsize=self.stepsize, sigma=self.k.sigma, tau=self.k.tau,
            sigma_p=self.k.sigma_p, norming=self.k.norming,
            alpha_inv0=self.k.alpha_inv0, alpha_inv1=self.k.alpha_inv1)
        ptr = lambda t: t.data_ptr() if t is not None else None
        self.state = _lib.TgvState(
            u=ptr(self.u), ubar=ptr(self.ubar), p=ptr(self.p), pbar=ptr(self.pbar),
            v=ptr(self.v), w=ptr(self.w), lam=ptr(self.lam), ktl=ptr(self.ktl),
            u2=ptr(self.u2), v2=ptr(self.v2), w2=ptr(self.w2),
            ku=ptr(self.ku), f=ptr(self.f), mfac=ptr(self.mfac))
        self.iterations = 0

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def iterate(self, n: int, ku_on_last: bool = False) -> None:
        """Advance ``n`` iterations asynchronously on the current stream."""
        if n <= 0:
            return
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().gr_tgv_iterate(
                self.geometry.handle, ctypes.byref(self.params), ctypes.byref(self.state),
                int(n), 1 if ku_on_last else 0, self._stream()), "tgv_iterate")
        self.iterations += n

    def profile(self, n: int):
        """Run ``n`` iterations with HIP events around each kernel; returns the
        summed milliseconds of (standalone dual steps, forward projections +
        prox, backprojections [+ primal step when the geometry fuses it],
        standalone primal steps, fused primal + next-dual steps)."""
        out = (ctypes.c_double * 5)()
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().gr_tgv_profile(
                self.geometry.handle, ctypes.byref(self.params), ctypes.byref(self.state),
                int(n), out, self._stream()), "tgv_profile")
        self.iterations += n
        return [float(x) for x in out]

    def energy(self):
        """(data, regulariser, total) of the primal energy at (u, p), FP64.

        Overwrites ``ku`` with R u."""
        out = (ctypes.c_double * 3)()
        mu = np.ascontiguousarray(self.mu, dtype=np.float64)
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().gr_tgv_energy(
                self.geometry.handle, ctypes.byref(self.params), ctypes.byref(self.state),
                mu.ctypes.data, self.alpha[0], self.alpha[1], self.k.L, out, self._stream()),
                "tgv_energy")
        return float(out[0]), float(out[1]), float(out[2])

    def voxels(self) -> int:
        """Voxel-channels per iteration (the unit of voxel-iters/s)."""
        return self.C * self.Nz * self.nx * self.ny

    def clone(self) -> "TgvProblem":
        """An independent copy of the problem and its current state (same
        geometry and constants); used for tuning on scratch state."""
        other = object.__new__(TgvProblem)
        other.__dict__.update({k: v for k, v in self.__dict__.items()
                               if k not in ("params", "state")})
        for name in ("f", "u", "ubar", "p", "pbar", "v", "w", "lam", "ktl", "ku", "mfac"):
            t = getattr(self, name)
            setattr(other, name, None if t is None else t.clone())
        for name in ("u2", "v2", "w2"):  # scratch: contents never carried over
            t = getattr(self, name, None)
            setattr(other, name, None if t is None else torch.zeros_like(t))
        other.params = _lib.TgvParams.from_buffer_copy(self.params)
        ptr = lambda t: t.data_ptr() if t is not None else None
        other.state = _lib.TgvState(
            u=ptr(other.u), ubar=ptr(other.ubar), p=ptr(other.p), pbar=ptr(other.pbar),
            v=ptr(other.v), w=ptr(other.w), lam=ptr(other.lam), ktl=ptr(other.ktl),
            ku=ptr(other.ku), f=ptr(other.f), mfac=ptr(other.mfac),
            u2=ptr(other.u2), v2=ptr(other.v2), w2=ptr(other.w2))
        return other


def progress_line(start, width, i, maxiter, label):
    """The reference's progress line (2D:809, 3D:1237), parsed by its GUI."""
    prog = start + width * float(i) / maxiter
    return ("Overall progress at {:3.0%}".format(prog)
            + ". Progress of current computation {:3.0%}".format(float(i) / maxiter)
            + label)


def run_with_progress(prob: TgvProblem, maxiter: int, start=0.0, width=1.0, label="",
                      printer=print):
    """Run ``maxiter`` iterations in blocks of ceil(maxiter/100), printing the
    reference's progress line after each block's synchronisation (the
    reference synchronises and prints at i % ceil(maxiter/100) == 0,
    2D:803-810).  The last iteration also stores R(ubar) into ``ku``."""
    step = max(1, math.ceil(maxiter / 100.0))
    i = 0
    while i < maxiter:
        if printer is not None and i % step == 0:
            torch.cuda.synchronize(prob.device)
            printer(progress_line(start, width, i, maxiter, label))
        n = min(step - (i % step), maxiter - i)
        prob.iterate(n, ku_on_last=(i + n == maxiter))
        i += n
    torch.cuda.synchronize(prob.device)
