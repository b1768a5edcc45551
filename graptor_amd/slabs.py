"""A 3D volume as z-slabs with halo planes: the part of a slab-sharded solve
that each rank runs, exercised on one GPU.

Parallel-beam projections are slice-independent, and the TGV stencils
couple only adjacent z planes, so a volume splits into slabs along z (the
channels stay together: the coupled projections are pointwise across
channels).  Per iteration a slab's dual step needs ubar of the plane above
and pbar of the plane below it, its primal step v_z of the plane below and
(w_xz, w_yz, w_zz) of the plane above: the halo planes of gr_tgv_state.
Ranks would exchange these planes with their neighbours twice per iteration
(send / receive pairs over xGMI) as buffers packed by gr_tgv_pack_halos;
here all slabs share one GPU and each slab packs its planes straight into
its neighbours' halo fields, so the result equals the unsplit volume's solve
bit for bit.

The iteration is composed from the single-step entry points (dual step,
forward projection, prox, backprojection, primal step) so that the halo
exchanges can sit between the dual and the primal steps; the reference solves
slabs as independent problems instead (RC:408-458, with overlapping slabs).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .projector import RadonGeometry, radon, radon_ad
from .solver import TgvProblem

class SlabSolve:
    """``nslabs`` z-slabs of a (C, Nz, na, nd) data stack, each a
    TgvProblem with halo planes from its neighbours; ``L`` (the operator
    norm of the whole volume's projector, i.e. of one slice) is injected."""

    def __init__(self, f, geometry: RadonGeometry, alpha, mu, nslabs, L, coupling="frobenius",
                 discrepancy="KL", regularisation="TGV", stepsize=1.0, device=None):
        ft = f if isinstance(f, torch.Tensor) else torch.as_tensor(np.asarray(f, np.float32))
        C, Nz, na, nd = ft.shape
        if not 1 <= nslabs <= Nz:
            raise ValueError(f"nslabs must be in [1, {Nz}]")
        bounds = np.linspace(0, Nz, nslabs + 1).round().astype(int)
        self.slabs = [TgvProblem(ft[:, z0:z1].contiguous(), geometry, alpha, mu, dim=3,
                                 coupling=coupling, discrepancy=discrepancy,
                                 regularisation=regularisation, stepsize=stepsize, L=L,
                                 device=device)
                      for z0, z1 in zip(bounds[:-1], bounds[1:])]
        self.bounds = [(int(a), int(b)) for a, b in zip(bounds[:-1], bounds[1:])]
        self.geometry = geometry
        self.tgv = regularisation.upper() == "TGV"
        self.halos = []
        for s, p in enumerate(self.slabs):
            plane = (p.C, p.ny, p.nx)
            z = lambda *shape: torch.zeros(shape, dtype=torch.float32, device=p.device)
            h = {}
            if s > 0:  # a slab below
                h["v_lo"] = z(*plane)
                if self.tgv:
                    h["pbar_lo"] = z(3, *plane)
            if s < len(self.slabs) - 1:  # a slab above
                h["ubar_hi"] = z(*plane)
                if self.tgv:
                    h["w_hi"] = z(3, *plane)
            ptr = lambda k: h[k].data_ptr() if k in h else None
            p.state.halo_v_lo = ptr("v_lo")
            p.state.halo_pbar_lo = ptr("pbar_lo")
            p.state.halo_ubar_hi = ptr("ubar_hi")
            p.state.halo_w_hi = ptr("w_hi")
            self.halos.append(h)

    # halo exchanges: each slab packs its outgoing planes with one launch
    # (gr_tgv_pack_halos); a rank would send those buffers to its neighbours and
    # receive into its halo fields.  Slabs sharing one GPU pack straight into
    # the neighbours' halo fields.
    def _pack(self, phase: str, below: str, above: str) -> None:
        lib = _lib.load()
        for s, p in enumerate(self.slabs):
            to_below = self.halos[s - 1].get(below) if s > 0 else None
            to_above = self.halos[s + 1].get(above) if s + 1 < len(self.slabs) else None
            if to_below is None and to_above is None:
                continue
            _lib.check(lib.gr_tgv_pack_halos(
                ctypes.byref(p.params), ctypes.byref(p.state), _lib.HALO_PHASE[phase],
                to_below.data_ptr() if to_below is not None else None,
                to_above.data_ptr() if to_above is not None else None, p._stream()),
                "pack_halos")

    def exchange_dual_halos(self):
        """ubar of the plane above, pbar of the plane below (read by the dual step)."""
        self._pack("dual", "ubar_hi", "pbar_lo")

    def exchange_primal_halos(self):
        """v_z of the plane below, (w_xz, w_yz, w_zz) of the plane above (read
        by the primal step)."""
        self._pack("primal", "w_hi", "v_lo")

    def iterate(self, n: int) -> None:
        lib = _lib.load()
        for _ in range(int(n)):
            self.exchange_dual_halos()
            for p in self.slabs:
                st = p._stream()
                _lib.check(lib.gr_tgv_dual_step(self.geometry.handle, ctypes.byref(p.params),
                                                ctypes.byref(p.state), st), "dual_step")
            for p in self.slabs:
                S = p.C * p.Nz
                radon(p.ku.view(S, p.na, p.nd), p.ubar.view(S, p.ny, p.nx), self.geometry)
                _lib.check(lib.gr_prox(_lib.DISCREPANCY[p.discrepancy], float(p.params.sigma_p),
                                       p.mfac.data_ptr(), p.C, p.Nz * p.na * p.nd,
                                       p.ku.data_ptr(), p.f.data_ptr(), p.lam.data_ptr(),
                                       p._stream()), "prox")
                radon_ad(p.ktl.view(S, p.ny, p.nx), p.lam.view(S, p.na, p.nd), self.geometry)
            self.exchange_primal_halos()
            for p in self.slabs:
                _lib.check(lib.gr_tgv_primal_step(self.geometry.handle, ctypes.byref(p.params),
                                                  ctypes.byref(p.state), p._stream()),
                           "primal_step")
        for p in self.slabs:
            p.iterations += int(n)

    def gather(self, name: str) -> torch.Tensor:
        """A state array of the whole volume (z concatenated)."""
        parts = [getattr(p, name) for p in self.slabs]
        axis = 1 if parts[0].dim() == 4 else 2
        return torch.cat(parts, dim=axis)
