"""Batched solver: many independent reconstruction problems of uneven image,
angle and channel counts advanced together.

The reference solves one problem per call (Reconstruction_coupled.py calls
Reconstructionapproach2d once per slice, RC:188-233); here a ``gr_plan``
holds a descriptor per problem and one work-item table per kernel class, so
that every phase of an iteration is one launch for the whole batch, and
``iterate_graph`` replays a captured hipGraph of several iterations.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .solver import TgvProblem


class BatchPlan:
    """A ``gr_plan`` over a list of ``TgvProblem`` (all on one device)."""

    def __init__(self, problems):
        self.problems = list(problems)
        if not self.problems:
            raise ValueError("empty batch")
        dev = self.problems[0].device
        if any(p.device != dev for p in self.problems):
            raise ValueError("all problems of a batch must live on one device")
        self.device = dev
        n = len(self.problems)
        self._geoms = (ctypes.c_void_p * n)(*[p.geometry.handle.value for p in self.problems])
        self._prms = (_lib.TgvParams * n)(*[p.params for p in self.problems])
        self._states = (_lib.TgvState * n)(*[p.state for p in self.problems])
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(_lib.load().gr_plan_create(n, self._geoms, self._prms, self._states,
                                                  ctypes.byref(h)), "plan_create")
        self._h = h

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def iterate(self, n: int, ku_on_last: bool = False) -> None:
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().gr_plan_iterate(self._h, int(n), 1 if ku_on_last else 0,
                                                   self._stream()), "plan_iterate")
        for p in self.problems:
            p.iterations += n

    def iterate_graph(self, n: int, block: int = 10) -> None:
        """``n`` iterations through a captured graph of ``block`` iterations."""
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().gr_plan_iterate_graph(self._h, int(n), int(block),
                                                         self._stream()), "plan_iterate_graph")
        for p in self.problems:
            p.iterations += n

    def profile(self, n: int):
        """Summed ms of (dual, forward+prox, backprojection, primal) phases."""
        out = (ctypes.c_double * 4)()
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().gr_plan_profile(self._h, int(n), out, self._stream()),
                       "plan_profile")
        for p in self.problems:
            p.iterations += n
        return [float(x) for x in out]

    def retune(self) -> None:
        """Pick up the geometries' current tile shapes."""
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().gr_plan_retune(self._h, self._geoms), "plan_retune")

    def voxels(self) -> int:
        return sum(p.voxels() for p in self.problems)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().gr_plan_destroy(h)
            except Exception:
                pass
            self._h = None
