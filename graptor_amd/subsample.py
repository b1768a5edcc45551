"""Two-resolution solver: channels recorded on a coarser detector / slice
grid reconstructed jointly with full-resolution channels
(radon_tgv_primal_dual_3dsubsample.py = 3DS:).

Channels 0..C0-1 are measured on the fine grid (nd bins, Nz slices, na0
angles); channels C0.. on a coarse grid (nd1 bins, Nz1 slices, na1 angles).
Their forward model is S R1 u, where S averages the fine projection over the
footprint of each coarse bin (a box of p x p/Stepsize fine pixels placed by
the reference points Refold -> Refnew), applied per angle over the
(bin, slice) plane.  The iteration is the coupled TGV primal-dual one with
Lsqr = 36 (3DS:1003) and the operator norm of the joint operator
(3DS:532-575).  Everything runs on the GPU through libgraptor_amd: the dual
and primal steps, the projector pair per group, S and S^T as CSR products
(``gr_csr_apply``), the prox per group (``gr_prox``) and the FP64 norms of
the power iteration (``gr_sumsq_f64``).
"""
from __future__ import annotations

import ctypes
import math
import sys

import numpy as np
import torch

from . import _lib
from .projector import RadonGeometry, make_function_feasible, radon, radon_ad
from .solver import progress_line, resolve_device

LSQR = 36.0
METHODS = {"Frobenius 3D Subsampling Reconstruction": "frobenius",
           "Uncorrelated 3D Subsampling Reconstruction": "uncorrelated"}


def interpolation_points(refold, refnew, nx, nz, ratio):
    """Fine-grid (bin, slice) coordinates of every coarse point, slice-major:
    ratio * (coarse - refold) + refnew (3DS:578-585).  Returns (P, 2)."""
    j, i = np.mgrid[0:nz, 0:nx]
    x = ratio[0] * (i.reshape(-1) - refold[0]) + refnew[0]
    y = ratio[1] * (j.reshape(-1) - refold[1]) + refnew[1]
    return np.stack([x, y], axis=1).astype(np.float64)


def box_average_csr(points, size, nx, ny):
    """CSR of the box-average interpolation: row r averages the fine pixels
    under a box of ``size`` centred at points[r], weight = overlap area /
    box area, over the candidate pixels ceil(c - r - 0.499999) ..
    floor(c + r + 0.499999) per axis, indices clipped into the grid (pixels
    beyond the border fold onto it) (3DS:587-651).  Returns (rowstart int32,
    index int32, value float32); entries of a row in y-major, x-minor order."""
    r0, r1 = size[0] / 2.0, size[1] / 2.0
    rs, ix, vs = [0], [], []
    for x, y in points:
        i0 = int(math.ceil(x - r0 - 0.499999))
        i1 = int(math.floor(x + r0 + 0.499999))
        j0 = int(math.ceil(y - r1 - 0.499999))
        j1 = int(math.floor(y + r1 + 0.499999))
        ii = np.arange(i0, i1 + 1, dtype=np.float64)
        jj = np.arange(j0, j1 + 1, dtype=np.float64)
        ox = np.clip(x + r0 - ii, -0.5, 0.5) + np.clip(ii - (x - r0), -0.5, 0.5)
        oy = np.clip(y + r1 - jj, -0.5, 0.5) + np.clip(jj - (y - r1), -0.5, 0.5)
        w = (oy[:, None] * ox[None, :]) / (4.0 * r0 * r1)
        col = (np.clip(ii, 0, nx - 1)[None, :] + nx * np.clip(jj, 0, ny - 1)[:, None])
        vs.append(w.reshape(-1))
        ix.append(col.reshape(-1).astype(np.int64))
        rs.append(rs[-1] + w.size)
    index = np.concatenate(ix) if ix else np.zeros(0, np.int64)
    value = np.concatenate(vs) if vs else np.zeros(0)
    return (np.asarray(rs, np.int32), index.astype(np.int32), value.astype(np.float32))


def transpose_csr(rowstart, index, value, ncols):
    """CSR of the transpose (rows ordered by column, entries of a row in
    ascending original-row order, duplicates kept)."""
    nrows = len(rowstart) - 1
    rows = np.repeat(np.arange(nrows, dtype=np.int64), np.diff(rowstart))
    order = np.lexsort((rows, index))  # by column, then original row
    counts = np.bincount(index, minlength=ncols)
    rs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return rs, rows[order].astype(np.int32), value[order].astype(np.float32)


class _Csr:
    """A CSR matrix on the device (plane operator nx_in*nz_in -> nx_out*nz_out)."""

    def __init__(self, rs, ix, vs, nx_out, nz_out, nx_in, nz_in, device):
        if len(rs) != nx_out * nz_out + 1:
            raise ValueError("row count does not match the output plane")
        if len(ix) and (ix.min() < 0 or ix.max() >= nx_in * nz_in):
            raise ValueError("column index outside the input plane")
        self.rs = torch.as_tensor(rs, device=device)
        self.ix = torch.as_tensor(ix, device=device)
        self.vs = torch.as_tensor(vs, device=device)
        self.dims = (nx_out, nz_out, nx_in, nz_in)

    def apply(self, src, dst, na, nch, stream):
        nx_out, nz_out, nx_in, nz_in = self.dims
        _lib.check(_lib.load().gr_csr_apply(self.rs.data_ptr(), self.ix.data_ptr(),
                                            self.vs.data_ptr(), nx_out, nz_out, nx_in, nz_in,
                                            na, nch, src.data_ptr(), dst.data_ptr(), stream),
                   "csr_apply")


class SubsampledProblem:
    """Joint reconstruction of fine-grid channels f0 (C0, Nz, na0, nd) and
    coarse-grid channels f1 (C1, Nz1, na1, nd1) on an nd x nd x Nz grid.

    topology = (p, refold, refnew): coarse pixel size in fine bins, and a
    coarse reference point (bin, slice) with its fine-grid position
    (3DS:1211, 940-944; refnew's slice coordinate is divided by stepsize
    here as there)."""

    def __init__(self, f0, f1, angles0, angles1, topology, alpha, mu, coupling="frobenius",
                 discrepancy="KL", regularisation="TGV", stepsize=1.0, L=None, normest_seed=0,
                 device=None):
        self.device = resolve_device(device)
        dev = self.device
        t = lambda a: torch.as_tensor(np.asarray(a, np.float32) if not isinstance(a, torch.Tensor)
                                      else a, dtype=torch.float32, device=dev).contiguous()
        f0, f1 = t(f0), t(f1)
        C0, nz, na0, nd = f0.shape
        C1, nz1, na1, nd1 = f1.shape
        self.C0, self.C1, self.C = C0, C1, C0 + C1
        self.nz, self.nz1, self.nd, self.nd1, self.na0, self.na1 = nz, nz1, nd, nd1, na0, na1
        self.stepsize = float(stepsize)
        self.g0 = RadonGeometry((nd, nd), angles0, n_detectors=nd)
        self.g1 = RadonGeometry((nd, nd), angles1, n_detectors=nd)
        if (self.g0.na, self.g1.na) != (na0, na1):
            raise ValueError("angle counts do not match the sinograms")
        p, refold, refnew = topology
        refnew = (float(refnew[0]), float(refnew[1]) / self.stepsize)
        ratio = (float(p), float(p) / self.stepsize)
        pts = interpolation_points(refold, refnew, nd1, nz1, ratio)
        rs, ix, vs = box_average_csr(pts, ratio, nd, nz)
        self.S = _Csr(rs, ix, vs, nd1, nz1, nd, nz, dev)
        rsT, ixT, vsT = transpose_csr(rs, ix, vs, nd * nz)
        self.ST = _Csr(rsT, ixT, vsT, nd, nz, nd1, nz1, dev)
        with torch.cuda.device(dev):
            make_function_feasible(f0.view(C0 * nz, na0, nd), self.g0)
            # the coarse channels are masked with the fine geometry's R 1 of
            # their angles (3DS:998, 466-476): bins (i < nd1, angle) where no
            # fine pixel projects are zeroed in every slice
            ones = torch.ones((1, nd, nd), device=dev)
            r1 = torch.zeros((1, na1, nd), device=dev)
            radon(r1, ones, self.g1)
            dead = (r1[0, :, :min(nd, nd1)] <= 0)
            if bool(dead.any()):
                f1[:, :, :, :min(nd, nd1)].masked_fill_(dead.view(1, 1, na1, -1), 0.0)
            self.f0, self.f1 = f0, f1
            self._alloc()
            self.L = float(L) if L is not None else self.normest(seed=normest_seed)
        self.alpha = tuple(float(a) for a in alpha)
        self.mu = np.asarray(mu, np.float64).reshape(-1)
        if self.mu.size != self.C:
            raise ValueError(f"mu has {self.mu.size} entries for {self.C} channels")
        self.discrepancy, self.coupling = discrepancy, coupling
        self.regularisation = regularisation.upper()
        sigma = 1.0 / math.sqrt(LSQR)
        self.sigma = self.tau = sigma
        mf = self.mu if discrepancy == "KL" else self.mu / (sigma + self.mu)
        self.mfac = torch.as_tensor(mf.astype(np.float32), device=dev)
        at = np.asarray(self.alpha) / self.L ** 2
        self.params = _lib.TgvParams(
            dim=3, nx=nd, ny=nd, nz=nz, nch=self.C, coupling=_lib.COUPLING[coupling],
            discrepancy=_lib.DISCREPANCY[discrepancy],
            regularisation=_lib.REGULARISATION[self.regularisation], stepsize=self.stepsize,
            sigma=sigma, tau=sigma, sigma_p=sigma / self.L, norming=1.0 / self.L,
            alpha_inv0=1.0 / at[0], alpha_inv1=1.0 / at[1])
        ptr = lambda x: x.data_ptr() if x is not None else None
        self.state = _lib.TgvState(u=ptr(self.u), ubar=ptr(self.ubar), p=ptr(self.p),
                                   pbar=ptr(self.pbar), v=ptr(self.v), w=ptr(self.w),
                                   lam=None, ktl=ptr(self.ktl), ku=None, f=ptr(self.f0),
                                   mfac=ptr(self.mfac))
        self.iterations = 0

    def _alloc(self):
        z = lambda *s: torch.zeros(s, dtype=torch.float32, device=self.device)
        C, nz, nd = self.C, self.nz, self.nd
        img = (C, nz, nd, nd)
        self.u, self.ubar, self.ktl = z(*img), z(*img), z(*img)
        self.p, self.pbar, self.v = z(3, *img), z(3, *img), z(3, *img)
        self.w = z(6, *img)
        self.lam0 = torch.zeros_like(self.f0)
        self.lam1 = torch.zeros_like(self.f1)
        self.ku0 = torch.zeros_like(self.f0)
        self.ku1_fine = z(self.C1, nz, self.na1, nd)
        self.ku1 = torch.zeros_like(self.f1)
        self.lam1_fine = z(self.C1, nz, self.na1, nd)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    # K = (R0 u0, S R1 u1) and its adjoint, on device tensors
    def _forward(self, img, k0, k1_fine, k1):
        C0, C1, nz, nd = self.C0, self.C1, self.nz, self.nd
        radon(k0.view(C0 * nz, self.na0, nd), img[:C0].reshape(C0 * nz, nd, nd), self.g0)
        radon(k1_fine.view(C1 * nz, self.na1, nd), img[C0:].reshape(C1 * nz, nd, nd), self.g1)
        self.S.apply(k1_fine, k1, self.na1, C1, self._stream())

    def _adjoint(self, l0, l1, l1_fine, out):
        C0, C1, nz, nd = self.C0, self.C1, self.nz, self.nd
        self.ST.apply(l1, l1_fine, self.na1, C1, self._stream())
        radon_ad(out[:C0].view(C0 * nz, nd, nd), l0.view(C0 * nz, self.na0, nd), self.g0)
        radon_ad(out[C0:].view(C1 * nz, nd, nd), l1_fine.view(C1 * nz, self.na1, nd), self.g1)

    def _norm(self, x):
        out = ctypes.c_double()
        _lib.check(_lib.load().gr_sumsq_f64(x.data_ptr(), x.numel(), ctypes.byref(out),
                                            self._stream()), "sumsq_f64")
        return math.sqrt(out.value)

    def normest(self, rounds=25, seed=0):
        """Operator norm of K by 25 L2-normalised rounds of K^T K from a
        seeded standard-normal start (3DS:532-575; the reference's start is
        unseeded); the estimate is the square root of the last norm taken."""
        g = torch.Generator(device=self.device)
        g.manual_seed(int(seed))
        x = torch.randn((self.C, self.nz, self.nd, self.nd), device=self.device, generator=g)
        k0, k1f, k1 = torch.zeros_like(self.f0), torch.zeros_like(self.ku1_fine), \
            torch.zeros_like(self.f1)
        l1f = torch.zeros_like(self.ku1_fine)
        y = torch.zeros_like(x)
        nrm = 1.0
        for _ in range(rounds):
            nrm = self._norm(x)
            x.mul_(1.0 / nrm)
            self._forward(x, k0, k1f, k1)
            self._adjoint(k0, k1, l1f, y)
            x, y = y, x
        return math.sqrt(nrm)

    def iterate(self, n):
        """n iterations of 3DS:1061-1106 on the current stream."""
        lib, st = _lib.load(), self._stream()
        prm, state = ctypes.byref(self.params), ctypes.byref(self.state)
        disc = _lib.DISCREPANCY[self.discrepancy]
        sp = float(self.params.sigma_p)
        per0 = self.nz * self.na0 * self.nd
        per1 = self.nz1 * self.na1 * self.nd1
        mf0 = self.mfac[:self.C0]
        mf1 = self.mfac[self.C0:]
        with torch.cuda.device(self.device):
            for _ in range(int(n)):
                _lib.check(lib.gr_tgv_dual_step(self.g0.handle, prm, state, st), "dual_step")
                self._forward(self.ubar, self.ku0, self.ku1_fine, self.ku1)
                _lib.check(lib.gr_prox(disc, sp, mf0.data_ptr(), self.C0, per0,
                                       self.ku0.data_ptr(), self.f0.data_ptr(),
                                       self.lam0.data_ptr(), st), "prox")
                _lib.check(lib.gr_prox(disc, sp, mf1.data_ptr(), self.C1, per1,
                                       self.ku1.data_ptr(), self.f1.data_ptr(),
                                       self.lam1.data_ptr(), st), "prox")
                self._adjoint(self.lam0, self.lam1, self.lam1_fine, self.ktl)
                _lib.check(lib.gr_tgv_primal_step(self.g0.handle, prm, state, st),
                           "primal_step")
        self.iterations += int(n)


def Reconstructionapproach3dsubsmapling(sino, angles, Topology, Parameter, mu, maxiter, ctx,
                                        plott, Stepsize, discrepancy, regularisationmethod,
                                        regularisation, Info, *, L=None, normest_seed=0,
                                        return_problem=False):
    """Drop-in for the reference's entry (3DS:51; its name is kept).

    sino: list of C arrays [angles, slices, bins]; channels
    0..Topology[0]-1 on the fine grid, the rest on the coarse grid.
    angles: list of C angle arrays (radians, already +pi/2); the first of
    each group is used (3DS:1213).  Topology = [first coarse channel, p,
    Refold, Refnew].  Returns (Solution float64 [C, nd, nd, Nz],
    [sinograms fine (C0, na0, Nz, nd), coarse channels on the fine grid
    (C1, na1, Nz, nd), coarse channels (C1, na1, Nz1, nd1)])."""
    if regularisationmethod not in METHODS:
        raise ValueError(f"unknown method {regularisationmethod!r}")
    start, width, current = Info[0], Info[1], Info[2]
    c0, p, refold, refnew = Topology
    C = len(sino)
    f0 = np.stack([np.asarray(s, np.float32) for s in sino[:c0]]).transpose(0, 2, 1, 3)
    f1 = np.stack([np.asarray(s, np.float32) for s in sino[c0:]]).transpose(0, 2, 1, 3)
    print("########## Starting\t" + str(discrepancy) + " " + str(regularisation) + " "
          + str(regularisationmethod) + " ############")
    prob = SubsampledProblem(np.ascontiguousarray(f0), np.ascontiguousarray(f1), angles[0],
                             angles[c0], (p, refold, refnew), Parameter, mu,
                             coupling=METHODS[regularisationmethod], discrepancy=discrepancy,
                             regularisation=regularisation, stepsize=float(Stepsize), L=L,
                             normest_seed=normest_seed, device=resolve_device(ctx))
    step = max(1, int(math.ceil(maxiter / 100.0)))
    label = (f"{current[0]}, {current[1]}...{current[-1]}"
             if isinstance(current, (list, tuple)) and len(current) > 3 else str(current))
    done = 0
    while done < maxiter:
        n = min(step, maxiter - done)
        prob.iterate(n)
        torch.cuda.synchronize(prob.device)
        print(progress_line(start, width, done, maxiter, "") + " for Section " + label)
        sys.stdout.flush()
        done += n
    print("iterations done")
    u = prob.u.double().cpu().numpy()
    solution = np.ascontiguousarray(u.transpose(0, 3, 2, 1))
    sinos = [np.ascontiguousarray(prob.ku0.cpu().numpy().transpose(0, 2, 1, 3)),
             np.ascontiguousarray(prob.ku1_fine.cpu().numpy().transpose(0, 2, 1, 3)),
             np.ascontiguousarray(prob.ku1.cpu().numpy().transpose(0, 2, 1, 3))]
    if return_problem:
        return solution, sinos, prob
    return solution, sinos
