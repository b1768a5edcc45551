"""Drop-in for ``radon_tgv_primal_dual_2d`` (reference 2D:69-876).

``Reconstructionapproach2d`` keeps the reference's signature, argument
meaning, return values and progress output; the computation runs on the GPU
through libgraptor_amd.  Extra keyword arguments (all optional, absent from
the reference): ``L`` injects the operator norm, ``normest_mode`` /
``normest_seed`` control its estimate, ``u0`` warm-starts the primal
variable, ``return_problem`` also returns the device-resident TgvProblem.
"""
from __future__ import annotations

import sys

import numpy as np

from .projector import RadonGeometry, radon, radon_ad, radon_normest, radon_struct  # noqa: F401
from .projector import make_function_feasible  # noqa: F401
from .solver import METHODS, TgvProblem, resolve_device, run_with_progress


def Reconstructionapproach2d(sino, angles, Parameter, mu, maxiter, ctx, plott, discrepancy,
                             regularisationmethod, regularisation, Info, *, L=None,
                             normest_mode="reference", normest_seed=None, u0=None,
                             return_problem=False):
    """TGV/TV reconstruction of one slice from C sinograms (2D:69).

    sino      : list of C arrays [Number_of_angles, Nx] (Nx detector bins)
    angles    : radians (callers pass theta + pi/2, Reconstruction_coupled.py:233)
    Parameter : (alpha0, alpha1);  mu: C channel weights
    maxiter   : iterations;  ctx: device selector (int / torch.device / None)
    plott     : [0/1, frequency] (live plotting is not reproduced)
    discrepancy: 'KL' or 'L2';  regularisationmethod: 'Uncorrelated 2D
    Reconstruction' | 'Frobenius 2D Reconstruction' | 'Nuclear 2D
    Reconstruction';  regularisation: 'TGV' or 'TV'
    Info      : [start fraction, width fraction, slice id] for the progress line
    Returns (Solution float64 [C, Nx, Nx], list of C float32 sinograms [na, Nx]).
    """
    start, width, current = Info[0], Info[1], Info[2]
    if regularisationmethod not in METHODS or METHODS[regularisationmethod][1] != 2:
        raise ValueError(f"unknown 2D method {regularisationmethod!r}")
    coupling = METHODS[regularisationmethod][0]
    C = len(sino)
    na, nd = np.asarray(sino[0]).shape
    # f[c][0][angle][bin] = sino[c][angle, bin]  (2D:856-871 reorder, F layout)
    f = np.stack([np.asarray(s, dtype=np.float64) for s in sino]).astype(np.float32)
    f = f.reshape(C, 1, na, nd)
    if angles is None:
        angles = na
    dev = resolve_device(ctx)
    geom = RadonGeometry((nd, nd), angles, n_detectors=nd)
    prob = TgvProblem(f, geom, Parameter, mu, dim=2, coupling=coupling,
                      discrepancy=discrepancy, regularisation=regularisation, stepsize=1.0,
                      L=L, normest_mode=normest_mode, normest_seed=normest_seed, u0=u0,
                      device=dev)

    def printer(line):
        print(line + " for Slice " + str(current))
        sys.stdout.flush()

    run_with_progress(prob, int(maxiter), start, width, "", printer)
    u = prob.u.double().cpu().numpy()           # (C, 1, Ny, Nx)
    ku = prob.ku.cpu().numpy()                   # (C, 1, na, nd)
    solution = np.ascontiguousarray(u[:, 0].transpose(0, 2, 1))  # [C, x, y]
    sinograms = [np.ascontiguousarray(ku[c, 0]) for c in range(C)]
    if return_problem:
        return solution, sinograms, prob
    return solution, sinograms
