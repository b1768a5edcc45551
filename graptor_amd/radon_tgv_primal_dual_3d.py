"""Drop-in for ``radon_tgv_primal_dual_3d`` (reference 3D:69-1318).

``Reconstructionapproach3d`` keeps the reference's signature, argument
meaning, return values and progress output; the computation runs on the GPU
through libgraptor_amd.  The optional keyword arguments are those of the 2D
module (``L``, ``normest_mode``, ``normest_seed``, ``u0``, ``return_problem``).
"""
from __future__ import annotations

import sys

import numpy as np

from .projector import RadonGeometry, radon, radon_ad, radon_normest, radon_struct  # noqa: F401
from .projector import make_function_feasible  # noqa: F401
from .solver import METHODS, TgvProblem, resolve_device, run_with_progress


def _section_label(current):
    if isinstance(current, (list, tuple)) and len(current) > 3:
        return f"{current[0]}, {current[1]}...{current[len(current) - 1]}"
    return str(current)


def Reconstructionapproach3d(sino, angles, Parameter, mu, maxiter, ctx, plott, Stepsize,
                             discrepancy, regularisationmethod, regularisation, Info, *,
                             L=None, normest_mode="reference", normest_seed=None, u0=None,
                             return_problem=False):
    """TGV/TV reconstruction of a slab of slices from C sinogram stacks (3D:69).

    sino      : list of C arrays [Number_of_angles, Nz slices, Nx bins]
    Stepsize  : slice spacing h dividing every z-difference (3D:631)
    regularisationmethod: 'Uncorrelated 3D Reconstruction' | 'Frobenius 3D
    Reconstruction' | 'Nuclear 3D Reconstruction'
    Info      : [start fraction, width fraction, list of slice ids]
    Other arguments as Reconstructionapproach2d.
    Returns (Solution float64 [C, Nx, Nx, Nz], Sinograms float64 [C, na, Nz, Nx]).
    """
    start, width, current = Info[0], Info[1], Info[2]
    if regularisationmethod not in METHODS or METHODS[regularisationmethod][1] != 3:
        raise ValueError(f"unknown 3D method {regularisationmethod!r}")
    coupling = METHODS[regularisationmethod][0]
    print("########## Starting\t" + str(discrepancy) + " " + str(regularisation) + " "
          + str(regularisationmethod) + " ############")
    C = len(sino)
    na, nz, nd = np.asarray(sino[0]).shape
    # f[c][z][angle][bin] = sino[c][angle, z, bin]  (3D:1299-1311 reorder)
    f = np.stack([np.asarray(s, dtype=np.float32) for s in sino]).transpose(0, 2, 1, 3)
    f = np.ascontiguousarray(f)
    if angles is None:
        angles = na
    dev = resolve_device(ctx)
    geom = RadonGeometry((nd, nd), angles, n_detectors=nd)
    prob = TgvProblem(f, geom, Parameter, mu, dim=3, coupling=coupling,
                      discrepancy=discrepancy, regularisation=regularisation,
                      stepsize=float(Stepsize), L=L, normest_mode=normest_mode,
                      normest_seed=normest_seed, u0=u0, device=dev)
    label = _section_label(current)

    def printer(line):
        print(line + " for Section " + label)
        sys.stdout.flush()

    run_with_progress(prob, int(maxiter), start, width, "", printer)
    print("Progress at {:3.0%}".format(1.0))
    sys.stdout.flush()
    print("Iterations done for section " + label)
    u = prob.u.double().cpu().numpy()            # (C, Nz, Ny, Nx)
    ku = prob.ku.double().cpu().numpy()          # (C, Nz, na, nd)
    solution = np.ascontiguousarray(u.transpose(0, 3, 2, 1))     # [C, x, y, z]
    sinograms = np.ascontiguousarray(ku.transpose(0, 2, 1, 3))   # [C, na, z, nd]
    if return_problem:
        return solution, sinograms, prob
    return solution, sinograms
