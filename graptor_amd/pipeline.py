"""Caller counterpart: data preparation, slab scheduling and output of a
coupled multi-channel reconstruction (Reconstruction_coupled.py = RC:,
h5utilities.py = H5:).

The reference's caller reads each channel's tilt series, removes bad
projections, cuts detector edges, corrects brightness, scales the data by
percentiles, applies a baseline rule, solves (2D: one slice per call; 3D: in
slabs sized to the device's memory, with overlap, shrinking on
out-of-memory), rescales the result by the data's mass ("Faktor") and writes
MRC files.  The same steps are kept here with the same argument meaning; the
solver is this package's (2D slices are solved together in one batch plan
instead of one call per slice).  HDF5 input is not supported (h5py is not
available); MRC + .rawtlt/.csv is.
"""
from __future__ import annotations

import math
import sys
import time

import numpy as np

from . import io as gio
from ._lib import GraptorError, GraptorMemoryError

# --------------------------------------------------------------------------
# pre-processing (H5:56-313)
# --------------------------------------------------------------------------


def read_bad_file(path: str | None) -> list[int]:
    """1-based indices of bad projections listed in a text file (H5:148-153)."""
    if not path:
        return []
    with open(path) as fh:
        return [int(t) for t in fh.read().split()]


def remove_projections(data, angles, bad):
    """Drop the projections whose 1-based index is in ``bad`` (H5:67-76)."""
    bad = set(int(b) for b in bad)
    keep = [i for i in range(data.shape[0]) if i + 1 not in bad]
    return data[keep], np.asarray(angles)[keep]


def cut_edges(data, cutedges):
    """Remove ceil(Nx*cutedges) detector bins on both sides; the background
    is the mean of the outer half of each removed band (average of the left
    and right means) when the cut is at least 2 bins, else 0
    (H5:78-92).  Returns (data, background)."""
    Nx = data.shape[2]
    cut = int(math.ceil(Nx * float(cutedges)))
    # the reference's test `cut/2 != cut and cut/2 > 0` (Python 2 integer
    # division) holds exactly for cut >= 2
    if cut >= 2:
        bg = 0.5 * (np.mean(data[:, :, cut // 2:cut]) + np.mean(data[:, :, Nx - cut // 2:Nx]))
    else:
        bg = 0.0
    return data[:, :, cut:Nx - cut], bg


def brightness_rescale(data):
    """Scale each projection so its mean equals the mean of all projection
    means (H5:94-103, Brightness_correction2 H5:256-268)."""
    data = np.array(data, dtype=np.float64, copy=True)
    C = data.reshape(data.shape[0], -1).mean(axis=1)
    W = C.mean()
    data *= (W / C)[:, None, None] if data.ndim == 3 else (W / C)[:, None]
    return data


def load_preprocessed(name, brightness="", badfile="", cutedges=0.0, bad_projections=()):
    """Tilt series (tilts, slices, bins) and angles (radians) after the
    caller's preparation (RC:144-152 / 334-339):

    * 'bright': remove bad projections, cut edges minus background,
      brightness rescale (H5:144-198); noise estimate 0;
    * 'basic' : shift to min 0, remove bad projections, shift again, cut
      edges; returns the edge background as the noise estimate (H5:273-313);
    * otherwise the raw series (H5:358).
    Returns (data, angles, averagenoise)."""
    data, angles = gio.read_tilt_series(name)
    bad = read_bad_file(badfile) + list(bad_projections)
    mode = (brightness or "").upper()
    if mode == "BRIGHT":
        data, angles = remove_projections(data, angles, bad)
        data, bg = cut_edges(data, cutedges)
        data = brightness_rescale(data - bg)
        return data, angles, 0.0
    if mode == "BASIC":
        data = data - data.min()
        data, angles = remove_projections(data, angles, bad)
        data = data - data.min()
        data, bg = cut_edges(data, cutedges)
        return data, angles, bg
    return data, angles, 0.0


def scale_data(data, section, scalepars, dim):
    """Percentile scaling (RC:159-162 for 2D, RC:348-350 for 3D): the mass
    ratio sum(data[:, section]) / n_angles is taken after the low-percentile
    shift in 2D and before it in 3D, as the reference does.  Returns
    (scaled data, ratio)."""
    data = np.array(data, dtype=np.float64, copy=True)
    na = data.shape[0]
    if dim == 3:
        ratio = float(np.sum(data[:, list(section)]) / na)
        data -= np.percentile(data, scalepars[0])
    else:
        data -= np.percentile(data, scalepars[0])
        ratio = float(np.sum(data[:, list(section)]) / na)
    data /= np.percentile(data, scalepars[1])
    return data, ratio


def apply_baseline(sino, mode, threshold, averagenoise, dim):
    """Baseline rules (2D RC:193-208: 'mean', 'meanthres', 'thresholding',
    'mean_threshold_bright'; 3D RC:364-378: 'mean1', 'mean2', 'thresholding',
    'mean_threshold_bright').  'mean_threshold_bright' zeroes values below
    the threshold and then scales every projection (first axis) so that its
    mean equals the mean of the projection means (Brightness_correction2,
    H5:256-269).  The reference also re-reads the last channel's file there
    (H5:273-312) and keeps only its angle list, which is the loaded one
    whenever the brightness mode removed the same bad projections; the
    loaded angles are kept here."""
    m = (mode or "").upper()
    sino = np.array(sino, dtype=np.float64, copy=True)
    if (dim == 2 and m == "MEAN") or (dim == 3 and m == "MEAN1"):
        return sino - averagenoise
    if (dim == 2 and m == "MEANTHRES") or (dim == 3 and m == "MEAN2"):
        return sino - np.mean(sino[sino < threshold])
    if m == "THRESHOLDING":
        sino[sino < threshold] = 0.0
        return sino
    if m == "MEAN_THRESHOLD_BRIGHT":
        sino[sino < threshold] = 0.0
        means = sino.reshape(sino.shape[0], -1).mean(axis=1)
        scale = means.mean() / means
        return sino * scale.reshape((-1,) + (1,) * (sino.ndim - 1))
    return sino


# --------------------------------------------------------------------------
# slabs (RC:392-540)
# --------------------------------------------------------------------------


def slices_per_device(nd, na, nch, free_bytes, fraction=0.7):
    """Slices of an nd x nd grid with na angles and nch channels that fit in
    ``fraction`` of ``free_bytes`` (the reference's Allows_number_slices
    RC:96-106 for its own layout; here the compact state: 18 image arrays +
    3 sinogram arrays per voxel-channel, FP32)."""
    per_slice = 4 * (18 * nd * nd + 3 * na * nd) * nch
    return int(fraction * free_bytes // per_slice)


def slab_schedule(n, possible, overlay):
    """Slabs of the reference's loop RC:408-458 over n slices: each slab is
    (first, last) indices (inclusive) plus the number of its leading and
    trailing slices that are overlap only (solved, not stored).  ``overlay``
    is capped at possible // 4 as RC:409-412 does."""
    if possible < 1:
        raise ValueError("possible must be >= 1")
    out = []
    remaining = list(range(n))
    while remaining:
        if overlay > possible / 4.0:
            overlay = int(possible / 4)
        current = remaining[:min(possible, len(remaining))]
        start, end = current[0], current[-1]
        localending = possible - 2 * overlay if possible < len(remaining) else len(remaining)
        ov_s = 0 if len(remaining) == n else overlay
        ov_e = len(current) - len(remaining) if possible >= len(remaining) else overlay
        out.append((start, end, ov_s, ov_e))
        if localending <= 0:
            raise ValueError("slab step is zero: overlay too large for the slab size")
        remaining = remaining[localending:]
    return out


def _device_oom_types():
    try:
        import torch
        return (GraptorMemoryError, torch.cuda.OutOfMemoryError)
    except ImportError:  # pragma: no cover - torch is a dependency of the solver
        return (GraptorMemoryError,)


def solve_slabs(solve, nz, possible, overlay, C, nd, na, log=None):
    """The slab loop of RC:405-536 around a solver callable.

    ``solve(start, end, ov_s, ov_e)`` solves slices start..end (inclusive) and
    returns (u (C, nd, nd, n), ku (C, na, n, nd)); the interior of each slab
    (its overlap slices excluded) is stored.  A device out-of-memory error
    (GraptorMemoryError / torch OOM, the reference's cl.MemoryError) and any
    other library failure (GraptorError, its cl.Error) shrink the slab to
    int(0.7 * possible) slices and restart the schedule (RC:510-536); below
    one slice the memory case raises MemoryError and the other case
    GraptorError.  Returns (U (C, nd, nd, nz), sinogram (C, na, nz, nd), the
    slab size that succeeded)."""
    log = sys.stderr if log is None else log
    oom = _device_oom_types()
    U = np.zeros((C, nd, nd, nz))
    sino_out = np.zeros((C, na, nz, nd))
    while True:
        try:
            for start, end, ov_s, ov_e in slab_schedule(nz, possible, overlay):
                u, ku = solve(start, end, ov_s, ov_e)
                n = end + 1 - start
                U[:, :, :, start + ov_s:end + 1 - ov_e] = u[:, :, :, ov_s:n - ov_e]
                sino_out[:, :, start + ov_s:end + 1 - ov_e, :] = ku[:, :, ov_s:n - ov_e, :]
            return U, sino_out, possible
        except oom as err:
            possible = int(possible * 0.7)
            print("Warning: Computation requires too much memory, the problem is split, "
                  f"trying to compute {possible} slices at once.", file=log)
            if possible < 1:
                raise MemoryError("Even a single slice requires too much memory.") from err
        except GraptorError as err:
            possible = int(possible * 0.7)
            print(f"An exception of type {type(err).__name__} occurred. Arguments:\n{err.args!r}",
                  file=log)
            print("Warning: unexpected device error occurred, trying to restart with "
                  f"{possible} slices at once.", file=log)
            if possible < 1:
                raise GraptorError("Even a single slice creates an unknown error.") from err


def faktor_rescale(solution, sinogram, ratio, na):
    """Divide each channel of the solution and of the model sinogram by
    Faktor / na with Faktor = sum(sinogram_c) / ratio_c (RC:477-480)."""
    solution = np.array(solution, dtype=np.float32, copy=True)
    sinogram = np.array(sinogram, dtype=np.float32, copy=True)
    for c in range(solution.shape[0]):
        faktor = float(np.sum(sinogram[c], dtype=np.float64)) / ratio[c]
        solution[c] /= faktor / na
        sinogram[c] /= faktor / na
    return solution, sinogram


def write_results(result, channelnames, solution, sinogram):
    """MRC files per channel (solution and sinogram) and, for several
    channels, the side-by-side collections, plain and per-channel normalised
    (RC:482-507).  solution: (C, Nz, Nx, Ny)."""
    C, Nz, Nx, Ny = solution.shape
    for c in range(C):
        gio.write_mrc(f"{result}_{channelnames[c]}.mrc", solution[c])
        gio.write_mrc(f"{result}_{channelnames[c]}_sinogram.mrc", sinogram[c])
    if C > 1:
        coll = np.concatenate([solution[c] for c in range(C)], axis=2)
        gio.write_mrc(f"{result}_all_channels.mrc", coll)
        norm = np.concatenate([solution[c] / np.max(solution[c]) for c in range(C)], axis=2)
        gio.write_mrc(f"{result}_all_channels_normalized.mrc", norm)


def _prepare(names, section, datahandling, scalepars, dim, reductionparameter=(0, None)):
    brightness, baseline, cutedges, badfile, threshold, bad = datahandling
    bad = list(bad)
    fraction, radius = reductionparameter
    if fraction != 0:
        # RC:319-320 appends the detected list as one element of the list,
        # which the membership test of H5:70 then never matches, so the
        # reference drops nothing; the detected projections are removed here
        from .badprojections import find_bad_projections
        bad.extend(find_bad_projections(names[0], radius, fraction))
    coll, ratio, noise, angles = [], [], [], None
    for name in names:
        data, angles, avg = load_preprocessed(name, brightness, badfile, float(cutedges), bad)
        if min(section) < 0 or max(section) >= data.shape[1]:
            raise IndexError(f"slice levels must lie between 0 and {data.shape[1] - 1}")
        data, r = scale_data(data, section, scalepars, dim)
        coll.append(data)
        ratio.append(r)
        noise.append(avg)
    return coll, ratio, noise, angles, baseline, float(threshold)


def coupled_reconstruction_3d(names, result, mu, alpha, maxiter, method, sinogram_level,
                              datahandling, channelnames, regularisation="TGV",
                              discrepancy="KL", overlay=1, chunksize=0, scalepars=(1, 98),
                              device=None, write=True, L=None, reductionparameter=(0, None)):
    """Counterpart of Coupled_Reconstruction3d (RC:294-540).  sinogram_level
    = (first, last, step); datahandling = [brightness, baseline, cutedges,
    badfile, threshold, bad_projections]; reductionparameter = (fraction or
    count, search radius) of the bad-projection detection (RC:319-320).  Returns U (C, Nx, Nx, nslices)
    float64 and the rescaled (Solution, Sinogram) arrays."""
    from . import radon_tgv_primal_dual_3d as r3d
    from .solver import resolve_device

    section = list(range(sinogram_level[0], sinogram_level[1] + 1, sinogram_level[2]))
    coll, ratio, noise, angles, baseline, threshold = _prepare(names, section, datahandling,
                                                               scalepars, 3, reductionparameter)
    sinos = [apply_baseline(d, baseline, threshold, noise[i], 3)[:, section, :]
             for i, d in enumerate(coll)]
    C = len(sinos)
    na, nz, nd = sinos[0].shape
    dev = resolve_device(device)
    import torch
    free = torch.cuda.mem_get_info(dev)[0]
    possible = min(nz, max(1, slices_per_device(nd, na, C, free)))
    if chunksize >= 1:
        possible = min(possible, int(chunksize))
    t0 = time.time()

    def solve(start, end, ov_s, ov_e):
        cur = [section[i] for i in range(start, end + 1)]
        info = [start / float(nz), (end + 1 - start - ov_e) / float(nz), cur]
        sc = [s[:, start:end + 1, :] for s in sinos]
        return r3d.Reconstructionapproach3d(sc, angles + np.pi / 2, tuple(alpha), mu, maxiter,
                                            dev, [0, 0], sinogram_level[2], discrepancy, method,
                                            regularisation, info, L=L)

    U, sino_out, _ = solve_slabs(solve, nz, possible, overlay, C, nd, na)
    print("\nTime required for reconstruction:", int(10.0 * (time.time() - t0)) / 10.0,
          " seconds.\n")
    solution = np.ascontiguousarray(U.transpose(0, 3, 1, 2))      # (C, Nz, Nx, Ny)
    solution, newsino = faktor_rescale(solution, sino_out, ratio, len(angles))
    if write:
        write_results(result, channelnames, solution, newsino)
    return U, solution, newsino


def coupled_reconstruction_2d(names, result, mu, alpha, maxiter, method, slice_levels,
                              datahandling, channelnames, regularisation="TGV",
                              discrepancy="KL", scalepars=(1, 98), device=None, write=True,
                              L=None, batch=True, reductionparameter=(0, None)):
    """Counterpart of Coupled_Reconstruction2d (RC:109-292): every selected
    slice is an independent 2D problem.  With ``batch`` they are solved
    together in one batch plan (one launch per kernel class per phase);
    otherwise one Reconstructionapproach2d call per slice as the reference
    does.  Returns the rescaled (Solution (C, Nz, Nx, Nx), Sinogram
    (C, na, Nz, Nx))."""
    import torch

    from . import radon_tgv_primal_dual_2d as r2d
    from .batch import BatchPlan
    from .projector import RadonGeometry, radon_normest
    from .solver import METHODS, TgvProblem, resolve_device

    considered = list(range(slice_levels[0], slice_levels[1] + 1, slice_levels[2]))
    coll, ratio, noise, angles, baseline, threshold = _prepare(names, considered,
                                                               datahandling, scalepars, 2,
                                                               reductionparameter)
    C = len(coll)
    na, _, nd = coll[0].shape
    Nz = len(considered)
    if C == 1:
        method = "Uncorrelated 2D Reconstruction"  # RC:233
    dev = resolve_device(device)
    solution = np.zeros((C, Nz, nd, nd), np.float32)
    sino_out = np.zeros((C, na, Nz, nd), np.float32)
    t0 = time.time()
    sinos = [[apply_baseline(coll[i][:, lvl, :], baseline, threshold, noise[i], 2)
              for i in range(C)] for lvl in considered]
    if not batch:
        for j, lvl in enumerate(considered):
            info = [j / float(Nz), 1.0 / Nz, lvl]
            u, ks = r2d.Reconstructionapproach2d(sinos[j], angles + np.pi / 2, tuple(alpha), mu,
                                                maxiter, dev, [0, 0], discrepancy, method,
                                                regularisation, info, L=L)
            solution[:, j] = u
            for c in range(C):
                sino_out[c, :, j, :] = ks[c]
    else:
        coupling = METHODS[method][0]
        geom = RadonGeometry((nd, nd), angles + np.pi / 2, n_detectors=nd)
        if L is None:
            L = radon_normest(None, geom, mode="power", device=dev)
        probs = []
        for j in range(Nz):
            f = np.stack([np.asarray(s, np.float32) for s in sinos[j]])[:, None]
            probs.append(TgvProblem(f, geom, alpha, mu, dim=2, coupling=coupling,
                                    discrepancy=discrepancy, regularisation=regularisation,
                                    L=L, device=dev, fused=False))
        plan = BatchPlan(probs)
        plan.iterate(int(maxiter), ku_on_last=True)
        torch.cuda.synchronize(dev)
        for j, p in enumerate(probs):
            # 2D solution layout [C, x, y] as Reconstructionapproach2d returns it
            solution[:, j] = p.u.view(C, nd, nd).transpose(1, 2).cpu().numpy()
            sino_out[:, :, j, :] = p.ku.view(C, na, nd).cpu().numpy()
    print("\nTime required for reconstruction:", int(10.0 * (time.time() - t0)) / 10.0,
          " seconds.\n")
    solution, newsino = faktor_rescale(solution, sino_out, ratio, len(angles))
    if write:
        write_results(result, channelnames, solution, newsino)
    return solution, newsino
