"""Detection of unsuitable projections in a tilt series (find_bad_ele.py =
FB:), CPU pre-processing of the caller.

Each projection is smoothed, turned into an edge map (largest slope to any
of the 8 neighbours times the smallest neighbour value), and compared with
its neighbours one and two tilts away by the best match over small shifts.
Projections whose combined mismatch is largest are reported (1-based), the
way the caller feeds them to the bad-projection removal.
"""
from __future__ import annotations

import math

import numpy as np

from . import io as gio
from .pipeline import brightness_rescale


def reduce_noise(U, steps):
    """Box sum over the (2 steps + 1)^2 window divided by 4 steps^2, on the
    interior (FB:68-75; the divisor is the reference's)."""
    Nx, Ny = U.shape
    E = np.zeros((Nx - 2 * steps, Ny - 2 * steps))
    for k in range(-steps, steps + 1):
        for l in range(-steps, steps + 1):
            E += U[steps + k:Nx - steps + k, steps + l:Ny - steps + l]
    return E / (4 * steps ** 2)


def compare_shift(U, V, steps):
    """Pointwise minimum over shifts (k, l) in [-steps, steps]^2 of
    |U - V shifted| on U's interior; returns (Frobenius norm, map) (FB:76-99)."""
    Nx0, Ny0 = U.shape
    Uc = U[steps:Nx0 - steps, steps:Ny0 - steps]
    E = None
    for k in range(-steps, steps + 1):
        for l in range(-steps, steps + 1):
            d = np.abs(Uc - V[steps + k:Nx0 - steps + k, steps + l:Ny0 - steps + l])
            E = d if E is None else np.minimum(E, d)
    return float(np.linalg.norm(E)), E


def compare_shift_twosided(U, V, steps):
    """max of the one-sided maps of (U, V) and (V, U); returns its l1 norm
    and the map (FB:53-66; the square root there is computed and dropped)."""
    _, E1 = compare_shift(U, V, steps)
    _, E2 = compare_shift(V, U, steps)
    E = np.maximum(E1, E2)
    return float(np.abs(E).sum()), E


def weighted_gradient(u):
    """max over the 8 neighbour directions of |u - u_shift| / distance, times
    the min over the directions of |u_shift|, on the interior (FB:102-131)."""
    Nx, Ny = u.shape
    c = u[1:Nx - 1, 1:Ny - 1]
    dirs = [(i, j) for i in (-1, 0, 1) for j in (-1, 1)] + [(-1, 0), (1, 0)]
    E = None
    Nb = None
    for i, j in dirs:
        sh = u[1 + i:Nx - 1 + i, 1 + j:Ny - 1 + j]
        e = np.abs(c - sh) / math.sqrt(i * i + j * j)
        n = np.abs(sh)
        E = e if E is None else np.maximum(E, e)
        Nb = n if Nb is None else np.minimum(Nb, n)
    value = E * Nb
    return float(np.linalg.norm(value)), value


def combined_errors(error1, error2):
    """Per projection i the smaller of the two pairings (i, i+1)/(i, i+2)
    and (i-1, i)/(i-2, i), indices clipped at the ends (FB:179-193)."""
    N = len(error1) + 1
    out = []
    for i in range(N):
        j = min(max(i, 0), N - 2)
        jj = min(max(i - 1, 0), N - 2)
        jjj = min(max(i, 0), N - 3)
        jjjj = min(max(i - 2, 0), N - 3)
        a = error1[j] + 0.5 * error2[jjj]
        b = error1[jj] + 0.5 * error2[jjjj]
        out.append(min(a, b))
    return out


def _round_half_away(x):
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def select_worst(errors, fraction):
    """1-based indices of the round(N * fraction) largest errors, largest
    first, ties to the lower index; fraction >= 1 is a count (FB:196-207)."""
    N = len(errors)
    if fraction >= 1:
        fraction = float(fraction) / N
    order = sorted(range(N), key=lambda i: (-errors[i], i))
    return [i + 1 for i in order][:_round_half_away(N * fraction)]


def find_bad_projections_array(data, searchradius=None, fraction=0.1, smoothing_step=3,
                               target_width=250):
    """Bad projections of a brightness-corrected series (tilts, slices,
    bins) (FB:134-209).  fraction == -1 means 0.1."""
    if fraction == -1:
        fraction = 0.1
    NA, Nz, Nx = data.shape
    Nz_new = int(target_width * Nz / float(Nx))
    if Nx > target_width:
        # the reference's integer strides (Python 2 division)
        data = data[:, ::max(1, Nz // max(Nz_new, 1))][:, :, ::max(1, Nx // target_width)]
    data = np.stack([reduce_noise(d, smoothing_step) for d in data])
    NA, Nz, Nx = data.shape
    if searchradius is None:
        searchradius = min(Nx, Nz) / 10
    searchradius = int(searchradius)
    W = np.stack([weighted_gradient(d)[1] for d in data])
    e1 = [compare_shift_twosided(W[i], W[i + 1], searchradius)[0] for i in range(NA - 1)]
    e2 = [compare_shift_twosided(W[i], W[i + 2], int(searchradius * 1.5))[0]
          for i in range(NA - 2)]
    return select_worst(combined_errors(e1, e2), fraction)


def find_bad_projections(name, searchradius=None, fraction=0.1):
    """Find_Bad_projections (FB:134): the series is read and brightness
    corrected without removals or edge cuts (FB:136)."""
    data, _ = gio.read_tilt_series(name)
    return find_bad_projections_array(brightness_rescale(data), searchradius, fraction)
