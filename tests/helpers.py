"""Shared builders of seeded test problems (tests only)."""
import numpy as np

from oracle import tgv_oracle as O


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def angles_deg(lo, hi, n, endpoint=False):
    return np.deg2rad(np.linspace(lo, hi, n, endpoint=endpoint)) + np.pi / 2


def blob_phantom(rng, C, Nz, N, scale=1.0):
    """Piecewise-constant blobs, non-negative, (C, Nz, N, N)."""
    yy, xx = np.mgrid[0:N, 0:N] - (N - 1) / 2.0
    out = np.zeros((C, Nz, N, N))
    for c in range(C):
        for z in range(Nz):
            img = np.zeros((N, N))
            for _ in range(rng.integers(3, 7)):
                cx, cy = rng.uniform(-0.3, 0.3, 2) * N
                a, b = rng.uniform(0.08, 0.25, 2) * N
                img += rng.uniform(0.2, 1.0) * (((xx - cx) / a) ** 2 + ((yy - cy) / b) ** 2 < 1)
            out[c, z] = img * scale
    return out


def make_problem(seed, C, Nz, N, na, lo=-70, hi=70, poisson=None):
    """Seeded phantom and its (noisy) data; returns (u_true, f, ofs)."""
    rng = np.random.default_rng(seed)
    ofs, _, _ = O.radon_struct((N, N), angles_deg(lo, hi, na), n_detectors=N)
    u = blob_phantom(rng, C, Nz, N)
    f = O.radon_fwd(u, ofs, N)
    if poisson is not None:
        f = rng.poisson(np.maximum(f, 0) * poisson) / poisson
    f = f * O.feasible_mask(ofs, N, N, N)[None, None]
    return u, f, ofs
