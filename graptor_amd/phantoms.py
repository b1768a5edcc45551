"""Seeded synthetic phantoms and data for the benchmark configurations.

All generators use ``numpy.random.default_rng(seed)``; projections of the
phantoms are computed on the GPU with this package's own projector.
"""
from __future__ import annotations

import numpy as np
import torch

from .projector import RadonGeometry, radon


def tilt_angles(lo_deg, hi_deg, n, endpoint=True):
    """Tilt series as solver angles: theta + pi/2, as the reference's caller
    passes them (Reconstruction_coupled.py:233, 448)."""
    return np.deg2rad(np.linspace(lo_deg, hi_deg, n, endpoint=endpoint)) + np.pi / 2


def ellipse_phantom(rng, n_support, n_grid, n_min=8, n_max=15):
    """2D phantom: n_min..n_max random ellipses with intensities U[0.1, 1],
    additive, clipped >= 0, on an n_support square centred in n_grid."""
    img = np.zeros((n_grid, n_grid))
    off = (n_grid - n_support) // 2
    yy, xx = np.mgrid[0:n_support, 0:n_support] - (n_support - 1) / 2.0
    sub = np.zeros((n_support, n_support))
    for _ in range(rng.integers(n_min, n_max + 1)):
        cx, cy = rng.uniform(-0.3, 0.3, 2) * n_support
        a, b = rng.uniform(0.05, 0.3, 2) * n_support
        th = rng.uniform(0, np.pi)
        xr = (xx - cx) * np.cos(th) + (yy - cy) * np.sin(th)
        yr = -(xx - cx) * np.sin(th) + (yy - cy) * np.cos(th)
        sub += rng.uniform(0.1, 1.0) * ((xr / a) ** 2 + (yr / b) ** 2 < 1.0)
    img[off:off + n_support, off:off + n_support] = np.maximum(sub, 0.0)
    return img


def granule_volume(rng, N, Nz, n_maps=3, n_min=20, n_max=60):
    """3D granule phantom: a cylindrical matrix holding ``n_maps`` element
    maps of n_min..n_max ellipsoidal granules each.  Returns (matrix, maps)
    with maps shaped (n_maps, Nz, N, N), piecewise constant."""
    z, y, x = np.ogrid[0:Nz, 0:N, 0:N]
    cy = cx = (N - 1) / 2.0
    rad = 0.42 * N
    matrix = (((x - cx) ** 2 + (y - cy) ** 2) < rad ** 2) * np.ones((Nz, 1, 1))
    maps = np.zeros((n_maps, Nz, N, N))
    for k in range(n_maps):
        for _ in range(rng.integers(n_min, n_max + 1)):
            r = rad * np.sqrt(rng.uniform(0, 0.8))
            phi = rng.uniform(0, 2 * np.pi)
            gx, gy = cx + r * np.cos(phi), cy + r * np.sin(phi)
            gz = rng.uniform(0, Nz - 1)
            ax = rng.uniform(0.02, 0.08, 3) * np.array([N, N, max(Nz, 8)])
            lo = [max(0, int(gz - ax[2]) - 1), max(0, int(gy - ax[1]) - 1),
                  max(0, int(gx - ax[0]) - 1)]
            hi = [min(Nz, int(gz + ax[2]) + 2), min(N, int(gy + ax[1]) + 2),
                  min(N, int(gx + ax[0]) + 2)]
            zz = np.arange(lo[0], hi[0])[:, None, None]
            yy = np.arange(lo[1], hi[1])[None, :, None]
            xx = np.arange(lo[2], hi[2])[None, None, :]
            inside = (((xx - gx) / ax[0]) ** 2 + ((yy - gy) / ax[1]) ** 2
                      + ((zz - gz) / ax[2]) ** 2) < 1.0
            val = rng.uniform(0.5, 1.0)
            blk = maps[k, lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
            np.maximum(blk, val * inside, out=blk)
    return matrix, maps


def coupled_granule_problem(N=256, Nz=64, na=60, tilt=(-75.0, 75.0), seed_phantom=2001,
                            seed_noise=2002, haadf_counts=1e4, edx_counts=5.0,
                            device="cuda"):
    """4-channel HAADF + 3 x EDX problem of granules in a matrix.

    Channel 0 = weighted sum of matrix and element maps scaled to a mean of
    ``haadf_counts`` per illuminated bin, channels 1..3 = the element maps
    scaled to ``edx_counts``; Poisson noise.  Returns (f (4, Nz, na, N)
    float32 CUDA tensor, angles, phantom (4, Nz, N, N) float32 numpy).
    """
    rng = np.random.default_rng(seed_phantom)
    matrix, maps = granule_volume(rng, N, Nz)
    haadf = 0.25 * matrix + 1.0 * maps[0] + 0.6 * maps[1] + 0.8 * maps[2]
    phantom = np.stack([haadf, maps[0], maps[1], maps[2]]).astype(np.float32)
    angles = tilt_angles(tilt[0], tilt[1], na)
    geom = RadonGeometry((N, N), angles, n_detectors=N)
    img = torch.as_tensor(phantom, device=device)
    sino = torch.zeros((4, Nz, na, N), device=device)
    radon(sino.view(4 * Nz, na, N), img.view(4 * Nz, N, N), geom)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed_noise)
    f = torch.empty_like(sino)
    targets = [haadf_counts, edx_counts, edx_counts, edx_counts]
    scales = []
    for c in range(4):
        s = sino[c]
        mean = s[s > 0].mean()
        scale = float(targets[c] / mean) if mean > 0 else 1.0
        scales.append(scale)
        f[c] = torch.poisson(torch.clamp(s * scale, min=0), generator=gen)
    phantom_scaled = phantom * np.array(scales, np.float32)[:, None, None, None]
    return f, angles, phantom_scaled, geom


def ellipse_problem_2d(N_support=128, N_grid=183, na=90, lo=-70.0, hi=70.0, seed_phantom=1001,
                       seed_noise=1002, noise=0.01, device="cuda"):
    """Single-channel 2D problem: ellipse phantom, Gaussian noise of
    ``noise`` x max.  Returns (f (1, 1, na, N_grid) CUDA tensor, angles,
    phantom, geometry)."""
    rng = np.random.default_rng(seed_phantom)
    ph = ellipse_phantom(rng, N_support, N_grid).astype(np.float32)
    angles = tilt_angles(lo, hi, na, endpoint=False)
    geom = RadonGeometry((N_grid, N_grid), angles, n_detectors=N_grid)
    img = torch.as_tensor(ph, device=device).view(1, N_grid, N_grid)
    sino = torch.zeros((1, na, N_grid), device=device)
    radon(sino, img, geom)
    nrng = np.random.default_rng(seed_noise)
    s = sino.cpu().numpy()
    s = s + nrng.normal(0.0, noise * s.max(), s.shape)
    return (torch.as_tensor(s.astype(np.float32), device=device).view(1, 1, na, N_grid), angles,
            ph, geom)


def multichannel_params(rng, N, C, n_min=5, n_max=12):
    """Random draws of one C-channel slice: per element map a list of
    ellipses (cx, cy, a, b, theta, value) in pixels about the centre, and the
    map weights of channel 0."""
    nmaps = max(C - 1, 1)
    maps = []
    for _ in range(nmaps):
        ell = []
        for _ in range(rng.integers(n_min, n_max + 1)):
            cx, cy = rng.uniform(-0.3, 0.3, 2) * N
            a, b = rng.uniform(0.03, 0.15, 2) * N
            th = rng.uniform(0, np.pi)
            ell.append((cx, cy, a, b, th, rng.uniform(0.5, 1.0)))
        maps.append(ell)
    w = rng.uniform(0.5, 1.0, nmaps) if C > 1 else None
    return dict(N=N, C=C, maps=maps, w=w)


def multichannel_slice(rng, N, C, n_min=5, n_max=12):
    """C-channel 2D slice: channel 0 a weighted sum of the C-1 element maps
    (or one map when C == 1), the others the maps; each map holds
    n_min..n_max random ellipses with values U[0.5, 1] inside a disc."""
    prm = multichannel_params(rng, N, C, n_min, n_max)
    yy, xx = np.mgrid[0:N, 0:N] - (N - 1) / 2.0
    maps = np.zeros((len(prm["maps"]), N, N))
    for k, ell in enumerate(prm["maps"]):
        for cx, cy, a, b, th, val in ell:
            xr = (xx - cx) * np.cos(th) + (yy - cy) * np.sin(th)
            yr = -(xx - cx) * np.sin(th) + (yy - cy) * np.cos(th)
            np.maximum(maps[k], val * ((xr / a) ** 2 + (yr / b) ** 2 < 1.0), out=maps[k])
    matrix = (xx ** 2 + yy ** 2 < (0.42 * N) ** 2).astype(np.float64)
    if C == 1:
        return (0.25 * matrix + maps[0])[None]
    haadf = 0.25 * matrix + np.tensordot(prm["w"], maps, axes=1)
    return np.concatenate([haadf[None], maps], axis=0)


def multichannel_slices(seeds, N, C, device="cuda"):
    """The slices multichannel_slice(default_rng(seed), N, C) for many seeds
    at once, rasterised on the GPU in float64 (same draws and the same inside
    tests).  Returns a float32 tensor (len(seeds), C, N, N)."""
    prms = [multichannel_params(np.random.default_rng(int(s)), N, C) for s in seeds]
    B, nmaps = len(prms), max(C - 1, 1)
    E = max(len(ell) for p in prms for ell in p["maps"])
    P = np.zeros((B, nmaps, E, 6))
    for i, p in enumerate(prms):
        for k, ell in enumerate(p["maps"]):
            if ell:
                P[i, k, :len(ell)] = np.asarray(ell)
    Pt = torch.as_tensor(P, device=device)
    c = (torch.arange(N, dtype=torch.float64, device=device) - (N - 1) / 2.0)
    yy, xx = c.view(1, 1, N, 1), c.view(1, 1, 1, N)
    maps = torch.zeros((B, nmaps, N, N), dtype=torch.float64, device=device)
    for e in range(E):
        cx, cy, a, b, th, val = (Pt[:, :, e, j].view(B, nmaps, 1, 1) for j in range(6))
        valid = a > 0
        a = torch.where(valid, a, torch.ones_like(a))
        b = torch.where(valid, b, torch.ones_like(b))
        ct, st = torch.cos(th), torch.sin(th)
        xr = (xx - cx) * ct + (yy - cy) * st
        yr = -(xx - cx) * st + (yy - cy) * ct
        inside = ((xr / a) ** 2 + (yr / b) ** 2 < 1.0) & valid
        maps = torch.maximum(maps, val * inside)
    matrix = ((xx ** 2 + yy ** 2) < (0.42 * N) ** 2).to(torch.float64).view(1, N, N)
    if C == 1:
        return (0.25 * matrix + maps[:, 0]).unsqueeze(1).float()
    w = torch.as_tensor(np.stack([p["w"] for p in prms]), device=device)  # (B, nmaps)
    haadf = 0.25 * matrix + (w.view(B, nmaps, 1, 1) * maps).sum(dim=1)
    return torch.cat([haadf.unsqueeze(1), maps], dim=1).float()


def poisson_sinograms(img, geom, counts=(1e3, 20.0), seed=0):
    """Poisson data of a (B, C, N, N) image stack: channel 0 scaled to
    ``counts[0]`` mean counts over the bins its projection reaches, the other
    channels to ``counts[1]``; one torch generator seeded with ``seed``.
    Returns (B, C, na, nd) float32."""
    B, C = img.shape[:2]
    sino = torch.zeros((B * C, geom.na, geom.nd), device=img.device)
    radon(sino, img.reshape(B * C, *img.shape[2:]).contiguous(), geom)
    sino = sino.view(B, C, geom.na, geom.nd)
    pos = (sino > 0).to(sino.dtype)
    mean = (sino * pos).sum(dim=(2, 3)) / pos.sum(dim=(2, 3)).clamp(min=1)
    target = torch.full((B, C), float(counts[1]), device=img.device)
    target[:, 0] = float(counts[0])
    lam = torch.clamp(sino * (target / mean).view(B, C, 1, 1), min=0)
    gen = torch.Generator(device=img.device)
    gen.manual_seed(int(seed))
    return torch.poisson(lam, generator=gen)


def batched_2d_problem_set(n=64, seed=3000, counts=(1e3, 20.0), device="cuda"):
    """The 2D batch of uneven problems: for problem k (seed + k) the image
    size N ~ U{96..384} (grid nd = N), angles ~ U{30..180} over [-70, 70)
    degrees, channels ~ U{1..4}; channel 0 Poisson at ``counts[0]`` mean
    counts per bin, the others at ``counts[1]``.  Returns a list of dicts
    (f (C, 1, na, N) CUDA tensor, geometry, C, N, na, phantom)."""
    out = []
    for k in range(n):
        rng = np.random.default_rng(seed + k)
        N = int(rng.integers(96, 385))
        na = int(rng.integers(30, 181))
        C = int(rng.integers(1, 5))
        ph = multichannel_slice(rng, N, C).astype(np.float32)
        angles = tilt_angles(-70.0, 70.0, na, endpoint=False)
        geom = RadonGeometry((N, N), angles, n_detectors=N)
        img = torch.as_tensor(ph, device=device)
        sino = torch.zeros((C, na, N), device=device)
        radon(sino, img, geom)
        gen = torch.Generator(device=device)
        gen.manual_seed(seed + 100000 + k)
        f = torch.empty_like(sino)
        for c in range(C):
            s = sino[c]
            mean = s[s > 0].mean()
            target = counts[0] if c == 0 else counts[1]
            f[c] = torch.poisson(torch.clamp(s * (target / mean), min=0), generator=gen)
        out.append(dict(f=f.view(C, 1, na, N), geometry=geom, C=C, N=N, na=na, phantom=ph))
    return out
