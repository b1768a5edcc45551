"""FP64 NumPy oracle of Graptor's multi-channel TGV primal-dual reconstruction.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``graptor_amd``) imports
this module; it is the checker that ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` compare the HIP path against.

It restates, in float64 and with plain NumPy, the discretisation that the
reference implements in OpenCL C inside ``radon_tgv_primal_dual_2d.py`` (2D:)
and ``radon_tgv_primal_dual_3d.py`` (3D:).  Every function cites the lines it
follows.  Array conventions (identical memory order to the reference's
pyopencl ``order='F'`` buffers, first index fastest):

* image / primal ``u``      : ``(C, Nz, Ny, Nx)``      (2D: Nz == 1)
* vector fields ``p``, ``v``: ``(d, C, Nz, Ny, Nx)``   d = 2 (2D) or 3 (3D)
* tensor field ``w``        : ``(m, C, Nz, Ny, Nx)``   m = 3 (2D) or 6 (3D);
  2D order (xx, yy, xy), 3D order (xx, yy, zz, xy, xz, yz)
* sinograms ``f``, ``lam``  : ``(C, Nz, na, nd)``

How each function is pinned (tests/test_oracle.py): the projector against an
explicit brute-force hat matrix and the dot-product identity; the stencils
against their adjoint identities; the projections against closed forms and
``numpy.linalg.svd``; the proximal maps against the Moreau identity and a
scalar minimisation; the solver against its own energy/gap behaviour.
"""
from __future__ import annotations

import math
import numpy as np

# Step-size bound sigma*tau*||K||^2 <= 1 with the operator stack of the 2D and
# 3D solvers (2D:754, 3D:1181).
LSQR_2D = 0.5 * (18.0 + math.sqrt(33.0))
LSQR_3D = 17.0
# The reference weights off-diagonal tensor entries with this float literal
# rather than sqrt(2) (2D:254, 3D:582, 3D:606).
SQRT2_REF = float(np.float32(1.4142))

COUPLINGS = ("uncorrelated", "frobenius", "nuclear")


# --------------------------------------------------------------------------
# geometry  (2D:503-535, 3D:947-979)
# --------------------------------------------------------------------------
def radon_struct(shape, angles, n_detectors=None, detector_width=1.0,
                 detector_shift=0.0):
    """Per-angle table ``ofs[4, na]`` = (cos/dw, sin/dw, offset, 1/cos), FP32.

    Follows 2D:503-535: scalar ``angles`` means that many equispaced angles
    in [0, pi); the default detector count is ceil(hypot(Nx, Ny)); rays with
    |1/cos| > 10*nd are snapped to vertical (cos = 0, sin rounded, 1/cos = 0).
    The table is computed in float64 and stored in float32, as the reference
    does; the oracle then computes with the float32 values widened to float64
    so that it discretises exactly the same geometry.
    Returns ``(ofs, shape, (nd, na))``.
    """
    if np.isscalar(angles):
        angles = np.linspace(0, np.pi, int(angles) + 1)[:-1]
    angles = np.asarray(angles, dtype=np.float64)
    if n_detectors is None:
        nd = int(math.ceil(math.hypot(shape[0], shape[1])))
    else:
        nd = int(n_detectors)
    mid_x = (shape[0] - 1) / 2.0
    mid_y = (shape[1] - 1) / 2.0
    mid_d = (nd - 1.0) / 2.0
    cx = np.cos(angles) / detector_width
    sy = np.sin(angles) / detector_width
    with np.errstate(divide="ignore"):
        cinv = 1.0 / cx
    vertical = np.abs(cinv) > 10 * nd
    cx[vertical] = 0.0
    sy[vertical] = np.round(np.sin(angles[vertical])) / detector_width
    cinv[vertical] = 0.0
    offset = mid_d - cx * mid_x - sy * mid_y + detector_shift / detector_width
    ofs = np.zeros((4, len(angles)), dtype=np.float32)
    ofs[0], ofs[1], ofs[2], ofs[3] = cx, sy, offset, cinv
    return ofs, tuple(shape), (nd, len(angles))


# --------------------------------------------------------------------------
# projector  (kernels radon 2D:391-434 / 3D:773-816, radon_ad 2D:436-462)
# --------------------------------------------------------------------------
def _detector_coords(ofs, nx, ny):
    """t[j, y, x] = x*cos_j + y*sin_j + offset_j in float64."""
    o = ofs.astype(np.float64)
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    return (o[0][:, None, None] * x[None, None, :]
            + o[1][:, None, None] * y[None, :, None] + o[2][:, None, None])


def radon_fwd(img, ofs, nd):
    """Forward projection ``img (..., Ny, Nx) -> sino (..., na, nd)``.

    The reference (2D:391-434) walks each ray and sums hat weights
    max(0, 1 - |x c + y s + off - i|) over the pixels of every row.  A pixel
    at detector coordinate t has positive hat weight only for bins
    floor(t) (weight 1 - frac t) and floor(t) + 1 (weight frac t), so the
    same sum is formed here by splatting every pixel into those two bins
    (bins outside [0, nd) dropped, as the kernel is only run over [0, nd)).
    The brute-force matrix ``radon_matrix`` pins this equality.
    """
    img = np.asarray(img, dtype=np.float64)
    lead = img.shape[:-2]
    ny, nx = img.shape[-2:]
    flat = img.reshape(-1, ny * nx)
    nsl = flat.shape[0]
    na = ofs.shape[1]
    out = np.zeros((nsl, na, nd))
    t = _detector_coords(ofs, nx, ny).reshape(na, -1)
    for j in range(na):
        i0 = np.floor(t[j])
        fr = t[j] - i0
        i0 = i0.astype(np.int64)
        for k, wgt in ((0, 1.0 - fr), (1, fr)):
            b = i0 + k
            ok = (b >= 0) & (b < nd)
            if not ok.any():
                continue
            for s in range(nsl):
                out[s, j] += np.bincount(b[ok], weights=wgt[ok] * flat[s, ok],
                                         minlength=nd)
    return out.reshape(lead + (na, nd))


def radon_adj(sino, ofs, nx, ny):
    """Backprojection ``sino (..., na, nd) -> img (..., Ny, Nx)``.

    Follows 2D:436-462: for every pixel and angle, t = x c + y s + off; if
    -1 < t < nd, add (1 - frac t) sino[floor t] (when floor t >= 0) and
    frac t * sino[floor t + 1] (when floor t <= nd - 2).  Exactly the
    transpose of ``radon_fwd``.
    """
    sino = np.asarray(sino, dtype=np.float64)
    lead = sino.shape[:-2]
    na, nd = sino.shape[-2:]
    flat = sino.reshape(-1, na, nd)
    out = np.zeros((flat.shape[0], ny * nx))
    t = _detector_coords(ofs, nx, ny).reshape(na, -1)
    for j in range(na):
        tj = t[j]
        inside = (tj > -1) & (tj < nd)
        i0 = np.floor(tj)
        fr = tj - i0
        i0 = i0.astype(np.int64)
        m0 = inside & (i0 >= 0)
        m1 = inside & (i0 <= nd - 2)
        out[:, m0] += (1.0 - fr[m0]) * flat[:, j, i0[m0]]
        out[:, m1] += fr[m1] * flat[:, j, i0[m1] + 1]
    return out.reshape(lead + (ny, nx))


def radon_matrix(ofs, nx, ny, nd):
    """Explicit ``(na*nd, ny*nx)`` matrix of hat weights max(0, 1-|t-i|).

    Brute force straight from the ray-sum definition of 2D:405-432 (every
    pixel, every bin, no bounds logic); only for tiny grids.
    """
    na = ofs.shape[1]
    t = _detector_coords(ofs, nx, ny).reshape(na, 1, ny * nx)
    i = np.arange(nd, dtype=np.float64).reshape(1, nd, 1)
    return np.maximum(0.0, 1.0 - np.abs(t - i)).reshape(na * nd, ny * nx)


def feasible_mask(ofs, nx, ny, nd):
    """Bins reached by any pixel: R 1 > 0 on one slice (2D:476-486)."""
    return radon_fwd(np.ones((ny, nx)), ofs, nd) > 0


def normest_reference(ofs, nx, ny, nd, x0, iters=10):
    """Sum-normalised power iteration on R^T R (2D:576-588).

    ``x0`` is the (ny, nx) start image (the reference draws it from an
    unseeded ``randn``).  Each round: n = sum(x); x /= n; x = R^T R x.
    Returns sqrt(n) of the last round.
    """
    x = np.asarray(x0, dtype=np.float64).copy()
    n = 1.0
    for _ in range(iters):
        n = float(x.sum())
        x = x / n
        x = radon_adj(radon_fwd(x, ofs, nd), ofs, nx, ny)
    return math.sqrt(n)


def normest_power(ofs, nx, ny, nd, iters=200, tol=1e-12, seed=0):
    """||R||_2 by L2-normalised power iteration on R^T R (converged)."""
    rng = np.random.default_rng(seed)
    x = np.abs(rng.standard_normal((ny, nx))) + 0.1
    lam = 0.0
    for _ in range(iters):
        x /= np.linalg.norm(x)
        y = radon_adj(radon_fwd(x, ofs, nd), ofs, nx, ny)
        new = float(np.vdot(x, y))
        x = y
        if abs(new - lam) <= tol * abs(new):
            lam = new
            break
        lam = new
    return math.sqrt(lam)


# --------------------------------------------------------------------------
# stencils  (update_v 2D:265-284 / 3D:615-634, update_w 2D:286-304 /
#            3D:637-660, update_u 2D:340-357 / 3D:706-729,
#            update_p 2D:359-376 / 3D:734-757)
# --------------------------------------------------------------------------
def _fwd_diff(a, axis):
    out = np.zeros_like(a)
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    src[axis] = slice(1, None)
    dst[axis] = slice(0, -1)
    out[tuple(dst)] = a[tuple(src)] - a[tuple(dst)]
    return out


def _bwd_diff(a, axis):
    """a[k] - a[k-1] for k > 0, zero at k == 0."""
    out = np.zeros_like(a)
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    src[axis] = slice(0, -1)
    dst[axis] = slice(1, None)
    out[tuple(dst)] = a[tuple(dst)] - a[tuple(src)]
    return out


def _div_bwd(a, axis):
    """-(forward difference)^T: a[k] (0 at the last k) - a[k-1] (k > 0)."""
    out = a.copy()
    last = [slice(None)] * a.ndim
    last[axis] = -1
    out[tuple(last)] = 0.0
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    src[axis] = slice(0, -1)
    dst[axis] = slice(1, None)
    out[tuple(dst)] -= a[tuple(src)]
    return out


def _div_fwd(a, axis):
    """-(backward difference)^T: a[k+1] (k < n-1) - a[k] (0 at k == 0)."""
    out = -a
    first = [slice(None)] * a.ndim
    first[axis] = 0
    out[tuple(first)] = 0.0
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    src[axis] = slice(1, None)
    dst[axis] = slice(0, -1)
    out[tuple(dst)] += a[tuple(src)]
    return out


# axes of a (C, Nz, Ny, Nx) array
_AX = {"x": 3, "y": 2, "z": 1}


def grad(u, dim, h=1.0):
    """Forward-difference gradient, zero at the last index, z-part / h."""
    comps = [_fwd_diff(u, _AX["x"]), _fwd_diff(u, _AX["y"])]
    if dim == 3:
        comps.append(_fwd_diff(u, _AX["z"]) / h)
    return np.stack(comps)


def div(v, dim, h=1.0):
    """Backward-difference divergence, the negative adjoint of ``grad``."""
    out = _div_bwd(v[0], _AX["x"]) + _div_bwd(v[1], _AX["y"])
    if dim == 3:
        out = out + _div_bwd(v[2], _AX["z"]) / h
    return out


def symgrad(p, dim, h=1.0):
    """Symmetrised backward-difference gradient E p (2D:290-297, 3D:647-654)."""
    dx = lambda a: _bwd_diff(a, _AX["x"])
    dy = lambda a: _bwd_diff(a, _AX["y"])
    if dim == 2:
        return np.stack([dx(p[0]), dy(p[1]), 0.5 * (dx(p[1]) + dy(p[0]))])
    dz = lambda a: _bwd_diff(a, _AX["z"]) / h
    return np.stack([dx(p[0]), dy(p[1]), dz(p[2]),
                     0.5 * (dx(p[1]) + dy(p[0])),
                     0.5 * (dx(p[2]) + dz(p[0])),
                     0.5 * (dy(p[2]) + dz(p[1]))])


def div2(w, dim, h=1.0):
    """Tensor divergence, the negative adjoint of ``symgrad`` when the
    off-diagonal entries are paired twice (2D:363-373, 3D:744-754)."""
    fx = lambda a: _div_fwd(a, _AX["x"])
    fy = lambda a: _div_fwd(a, _AX["y"])
    if dim == 2:
        return np.stack([fx(w[0]) + fy(w[2]), fx(w[2]) + fy(w[1])])
    fz = lambda a: _div_fwd(a, _AX["z"]) / h
    return np.stack([fx(w[0]) + fy(w[3]) + fz(w[4]),
                     fx(w[3]) + fy(w[1]) + fz(w[5]),
                     fx(w[4]) + fy(w[5]) + fz(w[2])])


def sym_pair_weights(dim):
    """Pairing weights b_k of <E p, w>: off-diagonals count twice."""
    return np.array([1, 1, 2.0]) if dim == 2 else np.array([1, 1, 1, 2, 2, 2.0])


# --------------------------------------------------------------------------
# coupled projections  (2D:98-262, 3D:290-613; dispatch 2D:640-690,
#                       3D:1079-1130)
# --------------------------------------------------------------------------
def w_norm_weights(dim, coupling):
    """Squared per-component weights a_k of the W-ball norm.

    2D Frobenius weighs the off-diagonal by 2 (2D:232), 2D uncorrelated by
    1.4142 (2D:254); 3D Frobenius weighs all three off-diagonals by 1.4142
    (3D:582); 3D uncorrelated weighs xy and xz by 1.4142 but not yz (3D:606).
    Nuclear coupling projects W with the Frobenius rule (2D:664, 3D:1121).
    """
    r2 = SQRT2_REF ** 2
    if dim == 2:
        if coupling == "uncorrelated":
            return np.array([1.0, 1.0, r2])
        return np.array([1.0, 1.0, 4.0])
    if coupling == "uncorrelated":
        return np.array([1.0, 1.0, 1.0, r2, r2, 1.0])
    return np.array([1.0, 1.0, 1.0, r2, r2, r2])


def proj_v(v, alpha_inv, coupling, dim, nuclear_mode="svd"):
    """Project v onto the alpha-ball of the coupling norm.

    Frobenius (2D:98-119, 3D:529-551): n = sqrt(sum over channels and
    components of v^2) * alpha_inv; divide by n where n > 1.
    Uncorrelated (2D:208-220, 3D:553-565): the same per channel.
    Nuclear (2D:124-204, 3D:290-520): clip the singular values of the C x d
    matrix [v_c]_c at 1/alpha_inv.  ``nuclear_mode='svd'`` is the exact
    spectral projection; ``'ref2d'`` restates the reference's closed-form
    2x2 Gram route including its diagonal-dominance shortcut (2D only).
    """
    v = np.asarray(v, dtype=np.float64)
    if coupling == "frobenius":
        n = np.sqrt((v ** 2).sum(axis=(0, 1))) * alpha_inv
        return v / np.maximum(n, 1.0)[None, None]
    if coupling == "uncorrelated":
        n = np.sqrt((v ** 2).sum(axis=0)) * alpha_inv
        return v / np.maximum(n, 1.0)[None]
    if coupling == "nuclear":
        if nuclear_mode == "ref2d":
            assert dim == 2
            return _proj_nuclear_ref2d(v, alpha_inv)
        return _proj_nuclear_svd(v, alpha_inv)
    raise ValueError(coupling)


def _proj_nuclear_svd(v, alpha_inv):
    d, C = v.shape[:2]
    A = np.moveaxis(v.reshape(d, C, -1), 2, 0).transpose(0, 2, 1)  # (N,C,d)
    U, S, Vt = np.linalg.svd(A, full_matrices=False)
    S = np.minimum(S, 1.0 / alpha_inv)
    P = np.einsum("nik,nk,nkj->nij", U, S, Vt)  # (N, C, d)
    return np.moveaxis(P.transpose(0, 2, 1), 0, 2).reshape(v.shape)


def _proj_nuclear_ref2d(v, alpha_inv, eps=1e-5):
    """Closed-form route of 2D:124-204 in float64.

    Gram G = A^T A = [[a, b], [b, c]].  When b^2 > eps * min(a, c)^2 (and
    min(a,c) != 0) the eigenvectors (g, b) and (h, b) with
    g, h = ((a-c) +- sqrt(4b^2 + (a-c)^2))/2 are used; otherwise G is
    treated as diagonal.  Singular values sigma_k = sqrt(eigenvalue_k) are
    clipped: scale S_k = min(1, 1/(alpha_inv sigma_k)); v_c <- v_c M with
    M = sum_k S_k e_k e_k^T.
    """
    a = (v[0] ** 2).sum(axis=0)
    b = (v[0] * v[1]).sum(axis=0)
    c = (v[1] ** 2).sum(axis=0)
    o = np.minimum(a, c)
    with np.errstate(divide="ignore", invalid="ignore"):
        full = (o * o > 0) & ((b * b) / (o * o) > eps)
    d_ = a + c
    e_ = a - c
    f_ = np.sqrt(np.maximum(4 * b * b + e_ * e_, 0.0))
    s1 = np.where(full, np.sqrt(np.maximum((d_ + f_) * 0.5, 0.0)), np.sqrt(a))
    s2 = np.where(full, np.sqrt(np.maximum((d_ - f_) * 0.5, 0.0)), np.sqrt(c))
    g = (e_ + f_) * 0.5
    h = (e_ - f_) * 0.5
    e1 = np.stack([g * a + b * b, g * b + c * b])
    e2 = np.stack([h * a + b * b, h * b + b * c])
    n1 = np.hypot(e1[0], e1[1])
    n2 = np.hypot(e2[0], e2[1])
    with np.errstate(divide="ignore", invalid="ignore"):
        e1 = np.where(n1 > 0, e1 / np.where(n1 > 0, n1, 1), e1)
        e2 = np.where(n2 > 0, e2 / np.where(n2 > 0, n2, 1), e2)
        S1 = np.where(s1 * alpha_inv > 1, 1.0 / (s1 * alpha_inv), 1.0)
        S2 = np.where(s2 * alpha_inv > 1, 1.0 / (s2 * alpha_inv), 1.0)
    m00 = np.where(full, e1[0] ** 2 * S1 + e2[0] ** 2 * S2, S1)
    m11 = np.where(full, e1[1] ** 2 * S1 + e2[1] ** 2 * S2, S2)
    m01 = np.where(full, S1 * e1[0] * e1[1] + S2 * e2[0] * e2[1], 0.0)
    out = np.empty_like(v)
    out[0] = v[0] * m00 + v[1] * m01
    out[1] = v[0] * m01 + v[1] * m11
    return out


def proj_w(w, alpha_inv, coupling, dim):
    """Project w onto the alpha-ball of the weighted tensor norm
    (2D:223-258, 3D:573-610); nuclear coupling uses the Frobenius rule."""
    w = np.asarray(w, dtype=np.float64)
    a = w_norm_weights(dim, coupling).reshape(-1, *([1] * (w.ndim - 1)))
    if coupling == "uncorrelated":
        n = np.sqrt((a * w ** 2).sum(axis=0)) * alpha_inv
        return w / np.maximum(n, 1.0)[None]
    n = np.sqrt((a * w ** 2).sum(axis=(0, 1))) * alpha_inv
    return w / np.maximum(n, 1.0)[None, None]


# --------------------------------------------------------------------------
# proximal maps  (update_lambda_KL 2D:316-328 / 3D:690-704,
#                 update_lambda_L2 2D:330-338 / 3D:677-688)
# --------------------------------------------------------------------------
def prox_kl(lam, ku, f, sigma_p, mu):
    """lam <- s - (z + sqrt(z^2 + 4 sigma' f mu))/2, s = lam + sigma' Ku,
    z = s - mu; ``mu`` broadcast over the channel axis 0."""
    mu = np.asarray(mu, dtype=np.float64).reshape(-1, *([1] * (lam.ndim - 1)))
    s = lam + sigma_p * ku
    z = s - mu
    return s - 0.5 * (z + np.sqrt(z * z + 4.0 * sigma_p * f * mu))


def prox_l2(lam, ku, f, sigma_p, mfac):
    """lam <- (lam + sigma'(Ku - f)) * m with m = mu/(sigma + mu)."""
    mfac = np.asarray(mfac, dtype=np.float64).reshape(-1, *([1] * (lam.ndim - 1)))
    return (lam + sigma_p * (ku - f)) * mfac


# --------------------------------------------------------------------------
# solver  (tgv_radon 2D:710-851, 3D:1135-1292)
# --------------------------------------------------------------------------
def solver_constants(dim, L, alpha, mu, discrepancy):
    """Step sizes and scaled parameters as the host code derives them
    (2D:752-762, 3D:1180-1189)."""
    lsqr = LSQR_2D if dim == 2 else LSQR_3D
    sigma = tau = 1.0 / math.sqrt(lsqr)
    alpha_t = np.asarray(alpha, dtype=np.float64) / L ** 2
    mu = np.asarray(mu, dtype=np.float64)
    mfac = mu if discrepancy == "KL" else mu / (sigma + mu)
    return dict(sigma=sigma, tau=tau, sigma_p=sigma / L, norming=1.0 / L,
                alpha_inv0=1.0 / alpha_t[0], alpha_inv1=1.0 / alpha_t[1],
                mfac=mfac)


def solve(f, ofs, alpha, mu, maxiter, L, coupling="frobenius",
          discrepancy="KL", regularisation="TGV", dim=3, stepsize=1.0,
          u0=None, nuclear_mode="svd", callback=None, mask=None):
    """Primal-dual iteration of Appendix-A form (3D:1194-1221).

    ``f``: (C, Nz, na, nd) data (already zeroed where R 1 <= 0 unless
    ``mask`` is given, in which case it is applied here).  The image grid
    is nd x nd.  Returns dict with u (final primal), ku (R of the last
    extrapolated iterate, the reference's returned sinogram), p, v, w, lam.
    """
    f = np.asarray(f, dtype=np.float64)
    C, Nz, na, nd = f.shape
    N = nd
    if mask is not None:
        f = f * mask[None, None]
    k = solver_constants(dim, L, alpha, mu, discrepancy)
    d = 2 if dim == 2 else 3
    m = 3 if dim == 2 else 6
    shape = (C, Nz, N, N)
    u = np.zeros(shape) if u0 is None else np.asarray(u0, np.float64).copy()
    ub = u.copy()
    p = np.zeros((d,) + shape)
    pb = p.copy()
    v = np.zeros((d,) + shape)
    w = np.zeros((m,) + shape)
    lam = np.zeros_like(f)
    ku = np.zeros_like(f)
    tgv = regularisation.upper() == "TGV"
    for it in range(maxiter):
        v = v + k["sigma"] * (grad(ub, dim, stepsize) - pb)
        if tgv:
            w = w + k["sigma"] * symgrad(pb, dim, stepsize)
        v = proj_v(v, k["alpha_inv1"], coupling, dim, nuclear_mode)
        if tgv:
            w = proj_w(w, k["alpha_inv0"], coupling, dim)
        ku = radon_fwd(ub, ofs, nd)
        if discrepancy == "KL":
            lam = prox_kl(lam, ku, f, k["sigma_p"], k["mfac"])
        else:
            lam = prox_l2(lam, ku, f, k["sigma_p"], k["mfac"])
        ktl = radon_adj(lam, ofs, N, N)
        un = np.maximum(0.0, u + k["tau"] * (div(v, dim, stepsize)
                                             - k["norming"] * ktl))
        if tgv:
            pn = p + k["tau"] * (v + div2(w, dim, stepsize))
        else:
            pn = np.zeros_like(p)
        ub = 2.0 * un - u
        pb = 2.0 * pn - p
        u, p = un, pn
        if callback is not None:
            callback(it, u, p)
    return dict(u=u, ku=ku, p=p, v=v, w=w, lam=lam, ubar=ub, pbar=pb)


# --------------------------------------------------------------------------
# energy (derived from the fixed points of the iteration; the reference has
# no energy code on this path, its Gap_TGV_KL 3D:882-927 is unused)
# --------------------------------------------------------------------------
def kl_divergence(y, f):
    """sum of y - f + f log(f/y) (f > 0), y (f == 0); +inf if y <= 0 < f."""
    y = np.asarray(y, np.float64)
    f = np.asarray(f, np.float64)
    pos = f > 0
    if np.any(pos & (y <= 0)):
        return np.inf
    out = np.where(pos, y - f + f * np.log(np.where(pos, f, 1.0) /
                                           np.where(pos, y, 1.0)), y)
    return float(out.sum())


def v_dual_norm(q, coupling):
    """Pointwise norm dual to the V-ball of ``proj_v``, summed over voxels."""
    if coupling == "frobenius":
        return float(np.sqrt((q ** 2).sum(axis=(0, 1))).sum())
    if coupling == "uncorrelated":
        return float(np.sqrt((q ** 2).sum(axis=0)).sum())
    d, C = q.shape[:2]
    A = np.moveaxis(q.reshape(d, C, -1), 2, 0).transpose(0, 2, 1)
    return float(np.linalg.svd(A, compute_uv=False).sum())


def w_dual_norm(q, coupling, dim):
    """Norm dual to the W-ball under the off-diagonal-twice pairing:
    sqrt(sum b_k^2 q_k^2 / a_k), summed over voxels."""
    a = w_norm_weights(dim, coupling)
    b = sym_pair_weights(dim)
    c = (b ** 2 / a).reshape(-1, *([1] * (q.ndim - 1)))
    if coupling == "uncorrelated":
        return float(np.sqrt((c * q ** 2).sum(axis=0)).sum())
    return float(np.sqrt((c * q ** 2).sum(axis=(0, 1))).sum())


def energy(u, p, f, ofs, alpha, mu, L, coupling="frobenius",
           discrepancy="KL", regularisation="TGV", dim=3, stepsize=1.0):
    """Primal objective at the iterate (u, p).

    KL: sum_c mu_c KL(R u_c, f_c) + TGV_{alpha/L}(u);
    L2: sum_c mu_c/2 ||R u_c - f_c||^2 + TGV_alpha(u).
    TGV(u; p) = a1 |grad u - p|_V* + a0 |E p|_W*.
    """
    u = np.asarray(u, np.float64)
    nd = f.shape[-1]
    ku = radon_fwd(u, ofs, nd)
    data = 0.0
    for c in range(u.shape[0]):
        if discrepancy == "KL":
            data += mu[c] * kl_divergence(ku[c], f[c])
        else:
            data += 0.5 * mu[c] * float(((ku[c] - f[c]) ** 2).sum())
    scale = 1.0 / L if discrepancy == "KL" else 1.0
    a0, a1 = alpha[0] * scale, alpha[1] * scale
    if regularisation.upper() == "TGV":
        reg = (a1 * v_dual_norm(grad(u, dim, stepsize) - p, coupling)
               + a0 * w_dual_norm(symgrad(p, dim, stepsize), coupling, dim))
    else:
        reg = a1 * v_dual_norm(grad(u, dim, stepsize), coupling)
    return data + reg
