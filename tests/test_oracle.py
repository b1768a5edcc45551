"""Pins the FP64 oracle and the FP32 C restatement against mathematics that
fixes them independently (closed forms, brute force, adjoint identities,
SVD, scalar minimisation), so that a dropped term, a wrong sign or index or a
transposed operand fails here.  CPU only."""
import math

import numpy as np
import pytest
from scipy.optimize import minimize_scalar

from oracle import cref
from oracle import tgv_oracle as O
from tests.helpers import angles_deg, make_problem, rel

RNG = np.random.default_rng(12345)


# ---------------------------------------------------------------- geometry
def test_radon_struct_closed_form():
    ang = np.array([0.1, 0.7, 1.3, np.pi / 2, 2.9])
    ofs, shape, (nd, na) = O.radon_struct((20, 12), ang, n_detectors=25)
    assert (nd, na) == (25, 5)
    c, s = np.cos(ang), np.sin(ang)
    # pi/2 is snapped to a vertical ray: |1/cos| > 10 nd
    c[3], s[3] = 0.0, 1.0
    off = 12.0 - c * 9.5 - s * 5.5
    np.testing.assert_array_equal(ofs[0], c.astype(np.float32))
    np.testing.assert_array_equal(ofs[1], s.astype(np.float32))
    np.testing.assert_allclose(ofs[2], off.astype(np.float32), rtol=0, atol=2e-6)
    assert ofs[3][3] == 0.0
    # default detector count and scalar angles
    _, _, (nd2, na2) = O.radon_struct((20, 12), 7)
    assert nd2 == math.ceil(math.hypot(20, 12)) and na2 == 7


# --------------------------------------------------------------- projector
@pytest.mark.parametrize("shape,nd", [((9, 9), 9), ((12, 7), 15), ((16, 16), 22)])
def test_forward_equals_bruteforce_matrix(shape, nd):
    nx, ny = shape
    ofs, _, _ = O.radon_struct(shape, np.linspace(0.05, 3.1, 13), n_detectors=nd)
    M = O.radon_matrix(ofs, nx, ny, nd)
    img = RNG.random((2, ny, nx))
    s = O.radon_fwd(img, ofs, nd)
    ref = (M @ img.reshape(2, -1).T).T.reshape(s.shape)
    assert np.abs(s - ref).max() <= 1e-12 * np.abs(ref).max()
    sino = RNG.random((2, 13, nd))
    a = O.radon_adj(sino, ofs, nx, ny)
    refa = (M.T @ sino.reshape(2, -1).T).T.reshape(a.shape)
    assert np.abs(a - refa).max() <= 1e-12 * np.abs(refa).max()


def test_projector_axis_aligned_cases():
    """Vertical (snapped) rays give row sums; theta' = 0 gives column sums;
    with nd == N the offset is an integer and the sums are exact."""
    N = 10
    img = RNG.random((N, N))
    ofs, _, _ = O.radon_struct((N, N), np.array([np.pi / 2, 0.0]), n_detectors=N)
    s = O.radon_fwd(img, ofs, N)
    np.testing.assert_allclose(s[0], img.sum(axis=1), rtol=1e-13)   # t = y
    np.testing.assert_allclose(s[1], img.sum(axis=0), rtol=1e-13)   # t = x
    # odd nd - N: half-integer offset gives 0.5/0.5 blends of neighbours
    ofs2, _, _ = O.radon_struct((N, N), np.array([0.0]), n_detectors=N + 1)
    s2 = O.radon_fwd(img, ofs2, N + 1)[0]
    col = img.sum(axis=0)
    expect = np.zeros(N + 1)
    expect[:-1] += 0.5 * col
    expect[1:] += 0.5 * col
    np.testing.assert_allclose(s2, expect, rtol=1e-13)


def test_adjoint_dot_identity():
    ofs, _, _ = O.radon_struct((31, 31), angles_deg(-75, 75, 17), n_detectors=31)
    x = RNG.standard_normal((3, 31, 31))
    y = RNG.standard_normal((3, 17, 31))
    lhs = np.vdot(O.radon_fwd(x, ofs, 31), y)
    rhs = np.vdot(x, O.radon_adj(y, ofs, 31, 31))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_normest_matches_svd():
    ofs, _, _ = O.radon_struct((14, 14), angles_deg(-60, 60, 9), n_detectors=14)
    M = O.radon_matrix(ofs, 14, 14, 14)
    smax = np.linalg.svd(M, compute_uv=False)[0]
    assert abs(O.normest_power(ofs, 14, 14, 14) - smax) <= 1e-9 * smax
    # the reference's 10 sum-normalised rounds from a random start approach it
    x0 = np.random.default_rng(3).standard_normal((14, 14))
    est = O.normest_reference(ofs, 14, 14, 14, x0)
    assert abs(est - smax) <= 0.05 * smax


def test_feasible_mask_is_support_of_R1():
    ofs, _, _ = O.radon_struct((12, 12), angles_deg(-70, 70, 7), n_detectors=20)
    m = O.feasible_mask(ofs, 12, 12, 20)
    M = O.radon_matrix(ofs, 12, 12, 20)
    np.testing.assert_array_equal(m.reshape(-1), M.sum(axis=1) > 0)
    assert m[:, 0].sum() == 0 and m[:, -1].sum() == 0  # edge bins unreachable


# ---------------------------------------------------------------- stencils
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("h", [1.0, 2.0])
def test_stencil_adjoints(dim, h):
    Nz = 1 if dim == 2 else 5
    shp = (2, Nz, 7, 6)
    d = 2 if dim == 2 else 3
    m = 3 if dim == 2 else 6
    u = RNG.standard_normal(shp)
    v = RNG.standard_normal((d,) + shp)
    lhs = np.vdot(O.grad(u, dim, h), v)
    rhs = -np.vdot(u, O.div(v, dim, h))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
    p = RNG.standard_normal((d,) + shp)
    w = RNG.standard_normal((m,) + shp)
    b = O.sym_pair_weights(dim).reshape(-1, 1, 1, 1, 1)
    lhs2 = float((b * O.symgrad(p, dim, h) * w).sum())
    rhs2 = -np.vdot(p, O.div2(w, dim, h))
    assert abs(lhs2 - rhs2) <= 1e-12 * max(1.0, abs(lhs2))


def test_gradient_values_and_bound():
    u = np.zeros((1, 3, 4, 5))
    u[0, 1, 2, 3] = 1.0
    g = O.grad(u, 3, 2.0)
    assert g[0, 0, 1, 2, 2] == 1.0 and g[0, 0, 1, 2, 3] == -1.0
    assert g[1, 0, 1, 1, 3] == 1.0 and g[1, 0, 1, 2, 3] == -1.0
    assert g[2, 0, 0, 2, 3] == 0.5 and g[2, 0, 1, 2, 3] == -0.5
    assert np.all(g[0, :, :, :, -1] == 0)  # zero at the last index
    # |grad|^2 <= 8 in 2D (power iteration on -div grad)
    x = RNG.standard_normal((1, 1, 16, 16))
    for _ in range(300):
        x = -O.div(O.grad(x, 2), 2)
        lam = np.linalg.norm(x)
        x /= lam
    assert lam <= 8.0 + 1e-9 and lam > 7.0


def test_symgrad_values():
    p = np.zeros((3, 1, 3, 4, 4))
    p[1, 0, 1, 2, 2] = 1.0  # y-component
    e = O.symgrad(p, 3)
    assert e[1, 0, 1, 2, 2] == 1.0 and e[1, 0, 1, 3, 2] == -1.0   # d_y p_y
    assert e[3, 0, 1, 2, 2] == 0.5 and e[3, 0, 1, 2, 3] == -0.5   # (d_x p_y)/2
    assert e[5, 0, 1, 2, 2] == 0.5 and e[5, 0, 2, 2, 2] == -0.5   # (d_z p_y)/2
    assert np.all(e[0] == 0) and np.all(e[2] == 0) and np.all(e[4] == 0)


# ------------------------------------------------------------- projections
@pytest.mark.parametrize("coupling", ["frobenius", "uncorrelated"])
def test_v_projection_ball(coupling):
    v = RNG.standard_normal((3, 4, 1, 5, 5)) * 3
    ainv = 1.7
    pv = O.proj_v(v, ainv, coupling, 3)
    if coupling == "frobenius":
        n = np.sqrt((pv ** 2).sum(axis=(0, 1)))
    else:
        n = np.sqrt((pv ** 2).sum(axis=0))
    assert np.all(n <= 1 / ainv * (1 + 1e-12))
    np.testing.assert_allclose(O.proj_v(pv, ainv, coupling, 3), pv, rtol=1e-13)  # idempotent
    small = v * 1e-3
    np.testing.assert_array_equal(O.proj_v(small, ainv, coupling, 3), small)     # identity inside
    # projection = radial scaling with factor 1/(|v| ainv) where outside
    if coupling == "frobenius":
        n0 = np.sqrt((v ** 2).sum(axis=(0, 1))) * ainv
        np.testing.assert_allclose(pv, v / np.maximum(n0, 1)[None, None], rtol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_nuclear_projection_svd_properties(dim):
    v = RNG.standard_normal((dim, 4, 1, 3, 3)) * 2
    ainv = 0.8
    pv = O.proj_v(v, ainv, "nuclear", dim)
    A = np.moveaxis(pv.reshape(dim, 4, -1), 2, 0).transpose(0, 2, 1)
    S = np.linalg.svd(A, compute_uv=False)
    assert S.max() <= 1 / ainv * (1 + 1e-12)
    # singular vectors preserved: projection commutes with the SVD of v
    A0 = np.moveaxis(v.reshape(dim, 4, -1), 2, 0).transpose(0, 2, 1)
    U, S0, Vt = np.linalg.svd(A0, full_matrices=False)
    ref = np.einsum("nik,nk,nkj->nij", U, np.minimum(S0, 1 / ainv), Vt)
    np.testing.assert_allclose(A, ref, atol=1e-12)


def test_nuclear_2d_reference_route_matches_svd():
    """The reference's closed-form 2x2 Gram route equals the exact SVD
    projection when its diagonal-dominance shortcut is inactive."""
    v = RNG.standard_normal((2, 3, 1, 6, 6))
    for ainv in (0.3, 1.0, 4.0):
        np.testing.assert_allclose(O._proj_nuclear_ref2d(v, ainv),
                                   O._proj_nuclear_svd(v, ainv), atol=1e-12)
    # shortcut: a diagonal Gram matrix is treated as diagonal exactly
    w = np.zeros((2, 2, 1, 1, 1))
    w[0, 0] = 3.0
    w[1, 1] = 0.5
    np.testing.assert_allclose(O._proj_nuclear_ref2d(w, 1.0), O._proj_nuclear_svd(w, 1.0),
                               atol=1e-15)


def test_w_norm_weights_reference_quirks():
    assert list(O.w_norm_weights(2, "frobenius")) == [1, 1, 4]
    r2 = float(np.float32(1.4142)) ** 2
    np.testing.assert_allclose(O.w_norm_weights(2, "uncorrelated"), [1, 1, r2])
    np.testing.assert_allclose(O.w_norm_weights(3, "frobenius"), [1, 1, 1, r2, r2, r2])
    np.testing.assert_allclose(O.w_norm_weights(3, "uncorrelated"), [1, 1, 1, r2, r2, 1])
    w = RNG.standard_normal((6, 2, 1, 4, 4)) * 5
    pw = O.proj_w(w, 2.0, "frobenius", 3)
    a = O.w_norm_weights(3, "frobenius").reshape(-1, 1, 1, 1, 1)
    n = np.sqrt((a * pw ** 2).sum(axis=(0, 1)))
    assert np.all(n <= 0.5 * (1 + 1e-12))


# ------------------------------------------------------------------ prox
def test_kl_prox_minimises_prox_objective():
    """prox_{s F*}(q) for F(y) = mu (y - f log y): minimise
    F*(l) + (l - q)^2 / (2 s), F*(l) = -mu f log(mu - l) + const, l < mu."""
    for (lam, ku, f, sp, mu) in [(0.1, 2.0, 3.0, 0.4, 0.7), (-1.0, 0.3, 0.5, 1.3, 2.0),
                                 (0.0, 5.0, 1e-3, 0.2, 0.01)]:
        q = lam + sp * ku
        got = float(O.prox_kl(np.array([lam]), np.array([ku]), np.array([f]), sp,
                              np.array([mu]))[0])
        obj = lambda l: -mu * f * np.log(mu - l) + (l - q) ** 2 / (2 * sp)
        res = minimize_scalar(obj, bounds=(q - 50, mu - 1e-14), method="bounded",
                              options={"xatol": 1e-13})
        assert abs(got - res.x) < 1e-7 * max(1, abs(res.x))
    # f = 0: the conjugate is the indicator of l <= mu, prox = min(q, mu)
    out = O.prox_kl(np.array([0.5, 3.0]), np.array([1.0, 1.0]), np.zeros(2), 1.0, np.array([2.0]))
    np.testing.assert_allclose(out, [1.5, 2.0])


def test_l2_prox_closed_form():
    sp, m = 0.3, 0.8
    mfac = m / (sp + m)
    lam, ku, f = 0.2, 1.5, 0.7
    q = lam + sp * ku
    got = float(O.prox_l2(np.array([lam]), np.array([ku]), np.array([f]), sp,
                          np.array([mfac]))[0])
    obj = lambda l: l * l / (2 * m) + l * f + (l - q) ** 2 / (2 * sp)
    res = minimize_scalar(obj, bracket=(-10, 10), tol=1e-14)
    assert abs(got - res.x) < 1e-7


# ---------------------------------------------------------------- solver
def test_solver_converges_and_decreases_energy():
    """3D Frobenius TGV-KL: the energy at 1500 iterations is below the one at
    150, and the iterate is a local minimiser along random feasible
    perturbations."""
    u_true, f, ofs = make_problem(7, 2, 3, 16, 12, poisson=200.0)
    L = O.normest_power(ofs, 16, 16, 16)
    kw = dict(alpha=(4.0, 1.0), mu=[1.0, 0.5], L=L, coupling="frobenius",
              discrepancy="KL", dim=3)
    e = {}
    r = None
    for it in (150, 1500):
        r = O.solve(f, ofs, maxiter=it, **kw)
        e[it] = O.energy(r["u"], r["p"], f, ofs, (4.0, 1.0), [1.0, 0.5], L, "frobenius",
                         "KL", "TGV", 3)
    assert e[1500] < e[150]
    rng = np.random.default_rng(0)
    for _ in range(3):
        du = rng.standard_normal(r["u"].shape) * 1e-3
        up = np.maximum(r["u"] + du, 0)
        ep = O.energy(up, r["p"], f, ofs, (4.0, 1.0), [1.0, 0.5], L, "frobenius", "KL",
                      "TGV", 3)
        assert ep >= e[1500] - 1e-6 * abs(e[1500])


def test_solver_tv_is_tgv_with_zero_p():
    """TV mode equals the TGV iteration with p and w pinned at zero."""
    u_true, f, ofs = make_problem(8, 1, 1, 12, 8)
    L = O.normest_power(ofs, 12, 12, 12)
    r = O.solve(f, ofs, (2.0, 0.5), [1.0], 40, L, "uncorrelated", "L2", "TV", dim=2)
    assert np.all(r["p"] == 0) and np.all(r["w"] == 0)
    assert np.isfinite(r["u"]).all() and r["u"].min() >= 0


@pytest.mark.parametrize("dim,coupling,disc", [
    (2, "uncorrelated", "L2"), (2, "frobenius", "KL"), (3, "frobenius", "KL"),
    (3, "uncorrelated", "L2"), (3, "nuclear", "KL")])
def test_cref_restatement_matches_fp64_oracle(dim, coupling, disc):
    """The FP32 C restatement (the stand-in for the reference's OpenCL path)
    agrees with the FP64 oracle to FP32 accuracy."""
    Nz = 1 if dim == 2 else 3
    u_true, f, ofs = make_problem(21, 3, Nz, 18, 10)
    L = O.normest_power(ofs, 18, 18, 18)
    kw = dict(alpha=(4.0, 1.0), mu=[1.0, 0.4, 0.2], maxiter=50, L=L, coupling=coupling,
              discrepancy=disc, dim=dim)
    r64 = O.solve(f, ofs, **kw)
    r32 = cref.solve(f.astype(np.float32), ofs, **kw)
    assert rel(r32["u"], r64["u"]) < 2e-5
    assert rel(r32["ku"], r64["ku"]) < 2e-5
    assert rel(r32["p"], r64["p"]) < 1e-4


def test_cref_projector_matches_oracle():
    ofs, _, _ = O.radon_struct((40, 40), angles_deg(-75, 75, 21), n_detectors=40)
    img = RNG.random((3, 40, 40))
    assert rel(cref.radon(img, ofs, 40), O.radon_fwd(img, ofs, 40)) < 2e-6
    sino = RNG.random((3, 21, 40))
    assert rel(cref.radon_ad(sino, ofs, 40, 40), O.radon_adj(sino, ofs, 40, 40)) < 2e-6


def test_cref_jacobi_eigen():
    A = RNG.standard_normal((3, 3))
    G = A @ A.T
    ev, V = cref.jacobi3(np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1], G[0, 2], G[1, 2]]))
    np.testing.assert_allclose(np.sort(ev), np.linalg.eigvalsh(G), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(G @ V, V * ev[None, :], atol=1e-4)


def test_reference_2d_nuclear_fp32_route_is_ill_conditioned():
    """The reference's closed-form 2D nuclear projection (2D:124-204) is exact
    in FP64 (it equals the SVD projection) but cancels catastrophically in
    FP32: a whole FP32 iteration drifts from its FP64 evaluation by > 1e-3
    within 150 steps, and two FP32 compilations of the same text (multiply-adds
    rounded separately or contracted into FMAs, as OpenCL and HIP compilers
    do by default) drift apart by as much.  Frobenius coupling, the control,
    stays at FP32 accuracy in all three comparisons."""
    import os
    _, f, ofs = make_problem(31, 3, 1, 40, 24, poisson=50.0)
    f = f.astype(np.float32)
    L = O.normest_power(ofs, 40, 40, 40)
    cref.load()
    if not os.path.exists(cref.LIB_FMA):
        cref.build()
    fma = cref.load(cref.LIB_FMA)
    out = {}
    for cpl in ("nuclear", "frobenius"):
        kw = dict(alpha=(4.0, 1.0), mu=[1.0, 0.3, 0.1], maxiter=150, L=L, coupling=cpl,
                  discrepancy="KL", dim=2)
        plain = cref.solve(f, ofs, **kw)["u"]
        contracted = cref.solve(f, ofs, lib=fma, **kw)["u"]
        r64 = O.solve(f.astype(np.float64), ofs, nuclear_mode="ref2d", **kw)["u"]
        out[cpl] = (rel(plain, contracted), rel(plain, r64), rel(contracted, r64))
        if cpl == "nuclear":
            s64 = O.solve(f.astype(np.float64), ofs, nuclear_mode="svd", **kw)["u"]
            assert rel(r64, s64) < 1e-6
    print(f"2D nuclear: FP32 builds apart {out['nuclear'][0]:.1e}, vs FP64 "
          f"{out['nuclear'][1]:.1e} / {out['nuclear'][2]:.1e}; Frobenius {out['frobenius']}")
    assert min(out["nuclear"]) > 1e-3
    assert max(out["frobenius"]) < 2e-5
