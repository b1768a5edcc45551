"""GPU parity of the full primal-dual iteration against the FP32 restatement
of the reference kernels and the FP64 oracle, with the same injected
operator norm L on every side.  Bars: reconstruction after N iterations
within 1e-3 relative L2 (typically ~1e-6), primal energy within 1e-4."""
import numpy as np
import pytest

from oracle import cref
from oracle import tgv_oracle as O
from tests.helpers import make_problem, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

COMBOS = [
    (3, "frobenius", "KL", "TGV"),
    (3, "uncorrelated", "L2", "TGV"),
    (3, "nuclear", "KL", "TGV"),
    (3, "frobenius", "L2", "TV"),
    (2, "frobenius", "KL", "TGV"),
    (2, "uncorrelated", "L2", "TGV"),
    (2, "nuclear", "L2", "TGV"),
    (2, "uncorrelated", "KL", "TV"),
]


def _problem(dim, C=3, N=40, na=24, seed=31):
    Nz = 1 if dim == 2 else 5
    u_true, f, ofs = make_problem(seed, C, Nz, N, na, poisson=50.0)
    return u_true, f.astype(np.float32), ofs


# more than 8 channels: the dual step's two-sweep form
WIDE = [
    (3, "frobenius", "KL", "TGV"),
    (2, "nuclear", "KL", "TGV"),
    (3, "uncorrelated", "L2", "TGV"),
]


@pytest.mark.parametrize("combo", COMBOS, ids=["-".join(map(str, c)) for c in COMBOS])
def test_iteration_parity(combo):
    _parity(combo)


@pytest.mark.parametrize("combo", WIDE, ids=["-".join(map(str, c)) for c in WIDE])
def test_iteration_parity_more_than_eight_channels(combo):
    _parity(combo, C=10, iters=60)


def _parity(combo, C=3, iters=150):
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    dim, coupling, disc, reg = combo
    print(f"solver parity {combo}, {C} channels", flush=True)
    _, f, ofs = _problem(dim, C=C)
    C, Nz, na, N = f.shape
    L = O.normest_power(ofs, N, N, N)
    alpha = (4.0, 1.0)
    mu = [1.0, 0.3, 0.1] if C == 3 else [float(m) for m in np.linspace(1.0, 0.1, C)]
    # the device geometry must be built from the same float32 table
    g = RadonGeometry((N, N), np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2,
                      n_detectors=N)
    np.testing.assert_array_equal(g.ofs.T, ofs)
    prob = TgvProblem(f, g, alpha, mu, dim=dim, coupling=coupling, discrepancy=disc,
                      regularisation=reg, L=L)
    prob.iterate(iters, ku_on_last=True)
    torch.cuda.synchronize()
    u = prob.u.cpu().numpy()
    ku = prob.ku.cpu().numpy()
    kw = dict(alpha=alpha, mu=mu, maxiter=iters, L=L, coupling=coupling, discrepancy=disc,
              regularisation=reg, dim=dim)
    r32 = cref.solve(f, ofs, **kw)
    nm = "ref2d" if (coupling == "nuclear" and dim == 2) else "svd"
    r64 = O.solve(f.astype(np.float64), ofs, nuclear_mode=nm, **kw)
    e32 = rel(u, r32["u"])
    e64 = rel(u, r64["u"])
    print(f"  rel L2 u: vs FP32 restatement {e32:.2e}, vs FP64 oracle {e64:.2e}; "
          f"ku vs FP32 {rel(ku, r32['ku']):.2e}", flush=True)
    assert e64 <= 1e-4
    if coupling == "nuclear" and dim == 2:
        # the reference's FP32 closed form of the 2D nuclear projection is
        # itself off its FP64 evaluation by 1e-2 after 150 steps (and two FP32
        # compilations of it differ by as much); the GPU evaluates that route
        # in FP64, so it is held to the FP64 evaluation above, and to the FP32
        # restatement only within the restatement's own distance from it
        d32 = rel(r32["u"], r64["u"])
        print(f"  2D nuclear: FP32 restatement vs FP64 evaluation of its route {d32:.2e}",
              flush=True)
        assert e32 <= d32 + 1e-4
    else:
        assert e32 <= 1e-4
        assert rel(ku, r32["ku"]) <= 1e-4
    # primal energy: GPU FP64 reduction vs oracle at the GPU iterate
    data, regv, total = prob.energy()
    p = prob.p.cpu().numpy() if prob.p is not None else np.zeros((dim,) + u.shape)
    ref_e = O.energy(u, p, f.astype(np.float64), ofs, alpha, mu, L, coupling, disc, reg, dim)
    e_ref32 = O.energy(r32["u"], r32["p"], f.astype(np.float64), ofs, alpha, mu, L, coupling,
                       disc, reg, dim)
    if np.isinf(ref_e):
        # KL is +inf while some bin has R u = 0 and f > 0: every side must say so
        assert total == ref_e == e_ref32, (total, ref_e, e_ref32)
    else:
        assert abs(total - ref_e) <= 1e-5 * abs(ref_e), (total, ref_e)
        assert abs(total - e_ref32) <= 1e-4 * abs(e_ref32)


def test_dropin_3d_shapes_and_progress(capsys):
    from graptor_amd.radon_tgv_primal_dual_3d import Reconstructionapproach3d
    _, f, ofs = _problem(3, C=2, N=32, na=16)
    C, Nz, na, nd = f.shape
    angles = np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2
    sino = [np.ascontiguousarray(f[c].transpose(1, 0, 2)) for c in range(C)]  # [na, Nz, nd]
    L = O.normest_power(ofs, nd, nd, nd)
    sol, sinos = Reconstructionapproach3d(sino, angles, (4, 1), [1.0, 0.5], 50, 0, [0, 0.1], 1,
                                          "KL", "Frobenius 3D Reconstruction", "TGV",
                                          [0.0, 1.0, list(range(Nz))], L=L)
    out = capsys.readouterr().out
    assert "Overall progress at" in out and "for Section" in out
    assert sol.shape == (C, nd, nd, Nz) and sol.dtype == np.float64
    assert sinos.shape == (C, na, Nz, nd) and sinos.dtype == np.float64
    r = cref.solve(f, ofs, (4, 1), [1.0, 0.5], 50, L, "frobenius", "KL", dim=3)
    assert rel(sol, r["u"].transpose(0, 3, 2, 1)) <= 1e-4
    assert rel(sinos, r["ku"].transpose(0, 2, 1, 3)) <= 1e-4


def test_dropin_2d_shapes(capsys):
    from graptor_amd.radon_tgv_primal_dual_2d import Reconstructionapproach2d
    _, f, ofs = _problem(2, C=1, N=36, na=18)
    C, _, na, nd = f.shape
    angles = np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2
    L = O.normest_power(ofs, nd, nd, nd)
    sol, sinos = Reconstructionapproach2d([f[0, 0]], angles, (4, 1), [1.0], 40, None, [0, 0.1],
                                          "L2", "Uncorrelated 2D Reconstruction", "TGV",
                                          [0.0, 1.0, 7], L=L)
    assert "for Slice 7" in capsys.readouterr().out
    assert sol.shape == (1, nd, nd) and isinstance(sinos, list) and sinos[0].shape == (na, nd)
    r = cref.solve(f, ofs, (4, 1), [1.0], 40, L, "uncorrelated", "L2", dim=2)
    assert rel(sol, r["u"][:, 0].transpose(0, 2, 1)) <= 1e-4


def test_warm_start_from_solution_is_stationary():
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    _, f, ofs = _problem(3, C=2, N=32, na=16)
    C, Nz, na, N = f.shape
    g = RadonGeometry((N, N), np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2,
                      n_detectors=N)
    L = O.normest_power(ofs, N, N, N)
    a = TgvProblem(f, g, (4, 1), [1.0, 0.5], dim=3, L=L)
    a.iterate(1500)
    e_a = a.energy()[2]
    b = TgvProblem(f, g, (4, 1), [1.0, 0.5], dim=3, L=L, u0=a.u)
    b.iterate(50)
    e_b = b.energy()[2]
    cold = TgvProblem(f, g, (4, 1), [1.0, 0.5], dim=3, L=L)
    cold.iterate(50)
    e_c = cold.energy()[2]
    assert e_b < e_c  # the warm start is far ahead of the cold start
    assert rel(b.u.cpu().numpy(), a.u.cpu().numpy()) < 0.05


def test_reference_example_data_parity():
    """Real HAADF + 3 EDX tilt series from the reference's example/ (two
    slices, 36 tilts over 180 degrees, 433 bins), scaled as the reference's
    caller scales them (percentiles 1 -> 0, 98 -> 1, threshold 0.05,
    Reconstruction_coupled.py:349-372), Frobenius 3D TGV-KL with the README's
    parameters, GPU vs the FP32 restatement."""
    import os
    from graptor_amd.radon_tgv_primal_dual_3d import Reconstructionapproach3d
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "example_2slices.npz"))
    sino, angles = z["sino"].astype(np.float64), z["angles"]
    chans = []
    for c in range(sino.shape[0]):
        d = sino[c] - np.percentile(sino[c], 1)
        d = d / np.percentile(d, 98)
        d[d < 0.05] = 0
        chans.append(d.astype(np.float32))          # [na, Nz, nd]
    na, nz, nd = chans[0].shape
    ofs, _, _ = O.radon_struct((nd, nd), angles + np.pi / 2, n_detectors=nd)
    L = O.normest_power(ofs, nd, nd, nd, iters=100, tol=1e-10)
    mu = [0.2, 5e-4, 4e-3, 5e-4]
    sol, sinos = Reconstructionapproach3d(chans, angles + np.pi / 2, (4, 1), mu, 60, 0,
                                          [0, 0.1], 1, "KL", "Frobenius 3D Reconstruction",
                                          "TGV", [0.0, 1.0, [28, 29]], L=L)
    f = np.ascontiguousarray(np.stack(chans).transpose(0, 2, 1, 3))
    f = f * O.feasible_mask(ofs, nd, nd, nd)[None, None]
    r = cref.solve(f.astype(np.float32), ofs, (4, 1), mu, 60, L, "frobenius", "KL", dim=3)
    err = rel(sol, r["u"].transpose(0, 3, 2, 1))
    print(f"example data: rel L2 vs FP32 restatement {err:.2e}", flush=True)
    assert err <= 1e-4


@pytest.mark.parametrize("dim", [2, 3])
def test_fused_primal_matches_unfused(dim):
    """The primal step inside the backprojection kernel gives the same
    iterates as the separate primal kernel."""
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    _, f, ofs = _problem(dim, C=3, N=48, na=20)
    C, Nz, na, N = f.shape
    ang = np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2
    L = O.normest_power(ofs, N, N, N)
    out = []
    for fuse in (False, True):
        g = RadonGeometry((N, N), ang, n_detectors=N)
        g.set_fusion(fuse)
        prob = TgvProblem(f, g, (4, 1), [1.0, 0.3, 0.1], dim=dim, L=L)
        prob.iterate(60, ku_on_last=True)
        torch.cuda.synchronize()
        out.append((prob.u.cpu().numpy(), prob.ku.cpu().numpy()))
    assert rel(out[1][0], out[0][0]) <= 1e-6
    assert rel(out[1][1], out[0][1]) <= 1e-6


@pytest.mark.parametrize("combo", COMBOS[:4] + COMBOS[4:5],
                         ids=["-".join(map(str, c)) for c in COMBOS[:4] + COMBOS[4:5]])
def test_two_stream_iteration_matches_one_stream(combo):
    """Untimed iterations run the dual step and the p half of the primal step
    on a side stream; the profiled form runs everything on the caller's
    stream with the primal step as one kernel.  Both must give the same bits,
    and work queued on the caller's stream right after the call (no host
    synchronisation) must see the side stream's results."""
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    dim, coupling, disc, reg = combo
    print(f"two streams vs one {combo}", flush=True)
    _, f, ofs = _problem(dim, C=3, N=48, na=20)
    C, Nz, na, N = f.shape
    ang = np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2
    L = O.normest_power(ofs, N, N, N)
    g = RadonGeometry((N, N), ang, n_detectors=N)
    kw = dict(dim=dim, coupling=coupling, discrepancy=disc, regularisation=reg, L=L)
    two = TgvProblem(f, g, (4, 1), [1.0, 0.3, 0.1], **kw)
    one = TgvProblem(f, g, (4, 1), [1.0, 0.3, 0.1], **kw)
    for n in (1, 7, 12):
        two.iterate(n)
        # queued on the caller's stream before any host synchronisation
        early = {k: getattr(two, k).sum() for k in ("u", "p", "v", "w")
                 if getattr(two, k) is not None}
        one.profile(n)
        torch.cuda.synchronize()
        for k, t in early.items():
            assert float(t) == float(getattr(two, k).sum()), (n, k)
        for name in ("u", "ubar", "p", "pbar", "v", "w", "lam"):
            a, b = getattr(two, name), getattr(one, name)
            if b is None:
                continue
            assert torch.equal(a, b), (n, name)


def test_two_streams_with_many_slice_groups():
    """2048 slices (512 slice groups of 4, a forward grid of over a thousand
    work-groups beside the dual step): the two-stream iterations equal the
    one-stream order bit for bit."""
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    C, Nz, N, na = 4, 512, 64, 60
    print(f"many slice groups C={C} Nz={Nz} N={N} na={na}", flush=True)
    rng = np.random.default_rng(71)
    f = rng.poisson(20.0, size=(C, Nz, na, N)).astype(np.float32)
    ang = np.deg2rad(np.linspace(-75, 75, na)) + np.pi / 2
    g = RadonGeometry((N, N), ang, n_detectors=N)
    kw = dict(dim=3, coupling="frobenius", discrepancy="KL", regularisation="TGV", L=90.0)
    two = TgvProblem(f, g, (4, 1), [0.2, 5e-4, 4e-3, 5e-4], **kw)
    one = TgvProblem(f, g, (4, 1), [0.2, 5e-4, 4e-3, 5e-4], **kw)
    two.iterate(3)
    one.profile(3)
    torch.cuda.synchronize()
    for name in ("u", "p", "v", "w", "lam"):
        assert torch.equal(getattr(two, name), getattr(one, name)), name


HINT_CASES = [
    # 4096 voxel tiles: the two-stream u half takes the per-voxel kernel
    (3, "frobenius", "KL", "TGV", 4, 256, 64),
    (3, "nuclear", "L2", "TGV", 3, 5, 40),
    # more than 8 channels: the two-sweep dual step
    (2, "nuclear", "KL", "TGV", 10, 1, 48),
    (2, "uncorrelated", "KL", "TV", 3, 1, 40),
]


@pytest.mark.parametrize("case", HINT_CASES, ids=["-".join(map(str, c)) for c in HINT_CASES])
def test_stream_hints_do_not_change_iterates(case, monkeypatch):
    """The stencil kernels with and without the non-temporal load / store
    hints (GRAPTOR_STREAM_HINTS forces the size-based choice) give the same
    bits, on two streams and on one."""
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    dim, coupling, disc, reg, C, Nz, N = case
    na = 20
    print(f"stream hints {case}", flush=True)
    rng = np.random.default_rng(83)
    f = rng.poisson(20.0, size=(C, Nz, na, N)).astype(np.float32)
    ang = np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2
    g = RadonGeometry((N, N), ang, n_detectors=N)
    kw = dict(dim=dim, coupling=coupling, discrepancy=disc, regularisation=reg, L=float(N))
    mu = list(np.geomspace(1.0, 1e-2, C))
    out = {}
    for hint in ("0", "1"):
        monkeypatch.setenv("GRAPTOR_STREAM_HINTS", hint)
        prob = TgvProblem(f, g, (4, 1), mu, **kw)
        prob.iterate(9)
        prob.profile(3)
        torch.cuda.synchronize()
        out[hint] = {k: getattr(prob, k).clone() for k in ("u", "p", "v", "w", "lam")
                     if getattr(prob, k) is not None}
    for k, t in out["0"].items():
        assert torch.equal(t, out["1"][k]), k


def test_two_streams_from_two_host_threads():
    """Each host thread gets its own side stream: two problems iterated from
    two threads on their own torch streams give the same bits as the same
    problems iterated one after the other."""
    import threading
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    _, f, ofs = _problem(3, C=3, N=48, na=20)
    C, Nz, na, N = f.shape
    ang = np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2
    L = O.normest_power(ofs, N, N, N)
    g = RadonGeometry((N, N), ang, n_detectors=N)

    def make(scale):
        return TgvProblem(f * scale, g, (4, 1), [1.0, 0.3, 0.1], dim=3, L=L)

    ref = [make(1.0), make(2.0)]
    for p in ref:
        p.iterate(25)
    torch.cuda.synchronize()
    thr = [make(1.0), make(2.0)]
    torch.cuda.synchronize()  # set-up work of the default stream is done
    errors = []

    def run(p):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(5):
                    p.iterate(5)
            s.synchronize()
        except Exception as e:  # reported below
            errors.append(e)

    ts = [threading.Thread(target=run, args=(p,)) for p in thr]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for a, b in zip(thr, ref):
        for name in ("u", "p", "v", "w", "lam"):
            assert torch.equal(getattr(a, name), getattr(b, name)), name


@pytest.mark.parametrize("combo", COMBOS, ids=["-".join(map(str, c)) for c in COMBOS])
def test_fused_primal_dual_kernel_matches_separate_kernels(combo):
    """The primal step of each iteration fused with the next dual step (two
    alternating state sets) gives the same state as the separate kernels,
    over call sequences with odd and even iteration counts."""
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    dim, coupling, disc, reg = combo
    print(f"fused primal+dual {combo}", flush=True)
    _, f, ofs = _problem(dim, C=3, N=40, na=20)
    C, Nz, na, N = f.shape
    ang = np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2
    L = O.normest_power(ofs, N, N, N)
    g = RadonGeometry((N, N), ang, n_detectors=N)
    kw = dict(dim=dim, coupling=coupling, discrepancy=disc, regularisation=reg, L=L)
    ref = TgvProblem(f, g, (4, 1), [1.0, 0.3, 0.1], fused=False, **kw)
    fus = TgvProblem(f, g, (4, 1), [1.0, 0.3, 0.1], fused=True, **kw)
    assert fus.u2 is not None and ref.u2 is None
    for n in (1, 2, 5, 8, 3):
        ref.iterate(n, ku_on_last=True)
        fus.iterate(n, ku_on_last=True)
        torch.cuda.synchronize()
        for name in ("u", "ubar", "p", "pbar", "v", "w", "lam", "ku"):
            a, b = getattr(fus, name), getattr(ref, name)
            if b is None:
                continue
            err = rel(a.cpu().numpy(), b.cpu().numpy())
            assert err <= 1e-6, (n, name, err)


@pytest.mark.parametrize("coupling,reg", [("frobenius", "TGV"), ("nuclear", "TGV"),
                                          ("uncorrelated", "TV")])
def test_fused_primal_dual_z_runs_and_partial_tiles(coupling, reg):
    """3D grid with partial 64 x 4 tiles and several runs of planes: the fused
    form equals the separate kernels on every state array."""
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    print(f"fused z-runs {coupling} {reg}", flush=True)
    u_true, f, ofs = make_problem(41, 2, 19, 70, 18, poisson=50.0)
    f = f.astype(np.float32)
    N = 70
    ang = np.deg2rad(np.linspace(-70, 70, 18, endpoint=False)) + np.pi / 2
    L = O.normest_power(ofs, N, N, N)
    g = RadonGeometry((N, N), ang, n_detectors=N)
    kw = dict(dim=3, coupling=coupling, discrepancy="KL", regularisation=reg, L=L,
              stepsize=2.0)
    ref = TgvProblem(f, g, (4, 1), [1.0, 0.3], fused=False, **kw)
    fus = TgvProblem(f, g, (4, 1), [1.0, 0.3], fused=True, **kw)
    for n in (3, 4):
        ref.iterate(n)
        fus.iterate(n)
        torch.cuda.synchronize()
        for name in ("u", "ubar", "p", "pbar", "v", "w", "lam"):
            a, b = getattr(fus, name), getattr(ref, name)
            if b is None:
                continue
            err = rel(a.cpu().numpy(), b.cpu().numpy())
            assert err <= 1e-6, (n, name, err)


@pytest.mark.parametrize("dim,C", [(3, 6), (2, 8)])
def test_many_channel_parity(dim, C):
    """Channel counts above 4 (the 8-channel register layout of the dual step)."""
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    print(f"many channels dim={dim} C={C}", flush=True)
    Nz = 1 if dim == 2 else 3
    _, f, ofs = make_problem(53, C, Nz, 32, 16, poisson=50.0)
    f = f.astype(np.float32)
    N = 32
    ang = np.deg2rad(np.linspace(-70, 70, 16, endpoint=False)) + np.pi / 2
    L = O.normest_power(ofs, N, N, N)
    g = RadonGeometry((N, N), ang, n_detectors=N)
    mu = list(np.linspace(1.0, 0.1, C))
    prob = TgvProblem(f, g, (4, 1), mu, dim=dim, coupling="frobenius", discrepancy="KL", L=L)
    prob.iterate(40)
    torch.cuda.synchronize()
    r32 = cref.solve(f, ofs, alpha=(4, 1), mu=mu, maxiter=40, L=L, coupling="frobenius",
                     discrepancy="KL", dim=dim)
    e = rel(prob.u.cpu().numpy(), r32["u"])
    print(f"  rel L2 vs FP32 restatement {e:.2e}", flush=True)
    assert e <= 1e-4


def test_ellipse_2d_l2_configuration_parity():
    """The 2D single-channel TGV-L2 configuration (128x128 ellipse phantom in a
    183 grid, 90 angles in [-70, 70), 1% Gaussian noise, alpha (4, 1), mu 1,
    500 iterations) through the drop-in entry, against the FP32 restatement
    and the FP64 oracle with the same L."""
    from graptor_amd.phantoms import ellipse_problem_2d
    from graptor_amd.radon_tgv_primal_dual_2d import Reconstructionapproach2d
    print("ellipse 2D L2 500 iterations", flush=True)
    f, angles, ph, geom = ellipse_problem_2d()
    fh = f.cpu().numpy()                               # (1, 1, 90, 183)
    ofs = geom.ofs.T.copy()
    fh = fh * O.feasible_mask(ofs, 183, 183, 183)[None, None]
    L = O.normest_power(ofs, 183, 183, 183)
    sol, sinos = Reconstructionapproach2d([fh[0, 0]], angles, (4.0, 1.0), [1.0], 500, None,
                                          [0, 0], "L2", "Uncorrelated 2D Reconstruction", "TGV",
                                          [0.0, 1.0, 0], L=L)
    u = sol[0].T                                       # [x, y] -> [y, x]
    r32 = cref.solve(fh.astype(np.float32), ofs, alpha=(4.0, 1.0), mu=[1.0], maxiter=500, L=L,
                     coupling="uncorrelated", discrepancy="L2", dim=2)
    e32 = rel(u, r32["u"][0, 0])
    # the FP64 oracle (NumPy) on the first 20 iterations
    s20, _ = Reconstructionapproach2d([fh[0, 0]], angles, (4.0, 1.0), [1.0], 20, None, [0, 0],
                                      "L2", "Uncorrelated 2D Reconstruction", "TGV",
                                      [0.0, 1.0, 0], L=L)
    r64 = O.solve(fh.astype(np.float64), ofs, (4.0, 1.0), [1.0], 20, L, "uncorrelated", "L2",
                  dim=2)
    e64 = rel(s20[0].T, r64["u"][0, 0])
    print(f"  rel L2 vs FP32 restatement (500 it) {e32:.2e}, vs FP64 oracle (20 it) {e64:.2e}",
          flush=True)
    assert e32 <= 1e-4 and e64 <= 1e-4
