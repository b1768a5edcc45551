"""GPU tests of the batched solver: a plan over problems of uneven size,
angle and channel counts must reproduce each problem solved on its own, its
captured hipGraph must reproduce the eager launches on every replay, and the
autotuner must leave results unchanged."""
import numpy as np
import pytest

from oracle import cref
from oracle import tgv_oracle as O
from tests.helpers import angles_deg, make_problem, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SPECS = [
    # (dim, N, na, C, Nz, coupling, discrepancy)
    (2, 40, 30, 1, 1, "uncorrelated", "KL"),
    (2, 56, 45, 3, 1, "frobenius", "KL"),
    (2, 72, 33, 2, 1, "frobenius", "L2"),
    (2, 48, 60, 4, 1, "nuclear", "KL"),
    (3, 36, 20, 2, 5, "frobenius", "KL"),
    (2, 44, 28, 6, 1, "frobenius", "KL"),       # 5..8 channels: the 8-channel kernels
    (3, 32, 18, 8, 3, "uncorrelated", "L2"),
    (2, 36, 24, 10, 1, "nuclear", "KL"),        # more than 8 channels: the two-sweep kernels
    (3, 28, 16, 9, 2, "frobenius", "KL"),
]
MU = [1.0, 0.3, 0.2, 0.1, 0.5, 0.4, 0.25, 0.15, 0.35, 0.45]


def _make(seed, spec):
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    dim, N, na, C, Nz, cpl, disc = spec
    _, f, ofs = make_problem(seed, C, Nz, N, na, poisson=20.0)
    g = RadonGeometry((N, N), angles_deg(-70, 70, na), n_detectors=N)
    L = O.normest_power(ofs, N, N, N, iters=60, tol=1e-9)
    mu = MU[:C]
    prob = TgvProblem(f.astype(np.float32), g, (4.0, 1.0), mu, dim=dim, coupling=cpl,
                      discrepancy=disc, L=L)
    return prob, f.astype(np.float32), ofs, L, mu


def _reference(spec, f, ofs, L, mu, iters):
    """The FP32 restatement of the reference kernels; for 2D nuclear coupling
    the FP64 evaluation of the reference's route (its FP32 closed form is off
    its own FP64 evaluation by up to 1e-2, see test_gpu_solver)."""
    dim, N, na, C, Nz, cpl, disc = spec
    if cpl == "nuclear" and dim == 2:
        r = O.solve(f.astype(np.float64), ofs, (4.0, 1.0), mu, iters, L, cpl, disc, dim=dim,
                    nuclear_mode="ref2d")
        r["what"] = "FP64 oracle"
        return r
    r = cref.solve(f, ofs, (4.0, 1.0), mu, iters, L, cpl, disc, dim=dim)
    r["what"] = "FP32 restatement"
    return r


def test_batch_matches_individual_solves():
    from graptor_amd.batch import BatchPlan
    batch = [_make(500 + i, s) for i, s in enumerate(SPECS)]
    single = [_make(500 + i, s) for i, s in enumerate(SPECS)]
    plan = BatchPlan([b[0] for b in batch])
    plan.iterate(40, ku_on_last=True)
    for s in single:
        s[0].iterate(40, ku_on_last=True)
    torch.cuda.synchronize()
    for i, (b, s) in enumerate(zip(batch, single)):
        eu = rel(b[0].u.cpu().numpy(), s[0].u.cpu().numpy())
        ek = rel(b[0].ku.cpu().numpy(), s[0].ku.cpu().numpy())
        print(f"problem {i} {SPECS[i]}: batch vs single u {eu:.1e} ku {ek:.1e}", flush=True)
        assert eu <= 1e-6 and ek <= 1e-6
    # and every one of them against the reference's iteration
    for i, (prob, f, ofs, L, mu) in enumerate(batch):
        r = _reference(SPECS[i], f, ofs, L, mu, 40)
        eu = rel(prob.u.cpu().numpy(), r["u"])
        ek = rel(prob.ku.cpu().numpy(), r["ku"])
        print(f"problem {i} {SPECS[i]}: batch vs {r['what']} u {eu:.1e} ku {ek:.1e}",
              flush=True)
        assert eu <= 1e-4 and ek <= 1e-4


def test_fused_form_refuses_more_than_eight_channels():
    """The fused primal + next-dual kernel holds at most 8 channels of a
    voxel; every other path (the batches above included) takes any count."""
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    _, f, ofs = make_problem(481, 9, 1, 32, 16, poisson=20.0)
    g = RadonGeometry((32, 32), angles_deg(-70, 70, 16), n_detectors=32)
    with pytest.raises(ValueError, match="at most 8"):
        TgvProblem(f.astype(np.float32), g, (4.0, 1.0), MU[:9], dim=2, L=10.0, fused=True)


def test_graph_replay_matches_eager():
    from graptor_amd.batch import BatchPlan
    eager = [_make(700 + i, s)[0] for i, s in enumerate(SPECS)]
    graph = [_make(700 + i, s)[0] for i, s in enumerate(SPECS)]
    pe = BatchPlan(eager)
    pg = BatchPlan(graph)
    for rep in range(3):  # several replays with eager host/device work between
        pe.iterate(12)
        pg.iterate_graph(12, block=5)  # 2 graph replays + 2 eager iterations
        torch.cuda.synchronize()
        _ = [p.u.sum().item() for p in graph]
        for i, (a, b) in enumerate(zip(eager, graph)):
            assert rel(b.u.cpu().numpy(), a.u.cpu().numpy()) <= 1e-6, (rep, i)
            assert rel(b.lam.cpu().numpy(), a.lam.cpu().numpy()) <= 1e-6, (rep, i)
    # the replayed iterations against the FP32 restatement, every problem
    for i, s in enumerate(SPECS):
        _, f, ofs, L, mu = _make(700 + i, s)
        r = _reference(s, f, ofs, L, mu, 36)
        err = rel(graph[i].u.cpu().numpy(), r["u"])
        print(f"problem {i} {s}: graph replay vs {r['what']} u {err:.1e}", flush=True)
        assert err <= 1e-4, i


def test_batched_2d_shape_against_restatement():
    """A seeded subset of the 64-problem batch (N in 96..384, 30..180 angles
    over [-70, 70) degrees, 1..4 channels, Poisson data), with tiles from the
    autotuner, advanced through the captured graph, against the FP32
    restatement problem by problem."""
    from graptor_amd.autotune import autotune
    from graptor_amd.batch import BatchPlan
    from graptor_amd.phantoms import batched_2d_problem_set
    from graptor_amd.solver import TgvProblem
    dev = torch.device("cuda", 0)
    specs = batched_2d_problem_set(n=8, device=dev)
    probs = []
    for sp in specs:
        cpl = "frobenius" if sp["C"] > 1 else "uncorrelated"
        mu = ([0.02] + [0.5] * 3)[:sp["C"]]
        probs.append(TgvProblem(sp["f"], sp["geometry"], (4.0, 1.0), mu, dim=2, coupling=cpl,
                                discrepancy="KL", normest_mode="power", device=dev, fused=False))
    plan = BatchPlan(probs)
    tune = autotune(probs, iters=2, plan=plan)
    print(f"C3 subset: {[(sp['N'], sp['na'], sp['C']) for sp in specs]} tiles {tune['fwd_band']}"
          f"/{tune['adj_tile_y']}", flush=True)
    iters = 30
    plan.iterate_graph(iters, block=10)
    torch.cuda.synchronize()
    for i, (sp, q) in enumerate(zip(specs, probs)):
        ofs, _, _ = O.radon_struct((q.nx, q.ny), q.geometry.angles, n_detectors=q.nd)
        r = cref.solve(q.f.cpu().numpy(), ofs, (4.0, 1.0), q.mu, iters, q.k.L, q.coupling, "KL",
                       dim=2)
        err = rel(q.u.cpu().numpy(), r["u"])
        print(f"problem {i} N={sp['N']} na={sp['na']} C={sp['C']}: rel L2 {err:.1e}", flush=True)
        assert err <= 1e-4, i


def test_autotune_keeps_results():
    from graptor_amd.autotune import autotune
    from graptor_amd.batch import BatchPlan
    probs = [_make(900 + i, s)[0] for i, s in enumerate(SPECS[:3])]
    ref = [_make(900 + i, s)[0] for i, s in enumerate(SPECS[:3])]
    plan = BatchPlan(probs)
    res = autotune(probs, iters=2, plan=plan)
    print(f"autotune: {res}", flush=True)
    assert res["fwd_band"] in (4, 8, 12, 16, 24, 32) and res["adj_tile_y"] in (8, 16, 32)
    # tuning ran on scratch copies: the real state is untouched
    assert all(float(p.u.abs().sum()) == 0.0 for p in probs)
    plan.iterate(20)
    for p in ref:
        p.iterate(20)
    torch.cuda.synchronize()
    for a, b in zip(probs, ref):
        assert rel(a.u.cpu().numpy(), b.u.cpu().numpy()) <= 1e-5


def test_batch_mixes_tv_and_fused_primal():
    """A plan whose problems take different primal-step routes: TV (no p
    half), the primal step fused into the backprojection (no primal launch),
    2D and 3D TGV, large enough in total (over 2^20 voxel-channels) that the
    p half of the primal step runs beside the backprojection and the dual
    step beside the forward projection; eager and graph-replayed iterations
    must reproduce each problem solved on its own."""
    from graptor_amd.batch import BatchPlan
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    specs = [
        # (dim, N, na, C, Nz, coupling, discrepancy, regularisation, fused primal)
        (2, 40, 26, 2, 1, "frobenius", "KL", "TV", False),
        (2, 48, 30, 3, 1, "frobenius", "KL", "TGV", True),
        (3, 32, 18, 2, 4, "nuclear", "L2", "TGV", False),
        (2, 36, 22, 1, 1, "uncorrelated", "L2", "TGV", False),
        (3, 28, 16, 3, 3, "uncorrelated", "KL", "TV", True),
        # 2^20 voxel-channels: the plan's total reaches the size from which it
        # overlaps projections and stencil steps
        (3, 128, 24, 2, 32, "frobenius", "KL", "TGV", False),
    ]

    def make(seed, sp):
        dim, N, na, C, Nz, cpl, disc, reg, fuse = sp
        _, f, ofs = make_problem(seed, C, Nz, N, na, poisson=20.0)
        g = RadonGeometry((N, N), angles_deg(-70, 70, na), n_detectors=N)
        g.set_fusion(fuse)
        return TgvProblem(f.astype(np.float32), g, (4.0, 1.0), MU[:C], dim=dim, coupling=cpl,
                          discrepancy=disc, regularisation=reg, L=float(N))

    eager = [make(1100 + i, s) for i, s in enumerate(specs)]
    graph = [make(1100 + i, s) for i, s in enumerate(specs)]
    single = [make(1100 + i, s) for i, s in enumerate(specs)]
    pe, pg = BatchPlan(eager), BatchPlan(graph)
    for rep in range(2):
        pe.iterate(11)
        pg.iterate_graph(11, block=4)
        for s in single:
            s.iterate(11)
        torch.cuda.synchronize()
        for i, (a, b, s) in enumerate(zip(eager, graph, single)):
            for name in ("u", "p", "v", "w", "lam"):
                ref = getattr(s, name)
                if ref is None:
                    continue
                ea = rel(getattr(a, name).cpu().numpy(), ref.cpu().numpy())
                eb = rel(getattr(b, name).cpu().numpy(), ref.cpu().numpy())
                print(f"rep {rep} problem {i} {specs[i]} {name}: eager {ea:.1e} graph {eb:.1e}",
                      flush=True)
                assert ea <= 1e-6 and eb <= 1e-6, (rep, i, name)
