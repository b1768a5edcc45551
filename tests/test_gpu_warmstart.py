"""The learned warm start end to end on small slices: solver-generated
training pairs, training on the MFMA kernels, and the gain over plain
primal-dual on unseen slices (the plain iteration still runs in both arms)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_warm_start_gains_over_plain_primal_dual():
    from graptor_amd import warmstart as ws
    cfg = ws.SliceConfig(N=64, na=30)
    r = ws.study(cfg, n_train=512, n_eval=32, target_iters=1000, ustar_iters=5000,
                 train_steps=600, budget_iters=500, max_iters=2000, device="cuda")
    # training converged to something useful
    assert r["losses"][-1][1] < 0.3 * r["losses"][0][1]
    plain0, warm0 = r["error_median"][0]
    assert plain0 == pytest.approx(1.0)
    assert warm0 < 0.4, warm0
    # u* is a converged reference at the tolerances compared below
    assert r["ustar_last1000_rel_change_max"] < 2e-3
    row = r["targets"]["plain_at_budget"]
    assert row["plain_reached"] == row["warm_reached"] == 32
    # the warm start reaches plain's 500-iteration error in well under 500 iterations
    assert row["iter_ratio_median"] >= 1.5, row
    assert row["warm_iters_median"] < row["plain_iters_median"]
    # both arms reach 1e-2; the warm start needs fewer iterations (median)
    row = r["targets"]["0.01"]
    assert row["plain_reached"] == row["warm_reached"] == 32
    assert row["warm_iters_median"] < row["plain_iters_median"], row


def test_flip_augmentation_is_a_symmetry_of_the_data():
    """Mirroring a slice in x or y mirrors the backprojection features for a
    tilt series symmetric about 0 degrees (the premise of the training
    augmentation)."""
    from graptor_amd import warmstart as ws
    from graptor_amd.phantoms import multichannel_slices
    from graptor_amd.projector import radon
    cfg = ws.SliceConfig(N=64, na=30)
    geom = cfg.geometry()
    img = multichannel_slices([7, 8], 64, 4, device="cuda")
    for dims in ([2], [3], [2, 3]):
        fl = img.flip(dims).contiguous()
        s1 = torch.zeros((8, cfg.na, cfg.N), device="cuda")
        s2 = torch.zeros_like(s1)
        radon(s1, img.view(8, 64, 64), geom)
        radon(s2, fl.view(8, 64, 64), geom)
        x1, _ = ws.features(s1.view(2, 4, cfg.na, cfg.N), geom)
        x2, _ = ws.features(s2.view(2, 4, cfg.na, cfg.N), geom)
        # features are NHWC: image dims 2, 3 of (B, C, H, W) are 1, 2 here
        x1f = x1.flip([d - 1 for d in dims])
        err = (x1f - x2).norm() / x2.norm()
        assert err < 1e-5, (dims, float(err))


def test_step_ratio_runs_share_the_fixed_point():
    """SliceBatch(step_ratio=s) keeps sigma * tau (so the convergence
    condition) and moves toward the same u as the reference steps: after many
    iterations the runs with s = 1 and s = 2 are far closer to each other than
    either is to its start or to its 200-iteration iterate."""
    from graptor_amd import warmstart as ws
    cfg = ws.SliceConfig(N=32, na=20)
    geom = cfg.geometry()
    f = ws.make_data(cfg, geom, [60000, 60001], noise_seed=7, device="cuda")
    L = 40.0
    one = ws.SliceBatch(cfg, geom, f, L)
    two = ws.SliceBatch(cfg, geom, f, L, step_ratio=2.0)
    for a, b in zip(one.problems, two.problems):
        assert b.params.sigma == pytest.approx(2 * a.params.sigma, rel=1e-6)
        assert b.params.tau == pytest.approx(a.params.tau / 2, rel=1e-6)
        assert b.params.sigma * b.params.tau == pytest.approx(a.params.sigma * a.params.tau,
                                                             rel=1e-6)
        assert b.params.sigma_p == pytest.approx(b.params.sigma / L, rel=1e-6)
    one.iterate(200)
    two.iterate(200)
    early = float(ws.rel_errors(one.u(), two.u()).max())
    one.iterate(20000)
    two.iterate(20000)
    late = float(ws.rel_errors(one.u(), two.u()).max())
    print(f"distance s=1 vs s=2: {early:.2e} after 200, {late:.2e} after 20200", flush=True)
    assert torch.isfinite(one.u()).all() and torch.isfinite(two.u()).all()
    assert early > 1e-3 and late < 0.2 * early
    with pytest.raises(ValueError):
        ws.SliceBatch(ws.SliceConfig(N=32, na=20, discrepancy="L2"), geom, f, L, step_ratio=2.0)
