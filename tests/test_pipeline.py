"""Caller counterpart (Reconstruction_coupled.py / h5utilities.py steps):
pre-processing, slab scheduling, rescaling and output, on the CPU; the
end-to-end 2D/3D pipelines on the GPU."""
import os

import numpy as np
import pytest

from graptor_amd import io as gio
from graptor_amd import pipeline as pl

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("n,possible,overlay", [(10, 4, 1), (10, 10, 2), (64, 16, 3), (7, 3, 0),
                                                (33, 8, 5), (1, 1, 0), (100, 37, 4)])
def test_slab_schedule_partitions_the_slices(n, possible, overlay):
    slabs = pl.slab_schedule(n, possible, overlay)
    stored = []
    ov = overlay if overlay <= possible / 4.0 else int(possible / 4)
    for k, (a, b, s, e) in enumerate(slabs):
        assert 0 <= a <= b < n and b - a + 1 <= possible
        assert s == (0 if k == 0 else ov)
        if k == len(slabs) - 1:
            assert e == 0 and b == n - 1
        else:
            assert e == ov
        stored.extend(range(a + s, b + 1 - e))
    assert stored == list(range(n))  # every slice stored exactly once, in order


def test_slab_schedule_example():
    # 10 slices, 4 per slab, overlap 1: slabs start every 4 - 2 = 2 slices
    assert pl.slab_schedule(10, 4, 1) == [(0, 3, 0, 1), (2, 5, 1, 1), (4, 7, 1, 1), (6, 9, 1, 0)]


def test_cut_edges_and_background():
    data = np.zeros((2, 3, 20))
    data[:, :, :4] = 5.0      # left band: cut = ceil(20 * 0.2) = 4, outer half [2, 4)
    data[:, :, 16:] = 3.0     # right band, outer half [18, 20)
    data[:, :, 4:16] = 1.0
    out, bg = pl.cut_edges(data, 0.2)
    assert out.shape == (2, 3, 12) and (out == 1.0).all()
    assert bg == pytest.approx(0.5 * (5.0 + 3.0))
    out, bg = pl.cut_edges(data, 0.04)  # cut = 1: no background estimate
    assert out.shape == (2, 3, 18) and bg == 0.0


def test_brightness_rescale_equalises_projection_means():
    rng = np.random.default_rng(0)
    data = rng.uniform(0.5, 1.0, (5, 4, 6)) * np.arange(1, 6)[:, None, None]
    out = pl.brightness_rescale(data)
    means = out.reshape(5, -1).mean(axis=1)
    assert np.allclose(means, data.reshape(5, -1).mean(axis=1).mean())
    # each projection only scaled
    ratios = out / data
    assert np.allclose(ratios, ratios[:, :1, :1])


def test_remove_projections_is_one_based():
    data = np.arange(5)[:, None, None] * np.ones((5, 2, 2))
    d, a = pl.remove_projections(data, np.arange(5.0), [1, 4])
    assert list(a) == [1.0, 2.0, 4.0] and list(d[:, 0, 0]) == [1.0, 2.0, 4.0]


def test_scale_data_ratio_order_2d_vs_3d():
    rng = np.random.default_rng(1)
    data = rng.uniform(0.0, 10.0, (6, 5, 8)) + 2.0
    section = [1, 3]
    lo, hi_shifted = np.percentile(data, 1), None
    d3, r3 = pl.scale_data(data, section, (1, 98), dim=3)
    d2, r2 = pl.scale_data(data, section, (1, 98), dim=2)
    assert r3 == pytest.approx(np.sum(data[:, section]) / 6)
    assert r2 == pytest.approx(np.sum(data[:, section] - lo) / 6)
    shifted = data - lo
    hi_shifted = np.percentile(shifted, 98)
    assert np.allclose(d3, shifted / hi_shifted) and np.allclose(d2, d3)


def test_baselines():
    s = np.array([[0.1, 0.5, 2.0], [0.05, 1.0, 3.0]])
    assert np.allclose(pl.apply_baseline(s, "thresholding", 0.6, 0.0, 3), np.where(s < 0.6, 0, s))
    assert np.allclose(pl.apply_baseline(s, "mean1", 0.0, 0.2, 3), s - 0.2)
    assert np.allclose(pl.apply_baseline(s, "mean", 0.0, 0.2, 2), s - 0.2)
    m = np.mean(s[s < 0.6])
    assert np.allclose(pl.apply_baseline(s, "mean2", 0.6, 0.0, 3), s - m)
    assert np.allclose(pl.apply_baseline(s, "meanthres", 0.6, 0.0, 2), s - m)
    assert np.allclose(pl.apply_baseline(s, "", 0.6, 0.0, 2), s)


@pytest.mark.parametrize("dim", [2, 3])
def test_mean_threshold_bright_equalises_projection_means(dim):
    """Threshold, then every projection scaled to the mean of the projection
    means (H5:256-269): checked by its defining properties, not its formula —
    values below the threshold become 0, each projection is a positive
    multiple of its thresholded self, and all projection means are equal to
    the mean of the thresholded projection means."""
    rng = np.random.default_rng(11)
    shape = (7, 5) if dim == 2 else (7, 3, 5)
    s = rng.uniform(0.0, 2.0, shape) * rng.uniform(0.2, 3.0, (7,) + (1,) * (dim - 1))
    out = pl.apply_baseline(s, "mean_threshold_bright", 0.4, 0.0, dim)
    t = np.where(s < 0.4, 0.0, s)
    assert np.array_equal(out == 0, t == 0)
    ratio = out.reshape(7, -1)[t.reshape(7, -1) > 0] / t.reshape(7, -1)[t.reshape(7, -1) > 0]
    per = [np.unique(np.round((out[i] / np.where(t[i] > 0, t[i], 1))[t[i] > 0], 12))
           for i in range(7)]
    assert all(len(p) == 1 and p[0] > 0 for p in per) and ratio.min() > 0
    target = np.mean([t[i].mean() for i in range(7)])
    assert np.allclose([out[i].mean() for i in range(7)], target, rtol=1e-12)


def test_faktor_rescale_restores_mass_ratio():
    rng = np.random.default_rng(2)
    sol = rng.uniform(size=(2, 3, 4, 4)).astype(np.float32)
    sino = rng.uniform(size=(2, 5, 3, 4)).astype(np.float32)
    ratio = [7.0, 11.0]
    s2, k2 = pl.faktor_rescale(sol, sino, ratio, na=5)
    for c in range(2):
        # after rescaling, sum(sinogram_c) / na equals the data's ratio
        assert np.sum(k2[c], dtype=np.float64) / 5 == pytest.approx(ratio[c], rel=1e-5)
        assert np.allclose(s2[c] / sol[c], k2[c].sum() / sino[c].sum(), rtol=1e-5)


def test_write_results_layout(tmp_path):
    sol = np.arange(2 * 3 * 4 * 4, dtype=np.float32).reshape(2, 3, 4, 4) + 1
    sino = np.ones((2, 5, 3, 4), np.float32)
    base = str(tmp_path / "rec")
    pl.write_results(base, ["A", "B"], sol, sino)
    assert np.array_equal(gio.read_mrc(base + "_A.mrc"), sol[0])
    assert np.array_equal(gio.read_mrc(base + "_B_sinogram.mrc"), sino[1])
    coll = gio.read_mrc(base + "_all_channels.mrc")
    assert coll.shape == (3, 4, 8) and np.array_equal(coll[:, :, 4:], sol[1])
    norm = gio.read_mrc(base + "_all_channels_normalized.mrc")
    assert np.allclose(norm[:, :, :4], sol[0] / sol[0].max())


def _fixture_files(tmp_path):
    """The example fixture's four channels as MRC tilt series + .rawtlt,
    and the operator norm of its geometry (power iteration)."""
    from graptor_amd.projector import RadonGeometry, radon_normest
    d = np.load(os.path.join(HERE, "golden", "example_2slices.npz"))
    names = []
    for c, ch in enumerate(d["channels"]):
        p = str(tmp_path / f"{ch}.mrc")
        gio.write_mrc(p, d["sino"][c])  # (tilts, slices, bins)
        with open(p[:-4] + ".rawtlt", "w") as fh:
            fh.write("\n".join(f"{a:.6f}" for a in np.rad2deg(d["angles"])))
        names.append(p)
    angles = gio.read_angles(names[0])  # as the pipeline reads them back
    nd = d["sino"].shape[-1]
    L = radon_normest(None, RadonGeometry((nd, nd), angles + np.pi / 2, n_detectors=nd),
                      mode="power")
    return names, list(d["channels"]), L


@pytest.mark.gpu
def test_pipeline_2d_batched_equals_per_slice_calls(tmp_path):
    names, chans, L = _fixture_files(tmp_path)
    dh = ["", "thresholding", 0.0, "", 0.05, []]
    kw = dict(mu=[0.2, 5e-4, 4e-3, 5e-4], alpha=(4.0, 1.0), maxiter=60,
              method="Frobenius 2D Reconstruction", slice_levels=(0, 1, 1), datahandling=dh,
              channelnames=chans, write=False, L=L)
    s_b, k_b = pl.coupled_reconstruction_2d(names, str(tmp_path / "b"), batch=True, **kw)
    s_s, k_s = pl.coupled_reconstruction_2d(names, str(tmp_path / "s"), batch=False, **kw)
    assert np.isfinite(s_b).all() and s_b.shape == s_s.shape
    assert np.linalg.norm(s_b - s_s) / np.linalg.norm(s_s) < 1e-6
    assert np.linalg.norm(k_b - k_s) / np.linalg.norm(k_s) < 1e-6


@pytest.mark.gpu
def test_pipeline_3d_slabs_cover_the_volume(tmp_path):
    names, chans, L = _fixture_files(tmp_path)
    dh = ["", "thresholding", 0.0, "", 0.05, []]
    kw = dict(mu=[0.2, 5e-4, 4e-3, 5e-4], alpha=(4.0, 1.0), maxiter=40,
              method="Frobenius 3D Reconstruction", sinogram_level=(0, 1, 1), datahandling=dh,
              channelnames=chans, L=L)
    U1, sol1, k1 = pl.coupled_reconstruction_3d(names, str(tmp_path / "one"), chunksize=0,
                                                overlay=0, **kw)
    # one slab per slice: each slice solved as its own 1-slice volume
    U2, sol2, k2 = pl.coupled_reconstruction_3d(names, str(tmp_path / "two"), chunksize=1,
                                                overlay=0, **kw)
    assert U1.shape == U2.shape == (4, 433, 433, 2)
    assert np.isfinite(U2).all() and (np.abs(U2).sum(axis=(0, 1, 2)) > 0).all()
    # the data the pipeline hands the solver, rebuilt from its own steps
    from oracle import cref
    from oracle import tgv_oracle as O
    section = [0, 1]
    coll, ratio, noise, angles, baseline, threshold = pl._prepare(names, section, dh, (1, 98), 3)
    f = np.stack([pl.apply_baseline(d, baseline, threshold, noise[i], 3)[:, section, :]
                  for i, d in enumerate(coll)]).transpose(0, 2, 1, 3)   # (C, Nz, na, nd)
    nd = f.shape[-1]
    ofs, _, _ = O.radon_struct((nd, nd), angles + np.pi / 2, n_detectors=nd)
    f = np.ascontiguousarray(f * O.feasible_mask(ofs, nd, nd, nd)[None, None], np.float32)
    rk = dict(alpha=(4.0, 1.0), mu=kw["mu"], maxiter=40, L=L, coupling="frobenius",
              discrepancy="KL", dim=3)
    # one slab holding both slices: the whole volume
    whole = cref.solve(f, ofs, **rk)["u"].transpose(0, 3, 2, 1)          # (C, x, y, z)
    e1 = np.linalg.norm(U1 - whole) / np.linalg.norm(whole)
    # one slab per slice: each slice's interior is that slice solved alone
    e2 = []
    for z in section:
        one = cref.solve(np.ascontiguousarray(f[:, z:z + 1]), ofs, **rk)["u"][:, 0]
        ref = one.transpose(0, 2, 1)
        e2.append(np.linalg.norm(U2[..., z] - ref) / np.linalg.norm(ref))
    print(f"3D pipeline vs FP32 restatement: whole volume {e1:.1e}, slabs {e2}")
    assert e1 <= 1e-4 and max(e2) <= 1e-4
    # and the two differ (the z coupling of TGV is lost between slabs)
    assert np.linalg.norm(U1 - U2) / np.linalg.norm(U1) > 1e-3
    # files written for every channel
    for ch in chans:
        assert os.path.isfile(str(tmp_path / f"one_{ch}.mrc"))
    assert gio.read_mrc(str(tmp_path / "one_all_channels.mrc")).shape == (2, 433, 4 * 433)


class _FakeSolver:
    """Stands in for Reconstructionapproach3d: each slice's solution carries
    its own index, and slabs above ``limit`` slices fail with ``exc``."""

    def __init__(self, C, nd, na, limit, exc):
        self.C, self.nd, self.na, self.limit, self.exc = C, nd, na, limit, exc
        self.calls = []

    def __call__(self, start, end, ov_s, ov_e):
        n = end + 1 - start
        self.calls.append((start, end, ov_s, ov_e))
        if n > self.limit:
            raise self.exc("fake device failure")
        z = np.arange(start, end + 1, dtype=np.float64)
        c = 1000.0 * np.arange(self.C)[:, None]
        u = np.broadcast_to((z[None, :] + c)[:, None, None, :], (self.C, self.nd, self.nd, n))
        ku = np.broadcast_to((z[None, :] + c)[:, None, :, None], (self.C, self.na, n, self.nd))
        return u.copy(), ku.copy()


@pytest.mark.parametrize("exc_name", ["GraptorMemoryError", "GraptorError", "OutOfMemoryError"])
def test_slab_loop_shrinks_and_stores_every_slice(exc_name):
    """RC:510-536: a device out-of-memory error and a generic device error
    both shrink the slab to int(0.7 * possible) and restart the schedule;
    every slice is stored once, from a slab that holds it as interior."""
    import io

    import torch

    from graptor_amd import _lib
    exc = torch.cuda.OutOfMemoryError if exc_name == "OutOfMemoryError" else getattr(_lib,
                                                                                        exc_name)
    C, nd, na, nz = 2, 3, 4, 23
    fake = _FakeSolver(C, nd, na, limit=5, exc=exc)
    log = io.StringIO()
    U, S, possible = pl.solve_slabs(fake, nz, 10, 1, C, nd, na, log=log)
    assert possible == 4  # 10 -> 7 -> 4 (int(0.7 * 7)); 4 <= 5 succeeds
    want = np.arange(nz)[None, :] + 1000.0 * np.arange(C)[:, None]
    assert np.array_equal(U, np.broadcast_to(want[:, None, None, :], U.shape))
    assert np.array_equal(S, np.broadcast_to(want[:, None, :, None], S.shape))
    # the failing attempts stop at their first slab; the last attempt runs the
    # schedule of 4 slices with overlap 1
    sizes = [e - s + 1 for s, e, _, _ in fake.calls]
    assert sizes[:2] == [10, 7]
    assert fake.calls[2:] == pl.slab_schedule(nz, 4, 1)
    text = log.getvalue()
    assert text.count("trying to") == 2
    if exc_name == "GraptorError":
        assert "unexpected device error" in text and "An exception of type GraptorError" in text
    else:
        assert "requires too much memory" in text


@pytest.mark.parametrize("exc_name,raised,msg", [
    ("GraptorMemoryError", MemoryError, "single slice requires too much memory"),
    ("GraptorError", "GraptorError", "single slice creates an unknown error"),
])
def test_slab_loop_gives_up_below_one_slice(exc_name, raised, msg):
    import io

    from graptor_amd import _lib
    exc = getattr(_lib, exc_name)
    raised = getattr(_lib, raised) if isinstance(raised, str) else raised
    fake = _FakeSolver(1, 2, 2, limit=0, exc=exc)
    with pytest.raises(raised, match=msg):
        pl.solve_slabs(fake, 6, 3, 0, 1, 2, 2, log=io.StringIO())
    # 3 -> 2 -> 1 -> 0: three attempts, each failing on its first slab
    assert [e - s + 1 for s, e, _, _ in fake.calls] == [3, 2, 1]
