"""GPU parity of the projector pair, the operator-norm estimate and the
feasibility mask against the FP64 oracle and the FP32 restatement of the
reference kernels (bar: relative L2 <= 1e-5 in FP32)."""
import numpy as np
import pytest

from oracle import cref
from oracle import tgv_oracle as O
from tests.helpers import angles_deg, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [
    # (nx, ny, nd, angles, nslices)
    (64, 64, 64, angles_deg(-70, 70, 20), 1),
    (64, 64, 64, angles_deg(-75, 75, 30), 2),
    (96, 96, 96, angles_deg(-75, 75, 45), 5),
    (70, 50, 90, np.linspace(0.0, np.pi, 33, endpoint=False), 3),
    (183, 183, 183, angles_deg(-70, 70, 90), 1),
    (40, 40, 57, np.array([0.0, np.pi / 4, np.pi / 2, 3 * np.pi / 4, np.pi, -0.3, 2.0]), 4),
    (256, 256, 256, angles_deg(-75, 75, 60), 8),
    (70, 50, 90, np.linspace(0.0, np.pi, 33, endpoint=False), 11),   # 8 + 3 slices
    (96, 96, 96, angles_deg(-75, 75, 45), 16),
]


def _geom(nx, ny, nd, ang):
    from graptor_amd.projector import RadonGeometry
    return RadonGeometry((nx, ny), ang, n_detectors=nd)


def _cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")


@pytest.mark.parametrize("case", range(len(CASES)))
def test_forward_matches_oracle(case):
    from graptor_amd import radon
    nx, ny, nd, ang, ns = CASES[case]
    print(f"forward case {case}: {nx}x{ny} nd={nd} na={ang.size} slices={ns}", flush=True)
    g = _geom(nx, ny, nd, ang)
    rng = np.random.default_rng(100 + case)
    img = rng.random((ns, ny, nx)).astype(np.float32)
    sino = torch.zeros((ns, ang.size, nd), device="cuda")
    radon(sino, _cuda(img), (g, (nx, ny), (nd, ang.size)))
    got = sino.cpu().numpy()
    ofs = g.ofs.T.copy()
    ref64 = O.radon_fwd(img.astype(np.float64), ofs, nd)
    ref32 = cref.radon(img, ofs, nd)
    assert rel(got, ref64) <= 1e-5, rel(got, ref64)
    assert rel(got, ref32) <= 1e-5, rel(got, ref32)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_backprojection_matches_oracle(case):
    from graptor_amd import radon_ad
    nx, ny, nd, ang, ns = CASES[case]
    print(f"backprojection case {case}: {nx}x{ny} nd={nd} na={ang.size} slices={ns}",
          flush=True)
    g = _geom(nx, ny, nd, ang)
    rng = np.random.default_rng(200 + case)
    sino = rng.random((ns, ang.size, nd)).astype(np.float32)
    img = torch.zeros((ns, ny, nx), device="cuda")
    radon_ad(img, _cuda(sino), g)
    got = img.cpu().numpy()
    ofs = g.ofs.T.copy()
    ref64 = O.radon_adj(sino.astype(np.float64), ofs, nx, ny)
    ref32 = cref.radon_ad(sino, ofs, nx, ny)
    assert rel(got, ref64) <= 1e-5, rel(got, ref64)
    assert rel(got, ref32) <= 1e-5, rel(got, ref32)


def test_adjoint_dot_identity_on_gpu():
    from graptor_amd import radon, radon_ad
    g = _geom(128, 128, 128, angles_deg(-75, 75, 60))
    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 128, 128)).astype(np.float32)
    y = rng.standard_normal((4, 60, 128)).astype(np.float32)
    rx = torch.zeros((4, 60, 128), device="cuda")
    aty = torch.zeros((4, 128, 128), device="cuda")
    radon(rx, _cuda(x), g)
    radon_ad(aty, _cuda(y), g)
    lhs = np.vdot(rx.cpu().numpy().astype(np.float64), y.astype(np.float64))
    rhs = np.vdot(x.astype(np.float64), aty.cpu().numpy().astype(np.float64))
    scale = np.linalg.norm(rx.cpu().numpy()) * np.linalg.norm(y)
    assert abs(lhs - rhs) <= 1e-6 * scale


def test_fortran_order_views_accepted():
    """The reference's F-ordered (Nx, Ny, S) / (nd, na, S) arrays work as-is."""
    from graptor_amd import radon
    g = _geom(48, 48, 48, angles_deg(-70, 70, 16))
    rng = np.random.default_rng(9)
    img = rng.random((3, 48, 48)).astype(np.float32)
    f_img = _cuda(img).permute(2, 1, 0)                   # (Nx, Ny, S), F-contiguous
    f_sino = torch.zeros((3, 16, 48), device="cuda").permute(2, 1, 0)
    radon(f_sino, f_img, g)
    got = f_sino.permute(2, 1, 0).cpu().numpy()
    assert rel(got, O.radon_fwd(img.astype(np.float64), g.ofs.T.copy(), 48)) <= 1e-5


@pytest.mark.parametrize("band,tile", [(4, 4), (8, 8), (16, 32), (32, 16)])
def test_tile_shapes_agree(band, tile):
    from graptor_amd import radon, radon_ad
    g = _geom(100, 100, 110, angles_deg(-75, 75, 25))
    g.set_tuning(fwd_band=band, adj_tile_y=tile)
    assert g.get_tuning() == (band, tile)
    rng = np.random.default_rng(11)
    img = rng.random((4, 100, 100)).astype(np.float32)
    sino = torch.zeros((4, 25, 110), device="cuda")
    radon(sino, _cuda(img), g)
    back = torch.zeros((4, 100, 100), device="cuda")
    radon_ad(back, sino, g)
    ofs = g.ofs.T.copy()
    s64 = O.radon_fwd(img.astype(np.float64), ofs, 110)
    assert rel(sino.cpu().numpy(), s64) <= 1e-5
    assert rel(back.cpu().numpy(), O.radon_adj(s64, ofs, 100, 100)) <= 1e-5


@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("nslices", [1, 2, 3, 8, 11])
def test_tiles_every_slice_group(nslices, tile):
    """Tiles of <= 16 rows take the backprojection with 4 accumulator rows
    per lane, 32-row tiles with 8; slice groups of 1, 2, 4 (3 slices: one
    padded group) and 8 (two LDS planes; 11 slices: one padded group)."""
    from graptor_amd import radon_ad
    g = _geom(70, 70, 80, angles_deg(-70, 70, 19))
    g.set_tuning(fwd_band=8, adj_tile_y=tile)
    rng = np.random.default_rng(12 + nslices)
    sino = rng.random((nslices, 19, 80)).astype(np.float32)
    back = torch.zeros((nslices, 70, 70), device="cuda")
    radon_ad(back, _cuda(sino), g)
    ofs = g.ofs.T.copy()
    assert rel(back.cpu().numpy(), O.radon_adj(sino.astype(np.float64), ofs, 70, 70)) <= 1e-5


def test_normest_fp64():
    from graptor_amd import radon_normest
    g = _geom(40, 40, 40, angles_deg(-70, 70, 24))
    ofs = g.ofs.T.copy()
    x0 = np.random.default_rng(2).standard_normal((40, 40))
    got = radon_normest(None, g, mode="reference", x0=x0)
    ref = O.normest_reference(ofs, 40, 40, 40, x0.T)  # oracle takes (ny, nx)
    assert abs(got - ref) <= 1e-10 * ref
    gp = radon_normest(None, g, mode="power", tol=1e-13, iters=500)
    M = O.radon_matrix(ofs, 40, 40, 40)
    smax = np.linalg.svd(M, compute_uv=False)[0]
    assert abs(gp - smax) <= 1e-7 * smax


def test_make_feasible_matches_mask():
    from graptor_amd import make_function_feasible
    g = _geom(30, 30, 44, angles_deg(-70, 70, 12))
    f = torch.ones((3, 12, 44), device="cuda")
    make_function_feasible(f, g)
    mask = O.feasible_mask(g.ofs.T.copy(), 30, 30, 44)
    np.testing.assert_array_equal(f.cpu().numpy(), np.broadcast_to(mask, (3, 12, 44)) * 1.0)


def test_large_detector_count_chunks_bins():
    """nd > 2048 splits each angle's bins over several chunks."""
    from graptor_amd import radon
    g = _geom(64, 64, 2300, angles_deg(-60, 60, 3))
    img = np.random.default_rng(4).random((1, 64, 64)).astype(np.float32)
    sino = torch.zeros((1, 3, 2300), device="cuda")
    radon(sino, _cuda(img), g)
    assert rel(sino.cpu().numpy(), O.radon_fwd(img.astype(np.float64), g.ofs.T.copy(),
                                               2300)) <= 1e-5
