"""GPU tests of the z-slab decomposition: a volume solved as slabs with halo
planes exchanged between the dual and the primal steps must give the same
bits as the same volume solved whole, for uneven slab sizes, every coupling,
TV and TGV, and a z step size other than one."""
import ctypes

import numpy as np
import pytest

from oracle import tgv_oracle as O
from tests.helpers import make_problem, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [
    ("frobenius", "KL", "TGV", 1.0),
    ("nuclear", "L2", "TGV", 2.0),
    ("uncorrelated", "KL", "TV", 1.0),
]


def _setup(seed=61, C=3, Nz=10, N=40, na=20):
    from graptor_amd.projector import RadonGeometry
    _, f, ofs = make_problem(seed, C, Nz, N, na, poisson=40.0)
    ang = np.deg2rad(np.linspace(-70, 70, na, endpoint=False)) + np.pi / 2
    g = RadonGeometry((N, N), ang, n_detectors=N)
    L = O.normest_power(ofs, N, N, N)
    return f.astype(np.float32), g, L


@pytest.mark.parametrize("case", CASES, ids=["-".join(map(str, c)) for c in CASES])
def test_slabs_with_halos_equal_the_whole_volume(case):
    from graptor_amd.slabs import SlabSolve
    coupling, disc, reg, h = case
    print(f"slabs {case}", flush=True)
    f, g, L = _setup()
    kw = dict(coupling=coupling, discrepancy=disc, regularisation=reg, stepsize=h)
    whole = SlabSolve(f, g, (4, 1), [1.0, 0.3, 0.1], 1, L, **kw)
    parts = SlabSolve(f, g, (4, 1), [1.0, 0.3, 0.1], 3, L, **kw)
    assert [b - a for a, b in parts.bounds] == [3, 4, 3]
    for n in (1, 6, 5):
        whole.iterate(n)
        parts.iterate(n)
        torch.cuda.synchronize()
        for name in ("u", "ubar", "p", "pbar", "v", "w", "lam"):
            if getattr(whole.slabs[0], name) is None:
                continue
            a, b = parts.gather(name), getattr(whole.slabs[0], name)
            assert torch.equal(a, b), (n, name, rel(a.cpu().numpy(), b.cpu().numpy()))


def test_slab_steps_match_the_fused_iteration():
    """The single-step composition (dual step, projection, prox,
    backprojection, primal step) equals gr_tgv_iterate's kernels."""
    from graptor_amd.slabs import SlabSolve
    from graptor_amd.solver import TgvProblem
    f, g, L = _setup(seed=62, Nz=6)
    whole = SlabSolve(f, g, (4, 1), [1.0, 0.3, 0.1], 1, L)
    ref = TgvProblem(f, g, (4, 1), [1.0, 0.3, 0.1], dim=3, L=L)
    whole.iterate(15)
    ref.iterate(15)
    torch.cuda.synchronize()
    for name in ("u", "p", "v", "w", "lam"):
        a = getattr(whole.slabs[0], name).cpu().numpy()
        b = getattr(ref, name).cpu().numpy()
        assert rel(a, b) <= 1e-6, name


def test_halo_state_validation():
    """Halo planes only in 3D, all planes of a side together, and not with
    the alternate state buffers of the fused form."""
    from graptor_amd import _lib
    from graptor_amd.solver import TgvProblem
    f, g, L = _setup(seed=63, Nz=4)
    p = TgvProblem(f, g, (4, 1), [1.0, 0.3, 0.1], dim=3, L=L)
    plane = torch.zeros((3, 3, 40, 40), device=p.device)
    p.state.halo_v_lo = plane.data_ptr()  # pbar_lo missing with TGV
    with pytest.raises(ValueError):
        p.iterate(1)
    p.state.halo_pbar_lo = plane.data_ptr()
    # a complete halo side: the single steps take it, the whole iteration
    # refuses it (the halos must be exchanged between its dual and primal step)
    lib = _lib.load()
    _lib.check(lib.gr_tgv_dual_step(g.handle, ctypes.byref(p.params), ctypes.byref(p.state),
                                    p._stream()))
    torch.cuda.synchronize()
    with pytest.raises(NotImplementedError):
        p.iterate(1)
    with pytest.raises(NotImplementedError):
        p.profile(1)
    p.state.halo_w_hi = plane.data_ptr()  # w_hi without ubar_hi
    with pytest.raises(ValueError):
        p.iterate(1)
    with pytest.raises(ValueError):
        _lib.check(lib.gr_tgv_dual_step(g.handle, ctypes.byref(p.params), ctypes.byref(p.state),
                                        p._stream()))


def _pack(prob, phase, below, above):
    import ctypes

    from graptor_amd import _lib
    lib = _lib.load()
    return lib.gr_tgv_pack_halos(ctypes.byref(prob.params), ctypes.byref(prob.state),
                                 _lib.HALO_PHASE[phase],
                                 below.data_ptr() if below is not None else None,
                                 above.data_ptr() if above is not None else None, prob._stream())


@pytest.mark.parametrize("N,offset", [(40, 0), (39, 0), (40, 1)],
                         ids=["float4", "odd-plane", "unaligned"])
def test_pack_halos_equal_the_edge_planes(N, offset):
    """gr_tgv_pack_halos writes exactly the planes a neighbour's halo fields
    hold (ubar of plane 0, pbar of the last plane; w_xz, w_yz, w_zz of plane
    0, v_z of the last plane), in the [component][channel] order, for the
    float4 and the scalar copy paths; a NULL side leaves its buffer alone."""
    from graptor_amd import _lib
    from graptor_amd.solver import TgvProblem
    print(f"pack_halos N={N} offset={offset}", flush=True)
    f, g, L = _setup(seed=64, Nz=5, N=N)
    prob = TgvProblem(f, g, (4, 1), [1.0, 0.3, 0.1], dim=3, L=L)
    gen = torch.Generator(device=prob.device).manual_seed(7)
    for name in ("ubar", "pbar", "v", "w"):
        t = getattr(prob, name)
        t.copy_(torch.randn(t.shape, generator=gen, device=t.device))
    C = prob.C

    def buf(ncomp):
        n = ncomp * C * N * N
        raw = torch.full((n + offset,), float("nan"), device=prob.device)
        return raw[offset:].view(ncomp, C, N, N)

    d_lo, d_hi, p_lo, p_hi = buf(1), buf(3), buf(3), buf(1)
    _lib.check(_pack(prob, "dual", d_lo, d_hi), "pack dual")
    _lib.check(_pack(prob, "primal", p_lo, p_hi), "pack primal")
    torch.cuda.synchronize()
    assert torch.equal(d_lo[0], prob.ubar[:, 0])
    assert torch.equal(d_hi, prob.pbar[:, :, -1])
    assert torch.equal(p_lo, torch.stack([prob.w[m, :, 0] for m in (4, 5, 2)]))
    assert torch.equal(p_hi[0], prob.v[2, :, -1])
    # one side only (a slab at the volume's edge)
    lone = buf(1)
    _lib.check(_pack(prob, "primal", None, lone), "pack primal above")
    _lib.check(_pack(prob, "dual", None, None), "pack nothing")
    torch.cuda.synchronize()
    assert torch.equal(lone[0], prob.v[2, :, -1])
