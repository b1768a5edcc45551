"""The allocation-free energy entry point (gr_tgv_energy_noalloc) against the
allocating one (gr_tgv_energy), through the C ABI: the same FP64 values bit
for bit on consecutive calls that share the library's fixed scratch array,
the second with a smaller problem, and a clear refusal beyond its channel
limit."""
import ctypes

import numpy as np
import pytest

from oracle import tgv_oracle as O
from tests.helpers import angles_deg, make_problem

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _problem(seed, dim, C, N, na, Nz, coupling, disc, reg):
    from graptor_amd.projector import RadonGeometry
    from graptor_amd.solver import TgvProblem
    _, f, ofs = make_problem(seed, C, Nz, N, na, poisson=30.0)
    g = RadonGeometry((N, N), angles_deg(-70, 70, na), n_detectors=N)
    L = O.normest_power(ofs, N, N, N)
    mu = list(np.linspace(1.0, 0.2, C))
    return TgvProblem(f.astype(np.float32), g, (4.0, 1.0), mu, dim=dim, coupling=coupling,
                      discrepancy=disc, regularisation=reg, L=L)


def _energy(prob, fn):
    from graptor_amd import _lib
    out = (ctypes.c_double * 3)()
    mu = np.ascontiguousarray(prob.mu, dtype=np.float64)
    with torch.cuda.device(prob.device):
        rc = getattr(_lib.load(), fn)(prob.geometry.handle, ctypes.byref(prob.params),
                                      ctypes.byref(prob.state), mu.ctypes.data,
                                      prob.alpha[0], prob.alpha[1], prob.k.L, out,
                                      prob._stream())
    _lib.check(rc, fn)
    return np.array(out[:], dtype=np.float64)


def test_noalloc_energy_equals_allocating_energy():
    cases = [
        # a larger problem first, then a smaller one that reuses the same scratch
        (71, 3, 4, 48, 24, 6, "frobenius", "KL", "TGV"),
        (72, 2, 2, 24, 14, 1, "nuclear", "L2", "TGV"),
        (73, 3, 3, 32, 16, 4, "uncorrelated", "KL", "TV"),
    ]
    for case in cases:
        print(f"energy noalloc case {case}", flush=True)
        prob = _problem(*case)
        prob.iterate(25)
        # a positive shift keeps R u > 0 wherever f > 0, so the KL term is finite
        prob.u.add_(0.05)
        torch.cuda.synchronize()
        ref = _energy(prob, "gr_tgv_energy")
        got = _energy(prob, "gr_tgv_energy_noalloc")
        print(f"  allocating {ref.tolist()} noalloc {got.tolist()}", flush=True)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got, ref), (got, ref)
        # and again: the scratch keeps the last call's contents, which must not leak
        again = _energy(prob, "gr_tgv_energy_noalloc")
        assert np.array_equal(again, ref)


def test_noalloc_energy_refuses_too_many_channels():
    from graptor_amd import _lib
    C = _lib.ENERGY_NOALLOC_MAX_CHANNELS + 1
    prob = _problem(74, 2, C, 16, 8, 1, "uncorrelated", "L2", "TGV")
    with pytest.raises(ValueError, match="exceed the 32"):
        _energy(prob, "gr_tgv_energy_noalloc")
    # the allocating form has no such limit
    assert np.all(np.isfinite(_energy(prob, "gr_tgv_energy")))
