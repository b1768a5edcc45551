"""The C-ABI library loads, exports every symbol include/graptor_amd.h
declares, and its host-only paths behave (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from graptor_amd import _lib
from oracle import tgv_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "graptor_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 13
    for s in syms:
        assert hasattr(lib, s), f"missing export {s}"
    assert set(syms) == set(_lib.EXPORTS)
    assert lib.gr_abi_version() == 1


def test_radon_struct_matches_oracle():
    lib = _lib.load()
    for shape, nd, ang in [((16, 16), 16, np.linspace(0, 3, 7)),
                           ((30, 20), 41, np.deg2rad(np.linspace(-75, 75, 60)) + np.pi / 2),
                           ((8, 8), 8, np.array([np.pi / 2, 0.0, np.pi]))]:
        ofs = np.zeros((ang.size, 4), np.float32)
        a = np.ascontiguousarray(ang, np.float64)
        assert lib.gr_radon_struct(a.ctypes.data, ang.size, shape[0], shape[1], nd, 1.0, 0.0,
                                   ofs.ctypes.data) == 0
        ref, _, _ = O.radon_struct(shape, ang, n_detectors=nd)
        np.testing.assert_array_equal(ofs.T, ref)


def test_invalid_arguments_are_reported_not_raised():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.gr_geometry_create(None, 0, 4, 4, 4, ctypes.byref(h)) == _lib.GR_ERR_INVALID
    assert b"invalid" in lib.gr_last_error()
    assert lib.gr_radon(None, None, None, 1, None) == _lib.GR_ERR_INVALID
    assert lib.gr_radon_ad(None, None, None, 1, None) == _lib.GR_ERR_INVALID
    out = ctypes.c_double()
    assert lib.gr_normest(None, None, 10, 0, 0.0, ctypes.byref(out), None) == _lib.GR_ERR_INVALID
    assert lib.gr_tgv_iterate(None, None, None, 1, 0, None) == _lib.GR_ERR_INVALID
    with pytest.raises(ValueError):
        _lib.check(_lib.GR_ERR_INVALID, "x")
    with pytest.raises(MemoryError):
        _lib.check(_lib.GR_ERR_OOM, "x")


def test_pack_halos_rejects_what_a_slab_does_not_send():
    """gr_tgv_pack_halos checks its arguments on the host before any launch:
    3D only, a known phase, no pbar / w message without TGV; with both
    destinations NULL it has nothing to do."""
    lib = _lib.load()
    prm = _lib.TgvParams(dim=3, nx=8, ny=8, nz=4, nch=2, coupling=1, discrepancy=0,
                         regularisation=_lib.REGULARISATION["TV"], stepsize=1.0)
    fake = 1 << 20  # never dereferenced: every call below returns before a launch
    st = _lib.TgvState(ubar=fake, v=fake)
    P, S = ctypes.byref(prm), ctypes.byref(st)
    dual, primal = _lib.HALO_PHASE["dual"], _lib.HALO_PHASE["primal"]
    assert lib.gr_tgv_pack_halos(P, S, dual, None, fake, None) == _lib.GR_ERR_INVALID
    assert b"TGV only" in lib.gr_last_error()
    assert lib.gr_tgv_pack_halos(P, S, primal, fake, None, None) == _lib.GR_ERR_INVALID
    assert lib.gr_tgv_pack_halos(P, S, 2, None, None, None) == _lib.GR_ERR_INVALID
    assert lib.gr_tgv_pack_halos(P, S, dual, None, None, None) == _lib.GR_OK
    prm.dim, prm.nz = 2, 1
    assert lib.gr_tgv_pack_halos(P, S, dual, None, None, None) == _lib.GR_ERR_INVALID
    assert lib.gr_tgv_pack_halos(None, S, dual, None, None, None) == _lib.GR_ERR_INVALID


def test_header_is_plain_c():
    """include/graptor_amd.h compiles as C99 (no torch or C++ types at the
    boundary), as a cgo or ctypes binding would read it."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    r = subprocess.run([gcc, "-fsyntax-only", "-std=c99", "-Wall", "-Werror", "-x", "c",
                        os.path.join(ROOT, "include", "graptor_amd.h")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_param_struct_layout_matches_header():
    # 8 ints then 7 floats, no padding
    assert ctypes.sizeof(_lib.TgvParams) == 15 * 4
    assert ctypes.sizeof(_lib.TgvState) == 18 * ctypes.sizeof(ctypes.c_void_p)
