"""ctypes binding of the FP32 C restatement ``cpu_ref.c`` (test infrastructure).

Arrays use the compact layouts documented in ``tgv_oracle``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libcpuref.so")
# the same source built with FMA contraction (a second FP32 compilation)
LIB_FMA = os.path.join(_HERE, "_build", "libcpuref_fma.so")
_lib = None

CPL = {"uncorrelated": 0, "frobenius": 1, "nuclear": 2}
DISC = {"KL": 0, "L2": 1}


def build(arch_flags: str | None = None, out: str | None = None) -> str:
    """Compile cpu_ref.c (make); returns the library path."""
    env = dict(os.environ)
    cmd = ["make", "-s", "-C", _HERE]
    if arch_flags is not None:
        cmd.append(f"ARCHFLAGS={arch_flags}")
    subprocess.run(cmd, check=True, env=env)
    return _LIB


def load(path: str | None = None):
    global _lib
    if path is not None:
        lib = ctypes.CDLL(path)
    else:
        if _lib is not None:
            return _lib
        if not os.path.exists(_LIB):
            build()
        lib = ctypes.CDLL(_LIB)
    P = ctypes.c_void_p
    I = ctypes.c_int
    D = ctypes.c_double
    lib.cref_radon.argtypes = [P, P, P, I, I, I, I, I]
    lib.cref_radon_ad.argtypes = [P, P, P, I, I, I, I, I]
    lib.cref_tgv_solve.argtypes = [I, I, I, I, I, P, P, D, D, D, P, I, I, I, I,
                                   D, P, P, P, P]
    lib.cref_tgv_solve.restype = I
    lib.cref_jacobi3.argtypes = [P, P, P]
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ofs_aos(ofs):
    """(4, na) table -> na float4 records, as the kernels read it."""
    return np.ascontiguousarray(np.asarray(ofs, np.float32).T)


def radon(img, ofs, nd, lib=None):
    """img (..., Ny, Nx) float32 -> sino (..., na, nd) float32."""
    lib = lib or load()
    img = np.ascontiguousarray(img, dtype=np.float32)
    ny, nx = img.shape[-2:]
    ns = int(np.prod(img.shape[:-2], dtype=np.int64)) if img.ndim > 2 else 1
    na = ofs.shape[1]
    out = np.zeros(img.shape[:-2] + (na, nd), np.float32)
    o = _ofs_aos(ofs)
    lib.cref_radon(_ptr(out), _ptr(img), _ptr(o), nx, ny, nd, na, ns)
    return out


def radon_ad(sino, ofs, nx, ny, lib=None):
    lib = lib or load()
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    na, nd = sino.shape[-2:]
    ns = int(np.prod(sino.shape[:-2], dtype=np.int64)) if sino.ndim > 2 else 1
    out = np.zeros(sino.shape[:-2] + (ny, nx), np.float32)
    o = _ofs_aos(ofs)
    lib.cref_radon_ad(_ptr(out), _ptr(sino), _ptr(o), nx, ny, nd, na, ns)
    return out


def solve(f, ofs, alpha, mu, maxiter, L, coupling="frobenius", discrepancy="KL",
          regularisation="TGV", dim=3, stepsize=1.0, u0=None, lib=None):
    """FP32 restatement of the full iteration; f (C, Nz, na, nd) already
    feasibility-masked.  Returns dict(u, ku, p)."""
    lib = lib or load()
    f = np.ascontiguousarray(f, dtype=np.float32)
    C, Nz, na, nd = f.shape
    d = 2 if dim == 2 else 3
    u = np.zeros((C, Nz, nd, nd), np.float32)
    ku = np.zeros_like(f)
    p = np.zeros((d, C, Nz, nd, nd), np.float32)
    mu = np.ascontiguousarray(np.asarray(mu, np.float64).reshape(-1))
    o = _ofs_aos(ofs)
    u0p = None
    if u0 is not None:
        u0 = np.ascontiguousarray(u0, dtype=np.float32)
        u0p = _ptr(u0)
    rc = lib.cref_tgv_solve(dim, nd, na, Nz, C, _ptr(o), _ptr(f), float(L),
                            float(alpha[0]), float(alpha[1]), _ptr(mu), int(maxiter),
                            CPL[coupling], DISC[discrepancy],
                            1 if regularisation.upper() == "TGV" else 0,
                            float(stepsize), u0p, _ptr(u), _ptr(ku), _ptr(p))
    if rc != 0:
        raise RuntimeError(f"cref_tgv_solve failed with status {rc}")
    return dict(u=u, ku=ku, p=p)


def jacobi3(G6):
    lib = load()
    G = np.ascontiguousarray(G6, np.float32)
    ev = np.zeros(3, np.float32)
    V = np.zeros(9, np.float32)
    lib.cref_jacobi3(_ptr(G), _ptr(ev), _ptr(V))
    return ev, V.reshape(3, 3)
