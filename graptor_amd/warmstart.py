"""Learned warm start: a 4-level U-Net on hand-written BF16 MFMA kernels.

The network maps the channel-wise backprojections R^T f of a 2D slice
(4 channels, 256 x 256) to an estimate of its TGV reconstruction, which
initialises u of the primal-dual solver; the plain iteration then runs to
tolerance from there (the network never replaces the solver).

Everything numeric runs in libgraptor_amd (csrc/cnn.hip): implicit-GEMM
convolutions on v_mfma_f32_32x32x16_bf16 with FP32 accumulation (forward,
data gradient through flipped weights, weight gradient with split-K),
2x2 average pooling, nearest upsampling, ReLU backward, MSE gradient and
Adam.  Activations are NHWC bf16; the skip concatenations are channel slices
of shared buffers, so no copies are made.  torch only allocates memory.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .batch import BatchPlan
from .phantoms import multichannel_slices, poisson_sinograms, tilt_angles
from .projector import RadonGeometry, make_function_feasible, radon_ad, radon_normest
from .solver import TgvProblem

CIN_PAD = 16  # input channels padded to a multiple of 16 for the MFMA K step
# weights trained by this package on its own solver's reconstructions of
# 256 x 256 x 4 slices (bf16 bits, as inference uses them)
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                            "warmstart_unet_256x4.npz")


def _lib_cnn():
    return _lib.load()


class View:
    """A channel slice [c0, c0 + c) of an NHWC buffer with ``ld`` channels."""

    def __init__(self, buf: torch.Tensor, c0: int, c: int):
        self.buf, self.c0, self.c = buf, c0, c
        self.ld = buf.shape[-1]
        self.n, self.h, self.w = buf.shape[0], buf.shape[1], buf.shape[2]

    @property
    def ptr(self):
        return self.buf.data_ptr()

    def npix(self):
        return self.n * self.h * self.w


class Conv:
    """One convolution layer: FP32 master weights W[co][tap][ci] and bias,
    their bf16 copy, Adam state and FP32 gradient accumulators, all views of
    the network's flat parameter buffers."""

    def __init__(self, name, cin, cout, taps, relu, out_f32=False):
        self.name, self.cin, self.cout, self.taps = name, cin, cout, taps
        self.relu, self.out_f32 = relu, out_f32
        self.nw = cout * taps * cin
        self.nb = cout


class UNet:
    """4-level U-Net (base, 2 base, 4 base, 8 base channels; two 3x3 convs per
    level, 2x2 average pooling, nearest upsampling with skip concatenation, a
    1x1 head).  ``cin`` real input and ``cout`` real output channels."""

    def __init__(self, cin=4, cout=4, base=32, device="cuda", seed=0):
        self.dev = torch.device(device)
        self.cin, self.cout, self.base = cin, cout, base
        c0, c1, c2, c3 = base, 2 * base, 4 * base, 8 * base
        self.head_out = 32  # MFMA tile width; outputs beyond cout have zero weights
        specs = [("enc0a", CIN_PAD, c0, 9, True), ("enc0b", c0, c0, 9, True),
                 ("enc1a", c0, c1, 9, True), ("enc1b", c1, c1, 9, True),
                 ("enc2a", c1, c2, 9, True), ("enc2b", c2, c2, 9, True),
                 ("bota", c2, c3, 9, True), ("botb", c3, c3, 9, True),
                 ("dec2a", c2 + c3, c2, 9, True), ("dec2b", c2, c2, 9, True),
                 ("dec1a", c1 + c2, c1, 9, True), ("dec1b", c1, c1, 9, True),
                 ("dec0a", c0 + c1, c0, 9, True), ("dec0b", c0, c0, 9, True),
                 ("head", c0, self.head_out, 1, False)]
        self.layers = {n: Conv(n, ci, co, t, r, out_f32=(n == "head"))
                       for n, ci, co, t, r in specs}
        self.order = [s[0] for s in specs]
        total = sum(l.nw + l.nb for l in self.layers.values())
        self.nparams_real = sum(l.nw + l.nb for l in self.layers.values()) \
            - (self.head_out - cout) * (c0 + 1) - (CIN_PAD - cin) * 9 * c0
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.p = torch.zeros(total, **f32)
        self.g = torch.zeros(total, **f32)
        self.m = torch.zeros(total, **f32)
        self.v = torch.zeros(total, **f32)
        self.pb = torch.zeros(total, dtype=torch.bfloat16, device=self.dev)
        self.wflip = {}
        rng = np.random.default_rng(seed)
        off = 0
        host = np.zeros(total, np.float32)
        for n in self.order:
            l = self.layers[n]
            l.w_off, l.b_off = off, off + l.nw
            fan_in = l.taps * (cin if n == "enc0a" else l.cin)
            w = rng.standard_normal((l.cout, l.taps, l.cin)).astype(np.float32)
            w *= math.sqrt(2.0 / fan_in)
            if n == "enc0a":
                w[:, :, cin:] = 0.0
            if n == "head":
                w[cout:] = 0.0
                w *= 0.5
            host[l.w_off:l.w_off + l.nw] = w.reshape(-1)
            off += l.nw + l.nb
            self.wflip[n] = torch.zeros(l.nw, dtype=torch.bfloat16, device=self.dev)
        self.p.copy_(torch.as_tensor(host, device=self.dev))
        self.pb.copy_(self.p.to(torch.bfloat16))
        self.step = 0
        self._bufs = None
        self._batch = None

    # ---- buffers -------------------------------------------------------
    def _alloc(self, n, H, W):
        if self._batch == (n, H, W):
            return
        c0, c1, c2, c3 = self.base, 2 * self.base, 4 * self.base, 8 * self.base
        bf = dict(dtype=torch.bfloat16, device=self.dev)
        z = lambda h, w, c: torch.zeros((n, h, w, c), **bf)
        H1, W1, H2, W2, H3, W3 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        shapes = {
            "x": (H, W, CIN_PAD), "enc0a": (H, W, c0), "cat0": (H, W, c0 + c1),
            "p0": (H1, W1, c0), "enc1a": (H1, W1, c1), "cat1": (H1, W1, c1 + c2),
            "p1": (H2, W2, c1), "enc2a": (H2, W2, c2), "cat2": (H2, W2, c2 + c3),
            "p2": (H3, W3, c2), "bota": (H3, W3, c3), "botb": (H3, W3, c3),
            "dec2a": (H2, W2, c2), "dec2b": (H2, W2, c2), "dec1a": (H1, W1, c1),
            "dec1b": (H1, W1, c1), "dec0a": (H, W, c0), "dec0b": (H, W, c0),
        }
        self._bufs = {k: z(*v) for k, v in shapes.items()}
        self._grads = {k: z(*v) for k, v in shapes.items() if k != "x"}
        self._out = torch.zeros((n, H, W, self.head_out), dtype=torch.float32, device=self.dev)
        self._dout = torch.zeros((n, H, W, self.head_out), **bf)
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self._batch = (n, H, W)

    def _io(self):
        b, g = self._bufs, self._grads
        c0, c1, c2 = self.base, 2 * self.base, 4 * self.base
        V = View
        # (input view, output view) of each layer, and the gradient views
        io = {
            "enc0a": (V(b["x"], 0, CIN_PAD), V(b["enc0a"], 0, c0)),
            "enc0b": (V(b["enc0a"], 0, c0), V(b["cat0"], 0, c0)),
            "enc1a": (V(b["p0"], 0, c0), V(b["enc1a"], 0, c1)),
            "enc1b": (V(b["enc1a"], 0, c1), V(b["cat1"], 0, c1)),
            "enc2a": (V(b["p1"], 0, c1), V(b["enc2a"], 0, c2)),
            "enc2b": (V(b["enc2a"], 0, c2), V(b["cat2"], 0, c2)),
            "bota": (V(b["p2"], 0, c2), V(b["bota"], 0, 2 * c2)),
            "botb": (V(b["bota"], 0, 2 * c2), V(b["botb"], 0, 2 * c2)),
            "dec2a": (V(b["cat2"], 0, c2 + 2 * c2), V(b["dec2a"], 0, c2)),
            "dec2b": (V(b["dec2a"], 0, c2), V(b["dec2b"], 0, c2)),
            "dec1a": (V(b["cat1"], 0, c1 + c2), V(b["dec1a"], 0, c1)),
            "dec1b": (V(b["dec1a"], 0, c1), V(b["dec1b"], 0, c1)),
            "dec0a": (V(b["cat0"], 0, c0 + c1), V(b["dec0a"], 0, c0)),
            "dec0b": (V(b["dec0a"], 0, c0), V(b["dec0b"], 0, c0)),
        }
        gio = {
            "enc0a": (None, V(g["enc0a"], 0, c0)),
            "enc0b": (V(g["enc0a"], 0, c0), V(g["cat0"], 0, c0)),
            "enc1a": (V(g["p0"], 0, c0), V(g["enc1a"], 0, c1)),
            "enc1b": (V(g["enc1a"], 0, c1), V(g["cat1"], 0, c1)),
            "enc2a": (V(g["p1"], 0, c1), V(g["enc2a"], 0, c2)),
            "enc2b": (V(g["enc2a"], 0, c2), V(g["cat2"], 0, c2)),
            "bota": (V(g["p2"], 0, c2), V(g["bota"], 0, 2 * c2)),
            "botb": (V(g["bota"], 0, 2 * c2), V(g["botb"], 0, 2 * c2)),
            "dec2a": (V(g["cat2"], 0, 3 * c2), V(g["dec2a"], 0, c2)),
            "dec2b": (V(g["dec2a"], 0, c2), V(g["dec2b"], 0, c2)),
            "dec1a": (V(g["cat1"], 0, c1 + c2), V(g["dec1a"], 0, c1)),
            "dec1b": (V(g["dec1a"], 0, c1), V(g["dec1b"], 0, c1)),
            "dec0a": (V(g["cat0"], 0, c0 + c1), V(g["dec0a"], 0, c0)),
            "dec0b": (V(g["dec0a"], 0, c0), V(g["dec0b"], 0, c0)),
        }
        return io, gio

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    # ---- primitives --------------------------------------------------------
    def _conv(self, l: Conv, xin: View, out, ldo, co0, w_ptr=None, bias=True, relu=None,
              out_f32=None, cin=None, cout=None, taps=None):
        lib = _lib_cnn()
        wp = w_ptr if w_ptr is not None else self.pb.data_ptr() + 2 * l.w_off
        bp = (self.p.data_ptr() + 4 * l.b_off) if bias else None
        _lib.check(lib.gr_cnn_conv_fwd(
            xin.ptr, xin.ld, xin.c0, wp, bp, out, ldo, co0, xin.n, xin.h, xin.w,
            cin if cin is not None else l.cin, cout if cout is not None else l.cout,
            taps if taps is not None else l.taps,
            int(l.relu if relu is None else relu),
            int(l.out_f32 if out_f32 is None else out_f32), self._stream()), "conv_fwd")

    def forward(self, x_f32: torch.Tensor) -> torch.Tensor:
        """x_f32: (n, H, W, cin) float32 (already normalised).  Returns the
        (n, H, W, cout) float32 prediction (a view of an internal buffer)."""
        n, H, W, _ = x_f32.shape
        if H % 8 or W % 8:
            raise ValueError("image sizes must be multiples of 8")
        self._alloc(n, H, W)
        lib = _lib_cnn()
        b = self._bufs
        st = self._stream()
        xin = x_f32.contiguous()
        _lib.check(lib.gr_cnn_pack_bf16(xin.data_ptr(), self.cin, b["x"].data_ptr(), CIN_PAD, 0,
                                        n * H * W, 1.0, st), "pack")
        io, _ = self._io()
        c0, c1, c2, c3 = self.base, 2 * self.base, 4 * self.base, 8 * self.base
        L = self.layers

        def run(name):
            xi, yo = io[name]
            self._conv(L[name], xi, yo.ptr, yo.ld, yo.c0)

        def pool(src: View, dst: View):
            _lib.check(lib.gr_cnn_pool(src.ptr, src.ld, src.c0, dst.ptr, dst.ld, dst.c0, src.n,
                                       src.h, src.w, src.c, 0, 0, st), "pool")

        def up(src: View, dst: View):
            _lib.check(lib.gr_cnn_upsample(src.ptr, src.ld, src.c0, dst.ptr, dst.ld, dst.c0,
                                           src.n, src.h, src.w, src.c, 0, st), "up")

        V = View
        run("enc0a"); run("enc0b"); pool(V(b["cat0"], 0, c0), V(b["p0"], 0, c0))
        run("enc1a"); run("enc1b"); pool(V(b["cat1"], 0, c1), V(b["p1"], 0, c1))
        run("enc2a"); run("enc2b"); pool(V(b["cat2"], 0, c2), V(b["p2"], 0, c2))
        run("bota"); run("botb"); up(V(b["botb"], 0, c3), V(b["cat2"], c2, c3))
        run("dec2a"); run("dec2b"); up(V(b["dec2b"], 0, c2), V(b["cat1"], c1, c2))
        run("dec1a"); run("dec1b"); up(V(b["dec1b"], 0, c1), V(b["cat0"], c0, c1))
        run("dec0a"); run("dec0b")
        self._conv(L["head"], V(b["dec0b"], 0, c0), self._out.data_ptr(), self.head_out, 0)
        return self._out[..., :self.cout]

    def backward(self, target_f32: torch.Tensor, return_loss: bool = True):
        """MSE loss against ``target`` (n, H, W, cout) float32; accumulates the
        parameter gradients into self.g (zeroed first).  Returns the loss
        (which synchronises) when ``return_loss``, else None."""
        lib = _lib_cnn()
        st = self._stream()
        n, H, W = self._batch
        npix = n * H * W
        self.g.zero_()
        if return_loss:
            self._loss.zero_()
        tgt = target_f32.contiguous()
        total = npix * self.cout
        _lib.check(lib.gr_cnn_mse_grad(self._out.data_ptr(), self.head_out, tgt.data_ptr(),
                                       self.cout, npix, self._dout.data_ptr(), self.head_out,
                                       2.0 / total,
                                       self._loss.data_ptr() if return_loss else None, st),
                   "mse")
        io, gio = self._io()
        b, g = self._bufs, self._grads
        c0, c1, c2, c3 = self.base, 2 * self.base, 4 * self.base, 8 * self.base
        L = self.layers
        V = View

        def layer_back(name, xin: View, dy: View, yout: View, dx: View | None):
            l = L[name]
            if l.relu:
                _lib.check(lib.gr_cnn_relu_bwd(dy.ptr, dy.ld, dy.c0, yout.ptr, yout.ld, yout.c0,
                                               dy.npix(), l.cout, st), "relu_bwd")
            rows = n * xin.h
            # work-groups per chunk: 64 x 64 channel tiles x rows of taps
            blocks = ((l.cout + 63) // 64) * ((l.cin + 63) // 64) * (3 if l.taps == 9 else 1)
            chunks = max(1, min(rows, 1024 // max(blocks, 1)))
            rpc = max(1, (rows + chunks - 1) // chunks)
            _lib.check(lib.gr_cnn_conv_wgrad(xin.ptr, xin.ld, xin.c0, dy.ptr, dy.ld, dy.c0,
                                             self.g.data_ptr() + 4 * l.w_off, n, xin.h, xin.w,
                                             l.cin, l.cout, l.taps, rpc, st), "wgrad")
            _lib.check(lib.gr_cnn_bias_grad(dy.ptr, dy.ld, dy.c0, dy.npix(), l.cout,
                                            self.g.data_ptr() + 4 * l.b_off, st), "bias_grad")
            if dx is not None:
                wf = self.wflip[name]
                _lib.check(lib.gr_cnn_weight_flip(self.pb.data_ptr() + 2 * l.w_off,
                                                  wf.data_ptr(), l.cin, l.cout, l.taps, st),
                           "flip")
                self._conv(l, dy, dx.ptr, dx.ld, dx.c0, w_ptr=wf.data_ptr(), bias=False,
                           relu=False, out_f32=False, cin=l.cout, cout=l.cin)

        def pool_back(dcoarse: View, dfine: View):
            _lib.check(lib.gr_cnn_pool(dcoarse.ptr, dcoarse.ld, dcoarse.c0, dfine.ptr, dfine.ld,
                                       dfine.c0, dfine.n, dfine.h, dfine.w, dfine.c, 1, 1, st),
                       "pool_bwd")

        def up_back(dfine: View, dcoarse: View):
            _lib.check(lib.gr_cnn_upsample(dfine.ptr, dfine.ld, dfine.c0, dcoarse.ptr, dcoarse.ld,
                                           dcoarse.c0, dcoarse.n, dcoarse.h, dcoarse.w, dcoarse.c,
                                           1, st), "up_bwd")

        hb = View(b["dec0b"], 0, c0)
        layer_back("head", hb, View(self._dout, 0, self.head_out), None, View(g["dec0b"], 0, c0))
        for name in ("dec0b", "dec0a"):
            xi, yo = io[name]
            dxv, dyv = gio[name]
            layer_back(name, xi, dyv, yo, dxv)
        up_back(V(g["cat0"], c0, c1), V(g["dec1b"], 0, c1))
        for name in ("dec1b", "dec1a"):
            xi, yo = io[name]
            dxv, dyv = gio[name]
            layer_back(name, xi, dyv, yo, dxv)
        up_back(V(g["cat1"], c1, c2), V(g["dec2b"], 0, c2))
        for name in ("dec2b", "dec2a"):
            xi, yo = io[name]
            dxv, dyv = gio[name]
            layer_back(name, xi, dyv, yo, dxv)
        up_back(V(g["cat2"], c2, c3), V(g["botb"], 0, c3))
        for name in ("botb", "bota"):
            xi, yo = io[name]
            dxv, dyv = gio[name]
            layer_back(name, xi, dyv, yo, dxv)
        pool_back(V(g["p2"], 0, c2), V(g["cat2"], 0, c2))
        for name in ("enc2b", "enc2a"):
            xi, yo = io[name]
            dxv, dyv = gio[name]
            layer_back(name, xi, dyv, yo, dxv)
        pool_back(V(g["p1"], 0, c1), V(g["cat1"], 0, c1))
        for name in ("enc1b", "enc1a"):
            xi, yo = io[name]
            dxv, dyv = gio[name]
            layer_back(name, xi, dyv, yo, dxv)
        pool_back(V(g["p0"], 0, c0), V(g["cat0"], 0, c0))
        for name in ("enc0b", "enc0a"):
            xi, yo = io[name]
            dxv, dyv = gio[name]
            layer_back(name, xi, dyv, yo, dxv)
        return float(self._loss.item()) / total if return_loss else None

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        lib = _lib_cnn()
        _lib.check(lib.gr_cnn_adam(self.p.data_ptr(), self.g.data_ptr(), self.m.data_ptr(),
                                   self.v.data_ptr(), self.pb.data_ptr(), self.p.numel(), lr, b1,
                                   b2, eps, self.step, wd, self._stream()), "adam")
        # the padded input channels and output rows see zero gradients (their
        # activations / output gradients are zero), so they stay zero

    def params_numpy(self):
        """{name: (w (cout, taps, cin), b)} as float64 numpy (for the oracle)."""
        host = self.p.double().cpu().numpy()
        out = {}
        for n in self.order:
            l = self.layers[n]
            out[n] = (host[l.w_off:l.w_off + l.nw].reshape(l.cout, l.taps, l.cin),
                      host[l.b_off:l.b_off + l.nb])
        return out

    def grads_numpy(self):
        host = self.g.double().cpu().numpy()
        return {n: (host[l.w_off:l.w_off + l.nw].reshape(l.cout, l.taps, l.cin),
                    host[l.b_off:l.b_off + l.nb]) for n, l in self.layers.items()}


# ---------------------------------------------------------------------------
# Data, training and the warm-start evaluation
# ---------------------------------------------------------------------------
@dataclass
class SliceConfig:
    """The 2D slice problems of the warm start: C-channel N x N ellipse
    slices (channel 0 HAADF-like, the others element maps), na tilts over
    ``tilt`` degrees, N detector bins, Poisson counts, each slice's channels
    scaled as the reference's caller scales data (Reconstruction_coupled.py:
    349-350, Scalepars = [1, 98]: 1st percentile -> 0, 98th -> 1),
    Frobenius-coupled TGV-KL with the README's alpha and mu."""
    N: int = 256
    C: int = 4
    na: int = 60
    tilt: tuple = (-75.0, 75.0)
    counts: tuple = (1e3, 20.0)
    alpha: tuple = (4.0, 1.0)
    mu: tuple = (0.2, 5e-4, 4e-3, 5e-4)
    coupling: str = "frobenius"
    discrepancy: str = "KL"
    percentiles: tuple = (1.0, 98.0)

    def geometry(self):
        return RadonGeometry((self.N, self.N), tilt_angles(self.tilt[0], self.tilt[1], self.na),
                             n_detectors=self.N)


def make_data(cfg: SliceConfig, geom, seeds, noise_seed, device="cuda"):
    """Sinograms (B, C, na, nd) of the seeded slices: Poisson counts, zeroed
    where R 1 <= 0, then percentile-scaled per slice and channel (negative
    values after the shift clipped to 0)."""
    img = multichannel_slices(seeds, cfg.N, cfg.C, device=device)
    f = poisson_sinograms(img, geom, cfg.counts, seed=noise_seed)
    make_function_feasible(f.view(-1, cfg.na, cfg.N), geom)
    if cfg.percentiles is not None:
        qs = torch.tensor([cfg.percentiles[0] / 100.0, cfg.percentiles[1] / 100.0],
                          device=f.device)
        q = torch.quantile(f.flatten(2), qs, dim=2)  # (2, B, C)
        lo, hi = q[0][..., None, None], (q[1] - q[0]).clamp(min=1e-12)[..., None, None]
        f = ((f - lo) / hi).clamp(min=0.0).contiguous()
    return f


def features(f, geom):
    """Network input and channel scales of sinograms f (B, C, na, nd).

    scale[b, c] = sum(f[b, c]) / (na N^2) is the mean pixel value of a
    reconstruction that reproduces the data's mass (every parallel-beam
    projection of an image sums to the image's sum).  The input is the
    channel-wise backprojection R^T f divided by scale * na * N, so it is
    O(1) whatever the data's scale.  Returns (x (B, N, N, C) float32, scale)."""
    B, C, na, nd = f.shape
    N = geom.shape[0]
    bp = torch.zeros((B * C, N, N), device=f.device)
    radon_ad(bp, f.reshape(B * C, na, nd), geom)
    scale = (f.sum(dim=(2, 3)) / (na * N * N)).clamp(min=1e-12)
    x = bp.view(B, C, N, N) / (scale * (na * N)).view(B, C, 1, 1)
    return x.permute(0, 2, 3, 1).contiguous(), scale


class SliceBatch:
    """B independent 2D problems of one geometry advanced by one batch plan.
    ``step_ratio`` s != 1 replaces the reference's steps sigma = tau =
    1/sqrt(Lsqr) by sigma = s/sqrt(Lsqr), tau = 1/(s sqrt(Lsqr)) (the same
    product, so the same convergence condition and the same fixed point);
    only the u* check uses it."""

    def __init__(self, cfg: SliceConfig, geom, f, L, u0=None, step_ratio=1.0):
        B = f.shape[0]
        self.cfg, self.B = cfg, B
        self.problems = [
            TgvProblem(f[b].view(cfg.C, 1, cfg.na, cfg.N), geom, cfg.alpha, cfg.mu, dim=2,
                       coupling=cfg.coupling if cfg.C > 1 else "uncorrelated",
                       discrepancy=cfg.discrepancy, L=L, feasible=False,
                       u0=None if u0 is None else u0[b].view(cfg.C, 1, cfg.N, cfg.N),
                       device=f.device, fused=False)
            for b in range(B)]
        if step_ratio != 1.0:
            if cfg.discrepancy != "KL":
                raise ValueError("step_ratio: the L2 prox factor depends on sigma")
            for p in self.problems:
                s0 = p.params.sigma
                p.params.sigma, p.params.tau = s0 * step_ratio, s0 / step_ratio
                p.params.sigma_p = p.params.sigma / L
        self.plan = BatchPlan(self.problems)

    def energy(self):
        """FP64 primal energy of every slice at its current (u, p)."""
        return [p.energy()[2] for p in self.problems]

    def iterate(self, n, block=10):
        if n % block == 0:
            self.plan.iterate_graph(n, block)
        else:
            self.plan.iterate(n)

    def u(self):
        N, C = self.cfg.N, self.cfg.C
        return torch.stack([p.u.view(C, N, N) for p in self.problems])


def solve(cfg, geom, f, L, iters, u0=None, chunk=256):
    """Reconstructions (B, C, N, N) after ``iters`` iterations, in chunks of
    ``chunk`` problems per batch plan."""
    out = torch.empty((f.shape[0], cfg.C, cfg.N, cfg.N), device=f.device)
    for i in range(0, f.shape[0], chunk):
        sb = SliceBatch(cfg, geom, f[i:i + chunk], L,
                        None if u0 is None else u0[i:i + chunk])
        sb.iterate(iters)
        out[i:i + chunk] = sb.u()
        del sb
    return out


def train(net: UNet, X, T, steps, batch=16, lr=1e-3, lr_min=1e-5, seed=0, flips=True,
          log=None):
    """Adam on the MSE between net(X) and T (both (P, N, N, C) float32 on the
    device) with a cosine learning-rate decay.  With ``flips`` each batch is
    mirrored in x and/or y at random: for a tilt series symmetric about 0
    degrees (and N odd or even alike, the grid being centred) a mirrored
    slice's data and reconstruction are the mirrored originals, so the
    mirrors are further valid pairs.  ``log(step, loss)`` every 100 steps.
    Returns the losses logged."""
    P = X.shape[0]
    per_epoch = max(P // batch, 1)
    gen = torch.Generator()
    gen.manual_seed(seed)
    losses = []
    perm = None
    for step in range(steps):
        k = step % per_epoch
        if k == 0:
            perm = torch.randperm(P, generator=gen).to(X.device)
        idx = perm[k * batch:(k + 1) * batch]
        xb, tb = X.index_select(0, idx), T.index_select(0, idx)
        if flips:
            dims = [d for d, bit in ((1, 1), (2, 2)) if int(torch.randint(4, (1,), generator=gen)) & bit]
            if dims:
                xb, tb = xb.flip(dims), tb.flip(dims)
        net.forward(xb)
        want = step % 100 == 0 or step == steps - 1
        loss = net.backward(tb, return_loss=want)
        rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * step / max(steps, 1)))
        net.adam(lr=rate)
        if want:
            losses.append((step, loss))
            if log:
                log(step, loss)
    return losses


def predict(net: UNet, x, scale, batch=32):
    """Warm start u0 (B, C, N, N) from features (x, scale); negative values
    are clipped (the solver's u is non-negative)."""
    B, N, _, C = x.shape
    out = torch.empty((B, C, N, N), device=x.device)
    for i in range(0, B, batch):
        y = net.forward(x[i:i + batch])
        out[i:i + batch] = (y.permute(0, 3, 1, 2) * scale[i:i + batch].view(-1, C, 1, 1)).clamp(min=0)
    return out


def rel_errors(u, ustar):
    """||u - u*|| / ||u*|| per slice, over all channels."""
    return (u - ustar).flatten(1).norm(dim=1) / ustar.flatten(1).norm(dim=1).clamp(min=1e-30)


def error_curve(sb: SliceBatch, ustar, iters, check_every=50):
    """Iterate ``sb`` for ``iters`` iterations in blocks of ``check_every``;
    after each block record every slice's relative error to ``ustar`` and the
    accumulated HIP-event time of the iteration blocks (the error checks are
    not timed).  Returns (its (K,), ms (K,), errors (K, B)) with row 0 the
    start."""
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    its, ms, errs = [0], [0.0], [rel_errors(sb.u(), ustar).cpu().numpy()]
    it, t = 0, 0.0
    while it < iters:
        ev0.record()
        sb.iterate(check_every)
        ev1.record()
        ev1.synchronize()
        t += ev0.elapsed_time(ev1)
        it += check_every
        its.append(it)
        ms.append(t)
        errs.append(rel_errors(sb.u(), ustar).cpu().numpy())
    return np.asarray(its), np.asarray(ms), np.stack(errs)


def first_reach(its, errs, tol):
    """Per slice: the first recorded iteration with error <= tol (per-slice
    tol array or scalar), -1 where never."""
    tol = np.broadcast_to(np.asarray(tol, np.float64), errs.shape[1:])
    ok = errs <= tol[None]
    idx = np.where(ok.any(axis=0), ok.argmax(axis=0), -1)
    return np.where(idx >= 0, its[np.maximum(idx, 0)], -1)


def conv_flops(net: UNet, H: int, W: int) -> float:
    """Flops (2 per multiply-add) of one forward pass over an H x W image,
    counting the padded input channels and head rows the kernels compute."""
    level = {"enc0": 0, "dec0": 0, "head": 0, "enc1": 1, "dec1": 1, "enc2": 2, "dec2": 2,
             "bota": 3, "botb": 3}
    total = 0.0
    for name in net.order:
        l = net.layers[name]
        lvl = level[name[:4]]
        total += 2.0 * (H >> lvl) * (W >> lvl) * l.cin * l.cout * l.taps
    return total


def study(cfg: SliceConfig | None = None, n_train=4096, n_eval=256, target_iters=2500,
          ustar_iters=20000, train_steps=4000, batch=16, lr=1e-3, budget_iters=2500,
          max_iters=5000, check_every=50, tols=(1e-2, 3e-3, 1e-3), train_seed0=40000,
          eval_seed0=60000, chunk=256, device="cuda", log=None):
    """Train the warm-start network on solver outputs and compare warm start +
    primal-dual with plain primal-dual on unseen slices.

    Training pairs: slices seeded train_seed0 + i; targets = the solver's u
    after ``target_iters`` iterations.  Evaluation: slices seeded
    eval_seed0 + j; u* = ``ustar_iters`` iterations of the same solver (its
    last-1000-iteration relative change is reported as the convergence
    evidence).  Both runs are recorded every ``check_every`` iterations up to
    ``max_iters``; times are HIP-event times of the iteration blocks, plus the
    backprojection and network inference for the warm start.  Reported: the
    iterations and time each method needs to reach (a) plain PD's own error
    after ``budget_iters`` iterations, per slice, and (b) each tolerance in
    ``tols``."""
    cfg = cfg or SliceConfig()
    log = log or (lambda *a: None)
    dev = torch.device(device)
    N, C = cfg.N, cfg.C
    geom = cfg.geometry()
    L = float(radon_normest(None, geom, mode="power", device=dev))
    out = dict(L=L, n_train=n_train, n_eval=n_eval, target_iters=target_iters,
               ustar_iters=ustar_iters, budget_iters=budget_iters, max_iters=max_iters,
               check_every=check_every)

    # ---- training pairs from the solver ----
    t0 = time.perf_counter()
    X = torch.empty((n_train, N, N, C), device=dev)
    T = torch.empty((n_train, N, N, C), device=dev)
    for i in range(0, n_train, chunk):
        m = min(chunk, n_train - i)
        f = make_data(cfg, geom, range(train_seed0 + i, train_seed0 + i + m),
                      noise_seed=train_seed0 + 1_000_000 + i, device=dev)
        x, scale = features(f, geom)
        u = solve(cfg, geom, f, L, target_iters, chunk=chunk)
        X[i:i + m] = x
        T[i:i + m] = (u / scale.view(m, C, 1, 1)).permute(0, 2, 3, 1)
        log(f"training data {i + m}/{n_train} ({time.perf_counter() - t0:.1f} s)")
        del f, x, u
    torch.cuda.synchronize()
    out["data_s"] = time.perf_counter() - t0

    # ---- training ----
    net = UNet(cin=C, cout=C, base=32, device=dev, seed=0)
    t0 = time.perf_counter()
    losses = train(net, X, T, train_steps, batch=batch, lr=lr,
                   log=lambda s, l: log(f"step {s}/{train_steps} loss {l:.3e}"))
    torch.cuda.synchronize()
    out["train_s"] = time.perf_counter() - t0
    out["train_steps"] = train_steps
    out["train_tflops"] = 3.0 * conv_flops(net, N, N) * train_steps * batch / out["train_s"] / 1e12
    out["losses"] = losses
    del X, T

    # ---- evaluation ----
    f = make_data(cfg, geom, range(eval_seed0, eval_seed0 + n_eval),
                  noise_seed=eval_seed0 + 1_000_000, device=dev)
    t0 = time.perf_counter()
    ustar = torch.empty((n_eval, C, N, N), device=dev)
    change = []
    for i in range(0, n_eval, chunk):
        sb = SliceBatch(cfg, geom, f[i:i + chunk], L)
        sb.iterate(ustar_iters - 1000)
        prev = sb.u()
        sb.iterate(1000)
        cur = sb.u()
        change.append(rel_errors(prev, cur).max().item())
        ustar[i:i + chunk] = cur
        del sb, prev, cur
    out["ustar_s"] = time.perf_counter() - t0
    out["ustar_last1000_rel_change_max"] = max(change)
    log(f"u* done ({out['ustar_s']:.1f} s), last-1000 change {max(change):.2e}")

    sb = SliceBatch(cfg, geom, f, L)
    its, ms_p, err_p = error_curve(sb, ustar, max_iters, check_every)
    del sb
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    x, scale = features(f, geom)
    u0 = predict(net, x, scale, batch=32 if n_eval % 32 == 0 else 1)
    e1.record()
    e1.synchronize()
    infer_ms = e0.elapsed_time(e1)
    sb = SliceBatch(cfg, geom, f, L, u0=u0)
    _, ms_w, err_w = error_curve(sb, ustar, max_iters, check_every)
    del sb
    ms_w = ms_w + infer_ms
    per_iter_ms = ms_p[-1] / its[-1]

    def med(a):
        return float(np.median(a))

    k_budget = int(np.searchsorted(its, budget_iters))
    target = err_p[k_budget]                      # plain's own error at the budget
    it_p, it_w = first_reach(its, err_p, target), first_reach(its, err_w, target)
    rows = {}
    for name, tol in [("plain_at_budget", target)] + [(f"{t:g}", t) for t in tols]:
        a, b = first_reach(its, err_p, tol), first_reach(its, err_w, tol)
        both = (a >= 0) & (b >= 0)
        rows[name] = dict(
            plain_reached=int((a >= 0).sum()), warm_reached=int((b >= 0).sum()),
            plain_iters_median=med(a[a >= 0]) if (a >= 0).any() else None,
            warm_iters_median=med(b[b >= 0]) if (b >= 0).any() else None,
            iter_ratio_median=med(a[both] / np.maximum(b[both], check_every)) if both.any() else None,
            # whole batch: time until every slice has reached the target
            plain_ms_all=float(ms_p[its == a.max()][0]) if (a >= 0).all() else None,
            warm_ms_all=float(ms_w[its == b.max()][0]) if (b >= 0).all() else None)
        r = rows[name]
        r["time_ratio_all"] = (r["plain_ms_all"] / r["warm_ms_all"]
                               if r["plain_ms_all"] and r["warm_ms_all"] else None)
    out["targets"] = rows
    out["error_median"] = {int(k): (med(err_p[i]), med(err_w[i])) for i, k in enumerate(its)
                           if k in (0, 100, 250, 500, 1000, 2500, 5000, max_iters)}
    out["inference_ms"] = infer_ms
    out["inference_ms_per_slice"] = infer_ms / n_eval
    out["iteration_ms_batch"] = per_iter_ms
    out["inference_tflops"] = conv_flops(net, N, N) * n_eval / (infer_ms * 1e-3) / 1e12
    out["net"] = net
    return out


# ---------------------------------------------------------------------------
# Stages of the warm-start workflow: training (solver-generated pairs), saved
# weights, and the evaluation against a converged u*
# ---------------------------------------------------------------------------
def save_weights(net: UNet, path: str, meta: dict | None = None) -> None:
    """The bf16 weights inference runs with (as uint16 bits) and the shape."""
    bits = net.pb.detach().view(torch.int16).cpu().numpy().view(np.uint16)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, bf16=bits, cin=net.cin, cout=net.cout, base=net.base,
                        meta=np.array(json.dumps(meta or {})))


def load_weights(path: str = WEIGHTS_PATH, device="cuda") -> UNet:
    """A UNet holding saved bf16 weights (FP32 master copy = the same values).
    Loaded with numpy's default allow_pickle=False."""
    with np.load(path) as d:
        net = UNet(cin=int(d["cin"]), cout=int(d["cout"]), base=int(d["base"]), device=device)
        bits = d["bf16"]
        meta = json.loads(str(d["meta"]))
    if bits.size != net.pb.numel():
        raise ValueError(f"{path}: {bits.size} weights for a network of {net.pb.numel()}")
    w = torch.as_tensor(bits.view(np.int16), device=net.dev).view(torch.bfloat16)
    net.pb.copy_(w)
    net.p.copy_(w.float())
    net.meta = meta
    return net


def train_stage(cfg: SliceConfig | None = None, n_train=20000, target_iters=2500,
                train_steps=12000, batch=16, lr=1e-3, train_seed0=40000, chunk=256,
                device="cuda", log=None):
    """Training pairs from the solver (slices seeded train_seed0 + i, targets
    = u after ``target_iters`` iterations, divided by the data's channel
    scales) and Adam training.  Returns (net, info)."""
    cfg = cfg or SliceConfig()
    log = log or (lambda *a: None)
    dev = torch.device(device)
    N, C = cfg.N, cfg.C
    geom = cfg.geometry()
    L = float(radon_normest(None, geom, mode="power", device=dev))
    info = dict(L=L, n_train=n_train, target_iters=target_iters, train_seed0=train_seed0)
    t0 = time.perf_counter()
    X = torch.empty((n_train, N, N, C), device=dev)
    T = torch.empty((n_train, N, N, C), device=dev)
    for i in range(0, n_train, chunk):
        m = min(chunk, n_train - i)
        f = make_data(cfg, geom, range(train_seed0 + i, train_seed0 + i + m),
                      noise_seed=train_seed0 + 1_000_000 + i, device=dev)
        x, scale = features(f, geom)
        u = solve(cfg, geom, f, L, target_iters, chunk=chunk)
        X[i:i + m] = x
        T[i:i + m] = (u / scale.view(m, C, 1, 1)).permute(0, 2, 3, 1)
        log(f"training data {i + m}/{n_train} ({time.perf_counter() - t0:.1f} s)")
        del f, x, u
    torch.cuda.synchronize()
    info["data_s"] = time.perf_counter() - t0
    net = UNet(cin=C, cout=C, base=32, device=dev, seed=0)
    t0 = time.perf_counter()
    losses = train(net, X, T, train_steps, batch=batch, lr=lr,
                   log=lambda s_, l_: log(f"step {s_}/{train_steps} loss {l_:.3e}"))
    torch.cuda.synchronize()
    info["train_s"] = time.perf_counter() - t0
    info["train_steps"] = train_steps
    info["train_tflops"] = 3.0 * conv_flops(net, N, N) * train_steps * batch / info["train_s"] / 1e12
    info["final_loss"] = float(losses[-1][1]) if losses else None
    del X, T
    return net, info


def converge(sb: SliceBatch, min_iters, tol, cap, block=5000, log=None, snap_every=0,
             snap_until=0, anchor_every=25000, drift=None):
    """Iterate ``sb`` until the largest per-slice relative change of u over
    1000 iterations is <= tol (checked at the end of every ``block``
    iterations, after at least ``min_iters`` and at most about ``cap``
    iterations).  With ``snap_every`` the first ``snap_until`` iterations run
    in blocks of ``snap_every`` and every iterate between blocks is kept on
    the device, with the HIP-event time of the blocks up to it.  Iterates at
    multiples of ``anchor_every`` from min_iters / 2 on are kept as well; when
    ``drift`` is a dict it receives, per anchor iteration, the largest
    per-slice relative distance of the final u from it (a distance that stays
    at the size of the last change, instead of growing with the number of
    iterations between them, shows an iterate that fluctuates about a fixed
    point rather than drifting).  Returns (u, iterations, last change, (its,
    ms, snapshots))."""
    log = log or (lambda *a: None)
    anchors = {}
    its, ms, snaps = [0], [0.0], ([sb.u().clone()] if snap_every else [])
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    it, t, change = 0, 0.0, float("inf")
    while snap_every and it + snap_every <= snap_until:
        e0.record()
        sb.iterate(snap_every)
        e1.record()
        e1.synchronize()
        t += e0.elapsed_time(e1)
        it += snap_every
        its.append(it)
        ms.append(t)
        snaps.append(sb.u().clone())
        if it % 5000 == 0:
            log(f"plain run: {it} iterations")
    while it < cap:
        sb.iterate(max(block - 1000, 10))
        prev = sb.u().clone()
        sb.iterate(1000)
        it += max(block - 1000, 10) + 1000
        change = float(rel_errors(prev, sb.u()).max())
        log(f"u* run: {it} iterations, largest change over the last 1000 {change:.2e}")
        if it >= min_iters and change <= tol:
            break
        if anchor_every and it % anchor_every == 0 and 2 * it >= min_iters:
            anchors[it] = sb.u().clone()
    u = sb.u().clone()
    if drift is not None:
        drift.update({k: float(rel_errors(a, u).max()) for k, a in anchors.items() if k < it})
    return u, it, change, (np.asarray(its), np.asarray(ms), snaps)


def ustar_check(cfg: SliceConfig | None = None, seeds=range(60000, 60016), L=None,
                ustar_iters=300000, ratios=(0.25, 0.5, 2.0, 4.0), iters=120000, every=10000,
                device="cuda", log=None):
    """How close a plain-iteration u* is to the fixed point: u* = ``ustar_iters``
    plain iterations from zero.  A restart from u0 = u* with every dual
    variable at zero (a perfect u-only warm start) records its distance to u*
    after 1, 10, 100, 1000 and 10000 iterations.  Then, for each step ratio s, a run from zero
    with sigma = s/sqrt(Lsqr), tau = 1/(s sqrt(Lsqr)) (same fixed point),
    recording every ``every`` iterations its largest and median per-slice
    relative distance to u* and its change since the previous record, and
    at the end its FP64 energy minus u*'s, per slice."""
    cfg = cfg or SliceConfig()
    log = log or (lambda *a: None)
    dev = torch.device(device)
    geom = cfg.geometry()
    if L is None:
        L = float(radon_normest(None, geom, mode="power", device=dev))
    seeds = list(seeds)
    f = make_data(cfg, geom, seeds, noise_seed=seeds[0] + 1_000_000, device=dev)
    sb = SliceBatch(cfg, geom, f, L)
    for it in range(every, ustar_iters + 1, every):
        sb.iterate(every)
        if it % 50000 == 0:
            log(f"u* run: {it} iterations")
    ustar, e_star = sb.u().clone(), sb.energy()
    del sb
    out = dict(L=L, slices=len(seeds), seeds=[seeds[0], seeds[-1]], ustar_iters=ustar_iters,
               ustar_energy=e_star, runs={})
    # a perfect u-only warm start: u0 = u*, every dual variable at zero
    sb = SliceBatch(cfg, geom, f, L, u0=ustar)
    rows, done = [], 0
    for it in (1, 10, 100, 1000, 10000):
        sb.iterate(it - done)
        done = it
        d = rel_errors(sb.u(), ustar)
        rows.append(dict(iterations=it, max_dist=float(d.max()), median_dist=float(d.median())))
    out["restart_from_ustar"] = rows
    log("restart from u* with zero duals: " + ", ".join(
        f"{r['iterations']} it {r['median_dist']:.2e}" for r in rows))
    del sb
    for s in ratios:
        sb = SliceBatch(cfg, geom, f, L, step_ratio=s)
        rows, prev = [], None
        for it in range(every, iters + 1, every):
            sb.iterate(every)
            u = sb.u()
            d = rel_errors(u, ustar)
            rows.append(dict(iterations=it, max_dist=float(d.max()),
                             median_dist=float(d.median()),
                             max_change=None if prev is None
                             else float(rel_errors(prev, u).max())))
            prev = u.clone()
            log(f"step ratio {s:g}: {it} iterations, largest distance to u* {rows[-1]['max_dist']:.2e}")
        de = [a - b for a, b in zip(sb.energy(), e_star)]
        out["runs"][f"{s:g}"] = dict(rows=rows, energy_minus_ustar=de)
        del sb
    return out


def evaluate(net: UNet, cfg: SliceConfig | None = None, seeds=range(60000, 60016), L=None,
             tols=(1e-2, 3e-3, 1e-3), check_every=250, max_iters=60000, ustar_min=100000,
             ustar_tol=1e-5, ustar_cap=200000, device="cuda", log=None):
    """Warm start + primal-dual against plain primal-dual on the slices
    ``seeds``.  u* = the plain iteration run until its largest per-slice
    change over 1000 iterations is <= ustar_tol (at least ustar_min, at most
    ustar_cap iterations); the plain arm is that same run's iterates up to
    max_iters, the warm arm a second run from the network's u0 (checked every
    check_every iterations, stopped once every slice is within the smallest
    tolerance of u*).  Times are HIP-event times of the iteration blocks (error
    checks excluded), plus backprojection and inference for the warm arm.
    Returns a dict of per-tolerance iterations and times."""
    cfg = cfg or SliceConfig()
    log = log or (lambda *a: None)
    dev = torch.device(device)
    geom = cfg.geometry()
    if L is None:
        L = float(radon_normest(None, geom, mode="power", device=dev))
    seeds = list(seeds)
    B = len(seeds)
    f = make_data(cfg, geom, seeds, noise_seed=seeds[0] + 1_000_000, device=dev)
    out = dict(L=L, slices=B, seeds=[seeds[0], seeds[-1]], check_every=check_every,
               max_iters=max_iters)
    # u* and the plain arm
    t0 = time.perf_counter()
    sb = SliceBatch(cfg, geom, f, L)
    drift = {}
    ustar, n_star, change, (its_p, ms_p, snaps) = converge(
        sb, ustar_min, ustar_tol, ustar_cap, log=log, snap_every=check_every,
        snap_until=max_iters, drift=drift)
    del sb
    out.update(ustar_iters=n_star, ustar_last1000_rel_change_max=change,
               ustar_distance_to_earlier=drift, ustar_s=time.perf_counter() - t0)
    err_p = np.stack([rel_errors(u, ustar).cpu().numpy() for u in snaps])
    del snaps
    per_iter_ms = float(ms_p[-1] / its_p[-1])
    # warm arm
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    x, scale = features(f, geom)
    u0 = predict(net, x, scale, batch=16 if B % 16 == 0 else 1)
    e1.record()
    e1.synchronize()
    infer_ms = e0.elapsed_time(e1)
    sb = SliceBatch(cfg, geom, f, L, u0=u0)
    its_w, ms_w, errs_w = [0], [0.0], [rel_errors(sb.u(), ustar).cpu().numpy()]
    it, t = 0, 0.0
    while it < max_iters and errs_w[-1].max() > min(tols):
        e0.record()
        sb.iterate(check_every)
        e1.record()
        e1.synchronize()
        t += e0.elapsed_time(e1)
        it += check_every
        its_w.append(it)
        ms_w.append(t)
        errs_w.append(rel_errors(sb.u(), ustar).cpu().numpy())
        if it % 5000 == 0:
            log(f"warm arm: {it} iterations, largest error {errs_w[-1].max():.2e}")
    del sb
    its_w, ms_w, err_w = np.asarray(its_w), np.asarray(ms_w) + infer_ms, np.stack(errs_w)

    def med(a):
        return float(np.median(a)) if len(a) else None

    rows = {}
    for tol in tols:
        a, b = first_reach(its_p, err_p, tol), first_reach(its_w, err_w, tol)
        both = (a >= 0) & (b >= 0)
        r = dict(plain_reached=int((a >= 0).sum()), warm_reached=int((b >= 0).sum()),
                 plain_iters_median=med(a[a >= 0]), warm_iters_median=med(b[b >= 0]),
                 iter_ratio_median=med(a[both] / np.maximum(b[both], check_every)))
        r["plain_ms_all"] = float(ms_p[its_p == a.max()][0]) if (a >= 0).all() else None
        r["warm_ms_all"] = float(ms_w[its_w == b.max()][0]) if (b >= 0).all() else None
        r["time_ratio_all"] = (r["plain_ms_all"] / r["warm_ms_all"]
                               if r["plain_ms_all"] and r["warm_ms_all"] else None)
        rows[f"{tol:g}"] = r
    out["targets"] = rows
    marks = [0, 100, 500, 1000, 2500, 5000, 10000, 20000, 40000, max_iters]
    out["error_median"] = {
        int(k): (med(err_p[np.searchsorted(its_p, k)]) if k <= its_p[-1] else None,
                 med(err_w[np.searchsorted(its_w, k)]) if k <= its_w[-1] else None)
        for k in marks}
    out["inference_ms"] = infer_ms
    out["iteration_ms_batch"] = per_iter_ms
    out["per_slice"] = {"plain_iters_to_min_tol": first_reach(its_p, err_p, min(tols)).tolist(),
                        "warm_iters_to_min_tol": first_reach(its_w, err_w, min(tols)).tolist()}
    return out
