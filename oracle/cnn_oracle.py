"""NumPy oracle of the warm-start U-Net (float64, CPU).

TEST INFRASTRUCTURE ONLY: the checker for graptor_amd's BF16-MFMA CNN
kernels (forward values and gradients).  Tensors are NHWC; convolution
weights are W[co][tap][ci] with tap = 3*ky + kx (3x3, zero padding 1) or a
single tap (1x1), as in the HIP kernels.  Pinned in tests/test_oracle_cnn.py
against torch's CPU autograd.
"""
from __future__ import annotations

import numpy as np


def _cols(x, taps):
    """im2col: (N, H, W, C) -> (N, H, W, taps, C)."""
    if taps == 1:
        return x[:, :, :, None, :]
    N, H, W, C = x.shape
    xp = np.zeros((N, H + 2, W + 2, C), dtype=x.dtype)
    xp[:, 1:-1, 1:-1] = x
    return np.stack([xp[:, ky:ky + H, kx:kx + W] for ky in range(3) for kx in range(3)], axis=3)


def conv(x, w, b=None, relu=False):
    """y[n, h, w, co] = b[co] + sum_{tap, ci} w[co, tap, ci] x[n, h+dy, w+dx, ci]."""
    co, taps, ci = w.shape
    cols = _cols(x, taps)
    y = np.einsum("nhwtc,otc->nhwo", cols, w)
    if b is not None:
        y = y + b
    return np.maximum(y, 0.0) if relu else y


def conv_backward(x, w, dy):
    """Gradients of y = conv(x, w, b) given dy: (dx, dw, db)."""
    co, taps, ci = w.shape
    cols = _cols(x, taps)
    dw = np.einsum("nhwo,nhwtc->otc", dy, cols)
    db = dy.sum(axis=(0, 1, 2))
    dcols = np.einsum("nhwo,otc->nhwtc", dy, w)
    if taps == 1:
        dx = dcols[:, :, :, 0, :]
    else:
        N, H, W, C = x.shape
        dxp = np.zeros((N, H + 2, W + 2, C))
        t = 0
        for ky in range(3):
            for kx in range(3):
                dxp[:, ky:ky + H, kx:kx + W] += dcols[:, :, :, t]
                t += 1
        dx = dxp[:, 1:-1, 1:-1]
    return dx, dw, db


def pool(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_backward(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_backward(dy):
    N, H, W, C = dy.shape
    return dy.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


# U-Net topology: (name, input, cin, cout, taps, relu)
def layer_specs(cin=16, base=32, cout=32):
    c0, c1, c2, c3 = base, 2 * base, 4 * base, 8 * base
    return [
        ("enc0a", cin, c0, 9, True), ("enc0b", c0, c0, 9, True),
        ("enc1a", c0, c1, 9, True), ("enc1b", c1, c1, 9, True),
        ("enc2a", c1, c2, 9, True), ("enc2b", c2, c2, 9, True),
        ("bota", c2, c3, 9, True), ("botb", c3, c3, 9, True),
        ("dec2a", c2 + c3, c2, 9, True), ("dec2b", c2, c2, 9, True),
        ("dec1a", c1 + c2, c1, 9, True), ("dec1b", c1, c1, 9, True),
        ("dec0a", c0 + c1, c0, 9, True), ("dec0b", c0, c0, 9, True),
        ("head", c0, cout, 1, False),
    ]


def unet_forward(x, params, base=32):
    """x (N, H, W, 16); params[name] = (w, b).  Returns (out, cache)."""
    P = params
    a = {}
    a["enc0a"] = conv(x, *P["enc0a"], relu=True)
    a["enc0b"] = conv(a["enc0a"], *P["enc0b"], relu=True)
    a["p0"] = pool(a["enc0b"])
    a["enc1a"] = conv(a["p0"], *P["enc1a"], relu=True)
    a["enc1b"] = conv(a["enc1a"], *P["enc1b"], relu=True)
    a["p1"] = pool(a["enc1b"])
    a["enc2a"] = conv(a["p1"], *P["enc2a"], relu=True)
    a["enc2b"] = conv(a["enc2a"], *P["enc2b"], relu=True)
    a["p2"] = pool(a["enc2b"])
    a["bota"] = conv(a["p2"], *P["bota"], relu=True)
    a["botb"] = conv(a["bota"], *P["botb"], relu=True)
    a["cat2"] = np.concatenate([a["enc2b"], up(a["botb"])], axis=3)
    a["dec2a"] = conv(a["cat2"], *P["dec2a"], relu=True)
    a["dec2b"] = conv(a["dec2a"], *P["dec2b"], relu=True)
    a["cat1"] = np.concatenate([a["enc1b"], up(a["dec2b"])], axis=3)
    a["dec1a"] = conv(a["cat1"], *P["dec1a"], relu=True)
    a["dec1b"] = conv(a["dec1a"], *P["dec1b"], relu=True)
    a["cat0"] = np.concatenate([a["enc0b"], up(a["dec1b"])], axis=3)
    a["dec0a"] = conv(a["cat0"], *P["dec0a"], relu=True)
    a["dec0b"] = conv(a["dec0a"], *P["dec0b"], relu=True)
    out = conv(a["dec0b"], *P["head"], relu=False)
    a["x"] = x
    return out, a


def unet_backward(dout, params, a, base=32):
    """Gradients of every (w, b) given d(out); returns {name: (dw, db)}."""
    P = params
    g = {}
    c0, c1, c2 = base, 2 * base, 4 * base

    def back(name, xin, dy, relu_out=None):
        if relu_out is not None:
            dy = dy * (relu_out > 0)
        dx, dw, db = conv_backward(xin, P[name][0], dy)
        g[name] = (dw, db)
        return dx

    d = back("head", a["dec0b"], dout)
    d = back("dec0b", a["dec0a"], d, a["dec0b"])
    dcat0 = back("dec0a", a["cat0"], d, a["dec0a"])
    d = up_backward(dcat0[..., c0:])
    d = back("dec1b", a["dec1a"], d, a["dec1b"])
    dcat1 = back("dec1a", a["cat1"], d, a["dec1a"])
    d = up_backward(dcat1[..., c1:])
    d = back("dec2b", a["dec2a"], d, a["dec2b"])
    dcat2 = back("dec2a", a["cat2"], d, a["dec2a"])
    d = up_backward(dcat2[..., c2:])
    d = back("botb", a["bota"], d, a["botb"])
    d = back("bota", a["p2"], d, a["bota"])
    d = dcat2[..., :c2] + pool_backward(d)
    d = back("enc2b", a["enc2a"], d, a["enc2b"])
    d = back("enc2a", a["p1"], d, a["enc2a"])
    d = dcat1[..., :c1] + pool_backward(d)
    d = back("enc1b", a["enc1a"], d, a["enc1b"])
    d = back("enc1a", a["p0"], d, a["enc1a"])
    d = dcat0[..., :c0] + pool_backward(d)
    d = back("enc0b", a["enc0a"], d, a["enc0b"])
    back("enc0a", a["x"], d, a["enc0a"])
    return g
