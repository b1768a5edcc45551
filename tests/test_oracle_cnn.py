"""Pins the NumPy U-Net oracle against torch's CPU autograd (float64)."""
import numpy as np
import pytest

from oracle import cnn_oracle as C

torch = pytest.importorskip("torch")


def _torch_conv(x, w, b, taps):
    # NHWC / W[co][tap][ci] -> NCHW / OIHW
    xt = torch.tensor(x.transpose(0, 3, 1, 2), requires_grad=True)
    k = 3 if taps == 9 else 1
    wt = torch.tensor(w.reshape(w.shape[0], k, k, w.shape[2]).transpose(0, 3, 1, 2),
                      requires_grad=True)
    bt = torch.tensor(b, requires_grad=True)
    y = torch.nn.functional.conv2d(xt, wt, bt, padding=k // 2)
    return xt, wt, bt, y


@pytest.mark.parametrize("taps", [9, 1])
def test_conv_and_grads_match_torch(taps):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 6, 5, 4))
    w = rng.standard_normal((3, taps, 4))
    b = rng.standard_normal(3)
    xt, wt, bt, y = _torch_conv(x, w, b, taps)
    np.testing.assert_allclose(C.conv(x, w, b), y.detach().numpy().transpose(0, 2, 3, 1),
                               rtol=1e-12, atol=1e-12)
    dy = rng.standard_normal(y.shape)
    y.backward(torch.tensor(dy))
    dx, dw, db = C.conv_backward(x, w, dy.transpose(0, 2, 3, 1))
    k = 3 if taps == 9 else 1
    np.testing.assert_allclose(dx, xt.grad.numpy().transpose(0, 2, 3, 1), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(dw, wt.grad.numpy().transpose(0, 2, 3, 1).reshape(3, taps, 4),
                               rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-12)


def test_pool_up_adjoint_pairs():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 8, 6, 3))
    y = rng.standard_normal((1, 4, 3, 3))
    assert abs(np.vdot(C.pool(x), y) - np.vdot(x, C.pool_backward(y))) < 1e-12
    assert abs(np.vdot(C.up(y), x) - np.vdot(y, C.up_backward(x))) < 1e-12


def test_unet_gradients_match_torch():
    """Full U-Net backward against torch autograd on a tiny input."""
    rng = np.random.default_rng(2)
    base = 4
    specs = C.layer_specs(cin=3, base=base, cout=2)
    params = {n: (rng.standard_normal((co, t, ci)) * np.sqrt(2.0 / (t * ci)),
                  rng.standard_normal(co) * 0.1) for n, ci, co, t, _ in specs}
    x = rng.standard_normal((1, 16, 16, 3))
    out, cache = C.unet_forward(x, params, base)
    dout = rng.standard_normal(out.shape)
    g = C.unet_backward(dout, params, cache, base)
    # same network in torch
    tp = {n: [torch.tensor(w, requires_grad=True), torch.tensor(b, requires_grad=True)]
          for n, (w, b) in params.items()}

    def tconv(name, xin, relu):
        w, b = tp[name]
        co, t, ci = w.shape
        k = 3 if t == 9 else 1
        wt = w.reshape(co, k, k, ci).permute(0, 3, 1, 2)
        y = torch.nn.functional.conv2d(xin, wt, b, padding=k // 2)
        return torch.relu(y) if relu else y

    F = torch.nn.functional
    xt = torch.tensor(x.transpose(0, 3, 1, 2))
    e0 = tconv("enc0b", tconv("enc0a", xt, True), True)
    e1 = tconv("enc1b", tconv("enc1a", F.avg_pool2d(e0, 2), True), True)
    e2 = tconv("enc2b", tconv("enc2a", F.avg_pool2d(e1, 2), True), True)
    bt = tconv("botb", tconv("bota", F.avg_pool2d(e2, 2), True), True)
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    d2 = tconv("dec2b", tconv("dec2a", torch.cat([e2, up(bt)], 1), True), True)
    d1 = tconv("dec1b", tconv("dec1a", torch.cat([e1, up(d2)], 1), True), True)
    d0 = tconv("dec0b", tconv("dec0a", torch.cat([e0, up(d1)], 1), True), True)
    o = tconv("head", d0, False)
    np.testing.assert_allclose(out, o.detach().numpy().transpose(0, 2, 3, 1), rtol=1e-9,
                               atol=1e-9)
    o.backward(torch.tensor(dout.transpose(0, 3, 1, 2)))
    for n in params:
        np.testing.assert_allclose(g[n][0], tp[n][0].grad.numpy(), rtol=1e-8, atol=1e-9)
        np.testing.assert_allclose(g[n][1], tp[n][1].grad.numpy(), rtol=1e-8, atol=1e-9)
