"""GPU checks of the BF16 MFMA warm-start kernels against the NumPy oracle.

Inputs and weights are rounded to bf16 before they reach either side, so the
only differences left are FP32 accumulation order (conv, weight gradient) and
bf16 rounding of stored activations/gradients.  A transposed MFMA operand, a
wrong accumulator row mapping or a wrong tap shift gives O(1) errors, far
above the tolerances used here.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import cnn_oracle as orc

pytestmark = pytest.mark.gpu


def lib():
    from graptor_amd import _lib
    return _lib.load()


def chk(rc):
    from graptor_amd import _lib
    _lib.check(rc)


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def bf16_round(a):
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).double().numpy()


def to_dev_bf16(a):
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).cuda().contiguous()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("taps,cin,cout,W,relu", [
    (9, 32, 64, 40, False),   # partial 64-pixel segment, two co blocks
    (9, 16, 32, 64, True),    # K = 16 (one k-step per tap), relu
    (1, 32, 32, 72, False),   # 1x1 head shape
    (9, 48, 96, 24, False),   # co block with only one of two wave columns live
    # the tile shapes picked by row width and output channels
    (9, 32, 128, 70, False),  # 128-pixel x 64-channel tiles, partial segment
    (9, 64, 128, 50, True),   # 64 x 128 tiles, 16-channel chunks
    (9, 16, 256, 20, False),  # 32 x 256 tiles
    (1, 64, 128, 30, False),  # 32 x 128 tiles, 1x1
])
def test_conv_forward_f32(taps, cin, cout, W, relu):
    rng = np.random.default_rng(10 + taps + cin)
    n, H = 2, 6
    x = bf16_round(rng.standard_normal((n, H, W, cin)))
    w = bf16_round(rng.standard_normal((cout, taps, cin)) / np.sqrt(taps * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    xd, wd = to_dev_bf16(x), to_dev_bf16(w)
    bd = torch.as_tensor(b).cuda()
    y = torch.full((n, H, W, cout), float("nan"), device="cuda")
    chk(lib().gr_cnn_conv_fwd(xd.data_ptr(), cin, 0, wd.data_ptr(), bd.data_ptr(), y.data_ptr(),
                              cout, 0, n, H, W, cin, cout, taps, int(relu), 1, stream()))
    ref = orc.conv(x, w, b.astype(np.float64), relu=relu)
    got = y.cpu().double().numpy()
    assert np.isfinite(got).all()
    assert rel(got, ref) < 2e-6


def test_conv_forward_bf16_channel_slices():
    """Input and output as channel slices of wider buffers (skip concat)."""
    rng = np.random.default_rng(3)
    n, H, W, cin, cout = 1, 5, 33, 32, 64
    ldx, cx0, ldy, cy0 = 48, 16, 96, 32
    xfull = bf16_round(rng.standard_normal((n, H, W, ldx)))
    w = bf16_round(rng.standard_normal((cout, 9, cin)) / 17.0)
    xd, wd = to_dev_bf16(xfull), to_dev_bf16(w)
    y = torch.full((n, H, W, ldy), 7.0, device="cuda").to(torch.bfloat16)
    chk(lib().gr_cnn_conv_fwd(xd.data_ptr(), ldx, cx0, wd.data_ptr(), None, y.data_ptr(), ldy,
                              cy0, n, H, W, cin, cout, 9, 1, 0, stream()))
    got = y.float().cpu().double().numpy()
    ref = orc.conv(xfull[..., cx0:cx0 + cin], w, None, relu=True)
    assert rel(got[..., cy0:cy0 + cout], ref) < 8e-3  # bf16 output rounding
    untouched = np.concatenate([got[..., :cy0], got[..., cy0 + cout:]], axis=-1)
    assert (untouched == 7.0).all()


def test_conv_data_gradient_via_flipped_weights():
    rng = np.random.default_rng(4)
    n, H, W, cin, cout = 2, 7, 40, 32, 64
    x = bf16_round(rng.standard_normal((n, H, W, cin)))
    w = bf16_round(rng.standard_normal((cout, 9, cin)) / 17.0)
    dy = bf16_round(rng.standard_normal((n, H, W, cout)))
    wd, dyd = to_dev_bf16(w), to_dev_bf16(dy)
    wt = torch.zeros(cout * 9 * cin, dtype=torch.bfloat16, device="cuda")
    chk(lib().gr_cnn_weight_flip(wd.data_ptr(), wt.data_ptr(), cin, cout, 9, stream()))
    dx = torch.zeros((n, H, W, cin), device="cuda")
    chk(lib().gr_cnn_conv_fwd(dyd.data_ptr(), cout, 0, wt.data_ptr(), None, dx.data_ptr(), cin, 0,
                              n, H, W, cout, cin, 9, 0, 1, stream()))
    ref_dx, _, _ = orc.conv_backward(x, w, dy)
    assert rel(dx.cpu().double().numpy(), ref_dx) < 2e-6


@pytest.mark.parametrize("taps,cin,cout,rpc", [(9, 32, 64, 3), (9, 16, 128, 1), (1, 32, 32, 5)])
def test_conv_weight_and_bias_gradient(taps, cin, cout, rpc):
    rng = np.random.default_rng(5 + cin)
    n, H, W = 2, 6, 45
    x = bf16_round(rng.standard_normal((n, H, W, cin)))
    w = np.zeros((cout, taps, cin))
    dy = bf16_round(rng.standard_normal((n, H, W, cout)))
    xd, dyd = to_dev_bf16(x), to_dev_bf16(dy)
    dw = torch.zeros(cout * taps * cin, device="cuda")
    db = torch.zeros(cout, device="cuda")
    chk(lib().gr_cnn_conv_wgrad(xd.data_ptr(), cin, 0, dyd.data_ptr(), cout, 0, dw.data_ptr(), n,
                                H, W, cin, cout, taps, rpc, stream()))
    chk(lib().gr_cnn_bias_grad(dyd.data_ptr(), cout, 0, n * H * W, cout, db.data_ptr(),
                               stream()))
    _, ref_dw, ref_db = orc.conv_backward(x, w, dy)
    assert rel(dw.cpu().double().numpy().reshape(cout, taps, cin), ref_dw) < 2e-6
    assert rel(db.cpu().double().numpy(), ref_db) < 2e-6
    with pytest.raises(ValueError):  # channel counts must divide 256
        chk(lib().gr_cnn_bias_grad(dyd.data_ptr(), cout, 0, n * H * W, 96, db.data_ptr(),
                                   stream()))


def test_pool_upsample_relu_and_pack():
    rng = np.random.default_rng(6)
    n, H, W, c, ld = 2, 8, 12, 16, 24
    x = bf16_round(rng.standard_normal((n, H, W, ld)))
    xd = to_dev_bf16(x)
    L, s = lib(), stream()
    # pool forward of channels 8..23 into a dense buffer
    y = torch.zeros((n, H // 2, W // 2, c), dtype=torch.bfloat16, device="cuda")
    chk(L.gr_cnn_pool(xd.data_ptr(), ld, 8, y.data_ptr(), c, 0, n, H, W, c, 0, 0, s))
    assert rel(y.float().cpu().numpy(), orc.pool(x[..., 8:])) < 5e-3
    # pool backward, accumulating into the fine buffer's channels 8..23
    g = bf16_round(rng.standard_normal((n, H // 2, W // 2, c)))
    acc = to_dev_bf16(x)
    gd = to_dev_bf16(g)  # keep the device copy alive until the kernel has run
    chk(L.gr_cnn_pool(gd.data_ptr(), c, 0, acc.data_ptr(), ld, 8, n, H, W, c, 1, 1, s))
    ref = x.copy()
    ref[..., 8:] += orc.pool_backward(g)
    assert rel(acc.float().cpu().numpy(), ref) < 5e-3
    # upsample forward and backward
    up = torch.zeros((n, 2 * H, 2 * W, c), dtype=torch.bfloat16, device="cuda")
    chk(L.gr_cnn_upsample(xd.data_ptr(), ld, 8, up.data_ptr(), c, 0, n, H, W, c, 0, s))
    assert rel(up.float().cpu().numpy(), orc.up(x[..., 8:])) == 0.0
    fine = bf16_round(rng.standard_normal((n, 2 * H, 2 * W, c)))
    coarse = torch.zeros((n, H, W, c), dtype=torch.bfloat16, device="cuda")
    fd = to_dev_bf16(fine)
    chk(L.gr_cnn_upsample(fd.data_ptr(), c, 0, coarse.data_ptr(), c, 0, n, H, W, c, 1, s))
    assert rel(coarse.float().cpu().numpy(), orc.up_backward(fine)) < 5e-3
    # relu backward in place on a channel slice
    dyv = to_dev_bf16(x)
    yv = to_dev_bf16(np.abs(x) * np.sign(rng.standard_normal(x.shape)))
    chk(L.gr_cnn_relu_bwd(dyv.data_ptr(), ld, 4, yv.data_ptr(), ld, 4, n * H * W, 16, s))
    ref = x.copy()
    yh = yv.float().cpu().numpy()
    ref[..., 4:20] *= yh[..., 4:20] > 0
    assert rel(dyv.float().cpu().numpy(), ref) == 0.0
    # pack float -> bf16 into channels 2..5 of a 16-channel buffer
    xf = rng.standard_normal((n * H * W, 4)).astype(np.float32)
    pk = torch.zeros((n * H * W, 16), dtype=torch.bfloat16, device="cuda")
    xfd = torch.as_tensor(xf).cuda()
    chk(L.gr_cnn_pack_bf16(xfd.data_ptr(), 4, pk.data_ptr(), 16, 2, n * H * W, 2.0, s))
    got = pk.float().cpu().numpy()
    assert rel(got[:, 2:6], 2.0 * xf) < 5e-3 and (got[:, :2] == 0).all() and (got[:, 6:] == 0).all()


def test_mse_grad_and_adam():
    rng = np.random.default_rng(7)
    npix, c, ld = 1000, 4, 32
    pred = rng.standard_normal((npix, ld)).astype(np.float32)
    tgt = rng.standard_normal((npix, c)).astype(np.float32)
    g = torch.zeros((npix, ld), dtype=torch.bfloat16, device="cuda")
    loss = torch.zeros(1, device="cuda")
    L, s = lib(), stream()
    # device copies held in names: a temporary's memory could be reused by
    # the next allocation before the kernel reads it
    pd, td = torch.as_tensor(pred).cuda(), torch.as_tensor(tgt).cuda()
    chk(L.gr_cnn_mse_grad(pd.data_ptr(), ld, td.data_ptr(), c, npix, g.data_ptr(), ld, 0.5,
                          loss.data_ptr(), s))
    d = pred[:, :c].astype(np.float64) - tgt
    assert abs(loss.item() - (d ** 2).sum()) < 1e-4 * (d ** 2).sum()
    got = g.float().cpu().numpy()
    assert rel(got[:, :c], 0.5 * d) < 5e-3 and (got[:, c:] == 0).all()
    # Adam: two steps against the textbook update in float64
    n = 777
    p0 = rng.standard_normal(n)
    grads = [rng.standard_normal(n), rng.standard_normal(n)]
    P = torch.as_tensor(p0.astype(np.float32)).cuda()
    M, V = torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    PB = torch.zeros(n, dtype=torch.bfloat16, device="cuda")
    lr, b1, b2, eps, wd = 1e-2, 0.9, 0.999, 1e-8, 1e-3
    pr, m, v = p0.astype(np.float32).astype(np.float64), np.zeros(n), np.zeros(n)
    for t, gr in enumerate(grads, 1):
        G = torch.as_tensor(gr.astype(np.float32)).cuda()
        chk(L.gr_cnn_adam(P.data_ptr(), G.data_ptr(), M.data_ptr(), V.data_ptr(), PB.data_ptr(), n,
                          lr, b1, b2, eps, t, wd, s))
        gg = gr.astype(np.float32).astype(np.float64) + wd * pr
        m = b1 * m + (1 - b1) * gg
        v = b2 * v + (1 - b2) * gg * gg
        pr = pr - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    assert rel(P.cpu().double().numpy(), pr) < 1e-6
    assert rel(PB.float().cpu().double().numpy(), pr) < 5e-3


def test_unet_forward_and_gradients_match_oracle():
    """Whole network: forward output and every layer's weight/bias gradient of
    the MSE loss against the float64 oracle fed the same bf16 weights."""
    from graptor_amd.warmstart import UNet, CIN_PAD
    torch.manual_seed(0)
    net = UNet(cin=4, cout=4, base=32, device="cuda", seed=1)
    rng = np.random.default_rng(8)
    n, H, W = 2, 32, 48
    x = rng.standard_normal((n, H, W, 4)).astype(np.float32)
    tgt = rng.standard_normal((n, H, W, 4)).astype(np.float32)
    out = net.forward(torch.as_tensor(x).cuda())
    got = out.cpu().double().numpy()
    loss = net.backward(torch.as_tensor(tgt).cuda())
    # oracle with the bf16 weights the kernels used and float64 activations
    pb = net.pb.float().cpu().double().numpy()
    pf = net.p.cpu().double().numpy()
    params = {}
    for name in net.order:
        l = net.layers[name]
        params[name] = (pb[l.w_off:l.w_off + l.nw].reshape(l.cout, l.taps, l.cin),
                        pf[l.b_off:l.b_off + l.nb])
    xin = np.zeros((n, H, W, CIN_PAD))
    xin[..., :4] = bf16_round(x)
    ref_out, cache = orc.unet_forward(xin, params)
    assert rel(got, ref_out[..., :4]) < 3e-2
    d = ref_out[..., :4] - tgt
    total = d.size
    assert abs(loss - (d ** 2).mean()) < 3e-2 * (d ** 2).mean()
    dout = np.zeros_like(ref_out)
    dout[..., :4] = 2.0 * d / total
    ref_g = orc.unet_backward(dout, params, cache)
    grads = net.grads_numpy()
    for name in net.order:
        gw, gb = grads[name]
        rw, rb = ref_g[name]
        if name == "enc0a":
            gw, rw = gw[:, :, :4], rw[:, :, :4]
        if name == "head":
            gw, rw, gb, rb = gw[:4], rw[:4], gb[:4], rb[:4]
        cos = float((gw * rw).sum() / (np.linalg.norm(gw) * np.linalg.norm(rw)))
        assert cos > 0.998, (name, cos)
        assert rel(gw, rw) < 6e-2, (name, rel(gw, rw))
        assert rel(gb, rb) < 6e-2, (name, rel(gb, rb))


def test_unet_training_reduces_loss():
    """A few Adam steps on a fixed batch lower the loss (end-to-end wiring)."""
    from graptor_amd.warmstart import UNet
    net = UNet(cin=4, cout=4, base=32, device="cuda", seed=2)
    rng = np.random.default_rng(9)
    x = torch.as_tensor(rng.standard_normal((4, 32, 32, 4)).astype(np.float32)).cuda()
    tgt = torch.nn.functional.avg_pool2d(x.permute(0, 3, 1, 2), 3, 1, 1).permute(0, 2, 3, 1)
    tgt = tgt.contiguous()
    losses = []
    for _ in range(60):
        net.forward(x)
        losses.append(net.backward(tgt))
        net.adam(lr=2e-3)
    assert losses[-1] < 0.5 * losses[0], losses[::10]
