"""Pins the CPU oracle of the learned warm-start network (oracle/warm_net.py):
the convolution against a direct loop sum, pooling, upsampling and the
convolution's input gradient as adjoints of the forward maps, the block planes
against a loop over each block's cells, the bilinear resampling against
torch's, the whole backward pass against central finite differences of the
loss, and the bf16 rounding against its definition."""
import numpy as np
import pytest

from oracle import warm_net as wn


def test_parameter_count():
    assert wn.nparams() == 255489
    p = wn.init(1)
    assert p.size == wn.nparams()
    assert np.array_equal(wn.flatten(wn.unflatten(p)), p)


def test_conv_is_the_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 7, 3))
    W = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    want = np.zeros((2, 5, 7, 4))
    for n in range(2):
        for y in range(5):
            for xx in range(7):
                for co in range(4):
                    s = b[co]
                    for ci in range(3):
                        for ky in range(3):
                            for kx in range(3):
                                yy, xs = y + ky - 1, xx + kx - 1
                                if 0 <= yy < 5 and 0 <= xs < 7:
                                    s += x[n, yy, xs, ci] * W[co, ci, ky, kx]
                    want[n, y, xx, co] = s
    assert np.allclose(wn.conv(x, W, b), want, atol=1e-12)


@pytest.mark.parametrize("H,W", [(6, 8), (7, 5), (1, 3)])
def test_adjoints(H, W):
    rng = np.random.default_rng(H * 10 + W)
    x = rng.standard_normal((2, H, W, 3))
    h, w = (H + 1) // 2, (W + 1) // 2
    y = rng.standard_normal((2, h, w, 3))
    assert np.isclose(np.sum(wn.pool(x) * y), np.sum(x * wn.pool_bwd(y, x.shape)))
    assert np.isclose(np.sum(wn.up(y, (H, W)) * x), np.sum(y * wn.up_bwd(x, (h, w))))
    K = rng.standard_normal((4, 3, 3, 3))
    z = rng.standard_normal((2, H, W, 4))
    assert np.isclose(np.sum(wn.conv(x, K) * z), np.sum(x * wn.conv_dx(z, K)))
    # pooling averages: a constant stays constant on every cell, partial ones included
    assert np.allclose(wn.pool(np.ones((1, H, W, 2))), 1.0)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    B, H, W = 2, 41, 75  # block grid 6 x 10, partial blocks on both edges
    d = rng.uniform(0, 7, (B, H, W)).astype(np.float32)
    conf = rng.uniform(-0.1, 1, (B, H, W)).astype(np.float32)
    g = rng.uniform(0, 1, (B, H, W)).astype(np.float32)
    t = (d + rng.normal(0, 0.5, (B, H, W))).astype(np.float32)
    p = wn.init(3)
    value, grad = wn.loss_and_grads(p, d, conf, g, t, 7.0, sigma=0.2)
    f = lambda q: wn.loss_and_grads(q, d, conf, g, t, 7.0, sigma=0.2)[0]
    offs = np.cumsum([0] + [ci * co * 9 + co for ci, co, _ in wn.SPECS])
    idx = []
    for li in range(len(wn.SPECS)):  # two weights and a bias of every layer
        ci, co, _ = wn.SPECS[li]
        idx += [int(offs[li] + rng.integers(0, ci * co * 9)), int(offs[li] + rng.integers(0, ci * co * 9)),
                int(offs[li] + ci * co * 9 + rng.integers(0, co))]
    eps = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = eps
        num = (f(p + e) - f(p - e)) / (2 * eps)
        assert abs(num - grad[i]) <= 1e-6 * max(1.0, abs(num)) + 1e-9, (i, num, grad[i])
    assert value > 0


def test_loss_is_relative_per_system():
    """Scaling one system's target deviation and the prediction's error with it
    leaves the loss unchanged; a zero network predicts the weighted means, so
    its loss is 1 for every system above the floor."""
    rng = np.random.default_rng(6)
    d = rng.uniform(0, 7, (2, 24, 32)).astype(np.float32)
    conf = rng.uniform(0, 1, (2, 24, 32)).astype(np.float32)
    g = rng.uniform(0, 1, (2, 24, 32)).astype(np.float32)
    t = (d + rng.normal(0, 1.0, d.shape)).astype(np.float32)
    p = np.zeros(wn.nparams())
    loss, _ = wn.loss_and_grads(p, d, conf, g, t, 7.0)
    assert loss == pytest.approx(1.0, rel=1e-12)


def test_block_planes_are_the_block_loops():
    rng = np.random.default_rng(9)
    H, W, sig, ds = 19, 26, 0.2, 5.0  # blocks of 8 x 8, 8 x 2 and 3 x 8 cells at the edges
    d = rng.uniform(0, 5, (H, W)).astype(np.float32)
    conf = rng.uniform(-0.3, 1, (H, W)).astype(np.float32)
    g = rng.uniform(0, 1, (H, W)).astype(np.float32)
    t = rng.uniform(0, 5, (H, W)).astype(np.float32)
    X, dm, tc = wn.features(d, conf, g, ds, sig, t=t)
    assert X.shape == (1, 3, 4, 10)
    for by in range(3):
        for bx in range(4):
            sw = swd = sd = sax = say = st = 0.0
            n = 0
            for y in range(8 * by, min(8 * by + 8, H)):
                for x in range(8 * bx, min(8 * bx + 8, W)):
                    w = max(float(conf[y, x]), 0.0) + 1e-4
                    sw += w
                    swd += w * float(d[y, x])
                    sd += float(d[y, x])
                    st += float(t[y, x])
                    if x + 1 < W:
                        sax += np.exp(-abs(float(g[y, x]) - float(g[y, x + 1])) / sig)
                    if y + 1 < H:
                        say += np.exp(-abs(float(g[y, x]) - float(g[y + 1, x])) / sig)
                    n += 1
            want = [(sd / n / ds - swd / sw / ds) * 32, sw / n, np.log10(sw / n) / 4 + 1, sax / n, say / n]
            assert np.allclose(X[0, by, bx, [0, 4, 5, 6, 7]], want, rtol=1e-9, atol=1e-12)
            assert dm[0, by, bx] == pytest.approx(swd / sw / ds, rel=1e-12)
            assert tc[0, by, bx] == pytest.approx(st / n / ds - swd / sw / ds, rel=1e-9, abs=1e-12)
    # the push-pull levels over the 3 x 4 block grid: one cell of each holds
    # every block, so each level's mean is the constant weighted mean m
    pw = X[0, :, :, 4]
    m = np.sum(pw * dm[0]) / np.sum(pw)
    for c in (1, 2, 3):
        assert np.allclose(X[0, :, :, c], (m - dm[0]) * 32, rtol=1e-9, atol=1e-12)
    for c in (8, 9):
        assert np.allclose(X[0, :, :, c], np.log10(pw.mean()) / 4 + 1, rtol=1e-12)
    # depth enters only as differences: a constant added to it changes no input
    X2, dm2, tc2 = wn.features(d + 3.0, conf, g, ds, sig, t=t + 3.0)
    assert np.allclose(X2, X, rtol=1e-6, atol=1e-6) and np.allclose(tc2, tc, atol=1e-9)


@pytest.mark.parametrize("ih,iw,oh,ow", [(3, 4, 19, 26), (5, 9, 40, 72), (1, 1, 3, 5), (7, 3, 7, 3), (6, 10, 3, 4)])
def test_resample_is_torch_bilinear(ih, iw, oh, ow):
    import torch
    import torch.nn.functional as F

    rng = np.random.default_rng(ih * 100 + iw)
    p = rng.standard_normal((2, ih, iw))
    want = F.interpolate(torch.from_numpy(p)[:, None], size=(oh, ow), mode="bilinear", align_corners=False)
    assert np.allclose(wn.resample(p, oh, ow), want[:, 0].numpy(), rtol=1e-12, atol=1e-12)


def test_prediction_of_a_zero_network_is_the_prolonged_weighted_mean():
    rng = np.random.default_rng(4)
    d = rng.uniform(0, 7, (1, 30, 45)).astype(np.float32)
    conf = np.ones_like(d)
    g = rng.uniform(0, 1, d.shape).astype(np.float32)
    u = wn.predict(np.zeros(wn.nparams()), d, conf, g, 7.0)
    assert u.shape == d.shape
    # equal weights: the block means of d, resampled
    assert np.allclose(u, wn.resample(wn.block_means(d, 8), 30, 45), rtol=1e-12)
    # a constant field stays constant
    u = wn.predict(np.zeros(wn.nparams()), np.full_like(d, 3.0), conf, g, 7.0)
    assert np.allclose(u, 3.0, rtol=1e-12)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-30, 65504.0, 1.0 + 2 ** -9], np.float64)
    # bf16 keeps 8 significant bits: 1 + 2^-8 is a tie (to even: 1), 1 + 3*2^-8 rounds up to 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, None, 65536.0, 1.0], object)
    got = wn.bf16(vals)
    for gv, wv in zip(got, want):
        if wv is not None:
            assert gv == wv
    assert abs(got[4] - 1e-30) <= 1e-30 * 2 ** -8
