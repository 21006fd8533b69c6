"""Pins the CPU oracle of the learned warm-start network (oracle/warm_net.py):
the convolution against a direct loop sum, pooling, upsampling and the
convolution's input gradient as adjoints of the forward maps, the whole
backward pass against central finite differences of the loss, and the bf16
rounding against its definition."""
import numpy as np
import pytest

from oracle import warm_net as wn


def test_parameter_count():
    assert wn.nparams() == 254481
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


@pytest.mark.parametrize("loss", ["mse", "energy"])
def test_backward_matches_finite_differences(loss):
    rng = np.random.default_rng(5)
    B, H, W = 1, 10, 12
    d = rng.uniform(0, 7, (B, H, W)).astype(np.float32)
    conf = rng.uniform(0, 1, (B, H, W)).astype(np.float32)
    g = rng.uniform(0, 1, (B, H, W)).astype(np.float32)
    t = (d + rng.normal(0, 0.5, (B, H, W))).astype(np.float32)
    p = wn.init(3)
    kw = dict(loss=loss, lam=3.0, sigma=0.2)
    value, grad = wn.loss_and_grads(p, d, conf, g, t, 7.0, **kw)
    f = lambda q: wn.loss_and_grads(q, d, conf, g, t, 7.0, **kw)[0]
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


def test_energy_is_the_quadratic_form_of_the_wls_matrix():
    from tests.fixtures import wls_matrix

    rng = np.random.default_rng(8)
    B, H, W = 2, 7, 9
    e = rng.standard_normal((B, H, W))
    conf = rng.uniform(-0.2, 1.0, (B, H, W)).astype(np.float32)
    g = rng.uniform(0, 1, (B, H, W)).astype(np.float32)
    loss, grad = wn.energy(e, conf, g, 3.0, 0.2)
    want, wgrad = 0.0, []
    for b in range(B):
        A = wls_matrix(g[b], conf[b], 3.0, 0.2)
        v = e[b].ravel()
        want += v @ (A @ v)
        wgrad.append((2.0 * (A @ v) / e.size).reshape(H, W))
    assert loss == pytest.approx(want / e.size, rel=1e-6)
    # the fixture builds the edge weights in f32 (as the solver does), the oracle in f64
    assert np.allclose(grad, np.stack(wgrad), rtol=1e-6, atol=1e-7)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-30, 65504.0, 1.0 + 2 ** -9], np.float64)
    # bf16 keeps 8 significant bits: 1 + 2^-8 is a tie (to even: 1), 1 + 3*2^-8 rounds up to 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, None, 65536.0, 1.0], object)
    got = wn.bf16(vals)
    for gv, wv in zip(got, want):
        if wv is not None:
            assert gv == wv
    assert abs(got[4] - 1e-30) <= 1e-30 * 2 ** -8
