"""Pins for oracle functions the reference's own tests leave open, each by what
the mathematics fixes, chosen so that a plausible slip fails:

* upsample_bilinear (depth.rs:199-221): a constant stays; a ramp on the grid
  comes back at block centres, (x + 1/2)/k - 1/2 (a half-pixel slip moves it
  by half a block); borders clamp; k = 1 is the identity;
* the Huber IRLS step (depth.rs:1166-1172): the weights at closed-form points,
  and the depth pass's second solve equal to a direct f64 solve of the
  reweighted system built here from those closed forms (an inverted factor
  gives another system);
* SML (depth.rs:481-492) against a numpy sum-modified-Laplacian with
  reflect-101 borders, bit for bit;
* guided upsampling at k > 1 (depth.rs:1182-1191): the depth pass's output
  against the composition a(x) I(x) + b(x) of the guided coefficients, each
  pinned on its own against the textbook box means;
* the rotation's sin and cos (align.rs:243): correctly rounded against exact
  rational Taylor sums, and within one ulp of the C library's."""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse.linalg

from fixtures import wls_matrix
from oracle import ref


# ---------------------------------------------------------------- upsample_bilinear
def test_upsample_keeps_a_constant():
    for (dw, dh, k) in [(5, 4, 2), (7, 3, 4), (1, 1, 2)]:
        g = np.full((dh, dw), 0.375, np.float32)
        out = ref.upsample_bilinear(g, dw * k - 1, dh * k, k)
        assert np.all(out == np.float32(0.375))


@pytest.mark.parametrize("k", [2, 4])
def test_upsample_samples_a_ramp_at_block_centres(k):
    dw, dh = 9, 6
    w, h = dw * k, dh * k
    gx = np.arange(dw, dtype=np.float32) * np.float32(0.25) + np.float32(1.0)
    gy = np.arange(dh, dtype=np.float32) * np.float32(-0.5)
    g = (gx[None, :] + gy[:, None]).astype(np.float32)
    out = ref.upsample_bilinear(g, w, h, k).astype(np.float64)
    fx = np.clip((np.arange(w) + 0.5) / k - 0.5, 0, dw - 1)
    fy = np.clip((np.arange(h) + 0.5) / k - 0.5, 0, dh - 1)
    want = (1.0 + 0.25 * fx)[None, :] + (-0.5 * fy)[:, None]
    assert np.abs(out - want).max() < 1e-5
    # the first k/2 columns and rows sit at or before the first block centre: clamped to it
    assert np.all(out[:, : k // 2] == out[:, :1])
    assert np.all(out[: k // 2, :] == out[:1, :])
    # a half-pixel slip (x / k instead of (x + 1/2) / k - 1/2) would move every interior value by a quarter step
    slip = (1.0 + 0.25 * np.clip(np.arange(w) / k, 0, dw - 1))[None, :] + want[:, :1] - 1.0
    assert np.abs(out - slip).max() > 0.05


def test_upsample_at_k1_is_the_identity():
    g = np.random.default_rng(3).random((13, 17)).astype(np.float32)
    assert np.array_equal(ref.upsample_bilinear(g, 17, 13, 1), g)


# ---------------------------------------------------------------- Huber IRLS step
def test_huber_weights_at_closed_form_points():
    conf = np.array([0.8, 0.8, 0.8, 0.8, 0.0, 1.5], np.float32)
    d = np.array([1.0, 1.0, 1.0, 1.0, 2.0, 0.0], np.float32)
    u = np.array([1.0, 1.5, 3.0, -1.0, 0.0, 0.25], np.float32)
    got = ref.huber_weights(conf, d, u, 1.0)
    # |d - u| = 0 (min(inf, 1) = 1), 0.5 (inside tau), 2 (tau / 2), 2, any (conf 0), 0.25
    want = np.array([0.8, 0.8, 0.4, 0.4, 0.0, 1.5], np.float32)
    assert np.array_equal(got, want)
    assert np.array_equal(ref.huber_weights(conf, d, u, 0.5), np.array([0.8, 0.8, 0.2, 0.2, 0.0, 1.5], np.float32))


def _depth_case(seed, n=5, w=96, h=72):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    peak = (n - 1) * (0.5 + 0.4 * np.sin(xx / 11.0) * np.cos(yy / 9.0))
    slices = np.stack([np.exp(-0.5 * (i - peak) ** 2) + 0.02 * rng.random((h, w)) for i in range(n)]).astype(
        np.float32)
    # a few cells whose peak lies far from their neighbours': the Huber step's business
    slices[:, 30:33, 40:43] = np.stack([np.full((3, 3), 1.0 if i == 0 else 0.01) for i in range(n)])
    fused = np.stack([0.5 + 0.3 * np.sin(xx / 7.0 + c) for c in range(3)]).astype(np.float32)
    return slices, fused


def test_depth_pass_second_solve_is_the_reweighted_system():
    """scale 0 and bilinear output (k = 1: the identity), so the depth pass's
    output is its WLS solution clamped; both solves to 1e-9 here, so each is the
    direct solve of its system."""
    slices, fused = _depth_case(5)
    n = slices.shape[0]
    lam, sig, tau = 2.0, 0.04, 0.05
    p = ref.DepthParams.make(scale=0, agg_radius=2, lam=lam, sigma_c=sig, robust=tau, upsample="bilinear",
                             tol=1e-9, cg_iters=2000)
    g, d, c, _ = ref.depth_grid(list(slices), fused, p)
    h, w = d.shape
    A1 = wls_matrix(g, c, lam, sig)
    u1 = scipy.sparse.linalg.spsolve(A1.tocsc(), (np.maximum(c, 0) + 1e-4).astype(np.float64).ravel()
                                     * d.astype(np.float64).ravel()).reshape(h, w)
    tau32 = np.float32(tau)
    w2 = (c * np.minimum(tau32 / np.abs(d - u1.astype(np.float32)), np.float32(1.0))).astype(np.float32)
    A2 = wls_matrix(g, w2, lam, sig)
    u2 = scipy.sparse.linalg.spsolve(A2.tocsc(), (np.maximum(w2, 0) + 1e-4).astype(np.float64).ravel()
                                     * d.astype(np.float64).ravel()).reshape(h, w)
    depth, conf, _, info = ref.depth_from_slices(list(slices), fused, p)
    want = np.clip(u2, 0, n - 1)
    assert np.abs(depth - want).max() < 1e-4 * (n - 1)
    assert info.iters2 > 0
    # the reweighting matters here, and an inverted factor min(|d - u| / tau, 1) is another answer
    assert np.abs(np.clip(u1, 0, n - 1) - want).max() > 1e-2
    inv = (c * np.minimum(np.abs(d - u1.astype(np.float32)) / tau32, np.float32(1.0))).astype(np.float32)
    A3 = wls_matrix(g, inv, lam, sig)
    u3 = scipy.sparse.linalg.spsolve(A3.tocsc(), (np.maximum(inv, 0) + 1e-4).astype(np.float64).ravel()
                                     * d.astype(np.float64).ravel()).reshape(h, w)
    assert np.abs(np.clip(u3, 0, n - 1) - want).max() > 1e-2


# ---------------------------------------------------------------- SML
def _reflect(i, n):
    if n == 1:
        return np.zeros_like(i)
    p = 2 * (n - 1)
    i = np.abs(i) % p
    return np.where(i >= n, p - i, i)


@pytest.mark.parametrize("w,h,step", [(40, 31, 1), (40, 31, 2), (9, 7, 3), (3, 2, 1)])
def test_sml_matches_a_numpy_sum_modified_laplacian(w, h, step):
    y = np.random.default_rng(w + step).random((h, w)).astype(np.float32)
    got = ref.focus_measure(y, focus="sml", sml_step=step)
    xs, ys = np.arange(w), np.arange(h)
    l, r = y[:, _reflect(xs - step, w)], y[:, _reflect(xs + step, w)]
    u, d = y[_reflect(ys - step, h), :], y[_reflect(ys + step, h), :]
    two = np.float32(2.0)
    want = np.abs(two * y - l - r) + np.abs(two * y - u - d)
    assert got.dtype == np.float32 and np.array_equal(got, want.astype(np.float32))
    # a one-sided difference (|c - l| + |c - u|) is another measure
    assert not np.allclose(got, np.abs(y - l) + np.abs(y - u))


# ---------------------------------------------------------------- guided upsampling, k > 1
def _box_mean_naive(a, r):
    h, w = a.shape
    out = np.zeros_like(a, dtype=np.float64)
    for y in range(h):
        for x in range(w):
            out[y, x] = a[max(0, y - r):y + r + 1, max(0, x - r):x + r + 1].astype(np.float64).mean()
    return out


def test_guided_coefficients_match_the_textbook_box_means():
    rng = np.random.default_rng(8)
    I = rng.random((19, 23)).astype(np.float32)
    p = (0.6 * I + 0.2 * rng.random(I.shape)).astype(np.float32)
    r, eps = 2, 1e-2
    a, b = ref.guided_coeffs(I, p, r, eps)
    mI, mp = _box_mean_naive(I, r), _box_mean_naive(p, r)
    cov = _box_mean_naive(I.astype(np.float64) * p, r) - mI * mp
    var = _box_mean_naive(I.astype(np.float64) * I, r) - mI * mI
    a0 = cov / (var + eps)
    b0 = mp - a0 * mI
    # the coefficients of the reference are the box means of the per-window a and b
    assert np.abs(a - _box_mean_naive(a0, r)).max() < 1e-4
    assert np.abs(b - _box_mean_naive(b0, r)).max() < 1e-4


@pytest.mark.parametrize("scale", [1, 2])
def test_guided_upsampling_is_the_coefficients_times_the_full_luma(scale):
    slices, fused = _depth_case(11, w=101, h=77)
    n = slices.shape[0]
    k = 1 << scale
    sl = np.stack([ref.block_mean(s, k) for s in slices])
    p = ref.DepthParams.make(scale=scale, agg_radius=2, lam=0.0, upsample="guided", up_radius=2, up_eps=1e-2)
    g, d, c, _ = ref.depth_grid(list(sl), fused, p)
    depth, conf, _, _ = ref.depth_from_slices(list(sl), fused, p)
    _, h, w = fused.shape
    a, b = ref.guided_coeffs(g, d, 2, 1e-2)
    au, bu = ref.upsample_bilinear(a, w, h, k), ref.upsample_bilinear(b, w, h, k)
    y = ref.luma(fused)
    want = np.clip(au * y + bu, 0, n - 1)
    assert np.array_equal(depth, want.astype(np.float32))
    assert np.array_equal(conf, ref.upsample_bilinear(c, w, h, k))
    # the grid's guide in place of the full-resolution luma is another answer
    assert np.abs(np.clip(au * ref.upsample_bilinear(g, w, h, k) + bu, 0, n - 1) - depth).max() > 1e-3


# ---------------------------------------------------------------- sin and cos of the rotation
def _exact(x, odd):
    x = Fraction(x)
    t = x if odd else Fraction(1)
    s = t
    for k in range(1, 30):
        den = (2 * k) * (2 * k + 1) if odd else (2 * k - 1) * (2 * k)
        t = -t * x * x / den
        s += t
    return s


def _nearest(fr):
    v = float(fr)
    cands = [np.nextafter(v, -np.inf), v, np.nextafter(v, np.inf)]
    return min(cands, key=lambda c: abs(Fraction(float(c)) - fr))


def test_sincos_is_correctly_rounded():
    rng = np.random.default_rng(17)
    xs = np.concatenate([rng.uniform(-0.785, 0.785, 150), np.sign(rng.uniform(-1, 1, 150))
                         * np.exp(rng.uniform(math.log(1e-10), math.log(0.785), 150)), [0.0, 1e-300]])
    for x in xs:
        s, c = ref.sincos(float(x))
        assert s == _nearest(_exact(float(x), True)), x
        assert c == _nearest(_exact(float(x), False)), x


def test_sincos_is_within_an_ulp_of_the_c_library():
    libm = C.CDLL("libm.so.6")
    libm.sin.restype = libm.cos.restype = C.c_double
    libm.sin.argtypes = libm.cos.argtypes = [C.c_double]
    rng = np.random.default_rng(19)
    differ = 0
    for x in rng.uniform(-0.3, 0.3, 20000):
        s, c = ref.sincos(float(x))
        ls, lc = libm.sin(float(x)), libm.cos(float(x))
        assert abs(s - ls) <= np.spacing(abs(s)) and abs(c - lc) <= np.spacing(abs(c))
        differ += (s != ls) + (c != lc)
    assert differ < 0.02 * 2 * 20000
