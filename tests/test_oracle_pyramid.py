"""The oracle's pyramid.rs restatement, pinned by the reference's own tests
(pyramid.rs:252-301) and by dense-matrix brute force."""
import numpy as np
import pytest

from oracle import ref
from tests.fixtures import noise_img


def test_reflect_101():  # pyramid.rs:253-260
    assert ref.reflect(-1, 5) == 1
    assert ref.reflect(-2, 5) == 2
    assert ref.reflect(5, 5) == 3
    assert ref.reflect(6, 5) == 2
    assert ref.reflect(3, 2) == 1
    assert ref.reflect(-3, 1) == 0


def test_reduce_preserves_constant():  # pyramid.rs:271-279
    im = np.full((22, 37), 0.7, np.float32)
    r = ref.reduce(im)
    assert r.shape == (11, 19)
    assert np.all(np.abs(r - 0.7) < 1e-5)
    e = ref.expand(r, 37, 22)
    assert np.all(np.abs(e - 0.7) < 1e-5)


@pytest.mark.parametrize("w,h", [(64, 48), (63, 47), (1, 1), (2, 3), (129, 5)])
def test_perfect_reconstruction(w, h):  # pyramid.rs:282-295
    im = noise_img(w, h, 7 + w)
    levels = ref.auto_levels(w, h, 1)
    pyr = ref.pyr_build(im, levels)
    assert len(pyr) == levels + 1
    back = ref.pyr_collapse(pyr)
    assert np.all(np.abs(im - back) < 1e-5), f"{w}x{h}"


def test_auto_levels_targets_top_size():  # pyramid.rs:298-301
    assert ref.auto_levels(8280, 5520, 32) == 7
    assert ref.auto_levels(16, 16, 32) == 1
    # the config shapes: 1024x768 -> 4 with min_top 32 (5 levels are asked for explicitly)
    assert ref.auto_levels(8256, 5504, 32) == 7


def _reflect(i, n):
    if n == 1:
        return 0
    while True:
        if i < 0:
            i = -i
        elif i >= n:
            i = 2 * (n - 1) - i
        else:
            return i


def _reduce_matrix(n):
    k = np.array([1, 4, 6, 4, 1]) / 16.0
    m = np.zeros(((n + 1) // 2, n))
    for i in range((n + 1) // 2):
        for t in range(5):
            m[i, _reflect(2 * i + t - 2, n)] += k[t]
    return m


def _expand_matrix(cn, on):
    m = np.zeros((on, cn))
    for i in range(on):
        c = i // 2
        if i % 2 == 0:
            for off, wgt in ((-1, 0.125), (0, 0.75), (1, 0.125)):
                m[i, _reflect(c + off, cn)] += wgt
        else:
            m[i, _reflect(c, cn)] += 0.5
            m[i, _reflect(c + 1, cn)] += 0.5
    return m


@pytest.mark.parametrize("w,h", [(16, 16), (15, 9), (7, 13), (3, 2), (1, 5)])
def test_reduce_expand_match_dense_operators(w, h):
    """REDUCE = R_h X R_w^T and EXPAND = E_h X E_w^T with the reflect-101
    operators built independently from the Burt-Adelson definition."""
    x = noise_img(w, h, 3 + w * h)[0].astype(np.float64)
    want = _reduce_matrix(h) @ x @ _reduce_matrix(w).T
    got = ref.reduce(x.astype(np.float32))
    assert np.allclose(got, want, atol=2e-6)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    c = noise_img(cw, ch, 11)[1].astype(np.float64)
    want_e = _expand_matrix(ch, h) @ c @ _expand_matrix(cw, w).T
    got_e = ref.expand(c.astype(np.float32), w, h)
    assert np.allclose(got_e, want_e, atol=2e-6)


def test_laplacian_levels_are_g_minus_expand_reduce():
    im = noise_img(33, 21, 5)
    pyr = ref.pyr_build(im, 3)
    g = im
    for l in range(3):
        nxt = np.stack([ref.reduce(g[c]) for c in range(3)])
        lap = np.stack([g[c] - ref.expand(nxt[c], g.shape[2], g.shape[1]) for c in range(3)])
        assert np.array_equal(pyr[l], lap)
        g = nxt
    assert np.array_equal(pyr[3], g)
