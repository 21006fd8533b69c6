"""The oracle's align.rs and brightness.rs restatements, pinned by the
reference's tests (align.rs:961-1210, brightness.rs:145-213), textbook
Nelder-Mead problems and synthetic ground-truth registration."""
import math

import numpy as np
import pytest

from oracle import ref

W8 = ref.sim_vec


def test_reduce_keeps_a_constant_plane():  # align.rs:969-982
    for w, h in [(1, 1), (2, 3), (3, 4), (4, 5), (5, 5), (6, 7), (9, 65), (10, 66), (33, 130)]:
        d = ref.reduce_burt(np.full((h, w), 0.75, np.float32))
        assert d.shape == ((h + 1) // 2, (w + 1) // 2)
        assert np.all(np.abs(d - 0.75) < 1e-6), f"{w}x{h}"
    lv = ref.gauss_levels(np.full((300, 40), 0.25, np.float32))
    assert [(l.shape[1], l.shape[0]) for l in lv] == [(20, 150), (10, 75), (5, 38)]
    assert all(np.all(np.abs(l - 0.25) < 1e-6) for l in lv)


def test_cost_matches_the_two_pass_definition():  # align.rs:987-1011
    w, h = 96, 72
    i = np.arange(w * h)
    a = (((i * 7919) % 1000) / 1000.0).astype(np.float32).reshape(h, w)
    b = (((i * 104729 + 13) % 1000) / 1000.0 + 0.2).astype(np.float32).reshape(h, w)
    for s in [W8(), W8(xoff=0.03, yoff=-0.02, scale=1.02, rot=0.01), W8(px=0.03, py=-0.02, shear=0.01)]:
        bw, valid = ref.warp_plane(b, s, "spline4x4")
        m = valid != 0
        ma, mb = a[m].astype(np.float64).mean(), bw[m].astype(np.float64).mean()
        want = math.sqrt((((a[m] - ma) - (bw[m] - mb)) ** 2).mean())
        got = ref.warp_cost(a, b, s)
        assert abs(got - want) < 1e-5 * want
    assert ref.warp_cost(a, a, W8()) < 1e-9
    assert ref.warp_cost(a, a, W8(xoff=2.0)) == 1e9


def test_level_steps_move_the_edge_by_a_level_pixel():  # align.rs:1016-1031
    w, h = 8000, 6000
    step, tol = ref.level_steps(list(range(8)), 3, w, h)
    r = max(w, h) / 2
    gain = [w, h, r, r, h / 2, h / 2, w / 4, h / 4]
    for k in range(8):
        assert abs(step[k] * gain[k] - 8.0) < 1e-9
        assert abs(tol[k] * gain[k] - 0.8) < 1e-9
    assert step[0] * w == 8.0
    assert step[2] * w / 2 == 8.0
    s2, _ = ref.level_steps([0, 3], 0, w, h)
    assert len(s2) == 2 and s2[0] == 1.0 / w


def test_common_area_identity_is_full():  # align.rs:1034-1037
    assert ref.common_area([ref.SIM_ID, ref.SIM_ID], 640, 480) == (0, 0, 640, 480)
    assert ref.common_area([], 640, 480) == (0, 0, 640, 480)


def test_common_area_shift():  # align.rs:1040-1051
    w, h = 640, 480
    s = W8(xoff=10.0 / w)
    assert ref.common_area([ref.SIM_ID, s], w, h, "spline4x4") == (12, 2, w - 12, h - 4)
    assert ref.common_area([ref.SIM_ID, s], w, h, "lanczos3") == (13, 3, w - 13, h - 6)
    assert ref.common_area([ref.SIM_ID, s], w, h, "bilinear") == (11, 1, w - 11, h - 2)
    assert ref.common_area([ref.SIM_ID, s], w, h, "nearest") == (10, 0, w - 10, h)


@pytest.mark.parametrize("k", list(ref.INTERP))
def test_every_kernel_interpolates(k):  # align.rs:1056-1072
    w, h = 40, 30
    src = (((np.arange(w * h) * 7919) % 1000) / 1000.0).astype(np.float32).reshape(h, w)
    o, _ = ref.warp_plane(src, W8(), k)
    assert np.all(np.abs(o - src) < 1e-6)
    o, valid = ref.warp_plane(src, W8(xoff=3.0 / w, yoff=-2.0 / h), k)
    for y in range(6, h - 6):
        for x in range(6, w - 6):
            assert valid[y, x] == 1
            assert abs(o[y, x] - src[y + 2, x - 3]) < 1e-5


@pytest.mark.parametrize("k", list(ref.INTERP))
def test_kernels_reproduce_a_ramp(k):  # align.rs:1079-1094
    w, h = 48, 40
    i = np.arange(w * h)
    ramp = ((i % w) + 0.5 * (i // w)).astype(np.float32).reshape(h, w)
    o, _ = ref.warp_plane(ramp, W8(xoff=2.3 / w, yoff=1.6 / h), k)
    for y in range(8, h - 8):
        for x in range(8, w - 8):
            sx, sy = np.float32(x - 2.3), np.float32(y - 1.6)
            want = (np.round(sx) + 0.5 * np.round(sy)) if k == "nearest" else sx + 0.5 * sy
            tol = 0.03 if k == "lanczos3" else 1e-3
            assert abs(o[y, x] - want) < tol


def test_common_area_rotation_is_inside_every_frame():  # align.rs:1122-1148
    w, h = 640, 480
    sims = [ref.SIM_ID, W8(xoff=-0.01, yoff=0.02, scale=1.03, rot=0.02), W8(xoff=0.005, yoff=-0.01, scale=0.98, rot=-0.015)]
    m = 2.0
    x, y, rw, rh = ref.common_area(sims, w, h, "spline4x4")
    assert rw > w // 2 and rh > h // 2 and (x, y, rw, rh) != (0, 0, w, h)

    def sound(px, py):
        for s in sims[1:]:
            inv = ref.inverse(ref.sim_matrix(s, w, h))
            sx, sy = ref.map_point(inv, px, py)
            if not (m <= sx <= w - 1 - m and m <= sy <= h - 1 - m):
                return False
        return True

    for (px, py) in [(x, y), (x + rw - 1, y), (x, y + rh - 1), (x + rw - 1, y + rh - 1)]:
        assert sound(px, py)
    row_ok = lambda yy: all(sound(xx, yy) for xx in range(x, x + rw))
    col_ok = lambda xx: all(sound(xx, yy) for yy in range(y, y + rh))
    assert y == 0 or not row_ok(y - 1)
    assert y + rh == h or not row_ok(y + rh)
    assert x == 0 or not col_ok(x - 1)
    assert x + rw == w or not col_ok(x + rw)


def test_homography_round_trip_and_meaning():  # align.rs:1154-1182
    w, h = 640, 480
    s = W8(0.01, -0.02, 1.03, 0.02, 1.02, 0.01, 0.03, -0.02)
    m = ref.sim_matrix(s, w, h)
    inv = ref.inverse(m)
    assert inv[2, 2] == 1.0
    for (x, y) in [(0.0, 0.0), (100.0, 37.0), (639.0, 479.0), (320.0, 240.0)]:
        fx, fy = ref.map_point(m, x, y)
        bx, by = ref.map_point(inv, fx, fy)
        assert abs(bx - x) < 1e-9 and abs(by - y) < 1e-9
    cx, cy = w / 2, h / 2
    a = W8(aspect=1.5, shear=0.2)
    assert ref.map_point(ref.sim_matrix(a, w, h), cx, cy) == (cx, cy)
    assert ref.map_point(ref.sim_matrix(a, w, h), cx, cy + 10.0) == (cx + 2.0, cy + 15.0)
    p = W8(py=0.1)
    top_l, _ = ref.map_point(ref.sim_matrix(p, w, h), cx - 100.0, cy - 100.0)
    bot_l, _ = ref.map_point(ref.sim_matrix(p, w, h), cx - 100.0, cy + 100.0)
    assert top_l < cx - 100.0 and bot_l > cx - 100.0
    assert np.array_equal(ref.inverse(ref.sim_matrix(ref.SIM_ID, w, h)), np.eye(3))


def test_projective_warp_and_common_area():  # align.rs:1187-1210
    w, h = 64, 48
    i = np.arange(w * h)
    ramp = ((i % w) + 0.5 * (i // w)).astype(np.float32).reshape(h, w)
    s = W8(px=0.04, py=-0.03, shear=0.01)
    o, valid = ref.warp_plane(ramp, s, "bilinear")
    inv = ref.inverse(ref.sim_matrix(s, w, h))
    for y in range(4, h - 4):
        for x in range(4, w - 4):
            sx, sy = ref.map_point(inv, x, y)
            assert valid[y, x] == 1
            assert abs(float(o[y, x]) - (sx + 0.5 * sy)) < 1e-3
    x, y, rw, rh = ref.common_area([ref.SIM_ID, s], w, h, "bilinear")
    assert (x, y, rw, rh) != (0, 0, w, h) and rw > w // 2 and rh > h // 2
    for (px, py) in [(x, y), (x + rw - 1, y), (x, y + rh - 1), (x + rw - 1, y + rh - 1)]:
        sx, sy = ref.map_point(inv, px, py)
        assert 1.0 <= sx <= w - 2 and 1.0 <= sy <= h - 2


def test_nelder_mead_textbook_problems():
    # a convex quadratic: the minimum to within the tolerance
    f = lambda x: (x[0] - 1.0) ** 2 + 10 * (x[1] + 0.5) ** 2
    x, _ = ref.nelder_mead(f, [0, 0], [-5, -5], [5, 5], [0.5, 0.5], [1e-4, 1e-4])
    assert abs(x[0] - 1.0) < 1e-3 and abs(x[1] + 0.5) < 1e-3
    # Rosenbrock from the classic start (200 iterations is the cap: close, not exact)
    rb = lambda x: 100 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2
    x, _ = ref.nelder_mead(rb, [-1.2, 1.0], [-5, -5], [5, 5], [0.1, 0.1], [1e-6, 1e-6])
    assert rb(x) < rb([-1.2, 1.0]) * 1e-2
    # the bound is respected: the unconstrained minimum at 3 lies outside [-1, 2]
    x, _ = ref.nelder_mead(lambda x: (x[0] - 3) ** 2, [0], [-1], [2], [0.5], [1e-5])
    assert abs(x[0] - 2.0) < 1e-4
    # one dimension, and the zero-dimensional case
    x, ev = ref.nelder_mead(lambda x: abs(x[0] - 0.3), [0.0], [-1], [1], [0.1], [1e-6])
    assert abs(x[0] - 0.3) < 1e-3 and ev > 2


def _smooth_scene(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w))
    for _ in range(40):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        s = rng.uniform(3, 12)
        img += rng.uniform(-1, 1) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    img = 0.5 + 0.2 * img / np.abs(img).max()
    return img.astype(np.float32)


@pytest.mark.parametrize("coarsen", [0, 1])
def test_multiscale_align_recovers_a_known_similarity(coarsen):
    w, h = 256, 192
    rf = _smooth_scene(w, h, 3)
    truth = W8(xoff=2.4 / w, yoff=-1.7 / h, scale=1.004, rot=0.0015)
    # the target is the reference seen through the inverse transform: warping it by `truth` lands it back
    inv_truth = ref.inverse(ref.sim_matrix(truth, w, h))
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    mx = inv_truth  # target(x) = rf(M x) with M the forward matrix of truth
    m = ref.sim_matrix(truth, w, h)
    d = m[2, 0] * xs + m[2, 1] * ys + m[2, 2]
    sx = (m[0, 0] * xs + m[0, 1] * ys + m[0, 2]) / d
    sy = (m[1, 0] * xs + m[1, 1] * ys + m[1, 2]) / d
    from scipy.ndimage import map_coordinates
    tg = map_coordinates(rf.astype(np.float64), [sy, sx], order=3, mode="nearest").astype(np.float32)
    got, ev = ref.multiscale_align(rf, tg, coarsen=coarsen)
    assert ev > 10
    # the recovered transform moves the frame's corners within a few tenths of a pixel of the truth
    for (px, py) in [(0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1)]:
        a = ref.map_point(ref.sim_matrix(got, w, h), px, py)
        b = ref.map_point(m, px, py)
        tol = 0.15 * (1 << coarsen) + 0.1
        assert abs(a[0] - b[0]) < tol and abs(a[1] - b[1]) < tol, (got, truth)
    del mx


# ---------------------------------------------------------------- brightness.rs
def _scene(w, h):
    y, x = np.mgrid[0:h, 0:w]
    t = ((x * 7 + y * 3) % 23).astype(np.float32) / np.float32(23.0)
    r = np.float32(0.2) + np.float32(0.5) * t
    g = np.float32(0.3) + np.float32(0.3) * (x.astype(np.float32) / np.float32(w))
    b = np.float32(0.1) + np.float32(0.6) * (y.astype(np.float32) / np.float32(h))
    return np.stack([r, g, b]).astype(np.float32)


def _blur5(im):
    _, h, w = im.shape
    out = np.zeros_like(im)
    for c in range(3):
        for y in range(h):
            for x in range(w):
                s = np.float32(0)
                for dy in range(-2, 3):
                    for dx in range(-2, 3):
                        s += im[c, min(max(y + dy, 0), h - 1), min(max(x + dx, 0), w - 1)]
                out[c, y, x] = s / np.float32(25.0)
    return out


def test_same_frame_is_unity():  # brightness.rs:184-189
    a = _scene(160, 120)
    assert ref.bright_gains(a, a, ref.SIM_ID).tolist() == [1.0, 1.0, 1.0]
    assert ref.bright_gains(a, a, W8(xoff=0.05, yoff=-0.02, scale=1.01, rot=0.01)).tolist() == [1.0, 1.0, 1.0]


def test_darker_blurred_frame_is_brought_back():  # brightness.rs:191-206
    a = _scene(160, 120)
    b = ref.bright_apply(_blur5(a), [0.8, 0.9, 1.1])
    g = ref.bright_gains(a, b, ref.SIM_ID)
    for got, want in zip(g, [1.25, 1 / 0.9, 1 / 1.1]):
        assert abs(got / want - 1.0) < 0.01
    b = ref.bright_apply(b, g)
    for x, y in zip(ref.bright_means(b), ref.bright_means(a)):
        assert abs(x / y - 1.0) < 0.01
    assert ref.bright_gains_to(ref.bright_means(a), a).tolist() == [1.0, 1.0, 1.0]


def test_clamped():  # brightness.rs:208-213
    assert ref.bright_ratio([1.0] * 3, [0.01] * 3, 1000).tolist() == [4.0, 4.0, 4.0]
    assert ref.bright_ratio([1.0] * 3, [0.5] * 3, 63).tolist() == [1.0, 1.0, 1.0]
