"""The oracle's depth.rs restatement, pinned by the reference's tests
(depth.rs:1211-1405) and by textbook reductions: Thomas sweeps against
scipy.linalg.solve_banded, the WLS solution against a sparse direct solve in
f64, a symmetric V-cycle, and naive box/guided filters."""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse
import scipy.sparse.linalg

from oracle import ref
from tests.fixtures import blur, checker, lcg_seq, wls_matrix


def test_box_filter_matches_naive():  # depth.rs:1212-1230
    w, h, r = 23, 17, 3
    src = lcg_seq(5, w * h).reshape(h, w)
    m = ref.box_mean(src, r)
    for y in range(h):
        for x in range(w):
            win = src[max(y - r, 0):min(y + r + 1, h), max(x - r, 0):min(x + r + 1, w)]
            assert abs(m[y, x] - win.mean()) < 1e-4


def test_guided_filter_flat_guide_is_box_mean():  # depth.rs:1233-1245
    w, h = 40, 30
    p = lcg_seq(9, w * h).reshape(h, w)
    q = ref.guided_filter(np.full((h, w), 0.5, np.float32), p, 2, 1e-3)
    m = ref.box_mean(ref.box_mean(p, 2), 2)
    assert np.all(np.abs(q - m) < 1e-4)


def test_guided_filter_matches_the_textbook_formula():
    w, h, r, eps = 31, 22, 2, 1e-3
    I = lcg_seq(3, w * h).reshape(h, w).astype(np.float64)
    p = lcg_seq(4, w * h).reshape(h, w).astype(np.float64)

    def bm(a):
        out = np.zeros_like(a)
        for y in range(h):
            for x in range(w):
                out[y, x] = a[max(y - r, 0):y + r + 1, max(x - r, 0):x + r + 1].mean()
        return out

    mi, mp = bm(I), bm(p)
    a = (bm(I * p) - mi * mp) / (np.maximum(bm(I * I) - mi * mi, 0) + eps)
    b = mp - a * mi
    want = bm(a) * I + bm(b)
    got = ref.guided_filter(I.astype(np.float32), p.astype(np.float32), r, eps)
    assert np.allclose(got, want, atol=2e-4)


def test_rdf_taps_are_zero_mean():  # depth.rs:1248-1255
    taps = ref.rdf_taps(1, 3)
    assert abs(sum(t[2] for t in taps)) < 1e-6
    assert sum(1 for t in taps if t[2] > 0) == 5
    assert len(taps) == 29
    y = np.full((20, 20), 0.3, np.float32)
    assert np.all(np.abs(ref.focus_measure(y)) < 1e-6)


def test_focus_slice_matches_numpy():
    img = np.stack([lcg_seq(s, 37 * 29).reshape(29, 37) for s in (1, 2, 3)])
    y = (np.float32(0.299) * img[0] + np.float32(0.587) * img[1] + np.float32(0.114) * img[2]).astype(np.float64)
    taps = ref.rdf_taps(1, 3)
    ri = lambda i, n: abs(i) if i < 0 else (2 * (n - 1) - i if i >= n else i)
    f = np.zeros_like(y)
    for yy in range(29):
        for xx in range(37):
            f[yy, xx] = abs(sum(wt * y[ri(yy + dy, 29), ri(xx + dx, 37)] for dy, dx, wt in taps))
    want = np.zeros((15, 19))
    for oy in range(15):
        for ox in range(19):
            want[oy, ox] = f[2 * oy:2 * oy + 2, 2 * ox:2 * ox + 2].mean()
    got = ref.focus_slice(img.astype(np.float32))
    assert np.allclose(got, want, atol=1e-6)


def _track(profile):
    t = ref.PeakTracker(1)
    for v in profile:
        t.push(np.array([v], np.float32))
    d, c, _, _ = t.finish(2.0)
    return float(d[0]), float(c[0])


def test_peak_tracker_profiles():  # depth.rs:1267-1288
    d, c = _track([0.1, 0.5, 1.0, 0.5, 0.1])
    assert abs(d - 2.0) < 1e-5 and c > 0.5
    d, _ = _track([0.1, 0.4, 1.0, 0.8, 0.1])
    assert 2.0 < d < 2.5
    assert abs(_track([1.0, 0.5, 0.1])[0]) < 1e-6
    assert abs(_track([0.1, 0.5, 1.0])[0] - 2.0) < 1e-6
    assert abs(_track([0.1, 1.0, 1.0, 0.1])[0] - 1.5) < 1e-5
    _, c2 = _track([0.1, 1.0, 0.1, 0.95, 0.1])
    _, c1 = _track([0.1, 1.0, 0.1, 0.1, 0.1])
    assert c2 < 0.2 and c1 > 0.5
    assert _track([0.3, 0.3, 0.3, 0.3])[1] < 1e-6
    assert _track([0.7])[0] == 0.0


def test_peak_tracker_gaussian_fit_is_exact_on_a_gaussian():
    # a sampled Gaussian profile: the log-parabola through three samples is exact
    mu, s = 2.3, 0.9
    prof = [float(np.exp(-(i - mu) ** 2 / (2 * s * s))) for i in range(6)]
    d, _ = _track(prof)
    assert abs(d - mu) < 1e-4


def test_normalize_conf_and_median():
    conf = lcg_seq(7, 1000)
    out, p90 = ref.normalize_conf(conf)
    smp = np.sort(conf[::7])
    assert p90 == smp[min(len(smp) * 9 // 10, len(smp) - 1)]
    assert np.allclose(out, np.minimum(conf / p90, 1.0))
    src = lcg_seq(8, 12 * 9).reshape(9, 12)
    med = ref.median3(src)
    pad = np.pad(src, 1, mode="reflect")
    for y in range(9):
        for x in range(12):
            assert med[y, x] == np.median(pad[y:y + 3, x:x + 3])


def test_thomas_sweeps_match_solve_banded():
    w, h, lam = 17, 5, 2.5
    f = lcg_seq(11, w * h).reshape(h, w)
    wd = lcg_seq(12, w * h).reshape(h, w) + np.float32(0.1)
    a = lcg_seq(13, w * h).reshape(h, w)
    u = ref.solve_rows(f, wd, a, lam)
    for y in range(h):
        ai = a[y].astype(np.float64).copy()
        ai[-1] = 0
        prev = np.concatenate([[0.0], ai[:-1]])
        ab = np.zeros((3, w))
        ab[0, 1:] = -lam * ai[:-1]
        ab[1] = wd[y] + lam * (prev + ai)
        ab[2, :-1] = -lam * ai[:-1]
        want = scipy.linalg.solve_banded((1, 1), ab, wd[y].astype(np.float64) * f[y])
        assert np.allclose(u[y], want, rtol=1e-4, atol=1e-5)
    uc = ref.solve_cols(f.T.copy(), wd.T.copy(), a.T.copy(), lam)
    assert np.allclose(uc.T, u, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("w,h,lam", [(32, 24, 5.0), (64, 40, 3.0), (45, 33, 50.0)])
def test_wls_matches_a_direct_f64_solve(w, h, lam):
    guide = lcg_seq(21 + w, w * h).reshape(h, w) * np.float32(0.3)
    conf = lcg_seq(22 + w, w * h).reshape(h, w)
    conf[h // 3: h // 2, w // 4: w // 2] = 0  # a hole
    d = (lcg_seq(23 + w, w * h).reshape(h, w) * np.float32(7.0)).astype(np.float32)
    u, rel, it = ref.wls_solve(d, conf, guide, lam, 0.04, 300, tol=1e-6)
    assert rel < 1e-6 and it > 0
    A = wls_matrix(guide, conf, lam, 0.04)
    b = (np.maximum(conf, 0) + np.float32(1e-4)).astype(np.float64).ravel() * d.astype(np.float64).ravel()
    want = scipy.sparse.linalg.spsolve(A.tocsc(), b).reshape(h, w)
    # the oracle's f32 operator and the f64 matrix agree: residual of the oracle's answer in f64
    r64 = np.linalg.norm(A @ u.astype(np.float64).ravel() - b) / np.linalg.norm(b)
    assert r64 < 1e-5
    assert np.max(np.abs(u - want)) < 1e-3 * max(1.0, np.abs(want).max())


def test_vcycle_is_symmetric_and_apply_is_the_matrix():
    w, h = 37, 29
    guide = lcg_seq(31, w * h).reshape(h, w) * np.float32(0.2)
    conf = lcg_seq(32, w * h).reshape(h, w)
    s = ref.WlsSolver(guide, 3.0, 0.04)
    s.set_weights(conf)
    x = lcg_seq(33, w * h).reshape(h, w) - np.float32(0.5)
    y = lcg_seq(34, w * h).reshape(h, w) - np.float32(0.5)
    mx, my = s.vcycle(x), s.vcycle(y)
    a, b = float((mx.astype(np.float64) * y).sum()), float((x.astype(np.float64) * my).sum())
    assert abs(a - b) < 1e-4 * max(abs(a), abs(b))
    assert float((mx.astype(np.float64) * x).sum()) > 0  # positive definite on x
    A = wls_matrix(guide, conf, 3.0, 0.04)
    assert np.allclose(s.apply(x).ravel(), A @ x.astype(np.float64).ravel(), atol=1e-5)


def test_wls_keeps_confident_data_and_fills_holes():  # depth.rs:1291-1321
    w, h = 32, 24
    y, x = np.mgrid[0:h, 0:w]
    d = x.astype(np.float32)
    conf = np.ones((h, w), np.float32)
    hole = (x >= 10) & (x < 22) & (y >= 6) & (y < 18)
    conf[hole] = 0
    d[hole] = 100
    guide = np.full((h, w), 0.5, np.float32)
    u, rel, _ = ref.wls_solve(d, conf, guide, 5.0, 0.05, 500)
    assert rel < 1e-4
    assert np.all(np.abs(u[4:h - 4, 4:w - 4] - x[4:h - 4, 4:w - 4]) < 0.6)
    u0, _, _ = ref.wls_solve(d, np.ones((h, w), np.float32), guide, 0.0, 0.05, 10)
    assert np.all(np.abs(u0 - d) < 1e-3)


def test_wls_respects_guide_edges():  # depth.rs:1324-1341
    w, h = 40, 20
    x = np.arange(w)[None].repeat(h, 0)
    d = np.where(x >= w // 2, 10.0, 0.0).astype(np.float32)
    guide = np.where(x >= w // 2, 0.9, 0.1).astype(np.float32)
    u, _, _ = ref.wls_solve(d, np.full((h, w), 0.2, np.float32), guide, 50.0, 0.02, 300)
    assert u[10, w // 2 - 1] < 0.5 and u[10, w // 2] > 9.5
    v, _, _ = ref.wls_solve(d, np.full((h, w), 0.2, np.float32), np.full((h, w), 0.5, np.float32), 50.0, 0.02, 300)
    assert v[10, w // 2 - 1] > 2.0 and v[10, w // 2] < 8.0


def test_end_to_end_two_frames():  # depth.rs:1367-1405
    w, h = 128, 96
    sharp = checker(w, h, 5, factors=(1.0, 1.0, 1.0))
    soft = blur(sharp, 6)
    a, b = sharp.copy(), sharp.copy()
    a[:, :, w // 2:] = soft[:, :, w // 2:]
    b[:, :, : w // 2] = soft[:, :, : w // 2]
    slices = [ref.focus_slice(f) for f in (a, b)]
    depth, conf, _, _ = ref.depth_from_slices(slices, sharp, lam=5.0)
    assert depth.shape == (h, w)
    xs = np.arange(w)[None].repeat(h, 0)
    want = np.where(xs < w // 2, 0.0, 1.0)
    mask = np.abs(xs - w // 2) >= 6
    bad = np.count_nonzero((np.abs(depth - want) > 0.25) & mask)
    assert bad < w * h / 100
    assert conf.mean() > 0.2
