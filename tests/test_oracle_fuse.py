"""The oracle's fuse.rs restatement, pinned by the reference's own tests
(fuse.rs:527-722) and by brute force (naive windowed energy, histogram entropy,
an independent N-frame argmax fold)."""
import numpy as np
import pytest

from oracle import ref
from tests.fixtures import blur, checker, noise_img, rms


def test_winner_shares_count_detail_cells_only():  # fuse.rs:528-545
    win = np.array([0, 0, 0, 1, 1, 2], np.uint16)
    best = np.array([0.001, 0.002, 0.001, 1.0, 0.5, 0.2], np.float32)
    s = ref.winner_shares(win, best, 3)
    assert s.tolist() == [0.0, np.float32(2.0 / 3.0), np.float32(1.0 / 3.0)]
    assert np.all(ref.winner_shares(win, np.zeros(6, np.float32), 3) == 0.0)
    img = checker(64, 64, 8)
    f = ref.Fuser(64, 64)
    for _ in range(3):
        f.push(img)
    s = f.shares()
    assert len(s) == 3
    assert abs(s[0] - 1.0) < 1e-6 and s[1] == 0.0 and s[2] == 0.0


def test_binomial_weights():  # fuse.rs:582-589
    assert ref.binomial(0).tolist() == [1.0]
    assert ref.binomial(1).tolist() == [0.25, 0.5, 0.25]
    k = np.array([1, 4, 6, 4, 1], np.float32) / 16
    assert np.all(np.abs(ref.binomial(2) - k) < 1e-6)


def test_identical_frames_are_a_fixed_point():  # fuse.rs:592-601
    im = checker(70, 50, 7)
    f = ref.Fuser(70, 50)
    for _ in range(3):
        f.push(im)
    out, depth, _ = f.finish()
    assert rms(out, im) < 1e-5
    assert np.all(depth == 0.0), "ties keep the first frame"


@pytest.mark.parametrize("kw", [dict(), dict(energy_radius=0), dict(top_rule="average")])
def test_picks_the_sharp_half_from_each_frame(kw):  # fuse.rs:604-641
    w, h = 96, 64
    sharp = checker(w, h, 6)
    soft = blur(sharp, 6)
    a, b = sharp.copy(), sharp.copy()
    a[:, :, w // 2:] = soft[:, :, w // 2:]
    b[:, :, : w // 2] = soft[:, :, : w // 2]
    f = ref.Fuser(w, h, depth_level=0, **kw)
    f.push(a)
    f.push(b)
    out, depth, _ = f.finish()
    assert rms(out, sharp) < 0.6 * rms(a, sharp)
    n = depth.size
    xs = np.arange(w)[None, :].repeat(h, 0)
    left = np.count_nonzero((xs < w // 2 - 8) & (depth == 0.0))
    right = np.count_nonzero((xs >= w // 2 + 8) & (depth == 1.0))
    assert left > 0.9 * (n / 2 - 8 * h)
    assert right > 0.9 * (n / 2 - 8 * h)


def test_halo_weight_is_monotonic_and_bounded():  # fuse.rs:644-652
    assert abs(ref.halo_weight(1e-4 - 1e-8, 3.0) - 1.0) < 1e-5
    assert ref.halo_weight(1e-3, 2.0) > ref.halo_weight(1e-4, 2.0)
    assert ref.halo_weight(1e-4, 4.0) / ref.halo_weight(1e-5, 4.0) > 9e3
    assert 0.0 < ref.halo_weight(0.0, 8.0) < np.inf
    assert np.isfinite(ref.halo_weight(1e3, 8.0))


def test_halo_control_keeps_identical_frames():  # fuse.rs:654-665
    im = checker(70, 50, 7)
    f = ref.Fuser(70, 50, halo=2.0, levels=3, depth_level=1)
    for _ in range(3):
        f.push(im)
    out, depth, _ = f.finish()
    assert rms(out, im) < 1e-5
    assert np.all(depth == 0.0)


def test_halo_control_coarse_levels_follow_the_guide():  # fuse.rs:673-702
    w, h = 192, 144
    a = checker(w, h, 6)
    b = blur(a, 64)
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    d2 = (x - 96.0) ** 2 + (y - 72.0) ** 2
    bump = (np.float32(0.35) * np.exp(-d2 / (2.0 * 20.0 ** 2))).astype(np.float32)
    b = np.minimum(b + bump[None], 1.0).astype(np.float32)

    def fuse(halo, order):
        f = ref.Fuser(w, h, levels=5, depth_level=2, halo=halo)
        f.push(order[0])
        f.push(order[1])
        return f.finish()[0]

    off = rms(fuse(0.0, [a, b]), a)
    assert off > 0.05
    for halo in [1.0, 2.0, 4.0, 8.0]:
        assert rms(fuse(halo, [a, b]), a) < 0.15 * off
    assert rms(fuse(2.0, [b, a]), a) < 0.15 * off


def test_top_rule_two_frames_matches_eq12():  # fuse.rs:705-722
    a = checker(40, 40, 4)
    b = np.full_like(a, 0.5)
    ta, tb = ref.pyr_build(a, 1)[1], ref.pyr_build(b, 1)[1]
    top = ref.fuse_residuals([ta, tb], ref.FuseParams.make(levels=1))
    assert rms(top, ta) < 1e-6


def _reflect_idx(n, r):
    idx = np.arange(-r, n + r)
    if n == 1:
        return np.zeros_like(idx)
    idx = np.abs(idx)
    return np.where(idx >= n, 2 * (n - 1) - idx, idx)


def test_region_energy_matches_naive_window():
    l3 = noise_img(23, 17, 9) - 0.5
    got = ref.region_energy(l3, 1)
    y = (0.299 * l3[0].astype(np.float64) + 0.587 * l3[1] + 0.114 * l3[2])
    e = y * y
    k = np.array([0.25, 0.5, 0.25])
    ri, ci = _reflect_idx(17, 1), _reflect_idx(23, 1)
    pad = e[np.ix_(ri, ci)]
    want = np.zeros_like(e)
    for dy in range(3):
        for dx in range(3):
            want += k[dy] * k[dx] * pad[dy:dy + 17, dx:dx + 23]
    assert np.allclose(got, want, rtol=1e-5, atol=1e-8)
    chroma = ref.region_energy(l3, 1, use_chroma=True)
    e2 = (l3.astype(np.float64) ** 2).sum(0)
    pad = e2[np.ix_(ri, ci)]
    want2 = sum(k[dy] * k[dx] * pad[dy:dy + 17, dx:dx + 23] for dy in range(3) for dx in range(3))
    assert np.allclose(chroma, want2, rtol=1e-5, atol=1e-8)


def test_deviation_entropy_matches_numpy_histogram():
    y = noise_img(19, 13, 4)[0]
    y[2:5, 3:9] = 0.5  # flat patch: zero deviation, low entropy
    dev, ent = ref.deviation_entropy(y, 2, 256)
    ri, ci = _reflect_idx(13, 2), _reflect_idx(19, 2)
    pad = y[np.ix_(ri, ci)].astype(np.float64)
    for i in range(13):
        for j in range(19):
            win = pad[i:i + 5, j:j + 5].ravel()
            assert abs(dev[i, j] - win.var()) < 1e-5
            q = np.floor(np.clip(win, 0, 1) * 255 + 0.5).astype(int)
            p = np.bincount(q) / 25.0
            p = p[p > 0]
            assert abs(ent[i, j] - float(-(p * np.log(p)).sum())) < 1e-5


def test_fold_matches_an_independent_argmax():
    """The fold = per-level argmax of region energy over the frames (ties to the
    earliest), coefficients taken from the winner; residual = Pareto mean."""
    w, h, n = 48, 40, 4
    frames = [blur(noise_img(w, h, 100 + i), i) for i in range(n)]
    f = ref.Fuser(w, h, levels=3, depth_level=1)
    for fr in frames:
        f.push(fr)
    out, _, wl = f.finish()
    pyrs = [ref.pyr_build(fr, 3) for fr in frames]
    fused = []
    for l in range(3):
        en = np.stack([ref.region_energy(p[l], 1) for p in pyrs])
        win = np.zeros(en.shape[1:], np.int64)
        best = en[0].copy()
        for i in range(1, n):
            take = en[i] > best
            win[take] = i
            best[take] = en[i][take]
        if l == 1:
            assert np.array_equal(win, wl)
        lv = np.stack([p[l] for p in pyrs])
        fused.append(np.take_along_axis(lv, win[None, None], 0)[0])
    fused.append(ref.fuse_residuals([p[3] for p in pyrs]))
    want = np.clip(ref.pyr_collapse(fused), 0, 1)
    assert np.array_equal(out, want)
