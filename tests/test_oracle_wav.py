"""Pins the oracle's weighted average (wav.rs:88-161) and slab ranges
(stack.rs:76-92): the reference's own tests restated (wav.rs:168-265,
stack.rs:1125-1142), the weight at closed-form points, power 0 as the plain
mean of the frames, and identical frames as a fixed point."""
import numpy as np
import pytest

from oracle import ref


def test_slab_ranges_are_the_references():
    # stack.rs:1127-1132
    assert ref.slab_ranges(8, 3, 1) == [(0, 2), (2, 4), (4, 6), (6, 7)]
    assert ref.slab_ranges(20, 10, 2) == [(0, 9), (8, 17), (16, 19)]
    assert ref.slab_ranges(8, 8, 0) == [(0, 7)]
    assert ref.slab_ranges(8, 20, 5) == [(0, 7)]  # a slab larger than the stack is the stack
    assert ref.slab_ranges(5, 2, 7) == [(0, 1), (1, 2), (2, 3), (3, 4)]  # overlap clamped below the size
    assert ref.slab_ranges(0, 3, 1) == []
    # stack.rs:1133-1140: cover the stack, advance, leave no gap
    for count, size, overlap in [(100, 10, 2), (7, 3, 2), (9, 4, 0), (1, 5, 3)]:
        r = ref.slab_ranges(count, size, overlap)
        assert r[0][0] == 0 and r[-1][1] == count - 1
        for a, b in zip(r, r[1:]):
            assert b[0] > a[0] and b[0] <= a[1] + 1, r


def test_weight_closed_forms():
    f32 = np.float32
    # above the gate: (3 - 1.5)^2 + (0.01)^2; at or below it only the even weight
    assert ref.wav_weight(3.0, 1.0, 0.5, 2.0) == pytest.approx(2.25 + 1e-4, rel=1e-6)
    assert ref.wav_weight(1.5, 1.0, 0.5, 2.0) == pytest.approx(1e-4, rel=1e-5)
    assert ref.wav_weight(0.2, 1.0, 0.5, 2.0) == pytest.approx(1e-4, rel=1e-5)
    # power 1 is the contrast above the gate itself; gate 0 cuts at the floor
    assert ref.wav_weight(5.0, 2.0, 0.0, 1.0) == pytest.approx(3.0 + 0.02, rel=1e-6)
    # no floor: the contrast to the power (and the 1e-30 that keeps a zero cell a weight)
    assert ref.wav_weight(0.5, 0.0, 0.5, 2.0) == pytest.approx(0.25, rel=1e-6)
    assert ref.wav_weight(0.0, 0.0, 0.5, 2.0) == pytest.approx(1e-30, rel=1e-3)
    # power 0: every cell weighs 2 (x^0 = 1, 0^0 = 1)
    assert ref.wav_weight(0.0, 0.0, 0.5, 0.0) == f32(2.0)
    assert ref.wav_weight(7.0, 0.3, 0.5, 0.0) == f32(2.0)


def test_takes_the_sharper_frame():
    """wav.rs:168-195: two frames, each sharp on one half."""
    w = h = 64
    yy, xx = np.mgrid[0:h, 0:w]
    fine = np.where(((xx // 2 + yy // 2) % 2) == 0, 0.8, 0.2).astype(np.float32)
    va = np.where(xx < 32, fine, 0.5).astype(np.float32)
    vb = np.where(xx < 32, 0.5, fine).astype(np.float32)
    a, b = np.stack([va] * 3), np.stack([vb] * 3)
    floor = np.zeros((32, 32), np.float32)
    out = ref.weighted_average([a, b], floor, None, power=2.0, smooth=0, gate=0.5, edge=False, scale=1)
    assert abs(out[0, 16, 8] - a[0, 16, 8]) < 0.05
    assert abs(out[0, 16, 56] - b[0, 16, 56]) < 0.05


def _lcg_frames(w, h, n):
    """wav.rs:197-200's generator (6364136223846793005, 1442695040888963407), in the test's draw order."""
    mask = (1 << 64) - 1
    seed = 7
    frames = []
    for _ in range(n):
        f = np.zeros((3, h, w), np.float32)
        for c in range(3):
            v = np.zeros(w * h, np.float32)
            for i in range(w * h):
                seed = (seed * 6364136223846793005 + 1442695040888963407) & mask
                v[i] = np.float32(0.5) + np.float32(0.06) * (np.float32((seed >> 33) % 10000) / np.float32(10000.0)
                                                             - np.float32(0.5))
            f[c] = v.reshape(h, w)
        frames.append(f)
    return frames


def test_noise_floor_keeps_the_sharp_frame():
    """wav.rs:202-265: noise in every frame, a faint texture sharp in one of many."""
    w, h, n = 96, 64, 12
    frames = _lcg_frames(w, h, n)
    yy, xx = np.mgrid[0:h, 0:w // 2]
    t = np.where(((xx // 2 + yy // 2) % 2) == 0, 0.05, -0.05).astype(np.float32)
    for c in range(3):
        frames[5][c, :, : w // 2] += t
    fused = frames[5].copy()
    slices = [ref.focus_slice(f, scale=1) for f in frames]
    _, _, floor, _ = ref.depth_from_slices(slices, fused, scale=1)
    out = ref.weighted_average(frames, floor, fused, power=2.0, smooth=2, gate=0.5, edge=True, scale=1)
    sgn = np.where(((np.mgrid[8:h - 8, 8:w // 2 - 8][1] // 2 + np.mgrid[8:h - 8, 8:w // 2 - 8][0] // 2) % 2) == 0,
                   1.0, -1.0)

    def corr(img):
        return float(np.mean(sgn * (img[0, 8:h - 8, 8:w // 2 - 8] - 0.5)))

    kept = corr(out) / corr(frames[5])
    assert kept > 0.8, kept

    def sd(img):
        return float(np.std(img[0, 8:h - 8, w // 2 + 8:w - 8]))

    assert sd(out) / sd(frames[0]) < 0.5
    plain = ref.weighted_average(frames, np.zeros_like(floor), fused, power=2.0, smooth=2, gate=0.5, edge=True, scale=1)
    assert corr(plain) / corr(frames[5]) < 0.5


def test_power_zero_is_the_plain_mean():
    rng = np.random.default_rng(3)
    frames = [rng.uniform(0, 1, (3, 40, 52)).astype(np.float32) for _ in range(5)]
    floor = rng.uniform(0, 0.1, (20, 26)).astype(np.float32)
    out = ref.weighted_average(frames, floor, None, power=0.0, smooth=0, gate=0.5, edge=False, scale=1)
    acc, ws = np.zeros_like(frames[0]), np.zeros((40, 52), np.float32)
    for f in frames:  # every weight is 2 (+1e-30, below the f32 resolution of 2)
        acc = acc + np.float32(2.0) * f
        ws = ws + np.float32(2.0)
    assert np.array_equal(out, np.clip(acc / ws, 0, 1))


@pytest.mark.parametrize("smooth,edge", [(0, False), (2, False), (2, True)])
def test_identical_frames_are_a_fixed_point(smooth, edge):
    rng = np.random.default_rng(4)
    f = rng.uniform(0.1, 0.9, (3, 36, 44)).astype(np.float32)
    floor = np.full((18, 22), 0.01, np.float32)
    out = ref.weighted_average([f, f, f], floor, f, power=2.0, smooth=smooth, gate=0.5, edge=edge, scale=1)
    assert np.allclose(out, f, rtol=2e-6, atol=1e-7)
