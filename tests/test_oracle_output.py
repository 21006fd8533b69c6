"""The output side (SURVEY's next rows after the path): the oracle's resize
(prep.rs:63-71) pinned by closed forms, the 16-bit quantisation (io.rs:283-288)
at exact points, and the library's host common_area (align.rs:344-403)
against the oracle's, which test_oracle_align pins."""
import numpy as np
import pytest

import lapstack_amd as la
from oracle import ref

KERNELS = ["nearest", "bilinear", "bicubic", "spline4x4", "spline6x6", "lanczos3"]


@pytest.mark.parametrize("interp", KERNELS)
def test_resize_keeps_a_constant(interp):
    p = np.full((30, 40), 0.37, np.float32)
    for w, h in [(64, 48), (23, 17), (40, 61)]:
        assert np.allclose(ref.resize_plane(p, w, h, interp), 0.37, atol=2e-6)
    assert np.array_equal(ref.resize_plane(p, 40, 30, interp), p)  # the same size is a copy (prep.rs:64-66)


def test_resize_puts_pixel_centers_on_pixel_centers():
    # a bilinear ramp is reproduced exactly: output column o reads source column (o + 1/2) * pw / w - 1/2
    pw, ph, w, h = 40, 30, 64, 45
    ramp_x = np.tile(np.arange(pw, dtype=np.float32), (ph, 1))
    ramp_y = np.tile(np.arange(ph, dtype=np.float32)[:, None], (1, pw))
    ox = (np.arange(w) + 0.5) * pw / w - 0.5
    oy = (np.arange(h) + 0.5) * ph / h - 0.5
    rx = ref.resize_plane(ramp_x, w, h, "bilinear")
    ry = ref.resize_plane(ramp_y, w, h, "bilinear")
    inner_x = (ox >= 0) & (ox <= pw - 1)
    inner_y = (oy >= 0) & (oy <= ph - 1)
    assert np.allclose(rx[:, inner_x], ox[inner_x][None, :], atol=2e-5)
    assert np.allclose(ry[inner_y, :], oy[inner_y][:, None], atol=2e-5)
    # outside the source the taps repeat the edge
    assert np.allclose(rx[:, ~inner_x & (ox < 0)], 0.0, atol=1e-6)


def test_u16_quantisation_points():
    v = np.array([0.0, 1.0, 0.5, -1.0, 2.0, 1.0 / 65535.0, 1.5 / 65535.0, np.float32(0.99999)], np.float32)
    img = np.stack([v, v[::-1], v])[:, None, :]
    q = ref.planes_to_u16(img)[0]
    assert q[:, 0].tolist() == [0, 65535, 32768, 0, 65535, 1, 2, 65534]


@pytest.mark.parametrize("interp", KERNELS)
def test_library_common_area_is_the_oracles(interp):
    rng = np.random.default_rng(KERNELS.index(interp))
    for trial in range(12):
        w, h = int(rng.integers(60, 400)), int(rng.integers(40, 300))
        n = int(rng.integers(1, 6))
        sims = []
        for i in range(n):
            s = [rng.normal(0, 0.01), rng.normal(0, 0.01), 1 + rng.normal(0, 0.004), rng.normal(0, 0.003), 1.0, 0.0,
                 0.0, 0.0]
            if trial % 3 == 1:  # affine
                s[4], s[5] = 1 + rng.normal(0, 0.004), rng.normal(0, 0.003)
            if trial % 3 == 2:  # projective
                s[6], s[7] = rng.normal(0, 0.02), rng.normal(0, 0.02)
            sims.append(s if i or trial % 2 else list(la.SIM_ID))
        assert la.common_area(sims, w, h, interp) == ref.common_area(sims, w, h, interp), (trial, w, h)
    assert la.common_area([la.SIM_ID, la.SIM_ID], 77, 55, interp) == (0, 0, 77, 55)
    assert la.common_area([], 77, 55, interp) == (0, 0, 77, 55)
