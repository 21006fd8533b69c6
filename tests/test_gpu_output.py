"""The output side on the device, through the C ABI, against the oracle:
crop (Img3::crop), resize (prep.rs:46-71, the restretch of stack.rs:951-974)
and the 16-bit quantisation (io.rs:283-288).

Bars: crop and quantisation bit for bit; resize bit for bit for nearest,
bilinear and spline4x4 and within 1e-6 for the kernels whose weights go
through f64 sin or division (the warp's own bars); the crop rectangle of a
registered stack exactly the oracle's."""
import numpy as np
import pytest

import lapstack_amd as la
from lapstack_amd import synth
from oracle import ref

pytestmark = pytest.mark.gpu

EXACT = {"nearest", "bilinear", "spline4x4"}


@pytest.fixture(scope="module", autouse=True)
def _device():
    why = la.available()
    if why is not None:
        pytest.fail(f"HIP backend unavailable on a GPU run: {why}")


def test_crop_host_and_device():
    import torch

    rng = np.random.default_rng(1)
    img = rng.uniform(0, 1, (3, 37, 53)).astype(np.float32)
    r = (5, 3, 40, 30)
    assert np.array_equal(la.crop(img, r), img[:, 3:33, 5:45])
    d = torch.from_numpy(img).cuda()
    assert np.array_equal(la.crop(d, r).cpu().numpy(), img[:, 3:33, 5:45])
    assert np.array_equal(la.crop(img[0], r), img[0, 3:33, 5:45])
    with pytest.raises(la.HipError):
        la.crop(img, (20, 0, 40, 30))


@pytest.mark.parametrize("interp", ["nearest", "bilinear", "bicubic", "spline4x4", "spline6x6", "lanczos3"])
def test_resize_matches_the_oracle(interp):
    rng = np.random.default_rng(2)
    src = rng.uniform(0, 1, (3, 41, 57)).astype(np.float32)
    for w, h in [(64, 48), (57, 60), (30, 41), (23, 17)]:
        got = la.resize(src, w, h, interp)
        want = np.stack([ref.resize_plane(src[c], w, h, interp) for c in range(3)])
        if interp in EXACT:
            assert np.array_equal(got, want), (interp, w, h, np.abs(got - want).max())
        else:
            assert np.abs(got - want).max() <= 1e-6, (interp, w, h)
    assert np.array_equal(la.resize(src, 57, 41, interp), src)


def test_to_u16_matches_the_oracle():
    import torch

    rng = np.random.default_rng(3)
    img = rng.uniform(-0.2, 1.2, (3, 29, 31)).astype(np.float32)
    assert np.array_equal(la.to_u16(img), ref.planes_to_u16(img))
    out = torch.zeros((29, 31, 3), dtype=torch.uint16, device="cuda")
    la.to_u16(torch.from_numpy(img).cuda(), out)
    assert np.array_equal(out.cpu().numpy(), ref.planes_to_u16(img))


def test_output_chain_of_a_registered_stack():
    """A jittered stack through the stream, then stack.rs's output side: the
    common area of its Sims (exactly the oracle's), the fused image cropped to
    it, restretched to the frame's size, written as u16 -- each step against
    the oracle's on the same input."""
    w, h, n = 288, 208, 5
    b = synth.Bracket(w, h, n, seed=21, jitter=True, device="cuda")
    frames = [b.frame(i) for i in range(n)]
    ap = la.AlignParams(coarsen=1, interp="spline4x4")
    out = la.stack_stream(frames, la.FuseParams(levels=5), ap, None)
    sims = np.asarray(out["sims"])
    rect = la.common_area(sims, w, h, "spline4x4")
    assert rect == ref.common_area(sims, w, h, "spline4x4")
    assert rect != (0, 0, w, h)  # the jitter cuts the border
    img = np.asarray(out["image"], np.float32)
    cropped = la.crop(img, rect)
    x, y, rw, rh = rect
    assert np.array_equal(cropped, img[:, y:y + rh, x:x + rw])
    back = la.restretch(cropped, rect, w, h, "spline4x4")
    want = np.stack([ref.resize_plane(cropped[c], w, h, "spline4x4") for c in range(3)])
    assert back.shape == (3, h, w) and np.array_equal(back, want)
    assert np.array_equal(la.to_u16(back), ref.planes_to_u16(want))


@pytest.mark.parametrize("w,h", [(1000, 700), (8256, 5504), (130, 129), (257, 66), (17, 200), (9, 65), (3, 300)])
def test_registration_pyramid_matches_the_oracle(w, h):
    """align.rs:741-754 on the device (the pyramid every frame's search runs on), bit for bit."""
    rng = np.random.default_rng(w * 7 + h)
    y = rng.uniform(0, 1, (h, w)).astype(np.float32)
    if w > 100:
        yy, xx = np.mgrid[0:h, 0:w]
        y = (y * 0.2 + 0.8 * (0.5 + 0.5 * np.sin(xx / 13.0) * np.cos(yy / 7.0))).astype(np.float32)
    got, want = la.gauss_levels(y), ref.gauss_levels(y)
    assert len(got) == len(want)
    for l, (a, b) in enumerate(zip(got, want)):
        assert a.shape == b.shape and np.array_equal(a, b), (w, h, l)
