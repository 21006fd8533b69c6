"""The depth pass over a slice iterator with a log (include/lsh.h
lsh_depth_from_source; depth.rs:1119-1200): the same maps, bit for bit, as
over one contiguous array, from host slices and from one reused device
buffer; the reference's errors for a short source, a slice of the wrong size
and a failing source; the reference's progress lines, their numbers those of
the pass."""
import re

import numpy as np
import pytest

import lapstack_amd as la

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    why = la.available()
    if why is not None:
        pytest.fail(f"HIP backend unavailable on a GPU run: {why}")


def _case(w, h, n, p, seed):
    rng = np.random.default_rng(seed)
    dw, dh = p.grid(w, h)
    # a peak per cell at a frame that varies smoothly, plus noise
    peak = (np.add.outer(np.linspace(0, 1, dh), np.linspace(0, 1, dw)) / 2 * (n - 1)).astype(np.float32)
    m = np.arange(n, dtype=np.float32)[:, None, None]
    slices = (np.exp(-0.5 * (m - peak[None]) ** 2) + 0.05 * rng.random((n, dh, dw))).astype(np.float32)
    fused = rng.random((3, h, w)).astype(np.float32)
    return slices, fused


def _rust_e1(x):
    mant, ex = f"{x:.1e}".split("e")
    return f"{mant}e{int(ex)}"


def _same(a, b):
    assert np.array_equal(a.depth, b.depth)
    assert np.array_equal(a.conf, b.conf)
    assert np.array_equal(a.floor, b.floor)
    for k in ("dw", "dh", "iters1", "iters2", "rel1", "rel2", "p90", "floor"):
        assert a.info[k] == b.info[k], k


PARAMS = [
    la.DepthParams(),
    la.DepthParams(scale=0, focus="sml", agg_radius=0, lam=0.0, upsample="bilinear", median=False),
]


@pytest.mark.parametrize("pi", range(len(PARAMS)))
def test_host_slices_match_the_array_path(pi):
    p = PARAMS[pi]
    w, h, n = 320, 240, 5
    slices, fused = _case(w, h, n, p, 11 + pi)
    print(f"depth source, host slices, params {pi}: {n} x {w}x{h}", flush=True)
    want = la.depth_from_slices(slices, n, fused, p)
    lines = []
    got = la.depth_from_source(iter(list(slices)), n, fused, p, log=lines.append)
    _same(got, want)
    print("\n".join(lines), flush=True)
    dw, dh = p.grid(w, h)
    focus = "Rdf { r_in: 1, r_out: 3 }" if p.focus == "rdf" else "Sml { step: 1 }"
    eps = {1e-4: "0.0001"}[p.agg_eps]
    assert re.fullmatch(rf"depth: {n} frames, focus {re.escape(focus)}, working grid {dw}x{dh} \(1/{1 << p.scale}\), "
                        rf"aggregation r={p.agg_radius} eps={eps} \(GPU, \d+\.\d\ds to start\)", lines[0]), lines[0]
    assert re.fullmatch(rf"depth: {n} slices aggregated  \(\d+\.\ds\)", lines[1]), lines[1]
    m = re.fullmatch(r"depth: peaks found, confidence p90 (\d+\.\d{3}), mean \(normalized\) (\d+\.\d{3})  \(\d+\.\ds\)",
                     lines[2])
    assert m, lines[2]
    assert m.group(1) == f"{want.info['p90']:.3f}"
    if p.scale == 0:  # the returned confidence is the working grid's
        assert abs(float(m.group(2)) - float(np.mean(want.conf, dtype=np.float64))) <= 5.1e-4
    rest = lines[3:]
    if p.lam > 0 and p.robust > 0:
        assert rest[0] == (f"depth: robust reweighting (tau=1 frames), first solve {want.info['iters1']} iterations, "
                           f"residual {_rust_e1(want.info['rel1'])}"), rest[0]
        rest = rest[1:]
    lam = "3" if p.lam == 3.0 else "0"
    assert re.fullmatch(rf"depth: WLS lambda={lam} sigma_c=0\.04 solved in {want.info['iters2']} iterations, "
                        rf"residual {re.escape(_rust_e1(want.info['rel2']))}  \(\d+\.\ds\)", rest[0]), rest[0]
    up = "Guided { radius: 2, eps: 0.01 }" if p.upsample == "guided" else "Bilinear"
    assert re.fullmatch(rf"depth: upsampled to {w}x{h} \({re.escape(up)}\)  \(\d+\.\ds\)", rest[1]), rest[1]
    assert len(rest) == 2


def test_one_reused_device_buffer_matches_the_array_path():
    import torch

    p = la.DepthParams(agg_radius=8, agg_eps=1e-3)
    w, h, n = 400, 300, 6
    slices, fused = _case(w, h, n, p, 5)
    print(f"depth source, one device buffer: {n} x {w}x{h}", flush=True)
    want = la.depth_from_slices(slices, n, fused, p)
    buf = torch.empty(slices.shape[1:], dtype=torch.float32, device="cuda")

    def source():
        for s in slices:
            buf.copy_(torch.from_numpy(s))
            yield buf

    got = la.depth_from_source(source(), n, fused, p)
    _same(got, want)


def test_source_errors_are_the_references():
    p = la.DepthParams()
    w, h, n = 96, 64, 4
    slices, fused = _case(w, h, n, p, 3)
    dw, dh = p.grid(w, h)
    print("depth source errors", flush=True)
    with pytest.raises(la.HipError, match=r"^depth: slice 3 of 4 missing$"):
        la.depth_from_source(iter(list(slices[:3])), n, fused, p)
    bad = [slices[0], slices[1][:, :-1], slices[2], slices[3]]
    with pytest.raises(la.HipError, match=rf"^depth: slice 1 has {dh * (dw - 1)} cells, the working grid {dw}x{dh}$"):
        la.depth_from_source(bad, n, fused, p)
    err = ValueError("decode failed: frame 2")
    with pytest.raises(la.HipError, match=r"^decode failed: frame 2$") as ei:
        la.depth_from_source([slices[0], slices[1], err, slices[3]], n, fused, p)
    assert ei.value.__cause__ is err

    def failing():
        yield slices[0]
        raise OSError("read error on frame 1")

    with pytest.raises(la.HipError, match=r"^read error on frame 1$"):
        la.depth_from_source(failing(), n, fused, p)
    # the library is whole after the errors
    _same(la.depth_from_source(list(slices), n, fused, p), la.depth_from_slices(slices, n, fused, p))
