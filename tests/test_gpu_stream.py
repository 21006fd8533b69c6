"""The streamed stack (lsh_stack_stream): no host round trip per frame; each
frame's serial chain (search, gains, warp) runs beside the previous frame's
focus slice and fold, as two streams or as a replayed two-branch hipGraph.

Bars: the graph replay and the per-frame launches of the same step agree
bit for bit (the same kernels on the same data), and so does the ordinary
per-frame path (lsh_frames_process, which derives each frame's inverse map and
gains on the host): the same Sims, evaluations, fused image, winner map and
slices."""
import numpy as np
import pytest

import lapstack_amd as la
from lapstack_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    why = la.available()
    if why is not None:
        pytest.fail(f"HIP backend unavailable on a GPU run: {why}")


def _u16(img):
    return np.floor(np.clip(img, 0.0, 1.0) * 65535.0 + 0.5)


def _per_frame(frames, fp, ap, dp):
    h, w, _ = tuple(frames[0].shape)
    fr, fu = la.HipFrames(w, h, ap.interp), la.Fuser(w, h, fp)
    dw, dh = dp.grid(w, h)
    slices = np.zeros((len(frames), dh, dw), np.float32)
    last, sims, evals = la.SIM_ID, [], []
    for i, f in enumerate(frames):
        last, _, _ = fr.process(f, last, ap.free(), ap.coarsen, None, True, dp, slices[i])
        sims.append(last)
        evals.append(fr.last_evals)
        fu.push_frames(fr)
    image, winner = fu.finish()
    fr.close()
    fu.close()
    return dict(image=image, winner=winner, sims=np.array(sims), evals=np.array(evals), slices=slices)


# 5 and 4 frames: each parity's graph captured once; 9 frames: both replayed (frames 2..7 replay, 8 runs eagerly);
# ragged sizes (odd, no multiple of any tile) with the 6- and 4-tap kernels, through every edge tile
@pytest.mark.parametrize("w,h,n,interp", [(384, 288, 5, "bilinear"), (320, 240, 4, "spline4x4"),
                                          (256, 192, 9, "bilinear"), (301, 217, 6, "lanczos3"),
                                          (257, 163, 5, "bicubic")])
def test_stream_matches_the_per_frame_path(w, h, n, interp):
    print(f"stream {w}x{h}x{n} {interp}", flush=True)
    b = synth.Bracket(w, h, n, seed=17, jitter=True, device="cuda")
    frames = [b.frame(i) for i in range(n)]
    fp, ap, dp = la.FuseParams(levels=5), la.AlignParams(coarsen=1, interp=interp), la.DepthParams()
    ref = _per_frame(frames, fp, ap, dp)
    g = la.stack_stream(frames, fp, ap, dp, graph=True)
    p = la.stack_stream(frames, fp, ap, dp, graph=False)
    # graph replay and per-frame launches: the same kernels, bit for bit
    for k in ("image", "winner", "sims", "gains", "evals", "slices"):
        assert np.array_equal(np.asarray(g[k]), np.asarray(p[k])), k
    # against the host-steered path: the maps take the same correctly rounded
    # sin and cos on the device and on the host, so every value is the same
    assert np.array_equal(g["sims"], ref["sims"])
    assert np.array_equal(g["evals"][1:], ref["evals"][1:])
    assert np.array_equal(g["image"], ref["image"])
    assert np.array_equal(g["winner"], ref["winner"])
    assert np.array_equal(g["slices"], ref["slices"])
    print("evals stream", g["evals"].tolist(), "per-frame", ref["evals"].tolist(), flush=True)


def test_stream_device_buffers_and_single_frame():
    import torch

    w, h = 160, 120
    b = synth.Bracket(w, h, 1, seed=3, device="cuda")
    out = la.stack_stream([b.frame(0)], la.FuseParams(levels=4), la.AlignParams(coarsen=0, interp="bilinear"), None)
    assert out["sims"].shape == (1, 8) and out["evals"][0] == 0
    frames = [synth.Bracket(w, h, 3, seed=5, jitter=True, device="cuda").frame(i) for i in range(3)]
    img = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
    win = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    s = la.stack_stream(frames, la.FuseParams(levels=4), la.AlignParams(coarsen=0, interp="bilinear"), None,
                        img_out=img, winner_out=win)
    assert float(img.max()) > 0.0 and s["evals"][1] > 0


@pytest.mark.parametrize("n", [3, 8, 11])
def test_graph_reads_frames_through_the_table_and_writes_device_slices(n):
    """A replayed graph takes device frames through its device table of frame
    pointers and writes device slices at the fold's index; host frames go
    through the staging buffers.  All equal the launched path bit for bit."""
    import torch

    print(f"graph table {n}", flush=True)
    w, h = 288, 208
    b = synth.Bracket(w, h, n, seed=23, jitter=True, device="cuda")
    frames = [b.frame(i) for i in range(n)]
    fp, ap, dp = la.FuseParams(levels=5), la.AlignParams(coarsen=1, interp="bilinear"), la.DepthParams()
    dw, dh = dp.grid(w, h)
    ref = la.stack_stream(frames, fp, ap, dp, graph=False)
    sd = torch.full((n, dh, dw), -1.0, dtype=torch.float32, device="cuda")
    g = la.stack_stream(frames, fp, ap, dp, graph=True, slices_out=sd)
    # host frames: pageable numpy arrays and pinned tensors, uploaded through the ring on the copy stream
    hf = la.stack_stream([f.cpu().numpy() for f in frames], fp, ap, dp, graph=True)
    pinned = [f.cpu().pin_memory() for f in frames]
    hp = la.stack_stream(pinned, fp, ap, dp, graph=True)
    hl = la.stack_stream(pinned, fp, ap, dp, graph=False)
    for k in ("image", "winner", "sims", "gains", "evals"):
        assert np.array_equal(np.asarray(g[k]), np.asarray(ref[k])), k
        for other in (hf, hp, hl):
            assert np.array_equal(np.asarray(other[k]), np.asarray(ref[k])), k
    assert np.array_equal(sd.cpu().numpy(), np.asarray(ref["slices"]))
    for other in (hf, hp, hl):
        assert np.array_equal(np.asarray(other["slices"]), np.asarray(ref["slices"]))


@pytest.mark.parametrize("ext", ["tif", "png"])
def test_stream_from_files_equals_the_device_frames(tmp_path, ext):
    """lsh_stack_stream_files: frames written as 16-bit files, decoded by host
    threads ahead of their use into pinned buffers and uploaded beside the
    device's work -- the same result as the frames handed over in HBM."""
    print(f"stream from {ext} files", flush=True)
    w, h, n = 320, 224, 7
    b = synth.Bracket(w, h, n, seed=29, jitter=True, device="cuda")
    frames = [b.frame(i) for i in range(n)]
    paths = []
    for i, f in enumerate(frames):
        p = str(tmp_path / f"frame{i:03d}.{ext}")
        la.save_rgb_u16(f.cpu().numpy(), p)
        paths.append(p)
    fp, ap, dp = la.FuseParams(levels=5), la.AlignParams(coarsen=1, interp="bilinear"), la.DepthParams()
    ref = la.stack_stream(frames, fp, ap, dp, graph=True)
    for threads in (1, 4):
        got = la.stack_stream_files(paths, fp, ap, dp, graph=True, threads=threads)
        for k in ("image", "winner", "sims", "gains", "evals", "slices"):
            assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), (k, threads)
    launched = la.stack_stream_files(paths, fp, ap, dp, graph=False, threads=2)
    for k in ("image", "winner", "sims", "gains", "evals", "slices"):
        assert np.array_equal(np.asarray(launched[k]), np.asarray(ref[k])), k
    # a frame of another size ends the call with an error
    la.save_rgb_u16(np.zeros((h, w - 2, 3), np.uint16), paths[4])
    with pytest.raises(la.HipError, match="the stack is"):
        la.stack_stream_files(paths, fp, ap, dp, graph=True, threads=3)
    la.save_rgb_u16(np.zeros((h, w + 2, 3), np.uint16), paths[4])
    with pytest.raises(la.HipError, match="larger than the frame buffer"):
        la.stack_stream_files(paths, fp, ap, dp, graph=True, threads=3)


def test_graph_calls_return_their_output_memory():
    """Repeated graph-path calls, each with new device slices and image that the
    caller frees afterwards, give that memory back to the device: the graph's
    captured launches name no caller buffer (the focus slices find the slice
    stack through a device cell), so nothing the runtime keeps for a graph's
    arguments pins it.  With the stack's address as a captured argument every
    call lost its whole slice stack (36 MB here).  Also the library's kept
    blocks stay within the cache it reports."""
    import torch

    from lapstack_amd import _lib

    w, h, n = 4000, 3000, 3
    br = synth.Bracket(w, h, n, seed=17, jitter=True, device="cuda")
    frames = [br.frame(i) for i in range(n)]
    del br
    dp = la.DepthParams()
    dw, dh = dp.grid(w, h)

    def call():
        sl = torch.empty((n, dh, dw), dtype=torch.float32, device="cuda")
        img = torch.empty((3, h, w), dtype=torch.float32, device="cuda")
        la.stack_stream(frames, la.FuseParams(levels=5), la.AlignParams(coarsen=1), dp, graph=True, slices_out=sl,
                        img_out=img)
        del sl, img
        _lib.sync()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        return torch.cuda.mem_get_info()[0]

    call()
    first = call()
    for i in range(4):
        print(f"graph call {i}", flush=True)
        last = call()
    lost = first - last
    assert lost < 8 << 20, f"{lost / 2**20:.1f} MiB of device memory not returned over 4 calls"
    live, cached = _lib.mem_stats()
    total = torch.cuda.mem_get_info()[1]
    assert cached <= max(total // 8, 4 << 30)
    _lib.mem_release()
    assert _lib.mem_stats()[1] == 0
