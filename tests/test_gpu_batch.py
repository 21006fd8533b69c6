"""Many uneven stacks in one persistent launch (lsh_batch_fuse) against the
per-stack loop (lsh_frames_ingest + lsh_fuser_push_frames per frame, one stack
after another): the fused images and winner maps are identical, bit for bit,
for every fold tile height, grid and frames-per-round setting, and whatever
the autotuner picks (its trial runs replay round 0 before the real run)."""
import dataclasses

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


def _levels(w, h):
    return max(1, int(np.floor(np.log2(min(w, h)))) - 4)


def _per_stack(frames, levels, fp):
    h, w, _ = tuple(frames[0].shape)
    fr, fu = la.HipFrames(w, h, "bilinear"), la.Fuser(w, h, dataclasses.replace(fp, levels=levels))
    for f in frames:
        fr.ingest(f, False, None, None)
        fu.push_frames(fr)
    out = fu.finish()
    fr.close()
    fu.close()
    return out


# frame counts 1..5 (not multiples of 2 or 4: the last round of a stack is partial), odd sizes
SHAPES = [(160, 120, 3), (384, 256, 5), (200, 150, 1), (96, 72, 2), (320, 213, 4)]
PARAMS = {"default": la.FuseParams(), "r2_chroma": la.FuseParams(energy_radius=2, use_chroma=True)}


@pytest.fixture(scope="module")
def _stacks():
    stacks = []
    for k, (w, h, n) in enumerate(SHAPES):
        b = synth.Bracket(w, h, n, seed=100 + k, device="cuda")
        stacks.append([b.frame(i) for i in range(n)])
    levels = [_levels(w, h) for w, h, _ in SHAPES]
    refs = {name: [_per_stack(fr, levels[k], fp) for k, fr in enumerate(stacks)] for name, fp in PARAMS.items()}
    return stacks, levels, refs


# (params, tune, tile_h, blocks_per_cu, frames_per_round)
CASES = [
    ("default", 0, 16, 0, 1),
    ("default", 0, 8, 1, 2),
    ("default", 0, 16, 2, 4),
    ("default", 0, 8, 0, 3),
    ("default", 2, 0, 0, 0),
    ("r2_chroma", 0, 16, 0, 2),
    ("r2_chroma", 0, 8, 0, 4),
]


@pytest.mark.parametrize("name,tune,ty,pc,nf", CASES)
def test_batch_matches_the_per_stack_loop(_stacks, name, tune, ty, pc, nf):
    stacks, levels, refs = _stacks
    print(f"batch {name} tune={tune} tile_h={ty} blocks_per_cu={pc} frames_per_round={nf}", flush=True)
    fp = PARAMS[name]
    out, rep = la.batch_fuse(stacks, levels, fp, tune=tune, tile_h=ty, blocks_per_cu=pc, frames_per_round=nf)
    print(rep, flush=True)
    assert rep["tile_h"] in (8, 16) and 1 <= rep["frames_per_round"] <= 4 and rep["blocks_per_cu"] >= 1
    assert rep["rounds"] == -(-5 // rep["frames_per_round"])
    if tune == 0:
        assert (rep["tile_h"], rep["frames_per_round"]) == (ty, nf) and rep["ntried"] == 0
    else:
        assert rep["ntried"] >= 3 and rep["from_cache"] == 0
    for k in range(len(stacks)):
        img, win = refs[name][k]
        assert np.array_equal(out[k][0], img), f"stack {k}: image differs"
        assert np.array_equal(out[k][1], win), f"stack {k}: winner differs"
    if tune:
        out2, rep2 = la.batch_fuse(stacks, levels, fp, tune=1)
        assert rep2["from_cache"] == 1 and rep2["ntried"] == 0
        assert all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(out, out2))


def test_batch_errors():
    b = synth.Bracket(64, 48, 2, seed=1, device="cuda")
    frames = [b.frame(i) for i in range(2)]
    with pytest.raises(la.HipError, match="halo"):
        la.batch_fuse([frames], [3], la.FuseParams(halo=2.0))
    with pytest.raises(la.HipError, match="device tensors"):
        la.batch_fuse([[frames[0].cpu()]], [3])
    with pytest.raises(la.HipError, match="tile_h"):
        la.batch_fuse([frames], [3], tune=0, tile_h=12)
    with pytest.raises(la.HipError, match="frames_per_round"):
        la.batch_fuse([frames], [3], tune=0, frames_per_round=5)
