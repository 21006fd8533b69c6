"""The device search (k_nm_search) against the oracle's multiscale_align
(align.rs:796-942), evaluation by evaluation, through the C ABI.

Bars: the same evaluations per level; every counted cost the oracle's to a
relative 1e-9 (the device sums each cost's f64 terms in another order, and its
per-evaluation maps take the device library's sin and cos); the found Sim
bit for bit the oracle's (the iterates depend on the costs only through
comparisons).  Every block that took part folded the same sums (equal
digests), also when only some of the grid's blocks take part."""
import os

import numpy as np
import pytest

import lapstack_amd as la
from lapstack_amd import synth
from oracle import ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    why = la.available()
    if why is not None:
        pytest.fail(f"HIP backend unavailable on a GPU run: {why}")


def _pair(w, h, seed):
    b = synth.Bracket(w, h, 2, seed=seed, jitter=True, device="cuda")
    f0, f1 = b.frame(0).cpu().numpy(), b.frame(1).cpu().numpy()
    return f0, f1, b.sims[1]


def _device_search(f0, f1, coarsen, interp, free=(1, 1, 1, 1, 0, 0, 0, 0)):
    h, w, _ = f0.shape
    fr = la.HipFrames(w, h, interp)
    fr.trace_costs(1 << 16)
    fr.process(f0, la.SIM_ID, free, coarsen, None, False, None)
    sim, _, _ = fr.process(f1, la.SIM_ID, free, coarsen, None, False, None)
    out = dict(sim=np.array(sim), levels=np.array(fr.search_levels()), costs=fr.search_costs(),
               blocks=fr.search_blocks(), evals=fr.last_evals)
    fr.close()
    return out


# a config-2 frame (12 MP, coarsen 2, the bilinear resample), a smaller one
# searched to full resolution with the default Spline4x4, one at coarsen 1
@pytest.mark.parametrize("w,h,coarsen,interp", [(4000, 3000, 2, "bilinear"), (1024, 768, 1, "spline4x4"),
                                                (384, 288, 0, "spline4x4")])
def test_search_follows_the_oracle_evaluation_by_evaluation(w, h, coarsen, interp):
    print(f"search {w}x{h} coarsen {coarsen}", flush=True)
    f0, f1, truth = _pair(w, h, seed=41)
    dev = _device_search(f0, f1, coarsen, interp)
    y0, y1 = ref.luma(ref.u16_to_planes(f0)), ref.luma(ref.u16_to_planes(f1))
    sim, evals, lev, costs = ref.multiscale_align_trace(y0, y1, coarsen=coarsen)
    print("levels device", dev["levels"][:8].tolist(), "oracle", lev[:8].tolist(), flush=True)
    assert dev["levels"].tolist() == lev.tolist()
    assert dev["evals"] == evals
    assert len(dev["costs"]) == len(costs) == evals
    rel = np.abs(dev["costs"] - costs) / np.maximum(np.abs(costs), 1e-30)
    print(f"max relative cost difference {rel.max():.2e} over {evals} evaluations", flush=True)
    assert rel.max() <= 1e-9
    assert np.array_equal(dev["sim"], sim), (dev["sim"], sim)
    # the blocks that took part all folded the same sums
    b = dev["blocks"]
    assert len(b) >= 16 and np.all(b == b[0])
    # and the search found the generator's transform (to half a pixel of its finest level)
    scale = 1 << coarsen
    assert abs(dev["sim"][0] - truth[0]) * w < 0.5 * scale and abs(dev["sim"][1] - truth[1]) * h < 0.5 * scale


def test_search_with_part_of_the_grid():
    """A search takes the blocks running when it starts: with the participants
    capped below the grid (as when other work holds CUs), the result is still
    the oracle's search, and the participants agree."""
    w, h = 1024, 768
    f0, f1, _ = _pair(w, h, seed=43)
    full = _device_search(f0, f1, 1, "bilinear")
    os.environ["LSH_SEARCH_MAX_BLOCKS"] = "37"
    try:
        part = _device_search(f0, f1, 1, "bilinear")
    finally:
        del os.environ["LSH_SEARCH_MAX_BLOCKS"]
    assert len(part["blocks"]) == 37 and np.all(part["blocks"] == part["blocks"][0])
    assert len(full["blocks"]) > 37
    y0, y1 = ref.luma(ref.u16_to_planes(f0)), ref.luma(ref.u16_to_planes(f1))
    sim, evals, lev, costs = ref.multiscale_align_trace(y0, y1, coarsen=1)
    for r in (full, part):
        assert r["levels"].tolist() == lev.tolist()
        assert np.array_equal(r["sim"], sim)
        assert np.abs(r["costs"] - costs).max() <= 1e-9 * np.abs(costs).max()


# align.rs:115-119 free_mask: the affine model frees aspect and shear as well, the
# projective one the two perspective terms too (8 parameters, 9-point simplices;
# a projective point turns the tile windows off for its pass)
@pytest.mark.parametrize("model,free", [("affine", (1, 1, 1, 1, 1, 1, 0, 0)), ("projective", (1, 1, 1, 1, 1, 1, 1, 1))])
def test_search_models_follow_the_oracle(model, free):
    w, h, coarsen = 768, 512, 1
    print(f"search {model} {w}x{h} coarsen {coarsen}", flush=True)
    f0, f1, _ = _pair(w, h, seed=43)
    dev = _device_search(f0, f1, coarsen, "bilinear", free=free)
    y0, y1 = ref.luma(ref.u16_to_planes(f0)), ref.luma(ref.u16_to_planes(f1))
    sim, evals, lev, costs = ref.multiscale_align_trace(y0, y1, free=free, coarsen=coarsen)
    print("levels device", dev["levels"][:8].tolist(), "oracle", lev[:8].tolist(), flush=True)
    assert dev["levels"].tolist() == lev.tolist()
    assert dev["evals"] == evals
    rel = np.abs(np.array(dev["costs"]) - costs) / np.maximum(np.abs(costs), 1e-30)
    assert rel.max() <= 1e-9, rel.max()
    assert np.array_equal(dev["sim"], sim), (dev["sim"], sim)
    b = dev["blocks"]
    assert len(b) >= 1 and all(d == b[0] for d in b)
