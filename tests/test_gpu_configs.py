"""The benchmark configurations at their own settings and shapes, through the
C ABI, against the CPU oracle on the same seeded inputs.

* config 2: 4 jittered frames of 4000x3000, 7 levels, registration on the
  1/4-resolution luma pyramid (coarsen 2), Spline4x4 resample, guided filter
  r=8 eps=1e-3, MG-PCG to 1e-6 -- the whole stack through the replayed graph;
* config 3: 3 jittered frames of 8256x5504 with the benchmark's settings
  (bilinear resample), through the replayed graph;
* config 4: stacks of the batched configuration (its seeded plan, the frame
  generator of the benchmark) through one persistent launch;
* config 5: the WLS system of the 45 MP grid in the three precisions.

Bars: per frame the oracle's Sim bit for bit and its evaluation count, its
gains bit for bit; the winner maps and the fused floats identical (the fold
is the oracle's on identical frames); depth within 1e-3 of the frame-index
range and confidence within 1e-3 on >= 99.9 % of pixels; the batched stacks'
fused images and winner maps identical to the oracle's fold."""
import math

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


def _frames(w, h, n, seed, jitter=True):
    import torch

    b = synth.Bracket(w, h, n, seed=seed, jitter=jitter, device="cuda")
    dev = [b.frame(i) for i in range(n)]
    del b
    torch.cuda.empty_cache()
    return dev, np.stack([f.cpu().numpy() for f in dev])


def _run_both(w, h, n, seed, interp):
    import torch

    fp = la.FuseParams(levels=7)
    ap = la.AlignParams(coarsen=2, interp=interp)
    dp = la.DepthParams(agg_radius=8, agg_eps=1e-3, tol=1e-6)
    dev, host = _frames(w, h, n, seed)
    dw, dh = dp.grid(w, h)
    slices = torch.empty((n, dh, dw), dtype=torch.float32, device="cuda")
    out = la.stack_stream(dev, fp, ap, dp, graph=True, slices_out=slices)
    img = torch.from_numpy(np.ascontiguousarray(out["image"])).cuda()
    dm = la.depth_from_slices(slices, n, img, dp)
    got = dict(sims=out["sims"], gains=out["gains"], evals=out["evals"].astype(np.int64), image=out["image"],
               winner=out["winner"], depth=dm.depth, conf=dm.conf, info=dm.info, slices=slices, fused=img)
    want = ref.stack_run(host, ref.FuseParams.make(levels=7), ref.StackParams.make(align=True, coarsen=2, interp=interp),
                         ref.DepthParams.make(agg_radius=8, agg_eps=1e-3, tol=1e-6))
    return got, want


def _check(got, want, n):
    print("evals device", got["evals"].tolist(), "oracle", want["evals"].tolist(), flush=True)
    assert got["evals"][1:].tolist() == want["evals"][1:].tolist()
    assert np.array_equal(got["sims"], want["sims"]), (got["sims"], want["sims"])
    assert np.array_equal(got["gains"], want["gains"]), (got["gains"], want["gains"])
    wd = int((got["winner"] != want["winner"]).sum())
    fd = int((got["image"] != want["image"]).sum())
    print(f"winner pixels that differ: {wd}; fused values that differ: {fd}", flush=True)
    assert wd == 0 and fd == 0
    dd = np.abs(got["depth"] - want["depth"])
    cd = np.abs(got["conf"] - want["conf"])
    fr_d = (dd <= 1e-3 * (n - 1)).mean()
    fr_c = (cd <= 1e-3).mean()
    print(f"depth within 1e-3 (n-1): {fr_d:.6f} (max {dd.max():.2e}); conf within 1e-3: {fr_c:.6f}; "
          f"WLS iterations {got['info']['iters1']}/{got['info']['iters2']} against {want['info'].iters1}/"
          f"{want['info'].iters2}", flush=True)
    assert fr_d >= 0.999 and fr_c >= 0.999


def test_config2_stream_matches_the_oracle():
    w, h, n = 4000, 3000, 4
    print(f"config 2: {n} x {w}x{h}", flush=True)
    got, want = _run_both(w, h, n, seed=2, interp="spline4x4")
    _check(got, want, n)


@pytest.fixture(scope="module")
def config3():
    w, h, n = 8256, 5504, 3
    print(f"config 3: {n} x {w}x{h}", flush=True)
    got, want = _run_both(w, h, n, seed=3, interp="bilinear")
    return got, want, n


def test_config3_stream_matches_the_oracle(config3):
    got, want, n = config3
    _check(got, want, n)


def test_config5_wls_precisions_on_the_45mp_grid(config3):
    """The 45 MP stack's WLS system (4128 x 2752) solved to 1e-6 by the FP32
    reference MG-PCG, the FP64-refined one and the FP64 solver, each against an
    FP64 solve to 1e-12; the reference mode also against the oracle's solver on
    the same system."""
    got, _, n = config3
    dp = la.DepthParams(agg_radius=8, agg_eps=1e-3, tol=1e-6)
    g, d, c, info = la.depth_grid(got["slices"], n, got["fused"], dp)
    res = {}
    for mode in ("reference", "refined", "fp64"):
        u, rep, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=1e-6,
                                 max_iters=dp.cg_iters if mode == "reference" else 2000, want_err=True)
        res[mode] = (u, rep)
        print(f"WLS {mode}: iterations {rep['iters']} outer {rep['outer']} rel {rep['rel']:.2e} true_rel "
              f"{rep['true_rel']:.2e} err vs FP64 {rep['err_fp64']:.2e} ({rep['ms']:.1f} ms)", flush=True)
    scale = float(n - 1)
    assert res["reference"][1]["rel"] <= 1e-6
    assert res["refined"][1]["true_rel"] <= 1e-6 and res["fp64"][1]["true_rel"] <= 1e-6
    for mode in res:
        assert 0.0 <= res[mode][1]["err_fp64"] <= 1e-2 * scale
    # the oracle's MG-PCG (depth.rs:958-1078) on the same system
    s = ref.WlsSolver(g, dp.lam, dp.sigma_c)
    s.set_weights(c)
    uo, rel, it = s.solve(d, max_iters=dp.cg_iters, tol=1e-6)
    du = np.abs(res["reference"][0] - uo).max()
    print(f"oracle MG-PCG: {it} iterations, rel {rel:.2e}; max |device - oracle| {du:.2e}", flush=True)
    assert abs(int(res["reference"][1]["iters"]) - int(it)) <= 2
    assert du <= 1e-3 * scale


def test_config4_batched_stacks_match_the_oracle_fold():
    """The three smallest stacks (by pixel-frames) of the batched
    configuration's seeded plan (frame counts 4-40, 2-12 MP, 3:2 or 4:3, depth
    floor(log2(min side)) - 4), made by the benchmark's frame generator, in one
    persistent launch: each fused image and winner map the oracle's fold of
    the same frames, bit for bit."""
    from lapstack_amd import batch_bench

    plan = batch_bench.plan(64, 7)
    pick = sorted(range(len(plan)), key=lambda i: plan[i][0] * plan[i][1] * plan[i][2])[:3]
    stacks, host, levels = [], [], []
    for i in pick:
        w, h, n, lv = plan[i]
        dev, hf = _frames(w, h, n, 7 + 1000000 * i, jitter=False)
        stacks.append(dev)
        host.append(hf)
        levels.append(lv)
        print(f"stack {i}: {n} x {w}x{h}, {lv} levels", flush=True)
    out, rep = la.batch_fuse(stacks, levels, la.FuseParams(), tune=1)
    for (img, win), hf, lv in zip(out, host, levels):
        n, h, w, _ = hf.shape
        f = ref.Fuser(w, h, ref.FuseParams.make(levels=lv))
        for k in range(n):
            f.push(ref.u16_to_planes(hf[k]))
        wi, ww = f.finish()[:2]
        assert np.array_equal(img, wi) and np.array_equal(win, ww)
    assert math.isfinite(rep["ms"]) and rep["rounds"] >= 1


def test_config3_depth_pass_with_the_learned_start(config3):
    """The depth pass of the 45 MP stack with the shipped learned warm start
    (the first solve from its prediction, the reweighted one from the first's
    result), in the refined mode: both solves reach the tolerance (true
    relative residual), and the depth stands against the oracle's (FGS start,
    FP32 MG-PCG) at the depth bar, like the FGS-started pass."""
    import os

    from lapstack_amd import learn

    got, want, n = config3
    assert os.path.exists(learn.WEIGHTS)
    dp = la.DepthParams(agg_radius=8, agg_eps=1e-3, tol=1e-6, precision="refined")
    plain = la.depth_from_slices(got["slices"], n, got["fused"], dp)
    net = la.WarmNet()
    net.load(learn.WEIGHTS)
    net.set_sigma(dp.sigma_c)
    dp.warm = net
    dm = la.depth_from_slices(got["slices"], n, got["fused"], dp)
    info = dm.info
    print(f"learned start: iterations {info['iters1']}/{info['iters2']} (FGS {plain.info['iters1']}/"
          f"{plain.info['iters2']}), true residuals {info['true_rel1']:.2e}/{info['true_rel2']:.2e}", flush=True)
    assert info["true_rel1"] <= 1e-6 and info["true_rel2"] <= 1e-6
    for d in (dm.depth, plain.depth):
        dd = np.abs(d - want["depth"])
        assert (dd <= 1e-3 * (n - 1)).mean() >= 0.999, dd.max()
    assert np.abs(dm.depth - plain.depth).max() <= 1e-3 * (n - 1)
    net.close()
