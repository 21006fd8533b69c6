"""Training data for, training of, and the measured gain of the learned WLS warm start.

A sample is one seeded 45 MP focus bracket (8256x5504 RGB u16 from
lapstack_amd.synth, per-frame similarity jitter, 24 to 100 frames) taken
through the device's stacking path with the benchmark's settings
(lsh_stack_stream: alignment, 7-level fusion, focus slices) and the depth pass
up to the solve (depth_grid: guide, peak depth, confidence on the 4128x2752
grid), with its WLS solution from the FP64 mode to a 1e-11 relative residual
(the target).  The network (WarmNet) is trained on whole systems by its own
device kernels; the evaluation solves held-out stacks from the reference's FGS
guess and from the network's start, both to the same tolerance by the same
MG-PCG, and reports iterations, device milliseconds (inference included) and
the error against the FP64 solution -- and, as the ablation, the start of an
untrained network (the blocks' weighted mean depths, prolongated).

    python -m lapstack_amd.learn --stacks 128 --steps 6000 --lr 3e-4 --out lapstack_amd/warm_net.f32
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "warm_net.f32")
W45, H45 = 8256, 5504


def depth_params():
    """The depth pass of the 45 MP workload: guided-filter aggregation r=8
    eps=1e-3, WLS lambda 3 sigma_c 0.04, PCG to a 1e-6 relative residual."""
    import lapstack_amd as la

    return la.DepthParams(agg_radius=8, agg_eps=1e-3, tol=1e-6)


def stack_planes(frames, dp, fp=None, apm=None):
    """(guide, depth, conf) on the working grid of one stack, as device
    tensors, through the stacking path (alignment, fusion, focus slices)."""
    import torch

    import lapstack_amd as la

    n = len(frames)
    h, w, _ = tuple(frames[0].shape)
    dw, dh = dp.grid(w, h)
    slices = torch.empty((n, dh, dw), dtype=torch.float32, device="cuda")
    image = torch.empty((3, h, w), dtype=torch.float32, device="cuda")
    la.stack_stream(frames, fp or la.FuseParams(levels=7), apm or la.AlignParams(coarsen=2, interp="bilinear"), dp,
                    graph=True, slices_out=slices, img_out=image)
    g, d, c, _ = la.depth_grid(slices, n, image, dp, out_device=True)
    return g.clone(), d.clone(), c.clone()


def make_problem(seed, n, dp, w=W45, h=H45):
    """(guide, depth, conf) device planes of one seeded jittered stack of n frames."""
    import torch

    from lapstack_amd import synth

    br = synth.Bracket(w, h, n, seed=seed, jitter=True, device="cuda")
    frames = [br.frame(i) for i in range(n)]
    del br
    out = stack_planes(frames, dp)
    del frames
    torch.cuda.empty_cache()
    return out


def fp64_solution(g, d, c, dp):
    """The FP64 solve to a 1e-11 relative residual (float64 device tensor)."""
    import torch

    import lapstack_amd as la

    u = torch.empty_like(d)
    x = torch.empty(d.shape, dtype=torch.float64, device="cuda")
    la.wls_solve(g, c, d, dp.lam, dp.sigma_c, "fp64", tol=1e-11, max_iters=2000, u_out=u, u64_out=x)
    return x


def solve_pair(net, g, d, c, dscale, dp, mode="reference", oracle=None, ablation=None):
    """The plain solve (FGS guess) and the solve from the network's start, to
    the same tolerance by the same MG-PCG; errors against an FP64 solve.  With
    ``ablation`` (an untrained network) also the solve from its start."""
    import torch

    import lapstack_amd as la

    u = torch.empty_like(d)
    if oracle is None:
        oracle = fp64_solution(g, d, c, dp)
    res = {}
    _, rep0, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=dp.tol, max_iters=0, u_out=u)  # the FGS guess
    for _ in range(2):  # the second run is timed warm
        _, rep, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=dp.tol, max_iters=1000, u_out=u)
    res["plain"] = dict(iters=rep["iters"], ms=rep["ms"], true_rel=rep["true_rel"],
                        err=float((u.double() - oracle).abs().max()), guess_rel=rep0["true_rel"])
    for key, m in (("learned", net), ("ablation", ablation)):
        if m is None:
            continue
        warm = torch.empty_like(d)
        m.predict(d, c, g, dscale, out=warm)  # sizes the activations
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.predict(d, c, g, dscale, out=warm)  # returns once the start is on the device
        infer_ms = 1000.0 * (time.perf_counter() - t0)
        _, repw, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=dp.tol, max_iters=0, warm=warm, u_out=u)
        for _ in range(2):
            _, rep, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=dp.tol, max_iters=1000, warm=warm, u_out=u)
        res[key] = dict(guess_rel=repw["true_rel"], iters=rep["iters"], ms=rep["ms"] + infer_ms, solve_ms=rep["ms"],
                        infer_ms=infer_ms, true_rel=rep["true_rel"], err=float((u.double() - oracle).abs().max()))
    return res


def infer_timing(net, grid):
    """Per-kernel device time of one prediction on a grid of ``WxH`` cells."""
    import torch

    from lapstack_amd import _lib

    w, h = (int(v) for v in grid.split("x"))
    planes = [torch.rand((h, w), device="cuda") for _ in range(3)]
    out = torch.empty((h, w), device="cuda")
    net.predict(*planes, 7.0, out=out)  # sizes the activations
    _lib.sync()
    _lib.prof_reset()
    _lib.prof_enable(True)
    net.predict(*planes, 7.0, out=out)
    _lib.sync()
    _lib.prof_enable(False)
    prof = _lib.prof_read()
    total = sum(v[0] for v in prof.values())
    return {"grid": [w, h], "ms": round(total, 3),
            "kernels": {k: [round(v[0], 3), v[2]] for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}}


def train(args):
    import torch

    import lapstack_amd as la

    dp = depth_params()
    t0 = time.perf_counter()
    rng = np.random.default_rng(args.seed)
    G = D = Cf = U = None
    for s in range(args.stacks):
        n = int(rng.choice([24, 40, 64, 100]))
        g, d, c = make_problem(args.seed + s, n, dp)
        x = fp64_solution(g, d, c, dp)
        if G is None:
            G, D, Cf, U = (torch.empty((args.stacks,) + tuple(d.shape), dtype=torch.float32, device="cuda")
                           for _ in range(4))
        sc = float(n - 1)  # depth / dscale: every system in frame-range units
        G[s], D[s], Cf[s], U[s] = g, d / sc, c, (x / sc).float()
        del g, d, c, x
        if (s + 1) % 16 == 0 or s + 1 == args.stacks:
            print(f"samples {s + 1}/{args.stacks} ({time.perf_counter() - t0:.1f}s)", flush=True)
    N, H, W = D.shape
    net = la.WarmNet(seed=1)
    net.set_sigma(dp.sigma_c)
    gen = torch.Generator()
    gen.manual_seed(1)
    B = args.batch
    bufs = [torch.empty((B, H, W), dtype=torch.float32, device="cuda") for _ in range(4)]
    t1 = time.perf_counter()
    losses = []
    for step in range(args.steps):
        idx = torch.randint(0, N, (B,), generator=gen).cuda()
        flip = [d for d, on in ((-1, step % 2), (-2, (step // 2) % 2)) if on]
        for buf, src in zip(bufs, (D, Cf, G, U)):
            v = src.index_select(0, idx)
            buf.copy_(v.flip(flip) if flip else v)
        lr = args.lr * (0.5 * (1.0 + np.cos(np.pi * step / args.steps)))  # cosine decay
        losses.append(net.train_step(bufs[0], bufs[1], bufs[2], bufs[3], 1.0, lr))
        if (step + 1) % 250 == 0 or step + 1 == args.steps:
            print(f"step {step + 1}/{args.steps}: loss {np.mean(losses[-250:]):.4f} "
                  f"({time.perf_counter() - t1:.1f}s)", flush=True)
    net.save(args.out)
    print(f"saved {args.out}", flush=True)
    del G, D, Cf, U, bufs
    torch.cuda.empty_cache()
    base = la.WarmNet(seed=1)  # untrained: the weighted block means
    base.set_sigma(dp.sigma_c)
    rows = []
    for s in range(args.eval):  # held-out stacks: seeds after the training range
        n = (100, 40, 64, 24)[s % 4]
        g, d, c = make_problem(args.seed + args.stacks + 10_000 + s, n, dp)
        r = solve_pair(net, g, d, c, float(n - 1), dp, ablation=base)
        r["frames"] = n
        rows.append(r)
        print(json.dumps(r), flush=True)
    infer = infer_timing(net, args.infer_grid) if args.infer_grid else None
    keys = ("plain", "learned", "ablation")
    summary = {
        "infer": infer,
        "train": {"stacks": args.stacks, "grid": [W, H], "steps": args.steps, "batch": B,
                  "final_loss": float(np.mean(losses[-250:])), "train_s": round(time.perf_counter() - t1, 1)},
        "eval_stacks": args.eval,
        **{k: {f: float(np.mean([r[k][f] for r in rows])) for f in rows[0][k]} for k in keys},
        "wins": int(sum(r["learned"]["ms"] < r["plain"]["ms"] for r in rows)),
    }
    print(json.dumps(summary), flush=True)
    return summary


def evaluate(args):
    """The shipped weights on held-out 45 MP stacks (the seeds after the
    training range): plain, learned and untrained starts, each to the same
    tolerance by the same MG-PCG."""
    import lapstack_amd as la

    dp = depth_params()
    net = la.WarmNet()
    net.load(args.weights)
    net.set_sigma(dp.sigma_c)
    base = la.WarmNet(seed=1)
    base.set_sigma(dp.sigma_c)
    rows = []
    for s in range(args.eval):
        n = (100, 40, 64, 24)[s % 4]
        g, d, c = make_problem(args.seed + args.stacks + 10_000 + s, n, dp)
        r = solve_pair(net, g, d, c, float(n - 1), dp, ablation=base)
        r["frames"] = n
        rows.append(r)
        print(json.dumps(r), flush=True)
    keys = ("plain", "learned", "ablation")
    summary = {"eval_stacks": args.eval,
               **{k: {f: float(np.mean([r[k][f] for r in rows])) for f in rows[0][k]} for k in keys},
               "wins": int(sum(r["learned"]["ms"] < r["plain"]["ms"] for r in rows))}
    print(json.dumps(summary), flush=True)
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stacks", type=int, default=128)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--seed", type=int, default=20_000)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--infer-grid", default="4128x2752", help="time one prediction on this grid ('' to skip)")
    ap.add_argument("--eval-only", action="store_true",
                    help="no training: evaluate --weights on --eval held-out stacks")
    ap.add_argument("--weights", default=WEIGHTS)
    args = ap.parse_args()
    sys.exit(0 if (evaluate(args) if args.eval_only else train(args)) else 1)


if __name__ == "__main__":
    main()
