"""Training data for, training of, and the measured gain of the learned WLS warm start.

A sample is one seeded focus bracket (8 frames of 1024x768 RGB u16 from
lapstack_amd.synth, no jitter) taken through the device path: the fold and
focus slices (HipFrames.ingest + Fuser), the depth pass up to the solve
(depth_grid: peak depth, confidence, guide on the 512x384 working grid), and
the plain MG-PCG solve to a 1e-6 relative residual (the target).  The network
(WarmNet) is trained on random crops of the samples by its own device
kernels; the evaluation solves held-out stacks from the FGS guess and from the
network's prediction, both to the same tolerance by the same MG-PCG, and
reports iterations, device milliseconds and the error against an FP64 solve.

    python -m lapstack_amd.learn --stacks 2000 --steps 4000 --out lapstack_amd/warm_net.f32
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


def depth_params():
    """The depth pass of the 45 MP workload: guided-filter aggregation r=8
    eps=1e-3, WLS lambda 3 sigma_c 0.04, PCG to a 1e-6 relative residual."""
    import lapstack_amd as la

    return la.DepthParams(agg_radius=8, agg_eps=1e-3, tol=1e-6)


def stack_planes(frames, dp, fp=None):
    """(guide, depth, conf) on the working grid of one stack, as device tensors."""
    import torch

    import lapstack_amd as la

    n = len(frames)
    h, w, _ = tuple(frames[0].shape)
    dw, dh = dp.grid(w, h)
    slices = torch.empty((n, dh, dw), dtype=torch.float32, device="cuda")
    fr = la.HipFrames(w, h, "bilinear")
    fu = la.Fuser(w, h, fp or la.FuseParams(levels=5))
    for i, f in enumerate(frames):
        fr.ingest(f, True, dp, slices[i])
        fu.push_frames(fr)
    image = torch.empty((3, h, w), dtype=torch.float32, device="cuda")
    fu.finish(img_out=image)
    g, d, c, _ = la.depth_grid(slices, n, image, dp, out_device=True)
    fr.close()
    fu.close()
    return g, d, c


def make_sample(seed, dp, w=1024, h=768, n=8):
    """(guide, depth, conf, target) device planes of one seeded stack."""
    import torch

    import lapstack_amd as la
    from lapstack_amd import synth

    br = synth.Bracket(w, h, n, seed=seed, jitter=False, device="cuda")
    frames = [br.frame(i) for i in range(n)]
    del br
    g, d, c = stack_planes(frames, dp)
    u = torch.empty_like(d)
    la.wls_solve(g, c, d, dp.lam, dp.sigma_c, "reference", tol=dp.tol, max_iters=1000, u_out=u)
    return g, d, c, u


def solve_pair(net, g, d, c, dscale, dp, mode="reference"):
    """The plain solve (FGS guess) and the solve from the network's warm start,
    to the same tolerance by the same MG-PCG; errors against an FP64 solve to 1e-12."""
    import torch

    import lapstack_amd as la

    u = torch.empty_like(d)
    x = torch.empty(d.shape, dtype=torch.float64, device="cuda")
    la.wls_solve(g, c, d, dp.lam, dp.sigma_c, "fp64", tol=1e-12, max_iters=2000, u_out=u, u64_out=x)
    oracle = x.clone()
    res = {}
    _, rep0, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=dp.tol, max_iters=0, u_out=u)  # the FGS guess
    fgs_err = float((u.double() - oracle).abs().max())
    _, rep, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=dp.tol, max_iters=1000, u_out=u)
    res["plain"] = dict(iters=rep["iters"], ms=rep["ms"], true_rel=rep["true_rel"],
                        err=float((u.double() - oracle).abs().max()), guess_err=fgs_err, guess_rel=rep0["true_rel"])
    warm = torch.empty_like(d)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    net.predict(d, c, g, dscale, out=warm)
    e1.record()
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1)
    _, repw, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=dp.tol, max_iters=0, warm=warm, u_out=u)
    _, rep, _ = la.wls_solve(g, c, d, dp.lam, dp.sigma_c, mode, tol=dp.tol, max_iters=1000, warm=warm, u_out=u)
    res["learned"] = dict(guess_rel=repw["true_rel"], iters=rep["iters"], ms=rep["ms"] + infer_ms, solve_ms=rep["ms"], infer_ms=infer_ms,
                          true_rel=rep["true_rel"], err=float((u.double() - oracle).abs().max()),
                          guess_err=float((warm.double() - oracle).abs().max()))
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
    G, D, Cf, U = [], [], [], []
    for s in range(args.stacks):
        g, d, c, u = make_sample(args.seed + s, dp)
        G.append(g)
        D.append(d)
        Cf.append(c)
        U.append(u)
        if (s + 1) % 50 == 0 or s + 1 == args.stacks:
            print(f"samples {s + 1}/{args.stacks} ({time.perf_counter() - t0:.1f}s)", flush=True)
    G, D, Cf, U = (torch.stack(a) for a in (G, D, Cf, U))
    N, H, W = D.shape
    dscale = 7.0  # 8 frames: frame index range n - 1
    net = la.WarmNet(seed=args.seed)
    if args.loss == "energy":
        net.set_loss("energy", dp.lam, dp.sigma_c)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    cs, B = args.crop, args.batch
    bufs = [torch.empty((B, cs, cs), dtype=torch.float32, device="cuda") for _ in range(4)]
    t1 = time.perf_counter()
    losses = []
    for step in range(args.steps):
        idx = torch.randint(0, N, (B,), generator=gen, device="cuda")
        oy = torch.randint(0, H - cs + 1, (B,), generator=gen, device="cuda")
        ox = torch.randint(0, W - cs + 1, (B,), generator=gen, device="cuda")
        ar = torch.arange(cs, device="cuda")
        yy = (oy[:, None] + ar[None, :])[:, :, None]
        xx = (ox[:, None] + ar[None, :])[:, None, :]
        ii = idx[:, None, None]
        for buf, src in zip(bufs, (D, Cf, G, U)):
            buf.copy_(src[ii, yy, xx])
        lr = args.lr * (0.5 * (1.0 + np.cos(np.pi * step / args.steps)))  # cosine decay
        losses.append(net.train_step(bufs[0], bufs[1], bufs[2], bufs[3], dscale, lr))
        if (step + 1) % 250 == 0 or step + 1 == args.steps:
            print(f"step {step + 1}/{args.steps}: loss {np.mean(losses[-250:]):.3e} "
                  f"({time.perf_counter() - t1:.1f}s)", flush=True)
    net.save(args.out)
    print(f"saved {args.out}", flush=True)
    del G, D, Cf, U
    torch.cuda.empty_cache()
    # held-out stacks: seeds after the training range
    rows = []
    for s in range(args.eval):
        g, d, c, _ = make_sample(args.seed + args.stacks + 10_000 + s, dp)
        rows.append(solve_pair(net, g, d, c, dscale, dp))
    infer = infer_timing(net, args.infer_grid) if args.infer_grid else None
    summary = {
        "infer": infer,
        "train": {"loss": args.loss, "stacks": args.stacks, "grid": [W, H], "steps": args.steps, "batch": B, "crop": cs,
                  "final_loss": float(np.mean(losses[-250:])), "train_s": round(time.perf_counter() - t1, 1)},
        "eval_stacks": args.eval,
        "plain": {k: float(np.mean([r["plain"][k] for r in rows])) for k in rows[0]["plain"]},
        "learned": {k: float(np.mean([r["learned"][k] for r in rows])) for k in rows[0]["learned"]},
    }
    print(json.dumps(summary), flush=True)
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stacks", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=20_000)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--loss", choices=["mse", "energy"], default="energy",
                    help="energy: the error's energy norm under the WLS operator (what MG-PCG reduces)")
    ap.add_argument("--infer-grid", default="4128x2752", help="time one prediction on this grid ('' to skip)")
    args = ap.parse_args()
    sys.exit(0 if train(args) else 1)


if __name__ == "__main__":
    main()
