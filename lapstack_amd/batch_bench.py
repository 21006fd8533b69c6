"""Many uneven stacks in one persistent launch against a per-stack launch loop.

64 independent seeded stacks: frame counts uniform in [4, 40], frame sizes
uniform in [2, 12] MP with aspect 3:2 or 4:3, interleaved RGB u16 resident in
HBM; pyramid depth per stack floor(log2(min side)) - 4.  The fuse path only
(the frame as decoded, REDUCE chain, region-energy fold, residual store).

Each stack is a seeded focus bracket of the benchmark's generator
(synth.Bracket at the stack's own size, no jitter): its own 1/f scene and step
edges, its own depth map, frame i blurred per pixel by 6 px |depth - f_i|,
noise, u16 quantisation; stack k's seed is s0 + 1000000 k.

Timed, after a warm-up of both paths:
  loop   -- per stack, per frame: lsh_frames_ingest + lsh_fuser_push_frames
            (handles made beforehand), one stack after another;
  batch  -- one lsh_batch_fuse call (its set-up, the persistent launch and
            its synchronisation; fusers made beforehand), with the autotuner's
            cached choice; the autotuner's own run (tune=2) is reported.
Neither includes the collapse (lsh_fuser_finish), which is the same for both.
Every stack's fused image and winner map are compared between the two.

    python -m lapstack_amd.batch_bench [--stacks 64] [--seed 7]
"""
from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import hashlib
import json
import math
import time

import numpy as np


def plan(n_stacks, seed):
    """[(w, h, frames, levels)] of the seeded batch."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_stacks):
        frames = int(rng.integers(4, 41))
        mp = float(rng.uniform(2.0, 12.0)) * 1e6
        ar = (3, 2) if rng.random() < 0.5 else (4, 3)
        h = int(round(math.sqrt(mp * ar[1] / ar[0])))
        w = int(round(h * ar[0] / ar[1]))
        levels = max(1, int(math.floor(math.log2(min(w, h)))) - 4)
        out.append((w, h, frames, levels))
    return out


def make_stack(w, h, n, seed):
    """n frames (h, w, 3) torch.uint16 on cuda: synth.Bracket(w, h, n, seed)."""
    import torch

    from lapstack_amd import synth

    b = synth.Bracket(w, h, n, seed=seed, jitter=False, device="cuda")
    frames = [b.frame(i) for i in range(n)]
    del b
    torch.cuda.empty_cache()
    return frames


def _digest(img, win):
    return hashlib.sha1(np.ascontiguousarray(img).tobytes() + np.ascontiguousarray(win).tobytes()).hexdigest()


def main():
    import torch

    import lapstack_amd as la
    from lapstack_amd import _lib

    ap = argparse.ArgumentParser()
    ap.add_argument("--stacks", type=int, default=64)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--only-batch", metavar="TILE_H,BLOCKS_PER_CU,FRAMES_PER_ROUND", default=None,
                    help="one batched launch with this configuration, no loop and no tuning (for profiling)")
    args = ap.parse_args()
    shapes = plan(args.stacks, args.seed)
    t0 = time.perf_counter()
    stacks = []
    for k, (w, h, n, _) in enumerate(shapes):
        stacks.append(make_stack(w, h, n, args.seed + 1000000 * k))
        if (k + 1) % 8 == 0:
            print(f"generated {k + 1}/{len(shapes)} stacks ({time.perf_counter() - t0:.1f}s)", flush=True)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    mpf = sum(w * h * n for w, h, n, _ in shapes) / 1e6
    levels = [lv for _, _, _, lv in shapes]
    fp = la.FuseParams()
    lib = _lib.load()

    if args.only_batch:
        ty, pc, nf = (int(v) for v in args.only_batch.split(","))
        out, rep = la.batch_fuse(stacks, levels, fp, tune=0, tile_h=ty, blocks_per_cu=pc, frames_per_round=nf)
        print(json.dumps({"mpix_frames": round(mpf, 1), "batch_device_ms": round(rep["ms"], 2),
                          "batch_device_mpix_frames_per_s": round(mpf / (rep["ms"] / 1e3), 1),
                          "config": {k: rep[k] for k in ("tile_h", "blocks_per_cu", "resident_blocks",
                                                         "frames_per_round", "rounds")}}), flush=True)
        return

    # warm-up: both paths on two small stacks
    small = sorted(range(len(shapes)), key=lambda k: shapes[k][0] * shapes[k][1])[:2]
    la.batch_fuse([stacks[k][:4] for k in small], [levels[k] for k in small], fp, tune=0)
    for k in small:
        w, h, _, lv = shapes[k]
        fr, fu = la.HipFrames(w, h, "bilinear"), la.Fuser(w, h, dataclasses.replace(fp, levels=lv))
        for f in stacks[k][:4]:
            fr.ingest(f, False, None, None)
            fu.push_frames(fr)
        fu.finish()
        fr.close()
        fu.close()
    _lib.sync()

    # the per-stack loop
    want = []
    for k, ((w, h, n, lv), frames) in enumerate(zip(shapes, stacks)):
        fr, fu = la.HipFrames(w, h, "bilinear"), la.Fuser(w, h, dataclasses.replace(fp, levels=lv))
        _lib.sync()
        t = time.perf_counter()
        for f in frames:
            fr.ingest(f, False, None, None)
            fu.push_frames(fr)
        _lib.sync()
        dt = time.perf_counter() - t
        want.append((dt, _digest(*fu.finish())))
        fr.close()
        fu.close()
    t_loop = sum(dt for dt, _ in want)
    print(f"per-stack loop: {t_loop:.3f}s", flush=True)

    # the batched launch: one call over all stacks (fusers made beforehand, collapse after)
    def batched(tune):
        fusers = [la.Fuser(w, h, dataclasses.replace(fp, levels=lv)) for (w, h, _, lv) in shapes]
        descs, keep = (_lib.BatchStackC * len(shapes))(), []
        for i, frames in enumerate(stacks):
            ptrs = (C.c_void_p * len(frames))(*[f.data_ptr() for f in frames])
            keep.append(ptrs)
            descs[i].w, descs[i].h, descs[i].n = shapes[i][0], shapes[i][1], len(frames)
            descs[i].frames = C.cast(ptrs, C.c_void_p)
        handles = (C.c_void_p * len(shapes))(*[fu._h.value for fu in fusers])
        opt, rep = _lib.BatchOptionsC(tune, 0, 0, 0), _lib.BatchReportC()
        _lib.sync()
        t = time.perf_counter()
        _lib.check(lib.lsh_batch_fuse(handles, descs, len(shapes), C.byref(opt), C.byref(rep)))
        dt = time.perf_counter() - t
        digests = [_digest(*fu.finish()) for fu in fusers]
        for fu in fusers:
            fu.close()
        tried = [(c & 0xFF, (c >> 8) & 0xFF, c >> 16, round(float(u), 3))
                 for c, u in zip(list(rep.tried_cfg)[:rep.ntried], list(rep.tried_us_per_mpf)[:rep.ntried])]
        return dt, rep, digests, tried

    t_tune, rep_t, dig_t, tried = batched(2)
    print(f"batched launch with the autotuner: {t_tune:.3f}s, tried (tile_h, blocks/CU, frames/round, "
          f"us per MP-frame of round 0): {tried}", flush=True)
    t_batch, rep, dig, _ = batched(1)
    print(f"batched launch (cached choice): {t_batch:.3f}s (device {rep.ms:.1f} ms)", flush=True)
    same = sum(a == b == c for (_, a), b, c in zip(want, dig, dig_t))
    line = {
        "metric": "aggregate megapixel-frames/s, 64 uneven stacks, fuse path",
        "workload": f"{len(shapes)} stacks, 4-40 frames, 2-12 MP (3:2 / 4:3), RGB u16 resident in HBM, each a "
                    f"seeded synth.Bracket; levels floor(log2(min side)) - 4",
        "mpix_frames": round(mpf, 1),
        "batch_mpix_frames_per_s": round(mpf / t_batch, 1),
        "batch_device_mpix_frames_per_s": round(mpf / (rep.ms / 1e3), 1),
        "loop_mpix_frames_per_s": round(mpf / t_loop, 1),
        "speedup_vs_loop": round(t_loop / t_batch, 3),
        "batch_s": round(t_batch, 3), "batch_device_ms": round(rep.ms, 2), "loop_s": round(t_loop, 3),
        "batch_with_tuning_s": round(t_tune, 3),
        "config": {"tile_h": rep.tile_h, "blocks_per_cu": rep.blocks_per_cu, "resident_blocks": rep.resident_blocks,
                   "frames_per_round": rep.frames_per_round,
                   "rounds": rep.rounds, "from_cache": rep.from_cache},
        "autotuner_tried": [{"tile_h": a, "blocks_per_cu": b, "frames_per_round": c, "us_per_mpix_frame_round0": u}
                            for a, b, c, u in tried],
        "stacks_identical_to_loop": f"{same}/{len(shapes)}",
    }
    print(json.dumps(line), flush=True)
    if same != len(shapes):
        raise SystemExit("batched output differs from the per-stack loop")


if __name__ == "__main__":
    main()
