"""Trains the learned denoiser on the GPU and writes its weights.

    python -m retrocam_amd.train_denoiser [--scenes 1024] [--levels 4] [--frame-seeds 64]
        [--ref-spp 2048] [--steps 120000] [--batch 8] [--out retrocam_amd/weights/unet_denoiser.npy]

(the defaults are the settings of the committed weights)

Training data, rendered by this package's path tracer: seeded views from
workloads.training_scene, each giving --levels inputs at increasing sample
counts from 4..256 (cumulative reads of one accumulation) that share a target at
--ref-spp with FP64 accumulation; and 128x128 crops of full frames of the
benchmark scene generator (workloads.full_frame, seeds 2 .. 1 + --frame-seeds;
seed 1, the benchmark frame, is never used) with the same input levels. The
loss is minimised by the targets' mean, so their remaining noise averages out
over the dataset. The optimizer runs entirely in
libretrocam_hip.so (rc_net_train: forward, backward and Adam kernels).
Evaluation: relMSE on held-out views (seeds >= 1,000,000) against their
high-spp renders, for the noisy input, a-trous and the network.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _native as N
from . import workloads
from .batch import ViewBatch, ViewRequest
from .denoiser import WEIGHTS_PATH, NetDenoiser, make_dataset, make_frame_crops, rel_mse
from .post import TemporalInputs, make_params
from .renderer import RenderContext


def evaluate(ctx, net, seeds, size, spp_in, ref_spp):
    """relMSE of the noisy input, a-trous (5 passes, default settings) and the
    network on views the training never saw, against ref_spp renders."""
    seeds = list(seeds)
    raw, tgt = make_dataset(ctx, seeds, size=size, spp_in=(spp_in,), spp_ref=ref_spp)
    vb = ViewBatch(ctx)
    ws = [workloads.training_scene(int(s), size=size) for s in seeds]
    for k, w in enumerate(ws):
        vb.upload_scene(k, w.scene)
    vb.render([ViewRequest(k, w.camera, w.params, spp_in) for k, w in enumerate(ws)], reset=True)
    pp = make_params(atrous=True, iterations=5)
    rows = []
    for k in range(len(ws)):
        vb.post(k, pp, TemporalInputs())
        at = vb.read(k, N.RC_BUF_POST_HDR)[:, :3].reshape(size, size, 3)
        out = net.forward(raw[k:k + 1])[0]
        rows.append((rel_mse(raw[k, ..., 0:3], tgt[k]), rel_mse(at, tgt[k]), rel_mse(out, tgt[k])))
    return np.array(rows)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=1024)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--ref-spp", type=int, default=2048)
    ap.add_argument("--eval-ref-spp", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=120000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eval-scenes", type=int, default=16)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    ap.add_argument("--levels", type=int, default=4, help="inputs per training view (increasing spp)")
    ap.add_argument("--frame-seeds", type=int, default=64,
                    help="full-frame scenes (seeds 2..) cut into crops; seed 1 is the benchmark frame, never used")
    ap.add_argument("--summary", default=None, help="write a JSON summary here")
    ap.add_argument("--init", default=None, help="start from these weights (.npy) instead of a seeded init")
    ap.add_argument("--scene-offset", type=int, default=0, help="training views use seeds offset .. offset+scenes-1")
    ap.add_argument("--frame-offset", type=int, default=2,
                    help="full-frame seeds offset .. offset+frame-seeds-1 (seed 1, the benchmark frame, and the "
                         "held-out / validation frames 900..999 are refused)")
    ap.add_argument("--spp-in", default="4,8,16,32,64,128,256",
                    help="input sample counts the levels are drawn from (comma-separated)")
    ap.add_argument("--independent-levels", action="store_true",
                    help="each input level from its own run of samples (independent noise) instead of reads "
                         "of one accumulation")
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    ctx = RenderContext(0)
    t0 = time.perf_counter()
    spp_in = tuple(int(v) for v in args.spp_in.split(",") if v.strip())
    if len(spp_in) < args.levels:
        raise SystemExit("--spp-in needs at least --levels values")
    if args.frame_offset <= 1 < args.frame_offset + args.frame_seeds:
        raise SystemExit("full-frame seed 1 is the benchmark frame: use --frame-offset >= 2")
    if args.frame_seeds > 0 and args.frame_offset < 1000 and args.frame_offset + args.frame_seeds > 900:
        raise SystemExit("full-frame seeds 900..999 are the held-out and validation frames: use other seeds")
    scene_seeds = range(args.scene_offset, args.scene_offset + args.scenes)
    if any(s >= 1_000_000 for s in (scene_seeds[-1:] if args.scenes else [])):
        raise SystemExit("training view seeds >= 1000000 are the held-out evaluation views")
    raw, tgt = make_dataset(ctx, scene_seeds, size=args.size, spp_in=spp_in, spp_ref=args.ref_spp, log=log,
                            levels=args.levels, seed=args.seed, independent=args.independent_levels)
    if args.frame_seeds > 0:
        fr, ft = make_frame_crops(ctx, range(args.frame_offset, args.frame_offset + args.frame_seeds), size=args.size,
                                  spp_in=spp_in,
                                  spp_ref=args.ref_spp, levels=args.levels, log=log, seed=args.seed + 1,
                                  independent=args.independent_levels)
        raw, tgt = np.concatenate([raw, fr]), np.concatenate([tgt, ft])
        del fr, ft
    t_data = time.perf_counter() - t0
    log(f"[train] dataset: {raw.shape[0]} views {args.size}x{args.size} in {t_data:.1f} s")
    net = NetDenoiser(ctx).load(args.init) if args.init else NetDenoiser(ctx).init(args.seed)
    net.set_dataset(raw, tgt)
    t1 = time.perf_counter()
    chunk = 250
    losses = []
    done = 0
    while done < args.steps:
        k = min(chunk, args.steps - done)
        frac = done / args.steps
        lr = args.lr * (1.0 if frac < 0.6 else (0.3 if frac < 0.85 else 0.1))
        losses.append(net.train(k, batch=args.batch, lr=lr, seed=args.seed * 7919 + done))
        done += k
        log(f"[train] step {done}/{args.steps} lr {lr:.1e} loss {float(np.mean(losses[-1][-50:])):.5f}")
    ctx.sync()
    t_train = time.perf_counter() - t1
    losses = np.concatenate(losses)
    net.save(args.out)
    log(f"[train] {args.steps} steps in {t_train:.1f} s; weights -> {args.out}")
    ev = evaluate(ctx, net, range(1_000_000, 1_000_000 + args.eval_scenes), args.size, 8, args.eval_ref_spp) \
        if args.eval_scenes > 0 else np.full((1, 3), np.nan)
    summary = {
        "init": args.init, "scene_offset": args.scene_offset, "frame_offset": args.frame_offset,
        "scenes": args.scenes, "levels": args.levels, "spp_in": list(spp_in), "frame_seeds": args.frame_seeds,
        "independent_levels": bool(args.independent_levels),
        "pairs": int(raw.shape[0]),
        "size": args.size, "ref_spp": args.ref_spp, "steps": args.steps,
        "batch": args.batch, "lr": args.lr, "params": net.n,
        "dataset_s": round(t_data, 2), "train_s": round(t_train, 2),
        "ms_per_step": round(t_train / args.steps * 1e3, 3),
        "loss_first50": float(np.mean(losses[:50])), "loss_last50": float(np.mean(losses[-50:])),
        "heldout_8spp_relmse": {"input": float(ev[:, 0].mean()), "atrous": float(ev[:, 1].mean()),
                                "net": float(ev[:, 2].mean())},
    }
    log(f"[train] held-out relMSE at 8 spp: {summary['heldout_8spp_relmse']}")
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(summary, f, indent=1)
    print(json.dumps(summary))
    ctx.close()


if __name__ == "__main__":
    main()
