"""Trains the kernel-predicting denoiser on the GPU with the library's own
kernels (hr_denoiser_train_step), on pairs the tracer renders: random
256 x 256 views of seeded terrain scenes at 4 spp (input) and 64 spp
(target). Writes the BF16 weights and prints one JSON line per evaluation.

    python -m vulkanhybridraytracer_amd.train_denoiser --scenes 256 --views 4 --steps 20000
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import denoise, hr


def log(*a):
    print(*a, flush=True)


def lr_at(step, steps, lr, lr_final, warmup):
    """linear warm-up, then cosine decay from lr to lr_final"""
    if step < warmup:
        return lr * (step + 1) / warmup
    t = (step - warmup) / max(1, steps - warmup)
    return lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * t))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--views", type=int, default=4)
    ap.add_argument("--val-scenes", type=int, default=8)
    ap.add_argument("--val-views", type=int, default=2)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eval-every", type=int, default=1000)
    ap.add_argument("--time-limit", type=float, default=0.0, help="stop training after this many seconds")
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--loss", choices=["l1", "relmse"], default="relmse")
    ap.add_argument("--out", default=denoise.WEIGHTS_PATH)
    args = ap.parse_args(argv)

    ctx = hr.Context(0)
    t0 = time.time()
    progress = {"n": 0}

    def scene_done(msg):
        progress["n"] += 1
        if progress["n"] % 32 == 0:
            log(json.dumps({"event": "data_progress", "scenes": progress["n"],
                            "seconds": round(time.time() - t0, 1)}))

    train = denoise.make_dataset(ctx, range(1000, 1000 + args.scenes), args.views, log=scene_done)
    log(json.dumps({"event": "data", "train_samples": len(train), "seconds": round(time.time() - t0, 1)}))
    val = denoise.make_dataset(ctx, range(9000, 9000 + args.val_scenes), args.val_views)
    dn = denoise.Denoiser(ctx, weights=denoise.load_weights(args.init) if args.init else None, seed=args.seed)
    dn.set_loss(dn.LOSS_REL_MSE if args.loss == "relmse" else dn.LOSS_L1_TONEMAPPED)
    rng = np.random.default_rng(args.seed)
    ev = denoise.evaluate(dn, val)
    log(json.dumps({"event": "eval", "step": 0, **{k: round(v, 6) for k, v in ev.items()}}))
    t1 = time.time()
    step = 0
    last_print = t1
    for step in range(1, args.steps + 1):
        idx = rng.choice(len(train), size=args.batch, replace=False)
        lr = lr_at(step - 1, args.steps, args.lr, args.lr_final, args.warmup)
        want = step % 100 == 0
        loss = dn.train_step([train[i].image() for i in idx], lr, want_loss=want)
        if want and time.time() - last_print > 10:
            last_print = time.time()
            log(json.dumps({"event": "train", "step": step, "loss": round(loss, 6), "lr": lr,
                            "ms_per_step": round((time.time() - t1) * 1e3 / step, 3)}))
        if step % args.eval_every == 0 or step == args.steps:
            ev = denoise.evaluate(dn, val, with_filters=False)
            log(json.dumps({"event": "eval", "step": step, **{k: round(v, 6) for k, v in ev.items()}}))
        if args.time_limit and time.time() - t1 > args.time_limit:
            log(json.dumps({"event": "time_limit", "step": step}))
            break
    ctx.sync()
    train_s = time.time() - t1
    dn.save(args.out)
    ev = denoise.evaluate(dn, val)
    log(json.dumps({"event": "final", "steps": step, "train_seconds": round(train_s, 1),
                    "ms_per_step": round(train_s * 1e3 / max(step, 1), 3), "samples": len(train),
                    "out": os.path.relpath(args.out), **{k: round(v, 6) for k, v in ev.items()}}))
    dn.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
