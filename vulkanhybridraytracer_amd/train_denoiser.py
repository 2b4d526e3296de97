"""Trains the kernel-predicting denoiser on the GPU with the library's own
kernels (hr_denoiser_train_step), on pairs the tracer renders: random
256 x 256 views of seeded terrain scenes at 4 spp (input) and 64 spp
(target). Writes the BF16 weights and prints one JSON line per evaluation.
With --checkpoint the full training state (FP32 master weights, Adam
moments, step, batch sampler) is written at every evaluation; --resume
continues from such a file with the same data arguments, and gives the
weights an uninterrupted run gives.

    python -m vulkanhybridraytracer_amd.train_denoiser --scenes 2000 --views 1 --steps 20000 \
        --loss l1 --finetune-steps 6000

--finetune-steps adds a second phase on the same data: the optimiser starts
afresh from the first phase's weights with --finetune-loss at
--finetune-lr (same warm-up and cosine decay).
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
    ap.add_argument("--checkpoint", default=None, help="write the training state here at every evaluation")
    ap.add_argument("--resume", default=None, help="continue from this checkpoint")
    ap.add_argument("--finetune-steps", type=int, default=0)
    ap.add_argument("--finetune-lr", type=float, default=5e-5)
    ap.add_argument("--finetune-loss", choices=["l1", "relmse"], default="relmse")
    args = ap.parse_args(argv)
    total = args.steps + args.finetune_steps

    def loss_kind(name):
        return denoise.Denoiser.LOSS_REL_MSE if name == "relmse" else denoise.Denoiser.LOSS_L1_TONEMAPPED

    def schedule(step):
        """learning rate of global step `step` (1-based) and whether it opens the second phase"""
        if step <= args.steps:
            return lr_at(step - 1, args.steps, args.lr, args.lr_final, args.warmup), False
        k = step - args.steps
        return lr_at(k - 1, args.finetune_steps, args.finetune_lr, args.lr_final, args.warmup), k == 1

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
    rng = np.random.default_rng(args.seed)
    first = 1
    if args.resume:
        dn, meta = denoise.Denoiser.load_checkpoint(ctx, args.resume)
        rng.bit_generator.state = meta["rng"]
        first = meta["step"] + 1
        log(json.dumps({"event": "resume", "step": meta["step"], "from": os.path.relpath(args.resume)}))
    else:
        dn = denoise.Denoiser(ctx, weights=denoise.load_weights(args.init) if args.init else None, seed=args.seed)
        dn.set_loss(loss_kind(args.loss))
        ev = denoise.evaluate(dn, val)
        log(json.dumps({"event": "eval", "step": 0, **{k: round(v, 6) for k, v in ev.items()}}))
    t1 = time.time()
    step = first - 1
    last_print = t1
    for step in range(first, total + 1):
        idx = rng.choice(len(train), size=args.batch, replace=False)
        lr, opens = schedule(step)
        if opens:
            dn.set_weights(dn.get_weights())  # fresh optimiser state
            dn.set_loss(loss_kind(args.finetune_loss))
            log(json.dumps({"event": "finetune", "step": step, "loss": args.finetune_loss}))
        want = step % 100 == 0
        loss = dn.train_step([train[i].image() for i in idx], lr, want_loss=want)
        if want and time.time() - last_print > 10:
            last_print = time.time()
            log(json.dumps({"event": "train", "step": step, "loss": round(loss, 6), "lr": lr,
                            "ms_per_step": round((time.time() - t1) * 1e3 / (step - first + 1), 3)}))
        if step % args.eval_every == 0 or step in (args.steps, total):
            ev = denoise.evaluate(dn, val, with_filters=False)
            log(json.dumps({"event": "eval", "step": step, **{k: round(v, 6) for k, v in ev.items()}}))
            if args.checkpoint:
                dn.save_checkpoint(args.checkpoint, {"step": step, "rng": rng.bit_generator.state})
        if args.time_limit and time.time() - t1 > args.time_limit:
            log(json.dumps({"event": "time_limit", "step": step}))
            if args.checkpoint:
                dn.save_checkpoint(args.checkpoint, {"step": step, "rng": rng.bit_generator.state})
            break
    ctx.sync()
    train_s = time.time() - t1
    dn.save(args.out)
    ev = denoise.evaluate(dn, val)
    log(json.dumps({"event": "final", "steps": step, "train_seconds": round(train_s, 1),
                    "ms_per_step": round(train_s * 1e3 / max(step - first + 1, 1), 3), "samples": len(train),
                    "out": os.path.relpath(args.out), **{k: round(v, 6) for k, v in ev.items()}}))
    dn.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
