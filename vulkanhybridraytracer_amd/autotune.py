"""Autotuner of the launch shapes (hr_tuning) for a batch of frames.

It sweeps the persistent grid size (workgroups per CU), the shadow work tile
(tile_w x 64/tile_w pixels per wave) and the reflection-queue workspace per
slab. For each candidate it captures the batch frame in a hipGraph, replays
it, and keeps the fastest. Results are cached in a JSON file, keyed by the
workload signature.
"""
import hashlib
import json
import os

DEFAULT_CANDIDATES = {
    "blocks_per_cu": [4, 6, 8, 12],
    "tile_w": [4, 8, 16],
    "wave_budget_mb": [3072, 16384, 40960],
}


def signature(batch):
    parts = []
    for rd in batch.renderers:
        info = rd.hs.info()
        parts.append("%dx%d:%d:%d:%d:%d:%d" % (rd.width, rd.height, info["tri_count"], rd.nl, rd.params.shadow_samples,
                                              rd.params.refl_spp, rd.params.flags))
    return hashlib.sha1(";".join(parts).encode()).hexdigest()[:16]


def default_cache_path():
    return os.environ.get("HR_TUNING_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "hybridrt_tuning.json"))


def load_cache(path=None):
    path = path or default_cache_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(batch, candidates=None, reps=2, cache_path=None, use_cache=True, log=None):
    """Returns (best tuning dict, ms of its replay, all trials). Leaves the
    context tuned with the best candidate."""
    ctx = batch.ctx
    sig = signature(batch)
    cache = load_cache(cache_path) if use_cache else {}
    if sig in cache:
        t = cache[sig]["tuning"]
        ctx.set_tuning(**t)
        return t, cache[sig]["ms"], []
    cand = dict(DEFAULT_CANDIDATES)
    if candidates:
        cand.update(candidates)
    trials = []
    for bpc in cand["blocks_per_cu"]:
        for tw in cand["tile_w"]:
            for wb in cand["wave_budget_mb"]:
                t = dict(blocks_per_cu=bpc, tile_w=tw, wave_budget_mb=wb)
                ctx.set_tuning(**t)
                g = batch.capture()
                g.launch()
                ms = min(g.launch(stats=True)["ms_shade"] for _ in range(reps))
                g.close()
                trials.append((ms, t))
                if log:
                    log("autotune %s: %.3f ms" % (t, ms))
    ms, best = min(trials, key=lambda x: x[0])
    ctx.set_tuning(**best)
    if use_cache:
        path = cache_path or default_cache_path()
        cache = load_cache(path)
        cache[sig] = {"tuning": best, "ms": ms}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return best, ms, trials
