"""Benchmark of the hybrid pass on one MI355X.

A step is one hybrid pass over a resident G-buffer: the shadow trace
(shadowRaygen) and the hybrid trace (hybridRaygen + hybridHit: reflection
bounces and their hit shadow rays). The default workload is the 1920x1080,
1,048,576-triangle scene (terrain + 256 triangle soups, 4 rectangular area
lights with 16 stratified shadow samples each, GGX glossy reflections at
4 spp, 1 bounce). value = traced rays (shadow + reflection + hit shadow) per
second of the K timed steps; primary rays (the G-buffer) are excluded.

roofline: the shadow kernel (most of the rays) against the unit that binds
it, the vector-memory address path: its wave-level vector load instructions
per launch (rocprofv3 SQ_INSTS_VMEM_RD, from the committed profile of the same
workload) over its average launch time measured live with HIP events on the
library's stream, against 256 CUs x 2.4 GHz / 16 cycles (a CU's texture
addresser takes 4 lanes per clock: 16 clocks per 64-lane load). Beside it: the
measured TA busy fraction and VALU issue (a wave64 instruction per SIMD every
4 cycles) with the lane utilisation, the algorithmic HBM bytes of the traced
rays (32 B per ray + 4 B per shadow hit + 32 B per binary node visited + 48 B
per triangle tested, the counts from a counting launch that walks the oracle's
order, outside the timed region) and the HBM bytes rocprofv3 measured, both
against 8 TB/s.

cpu_baseline: the CPU oracle (test infrastructure; a scalar C++ port of the
same pass) timed on the same frame (a row subset for the larger workloads) on
the process's share of host threads, with the CPU model named.

parity: the same frame's shadow factors and lit / reflection colours against
the oracle's on those rows (bit-exact shadows, radiance within 1e-4), and the
ray counts of the timed pass against the oracle's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# wave-level VALU issue peak of an MI355X: 256 CUs x 4 SIMDs x 2.4 GHz, one
# wave64 VALU instruction per SIMD every 4 cycles
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4
# wave-level vector memory instruction peak: a CU's texture addresser takes the
# addresses of 4 lanes per clock, 16 clocks per 64-lane instruction
VMEM_PEAK_GINST = 256 * 2.4 / 16
RAD_TOL = 1e-4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain copy of known size: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
    this run gives the counter-to-bytes factor used for the traffic figure."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def shadow_kernel_name(params):
    """The kernel(s) of the timed shadow pass: sample-parallel for 4..64
    samples (a power of two), in stages with adaptive sampling."""
    if params["shadow_samples"] not in (4, 8, 16, 32, 64):
        return "k_shadows"
    if params.get("flags", 0) & 4:
        return "k_shadows_spl (probe stage and later stages, with k_adaptive_decide)"
    return "k_shadows_spl"


def load_profile(kernel, workload):
    """The per-launch counters of `kernel` on `workload` that rocprofv3 runs
    measured (profiles/<kernel>_<workload>.json): VALU instructions, lane
    utilisation, TA busy, HBM bytes, L2 hit rate (None when absent)."""
    p = os.path.join(ROOT, "profiles", "%s_%s.json" % (kernel, workload))
    if not os.path.exists(p):
        return None, None
    return json.load(open(p)), os.path.relpath(p, ROOT)


def roofline_of(kernel, workload, ms, rays, nodes, tris):
    """Roofline of one launch of a traversal kernel that ran `ms` on average:
    VALU issue (binding) with the HBM figures beside it."""
    prof, src = load_profile(kernel, workload)
    alg = 36 * rays + 32 * nodes + 48 * tris
    alg_gbs = alg / (ms * 1e-3) / 1e9
    out = {"kernel": kernel, "avg_launch_ms": round(ms, 3),
           "hbm_algorithmic": {"bytes": int(alg), "achieved": round(alg_gbs, 1), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(alg_gbs / HBM_PEAK_GBS, 4),
                               "per_ray": "32 B ray + 4 B hit + 32 B per binary node + 48 B per triangle",
                               "nodes_per_ray": round(nodes / max(rays, 1), 2),
                               "tris_per_ray": round(tris / max(rays, 1), 2)}}
    if prof is None:
        out.update(bound="hbm", achieved=out["hbm_algorithmic"]["achieved"], peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=out["hbm_algorithmic"]["frac"], traffic=None)
        return out
    vm = prof["vmem_read_insts_per_launch"]
    ach = vm / (ms * 1e-3) / 1e9
    traffic = prof.get("hbm_bytes_per_launch")
    vi = prof["valu_insts_per_launch"]
    valu = vi / (ms * 1e-3) / 1e9
    out.update(bound="vmem", achieved=round(ach, 2), peak=VMEM_PEAK_GINST, unit="Ginst/s",
               frac=round(ach / VMEM_PEAK_GINST, 4), traffic=traffic, profile=src,
               vmem_insts_per_launch=vm, ta_busy=prof.get("ta_busy"), l2_hit_rate=prof.get("l2_hit_rate"),
               valu={"insts_per_launch": vi, "achieved": round(valu, 1), "peak": VALU_PEAK_GINST, "unit": "Ginst/s",
                     "issue_frac": round(valu / VALU_PEAK_GINST, 4),
                     "lane_utilization": prof.get("lane_utilization"),
                     "useful_lane_frac": round(valu / VALU_PEAK_GINST * prof.get("lane_utilization", 0.0), 4)})
    if traffic:
        gbs = traffic / (ms * 1e-3) / 1e9
        out["hbm_measured"] = {"bytes": traffic, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(gbs / HBM_PEAK_GBS, 5)}
    return out


def frame_parity(sh, osh, lit, olit, refl, orefl):
    """mismatch figures of GPU buffers against the oracle's on the same pixels"""
    sbad = ~((sh == osh) | (np.isnan(sh) & np.isnan(osh)))
    nan_same = bool(np.array_equal(np.isnan(lit), np.isnan(olit)))
    dl = np.nan_to_num(np.abs(lit[..., :3] - olit[..., :3]))
    dr = np.nan_to_num(np.abs(refl[..., :3] - orefl[..., :3]))
    return {"shadow_mismatch_frac": float(sbad.mean()), "pixels": int(lit.shape[0] * lit.shape[1]),
            "lit_mismatch_frac": float((dl > RAD_TOL).any(-1).mean()), "lit_max_abs": float(dl.max()),
            "refl_mismatch_frac": float((dr > RAD_TOL).any(-1).mean()), "refl_max_abs": float(dr.max()),
            "nan_pattern_equal": nan_same, "tolerance": "shadow bit-exact, radiance %g absolute" % RAD_TOL}


def cpu_info(threads):
    from oracle import pyoracle
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count()
    return {"cores": threads, "cpu_model": pyoracle.cpu_model(), "host_threads": os.cpu_count(),
            "affinity_threads": aff, "thread_share": pyoracle.HOST_THREAD_SHARE}


def batch_main(args):
    """hybrid_batch32: 32 scenes (50k-2M triangles, 256x256 to 1080p, adaptive
    1-16 spp) traced by one persistent shadow kernel and their shading passes,
    captured in one hipGraph (tuned by the autotuner) and replayed."""
    from vulkanhybridraytracer_amd import autotune, hr, scenes
    t0 = time.time()
    scs = scenes.make_batch(count=args.batch_count)
    log("batch: %d scenes, %d triangles, %d pixels, generated in %.1fs" % (
        len(scs), sum(s.tri_count for s in scs), sum(s.width * s.height for s in scs), time.time() - t0))
    ctx = hr.Context(0)
    rds = [hr.HybridRenderer(ctx, hr.HybridScene(ctx, sc)) for sc in scs]
    for rd in rds:
        rd.gbuffer_primary()
    br = hr.BatchRenderer(ctx, rds)
    tuning, trials = ctx.get_tuning(), []
    if args.autotune:
        tuning, ms, trials = autotune.autotune(br, reps=1, use_cache=False, log=log)
        log("autotuned:", tuning, "%.2f ms" % ms)
    g = br.capture()
    for _ in range(args.warmup):
        g.launch()
    ctx.sync()
    rays, ms = 0, []
    t = time.perf_counter()
    for _ in range(args.steps):
        st = g.launch(stats=True)
        rays += st["shadow_rays"] + st["reflection_rays"] + st["hit_shadow_rays"]
        ms.append(st["ms_shade"])
    ctx.sync()
    elapsed = time.perf_counter() - t
    out = {"metric": "Grays/s (hybrid_batch32)", "value": round(rays / elapsed / 1e9, 4), "unit": "Grays/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(elapsed * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "hybrid_batch32", "scenes": len(scs),
                      "triangles": sum(s.tri_count for s in scs), "pixels": sum(s.width * s.height for s in scs),
                      "tuning": tuning, "autotune_trials": len(trials)},
           "ms_graph": round(float(np.mean(ms)), 3), "rays_per_step": int(rays // args.steps)}
    print(json.dumps(out), flush=True)


BF16_PEAK_TFLOPS = 2500.0


def denoise_measure(steps=3, log_fn=log):
    """4 spp + denoise against plain 64 spp on the 1920x1080, 4.2M-triangle
    scene (non-adaptive sampling): ms per frame of each, their ratio, the
    rMSE and max abs error of the denoised frame (and of the raw 4 spp frame)
    against the 64 spp frame, and the denoiser's MFMA roofline."""
    from vulkanhybridraytracer_amd import abi, denoise, hr, scenes
    if not os.path.exists(denoise.WEIGHTS_PATH):
        return {"skipped": "no trained weights at %s" % os.path.relpath(denoise.WEIGHTS_PATH, ROOT)}
    fr = scenes.config("denoise_1080p_4m")["frame"]
    lo, hi = fr["low_spp"], fr["high_spp"]
    sc = scenes.make_1080p_4m(samples=lo, spp=lo, adaptive=fr["adaptive"])
    ctx = hr.Context(0)
    hs = hr.HybridScene(ctx, sc)
    rd = hr.HybridRenderer(ctx, hs)
    rd.gbuffer_primary()
    W, H = sc.width, sc.height
    low = dict(sc.params, shadow_samples=lo, refl_spp=lo, flags=0)
    high = dict(sc.params, shadow_samples=hi, refl_spp=hi, flags=0)

    def frame(params, reps):
        rd.set_frame(sc.camera, sc.lights, sc.shapes, params)
        ms, rays = [], 0
        for _ in range(reps):
            s1, s2 = abi.hr_stats(), abi.hr_stats()
            rd.trace_shadows(s1)
            rd.shade_hybrid(s2)
            ms.append(s1.ms_shadow + s2.ms_shade)
            rays = s1.shadow_rays + s2.reflection_rays + s2.hit_shadow_rays
        return float(np.median(ms)), int(rays)

    ms4, rays4 = frame(low, steps + 1)
    noisy = rd.lit.download((H, W, 4))
    dn = denoise.Denoiser.load(ctx)
    out = ctx.alloc(W * H * 16)
    noisy_buf = ctx.alloc(W * H * 16)
    noisy_buf.upload(noisy)
    img = denoise.Denoiser.image(rd.gb, noisy_buf.ptr, None, out.ptr, _eye(sc))
    dn.denoise(img)
    ms_dn = float(np.median([dn.denoise(img, timing=True) for _ in range(steps + 2)]))
    den = out.download((H, W, 4))
    ms64, rays64 = frame(high, 1)
    ref = rd.lit.download((H, W, 4))
    flops = denoise.flops(H, W)
    tflops = flops / (ms_dn * 1e-3) / 1e12
    res = {"workload": "denoise_1080p_4m", "ms_4spp": round(ms4, 3), "ms_denoise": round(ms_dn, 3),
           "ms_4spp_plus_denoise": round(ms4 + ms_dn, 3), "ms_64spp": round(ms64, 3),
           "speedup_vs_64spp": round(ms64 / (ms4 + ms_dn), 2),
           "rmse_denoised": round(denoise.rmse(den, ref), 5), "rmse_4spp": round(denoise.rmse(noisy, ref), 5),
           "max_abs_denoised": round(float(np.abs(den[..., :3] - ref[..., :3]).max()), 4),
           "max_abs_4spp": round(float(np.abs(noisy[..., :3] - ref[..., :3]).max()), 4),
           "rays_4spp": rays4, "rays_64spp": rays64,
           "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tflops / BF16_PEAK_TFLOPS, 4), "flops_per_frame": int(flops),
                        "kernel": "denoiser forward (k_conv + features + kernel prediction)"}}
    log_fn("denoise", res)
    dn.close()
    ctx.close()
    return res


def _eye(sc):
    """Camera position: translation column of the inverse view matrix."""
    iv = np.asarray(sc.camera[:16], dtype=np.float64).reshape(4, 4)  # column-major
    return tuple(float(v) for v in iv[3, :3])


def denoise_main(args):
    res = denoise_measure(args.steps)
    out = {"metric": "ms/frame 4spp+denoise (1080p, 4M tris)", "value": res.get("ms_4spp_plus_denoise"),
           "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res.get("ms_4spp_plus_denoise"), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16", "data": "synthetic", "config": {"workload": "denoise_1080p_4m"},
           **{k: v for k, v in res.items() if k != "workload"}}
    print(json.dumps(out), flush=True)


def rt_main(args):
    """rt_1080p_1m: the pure ray tracing mode (raygen.rgen + closesthit.rchit)
    on the 1080p, 1M-triangle scene: 1 camera sample per pixel, up to 8 rays
    per chain (MAX_RECURSION), 16 stratified shadow rays per light at every
    hit. value = camera + bounce + hit shadow rays per second."""
    from vulkanhybridraytracer_amd import abi, hr, scenes
    sc = scenes.make_1080p_1m()
    p = scenes.config("rt_1080p_1m")["params"]
    ctx = hr.Context(0)
    rd = hr.HybridRenderer(ctx, hr.HybridScene(ctx, sc))
    rd.set_frame(sc.camera, sc.lights, sc.shapes, dict(sc.params, shadow_samples=p["shadow_samples"],
                                                       refl_spp=p["refl_spp"], max_depth=p["max_depth"]))
    acc = ctx.alloc(sc.width * sc.height * 16)
    for _ in range(args.warmup):
        rd.render_rt(acc)
    ctx.sync()
    rays, ms = 0, []
    t = time.perf_counter()
    for _ in range(args.steps):
        st = abi.hr_stats()
        rd.render_rt(acc, st)
        rays += st.primary_rays + st.reflection_rays + st.hit_shadow_rays
        ms.append(st.ms_shade)
    ctx.sync()
    elapsed = time.perf_counter() - t
    out = {"metric": "Grays/s (pure ray tracing, 1080p, 1M tris)", "value": round(rays / elapsed / 1e9, 4),
           "unit": "Grays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(elapsed * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "rt_1080p_1m", "shadow_samples": p["shadow_samples"], "camera_spp": p["refl_spp"],
                      "max_depth": p["max_depth"]},
           "ms_frame": round(float(np.mean(ms)), 3), "rays_per_step": int(rays // max(args.steps, 1))}
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hybrid_1080p_1m")
    ap.add_argument("--cpu-rows-step", type=int, default=0,
                    help="oracle rows y = 0, k, 2k ... (0: every row for the 1M-triangle frame, every 4th otherwise)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--leaf-max", type=int, default=2, help="largest subtree stored as one BVH leaf")
    ap.add_argument("--wave-budget-mb", type=int, default=0,
                    help="reflection workspace budget in MiB (0: the library default)")
    ap.add_argument("--batch-count", type=int, default=32)
    ap.add_argument("--autotune", action="store_true", help="tune launch shapes before timing (batch workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-denoise", action="store_true",
                    help="skip the 4 spp + denoise vs 64 spp measurement of the default line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the hybrid pass is built and measured on a single MI355X "
              "(frames would shard by screen tiles, batches by scene); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "hybrid_batch32":
        batch_main(args)
        return
    if args.workload == "denoise_1080p_4m":
        denoise_main(args)
        return
    if args.workload == "rt_1080p_1m":
        rt_main(args)
        return
    from vulkanhybridraytracer_amd import hr, scenes

    t0 = time.time()
    sc = scenes.WORKLOADS[args.workload]()
    log("scene %s: %d triangles, %dx%d, generated in %.1fs" % (sc.name, sc.tri_count, sc.width, sc.height,
                                                              time.time() - t0))
    ctx = hr.Context(0)
    if args.wave_budget_mb:
        t = ctx.get_tuning()
        ctx.set_tuning(t["blocks_per_cu"], t["tile_w"], args.wave_budget_mb)
    hs = hr.HybridScene(ctx, sc, leaf_max=args.leaf_max)
    info = hs.info()
    mem = hs.memory()
    log("LBVH build %.2f ms, depth %d, %d binary / %d 4-wide nodes, scene %.1f MiB on the device" % (
        info["ms_build"], info["depth"], info["node_count"], mem["wide_count"], mem["device_bytes"] / 2**20))
    rd = hr.HybridRenderer(ctx, hs)
    gst = rd.rasterize_hybrid(stats=True)  # G-buffer + one pass (also warms up)
    log("first pass", gst)

    # counting launches (node / triangle visits) outside the timed region
    flags_saved = rd.params.flags
    rd.params.flags = flags_saved | 1
    from vulkanhybridraytracer_amd import abi
    cs = abi.hr_stats()
    rd.trace_shadows(cs)
    shadow_nodes, shadow_tris, shadow_rays = cs.node_visits, cs.tri_tests, cs.shadow_rays
    rd.params.flags = flags_saved

    for _ in range(args.warmup):
        rd.trace_shadows()
        rd.shade_hybrid()
    ctx.sync()

    ms_shadow, ms_shade = [], []
    rays = 0
    st_s, st_h = abi.hr_stats(), abi.hr_stats()
    ctx.sync()
    t = time.perf_counter()
    for _ in range(args.steps):
        rd.trace_shadows(st_s)
        rd.shade_hybrid(st_h)
        ms_shadow.append(st_s.ms_shadow)
        ms_shade.append(st_h.ms_shade)
        rays += st_s.shadow_rays + st_h.reflection_rays + st_h.hit_shadow_rays
    ctx.sync()
    elapsed = time.perf_counter() - t
    ms_per_step = elapsed * 1e3 / args.steps
    value = rays / elapsed / 1e9
    per_step = dict(shadow_rays=int(st_s.shadow_rays), reflection_rays=int(st_h.reflection_rays),
                    hit_shadow_rays=int(st_h.hit_shadow_rays), occluded_rays=int(st_s.occluded_rays +
                                                                                  st_h.occluded_rays))
    W, H, nl = sc.width, sc.height, len(sc.lights)
    avg_shadow_ms = float(np.mean(ms_shadow))
    kname = shadow_kernel_name(sc.params)
    roofline = roofline_of(kname.split()[0], sc.name, avg_shadow_ms, shadow_rays, shadow_nodes, shadow_tris)
    if kname != roofline["kernel"]:
        roofline["kernels"] = kname

    cpu, parity = None, None
    if not args.no_cpu_baseline:
        from oracle import pyoracle
        threads = args.cpu_threads or pyoracle.default_threads()
        step = args.cpu_rows_step or (1 if sc.name == "hybrid_1080p_1m" else 4)
        gb = rd.gb.download()
        sh, lit, refl = rd.download_shadow(), rd.download_lit(), rd.download_refl()
        os_ = pyoracle.OracleScene(sc, leaf_max=args.leaf_max, threads=threads)
        rows = (0, step)
        t1 = time.perf_counter()
        osh, c1 = os_.trace_shadows(gb, rows=rows)
        olit, orefl, c2 = os_.shade_hybrid(gb, sh, rows=rows)
        tc = time.perf_counter() - t1
        crays = c1["shadow_rays"] + c2["reflection_rays"] + c2["hit_shadow_rays"]
        cpu = {"value": round(crays / tc / 1e9, 6), "unit": "Grays/s", **cpu_info(threads), "kind": "port",
               "sample": "%s of the timed frame (%d rays, %.1f s)" % (
                   "every row" if step == 1 else "rows 0::%d" % step, crays, tc)}
        log("cpu baseline", cpu)
        sel = slice(0, None, step)
        parity = frame_parity(sh[:, sel], osh[:, sel], lit[sel], olit[sel], refl[sel], orefl[sel])
        if step == 1:  # the whole frame: the timed pass's ray counts are the oracle's
            parity["counts_equal"] = bool(per_step["shadow_rays"] == c1["shadow_rays"] and
                                          per_step["reflection_rays"] == c2["reflection_rays"] and
                                          per_step["hit_shadow_rays"] == c2["hit_shadow_rays"] and
                                          per_step["occluded_rays"] == c1["occluded_rays"] + c2["occluded_rays"])
        parity["rows"] = "all" if step == 1 else "0::%d" % step
        log("parity", parity)

    metric = "Grays/s (1080p, 1M tris)" if sc.name == "hybrid_1080p_1m" else "Grays/s (%s)" % sc.name
    out = {"metric": metric, "value": round(value, 4), "unit": "Grays/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": sc.name, "width": W, "height": H, "triangles": sc.tri_count, "lights": nl,
                      "shadow_samples": sc.params["shadow_samples"], "refl_spp": sc.params["refl_spp"],
                      "max_depth": sc.params["max_depth"], "leaf_max": args.leaf_max},
           "rays_per_step": per_step, "ms_shadow": round(avg_shadow_ms, 3),
           "ms_shade": round(float(np.mean(ms_shade)), 3), "ms_gbuffer": round(gst["ms_gbuffer"], 3),
           "ms_lbvh_build": round(info["ms_build"], 3), "roofline": roofline, "cpu_baseline": cpu,
           "parity": parity,
           "scene_memory": {"binary_nodes": info["node_count"], "wide_nodes": mem["wide_count"],
                            "device_bytes": mem["device_bytes"]}}
    # the closest-hit queue kernel of the reflection stage, from its own profile
    # (it runs inside hr_shade_hybrid, so its launch time is rocprofv3's)
    prof_r, src_r = load_profile("k_wf_trace", sc.name)
    if prof_r:
        ms_r = prof_r["avg_launch_ms_traced"]
        ach = prof_r["vmem_read_insts_per_launch"] / (ms_r * 1e-3) / 1e9
        valu = prof_r["valu_insts_per_launch"] / (ms_r * 1e-3) / 1e9
        out["roofline_closest_hit"] = {"kernel": "k_wf_trace", "bound": "vmem", "achieved": round(ach, 2),
                                       "peak": VMEM_PEAK_GINST, "unit": "Ginst/s",
                                       "frac": round(ach / VMEM_PEAK_GINST, 4), "avg_launch_ms_rocprof": ms_r,
                                       "ta_busy": prof_r.get("ta_busy"),
                                       "valu_issue_frac": round(valu / VALU_PEAK_GINST, 4),
                                       "lane_utilization": prof_r.get("lane_utilization"), "profile": src_r}
    if sc.name == "hybrid_1080p_1m" and not args.no_denoise:
        del rd, hs
        ctx.close()
        out["denoise"] = denoise_measure(3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
