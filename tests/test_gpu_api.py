"""GPU: host-side behaviour of the C ABI that needs a device context.

- rejected inputs return their status with a message (never abort):
  shading models other than 0 / 3 / 4 (hybridRaygen.rgen:105-148), material
  indices out of range, non-finite lights, unknown trace flags;
- an empty light list (the reference's light buffer may be empty,
  shadowRaygen.rgen:52, hybridRaygen.rgen:73) traces nothing and shades
  ambient + emissive, as the oracle does;
- a captured graph refuses to replay once a setter replaced what it captured,
  and replays again after a new capture; a refit keeps it valid;
- repeated environment uploads of changing size do not leak device memory;
- hr_trace_rays walks the 4-wide nodes unless asked to count, with the same hits;
- module-scope library objects are released at interpreter exit (rc 0).
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from vulkanhybridraytracer_amd import abi, scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hr():
    from vulkanhybridraytracer_amd import hr
    return hr


def _status(exc):
    return exc.value.status


def test_rejects_unknown_shading_model(gpu_ctx):
    hr = _hr()
    sc = scenes.make_small(spheres=2)
    sc.materials = sc.materials.copy()
    sc.materials[1, 8] = 2.0  # spec-gloss (material.h:17) is not a shading model of the pass
    with pytest.raises(hr.HRError) as e:
        hr.HybridScene(gpu_ctx, sc)
    assert _status(e) == abi.HR_E_INVALID and b"shading model" in gpu_ctx.last_error()


def test_rejects_material_index_out_of_range(gpu_ctx):
    hr = _hr()
    sc = scenes.make_small(spheres=2)
    sc.materials = sc.materials.copy()
    sc.materials[0, 11] = float(len(sc.materials))
    with pytest.raises(hr.HRError) as e:
        hr.HybridScene(gpu_ctx, sc)
    assert _status(e) == abi.HR_E_INVALID and b"material index" in gpu_ctx.last_error()


@pytest.mark.parametrize("col", [0, 3, 4, 7, 8])
def test_rejects_non_finite_light(gpu_ctx, col):
    hr = _hr()
    sc = scenes.make_small(width=32, height=18, spheres=2)
    hs = hr.HybridScene(gpu_ctx, sc)
    rd = hr.HybridRenderer(gpu_ctx, hs)
    rd.gbuffer_primary()
    bad = sc.lights.copy()
    bad[0, col] = np.inf if col != 8 else np.nan
    rd.set_frame(sc.camera, bad, sc.shapes, sc.params)
    with pytest.raises(hr.HRError) as e:
        rd.trace_shadows()
    assert _status(e) == abi.HR_E_INVALID and b"non-finite" in gpu_ctx.last_error()
    with pytest.raises(hr.HRError):
        rd.shade_hybrid()


def test_rejects_unknown_trace_flags(gpu_ctx):
    hr = _hr()
    hs = hr.HybridScene(gpu_ctx, scenes.make_small(spheres=2))
    rays = np.zeros(4, dtype=abi.RAY_DTYPE)
    rays["dir"] = (0, -1, 0)
    rays["tmax"] = 10
    rb, hb = gpu_ctx.alloc(rays.nbytes), gpu_ctx.alloc(64)
    rb.upload(rays)
    st = gpu_ctx.L.hr_trace_rays(hs.h, rb.ptr, 4, 8, hb.ptr, None)
    assert st == abi.HR_E_INVALID and b"flags" in gpu_ctx.last_error()


def test_trace_rays_count_flag(gpu_ctx, oracle_lib):
    """Without HR_TRACE_COUNT_VISITS the 4-wide walk runs (no counts), with it
    the binary walk (the oracle's counts); the hits are the same."""
    hr = _hr()
    sc = scenes.make_small()
    hs = hr.HybridScene(gpu_ctx, sc)
    rng = np.random.default_rng(3)
    rays = np.zeros(5000, dtype=abi.RAY_DTYPE)
    org = rng.uniform([-45, 0.5, -45], [45, 12, 45], size=(5000, 3)).astype(np.float32)
    d = rng.normal(size=(5000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["org"], rays["dir"], rays["tmin"], rays["tmax"] = org, d.astype(np.float32), 1e-3, 1e4
    for any_hit in (False, True):
        fast, fst = hs.trace_rays(rays, any_hit=any_hit, stats=True, count=False)
        slow, sst = hs.trace_rays(rays, any_hit=any_hit, stats=True, count=True)
        o, ost = oracle_lib.OracleScene(sc).trace_rays(rays, any_hit=any_hit)
        assert fst["node_visits"] == 0 and sst["node_visits"] == ost["node_visits"] > 0
        assert np.array_equal(fast["tri"] != 0xFFFFFFFF, slow["tri"] != 0xFFFFFFFF)
        if not any_hit:
            assert np.array_equal(fast.view(np.uint32), slow.view(np.uint32))


def test_zero_lights(gpu_ctx, oracle_lib):
    """No lights: the shadow pass traces nothing; shading gives ambient +
    emissive with no bounce (refl = 1), equal to the oracle's."""
    hr = _hr()
    sc = scenes.make_small(width=160, height=90, spheres=16)
    sc.lights = np.zeros((0, 12), np.float32)
    hs = hr.HybridScene(gpu_ctx, sc)
    rd = hr.HybridRenderer(gpu_ctx, hs)
    s1, s2, s3 = abi.hr_stats(), abi.hr_stats(), abi.hr_stats()
    rd.gbuffer_primary(s1)
    rd.trace_shadows(s2)
    rd.shade_hybrid(s3)
    assert s2.shadow_rays == 0 and s3.reflection_rays == 0 and s3.hit_shadow_rays == 0
    gb = rd.gb.download()
    lit, refl = rd.download_lit(), rd.download_refl()
    os_ = oracle_lib.OracleScene(sc)
    olit, orefl, oc = os_.shade_hybrid(gb, np.zeros((0, sc.height, sc.width), np.float32))
    assert oc["reflection_rays"] == 0
    assert np.array_equal(refl, np.ones_like(refl))
    assert np.array_equal(orefl, refl)
    assert np.nanmax(np.abs(lit - olit)) < 1e-4
    mirror = (gb["material"][..., 2] == 1.0) & (gb["normal"][..., 3] != 0)
    assert mirror.any()  # shading-model 3 pixels exist and took no bounce


def test_graph_refuses_replay_after_setters(gpu_ctx):
    hr = _hr()
    batch = scenes.make_batch(count=2, tri_range=(8_000, 20_000), resolutions=((64, 48),), spp=4)
    rds = [hr.HybridRenderer(gpu_ctx, hr.HybridScene(gpu_ctx, sc)) for sc in batch]
    for rd in rds:
        rd.gbuffer_primary()
    br = hr.BatchRenderer(gpu_ctx, rds)
    g = br.capture()
    g.launch()
    rds[1].hs.refit(batch[1].prims["model"])  # same buffers: still replays
    g.launch()
    sky, irr = scenes.make_env(64, 32, 16, 8)
    setters = [
        lambda: rds[0].hs.set_materials(batch[0].materials),
        lambda: rds[1].hs.set_textures([scenes.texture_checker(8)]),
        lambda: rds[0].hs.set_env(sky, irr),
    ]
    for k, setter in enumerate(setters):
        setter()
        with pytest.raises(hr.HRError) as e:
            g.launch()
        assert _status(e) == abi.HR_E_INVALID and b"capture" in gpu_ctx.last_error(), k
        g.close()
        g = br.capture()
        g.launch()
    g.close()


def _device_free_bytes():
    """hipMemGetInfo of the HIP runtime libhr.so runs on (the device the
    library last made current on this thread)."""
    from vulkanhybridraytracer_amd import hr
    out = subprocess.check_output(["ldd", hr.LIB_PATH]).decode()
    path = next((ln.split("=>")[1].split()[0] for ln in out.splitlines() if "libamdhip64" in ln), None)
    assert path, "libhr.so does not link libamdhip64"
    hip = C.CDLL(path)
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_env_reupload_does_not_leak(gpu_ctx):
    hr = _hr()
    sc = scenes.make_small(width=32, height=18, spheres=2)
    hs = hr.HybridScene(gpu_ctx, sc)
    big = scenes.make_env(2048, 1024, 1024, 512)   # 32 MiB + 8 MiB
    small = scenes.make_env(1024, 512, 512, 256)
    hs.set_env(*big)
    gpu_ctx.sync()
    free0 = _device_free_bytes()
    for k in range(40):
        hs.set_env(*(small if k % 2 == 0 else big))
    gpu_ctx.sync()
    free1 = _device_free_bytes()
    # a leak would keep 40 x ~25 MiB; allow allocator noise
    assert free0 - free1 < (200 << 20), (free0 - free1) / 2**20


def test_interpreter_exit_with_module_scope_objects():
    """Library objects left at module scope (a timing script's habit) are
    released before the interpreter tears modules down: the process exits 0."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from vulkanhybridraytracer_amd import hr, scenes\n"
            "ctx = hr.Context(0)\n"
            "sc = scenes.make_small(width=64, height=36, spheres=4)\n"
            "hs = hr.HybridScene(ctx, sc)\n"
            "rd = hr.HybridRenderer(ctx, hs)\n"
            "rd.rasterize_hybrid()\n"
            "g = hr.BatchRenderer(ctx, [rd]).capture()\n"
            "g.launch()\n"
            "print('done', flush=True)\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    assert "done" in r.stdout
