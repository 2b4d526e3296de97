"""GPU: hr_trace_shadows_noalloc (the shadow pass with its scratch in a static
device array) against hr_trace_shadows, through the C ABI.

Bar: bit-exact shadow factors and equal ray counts, on calls in a row with
shrinking inputs (the array keeps the previous call's contents), on each
kernel path (sample-parallel, adaptive stages, group per lane); frames
whose scratch exceeds HR_NOALLOC_SCRATCH_BYTES are refused on the host.
"""
import ctypes as C

import numpy as np
import pytest

from vulkanhybridraytracer_amd import abi, scenes

pytestmark = pytest.mark.gpu

NOALLOC_BYTES = 128 << 20


def _noalloc(rd, stats):
    L = rd.ctx.L
    f = L.hr_trace_shadows_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.POINTER(abi.hr_gbuffer), C.c_void_p, C.c_void_p, C.c_uint32,
                  C.POINTER(abi.hr_camera), C.POINTER(abi.hr_trace_params), C.c_void_p, C.POINTER(abi.hr_stats)]
    return f(rd.hs.h, C.byref(rd.gb.c), C.cast(rd.lights, C.c_void_p), rd._shapes_ptr(), rd.nl, C.byref(rd.cam),
             C.byref(rd.params), rd.shadow.ptr, C.byref(stats) if stats is not None else None)


CASES = [  # (width, height, samples, adaptive): largest first, each call smaller than the one before
    (320, 180, 16, False),   # sample-parallel kernel
    (200, 120, 32, True),    # adaptive: probes, then stages of 16 and 32
    (160, 90, 8, False),
    (96, 54, 5, False),      # a count that is no power of two: group-per-lane kernel
]


def test_noalloc_equals_trace_shadows(gpu_ctx):
    from vulkanhybridraytracer_amd import hr
    for w, h, S, ad in CASES:
        print("case", w, h, S, ad, flush=True)
        sc = scenes.make_terrain_scene(seed=5, width=w, height=h, quads=96, instances=24, soup_tris=256,
                                       samples=S, spp=1, adaptive=ad)
        rd = hr.HybridRenderer(gpu_ctx, hr.HybridScene(gpu_ctx, sc))
        rd.gbuffer_primary()
        ref_st = abi.hr_stats()
        rd.trace_shadows(ref_st)
        ref = rd.download_shadow()
        assert ((ref > 0) & (ref < 1)).any()
        for rep in range(2):
            rd.shadow.zero()
            st = abi.hr_stats()
            rc = _noalloc(rd, st)
            assert rc == 0, gpu_ctx.last_error()
            got = rd.download_shadow()
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (w, h, S, ad, rep)
            assert st.shadow_rays == ref_st.shadow_rays and st.occluded_rays == ref_st.occluded_rays
        rd.shadow.zero()
        assert _noalloc(rd, None) == 0  # without stats
        assert np.array_equal(rd.download_shadow().view(np.uint32), ref.view(np.uint32))


def test_noalloc_refuses_oversize_frame(gpu_ctx):
    """2048x2048 with 4 lights at 64 samples needs 134 MB of sample bits: refused
    with HR_E_UNSUPPORTED before any device work (full-size buffers are passed
    all the same)."""
    from vulkanhybridraytracer_amd import hr
    sc = scenes.make_terrain_scene(seed=0, width=2048, height=2048, quads=32, instances=4, soup_tris=64,
                                   samples=64, spp=1)
    rd = hr.HybridRenderer(gpu_ctx, hr.HybridScene(gpu_ctx, sc))
    rc = _noalloc(rd, None)
    assert rc == abi.HR_E_UNSUPPORTED
    assert b"HR_NOALLOC_SCRATCH_BYTES" in gpu_ctx.last_error()
    # a smaller frame works again afterwards
    rd2 = hr.HybridRenderer(gpu_ctx, hr.HybridScene(gpu_ctx, scenes.make_terrain_scene(
        seed=1, width=64, height=48, quads=32, instances=4, soup_tris=64, samples=16, spp=1)))
    rd2.gbuffer_primary()
    rd2.trace_shadows()
    ref = rd2.download_shadow()
    rd2.shadow.zero()
    assert _noalloc(rd2, None) == 0
    assert np.array_equal(rd2.download_shadow().view(np.uint32), ref.view(np.uint32))
