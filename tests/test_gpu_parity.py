"""GPU parity: libhr.so (through the C ABI) against the CPU oracle.

Bars (same inputs on both sides):
  - LBVH arrays (Morton keys, sorted ids, Karras nodes, 64-byte traversal
    nodes, sorted triangles): bit-exact;
  - ray queries: hit/miss and triangle ids bit-exact, t/u/v bit-exact; a
    mismatch fraction below 1e-5 is tolerated for edge-grazing rays and
    reported;
  - G-buffer: bit-exact except the same grazing-ray fraction;
  - shadow factors: bit-exact except the grazing fraction;
  - lit / reflection radiance: within 1e-4 absolute per channel (libm vs
    device pow in albedo^2.2 and the Fresnel term), except the grazing
    fraction; NaN where the oracle has NaN;
  - ray counters equal to the oracle's.
"""
import numpy as np
import pytest

from vulkanhybridraytracer_amd import abi, scenes

pytestmark = pytest.mark.gpu

GRAZE_FRAC = 1e-5
RAD_TOL = 1e-4


def _hr():
    from vulkanhybridraytracer_amd import hr
    return hr


def _pair(gpu_ctx, oracle_lib, scene, leaf_max=4):
    hr = _hr()
    hs = hr.HybridScene(gpu_ctx, scene, leaf_max=leaf_max)
    os_ = oracle_lib.OracleScene(scene, leaf_max=leaf_max)
    return hs, os_


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _mismatch_frac(a, b, axis=None):
    a = np.asarray(a)
    b = np.asarray(b)
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    if axis is not None:
        bad = bad.any(axis=axis)
    return bad.mean(), bad


def _random_rays(rng, n, lo, hi):
    r = np.zeros(n, dtype=abi.RAY_DTYPE)
    org = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    tgt = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["org"] = org
    r["dir"] = d.astype(np.float32)
    r["tmin"] = 1e-3
    r["tmax"] = 1e4
    return r


@pytest.mark.parametrize("leaf_max", [1, 4])
def test_lbvh_bit_exact_small(gpu_ctx, oracle_lib, leaf_max):
    sc = scenes.make_small()
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc, leaf_max)
    info = hs.info()
    assert info["tri_count"] == os_.tri_count == 20482
    assert info["node_count"] == os_.node_count
    assert info["depth"] == os_.depth
    for which in (2, 3, 4, 0, 1):
        g, o = hs.export(which), os_.export(which)
        assert g.shape == o.shape
        assert np.array_equal(_bits(g), _bits(o)), "export %d differs" % which


def test_lbvh_bit_exact_terrain(gpu_ctx, oracle_lib):
    sc = scenes.make_terrain_scene(seed=3, width=64, height=36, quads=128, instances=32, soup_tris=512)
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc)
    assert hs.info()["depth"] == os_.depth
    for which in (2, 3, 4, 0, 1):
        assert np.array_equal(_bits(hs.export(which)), _bits(os_.export(which))), which


def test_lbvh_bit_exact_without_treelets(gpu_ctx, oracle_lib):
    """HR_BVH_NO_TREELETS (the plain Karras tree, also what a too-deep
    restructured tree falls back to) against the oracle built the same way"""
    sc = scenes.make_terrain_scene(seed=3, width=64, height=36, quads=128, instances=32, soup_tris=512)
    gpu_ctx.set_bvh_flags(abi.HR_BVH_NO_TREELETS)
    try:
        hs = _hr().HybridScene(gpu_ctx, sc, leaf_max=2)
    finally:
        gpu_ctx.set_bvh_flags(0)
    with oracle_lib.treelets(False):
        os_ = oracle_lib.OracleScene(sc, leaf_max=2)
    assert hs.info()["depth"] == os_.depth
    for which in (2, 3, 4, 0, 1):
        assert np.array_equal(_bits(hs.export(which)), _bits(os_.export(which))), which
    # and the restructured tree of the same scene is a different one
    assert not np.array_equal(_bits(_hr().HybridScene(gpu_ctx, sc, leaf_max=2).export(4)), _bits(os_.export(4)))


def test_lbvh_bit_exact_coincident(gpu_ctx, oracle_lib):
    """3000 copies of one triangle: every treelet split ties"""
    from test_oracle_geometry import coincident_scene
    sc = coincident_scene(3000)
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc, leaf_max=2)
    assert hs.info()["depth"] == os_.depth <= 64
    for which in (2, 3, 4, 0, 1):
        assert np.array_equal(_bits(hs.export(which)), _bits(os_.export(which))), which


def test_unknown_bvh_flags_are_rejected(gpu_ctx):
    hr = _hr()
    with pytest.raises(hr.HRError) as e:
        gpu_ctx.set_bvh_flags(2)
    assert e.value.status == abi.HR_E_INVALID
    assert gpu_ctx.bvh_flags() == 0


def test_lbvh_refit_bit_exact(gpu_ctx, oracle_lib):
    sc = scenes.make_small(spheres=16)
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc)
    rng = np.random.default_rng(5)
    models = sc.prims["model"].copy()
    models[1:, 12:15] += rng.uniform(-2, 2, size=(len(models) - 1, 3)).astype(np.float32)
    hs.refit(models)
    os_.refit(models)
    for which in (0, 1):
        assert np.array_equal(_bits(hs.export(which)), _bits(os_.export(which))), which


@pytest.mark.parametrize("any_hit", [False, True])
def test_trace_rays_bit_exact(gpu_ctx, oracle_lib, any_hit):
    sc = scenes.make_small()
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc)
    rays = _random_rays(np.random.default_rng(1), 20000, [-45, 0.5, -45], [45, 12, 45])
    g, gst = hs.trace_rays(rays, any_hit=any_hit, stats=True)
    o, ost = os_.trace_rays(rays, any_hit=any_hit)
    hit_g, hit_o = g["tri"] != 0xFFFFFFFF, o["tri"] != 0xFFFFFFFF
    assert hit_o.sum() > 1000
    frac = np.mean(hit_g != hit_o)
    print("hit/miss mismatch fraction", frac)
    assert frac < GRAZE_FRAC
    if not any_hit:
        assert np.array_equal(g["tri"], o["tri"])
        for f in ("t", "u", "v"):
            assert np.array_equal(g[f].view(np.uint32), o[f].view(np.uint32)), f
    assert gst["node_visits"] == ost["node_visits"]
    assert gst["tri_tests"] == ost["tri_tests"]


def test_gbuffer_shadow_shade_small(gpu_ctx, oracle_lib):
    hr = _hr()
    sc = scenes.make_small()
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc)
    rd = hr.HybridRenderer(gpu_ctx, hs)
    st = rd.rasterize_hybrid(stats=True)
    gb = rd.gb.download()
    ogb, oc = os_.gbuffer()
    assert st["primary_rays"] == oc["primary_rays"] == sc.width * sc.height
    for k in hr.PLANES:
        frac, _ = _mismatch_frac(gb[k], ogb[k], axis=-1)
        print("gbuffer", k, frac)
        assert frac < GRAZE_FRAC, k
    # shadows and shading from the GPU G-buffer on both sides
    sh = rd.download_shadow()
    osh, sc_cnt = os_.trace_shadows(gb)
    frac, _ = _mismatch_frac(sh, osh)
    print("shadow mismatch", frac)
    assert frac < GRAZE_FRAC
    assert st["shadow_rays"] == sc_cnt["shadow_rays"]
    lit, refl = rd.download_lit(), rd.download_refl()
    olit, orefl, hc = os_.shade_hybrid(gb, sh)
    assert np.array_equal(np.isnan(lit), np.isnan(olit))
    d = np.nan_to_num(np.abs(lit - olit))
    bad = (d > RAD_TOL).any(-1)
    print("lit max abs err", d.max(), "bad frac", bad.mean())
    assert bad.mean() < GRAZE_FRAC
    dr = np.nan_to_num(np.abs(refl - orefl))
    assert ((dr > RAD_TOL).any(-1)).mean() < GRAZE_FRAC
    assert st["reflection_rays"] == hc["reflection_rays"]
    assert st["hit_shadow_rays"] == hc["hit_shadow_rays"]


def test_adaptive_parity(gpu_ctx, oracle_lib):
    """4M-generator settings (adaptive 1/4/64 shadow rays per light, 1-64
    glossy paths by roughness) at 320x180 on a smaller scene: every 4th row."""
    hr = _hr()
    sc = scenes.make_terrain_scene(seed=1, width=320, height=180, quads=128, instances=48, samples=64,
                                   spp=64, rough_max=1.0, adaptive=True)
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc)
    rd = hr.HybridRenderer(gpu_ctx, hs)
    st = rd.rasterize_hybrid(stats=True)
    gb = rd.gb.download()
    rows = (1, 4)
    sel = slice(rows[0], None, rows[1])
    sh = rd.download_shadow()
    osh, oc = os_.trace_shadows(gb, rows=rows)
    frac, _ = _mismatch_frac(sh[:, sel], osh[:, sel])
    print("adaptive shadow mismatch", frac, "soft", float(((sh > 0) & (sh < 1)).mean()))
    assert frac < GRAZE_FRAC
    assert ((sh > 0) & (sh < 1)).any()
    lit = rd.download_lit()
    olit, orefl, hc = os_.shade_hybrid(gb, sh, rows=rows)
    d = np.nan_to_num(np.abs(lit[sel] - olit[sel]))
    print("adaptive lit max abs err", d.max())
    assert ((d > RAD_TOL).any(-1)).mean() < GRAZE_FRAC
    # the full-frame GPU counts equal the oracle's on the full frame
    _, oc_full = os_.trace_shadows(gb)
    assert st["shadow_rays"] == oc_full["shadow_rays"]


@pytest.mark.parametrize("samples,adaptive", [(4, False), (8, False), (32, False), (64, False),
                                              (8, True), (32, True)])
def test_shadow_sample_counts_parity(gpu_ctx, oracle_lib, samples, adaptive):
    """Every sample count the sample-parallel shadow kernels take (powers of
    two 4..64, fixed and adaptive; the group windows are parts of a word, one
    word or two words) against the oracle, with the exact shadow-ray count."""
    hr = _hr()
    sc = scenes.make_terrain_scene(seed=2, width=160, height=96, quads=96, instances=24, samples=samples,
                                   spp=1, adaptive=adaptive)
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc)
    rd = hr.HybridRenderer(gpu_ctx, hs)
    rd.gbuffer_primary()
    st = abi.hr_stats()
    rd.trace_shadows(st)
    gb = rd.gb.download()
    sh = rd.download_shadow()
    osh, oc = os_.trace_shadows(gb)
    frac, _ = _mismatch_frac(sh, osh)
    print("S=%d adaptive=%s shadow mismatch %g soft %g" % (samples, adaptive, frac,
                                                           float(((sh > 0) & (sh < 1)).mean())))
    assert frac < GRAZE_FRAC
    assert ((sh > 0) & (sh < 1)).any()
    assert st.shadow_rays == oc["shadow_rays"]


def test_terrain_rows_parity(gpu_ctx, oracle_lib):
    """1080p-generator scene at 480x270 with all features (rect lights, 16
    stratified samples, GGX 4 spp): every 8th row against the oracle."""
    hr = _hr()
    sc = scenes.make_terrain_scene(seed=0, width=480, height=270, quads=256, instances=64)
    hs, os_ = _pair(gpu_ctx, oracle_lib, sc)
    rd = hr.HybridRenderer(gpu_ctx, hs)
    rd.rasterize_hybrid()
    gb = rd.gb.download()
    rows = (3, 8)
    sel = slice(rows[0], None, rows[1])
    ogb, _ = os_.gbuffer(rows=rows)
    for k in hr.PLANES:
        frac, _ = _mismatch_frac(gb[k][sel], ogb[k][sel], axis=-1)
        assert frac < GRAZE_FRAC, k
    sh = rd.download_shadow()
    osh, _ = os_.trace_shadows(gb, rows=rows)
    frac, _ = _mismatch_frac(sh[:, sel], osh[:, sel])
    print("shadow mismatch", frac)
    assert frac < GRAZE_FRAC
    lit = rd.download_lit()
    olit, orefl, _ = os_.shade_hybrid(gb, sh, rows=rows)
    d = np.nan_to_num(np.abs(lit[sel] - olit[sel]))
    print("lit max abs err", d.max())
    assert ((d > RAD_TOL).any(-1)).mean() < GRAZE_FRAC


def _binary_children(nodes, i):
    n = nodes[i]
    u = n.view(np.uint32)
    return [(int(u[3]), n[0:3], n[4:7]), (int(u[7]), n[8:11], n[12:15])]


@pytest.mark.parametrize("leaf_max", [1, 4])
def test_wide_nodes_match_binary(gpu_ctx, oracle_lib, leaf_max):
    """the 4-wide nodes of the timed kernels: stored densely (the root first,
    then in the order of the binary nodes they are rooted at, fewer than 0.6
    as many as binary nodes); from the root they reach exactly the binary
    tree's leaves, each child box equal to the binary box of the same subtree,
    empty slots with inverted boxes, and every wide node is the collapse of
    its binary node that minimises the summed wide-node box area. A refit
    keeps the wide topology and the boxes follow the refitted binary nodes."""
    sc = scenes.make_terrain_scene(seed=2, width=32, height=18, quads=64, instances=24, soup_tris=256)
    hr = _hr()
    hs = hr.HybridScene(gpu_ctx, sc, leaf_max=leaf_max)
    LEAF = 0x80000000
    check_wide_nodes(hs, LEAF, check_collapse=True)
    # refit with moved instances: same wide topology, boxes of the new binary nodes
    w_before = hs.export(5).reshape(-1, 32)
    models = sc.prims["model"].copy()
    rng = np.random.default_rng(5)
    models[1:, 12:15] += rng.uniform(-3, 3, size=(len(models) - 1, 3)).astype(np.float32)
    hs.refit(models)
    w_after = hs.export(5).reshape(-1, 32)
    assert np.array_equal(w_before[:, 24:29], w_after[:, 24:29])  # codes and binary roots kept
    assert not np.array_equal(w_before[:, :24], w_after[:, :24])
    check_wide_nodes(hs, LEAF, check_collapse=False)


def check_wide_nodes(hs, LEAF, check_collapse):
    bn = hs.export(0).view(np.float32).reshape(-1, 16)
    w4 = hs.export(5).reshape(-1, 32)
    mem = hs.memory()
    assert mem["wide_count"] == len(w4) and 0 < len(w4) < 0.6 * len(bn)
    binary_of = w4[:, 28].astype(np.int64)
    assert binary_of[0] == 0 and np.all(np.diff(binary_of) > 0)
    dense_of = {int(b): d for d, b in enumerate(binary_of)}

    def wide(i):  # by binary root, internal codes mapped back to binary ids
        row = w4[dense_of[i]]
        f = row.view(np.float32)
        lo = np.stack([f[0:4], f[8:12], f[16:20]], -1)
        hi = np.stack([f[4:8], f[12:16], f[20:24]], -1)
        codes = [int(row[24 + k]) for k in range(4)]
        codes = [c if c == 0xFFFFFFFF or c & LEAF else int(binary_of[c]) for c in codes]
        return [(codes[k], lo[k], hi[k]) for k in range(4)]

    bin_leaves, wide_leaves = [], []
    stack = [0]
    while stack:  # binary: leaves with their boxes
        i = stack.pop()
        for code, lo, hi in _binary_children(bn, i):
            if code & LEAF:
                bin_leaves.append((code, tuple(lo), tuple(hi)))
            else:
                stack.append(code)
    stack = [0]
    internal_boxes = {}
    for i in range(len(bn)):
        for code, lo, hi in _binary_children(bn, i):
            if not code & LEAF:
                internal_boxes[code] = (tuple(lo), tuple(hi))
    while stack:
        i = stack.pop()
        for code, lo, hi in wide(i):
            if code == 0xFFFFFFFF:
                assert np.all(lo == np.inf) and np.all(hi == -np.inf)
            elif code & LEAF:
                wide_leaves.append((code, tuple(lo), tuple(hi)))
            else:
                assert internal_boxes[code] == (tuple(lo), tuple(hi))
                stack.append(code)
    assert sorted(bin_leaves) == sorted(wide_leaves)
    assert len(wide_leaves) > 100
    if not check_collapse:
        return

    # the collapse rule itself, restated in float32: F(n, k) (k = 1..4), the
    # least summed box area of the wide nodes below a cut of n's subtree into
    # at most k slots, bottom-up (F = 0 for a leaf; F(n, 1) = area(n) +
    # S(n, 4), F(n, k) = min(F(n, 1), S(n, k)), S(n, k) = min over k1 of
    # F(left, k1) + F(right, k - k1), the first k1 on ties); then from each
    # wide root its budget of 4 split by the same minima, a slot opened when
    # S(c, k) < F(c, 1), its first child taking the slot and the second
    # appended
    f32 = np.float32

    def area(lo, hi):
        e = [max(f32(hi[a]) - f32(lo[a]), f32(0)) for a in range(3)]
        return (e[0] * e[1] + e[1] * e[2]) + e[2] * e[0]

    F = {}
    zero = [f32(0)] * 4

    def child_f(i, side):
        code = _binary_children(bn, i)[side][0]
        return zero if code & LEAF else F[code]

    def split(l, r, k):
        best, k1 = f32(l[0] + r[k - 2]), 1
        for j in range(2, k):
            c = f32(l[j - 1] + r[k - j - 1])
            if c < best:
                best, k1 = c, j
        return best, k1

    order, todo = [], [0]
    while todo:
        i = todo.pop()
        order.append(i)
        todo += [c for c, _, _ in _binary_children(bn, i) if not c & LEAF]
    if len(bn) > 1:
        for i in reversed(order):
            (_, lo0, hi0), (_, lo1, hi1) = _binary_children(bn, i)
            lo = [min(f32(lo0[a]), f32(lo1[a])) for a in range(3)]
            hi = [max(f32(hi0[a]), f32(hi1[a])) for a in range(3)]
            l, r = child_f(i, 0), child_f(i, 1)
            s2, s3, s4 = (split(l, r, k)[0] for k in (2, 3, 4))
            f1 = f32(area(lo, hi) + s4)
            F[i] = [f1, min(f1, s2), min(f1, s3), min(f1, s4)]

    stack, roots = [0], 0
    while stack:
        i = stack.pop()
        roots += 1
        bc = _binary_children(bn, i)
        k1 = split(child_f(i, 0), child_f(i, 1), 4)[1]
        slots = [[k1, bc[0]], [4 - k1, bc[1]]]
        q = 0
        while q < len(slots):
            kb, (code, _, _) = slots[q]
            if kb > 1 and not code & LEAF and len(slots) < 4:
                sc, j1 = split(child_f(code, 0), child_f(code, 1), kb)
                if not F[code][0] <= sc:
                    cc = _binary_children(bn, code)
                    slots.append([kb - j1, cc[1]])
                    slots[q] = [j1, cc[0]]
                    continue
            q += 1
        ch = [c for _, c in slots]
        got = wide(i)
        for k in range(4):
            if k < len(ch):
                assert got[k][0] == ch[k][0], (i, k)
                assert np.array_equal(got[k][1], ch[k][1]) and np.array_equal(got[k][2], ch[k][2]), (i, k)
            else:
                assert got[k][0] == 0xFFFFFFFF, (i, k)
        stack += [c for c, _, _ in ch if not c & LEAF]
    assert roots > 10 and roots == len(w4)
