// lbvh.hip - on-GPU linear BVH over world-space triangles (replaces the
// driver BLAS/TLAS build of renderer.cpp:1643-1916).
//
// 1. k_transform   world-space triangles (model * position, vk_mesh.cpp:476;
//                  44-byte vertices, uint32 indices, vk_mesh.cpp:436-442),
//                  triangle boxes and the centroid bounds (ordered-uint atomics)
// 2. k_morton      30-bit Morton code of each box centre on a grid of cubic
//                  cells, 1024 along the longest extent of the centre bounds
// 3. radix sort    stable LSD sort of (code, triangle id), four 8-bit digits:
//                  per-tile digit histograms, one exclusive scan, and a scatter
//                  whose in-tile ranks come from 64-lane ballot matching, so
//                  equal codes keep triangle-id order
// 4. k_karras      Karras (HPG 2012) hierarchy from the sorted keys, ties
//                  broken by the sorted index
// 5. k_refit_level bottom-up boxes, one launch per depth
// 6. k_treelet_level  treelet restructuring (Karras and Aila, HPG 2013), one
//                  launch per depth from the deepest; then the leaves are
//                  renumbered depth first (k_relabel_*) so that every subtree
//                  covers a contiguous range of sorted triangles again
// 7. k_layout      64-byte two-child-box nodes, subtrees of <= leaf_max
//                  triangles collapsed into leaf ranges; k_gather writes the
//                  sorted 48-byte triangles
// Every step is integer work or exact min/max, so the node and triangle arrays
// are bit-identical to a CPU build of the same definitions.
#include "device_math.h"
#include "internal.h"

#include <algorithm>

namespace hr {
using namespace hrd;

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int SCAN_THREADS = 1024;

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_init_bounds(uint32_t* b) {
    if (threadIdx.x < 3) b[threadIdx.x] = 0xFFFFFFFFu;
    else if (threadIdx.x < 6) b[threadIdx.x] = 0u;
}

__device__ __forceinline__ uint32_t find_prim(const uint32_t* off, uint32_t np, uint32_t g) {
    uint32_t lo = 0, hi = np;  // largest p with off[p] <= g
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_transform(BuildBuffers b, int with_bounds) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = g < b.tri_count;
    uint32_t omin[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t omax[3] = {0u, 0u, 0u};
    if (valid) {
        uint32_t p = find_prim(b.prim_tri_offset, b.prim_count, g);
        const PrimInfo pi = b.prims[p];
        const float* m = b.models + 16 * p;
        float M[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) M[i] = m[i];
        uint32_t local = g - pi.tri_offset;
        float w[9];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = b.indices[pi.index_base + 3 * local + k];
            const float* pos = (const float*)(b.verts + (size_t)(pi.vertex_base + vi) * 44u);
            float o[4];
            mul4(M, pos[0], pos[1], pos[2], 1.0f, o);
            w[3 * k + 0] = o[0];
            w[3 * k + 1] = o[1];
            w[3 * k + 2] = o[2];
            b.wtri[3 * (size_t)g + k] = make_float4(o[0], o[1], o[2], 0.0f);
        }
        float lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = minsel(minsel(w[a], w[3 + a]), w[6 + a]);
            hi[a] = maxsel(maxsel(w[a], w[3 + a]), w[6 + a]);
        }
        b.tbox[2 * (size_t)g + 0] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        b.tbox[2 * (size_t)g + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        b.tri_prim[g] = p;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t oc = f2ord((lo[a] + hi[a]) * 0.5f);
            omin[a] = oc;
            omax[a] = oc;
        }
    }
    if (!with_bounds) return;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            omin[a] = min(omin[a], (uint32_t)__shfl_xor((int)omin[a], off, 64));
            omax[a] = max(omax[a], (uint32_t)__shfl_xor((int)omax[a], off, 64));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&b.bounds[a], omin[a]);
            atomicMax(&b.bounds[3 + a], omax[a]);
        }
    }
}

__global__ __launch_bounds__(256) void k_morton(BuildBuffers b) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= b.tri_count) return;
    // cubic cells, 1024 along the longest extent of the centre bounds: a
    // flat scene spends its leading key bits on its long axes, so the upper
    // levels split the long axes instead of cutting thin slabs
    float cmin[3], cmax[3], ext[3], scale[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        cmin[a] = ord2f(b.bounds[a]);
        cmax[a] = ord2f(b.bounds[3 + a]);
        ext[a] = cmax[a] - cmin[a];
    }
    const float em = fmaxf(ext[0], fmaxf(ext[1], ext[2]));
#pragma unroll
    for (int a = 0; a < 3; ++a) scale[a] = ext[a] > 0.0f ? 1024.0f / em : 0.0f;
    float4 lo = b.tbox[2 * (size_t)g], hi = b.tbox[2 * (size_t)g + 1];
    float cx = (lo.x + hi.x) * 0.5f, cy = (lo.y + hi.y) * 0.5f, cz = (lo.z + hi.z) * 0.5f;
    b.keys[g] = morton3(quant10(cx, cmin[0], scale[0]), quant10(cy, cmin[1], scale[1]),
                        quant10(cz, cmin[2], scale[2]));
    b.ids[g] = g;
}

// ---- radix sort -------------------------------------------------------------
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint32_t* keys, uint32_t n, int shift,
                                                        uint32_t* hist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        uint32_t idx = base + j * RS_THREADS + tid;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)tid * nblocks + blockIdx.x] = h[tid];
}

// exclusive scan of len words in one workgroup
__global__ __launch_bounds__(SCAN_THREADS) void k_rs_scan(uint32_t* a, uint32_t len) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    __shared__ uint32_t carry_s;
    int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < len; base += SCAN_THREADS * 4) {
        uint32_t v[4];
        uint32_t i0 = base + tid * 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (i0 + k < len) ? a[i0 + k] : 0u;
        uint32_t s = v[0] + v[1] + v[2] + v[3];
        uint32_t inc = s;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t t = (uint32_t)__shfl_up((int)inc, off, 64);
            if (lane >= off) inc += t;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (w == 0) {
            uint32_t x = lane < SCAN_THREADS / 64 ? wsum[lane] : 0u;
            uint32_t y = x;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                uint32_t t = (uint32_t)__shfl_up((int)y, off, 64);
                if (lane >= off) y += t;
            }
            if (lane < SCAN_THREADS / 64) wsum[lane] = y - x;  // exclusive wave offsets
        }
        __syncthreads();
        uint32_t run = carry_s + wsum[w] + inc - s;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < len) a[i0 + k] = run;
            run += v[k];
        }
        __syncthreads();
        if (tid == SCAN_THREADS - 1) carry_s = run;
        __syncthreads();
    }
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                           uint32_t* vout, uint32_t n, int shift,
                                                           const uint32_t* hist, uint32_t nblocks) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[RS_THREADS / 64][256];
    int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    base[tid] = hist[(size_t)tid * nblocks + blockIdx.x];
#pragma unroll
    for (int q = 0; q < RS_THREADS / 64; ++q) wcnt[q][tid] = 0;
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int j = 0; j < RS_ITEMS; ++j) {
        uint32_t idx = blockIdx.x * RS_TILE + j * RS_THREADS + tid;
        bool valid = idx < n;
        uint32_t key = valid ? kin[idx] : 0u;
        uint32_t val = valid ? vin[idx] : 0u;
        uint32_t d = (key >> shift) & 255u;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            bool on = (d >> bit) & 1u;
            unsigned long long bb = __ballot(on);
            m &= on ? bb : ~bb;
        }
        uint32_t r = (uint32_t)__popcll(m & lt);
        if (valid && r == 0) wcnt[w][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + r;
            for (int q = 0; q < w; ++q) pos += wcnt[q][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int q = 0; q < RS_THREADS / 64; ++q) {
            tot += wcnt[q][tid];
            wcnt[q][tid] = 0;
        }
        base[tid] += tot;
        __syncthreads();
    }
}

// ---- Karras hierarchy -------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint32_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

__global__ __launch_bounds__(256) void k_karras(BuildBuffers b) {
    int n = (int)b.tri_count;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const uint32_t* k = b.keys;
    int d = (kdelta(k, n, i, i + 1) - kdelta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(k, n, i, i - d);
    int lmax = 2;
    while (kdelta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(k, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(k, n, i, j);
    int sp = 0;
    int div = 2;
    for (;;) {
        int t = (l + div - 1) / div;
        if (kdelta(k, n, i, i + (sp + t) * d) > dnode) sp += t;
        if (t <= 1) break;
        div *= 2;
    }
    int gamma = i + sp * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    uint32_t left = (lo == gamma) ? (LEAF_BIT | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t right = (hi == gamma + 1) ? (LEAF_BIT | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    b.kl[i] = left;
    b.kr[i] = right;
    b.kfirst[i] = (uint32_t)lo;
    b.klast[i] = (uint32_t)hi;
    if (left & LEAF_BIT) b.lparent[left & ~LEAF_BIT] = (uint32_t)i; else b.kparent[left] = (uint32_t)i;
    if (right & LEAF_BIT) b.lparent[right & ~LEAF_BIT] = (uint32_t)i; else b.kparent[right] = (uint32_t)i;
    if (i == 0) b.kparent[0] = 0xFFFFFFFFu;
}

__device__ __forceinline__ void child_box(const BuildBuffers& b, uint32_t c, float4& lo, float4& hi) {
    if (c & LEAF_BIT) {
        uint32_t g = b.ids[c & ~LEAF_BIT];
        lo = b.tbox[2 * (size_t)g];
        hi = b.tbox[2 * (size_t)g + 1];
    } else {
        lo = b.kbox[2 * (size_t)c];
        hi = b.kbox[2 * (size_t)c + 1];
    }
}

// Box refit, level by level from the deepest internal nodes up (one launch
// per level): a node's box is the union of its children's, so each level only
// reads boxes the previous launches wrote. A bottom-up pass with arrival
// counters would need a device-scope acquire/release per node, which on a
// GPU with one L2 per XCD writes back and invalidates caches each time.
// Depths and the per-level node lists depend on the topology only, so the
// build computes them once and every refit reuses them.
__global__ __launch_bounds__(256) void k_node_depth(BuildBuffers b, uint32_t nn) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    uint32_t d = 0;
    for (uint32_t j = i; j != 0u && d < REFIT_LEVELS - 1u; ++d) j = b.kparent[j];
    b.flags[i] = d;
}

__global__ __launch_bounds__(256) void k_level_count(BuildBuffers b, uint32_t nn) {
    __shared__ uint32_t h[REFIT_LEVELS];
    for (uint32_t k = threadIdx.x; k < REFIT_LEVELS; k += blockDim.x) h[k] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) atomicAdd(&h[b.flags[i]], 1u);
    __syncthreads();
    uint32_t* counts = b.lvl + nn + REFIT_LEVELS + 1u;
    for (uint32_t k = threadIdx.x; k < REFIT_LEVELS; k += blockDim.x)
        if (h[k]) atomicAdd(&counts[k], h[k]);
}

// offsets (exclusive sums of the counts); the counts become scatter cursors
__global__ void k_level_offsets(BuildBuffers b, uint32_t nn) {
    if (threadIdx.x != 0) return;
    uint32_t* off = b.lvl + nn;
    uint32_t* counts = off + REFIT_LEVELS + 1u;
    uint32_t acc = 0;
    for (uint32_t k = 0; k < REFIT_LEVELS; ++k) {
        off[k] = acc;
        acc += counts[k];
        counts[k] = off[k];
    }
    off[REFIT_LEVELS] = acc;
}

__global__ __launch_bounds__(256) void k_level_scatter(BuildBuffers b, uint32_t nn) {
    __shared__ uint32_t h[REFIT_LEVELS], base[REFIT_LEVELS];
    for (uint32_t k = threadIdx.x; k < REFIT_LEVELS; k += blockDim.x) h[k] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t d = 0, rank = 0;
    if (i < nn) {
        d = b.flags[i];
        rank = atomicAdd(&h[d], 1u);
    }
    __syncthreads();
    uint32_t* cursor = b.lvl + nn + REFIT_LEVELS + 1u;
    for (uint32_t k = threadIdx.x; k < REFIT_LEVELS; k += blockDim.x)
        if (h[k]) base[k] = atomicAdd(&cursor[k], h[k]);
    __syncthreads();
    if (i < nn) b.lvl[base[d] + rank] = i;
}

// deepest depth in use (read back by the build)
__global__ void k_level_depth(BuildBuffers b, uint32_t nn, uint32_t* out) {
    if (threadIdx.x != 0) return;
    const uint32_t* off = b.lvl + nn;
    uint32_t L = 0;
    for (uint32_t k = 0; k < REFIT_LEVELS; ++k)
        if (off[k + 1] > off[k]) L = k + 1u;
    *out = L;
}

__global__ __launch_bounds__(256) void k_refit_level(BuildBuffers b, uint32_t nn, uint32_t level) {
    const uint32_t* off = b.lvl + nn;
    const uint32_t first = off[level], last = off[level + 1u];
    for (uint32_t t = first + blockIdx.x * blockDim.x + threadIdx.x; t < last; t += gridDim.x * blockDim.x) {
        const uint32_t p = b.lvl[t];
        float4 alo, ahi, blo, bhi;
        child_box(b, b.kl[p], alo, ahi);
        child_box(b, b.kr[p], blo, bhi);
        b.kbox[2 * (size_t)p] = make_float4(minsel(alo.x, blo.x), minsel(alo.y, blo.y), minsel(alo.z, blo.z), 0.0f);
        b.kbox[2 * (size_t)p + 1] =
            make_float4(maxsel(ahi.x, bhi.x), maxsel(ahi.y, bhi.y), maxsel(ahi.z, bhi.z), 0.0f);
    }
}

// ---- treelet restructuring (Karras and Aila, HPG 2013) ----------------------
// TREELET_PASSES bottom-up passes, the first over the Karras tree, a launch
// per depth from the deepest (the nodes of one depth root disjoint subtrees,
// and everything below them is done; the depths are recomputed between
// passes). Node N's treelet grows from N's two children by
// replacing, while it has fewer than TREELET_LEAVES leaves, its internal
// leaf of largest half box area (the first in list order on ties) by that
// node's two children (the first in its place, the second appended). Over the
// subsets S of the treelet leaves (bit j = leaf j), by increasing S:
//   cost(S) = SAH_TRI * count(S) * area(S)                     count(S) <= leaf_max
//           = SAH_NODE * area(S) + min_P cost(P) + cost(S \ P)   otherwise
// with P over the subsets of S that hold S's lowest leaf (submasks in
// decreasing order; the first minimum, ties to the split whose parts differ
// least in triangle count). A treelet leaf costs its subtree's cost. The treelet
// is rebuilt in its cheapest topology: N stays the root, the other internal
// nodes are the ones it removed, in removal order, given out in pre-order.
// area = dx*dy + dy*dz + dz*dx, boxes are unions in increasing leaf order.
// Each restructured node's cost and triangle count go to the radix sort's
// spare arrays (keys_tmp as float bits, ids_tmp), free after the sort.
constexpr int TREELET_LEAVES = 7, TREELET_PASSES = 2;
constexpr float SAH_NODE = 1.2f, SAH_TRI = 1.0f;

__device__ __forceinline__ float half_area(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// One lane per treelet. The subset tables (cost and chosen split per subset)
// are the lane's column of two LDS arrays, [subset][lane]; the treelet's
// leaves, boxes and counts stay in registers (loops unrolled to
// TREELET_LEAVES with the leaf count as a guard).
constexpr uint32_t TREELET_BLOCK = 64;
constexpr int TREELET_SETS = 1 << TREELET_LEAVES;

__device__ void restructure_treelet(const BuildBuffers& b, uint32_t N, float* cost, uint32_t* cnt, float* C,
                                    uint8_t* part) {
    // C[S * TREELET_BLOCK], part[S * TREELET_BLOCK]: this lane's subset tables
    const uint32_t leaf_max = b.leaf_max;
    uint32_t tl[TREELET_LEAVES], freed[TREELET_LEAVES - 2];
#pragma unroll
    for (int j = 0; j < TREELET_LEAVES; ++j) tl[j] = 0u;
#pragma unroll
    for (int j = 0; j < TREELET_LEAVES - 2; ++j) freed[j] = 0u;
    int m = 2;
    tl[0] = b.kl[N];
    tl[1] = b.kr[N];
#pragma unroll
    for (int step = 0; step < TREELET_LEAVES - 2; ++step) {
        int bi = -1;
        float ba = -1.0f;
#pragma unroll
        for (int i = 0; i < TREELET_LEAVES; ++i) {
            if (i >= m || (tl[i] & LEAF_BIT)) continue;
            const float4 l = b.kbox[2 * (size_t)tl[i]], h = b.kbox[2 * (size_t)tl[i] + 1];
            const float lo3[3] = {l.x, l.y, l.z}, hi3[3] = {h.x, h.y, h.z};
            const float ar = half_area(lo3, hi3);
            if (ar > ba) {
                ba = ar;
                bi = i;
            }
        }
        if (bi < 0) break;
        uint32_t x = 0u;
#pragma unroll
        for (int i = 0; i < TREELET_LEAVES; ++i)
            if (i == bi) x = tl[i];
        const uint32_t xl = b.kl[x], xr = b.kr[x];
        freed[step] = x;
#pragma unroll
        for (int i = 0; i < TREELET_LEAVES; ++i) {
            if (i == bi) tl[i] = xl;
            if (i == m) tl[i] = xr;
        }
        ++m;
    }
    const int ns = 1 << m;
    float lo[TREELET_LEAVES][3], hi[TREELET_LEAVES][3];
    uint32_t kc[TREELET_LEAVES];
#pragma unroll
    for (int j = 0; j < TREELET_LEAVES; ++j) {
        kc[j] = 0u;
#pragma unroll
        for (int q = 0; q < 3; ++q) lo[j][q] = hi[j][q] = 0.0f;
        if (j >= m) continue;
        float4 l, h;
        child_box(b, tl[j], l, h);
        lo[j][0] = l.x; lo[j][1] = l.y; lo[j][2] = l.z;
        hi[j][0] = h.x; hi[j][1] = h.y; hi[j][2] = h.z;
        const uint32_t k = (tl[j] & LEAF_BIT) ? 1u : cnt[tl[j]];
        kc[j] = k;
        C[(1 << j) * TREELET_BLOCK] = ((tl[j] & LEAF_BIT) || k <= leaf_max)
                                          ? SAH_TRI * (float)k * half_area(lo[j], hi[j])
                                          : cost[tl[j]];
        part[(1 << j) * TREELET_BLOCK] = 0;
    }
    // union box and triangle count of a subset, members in increasing order
    auto set_box = [&](int S, float* bl, float* bh) {
        uint32_t k = 0;
        bool firstm = true;
#pragma unroll
        for (int j = 0; j < TREELET_LEAVES; ++j) {
            if (j >= m || !((S >> j) & 1)) continue;
            k += kc[j];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                bl[q] = firstm ? lo[j][q] : minsel(bl[q], lo[j][q]);
                bh[q] = firstm ? hi[j][q] : maxsel(bh[q], hi[j][q]);
            }
            firstm = false;
        }
        return k;
    };
    auto tri_count = [&](int X) {
        uint32_t k = 0;
#pragma unroll
        for (int j = 0; j < TREELET_LEAVES; ++j)
            if (j < m && ((X >> j) & 1)) k += kc[j];
        return k;
    };
    for (int S = 1; S < ns; ++S) {
        if ((S & (S - 1)) == 0) continue;
        const int low = S & -S;
        float best = __uint_as_float(0x7f800000u);
        int bp = 0;
        // ties: the split whose parts differ least in triangle count
        auto imbalance = [&](int P) -> uint32_t {
            if (P == 0) return 0xFFFFFFFFu;
            const uint32_t a = tri_count(P), b = tri_count(S ^ P);
            return a > b ? a - b : b - a;
        };
        for (int P = (S - 1) & S; P > 0; P = (P - 1) & S) {
            if (!(P & low)) continue;
            const float c = C[P * TREELET_BLOCK] + C[(S ^ P) * TREELET_BLOCK];
            if (c < best || (c == best && imbalance(P) < imbalance(bp))) {
                best = c;
                bp = P;
            }
        }
        float bl[3], bh[3];
        const uint32_t k = set_box(S, bl, bh);
        const float A = half_area(bl, bh);
        C[S * TREELET_BLOCK] = k <= leaf_max ? SAH_TRI * (float)k * A : SAH_NODE * A + best;
        part[S * TREELET_BLOCK] = (uint8_t)bp;
    }
    // the internal subsets in pre-order, then their nodes
    int order[TREELET_LEAVES - 1], stk[TREELET_LEAVES];
#pragma unroll
    for (int j = 0; j < TREELET_LEAVES - 1; ++j) order[j] = 0;
    int no = 0, sp = 0;
    stk[sp++] = ns - 1;
    while (sp) {
        const int S = stk[--sp];
        order[no++] = S;
        const int ps = (int)part[S * TREELET_BLOCK];
        const int P = ps ? ps : (S & -S), Q = S ^ P;
        if (Q & (Q - 1)) stk[sp++] = Q;
        if (P & (P - 1)) stk[sp++] = P;
    }
    auto code_of = [&](int S) -> uint32_t {
        uint32_t code = 0xFFFFFFFFu;
        if ((S & (S - 1)) == 0) {
#pragma unroll
            for (int j = 0; j < TREELET_LEAVES; ++j)
                if (S == (1 << j)) code = tl[j];
            return code;
        }
#pragma unroll
        for (int t = 0; t < TREELET_LEAVES - 1; ++t) {
            if (t >= no || order[t] != S) continue;
            code = N;
#pragma unroll
            for (int f = 0; f < TREELET_LEAVES - 2; ++f)
                if (t == f + 1) code = freed[f];
        }
        return code;
    };
#pragma unroll
    for (int t = 0; t < TREELET_LEAVES - 1; ++t) {
        if (t >= no) continue;
        const int S = order[t];
        const int ps = (int)part[S * TREELET_BLOCK];
        const int P = ps ? ps : (S & -S), Q = S ^ P;
        uint32_t id = N;
#pragma unroll
        for (int f = 0; f < TREELET_LEAVES - 2; ++f)
            if (t == f + 1) id = freed[f];
        const uint32_t l = code_of(P), r = code_of(Q);
        b.kl[id] = l;
        b.kr[id] = r;
        float bl[3], bh[3];
        const uint32_t k = set_box(S, bl, bh);
        b.kbox[2 * (size_t)id] = make_float4(bl[0], bl[1], bl[2], 0.0f);
        b.kbox[2 * (size_t)id + 1] = make_float4(bh[0], bh[1], bh[2], 0.0f);
        if (l & LEAF_BIT) b.lparent[l & ~LEAF_BIT] = id; else b.kparent[l] = id;
        if (r & LEAF_BIT) b.lparent[r & ~LEAF_BIT] = id; else b.kparent[r] = id;
        cost[id] = C[S * TREELET_BLOCK];
        cnt[id] = k;
    }
}

__global__ __launch_bounds__(TREELET_BLOCK) void k_treelet_level(BuildBuffers b, uint32_t nn, uint32_t level) {
    __shared__ float sC[TREELET_SETS * TREELET_BLOCK];
    __shared__ uint8_t sPart[TREELET_SETS * TREELET_BLOCK];
    const uint32_t* off = b.lvl + nn;
    const uint32_t first = off[level], last = off[level + 1u];
    float* cost = reinterpret_cast<float*>(b.keys_tmp);
    for (uint32_t t = first + blockIdx.x * blockDim.x + threadIdx.x; t < last; t += gridDim.x * blockDim.x)
        restructure_treelet(b, b.lvl[t], cost, b.ids_tmp, sC + threadIdx.x, sPart + threadIdx.x);
}

// Depth-first numbering (first child first) from the subtree triangle counts
// (ids_tmp): a node's first leaf is the sum of the counts of the first
// children along its path whose second child the path takes. Internal nodes
// get their ranges, leaves their new positions (keys_tmp; the costs are dead).
__global__ __launch_bounds__(256) void k_relabel_walk(BuildBuffers b, uint32_t n) {
    const uint32_t nn = n - 1u;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn + n) return;
    const uint32_t* cnt = b.ids_tmp;
    const uint32_t self = i < nn ? i : (LEAF_BIT | (i - nn));
    uint32_t c = self, first = 0;
    // a tree that passes the depth check is far shallower than the bound
    for (uint32_t step = 0; c != 0u && step < 4096u; ++step) {
        const uint32_t p = (c & LEAF_BIT) ? b.lparent[c & ~LEAF_BIT] : b.kparent[c];
        if (b.kr[p] == c) {
            const uint32_t l = b.kl[p];
            first += (l & LEAF_BIT) ? 1u : cnt[l];
        }
        c = p;
    }
    if (i < nn) {
        b.kfirst[i] = first;
        b.klast[i] = first + cnt[i] - 1u;
    } else {
        b.keys_tmp[i - nn] = first;
    }
}

// leaf codes and sorted ids in the new order (ids_tmp; the counts are dead)
__global__ __launch_bounds__(256) void k_relabel_apply(BuildBuffers b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t* pos = b.keys_tmp;
    if (i + 1u < n) {
        const uint32_t l = b.kl[i], r = b.kr[i];
        if (l & LEAF_BIT) b.kl[i] = LEAF_BIT | pos[l & ~LEAF_BIT];
        if (r & LEAF_BIT) b.kr[i] = LEAF_BIT | pos[r & ~LEAF_BIT];
    }
    if (i < n) b.ids_tmp[pos[i]] = b.ids[i];
}

__global__ __launch_bounds__(256) void k_leaf_parents(BuildBuffers b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1u >= n) return;
    const uint32_t l = b.kl[i], r = b.kr[i];
    if (l & LEAF_BIT) b.lparent[l & ~LEAF_BIT] = i;
    if (r & LEAF_BIT) b.lparent[r & ~LEAF_BIT] = i;
}

__global__ __launch_bounds__(256) void k_layout(BuildBuffers b) {
    uint32_t n = b.tri_count;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 1) {
        if (i == 0) {
            uint32_t g = b.ids[0];
            float4 lo = b.tbox[2 * (size_t)g], hi = b.tbox[2 * (size_t)g + 1];
            TNode nd;
            nd.a = make_float4(lo.x, lo.y, lo.z, __uint_as_float(leaf_code(0, 1)));
            nd.b = make_float4(hi.x, hi.y, hi.z, __uint_as_float(leaf_code(0, 1)));
            nd.c = make_float4(lo.x, lo.y, lo.z, 0.0f);
            nd.d = make_float4(hi.x, hi.y, hi.z, 0.0f);
            b.nodes[0] = nd;
        }
        return;
    }
    if (i >= n - 1) return;
    uint32_t ch[2] = {b.kl[i], b.kr[i]};
    uint32_t code[2];
    float4 lo[2], hi[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        child_box(b, ch[c], lo[c], hi[c]);
        if (ch[c] & LEAF_BIT) {
            code[c] = leaf_code(ch[c] & ~LEAF_BIT, 1);
        } else {
            uint32_t sz = b.klast[ch[c]] - b.kfirst[ch[c]] + 1;
            code[c] = sz <= b.leaf_max ? leaf_code(b.kfirst[ch[c]], sz) : ch[c];
        }
    }
    TNode nd;
    nd.a = make_float4(lo[0].x, lo[0].y, lo[0].z, __uint_as_float(code[0]));
    nd.b = make_float4(hi[0].x, hi[0].y, hi[0].z, __uint_as_float(code[1]));
    nd.c = make_float4(lo[1].x, lo[1].y, lo[1].z, 0.0f);
    nd.d = make_float4(hi[1].x, hi[1].y, hi[1].z, 0.0f);
    b.nodes[i] = nd;
}

// 4-wide nodes, top-down, one launch per wide level. The wide node rooted at
// binary node i takes as its 2-4 child slots the cut of i's subtree below i
// that minimises the summed box area of the wide nodes the collapse then
// needs (the surface-area estimate of the wide steps a ray takes; every leaf
// is one slot of some wide node whatever the cut, so leaves cost nothing).
// Bottom-up, F(n, k) (k = 1..4) is the least such area below a cut of n's
// subtree into at most k slots: F(n, 1) = area(n) + S(n, 4), F(n, k) =
// min(F(n, 1), S(n, k)), S(n, k) = min over k1 of F(left, k1) + F(right,
// k - k1); a leaf has F = 0. Top-down, each wide root splits its budget of 4
// by the same minima (the first on ties, a single slot when F(n, 1) <= S(n,
// k)). On the 1M-triangle terrain this takes 20% fewer wide nodes and 3%
// fewer expected wide steps per ray than greedily opening the child of
// largest area.
__device__ __forceinline__ float box_area(float4 l, float4 h) {
    const float ex = fmaxf(h.x - l.x, 0.0f), ey = fmaxf(h.y - l.y, 0.0f), ez = fmaxf(h.z - l.z, 0.0f);
    return (ex * ey + ey * ez) + ez * ex;
}

// F of child `side` of binary node nd (zero for a leaf)
__device__ __forceinline__ float4 child_cost(const TNode& nd, int side, const float4* __restrict__ wcost) {
    const uint32_t code = __float_as_uint(side ? nd.b.w : nd.a.w);
    return (code & LEAF_BIT) ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : wcost[code];
}
// S(n, k) from the children's F, and its left budget k1 (the first minimum)
__device__ __forceinline__ float split_cost(const float4& l, const float4& r, uint32_t k, uint32_t& k1) {
    const float fl[4] = {l.x, l.y, l.z, l.w}, fr[4] = {r.x, r.y, r.z, r.w};
    float best = fl[0] + fr[k - 2u];
    k1 = 1u;
    for (uint32_t j = 2u; j < k; ++j) {
        const float c = fl[j - 1u] + fr[k - j - 1u];
        if (c < best) {
            best = c;
            k1 = j;
        }
    }
    return best;
}

// F of the internal nodes of one refit level (children first)
__global__ __launch_bounds__(256) void k_wide_cost_level(BuildBuffers b, uint32_t nn, uint32_t level,
                                                         float4* wcost) {
    const uint32_t* off = b.lvl + nn;
    const uint32_t first = off[level], last = off[level + 1u];
    for (uint32_t t = first + blockIdx.x * blockDim.x + threadIdx.x; t < last; t += gridDim.x * blockDim.x) {
        const uint32_t p = b.lvl[t];
        const TNode nd = b.nodes[p];
        const float4 l = child_cost(nd, 0, wcost), r = child_cost(nd, 1, wcost);
        const float4 lo = make_float4(fminf(nd.a.x, nd.c.x), fminf(nd.a.y, nd.c.y), fminf(nd.a.z, nd.c.z), 0.0f);
        const float4 hi = make_float4(fmaxf(nd.b.x, nd.d.x), fmaxf(nd.b.y, nd.d.y), fmaxf(nd.b.z, nd.d.z), 0.0f);
        uint32_t k1;
        const float s2 = split_cost(l, r, 2u, k1), s3 = split_cost(l, r, 3u, k1), s4 = split_cost(l, r, 4u, k1);
        const float f1 = box_area(lo, hi) + s4;
        wcost[p] = make_float4(f1, fminf(f1, s2), fminf(f1, s3), fminf(f1, s4));
    }
}

__global__ void k_collapse_begin(uint32_t* roots, uint32_t* counts) {
    if (threadIdx.x == 0) {
        roots[0] = 0u;  // the binary root
        counts[0] = 1u;
        counts[1] = 0u;
    }
}

__global__ void k_collapse_consumed(uint32_t* count) {
    if (threadIdx.x == 0) *count = 0u;
}

// The choice is recorded, not yet written as a node: wide roots are flagged
// in `flag` (by binary index) and each child slot's binary source (node << 1
// | side) goes to src[i]; lbvh_build_wide then numbers the wide roots densely
// in binary-index order and writes the nodes.
__global__ __launch_bounds__(256) void k_collapse_level(BuildBuffers b, const uint32_t* roots,
                                                        const uint32_t* nroots, uint32_t* next,
                                                        uint32_t* nnext, uint32_t* flag, uint4* src_out,
                                                        const float4* __restrict__ wcost) {
    const uint32_t n_in = *nroots;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_in; t += gridDim.x * blockDim.x) {
        const uint32_t i = roots[t];
        // slots as (binary node, side) with their budgets: the root's 4
        // split between its children, then each slot with budget k > 1 that
        // is an internal node opened when S(c, k) < F(c, 1)
        uint32_t node[4], side[4], kb[4], code[4], src[4];
        const TNode nd = b.nodes[i];
        uint32_t k1;
        split_cost(child_cost(nd, 0, wcost), child_cost(nd, 1, wcost), 4u, k1);
        node[0] = i;
        side[0] = 0u;
        kb[0] = k1;
        code[0] = __float_as_uint(nd.a.w);
        node[1] = i;
        side[1] = 1u;
        kb[1] = 4u - k1;
        code[1] = __float_as_uint(nd.b.w);
        uint32_t m = 2;
        for (uint32_t q = 0; q < m;) {
            const uint32_t c = code[q];
            if (kb[q] > 1u && !(c & LEAF_BIT) && m < 4u) {
                const TNode cn = b.nodes[c];
                uint32_t j1;
                const float sc = split_cost(child_cost(cn, 0, wcost), child_cost(cn, 1, wcost), kb[q], j1);
                if (!(wcost[c].x <= sc)) {  // opened: its first child takes the slot, the second is appended
                    node[m] = c;
                    side[m] = 1u;
                    kb[m] = kb[q] - j1;
                    code[m] = __float_as_uint(cn.b.w);
                    ++m;
                    node[q] = c;
                    side[q] = 0u;
                    kb[q] = j1;
                    code[q] = __float_as_uint(cn.a.w);
                    continue;
                }
            }
            ++q;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) src[k] = (uint32_t)k < m ? (node[k] << 1) | side[k] : EMPTY4;
        flag[i] = 1u;
        src_out[i] = make_uint4(src[0], src[1], src[2], src[3]);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if ((uint32_t)k < m && !(code[k] & LEAF_BIT)) next[atomicAdd(nnext, 1u)] = code[k];
    }
}

// A 4-wide node from the binary sources of its child slots: child k is side
// (src & 1) of binary node src >> 1 (its box and code as that node stores
// them; internal codes mapped to dense indices by `dense`), EMPTY4 slots get
// the inverted box. `binary` (the binary node the wide node is rooted at) is
// kept in the padding word.
__device__ TNode4 wide_node(const TNode* __restrict__ nodes, uint4 src, const uint32_t* __restrict__ dense,
                            uint32_t binary) {
    const uint32_t s[4] = {src.x, src.y, src.z, src.w};
    const float inf = __uint_as_float(0x7f800000u);
    float4 lo[4], hi[4];
    uint32_t code[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (s[k] == EMPTY4) {
            code[k] = EMPTY4;
            lo[k] = make_float4(inf, inf, inf, 0.0f);
            hi[k] = make_float4(-inf, -inf, -inf, 0.0f);
            continue;
        }
        const TNode nd = nodes[s[k] >> 1];
        const bool second = (s[k] & 1u) != 0u;
        lo[k] = second ? nd.c : nd.a;
        hi[k] = second ? nd.d : nd.b;
        code[k] = __float_as_uint(second ? nd.b.w : nd.a.w);
        if (dense && !(code[k] & LEAF_BIT)) code[k] = dense[code[k]];
    }
    TNode4 q;
    q.lox = make_float4(lo[0].x, lo[1].x, lo[2].x, lo[3].x);
    q.hix = make_float4(hi[0].x, hi[1].x, hi[2].x, hi[3].x);
    q.loy = make_float4(lo[0].y, lo[1].y, lo[2].y, lo[3].y);
    q.hiy = make_float4(hi[0].y, hi[1].y, hi[2].y, hi[3].y);
    q.loz = make_float4(lo[0].z, lo[1].z, lo[2].z, lo[3].z);
    q.hiz = make_float4(hi[0].z, hi[1].z, hi[2].z, hi[3].z);
    q.code = make_uint4(code[0], code[1], code[2], code[3]);
    q.pad = make_uint4(binary, 0u, 0u, 0u);
    return q;
}

// Dense numbering of the wide roots: an exclusive scan of the flags over
// tiles of SCAN_THREADS * 4 words (tile sums, their scan, then each tile's
// own scan from its offset); the total goes to *total.
__global__ __launch_bounds__(SCAN_THREADS) void k_tile_sums(const uint32_t* a, uint32_t len, uint32_t* sums) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const uint32_t base = blockIdx.x * SCAN_THREADS * 4u + threadIdx.x * 4u;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) s += base + k < len ? a[base + k] : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += (uint32_t)__shfl_down((int)s, off, 64);
    if ((threadIdx.x & 63u) == 0u) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < SCAN_THREADS / 64; ++w) t += wsum[w];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_tile_scan(uint32_t* a, uint32_t len, const uint32_t* offs,
                                                            uint32_t ntiles, uint32_t* total) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t i0 = blockIdx.x * SCAN_THREADS * 4u + tid * 4u;
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = i0 + k < len ? a[i0 + k] : 0u;
    const uint32_t s = v[0] + v[1] + v[2] + v[3];
    uint32_t inc = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < w; ++k) before += wsum[k];
    uint32_t run = offs[blockIdx.x] + before + inc - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (i0 + k < len) a[i0 + k] = run;
        run += v[k];
    }
    if (blockIdx.x == ntiles - 1u && tid == SCAN_THREADS - 1) *total = run;
}

// the 4-wide nodes at their dense slots (dense = the scanned flags: binary
// node i is a wide root when the prefix grows after it)
__global__ __launch_bounds__(256) void k_wide_fill(BuildBuffers b, uint32_t nn, const uint32_t* dense,
                                                   const uint4* src, const uint32_t* total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const uint32_t d = dense[i];
    const uint32_t d_next = i + 1u < nn ? dense[i + 1u] : *total;
    if (d_next == d) return;
    const uint4 s = src[i];
    b.nodes4[d] = wide_node(b.nodes, s, dense, i);
    b.wsrc[d] = s;
}

// refit: the same 4-wide nodes with boxes from the refitted binary nodes
// (codes unchanged)
__global__ __launch_bounds__(256) void k_wide_refit(BuildBuffers b) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= b.wide_count) return;
    const TNode4 old = b.nodes4[d];
    TNode4 q = wide_node(b.nodes, b.wsrc[d], nullptr, old.pad.x);
    q.code = old.code;
    b.nodes4[d] = q;
}

__global__ __launch_bounds__(256) void k_gather(BuildBuffers b) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.tri_count) return;
    uint32_t g = b.ids[k];
    float4 v0 = b.wtri[3 * (size_t)g], v1 = b.wtri[3 * (size_t)g + 1], v2 = b.wtri[3 * (size_t)g + 2];
    STri t;
    t.v0 = make_float4(v0.x, v0.y, v0.z, __uint_as_float(g));
    t.v1 = make_float4(v1.x, v1.y, v1.z, 0.0f);
    t.v2 = make_float4(v2.x, v2.y, v2.z, 0.0f);
    b.stris[k] = t;
}

// deepest leaf path counted in traversal nodes (the root and subtrees with
// more than leaf_max triangles)
__global__ __launch_bounds__(256) void k_depth(BuildBuffers b) {
    uint32_t n = b.tri_count;
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t d = 0;
    if (n >= 2) {
        uint32_t a = b.lparent[k];
        for (uint32_t step = 0; step < 4096u; ++step) {
            uint32_t sz = b.klast[a] - b.kfirst[a] + 1;
            if (a == 0 || sz > b.leaf_max) ++d;
            if (a == 0) break;
            a = b.kparent[a];
        }
    } else {
        d = 1;
    }
    atomicMax(b.depth, d);
}

inline uint32_t blocks(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

hipError_t radix_sort(BuildBuffers& b, hipStream_t st) {
    uint32_t n = b.tri_count;
    uint32_t nb = blocks(n, RS_TILE);
    uint32_t* kin = b.keys;
    uint32_t* vin = b.ids;
    uint32_t* kout = b.keys_tmp;
    uint32_t* vout = b.ids_tmp;
    for (int pass = 0; pass < 4; ++pass) {
        int shift = pass * 8;
        k_rs_hist<<<nb, RS_THREADS, 0, st>>>(kin, n, shift, b.hist, nb);
        k_rs_scan<<<1, SCAN_THREADS, 0, st>>>(b.hist, 256u * nb);
        k_rs_scatter<<<nb, RS_THREADS, 0, st>>>(kin, vin, kout, vout, n, shift, b.hist, nb);
        uint32_t* t = kin; kin = kout; kout = t;
        t = vin; vin = vout; vout = t;
    }
    return hipGetLastError();  // four passes end in keys/ids
}

// the refit order (depths and per-level node lists) of a new topology; reads
// back the number of levels, so it synchronises the stream
hipError_t refit_order(BuildBuffers& b, hipStream_t st) {
    const uint32_t nn = b.tri_count - 1u;
    k_node_depth<<<blocks(nn, 256), 256, 0, st>>>(b, nn);
    k_fill_u32<<<1, 256, 0, st>>>(b.lvl + nn + REFIT_LEVELS + 1u, REFIT_LEVELS, 0u);
    k_level_count<<<blocks(nn, 256), 256, 0, st>>>(b, nn);
    k_level_offsets<<<1, 64, 0, st>>>(b, nn);
    k_level_scatter<<<blocks(nn, 256), 256, 0, st>>>(b, nn);
    k_level_depth<<<1, 64, 0, st>>>(b, nn, b.depth);
    hipError_t e = hipMemcpyAsync(&b.refit_levels, b.depth, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

void refit_boxes(BuildBuffers& b, hipStream_t st) {
    const uint32_t nn = b.tri_count - 1u;
    const uint32_t grid = std::min(blocks(nn, 256), 1024u);
    for (uint32_t L = b.refit_levels; L-- > 0u;) k_refit_level<<<grid, 256, 0, st>>>(b, nn, L);
}

// the treelet pass over the Karras tree (boxes and refit order in place),
// the depth-first renumbering, and the refit order of the new topology
hipError_t restructure(BuildBuffers& b, hipStream_t st) {
    const uint32_t n = b.tri_count, nn = n - 1u;
    const uint32_t tgrid = std::min(blocks(nn, TREELET_BLOCK), 4096u);
    for (int pass = 0; pass < TREELET_PASSES; ++pass) {
        if (pass > 0) {
            hipError_t e = refit_order(b, st);  // the levels of the restructured tree
            if (e != hipSuccess) return e;
        }
        for (uint32_t L = b.refit_levels; L-- > 0u;) k_treelet_level<<<tgrid, TREELET_BLOCK, 0, st>>>(b, nn, L);
    }
    k_relabel_walk<<<blocks(nn + n, 256), 256, 0, st>>>(b, n);
    k_relabel_apply<<<blocks(n, 256), 256, 0, st>>>(b, n);
    hipError_t e = hipMemcpyAsync(b.ids, b.ids_tmp, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    k_leaf_parents<<<blocks(nn, 256), 256, 0, st>>>(b, n);
    return refit_order(b, st);
}

// binary boxes bottom-up, the binary nodes, the sorted triangles and the depth
hipError_t finish_binary(BuildBuffers& b, hipStream_t st) {
    uint32_t n = b.tri_count;
    if (n >= 2) refit_boxes(b, st);
    uint32_t nn = n >= 2 ? n - 1 : 1;
    k_layout<<<blocks(nn, 256), 256, 0, st>>>(b);
    k_gather<<<blocks(n, 256), 256, 0, st>>>(b);
    k_fill_u32<<<1, 64, 0, st>>>(b.depth, 1, 0u);
    k_depth<<<blocks(n, 256), 256, 0, st>>>(b);
    return hipGetLastError();
}

// Scratch of the wide-node choice, free once the binary nodes are laid out:
// the depth flags (node_count words) hold the wide-root flags and then their
// dense numbers, the internal-node boxes (32 bytes per node) the child
// sources (16 bytes per node) and F (16 bytes per node after them), the
// radix histogram the tile sums and total.
uint32_t* wide_dense(BuildBuffers& b) { return b.flags; }
uint4* wide_src(BuildBuffers& b) { return reinterpret_cast<uint4*>(b.kbox); }
float4* wide_cost(BuildBuffers& b) { return b.kbox + (b.tri_count >= 2u ? b.tri_count - 1u : 1u); }
uint32_t* wide_total(BuildBuffers& b) { return b.hist + 4; }
uint32_t* wide_sums(BuildBuffers& b) { return b.hist + 64; }

// The wide roots and their child sources, level by level from the root:
// roots alternate between the radix sort's spare key and id arrays (n entries
// each; a level has fewer than n roots), counts in the histogram buffer. A
// wide level descends at least one binary level, so as many levels as the
// binary tree has internal-node depths reach every leaf (one for a single
// triangle). Then the dense numbering; its total is read back.
hipError_t choose_wide(BuildBuffers& b, hipStream_t st) {
    const uint32_t n = b.tri_count;
    const uint32_t nn = n >= 2 ? n - 1 : 1;
    uint32_t* roots[2] = {b.keys_tmp, b.ids_tmp};
    uint32_t* counts = b.hist;
    uint32_t* dense = wide_dense(b);
    k_fill_u32<<<blocks(nn, 256), 256, 0, st>>>(dense, nn, 0u);
    const uint32_t grid = std::min(blocks(nn, 256), 1024u);
    const uint32_t levels = std::max(b.refit_levels, 1u);
    // F of every internal node, children first (a single triangle's root
    // has only leaf children and reads none)
    float4* wcost = wide_cost(b);
    if (n >= 2)
        for (uint32_t L = b.refit_levels; L-- > 0u;) k_wide_cost_level<<<grid, 256, 0, st>>>(b, nn, L, wcost);
    k_collapse_begin<<<1, 64, 0, st>>>(roots[0], counts);
    for (uint32_t level = 0; level < levels; ++level) {
        const uint32_t a = level & 1u;
        k_collapse_level<<<grid, 256, 0, st>>>(b, roots[a], counts + a, roots[a ^ 1u], counts + (a ^ 1u), dense,
                                                wide_src(b), wcost);
        k_collapse_consumed<<<1, 64, 0, st>>>(counts + a);
    }
    const uint32_t ntiles = blocks(nn, SCAN_THREADS * 4);
    k_tile_sums<<<ntiles, SCAN_THREADS, 0, st>>>(dense, nn, wide_sums(b));
    k_rs_scan<<<1, SCAN_THREADS, 0, st>>>(wide_sums(b), ntiles);
    k_tile_scan<<<ntiles, SCAN_THREADS, 0, st>>>(dense, nn, wide_sums(b), ntiles, wide_total(b));
    hipError_t e = hipMemcpyAsync(&b.wide_count, wide_total(b), sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

}  // namespace

size_t radix_hist_words(uint32_t n) { return 256u * (size_t)blocks(n > 0 ? n : 1, RS_TILE); }

hipError_t lbvh_build(BuildBuffers& b, hipStream_t st, bool treelets) {
    uint32_t n = b.tri_count;
    b.wide_count = 0;
    if (n == 0) return hipSuccess;
    k_init_bounds<<<1, 64, 0, st>>>(b.bounds);
    k_transform<<<blocks(n, 256), 256, 0, st>>>(b, 1);
    k_morton<<<blocks(n, 256), 256, 0, st>>>(b);
    hipError_t e = radix_sort(b, st);
    if (e != hipSuccess) return e;
    if (n >= 2) {
        k_karras<<<blocks(n - 1, 256), 256, 0, st>>>(b);
        e = refit_order(b, st);
        if (e != hipSuccess) return e;
    }
    const bool restructured = treelets && n >= 3;
    if (restructured) {
        refit_boxes(b, st);
        e = restructure(b, st);
        if (e != hipSuccess) return e;
    }
    e = finish_binary(b, st);
    if (e != hipSuccess) return e;
    if (restructured) {
        // the treelet cost ignores depth: a tree too deep for the traversal
        // stack is rebuilt without it (the plain tree's own depth check
        // follows in the caller)
        uint32_t depth = 0;
        e = hipMemcpyAsync(&depth, b.depth, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        if (depth > MAX_DEPTH) return lbvh_build(b, st, false);
    }
    return choose_wide(b, st);
}

hipError_t lbvh_build_wide(BuildBuffers& b, hipStream_t st) {
    const uint32_t n = b.tri_count;
    if (n == 0) return hipSuccess;
    const uint32_t nn = n >= 2 ? n - 1 : 1;
    k_wide_fill<<<blocks(nn, 256), 256, 0, st>>>(b, nn, wide_dense(b), wide_src(b), wide_total(b));
    return hipGetLastError();
}

hipError_t lbvh_refit(BuildBuffers& b, hipStream_t st) {
    uint32_t n = b.tri_count;
    if (n == 0) return hipSuccess;
    k_transform<<<blocks(n, 256), 256, 0, st>>>(b, 0);
    hipError_t e = finish_binary(b, st);
    if (e != hipSuccess) return e;
    if (b.wide_count) k_wide_refit<<<blocks(b.wide_count, 256), 256, 0, st>>>(b);
    return hipGetLastError();
}

}  // namespace hr
