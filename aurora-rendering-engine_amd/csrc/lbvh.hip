// lbvh.hip — LBVH build on gfx950.
//
// Pipeline (one stream, one host read of the per-depth node counts):
//   k_prim_bounds (padded primitive boxes, per-block scene/centroid bounds)
//   -> k_bounds_final (one block reduces the block bounds; no atomics)
//   -> k_morton (30-bit codes) -> 4 x {k_hist, scan, k_scatter} (stable LSD
//   radix sort of (code, id) with 8-bit digits; the scatter ranks keys with
//   wave64 ballots and LDS per-wave digit counts) -> k_karras (Karras 2012
//   radix tree) -> depth of every internal node by pointer jumping
//   (k_jump_init, 7 x k_jump, k_jump_out) -> internal nodes sorted by depth
//   (one radix pass) -> k_leaves, then one k_refit_level per depth, deepest
//   first -> four-wide tree: one k_wide_mark per depth, shallowest first, a
//   scan of the wide-node flags, k_wide_fill.  Each level reads child boxes written by earlier launches, so the
//   refit needs no arrival counters: on gfx950 an agent-scope acquire or
//   release writes back or invalidates the XCD's L2, and one pair per node
//   cost 19 ms at 4M primitives.
//
// Layout: node 0 is the root, node 1 padding; the children of the internal
// node with split gamma sit at 2+2*gamma and 3+2*gamma, so every child pair
// is one aligned 64-byte line.  Each node is two float4: (lo, child) and
// (hi, parent).
//
// No reference counterpart: the reference scans every triangle for every ray
// (experiments/rt.cpp:209-218).
#include <algorithm>
#include <vector>

#include "internal.hpp"
#include "traverse.hpp"

namespace aur {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;
constexpr uint32_t kNone = 0xffffffffu;

__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float o2f(uint32_t o) {
    union {
        uint32_t u;
        float f;
    } c;
    c.u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return c.f;
}

__device__ __forceinline__ uint32_t wave_umin(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o));
    return v;
}

// bounds: [0..2] scene lo, [3..5] scene hi, [6..8] centroid lo, [9..11] hi.
// Reduced as ordered u32 (a total order: the result does not depend on the
// grouping, -0 sorts below +0) in two stages: each block writes its 12
// values to part, k_bounds_final reduces those.
constexpr int kBoundBlocks = 1024;

__device__ __forceinline__ bool is_lo(int k) { return ((k / 3) % 2) == 0; }

__global__ __launch_bounds__(kBlock) void k_prim_bounds(DevScene sc, uint32_t n, float4* box_lo, float4* box_hi,
                                                        uint32_t* part) {
    __shared__ uint32_t red[4][12];
    uint32_t r[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) r[k] = is_lo(k) ? 0xffffffffu : 0u;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        v3 lo, hi;
        prim_box(sc, i, lo, hi);
        box_lo[i] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        box_hi[i] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        const float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t c = f2o((l[k] + h[k]) * 0.5f);
            r[k] = min(r[k], f2o(l[k]));
            r[3 + k] = max(r[3 + k], f2o(h[k]));
            r[6 + k] = min(r[6 + k], c);
            r[9 + k] = max(r[9 + k], c);
        }
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) r[k] = is_lo(k) ? wave_umin(r[k]) : wave_umax(r[k]);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) red[wave][k] = r[k];
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        uint32_t v = red[0][k];
        for (int w = 1; w < kBlock / 64; ++w) v = is_lo(k) ? min(v, red[w][k]) : max(v, red[w][k]);
        part[(size_t)blockIdx.x * 12 + k] = v;
    }
}

// one block: thread t reduces the partials of blocks t, t + 256, ...
__global__ __launch_bounds__(kBlock) void k_bounds_final(const uint32_t* part, uint32_t nparts, uint32_t* bounds) {
    __shared__ uint32_t red[4][12];
    uint32_t r[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) r[k] = is_lo(k) ? 0xffffffffu : 0u;
    for (uint32_t b = threadIdx.x; b < nparts; b += kBlock) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const uint32_t v = part[(size_t)b * 12 + k];
            r[k] = is_lo(k) ? min(r[k], v) : max(r[k], v);
        }
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) r[k] = is_lo(k) ? wave_umin(r[k]) : wave_umax(r[k]);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) red[wave][k] = r[k];
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        uint32_t v = red[0][k];
        for (int w = 1; w < kBlock / 64; ++w) v = is_lo(k) ? min(v, red[w][k]) : max(v, red[w][k]);
        bounds[k] = v;
    }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t quantise(float c, float lo, float hi) {
    float e = hi - lo;
    float x = e > 0.0f ? (c - lo) / e : 0.0f;
    float q = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    return (uint32_t)q;
}

__global__ __launch_bounds__(kBlock) void k_morton(uint32_t n, const float4* box_lo, const float4* box_hi,
                                                   const uint32_t* bounds, uint32_t* keys, uint32_t* vals) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float4 lo = box_lo[i], hi = box_hi[i];
    float cx = (lo.x + hi.x) * 0.5f, cy = (lo.y + hi.y) * 0.5f, cz = (lo.z + hi.z) * 0.5f;
    uint32_t qx = quantise(cx, o2f(bounds[6]), o2f(bounds[9]));
    uint32_t qy = quantise(cy, o2f(bounds[7]), o2f(bounds[10]));
    uint32_t qz = quantise(cz, o2f(bounds[8]), o2f(bounds[11]));
    keys[i] = (expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz);
    vals[i] = i;
}

// ---- radix sort: per-tile digit histograms (digit-major layout) ----
__global__ __launch_bounds__(kBlock) void k_hist(const uint32_t* keys, uint32_t n, int shift, uint32_t* hist,
                                                 uint32_t nblocks) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        size_t idx = base + (size_t)it * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of `total` counts in three coalesced kernels: per-tile
// sums (kScanTile entries per block), a scan of the tile sums by one block,
// then each tile rescanned through LDS with its offset added.
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* red) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const uint32_t t = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return t;
}

// exclusive scan across the 256 threads of a block; returns the block total
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t& excl, uint32_t* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    if (lane == 63) red[wave] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        before += w < wave ? red[w] : 0u;
        tot += red[w];
    }
    __syncthreads();
    excl = before + inc - v;
    return tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* h, uint32_t total, uint32_t* part) {
    __shared__ uint32_t red[4];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t v = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        const size_t i = base + (size_t)it * kBlock + threadIdx.x;
        if (i < total) v += h[i];
    }
    v = block_sum(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// one block: exclusive scan of the np tile sums, 256 at a time with a carry
__global__ __launch_bounds__(kBlock) void k_scan_parts(uint32_t* part, uint32_t np) {
    __shared__ uint32_t red[4];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < np; b += kBlock) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < np ? part[i] : 0u;
        uint32_t ex;
        const uint32_t tot = block_exclusive(v, ex, red);
        if (i < np) part[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(uint32_t* h, uint32_t total, const uint32_t* part) {
    // row r of the tile at tile[r * 17]: the pad spreads a thread's 16
    // consecutive entries over distinct banks
    __shared__ uint32_t tile[kBlock * (kScanItems + 1)];
    __shared__ uint32_t red[4];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    const int tid = threadIdx.x;
    auto at = [](uint32_t k) { return (k / kScanItems) * (kScanItems + 1) + k % kScanItems; };
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        const uint32_t k = it * kBlock + tid;
        const size_t i = base + k;
        tile[at(k)] = i < total ? h[i] : 0u;
    }
    __syncthreads();
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) sum += tile[tid * (kScanItems + 1) + j];
    uint32_t run;
    block_exclusive(sum, run, red);
    run += part[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t c = tile[tid * (kScanItems + 1) + j];
        tile[tid * (kScanItems + 1) + j] = run;
        run += c;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
        const uint32_t k = it * kBlock + tid;
        const size_t i = base + k;
        if (i < total) h[i] = tile[at(k)];
    }
}

// stable scatter: ranks within a wave from 8 ballots over the digit bits,
// across the 4 waves of a 256-key chunk through LDS counts
__global__ __launch_bounds__(kBlock) void k_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                    uint32_t* vout, uint32_t n, int shift, const uint32_t* hist,
                                                    uint32_t nblocks) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    run[tid] = hist[(size_t)tid * nblocks + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    size_t base = (size_t)blockIdx.x * kTile;
    for (int it = 0; it < kItems; ++it) {
        size_t idx = base + (size_t)it * kBlock + tid;
        bool act = idx < n;
        uint32_t key = act ? kin[idx] : 0u, val = act ? vin[idx] : 0u;
        uint32_t dg = (key >> shift) & 255u;
#pragma unroll
        for (int w = 0; w < 4; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        uint64_t peers = __ballot(act);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (dg >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rank = __popcll(peers & lt);
        uint32_t tot = __popcll(peers);
        if (act && rank == tot - 1) wcnt[wave][dg] = tot;
        __syncthreads();
        {
            uint32_t r = run[tid];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = r;
                r += c;
            }
            run[tid] = r;
        }
        __syncthreads();
        if (act) {
            uint32_t dst = wcnt[wave][dg] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int delta(const uint32_t* k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = ((uint64_t)k[i] << 32) | (uint64_t)i;
    uint64_t b = ((uint64_t)k[j] << 32) | (uint64_t)j;
    return __clzll(a ^ b);
}

__global__ __launch_bounds__(kBlock) void k_karras(const uint32_t* keys, uint32_t n, uint32_t* gamma,
                                                   uint32_t* flags, uint32_t* int_parent, uint32_t* leaf_parent,
                                                   uint32_t* int_pos, uint32_t* leaf_pos, float4* nodes) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t N = n;
    if (i >= N - 1) return;
    if (i == 0) {
        int_pos[0] = 0;
        int_parent[0] = kNone;
        nodes[2] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));  // node 1: padding
        nodes[3] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kNone));
    }
    const int d = (delta(keys, N, i, i + 1) - delta(keys, N, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, N, i, i - d);
    int64_t lmax = 2;
    while (delta(keys, N, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, N, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(keys, N, i, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) / 2;
        if (delta(keys, N, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t g = i + s * d + (d < 0 ? -1 : 0);
    const bool ll = (i < j ? i : j) == g;
    const bool rl = (i > j ? i : j) == g + 1;
    gamma[i] = (uint32_t)g;
    flags[i] = (ll ? 1u : 0u) | (rl ? 2u : 0u);
    const uint32_t ug = (uint32_t)g, ui = (uint32_t)i;
    if (ll) { leaf_parent[ug] = ui; leaf_pos[ug] = 2 + 2 * ug; }
    else { int_parent[ug] = ui; int_pos[ug] = 2 + 2 * ug; }
    if (rl) { leaf_parent[ug + 1] = ui; leaf_pos[ug + 1] = 3 + 2 * ug; }
    else { int_parent[ug + 1] = ui; int_pos[ug + 1] = 3 + 2 * ug; }
}

// ---- refit, one launch per depth ----
// Every node record is written once: a leaf by k_leaves, an internal node by
// the k_refit_level launch of its depth, after both of its children.

// leaves: the primitive's box, its id, and the position of its parent
__global__ __launch_bounds__(kBlock) void k_leaves(uint32_t n, const uint32_t* order, const float4* box_lo,
                                                   const float4* box_hi, const uint32_t* leaf_pos,
                                                   const uint32_t* leaf_parent, const uint32_t* int_pos,
                                                   float4* nodes) {
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= n) return;
    const uint32_t prim = order[m];
    const float4 blo = box_lo[prim], bhi = box_hi[prim];
    if (n == 1) {
        nodes[0] = make_float4(blo.x, blo.y, blo.z, __uint_as_float(AURORA_NODE_LEAF | prim));
        nodes[1] = make_float4(bhi.x, bhi.y, bhi.z, __uint_as_float(kNone));
        return;
    }
    const uint32_t pos = leaf_pos[m];
    nodes[2 * (size_t)pos] = make_float4(blo.x, blo.y, blo.z, __uint_as_float(AURORA_NODE_LEAF | prim));
    nodes[2 * (size_t)pos + 1] = make_float4(bhi.x, bhi.y, bhi.z, __uint_as_float(int_pos[leaf_parent[m]]));
}

// the flat traversal array of a small scene: the leaf records in Morton order
__global__ __launch_bounds__(kBlock) void k_flat_fill(uint32_t n, const uint32_t* leaf_pos, const float4* nodes,
                                                      float4* flat) {
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= n) return;
    const size_t pos = n == 1 ? 0 : leaf_pos[m];
    flat[2 * (size_t)m] = nodes[2 * pos];
    flat[2 * (size_t)m + 1] = nodes[2 * pos + 1];
}

// the internal nodes of one depth: union of the two child records (written
// by earlier launches), child index, parent position
__global__ __launch_bounds__(kBlock) void k_refit_level(const uint32_t* lvl_node, uint32_t count,
                                                        const uint32_t* gamma, const uint32_t* int_pos,
                                                        const uint32_t* int_parent, float4* nodes) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t i = lvl_node[k];
    const uint32_t c = 2 + 2 * gamma[i];
    const float4 a0 = nodes[2 * (size_t)c], a1 = nodes[2 * (size_t)c + 1];
    const float4 b0 = nodes[2 * (size_t)c + 2], b1 = nodes[2 * (size_t)c + 3];
    const uint32_t gp = i == 0 ? kNone : int_pos[int_parent[i]];
    const uint32_t pos = int_pos[i];
    nodes[2 * (size_t)pos] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), __uint_as_float(c));
    nodes[2 * (size_t)pos + 1] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), __uint_as_float(gp));
}

// ---- depth of every internal node by pointer jumping ----
// J[i] = (ancestor a of i, number of edges from i up to a); kNone marks the
// root's parent.  A round replaces (a, d) by (ancestor of a, d + its
// distance), doubling the span, so after r rounds every node within 2^r
// edges of the root is resolved.  A Karras tree over distinct 64-bit keys
// (30-bit code, 32-bit index) is at most 64 internal levels deep: each
// child's common prefix is longer than its parent's.  Seven rounds cover 128.
constexpr int kJumpRounds = 7;

__global__ __launch_bounds__(kBlock) void k_jump_init(uint32_t ni, const uint32_t* int_parent, uint2* J) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ni) return;
    const uint32_t p = int_parent[i];
    J[i] = make_uint2(p, p == kNone ? 0u : 1u);
}

__global__ __launch_bounds__(kBlock) void k_jump(uint32_t ni, const uint2* A, uint2* B) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ni) return;
    const uint2 a = A[i];
    if (a.x == kNone) {
        B[i] = a;
        return;
    }
    const uint2 q = A[a.x];
    B[i] = make_uint2(q.x, a.y + q.y);
}

// depth keys for the level sort (255 marks a node left unresolved, which the
// host reports), node ids, and the wide-node flags cleared but for the root
// (k_wide_mark sets the others); flag[ni] = 0 so the exclusive scan leaves
// the wide-node count at [ni]
__global__ __launch_bounds__(kBlock) void k_jump_out(uint32_t ni, const uint2* J, uint32_t* key, uint32_t* id,
                                                     uint32_t* wflag) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i > ni) return;
    wflag[i] = i == 0 && ni > 0 ? 1u : 0u;  // wide level of the root
    if (i == ni) return;
    const uint2 a = J[i];
    const uint32_t d = (a.x == kNone && a.y < 255u) ? a.y : 255u;
    key[i] = d;
    id[i] = i;
}

// first sorted index of each depth: the scanned digit-major histogram of the
// one-pass sort holds it at hist[d * nblocks]
__global__ void k_level_starts(const uint32_t* hist, uint32_t nblocks, uint32_t total, uint32_t* starts) {
    const uint32_t d = threadIdx.x;
    if (d < 256) starts[d] = hist[(size_t)d * nblocks];
    if (d == 0) starts[256] = total;
}

// ---- four-wide traversal tree ----
// A wide node is made for the root and for every internal binary node that a
// wide node's slots refer to, numbered in Karras index order.  Its slots are
// a cut below the binary node, in one of two forms:
//  - greedy (scenes of at least kWideGreedyMin primitives): start with the two
//    children; while fewer than four slots, replace the internal slot with
//    the largest box surface area (dx*dy + dy*dz + dz*dx; the first in slot
//    order on ties) in place by its two children.  Opening the largest boxes
//    first suits rays that enter boxes from outside, as incoherent rays among
//    many small primitives do: 18% fewer node records per ray than the
//    pairwise form on 2M primitives;
//  - pairwise (smaller scenes): each internal child replaced by its two
//    children, so the wide nodes are the even-depth binary nodes.  Where rays
//    start inside most boxes (a closed room of a few large quads) the greedy
//    form visits more nodes (Cornell box: 4.3 against 3.4 per ray).
// The set of wide nodes is found top down: k_wide_mark runs once per depth,
// shallowest first; a flagged node's cut flags the internal nodes it refers
// to, all deeper, so each flag is final before its depth runs.
constexpr uint32_t kWideGreedyMin = 4096;

// a slot of a cut: binary node position, internal index, leaf or not
struct WideRef {
    uint32_t pos, j;
    bool leaf;
};

__device__ __forceinline__ float box_area(const float4* nodes, uint32_t pos) {
    const float4 lo = nodes[2 * (size_t)pos], hi = nodes[2 * (size_t)pos + 1];
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// the slots of internal node i (children of internal i: positions 2+2g,
// internal index g, and 3+2g, index g+1; flags bit 0/1: left/right is a
// leaf).  Every array index below is a constant after unrolling, so the cut
// stays in registers.
__device__ __forceinline__ int wide_cut(uint32_t i, const uint32_t* gamma, const uint32_t* flags,
                                        const float4* nodes, bool greedy, WideRef* out) {
    const uint32_t g = gamma[i], fl = flags[i];
    if (!greedy) {
        int k = 0;
#pragma unroll
        for (uint32_t c = 0; c < 2; ++c) {
            const uint32_t j = g + c;
            const bool leaf = (fl & (1u << c)) != 0u;
            // a leaf's j may be one past the internal nodes: read at g instead
            const uint32_t jr = leaf ? g : j;
            const uint32_t gj = gamma[jr], fj = leaf ? 0u : flags[jr];
            // a leaf child stays itself; an internal child gives its two children
            const WideRef first = leaf ? WideRef{2 + 2 * g + c, j, true} : WideRef{2 + 2 * gj, gj, (fj & 1u) != 0u};
            const WideRef second = WideRef{3 + 2 * gj, gj + 1, (fj & 2u) != 0u};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q == k) out[q] = first;
                if (!leaf && q == k + 1) out[q] = second;
            }
            k += leaf ? 1 : 2;
        }
        return k;
    }
    out[0] = WideRef{2 + 2 * g, g, (fl & 1u) != 0u};
    out[1] = WideRef{3 + 2 * g, g + 1, (fl & 2u) != 0u};
    out[2] = out[3] = WideRef{0u, 0u, true};
    int k = 2;
#pragma unroll
    for (int round = 0; round < 2; ++round) {
        int best = -1;
        float ba = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < k && !out[q].leaf) {
                const float a = box_area(nodes, out[q].pos);
                if (best < 0 || a > ba) {
                    best = q;
                    ba = a;
                }
            }
        if (best < 0) break;
        uint32_t j = out[0].j;
#pragma unroll
        for (int q = 1; q < 4; ++q) j = best == q ? out[q].j : j;
        const uint32_t gj = gamma[j], fj = flags[j];
        const WideRef lo = WideRef{2 + 2 * gj, gj, (fj & 1u) != 0u};
        const WideRef hi = WideRef{3 + 2 * gj, gj + 1, (fj & 2u) != 0u};
        // slot best becomes (lo, hi); the slots after it move up one
#pragma unroll
        for (int q = 3; q >= 1; --q) {
            if (q > best + 1) out[q] = out[q - 1];
            if (q == best + 1) out[q] = hi;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q == best) out[q] = lo;
        ++k;
    }
    return k;
}

// the flagged internal nodes of one depth flag the internal nodes of their
// cuts with their own wide level plus one (the root's level is 1)
__global__ __launch_bounds__(kBlock) void k_wide_mark(const uint32_t* lvl_node, uint32_t count,
                                                      const uint32_t* gamma, const uint32_t* flags,
                                                      const float4* nodes, bool greedy, uint32_t* wflag) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t i = lvl_node[k];
    const uint32_t lv = wflag[i];
    if (!lv) return;
    WideRef r[4];
    const int m = wide_cut(i, gamma, flags, nodes, greedy, r);
    for (int q = 0; q < m; ++q)
        if (!r[q].leaf) wflag[r[q].j] = lv + 1u;
}

// the deepest wide level into *maxlv (0 before the launch), and the levels
// turned into the 0/1 flags that the scan numbers.  Grid-stride over at
// most kLevelsGrid blocks, reduced in the block: one atomic per block (one
// per wave on a single address serialised 62 K atomics, 0.7 ms, on 4 M
// primitives)
constexpr uint32_t kLevelsGrid = 1024;
__global__ __launch_bounds__(kBlock) void k_wide_levels(uint32_t n, uint32_t* wflag, uint32_t* maxlv) {
    __shared__ uint32_t wave_max[kBlock / 64];
    uint32_t lv = 0u;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t v = wflag[i];
        wflag[i] = v ? 1u : 0u;
        lv = max(lv, v);
    }
    for (int o = 32; o > 0; o >>= 1) lv = max(lv, (uint32_t)__shfl_xor((int)lv, o));
    if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = lv;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0u;
        for (int w = 0; w < kBlock / 64; ++w) m = max(m, wave_max[w]);
        if (m) atomicMax(maxlv, m);
    }
}

// back to the root's flag alone (a second collapse, pairwise)
__global__ __launch_bounds__(kBlock) void k_wide_reset(uint32_t n, uint32_t* wflag, uint32_t* maxlv) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) wflag[i] = i == 0 ? 1u : 0u;
    if (i == 0) *maxlv = 0u;
}

__device__ __forceinline__ void wide_slot(float4* w, int k, const float4* nodes, uint32_t pos, uint32_t ref) {
    const float4 lo = nodes[2 * (size_t)pos], hi = nodes[2 * (size_t)pos + 1];
    w[2 * k] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(ref));
    w[2 * k + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
}

// one wide node per flagged internal node i, at widx[i] (the exclusive scan
// of the flags): the slots of its cut (a leaf slot refers to its primitive,
// an internal one to that node's wide index); unused slots are empty
__global__ __launch_bounds__(kBlock) void k_wide_fill(uint32_t ni, const uint32_t* gamma, const uint32_t* flags,
                                                      const uint32_t* widx, const float4* nodes, bool greedy,
                                                      float4* wide) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ni || widx[i + 1] == widx[i]) return;
    float4* w = wide + 8 * (size_t)widx[i];
    WideRef r[4];
    const int m = wide_cut(i, gamma, flags, nodes, greedy, r);
    for (int q = 0; q < m; ++q)
        wide_slot(w, q, nodes, r[q].pos, r[q].leaf ? __float_as_uint(nodes[2 * (size_t)r[q].pos].w) : widx[r[q].j]);
    for (int q = m; q < 4; ++q) {
        w[2 * q] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(AURORA_BVH4_EMPTY));
        w[2 * q + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// a single primitive: one wide node holding the leaf
__global__ void k_wide_single(const float4* nodes, float4* wide) {
    const int k = threadIdx.x;
    if (k >= 4) return;
    if (k == 0) {
        wide_slot(wide, 0, nodes, 0, __float_as_uint(nodes[0].w));
    } else {
        wide[2 * k] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(AURORA_BVH4_EMPTY));
        wide[2 * k + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

// histogram entries of a sort of n keys, and the tile sums of its scan after them
inline uint32_t hist_entries(uint32_t n) { return 256u * blocks_for(n, kTile); }

void scan_exclusive(uint32_t* h, uint32_t total, hipStream_t st) {
    uint32_t* part = h + total;
    const uint32_t np = blocks_for(total, kScanTile);
    k_scan_reduce<<<np, kBlock, 0, st>>>(h, total, part);
    k_scan_parts<<<1, kBlock, 0, st>>>(part, np);
    k_scan_apply<<<np, kBlock, 0, st>>>(h, total, part);
}

}  // namespace

size_t radix_hist_bytes(uint32_t n) {
    const uint32_t e = hist_entries(n);
    return ((size_t)e + blocks_for(e, kScanTile)) * 4;
}

Status radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t* hist, uint32_t n,
                        int passes, hipStream_t st) {
    if (n == 0) return Status::ok();
    const uint32_t nb = blocks_for(n, kTile);
    uint32_t* ks[2] = {k0, k1};
    uint32_t* vs[2] = {v0, v1};
    for (int pass = 0; pass < passes; ++pass) {
        const int a = pass & 1, b = a ^ 1;
        k_hist<<<nb, kBlock, 0, st>>>(ks[a], n, 8 * pass, hist, nb);
        scan_exclusive(hist, 256u * nb, st);
        k_scatter<<<nb, kBlock, 0, st>>>(ks[a], vs[a], ks[b], vs[b], n, 8 * pass, hist, nb);
    }
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

// The flat array holds the leaf records ordered by box surface area, largest
// first (stable: Morton order among equal areas): a flat walk meets the big
// occluders first, so its best t falls early and later boxes are culled
// (1% on the Cornell box).  The order changes no hit.
Status flat_build(uint32_t n, const DevBuf& nodes, const BvhScratch& s, DevBuf& flat, hipStream_t st) {
    if (n == 0 || n > kFlatMaxPrims) return Status::ok();
    AUR_TRY(flat.alloc((size_t)n * 32));
    k_flat_fill<<<blocks_for(n, kBlock), kBlock, 0, st>>>(n, s.leaf_pos.as<uint32_t>(), nodes.as<float4>(),
                                                          flat.as<float4>());
    AUR_HIP(hipGetLastError());
    std::vector<float4> rec((size_t)2 * n), sorted((size_t)2 * n);
    AUR_HIP(hipMemcpyAsync(rec.data(), flat.p, (size_t)n * 32, hipMemcpyDeviceToHost, st));
    AUR_HIP(hipStreamSynchronize(st));
    auto area = [&](uint32_t i) {
        const float4 lo = rec[2 * (size_t)i], hi = rec[2 * (size_t)i + 1];
        const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
        return dx * dy + dy * dz + dz * dx;
    };
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return area(a) > area(b); });
    for (uint32_t i = 0; i < n; ++i) {
        sorted[2 * (size_t)i] = rec[2 * (size_t)idx[i]];
        sorted[2 * (size_t)i + 1] = rec[2 * (size_t)idx[i] + 1];
    }
    AUR_HIP(hipMemcpyAsync(flat.p, sorted.data(), (size_t)n * 32, hipMemcpyHostToDevice, st));
    AUR_HIP(hipStreamSynchronize(st));
    return Status::ok();
}

Status lbvh_build(const DevScene& sc_in, uint32_t n, DevBuf& nodes, DevBuf& wide, DevBuf& morton, DevBuf& order,
                  BvhScratch& s, hipStream_t st, aurora_bvh_stats* stats) {
    DevScene sc = sc_in;
    aurora_bvh_stats local{};
    local.n_prims = n;
    if (n == 0) {
        if (stats) *stats = local;
        return Status::ok();
    }
    const uint32_t n_nodes = n == 1 ? 1u : 2u * n;
    const uint32_t nb = blocks_for(n, kTile);
    const uint32_t nbound = std::min<uint32_t>(blocks_for(n, kBlock), kBoundBlocks);
    AUR_TRY(nodes.alloc((size_t)n_nodes * 32));
    AUR_TRY(s.box_lo.alloc((size_t)n * 16));
    AUR_TRY(s.box_hi.alloc((size_t)n * 16));
    AUR_TRY(s.bounds.alloc(64));
    AUR_TRY(s.bpart.alloc((size_t)kBoundBlocks * 12 * 4));
    for (int k = 0; k < 2; ++k) {
        AUR_TRY(s.keys[k].alloc((size_t)n * 4));
        AUR_TRY(s.vals[k].alloc((size_t)n * 4));
    }
    AUR_TRY(s.hist.alloc(radix_hist_bytes(n)));
    const size_t ni = n > 1 ? n - 1 : 1;
    AUR_TRY(s.gamma.alloc(ni * 4));
    AUR_TRY(s.flags.alloc(ni * 4));
    AUR_TRY(s.int_parent.alloc(ni * 4));
    AUR_TRY(s.int_pos.alloc(ni * 4));
    AUR_TRY(s.leaf_parent.alloc((size_t)n * 4));
    AUR_TRY(s.leaf_pos.alloc((size_t)n * 4));
    AUR_TRY(s.jump[0].alloc(ni * 8));
    AUR_TRY(s.jump[1].alloc(ni * 8));
    AUR_TRY(s.lvl_key.alloc(ni * 4));
    AUR_TRY(s.lvl_node.alloc(ni * 4));
    AUR_TRY(s.lvl_start.alloc(257 * 4));
    const uint32_t n_flags = (uint32_t)ni + 1;
    AUR_TRY(s.widx.alloc(((size_t)n_flags + blocks_for(n_flags, kScanTile)) * 4));
    // the wide-node count is known only after the scan: room for its bound
    // (every internal node)
    AUR_TRY(wide.alloc(ni * 128));

    EventPair ev;
    AUR_TRY(ev.create());
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    AUR_HIP(hipEventRecord(e0, st));
    k_prim_bounds<<<nbound, kBlock, 0, st>>>(sc, n, s.box_lo.as<float4>(), s.box_hi.as<float4>(),
                                             s.bpart.as<uint32_t>());
    k_bounds_final<<<1, kBlock, 0, st>>>(s.bpart.as<uint32_t>(), nbound, s.bounds.as<uint32_t>());
    k_morton<<<blocks_for(n, kBlock), kBlock, 0, st>>>(n, s.box_lo.as<float4>(), s.box_hi.as<float4>(),
                                                       s.bounds.as<uint32_t>(), s.keys[0].as<uint32_t>(),
                                                       s.vals[0].as<uint32_t>());
    for (int pass = 0; pass < 4; ++pass) {
        const int a = pass & 1, b = a ^ 1;
        k_hist<<<nb, kBlock, 0, st>>>(s.keys[a].as<uint32_t>(), n, 8 * pass, s.hist.as<uint32_t>(), nb);
        scan_exclusive(s.hist.as<uint32_t>(), 256u * nb, st);
        k_scatter<<<nb, kBlock, 0, st>>>(s.keys[a].as<uint32_t>(), s.vals[a].as<uint32_t>(), s.keys[b].as<uint32_t>(),
                                         s.vals[b].as<uint32_t>(), n, 8 * pass, s.hist.as<uint32_t>(), nb);
    }
    // four passes: sorted data is back in buffer 0
    uint32_t depth = 0, n_wide = 1, wide_depth = 1;
    bool wide_greedy = false;
    if (n > 1) {
        const uint32_t nint = n - 1;
        const uint32_t gi = blocks_for(nint, kBlock);
        k_karras<<<gi, kBlock, 0, st>>>(s.keys[0].as<uint32_t>(), n, s.gamma.as<uint32_t>(), s.flags.as<uint32_t>(),
                                        s.int_parent.as<uint32_t>(), s.leaf_parent.as<uint32_t>(),
                                        s.int_pos.as<uint32_t>(), s.leaf_pos.as<uint32_t>(), nodes.as<float4>());
        uint2* J[2] = {s.jump[0].as<uint2>(), s.jump[1].as<uint2>()};
        k_jump_init<<<gi, kBlock, 0, st>>>(nint, s.int_parent.as<uint32_t>(), J[0]);
        for (int r = 0; r < kJumpRounds; ++r) k_jump<<<gi, kBlock, 0, st>>>(nint, J[r & 1], J[(r + 1) & 1]);
        k_jump_out<<<blocks_for(nint + 1, kBlock), kBlock, 0, st>>>(nint, J[kJumpRounds & 1],
                                                                    s.keys[1].as<uint32_t>(),
                                                                    s.vals[1].as<uint32_t>(), s.widx.as<uint32_t>());
        // one 8-bit pass: the sorted pairs land in (lvl_key, lvl_node)
        AUR_TRY(radix_sort_pairs(s.keys[1].as<uint32_t>(), s.vals[1].as<uint32_t>(), s.lvl_key.as<uint32_t>(),
                                 s.lvl_node.as<uint32_t>(), s.hist.as<uint32_t>(), nint, 1, st));
        k_level_starts<<<1, 256, 0, st>>>(s.hist.as<uint32_t>(), blocks_for(nint, kTile), nint,
                                          s.lvl_start.as<uint32_t>());
        uint32_t starts[257];
        AUR_HIP(hipMemcpyAsync(starts, s.lvl_start.p, sizeof(starts), hipMemcpyDeviceToHost, st));
        AUR_HIP(hipStreamSynchronize(st));
        if (starts[255] != nint) return Status{AURORA_ERR_RUNTIME, "LBVH build: node depths did not resolve"};
        uint32_t maxd = 0;
        for (uint32_t d = 0; d < 255; ++d)
            if (starts[d + 1] > starts[d]) maxd = d;
        depth = maxd + 1;  // a deepest internal node has two leaf children
        k_leaves<<<blocks_for(n, kBlock), kBlock, 0, st>>>(n, s.vals[0].as<uint32_t>(), s.box_lo.as<float4>(),
                                                           s.box_hi.as<float4>(), s.leaf_pos.as<uint32_t>(),
                                                           s.leaf_parent.as<uint32_t>(), s.int_pos.as<uint32_t>(),
                                                           nodes.as<float4>());
        for (int d = (int)maxd; d >= 0; --d) {
            const uint32_t cnt = starts[d + 1] - starts[d];
            if (cnt == 0) continue;
            k_refit_level<<<blocks_for(cnt, kBlock), kBlock, 0, st>>>(
                s.lvl_node.as<uint32_t>() + starts[d], cnt, s.gamma.as<uint32_t>(), s.int_pos.as<uint32_t>(),
                s.int_parent.as<uint32_t>(), nodes.as<float4>());
        }
        // four-wide traversal tree: levels top down, then the flags' scan and
        // the fill.  The walk pushes at most three nodes per wide level above
        // the deepest, so the greedy form is kept only while that fits the
        // traversal stack; the pairwise form has ceil(depth / 2) levels, which
        // fits every tree of depth <= 64
        if (depth > (uint32_t)kTravMaxStack)
            return Status{AURORA_ERR_RUNTIME, "LBVH build: tree deeper than the traversal stack"};
        bool greedy = n >= kWideGreedyMin;
        uint32_t* maxlv = s.bounds.as<uint32_t>() + 12;
        for (int attempt = 0; attempt < 2; ++attempt) {
            k_wide_reset<<<blocks_for(nint + 1, kBlock), kBlock, 0, st>>>(nint + 1, s.widx.as<uint32_t>(), maxlv);
            for (uint32_t d = 0; d <= maxd; ++d) {
                const uint32_t cnt = starts[d + 1] - starts[d];
                if (cnt == 0) continue;
                k_wide_mark<<<blocks_for(cnt, kBlock), kBlock, 0, st>>>(s.lvl_node.as<uint32_t>() + starts[d], cnt,
                                                                         s.gamma.as<uint32_t>(),
                                                                         s.flags.as<uint32_t>(), nodes.as<float4>(),
                                                                         greedy, s.widx.as<uint32_t>());
            }
            k_wide_levels<<<std::min<uint32_t>(blocks_for(nint + 1, kBlock), kLevelsGrid), kBlock, 0, st>>>(
                nint + 1, s.widx.as<uint32_t>(), maxlv);
            AUR_HIP(hipMemcpyAsync(&wide_depth, maxlv, 4, hipMemcpyDeviceToHost, st));
            AUR_HIP(hipStreamSynchronize(st));
            if (3u * (wide_depth - 1u) <= (uint32_t)kTravMaxStack) break;
            if (!greedy) return Status{AURORA_ERR_RUNTIME, "LBVH build: four-wide tree deeper than the traversal stack"};
            greedy = false;
        }
        wide_greedy = greedy;
        scan_exclusive(s.widx.as<uint32_t>(), nint + 1, st);
        k_wide_fill<<<gi, kBlock, 0, st>>>(nint, s.gamma.as<uint32_t>(), s.flags.as<uint32_t>(),
                                           s.widx.as<uint32_t>(), nodes.as<float4>(), greedy, wide.as<float4>());
    } else {
        k_leaves<<<1, kBlock, 0, st>>>(n, s.vals[0].as<uint32_t>(), s.box_lo.as<float4>(), s.box_hi.as<float4>(),
                                       s.leaf_pos.as<uint32_t>(), s.leaf_parent.as<uint32_t>(),
                                       s.int_pos.as<uint32_t>(), nodes.as<float4>());
        k_wide_single<<<1, 64, 0, st>>>(nodes.as<float4>(), wide.as<float4>());
    }
    AUR_HIP(hipGetLastError());
    AUR_HIP(hipEventRecord(e1, st));
    std::swap(morton, s.keys[0]);
    std::swap(order, s.vals[0]);

    uint32_t b[12];
    AUR_HIP(hipMemcpyAsync(b, s.bounds.p, 48, hipMemcpyDeviceToHost, st));
    if (n > 1) AUR_HIP(hipMemcpyAsync(&n_wide, s.widx.as<uint32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, st));
    AUR_HIP(hipStreamSynchronize(st));
    float ms = 0.0f;
    AUR_HIP(hipEventElapsedTime(&ms, e0, e1));
    local.n_nodes = n_nodes;
    local.n_wide = n_wide;
    local.depth = depth;
    local.wide_depth = wide_depth;
    local.wide_form = wide_greedy ? AURORA_BVH4_GREEDY : AURORA_BVH4_PAIRWISE;
    for (int k = 0; k < 3; ++k) {
        local.scene_lo[k] = o2f(b[k]);
        local.scene_hi[k] = o2f(b[3 + k]);
        local.centroid_lo[k] = o2f(b[6 + k]);
        local.centroid_hi[k] = o2f(b[9 + k]);
    }
    local.build_ms = ms;
    if (stats) *stats = local;
    return Status::ok();
}

}  // namespace aur
