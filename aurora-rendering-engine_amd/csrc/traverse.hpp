// traverse.hpp — single-ray LBVH closest hit, for kernels that trace a ray
// to completion in one thread (the fused tail of the wavefront loop).  Same
// algorithm and operation order as k_extend's step (trace.hip): root test,
// then per step the aligned child pair, both slabs against the current best
// t, leaf primitives tested at once, descend to the nearer child and push
// the farther; ties in t go to the lower primitive id.
#pragma once

#include "dev_common.hpp"

namespace aur {

constexpr int kTravLds = 16;       // stack levels kept in LDS
constexpr int kTravMaxStack = 64;  // Karras trees are shallower than this

// stk: this thread's LDS column (level k at stk[k * col_stride]); spill: its
// global column for levels >= kTravLds (level k at spill[(k-kTravLds)*sstride]).
__device__ __forceinline__ void traverse_closest(const DevScene& sc, v3 o, v3 d, float tmin, float tmax,
                                                 uint32_t* stk, int col_stride, uint32_t* spill, size_t sstride,
                                                 float& t_out, int& prim_out) {
    // scene fields in registers: the spill stores below could otherwise
    // force a reload of every field read through `sc` after them
    const float4* const N = sc.nodes;
    const float4* const S = sc.spheres;
    const float4* const Qd = sc.quads;
    const uint32_t n_spheres = sc.n_spheres, n_nodes = sc.n_nodes;
    const uint32_t n_prims = sc.n_spheres + sc.n_quads;
    const v3 inv = V(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float best_t = tmax;
    uint32_t best_id = 0x7fffffffu;
    auto test_prim = [&](uint32_t id) {
        if (id >= n_prims) return;
        float t;
        bool h;
        if (id < n_spheres) {
            h = hit_sphere(ldg4(S, id), o, d, tmin, tmax, t);
        } else {
            const size_t q = 3 * (size_t)(id - n_spheres);
            float al, be;
            h = hit_quad(ldg4(Qd, q), ldg4(Qd, q + 1), ldg4(Qd, q + 2), o, d, tmin, tmax, t, al, be);
        }
        if (h && (t < best_t || (t == best_t && id < best_id))) {
            best_t = t;
            best_id = id;
        }
    };
    if (n_nodes > 0) {
        const float4 r0 = ldg4(N, 0), r1 = ldg4(N, 1);
        float tr;
        const uint32_t rc = __float_as_uint(r0.w);
        if (slab(r0, r1, o, inv, tmin, best_t, tr)) {
            if (rc & AURORA_NODE_LEAF) {
                test_prim(rc & ~AURORA_NODE_LEAF);
            } else {
                uint32_t cur = rc;
                int sp = 0;
                for (;;) {
                    if ((cur & 1u) || cur + 1 >= n_nodes) break;
                    const float4 a0 = ldg4(N, 2 * (size_t)cur), a1 = ldg4(N, 2 * (size_t)cur + 1);
                    const float4 b0 = ldg4(N, 2 * (size_t)cur + 2), b1 = ldg4(N, 2 * (size_t)cur + 3);
                    float ta, tb;
                    bool ha = slab(a0, a1, o, inv, tmin, best_t, ta);
                    bool hb = slab(b0, b1, o, inv, tmin, best_t, tb);
                    const uint32_t ca = __float_as_uint(a0.w), cb = __float_as_uint(b0.w);
                    if (ha && (ca & AURORA_NODE_LEAF)) {
                        test_prim(ca & ~AURORA_NODE_LEAF);
                        ha = false;
                    }
                    if (hb && (cb & AURORA_NODE_LEAF)) {
                        test_prim(cb & ~AURORA_NODE_LEAF);
                        hb = false;
                    }
                    if (ha && hb) {
                        const bool bnear = tb < ta;
                        const uint32_t far_ = bnear ? ca : cb;
                        cur = bnear ? cb : ca;
                        if (sp < kTravLds) stk[sp * col_stride] = far_;
                        else if (sp < kTravMaxStack) spill[(size_t)(sp - kTravLds) * sstride] = far_;
                        if (sp < kTravMaxStack) ++sp;
                    } else if (ha) {
                        cur = ca;
                    } else if (hb) {
                        cur = cb;
                    } else if (sp == 0) {
                        break;
                    } else {
                        --sp;
                        if (sp < kTravLds) cur = stk[sp * col_stride];
                        else cur = ldg1(spill, (size_t)(sp - kTravLds) * sstride);
                    }
                }
            }
        }
    }
    const bool miss = best_id == 0x7fffffffu;
    t_out = miss ? tmax : best_t;
    prim_out = miss ? AURORA_MISS : (int)best_id;
}

}  // namespace aur
