// LDS ring traversal stacks of the persistent traversal kernels (the
// wavefront's k_wf_trace) and of the megakernel with path regeneration
// (k_render_plus), for one-wave (64-thread) blocks.
#pragma once

#include "dev_bvh.h"

namespace shiro_hip {

// Traversal stack of the persistent kernels.  The newest kRing entries of
// each lane live in an LDS ring (slot k of lane i at [k * 64 + i]); older
// entries move to a private overflow array, touched only while a lane holds
// more than kRing entries.  Entries [0, base) are in the overflow array,
// [base, sp) in the ring.  Any-hit traversal keeps node words only.
constexpr int kRing = 8;
static __shared__ uint32_t gRingNode[kRing * 64];
static __shared__ float gRingT[kRing * 64];

DEV uint32_t ring_slot(int i, uint32_t lane) { return static_cast<uint32_t>(i & (kRing - 1)) * 64u + lane; }

template <bool KEEP_T>
DEV void ring_push(uint32_t n, float t, uint32_t lane, uint32_t* ovNode, float* ovT, int& sp, int& base) {
    if (sp - base == kRing) {
        const uint32_t old = ring_slot(base, lane);
        ovNode[base] = gRingNode[old];
        if (KEEP_T) ovT[base] = gRingT[old];
        ++base;
    }
    const uint32_t at = ring_slot(sp, lane);
    gRingNode[at] = n;
    if (KEEP_T) gRingT[at] = t;
    ++sp;
}

// push_sorted (dev_bvh.h) onto the ring stack, for W-wide nodes.
template <bool KEEP_T, int W>
DEV uint32_t ring_push_sorted(uint32_t mask, const float* tn, const uint32_t* w, uint32_t lane, uint32_t* ovNode,
                              float* ovT, int& sp, int& base) {
    const int nh = __popc(mask);
    if constexpr (W == 4) {
        float k[4] = {(mask & 1u) ? tn[0] : kInf, (mask & 2u) ? tn[1] : kInf, (mask & 4u) ? tn[2] : kInf,
                      (mask & 8u) ? tn[3] : kInf};
        uint32_t v[4] = {w[0], w[1], w[2], w[3]};
        cswap(k[0], v[0], k[1], v[1]);
        cswap(k[2], v[2], k[3], v[3]);
        cswap(k[0], v[0], k[2], v[2]);
        cswap(k[1], v[1], k[3], v[3]);
        cswap(k[1], v[1], k[2], v[2]);
        // Pushes v[nh-1] .. v[1] (farthest first) without a loop: entry i
        // goes to stack position sp + nh - 1 - i.  Room is made first by
        // moving the oldest ring entries to the overflow array (rare).
        const int m = nh - 1;
        while (sp + m - base > kRing) {
            const uint32_t old = ring_slot(base, lane);
            ovNode[base] = gRingNode[old];
            if (KEEP_T) ovT[base] = gRingT[old];
            ++base;
        }
#pragma unroll
        for (int i = 1; i < 4; ++i) {
            if (i <= m) {
                const uint32_t at = ring_slot(sp + m - i, lane);
                gRingNode[at] = v[i];
                if (KEEP_T) gRingT[at] = k[i];
            }
        }
        sp += m;
        return v[0];
    } else {
        float k[8];
        uint32_t v[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            k[c] = (mask >> c) & 1u ? tn[c] : kInf;
            v[c] = w[c];
        }
        sort8(k, v);
#pragma unroll
        for (int i = 7; i >= 1; --i)
            if (i < nh) ring_push<KEEP_T>(v[i], k[i], lane, ovNode, ovT, sp, base);
        return v[0];
    }
}

// push_in_order (dev_bvh.h) onto the ring stack: the hit children in node
// order, the first returned, the others pushed last first, written to
// their final ring slots after room is made (any-hit traversal keeps node
// words only).  `mask` is non-zero.
template <int W>
DEV uint32_t ring_push_in_order(uint32_t mask, const uint32_t* w, uint32_t lane, uint32_t* ovNode, int& sp,
                                int& base) {
    const int m = __popc(mask) - 1;
    while (sp + m - base > kRing) {
        const uint32_t old = ring_slot(base, lane);
        ovNode[base] = gRingNode[old];
        ++base;
    }
    uint32_t cur = kEmpty;
    int i = sp;
#pragma unroll
    for (int k = W - 1; k >= 0; --k) {
        if ((mask >> k) & 1u) {
            if (cur != kEmpty) gRingNode[ring_slot(i++, lane)] = cur;
            cur = w[k];
        }
    }
    sp += m;
    return cur;
}

// pop_entry (dev_bvh.h) from the ring stack.  Entries are always read from
// the ring: when it runs empty while older entries wait in the overflow
// array, up to kRing of them move back first (rare).
template <bool ANY>
DEV uint32_t ring_pop(float tcull, uint32_t lane, const uint32_t* ovNode, const float* ovT, int& sp, int& base) {
    while (sp > 0) {
        if (sp == base) {
            const int k = base < kRing ? base : kRing;
            for (int i = base - k; i < base; ++i) {
                const uint32_t at = ring_slot(i, lane);
                gRingNode[at] = ovNode[i];
                if (!ANY) gRingT[at] = ovT[i];
            }
            base -= k;
        }
        --sp;
        const uint32_t at = ring_slot(sp, lane);
        const uint32_t n = gRingNode[at];
        if (ANY) return n;
        if (gRingT[at] <= tcull) return n;
    }
    return kEmpty;
}

// intersect_closest (dev_bvh.h: same while-while loop, node test, child
// order, tie rule and triangle test) on the ring stack: the newest kRing
// entries of the lane's stack in LDS, older ones in `ovNode`/`ovT`
// (kStackSize entries, private).
DEV HitRec intersect_closest_ring(const DevScene& S, const Ray& ray, uint32_t lane, uint32_t* ovNode, float* ovT) {
    HitRec best;
    best.t = ray.tmax;
    best.b1 = best.b2 = 0.0f;
    best.slot = kNoHit;
    best.prim = kNoHit;
    if (S.nodeCount == 0) return best;
    const RayPre r = precompute(ray);
    const TravCount<false> cnt{nullptr, nullptr};
    int sp = 0, base = 0;
    uint32_t cur = 0u;
    for (;;) {
        while (cur != kEmpty && !(cur & kLeafBit)) {
            float tn[4];
            uint32_t w[4];
            const uint32_t mask = node_boxes(S.nodes + static_cast<size_t>(cur) * 4u, r, best.t, tn, w);
            cur = mask ? ring_push_sorted<true, 4>(mask, tn, w, lane, ovNode, ovT, sp, base)
                       : ring_pop<false>(best.t, lane, ovNode, ovT, sp, base);
        }
        if (cur == kEmpty) break;
        leaf_closest(S, r, cur & ~kLeafBit, best, cnt);
        cur = ring_pop<false>(best.t, lane, ovNode, ovT, sp, base);
        if (cur == kEmpty) break;
    }
    return best;
}

// intersect_any (dev_bvh.h) on the ring stack.
DEV bool intersect_any_ring(const DevScene& S, const Ray& ray, uint32_t lane, uint32_t* ovNode) {
    if (S.nodeCount == 0) return false;
    const RayPre r = precompute(ray);
    int sp = 0, base = 0;
    uint32_t cur = 0u;
    for (;;) {
        while (cur != kEmpty && !(cur & kLeafBit)) {
            float tn[4];
            uint32_t w[4];
            const uint32_t mask = node_boxes(S.nodes + static_cast<size_t>(cur) * 4u, r, ray.tmax, tn, w);
            cur = mask ? ring_push_in_order<4>(mask, w, lane, ovNode, sp, base)
                       : ring_pop<true>(0.0f, lane, ovNode, nullptr, sp, base);
        }
        if (cur == kEmpty) break;
        uint32_t slot = cur & ~kLeafBit;
        for (;;) {
            const float4* tp = S.tris + static_cast<size_t>(slot) * 3u;
            const float4 a = tp[0], b = tp[1], c = tp[2];
            float t, b1, b2;
            if (tri_test(r, xyz(a), xyz(b), xyz(c), ray.tmax, &t, &b1, &b2)) return true;
            if (__float_as_uint(b.w) & kLeafEndFlag) break;
            ++slot;
        }
        cur = ring_pop<true>(0.0f, lane, ovNode, nullptr, sp, base);
    }
    return false;
}

}  // namespace shiro_hip
