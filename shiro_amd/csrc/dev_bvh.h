// BVH traversal on gfx950 with a watertight ray/triangle test.
//
// Replaces Embree's rtcIntersect1 / rtcOccluded1 (reference
// CpuPathTracer.cpp:601-681) and the brute-force fallback (:386-475).
//  - Nodes are 64 B: an fp32 origin, per-axis power-of-two scales, child
//    words and the child boxes quantised to 8 bits per plane (layout.h).  A
//    thread loads a node with four 16-byte loads.
//  - Closest hit: front-to-back ordered traversal; the hit children of a
//    node are sorted by entry distance with a five-comparator network, the
//    nearest is visited next and the rest are pushed (farthest first) with
//    their entry distances, so popped entries beyond the current hit are
//    skipped.  Exact ties in distance resolve to the lower global triangle
//    index, which is the first-in-scene-order rule of the reference's scan.
//  - Any hit: hit children in node order (no sort), returns at the first
//    accepted triangle.
//  - The triangle test is the watertight test of Woop, Benthin and Wald
//    (JCGT 2013): shear to the ray's dominant axis, edge functions U, V, W
//    (recomputed in fp64 when one is exactly zero), so rays through shared
//    edges and vertices of a closed mesh cannot slip through.
//  - The stack holds kStackSize entries per thread; the uploader rejects
//    trees deeper than kMaxBvhDepth so that it cannot overflow.
//  - The persistent traversal kernels can walk an 8-wide tree over the
//    same leaves instead (node_boxes8, sort8): fewer, wider node steps.
#pragma once

#include "dev_math.h"
#include "layout.h"

namespace shiro_hip {

struct Ray {
    V3 o, d;
    float tmin, tmax;
};

// Per-ray constants of the watertight test and the slab test.
struct RayPre {
    V3 o;
    V3 inv;            // 1/d with zero components replaced by tiny values
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float tmin;
};

DEV float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

DEV RayPre precompute(const Ray& r) {
    RayPre p;
    p.o = r.o;
    const float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dz = comp(r.d, kz);
    if (dz < 0.0f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    p.kx = kx;
    p.ky = ky;
    p.kz = kz;
    p.Sx = comp(r.d, kx) / dz;
    p.Sy = comp(r.d, ky) / dz;
    p.Sz = 1.0f / dz;
    const float tiny = 1.0e-30f;
    p.inv = V3{1.0f / (ax > tiny ? r.d.x : copysignf(tiny, r.d.x)), 1.0f / (ay > tiny ? r.d.y : copysignf(tiny, r.d.y)),
               1.0f / (az > tiny ? r.d.z : copysignf(tiny, r.d.z))};
    p.tmin = r.tmin;
    return p;
}

// Watertight test; on success t in [tmin, tmax] and barycentrics b1, b2 of
// vertices 1 and 2.
DEV bool tri_test(const RayPre& r, V3 v0, V3 v1, V3 v2, float tmax, float* tOut, float* b1Out, float* b2Out) {
    const V3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.Sx * Akz;
    const float Ay = comp(A, r.ky) - r.Sy * Akz;
    const float Bx = comp(B, r.kx) - r.Sx * Bkz;
    const float By = comp(B, r.ky) - r.Sy * Bkz;
    const float Cx = comp(C, r.kx) - r.Sx * Ckz;
    const float Cy = comp(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = static_cast<float>(static_cast<double>(Cx) * By - static_cast<double>(Cy) * Bx);
        V = static_cast<float>(static_cast<double>(Ax) * Cy - static_cast<double>(Ay) * Cx);
        W = static_cast<float>(static_cast<double>(Bx) * Ay - static_cast<double>(By) * Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
    const float inv = 1.0f / det;
    const float t = T * inv;
    if (!(t >= r.tmin && t <= tmax)) return false;
    *tOut = t;
    *b1Out = V * inv;
    *b2Out = W * inv;
    return true;
}

struct HitRec {
    float t;
    float b1, b2;
    uint32_t slot;   // leaf-order triangle slot, kNoHit on a miss
    uint32_t prim;   // global triangle index
};

template <bool COUNT>
struct TravCount {
    unsigned long long* nodes;
    unsigned long long* tris;
    DEV void node() const {
        if (COUNT) *nodes += 1;
    }
    DEV void tri() const {
        if (COUNT) *tris += 1;
    }
    DEV void node(uint32_t) const { node(); }
    DEV void tri(uint32_t) const { tri(); }
};

DEV float byte_as_float(uint32_t w, int k) { return static_cast<float>((w >> (8 * k)) & 0xffu); }

// Slab test of the four children of a quantised node (layout.h); returns a
// 4-bit hit mask, the entry distances and the child words.  Per axis the
// plane distance is q * (2^e / d) + (origin - o) / d: one byte-to-float
// conversion and one FMA per plane.  Near and far planes are chosen by the
// sign of the ray direction; empty children are masked out by the node's
// valid bits.  The boxes are widened on the host by more than the rounding
// of these expressions, so the test is conservative and the triangle test
// alone decides hits.
DEV uint32_t node_boxes(const float4* __restrict__ nd, const RayPre& r, float tmax, float tn[4], uint32_t w[4]) {
    const float4 h = nd[0];
    const float4 cw = nd[1];
    const float4 qa = nd[2];
    const float4 qb = nd[3];
    const uint32_t ex = __float_as_uint(h.w);
    const float ax = ldexpf(r.inv.x, static_cast<int>(ex & 0xffu) - 128);
    const float ay = ldexpf(r.inv.y, static_cast<int>((ex >> 8) & 0xffu) - 128);
    const float az = ldexpf(r.inv.z, static_cast<int>((ex >> 16) & 0xffu) - 128);
    const float bx = (h.x - r.o.x) * r.inv.x;
    const float by = (h.y - r.o.y) * r.inv.y;
    const float bz = (h.z - r.o.z) * r.inv.z;
    const uint32_t lx = __float_as_uint(qa.x), hx = __float_as_uint(qa.y);
    const uint32_t ly = __float_as_uint(qa.z), hy = __float_as_uint(qa.w);
    const uint32_t lz = __float_as_uint(qb.x), hz = __float_as_uint(qb.y);
    const uint32_t nx = r.inv.x >= 0.0f ? lx : hx, fx = r.inv.x >= 0.0f ? hx : lx;
    const uint32_t ny = r.inv.y >= 0.0f ? ly : hy, fy = r.inv.y >= 0.0f ? hy : ly;
    const uint32_t nz = r.inv.z >= 0.0f ? lz : hz, fz = r.inv.z >= 0.0f ? hz : lz;
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float t0 = fmaxf(fmaxf(__builtin_fmaf(byte_as_float(nx, c), ax, bx),
                                     __builtin_fmaf(byte_as_float(ny, c), ay, by)),
                               fmaxf(__builtin_fmaf(byte_as_float(nz, c), az, bz), r.tmin));
        const float t1 = fminf(fminf(__builtin_fmaf(byte_as_float(fx, c), ax, bx),
                                     __builtin_fmaf(byte_as_float(fy, c), ay, by)),
                               fminf(__builtin_fmaf(byte_as_float(fz, c), az, bz), tmax));
        tn[c] = t0;
        mask |= (t0 <= t1 ? 1u : 0u) << c;
    }
    w[0] = __float_as_uint(cw.x);
    w[1] = __float_as_uint(cw.y);
    w[2] = __float_as_uint(cw.z);
    w[3] = __float_as_uint(cw.w);
    // The child words are used only when a child is hit; without this the
    // compiler defers their load past the box test, which costs a second
    // memory round trip per node.  The empty asm makes them operands here so
    // the load is issued together with the boxes'.
    asm volatile("" : "+v"(w[0]), "+v"(w[1]), "+v"(w[2]), "+v"(w[3]));
    return mask & (ex >> 24);
}

DEV void cswap(float& ka, uint32_t& va, float& kb, uint32_t& vb) {
    if (kb < ka) {
        const float tk = ka;
        ka = kb;
        kb = tk;
        const uint32_t tv = va;
        va = vb;
        vb = tv;
    }
}

// Orders the hit children of a node by entry distance (five-comparator
// network), pushes all but the nearest (farthest first) and returns the
// nearest.  `mask` is non-zero.
DEV uint32_t push_sorted(uint32_t mask, const float tn[4], const uint32_t w[4], uint32_t* stackNode, float* stackT,
                         int& sp) {
    float k[4] = {(mask & 1u) ? tn[0] : kInf, (mask & 2u) ? tn[1] : kInf, (mask & 4u) ? tn[2] : kInf,
                  (mask & 8u) ? tn[3] : kInf};
    uint32_t v[4] = {w[0], w[1], w[2], w[3]};
    cswap(k[0], v[0], k[1], v[1]);
    cswap(k[2], v[2], k[3], v[3]);
    cswap(k[0], v[0], k[2], v[2]);
    cswap(k[1], v[1], k[3], v[3]);
    cswap(k[1], v[1], k[2], v[2]);
    const int nh = __popc(mask);
    for (int i = nh - 1; i >= 1; --i) {
        stackNode[sp] = v[i];
        stackT[sp] = k[i];
        ++sp;
    }
    return v[0];
}

// Any-hit child order: the hit children in node order, the first visited
// next and the others pushed (the last first), with no sort: an occlusion
// test needs any hit, not the nearest, and the sort costs about a fifth of
// a node step's instructions.  `mask` is non-zero.
template <int W>
DEV uint32_t push_in_order(uint32_t mask, const uint32_t* w, uint32_t* stackNode, int& sp) {
    uint32_t cur = kEmpty;
#pragma unroll
    for (int k = W - 1; k >= 0; --k) {
        if ((mask >> k) & 1u) {
            if (cur != kEmpty) stackNode[sp++] = cur;
            cur = w[k];
        }
    }
    return cur;
}

// Slab test of the eight children of an 8-wide node (layout.h, nodes8):
// per child the same arithmetic as node_boxes, so both widths accept the
// same (conservative) boxes.
DEV uint32_t node_boxes8(const float4* __restrict__ nd, const RayPre& r, float tmax, float tn[8], uint32_t w[8]) {
    const float4 h = nd[0];
    const float4 c0 = nd[1], c1 = nd[2];
    const float4 qx = nd[3], qy = nd[4], qz = nd[5];
    const uint32_t ex = __float_as_uint(h.w);
    const float ax = ldexpf(r.inv.x, static_cast<int>(ex & 0xffu) - 128);
    const float ay = ldexpf(r.inv.y, static_cast<int>((ex >> 8) & 0xffu) - 128);
    const float az = ldexpf(r.inv.z, static_cast<int>((ex >> 16) & 0xffu) - 128);
    const float bx = (h.x - r.o.x) * r.inv.x;
    const float by = (h.y - r.o.y) * r.inv.y;
    const float bz = (h.z - r.o.z) * r.inv.z;
    // Plane words: [lo 0-3, lo 4-7, hi 0-3, hi 4-7] per axis; near = lo
    // for a non-negative direction component.
    const bool px = r.inv.x >= 0.0f, py = r.inv.y >= 0.0f, pz = r.inv.z >= 0.0f;
    const uint32_t nx[2] = {__float_as_uint(px ? qx.x : qx.z), __float_as_uint(px ? qx.y : qx.w)};
    const uint32_t fx[2] = {__float_as_uint(px ? qx.z : qx.x), __float_as_uint(px ? qx.w : qx.y)};
    const uint32_t ny[2] = {__float_as_uint(py ? qy.x : qy.z), __float_as_uint(py ? qy.y : qy.w)};
    const uint32_t fy[2] = {__float_as_uint(py ? qy.z : qy.x), __float_as_uint(py ? qy.w : qy.y)};
    const uint32_t nz[2] = {__float_as_uint(pz ? qz.x : qz.z), __float_as_uint(pz ? qz.y : qz.w)};
    const uint32_t fz[2] = {__float_as_uint(pz ? qz.z : qz.x), __float_as_uint(pz ? qz.w : qz.y)};
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int hw = c >> 2, b = c & 3;
        const float t0 = fmaxf(fmaxf(__builtin_fmaf(byte_as_float(nx[hw], b), ax, bx),
                                     __builtin_fmaf(byte_as_float(ny[hw], b), ay, by)),
                               fmaxf(__builtin_fmaf(byte_as_float(nz[hw], b), az, bz), r.tmin));
        const float t1 = fminf(fminf(__builtin_fmaf(byte_as_float(fx[hw], b), ax, bx),
                                     __builtin_fmaf(byte_as_float(fy[hw], b), ay, by)),
                               fminf(__builtin_fmaf(byte_as_float(fz[hw], b), az, bz), tmax));
        tn[c] = t0;
        mask |= (t0 <= t1 ? 1u : 0u) << c;
    }
    w[0] = __float_as_uint(c0.x);
    w[1] = __float_as_uint(c0.y);
    w[2] = __float_as_uint(c0.z);
    w[3] = __float_as_uint(c0.w);
    w[4] = __float_as_uint(c1.x);
    w[5] = __float_as_uint(c1.y);
    w[6] = __float_as_uint(c1.z);
    w[7] = __float_as_uint(c1.w);
    // Issued with the boxes' loads, as in node_boxes.
    asm volatile("" : "+v"(w[0]), "+v"(w[1]), "+v"(w[2]), "+v"(w[3]));
    asm volatile("" : "+v"(w[4]), "+v"(w[5]), "+v"(w[6]), "+v"(w[7]));
    return mask & (ex >> 24);
}

// cswap on integer keys: the bit patterns of non-negative floats (entry
// distances, FLT_MAX, +inf) order as the floats do, and on wave-uniform
// operands the compiler keeps the comparison on the scalar unit.
DEV void cswap(uint32_t& ka, uint32_t& va, uint32_t& kb, uint32_t& vb) {
    if (kb < ka) {
        const uint32_t tk = ka;
        ka = kb;
        kb = tk;
        const uint32_t tv = va;
        va = vb;
        vb = tv;
    }
}

// Sorts eight (key, value) pairs by key: the 19-comparator network of
// depth 6 for eight inputs (Knuth, TAOCP vol. 3, 5.3.4).
template <typename K>
DEV void sort8(K k[8], uint32_t v[8]) {
    cswap(k[0], v[0], k[2], v[2]);
    cswap(k[1], v[1], k[3], v[3]);
    cswap(k[4], v[4], k[6], v[6]);
    cswap(k[5], v[5], k[7], v[7]);
    cswap(k[0], v[0], k[4], v[4]);
    cswap(k[1], v[1], k[5], v[5]);
    cswap(k[2], v[2], k[6], v[6]);
    cswap(k[3], v[3], k[7], v[7]);
    cswap(k[0], v[0], k[1], v[1]);
    cswap(k[2], v[2], k[3], v[3]);
    cswap(k[4], v[4], k[5], v[5]);
    cswap(k[6], v[6], k[7], v[7]);
    cswap(k[2], v[2], k[4], v[4]);
    cswap(k[3], v[3], k[5], v[5]);
    cswap(k[1], v[1], k[4], v[4]);
    cswap(k[3], v[3], k[6], v[6]);
    cswap(k[1], v[1], k[2], v[2]);
    cswap(k[3], v[3], k[4], v[4]);
    cswap(k[5], v[5], k[6], v[6]);
}

// Node access by tree width: W = 4 walks `nodes` (64 B nodes), W = 8
// walks `nodes8` (128 B nodes) over the same leaves.
template <int W>
DEV const float4* node_at(const DevScene& S, uint32_t id) {
    return W == 8 ? S.nodes8 + static_cast<size_t>(id) * 8u : S.nodes + static_cast<size_t>(id) * 4u;
}

template <int W>
DEV uint32_t node_test(const float4* __restrict__ nd, const RayPre& r, float tmax, float* tn, uint32_t* w) {
    if constexpr (W == 8) {
        return node_boxes8(nd, r, tmax, tn, w);
    } else {
        return node_boxes(nd, r, tmax, tn, w);
    }
}

// Pops the next stack entry; a closest-hit traversal drops entries whose
// entry distance is beyond the current hit.  kEmpty when none is left.
template <bool ANY>
DEV uint32_t pop_entry(const uint32_t* stackNode, const float* stackT, int& sp, float tcull) {
    while (sp > 0) {
        --sp;
        if (ANY || stackT[sp] <= tcull) return stackNode[sp];
    }
    return kEmpty;
}

// Tests the triangles of the leaf starting at `slot` (the last one carries
// the end-of-leaf flag in tris[3*slot+1].w) against the current best hit.
// CNT: a fetch counter (TravCount, or the wavefront kernels' WaveCount).
template <class CNT>
DEV void leaf_closest(const DevScene& S, const RayPre& r, uint32_t slot, HitRec& best, CNT& cnt) {
    for (;;) {
        const float4* tp = S.tris + static_cast<size_t>(slot) * 3u;
        const float4 a = tp[0], b = tp[1], c = tp[2];
        cnt.tri(slot);
        float t, b1, b2;
        if (tri_test(r, xyz(a), xyz(b), xyz(c), best.t, &t, &b1, &b2)) {
            const uint32_t prim = __float_as_uint(a.w);
            // t <= best.t here; an exact tie goes to the lower index.
            if (t < best.t || prim < best.prim) {
                best.t = t;
                best.b1 = b1;
                best.b2 = b2;
                best.slot = slot;
                best.prim = prim;
            }
        }
        if (__float_as_uint(b.w) & kLeafEndFlag) break;
        ++slot;
    }
}

// Closest hit.  "While-while" loop: each lane descends internal nodes
// until it reaches a leaf (or runs out of work), and only then do the lanes
// intersect leaf triangles, so node steps and triangle tests of different
// lanes are not interleaved within one iteration.
template <bool COUNT>
DEV HitRec intersect_closest(const DevScene& S, const Ray& ray, const TravCount<COUNT>& cnt) {
    HitRec best;
    best.t = ray.tmax;
    best.b1 = best.b2 = 0.0f;
    best.slot = kNoHit;
    best.prim = kNoHit;
    if (S.nodeCount == 0) return best;
    const RayPre r = precompute(ray);
    uint32_t stackNode[kStackSize];
    float stackT[kStackSize];
    int sp = 0;
    uint32_t cur = 0u;  // the root is always an internal node
    for (;;) {
        // Node phase.
        while (cur != kEmpty && !(cur & kLeafBit)) {
            const float4* nd = S.nodes + static_cast<size_t>(cur) * 4u;
            cnt.node();
            float tn[4];
            uint32_t w[4];
            const uint32_t mask = node_boxes(nd, r, best.t, tn, w);
            if (mask) {
                cur = push_sorted(mask, tn, w, stackNode, stackT, sp);
            } else {
                cur = kEmpty;
                while (sp > 0) {
                    --sp;
                    if (stackT[sp] <= best.t) {
                        cur = stackNode[sp];
                        break;
                    }
                }
            }
        }
        if (cur == kEmpty) break;
        // Leaf phase.
        leaf_closest(S, r, cur & ~kLeafBit, best, cnt);
        cur = kEmpty;
        while (sp > 0) {
            --sp;
            if (stackT[sp] <= best.t) {
                cur = stackNode[sp];
                break;
            }
        }
        if (cur == kEmpty) break;
    }
    return best;
}

// Closest hit of nearly parallel rays as a packet: every active lane of
// the wave calls this together (lanes with valid = false only take part in
// the wave's decisions).  The wave holds one stack (`wstack`, kStackSize
// entries in LDS) and one current node; every lane tests its own ray
// against the node's boxes (culled by its own closest hit), the wave
// descends into the children any lane hits, nearest first by the entry
// distances of the first lane that hits one, and at a leaf every lane tests
// the leaf's triangles.  Camera rays of an 8x8 tile barely diverge, so the
// union of their paths through the tree is barely larger than one ray's and
// every node test keeps all lanes busy, where the per-lane traversal loses
// most lanes to divergence.  Results equal intersect_closest's: each lane
// still meets every triangle whose box its ray enters with an entry
// distance up to its closest hit, and the closest-hit rule (smallest t,
// then smallest primitive index) does not depend on the order.
// CNT counts the node and triangle fetches (node(id) / tri(slot) per lane).
template <int W, class CNT>
DEV HitRec intersect_closest_packet(const DevScene& S, const Ray& ray, bool valid, uint32_t* wstack, CNT& cnt) {
    HitRec best;
    best.t = valid ? ray.tmax : kFltMax;
    best.b1 = best.b2 = 0.0f;
    best.slot = best.prim = kNoHit;
    RayPre r{};
    if (valid) r = precompute(ray);
    if (S.nodeCount == 0u || !__any(valid)) return best;
    int sp = 0;
    uint32_t cur = 0u;
    for (;;) {
        if (cur & kLeafBit) {
            if (valid) leaf_closest(S, r, cur & ~kLeafBit, best, cnt);
        } else {
            const float4* nd = node_at<W>(S, cur);
            if (valid) cnt.node(cur);
            float tn[W];
            uint32_t wd[W];
#pragma unroll
            for (int k = 0; k < W; ++k) {
                tn[k] = kInf;
                wd[k] = kEmpty;
            }
            const uint32_t mask = valid ? node_test<W>(nd, r, best.t, tn, wd) : 0u;
            const unsigned long long hitters = __ballot(mask != 0u);
            if (hitters != 0ull) {
                uint32_t any = 0u;
                for (uint32_t k = 0; k < static_cast<uint32_t>(W); ++k)
                    if (__ballot((mask >> k) & 1u)) any |= 1u << k;
                // Children ordered by the first hitting lane's entry
                // distances; children only other lanes hit follow (key
                // FLT_MAX), unhit ones sort last (key +inf).
                // The leader's values are read into scalar registers
                // (readlane) and sorted there as integer keys: the sort is
                // the same for every lane, so it runs on the scalar unit
                // beside the vector instructions of other waves.
                const int leader = __ffsll(static_cast<long long>(hitters)) - 1;
                const uint32_t lm = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mask), leader));
                uint32_t key[W], val[W];
                for (int k = 0; k < W; ++k) {
                    const uint32_t lt = static_cast<uint32_t>(
                        __builtin_amdgcn_readlane(static_cast<int>(__float_as_uint(tn[k])), leader));
                    key[k] = (lm >> k) & 1u ? lt : ((any >> k) & 1u ? __float_as_uint(kFltMax) : __float_as_uint(kInf));
                    val[k] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(wd[k]), leader));
                }
                if constexpr (W == 8) {
                    sort8(key, val);
                } else {
                    cswap(key[0], val[0], key[1], val[1]);
                    cswap(key[2], val[2], key[3], val[3]);
                    cswap(key[0], val[0], key[2], val[2]);
                    cswap(key[1], val[1], key[3], val[3]);
                    cswap(key[1], val[1], key[2], val[2]);
                }
                const int n = __popc(any);
                for (int k = n - 1; k >= 1; --k) {
                    if (sp < kStackSize) wstack[sp] = val[k];
                    ++sp;
                }
                cur = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(val[0])));
                continue;
            }
        }
        if (sp == 0) break;
        --sp;
        cur = static_cast<uint32_t>(
            __builtin_amdgcn_readfirstlane(static_cast<int>(wstack[sp < kStackSize ? sp : kStackSize - 1])));
    }
    return best;
}

// Any hit in [tmin, tmax]: Occluded (CpuPathTracer.cpp:659-681, 697-709).
// Same while-while structure as the closest-hit traversal, hit children in
// node order (push_in_order), no distance culling of popped entries, first
// accepted triangle ends it.
template <bool COUNT>
DEV bool intersect_any(const DevScene& S, const Ray& ray, const TravCount<COUNT>& cnt) {
    if (S.nodeCount == 0) return false;
    const RayPre r = precompute(ray);
    uint32_t stackNode[kStackSize];
    float stackT[kStackSize];
    int sp = 0;
    uint32_t cur = 0u;
    for (;;) {
        while (cur != kEmpty && !(cur & kLeafBit)) {
            const float4* nd = S.nodes + static_cast<size_t>(cur) * 4u;
            cnt.node();
            float tn[4];
            uint32_t w[4];
            const uint32_t mask = node_boxes(nd, r, ray.tmax, tn, w);
            if (mask) {
                cur = push_in_order<4>(mask, w, stackNode, sp);
            } else {
                cur = sp > 0 ? stackNode[--sp] : kEmpty;
            }
        }
        if (cur == kEmpty) break;
        uint32_t slot = cur & ~kLeafBit;
        for (;;) {
            const float4* tp = S.tris + static_cast<size_t>(slot) * 3u;
            const float4 a = tp[0], b = tp[1], c = tp[2];
            cnt.tri();
            float t, b1, b2;
            if (tri_test(r, xyz(a), xyz(b), xyz(c), ray.tmax, &t, &b1, &b2)) return true;
            if (__float_as_uint(b.w) & kLeafEndFlag) break;
            ++slot;
        }
        cur = sp > 0 ? stackNode[--sp] : kEmpty;
    }
    return false;
}

}  // namespace shiro_hip
