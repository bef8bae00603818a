// Host BVH builder (4- and 8-wide); see bvh_build.h.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>

#include "layout.h"

namespace shiro_hip {

namespace {

constexpr int kBins = 32;
// Leaves of at most kLeafMax triangles, or up to twice that where the SAH
// finds splitting does not pay.  Small leaves suit the traversal kernels,
// whose node steps take about twice the instructions of a triangle test
// (c3, 5M triangles: leaves of <= 2/4 triangles against <= 4/8 cut triangle
// tests per sample from 10.2 to 5.7 and raise node steps from 24.8 to 26.1;
// closest-hit 25.1 -> 23.4 ms, any-hit 24.9 -> 23.9 ms per frame).
constexpr uint32_t kLeafMax = 2;
constexpr uint32_t kMaxBinaryDepth = 58;
constexpr uint32_t kParallelThreshold = 32768;

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Box& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Ref {
    Box box;
    float c[3];
    uint32_t tri;
};

struct Node2 {
    Box box;
    uint32_t left = 0;   // internal: left child index (right = left+1); leaf: first ref
    uint32_t count = 0;  // 0: internal
};

struct Builder {
    std::vector<Ref> refs;
    std::vector<Node2> nodes;
    std::atomic<uint32_t> nodeCount{1};
    std::atomic<int> threads{1};
    int maxThreads = 1;
    uint32_t maxBinaryDepth = kMaxBinaryDepth;  // nodes at this depth become leaves
};

void build_rec(Builder& B, uint32_t ni, uint32_t begin, uint32_t end, uint32_t depth) {
    Node2& node = B.nodes[ni];
    Box box, cbox;
    for (uint32_t i = begin; i < end; ++i) {
        box.grow(B.refs[i].box);
        for (int a = 0; a < 3; ++a) {
            cbox.lo[a] = std::min(cbox.lo[a], B.refs[i].c[a]);
            cbox.hi[a] = std::max(cbox.hi[a], B.refs[i].c[a]);
        }
    }
    node.box = box;
    const uint32_t count = end - begin;
    auto leaf = [&]() {
        node.left = begin;
        node.count = count;
    };
    if (count <= kLeafMax || depth >= B.maxBinaryDepth) {
        leaf();
        return;
    }
    // Binned SAH over the three axes.
    int bestAxis = -1, bestSplit = -1;
    float bestCost = FLT_MAX;
    float scale[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = cbox.hi[a] - cbox.lo[a];
        scale[a] = ext > 0.0f ? static_cast<float>(kBins) * (1.0f - 1e-6f) / ext : 0.0f;
    }
    for (int a = 0; a < 3; ++a) {
        if (scale[a] == 0.0f) continue;
        Box bb[kBins];
        uint32_t bc[kBins] = {};
        for (uint32_t i = begin; i < end; ++i) {
            int b = static_cast<int>((B.refs[i].c[a] - cbox.lo[a]) * scale[a]);
            b = std::min(std::max(b, 0), kBins - 1);
            bb[b].grow(B.refs[i].box);
            ++bc[b];
        }
        float rightArea[kBins];
        uint32_t rightCount[kBins];
        Box acc;
        uint32_t rc = 0;
        for (int b = kBins - 1; b > 0; --b) {
            acc.grow(bb[b]);
            rc += bc[b];
            rightArea[b] = acc.area();
            rightCount[b] = rc;
        }
        Box lacc;
        uint32_t lc = 0;
        for (int b = 0; b < kBins - 1; ++b) {
            lacc.grow(bb[b]);
            lc += bc[b];
            if (lc == 0 || rightCount[b + 1] == 0) continue;
            const float cost = lacc.area() * static_cast<float>(lc) + rightArea[b + 1] * static_cast<float>(rightCount[b + 1]);
            if (cost < bestCost) {
                bestCost = cost;
                bestAxis = a;
                bestSplit = b;
            }
        }
    }
    uint32_t mid;
    if (bestAxis < 0) {
        // All centroids coincide: split the range in the middle.
        mid = begin + count / 2;
    } else {
        // Leaf if splitting does not pay (intersection cost 1, traversal 1).
        const float leafCost = box.area() * static_cast<float>(count);
        if (count <= 2 * kLeafMax && leafCost <= bestCost + box.area()) {
            leaf();
            return;
        }
        const int a = bestAxis;
        const float lo = cbox.lo[a], s = scale[a];
        Ref* it = std::partition(B.refs.data() + begin, B.refs.data() + end, [&](const Ref& r) {
            int b = static_cast<int>((r.c[a] - lo) * s);
            b = std::min(std::max(b, 0), kBins - 1);
            return b <= bestSplit;
        });
        mid = static_cast<uint32_t>(it - B.refs.data());
        if (mid == begin || mid == end) mid = begin + count / 2;
    }
    const uint32_t l = B.nodeCount.fetch_add(2);
    node.left = l;
    node.count = 0;
    if (count > kParallelThreshold && B.threads.load() < B.maxThreads) {
        B.threads.fetch_add(1);
        std::thread t([&B, l, begin, mid, depth]() { build_rec(B, l, begin, mid, depth + 1); });
        build_rec(B, l + 1, mid, end, depth + 1);
        t.join();
        B.threads.fetch_sub(1);
    } else {
        build_rec(B, l, begin, mid, depth + 1);
        build_rec(B, l + 1, mid, end, depth + 1);
    }
}

void pad(Box& b) {
    for (int a = 0; a < 3; ++a) {
        const float p = (std::fabs(b.lo[a]) + std::fabs(b.hi[a])) * 2.0e-6f + 1.0e-9f;
        b.lo[a] -= p;
        b.hi[a] += p;
    }
}

// Node layouts of the two widths (layout.h): floats per node, where the
// child words start, and the uint holding plane byte k of axis a (hi: the
// hi plane) with the byte's position in it.
template <int W>
struct NodeLayout;
template <>
struct NodeLayout<4> {
    static constexpr size_t kFloats = kNodeFloats;
    static int plane(int a, int hi, int k, int* shift) {
        *shift = 8 * k;
        return 8 + 2 * a + hi;
    }
};
template <>
struct NodeLayout<8> {
    static constexpr size_t kFloats = kNode8Floats;
    static int plane(int a, int hi, int k, int* shift) {
        *shift = 8 * (k & 3);
        return 12 + 4 * a + 2 * hi + (k >> 2);
    }
};

// Quantises a node's child boxes to 8 bits per plane relative to their
// union (layout.h): origin = union.lo, per-axis scale 2^e with
// 255 * 2^e >= extent, lo rounded down and hi rounded up, so every
// dequantised box contains the exact one.  Empty children get inverted
// planes (lo 255, hi 0) and are masked out by the valid bits.
template <int W>
void quantize_node(const Box* boxes, int nk, const uint32_t* words, float* nd) {
    using L = NodeLayout<W>;
    Box u;
    for (int k = 0; k < nk; ++k) u.grow(boxes[k]);
    std::fill(nd, nd + L::kFloats, 0.0f);
    uint32_t q[L::kFloats] = {};
    uint32_t expo = 0;
    for (int a = 0; a < 3; ++a) {
        const double origin = u.lo[a];
        const double extent = static_cast<double>(u.hi[a]) - origin;
        int e = extent > 0.0 ? static_cast<int>(std::ceil(std::log2(extent / 255.0))) : -100;
        e = std::max(-100, std::min(100, e));
        while (std::ldexp(255.0, e) < extent) ++e;
        const double s = std::ldexp(1.0, e);
        for (int k = 0; k < W; ++k) {
            uint32_t lo = 0xffu, hi = 0u;
            if (k < nk) {
                double ql = std::floor((static_cast<double>(boxes[k].lo[a]) - origin) / s);
                double qh = std::ceil((static_cast<double>(boxes[k].hi[a]) - origin) / s);
                ql = std::max(0.0, std::min(255.0, ql));
                qh = std::max(0.0, std::min(255.0, qh));
                while (ql > 0.0 && origin + ql * s > boxes[k].lo[a]) ql -= 1.0;
                while (qh < 255.0 && origin + qh * s < boxes[k].hi[a]) qh += 1.0;
                lo = static_cast<uint32_t>(ql);
                hi = static_cast<uint32_t>(qh);
            }
            int sh;
            const int il = L::plane(a, 0, k, &sh);
            q[il] |= lo << sh;
            const int ih = L::plane(a, 1, k, &sh);
            q[ih] |= hi << sh;
        }
        expo |= static_cast<uint32_t>(e + 128) << (8 * a);
        nd[a] = u.lo[a];
    }
    uint32_t valid = 0;
    for (int k = 0; k < nk; ++k) valid |= 1u << k;
    expo |= valid << 24;
    std::memcpy(&nd[3], &expo, 4);
    uint32_t w[W];
    for (int k = 0; k < W; ++k) w[k] = k < nk ? words[k] : kEmpty;
    std::memcpy(&nd[4], w, sizeof(w));
    for (int a = 0; a < 3; ++a)
        for (int hi = 0; hi < 2; ++hi)
            for (int k = 0; k < W; k += 4) {
                int sh;
                const int i = L::plane(a, hi, k, &sh);
                std::memcpy(&nd[i], &q[i], 4);
            }
}

// Child words of a node (W of them; kEmpty past the valid ones) and its
// valid mask.
template <int W>
uint32_t node_words(const float* nd, uint32_t* words) {
    uint32_t ex;
    std::memcpy(&ex, &nd[3], 4);
    std::memcpy(words, &nd[4], 4 * W);
    return ex >> 24;
}

// Collapses the binary tree into W-wide nodes over the binary tree's leaves
// (so every width shares the triangle order).  Which binary nodes become
// wide nodes is chosen by dynamic programming over the binary tree (the
// surface-area cost of Ylitie, Karras and Laine, "Efficient incoherent ray
// traversal on GPUs through compressed wide BVHs", HPG 2017, with the
// leaves fixed): with every leaf kept, the expected traversal cost of a
// random ray is, up to a constant, the sum of the surface areas of the wide
// nodes, so for binary node n and i = 1..W
//   cost(n, i) = least area sum of representing n's subtree by at most i
//                roots (a leaf is one root of cost 0),
//   split(n, j) = min over k of cost(left, k) + cost(right, j - k),
//   cost(n, 1) = A(n) + split(n, W)           (n becomes one wide node),
//   cost(n, i) = min(cost(n, 1), split(n, i)) for i >= 2,
// and the root is cost(root, 1).
template <int W>
struct Collapser {
    const Builder& B;
    std::vector<float>* nodes;
    Bvh* out;              // leaf statistics (filled by the 4-wide pass only)
    float slack = 0.0f;    // absolute widening of every box (scene-scale rounding margin)
    // Per binary node and slot count i (index i - 1): the cost, and how it
    // is reached: 0 = the node itself is one root, k > 0 = k slots to the
    // left child and i - k to the right.
    std::vector<float> cost;
    std::vector<uint8_t> how;

    uint32_t alloc() {
        const uint32_t id = static_cast<uint32_t>(nodes->size() / NodeLayout<W>::kFloats);
        nodes->resize(nodes->size() + NodeLayout<W>::kFloats, 0.0f);
        return id;
    }

    void plan() {
        const uint32_t nb = B.nodeCount.load();
        cost.assign(static_cast<size_t>(nb) * W, 0.0f);
        how.assign(static_cast<size_t>(nb) * W, 0);
        // Children have larger indices than their parent.
        for (uint32_t n = nb; n-- > 0;) {
            const Node2& nd = B.nodes[n];
            float* c = &cost[static_cast<size_t>(n) * W];
            uint8_t* h = &how[static_cast<size_t>(n) * W];
            if (nd.count > 0) continue;  // a leaf: cost 0 with any slot count
            const float* cl = &cost[static_cast<size_t>(nd.left) * W];
            const float* cr = &cost[static_cast<size_t>(nd.left + 1) * W];
            float split[W + 1];
            uint8_t arg[W + 1];
            for (int j = 2; j <= W; ++j) {
                split[j] = FLT_MAX;
                arg[j] = 1;
                for (int k = 1; k < j; ++k) {
                    const float v = cl[k - 1] + cr[j - k - 1];
                    if (v < split[j]) {
                        split[j] = v;
                        arg[j] = static_cast<uint8_t>(k);
                    }
                }
            }
            c[0] = nd.box.area() + split[W];
            h[0] = 0;
            for (int i = 2; i <= W; ++i) {
                if (split[i] < c[0]) {
                    c[i - 1] = split[i];
                    h[i - 1] = arg[i];
                } else {
                    c[i - 1] = c[0];
                    h[i - 1] = 0;
                }
            }
        }
    }

    // Roots representing binary node m with i slots.
    void gather(uint32_t m, int i, uint32_t* kids, int* nk) const {
        const Node2& nd = B.nodes[m];
        const uint8_t h = how[static_cast<size_t>(m) * W + (i - 1)];
        if (nd.count > 0 || h == 0) {
            kids[(*nk)++] = m;
            return;
        }
        gather(nd.left, h, kids, nk);
        gather(nd.left + 1, i - h, kids, nk);
    }

    // Children of the wide node made from internal binary node bi.
    int children(uint32_t bi, uint32_t* kids) const {
        const Node2& nd = B.nodes[bi];
        const float* cl = &cost[static_cast<size_t>(nd.left) * W];
        const float* cr = &cost[static_cast<size_t>(nd.left + 1) * W];
        int bestK = 1;
        float best = FLT_MAX;
        for (int k = 1; k < W; ++k) {
            const float v = cl[k - 1] + cr[W - k - 1];
            if (v < best) {
                best = v;
                bestK = k;
            }
        }
        int nk = 0;
        gather(nd.left, bestK, kids, &nk);
        gather(nd.left + 1, W - bestK, kids, &nk);
        return nk;
    }

    // Fills node `id` from binary node `bi`; returns the subtree depth and,
    // in *stack, the most stack entries a traversal below it holds (each
    // node on the way down leaves at most its other children pushed).
    uint32_t emit(uint32_t id, uint32_t bi, uint32_t* stack) {
        uint32_t kids[W];
        int nk = 0;
        if (B.nodes[bi].count > 0) {
            kids[nk++] = bi;  // a leaf root: one leaf child
        } else {
            nk = children(bi, kids);
        }
        // Allocate the internal children first so they are contiguous.
        uint32_t childId[W];
        for (int k = 0; k < W; ++k) childId[k] = kEmpty;
        for (int k = 0; k < nk; ++k)
            if (B.nodes[kids[k]].count == 0) childId[k] = alloc();
        Box boxes[W];
        uint32_t words[W];
        for (int k = 0; k < nk; ++k) {
            const Node2& c = B.nodes[kids[k]];
            boxes[k] = c.box;
            pad(boxes[k]);
            for (int a = 0; a < 3; ++a) {
                boxes[k].lo[a] -= slack;
                boxes[k].hi[a] += slack;
            }
            if (c.count > 0) {
                words[k] = kLeafBit | c.left;
                if (W == 4) {
                    out->leafCount += 1;
                    out->maxLeafSize = std::max(out->maxLeafSize, c.count);
                    out->leafEnd[c.left + c.count - 1] = 1;
                }
            } else {
                words[k] = childId[k];
            }
        }
        quantize_node<W>(boxes, nk, words, &(*nodes)[static_cast<size_t>(id) * NodeLayout<W>::kFloats]);
        uint32_t sub = 0, below = 0;
        for (int k = 0; k < nk; ++k)
            if (childId[k] != kEmpty) {
                uint32_t st = 0;
                sub = std::max(sub, emit(childId[k], kids[k], &st));
                below = std::max(below, st);
            }
        *stack = static_cast<uint32_t>(nk - 1) + below;
        return sub + 1;
    }
};

}  // namespace

bool build_bvh(const float* tris, size_t n, uint32_t maxDepth, Bvh* out, std::string* err) {
    out->nodes.clear();
    out->nodes8.clear();
    out->order.clear();
    out->leafEnd.clear();
    out->depth = out->depth8 = 0;
    out->stack = out->stack8 = 0;
    out->leafCount = 0;
    out->maxLeafSize = 0;
    if (n == 0) return true;
    if (n >= static_cast<size_t>(kLeafBit)) {
        *err = "too many triangles for the 31-bit leaf slot index";
        return false;
    }
    if (build_bvh_capped(tris, n, maxDepth, kMaxBinaryDepth, out, err)) return true;
    // A tree deeper than the traversal stack allows (inputs whose SAH splits
    // peel off a few triangles at a time, such as geometrically spaced
    // ones): build again with binary nodes at depth maxDepth made leaves.
    // Collapsing never deepens a tree, so the 4-wide tree then fits; its
    // deepest leaves hold more triangles.
    err->clear();
    return build_bvh_capped(tris, n, maxDepth, maxDepth > 1u ? maxDepth - 1u : 1u, out, err);
}

bool build_bvh_capped(const float* tris, size_t n, uint32_t maxDepth, uint32_t binaryCap, Bvh* out,
                      std::string* err) {
    out->nodes.clear();
    out->nodes8.clear();
    out->order.clear();
    out->leafEnd.clear();
    out->depth = out->depth8 = 0;
    out->stack = out->stack8 = 0;
    out->leafCount = 0;
    out->maxLeafSize = 0;
    Builder B;
    B.maxBinaryDepth = binaryCap;
    B.refs.resize(n);
    for (size_t i = 0; i < n; ++i) {
        Ref& r = B.refs[i];
        const float* t = tris + 9 * i;
        for (int a = 0; a < 3; ++a) {
            r.box.lo[a] = std::min(t[a], std::min(t[3 + a], t[6 + a]));
            r.box.hi[a] = std::max(t[a], std::max(t[3 + a], t[6 + a]));
            r.c[a] = r.box.lo[a] + r.box.hi[a];
        }
        r.tri = static_cast<uint32_t>(i);
    }
    B.nodes.resize(2 * n + 1);
    B.maxThreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    build_rec(B, 0, 0, static_cast<uint32_t>(n), 0);
    out->order.resize(n);
    for (size_t i = 0; i < n; ++i) out->order[i] = B.refs[i].tri;
    out->leafEnd.assign(n, 0);
    // Rounding margin of the device slab test, which forms plane distances
    // from differences of scene coordinates: 16 ulps of the largest one.
    float scale = 0.0f;
    for (int a = 0; a < 3; ++a)
        scale = std::max(scale, std::max(std::fabs(B.nodes[0].box.lo[a]), std::fabs(B.nodes[0].box.hi[a])));
    const float slack = scale * 16.0f * FLT_EPSILON;
    Collapser<4> C{B, &out->nodes, out, slack};
    C.plan();
    out->nodes.reserve(kNodeFloats * (n / 2 + 4));
    out->depth = C.emit(C.alloc(), 0, &out->stack);
    if (out->depth > maxDepth) {
        *err = "BVH4 depth " + std::to_string(out->depth) + " exceeds the traversal stack bound " +
               std::to_string(maxDepth);
        return false;
    }
    C.cost = std::vector<float>();
    C.how = std::vector<uint8_t>();
    Collapser<8> C8{B, &out->nodes8, out, slack};
    C8.plan();
    out->depth8 = C8.emit(C8.alloc(), 0, &out->stack8);
    if (out->stack8 + 1 > static_cast<uint32_t>(kStackSize)) {
        out->nodes8.clear();
        out->nodes8.shrink_to_fit();
    }
    return true;
}

namespace {

template <int W>
struct Refitter {
    Bvh& b;
    std::vector<float>& nodes;
    const float* tris;
    float slack;

    Box tri_box(uint32_t slot) const {
        Box r;
        const float* t = tris + 9u * static_cast<size_t>(b.order[slot]);
        for (int a = 0; a < 3; ++a) {
            r.lo[a] = std::min(t[a], std::min(t[3 + a], t[6 + a]));
            r.hi[a] = std::max(t[a], std::max(t[3 + a], t[6 + a]));
        }
        return r;
    }

    // Requantises node id from its children's exact boxes; returns the union
    // of those boxes (the node's exact box, before padding).
    Box fit(uint32_t id) {
        float* nd = &nodes[static_cast<size_t>(id) * NodeLayout<W>::kFloats];
        uint32_t words[W];
        const uint32_t valid = node_words<W>(nd, words);
        int nk = 0;
        while (nk < W && (valid >> nk) & 1u) ++nk;
        Box boxes[W], u;
        for (int k = 0; k < nk; ++k) {
            Box c;
            if (words[k] & kLeafBit) {
                for (uint32_t slot = words[k] & ~kLeafBit;; ++slot) {
                    c.grow(tri_box(slot));
                    if (b.leafEnd[slot]) break;
                }
            } else {
                c = fit(words[k]);
            }
            u.grow(c);
            boxes[k] = c;
            pad(boxes[k]);
            for (int a = 0; a < 3; ++a) {
                boxes[k].lo[a] -= slack;
                boxes[k].hi[a] += slack;
            }
        }
        quantize_node<W>(boxes, nk, words, nd);
        return u;
    }
};

}  // namespace

bool refit_bvh(Bvh* b, const float* tris, size_t n, std::string* err) {
    if (n != b->order.size()) {
        *err = "refit: triangle count differs from the built tree's";
        return false;
    }
    if (n == 0) return true;
    Box all;
    for (size_t i = 0; i < 3 * n; ++i)
        for (int a = 0; a < 3; ++a) {
            all.lo[a] = std::min(all.lo[a], tris[3 * i + a]);
            all.hi[a] = std::max(all.hi[a], tris[3 * i + a]);
        }
    float scale = 0.0f;
    for (int a = 0; a < 3; ++a) scale = std::max(scale, std::max(std::fabs(all.lo[a]), std::fabs(all.hi[a])));
    const float slack = scale * 16.0f * FLT_EPSILON;
    Refitter<4> R{*b, b->nodes, tris, slack};
    R.fit(0);
    if (!b->nodes8.empty()) {
        Refitter<8> R8{*b, b->nodes8, tris, slack};
        R8.fit(0);
    }
    return true;
}

namespace {

// Dequantised box of child k of a node, in double precision.
template <int W>
void child_box(const float* nd, int k, double lo[3], double hi[3]) {
    uint32_t ex;
    std::memcpy(&ex, &nd[3], 4);
    for (int a = 0; a < 3; ++a) {
        const double s = std::ldexp(1.0, static_cast<int>((ex >> (8 * a)) & 0xffu) - 128);
        int sh;
        uint32_t ql, qh;
        std::memcpy(&ql, &nd[NodeLayout<W>::plane(a, 0, k, &sh)], 4);
        std::memcpy(&qh, &nd[NodeLayout<W>::plane(a, 1, k, &sh)], 4);
        lo[a] = nd[a] + static_cast<double>((ql >> sh) & 0xffu) * s;
        hi[a] = nd[a] + static_cast<double>((qh >> sh) & 0xffu) * s;
    }
}

template <int W>
struct Checker {
    const Bvh& b;
    const std::vector<float>& nodes;
    const float* tris;
    size_t n;
    uint32_t maxDepth;
    uint64_t nn;
    std::vector<uint8_t> seen, slotSeen;
    std::string* err;

    bool fail(const char* m) {
        *err = m;
        return false;
    }

    // Visits node `id`; grows lo/hi by the vertices of every triangle below
    // it and checks that each child's box contains its triangles.
    bool visit(uint32_t id, uint32_t depth, double lo[3], double hi[3]) {
        if (depth > maxDepth) return fail("depth bound exceeded");
        const float* nd = &nodes[static_cast<size_t>(id) * NodeLayout<W>::kFloats];
        uint32_t words[W];
        const uint32_t validMask = node_words<W>(nd, words);
        for (int k = 0; k < W; ++k) {
            const uint32_t w = words[k];
            const bool valid = (validMask >> k) & 1u;
            if (w == kEmpty) {
                if (valid) return fail("empty child marked valid");
                continue;
            }
            if (!valid) return fail("child not marked valid");
            double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
            if (w & kLeafBit) {
                uint32_t slot = w & ~kLeafBit;
                for (;;) {
                    if (slot >= n) return fail("leaf runs past the triangle array");
                    if (slotSeen[slot]) return fail("triangle slot referenced twice");
                    slotSeen[slot] = 1;
                    const float* t = tris + 9 * static_cast<size_t>(b.order[slot]);
                    for (int v = 0; v < 3; ++v)
                        for (int a = 0; a < 3; ++a) {
                            clo[a] = std::min(clo[a], static_cast<double>(t[3 * v + a]));
                            chi[a] = std::max(chi[a], static_cast<double>(t[3 * v + a]));
                        }
                    if (b.leafEnd[slot]) break;
                    ++slot;
                }
            } else {
                if (w == 0u || w >= nn || seen[w]) return fail("bad or shared child node index");
                seen[w] = 1;
                if (!visit(w, depth + 1, clo, chi)) return false;
            }
            double qlo[3], qhi[3];
            child_box<W>(nd, k, qlo, qhi);
            for (int a = 0; a < 3; ++a) {
                if (!(qlo[a] <= clo[a] && qhi[a] >= chi[a])) return fail("quantised child box does not contain its triangles");
                lo[a] = std::min(lo[a], clo[a]);
                hi[a] = std::max(hi[a], chi[a]);
            }
        }
        return true;
    }
};

template <int W>
bool validate_width(const Bvh& b, const std::vector<float>& nodes, const float* tris, size_t n, uint32_t maxDepth,
                    std::string* err) {
    const uint64_t nn = nodes.size() / NodeLayout<W>::kFloats;
    if (n == 0) return nn == 0 || (*err = "nodes without triangles", false);
    if (nn == 0) {
        *err = "no nodes";
        return false;
    }
    Checker<W> c{b, nodes, tris, n, maxDepth, nn, std::vector<uint8_t>(nn, 0), std::vector<uint8_t>(n, 0), err};
    c.seen[0] = 1;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (!c.visit(0u, 1u, lo, hi)) return false;
    for (size_t i = 0; i < n; ++i)
        if (!c.slotSeen[i]) {
            *err = "triangle slot not referenced by any leaf";
            return false;
        }
    return true;
}

}  // namespace

bool validate_bvh(const Bvh& b, const float* tris, size_t n, uint32_t maxDepth, std::string* err) {
    if (!validate_width<4>(b, b.nodes, tris, n, maxDepth, err)) return false;
    if (b.nodes8.empty()) return true;
    if (b.stack8 + 1 > static_cast<uint32_t>(kStackSize)) {
        *err = "8-wide tree: traversal stack bound exceeded";
        return false;
    }
    std::string e8;
    if (!validate_width<8>(b, b.nodes8, tris, n, 1000u, &e8)) {
        *err = "8-wide tree: " + e8;
        return false;
    }
    return true;
}

}  // namespace shiro_hip
