// Host BVH4 builder; see bvh_build.h.
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
constexpr uint32_t kLeafMax = 4;
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
    if (count <= kLeafMax || depth >= kMaxBinaryDepth) {
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

struct Collapser {
    const Builder& B;
    Bvh4* out;
    uint32_t maxDepth = 0;

    uint32_t alloc() {
        const uint32_t id = static_cast<uint32_t>(out->nodes.size() / 32u);
        out->nodes.resize(out->nodes.size() + 32u, 0.0f);
        return id;
    }

    // Fills 4-wide node `id` from binary node `bi`; returns the subtree depth.
    uint32_t emit(uint32_t id, uint32_t bi) {
        uint32_t kids[4];
        int nk = 0;
        const Node2& n = B.nodes[bi];
        if (n.count > 0) {
            kids[nk++] = bi;  // a leaf root: one leaf child
        } else {
            kids[nk++] = n.left;
            kids[nk++] = n.left + 1;
            while (nk < 4) {
                int best = -1;
                float bestArea = -1.0f;
                for (int k = 0; k < nk; ++k) {
                    const Node2& c = B.nodes[kids[k]];
                    if (c.count == 0 && c.box.area() > bestArea) {
                        bestArea = c.box.area();
                        best = k;
                    }
                }
                if (best < 0) break;
                const uint32_t open = kids[best];
                kids[best] = B.nodes[open].left;
                kids[nk++] = B.nodes[open].left + 1;
            }
        }
        // Allocate the internal children first so they are contiguous.
        uint32_t childId[4] = {kEmpty, kEmpty, kEmpty, kEmpty};
        for (int k = 0; k < nk; ++k)
            if (B.nodes[kids[k]].count == 0) childId[k] = alloc();
        uint32_t sub = 0;
        float* nd = &out->nodes[static_cast<size_t>(id) * 32u];
        for (int k = 0; k < 4; ++k) {
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            uint32_t word = kEmpty, cnt = 0;
            if (k < nk) {
                const Node2& c = B.nodes[kids[k]];
                Box b = c.box;
                pad(b);
                for (int a = 0; a < 3; ++a) {
                    lo[a] = b.lo[a];
                    hi[a] = b.hi[a];
                }
                if (c.count > 0) {
                    word = kLeafBit | c.left;
                    cnt = c.count;
                    out->leafCount += 1;
                    out->maxLeafSize = std::max(out->maxLeafSize, c.count);
                    out->leafEnd[c.left + c.count - 1] = 1;
                } else {
                    word = childId[k];
                }
            }
            nd = &out->nodes[static_cast<size_t>(id) * 32u];
            nd[0 + k] = lo[0];
            nd[4 + k] = hi[0];
            nd[8 + k] = lo[1];
            nd[12 + k] = hi[1];
            nd[16 + k] = lo[2];
            nd[20 + k] = hi[2];
            std::memcpy(&nd[24 + k], &word, 4);
            std::memcpy(&nd[28 + k], &cnt, 4);
        }
        for (int k = 0; k < nk; ++k)
            if (childId[k] != kEmpty) sub = std::max(sub, emit(childId[k], kids[k]));
        return sub + 1;
    }
};

}  // namespace

bool build_bvh4(const float* tris, size_t n, uint32_t maxDepth, Bvh4* out, std::string* err) {
    out->nodes.clear();
    out->order.clear();
    out->leafEnd.clear();
    out->depth = 0;
    out->leafCount = 0;
    out->maxLeafSize = 0;
    if (n == 0) return true;
    if (n >= static_cast<size_t>(kLeafBit)) {
        *err = "too many triangles for the 31-bit leaf slot index";
        return false;
    }
    Builder B;
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
    Collapser C{B, out};
    out->nodes.reserve(32u * (n / 2 + 4));
    const uint32_t root = C.alloc();
    out->depth = C.emit(root, 0);
    if (out->depth > maxDepth) {
        *err = "BVH4 depth " + std::to_string(out->depth) + " exceeds the traversal stack bound " +
               std::to_string(maxDepth);
        return false;
    }
    return true;
}

bool validate_bvh4(const Bvh4& b, size_t n, uint32_t maxDepth, std::string* err) {
    const uint64_t nn = b.nodeCount();
    if (n == 0) return nn == 0 || (*err = "nodes without triangles", false);
    if (nn == 0) {
        *err = "no nodes";
        return false;
    }
    std::vector<uint8_t> seen(nn, 0);
    std::vector<uint8_t> slotSeen(n, 0);
    struct Item {
        uint32_t node, depth;
    };
    std::vector<Item> todo{{0u, 1u}};
    seen[0] = 1;
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        if (it.depth > maxDepth) {
            *err = "depth bound exceeded";
            return false;
        }
        const float* nd = &b.nodes[static_cast<size_t>(it.node) * 32u];
        for (int k = 0; k < 4; ++k) {
            uint32_t w;
            std::memcpy(&w, &nd[24 + k], 4);
            if (w == kEmpty) continue;
            if (w & kLeafBit) {
                uint32_t slot = w & ~kLeafBit;
                for (;;) {
                    if (slot >= n) {
                        *err = "leaf runs past the triangle array";
                        return false;
                    }
                    if (slotSeen[slot]) {
                        *err = "triangle slot referenced twice";
                        return false;
                    }
                    slotSeen[slot] = 1;
                    if (b.leafEnd[slot]) break;
                    ++slot;
                }
            } else {
                if (w == 0u || w >= nn || seen[w]) {
                    *err = "bad or shared child node index";
                    return false;
                }
                seen[w] = 1;
                todo.push_back({w, it.depth + 1u});
            }
        }
    }
    for (size_t i = 0; i < n; ++i)
        if (!slotSeen[i]) {
            *err = "triangle slot not referenced by any leaf";
            return false;
        }
    return true;
}

}  // namespace shiro_hip
