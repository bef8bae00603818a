// Host check of the BVH builder (shiro_amd/csrc/bvh_build.cpp): builds
// both tree widths over seeded triangle sets, runs the structural checks,
// refits, and prints one line of figures per case for the Python test.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "layout.h"

using shiro_hip::Bvh;

static std::vector<float> make_case(int kind, size_t n, uint32_t seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    std::vector<float> t(9 * n);
    for (size_t i = 0; i < n; ++i) {
        float c[3];
        for (int a = 0; a < 3; ++a) c[a] = kind == 1 ? 0.25f : U(rng) * 50.0f;  // kind 1: coincident centroids
        if (kind == 2) c[1] = 0.0f;                                           // kind 2: a flat layer
        const float s = kind == 3 ? 1e-3f : 0.5f;                              // kind 3: tiny triangles
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) t[9 * i + 3 * v + a] = c[a] + (kind == 2 && a == 1 ? 0.0f : s * U(rng));
    }
    return t;
}

int main() {
    const struct { int kind; size_t n; } cases[] = {{0, 1}, {0, 2}, {0, 7}, {0, 100}, {0, 20000}, {1, 300},
                                                    {2, 5000}, {3, 5000}, {0, 150000}};
    int failures = 0;
    for (const auto& c : cases) {
        std::vector<float> tris = make_case(c.kind, c.n, 1234u + static_cast<uint32_t>(c.n));
        Bvh b;
        std::string err;
        bool ok = shiro_hip::build_bvh(tris.data(), c.n, shiro_hip::kMaxBvhDepth, &b, &err);
        if (ok) ok = shiro_hip::validate_bvh(b, tris.data(), c.n, shiro_hip::kMaxBvhDepth, &err);
        // Refit of the unmoved triangles reproduces both node arrays.
        bool same = false;
        if (ok) {
            Bvh r = b;
            ok = shiro_hip::refit_bvh(&r, tris.data(), c.n, &err);
            // Bitwise: empty child words are NaN patterns as floats.
            same = ok && r.nodes.size() == b.nodes.size() && r.nodes8.size() == b.nodes8.size() &&
                   std::memcmp(r.nodes.data(), b.nodes.data(), 4 * b.nodes.size()) == 0 &&
                   std::memcmp(r.nodes8.data(), b.nodes8.data(), 4 * b.nodes8.size()) == 0;
        }
        // Refit after moving every vertex still passes the checks.
        bool moved = false;
        if (ok) {
            std::vector<float> m = tris;
            std::mt19937 rng(7u);
            std::uniform_real_distribution<float> U(-0.3f, 0.3f);
            for (float& v : m) v += U(rng);
            Bvh r = b;
            moved = shiro_hip::refit_bvh(&r, m.data(), c.n, &err) && shiro_hip::validate_bvh(r, m.data(), c.n,
                                                                                         shiro_hip::kMaxBvhDepth, &err);
        }
        if (!ok || !same || !moved) ++failures;
        std::printf("case kind=%d n=%zu ok=%d refit_same=%d refit_moved=%d nodes=%llu nodes8=%llu depth=%u depth8=%u "
                    "stack=%u stack8=%u leaves=%llu err=%s\n",
                    c.kind, c.n, ok ? 1 : 0, same ? 1 : 0, moved ? 1 : 0,
                    static_cast<unsigned long long>(b.nodeCount()), static_cast<unsigned long long>(b.node8Count()),
                    b.depth, b.depth8, b.stack, b.stack8, static_cast<unsigned long long>(b.leafCount),
                    err.empty() ? "-" : err.c_str());
    }
    // A depth bound the SAH tree exceeds: build_bvh builds again with
    // binary nodes at the bound made leaves, and the result passes the
    // checks at that bound (deeper leaves hold more triangles).
    for (uint32_t bound : {3u, 5u}) {
        const size_t n = 20000;
        std::vector<float> tris = make_case(0, n, 99u + bound);
        Bvh b;
        std::string err;
        bool ok = shiro_hip::build_bvh(tris.data(), n, bound, &b, &err);
        if (ok) ok = shiro_hip::validate_bvh(b, tris.data(), n, bound, &err);
        if (!ok || b.depth > bound) ++failures;
        std::printf("capped bound=%u ok=%d depth=%u stack=%u leaves=%llu maxleaf=%u err=%s\n", bound, ok ? 1 : 0,
                    b.depth, b.stack, static_cast<unsigned long long>(b.leafCount), b.maxLeafSize,
                    err.empty() ? "-" : err.c_str());
    }
    return failures == 0 ? 0 : 1;
}
