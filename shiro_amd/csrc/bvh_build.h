// Host BVH4 builder for the device traversal (dev_bvh.h).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace shiro_hip {

struct Bvh4 {
    std::vector<float> nodes;       // kNodeFloats (64 B) per node, layout.h
    std::vector<uint32_t> order;    // leaf slot -> triangle index
    std::vector<uint8_t> leafEnd;   // leaf slot -> 1 if last of its leaf
    uint32_t depth = 0;             // BVH4 levels
    uint64_t leafCount = 0;
    uint32_t maxLeafSize = 0;
    uint64_t nodeCount() const { return nodes.size() / 16u; }
};

// Builds a BVH4 over n triangles given as 9 floats each (v0, v1, v2).
// Binned SAH (32 bins) on centroids builds a binary tree with leaves of at
// most 4 triangles (more only where centroids coincide or at the depth
// limit); subtrees above 32K triangles are built on worker threads.  The
// binary tree is then collapsed: each 4-wide node adopts its two children
// and repeatedly opens the internal child of largest surface area until it
// has four.  Boxes are padded by a few ulps (and by a rounding margin of the
// scene's coordinate scale) so that the device slab test is conservative,
// then quantised to 8 bits per plane relative to their parent node.
// Returns false (with a message) if the tree is deeper than maxDepth 4-wide
// levels.
bool build_bvh4(const float* tris, size_t n, uint32_t maxDepth, Bvh4* out, std::string* err);

// Refit: recomputes every box of a built tree for moved vertices of the
// same n triangles (same order as given to build_bvh4), keeping the tree's
// topology and leaf order.  Boxes are padded and quantised as in the build,
// so refitting unmoved triangles reproduces the built node array exactly.
bool refit_bvh4(Bvh4* b, const float* tris, size_t n, std::string* err);

// Check of a built tree before upload: every internal child word names an
// existing node other than the root, every leaf word names a triangle slot
// whose run ends with an end-of-leaf flag inside the array, every node is
// reached once, the depth is within maxDepth, the valid bits match the
// child words, and every dequantised child box contains all vertices of the
// triangles below it (tris as given to build_bvh4).
bool validate_bvh4(const Bvh4& b, const float* tris, size_t n, uint32_t maxDepth, std::string* err);

}  // namespace shiro_hip
