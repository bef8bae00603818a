// Host BVH builder for the device traversal (dev_bvh.h): one binary SAH
// tree collapsed into a 4-wide and an 8-wide tree over the same leaves.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace shiro_hip {

struct Bvh {
    std::vector<float> nodes;       // 4-wide: kNodeFloats (64 B) per node, layout.h
    std::vector<float> nodes8;      // 8-wide: kNode8Floats (128 B) per node; empty if its stack bound is too deep
    std::vector<uint32_t> order;    // leaf slot -> triangle index (both widths)
    std::vector<uint8_t> leafEnd;   // leaf slot -> 1 if last of its leaf
    uint32_t depth = 0, depth8 = 0; // levels of each tree
    uint32_t stack = 0, stack8 = 0; // most entries a traversal of each tree keeps on its stack
    uint64_t leafCount = 0;
    uint32_t maxLeafSize = 0;
    uint64_t nodeCount() const { return nodes.size() / 16u; }
    uint64_t node8Count() const { return nodes8.size() / 32u; }
};

// Builds the trees over n triangles given as 9 floats each (v0, v1, v2).
// Binned SAH (32 bins) on centroids builds a binary tree with leaves of at
// most 4 triangles (more only where centroids coincide or at the depth
// limit); subtrees above 32K triangles are built on worker threads.  The
// binary tree is then collapsed twice, into W-wide nodes (W = 4, 8) over its
// leaves, choosing which binary nodes become wide nodes by dynamic
// programming to the least sum of wide-node surface areas (the SAH cost
// with the leaves fixed).  Boxes are padded by a few ulps (and by a
// rounding margin of the scene's coordinate scale) so that the device slab
// test is conservative, then quantised to 8 bits per plane relative to their
// parent node.  Returns false (with a message) if the 4-wide tree is deeper
// than maxDepth levels; the 8-wide tree is dropped (nodes8 empty) if a
// traversal of it could hold more than kStackSize - 1 stack entries.
bool build_bvh(const float* tris, size_t n, uint32_t maxDepth, Bvh* out, std::string* err);
// build_bvh with binary nodes at depth binaryCap made leaves; build_bvh
// falls back to binaryCap = maxDepth when its tree is too deep.
bool build_bvh_capped(const float* tris, size_t n, uint32_t maxDepth, uint32_t binaryCap, Bvh* out,
                      std::string* err);

// Refit: recomputes every box of both built trees for moved vertices of
// the same n triangles (same order as given to build_bvh), keeping the
// topology and leaf order.  Boxes are padded and quantised as in the build,
// so refitting unmoved triangles reproduces the built node arrays exactly.
bool refit_bvh(Bvh* b, const float* tris, size_t n, std::string* err);

// Check of the built trees before upload: every internal child word names
// an existing node other than the root, every leaf word names a triangle
// slot whose run ends with an end-of-leaf flag inside the array, every node
// is reached once and every triangle slot once, the depth is within maxDepth
// (4-wide) and the stack bound within kStackSize (8-wide), the valid bits
// match the child words, and every dequantised child box contains all
// vertices of the triangles below it (tris as given to build_bvh).
bool validate_bvh(const Bvh& b, const float* tris, size_t n, uint32_t maxDepth, std::string* err);

}  // namespace shiro_hip
