// Mesh ingestion helpers for hosts that feed the backend from polygon
// meshes, as Shiro's Hydra scene bridge does before it hands triangle
// meshes to a render backend (reference src/hydra/SceneBridge.cpp).  Plain
// host code; fp32 operations in the order of the reference's Vec3f helpers
// (include/shiro/render/Types.h) with contraction off, so results equal the
// bridge's bit for bit.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/shiro_hip.h"

namespace {

struct F3 {
    float x, y, z;
};

F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 add(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length(F3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
F3 load(const float* p, uint32_t i) { return F3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

}  // namespace

extern "C" int shiro_hip_triangulate(const int32_t* faceVertexCounts, uint32_t faceCount,
                                     const int32_t* faceVertexIndices, uint32_t indexCount, uint32_t* outIndices,
                                     uint64_t outCapacity, uint64_t* outCount) {
    if ((faceCount && !faceVertexCounts) || (indexCount && !faceVertexIndices) || !outCount) return SHIRO_HIP_INVALID;
    uint64_t need = 0, off = 0;
    for (uint32_t f = 0; f < faceCount; ++f) {
        const int32_t c = faceVertexCounts[f];
        if (c < 0 || off + static_cast<uint64_t>(c) > indexCount) return SHIRO_HIP_INVALID;
        if (c >= 3) need += 3u * static_cast<uint64_t>(c - 2);
        off += static_cast<uint64_t>(c);
    }
    *outCount = need;
    if (!outIndices) return SHIRO_HIP_OK;
    if (outCapacity < need) return SHIRO_HIP_INVALID;
    uint64_t w = 0;
    off = 0;
    for (uint32_t f = 0; f < faceCount; ++f) {
        const int32_t c = faceVertexCounts[f];
        if (c >= 3) {
            const uint32_t root = static_cast<uint32_t>(faceVertexIndices[off]);
            for (int32_t k = 1; k + 1 < c; ++k) {
                outIndices[w++] = root;
                outIndices[w++] = static_cast<uint32_t>(faceVertexIndices[off + k]);
                outIndices[w++] = static_cast<uint32_t>(faceVertexIndices[off + k + 1]);
            }
        }
        off += static_cast<uint64_t>(c);
    }
    return SHIRO_HIP_OK;
}

extern "C" int shiro_hip_smooth_normals(const float* positions, uint32_t vertexCount, const uint32_t* indices,
                                        uint32_t indexCount, float* outNormals) {
    if ((vertexCount && (!positions || !outNormals)) || (indexCount && !indices)) return SHIRO_HIP_INVALID;
    std::vector<F3> acc(vertexCount, F3{0.0f, 0.0f, 0.0f});
    for (uint64_t t = 0; t + 2 < indexCount; t += 3) {
        const uint32_t i0 = indices[t], i1 = indices[t + 1], i2 = indices[t + 2];
        if (i0 >= vertexCount || i1 >= vertexCount || i2 >= vertexCount) continue;
        const F3 p0 = load(positions, i0);
        const F3 fn = cross(sub(load(positions, i1), p0), sub(load(positions, i2), p0));
        if (length(fn) <= 1.0e-8f) continue;  // degenerate triangle
        acc[i0] = add(acc[i0], fn);
        acc[i1] = add(acc[i1], fn);
        acc[i2] = add(acc[i2], fn);
    }
    for (uint32_t v = 0; v < vertexCount; ++v) {
        const float l = length(acc[v]);
        const F3 n = l > 0.0f ? F3{acc[v].x / l, acc[v].y / l, acc[v].z / l} : F3{0.0f, 1.0f, 0.0f};
        std::memcpy(outNormals + 3u * v, &n, sizeof(n));
    }
    return SHIRO_HIP_OK;
}
