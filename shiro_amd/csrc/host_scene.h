// Host-side scene preparation for the device layout (layout.h).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "layout.h"

namespace shiro_hip {

// Reference PbrMaterial defaults (Renderer.h:56-102).
shiro_hip_material default_material();

// Material-only constants with the reference's expressions
// (CpuPathTracer.cpp:841-1441): lobe weights, F0s, tints, isotropic
// sampling roughnesses, medium extinction.
MaterialDerived derive_material(const shiro_hip_material& m);
// Assigns MaterialDerived::shadeClass: materials whose shading takes the
// same branches (feature flags, which lobes have weight, which lobes are
// delta) share a class, numbered in order of first appearance; classes past
// kShadeClasses - 2 share the last material class.
void assign_shade_classes(const shiro_hip_material* m, MaterialDerived* d, size_t n);

// Row weight factor of the importance table, max(sin(theta_y), 1e-6) at
// the row centre (EnvironmentMap.cpp:197-199); the device rebuild of the
// tables takes these host values so that it matches build_env_tables.
float env_row_sine(uint32_t y, uint32_t h);
// EnvironmentMap::BuildImportanceTable (EnvironmentMap.cpp:185-236):
// sequential fp32 prefix sums in the reference's order.  pixels is w*h RGB.
// Returns false (tables empty) when the layout is not lat-long.
bool build_env_tables(const float* pixels, uint32_t w, uint32_t h, uint32_t layout, std::vector<float>* rowCdf,
                      std::vector<float>* condCdf);

// Search guides of a CDF table for the device's lower_bound: for a table
// cdf[begin .. end] and G buckets, guide[g] = the lower_bound position of
// key g / G over cdf[begin + 1 .. end] (end + 1 if none), g = 0..G.  A key
// in [a / G, b / G] then has its lower_bound in [guide[a], guide[b]].  If
// the table is not non-decreasing (NaN weights), the guide brackets the
// whole range instead.  `out` receives G + 1 entries per table.
void build_cdf_guide(const float* cdf, uint32_t begin, uint32_t end, uint32_t G, uint32_t* out);

// Flattened triangle soup of a scene in scene order (meshes concatenated,
// full index triples only, CpuPathTracer.cpp:434).
struct FlatScene {
    std::vector<float> tri;        // 9 floats per triangle
    std::vector<float> nrm;        // 9 floats per triangle (vertex normals), when hasNormals
    std::vector<uint8_t> hasNormals;
    std::vector<uint32_t> material; // material index per triangle (default material = materialCount)
    std::vector<float> uv;          // 6 floats per triangle (vertex UVs, 0 where a mesh has none), when hasUvs
    bool hasUvs = false;            // any mesh with UVs
    bool textured = false;          // any material with a texture
    bool plainBsdf = true;          // no material with anisotropy, thin film, sheen or subsurface wrap
    bool emissive = false;          // any emissive material in use
};

// Emissive-triangle light list of BuildAccelerationScene
// (CpuPathTracer.cpp:482-522): meshes in scene order whose index count is a
// multiple of 3 and whose material emits (luminance of emissionColor *
// strength > 0), their triangles of non-zero area in order; weight = area *
// luminance, cumulative = running fp32 sum of the weights.
struct EmissiveEntry {
    uint32_t tri;          // global triangle index (scene order)
    float area, weight, cumulative;
    float normal[3], radiance[3];
    float p[9];            // vertices
};
void build_emissive(const shiro_hip_scene& s, std::vector<EmissiveEntry>* out, float* sum);

// Validates and flattens; returns false with a message on invalid input
// (index out of range, null arrays, too many lights, a material naming a
// missing texture, an empty texture).
bool flatten_scene(const shiro_hip_scene& s, FlatScene* out, std::string* err);

}  // namespace shiro_hip
