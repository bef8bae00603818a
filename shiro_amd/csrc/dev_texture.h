// Texture extension on the device (no counterpart in the reference): UV
// interpolation at a hit, bilinear lookups with repeat addressing, and the
// per-hit re-derivation of the material constants that depend on the base
// colour or the roughness, with the expressions of derive_material
// (host_scene.cpp, reference CpuPathTracer.cpp:841-1441) in the same fp32
// order, so that a texture holding a material's own values reproduces the
// untextured material bit for bit.
#pragma once

#include "dev_math.h"
#include "layout.h"

namespace shiro_hip {

// Bilinear footprint of one lookup: four texel indices (y * w + x) and their
// weights, in the order the lookup sums them.
struct TexTap {
    uint32_t i[4];
    float w[4];     // bilinear weights (the derivatives of the lookup)
    float tx, ty;   // fractional position between texel centres
};

DEV uint32_t wrap_index(int i, uint32_t n) {
    const int m = static_cast<int>(n);
    const int r = i % m;
    return static_cast<uint32_t>(r < 0 ? r + m : r);
}

// Texel centres at (i + 0.5) / size; coordinates beyond +-2^24 (where floats
// are integers) and non-finite ones are clamped first.
DEV TexTap tex_tap(uint32_t w, uint32_t h, V2 uv) {
    float x = uv.x * static_cast<float>(w) - 0.5f;
    float y = uv.y * static_cast<float>(h) - 0.5f;
    x = x == x ? clampf(x, -16777216.0f, 16777216.0f) : 0.0f;
    y = y == y ? clampf(y, -16777216.0f, 16777216.0f) : 0.0f;
    const float fx = floorf(x), fy = floorf(y);
    const float tx = x - fx, ty = y - fy;
    const uint32_t x0 = wrap_index(static_cast<int>(fx), w), y0 = wrap_index(static_cast<int>(fy), h);
    const uint32_t x1 = x0 + 1u == w ? 0u : x0 + 1u, y1 = y0 + 1u == h ? 0u : y0 + 1u;
    TexTap t;
    t.i[0] = y0 * w + x0;
    t.i[1] = y0 * w + x1;
    t.i[2] = y1 * w + x0;
    t.i[3] = y1 * w + x1;
    t.w[0] = (1.0f - tx) * (1.0f - ty);
    t.w[1] = tx * (1.0f - ty);
    t.w[2] = (1.0f - tx) * ty;
    t.w[3] = tx * ty;
    t.tx = tx;
    t.ty = ty;
    return t;
}

// Two lerps along x, one along y: exact for equal texels, so a constant
// texture returns its value unchanged.
DEV float tex_channel(const float* texels, uint32_t channels, const TexTap& t, uint32_t c) {
    const float v0 = texels[t.i[0] * channels + c], v1 = texels[t.i[1] * channels + c];
    const float v2 = texels[t.i[2] * channels + c], v3 = texels[t.i[3] * channels + c];
    const float a = v0 + (v1 - v0) * t.tx;
    const float b = v2 + (v3 - v2) * t.tx;
    return a + (b - a) * t.ty;
}

// Base colour from channels 0..2 (grey for a one-channel texture).
DEV V3 tex_rgb(const float* texels, uint32_t channels, const TexTap& t) {
    const float r = tex_channel(texels, channels, t, 0u);
    if (channels < 3u) return V3{r, r, r};
    return V3{r, tex_channel(texels, channels, t, 1u), tex_channel(texels, channels, t, 2u)};
}

// Interpolated UV of a closest hit (same barycentric order as the shading
// normal); (0, 0) when the scene has no UVs.
DEV V2 hit_uv(const DevScene& S, uint32_t slot, float b1, float b2) {
    if (!S.uvs) return V2{0.0f, 0.0f};
    const float2* q = S.uvs + static_cast<size_t>(slot) * 3u;
    const float2 a = q[0], b = q[1], c = q[2];
    const float b0 = 1.0f - b1 - b2;
    return V2{a.x * b0 + b.x * b1 + c.x * b2, a.y * b0 + b.y * b1 + c.y * b2};
}

DEV float f0_of_ior(float ior) {
    const float i = fmaxs(ior, 1.0f);
    const float r = (i - 1.0f) / (i + 1.0f);
    return r * r;
}

DEV V3 one_minus_clamped3(V3 v) {
    return V3{clampf(1.0f - v.x, 0.0f, 1.0f), clampf(1.0f - v.y, 0.0f, 1.0f), clampf(1.0f - v.z, 0.0f, 1.0f)};
}

DEV void st3(float* d, V3 v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}

// derive_material's base-colour and roughness terms for a material whose
// baseColor / roughness are replaced by `base` / `rough`; the other fields of
// `d` (coat F0 and tint, emission, medium, flags, sampling scales) are the
// material's own.
DEV void derive_textured(const shiro_hip_material& m, V3 base, float rough, MaterialDerived& d) {
    const float metal = clampf(m.metallic, 0.0f, 1.0f);
    const float trans = clampf(m.transmission, 0.0f, 1.0f);
    const float sss = clampf(m.subsurface, 0.0f, 1.0f);
    const float blend = clampf(max3(ld3(m.subsurfaceRadius)) / 3.0f, 0.0f, 1.0f);
    const V3 tint = lerp(base, ld3(m.subsurfaceColor), blend);
    const V3 b = lerp(base, tint, sss);
    const V3 diffuse = b * (clampf(m.baseWeight, 0.0f, 1.0f) * (1.0f - metal) * (1.0f - trans));
    st3(d.diffuseColor, diffuse);
    const V3 diel = ld3(m.specularColor) * (clampf(m.specularWeight, 0.0f, 1.0f) * f0_of_ior(m.ior));
    const V3 f0b = lerp(diel, base, metal);
    st3(d.specularF0, f0b);
    const V3 uc = ld3(d.coatTint) * one_minus_clamped3(ld3(d.coatF0));
    const V3 ub = one_minus_clamped3(f0b);
    d.wDiffuse = fmaxs(0.0f, max3(diffuse * uc * ub));
    d.wBaseSpecular = fmaxs(0.0f, max3(uc * f0b));
    if (trans <= 0.0f || metal >= 0.999f) {
        d.wTransmission = 0.0f;
    } else {
        const V3 tt = clamp3(ld3(m.transmissionColor), 0.0f, 1.0f);
        d.wTransmission = fmaxs(0.0f, max3(tt * uc * ub)) * trans * (1.0f - metal);
    }
    d.baseRoughness = clampf(rough + m.coatAffectRoughness * m.coatWeight * m.coatRoughness, 0.0f, 1.0f);
    d.transmissionRoughness = clampf(rough + m.transmissionExtraRoughness, 0.0f, 1.0f);
    d.sampleRoughnessBase = clampf(d.baseRoughness * d.sampleScaleBase, 0.02f, 1.0f);
    d.sampleRoughnessTrans = clampf(d.transmissionRoughness * d.sampleScaleTrans, 0.02f, 1.0f);
}

// Per-hit constants of a textured material: `eval` from the evaluation
// texels, `samp` from the sampling texels (layout.h DevTexture).
struct TexState {
    MaterialDerived eval, samp;
};

DEV void texture_material(const DevScene& S, const shiro_hip_material& m, const MaterialDerived& table, V2 uv,
                          TexState& ts) {
    V3 baseE = ld3(m.baseColor), baseS = baseE;
    float roughE = m.roughness, roughS = roughE;
    if (m.baseColorTexture >= 0) {
        const DevTexture& t = S.textures[m.baseColorTexture];
        const TexTap tap = tex_tap(t.w, t.h, uv);
        baseE = tex_rgb(t.eval, t.channels, tap);
        baseS = t.sampling == t.eval ? baseE : tex_rgb(t.sampling, t.channels, tap);
    }
    if (m.roughnessTexture >= 0) {
        const DevTexture& t = S.textures[m.roughnessTexture];
        const TexTap tap = tex_tap(t.w, t.h, uv);
        roughE = tex_channel(t.eval, t.channels, tap, 0u);
        roughS = t.sampling == t.eval ? roughE : tex_channel(t.sampling, t.channels, tap, 0u);
    }
    ts.eval = table;
    ts.samp = table;
    derive_textured(m, baseE, roughE, ts.eval);
    derive_textured(m, baseS, roughS, ts.samp);
}

}  // namespace shiro_hip
