// Host-side scene preparation; see host_scene.h.  The material expressions
// are evaluated in the same fp32 order as Shiro's CPU path so that the
// per-material constants uploaded to the device equal what the CPU path
// computes inline (this file is built with -ffp-contract=off).
#include "host_scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace shiro_hip {

namespace {

constexpr float kPiH = 3.14159265358979323846f;

struct H3 {
    float x, y, z;
};
inline H3 h3(const float* p) { return {p[0], p[1], p[2]}; }
inline H3 operator+(H3 a, H3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline H3 operator-(H3 a, H3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline H3 operator*(H3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline H3 operator*(H3 a, H3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float clampf(float v, float lo, float hi) { return std::clamp(v, lo, hi); }
inline H3 clamp3(H3 v, float lo, float hi) { return {clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)}; }
inline H3 lerp(H3 a, H3 b, float t) { return a * (1.0f - t) + b * t; }
inline float max3(H3 v) { return std::max(v.x, std::max(v.y, v.z)); }
inline void st3(float* d, H3 v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}

float f0_scalar(float ior) {
    const float i = std::max(ior, 1.0f);
    const float r = (i - 1.0f) / (i + 1.0f);
    return r * r;
}

H3 one_minus_clamped(H3 v) {
    return {clampf(1.0f - v.x, 0.0f, 1.0f), clampf(1.0f - v.y, 0.0f, 1.0f), clampf(1.0f - v.z, 0.0f, 1.0f)};
}

// The factor of the isotropic AdjustedRoughness (CpuPathTracer.cpp:845-850):
// AdjustedRoughness = clamp(r * factor, 0.02, 1).
float sampling_scale(float aniso, float rot) {
    const float a = clampf(aniso, -0.95f, 0.95f);
    const float phase = std::cos(rot * 2.0f * kPiH);
    return 1.0f - 0.35f * a * phase;
}

float sampling_roughness(float r, float aniso, float rot) { return clampf(r * sampling_scale(aniso, rot), 0.02f, 1.0f); }

}  // namespace

shiro_hip_material default_material() {
    shiro_hip_material m;
    std::memset(&m, 0, sizeof(m));
    auto set3 = [](float* d, float a, float b, float c) {
        d[0] = a;
        d[1] = b;
        d[2] = c;
    };
    set3(m.baseColor, 0.8f, 0.8f, 0.8f);
    set3(m.specularColor, 1.0f, 1.0f, 1.0f);
    set3(m.transmissionColor, 1.0f, 1.0f, 1.0f);
    set3(m.coatColor, 1.0f, 1.0f, 1.0f);
    set3(m.subsurfaceColor, 1.0f, 1.0f, 1.0f);
    set3(m.subsurfaceRadius, 1.0f, 1.0f, 1.0f);
    set3(m.sheenColor, 1.0f, 1.0f, 1.0f);
    m.baseWeight = 1.0f;
    m.specularWeight = 1.0f;
    m.roughness = 0.5f;
    m.opacity = 1.0f;
    m.coatRoughness = 0.03f;
    m.coatIor = 1.5f;
    m.sheenRoughness = 0.3f;
    m.subsurfaceScale = 1.0f;
    m.ior = 1.5f;
    m.thinFilmIor = 1.5f;
    m.baseColorTexture = -1;
    m.roughnessTexture = -1;
    return m;
}

MaterialDerived derive_material(const shiro_hip_material& m) {
    MaterialDerived d;
    std::memset(&d, 0, sizeof(d));
    const float metal = clampf(m.metallic, 0.0f, 1.0f);
    const float trans = clampf(m.transmission, 0.0f, 1.0f);
    // DiffuseColor (:900-908).
    const float sss = clampf(m.subsurface, 0.0f, 1.0f);
    const float blend = clampf(max3(h3(m.subsurfaceRadius)) / 3.0f, 0.0f, 1.0f);
    const H3 tint = lerp(h3(m.baseColor), h3(m.subsurfaceColor), blend);
    const H3 base = lerp(h3(m.baseColor), tint, sss);
    const H3 diffuse = base * (clampf(m.baseWeight, 0.0f, 1.0f) * (1.0f - metal) * (1.0f - trans));
    st3(d.diffuseColor, diffuse);
    // SpecularF0 / CoatF0 (:918-928).
    const H3 diel = h3(m.specularColor) * (clampf(m.specularWeight, 0.0f, 1.0f) * f0_scalar(m.ior));
    const H3 f0b = lerp(diel, h3(m.baseColor), metal);
    st3(d.specularF0, f0b);
    const H3 f0c = h3(m.coatColor) * (clampf(m.coatWeight, 0.0f, 1.0f) * f0_scalar(m.coatIor));
    st3(d.coatF0, f0c);
    // CoatUnderlayerTint (:1042-1045).
    const float k = clampf(m.coatAffectColor * m.coatWeight, 0.0f, 1.0f);
    const H3 ctint = lerp(H3{1.0f, 1.0f, 1.0f}, clamp3(h3(m.coatColor), 0.0f, 1.0f), k);
    st3(d.coatTint, ctint);
    // Lobe selection weights (:1376-1403).
    const H3 uc = ctint * one_minus_clamped(f0c);
    const H3 ub = one_minus_clamped(f0b);
    d.wDiffuse = std::max(0.0f, max3(diffuse * uc * ub));
    d.wBaseSpecular = std::max(0.0f, max3(uc * f0b));
    d.wCoat = std::max(0.0f, max3(f0c));
    if (trans <= 0.0f || metal >= 0.999f) {
        d.wTransmission = 0.0f;
    } else {
        const H3 tt = clamp3(h3(m.transmissionColor), 0.0f, 1.0f);
        d.wTransmission = std::max(0.0f, max3(tt * uc * ub)) * trans * (1.0f - metal);
    }
    // Roughnesses (:911-916, 930-932).
    d.baseRoughness = clampf(m.roughness + m.coatAffectRoughness * m.coatWeight * m.coatRoughness, 0.0f, 1.0f);
    d.transmissionRoughness = clampf(m.roughness + m.transmissionExtraRoughness, 0.0f, 1.0f);
    d.sampleRoughnessBase = sampling_roughness(d.baseRoughness, m.specularAnisotropy, m.specularRotation);
    d.sampleRoughnessCoat = sampling_roughness(m.coatRoughness, m.coatAnisotropy, m.coatRotation);
    d.sampleRoughnessTrans = sampling_roughness(d.transmissionRoughness, m.specularAnisotropy, m.specularRotation);
    d.sampleScaleBase = sampling_scale(m.specularAnisotropy, m.specularRotation);
    d.sampleScaleTrans = d.sampleScaleBase;
    // Emission (:1728-1733).
    d.emission[0] = std::max(m.emissionColor[0] * m.emissionStrength, 0.0f);
    d.emission[1] = std::max(m.emissionColor[1] * m.emissionStrength, 0.0f);
    d.emission[2] = std::max(m.emissionColor[2] * m.emissionStrength, 0.0f);
    // MediumExtinction (:1420-1437).
    H3 ext{std::max(m.transmissionScatter[0], 0.0f), std::max(m.transmissionScatter[1], 0.0f),
           std::max(m.transmissionScatter[2], 0.0f)};
    if (!m.thinWalled && m.transmissionDepth > 0.0f) {
        const float depth = std::max(m.transmissionDepth, 1.0e-3f);
        const H3 t = clamp3(h3(m.transmissionColor), 1.0e-4f, 1.0f);
        ext.x += -std::log(t.x) / depth;
        ext.y += -std::log(t.y) / depth;
        ext.z += -std::log(t.z) / depth;
    }
    st3(d.mediumExtinction, ext);
    // TransmissionSurfaceTint (:1405-1410).
    if (!m.thinWalled && m.transmissionDepth > 0.0f) {
        st3(d.transmissionTint, H3{1.0f, 1.0f, 1.0f});
    } else {
        st3(d.transmissionTint, clamp3(h3(m.transmissionColor), 0.0f, 1.0f));
    }
    uint32_t f = 0;
    if (clampf(m.specularAnisotropy, -0.95f, 0.95f) != 0.0f || clampf(m.coatAnisotropy, -0.95f, 0.95f) != 0.0f)
        f |= kMatAnisotropic;
    if (m.thinFilmThickness > 0.0f) f |= kMatThinFilm;
    if (m.sheen > 0.0f) f |= kMatSheen;
    if (m.hasNormalOverride) f |= kMatNormalOverride;
    if (m.hasCoatNormalOverride) f |= kMatCoatNormalOverride;
    if (m.hasTangentOverride) f |= kMatTangentOverride;
    if (m.emissionStrength > 0.0f) f |= kMatEmissive;
    if (clampf(m.opacity, 0.0f, 1.0f) < 0.999f) f |= kMatOpacity;
    if (m.baseColorTexture >= 0 || m.roughnessTexture >= 0) f |= kMatTextured;
    d.flags = f;
    return d;
}

void assign_shade_classes(const shiro_hip_material* m, MaterialDerived* d, size_t n) {
    std::vector<uint32_t> seen;
    for (size_t i = 0; i < n; ++i) {
        const MaterialDerived& x = d[i];
        uint32_t sig = x.flags;
        sig |= (x.wDiffuse > 0.0f ? 1u : 0u) << 16;
        sig |= (x.wBaseSpecular > 0.0f ? 1u : 0u) << 17;
        sig |= (x.wCoat > 0.0f ? 1u : 0u) << 18;
        sig |= (x.wTransmission > 0.0f ? 1u : 0u) << 19;
        sig |= (x.sampleRoughnessBase <= 0.02f ? 1u : 0u) << 20;
        sig |= (x.sampleRoughnessCoat <= 0.02f ? 1u : 0u) << 21;
        sig |= (x.sampleRoughnessTrans <= 0.02f ? 1u : 0u) << 22;
        sig |= (m[i].thinWalled ? 1u : 0u) << 23;
        size_t k = 0;
        while (k < seen.size() && seen[k] != sig) ++k;
        if (k == seen.size()) seen.push_back(sig);
        d[i].shadeClass = static_cast<uint32_t>(std::min<size_t>(k, kShadeMissClass - 1u));
    }
}

float env_row_sine(uint32_t y, uint32_t h) {
    const float theta = (static_cast<float>(y) + 0.5f) * kPiH / static_cast<float>(h);
    return std::max(std::sin(theta), 1.0e-6f);
}

bool build_env_tables(const float* px, uint32_t w, uint32_t h, uint32_t layout, std::vector<float>* rowCdf,
                      std::vector<float>* condCdf) {
    rowCdf->clear();
    condCdf->clear();
    if (!px || w == 0 || h == 0 || layout != SHIRO_HIP_ENV_LATLONG) return false;
    rowCdf->assign(h + 1u, 0.0f);
    condCdf->assign(static_cast<size_t>(h) * (w + 1u), 0.0f);
    float total = 0.0f;
    for (uint32_t y = 0; y < h; ++y) {
        const float st = env_row_sine(y, h);
        float* row = condCdf->data() + static_cast<size_t>(y) * (w + 1u);
        float sum = 0.0f;
        for (uint32_t x = 0; x < w; ++x) {
            const float* c = px + 3 * (static_cast<size_t>(y) * w + x);
            const float lum = c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f;
            sum += std::max(0.0f, lum) * st;
            row[x + 1u] = sum;
        }
        if (sum > 0.0f) {
            for (uint32_t x = 1; x <= w; ++x) row[x] /= sum;
        } else {
            for (uint32_t x = 1; x <= w; ++x) row[x] = static_cast<float>(x) / static_cast<float>(w);
            sum = 1.0f;
        }
        total += sum;
        (*rowCdf)[y + 1u] = total;
    }
    if (total <= 0.0f) {
        for (uint32_t y = 1; y <= h; ++y) (*rowCdf)[y] = static_cast<float>(y) / static_cast<float>(h);
        return true;
    }
    for (uint32_t y = 1; y <= h; ++y) (*rowCdf)[y] /= total;
    return true;
}

void build_cdf_guide(const float* cdf, uint32_t begin, uint32_t end, uint32_t G, uint32_t* out) {
    bool monotone = true;
    for (uint32_t i = begin; i < end && monotone; ++i) monotone = cdf[i] <= cdf[i + 1u];
    for (uint32_t g = 0; g <= G; ++g) {
        if (!monotone) {
            out[g] = g == G ? end + 1u : begin + 1u;
            continue;
        }
        const float key = static_cast<float>(g) / static_cast<float>(G);
        uint32_t first = begin + 1u, count = end - begin;
        while (count > 0u) {
            const uint32_t half = count >> 1;
            const uint32_t mid = first + half;
            if (cdf[mid] < key) {
                first = mid + 1u;
                count -= half + 1u;
            } else {
                count = half;
            }
        }
        out[g] = first;
    }
}

void build_emissive(const shiro_hip_scene& s, std::vector<EmissiveEntry>* out, float* sum) {
    out->clear();
    *sum = 0.0f;
    uint32_t base = 0;  // global index of the mesh's first triangle
    for (uint32_t mi = 0; mi < s.meshCount; ++mi) {
        const shiro_hip_mesh& m = s.meshes[mi];
        const uint32_t nt = m.indexCount / 3u;
        const uint32_t first = base;
        base += nt;
        if (m.indexCount < 3u || m.indexCount % 3u != 0u || !m.positions || !m.indices) continue;
        const shiro_hip_material mat = m.materialIndex < s.materialCount ? s.materials[m.materialIndex] : default_material();
        const H3 rad{std::max(mat.emissionColor[0] * mat.emissionStrength, 0.0f),
                     std::max(mat.emissionColor[1] * mat.emissionStrength, 0.0f),
                     std::max(mat.emissionColor[2] * mat.emissionStrength, 0.0f)};
        const float rw = std::max(0.2126f * rad.x + 0.7152f * rad.y + 0.0722f * rad.z, 0.0f);
        if (rw <= 0.0f) continue;
        for (uint32_t p = 0; p < nt; ++p) {
            const uint32_t* ix = m.indices + 3u * p;
            if (ix[0] >= m.vertexCount || ix[1] >= m.vertexCount || ix[2] >= m.vertexCount) continue;
            const H3 p0 = h3(m.positions + 3u * ix[0]), p1 = h3(m.positions + 3u * ix[1]),
                     p2 = h3(m.positions + 3u * ix[2]);
            const H3 e1 = p1 - p0, e2 = p2 - p0;
            const H3 av{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
            const float twice = std::sqrt(av.x * av.x + av.y * av.y + av.z * av.z);
            if (twice <= 1.0e-8f) continue;
            EmissiveEntry e;
            e.tri = first + p;
            e.area = twice * 0.5f;
            e.weight = e.area * rw;
            *sum += e.weight;
            e.cumulative = *sum;
            e.normal[0] = av.x / twice;
            e.normal[1] = av.y / twice;
            e.normal[2] = av.z / twice;
            st3(e.radiance, rad);
            st3(e.p, p0);
            st3(e.p + 3, p1);
            st3(e.p + 6, p2);
            out->push_back(e);
        }
    }
}

// True if the BSDF evaluation of m needs none of its optional terms: no
// specular or coat anisotropy, no thin film, no sheen, and a subsurface wrap
// of exactly 0 (subsurface 0 with finite wrap parameters), so that the
// device's plain evaluation gives the general one's results.
bool plain_bsdf(const shiro_hip_material& m) {
    const bool finite = std::isfinite(m.subsurfaceScale) && std::isfinite(m.subsurfaceAnisotropy) &&
                        std::isfinite(m.subsurfaceRadius[0]) && std::isfinite(m.subsurfaceRadius[1]) &&
                        std::isfinite(m.subsurfaceRadius[2]);
    return m.specularAnisotropy == 0.0f && m.coatAnisotropy == 0.0f && m.thinFilmThickness <= 0.0f &&
           !(m.sheen > 0.0f) && m.subsurface == 0.0f && finite;
}

bool flatten_scene(const shiro_hip_scene& s, FlatScene* out, std::string* err) {
    *out = FlatScene();
    if (s.materialCount > 0 && !s.materials) {
        *err = "materials is null";
        return false;
    }
    if (s.meshCount > 0 && !s.meshes) {
        *err = "meshes is null";
        return false;
    }
    if (s.distantLightCount > kMaxSceneDirLights) {
        *err = "too many distant lights (max " + std::to_string(kMaxSceneDirLights) + ")";
        return false;
    }
    if (s.domeLightCount > kMaxSceneDomes) {
        *err = "too many dome lights (max " + std::to_string(kMaxSceneDomes) + ")";
        return false;
    }
    if (s.distantLightCount > 0 && !s.distantLights) {
        *err = "distantLights is null";
        return false;
    }
    if (s.domeLightCount > 0 && !s.domeLights) {
        *err = "domeLights is null";
        return false;
    }
    size_t total = 0;
    for (uint32_t i = 0; i < s.meshCount; ++i) total += s.meshes[i].indexCount / 3u;
    out->tri.reserve(total * 9);
    out->material.reserve(total);
    out->hasNormals.reserve(total);
    out->nrm.reserve(total * 9);
    for (uint32_t mi = 0; mi < s.meshCount; ++mi)
        if (s.meshes[mi].uvs && s.meshes[mi].indexCount >= 3u) out->hasUvs = true;
    if (out->hasUvs) out->uv.reserve(total * 6);
    if (s.textureCount > 0 && !s.textures) {
        *err = "textures is null";
        return false;
    }
    for (uint32_t ti = 0; ti < s.textureCount; ++ti) {
        const shiro_hip_texture& t = s.textures[ti];
        if (!t.texels || t.width == 0 || t.height == 0 || t.channels == 0 || t.channels > 4) {
            *err = "texture " + std::to_string(ti) + " is empty or has 0 or more than 4 channels";
            return false;
        }
    }
    for (uint32_t i = 0; i < s.materialCount; ++i) {
        const shiro_hip_material& m = s.materials[i];
        if (m.baseColorTexture >= static_cast<int32_t>(s.textureCount) ||
            m.roughnessTexture >= static_cast<int32_t>(s.textureCount)) {
            *err = "material " + std::to_string(i) + " names a missing texture";
            return false;
        }
        if (m.baseColorTexture >= 0 || m.roughnessTexture >= 0) out->textured = true;
        if (!plain_bsdf(m)) out->plainBsdf = false;
    }
    if (!plain_bsdf(default_material())) out->plainBsdf = false;
    for (uint32_t mi = 0; mi < s.meshCount; ++mi) {
        const shiro_hip_mesh& m = s.meshes[mi];
        const uint32_t nt = m.indexCount / 3u;
        if (nt == 0) continue;
        if (!m.positions || !m.indices) {
            *err = "mesh " + std::to_string(mi) + " has null positions or indices";
            return false;
        }
        const bool normals = m.normals && m.normalCount == m.vertexCount;
        const uint32_t mat = m.materialIndex < s.materialCount ? m.materialIndex : s.materialCount;
        const shiro_hip_material mm = m.materialIndex < s.materialCount ? s.materials[m.materialIndex] : default_material();
        if (mm.emissionStrength > 0.0f) out->emissive = true;
        for (uint32_t p = 0; p < nt; ++p) {
            const uint32_t* ix = m.indices + 3u * p;
            for (int k = 0; k < 3; ++k) {
                if (ix[k] >= m.vertexCount) {
                    *err = "mesh " + std::to_string(mi) + " index out of range";
                    return false;
                }
                const float* v = m.positions + 3u * ix[k];
                out->tri.insert(out->tri.end(), v, v + 3);
                if (out->hasUvs) {
                    if (m.uvs) {
                        out->uv.push_back(m.uvs[2u * ix[k]]);
                        out->uv.push_back(m.uvs[2u * ix[k] + 1u]);
                    } else {
                        out->uv.insert(out->uv.end(), {0.0f, 0.0f});
                    }
                }
                if (normals) {
                    const float* n = m.normals + 3u * ix[k];
                    out->nrm.insert(out->nrm.end(), n, n + 3);
                } else {
                    out->nrm.insert(out->nrm.end(), {0.0f, 0.0f, 0.0f});
                }
            }
            out->hasNormals.push_back(normals ? 1 : 0);
            out->material.push_back(mat);
        }
    }
    return true;
}

}  // namespace shiro_hip
