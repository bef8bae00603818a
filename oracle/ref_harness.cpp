// TEST INFRASTRUCTURE ONLY -- builds oracle/_ref/libshiro_ref.so.
//
// A C ABI over the reference's own CPU path-tracer sources, compiled where
// they lie under the reference checkout (see oracle/Makefile; nothing of the
// reference is copied into this repository).  The reference is built in its
// documented no-dependency configuration (SHIRO_HAVE_EMBREE=0,
// SHIRO_HAVE_OIIO=0, SHIRO_WITH_OPENQMC=0, no CUDA/OptiX): intersection is
// its brute-force Moller-Trumbore path and the sampler is its fallback
// branch.  The reference .cpp files are included into this translation unit
// so that their internal-linkage functions (EvaluateBsdf, LowDiscrepancySample,
// ...) can be called directly.
//
// EnvironmentMap has a private constructor and is normally created only by
// Load() through OpenImageIO, which is not installed.  The access specifiers
// of that one header are lifted with a preprocessor define so that the
// harness can fill a map from a pixel array and run the reference's own
// BuildImportanceTable / Sample / Pdf / Evaluate on it.

#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <string_view>
#include <thread>
#include <utility>
#include <vector>

#include "shiro/core/Config.h"
#include "shiro/render/Types.h"
#include "shiro/render/FrameBuffer.h"
#include "shiro/render/Renderer.h"
#include "shiro/backend/RenderBackend.h"
#include "shiro/backend/cpu/CpuPathTracer.h"
#include "shiro/render/OpenQmcSampler.h"

#define private public
#include "shiro/render/EnvironmentMap.h"
#undef private

#include "src/render/OpenQmcSampler.cpp"
#include "src/render/EnvironmentMap.cpp"
#include "src/render/FrameBuffer.cpp"
#include "src/backend/cpu/CpuPathTracer.cpp"

#include "../include/shiro_hip.h"

using namespace shiro::render;

namespace {

Vec3f V(const float* p) { return {p[0], p[1], p[2]}; }

PbrMaterial ToMaterial(const shiro_hip_material& m) {
    PbrMaterial o;
    o.baseColor = V(m.baseColor);
    o.emissionColor = V(m.emissionColor);
    o.specularColor = V(m.specularColor);
    o.transmissionColor = V(m.transmissionColor);
    o.transmissionScatter = V(m.transmissionScatter);
    o.coatColor = V(m.coatColor);
    o.subsurfaceColor = V(m.subsurfaceColor);
    o.subsurfaceRadius = V(m.subsurfaceRadius);
    o.sheenColor = V(m.sheenColor);
    o.normalOverride = V(m.normalOverride);
    o.coatNormalOverride = V(m.coatNormalOverride);
    o.tangentOverride = V(m.tangentOverride);
    o.baseWeight = m.baseWeight;
    o.emissionStrength = m.emissionStrength;
    o.specularWeight = m.specularWeight;
    o.metallic = m.metallic;
    o.roughness = m.roughness;
    o.diffuseRoughness = m.diffuseRoughness;
    o.opacity = m.opacity;
    o.specularAnisotropy = m.specularAnisotropy;
    o.specularRotation = m.specularRotation;
    o.coatWeight = m.coatWeight;
    o.coatRoughness = m.coatRoughness;
    o.coatIor = m.coatIor;
    o.coatAnisotropy = m.coatAnisotropy;
    o.coatRotation = m.coatRotation;
    o.sheen = m.sheen;
    o.sheenRoughness = m.sheenRoughness;
    o.subsurface = m.subsurface;
    o.subsurfaceScale = m.subsurfaceScale;
    o.subsurfaceAnisotropy = m.subsurfaceAnisotropy;
    o.transmission = m.transmission;
    o.transmissionDepth = m.transmissionDepth;
    o.transmissionScatterAnisotropy = m.transmissionScatterAnisotropy;
    o.transmissionDispersion = m.transmissionDispersion;
    o.transmissionExtraRoughness = m.transmissionExtraRoughness;
    o.ior = m.ior;
    o.coatAffectColor = m.coatAffectColor;
    o.coatAffectRoughness = m.coatAffectRoughness;
    o.thinFilmThickness = m.thinFilmThickness;
    o.thinFilmIor = m.thinFilmIor;
    o.thinWalled = m.thinWalled != 0;
    o.hasNormalOverride = m.hasNormalOverride != 0;
    o.hasCoatNormalOverride = m.hasCoatNormalOverride != 0;
    o.hasTangentOverride = m.hasTangentOverride != 0;
    return o;
}

std::shared_ptr<EnvironmentMap> MakeEnv(const shiro_hip_dome_light& d) {
    auto env = std::shared_ptr<EnvironmentMap>(new EnvironmentMap());
    env->width_ = d.width;
    env->height_ = d.height;
    env->layout_ = ResolveLayout(static_cast<EnvironmentMapLayout>(d.layout));
    env->pixels_.resize(static_cast<size_t>(d.width) * d.height);
    for (size_t p = 0; p < env->pixels_.size(); ++p) env->pixels_[p] = V(d.pixels + 3 * p);
    env->BuildImportanceTable();
    return env;
}

struct RefScene {
    Scene scene;
    shiro::backend::cpu::CpuPathTracer tracer;
};

RenderSettings ToSettings(const shiro_hip_settings& s) {
    RenderSettings o;
    o.width = s.width;
    o.height = s.height;
    o.samplesPerPixel = s.samplesPerPixel;
    o.samplesPerUpdate = s.samplesPerUpdate;
    o.domeLightSamples = s.domeLightSamples;
    o.maxDepth = s.maxDepth;
    o.diffuseDepth = s.diffuseDepth;
    o.specularDepth = s.specularDepth;
    o.threadLimit = s.threadLimit;
    o.maxSubdivLevel = s.maxSubdivLevel;
    o.backend = BackendKind::Cpu;
    o.backgroundVisible = s.backgroundVisible != 0;
    o.enableHeadlight = s.enableHeadlight != 0;
    o.enableCaustics = s.enableCaustics != 0;
    o.enableSss = s.enableSss != 0;
    return o;
}

Camera ToCamera(const shiro_hip_camera& c) {
    Camera o;
    o.position = V(c.position);
    o.forward = V(c.forward);
    o.right = V(c.right);
    o.up = V(c.up);
    o.verticalFovDegrees = c.verticalFovDegrees;
    o.aspectRatio = c.aspectRatio;
    return o;
}

void WriteFrame(const FrameBuffer& f, shiro_hip_frame* out) {
    const size_t n = f.Beauty().size();
    for (size_t p = 0; p < n; ++p) {
        out->beauty[4 * p + 0] = f.Beauty()[p].x;
        out->beauty[4 * p + 1] = f.Beauty()[p].y;
        out->beauty[4 * p + 2] = f.Beauty()[p].z;
        out->beauty[4 * p + 3] = f.Beauty()[p].w;
        out->albedo[3 * p + 0] = f.Albedo()[p].x;
        out->albedo[3 * p + 1] = f.Albedo()[p].y;
        out->albedo[3 * p + 2] = f.Albedo()[p].z;
        out->normal[3 * p + 0] = f.Normal()[p].x;
        out->normal[3 * p + 1] = f.Normal()[p].y;
        out->normal[3 * p + 2] = f.Normal()[p].z;
        out->depth[p] = f.Depth()[p];
    }
}

}  // namespace

extern "C" {

float ref_qmc(uint32_t px, uint32_t py, uint32_t s, uint32_t dim) { return LowDiscrepancySample(px, py, s, dim); }

float ref_radical_inverse(uint32_t base, uint32_t n) { return RadicalInverse(base, n); }

void ref_next2d(uint32_t px, uint32_t py, uint32_t s, uint32_t branch, float* out) {
    OpenQmcSampler smp({px, py, s});
    const Vec2f v = smp.Next2D(branch);
    out[0] = v.x;
    out[1] = v.y;
}

void* ref_scene_create(const shiro_hip_scene* in) {
    auto* r = new RefScene();
    Scene& s = r->scene;
    for (uint32_t i = 0; i < in->materialCount; ++i) s.materials.push_back(ToMaterial(in->materials[i]));
    for (uint32_t i = 0; i < in->meshCount; ++i) {
        const shiro_hip_mesh& m = in->meshes[i];
        TriangleMesh tm;
        for (uint32_t v = 0; v < m.vertexCount; ++v) tm.positions.push_back(V(m.positions + 3 * v));
        if (m.normals)
            for (uint32_t v = 0; v < m.normalCount; ++v) tm.normals.push_back(V(m.normals + 3 * v));
        tm.indices.assign(m.indices, m.indices + m.indexCount);
        tm.materialIndex = m.materialIndex;
        s.meshes.push_back(std::move(tm));
    }
    for (uint32_t i = 0; i < in->distantLightCount; ++i)
        s.distantLights.push_back({V(in->distantLights[i].direction), V(in->distantLights[i].radiance)});
    for (uint32_t i = 0; i < in->domeLightCount; ++i) {
        const shiro_hip_dome_light& d = in->domeLights[i];
        DomeLight dl;
        dl.radiance = V(d.radiance);
        dl.right = V(d.right);
        dl.up = V(d.up);
        dl.forward = V(d.forward);
        dl.layout = static_cast<EnvironmentMapLayout>(d.layout);
        if (d.pixels && d.width > 0 && d.height > 0) dl.environment = MakeEnv(d);
        s.domeLights.push_back(dl);
    }
    s.environmentTop = V(in->environmentTop);
    s.environmentBottom = V(in->environmentBottom);
    return r;
}

void ref_scene_destroy(void* h) { delete static_cast<RefScene*>(h); }

// CpuPathTracer::RenderSampleBatch (CpuPathTracer.cpp:2043-2170).
int ref_render(void* h, const shiro_hip_camera* camera, const shiro_hip_settings* settings, uint32_t sampleStart,
               uint32_t sampleCount, shiro_hip_frame* out) {
    RefScene& r = *static_cast<RefScene*>(h);
    const RenderSettings st = ToSettings(*settings);
    const Camera cam = ToCamera(*camera);
    shiro::backend::RenderRequest req{r.scene, cam, st, sampleStart, sampleCount, nullptr};
    const FrameBuffer f = r.tracer.RenderSampleBatch(req);
    WriteFrame(f, out);
    return 0;
}

// Renderer::FrameAccumulator over a sequence of batches rendered by the
// reference (Renderer.cpp:47-110): batches[i] samples each, in order.
int ref_render_accumulated(void* h, const shiro_hip_camera* camera, const shiro_hip_settings* settings,
                           const uint32_t* batches, uint32_t batchCount, shiro_hip_frame* out) {
    RefScene& r = *static_cast<RefScene*>(h);
    const RenderSettings st = ToSettings(*settings);
    const Camera cam = ToCamera(*camera);
    Renderer::FrameAccumulator acc;
    acc.Reset(st.width, st.height);
    uint32_t start = 0;
    for (uint32_t b = 0; b < batchCount; ++b) {
        shiro::backend::RenderRequest req{r.scene, cam, st, start, batches[b], nullptr};
        const FrameBuffer f = r.tracer.RenderSampleBatch(req);
        acc.Accumulate(f, batches[b]);
        start += batches[b];
    }
    WriteFrame(acc.Resolve(), out);
    return 0;
}

void ref_eval_bsdf(const shiro_hip_material* mats, const uint32_t* matIndex, const float* n, const float* cn,
                   const float* v, const float* l, uint32_t count, float* value, float* pdf) {
    for (uint32_t i = 0; i < count; ++i) {
        const PbrMaterial m = ToMaterial(mats[matIndex[i]]);
        const Vec3f f = shiro::backend::cpu::EvaluateBsdf(m, V(n + 3 * i), V(cn + 3 * i), V(v + 3 * i), V(l + 3 * i));
        value[3 * i] = f.x;
        value[3 * i + 1] = f.y;
        value[3 * i + 2] = f.z;
        pdf[i] = shiro::backend::cpu::EvaluateBsdfPdf(m, V(n + 3 * i), V(cn + 3 * i), V(v + 3 * i), V(l + 3 * i));
    }
}

void ref_lobe_weights(const shiro_hip_material* mat, float* w4, float* diffuse3) {
    const PbrMaterial m = ToMaterial(*mat);
    w4[0] = shiro::backend::cpu::DiffuseSampleWeight(m);
    w4[1] = shiro::backend::cpu::BaseSpecularSampleWeight(m);
    w4[2] = shiro::backend::cpu::CoatSampleWeight(m);
    w4[3] = shiro::backend::cpu::TransmissionSampleWeight(m);
    const Vec3f d = shiro::backend::cpu::DiffuseColor(m);
    diffuse3[0] = d.x;
    diffuse3[1] = d.y;
    diffuse3[2] = d.z;
}

float ref_ggx_d(float nDotH, float r) { return shiro::backend::cpu::GgxDistribution(nDotH, r); }
float ref_ggx_g2(float nv, float nl, float r) { return shiro::backend::cpu::GgxSmithVisibility(nv, nl, r); }
float ref_fresnel_dielectric(float c, float ei, float et) { return shiro::backend::cpu::FresnelDielectric(c, ei, et); }

void ref_sample_ggx_half(const float* n, float u0, float u1, float r, float* h) {
    const Vec3f o = shiro::backend::cpu::GgxSampleHalfVector(V(n), {u0, u1}, r);
    h[0] = o.x;
    h[1] = o.y;
    h[2] = o.z;
}

void ref_sample_cosine(const float* n, float u0, float u1, float* d) {
    const Vec3f o = shiro::backend::cpu::CosineSampleHemisphere(V(n), {u0, u1});
    d[0] = o.x;
    d[1] = o.y;
    d[2] = o.z;
}

int ref_env_tables(void* h, uint32_t dome, float* rowCdf, float* condCdf) {
    RefScene& r = *static_cast<RefScene*>(h);
    if (dome >= r.scene.domeLights.size() || !r.scene.domeLights[dome].environment) return 1;
    const EnvironmentMap& e = *r.scene.domeLights[dome].environment;
    if (e.RowCdf().empty()) return 2;
    std::copy(e.RowCdf().begin(), e.RowCdf().end(), rowCdf);
    std::copy(e.ConditionalCdf().begin(), e.ConditionalCdf().end(), condCdf);
    return 0;
}

int ref_env_sample(void* h, uint32_t dome, const float* u, uint32_t n, float* dir, float* pdf, float* rgb) {
    RefScene& r = *static_cast<RefScene*>(h);
    if (dome >= r.scene.domeLights.size() || !r.scene.domeLights[dome].environment) return 1;
    const EnvironmentMap& e = *r.scene.domeLights[dome].environment;
    for (uint32_t i = 0; i < n; ++i) {
        Vec3f d;
        float p;
        const Vec3f c = e.Sample({u[2 * i], u[2 * i + 1]}, &d, &p);
        dir[3 * i] = d.x;
        dir[3 * i + 1] = d.y;
        dir[3 * i + 2] = d.z;
        pdf[i] = p;
        rgb[3 * i] = c.x;
        rgb[3 * i + 1] = c.y;
        rgb[3 * i + 2] = c.z;
    }
    return 0;
}

int ref_env_eval(void* h, uint32_t dome, const float* dir, uint32_t n, float* pdf, float* rgb) {
    RefScene& r = *static_cast<RefScene*>(h);
    if (dome >= r.scene.domeLights.size() || !r.scene.domeLights[dome].environment) return 1;
    const EnvironmentMap& e = *r.scene.domeLights[dome].environment;
    for (uint32_t i = 0; i < n; ++i) {
        const Vec3f d = V(dir + 3 * i);
        pdf[i] = e.Pdf(d);
        const Vec3f c = e.Evaluate(d);
        rgb[3 * i] = c.x;
        rgb[3 * i + 1] = c.y;
        rgb[3 * i + 2] = c.z;
    }
    return 0;
}

// IntersectSceneBruteForce (CpuPathTracer.cpp:431-466); prim is the global
// triangle index (meshes concatenated).
void ref_intersect(void* h, const float* rays, uint32_t n, uint32_t* prim, float* t) {
    RefScene& r = *static_cast<RefScene*>(h);
    std::vector<uint32_t> offs(r.scene.meshes.size() + 1, 0);
    for (size_t i = 0; i < r.scene.meshes.size(); ++i)
        offs[i + 1] = offs[i] + static_cast<uint32_t>(r.scene.meshes[i].indices.size() / 3u);
    for (uint32_t i = 0; i < n; ++i) {
        shiro::backend::cpu::Ray ray;
        ray.origin = V(rays + 8 * i);
        ray.direction = V(rays + 8 * i + 3);
        ray.tMin = rays[8 * i + 6];
        ray.tMax = rays[8 * i + 7];
        shiro::backend::cpu::Hit hit;
        if (shiro::backend::cpu::IntersectSceneBruteForce(r.scene, ray, &hit)) {
            prim[i] = offs[hit.meshIndex] + hit.primitiveIndex;
            t[i] = hit.distance;
        } else {
            prim[i] = 0xffffffffu;
            t[i] = FLT_MAX;
        }
    }
}

}  // extern "C"
