// HipBackend over libshiro_hip.so; see include/shiro/backend/hip/HipBackend.h.
// Compiled inside a Shiro tree (it uses Shiro's render types).
#include "shiro/backend/hip/HipBackend.h"

#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "shiro/render/EnvironmentMap.h"
#include "shiro_hip.h"

namespace shiro::backend::hip {

using namespace shiro::render;

namespace {

void Set3(float* d, const Vec3f& v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}

shiro_hip_material ToMaterial(const PbrMaterial& m) {
    shiro_hip_material o;
    shiro_hip_default_material(&o);
    Set3(o.baseColor, m.baseColor);
    Set3(o.emissionColor, m.emissionColor);
    Set3(o.specularColor, m.specularColor);
    Set3(o.transmissionColor, m.transmissionColor);
    Set3(o.transmissionScatter, m.transmissionScatter);
    Set3(o.coatColor, m.coatColor);
    Set3(o.subsurfaceColor, m.subsurfaceColor);
    Set3(o.subsurfaceRadius, m.subsurfaceRadius);
    Set3(o.sheenColor, m.sheenColor);
    Set3(o.normalOverride, m.normalOverride);
    Set3(o.coatNormalOverride, m.coatNormalOverride);
    Set3(o.tangentOverride, m.tangentOverride);
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
    o.thinWalled = m.thinWalled ? 1u : 0u;
    o.hasNormalOverride = m.hasNormalOverride ? 1u : 0u;
    o.hasCoatNormalOverride = m.hasCoatNormalOverride ? 1u : 0u;
    o.hasTangentOverride = m.hasTangentOverride ? 1u : 0u;
    return o;
}

// C views of a Scene; the vectors own the converted arrays.
struct SceneView {
    std::vector<shiro_hip_material> materials;
    std::vector<shiro_hip_mesh> meshes;
    std::vector<shiro_hip_directional_light> dirs;
    std::vector<shiro_hip_dome_light> domes;
    std::vector<std::vector<float>> domePixels;
    shiro_hip_scene scene{};
};

void BuildView(const Scene& s, SceneView* v) {
    for (const PbrMaterial& m : s.materials) v->materials.push_back(ToMaterial(m));
    for (const TriangleMesh& m : s.meshes) {
        shiro_hip_mesh c{};
        c.positions = m.positions.empty() ? nullptr : &m.positions[0].x;
        c.normals = m.normals.empty() ? nullptr : &m.normals[0].x;
        c.uvs = nullptr;
        c.indices = m.indices.empty() ? nullptr : m.indices.data();
        c.vertexCount = static_cast<uint32_t>(m.positions.size());
        c.normalCount = static_cast<uint32_t>(m.normals.size());
        c.indexCount = static_cast<uint32_t>(m.indices.size());
        c.materialIndex = m.materialIndex;
        v->meshes.push_back(c);
    }
    for (const DirectionalLight& d : s.distantLights) {
        shiro_hip_directional_light c{};
        Set3(c.direction, d.direction);
        Set3(c.radiance, d.radiance);
        v->dirs.push_back(c);
    }
    v->domePixels.resize(s.domeLights.size());
    for (size_t i = 0; i < s.domeLights.size(); ++i) {
        const DomeLight& d = s.domeLights[i];
        shiro_hip_dome_light c{};
        Set3(c.radiance, d.radiance);
        Set3(c.right, d.right);
        Set3(c.up, d.up);
        Set3(c.forward, d.forward);
        if (d.environment && !d.environment->Pixels().empty()) {
            const auto& px = d.environment->Pixels();
            v->domePixels[i].resize(px.size() * 3);
            for (size_t p = 0; p < px.size(); ++p) {
                v->domePixels[i][3 * p] = px[p].x;
                v->domePixels[i][3 * p + 1] = px[p].y;
                v->domePixels[i][3 * p + 2] = px[p].z;
            }
            c.pixels = v->domePixels[i].data();
            c.width = d.environment->Width();
            c.height = d.environment->Height();
            c.layout = d.environment->Layout() == EnvironmentMapLayout::Angular ? SHIRO_HIP_ENV_ANGULAR
                                                                                 : SHIRO_HIP_ENV_LATLONG;
        }
        v->domes.push_back(c);
    }
    shiro_hip_scene& o = v->scene;
    o.materials = v->materials.data();
    o.materialCount = static_cast<uint32_t>(v->materials.size());
    o.meshes = v->meshes.data();
    o.meshCount = static_cast<uint32_t>(v->meshes.size());
    o.distantLights = v->dirs.data();
    o.distantLightCount = static_cast<uint32_t>(v->dirs.size());
    o.domeLights = v->domes.data();
    o.domeLightCount = static_cast<uint32_t>(v->domes.size());
    Set3(o.environmentTop, s.environmentTop);
    Set3(o.environmentBottom, s.environmentBottom);
}

// FNV-1a over the scene content, like the reference's scene signature
// (OptixBackend.cpp:150-232): a re-upload happens only when it changes.
uint64_t Signature(const Scene& s) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    };
    for (const PbrMaterial& m : s.materials) {
        const shiro_hip_material c = ToMaterial(m);
        mix(&c, sizeof(c));
    }
    for (const TriangleMesh& m : s.meshes) {
        mix(m.positions.data(), m.positions.size() * sizeof(Vec3f));
        mix(m.normals.data(), m.normals.size() * sizeof(Vec3f));
        mix(m.indices.data(), m.indices.size() * sizeof(uint32_t));
        mix(&m.materialIndex, sizeof(m.materialIndex));
    }
    for (const DirectionalLight& d : s.distantLights) mix(&d, sizeof(d));
    for (const DomeLight& d : s.domeLights) {
        mix(&d.radiance, sizeof(Vec3f));
        mix(&d.right, sizeof(Vec3f) * 3);
        const void* env = d.environment.get();
        mix(&env, sizeof(env));
    }
    mix(&s.environmentTop, sizeof(Vec3f));
    mix(&s.environmentBottom, sizeof(Vec3f));
    return h;
}

// FNV-1a over the meshes' vertex counts and index arrays: scenes that agree
// on it differ only in vertex data, materials or lights, so the device BVH
// can be refitted instead of rebuilt.
uint64_t Topology(const Scene& s) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    };
    const size_t meshes = s.meshes.size();
    mix(&meshes, sizeof(meshes));
    for (const TriangleMesh& m : s.meshes) {
        const size_t n = m.positions.size();
        mix(&n, sizeof(n));
        mix(m.indices.data(), m.indices.size() * sizeof(uint32_t));
    }
    return h;
}

}  // namespace

struct HipBackend::RuntimeState {
    shiro_hip_ctx* ctx = nullptr;
    uint64_t signature = 0;
    uint64_t topology = 0;
    bool hasScene = false;
    ~RuntimeState() {
        if (ctx) shiro_hip_destroy(ctx);
    }
};

HipBackend::HipBackend() = default;
HipBackend::~HipBackend() = default;

bool HipBackend::EnsureRuntime() const {
    if (runtime_) return true;
    if (runtimeAttempted_) return false;
    runtimeAttempted_ = true;
    auto rt = std::make_unique<RuntimeState>();
    const int rc = shiro_hip_create(0, &rt->ctx);
    if (rc != SHIRO_HIP_OK) {
        std::fprintf(stderr, "shiro hip backend unavailable (code %d)\n", rc);
        return false;
    }
    runtime_ = std::move(rt);
    return true;
}

BackendStatus HipBackend::GetStatus() const {
    std::scoped_lock lock(mutex_);
    BackendStatus status;
    status.kind = BackendKind::Gpu;
    status.name = "hip";
    status.usesGpu = true;
    status.available = EnsureRuntime();
    return status;
}

bool HipBackend::SupportsScene(const Scene& scene) const {
    SceneView view;
    BuildView(scene, &view);
    return shiro_hip_supports_scene(&view.scene) != 0;
}

FrameBuffer HipBackend::RenderSampleBatch(const RenderRequest& request) const {
    std::scoped_lock lock(mutex_);
    FrameBuffer frame;
    frame.Resize(request.settings.width, request.settings.height);
    frame.Clear();
    if (request.sampleCount == 0 || !EnsureRuntime()) return frame;
    const uint64_t sig = Signature(request.scene);
    if (!runtime_->hasScene || sig != runtime_->signature) {
        SceneView view;
        BuildView(request.scene, &view);
        // Moved vertices on unchanged meshes (deformation): refit the BVH;
        // anything else, or a failed refit: upload and rebuild.
        const uint64_t topo = Topology(request.scene);
        const bool refitted = runtime_->hasScene && topo == runtime_->topology &&
                              shiro_hip_refit_scene(runtime_->ctx, &view.scene) == SHIRO_HIP_OK;
        if (!refitted && shiro_hip_upload_scene(runtime_->ctx, &view.scene) != SHIRO_HIP_OK) {
            std::fprintf(stderr, "shiro hip upload failed: %s\n", shiro_hip_last_error(runtime_->ctx));
            runtime_->hasScene = false;
            return frame;
        }
        runtime_->signature = sig;
        runtime_->topology = topo;
        runtime_->hasScene = true;
    }
    const Camera& c = request.camera;
    shiro_hip_camera cam{};
    Set3(cam.position, c.position);
    Set3(cam.forward, c.forward);
    Set3(cam.right, c.right);
    Set3(cam.up, c.up);
    cam.verticalFovDegrees = c.verticalFovDegrees;
    cam.aspectRatio = c.aspectRatio;
    const RenderSettings& r = request.settings;
    shiro_hip_settings st;
    shiro_hip_default_settings(&st);
    st.width = r.width;
    st.height = r.height;
    st.samplesPerPixel = r.samplesPerPixel;
    st.samplesPerUpdate = r.samplesPerUpdate;
    st.domeLightSamples = r.domeLightSamples;
    st.maxDepth = r.maxDepth;
    st.diffuseDepth = r.diffuseDepth;
    st.specularDepth = r.specularDepth;
    st.threadLimit = r.threadLimit;
    st.maxSubdivLevel = r.maxSubdivLevel;
    st.backgroundVisible = r.backgroundVisible ? 1u : 0u;
    st.enableHeadlight = r.enableHeadlight ? 1u : 0u;
    st.enableCaustics = r.enableCaustics ? 1u : 0u;
    st.enableSss = r.enableSss ? 1u : 0u;
    const size_t np = static_cast<size_t>(r.width) * r.height;
    std::vector<float> beauty(np * 4), albedo(np * 3), normal(np * 3), depth(np);
    shiro_hip_frame out{beauty.data(), albedo.data(), normal.data(), depth.data(), r.width, r.height};
    // The library polls the request's std::atomic<bool> cancel flag between
    // launches through its one-byte object representation.
    static_assert(sizeof(std::atomic<bool>) == 1 && std::atomic<bool>::is_always_lock_free, "atomic<bool> layout");
    const volatile uint8_t* cancel =
        request.cancel ? reinterpret_cast<const volatile uint8_t*>(request.cancel) : nullptr;
    const int rc = shiro_hip_render_batch(runtime_->ctx, &cam, &st, request.sampleStart, request.sampleCount, cancel,
                                          &out);
    if (rc != SHIRO_HIP_OK) {
        std::fprintf(stderr, "shiro hip render failed: %s\n", shiro_hip_last_error(runtime_->ctx));
        return frame;
    }
    if (request.cancel && request.cancel->load(std::memory_order_relaxed)) return frame;
    for (uint32_t y = 0; y < r.height; ++y) {
        for (uint32_t x = 0; x < r.width; ++x) {
            const size_t p = static_cast<size_t>(y) * r.width + x;
            frame.SetBeauty(x, y, {beauty[4 * p], beauty[4 * p + 1], beauty[4 * p + 2], beauty[4 * p + 3]});
            frame.SetAlbedo(x, y, {albedo[3 * p], albedo[3 * p + 1], albedo[3 * p + 2]});
            frame.SetNormal(x, y, {normal[3 * p], normal[3 * p + 1], normal[3 * p + 2]});
            frame.SetDepth(x, y, depth[p]);
        }
    }
    return frame;
}

}  // namespace shiro::backend::hip
