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

// Content hash of a scene, like the reference's scene signature
// (OptixBackend.cpp:150-232): a re-upload happens only when it changes.
// FNV-1a over 8-byte words (the tail bytewise), with every vector's length
// mixed in before its contents so that split and merged meshes or lights
// holding the same bytes differ, and the dome maps hashed by size, layout
// and pixels (not by address: a new map may reuse a freed map's address).
class Hasher {
public:
    void Bytes(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            h_ = (h_ ^ w) * 1099511628211ull;
        }
        for (; i < n; ++i) h_ = (h_ ^ b[i]) * 1099511628211ull;
    }
    void Size(size_t n) {
        const uint64_t v = n;
        Bytes(&v, sizeof(v));
    }
    template <typename T>
    void Vector(const std::vector<T>& v) {
        Size(v.size());
        Bytes(v.data(), v.size() * sizeof(T));
    }
    uint64_t Value() const { return h_; }

private:
    uint64_t h_ = 1469598103934665603ull;
};

uint64_t Signature(const Scene& s) {
    Hasher h;
    h.Size(s.materials.size());
    for (const PbrMaterial& m : s.materials) {
        const shiro_hip_material c = ToMaterial(m);
        h.Bytes(&c, sizeof(c));
    }
    h.Size(s.meshes.size());
    for (const TriangleMesh& m : s.meshes) {
        h.Vector(m.positions);
        h.Vector(m.normals);
        h.Vector(m.indices);
        h.Bytes(&m.materialIndex, sizeof(m.materialIndex));
    }
    h.Size(s.distantLights.size());
    for (const DirectionalLight& d : s.distantLights) {
        h.Bytes(&d.direction, sizeof(Vec3f));
        h.Bytes(&d.radiance, sizeof(Vec3f));
    }
    h.Size(s.domeLights.size());
    for (const DomeLight& d : s.domeLights) {
        h.Bytes(&d.radiance, sizeof(Vec3f));
        h.Bytes(&d.right, sizeof(Vec3f));
        h.Bytes(&d.up, sizeof(Vec3f));
        h.Bytes(&d.forward, sizeof(Vec3f));
        const EnvironmentMap* env = d.environment.get();
        if (env && !env->Pixels().empty()) {
            const uint32_t dims[3] = {env->Width(), env->Height(),
                                      env->Layout() == EnvironmentMapLayout::Angular ? 1u : 0u};
            h.Bytes(dims, sizeof(dims));
            h.Vector(env->Pixels());
        } else {
            h.Size(0);
        }
    }
    h.Bytes(&s.environmentTop, sizeof(Vec3f));
    h.Bytes(&s.environmentBottom, sizeof(Vec3f));
    return h.Value();
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

// Batches of at least this many samples (pixels x samples) render with the
// wavefront transport, smaller ones with the megakernel, whose single
// launch per slice has less fixed cost.  Both transports render the same
// frame bit for bit, so the choice changes only the time.
constexpr uint64_t kWavefrontMinSamples = uint64_t{1} << 18;

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

bool HipBackend::PrepareScene(const Scene& scene) const {
    const uint64_t sig = Signature(scene);
    if (runtime_->hasScene && sig == runtime_->signature) return true;
    SceneView view;
    BuildView(scene, &view);
    // Moved vertices on unchanged meshes (deformation): refit the BVH;
    // anything else, or a failed refit: upload and rebuild.
    const uint64_t topo = Topology(scene);
    const bool refitted = runtime_->hasScene && topo == runtime_->topology &&
                          shiro_hip_refit_scene(runtime_->ctx, &view.scene) == SHIRO_HIP_OK;
    if (!refitted && shiro_hip_upload_scene(runtime_->ctx, &view.scene) != SHIRO_HIP_OK) {
        std::fprintf(stderr, "shiro hip upload failed: %s\n", shiro_hip_last_error(runtime_->ctx));
        runtime_->hasScene = false;
        return false;
    }
    runtime_->signature = sig;
    runtime_->topology = topo;
    runtime_->hasScene = true;
    return true;
}

namespace {

shiro_hip_camera ToCamera(const Camera& c) {
    shiro_hip_camera cam{};
    Set3(cam.position, c.position);
    Set3(cam.forward, c.forward);
    Set3(cam.right, c.right);
    Set3(cam.up, c.up);
    cam.verticalFovDegrees = c.verticalFovDegrees;
    cam.aspectRatio = c.aspectRatio;
    return cam;
}

shiro_hip_settings ToSettings(const RenderSettings& r, uint32_t sampleCount) {
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
    const uint64_t samples = static_cast<uint64_t>(r.width) * r.height * sampleCount;
    st.transport = samples >= kWavefrontMinSamples ? SHIRO_HIP_TRANSPORT_WAVEFRONT : SHIRO_HIP_TRANSPORT_MEGAKERNEL;
    return st;
}

// Host AOV buffers in the ABI's layout, copied into a FrameBuffer through
// its setters (the only way to fill one).
struct HostFrame {
    std::vector<float> beauty, albedo, normal, depth;
    shiro_hip_frame view{};
    HostFrame(uint32_t w, uint32_t h) {
        const size_t np = static_cast<size_t>(w) * h;
        beauty.resize(np * 4);
        albedo.resize(np * 3);
        normal.resize(np * 3);
        depth.resize(np);
        view = shiro_hip_frame{beauty.data(), albedo.data(), normal.data(), depth.data(), w, h};
    }
    void CopyTo(FrameBuffer* frame) const {
        for (uint32_t y = 0; y < view.height; ++y) {
            for (uint32_t x = 0; x < view.width; ++x) {
                const size_t p = static_cast<size_t>(y) * view.width + x;
                frame->SetBeauty(x, y, {beauty[4 * p], beauty[4 * p + 1], beauty[4 * p + 2], beauty[4 * p + 3]});
                frame->SetAlbedo(x, y, {albedo[3 * p], albedo[3 * p + 1], albedo[3 * p + 2]});
                frame->SetNormal(x, y, {normal[3 * p], normal[3 * p + 1], normal[3 * p + 2]});
                frame->SetDepth(x, y, depth[p]);
            }
        }
    }
};

// The library polls the request's std::atomic<bool> cancel flag between
// launches through its one-byte object representation.
const volatile uint8_t* CancelByte(const std::atomic<bool>* cancel) {
    static_assert(sizeof(std::atomic<bool>) == 1 && std::atomic<bool>::is_always_lock_free, "atomic<bool> layout");
    return cancel ? reinterpret_cast<const volatile uint8_t*>(cancel) : nullptr;
}

}  // namespace

FrameBuffer HipBackend::RenderSampleBatch(const RenderRequest& request) const {
    std::scoped_lock lock(mutex_);
    FrameBuffer frame;
    frame.Resize(request.settings.width, request.settings.height);
    frame.Clear();
    if (request.sampleCount == 0 || !EnsureRuntime() || !PrepareScene(request.scene)) return frame;
    const shiro_hip_camera cam = ToCamera(request.camera);
    const shiro_hip_settings st = ToSettings(request.settings, request.sampleCount);
    HostFrame out(st.width, st.height);
    const int rc = shiro_hip_render_batch(runtime_->ctx, &cam, &st, request.sampleStart, request.sampleCount,
                                          CancelByte(request.cancel), &out.view);
    if (rc != SHIRO_HIP_OK) {
        if (rc != SHIRO_HIP_CANCELLED)
            std::fprintf(stderr, "shiro hip render failed: %s\n", shiro_hip_last_error(runtime_->ctx));
        return frame;
    }
    if (request.cancel && request.cancel->load(std::memory_order_relaxed)) return frame;
    out.CopyTo(&frame);
    return frame;
}

FrameBuffer HipBackend::RenderProgressiveBatch(const RenderRequest& request, bool reset, bool resolve,
                                               uint32_t* accumulatedSamples) const {
    std::scoped_lock lock(mutex_);
    FrameBuffer frame;
    if (accumulatedSamples) *accumulatedSamples = 0;
    if (!EnsureRuntime() || !PrepareScene(request.scene)) return frame;
    const shiro_hip_camera cam = ToCamera(request.camera);
    const shiro_hip_settings st = ToSettings(request.settings, request.sampleCount);
    std::unique_ptr<HostFrame> out;
    if (resolve) out = std::make_unique<HostFrame>(st.width, st.height);
    const int rc = shiro_hip_render_progressive(runtime_->ctx, &cam, &st, request.sampleStart, request.sampleCount,
                                                reset ? SHIRO_HIP_PROGRESSIVE_RESET : 0u, CancelByte(request.cancel),
                                                out ? &out->view : nullptr, accumulatedSamples);
    if (rc != SHIRO_HIP_OK) {
        if (rc != SHIRO_HIP_CANCELLED)
            std::fprintf(stderr, "shiro hip render failed: %s\n", shiro_hip_last_error(runtime_->ctx));
        return frame;
    }
    if (out) {
        frame.Resize(st.width, st.height);
        out->CopyTo(&frame);
    }
    return frame;
}

}  // namespace shiro::backend::hip
