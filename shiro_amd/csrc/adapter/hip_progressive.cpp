// Progressive rendering through HipBackend, the way Shiro's Hydra worker
// drives a backend (src/hydra/RenderParam.cpp:744-836): batches of
// samplesPerUpdate samples, each accumulated and resolved for display.
//
// Three routes over the same batches:
//   host    RenderSampleBatch per batch (frame read back), accumulated and
//           resolved on the host with FrameAccumulator's arithmetic
//           (src/render/Renderer.cpp:47-110, restated below)
//   device  RenderProgressiveBatch with the accumulator on the device, the
//           resolved frame read back every batch
//   nodisp  RenderProgressiveBatch without readback, one resolve at the end
// The resolved frames of all three must be identical bit for bit.  Prints
// one line: "progressive ok WxH batches N spp S: host A ms, device B ms,
// nodisp C ms per batch" (exit 0), or the first mismatch (exit 1).
//
//   hip_progressive [width height batches samplesPerUpdate]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "shiro/backend/RenderBackend.h"
#include "shiro/backend/hip/HipBackend.h"
#include "shiro/render/Renderer.h"

using namespace shiro::render;

namespace {

// FrameAccumulator's running sums (Renderer.cpp:47-83) and Resolve
// (:85-110): weighted sums of the batch means, depth summed only where a
// batch hit, resolved as sum / total (depth / its own weight, +inf if none).
struct HostAccumulator {
    std::vector<Vec4f> beauty;
    std::vector<Vec3f> albedo, normal;
    std::vector<float> depth;
    std::vector<uint32_t> depthSamples;
    uint32_t total = 0;

    void Reset(size_t np) {
        beauty.assign(np, Vec4f{});
        albedo.assign(np, Vec3f{});
        normal.assign(np, Vec3f{});
        depth.assign(np, 0.0f);
        depthSamples.assign(np, 0u);
        total = 0;
    }
    void Accumulate(const FrameBuffer& f, uint32_t count) {
        const float w = static_cast<float>(count);
        for (size_t i = 0; i < beauty.size(); ++i) {
            beauty[i] = beauty[i] + f.Beauty()[i] * w;
            albedo[i] = albedo[i] + f.Albedo()[i] * w;
            normal[i] = normal[i] + f.Normal()[i] * w;
            if (f.Depth()[i] < std::numeric_limits<float>::infinity()) {
                depth[i] += f.Depth()[i] * w;
                depthSamples[i] += count;
            }
        }
        total += count;
    }
    FrameBuffer Resolve(uint32_t width, uint32_t height) const {
        FrameBuffer f;
        f.Resize(width, height);
        const float inv = 1.0f / static_cast<float>(total);
        for (uint32_t y = 0; y < height; ++y) {
            for (uint32_t x = 0; x < width; ++x) {
                const size_t i = static_cast<size_t>(y) * width + x;
                f.SetBeauty(x, y, beauty[i] * inv);
                f.SetAlbedo(x, y, albedo[i] * inv);
                f.SetNormal(x, y, normal[i] * inv);
                f.SetDepth(x, y,
                           depthSamples[i] > 0 ? depth[i] / static_cast<float>(depthSamples[i])
                                               : std::numeric_limits<float>::infinity());
            }
        }
        return f;
    }
};

// A ground quad and a UV sphere (48 x 24 quads, 2,304 triangles) under a
// directional light and a gradient sky; the sphere's silhouette and the
// horizon give pixels whose batch minima differ, so the depth AOV's mean of
// minima is exercised.
Scene MakeScene() {
    Scene scene;
    scene.environmentTop = {0.3f, 0.5f, 0.9f};
    scene.environmentBottom = {0.08f, 0.07f, 0.06f};
    PbrMaterial ground;
    ground.baseColor = {0.6f, 0.55f, 0.5f};
    ground.roughness = 0.6f;
    PbrMaterial ball;
    ball.baseColor = {0.8f, 0.25f, 0.2f};
    ball.roughness = 0.3f;
    ball.metallic = 0.2f;
    scene.materials = {ground, ball};
    TriangleMesh quad;
    quad.positions = {{-8.0f, 0.0f, -8.0f}, {8.0f, 0.0f, -8.0f}, {8.0f, 0.0f, 8.0f}, {-8.0f, 0.0f, 8.0f}};
    quad.normals = {{0.0f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}};
    quad.indices = {0u, 2u, 1u, 0u, 3u, 2u};
    quad.materialIndex = 0;
    TriangleMesh sphere;
    const int nu = 48, nv = 24;
    for (int j = 0; j <= nv; ++j) {
        for (int i = 0; i <= nu; ++i) {
            const float th = 3.14159265f * static_cast<float>(j) / nv;
            const float ph = 6.28318531f * static_cast<float>(i) / nu;
            const Vec3f n{std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph)};
            sphere.positions.push_back({n.x, n.y + 1.05f, n.z});
            sphere.normals.push_back(n);
        }
    }
    for (int j = 0; j < nv; ++j) {
        for (int i = 0; i < nu; ++i) {
            const uint32_t a = j * (nu + 1) + i, b = a + 1, c = a + (nu + 1), d = c + 1;
            sphere.indices.insert(sphere.indices.end(), {a, c, b, b, c, d});
        }
    }
    sphere.materialIndex = 1;
    scene.meshes = {quad, sphere};
    scene.distantLights.push_back({{-0.4f, -1.0f, -0.3f}, {2.5f, 2.5f, 2.5f}});
    return scene;
}

bool Same(const FrameBuffer& a, const FrameBuffer& b, const char* what) {
    auto eq = [](const void* x, const void* y, size_t n) { return std::memcmp(x, y, n) == 0; };
    const size_t np = static_cast<size_t>(a.Width()) * a.Height();
    if (a.Width() != b.Width() || a.Height() != b.Height() || a.Beauty().size() != np || b.Beauty().size() != np) {
        std::printf("progressive mismatch (%s): frame sizes\n", what);
        return false;
    }
    if (!eq(a.Beauty().data(), b.Beauty().data(), np * sizeof(Vec4f)) ||
        !eq(a.Albedo().data(), b.Albedo().data(), np * sizeof(Vec3f)) ||
        !eq(a.Normal().data(), b.Normal().data(), np * sizeof(Vec3f)) ||
        !eq(a.Depth().data(), b.Depth().data(), np * sizeof(float))) {
        std::printf("progressive mismatch (%s): AOVs differ\n", what);
        return false;
    }
    return true;
}

double Now() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char** argv) {
    const uint32_t width = argc > 1 ? static_cast<uint32_t>(std::atoi(argv[1])) : 3840u;
    const uint32_t height = argc > 2 ? static_cast<uint32_t>(std::atoi(argv[2])) : 2160u;
    const uint32_t batches = argc > 3 ? static_cast<uint32_t>(std::atoi(argv[3])) : 4u;
    const uint32_t perUpdate = argc > 4 ? static_cast<uint32_t>(std::atoi(argv[4])) : 1u;
    shiro::backend::hip::HipBackend backend;
    if (!backend.GetStatus().available) {
        std::fprintf(stderr, "hip progressive failed: backend unavailable\n");
        return 3;
    }
    RenderSettings settings;
    settings.width = width;
    settings.height = height;
    settings.samplesPerPixel = batches * perUpdate;
    settings.samplesPerUpdate = perUpdate;
    settings.maxDepth = 4;
    settings.diffuseDepth = 4;
    settings.specularDepth = 4;
    const Scene scene = MakeScene();
    Camera camera;
    camera.position = {0.0f, 2.4f, 6.0f};
    const Vec3f f = Normalize(Vec3f{0.0f, 0.9f, 0.0f} - camera.position);
    const Vec3f r = Normalize(Cross(f, Vec3f{0.0f, 1.0f, 0.0f}));
    camera.forward = f;
    camera.right = r;
    camera.up = Cross(r, f);
    camera.aspectRatio = static_cast<float>(width) / static_cast<float>(height);

    // Warm-up: scene upload and first-use allocations of both routes.
    {
        const shiro::backend::RenderRequest q{scene, camera, settings, 0, perUpdate, nullptr};
        (void)backend.RenderSampleBatch(q);
        (void)backend.RenderProgressiveBatch(q, true, true);
    }
    HostAccumulator acc;
    acc.Reset(static_cast<size_t>(width) * height);
    FrameBuffer host, device, nodisp;
    double tHost = 0.0, tDevice = 0.0, tNoDisp = 0.0;
    for (uint32_t b = 0; b < batches; ++b) {
        const shiro::backend::RenderRequest q{scene, camera, settings, b * perUpdate, perUpdate, nullptr};
        double t0 = Now();
        const FrameBuffer batch = backend.RenderSampleBatch(q);
        acc.Accumulate(batch, perUpdate);
        host = acc.Resolve(width, height);
        tHost += Now() - t0;
        t0 = Now();
        uint32_t n = 0;
        device = backend.RenderProgressiveBatch(q, b == 0, true, &n);
        tDevice += Now() - t0;
        if (n != (b + 1) * perUpdate) {
            std::printf("progressive mismatch: accumulator holds %u samples after batch %u\n", n, b);
            return 1;
        }
    }
    for (uint32_t b = 0; b < batches; ++b) {
        const shiro::backend::RenderRequest q{scene, camera, settings, b * perUpdate, perUpdate, nullptr};
        const double t0 = Now();
        nodisp = backend.RenderProgressiveBatch(q, b == 0, b + 1 == batches);
        tNoDisp += Now() - t0;
    }
    if (!Same(host, device, "host vs device") || !Same(device, nodisp, "device vs no display")) return 1;
    std::printf("progressive ok %ux%u batches %u spp %u: host %.2f ms, device %.2f ms, nodisp %.2f ms per batch\n",
                width, height, batches, perUpdate, tHost / batches, tDevice / batches, tNoDisp / batches);
    return 0;
}
