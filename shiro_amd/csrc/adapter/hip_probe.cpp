// Smoke probe of the HIP backend through Shiro's RenderBackend interface:
// the check shape of the reference's tools/optix_probe.cpp:7-80 (one
// triangle, one directional light, gradient sky, 8x8 at 1 spp; the centre
// pixel must be lit, covered and at finite depth), run on HipBackend.
#include <cstdio>

#include "shiro/backend/RenderBackend.h"
#include "shiro/backend/hip/HipBackend.h"
#include "shiro/render/Renderer.h"

int main() {
    shiro::backend::hip::HipBackend backend;
    const shiro::backend::BackendStatus status = backend.GetStatus();
    if (!status.available) {
        std::fprintf(stderr, "hip probe failed: backend unavailable\n");
        return 3;
    }
    shiro::render::RenderSettings settings;
    settings.width = 8;
    settings.height = 8;
    settings.samplesPerPixel = 1;
    settings.samplesPerUpdate = 1;
    settings.backend = shiro::render::BackendKind::Gpu;

    shiro::render::Scene scene;
    scene.environmentTop = {0.2f, 0.4f, 0.8f};
    scene.environmentBottom = {0.05f, 0.05f, 0.05f};
    shiro::render::PbrMaterial material;
    material.baseColor = {0.8f, 0.2f, 0.2f};
    material.roughness = 0.4f;
    scene.materials.push_back(material);
    shiro::render::TriangleMesh mesh;
    mesh.positions = {{-1.0f, -1.0f, 0.0f}, {1.0f, -1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}};
    mesh.normals = {{0.0f, 0.0f, 1.0f}, {0.0f, 0.0f, 1.0f}, {0.0f, 0.0f, 1.0f}};
    mesh.indices = {0u, 1u, 2u};
    scene.meshes.push_back(mesh);
    scene.distantLights.push_back({{0.0f, -1.0f, -1.0f}, {2.0f, 2.0f, 2.0f}});
    if (!backend.SupportsScene(scene)) {
        std::fprintf(stderr, "hip probe failed: scene not supported\n");
        return 5;
    }
    shiro::render::Camera camera;
    camera.aspectRatio = static_cast<float>(settings.width) / static_cast<float>(settings.height);

    const shiro::backend::RenderRequest request{scene, camera, settings, 0, 1, nullptr};
    const shiro::render::FrameBuffer frame = backend.RenderSampleBatch(request);
    if (frame.Width() == 0 || frame.Height() == 0 || frame.Beauty().empty()) {
        std::fprintf(stderr, "hip probe failed: empty frame\n");
        return 1;
    }
    const size_t centre = static_cast<size_t>(frame.Height() / 2u) * frame.Width() + frame.Width() / 2u;
    const shiro::render::Vec4f c = frame.Beauty()[centre];
    if (c.w <= 0.0f || (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f)) {
        std::fprintf(stderr, "hip probe failed: blank frame\n");
        return 2;
    }
    if (!(frame.Depth()[centre] < 1.0e10f)) {
        std::fprintf(stderr, "hip probe failed: no geometry hit\n");
        return 4;
    }
    std::printf("probe center %.4f %.4f %.4f %.4f depth %.4f\n", c.x, c.y, c.z, c.w, frame.Depth()[centre]);
    return 0;
}
