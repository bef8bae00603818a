// HipBackend: Shiro's GPU backend slot filled by the MI355X path tracer.
//
// Drop-in for the reference's optix::OptixBackend (include/shiro/backend/
// optix/OptixBackend.h:10-27): it implements the virtual RenderBackend
// interface (include/shiro/backend/RenderBackend.h:28-34) plus the
// non-virtual SupportsScene() that Renderer::Impl::ResolveBackend consults
// (src/render/Renderer.cpp:16-31).  It is a thin adapter over the C ABI in
// include/shiro_hip.h: scene flattening and a content-hash scene cache on
// the host, everything else in libshiro_hip.so.
//
// Semantics kept from the reference backends: methods are const and
// serialised by a mutex; failures print to stderr and return a cleared
// frame (beauty 0, depth +inf); an unavailable device reports
// available = false so that the renderer falls back to the CPU backend.
#pragma once

#include <memory>
#include <mutex>

#include "shiro/backend/RenderBackend.h"

namespace shiro::backend::hip {

class HipBackend final : public RenderBackend {
public:
    HipBackend();
    ~HipBackend() override;

    BackendStatus GetStatus() const override;
    bool SupportsScene(const render::Scene& scene) const;
    render::FrameBuffer RenderSampleBatch(const RenderRequest& request) const override;

    // Progressive rendering with the accumulator on the device: one call
    // stands for the Hydra worker's RenderSampleBatch, FrameAccumulator::
    // Accumulate and Resolve (src/hydra/RenderParam.cpp:789-803,
    // src/render/Renderer.cpp:47-110), with the same results.  `reset`
    // clears the accumulator first (the worker resets at sampleStart == 0);
    // a size change also clears it.  With `resolve` the returned frame is
    // the resolved running frame; without it the call reads nothing back and
    // returns an empty frame.  A cancelled batch is not accumulated and
    // returns an empty frame.  accumulatedSamples (may be null) receives the
    // accumulator's sample count (0 on failure).
    render::FrameBuffer RenderProgressiveBatch(const RenderRequest& request, bool reset, bool resolve = true,
                                               uint32_t* accumulatedSamples = nullptr) const;

    struct RuntimeState;

private:
    bool EnsureRuntime() const;
    bool PrepareScene(const render::Scene& scene) const;

    mutable std::mutex mutex_;
    mutable std::unique_ptr<RuntimeState> runtime_;
    mutable bool runtimeAttempted_ = false;
};

}  // namespace shiro::backend::hip
