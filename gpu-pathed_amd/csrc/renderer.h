// renderer.h -- the reference's renderer plugin interface (include/renderers/renderer.h:14-35,
// include/render_record.h:7-11, include/cuda_globals.h:11-38) for the HIP wavefront tracer.
//
// What changes at this boundary:
//   cudaEvent_t                 -> hipEvent_t (RenderRecord)
//   cudaGraphicsResource *      -> HeadlessPbo *: a pixel buffer in host memory that a finished
//                                  pass fills with its uchar4 sRGB framebuffer (no GL here)
//   CUDAGlobals                 -> DeviceGlobals: the scene on the device (world, BVH, emitters,
//                                  environment tables, camera, materials) in one libpathed
//                                  context, shared by whichever renderer drives it
//   PathTracer / OptixTracer    -> WavefrontRenderer (RendererType CUDA, Optix: plain NEE + MIS;
//   GBuffer(Normals)               Normals: the normals view; Guided: the learned guiding)
// Errors are std::runtime_error carrying pt_last_error(), where the reference exit()s.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/pathed.h"
#include "scene.h"

namespace rays {

// render_record.h:7-11: the events bracketing one pass on the renderer's stream, and its spp.
struct RenderRecord {
    hipEvent_t beginEvent = nullptr;
    hipEvent_t endEvent = nullptr;
    int spp = 0;
};

// A pixel buffer without GL (the PBOs of render_session.cu:15-53): `id` is the caller's handle,
// `rgba` receives width * height uchar4 sRGB pixels (row 0 = bottom) when a pass finishes.
struct HeadlessPbo {
    unsigned int id = 0;
    std::vector<uint8_t> rgba;
};
using PboResource = HeadlessPbo *;

// cuda_globals.h:11-38: the device-side scene every renderer reads.  mallocWorld uploads the
// world of a libpathed scene (building the BVH and the environment tables on the GPU) into a
// context for a width x height image; copyCamera and updateMaterial(s) are the edits of
// cuda_globals.cu.  One renderer renders through it at a time.
class DeviceGlobals {
public:
    DeviceGlobals() = default;
    ~DeviceGlobals();
    DeviceGlobals(const DeviceGlobals &) = delete;
    DeviceGlobals &operator=(const DeviceGlobals &) = delete;

    void mallocWorld(const pt_scene *world, int width, int height);
    void copyCamera(const Camera &camera);
    void updateMaterial(int materialID, const MaterialParams &params);
    void updateMaterials(const Scene &scene);
    // hit_test.cu:18-62: material of the closest hit of pixel (x, y)'s centre ray, -1 on a miss
    int hitTest(int x, int y);
    // waits for the pass in flight (render_session.cu:114 synchronises the device)
    void synchronize();

    bool allocated() const { return m_context != nullptr; }
    int width() const { return m_width; }
    int height() const { return m_height; }
    pt_renderer *context() const { return m_context; }
    hipStream_t stream() const;

    // The integrator settings of the context; only values that differ from the current ones
    // are pushed (each push restarts accumulation).  Called by the renderer that drives the
    // context; a different renderer taking it over also restarts accumulation.
    void bindIntegrator(const void *renderer, RendererType type, int maxDepth, bool nee) const;

private:
    pt_renderer *m_context = nullptr;
    int m_width = 0, m_height = 0;
    mutable const void *m_owner = nullptr;
    mutable int m_type = -1, m_maxDepth = -1, m_nee = -1;
};

// renderer.h:14-35
class Renderer {
public:
    virtual ~Renderer() = default;
    virtual void init(int width, int height, const Scene &scene);
    virtual RenderRecord renderAsync(int spp, PboResource pboResource, const Scene &scene, const DeviceGlobals &globals) = 0;
    virtual bool pollRender(PboResource pboResource, RenderRecord record) = 0;
    virtual void reset() = 0;
    virtual int getSpp() const = 0;
    // 3 * width * height floats, row-major, row 0 = bottom (renderer.cpp:20-42)
    virtual std::vector<float> getRadianceBuffer() const;

protected:
    int m_width = 0, m_height = 0;
};

// The HIP wavefront tracer behind the interface.  init() creates the pass events; renderAsync()
// pushes the scene's depth and NEE switch, restarts accumulation after reset(), and enqueues
// one progressive pass of `spp` samples per pixel between the record's events; pollRender()
// returns true once the end event completed, after copying the pass's framebuffer into the
// pixel buffer (when one is given).
class WavefrontRenderer : public Renderer {
public:
    explicit WavefrontRenderer(RendererType type = RendererType::Optix) : m_type(type) {}
    ~WavefrontRenderer() override;
    WavefrontRenderer(const WavefrontRenderer &) = delete;
    WavefrontRenderer &operator=(const WavefrontRenderer &) = delete;

    void init(int width, int height, const Scene &scene) override;
    RenderRecord renderAsync(int spp, PboResource pboResource, const Scene &scene, const DeviceGlobals &globals) override;
    bool pollRender(PboResource pboResource, RenderRecord record) override;
    void reset() override;
    int getSpp() const override { return m_currentSamples; }
    std::vector<float> getRadianceBuffer() const override;

    RendererType type() const { return m_type; }

private:
    RendererType m_type;
    int m_currentSamples = 0;
    bool m_shouldReset = true;
    hipEvent_t m_begin = nullptr, m_end = nullptr;
    const DeviceGlobals *m_globals = nullptr;
};

// The renderer RenderSession installs for a RendererType (render_session.cu:60-66,
// render_session.h:103-114): every type is served by the wavefront tracer.
std::unique_ptr<Renderer> makeRenderer(RendererType type);

}  // namespace rays
