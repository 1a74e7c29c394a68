// renderer.cpp -- DeviceGlobals, Renderer and WavefrontRenderer over the libpathed C ABI.
#include "renderer.h"

#include <stdexcept>

namespace rays {

namespace {

void check(int status) {
    if (status != PT_OK) throw std::runtime_error(pt_last_error());
}

void checkHip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// DeviceGlobals (cuda_globals.cu:10-139)
// ---------------------------------------------------------------------------------------------
DeviceGlobals::~DeviceGlobals() {
    if (m_context) pt_renderer_destroy(m_context);
}

void DeviceGlobals::mallocWorld(const pt_scene *world, int width, int height) {
    if (!world) throw std::invalid_argument("DeviceGlobals::mallocWorld: no scene");
    if (m_context && (width != m_width || height != m_height)) {
        pt_renderer_destroy(m_context);
        m_context = nullptr;
    }
    if (!m_context) {
        check(pt_renderer_create(width, height, &m_context));
        m_width = width;
        m_height = height;
    }
    check(pt_renderer_set_scene(m_context, world));
    // the upload restarts accumulation with the context's defaults; renderers push their own
    // settings at their next pass
    m_owner = nullptr;
    m_type = m_maxDepth = m_nee = -1;
}

void DeviceGlobals::copyCamera(const Camera &c) {
    if (!m_context) throw std::runtime_error("DeviceGlobals::copyCamera: call mallocWorld first");
    const float o[3] = {c.getOrigin().x(), c.getOrigin().y(), c.getOrigin().z()};
    const float t[3] = {c.getTarget().x(), c.getTarget().y(), c.getTarget().z()};
    const float u[3] = {c.getUp().x(), c.getUp().y(), c.getUp().z()};
    check(pt_renderer_set_camera(m_context, o, t, u, c.getVerticalFOV(), c.getFlipHandedness() ? 1 : 0));
}

void DeviceGlobals::updateMaterial(int id, const MaterialParams &m) {
    if (!m_context) throw std::runtime_error("DeviceGlobals::updateMaterial: call mallocWorld first");
    const float albedo[3] = {m.albedo.x(), m.albedo.y(), m.albedo.z()};
    const float emit[3] = {m.emit.x(), m.emit.y(), m.emit.z()};
    check(pt_renderer_set_material(m_context, id, (int)m.type, albedo, emit, m.ior, m.alpha));
}

void DeviceGlobals::updateMaterials(const Scene &scene) {
    for (int id = 0; id < scene.materialCount(); id++) updateMaterial(id, scene.getMaterial(id));
}

int DeviceGlobals::hitTest(int x, int y) {
    if (!m_context) throw std::runtime_error("DeviceGlobals::hitTest: call mallocWorld first");
    synchronize();
    int id = -1, hit = 0;
    check(pt_renderer_hit_test(m_context, x, y, &id, &hit));
    return hit ? id : -1;
}

void DeviceGlobals::synchronize() {
    if (m_context) check(pt_renderer_synchronize(m_context));
}

hipStream_t DeviceGlobals::stream() const {
    if (!m_context) return nullptr;
    void *s = nullptr;
    check(pt_renderer_get_stream(m_context, &s));
    return (hipStream_t)s;
}

void DeviceGlobals::bindIntegrator(const void *renderer, RendererType type, int maxDepth, bool nee) const {
    if (!m_context) throw std::runtime_error("DeviceGlobals: call mallocWorld first");
    if ((int)type != m_type) {
        check(pt_renderer_set_type(m_context, (int)type));
        m_type = (int)type;
    }
    if (maxDepth != m_maxDepth) {
        check(pt_renderer_set_max_depth(m_context, maxDepth));
        m_maxDepth = maxDepth;
    }
    if ((int)nee != m_nee) {
        check(pt_renderer_set_nee(m_context, nee ? 1 : 0));
        m_nee = (int)nee;
    }
    if (renderer != m_owner) {
        check(pt_renderer_reset(m_context));
        m_owner = renderer;
    }
}

// ---------------------------------------------------------------------------------------------
// Renderer (renderer.cpp:10-42)
// ---------------------------------------------------------------------------------------------
void Renderer::init(int width, int height, const Scene &) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("Renderer::init: width and height must be positive");
    m_width = width;
    m_height = height;
}

std::vector<float> Renderer::getRadianceBuffer() const { return std::vector<float>(3 * (size_t)m_width * m_height, 0.f); }

// ---------------------------------------------------------------------------------------------
// WavefrontRenderer (path_tracer.cu:266-356 for the plain integrator, g_buffer.cu for Normals)
// ---------------------------------------------------------------------------------------------
WavefrontRenderer::~WavefrontRenderer() {
    if (m_begin) (void)hipEventDestroy(m_begin);
    if (m_end) (void)hipEventDestroy(m_end);
}

void WavefrontRenderer::init(int width, int height, const Scene &scene) {
    Renderer::init(width, height, scene);
    if (!m_begin) checkHip(hipEventCreate(&m_begin), "hipEventCreate");
    if (!m_end) checkHip(hipEventCreate(&m_end), "hipEventCreate");
    m_currentSamples = 0;
    m_shouldReset = true;
    m_globals = nullptr;
}

RenderRecord WavefrontRenderer::renderAsync(int spp, PboResource, const Scene &scene, const DeviceGlobals &globals) {
    if (!m_begin) throw std::runtime_error("WavefrontRenderer::renderAsync: call init() first");
    if (!globals.allocated()) throw std::runtime_error("WavefrontRenderer::renderAsync: the device globals hold no world");
    if (globals.width() != m_width || globals.height() != m_height)
        throw std::invalid_argument("WavefrontRenderer::renderAsync: image size differs from the device globals'");
    if (spp <= 0) throw std::invalid_argument("WavefrontRenderer::renderAsync: spp must be positive");
    pt_renderer *ctx = globals.context();
    globals.bindIntegrator(this, m_type, scene.getMaxDepth(), scene.getNextEventEstimation());
    if (m_shouldReset) {
        // path_tracer.cu:307-310: the pass after a reset starts the accumulation again
        check(pt_renderer_reset(ctx));
        m_currentSamples = 0;
        m_shouldReset = false;
    }
    m_globals = &globals;
    const hipStream_t stream = globals.stream();
    checkHip(hipEventRecord(m_begin, stream), "hipEventRecord");
    check(pt_renderer_render_async(ctx, spp));
    checkHip(hipEventRecord(m_end, stream), "hipEventRecord");
    RenderRecord rec;
    rec.beginEvent = m_begin;
    rec.endEvent = m_end;
    rec.spp = spp;
    return rec;
}

bool WavefrontRenderer::pollRender(PboResource pbo, RenderRecord record) {
    if (!m_globals) return true;  // nothing in flight
    const hipError_t q = hipEventQuery(record.endEvent);
    if (q == hipErrorNotReady) return false;
    checkHip(q, "hipEventQuery");
    pt_renderer *ctx = m_globals->context();
    int finished = 0;
    float ms = 0.f;
    check(pt_renderer_poll(ctx, &finished, &ms));
    if (!finished) return false;  // the end event follows the pass on one stream
    int64_t spp = 0;
    check(pt_renderer_get_spp(ctx, &spp));
    m_currentSamples = (int)spp;
    if (pbo) {
        // path_tracer.cu:349-352: the pass's framebuffer goes to the pixel buffer
        pbo->rgba.resize(4 * (size_t)m_width * m_height);
        check(pt_renderer_get_rgba8(ctx, pbo->rgba.data()));
    }
    return true;
}

void WavefrontRenderer::reset() { m_shouldReset = true; }

std::vector<float> WavefrontRenderer::getRadianceBuffer() const {
    if (!m_globals) return Renderer::getRadianceBuffer();
    std::vector<float> out(3 * (size_t)m_width * m_height);
    check(pt_renderer_get_radiance(m_globals->context(), out.data()));
    return out;
}

std::unique_ptr<Renderer> makeRenderer(RendererType type) { return std::make_unique<WavefrontRenderer>(type); }

}  // namespace rays
