// session.cpp -- RenderSession / Scene / SceneModel / SppOptimizer over the C ABI.
#include "session.h"

#include <math.h>

#include <algorithm>

namespace rays {

// ---------------------------------------------------------------------------------------
// SppOptimizer (spp_optimizer.cpp:8-29)
// ---------------------------------------------------------------------------------------
static constexpr float budgetMilliseconds = 80.f;

void SppOptimizer::track(int spp, float milliseconds) {
    const float samplesPerMs = spp / milliseconds;
    m_nextEstimate = std::max(1, (int)std::floor(budgetMilliseconds * samplesPerMs));
}
int SppOptimizer::estimateSpp() {
    if (m_reset) {
        m_reset = false;
        return 1;
    }
    return m_nextEstimate;
}
void SppOptimizer::reset() { m_reset = true; }

// ---------------------------------------------------------------------------------------
// SceneModel (scene_model.cu:6-347)
// ---------------------------------------------------------------------------------------
SceneModel::SceneModel(const Scene *scene, float lightPosition, RendererType rendererType)
    : m_scene(scene), m_rendererType(rendererType), m_lightPosition(lightPosition) {}

SceneModelAttributes SceneModel::current() const {
    return SceneModelAttributes{getColor(), getEmit(), getIOR(), getAlpha(), m_scene->getCamera(), m_rendererType,
                                getMaxDepth(), getNextEventEstimation(), getMaterialType()};
}

void SceneModel::setRendererType(RendererType t) {
    m_rendererType = t;
    fire(current());
}
void SceneModel::setColor(float r, float g, float b) {
    SceneModelAttributes a = current();
    a.albedo = Vec3(r, g, b);
    fire(a);
}
Vec3 SceneModel::getColor() const { return m_materialID == -1 ? Vec3(0.f) : m_scene->getColor(m_materialID); }
void SceneModel::setEmit(float r, float g, float b) {
    SceneModelAttributes a = current();
    a.emitted = Vec3(r, g, b);
    fire(a);
}
Vec3 SceneModel::getEmit() const { return m_materialID == -1 ? Vec3(0.f) : m_scene->getEmit(m_materialID); }
void SceneModel::setIOR(float ior) {
    SceneModelAttributes a = current();
    a.ior = ior;
    fire(a);
}
float SceneModel::getIOR() const { return m_materialID == -1 ? -1.f : m_scene->getIOR(m_materialID); }
void SceneModel::setAlpha(float alpha) {
    SceneModelAttributes a = current();
    a.alpha = alpha;
    fire(a);
}
float SceneModel::getAlpha() const { return m_materialID == -1 ? -1.f : m_scene->getAlpha(m_materialID); }
MaterialType SceneModel::getMaterialType() const { return m_materialID == -1 ? MaterialType::None : m_scene->getMaterialType(m_materialID); }
void SceneModel::setMaterialType(MaterialType t) {
    SceneModelAttributes a = current();
    a.materialType = t;
    fire(a);
}
void SceneModel::setLightPosition(float p) {
    m_lightPosition = p;
    fire(current());
}
void SceneModel::setCameraOrigin(float x, float y, float z) {
    const Camera &c = m_scene->getCamera();
    SceneModelAttributes a = current();
    a.camera = Camera(Vec3(x, y, z), c.getTarget(), c.getUp(), c.getVerticalFOV(), c.getResolution(), c.getFlipHandedness());
    fire(a);
}
void SceneModel::setCameraTarget(float x, float y, float z) {
    const Camera &c = m_scene->getCamera();
    SceneModelAttributes a = current();
    a.camera = Camera(c.getOrigin(), Vec3(x, y, z), c.getUp(), c.getVerticalFOV(), c.getResolution(), c.getFlipHandedness());
    fire(a);
}
void SceneModel::setCameraUp(float x, float y, float z) {
    const Camera &c = m_scene->getCamera();
    SceneModelAttributes a = current();
    a.camera = Camera(c.getOrigin(), c.getTarget(), Vec3(x, y, z), c.getVerticalFOV(), c.getResolution(), c.getFlipHandedness());
    fire(a);
}
void SceneModel::setMaxDepth(int d) {
    SceneModelAttributes a = current();
    a.maxDepth = d;
    fire(a);
}
void SceneModel::setNextEventEstimation(bool v) {
    SceneModelAttributes a = current();
    a.nextEventEstimation = v;
    fire(a);
}
// scene_model.cu:296-318
void SceneModel::zoomCamera(float ticks) {
    const Camera &c = m_scene->getCamera();
    const Vec3 o = c.getOrigin(), t = c.getTarget();
    float d[3] = {t.x() - o.x(), t.y() - o.y(), t.z() - o.z()};
    const float l = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (float &v : d) v /= l;
    SceneModelAttributes a = current();
    a.camera = Camera(Vec3(o.x() + d[0] * ticks, o.y() + d[1] * ticks, o.z() + d[2] * ticks), t, c.getUp(), c.getVerticalFOV(),
                      c.getResolution(), c.getFlipHandedness());
    fire(a);
}

// ---------------------------------------------------------------------------------------
// RenderSession (render_session.cu:55-141, render_session.h:59-118)
// ---------------------------------------------------------------------------------------
namespace {
void check(int status) {
    if (status != PT_OK) throw std::runtime_error(pt_last_error());
}
}  // namespace

RenderSession::RenderSession(int width, int height) : m_width(width), m_height(height), m_rendererType(RendererType::Optix) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("RenderSession: width and height must be positive");
    m_pathTracer = makeRenderer(m_rendererType);
    m_deviceGlobals = std::make_unique<DeviceGlobals>();
    check(pt_scene_create(&m_hostScene));
    // render_session.cu:69-71: scene preset 0 and its camera
    check(pt_scene_preset(m_hostScene, 0, 0, 0));
    buildSceneObjects();
}

RenderSession::~RenderSession() {
    m_pathTracer.reset();
    m_deviceGlobals.reset();
    if (m_hostScene) pt_scene_destroy(m_hostScene);
}

void RenderSession::buildSceneObjects() {
    pt_scene_info info;
    check(pt_scene_get_info(m_hostScene, &info));
    std::vector<int> type(info.n_materials);
    std::vector<float> albedo(3 * info.n_materials), emit(3 * info.n_materials), ior(info.n_materials), alpha(info.n_materials);
    check(pt_scene_export(m_hostScene, nullptr, nullptr, nullptr, nullptr, nullptr, type.data(), albedo.data(), emit.data(), ior.data(),
                          alpha.data(), nullptr));
    std::vector<MaterialParams> mats(info.n_materials);
    for (int64_t i = 0; i < info.n_materials; i++) {
        MaterialParams &m = mats[i];
        m.type = (MaterialType)type[i];
        m.albedo = Vec3(albedo[3 * i], albedo[3 * i + 1], albedo[3 * i + 2]);
        m.emit = Vec3(emit[3 * i], emit[3 * i + 1], emit[3 * i + 2]);
        m.ior = ior[i];
        m.alpha = alpha[i];
    }
    const Camera camera(Vec3(info.camera_origin[0], info.camera_origin[1], info.camera_origin[2]),
                        Vec3(info.camera_target[0], info.camera_target[1], info.camera_target[2]),
                        Vec3(info.camera_up[0], info.camera_up[1], info.camera_up[2]), info.camera_vfov, Resolution{m_width, m_height},
                        info.camera_flip != 0);
    m_scene = std::make_unique<Scene>(camera, std::move(mats));
    m_sceneModel = std::make_unique<SceneModel>(m_scene.get(), defaultLightPosition, m_rendererType);
    m_sceneModel->subscribe([this](const SceneModelAttributes &a) { applyAttributes(a); });
}

// render_session.cu:82-116
void RenderSession::applyAttributes(const SceneModelAttributes &a) {
    if (a.rendererType != m_rendererType) {
        m_rendererType = a.rendererType;
        m_resetRenderer = true;
        return;
    }
    m_scene->setMaxDepth(a.maxDepth);
    m_scene->setNextEventEstimation(a.nextEventEstimation);
    const int id = m_sceneModel->getMaterialID();
    if (id >= 0) {
        if (m_scene->getMaterialType(id) != a.materialType) {
            m_scene->setMaterialType(id, a.materialType);
        } else {
            m_scene->setColor(id, a.albedo);
            m_scene->setEmit(id, a.emitted);
            m_scene->setIOR(id, a.ior);
            m_scene->setAlpha(id, a.alpha);
        }
    }
    m_scene->setCamera(a.camera);
    m_pathTracer->reset();
    m_sppOptimizer.reset();
    if (!m_deviceGlobals->allocated()) {
        if (id >= 0) m_dirtyMaterials.push_back(id);
        return;
    }
    // the reference synchronises the device inside the callback (render_session.cu:114); the
    // edits below wait for the pass in flight the same way
    m_deviceGlobals->synchronize();
    m_deviceGlobals->copyCamera(m_scene->getCamera());
    if (id >= 0) m_deviceGlobals->updateMaterial(id, m_scene->getMaterial(id));
}

// render_session.cu:119-136
RenderState RenderSession::init(unsigned int pbo1, unsigned int pbo2) {
    if (m_rendering) {
        m_deviceGlobals->synchronize();
        m_rendering = false;
    }
    m_pbo[0].id = pbo1;
    m_pbo[1].id = pbo2;
    m_computeOnPrimary = false;
    m_deviceGlobals->mallocWorld(m_hostScene, m_width, m_height);
    // the scene model's state may have been edited before init
    for (int id : m_dirtyMaterials) m_deviceGlobals->updateMaterial(id, m_scene->getMaterial(id));
    m_dirtyMaterials.clear();
    m_deviceGlobals->copyCamera(m_scene->getCamera());
    m_pathTracer->init(m_width, m_height, *m_scene);
    return RenderState{false, pbo1};
}

void RenderSession::hitTest(int x, int y) {
    if (!m_deviceGlobals->allocated()) throw std::runtime_error("hitTest: call init() first");
    // hit_test.cu:18-62 sets the scene model's selected material
    m_sceneModel->setMaterialID(m_deviceGlobals->hitTest(x, y));
}

// render_session.h:59-70
RenderState RenderSession::renderAsync() {
    if (!m_deviceGlobals->allocated()) throw std::runtime_error("renderAsync: call init() first");
    if (m_rendering) throw std::runtime_error("renderAsync: previous pass not finished; call pollRender()");
    const int spp = m_sppOptimizer.estimateSpp();
    m_currentRecord = m_pathTracer->renderAsync(spp, renderPbo(), *m_scene, *m_deviceGlobals);
    m_rendering = true;
    return RenderState{true, displayPbo()};
}

void RenderSession::installRenderer() {
    m_pathTracer = makeRenderer(m_rendererType);
    m_pathTracer->init(m_width, m_height, *m_scene);
}

// render_session.h:72-117
RenderState RenderSession::pollRender() {
    if (!m_deviceGlobals->allocated() || !m_rendering) return RenderState{false, displayPbo()};
    const bool finished = m_pathTracer->pollRender(renderPbo(), m_currentRecord);
    if (finished) {
        m_rendering = false;
        float milliseconds = 0.f;
        if (hipEventElapsedTime(&milliseconds, m_currentRecord.beginEvent, m_currentRecord.endEvent) != hipSuccess) milliseconds = 0.f;
        m_sppOptimizer.track(m_currentRecord.spp, std::max(milliseconds, 1e-3f));
        m_computeOnPrimary = !m_computeOnPrimary;  // swapPBOs
        const int spp = m_pathTracer->getSpp();
        m_sceneModel->updateSpp(spp);
        // render_session.h:87-101: dump the radiance when this pass crossed a power of two
        if (!m_exrDumpPath.empty()) {
            for (int64_t t = 2; t <= (1 << 20); t <<= 1) {
                if (spp >= t && spp - m_currentRecord.spp < t) {
                    const std::vector<float> radiance = m_pathTracer->getRadianceBuffer();
                    check(pt_exr_write_rgb(m_exrDumpPath.c_str(), radiance.data(), m_width, m_height, 1));
                    break;
                }
            }
        }
        if (m_resetRenderer) {
            // render_session.h:103-114: a new renderer of the selected type, accumulation
            // restarts
            installRenderer();
            m_sppOptimizer.reset();
            m_resetRenderer = false;
        }
    }
    return RenderState{!finished, displayPbo()};
}

std::vector<float> RenderSession::getRadianceBuffer() {
    if (!m_deviceGlobals->allocated()) throw std::runtime_error("getRadianceBuffer: call init() first");
    if (m_rendering) m_deviceGlobals->synchronize();
    return m_pathTracer->getRadianceBuffer();
}

std::vector<uint8_t> RenderSession::getFramebuffer() {
    if (!m_deviceGlobals->allocated()) throw std::runtime_error("getFramebuffer: call init() first");
    // the display pixel buffer holds the last finished pass (black before the first)
    const HeadlessPbo &p = m_computeOnPrimary ? m_pbo[1] : m_pbo[0];
    if (p.rgba.empty()) return std::vector<uint8_t>(4 * (size_t)m_width * m_height, 0);
    return p.rgba;
}

int64_t RenderSession::getAccumulatedSpp() {
    if (!m_deviceGlobals->allocated()) return 0;
    return m_pathTracer->getSpp();
}

void RenderSession::loadReferenceScene(int index, const std::string &scenesDir) {
    if (m_rendering) throw std::runtime_error("loadReferenceScene: a pass is still running");
    pt_scene *fresh = nullptr;
    check(pt_scene_create(&fresh));
    const int rc = pt_scene_reference(fresh, index, scenesDir.c_str());
    if (rc != PT_OK) {
        const std::string msg = pt_last_error();
        pt_scene_destroy(fresh);
        throw std::runtime_error(msg);
    }
    pt_scene_destroy(m_hostScene);
    m_hostScene = fresh;
    m_dirtyMaterials.clear();
    buildSceneObjects();
    if (m_deviceGlobals->allocated()) init(m_pbo[0].id, m_pbo[1].id);
}

void RenderSession::loadPreset(int preset, uint64_t seed, int64_t param) {
    if (m_rendering) throw std::runtime_error("loadPreset: a pass is still running");
    check(pt_scene_preset(m_hostScene, preset, seed, param));
    m_dirtyMaterials.clear();
    buildSceneObjects();
    if (m_deviceGlobals->allocated()) init(m_pbo[0].id, m_pbo[1].id);
}

}  // namespace rays
