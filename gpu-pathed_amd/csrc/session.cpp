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
// Scene (scene.h:38-61)
// ---------------------------------------------------------------------------------------
void Scene::setMaterialType(int id, MaterialType t) {
    MaterialParams &m = m_materials.at(id);
    if (m.type == t) return;
    MaterialParams n;
    n.type = t;
    switch (t) {
    case MaterialType::Lambertian: n.albedo = Vec3(0.f); n.emit = Vec3(0.f); break;
    case MaterialType::Mirror: break;
    case MaterialType::Glass: n.ior = 1.4f; break;
    case MaterialType::Microfacet: n.alpha = 0.1f; break;
    default: return;
    }
    m = n;
}

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
void RenderSession::check(int status) const {
    if (status != PT_OK) throw std::runtime_error(pt_last_error());
}

RenderSession::RenderSession(int width, int height) : m_width(width), m_height(height), m_rendererType(RendererType::Optix) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("RenderSession: width and height must be positive");
    check(pt_scene_create(&m_hostScene));
    // render_session.cu:69-71: scene preset 0 and its camera
    check(pt_scene_preset(m_hostScene, 0, 0, 0));
    buildSceneObjects();
}

RenderSession::~RenderSession() {
    if (m_renderer) pt_renderer_destroy(m_renderer);
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

void RenderSession::pushMaterial(int id) {
    const MaterialParams &m = m_scene->getMaterial(id);
    const float albedo[3] = {m.albedo.x(), m.albedo.y(), m.albedo.z()};
    const float emit[3] = {m.emit.x(), m.emit.y(), m.emit.z()};
    check(pt_renderer_set_material(m_renderer, id, (int)m.type, albedo, emit, m.ior, m.alpha));
}

void RenderSession::pushCamera() {
    const Camera &c = m_scene->getCamera();
    const float o[3] = {c.getOrigin().x(), c.getOrigin().y(), c.getOrigin().z()};
    const float t[3] = {c.getTarget().x(), c.getTarget().y(), c.getTarget().z()};
    const float u[3] = {c.getUp().x(), c.getUp().y(), c.getUp().z()};
    check(pt_renderer_set_camera(m_renderer, o, t, u, c.getVerticalFOV(), c.getFlipHandedness() ? 1 : 0));
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
    m_sppOptimizer.reset();
    if (!m_renderer) {
        if (id >= 0) m_dirtyMaterials.push_back(id);
        return;
    }
    if (m_rendering) {
        // the reference synchronises the device inside the callback (render_session.cu:114)
        check(pt_renderer_synchronize(m_renderer));
        float ms = 0.f;
        int finished = 0;
        check(pt_renderer_poll(m_renderer, &finished, &ms));
    }
    check(pt_renderer_set_max_depth(m_renderer, a.maxDepth));
    check(pt_renderer_set_nee(m_renderer, a.nextEventEstimation ? 1 : 0));
    if (id >= 0) pushMaterial(id);
    pushCamera();
    check(pt_renderer_reset(m_renderer));
}

RenderState RenderSession::init(unsigned int pbo1, unsigned int pbo2) {
    m_pbo1 = pbo1;
    m_pbo2 = pbo2;
    m_computeOnPrimary = false;
    if (!m_renderer) check(pt_renderer_create(m_width, m_height, &m_renderer));
    check(pt_renderer_set_scene(m_renderer, m_hostScene));
    // the scene model's state may have been edited before init
    check(pt_renderer_set_max_depth(m_renderer, m_scene->getMaxDepth()));
    check(pt_renderer_set_nee(m_renderer, m_scene->getNextEventEstimation() ? 1 : 0));
    for (int id : m_dirtyMaterials) pushMaterial(id);
    m_dirtyMaterials.clear();
    pushCamera();
    check(pt_renderer_set_type(m_renderer, (int)m_rendererType));
    check(pt_renderer_reset(m_renderer));
    m_initialized = true;
    return RenderState{false, pbo1};
}

void RenderSession::hitTest(int x, int y) {
    if (!m_renderer) throw std::runtime_error("hitTest: call init() first");
    if (m_rendering) {
        check(pt_renderer_synchronize(m_renderer));
    }
    int id = -1, hit = 0;
    check(pt_renderer_hit_test(m_renderer, x, y, &id, &hit));
    m_sceneModel->setMaterialID(hit ? id : -1);
}

RenderState RenderSession::renderAsync() {
    if (!m_renderer) throw std::runtime_error("renderAsync: call init() first");
    if (m_rendering) throw std::runtime_error("renderAsync: previous pass not finished; call pollRender()");
    const int spp = m_sppOptimizer.estimateSpp();
    check(pt_renderer_render_async(m_renderer, spp));
    m_currentSpp = spp;
    m_rendering = true;
    return RenderState{true, m_computeOnPrimary ? m_pbo2 : m_pbo1};
}

RenderState RenderSession::pollRender() {
    const unsigned int display = m_computeOnPrimary ? m_pbo2 : m_pbo1;
    if (!m_renderer || !m_rendering) return RenderState{false, display};
    int finished = 0;
    float ms = 0.f;
    check(pt_renderer_poll(m_renderer, &finished, &ms));
    if (finished) {
        m_rendering = false;
        m_sppOptimizer.track(m_currentSpp, std::max(ms, 1e-3f));
        m_computeOnPrimary = !m_computeOnPrimary;
        int64_t spp = 0;
        check(pt_renderer_get_spp(m_renderer, &spp));
        m_sceneModel->updateSpp((int)spp);
        // render_session.h:87-101: dump the radiance when this pass crossed a power of two
        if (!m_exrDumpPath.empty()) {
            for (int64_t t = 2; t <= (1 << 20); t <<= 1) {
                if (spp >= t && spp - m_currentSpp < t) {
                    const std::vector<float> radiance = getRadianceBuffer();
                    check(pt_exr_write_rgb(m_exrDumpPath.c_str(), radiance.data(), m_width, m_height, 1));
                    break;
                }
            }
        }
        if (m_resetRenderer) {
            // render_session.h:103-114: switch the integrator and restart accumulation
            check(pt_renderer_set_type(m_renderer, (int)m_rendererType));
            check(pt_renderer_reset(m_renderer));
            m_sppOptimizer.reset();
            m_resetRenderer = false;
        }
    }
    return RenderState{!finished, m_computeOnPrimary ? m_pbo2 : m_pbo1};
}

std::vector<float> RenderSession::getRadianceBuffer() {
    if (!m_renderer) throw std::runtime_error("getRadianceBuffer: call init() first");
    std::vector<float> out(3 * (size_t)m_width * m_height);
    check(pt_renderer_get_radiance(m_renderer, out.data()));
    return out;
}

std::vector<uint8_t> RenderSession::getFramebuffer() {
    if (!m_renderer) throw std::runtime_error("getFramebuffer: call init() first");
    std::vector<uint8_t> out(4 * (size_t)m_width * m_height);
    check(pt_renderer_get_rgba8(m_renderer, out.data()));
    return out;
}

int64_t RenderSession::getAccumulatedSpp() {
    if (!m_renderer) return 0;
    int64_t s = 0;
    check(pt_renderer_get_spp(m_renderer, &s));
    return s;
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
    if (m_renderer) init(m_pbo1, m_pbo2);
}

void RenderSession::loadPreset(int preset, uint64_t seed, int64_t param) {
    if (m_rendering) throw std::runtime_error("loadPreset: a pass is still running");
    check(pt_scene_preset(m_hostScene, preset, seed, param));
    m_dirtyMaterials.clear();
    buildSceneObjects();
    if (m_renderer) init(m_pbo1, m_pbo2);
}

}  // namespace rays
