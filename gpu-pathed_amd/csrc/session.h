// session.h -- the reference's host interface (RenderSession, Scene, SceneModel,
// SppOptimizer) restated over the libpathed C ABI.  Same class and method names and the
// same edit semantics as include/render_session.h, include/scene.h and
// include/scene_model.h of the reference, without GL interop: PBO handles are opaque
// integers and the rendered image is read back through getRadianceBuffer/getFramebuffer.
#pragma once

#include <stdint.h>

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pathed.h"

namespace rays {

// core/vec3.h:8-61 (value class exposed to Python)
class Vec3 {
public:
    Vec3() : e{0.f, 0.f, 0.f} {}
    Vec3(float a) : e{a, a, a} {}
    Vec3(float a, float b, float c) : e{a, b, c} {}
    float x() const { return e[0]; }
    float y() const { return e[1]; }
    float z() const { return e[2]; }
    float r() const { return e[0]; }
    float g() const { return e[1]; }
    float b() const { return e[2]; }
    float e[3];
};

// materials/types.h:5-16
enum class MaterialType { Lambertian = 0, Mirror = 1, Glass = 2, Microfacet = 3, None = 4 };
struct MaterialIndex {
    MaterialType materialType;
    size_t index;
};

// scene_model.h:15-19 (+ the guided wavefront integrator)
enum class RendererType { CUDA = 0, Optix = 1, Normals = 2, Guided = 3 };

// render_session.h:36-39 (pbo is an opaque handle here)
struct RenderState {
    bool isRendering;
    unsigned int pbo;
};

struct Resolution {
    int x, y;
};

// core/camera.h:16-72 (parameters only; the ray generation runs on the GPU)
class Camera {
public:
    Camera() = default;
    Camera(const Vec3 &origin, const Vec3 &target, const Vec3 &up, float verticalFOV, Resolution resolution, bool flipHandedness)
        : m_origin(origin), m_target(target), m_up(up), m_verticalFOV(verticalFOV), m_resolution(resolution), m_flipHandedness(flipHandedness) {}
    Vec3 getOrigin() const { return m_origin; }
    Vec3 getTarget() const { return m_target; }
    Vec3 getUp() const { return m_up; }
    float getVerticalFOV() const { return m_verticalFOV; }
    Resolution getResolution() const { return m_resolution; }
    bool getFlipHandedness() const { return m_flipHandedness; }

private:
    Vec3 m_origin, m_target, m_up;
    float m_verticalFOV = 0.f;
    Resolution m_resolution{0, 0};
    bool m_flipHandedness = false;
};

// spp_optimizer.cpp:8-29: next pass = floor(80 ms * measured spp/ms), at least 1; 1 after
// a reset
class SppOptimizer {
public:
    void track(int spp, float milliseconds);
    int estimateSpp();
    void reset();

private:
    int m_nextEstimate = 1;
    bool m_reset = false;
};

// One material's parameters with the reference's per-type applicability (params.h:15-107):
// albedo/emit only for Lambertian, ior only for Glass, alpha only for Microfacet.
struct MaterialParams {
    MaterialType type = MaterialType::Lambertian;
    Vec3 albedo, emit;
    float ior = 0.f, alpha = 0.f;
    Vec3 getAlbedo() const { return type == MaterialType::Lambertian ? albedo : Vec3(0.f); }
    Vec3 getEmit() const { return type == MaterialType::Lambertian ? emit : Vec3(0.f); }
    float getIOR() const { return type == MaterialType::Glass ? ior : 0.f; }
    float getAlpha() const { return type == MaterialType::Microfacet ? alpha : 0.f; }
};

constexpr float defaultLightPosition = -0.6f;  // scene.h:17
constexpr int defaultMaxDepth = 3;             // scene.h:18

// scene.h:20-118
class Scene {
public:
    Scene(const Camera &camera, std::vector<MaterialParams> materials)
        : m_camera(camera), m_materials(std::move(materials)), m_maxDepth(defaultMaxDepth), m_nextEventEstimation(true) {}
    const Camera &getCamera() const { return m_camera; }
    void setCamera(const Camera &camera) { m_camera = camera; }
    // scene.h:38-61: a type change installs the new type's default parameters
    void setMaterialType(int materialID, MaterialType materialType);
    Vec3 getColor(int id) const { return m_materials.at(id).getAlbedo(); }
    Vec3 getEmit(int id) const { return m_materials.at(id).getEmit(); }
    void setColor(int id, Vec3 c) { if (m_materials.at(id).type == MaterialType::Lambertian) m_materials[id].albedo = c; }
    void setEmit(int id, Vec3 c) { if (m_materials.at(id).type == MaterialType::Lambertian) m_materials[id].emit = c; }
    float getIOR(int id) const { return m_materials.at(id).getIOR(); }
    void setIOR(int id, float v) { if (m_materials.at(id).type == MaterialType::Glass) m_materials[id].ior = v; }
    float getAlpha(int id) const { return m_materials.at(id).getAlpha(); }
    void setAlpha(int id, float v) { if (m_materials.at(id).type == MaterialType::Microfacet) m_materials[id].alpha = v; }
    MaterialType getMaterialType(int id) const { return m_materials.at(id).type; }
    const MaterialParams &getMaterial(int id) const { return m_materials.at(id); }
    int materialCount() const { return (int)m_materials.size(); }
    int getMaxDepth() const { return m_maxDepth; }
    void setMaxDepth(int d) { m_maxDepth = d; }
    bool getNextEventEstimation() const { return m_nextEventEstimation; }
    void setNextEventEstimation(bool v) { m_nextEventEstimation = v; }

private:
    Camera m_camera;
    std::vector<MaterialParams> m_materials;
    int m_maxDepth;
    bool m_nextEventEstimation;
};

// scene_model.h:21-31
struct SceneModelAttributes {
    Vec3 albedo;
    Vec3 emitted;
    float ior;
    float alpha;
    Camera camera;
    RendererType rendererType;
    int maxDepth;
    bool nextEventEstimation;
    MaterialType materialType;
};

// scene_model.h:33-99: every setter builds the attribute set and fires the one callback
class SceneModel {
public:
    SceneModel(const Scene *scene, float lightPosition, RendererType rendererType);
    SceneModel(const SceneModel &) = delete;
    SceneModel(SceneModel &&) = delete;

    void subscribe(std::function<void(const SceneModelAttributes &)> callback) { m_callback = std::move(callback); }

    RendererType getRendererType() const { return m_rendererType; }
    void setRendererType(RendererType rendererType);
    void setColor(float r, float g, float b);
    Vec3 getColor() const;
    void setEmit(float r, float g, float b);
    Vec3 getEmit() const;
    void setIOR(float ior);
    float getIOR() const;
    void setAlpha(float alpha);
    float getAlpha() const;
    int getMaterialID() const { return m_materialID; }
    void setMaterialID(int materialID) { m_materialID = materialID; }
    MaterialType getMaterialType() const;
    void setMaterialType(MaterialType materialType);
    void setLightPosition(float lightPosition);
    float getLightPosition() const { return m_lightPosition; }
    Vec3 getCameraOrigin() const { return m_scene->getCamera().getOrigin(); }
    void setCameraOrigin(float x, float y, float z);
    Vec3 getCameraTarget() const { return m_scene->getCamera().getTarget(); }
    void setCameraTarget(float x, float y, float z);
    Vec3 getCameraUp() const { return m_scene->getCamera().getUp(); }
    void setCameraUp(float x, float y, float z);
    void updateSpp(int spp) { m_spp = spp; }
    int getSpp() const { return m_spp; }
    int getMaxDepth() const { return m_scene->getMaxDepth(); }
    void setMaxDepth(int maxDepth);
    bool getNextEventEstimation() const { return m_scene->getNextEventEstimation(); }
    void setNextEventEstimation(bool nee);
    void zoomCamera(float ticks);

private:
    SceneModelAttributes current() const;
    void fire(const SceneModelAttributes &a) {
        if (m_callback) m_callback(a);
    }

    const Scene *m_scene;
    RendererType m_rendererType;
    int m_materialID = -1;
    int m_spp = 0;
    float m_lightPosition;
    std::function<void(const SceneModelAttributes &)> m_callback;
};

// render_session.h:41-134
class RenderSession {
public:
    RenderSession(int width, int height);
    ~RenderSession();
    RenderSession(const RenderSession &) = delete;
    RenderSession(RenderSession &&) = delete;

    SceneModel &getSceneModel() { return *m_sceneModel; }
    RenderState init(unsigned int pbo1, unsigned int pbo2);
    int getWidth() const { return m_width; }
    int getHeight() const { return m_height; }
    void hitTest(int x, int y);
    RenderState renderAsync();
    RenderState pollRender();

    // additions for headless use (not in the reference)
    std::vector<float> getRadianceBuffer();  // renderer.cpp:20-42 layout
    std::vector<uint8_t> getFramebuffer();   // uchar4 sRGB of framebuffer.cu:35-40
    int64_t getAccumulatedSpp();
    void loadPreset(int preset, uint64_t seed, int64_t param);
    // the reference's scene `index` (scene.cu:14-176) from its scenes/ folder layout
    void loadReferenceScene(int index, const std::string &scenesDir);
    // render_session.h:87-101 writes out.exr whenever the spp count crosses a power of two;
    // here that dump is opt-in: a non-empty path enables it
    void setExrDumpPath(const std::string &path) { m_exrDumpPath = path; }
    pt_renderer *renderer() { return m_renderer; }

private:
    void check(int status) const;
    void applyAttributes(const SceneModelAttributes &attributes);
    void buildSceneObjects();
    void pushMaterial(int id);
    void pushCamera();

    int m_width, m_height;
    SppOptimizer m_sppOptimizer;
    unsigned int m_pbo1 = 0, m_pbo2 = 0;
    bool m_computeOnPrimary = false;
    bool m_initialized = false;
    bool m_rendering = false;
    int m_currentSpp = 0;
    bool m_resetRenderer = false;
    RendererType m_rendererType;
    std::vector<int> m_dirtyMaterials;  // edited before init()
    std::string m_exrDumpPath;
    pt_scene *m_hostScene = nullptr;
    pt_renderer *m_renderer = nullptr;
    std::unique_ptr<Scene> m_scene;
    std::unique_ptr<SceneModel> m_sceneModel;
};

}  // namespace rays
