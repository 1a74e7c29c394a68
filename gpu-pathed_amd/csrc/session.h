// session.h -- the reference's host interface (RenderSession, SceneModel, SppOptimizer)
// restated over the renderer plugin (renderer.h) and the libpathed C ABI.  Same class and
// method names and the same edit semantics as include/render_session.h and
// include/scene_model.h of the reference, without GL interop: PBO handles are opaque integers
// naming two host pixel buffers, read back through getFramebuffer / getRadianceBuffer.
#pragma once

#include <stdint.h>

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pathed.h"
#include "renderer.h"
#include "scene.h"

namespace rays {

// render_session.h:36-39 (pbo is an opaque handle here)
struct RenderState {
    bool isRendering;
    unsigned int pbo;
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
    std::vector<uint8_t> getFramebuffer();   // the display pixel buffer: uchar4 sRGB of framebuffer.cu:35-40
    int64_t getAccumulatedSpp();
    void loadPreset(int preset, uint64_t seed, int64_t param);
    // the reference's scene `index` (scene.cu:14-176) from its scenes/ folder layout
    void loadReferenceScene(int index, const std::string &scenesDir);
    // render_session.h:87-101 writes out.exr whenever the spp count crosses a power of two;
    // here that dump is opt-in: a non-empty path enables it
    void setExrDumpPath(const std::string &path) { m_exrDumpPath = path; }
    Renderer &renderer() { return *m_pathTracer; }
    DeviceGlobals &deviceGlobals() { return *m_deviceGlobals; }

private:
    void applyAttributes(const SceneModelAttributes &attributes);
    void buildSceneObjects();
    void installRenderer();
    // render_session.cu:15-53 (PBOManager) without GL: two host pixel buffers, one written by
    // the pass in flight, the other displayed
    PboResource renderPbo() { return m_computeOnPrimary ? &m_pbo[0] : &m_pbo[1]; }
    unsigned int displayPbo() const { return m_computeOnPrimary ? m_pbo[1].id : m_pbo[0].id; }

    int m_width, m_height;
    SppOptimizer m_sppOptimizer;
    HeadlessPbo m_pbo[2];
    bool m_computeOnPrimary = false;
    bool m_rendering = false;
    bool m_resetRenderer = false;
    RendererType m_rendererType;
    RenderRecord m_currentRecord;
    std::vector<int> m_dirtyMaterials;  // edited before init()
    std::string m_exrDumpPath;
    pt_scene *m_hostScene = nullptr;
    std::unique_ptr<Renderer> m_pathTracer;
    std::unique_ptr<DeviceGlobals> m_deviceGlobals;
    std::unique_ptr<Scene> m_scene;
    std::unique_ptr<SceneModel> m_sceneModel;
};

}  // namespace rays
