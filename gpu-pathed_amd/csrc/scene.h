// scene.h -- the reference's scene-side value classes (core/vec3.h, materials/types.h,
// core/camera.h, materials/params.h) and its Scene (include/scene.h:20-118): camera, per-
// material parameters, maximum depth and the NEE switch.  Geometry stays in the libpathed
// scene handle (pt_scene); the device copy is DeviceGlobals' (renderer.h).
#pragma once

#include <stddef.h>

#include <vector>

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

}  // namespace rays
