// scene.cpp -- Scene::setMaterialType (scene.h:38-61).
#include "scene.h"

namespace rays {

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

}  // namespace rays
