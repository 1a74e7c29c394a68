// pt_internal.h -- host-side internals of libpathed shared between translation units.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pathed.h"

namespace pt {

// Per-thread last error (pt_last_error).
void setError(const std::string &msg);
int fail(int code, const std::string &msg);

struct HostMaterial {
    int type;
    float albedo[3];
    float emit[3];
    float ior;
    float alpha;
};

struct HostCamera {
    float origin[3] = {0.f, 1.f, 6.8f};
    float target[3] = {0.f, 1.f, 0.f};
    float up[3] = {0.f, 1.f, 0.f};
    float vfov = 19.5f / 180.f * 3.14159265358979f;
    int flip = 0;
};

}  // namespace pt

// The opaque scene handle of the C ABI: flat host arrays.
struct pt_scene {
    std::vector<float> pos;      // 9 per triangle
    std::vector<float> nrm;      // 9 per triangle when hasNormals
    bool hasNormals = false;
    std::vector<int> triMat;
    std::vector<float> sph;      // cx cy cz r
    std::vector<int> sphMat;
    std::vector<pt::HostMaterial> mats;
    std::vector<float> env;      // rgba
    int envW = 0, envH = 0;
    int envSinWeight = 0;
    pt::HostCamera cam;
};

namespace pt {
// Emitter list in the reference order (scene_data.cu:92-115): emissive triangles by index,
// then emissive spheres.  Entries >= 0 are triangles, < 0 are ~sphere.
std::vector<int> emitterList(const pt_scene &s);
bool isEmissive(const HostMaterial &m);
// Camera-to-world matrix (transform.cu:7-33); false if the view direction equals up.
bool lookAt(const float o[3], const float t[3], const float u[3], int flip, float m[3][4]);
}  // namespace pt
