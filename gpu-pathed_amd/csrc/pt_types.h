// pt_types.h -- plain data shared by host code and kernels (no HIP or torch types in the
// public ABI; these views are internal to libpathed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

// bit casts usable from host and device code
__host__ __device__ __forceinline__ float asFloat(int v) { return __builtin_bit_cast(float, v); }
__host__ __device__ __forceinline__ int asInt(float v) { return __builtin_bit_cast(int, v); }

enum { PT_MAT_LAMBERTIAN = 0, PT_MAT_MIRROR = 1, PT_MAT_GLASS = 2, PT_MAT_MICROFACET = 3, PT_MAT_NONE = 4 };

// Primitive tags (p1.w of a primitive record)
enum { PT_TAG_FLAT = -1, PT_TAG_SPHERE = -2 };

// 32-byte material record: a = (albedo.rgb, type), b = (emit.rgb, ior for glass / alpha for
// microfacet).  Emission is kept only for Lambertians (the reference's only emitters).
struct PtMaterial {
    float4 a;
    float4 b;
};

struct PtEnv {
    const float4 *rgba;       // texels, row-major, theta rows from the +y pole
    const float2 *rows;       // per-row alias tables: (prob, alias bits), width*height
    const float2 *marginal;   // alias table over rows, height
    const float *texelPdf;    // pmf * W*H / (2 pi^2)
    int width, height;
    int present;
};

// BVH node: four float4 (64 bytes)
//   n0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   n1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   n2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   n3 = (child0, child1, 0, 0) as int bits; child >= 0: internal node index,
//        child < 0: leaf ~((count-1) << 28 | first primitive)
static constexpr int PT_LEAF_COUNT_SHIFT = 28;
static constexpr int PT_LEAF_MAX = 4;
// 4-wide BVH node: 8 float4 = 128 bytes (one L2 line): lo.x, hi.x, lo.y, hi.y, lo.z, hi.z of
// the four child boxes (one float4 per bound and axis), then the four child references as int
// bits (>= 0 node index, < 0 leaf ~((count - 1) << PT_LEAF_COUNT_SHIFT | first), or
// PT_EMPTY_CHILD for an unused slot), then one unused float4.
static constexpr int PT_NODE_F4 = 8;
static constexpr int PT_EMPTY_CHILD = 0x7fffffff;

struct PtSceneView {
    const float4 *nodes;      // PT_NODE_F4 per 4-wide node
    const float4 *prims;      // 3 per primitive, BVH order
    const float4 *normals;    // 3 per triangle with per-vertex normals
    const PtMaterial *materials;
    const float4 *emitters;   // 3 per emitter (primitive record copy), reference light order
    int numPrims;
    int numNodes;
    int numEmitters;
    int numMaterials;
    PtEnv env;
};

struct PtCamera {
    float m[3][4];    // camera-to-world (columns: x axis, y axis, view, origin)
    float top, width, height, right;  // image plane extents at z = -1
    int W, H;
};

}  // namespace pt
