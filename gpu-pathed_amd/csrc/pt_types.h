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
    const float4 *rgba;       // texels, row-major, theta rows from the +y pole: (r, g, b,
                              // texel pdf = pmf * W*H / (2 pi^2)), one gather for both
    const float2 *rows;       // per-row alias tables: (prob, alias bits), width*height
    const float2 *marginal;   // alias table over rows, height
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
static constexpr int PT_LEAF_MAX = 2;
// 4-wide BVH node: 8 float4 = 128 bytes (one L2 line): lo.x, hi.x, lo.y, hi.y, lo.z, hi.z of
// the four child boxes (one float4 per bound and axis), then the four child references as int
// bits (>= 0 node index, < 0 leaf ~((count - 1) << PT_LEAF_COUNT_SHIFT | first), or
// PT_EMPTY_CHILD for an unused slot, whose box is inverted: lower bounds +1e30, upper -1e30),
// then one unused float4.
static constexpr int PT_NODE_F4 = 8;
static constexpr int PT_EMPTY_CHILD = 0x7fffffff;

// Segmented path queues.  Every queue of a chunk (the rays of a bounce, the shadow rays of a
// bounce) is cut into PT_QUEUE_SEGS segments of segLen slots, segLen a multiple of 64: segment
// s holds slots [s * segLen, s * segLen + count_s).  A path stays in its segment from bounce to
// bounce, so a segment never receives more records than it has slots.  Each segment has its
// own counters, PT_SEG_STRIDE ints (2 KiB) apart, and within a segment each field has its own
// 256-byte row (one int per bounce): device-scope atomics serialise per 64-byte block (about
// 88 per microsecond on one block, scaling with the number of blocks), so the work cursors and
// append counters that one kernel hits spread over PT_QUEUE_SEGS x fields blocks.
static constexpr int PT_MAX_BOUNCES = 64;
static constexpr int PT_QUEUE_SEGS = 64;
static constexpr int PT_QF = 8;  // counter fields per bounce
static constexpr int PT_SEG_STRIDE = PT_QF * PT_MAX_BOUNCES;
enum {
    PT_Q_RAYS = 0,          // rays of the bounce in this segment
    PT_Q_SHADOW = 1,        // shadow rays of the bounce
    PT_Q_FETCH_EXT = 2,     // work cursors of extend, shade and shadow
    PT_Q_FETCH_SHADE = 3,
    PT_Q_FETCH_SHADOW = 4,
    PT_Q_LADD_SHADE = 5,    // radiance additions (statistics)
    PT_Q_LADD_SHADOW = 6,
    PT_Q_RECORDS = 7        // guided: training records of the chunk (bounce 0's entry)
};
__host__ __device__ __forceinline__ int ptQueueField(int seg, int bounce, int field) {
    return seg * PT_SEG_STRIDE + field * PT_MAX_BOUNCES + bounce;
}
// Segment length for a queue of n records: the smallest multiple of 64 with
// PT_QUEUE_SEGS * segLen >= n.  A linear queue of n records fills the segments in order; the
// camera queue deals runs of 64 paths to the segments in turn (ptCameraSlot).
__host__ __device__ __forceinline__ int ptSegLen(int64_t n) {
    const int64_t per = (n + PT_QUEUE_SEGS - 1) / PT_QUEUE_SEGS;
    return (int)(per < 64 ? 64 : (per + 63) / 64 * 64);
}
// Camera queue: path p belongs to run p / 64, dealt to segment (p / 64) % PT_QUEUE_SEGS, so
// every segment holds runs from the whole chunk (similar work per segment) and a run of 64
// consecutive paths stays together (with the renderer's pixel-major path numbering, the
// samples of one pixel, or of a few neighbouring pixels).
__host__ __device__ __forceinline__ int ptCameraSlot(int p, int segLen) {
    const int run = p >> 6;
    return (run % PT_QUEUE_SEGS) * segLen + (run / PT_QUEUE_SEGS) * 64 + (p & 63);
}
// Path of camera-queue slot `off` of segment `seg` (the inverse of ptCameraSlot).
__host__ __device__ __forceinline__ int ptCameraPath(int seg, int off) { return ((off >> 6) * PT_QUEUE_SEGS + seg) * 64 + (off & 63); }
__host__ __device__ __forceinline__ int ptCameraCount(int seg, int n) {
    const int full = n / (64 * PT_QUEUE_SEGS);
    const int rest = n - full * 64 * PT_QUEUE_SEGS - seg * 64;
    return full * 64 + (rest < 0 ? 0 : rest > 64 ? 64 : rest);
}
// Slots a pool of `paths` records needs for any queue laid out in segments.
__host__ __device__ __forceinline__ int64_t ptQueueSlots(int64_t paths) { return (int64_t)PT_QUEUE_SEGS * ptSegLen(paths); }
// Path capacity allocated for `need` paths under a configured capacity of `wanted` paths:
// `need` rounded up to a power of two (at least 4096), capped at `wanted`, and rounded up to a
// multiple of PT_QUEUE_SEGS * 64, so that for every chunk of n <= capacity paths
// ptQueueSlots(n) <= capacity: per-slot buffers of `capacity` entries hold every queue.
__host__ __device__ __forceinline__ int64_t ptPoolCapacity(int64_t need, int64_t wanted) {
    int64_t want = 4096;
    while (want < need) want <<= 1;
    const int64_t cap = wanted > 4096 ? wanted : 4096;
    if (want > cap) want = cap;
    const int64_t unit = (int64_t)PT_QUEUE_SEGS * 64;
    return (want + unit - 1) / unit * unit;
}

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
