// pt_render.hip -- queue-driven wavefront path tracer for one MI355X (gfx950).
//
// One progressive pass of `spp` samples per pixel is cut into chunks of at most `pool`
// paths (whole sample rows of the image, or pixel ranges when the image alone exceeds the
// pool).  A chunk runs:
//
//   raygen            one thread per path: PCG32 stream, camera ray, path record
//   for each bounce:
//     extend          persistent waves fetch 64 rays at a time; closest hit through the BVH
//                     with a traversal stack in LDS
//     shade           emission with MIS weight, BSDF sample, light sample; appends the
//                     continuation ray and the shadow ray to compacted queues (one atomic
//                     per wave, slots from the wave's ballot prefix)
//     shadow          persistent any-hit traversal; adds the unoccluded light contribution
//   accumulate        per pixel: FP64 sum of the chunk's sample radiances
//
// Path state is structure-of-arrays in HBM, double-buffered between bounces so every queue
// kernel reads and writes contiguously:
//   rayO  float4 (origin, tmin)        rayD  float4 (direction, tmax)
//   state uint4  (rng lo, rng hi, path id, depth)
//   beta  float4 (throughput, pdf of the BSDF sample that led here; < 0: camera or delta)
//   hit   float4 (t, u, v, primitive slot bits)
//   shadow: sO float4 (origin, tmax), sD float4 (direction, path id), sL float4 (contribution)
//   Lpath float4 per path id (radiance of the path so far)
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "pt_device.h"
#include "pt_gpu.h"
#include "pt_guide.h"
#include "pt_internal.h"

namespace pt {

static constexpr int kMaxBounces = PT_MAX_BOUNCES;
// Queue counters: PT_QUEUE_SEGS blocks of PT_SEG_STRIDE ints (ptQueueField), then a few
// chunk-wide counters.
static constexpr int kSegCounters = PT_QUEUE_SEGS * PT_SEG_STRIDE;
// guided mode: records kept, sum of their weights (float), and the exclusive prefix sums of
// the per-segment record counts (PT_QUEUE_SEGS + 1 entries)
static constexpr int kRecValidSlot = kSegCounters + 2;
static constexpr int kWsumSlot = kSegCounters + 3;
static constexpr int kRecPrefixSlot = kSegCounters + 8;
static constexpr int kNumCounters = kRecPrefixSlot + PT_QUEUE_SEGS + 8;
// device statistics (uint64): per bounce rays, shadows, shade adds, shadow adds; totals
enum {
    S_RAYS = 0,
    S_SHADOW = kMaxBounces,
    S_LADD_SHADE = 2 * kMaxBounces,
    S_LADD_SHADOW = 3 * kMaxBounces,
    S_PATHS = 4 * kMaxBounces,
    S_TRACE = 5 * kMaxBounces,   // traversal counts (PT_OPT_TRACE_STATS): rays, nodes, primitives; closest hit, then shadow
    S_COUNT = S_TRACE + 6
};

static constexpr int kTraceThreads = 128;
// Default path-pool capacity: 2^26 paths (about 13 GB of SoA path state), a whole 1024^2 x
// 64 spp pass in one chunk, so that every bounce's queues stay long.
static constexpr int64_t kDefaultPool = 1ll << 26;
// Trace blocks of 128 threads per CU: 14 would fill the 7 waves per SIMD that the trace
// kernels' 72 VGPRs and 8 KiB of LDS stack per block allow; 11 measured best on C1 (10: 4156,
// 11: 4185, 12: 4123, 14: 4095 Mrays/s), fewer waves sharing each CU's caches.
static constexpr int kDefaultTraceBlocks = 11;
// Binary tree under the 4-wide BVH.
// LBVH measured faster than PLOC on the displaced icosphere at 1024^2, 64 spp (4467 vs 4348
// Mrays/s; extend 2053 vs 2160 us per launch, although PLOC's tree has 12% fewer 4-wide nodes).
static constexpr int kDefaultBvhBuilder = PT_BVH_LBVH;
static constexpr float kShadowEps = 1e-4f;   // path_tracer.cu:104-110
static constexpr float kBounceEps = 1e-3f;   // path_tracer.cu:177

// One render of a batch wavefront (renders of different scenes sharing one set of queues):
// where its BVH starts in the merged node array, its lights, environment and camera, its pixel
// grid and its accumulator.
struct BatchRender {
    int root;                 // 4-wide root node in the merged node array
    int emitBegin;            // first of its emitters in the merged emitter array
    int numEmitters;
    int maxDepth;
    int W, H, npix;
    uint64_t seed;
    PtEnv env;
    PtCamera cam;
    double *accum;            // 3 doubles per pixel
};

// A run of sample rows of one render inside a chunk of a batch wavefront: chunk paths
// [pathBegin, pathBegin + count * npix), sample-major; accumulate threads [accBegin,
// accBegin + npix).
struct BatchGroup {
    int pathBegin;
    int accBegin;
    int render;
    int npix;
    int sampleBegin;
    int count;
};

struct ChunkParams {
    int pixBegin;     // first pixel of the chunk
    int npix;         // pixels per sample row of the chunk
    int samples;      // sample rows in the chunk
    int sampleBase;   // global index of the first sample row
    int nPaths;       // npix * samples
    int maxDepth;
    int nee;
    int misSample;    // 1: light-sample MIS weight uses the BSDF sampling density (consistent MIS)
    int W, H;
    uint64_t seed;
    int segLen;       // queue segment length, ptSegLen(nPaths)
    // Path numbering of the chunk (the path index in state.z, the Lpath slot, the camera queue's
    // order): 1: pixel-major, path = pixel * samples + sample row, so the camera bounce's waves
    // trace the samples of one pixel together; 0: sample-major, path = sample row * npix + pixel
    int pixelMajor = 0;
    // batch wavefront only (nullptr otherwise): the chunk's groups, the batch's renders, and
    // per render the rays its paths spawned (continuation + shadow)
    const BatchGroup *groups = nullptr;
    int numGroups = 0;
    const struct BatchBucket *buckets = nullptr;   // per bucket of kBatchBucketShift paths
    const BatchRender *renders = nullptr;
    unsigned long long *renderRays = nullptr;
};

// Paths per bucket of the batch wavefront's path-to-group table.
static constexpr int kBatchBucketShift = 12;

// The group holding a bucket's first path, the first path of the group after it (a bucket
// rarely reaches past its group: one sample row of 64 x 64 pixels is a bucket already) and
// the group's render's PCG32 seed: one 32-byte load gives a path's render, pixel, sample
// and stream.
struct BatchBucket {
    int group, pathBegin, render, count;
    int sampleBegin, nextBegin;
    uint64_t seed;
};

// The group of a chunk path that lies past its bucket's group (in a group that starts inside
// the bucket).
__device__ __forceinline__ int batchLaterGroup(const ChunkParams &cp, int bucketGroup, int path) {
    int g = bucketGroup + 1;
    while (g + 1 < cp.numGroups && cp.groups[g + 1].pathBegin <= path) g++;
    return g;
}

// The render of a chunk path of a batch wavefront.
__device__ __forceinline__ int batchRenderOf(const ChunkParams &cp, int path) {
    const BatchBucket bk = cp.buckets[path >> kBatchBucketShift];
    return path < bk.nextBegin ? bk.render : cp.groups[batchLaterGroup(cp, bk.group, path)].render;
}

// Render, pixel, sample row and PCG32 seed of a chunk path of a batch wavefront (a group's
// paths are numbered pixel-major, the samples of a pixel side by side).
struct PathSource {
    int render, pixel, sample;
    uint64_t seed;
};
__device__ __forceinline__ PathSource batchPathSource(const ChunkParams &cp, int path) {
    const BatchBucket bk = cp.buckets[path >> kBatchBucketShift];
    PathSource ps;
    int begin = bk.pathBegin, count = bk.count, sample0 = bk.sampleBegin;
    ps.render = bk.render;
    ps.seed = bk.seed;
    if (path >= bk.nextBegin) {
        const BatchGroup gr = cp.groups[batchLaterGroup(cp, bk.group, path)];
        begin = gr.pathBegin;
        count = gr.count;
        sample0 = gr.sampleBegin;
        ps.render = gr.render;
        ps.seed = cp.renders[gr.render].seed;
    }
    // pixel-major inside a group: path - begin = pixel * count + sample row
    int s;
    ps.pixel = divmodEst(path - begin, count, &s);
    ps.sample = sample0 + s;
    return ps;
}

// Pixel within the chunk and sample row of chunk path `path`.  Pixel-major numbering puts the
// samples of one pixel (nearly the same ray: the same BVH nodes, the same primitive) in
// neighbouring lanes of the camera bounce's waves, and their radiance in neighbouring slots.
__device__ __forceinline__ int chunkPixelOf(const ChunkParams &cp, int path, int *sample) {
    if (cp.pixelMajor) return divmodEst(path, cp.samples, sample);
    int px;
    *sample = divmodEst(path, cp.npix, &px);
    return px;
}

// Batch wavefront path state: depth in the low 16 bits of state.w, render above.
__device__ __forceinline__ uint32_t packDepth(int depth, int render) { return (uint32_t)depth | ((uint32_t)render << 16); }

// ------------------------------------------------------------------------------------------
// wave helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int laneId() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ uint32_t lanesBelow(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
// Appends for every lane with `want`: one atomic per wave; returns this lane's slot.
__device__ __forceinline__ int waveAppend(bool want, int *counter) {
    const uint64_t m = __ballot(want);
    if (m == 0) return -1;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (laneId() == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    return want ? base + (int)lanesBelow(m) : -1;
}
// Appends to segment `seg` of a segmented queue: the record's slot, or -1 for lanes without one.
__device__ __forceinline__ int segAppend(bool want, int *counters, int seg, int bounce, int field, int segLen) {
    const int k = waveAppend(want, counters + ptQueueField(seg, bounce, field));
    return k < 0 ? -1 : seg * segLen + k;
}
// Appends to each lane's own segment `seg` (the lanes of a wave may hold records of different
// segments): one atomic per distinct segment among the appending lanes; the slot
// seg * segLen + index, or -1 for lanes without one.
__device__ __forceinline__ int segAppendLanes(bool want, int *counters, int seg, int bounce, int field, int segLen) {
    int k = -1;
    uint64_t todo = __ballot(want);
    while (todo != 0) {
        const int s = __shfl(seg, __ffsll((unsigned long long)todo) - 1);
        const bool mine = want && seg == s;
        const int r = waveAppend(mine, counters + ptQueueField(s, bounce, field));
        if (mine) k = s * segLen + r;
        todo &= ~__ballot(mine);
    }
    return k;
}
// segAppendLanes in two steps, so that the atomic's round trip overlaps other work: the issue
// (the same atomics, made by the first appending lane of each segment) and, later, the slot.
struct PendingAppend {
    int base;     // the atomic's result, in the leader lanes
    int leader;   // this lane's leader, or -1 for a lane without a record
    int below;    // this lane's index among its segment's appending lanes
};
__device__ __forceinline__ PendingAppend segAppendIssue(bool want, int *counters, int seg, int bounce, int field) {
    PendingAppend a{0, -1, 0};
    uint64_t todo = __ballot(want);
    while (todo != 0) {
        const int lead = __ffsll((unsigned long long)todo) - 1;
        const int s = __shfl(seg, lead);
        const bool mine = want && seg == s;
        const uint64_t m = __ballot(mine);
        if (laneId() == lead) a.base = atomicAdd(counters + ptQueueField(s, bounce, field), __popcll(m));
        if (mine) {
            a.leader = lead;
            a.below = (int)lanesBelow(m);
        }
        todo &= ~m;
    }
    return a;
}
// The slot (seg * segLen + index, or -1) of an issued append.  Every lane that was active at
// the issue must be active here: the leaders hand their atomic's result to their lanes.
__device__ __forceinline__ int segAppendSlot(const PendingAppend &a, int seg, int segLen) {
    const int base = __shfl(a.base, a.leader < 0 ? laneId() : a.leader);
    return a.leader < 0 ? -1 : seg * segLen + base + a.below;
}
// Adds each lane's v (> 0) to counter[key]: one atomic per distinct key among those lanes.
__device__ __forceinline__ void waveAddKeyed(int v, int key, unsigned long long *counter) {
    uint64_t todo = __ballot(v > 0);
    while (todo != 0) {
        const int k = __shfl(key, __ffsll((unsigned long long)todo) - 1);
        const bool mine = v > 0 && key == k;
        int sum = mine ? v : 0;
        for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off);
        if (laneId() == 0) atomicAdd(&counter[k], (unsigned long long)sum);
        todo &= ~__ballot(mine);
    }
}
// Adds a per-lane count once per wave (statistics counted over a persistent wave's lifetime
// cost one atomic instead of one per iteration).
__device__ __forceinline__ void waveSum(int v, int *counter) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (laneId() == 0 && v != 0) atomicAdd(counter, v);
}

// ------------------------------------------------------------------------------------------
// ray generation (camera.h:40-61, PCG32 stream per (pixel, sample))
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void cameraRay(const PtCamera &c, int row, int col, float sx, float sy, f3 *o, f3 *d) {
    const float xc = fdiv(col + sx, (float)c.W);
    const float yc = fdiv(row + sy, (float)c.H);
    const float y = yc * c.height - c.top;
    const float x = xc * c.width - c.right;
    const f3 dl = normalize(mk(x, y, -1.f));
    *o = mk(c.m[0][3], c.m[1][3], c.m[2][3]);
    *d = mk(c.m[0][0] * dl.x + c.m[0][1] * dl.y + c.m[0][2] * dl.z, c.m[1][0] * dl.x + c.m[1][1] * dl.y + c.m[1][2] * dl.z,
            c.m[2][0] * dl.x + c.m[2][1] * dl.y + c.m[2][2] * dl.z);
}

// Clears every queue counter and makes field `field` of bounce 0 count a queue of n records
// with segment length segLen: the camera layout (ptCameraSlot) or, with camera = 0, the linear
// one (segment s holds min(segLen, n - s segLen) records).  One block per segment plus one for
// the chunk-wide counters.
__global__ __launch_bounds__(256) void k_queue_reset(int *counters, int field, int n, int segLen, int camera) {
    if (blockIdx.x < PT_QUEUE_SEGS) {
        const int seg = (int)blockIdx.x;
        int *q = counters + seg * PT_SEG_STRIDE;
        const int cnt = camera ? ptCameraCount(seg, n) : max(0, min(segLen, n - seg * segLen));
        const int at = ptQueueField(0, 0, field);
        for (int k = threadIdx.x; k < PT_SEG_STRIDE; k += blockDim.x) q[k] = k == at ? cnt : 0;
    } else {
        for (int k = kSegCounters + threadIdx.x; k < kNumCounters; k += blockDim.x) counters[k] = 0;
    }
}

// Writes only the camera rays' directions: their origin (the camera's), throughput (1),
// previous pdf (-1, none), path index, depth (0) and RNG state (the stream after the two jitter
// draws) are recomputed by the bounce-0 kernels, and shade writes the paths' radiance at bounce 0.
__global__ __launch_bounds__(256) void k_raygen(ChunkParams cp, PtCamera cam, float4 *rayD) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cp.nPaths) return;
    int col, s;
    const int rem = chunkPixelOf(cp, i, &s);
    const int pixel = cp.pixBegin + rem;
    const int sample = cp.sampleBase + s;
    const int row = divmodEst(pixel, cp.W, &col);
    Rng rng = seedRng(cp.seed, (uint32_t)pixel, (uint32_t)sample);
    const float sx = rng.uniform(), sy = rng.uniform();
    f3 o, d;
    cameraRay(cam, row, col, sx, sy, &o, &d);
    rayD[ptCameraSlot(i, cp.segLen)] = f4(d, FLT_MAX);
}

// Batch wavefront: every path of the chunk from its render's camera; like k_raygen it stores
// only the direction (the bounce-0 kernels find the render, its camera origin and root node
// from the path index).
__global__ __launch_bounds__(256) void k_raygen_batch(ChunkParams cp, float4 *rayD) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cp.nPaths) return;
    const PathSource ps = batchPathSource(cp, i);
    const BatchRender &R = cp.renders[ps.render];
    int col;
    const int row = divmodEst(ps.pixel, R.W, &col);
    Rng rng = seedRng(ps.seed, (uint32_t)ps.pixel, (uint32_t)ps.sample);
    const float sx = rng.uniform(), sy = rng.uniform();
    f3 o, d;
    cameraRay(R.cam, row, col, sx, sy, &o, &d);
    rayD[ptCameraSlot(i, cp.segLen)] = f4(d, FLT_MAX);
}

// The camera origin of a batch render.
__device__ __forceinline__ float4 batchCameraOrigin(const BatchRender &R) { return make_float4(R.cam.m[0][3], R.cam.m[1][3], R.cam.m[2][3], 0.f); }

// The camera-ray state k_raygen does not store: the RNG after the pixel's two jitter draws.
__device__ __forceinline__ Rng cameraRng(const ChunkParams &cp, int path) {
    int s;
    const int pixel = cp.pixBegin + chunkPixelOf(cp, path, &s);
    Rng rng = seedRng(cp.seed, (uint32_t)pixel, (uint32_t)(cp.sampleBase + s));
    rng.next();
    rng.next();
    return rng;
}

// ------------------------------------------------------------------------------------------
// BVH traversal (closest hit / any hit) over the 4-wide BVH: one 128-byte node fetch tests
// four child boxes; the hit children are ordered near to far (a 5-exchange sorting network),
// the nearest is visited next and the others are pushed far-first on a per-lane stack (any-hit
// queries too: nearer occluders are found sooner).
//
// The stack has two levels: the top LDS entries of each lane (stack[level * blockDim + lane],
// LDS entries per lane) and, below them, a spill area in global memory interleaved by lane
// (spill[(level - LDS) * lanes + lane]), sized for the deepest stack the tree can need.  The
// small LDS part keeps the trace kernels' occupancy high; deep stacks are rare.
//
// Scheduling: a wave runs traversal rounds (descent to a leaf, the leaf's primitives, one pop)
// on every lane that holds a ray, and lanes whose ray has finished take new rays from the
// queue once enough of them are idle.  Incoherent rays need very different numbers of
// rounds; refilling idle lanes keeps most of the wave busy instead of waiting for the longest
// ray of a fixed batch of 64.  Every ray visits nodes and primitives in the same order as a
// recursive front-to-back traversal, so hits do not depend on the schedule.
// ------------------------------------------------------------------------------------------
struct TraceHit {
    float t, u, v;
    int prim;
};

typedef __attribute__((address_space(3))) int LdsInt;

struct TraceStack {
    LdsInt *lds;    // this lane's first LDS entry; entries blockDim.x apart
    int *spill;     // this lane's first spill entry; entries spillStride apart
    int spillCap;   // spill entries per lane
    int spillStride;
};

__device__ __forceinline__ void cswap(float &ta, int &ra, float &tb, int &rb) {
    const bool s = tb < ta;
    const float t = s ? tb : ta;
    const int r = s ? rb : ra;
    tb = s ? ta : tb;
    rb = s ? ra : rb;
    ta = t;
    ra = r;
}

// One ray in flight on a lane: the slab-test and watertight-test set-ups, the search interval
// (tmax is the closest hit's distance once prim >= 0), the closest hit's primitive and
// barycentrics, and the traversal position (node >= 0: a 4-wide node; < 0: the rest of a
// leaf, ~((remaining - 1) << PT_LEAF_COUNT_SHIFT | next primitive); sp: stack depth).
struct LaneRay {
    WtRay wr;
    f3 d;
    float ix, iy, iz, ox, oy, oz;
    uint32_t nX, nY, nZ;
    float tmin, tmax;
    float u, v;
    int prim;
    int node, sp;
};

__device__ __forceinline__ void laneRayStart(LaneRay &r, f3 o, f3 d, float tmin, float tmax, int root = 0) {
    r.wr = wtSetup(o, d);
    r.d = d;
    r.ix = 1.f / (fabsf(d.x) > 1e-20f ? d.x : copysignf(1e-20f, d.x));
    r.iy = 1.f / (fabsf(d.y) > 1e-20f ? d.y : copysignf(1e-20f, d.y));
    r.iz = 1.f / (fabsf(d.z) > 1e-20f ? d.z : copysignf(1e-20f, d.z));
    r.ox = o.x * r.ix;
    r.oy = o.y * r.iy;
    r.oz = o.z * r.iz;
    // Per axis the near plane of every box is its lower bound when the direction is positive:
    // the slab test reads the near and far float4 of each axis directly (byte offsets within
    // the 128-byte node), which gives the same entry and exit distances as min/max of both.
    r.nX = r.ix < 0.f ? 16u : 0u;
    r.nY = r.iy < 0.f ? 48u : 32u;
    r.nZ = r.iz < 0.f ? 80u : 64u;
    r.tmin = tmin;
    r.tmax = tmax;
    r.u = 0.f;
    r.v = 0.f;
    r.prim = INT32_MIN;
    r.node = root;
    r.sp = 0;
}

__device__ __forceinline__ TraceHit laneRayHit(const LaneRay &r) {
    TraceHit th;
    th.t = r.prim >= 0 ? r.tmax : FLT_MAX;
    th.u = r.u;
    th.v = r.v;
    th.prim = r.prim;
    return th;
}

// One traversal round of a lane's ray.  The wave descends through 4-wide nodes, nearest hit
// child first, popping the stack after a node whose children all miss, until (nearly) every
// lane has found a leaf or exhausted its stack; a lane that finds a leaf while others are
// still searching parks it and keeps descending (speculatively, with the interval it has), up
// to a second leaf.  Then the parked leaves' primitives are tested together.  Leaves are tested in
// the order a depth-first traversal reaches them, and a node visited only because its leaf
// was parked holds no primitive at or below the final tmax, so the hit does not depend on the
// schedule.  Returns true when the ray is finished: the stack is exhausted, or an any-hit query
// found an occluder.
static constexpr int kTraceDone = INT32_MIN;   // LaneRay::node once the stack is exhausted
// Searching lanes a wave leaves behind to test its parked leaves: waiting for the last few
// lanes of a round costs more than another round (C1: 4 -> 3398, 12 -> 3413, 20 -> 3347,
// 32 -> 3250 Mrays/s; 0, wait for every lane: 3207; on the current tree 8 -> 4395,
// 12 -> 4420, 16 -> 4408)
static constexpr int kDescentTail = 12;
// Node steps per check of the descent's exit conditions: two steps save the ballots of every
// other step (C1 4452 -> 4498 Mrays/s with the trace kernels' register cap raised from 72 to
// 80, six waves per SIMD, which the default 11 blocks per CU do not exceed; at 72 registers the
// second step spills).
static constexpr int kDescentSteps = 2;

// Per-lane traversal counts of the instrumented trace kernels (STATS).
struct TraceCount {
    int rays = 0, nodes = 0, prims = 0;
};

template <bool ANY, int STACK, bool STATS = false>
__device__ __forceinline__ bool traceRound(const PtSceneView &sc, LaneRay &r, const TraceStack &ts, int *overflow, TraceCount *tc = nullptr) {
    const int stride = blockDim.x;
    // general push (LDS, then the spill area); the common closest-hit case writes LDS directly
    auto push = [&](int ref) {
        if (r.sp < STACK) {
            ts.lds[r.sp * stride] = ref;
            r.sp++;
        } else if (r.sp - STACK < ts.spillCap) {
            ts.spill[(size_t)(r.sp - STACK) * ts.spillStride] = ref;
            r.sp++;
        } else {
            atomicAdd(overflow, 1);
        }
    };
    auto pop = [&]() {
        if (r.sp == 0) {
            r.node = kTraceDone;
        } else {
            r.sp--;
            // an LDS read at a clamped level and, only below the LDS part, the spill read (the
            // LDS pointer's address space keeps the two from becoming one flat load)
            int v = ts.lds[min(r.sp, STACK - 1) * stride];
            if (r.sp >= STACK) v = ts.spill[(size_t)(r.sp - STACK) * ts.spillStride];
            r.node = v;
        }
    };
    const char *nodes = (const char *)sc.nodes;
    int leaf = 0;   // parked leaf reference (< 0), or 0
    while (true) {
        // the descent ends when no lane is still looking for a leaf, or when at most
        // kDescentTail lanes are and at least one lane holds a parked leaf (the searching
        // lanes carry on next round); every pass through this loop moves a searching lane on
        const uint64_t searching = __ballot(leaf == 0 && r.node != kTraceDone);
        if (searching == 0) break;
        if (__popcll(searching) <= kDescentTail && __ballot(leaf != 0) != 0) break;
        // kDescentSteps node steps per check of the exit conditions: a lane that parks a leaf
        // in the first step keeps descending internal nodes in the next (as it would while
        // other lanes search) and idles at its second leaf, so only the checks' cost changes
#pragma unroll
        for (int step = 0; step < kDescentSteps; step++) {
        if (r.node >= 0) {
            if constexpr (STATS) tc->nodes++;
            // 32-bit offsets from the array base (a node array stays below 4 GiB); a node is
            // 128 bytes, so the near bound's offset is an OR and the far bound of the same axis
            // (the other float4 of the pair) is that offset XOR 16
            const uint32_t base = (uint32_t)r.node * (uint32_t)(PT_NODE_F4 * 16);
            const uint32_t ax = base | r.nX, ay = base | r.nY, az = base | r.nZ;
            const float4 nx = *(const float4 *)(nodes + ax), fx = *(const float4 *)(nodes + (ax ^ 16u));
            const float4 ny = *(const float4 *)(nodes + ay), fy = *(const float4 *)(nodes + (ay ^ 16u));
            const float4 nz = *(const float4 *)(nodes + az), fz = *(const float4 *)(nodes + (az ^ 16u));
            const int4 rf = *(const int4 *)(nodes + (base + 96u));
            float tn[4];
            int ref[4] = {rf.x, rf.y, rf.z, rf.w};
            const float bnx[4] = {nx.x, nx.y, nx.z, nx.w}, bfx[4] = {fx.x, fx.y, fx.z, fx.w};
            const float bny[4] = {ny.x, ny.y, ny.z, ny.w}, bfy[4] = {fy.x, fy.y, fy.z, fy.w};
            const float bnz[4] = {nz.x, nz.y, nz.z, nz.w}, bfz[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                // an unused slot holds an inverted box (lower bounds +1e30, upper -1e30): it misses
                const float lo = fmaxf(fmaxf(bnx[k] * r.ix - r.ox, bny[k] * r.iy - r.oy), fmaxf(bnz[k] * r.iz - r.oz, r.tmin));
                const float hi = fminf(fminf(bfx[k] * r.ix - r.ox, bfy[k] * r.iy - r.oy), fminf(bfz[k] * r.iz - r.oz, r.tmax));
                tn[k] = lo <= hi ? lo : INFINITY;
            }
            // the references are used only when a child is hit; this empty statement keeps
            // their load beside the box loads instead of after the box test, a second round
            // trip on the common path
            asm volatile("" : "+v"(ref[0]), "+v"(ref[1]), "+v"(ref[2]), "+v"(ref[3]));
            // hit children nearest first (any-hit queries too: occluders nearer the origin are
            // found sooner)
            cswap(tn[0], ref[0], tn[1], ref[1]);
            cswap(tn[2], ref[2], tn[3], ref[3]);
            cswap(tn[0], ref[0], tn[2], ref[2]);
            cswap(tn[1], ref[1], tn[3], ref[3]);
            cswap(tn[1], ref[1], tn[2], ref[2]);
            if (tn[0] == INFINITY) {
                pop();
            } else {
                // hit children are a prefix of the sorted slots: push all but the nearest,
                // the second nearest on top
                const int more = (tn[1] != INFINITY) + (tn[2] != INFINITY) + (tn[3] != INFINITY);
                if (r.sp + 3 <= STACK) {
                    if (more >= 3) ts.lds[r.sp * stride] = ref[3];
                    if (more >= 2) ts.lds[(r.sp + more - 2) * stride] = ref[2];
                    if (more >= 1) ts.lds[(r.sp + more - 1) * stride] = ref[1];
                    r.sp += more;
                } else {
                    if (more >= 3) push(ref[3]);
                    if (more >= 2) push(ref[2]);
                    if (more >= 1) push(ref[1]);
                }
                r.node = ref[0];
            }
        } else if (leaf == 0 && r.node != kTraceDone) {
            leaf = r.node;
            pop();
        }
        // lanes holding a parked leaf and a second one, or done, wait for the others
        }
    }
    if (leaf != 0) {
        const int e = ~leaf;
        const int first = e & ((1 << PT_LEAF_COUNT_SHIFT) - 1);
        const int cnt = (e >> PT_LEAF_COUNT_SHIFT) + 1;
        for (int j = 0; j < cnt; j++) {
            // all three records of the slot (spheres have three too) are fetched before the
            // type test, so the loads overlap
            // 32-bit byte offsets from the array base (the primitive records stay below 4 GiB:
            // pt_renderer_set_scene's limit), like the node loads
            const uint32_t off = (uint32_t)(first + j) * 48u;
            if constexpr (STATS) tc->prims++;
            const char *pr = (const char *)sc.prims;
            const float4 p0 = *(const float4 *)(pr + off);
            float4 p1 = *(const float4 *)(pr + (off + 16u)), p2 = *(const float4 *)(pr + (off + 32u));
            // keeps the triangle-only components' loads ahead of the type test (the compiler
            // would otherwise issue them inside the triangle branch, after the first ones return)
            asm volatile("" : "+v"(p1.y), "+v"(p1.z), "+v"(p2.x), "+v"(p2.y), "+v"(p2.z));
            float t, u = 0.f, v = 0.f;
            bool h;
            if (asInt(p1.w) == PT_TAG_SPHERE) {
                h = sphereHit(xyz(p0), p1.x, r.wr.o, r.d, r.tmin, r.tmax, &t);
            } else {
                h = wtHit(r.wr, xyz(p0), xyz(p1), xyz(p2), r.tmin, r.tmax, &t, &u, &v);
            }
            if (h) {
                r.tmax = t;
                r.u = u;
                r.v = v;
                r.prim = first + j;
                if (ANY) return true;
            }
        }
    }
    return r.node == kTraceDone;
}

__device__ __forceinline__ TraceStack traceStack(int *sstack, int *spill, int spillCap) {
    TraceStack ts;
    ts.lds = (LdsInt *)(sstack + threadIdx.x);
    ts.spillStride = gridDim.x * blockDim.x;
    ts.spill = spill + blockIdx.x * blockDim.x + threadIdx.x;
    ts.spillCap = spillCap;
    return ts;
}

// Work distribution over a segmented queue (the trace kernels and shade).  Wave w's home
// segment is w % PT_QUEUE_SEGS; it claims records there with one atomic on that segment's
// cursor and, once its home is exhausted, helps up to kQueueSteal following segments before
// it leaves.  Every segment has home waves (each kernel launches at least PT_QUEUE_SEGS
// waves), so every record is claimed; stealing only evens out the segments' ends.
static constexpr int kQueueSteal = 8;
// Queue records a shade wave claims at a time, per lane.
static constexpr int kShadeClaimWords = 2;

struct SegCursor {
    int seg;    // segment being drained (wave-uniform)
    int left;   // segments this wave may still visit
    int n;      // the segment's record count once read (-1: not yet)
};

__device__ __forceinline__ int homeSegment() { return (int)(((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % PT_QUEUE_SEGS); }
__device__ __forceinline__ SegCursor segStart() { return SegCursor{homeSegment(), 1 + kQueueSteal, -1}; }

// Claims up to `want` records of the bounce's queue (count field `countField`, cursor field
// `fetchField`): returns how many were granted, the first at slot *slot, or 0 when this wave
// has no segment left.  Wave-uniform.  A queue being drained is complete (its kernel appends
// to other queues), so a segment's count is read once per visit.
__device__ __forceinline__ int segClaim(int *counters, int bounce, int countField, int fetchField, int segLen, int want, SegCursor &c,
                                        int *slot) {
    while (c.left > 0) {
        int *seg = counters + c.seg * PT_SEG_STRIDE;
        if (c.n < 0) c.n = seg[ptQueueField(0, bounce, countField)];
        if (c.n > 0) {
            int b = 0;
            if (laneId() == 0) b = atomicAdd(&seg[ptQueueField(0, bounce, fetchField)], want);
            b = __shfl(b, 0);
            if (b < c.n) {
                *slot = c.seg * segLen + b;
                return min(want, c.n - b);
            }
        }
        c.seg = c.seg + 1 == PT_QUEUE_SEGS ? 0 : c.seg + 1;
        c.left--;
        c.n = -1;
    }
    return 0;
}

// Traces the bounce's queue (count field countField, cursor field fetchField): load(i, o, d,
// tmin, tmax, root) reads the ray in slot i (root: the BVH node its traversal starts at),
// finish(i, hit, th) stores its result.  A wave takes new
// rays once `refill` or more of its lanes are idle, as a contiguous run of one segment so that
// the rays in flight on a CU stay neighbours, and leaves once its segments are exhausted and
// its lanes idle.
template <bool ANY, int STACK, bool STATS = false, class Load, class Finish>
__device__ __forceinline__ void traceQueue(const PtSceneView &sc, int *counters, int bounce, int countField, int fetchField, int segLen,
                                           int refill, const TraceStack &ts, int *overflow, Load load, Finish finish,
                                           unsigned long long *tstat = nullptr) {
    SegCursor cur = segStart();
    if (sc.numNodes == 0) {
        TraceHit th;
        th.t = FLT_MAX; th.u = 0.f; th.v = 0.f; th.prim = INT32_MIN;
        int slot = 0, got;
        while ((got = segClaim(counters, bounce, countField, fetchField, segLen, 64, cur, &slot)) > 0)
            if (laneId() < got) finish(slot + laneId(), false, th);
        return;
    }
    LaneRay r;
    int ray = -1;
    TraceCount tc;
    while (true) {
        const uint64_t idle = __ballot(ray < 0);
        if (cur.left > 0 && __popcll(idle) >= refill) {
            int slot = 0;
            const int got = segClaim(counters, bounce, countField, fetchField, segLen, __popcll(idle), cur, &slot);
            if (ray < 0) {
                const int k = (int)lanesBelow(idle);
                if (k < got) {
                    ray = slot + k;
                    f3 o, d;
                    float tmin, tmax;
                    int root = 0;
                    load(ray, o, d, tmin, tmax, root);
                    laneRayStart(r, o, d, tmin, tmax, root);
                    if constexpr (STATS) tc.rays++;
                }
            }
        }
        // all lanes idle after a refill attempt means this wave's segments are exhausted
        if (__ballot(ray >= 0) == 0) break;
        if (ray >= 0 && traceRound<ANY, STACK, STATS>(sc, r, ts, overflow, &tc)) {
            finish(ray, r.prim >= 0, laneRayHit(r));
            ray = -1;
        }
    }
    if constexpr (STATS) {
        int v[3] = {tc.rays, tc.nodes, tc.prims};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            for (int off = 32; off >= 1; off >>= 1) v[k] += __shfl_xor(v[k], off);
            if (laneId() == 0) atomicAdd(&tstat[k], (unsigned long long)v[k]);
        }
    }
}

// Persistent closest-hit kernel over the rays of `bounce`.  camO.w != 0: the rays start at
// camO (camera rays, tmin 0) and rayO is not read.  BATCH (a batch wavefront): the traversal
// starts at the root node of the ray's render; at bounce 0 the render comes from the path
// index (chunk groups) and the origin is its camera's, later rayO.w holds the render.
// STATS: the instrumented build that adds its rays, node visits and primitive tests to
// tstat[0..2].
template <int STACK, bool BATCH = false, bool STATS = false>
__global__ __launch_bounds__(kTraceThreads, STACK <= 16 ? 6 : 5) void k_extend(PtSceneView sc, int *counters, int bounce, int segLen,
                                                                              const float4 *rayO, const float4 *rayD, float4 camO, float4 *hits,
                                                                              int refill, int *spill, int spillCap, int *overflow,
                                                                              ChunkParams cp = ChunkParams{}, unsigned long long *tstat = nullptr) {
    __shared__ int sstack[STACK * kTraceThreads];
    const TraceStack ts = traceStack(sstack, spill, spillCap);
    traceQueue<false, STACK, STATS>(
        sc, counters, bounce, PT_Q_RAYS, PT_Q_FETCH_EXT, segLen, refill, ts, overflow,
        [&](int i, f3 &o, f3 &d, float &tmin, float &tmax, int &root) {
            const float4 rd = rayD[i];
            if constexpr (BATCH) {
                if (bounce == 0) {
                    int off;
                    const int seg = divmodEst(i, segLen, &off);
                    const int p = ptCameraPath(seg, off);
                    const BatchRender &R = cp.renders[batchRenderOf(cp, p)];
                    o = xyz(batchCameraOrigin(R));
                    tmin = 0.f;
                    root = R.root;
                } else {
                    const float4 ro = rayO[i];
                    o = xyz(ro);
                    tmin = kBounceEps;
                    root = cp.renders[asInt(ro.w)].root;
                }
            } else if (camO.w != 0.f) {
                o = xyz(camO);
                tmin = 0.f;
            } else {
                const float4 ro = rayO[i];
                o = xyz(ro);
                tmin = ro.w;
            }
            d = xyz(rd);
            tmax = rd.w;
        },
        [&](int i, bool, const TraceHit &th) { hits[i] = make_float4(th.t, th.u, th.v, asFloat(th.prim)); }, tstat);
}

// Wave stack entries of the packet traversal, and its blocks per CU.
static constexpr int kPacketStack = 64;
static constexpr int kPacketBlocksPerCU = 16;

// Camera bounce of a pixel-major chunk whose runs of 64 camera rays are one pixel's samples
// (or a few neighbouring pixels' samples),
// nearly the same ray: a wave traverses the BVH once for the whole run (packet traversal on a
// 64-wide wave, Wald et al. 2001, Aila and Laine 2009): one node sequence and one stack for
// the wave, each lane testing a node's child boxes and a leaf's primitives against its own ray
// and closest hit so far.  A child is visited when some lane's ray enters it before that
// lane's closest hit; the visited children are taken nearest-first for the run's first ray,
// those it misses last.  Every lane sees every leaf whose box any lane enters, so each ray
// gets the closest hit among all primitives of the leaves its own box tests reach (a tie of
// equal distances may resolve to another primitive than in the per-lane traversal).  Runs
// are claimed 64 at a time; the wave stack holds kPacketStack entries (launched only when the
// tree's 3 x depth + 1 fits, with an overflow counter besides).
__global__ __launch_bounds__(kTraceThreads, 6) void k_extend_packet(PtSceneView sc, int *counters, int bounce, int segLen, const float4 *rayO,
                                                                   const float4 *rayD, float4 camO, float4 *hits, int *overflow) {
    __shared__ int wstack[kTraceThreads / 64][kPacketStack];
    int *ws = wstack[threadIdx.x >> 6];
    const char *nodes = (const char *)sc.nodes;
    SegCursor cur = segStart();
    while (true) {
        int slot = 0;
        const int got = segClaim(counters, bounce, PT_Q_RAYS, PT_Q_FETCH_EXT, segLen, 64, cur, &slot);
        if (got == 0) break;
        const bool valid = laneId() < got;
        const int i = slot + (valid ? laneId() : 0);
        const float4 rd = rayD[i];
        // camera rays (camO.w != 0) start at the camera; later ones carry origin and tmin
        const float4 ro = camO.w != 0.f ? make_float4(camO.x, camO.y, camO.z, 0.f) : rayO[i];
        LaneRay r;
        // a lane without a ray gets an empty interval: every box and primitive test fails
        laneRayStart(r, xyz(ro), xyz(rd), ro.w, valid ? rd.w : -1.f, 0);
        int node = 0, sp = 0;   // wave-uniform
        while (true) {
            if (node >= 0) {
                const uint32_t base = (uint32_t)node * (uint32_t)(PT_NODE_F4 * 16);
                const uint32_t ax = base | r.nX, ay = base | r.nY, az = base | r.nZ;
                const float4 nx = *(const float4 *)(nodes + ax), fx = *(const float4 *)(nodes + (ax ^ 16u));
                const float4 ny = *(const float4 *)(nodes + ay), fy = *(const float4 *)(nodes + (ay ^ 16u));
                const float4 nz = *(const float4 *)(nodes + az), fz = *(const float4 *)(nodes + (az ^ 16u));
                const int4 rf = *(const int4 *)(nodes + (base + 96u));
                const float bnx[4] = {nx.x, nx.y, nx.z, nx.w}, bfx[4] = {fx.x, fx.y, fx.z, fx.w};
                const float bny[4] = {ny.x, ny.y, ny.z, ny.w}, bfy[4] = {fy.x, fy.y, fy.z, fy.w};
                const float bnz[4] = {nz.x, nz.y, nz.z, nz.w}, bfz[4] = {fz.x, fz.y, fz.z, fz.w};
                int ref[4] = {__builtin_amdgcn_readfirstlane(rf.x), __builtin_amdgcn_readfirstlane(rf.y),
                              __builtin_amdgcn_readfirstlane(rf.z), __builtin_amdgcn_readfirstlane(rf.w)};
                float key[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float lo = fmaxf(fmaxf(bnx[k] * r.ix - r.ox, bny[k] * r.iy - r.oy), fmaxf(bnz[k] * r.iz - r.oz, r.tmin));
                    const float hi = fminf(fminf(bfx[k] * r.ix - r.ox, bfy[k] * r.iy - r.oy), fminf(bfz[k] * r.iz - r.oz, r.tmax));
                    const bool in = lo <= hi;
                    // the run's first lane orders the children; one that only other lanes enter
                    // comes after its hits, one that no lane enters is not visited
                    const float first = __builtin_amdgcn_readfirstlane(__float_as_int(in ? lo : INFINITY)) == __float_as_int(INFINITY)
                                            ? FLT_MAX
                                            : __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(in ? lo : INFINITY)));
                    key[k] = __ballot(in) != 0 ? first : INFINITY;
                }
                cswap(key[0], ref[0], key[1], ref[1]);
                cswap(key[2], ref[2], key[3], ref[3]);
                cswap(key[0], ref[0], key[2], ref[2]);
                cswap(key[1], ref[1], key[3], ref[3]);
                cswap(key[1], ref[1], key[2], ref[2]);
                if (key[0] != INFINITY) {
                    const int more = (key[1] != INFINITY) + (key[2] != INFINITY) + (key[3] != INFINITY);
                    if (sp + more <= kPacketStack) {
                        // the farthest lowest, the second nearest on top
                        if (more >= 3) ws[sp] = ref[3];
                        if (more >= 2) ws[sp + more - 2] = ref[2];
                        if (more >= 1) ws[sp + more - 1] = ref[1];
                        sp += more;
                    } else if (laneId() == 0) {
                        atomicAdd(overflow, 1);
                    }
                    node = ref[0];
                    continue;
                }
            } else {
                // a leaf: every lane tests its primitives (as traceRound's leaf phase)
                const int e = ~node;
                const int first = e & ((1 << PT_LEAF_COUNT_SHIFT) - 1);
                const int cnt = (e >> PT_LEAF_COUNT_SHIFT) + 1;
                for (int j = 0; j < cnt; j++) {
                    const uint32_t off = (uint32_t)(first + j) * 48u;
                    const char *pr = (const char *)sc.prims;
                    const float4 p0 = *(const float4 *)(pr + off);
                    float4 p1 = *(const float4 *)(pr + (off + 16u)), p2 = *(const float4 *)(pr + (off + 32u));
                    asm volatile("" : "+v"(p1.y), "+v"(p1.z), "+v"(p2.x), "+v"(p2.y), "+v"(p2.z));
                    float t, u = 0.f, v = 0.f;
                    bool h;
                    if (asInt(p1.w) == PT_TAG_SPHERE) {
                        h = sphereHit(xyz(p0), p1.x, r.wr.o, r.d, r.tmin, r.tmax, &t);
                    } else {
                        h = wtHit(r.wr, xyz(p0), xyz(p1), xyz(p2), r.tmin, r.tmax, &t, &u, &v);
                    }
                    if (h) {
                        r.tmax = t;
                        r.u = u;
                        r.v = v;
                        r.prim = first + j;
                    }
                }
            }
            if (sp == 0) break;
            sp--;
            node = __builtin_amdgcn_readfirstlane(ws[sp]);
        }
        if (valid) {
            const TraceHit th = laneRayHit(r);
            hits[i] = make_float4(th.t, th.u, th.v, asFloat(th.prim));
        }
    }
}

// Persistent any-hit kernel for the shadow rays of `bounce` (interval [1e-4, tmax]); an
// unoccluded ray adds its light contribution to its path.
// BATCH: sL.w holds the ray's render, whose root node starts the traversal.
template <int STACK, bool BATCH = false, bool STATS = false>
__global__ __launch_bounds__(kTraceThreads, STACK <= 16 ? 6 : 5) void k_shadow(PtSceneView sc, int *counters, int bounce, int segLen,
                                                                              const float4 *sO, const float4 *sD, const float4 *sL, float4 *Lpath,
                                                                              int refill, int *spill, int spillCap, int *overflow,
                                                                              const BatchRender *renders = nullptr, unsigned long long *tstat = nullptr) {
    __shared__ int sstack[STACK * kTraceThreads];
    // each lane's path index, kept from the ray's load (a re-read after the traversal would be a
    // dependent round trip ahead of the radiance update)
    __shared__ int sPath[kTraceThreads];
    const TraceStack ts = traceStack(sstack, spill, spillCap);
    int adds = 0;
    traceQueue<true, STACK, STATS>(
        sc, counters, bounce, PT_Q_SHADOW, PT_Q_FETCH_SHADOW, segLen, refill, ts, overflow,
        [&](int i, f3 &o, f3 &d, float &tmin, float &tmax, int &root) {
            const float4 so = sO[i], sd = sD[i];
            o = xyz(so);
            d = xyz(sd);
            tmin = kShadowEps;
            tmax = so.w;
            sPath[threadIdx.x] = asInt(sd.w);
            if constexpr (BATCH) root = renders[asInt(((const float *)&sL[i])[3])].root;
        },
        [&](int i, bool occluded, const TraceHit &) {
            if (!occluded) {
                const int pid = sPath[threadIdx.x];
                const float4 l = sL[i];
                float4 acc = Lpath[pid];
                acc.x += l.x; acc.y += l.y; acc.z += l.z;
                Lpath[pid] = acc;
                adds++;
            }
        },
        tstat);
    waveSum(adds, &counters[ptQueueField(homeSegment(), bounce, PT_Q_LADD_SHADOW)]);
}

// ------------------------------------------------------------------------------------------
// shading (path_tracer.cu:38-184 NEE + MIS, :186-233 naive), one path vertex per lane
// ------------------------------------------------------------------------------------------
struct SurfacePoint {
    f3 p, n;
    int mat;
    float area;  // area of the primitive (for the light pdf of BSDF-sampled emitter hits)
};

__device__ __forceinline__ SurfacePoint surfaceAt(const PtSceneView &sc, int slot, f3 o, f3 d, float t, float u, float v) {
    SurfacePoint s;
    const float4 p0 = sc.prims[3 * (size_t)slot], p1 = sc.prims[3 * (size_t)slot + 1];
    float4 p2 = sc.prims[3 * (size_t)slot + 2];
    // issues the third record's load with the first two instead of after the type test
    asm volatile("" : "+v"(p2.x), "+v"(p2.y), "+v"(p2.z));
    s.mat = asInt(p0.w);
    const int tag = asInt(p1.w);
    if (tag == PT_TAG_SPHERE) {
        const f3 c = xyz(p0);
        const float r = p1.x;
        s.p = o + d * t;
        s.n = normalize((s.p - c) * frcp(r));
        s.area = sphereArea(r);
    } else {
        const f3 a = xyz(p0), b = xyz(p1), c = xyz(p2);
        s.p = a * (1.f - u - v) + b * u + c * v;  // triangle.cu:36
        if (tag >= 0) {
            const f3 na = xyz(sc.normals[3 * (size_t)tag]), nb = xyz(sc.normals[3 * (size_t)tag + 1]), nc = xyz(sc.normals[3 * (size_t)tag + 2]);
            s.n = normalize(na * (1.f - u - v) + nb * u + nc * v);
        } else {
            s.n = normalize(cross(b - a, c - a));
        }
        s.area = triArea(a, b, c);
    }
    return s;
}

struct LightSample {
    f3 wi, n, emitted;
    float dist, pdf;
};

// sampler.h:142-183 uniform choice over emitters (+ environment), triangle.h:66-90,
// sphere.h:30-44, surface_sample.h:13-24
__device__ __forceinline__ LightSample sampleLights(const PtSceneView &sc, f3 p, float l1, float l2, float l3) {
    LightSample ls;
    const int nl = sc.numEmitters;
    const int choiceSize = nl + (sc.env.present ? 1 : 0);
    const int choice = (int)fminf((float)(choiceSize - 1), floorf(l1 * (float)choiceSize));
    const float choicePdf = frcp((float)choiceSize);
    if (choice < nl) {
        const float4 e0 = sc.emitters[3 * choice];
        float4 e1 = sc.emitters[3 * choice + 1], e2 = sc.emitters[3 * choice + 2];
        // all three records load before the type test (the triangle-only components would
        // otherwise wait for the tag)
        asm volatile("" : "+v"(e1.x), "+v"(e1.y), "+v"(e1.z), "+v"(e2.x), "+v"(e2.y), "+v"(e2.z));
        f3 pt, n;
        float pdfA;
        if (asInt(e1.w) == PT_TAG_SPHERE) {
            const f3 pl = uniformSphere(l2, l3);
            pt = pl * e1.x + xyz(e0);
            n = pl;
            pdfA = frcp(sphereArea(e1.x));
        } else {
            const f3 a = xyz(e0), b = xyz(e1), c = xyz(e2);
            const float sr = fsqrt(l2);
            const float wa = 1.f - sr, wb = sr * (1.f - l3);
            const float wc = 1.f - wa - wb;
            pt = a * wa + b * wb + c * wc;
            n = normalize(cross(b - a, c - a));
            pdfA = frcp(triArea(a, b, c));
        }
        const f3 dir = pt - p;
        const float dist = length(dir);
        const f3 wdir = normalize(dir);
        const float proj = fabsf(dot(n, -wdir));
        const float sa = proj == 0.f ? 0.f : fdiv(pdfA * (dist * dist), proj);
        const PtMaterial m = sc.materials[asInt(e0.w)];
        ls.wi = wdir;
        ls.n = n;
        ls.dist = dist;
        ls.pdf = sa * choicePdf;
        ls.emitted = mk(m.b.x, m.b.y, m.b.z);
    } else {
        float pdf;
        f3 em;
        const f3 w = envSample(sc.env, l2, l3, &pdf, &em);
        ls.wi = normalize(w);
        ls.n = -ls.wi;
        ls.dist = 1e16f;
        ls.pdf = pdf * choicePdf;
        ls.emitted = em;
    }
    return ls;
}

// sampler.h:89-125 light pdf (solid angle) of a BSDF-sampled emitter hit, max(0, cos)
__device__ __forceinline__ float pdfSceneLight(const PtSceneView &sc, f3 ref, const SurfacePoint &s) {
    const f3 dir = ref - s.p;
    const f3 wo = normalize(dir);
    const float dist = length(dir);
    const float proj = fmaxf(0.f, dot(s.n, wo));
    const float pdfSA = fdiv(frcp(s.area) * (dist * dist), proj);
    return fdiv(pdfSA, (float)(sc.numEmitters + (sc.env.present ? 1 : 0)));
}

// Guided-vertex inputs of the shade kernel (ignored by the plain integrator)
struct GuideShadeArgs {
    const float4 *mix;      // GUIDE_REC_F4 float4 per queue slot (inference output)
    const float4 *feat;     // the slot's raw features (k_guide_features: even, then odd ones)
    float4 *records;        // GUIDE_TREC_F4 float4 per training record, recSeg per queue segment
    int recSeg;
    const int *recIn;       // training record of the vertex that spawned each queued ray (-1: none)
    int *recOut;
    float fracCap;          // renderer-side cap on the guided-sampling probability (defensive share)
};

__device__ __forceinline__ float lum(float4 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ float lum(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// A guided vertex's mixture record (16 words, layout at GUIDE_REC_F4) in registers.
struct GuideRec {
    uint32_t w[16];
};
__device__ __forceinline__ GuideRec loadGuideRec(const float4 *G) {
    GuideRec r;
    const uint4 *q = (const uint4 *)G;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint4 v = q[i];
        r.w[4 * i] = v.x;
        r.w[4 * i + 1] = v.y;
        r.w[4 * i + 2] = v.z;
        r.w[4 * i + 3] = v.w;
    }
    return r;
}
// vMF mixture density of direction w (world space)
__device__ __forceinline__ float mixturePdf(const GuideRec &g, f3 w) {
    float pi[GUIDE_LOBES];
    guideLobeWeights(g.w, pi);
    float q = 0.f;
#pragma unroll
    for (int k = 0; k < GUIDE_LOBES; k++) q += pi[k] * vmfDensity(guideOctDecode(g.w[2 * k]), guideHalfLo(g.w[2 * k + 1]), w);
    return q;
}

// One path vertex per lane.  Plain (GUIDED = false): the reference's NEE + MIS estimator
// (path_tracer.cu:142-184) with the BSDF-sampled emitter term folded into the next vertex's
// emission, or the naive estimator (:186-233).  Guided: at Lambert and Beckmann vertices the
// direction comes from the one-sample mixture c q + (1 - c) p_bsdf of the learned vMF lobes and
// the BSDF, the continuation is weighted by f cos / (c q + (1 - c) p_bsdf), and the same mixture
// density is the directional technique's pdf in both MIS weights.
// Plain shading fits 96 VGPRs without spilling: 5 waves per SIMD.
template <bool GUIDED, bool BATCH = false>
__global__ __launch_bounds__(256, GUIDED ? 3 : 4) void k_shade(PtSceneView sc, ChunkParams cp, int bounce, int *counters, const float4 *rayO, const float4 *rayD,
                                               const uint4 *state, const float4 *beta, const float4 *hits, float4 *nRayO, float4 *nRayD,
                                               uint4 *nState, float4 *nBeta, float4 *sO, float4 *sD, float4 *sL, float4 *Lpath,
                                               GuideShadeArgs ga, float4 camO) {
    // Waves claim 64 rays at a time from one queue segment and append the continuation and
    // shadow rays to the same segment of the next queues.  Appends happen where each record is
    // complete (one atomic per wave and branch: the lanes taking a branch append together), so
    // no output record stays live across the rest of the vertex.
    //
    // camO.w != 0 (bounce 0): the camera rays' origin, throughput, previous pdf, path index and
    // RNG state are recomputed here instead of read (k_raygen stores only the directions), and
    // the path's radiance is written instead of added to.
    //
    // BATCH (a batch wavefront): every path belongs to one render of the batch (state.w above
    // the depth; at bounce 0 found from the path index), whose camera, lights, environment,
    // depth limit and PCG32 streams it uses; continuation rays carry the render in rayO.w and
    // shadow rays in sL.w, and each render counts the rays its paths spawn.
    //
    // Misses (a fifth to most of a queue) take a few instructions, hits the whole BSDF and light
    // code.  A wave claims 64 x kShadeClaimWords records at a time (their hit words loaded
    // together), parks the slots of hits and of misses in two LDS lists and shades either 64
    // at a time, so both codes run on full waves.  Every vertex is shaded exactly once (the
    // last partial batches when the queue is exhausted); results do not depend on the order.
    const bool camera = camO.w != 0.f;
    // per wave: parked slots (the segment is slot / segLen), each list below 64 + one claim
    // (at most 63 left over plus the claim)
    __shared__ int sHit[4][64 * (kShadeClaimWords + 1)], sMiss[4][64 * (kShadeClaimWords + 1)];
    int *pHit = sHit[threadIdx.x >> 6], *pMiss = sMiss[threadIdx.x >> 6];
    SegCursor cur = segStart();
    int adds = 0;
    int nHit = 0, nMiss = 0;   // parked records (wave-uniform)
    bool claiming = true;      // the wave's segments still have rays
    while (true) {
        int i = 0, seg = 0;
        bool active;
        const bool hitBatch = nHit >= 64 || (!claiming && nHit > 0);
        if (hitBatch || nMiss >= 64 || (!claiming && nMiss > 0)) {
            const int n = hitBatch ? nHit : nMiss;
            const int nb = min(n, 64);
            const int at = n - nb;
            if (hitBatch) nHit = at;
            else nMiss = at;
            active = laneId() < nb;
            if (active) {
                i = (hitBatch ? pHit : pMiss)[at + laneId()];
                int k;
                seg = divmodEst(i, cp.segLen, &k);
            }
        } else if (claiming) {
            int base = 0;
            const int got = segClaim(counters, bounce, PT_Q_RAYS, PT_Q_FETCH_SHADE, cp.segLen, 64 * kShadeClaimWords, cur, &base);
            if (got == 0) {
                claiming = false;
                continue;
            }
            // the lane's hit words all in flight at once (indices clamped to the claimed run)
            const int last = base + got - 1;
            int w[kShadeClaimWords];
#pragma unroll
            for (int q = 0; q < kShadeClaimWords; q++) w[q] = asInt(hits[min(base + laneId() + 64 * q, last)].w);
#pragma unroll
            for (int q = 0; q < kShadeClaimWords; q++) {
                const int iq = base + laneId() + 64 * q;
                const bool hq = iq <= last && w[q] != INT32_MIN, mq = iq <= last && !hq;
                const uint64_t bh = __ballot(hq), bm = __ballot(mq);
                if (hq) pHit[nHit + (int)lanesBelow(bh)] = iq;
                if (mq) pMiss[nMiss + (int)lanesBelow(bm)] = iq;
                nHit += __popcll(bh);
                nMiss += __popcll(bm);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            continue;
        } else {
            break;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool addL = false;
        int rid = 0, spawned = 0;   // batch wavefront: the path's render, rays its vertex spawned
        if (active) {
            const float4 rd = rayD[i], h = hits[i];
            float4 ro, bp;
            uint4 st;
            if (camera) {
                const int p = ptCameraPath(seg, i - seg * cp.segLen);
                bp = make_float4(1.f, 1.f, 1.f, -1.f);
                if constexpr (BATCH) {
                    const PathSource ps = batchPathSource(cp, p);
                    const BatchRender &R = cp.renders[ps.render];
                    Rng g = seedRng(ps.seed, (uint32_t)ps.pixel, (uint32_t)ps.sample);
                    g.next();
                    g.next();
                    ro = batchCameraOrigin(R);
                    st = make_uint4((uint32_t)g.state, (uint32_t)(g.state >> 32), (uint32_t)p, packDepth(0, ps.render));
                } else {
                    const Rng g = cameraRng(cp, p);
                    ro = camO;
                    st = make_uint4((uint32_t)g.state, (uint32_t)(g.state >> 32), (uint32_t)p, 0u);
                }
            } else {
                ro = rayO[i];
                bp = beta[i];
                st = state[i];
            }
            const int pathId = (int)st.z;
            const int depth = BATCH ? (int)(st.w & 0xffffu) : (int)st.w;
            rid = BATCH ? (int)(st.w >> 16) : 0;
            // the scene as this path sees it: a batch render's lights, environment, depth limit
            PtSceneView lv = sc;
            int maxDepth = cp.maxDepth;
            if constexpr (BATCH) {
                const BatchRender &R = cp.renders[rid];
                lv.emitters = sc.emitters + 3 * (size_t)R.emitBegin;
                lv.numEmitters = R.numEmitters;
                lv.env = R.env;
                maxDepth = R.maxDepth;
            }
            const f3 o = xyz(ro), d = xyz(rd), thr = xyz(bp);
            const float prevPdf = bp.w;
            f3 L = mk(0.f);
            if constexpr (GUIDED) {
                // radiance of the path before this vertex: the baseline of the training record of
                // the vertex that sampled this ray
                const int rin = ga.recIn[i];
                if (rin >= 0) ((float *)&ga.records[(size_t)rin * GUIDE_TREC_F4 + 3])[2] = lum(Lpath[pathId]);
            }
            const int slot = asInt(h.w);
            if (slot == INT32_MIN) {
                // miss: environment (camera miss: environment_light.h:68-83 as is)
                if (depth == 0 || !cp.nee || prevPdf < 0.f || !lv.env.present) {
                    L = thr * envEmit(lv.env, d);
                } else {
                    // sampler.h:127-140: light-selection share of the environment density
                    float pe;
                    const f3 e = envEmitPdf(lv.env, d, &pe);
                    L = thr * e * balance(prevPdf, pe * frcp((float)(lv.numEmitters + 1)));
                }
            } else {
                const SurfacePoint sp = surfaceAt(lv, slot, o, d, h.x, h.y, h.z);
                const Frame fr = makeFrame(sp.n);
                const f3 wo = normalize(toLocal(fr, -d));
                const PtMaterial m = lv.materials[sp.mat];
                const int type = matType(m);
                // lambertian.cu:5-11: emission seen from the front only
                if (type == PT_MAT_LAMBERTIAN && wo.z >= 0.f && (m.b.x != 0.f || m.b.y != 0.f || m.b.z != 0.f)) {
                    const f3 e = mk(m.b.x, m.b.y, m.b.z);
                    if (depth == 0 || !cp.nee || prevPdf < 0.f) L = thr * e;
                    else L = thr * e * balance(prevPdf, pdfSceneLight(lv, o, sp));
                }
                if (depth + 1 < maxDepth) {
                    Rng rng;
                    if constexpr (BATCH) {
                        const PathSource ps = batchPathSource(cp, pathId);
                        rng.inc = streamInc(ps.seed, (uint32_t)ps.pixel, (uint32_t)ps.sample);
                    } else {
                        int s;
                        const int pixel = cp.pixBegin + chunkPixelOf(cp, pathId, &s);
                        rng.inc = streamInc(cp.seed, (uint32_t)pixel, (uint32_t)(cp.sampleBase + s));
                    }
                    rng.state = ((uint64_t)st.y << 32) | st.x;
                    GuideRec G;
                    bool guidedV = false;
                    float frac = 0.f;
                    if constexpr (GUIDED) {
                        if (ga.mix) {
                            G = loadGuideRec(ga.mix + (size_t)i * GUIDE_REC_F4);
                            const float f = guideHalfHi(G.w[15]);
                            guidedV = f > 0.f;
                            frac = fminf(f, ga.fracCap);
                        }
                    }
                    BsdfSample bs;
                    float pdfDir;  // density of the direction sampling technique at bs.wi
                    float pbRec = 0.f;
                    if (GUIDED && guidedV) {
                        const float usel = rng.uniform();
                        if (usel < frac) {
                            float pi[GUIDE_LOBES];
                            guideLobeWeights(G.w, pi);
                            const float ul = rng.uniform();
                            // lobe k: the first whose running weight sum exceeds ul (the last
                            // when rounding leaves none), its record words picked by selects
                            // (a dynamically indexed record would live in scratch memory)
                            uint32_t wMu = G.w[2 * (GUIDE_LOBES - 1)], wKappa = G.w[2 * (GUIDE_LOBES - 1) + 1];
                            float acc = 0.f;
                            bool found = false;
#pragma unroll
                            for (int j = 0; j < GUIDE_LOBES - 1; j++) {
                                acc += pi[j];
                                const bool take = !found && ul < acc;
                                wMu = take ? G.w[2 * j] : wMu;
                                wKappa = take ? G.w[2 * j + 1] : wKappa;
                                found = found || take;
                            }
                            const float u1 = rng.uniform(), u2 = rng.uniform();
                            const f3 ww = sampleVmf(guideOctDecode(wMu), guideHalfLo(wKappa), u1, u2);
                            bs.wi = toLocal(fr, ww);
                            bs.f = bsdfF(m, wo, bs.wi);
                            bs.delta = false;
                        } else {
                            const int dims = bsdfDims(type);
                            const float x1 = dims >= 1 ? rng.uniform() : 0.f;
                            const float x2 = dims >= 2 ? rng.uniform() : 0.f;
                            bs = bsdfSample(m, wo, x1, x2);
                        }
                        pbRec = bsdfSamplePdf(m, wo, bs.wi);
                        pdfDir = frac * mixturePdf(G, toWorld(fr, bs.wi)) + (1.f - frac) * pbRec;
                        bs.pdf = pdfDir;
                    } else {
                        const int dims = bsdfDims(type);
                        const float x1 = dims >= 1 ? rng.uniform() : 0.f;
                        const float x2 = dims >= 2 ? rng.uniform() : 0.f;
                        bs = bsdfSample(m, wo, x1, x2);
                        pdfDir = bs.pdf;
                    }
                    // the continuation's queue slot is claimed here and used after the light
                    // sample, whose work hides the atomic's round trip
                    bool pushRay = false;
                    f3 nb = mk(0.f);
                    if (!isZero(bs.f) && bs.pdf > 0.f) {
                        nb = thr * (bs.f * (fabsf(bs.wi.z) * frcp(bs.pdf)));
                        pushRay = !isZero(nb) && nb.x == nb.x;
                    }
                    const PendingAppend rayAppend = segAppendIssue(pushRay, counters, seg, bounce + 1, PT_Q_RAYS);
                    bool pushShadow = false;
                    f3 contrib = mk(0.f);
                    LightSample ls;
                    if (cp.nee && !bs.delta) {
                        // directSampleLights (path_tracer.cu:83-126); a scene without lights
                        // draws the three uniforms and adds nothing
                        const float l1 = rng.uniform(), l2 = rng.uniform(), l3 = rng.uniform();
                        const bool anyLight = lv.numEmitters > 0 || lv.env.present;
                        if (anyLight) ls = sampleLights(lv, sp.p, l1, l2, l3);
                        if (anyLight && dot(ls.n, ls.wi) < 0.f && ls.pdf > 0.f) {
                            const f3 wiL = toLocal(fr, ls.wi);
                            float pB;
                            if (GUIDED && guidedV) pB = frac * mixturePdf(G, ls.wi) + (1.f - frac) * bsdfSamplePdf(m, wo, wiL);
                            else pB = cp.misSample ? bsdfSamplePdf(m, wo, wiL) : bsdfPdf(m, wo, wiL);
                            const float w = balance(ls.pdf, pB);
                            contrib = thr * (ls.emitted * w * bsdfF(m, wo, wiL) * (fabsf(dot(sp.n, ls.wi)) * frcp(ls.pdf)));
                            pushShadow = !isZero(contrib) && !(contrib.x != contrib.x);
                        }
                    }
                    const PendingAppend shadowAppend = segAppendIssue(pushShadow, counters, seg, bounce, PT_Q_SHADOW);
                    spawned += (pushShadow ? 1 : 0) + (pushRay ? 1 : 0);
                    const int rs = segAppendSlot(rayAppend, seg, cp.segLen);
                    const f3 wworld = toWorld(fr, bs.wi);
                    if (pushRay) {
                        nRayO[rs] = f4(sp.p, BATCH ? asFloat(rid) : kBounceEps);
                        nRayD[rs] = f4(wworld, FLT_MAX);
                        nState[rs] = make_uint4((uint32_t)rng.state, (uint32_t)(rng.state >> 32), (uint32_t)pathId,
                                                BATCH ? packDepth(depth + 1, rid) : (uint32_t)(depth + 1));
                        nBeta[rs] = f4(nb, bs.delta ? -1.f : pdfDir);
                    }
                    const int ss = segAppendSlot(shadowAppend, seg, cp.segLen);
                    if (pushShadow) {
                        sO[ss] = f4(sp.p, ls.dist - 2e-4f);
                        sD[ss] = f4(ls.wi, asFloat(pathId));
                        sL[ss] = f4(contrib, BATCH ? asFloat(rid) : 0.f);
                    }
                    if constexpr (GUIDED) {
                        // a training record per guided vertex whose path continues
                        const float lt = lum(thr);
                        const bool wantRec = pushRay && guidedV && lt > 0.f;
                        // records go to the segment's own part of the record buffer
                        const int slotRec = segAppendLanes(wantRec, counters, seg, 0, PT_Q_RECORDS, 0);
                        if (pushRay) {
                            int ridx = -1;
                            if (wantRec && slotRec < ga.recSeg) {
                                ridx = seg * ga.recSeg + slotRec;
                                // the features the inference of this vertex used, in raw order
                                const float4 fe = ga.feat[2 * (size_t)i], fo = ga.feat[2 * (size_t)i + 1];
                                float4 *R = ga.records + (size_t)ridx * GUIDE_TREC_F4;
                                R[0] = make_float4(fe.x, fo.x, fe.y, fo.y);
                                R[1] = make_float4(fe.z, fo.z, fe.w, fo.w);
                                R[2] = f4(wworld, pbRec);
                                R[3] = make_float4(frcp(lt), asFloat(pathId), -1.f, 0.f);
                            }
                            ga.recOut[rs] = ridx;
                        }
                    }
                }
            }
            addL = !isZero(L);
            if (camera) {
                // the first contribution to a zero radiance (0 + L, as an addition would give)
                Lpath[pathId] = make_float4(0.f + L.x, 0.f + L.y, 0.f + L.z, 0.f);
            } else if (addL) {
                float4 acc = Lpath[pathId];
                acc.x += L.x; acc.y += L.y; acc.z += L.z;
                Lpath[pathId] = acc;
            }
        }
        adds += addL ? 1 : 0;
        if constexpr (BATCH) waveAddKeyed(spawned, rid, cp.renderRays);
    }
    waveSum(adds, &counters[ptQueueField(homeSegment(), bounce, PT_Q_LADD_SHADE)]);
}

// Training targets: weight = (radiance after the vertex's sampled direction) / luminance of the
// throughput at the vertex, i.e. the MC estimate of f cos L_i / p for the sampled direction.
// Records live in per-segment parts of recSeg entries; block 0 also writes the exclusive prefix
// sums of the kept counts, through which training reads the records as one sequence.
__global__ __launch_bounds__(256) void k_guide_finalize(float4 *records, int recSeg, int *counters, const float4 *Lpath) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int acc = 0;
        for (int s = 0; s < PT_QUEUE_SEGS; s++) {
            counters[kRecPrefixSlot + s] = acc;
            acc += min(counters[ptQueueField(s, 0, PT_Q_RECORDS)], recSeg);
        }
        counters[kRecPrefixSlot + PT_QUEUE_SEGS] = acc;
        counters[kRecValidSlot] = acc;
    }
    // grid-stride over the record slots; the weight sum is reduced per block so that the
    // shared accumulator takes one atomic per block
    float w = 0.f;
    const int total = PT_QUEUE_SEGS * recSeg;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        int off;
        const int seg = divmodEst(i, recSeg, &off);
        if (off >= min(counters[ptQueueField(seg, 0, PT_Q_RECORDS)], recSeg)) continue;
        float4 *r = records + (size_t)i * GUIDE_TREC_F4 + 3;
        const float4 v = *r;
        const int pid = asInt(v.y);
        float wi = 0.f;
        if (v.z >= 0.f) wi = fmaxf(0.f, lum(Lpath[pid]) - v.z) * v.x;
        if (!(wi == wi) || isinf(wi)) wi = 0.f;
        r->w = wi;
        w += wi;
    }
    for (int off = 32; off >= 1; off >>= 1) w += __shfl_down(w, off);
    __shared__ float part[4];
    if (laneId() == 0) part[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float t = part[0] + part[1] + part[2] + part[3];
        if (t != 0.f) atomicAdd((float *)&counters[kWsumSlot], t);
    }
}

// Per pixel of the chunk: FP64 sum over the chunk's sample rows (framebuffer.cu:7-41 keeps a
// running FP32 mean instead).  Block 0 also folds the chunk's queue counters into the stats.
__global__ __launch_bounds__(256) void k_accumulate(ChunkParams cp, const float4 *Lpath, double *accum, const int *counters, unsigned long long *stats) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0) {
        for (int b = threadIdx.x; b < kMaxBounces; b += blockDim.x) {
            unsigned long long v[4] = {0, 0, 0, 0};
            for (int seg = 0; seg < PT_QUEUE_SEGS; seg++) {
                v[0] += (unsigned long long)counters[ptQueueField(seg, b, PT_Q_RAYS)];
                v[1] += (unsigned long long)counters[ptQueueField(seg, b, PT_Q_SHADOW)];
                v[2] += (unsigned long long)counters[ptQueueField(seg, b, PT_Q_LADD_SHADE)];
                v[3] += (unsigned long long)counters[ptQueueField(seg, b, PT_Q_LADD_SHADOW)];
            }
            atomicAdd(&stats[S_RAYS + b], v[0]);
            atomicAdd(&stats[S_SHADOW + b], v[1]);
            atomicAdd(&stats[S_LADD_SHADE + b], v[2]);
            atomicAdd(&stats[S_LADD_SHADOW + b], v[3]);
        }
        if (threadIdx.x == 0) atomicAdd(&stats[S_PATHS], (unsigned long long)cp.nPaths);
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (cp.pixelMajor) {
        // A pixel's samples are contiguous: the block's 256 pixels are read in tiles of
        // kAccTile samples per pixel (each pixel's run of records loaded by neighbouring
        // lanes) into LDS, and each thread sums its pixel's samples in sample order, the same
        // FP64 sums as the sample-major layout's.
        constexpr int kAccTile = 8;
        __shared__ float tx[256][kAccTile + 1], ty[256][kAccTile + 1], tz[256][kAccTile + 1];
        const int p0 = blockIdx.x * blockDim.x;
        const int npb = min((int)blockDim.x, cp.npix - p0);
        for (int c0 = 0; c0 < cp.samples; c0 += kAccTile) {
            const int cn = min(kAccTile, cp.samples - c0);
            for (int e = threadIdx.x; e < npb * kAccTile; e += blockDim.x) {
                const int j = e / kAccTile, k = e % kAccTile;
                if (k < cn) {
                    const float4 l = Lpath[(size_t)(p0 + j) * cp.samples + c0 + k];
                    tx[j][k] = l.x;
                    ty[j][k] = l.y;
                    tz[j][k] = l.z;
                }
            }
            __syncthreads();
            if ((int)threadIdx.x < npb) {
                for (int k = 0; k < cn; k++) {
                    sx += (double)tx[threadIdx.x][k];
                    sy += (double)ty[threadIdx.x][k];
                    sz += (double)tz[threadIdx.x][k];
                }
            }
            __syncthreads();
        }
        if (p >= cp.npix) return;
    } else {
        if (p >= cp.npix) return;
        for (int s = 0; s < cp.samples; s++) {
            const float4 l = Lpath[(size_t)s * cp.npix + p];
            sx += (double)l.x;
            sy += (double)l.y;
            sz += (double)l.z;
        }
    }
    double *a = accum + 3 * (size_t)(cp.pixBegin + p);
    a[0] += sx;
    a[1] += sy;
    a[2] += sz;
}

// Batch wavefront: per (group, pixel) of the chunk, the FP64 sum over the group's sample rows
// into its render's accumulator; block 0 folds the chunk's queue counters into `stats`.
__global__ __launch_bounds__(256) void k_accumulate_batch(ChunkParams cp, int accThreads, const float4 *Lpath, const int *counters,
                                                          unsigned long long *stats) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0) {
        for (int b = threadIdx.x; b < kMaxBounces; b += blockDim.x) {
            unsigned long long v[2] = {0, 0};
            for (int seg = 0; seg < PT_QUEUE_SEGS; seg++) {
                v[0] += (unsigned long long)counters[ptQueueField(seg, b, PT_Q_RAYS)];
                v[1] += (unsigned long long)counters[ptQueueField(seg, b, PT_Q_SHADOW)];
            }
            atomicAdd(&stats[S_RAYS + b], v[0]);
            atomicAdd(&stats[S_SHADOW + b], v[1]);
        }
        if (threadIdx.x == 0) atomicAdd(&stats[S_PATHS], (unsigned long long)cp.nPaths);
    }
    // This thread's (group, pixel) and the run of records of its samples, [r0, r1).  The
    // block's entries are consecutive, and so are their runs (groups are numbered pixel-major
    // and laid out in order), so the block reads one contiguous range of Lpath in windows of
    // kAccWindow records through LDS; each thread adds the records of its run that a window
    // holds, in sample order (the sums of a per-thread loop over its samples).
    constexpr int kAccWindow = 2048;
    __shared__ float wx[kAccWindow], wy[kAccWindow], wz[kAccWindow];
    __shared__ int range[2];
    const bool valid = t < accThreads;
    int r0 = 0, r1 = 0, render = 0, p = 0;
    if (valid) {
        int lo = 0, hi = cp.numGroups - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cp.groups[mid].accBegin <= t) lo = mid;
            else hi = mid - 1;
        }
        const BatchGroup g = cp.groups[lo];
        p = t - g.accBegin;
        render = g.render;
        r0 = g.pathBegin + p * g.count;
        r1 = r0 + g.count;
    }
    const int lastValid = min((int)blockDim.x, accThreads - (int)(blockIdx.x * blockDim.x)) - 1;
    if ((int)threadIdx.x == 0) range[0] = r0;
    if ((int)threadIdx.x == lastValid) range[1] = r1;
    __syncthreads();
    const int b0 = range[0], b1 = range[1];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int w0 = b0; w0 < b1; w0 += kAccWindow) {
        const int wn = min(kAccWindow, b1 - w0);
        for (int e = threadIdx.x; e < wn; e += blockDim.x) {
            const float4 l = Lpath[(size_t)w0 + e];
            wx[e] = l.x;
            wy[e] = l.y;
            wz[e] = l.z;
        }
        __syncthreads();
        for (int k = max(r0, w0); k < min(r1, w0 + wn); k++) {
            sx += (double)wx[k - w0];
            sy += (double)wy[k - w0];
            sz += (double)wz[k - w0];
        }
        __syncthreads();
    }
    if (!valid) return;
    double *a = cp.renders[render].accum + 3 * (size_t)p;
    a[0] += sx;
    a[1] += sy;
    a[2] += sz;
}

// Merging one scene into a batch's arrays: node references move by the node and primitive
// offsets (leaf ranges keep their counts), primitive and emitter records by the material and
// normal offsets.
__global__ void k_rebase_nodes(float4 *nodes, int count, int nodeOff, int primOff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 &refs = nodes[(size_t)PT_NODE_F4 * i + 6];
    int r[4] = {asInt(refs.x), asInt(refs.y), asInt(refs.z), asInt(refs.w)};
    for (int k = 0; k < 4; k++) {
        if (r[k] == PT_EMPTY_CHILD) continue;
        if (r[k] >= 0) {
            r[k] += nodeOff;
        } else {
            const int v = ~r[k];
            const int mask = (1 << PT_LEAF_COUNT_SHIFT) - 1;
            r[k] = ~((v & ~mask) | ((v & mask) + primOff));
        }
    }
    refs = make_float4(asFloat(r[0]), asFloat(r[1]), asFloat(r[2]), asFloat(r[3]));
}

__global__ void k_rebase_prims(float4 *recs, int count, int matOff, int normOff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 &p0 = recs[3 * (size_t)i], &p1 = recs[3 * (size_t)i + 1];
    p0.w = asFloat(asInt(p0.w) + matOff);
    const int tag = asInt(p1.w);
    if (tag >= 0) p1.w = asFloat(tag + normOff);
}

// mean radiance + the uchar4 framebuffer of framebuffer.cu:35-40 (color.h:7-28 sRGB)
__global__ __launch_bounds__(256) void k_tonemap(const double *accum, int npix, double invSpp, float *radiance, uchar4 *rgba) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    unsigned char o[3];
    for (int k = 0; k < 3; k++) {
        const float v = (float)(accum[3 * (size_t)p + k] * invSpp);
        radiance[3 * (size_t)p + k] = v;
        o[k] = srgbByte(v);
    }
    rgba[p] = make_uchar4(o[0], o[1], o[2], 255);
}

// Normals view (g_buffer.cu:27-57): colour (n + 1) / 2 of the primary hit at the pixel centre
__global__ __launch_bounds__(256) void k_normals_shade(PtSceneView sc, int npix, const float4 *rayO, const float4 *rayD, const float4 *hits,
                                                       double *accum) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float4 h = hits[p];
    const int slot = asInt(h.w);
    f3 c = mk(0.f);
    if (slot != INT32_MIN) {
        const SurfacePoint s = surfaceAt(sc, slot, xyz(rayO[p]), xyz(rayD[p]), h.x, h.y, h.z);
        c = (s.n + mk(1.f)) / 2.f;
    }
    double *a = accum + 3 * (size_t)p;
    a[0] += c.x;
    a[1] += c.y;
    a[2] += c.z;
}

__global__ void k_centre_rays(PtCamera cam, int npix, int W, float4 *rayO, float4 *rayD) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    f3 o, d;
    cameraRay(cam, p / W, p % W, 0.5f, 0.5f, &o, &d);
    rayO[p] = f4(o, 0.f);
    rayD[p] = f4(d, FLT_MAX);
}

__global__ void k_map_prims(const PtSceneView sc, const float4 *hits, int n, float4 *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 h = hits[i];
    const int slot = asInt(h.w);
    if (slot != INT32_MIN) h.w = sc.prims[3 * (size_t)slot + 2].w;  // original index
    else h.x = -1.f;
    out[i] = h;
}

__global__ void k_zero_d(double *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
__global__ void k_fill_i32(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_zero_u64(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Completion stamp of the work queued before it on the stream: the device's constant-rate
// wall clock (hipDeviceAttributeWallClockRate kHz).
__global__ void k_stamp(unsigned long long *slot) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *slot = (unsigned long long)wall_clock64();
}

// The same completion stamp into slots[0 .. n): renders that complete together.
__global__ void k_stamp_range(unsigned long long *slots, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const unsigned long long t = wall_clock64();
        for (int k = 0; k < n; k++) slots[k] = t;
    }
}

}  // namespace pt

// ============================================================================================
// Renderer object
// ============================================================================================
using namespace pt;

namespace {
enum { K_RAYGEN = 0, K_EXTEND, K_SHADE, K_SHADOW, K_ACCUM, K_TONEMAP, K_GUIDE_INFER, K_GUIDE_TRAIN, K_NUM };
const char *kKernelNames[K_NUM] = {"raygen", "extend", "shade", "shadow", "accumulate", "tonemap", "guide_infer", "guide_train"};
}  // namespace

// SoA path state, traversal stack spill area and queue counters of up to `pool` paths.
struct PathPool {
    int64_t pool = 0;
    DevBuf rayO[2], rayD[2], st[2], beta[2], hits, sO, sD, sL, Lpath, counters, overflow, spill;
};

struct pt_renderer {
    int W = 0, H = 0;
    hipStream_t stream = nullptr;
    bool ownsStream = true;  // false for the renders of a batch, which share the batch's stream
    int cus = 256;
    // scene on the device
    DevBuf nodes, prims, primsIn, normals, materials, emitters, envRgba, envRows, envMarg, envPdf;
    PtSceneView view{};
    BvhInfo bvh;
    int bvhBuilder = kDefaultBvhBuilder;
    int stackNeed = 1;    // deepest traversal stack the tree can need (3 per 4-wide level)
    int ldsStack = 16;    // stack entries per lane kept in LDS; the rest spill to the pool's `spill`
    int refillLanes = 0;  // idle lanes of a trace wave that trigger taking new rays (0: per launch)
    bool packetCamera = true;  // camera runs of one pixel's samples traced as one packet
    bool traceStats = false;   // PT_OPT_TRACE_STATS: instrumented trace kernels
    bool hasScene = false;
    // host copies needed for edits
    std::vector<HostMaterial> mats;
    std::vector<int> triMat, sphMat;
    std::vector<float> triPos, sph;
    HostCamera camHost;
    PtCamera cam{};
    int maxDepth = 3;  // scene.h:18 defaultMaxDepth
    int nee = 1;
    int type = PT_RENDERER_OPTIX;
    uint64_t seed = 0;
    // path pool (owned, or shared by the renders of a batch) and its capacity
    std::shared_ptr<PathPool> pp = std::make_shared<PathPool>();
    int64_t poolWanted = kDefaultPool;
    int traceBlocksPerCU = kDefaultTraceBlocks, shadeBlocksPerCU = 8;
    // accumulation
    DevBuf accum, radiance, rgba, stats;
    int64_t spp = 0;
    bool resetPending = false;
    // pass tracking
    hipEvent_t evBegin = nullptr, evEnd = nullptr;
    bool inFlight = false;
    int passSpp = 0;
    int64_t lastPassPaths = 0;
    bool lastPassSingleChunk = false;
    int lastPassSamples = 0;        // sample rows of the last pass's (single) chunk
    bool lastPassPixelMajor = false;
    double gpuMs = 0.0;
    int64_t passes = 0;
    // kernel profiling
    bool profile = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> kevents;
    std::vector<hipEvent_t> eventPool;
    size_t eventNext = 0;
    pt_kernel_stat kstat[K_NUM] = {};
    double samplesAccumulated = 0.0, pixelsAccumulated = 0.0;
    std::vector<ChunkParams> lastChunks;
    int misSample = 0;
    // guiding network and its training state
    bool guideReady = false;
    DevBuf gParams, gM, gV, gGrad, gPartials, gLossPart, gLoss, gW, gWt, gB, gMix, gRecords, gRecIdx[2], gFeat;
    int64_t gRecCap = 0;
    int gTrainSteps = 16;   // Adam steps per chunk (room workload sweep: 2 -> 16 halves the MSE)
    int gBatch = 1 << 18;
    float gLr = 1e-2f;
    float gFracCap = 0.8f;  // at least 20% of guided-vertex directions come from the BSDF
    int gBounces = 0;       // guided vertices only at bounces < gBounces (0: every bounce)
    int gAdamStep = 0;
    int64_t gTrainTotal = 0;
    uint64_t gSeed = 1;
    float sceneLo[3] = {0.f, 0.f, 0.f}, sceneHi[3] = {1.f, 1.f, 1.f};
    double trainMs = 0.0;
    // the merged wavefront of a batch: its per-render table (device); null for a renderer
    const BatchRender *batchRenders = nullptr;
};

namespace {

hipEvent_t takeEvent(pt_renderer *r) {
    if (r->eventNext >= r->eventPool.size()) {
        hipEvent_t e;
        PT_HIP_TRY(hipEventCreate(&e));
        r->eventPool.push_back(e);
    }
    return r->eventPool[r->eventNext++];
}

struct KTimer {
    pt_renderer *r;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(pt_renderer *rr, int kk) : r(rr), k(kk) {
        if (r->profile) {
            a = takeEvent(r);
            b = takeEvent(r);
            PT_HIP_TRY(hipEventRecord(a, r->stream));
        }
    }
    ~KTimer() {
        if (r->profile) {
            (void)hipEventRecord(b, r->stream);
            r->kevents.push_back({k, {a, b}});
        }
    }
};

void computeCamera(pt_renderer *r) {
    PtCamera &c = r->cam;
    lookAt(r->camHost.origin, r->camHost.target, r->camHost.up, r->camHost.flip, c.m);
    // camera.h:41-46
    c.top = tanf(r->camHost.vfov / 2.f);
    c.height = c.top * 2.f;
    const float aspect = 1.f * r->W / r->H;
    c.width = c.height * aspect;
    c.right = c.width / 2.f;
    c.W = r->W;
    c.H = r->H;
}

void uploadMaterials(pt_renderer *r) {
    std::vector<PtMaterial> m(r->mats.size());
    for (size_t i = 0; i < m.size(); i++) {
        const HostMaterial &h = r->mats[i];
        const bool lamb = h.type == PT_LAMBERTIAN;
        m[i].a = make_float4(h.albedo[0], h.albedo[1], h.albedo[2], (float)h.type);
        const float param = h.type == PT_GLASS ? h.ior : (h.type == PT_MICROFACET ? h.alpha : 0.f);
        m[i].b = make_float4(lamb ? h.emit[0] : 0.f, lamb ? h.emit[1] : 0.f, lamb ? h.emit[2] : 0.f, param);
    }
    r->materials.alloc(std::max<size_t>(1, m.size()) * sizeof(PtMaterial));
    if (!m.empty()) PT_HIP_TRY(hipMemcpyAsync(r->materials.p, m.data(), m.size() * sizeof(PtMaterial), hipMemcpyHostToDevice, r->stream));
    r->view.materials = r->materials.as<PtMaterial>();
    r->view.numMaterials = (int)m.size();
}

// Emitter records in the reference light order (scene_data.cu:92-115); rebuilt after edits
// so that emission changes also change the light list.
void uploadEmitters(pt_renderer *r) {
    std::vector<float4> e;
    const int64_t nt = (int64_t)r->triMat.size();
    for (int64_t i = 0; i < nt; i++) {
        if (!isEmissive(r->mats[r->triMat[i]])) continue;
        const float *p = &r->triPos[i * 9];
        e.push_back(make_float4(p[0], p[1], p[2], asFloat(r->triMat[i])));
        e.push_back(make_float4(p[3], p[4], p[5], asFloat(PT_TAG_FLAT)));
        e.push_back(make_float4(p[6], p[7], p[8], asFloat((int)i)));
    }
    for (size_t i = 0; i < r->sphMat.size(); i++) {
        if (!isEmissive(r->mats[r->sphMat[i]])) continue;
        const float *s = &r->sph[i * 4];
        e.push_back(make_float4(s[0], s[1], s[2], asFloat(r->sphMat[i])));
        e.push_back(make_float4(s[3], 0.f, 0.f, asFloat(PT_TAG_SPHERE)));
        e.push_back(make_float4(0.f, 0.f, 0.f, asFloat(~(int)i)));
    }
    r->emitters.alloc(std::max<size_t>(1, e.size()) * sizeof(float4));
    if (!e.empty()) PT_HIP_TRY(hipMemcpyAsync(r->emitters.p, e.data(), e.size() * sizeof(float4), hipMemcpyHostToDevice, r->stream));
    r->view.emitters = r->emitters.as<float4>();
    r->view.numEmitters = (int)(e.size() / 3);
}

// Path pool sized to the work at hand: `need` paths rounded up to a power of two, at least
// 4096, at most the configured capacity, then rounded up to a multiple of PT_QUEUE_SEGS * 64;
// it grows, it does not shrink.  Every per-slot buffer holds `pool` entries, and a chunk of n
// <= pool paths addresses queue slots below ptQueueSlots(n) <= pool only because pool is such a
// multiple (a segment is a multiple of 64 slots long).
void ensurePool(pt_renderer *r, int64_t need = 4096) {
    const int64_t want = ptPoolCapacity(need, r->poolWanted);
    if (r->pp->pool >= want) return;
    const size_t n = (size_t)want;
    for (int k = 0; k < 2; k++) {
        r->pp->rayO[k].alloc(n * 16);
        r->pp->rayD[k].alloc(n * 16);
        r->pp->st[k].alloc(n * 16);
        r->pp->beta[k].alloc(n * 16);
    }
    r->pp->hits.alloc(n * 16);
    r->pp->sO.alloc(n * 16);
    r->pp->sD.alloc(n * 16);
    r->pp->sL.alloc(n * 16);
    r->pp->Lpath.alloc(n * 16);
    r->pp->pool = want;
}

void zeroAccum(pt_renderer *r) {
    const size_t n = 3 * (size_t)r->W * r->H;
    k_zero_d<<<(unsigned)((n + 255) / 256), 256, 0, r->stream>>>(r->accum.as<double>(), n);
    PT_HIP_TRY(hipGetLastError());
    r->spp = 0;
    r->resetPending = false;
}

// Spill area of the traversal stacks: entries beyond the LDS part for every lane of the grid.
int *ensureSpill(pt_renderer *r, int grid, int *cap) {
    *cap = std::max(0, r->stackNeed - r->ldsStack);
    const size_t bytes = (size_t)grid * kTraceThreads * std::max(*cap, 1) * sizeof(int);
    if (r->pp->spill.bytes < bytes) r->pp->spill.alloc(bytes);
    return r->pp->spill.as<int>();
}

// Idle lanes out of 64 at which a trace wave takes new rays.  A camera run of 64 holds the
// samples of one pixel (paths are numbered pixel-major), nearly the same ray, so a camera
// wave takes a whole run at a time: refilling early would mix in a pixel 64 runs away.  The
// incoherent queues refill earlier, which shortens the wait for a batch's longest ray.
// Per-launch optima on the 1.31M-triangle icosphere at 1024^2, 64 spp: camera rays 64
// (extend 1773 -> 1678 us per launch against 48; 40 and 56 as 48), the later closest-hit
// queues and the shadow rays 40 (whole batches of 64 take 1.9x as long on the first bounce's
// queue; 16 idle lanes cost 6% of the pass).
enum { TRACE_CAMERA = 0, TRACE_INCOHERENT = 1, TRACE_SHADOW = 2 };
int refillLanes(const pt_renderer *r, int kind) {
    if (r->refillLanes > 0) return r->refillLanes;
    return kind == TRACE_CAMERA ? 64 : 40;
}

// Trace grids hold at least PT_QUEUE_SEGS waves: every queue segment needs home waves.
int traceGrid(const pt_renderer *r) {
    return std::max(r->cus * r->traceBlocksPerCU, (PT_QUEUE_SEGS * 64 + kTraceThreads - 1) / kTraceThreads);
}

void launchExtend(pt_renderer *r, int bounce, int segLen, const float4 *ro, const float4 *rd, float4 *hits, int kind,
                  float4 camO = make_float4(0.f, 0.f, 0.f, 0.f), const ChunkParams *batchChunk = nullptr, bool packet = false) {
    const int refill = refillLanes(r, kind);
    const int grid = traceGrid(r);
    int *ctr = r->pp->counters.as<int>();
    int *ov = r->pp->overflow.as<int>();
    if (packet) {
        // the packet kernel holds 64 registers and a 512-byte LDS stack per block: 16 two-wave
        // blocks per CU fit (8 waves per SIMD); C1 6037-6050 against 5953-5972 Mrays/s with the
        // per-lane kernels' 11 blocks per CU
        k_extend_packet<<<std::max(grid, r->cus * kPacketBlocksPerCU), kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, ro, rd, camO, hits, ov);
        PT_HIP_TRY(hipGetLastError());
        return;
    }
    int cap = 0;
    int *sp = ensureSpill(r, grid, &cap);
    if (batchChunk) {
        const ChunkParams &cp = *batchChunk;
        if (r->ldsStack == 32) k_extend<32, true><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, ro, rd, camO, hits, refill, sp, cap, ov, cp);
        else k_extend<16, true><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, ro, rd, camO, hits, refill, sp, cap, ov, cp);
    } else if (r->traceStats) {
        unsigned long long *ts = r->stats.as<unsigned long long>() + S_TRACE;
        if (r->ldsStack == 32)
            k_extend<32, false, true><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, ro, rd, camO, hits, refill, sp, cap, ov,
                                                                            ChunkParams{}, ts);
        else
            k_extend<16, false, true><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, ro, rd, camO, hits, refill, sp, cap, ov,
                                                                            ChunkParams{}, ts);
    } else {
        if (r->ldsStack == 32) k_extend<32><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, ro, rd, camO, hits, refill, sp, cap, ov);
        else k_extend<16><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, ro, rd, camO, hits, refill, sp, cap, ov);
    }
    PT_HIP_TRY(hipGetLastError());
}
void launchShadow(pt_renderer *r, int bounce, int segLen) {
    const int refill = refillLanes(r, TRACE_SHADOW);
    const int grid = traceGrid(r);
    int *ctr = r->pp->counters.as<int>();
    int *ov = r->pp->overflow.as<int>();
    const float4 *o = r->pp->sO.as<float4>(), *d = r->pp->sD.as<float4>(), *l = r->pp->sL.as<float4>();
    float4 *L = r->pp->Lpath.as<float4>();
    int cap = 0;
    int *sp = ensureSpill(r, grid, &cap);
    const BatchRender *br = r->batchRenders;
    if (br) {
        if (r->ldsStack == 32) k_shadow<32, true><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, o, d, l, L, refill, sp, cap, ov, br);
        else k_shadow<16, true><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, o, d, l, L, refill, sp, cap, ov, br);
    } else if (r->traceStats) {
        unsigned long long *ts = r->stats.as<unsigned long long>() + S_TRACE + 3;
        if (r->ldsStack == 32)
            k_shadow<32, false, true><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, o, d, l, L, refill, sp, cap, ov, nullptr, ts);
        else
            k_shadow<16, false, true><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, o, d, l, L, refill, sp, cap, ov, nullptr, ts);
    } else {
        if (r->ldsStack == 32) k_shadow<32><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, o, d, l, L, refill, sp, cap, ov);
        else k_shadow<16><<<grid, kTraceThreads, 0, r->stream>>>(r->view, ctr, bounce, segLen, o, d, l, L, refill, sp, cap, ov);
    }
    PT_HIP_TRY(hipGetLastError());
}

// Clears the queue counters and describes a linear queue of n records (slots 0 .. n - 1) in
// field `field` of bounce 0; its segment length is ptSegLen(n).
void queueLinear(pt_renderer *r, int field, int n) {
    k_queue_reset<<<PT_QUEUE_SEGS + 1, 256, 0, r->stream>>>(r->pp->counters.as<int>(), field, n, ptSegLen(n), 0);
    PT_HIP_TRY(hipGetLastError());
}

void ensureGuide(pt_renderer *r) {
    if (r->guideReady && r->gRecCap == 2 * r->pp->pool) return;
    const size_t np = GUIDE_PARAMS;
    if (!r->guideReady) {
        r->gParams.alloc(np * 4);
        r->gM.alloc(np * 4);
        r->gV.alloc(np * 4);
        r->gGrad.alloc(np * 4);
        r->gLoss.alloc(4);
        r->gW.alloc(GUIDE_LAYERS * GUIDE_W * GUIDE_W * 2);
        r->gWt.alloc(GUIDE_LAYERS * GUIDE_W * GUIDE_W * 2);
        r->gB.alloc(GUIDE_LAYERS * GUIDE_W * 4);
        std::vector<float> init(np);
        guideInitParams(init.data(), r->gSeed);
        std::vector<float> zero(np, 0.f);
        PT_HIP_TRY(hipMemcpyAsync(r->gParams.p, init.data(), np * 4, hipMemcpyHostToDevice, r->stream));
        PT_HIP_TRY(hipMemcpyAsync(r->gM.p, zero.data(), np * 4, hipMemcpyHostToDevice, r->stream));
        PT_HIP_TRY(hipMemcpyAsync(r->gV.p, zero.data(), np * 4, hipMemcpyHostToDevice, r->stream));
        guideUploadBf16(r->gParams.as<float>(), r->gW.as<__bf16>(), r->gWt.as<__bf16>(), r->gB.as<float>(), r->stream);
        PT_HIP_TRY(hipStreamSynchronize(r->stream));
        r->gAdamStep = 0;
    }
    r->gMix.alloc((size_t)r->pp->pool * GUIDE_REC_F4 * sizeof(float4));
    r->gFeat.alloc((size_t)r->pp->pool * 2 * sizeof(float4));
    r->gRecCap = 2 * r->pp->pool;
    r->gRecords.alloc((size_t)r->gRecCap * GUIDE_TREC_F4 * sizeof(float4));
    r->gRecIdx[0].alloc((size_t)r->pp->pool * 4);
    r->gRecIdx[1].alloc((size_t)r->pp->pool * 4);
    r->guideReady = true;
}

GuideNet guideNet(pt_renderer *r) {
    GuideNet net;
    net.w = r->gW.as<__bf16>();
    net.wt = r->gWt.as<__bf16>();
    net.b = r->gB.as<float>();
    return net;
}

GuideSceneNorm guideNorm(pt_renderer *r) {
    GuideSceneNorm nm;
    for (int k = 0; k < 3; k++) {
        const float e = r->sceneHi[k] - r->sceneLo[k];
        nm.lo[k] = r->sceneLo[k];
        nm.invExt[k] = e > 0.f ? 1.f / e : 0.f;
    }
    return nm;
}

void runChunk(pt_renderer *r, const ChunkParams &cp) {
    int *ctr = r->pp->counters.as<int>();
    const bool guided = r->type == PT_RENDERER_GUIDED;
    {
        KTimer t(r, K_RAYGEN);
        k_queue_reset<<<PT_QUEUE_SEGS + 1, 256, 0, r->stream>>>(ctr, PT_Q_RAYS, cp.nPaths, cp.segLen, 1);
        k_raygen<<<(cp.nPaths + 255) / 256, 256, 0, r->stream>>>(cp, r->cam, r->pp->rayD[0].as<float4>());
        PT_HIP_TRY(hipGetLastError());
        if (guided) {
            // camera rays were not spawned by a guided vertex
            const int slots = PT_QUEUE_SEGS * cp.segLen;
            k_fill_i32<<<(slots + 255) / 256, 256, 0, r->stream>>>(r->gRecIdx[0].as<int>(), slots, -1);
            PT_HIP_TRY(hipGetLastError());
        }
    }
    // shade needs PT_QUEUE_SEGS waves or more (home waves for every segment)
    const int shadeGrid = (int)std::max<int64_t>(PT_QUEUE_SEGS / 4, std::min<int64_t>((cp.nPaths + 255) / 256, (int64_t)r->cus * r->shadeBlocksPerCU));
    // bounce 0 traces and shades the camera rays from the camera's origin (w = 1: in use)
    const float4 camO = make_float4(r->cam.m[0][3], r->cam.m[1][3], r->cam.m[2][3], 1.f);
    const float4 noCam = make_float4(0.f, 0.f, 0.f, 0.f);
    int cur = 0;
    for (int b = 0; b < cp.maxDepth; b++) {
        const float4 bo = b == 0 ? camO : noCam;
        {
            KTimer t(r, K_EXTEND);
            // camera runs of one pixel's samples, or of a few neighbouring pixels' (pixel-major
            // chunks of 16 sample rows or more): one traversal per run.  Runs of 64 pixels of a
            // row (one sample row) trace faster per lane (the room at 1 spp: 3444 against 3254
            // Mrays/s); the 48-render batch gains 0.5% from 16 rows against 64.
            const bool packet = b == 0 && cp.pixelMajor && cp.samples >= 16 && !r->traceStats && r->view.numNodes > 0 &&
                                r->stackNeed <= kPacketStack && r->packetCamera;
            launchExtend(r, b, cp.segLen, r->pp->rayO[cur].as<float4>(), r->pp->rayD[cur].as<float4>(), r->pp->hits.as<float4>(),
                         b == 0 ? TRACE_CAMERA : TRACE_INCOHERENT, bo, nullptr, packet);
        }
        GuideShadeArgs ga{};
        // vertices of bounces >= gBounces (when set) sample the BSDF only
        const bool guidedBounce = guided && b + 1 < cp.maxDepth && (r->gBounces <= 0 || b < r->gBounces);
        if (guidedBounce) {
            KTimer t(r, K_GUIDE_INFER);
            guideInfer(r->view, guideNorm(r), guideNet(r), ctr + ptQueueField(0, b, PT_Q_RAYS), PT_SEG_STRIDE, cp.segLen, r->pp->rayO[cur].as<float4>(),
                       r->pp->rayD[cur].as<float4>(), bo, r->pp->hits.as<float4>(), r->gFeat.as<float4>(), r->gMix.as<float4>(), r->cus, r->stream);
        }
        if (guided) {
            ga.mix = guidedBounce ? r->gMix.as<float4>() : nullptr;
            ga.feat = guidedBounce ? r->gFeat.as<float4>() : nullptr;
            ga.records = r->gRecords.as<float4>();
            ga.recSeg = (int)(r->gRecCap / PT_QUEUE_SEGS);
            ga.recIn = r->gRecIdx[cur].as<int>();
            ga.recOut = r->gRecIdx[1 - cur].as<int>();
            ga.fracCap = r->gFracCap;
        }
        {
            KTimer t(r, K_SHADE);
            if (guided)
                k_shade<true><<<shadeGrid, 256, 0, r->stream>>>(
                    r->view, cp, b, ctr, r->pp->rayO[cur].as<float4>(), r->pp->rayD[cur].as<float4>(), r->pp->st[cur].as<uint4>(), r->pp->beta[cur].as<float4>(),
                    r->pp->hits.as<float4>(), r->pp->rayO[1 - cur].as<float4>(), r->pp->rayD[1 - cur].as<float4>(), r->pp->st[1 - cur].as<uint4>(),
                    r->pp->beta[1 - cur].as<float4>(), r->pp->sO.as<float4>(), r->pp->sD.as<float4>(), r->pp->sL.as<float4>(), r->pp->Lpath.as<float4>(), ga, bo);
            else
                k_shade<false><<<shadeGrid, 256, 0, r->stream>>>(
                    r->view, cp, b, ctr, r->pp->rayO[cur].as<float4>(), r->pp->rayD[cur].as<float4>(), r->pp->st[cur].as<uint4>(), r->pp->beta[cur].as<float4>(),
                    r->pp->hits.as<float4>(), r->pp->rayO[1 - cur].as<float4>(), r->pp->rayD[1 - cur].as<float4>(), r->pp->st[1 - cur].as<uint4>(),
                    r->pp->beta[1 - cur].as<float4>(), r->pp->sO.as<float4>(), r->pp->sD.as<float4>(), r->pp->sL.as<float4>(), r->pp->Lpath.as<float4>(), ga, bo);
            PT_HIP_TRY(hipGetLastError());
        }
        if (cp.nee) {
            KTimer t(r, K_SHADOW);
            launchShadow(r, b, cp.segLen);
        }
        cur = 1 - cur;
    }
    {
        KTimer t(r, K_ACCUM);
        k_accumulate<<<(cp.npix + 255) / 256, 256, 0, r->stream>>>(cp, r->pp->Lpath.as<float4>(), r->accum.as<double>(), ctr,
                                                                  r->stats.as<unsigned long long>());
        PT_HIP_TRY(hipGetLastError());
    }
    if (guided && r->gTrainSteps > 0) {
        KTimer t(r, K_GUIDE_TRAIN);
        const int nrec = (int)r->gRecCap;
        const int fblocks = (int)std::min<int64_t>((nrec + 255) / 256, (int64_t)r->cus * 4);
        k_guide_finalize<<<fblocks, 256, 0, r->stream>>>(r->gRecords.as<float4>(), (int)(r->gRecCap / PT_QUEUE_SEGS), ctr,
                                                         r->pp->Lpath.as<float4>());
        PT_HIP_TRY(hipGetLastError());
        // training scratch for the current batch size (reallocated only when it changes)
        GuideTrainBufs tb;
        tb.blocks = guideTrainBlocks(r->gBatch, r->cus);
        r->gPartials.alloc((size_t)tb.blocks * GUIDE_PARAMS * 4);
        r->gLossPart.alloc((size_t)tb.blocks * 4);
        tb.partials = r->gPartials.as<float>();
        tb.lossPart = r->gLossPart.as<float>();
        for (int s = 0; s < r->gTrainSteps; s++) {
            const int base = (int)((r->gTrainTotal * (int64_t)r->gBatch) & 0x3fffffff);
            guideGradients(guideNet(r), r->gRecords.as<float4>(), &ctr[kRecValidSlot], &ctr[kRecPrefixSlot], (int)(r->gRecCap / PT_QUEUE_SEGS),
                           base, r->gBatch, (const float *)&ctr[kWsumSlot], tb, r->gGrad.as<float>(), r->gLoss.as<float>(), r->stream);
            guideAdam(r->gParams.as<float>(), r->gM.as<float>(), r->gV.as<float>(), r->gGrad.as<float>(), ++r->gAdamStep, r->gLr,
                      r->gW.as<__bf16>(), r->gWt.as<__bf16>(), r->gB.as<float>(), r->stream);
            r->gTrainTotal++;
        }
    }
}

// A pass of `spp` samples per pixel cut into chunks of at most the pool's capacity: whole
// sample rows of the image when a row fits, else runs of pixels of one sample.
void planChunks(const pt_renderer *r, int spp, std::vector<ChunkParams> &out) {
    const int npix = r->W * r->H;
    // a shared pool may be larger than this renderer's configured capacity
    const int64_t pool = std::min<int64_t>(r->pp->pool, std::max<int64_t>(4096, r->poolWanted));
    auto chunk = [&](int pixBegin, int n, int samples, int sampleBase) {
        ChunkParams cp;
        cp.pixBegin = pixBegin;
        cp.npix = n;
        cp.samples = samples;
        cp.sampleBase = sampleBase;
        cp.nPaths = n * samples;
        cp.segLen = ptSegLen(cp.nPaths);
        cp.maxDepth = r->maxDepth;
        cp.nee = r->nee;
        cp.misSample = r->misSample;
        cp.W = r->W;
        cp.H = r->H;
        cp.seed = r->seed;
        cp.pixelMajor = 1;
        out.push_back(cp);
    };
    if (npix <= pool) {
        const int rowsPerChunk = (int)std::min<int64_t>(spp, pool / npix);
        for (int s0 = 0; s0 < spp; s0 += rowsPerChunk) chunk(0, npix, std::min(rowsPerChunk, spp - s0), (int)(r->spp + s0));
    } else {
        for (int s0 = 0; s0 < spp; s0++)
            for (int64_t p0 = 0; p0 < npix; p0 += pool) chunk((int)p0, (int)std::min<int64_t>(pool, npix - p0), 1, (int)(r->spp + s0));
    }
}

void tonemap(pt_renderer *r) {
    KTimer t(r, K_TONEMAP);
    const int npix = r->W * r->H;
    const double inv = r->spp > 0 ? 1.0 / (double)r->spp : 0.0;
    k_tonemap<<<(npix + 255) / 256, 256, 0, r->stream>>>(r->accum.as<double>(), npix, inv, r->radiance.as<float>(), r->rgba.as<uchar4>());
    PT_HIP_TRY(hipGetLastError());
}

int guard(const char *fn, pt_renderer *r) {
    if (!r) return fail(PT_ERR_INVALID, std::string(fn) + ": renderer is NULL");
    return PT_OK;
}

#define PT_API_BEGIN try {
#define PT_API_END(fn)                                                        \
    }                                                                         \
    catch (const HipFailure &e) {                                             \
        return fail(PT_ERR_DEVICE, std::string(fn) + ": " + e.msg);           \
    }                                                                         \
    catch (const std::bad_alloc &) {                                          \
        return fail(PT_ERR_MEMORY, std::string(fn) + ": out of host memory"); \
    }

void finishPass(pt_renderer *r) {
    // called once the end event completed
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r->evBegin, r->evEnd);
    r->gpuMs += ms;
    r->passes++;
    if (r->profile) {
        // launch counts and times accumulate until pt_renderer_reset_stats
        for (auto &ke : r->kevents) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, ke.second.first, ke.second.second);
            r->kstat[ke.first].launches++;
            r->kstat[ke.first].ms += t;
        }
    }
    r->kevents.clear();
    r->eventNext = 0;
    r->inFlight = false;
}

}  // namespace

extern "C" {

int pt_device_count(int *count) {
    if (!count) return fail(PT_ERR_INVALID, "pt_device_count: count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return PT_OK;
}

int pt_renderer_create(int width, int height, pt_renderer **out) {
    if (!out || width <= 0 || height <= 0) return fail(PT_ERR_INVALID, "pt_renderer_create: bad arguments");
    if ((int64_t)width * height > (1ll << 28)) return fail(PT_ERR_INVALID, "pt_renderer_create: image too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PT_ERR_DEVICE, "pt_renderer_create: no HIP device");
    pt_renderer *r = new pt_renderer();
    PT_API_BEGIN
    r->W = width;
    r->H = height;
    PT_HIP_TRY(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    PT_HIP_TRY(hipEventCreate(&r->evBegin));
    PT_HIP_TRY(hipEventCreate(&r->evEnd));
    r->cus = deviceCUs();
    const size_t npix = (size_t)width * height;
    r->accum.alloc(npix * 3 * sizeof(double));
    r->radiance.alloc(npix * 3 * sizeof(float));
    r->rgba.alloc(npix * 4);
    r->stats.alloc(S_COUNT * sizeof(unsigned long long));
    r->pp->counters.alloc(kNumCounters * sizeof(int));
    r->pp->overflow.alloc(sizeof(int));
    PT_HIP_TRY(hipMemsetAsync(r->pp->overflow.p, 0, sizeof(int), r->stream));
    k_zero_u64<<<1, 512, 0, r->stream>>>(r->stats.as<unsigned long long>(), S_COUNT);
    zeroAccum(r);
    computeCamera(r);
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    *out = r;
    return PT_OK;
    }
    catch (const HipFailure &e) {
        delete r;
        return fail(PT_ERR_DEVICE, "pt_renderer_create: " + e.msg);
    }
}

int pt_renderer_destroy(pt_renderer *r) {
    if (!r) return PT_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (hipEvent_t e : r->eventPool) (void)hipEventDestroy(e);
    if (r->evBegin) (void)hipEventDestroy(r->evBegin);
    if (r->evEnd) (void)hipEventDestroy(r->evEnd);
    hipStream_t s = r->ownsStream ? r->stream : nullptr;
    delete r;
    if (s) (void)hipStreamDestroy(s);
    return PT_OK;
}

int pt_renderer_set_scene(pt_renderer *r, const pt_scene *s) {
    if (guard("pt_renderer_set_scene", r)) return PT_ERR_INVALID;
    if (!s) return fail(PT_ERR_INVALID, "pt_renderer_set_scene: scene is NULL");
    if (r->inFlight) return fail(PT_ERR_STATE, "pt_renderer_set_scene: a pass is still running");
    const int64_t nt = (int64_t)s->triMat.size(), ns = (int64_t)s->sphMat.size();
    // the trace kernels address primitive records (48 B) with 32-bit byte offsets
    if (nt + ns >= (1ll << PT_LEAF_COUNT_SHIFT) || (nt + ns) * 48 >= (1ll << 32))
        return fail(PT_ERR_INVALID, "pt_renderer_set_scene: too many primitives (at most 89,478,485)");
    if (s->mats.empty()) return fail(PT_ERR_INVALID, "pt_renderer_set_scene: scene has no materials");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    r->mats = s->mats;
    r->triMat = s->triMat;
    r->sphMat = s->sphMat;
    r->triPos = s->pos;
    r->sph = s->sph;
    r->camHost = s->cam;
    computeCamera(r);
    // primitive records in input order (triangles, then spheres)
    const int64_t n = nt + ns;
    std::vector<float4> recs((size_t)std::max<int64_t>(n, 1) * 3);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto grow = [&](float x, float y, float z) {
        lo[0] = std::min(lo[0], x); lo[1] = std::min(lo[1], y); lo[2] = std::min(lo[2], z);
        hi[0] = std::max(hi[0], x); hi[1] = std::max(hi[1], y); hi[2] = std::max(hi[2], z);
    };
    std::vector<float4> nrm;
    for (int64_t i = 0; i < nt; i++) {
        const float *p = &s->pos[i * 9];
        int tag = PT_TAG_FLAT;
        if (s->hasNormals) {
            // keep per-vertex normals only where they differ from the flat normal
            const float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]}, e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
            float g[3] = {e1[1] * e2[2] - e1[2] * e2[1], -(e1[0] * e2[2] - e1[2] * e2[0]), e1[0] * e2[1] - e1[1] * e2[0]};
            const float gl = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            bool flat = gl > 0.f;
            const float *nv = &s->nrm[i * 9];
            for (int v = 0; v < 3 && flat; v++)
                for (int k = 0; k < 3; k++)
                    if (fabsf(nv[v * 3 + k] - g[k] / gl) > 1e-6f) flat = false;
            if (!flat) {
                tag = (int)(nrm.size() / 3);
                for (int v = 0; v < 3; v++) nrm.push_back(make_float4(nv[v * 3], nv[v * 3 + 1], nv[v * 3 + 2], 0.f));
            }
        }
        recs[i * 3] = make_float4(p[0], p[1], p[2], asFloat(s->triMat[i]));
        recs[i * 3 + 1] = make_float4(p[3], p[4], p[5], asFloat(tag));
        recs[i * 3 + 2] = make_float4(p[6], p[7], p[8], asFloat((int)i));
        for (int v = 0; v < 3; v++) grow(p[v * 3], p[v * 3 + 1], p[v * 3 + 2]);
    }
    for (int64_t i = 0; i < ns; i++) {
        const float *q = &s->sph[i * 4];
        const int64_t k = nt + i;
        recs[k * 3] = make_float4(q[0], q[1], q[2], asFloat(s->sphMat[i]));
        recs[k * 3 + 1] = make_float4(q[3], 0.f, 0.f, asFloat(PT_TAG_SPHERE));
        recs[k * 3 + 2] = make_float4(0.f, 0.f, 0.f, asFloat(~(int)i));
        grow(q[0], q[1], q[2]);
    }
    r->primsIn.alloc(recs.size() * sizeof(float4));
    PT_HIP_TRY(hipMemcpyAsync(r->primsIn.p, recs.data(), recs.size() * sizeof(float4), hipMemcpyHostToDevice, r->stream));
    r->prims.alloc(recs.size() * sizeof(float4));
    r->normals.alloc(std::max<size_t>(nrm.size(), 1) * sizeof(float4));
    if (!nrm.empty()) PT_HIP_TRY(hipMemcpyAsync(r->normals.p, nrm.data(), nrm.size() * sizeof(float4), hipMemcpyHostToDevice, r->stream));
    for (int k = 0; k < 3; k++) {
        r->sceneLo[k] = std::isfinite(lo[k]) ? lo[k] : 0.f;
        r->sceneHi[k] = std::isfinite(hi[k]) ? hi[k] : 1.f;
    }
    buildBvh(r->primsIn.as<float4>(), (int)n, lo, hi, r->nodes, r->prims.as<float4>(), r->stream, &r->bvh, r->bvhBuilder);
    // and 4-wide nodes (128 B) with 32-bit byte offsets
    if ((int64_t)r->bvh.numNodes * PT_NODE_F4 * 16 >= (1ll << 32)) throw HipFailure("scene upload: BVH beyond 4 GiB of nodes");
    r->primsIn.release();
    // ordered 4-wide traversal pushes at most three references per level
    r->stackNeed = 3 * r->bvh.depth + 1;
    r->view.nodes = r->nodes.as<float4>();
    r->view.prims = r->prims.as<float4>();
    r->view.normals = r->normals.as<float4>();
    r->view.numPrims = (int)n;
    r->view.numNodes = r->bvh.numNodes;
    uploadMaterials(r);
    uploadEmitters(r);
    // environment
    r->view.env = PtEnv{};
    if (s->envW > 0 && s->envH > 0) {
        const size_t ne = (size_t)s->envW * s->envH;
        r->envRgba.alloc(ne * sizeof(float4));
        r->envRows.alloc(ne * sizeof(float2));
        r->envMarg.alloc((size_t)s->envH * sizeof(float2));
        r->envPdf.alloc(ne * sizeof(float));
        PT_HIP_TRY(hipMemcpyAsync(r->envRgba.p, s->env.data(), ne * sizeof(float4), hipMemcpyHostToDevice, r->stream));
        buildEnvTables(r->envRgba.as<float4>(), s->envW, s->envH, s->envSinWeight, r->envRows.as<float2>(), r->envMarg.as<float2>(),
                       r->envPdf.as<float>(), nullptr, nullptr, nullptr, nullptr, r->stream);
        r->view.env.rgba = r->envRgba.as<float4>();
        r->view.env.rows = r->envRows.as<float2>();
        r->view.env.marginal = r->envMarg.as<float2>();
        packEnvPdf(r->envRgba.as<float4>(), r->envPdf.as<float>(), (int)ne, r->stream);
        r->view.env.width = s->envW;
        r->view.env.height = s->envH;
        r->view.env.present = 1;
    } else {
        r->envRgba.release();
        r->envRows.release();
        r->envMarg.release();
        r->envPdf.release();
    }
    ensurePool(r);
    zeroAccum(r);
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    r->hasScene = true;
    return PT_OK;
    PT_API_END("pt_renderer_set_scene")
}

int pt_renderer_set_camera(pt_renderer *r, const float origin[3], const float target[3], const float up[3], float vfov, int flip) {
    if (guard("pt_renderer_set_camera", r)) return PT_ERR_INVALID;
    if (!origin || !target || !up) return fail(PT_ERR_INVALID, "pt_renderer_set_camera: NULL vector");
    float m[3][4];
    if (!lookAt(origin, target, up, flip, m)) return fail(PT_ERR_NUMERIC, "pt_renderer_set_camera: view direction equals the up vector");
    if (!(vfov > 0.f && vfov < 3.14159f)) return fail(PT_ERR_INVALID, "pt_renderer_set_camera: vfov out of (0, pi)");
    memcpy(r->camHost.origin, origin, 12);
    memcpy(r->camHost.target, target, 12);
    memcpy(r->camHost.up, up, 12);
    r->camHost.vfov = vfov;
    r->camHost.flip = flip ? 1 : 0;
    computeCamera(r);
    r->resetPending = true;
    return PT_OK;
}

int pt_renderer_set_material(pt_renderer *r, int id, int type, const float albedo[3], const float emit[3], float ior, float alpha) {
    if (guard("pt_renderer_set_material", r)) return PT_ERR_INVALID;
    if (!r->hasScene) return fail(PT_ERR_STATE, "pt_renderer_set_material: no scene");
    if (id < 0 || id >= (int)r->mats.size()) return fail(PT_ERR_INVALID, "pt_renderer_set_material: material id out of range");
    if (type < PT_LAMBERTIAN || type > PT_MICROFACET) return fail(PT_ERR_INVALID, "pt_renderer_set_material: bad material type");
    if (r->inFlight) return fail(PT_ERR_STATE, "pt_renderer_set_material: a pass is still running");
    PT_API_BEGIN
    HostMaterial &m = r->mats[id];
    const bool wasEmissive = isEmissive(m);
    m.type = type;
    if (albedo) memcpy(m.albedo, albedo, 12);
    if (emit) memcpy(m.emit, emit, 12);
    m.ior = ior;
    m.alpha = alpha;
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    uploadMaterials(r);
    if (wasEmissive || isEmissive(m)) uploadEmitters(r);
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    r->resetPending = true;
    return PT_OK;
    PT_API_END("pt_renderer_set_material")
}

int pt_renderer_set_max_depth(pt_renderer *r, int d) {
    if (guard("pt_renderer_set_max_depth", r)) return PT_ERR_INVALID;
    if (d < 0 || d >= kMaxBounces) return fail(PT_ERR_INVALID, "pt_renderer_set_max_depth: depth must be in [0, 63]");
    r->maxDepth = d;
    r->resetPending = true;
    return PT_OK;
}

int pt_renderer_set_nee(pt_renderer *r, int on) {
    if (guard("pt_renderer_set_nee", r)) return PT_ERR_INVALID;
    r->nee = on ? 1 : 0;
    r->resetPending = true;
    return PT_OK;
}

int pt_renderer_set_type(pt_renderer *r, int t) {
    if (guard("pt_renderer_set_type", r)) return PT_ERR_INVALID;
    if (t < PT_RENDERER_CUDA || t > PT_RENDERER_GUIDED) return fail(PT_ERR_INVALID, "pt_renderer_set_type: unknown renderer type");
    r->type = t;
    r->resetPending = true;
    return PT_OK;
}

int pt_renderer_set_seed(pt_renderer *r, uint64_t seed) {
    if (guard("pt_renderer_set_seed", r)) return PT_ERR_INVALID;
    r->seed = seed;
    r->resetPending = true;
    return PT_OK;
}

int pt_renderer_set_option(pt_renderer *r, int opt, int64_t v) {
    if (guard("pt_renderer_set_option", r)) return PT_ERR_INVALID;
    if (r->inFlight) return fail(PT_ERR_STATE, "pt_renderer_set_option: a pass is still running");
    PT_API_BEGIN
    switch (opt) {
    case PT_OPT_POOL_PATHS:
        if (v > (1ll << 27)) return fail(PT_ERR_INVALID, "pt_renderer_set_option: pool larger than 2^27 paths");
        PT_HIP_TRY(hipStreamSynchronize(r->stream));
        r->poolWanted = v > 0 ? v : kDefaultPool;
        // shrink: the next pass allocates at most the new capacity
        if (r->pp->pool > r->poolWanted) r->pp->pool = 0;
        break;
    case PT_OPT_TRACE_LDS_STACK:
        if (v != 0 && v != 16 && v != 32) return fail(PT_ERR_INVALID, "pt_renderer_set_option: LDS stack entries must be 16 or 32");
        r->ldsStack = v == 0 ? 16 : (int)v;
        break;
    case PT_OPT_TRACE_BLOCKS_PER_CU: r->traceBlocksPerCU = v > 0 ? (int)std::min<int64_t>(v, 64) : kDefaultTraceBlocks; break;
    case PT_OPT_TRACE_REFILL_LANES:
        if (v > 64) return fail(PT_ERR_INVALID, "pt_renderer_set_option: refill lanes must be in [1, 64]");
        r->refillLanes = v > 0 ? (int)v : 0;
        break;
    case PT_OPT_SHADE_BLOCKS_PER_CU: r->shadeBlocksPerCU = v > 0 ? (int)std::min<int64_t>(v, 64) : 8; break;
    case PT_OPT_PROFILE_KERNELS: r->profile = v != 0; break;
    case PT_OPT_TRACE_STATS: r->traceStats = v != 0; break;
    case PT_OPT_CONSISTENT_MIS: r->misSample = v != 0; r->resetPending = true; break;
    case PT_OPT_GUIDE_TRAIN_STEPS: r->gTrainSteps = v < 0 ? 16 : (int)std::min<int64_t>(v, 1024); break;
    case PT_OPT_GUIDE_BOUNCES: r->gBounces = v <= 0 ? 0 : (int)std::min<int64_t>(v, kMaxBounces); break;
    case PT_OPT_BVH_BUILDER:
        if (v < 0 || v > 2) return fail(PT_ERR_INVALID, "pt_renderer_set_option: BVH builder must be 0, 1 (LBVH) or 2 (PLOC)");
        r->bvhBuilder = v == 0 ? kDefaultBvhBuilder : v == 1 ? PT_BVH_LBVH : PT_BVH_PLOC;
        break;
    case PT_OPT_GUIDE_BATCH:
        if (v < 32 || v > (1 << 22)) return fail(PT_ERR_INVALID, "pt_renderer_set_option: guide batch must be in [32, 2^22]");
        r->gBatch = (int)v;
        break;
    case PT_OPT_GUIDE_MAX_FRACTION_MILLI:
        if (v <= 0 || v > 1000) return fail(PT_ERR_INVALID, "pt_renderer_set_option: guided fraction cap must be in (0, 1000] thousandths");
        r->gFracCap = (float)v * 1e-3f;
        break;
    case PT_OPT_GUIDE_LR_MICRO:
        if (v <= 0) return fail(PT_ERR_INVALID, "pt_renderer_set_option: learning rate must be positive");
        r->gLr = (float)v * 1e-6f;
        break;
    case PT_OPT_GUIDE_RESET:
        // re-initialise the network (seed = value) at the next guided pass
        PT_HIP_TRY(hipStreamSynchronize(r->stream));
        r->gSeed = (uint64_t)v;
        r->guideReady = false;
        r->gTrainTotal = 0;
        r->resetPending = true;
        break;
    default: return fail(PT_ERR_INVALID, "pt_renderer_set_option: unknown option");
    }
    return PT_OK;
    PT_API_END("pt_renderer_set_option")
}

int pt_renderer_get_option(pt_renderer *r, int opt, int64_t *v) {
    if (guard("pt_renderer_get_option", r) || !v) return fail(PT_ERR_INVALID, "pt_renderer_get_option: bad arguments");
    switch (opt) {
    case PT_OPT_POOL_PATHS: *v = r->poolWanted; break;
    case PT_OPT_TRACE_BLOCKS_PER_CU: *v = r->traceBlocksPerCU; break;
    case PT_OPT_TRACE_LDS_STACK: *v = r->ldsStack; break;
    case PT_OPT_TRACE_REFILL_LANES: *v = r->refillLanes; break;
    case PT_OPT_SHADE_BLOCKS_PER_CU: *v = r->shadeBlocksPerCU; break;
    case PT_OPT_PROFILE_KERNELS: *v = r->profile ? 1 : 0; break;
    case PT_OPT_TRACE_STATS: *v = r->traceStats ? 1 : 0; break;
    case PT_OPT_CONSISTENT_MIS: *v = r->misSample; break;
    case PT_OPT_GUIDE_TRAIN_STEPS: *v = r->gTrainSteps; break;
    case PT_OPT_GUIDE_BOUNCES: *v = r->gBounces; break;
    case PT_OPT_BVH_BUILDER: *v = r->bvhBuilder == PT_BVH_LBVH ? 1 : 2; break;
    case PT_OPT_GUIDE_BATCH: *v = r->gBatch; break;
    case PT_OPT_GUIDE_LR_MICRO: *v = (int64_t)llroundf(r->gLr * 1e6f); break;
    case PT_OPT_GUIDE_MAX_FRACTION_MILLI: *v = (int64_t)llroundf(r->gFracCap * 1e3f); break;
    case PT_OPT_GUIDE_RESET: *v = (int64_t)r->gTrainTotal; break;
    default: return fail(PT_ERR_INVALID, "pt_renderer_get_option: unknown option");
    }
    return PT_OK;
}

int pt_renderer_render_async(pt_renderer *r, int spp) {
    if (guard("pt_renderer_render_async", r)) return PT_ERR_INVALID;
    if (!r->hasScene) return fail(PT_ERR_STATE, "pt_renderer_render_async: no scene");
    if (r->inFlight) return fail(PT_ERR_STATE, "pt_renderer_render_async: previous pass not polled");
    if (spp <= 0) return fail(PT_ERR_INVALID, "pt_renderer_render_async: spp must be positive");
    PT_API_BEGIN
    if (r->resetPending) zeroAccum(r);
    PT_HIP_TRY(hipEventRecord(r->evBegin, r->stream));
    const int npix = r->W * r->H;
    r->lastChunks.clear();
    if (r->type == PT_RENDERER_NORMALS) {
        // one centre ray per pixel, added once per pass (the reference's framebuffer sees
        // spp = 1 each pass, g_buffer.cu:108-117)
        if (npix > r->poolWanted) r->poolWanted = npix;
        ensurePool(r, npix);
        k_centre_rays<<<(npix + 255) / 256, 256, 0, r->stream>>>(r->cam, npix, r->W, r->pp->rayO[0].as<float4>(), r->pp->rayD[0].as<float4>());
        queueLinear(r, PT_Q_RAYS, npix);
        launchExtend(r, 0, ptSegLen(npix), r->pp->rayO[0].as<float4>(), r->pp->rayD[0].as<float4>(), r->pp->hits.as<float4>(), TRACE_CAMERA);
        k_normals_shade<<<(npix + 255) / 256, 256, 0, r->stream>>>(r->view, npix, r->pp->rayO[0].as<float4>(), r->pp->rayD[0].as<float4>(),
                                                                  r->pp->hits.as<float4>(), r->accum.as<double>());
        PT_HIP_TRY(hipGetLastError());
        r->spp += 1;
        r->passSpp = spp;
    } else {
        ensurePool(r, (int64_t)npix * spp);
        if (r->type == PT_RENDERER_GUIDED) ensureGuide(r);
        planChunks(r, spp, r->lastChunks);
        if (r->maxDepth > 0) {
            for (const ChunkParams &cp : r->lastChunks) {
                runChunk(r, cp);
                r->pixelsAccumulated += cp.npix;
                r->samplesAccumulated += (double)cp.npix * cp.samples;
            }
        }
        r->spp += spp;
        r->passSpp = spp;
    }
    r->lastPassPaths = (int64_t)npix * spp;
    r->lastPassSingleChunk = r->lastChunks.size() == 1;
    r->lastPassSamples = r->lastChunks.empty() ? 0 : r->lastChunks[0].samples;
    r->lastPassPixelMajor = !r->lastChunks.empty() && r->lastChunks[0].pixelMajor;
    tonemap(r);
    PT_HIP_TRY(hipEventRecord(r->evEnd, r->stream));
    r->inFlight = true;
    return PT_OK;
    PT_API_END("pt_renderer_render_async")
}

int pt_renderer_poll(pt_renderer *r, int *finished, float *ms) {
    if (guard("pt_renderer_poll", r)) return PT_ERR_INVALID;
    PT_API_BEGIN
    if (!r->inFlight) {
        if (finished) *finished = 1;
        if (ms) *ms = 0.f;
        return PT_OK;
    }
    const hipError_t q = hipEventQuery(r->evEnd);
    if (q == hipErrorNotReady) {
        if (finished) *finished = 0;
        return PT_OK;
    }
    if (q != hipSuccess) return fail(PT_ERR_DEVICE, std::string("pt_renderer_poll: ") + hipGetErrorString(q));
    float t = 0.f;
    (void)hipEventElapsedTime(&t, r->evBegin, r->evEnd);
    finishPass(r);
    if (finished) *finished = 1;
    if (ms) *ms = t;
    return PT_OK;
    PT_API_END("pt_renderer_poll")
}

int pt_renderer_synchronize(pt_renderer *r) {
    if (guard("pt_renderer_synchronize", r)) return PT_ERR_INVALID;
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    if (r->inFlight) finishPass(r);
    int ov = 0;
    PT_HIP_TRY(hipMemcpy(&ov, r->pp->overflow.p, sizeof(int), hipMemcpyDeviceToHost));
    if (ov != 0) return fail(PT_ERR_NUMERIC, "pt_renderer_synchronize: traversal stack overflowed");
    return PT_OK;
    PT_API_END("pt_renderer_synchronize")
}

int pt_renderer_get_stream(pt_renderer *r, void **stream) {
    if (guard("pt_renderer_get_stream", r) || !stream) return fail(PT_ERR_INVALID, "pt_renderer_get_stream: bad arguments");
    *stream = (void *)r->stream;
    return PT_OK;
}

int pt_renderer_reset(pt_renderer *r) {
    if (guard("pt_renderer_reset", r)) return PT_ERR_INVALID;
    r->resetPending = true;
    return PT_OK;
}

int pt_renderer_get_spp(pt_renderer *r, int64_t *spp) {
    if (guard("pt_renderer_get_spp", r) || !spp) return fail(PT_ERR_INVALID, "pt_renderer_get_spp: bad arguments");
    *spp = r->resetPending ? 0 : r->spp;
    return PT_OK;
}

int pt_renderer_get_radiance(pt_renderer *r, float *out) {
    if (guard("pt_renderer_get_radiance", r) || !out) return fail(PT_ERR_INVALID, "pt_renderer_get_radiance: bad arguments");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    if (r->inFlight) finishPass(r);
    PT_HIP_TRY(hipMemcpy(out, r->radiance.p, (size_t)r->W * r->H * 3 * sizeof(float), hipMemcpyDeviceToHost));
    return PT_OK;
    PT_API_END("pt_renderer_get_radiance")
}

int pt_renderer_get_radiance_sum(pt_renderer *r, double *out) {
    if (guard("pt_renderer_get_radiance_sum", r) || !out) return fail(PT_ERR_INVALID, "pt_renderer_get_radiance_sum: bad arguments");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    if (r->inFlight) finishPass(r);
    PT_HIP_TRY(hipMemcpy(out, r->accum.p, (size_t)r->W * r->H * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return PT_OK;
    PT_API_END("pt_renderer_get_radiance_sum")
}

int pt_renderer_get_rgba8(pt_renderer *r, uint8_t *out) {
    if (guard("pt_renderer_get_rgba8", r) || !out) return fail(PT_ERR_INVALID, "pt_renderer_get_rgba8: bad arguments");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    if (r->inFlight) finishPass(r);
    PT_HIP_TRY(hipMemcpy(out, r->rgba.p, (size_t)r->W * r->H * 4, hipMemcpyDeviceToHost));
    return PT_OK;
    PT_API_END("pt_renderer_get_rgba8")
}

int pt_renderer_hit_test(pt_renderer *r, int x, int y, int *material_id, int *is_hit) {
    if (guard("pt_renderer_hit_test", r)) return PT_ERR_INVALID;
    if (!r->hasScene) return fail(PT_ERR_STATE, "pt_renderer_hit_test: no scene");
    if (x < 0 || y < 0 || x >= r->W || y >= r->H) {
        // hit_test.cu:25-33: no thread matches an off-image pixel; report a miss
        if (material_id) *material_id = -1;
        if (is_hit) *is_hit = 0;
        return PT_OK;
    }
    if (r->inFlight) return fail(PT_ERR_STATE, "pt_renderer_hit_test: a pass is still running");
    PT_API_BEGIN
    ensurePool(r);
    // one centre ray through the same extend kernel (hit_test.cu:25-47)
    const float4 *dummy = nullptr;
    (void)dummy;
    float4 ro, rd;
    {
        // camera.cu:24-27 generateRay(row, col) with samples (0.5, 0.5)
        const PtCamera &c = r->cam;
        const float xc = (x + 0.5f) / c.W, yc = (y + 0.5f) / c.H;
        const float py = yc * c.height - c.top, px = xc * c.width - c.right;
        const float l = sqrtf(px * px + py * py + 1.f);
        const float dl[3] = {px / l, py / l, -1.f / l};
        float d[3];
        for (int k = 0; k < 3; k++) d[k] = c.m[k][0] * dl[0] + c.m[k][1] * dl[1] + c.m[k][2] * dl[2];
        ro = make_float4(c.m[0][3], c.m[1][3], c.m[2][3], 0.f);
        rd = make_float4(d[0], d[1], d[2], FLT_MAX);
    }
    PT_HIP_TRY(hipMemcpyAsync(r->pp->rayO[0].p, &ro, 16, hipMemcpyHostToDevice, r->stream));
    PT_HIP_TRY(hipMemcpyAsync(r->pp->rayD[0].p, &rd, 16, hipMemcpyHostToDevice, r->stream));
    queueLinear(r, PT_Q_RAYS, 1);
    launchExtend(r, 0, ptSegLen(1), r->pp->rayO[0].as<float4>(), r->pp->rayD[0].as<float4>(), r->pp->hits.as<float4>(), TRACE_CAMERA);
    float4 h;
    PT_HIP_TRY(hipMemcpyAsync(&h, r->pp->hits.p, 16, hipMemcpyDeviceToHost, r->stream));
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    const int slot = asInt(h.w);
    if (slot == INT32_MIN) {
        if (material_id) *material_id = -1;
        if (is_hit) *is_hit = 0;
    } else {
        float4 p0;
        PT_HIP_TRY(hipMemcpy(&p0, r->prims.as<float4>() + 3 * (size_t)slot, 16, hipMemcpyDeviceToHost));
        if (material_id) *material_id = asInt(p0.w);
        if (is_hit) *is_hit = 1;
    }
    return PT_OK;
    PT_API_END("pt_renderer_hit_test")
}

int pt_renderer_get_stats(pt_renderer *r, pt_render_stats *out) {
    if (guard("pt_renderer_get_stats", r) || !out) return fail(PT_ERR_INVALID, "pt_renderer_get_stats: bad arguments");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    if (r->inFlight) finishPass(r);
    std::vector<unsigned long long> s(S_COUNT);
    PT_HIP_TRY(hipMemcpy(s.data(), r->stats.p, S_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    int ov = 0;
    PT_HIP_TRY(hipMemcpy(&ov, r->pp->overflow.p, sizeof(int), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof *out);
    for (int b = 0; b < kMaxBounces; b++) {
        out->extension_rays += (int64_t)s[S_RAYS + b];
        out->shadow_rays += (int64_t)s[S_SHADOW + b];
        if (b < 16) {
            out->bounce_rays[b] = (int64_t)s[S_RAYS + b];
            out->bounce_shadow[b] = (int64_t)s[S_SHADOW + b];
        }
    }
    out->rays_total = out->extension_rays + out->shadow_rays;
    for (int q = 0; q < 2; q++) {
        out->trace_rays[q] = (int64_t)s[S_TRACE + 3 * q];
        out->trace_nodes[q] = (int64_t)s[S_TRACE + 3 * q + 1];
        out->trace_prims[q] = (int64_t)s[S_TRACE + 3 * q + 2];
    }
    out->paths = (int64_t)s[S_PATHS];
    out->passes = r->passes;
    out->gpu_ms = r->gpuMs;
    out->bvh_depth = r->bvh.depth;
    out->bvh_nodes = r->bvh.numNodes;
    out->stack_overflows = ov;
    out->guide_steps = r->gTrainTotal;
    if (r->guideReady && r->gTrainTotal > 0) {
        float l = 0.f;
        PT_HIP_TRY(hipMemcpy(&l, r->gLoss.p, sizeof(float), hipMemcpyDeviceToHost));
        out->guide_loss = l;
    }
    return PT_OK;
    PT_API_END("pt_renderer_get_stats")
}

int pt_renderer_reset_stats(pt_renderer *r) {
    if (guard("pt_renderer_reset_stats", r)) return PT_ERR_INVALID;
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    if (r->inFlight) finishPass(r);
    k_zero_u64<<<1, 512, 0, r->stream>>>(r->stats.as<unsigned long long>(), S_COUNT);
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    r->gpuMs = 0.0;
    r->passes = 0;
    for (int k = 0; k < K_NUM; k++) {
        r->kstat[k].launches = 0;
        r->kstat[k].ms = 0.0;
        r->kstat[k].units = 0;
        r->kstat[k].bytes = 0.0;
    }
    r->samplesAccumulated = 0.0;
    r->pixelsAccumulated = 0.0;
    return PT_OK;
    PT_API_END("pt_renderer_reset_stats")
}

// Per-kernel statistics since the last pt_renderer_reset_stats: launch count and time from
// the HIP events recorded around each launch (profiling on); units and algorithmic bytes from
// the per-bounce queue counters the accumulate kernel folds into the device statistics.
int pt_renderer_get_kernel_stats(pt_renderer *r, pt_kernel_stat *out, int cap, int *count) {
    if (guard("pt_renderer_get_kernel_stats", r) || !out || !count) return fail(PT_ERR_INVALID, "pt_renderer_get_kernel_stats: bad arguments");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    if (r->inFlight) finishPass(r);
    std::vector<unsigned long long> s(S_COUNT);
    PT_HIP_TRY(hipMemcpy(s.data(), r->stats.p, S_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    double rays = 0, shadows = 0, addShade = 0, addShadow = 0, nextRays = 0;
    for (int b = 0; b < kMaxBounces; b++) {
        rays += (double)s[S_RAYS + b];
        shadows += (double)s[S_SHADOW + b];
        addShade += (double)s[S_LADD_SHADE + b];
        addShadow += (double)s[S_LADD_SHADOW + b];
        if (b > 0) nextRays += (double)s[S_RAYS + b];
    }
    const double paths = (double)s[S_PATHS];
    const double pixelsAcc = r->pixelsAccumulated, samplesAcc = r->samplesAccumulated;
    const double npix = (double)r->W * r->H * (double)r->kstat[K_TONEMAP].launches;
    // algorithmic bytes per unit (record sizes above)
    pt_kernel_stat *k = r->kstat;
    // camera rays (bounce 0) carry only a direction: no origin, state or throughput reads, and
    // shade writes their paths' radiance once instead of a read-modify-write
    const double camRays = (double)s[S_RAYS], camAdds = (double)s[S_LADD_SHADE];
    k[K_RAYGEN].units = (int64_t)paths;
    k[K_RAYGEN].bytes = 16.0 * paths;                    // writes rayD
    k[K_EXTEND].units = (int64_t)rays;
    k[K_EXTEND].bytes = 48.0 * rays - 16.0 * camRays;    // reads rayO (not for camera rays), rayD; writes hit
    k[K_SHADE].units = (int64_t)rays;
    k[K_SHADE].bytes = 80.0 * rays - 48.0 * camRays + 64.0 * nextRays + 48.0 * shadows + 32.0 * (addShade - camAdds) + 16.0 * camRays;
    k[K_SHADOW].units = (int64_t)shadows;
    k[K_SHADOW].bytes = 48.0 * shadows + 32.0 * addShadow;
    k[K_ACCUM].units = (int64_t)pixelsAcc;
    k[K_ACCUM].bytes = 16.0 * samplesAcc + 48.0 * pixelsAcc;
    k[K_TONEMAP].units = (int64_t)npix;
    k[K_TONEMAP].bytes = 40.0 * npix;
    const int n = std::min(cap, (int)K_NUM);
    for (int i = 0; i < n; i++) {
        out[i] = k[i];
        strncpy(out[i].name, kKernelNames[i], sizeof(out[i].name) - 1);
        out[i].name[sizeof(out[i].name) - 1] = 0;
    }
    *count = n;
    return PT_OK;
    PT_API_END("pt_renderer_get_kernel_stats")
}

int pt_renderer_get_path_radiance(pt_renderer *r, float *out, int64_t cap, int64_t *count) {
    if (guard("pt_renderer_get_path_radiance", r) || !out || !count) return fail(PT_ERR_INVALID, "pt_renderer_get_path_radiance: bad arguments");
    if (!r->lastPassSingleChunk) return fail(PT_ERR_STATE, "pt_renderer_get_path_radiance: last pass did not fit in one chunk");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    if (r->inFlight) finishPass(r);
    const int64_t n = std::min<int64_t>(cap, r->lastPassPaths);
    std::vector<float4> tmp((size_t)r->lastPassPaths);
    PT_HIP_TRY(hipMemcpy(tmp.data(), r->pp->Lpath.p, tmp.size() * sizeof(float4), hipMemcpyDeviceToHost));
    // the pass's paths in sample-major order (path = s * npix + pixel), whatever the numbering
    // the chunk used on the device
    const int64_t samples = std::max(1, r->lastPassSamples), npix = r->lastPassPaths / samples;
    for (int64_t i = 0; i < n; i++) {
        const int64_t j = r->lastPassPixelMajor ? (i % npix) * samples + i / npix : i;
        out[i * 3] = tmp[(size_t)j].x;
        out[i * 3 + 1] = tmp[(size_t)j].y;
        out[i * 3 + 2] = tmp[(size_t)j].z;
    }
    *count = n;
    return PT_OK;
    PT_API_END("pt_renderer_get_path_radiance")
}

int pt_pool_layout(int width, int height, int spp, int64_t pool_option, int64_t *capacity, int64_t *chunks, int64_t *max_slot) {
    if (width <= 0 || height <= 0 || spp <= 0 || pool_option > (1ll << 27) || !capacity || !max_slot)
        return fail(PT_ERR_INVALID, "pt_pool_layout: bad arguments");
    if ((int64_t)width * height > (1ll << 28)) return fail(PT_ERR_INVALID, "pt_pool_layout: image too large");
    PT_API_BEGIN
    // a renderer object without device state: the pass plan is the one render_async makes
    pt_renderer r;
    r.W = width;
    r.H = height;
    r.poolWanted = pool_option > 0 ? pool_option : kDefaultPool;
    r.pp->pool = ptPoolCapacity((int64_t)width * height * spp, r.poolWanted);
    std::vector<ChunkParams> plan;
    planChunks(&r, spp, plan);
    int64_t hi = -1;
    int seen = -1;
    for (const ChunkParams &cp : plan) {
        if (cp.nPaths == seen) continue;  // chunks of one size share their layout
        seen = cp.nPaths;
        // camera queue: each segment's last run lies among the chunk's last PT_QUEUE_SEGS runs
        for (int p = std::max(0, cp.nPaths - 64 * PT_QUEUE_SEGS); p < cp.nPaths; p++) hi = std::max<int64_t>(hi, ptCameraSlot(p, cp.segLen));
        // linear queues fill the segments in order; appended queues (continuation, shadow rays,
        // guided records) stay inside the segment of the vertex that spawned them
        hi = std::max<int64_t>(hi, ptQueueSlots(cp.nPaths) - 1);
    }
    *capacity = r.pp->pool;
    if (chunks) *chunks = (int64_t)plan.size();
    *max_slot = hi;
    r.pp->pool = 0;
    return PT_OK;
    PT_API_END("pt_pool_layout")
}

int pt_renderer_intersect(pt_renderer *r, const float *rays, int64_t n, int any_hit, float *out) {
    if (guard("pt_renderer_intersect", r) || (n > 0 && (!rays || !out)) || n < 0) return fail(PT_ERR_INVALID, "pt_renderer_intersect: bad arguments");
    if (!r->hasScene) return fail(PT_ERR_STATE, "pt_renderer_intersect: no scene");
    if (r->inFlight) return fail(PT_ERR_STATE, "pt_renderer_intersect: a pass is still running");
    PT_API_BEGIN
    ensurePool(r);
    for (int64_t done = 0; done < n;) {
        const int cnt = (int)std::min<int64_t>(n - done, r->pp->pool);
        std::vector<float4> o(cnt), d(cnt);
        for (int i = 0; i < cnt; i++) {
            const float *q = rays + (done + i) * 8;
            o[i] = make_float4(q[0], q[1], q[2], q[3]);
            d[i] = make_float4(q[4], q[5], q[6], q[7]);
        }
        if (!any_hit) {
            PT_HIP_TRY(hipMemcpyAsync(r->pp->rayO[0].p, o.data(), cnt * 16, hipMemcpyHostToDevice, r->stream));
            PT_HIP_TRY(hipMemcpyAsync(r->pp->rayD[0].p, d.data(), cnt * 16, hipMemcpyHostToDevice, r->stream));
            queueLinear(r, PT_Q_RAYS, cnt);
            launchExtend(r, 0, ptSegLen(cnt), r->pp->rayO[0].as<float4>(), r->pp->rayD[0].as<float4>(), r->pp->hits.as<float4>(), TRACE_INCOHERENT);
            k_map_prims<<<(cnt + 255) / 256, 256, 0, r->stream>>>(r->view, r->pp->hits.as<float4>(), cnt, r->pp->sL.as<float4>());
            std::vector<float4> h(cnt);
            PT_HIP_TRY(hipMemcpyAsync(h.data(), r->pp->sL.p, cnt * 16, hipMemcpyDeviceToHost, r->stream));
            PT_HIP_TRY(hipStreamSynchronize(r->stream));
            for (int i = 0; i < cnt; i++) memcpy(out + (done + i) * 4, &h[i], 16);
        } else {
            // shadow queue: occlusion over [tmin = 1e-4, tmax]; the contribution (1, 0, 0) lands
            // in Lpath[i] when the ray is unoccluded
            std::vector<float4> so(cnt), sd(cnt), sl(cnt, make_float4(1.f, 0.f, 0.f, 0.f)), z(cnt, make_float4(0.f, 0.f, 0.f, 0.f));
            for (int i = 0; i < cnt; i++) {
                so[i] = make_float4(o[i].x, o[i].y, o[i].z, d[i].w);
                sd[i] = make_float4(d[i].x, d[i].y, d[i].z, asFloat(i));
            }
            PT_HIP_TRY(hipMemcpyAsync(r->pp->sO.p, so.data(), cnt * 16, hipMemcpyHostToDevice, r->stream));
            PT_HIP_TRY(hipMemcpyAsync(r->pp->sD.p, sd.data(), cnt * 16, hipMemcpyHostToDevice, r->stream));
            PT_HIP_TRY(hipMemcpyAsync(r->pp->sL.p, sl.data(), cnt * 16, hipMemcpyHostToDevice, r->stream));
            PT_HIP_TRY(hipMemcpyAsync(r->pp->Lpath.p, z.data(), cnt * 16, hipMemcpyHostToDevice, r->stream));
            queueLinear(r, PT_Q_SHADOW, cnt);
            launchShadow(r, 0, ptSegLen(cnt));
            std::vector<float4> h(cnt);
            PT_HIP_TRY(hipMemcpyAsync(h.data(), r->pp->Lpath.p, cnt * 16, hipMemcpyDeviceToHost, r->stream));
            PT_HIP_TRY(hipStreamSynchronize(r->stream));
            for (int i = 0; i < cnt; i++) {
                const int prim = h[i].x > 0.5f ? INT32_MIN : 0;  // visible -> no occluder
                out[(done + i) * 4] = h[i].x > 0.5f ? -1.f : 0.f;
                out[(done + i) * 4 + 1] = 0.f;
                out[(done + i) * 4 + 2] = 0.f;
                memcpy(&out[(done + i) * 4 + 3], &prim, 4);
            }
        }
        done += cnt;
    }
    return PT_OK;
    PT_API_END("pt_renderer_intersect")
}

}  // extern "C"

// ============================================================================================
// Autotuner: coordinate descent over the path-pool capacity, trace and shade blocks per CU
// and the LDS part of the traversal stack, each candidate timed on a probe render.
// ============================================================================================
namespace {

struct TuneCfg {
    int64_t pool;
    int trace, shade, lds;
};

// A candidate replaces the current configuration only when it is faster by more than this
// factor: probe passes vary by 1-2% from run to run, and the defaults were tuned on full passes.
static constexpr double kTuneMargin = 1.02;

template <class Probe>
TuneCfg coordinateDescent(TuneCfg cur, const std::vector<int64_t> &pools, Probe probe, int *probes, double *bestRate) {
    double best = probe(cur);
    int n = 1;
    const std::vector<int> traces = {8, 11, 12, 14}, shades = {4, 8, 16}, ldss = {16, 32};
    for (int dim = 0; dim < 4; dim++) {
        const size_t count = dim == 0 ? pools.size() : dim == 1 ? traces.size() : dim == 2 ? ldss.size() : shades.size();
        for (size_t k = 0; k < count; k++) {
            TuneCfg c = cur;
            if (dim == 0) c.pool = pools[k];
            if (dim == 1) c.trace = traces[k];
            if (dim == 2) c.lds = ldss[k];
            if (dim == 3) c.shade = shades[k];
            if (c.pool == cur.pool && c.trace == cur.trace && c.lds == cur.lds && c.shade == cur.shade) continue;
            const double rate = probe(c);
            n++;
            if (rate > best * kTuneMargin) {
                best = rate;
                cur = c;
            }
        }
    }
    *probes = n;
    *bestRate = best;
    return cur;
}

void applyCfg(pt_renderer *r, const TuneCfg &c) {
    r->poolWanted = c.pool;
    r->traceBlocksPerCU = c.trace;
    r->shadeBlocksPerCU = c.shade;
    r->ldsStack = c.lds;
}

std::vector<int64_t> poolCandidates(int64_t maxPaths) {
    std::vector<int64_t> v;
    for (int64_t p = 1 << 20; p <= kDefaultPool; p <<= 2) {
        v.push_back(p);
        if (p >= maxPaths) break;
    }
    return v;
}

}  // namespace

// ============================================================================================
// Batch of independent renders in one captured graph
// ============================================================================================
// A chunk of the batch wavefront: its launch parameters (groups[firstGroup ..]), its
// accumulate grid, and the renders (batch order positions) whose last rows it holds.
struct WaveChunk {
    ChunkParams cp{};
    int firstGroup = 0;
    int accThreads = 0;
    std::vector<int> finishing;
};

struct pt_batch {
    hipStream_t stream = nullptr;
    std::shared_ptr<PathPool> pp = std::make_shared<PathPool>();
    std::vector<pt_renderer *> jobs;
    std::vector<int> spp;
    TuneCfg cfg{kDefaultPool, kDefaultTraceBlocks, 8, 16};
    bool useGraph = true;
    // one wavefront over all renders (paths of different renders in the same queues), or each
    // render's own chunks one render after the other
    bool wavefront = true;
    pt_renderer *wave = nullptr;   // merged scene + launch settings of the wavefront
    size_t waveJobs = 0;           // renders merged into `wave`
    DevBuf renderTable, groupTable, bucketTable, renderRays, waveStats;
    std::vector<WaveChunk> waveChunks;
    std::vector<BatchGroup> waveGroups;
    std::vector<int> waveLead;     // renders with no rows before the first chunk (depth 0)
    std::vector<int64_t> waveCamRays;   // per render: its camera rays in chunks of several renders
    std::vector<char> waveOwn;          // per render: some chunk runs the render's own kernels
    DevBuf stamps;
    double clockKHz = 100000.0;
    // last run
    std::vector<double> latencyMs;
    double wallMs = 0.0;
    int chunks = 0;
    size_t graphNodes = 0;
    std::vector<int> order;
    // captured graph of the last plan
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graphScale = -1;
    TuneCfg graphCfg{0, 0, 0, 0};
    size_t graphJobs = 0;
    void dropGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graphScale = -1;
    }
    ~pt_batch() { delete wave; }
};

namespace {

bool sameCfg(const TuneCfg &a, const TuneCfg &b) { return a.pool == b.pool && a.trace == b.trace && a.shade == b.shade && a.lds == b.lds; }

// Queues every render of the batch (shortest expected first: pixels x spp x log2 triangles)
// on the batch stream: accumulators and counters cleared, chunks traced, then a completion
// stamp per render.  spp of each render is divided by `scale` (autotuner probes).
// Device memory the batch's kernels need (shared pool, stack spill area, stamps) and the
// render order; allocates, so it runs before any stream capture.
void buildWave(pt_batch *b);
void planWave(pt_batch *b, int scale);

void prepareBatch(pt_batch *b, int scale) {
    for (pt_renderer *r : b->jobs) applyCfg(r, b->cfg);
    int64_t need = 0, total = 0;
    int deepest = 0;
    for (size_t j = 0; j < b->jobs.size(); j++) {
        pt_renderer *r = b->jobs[j];
        const int64_t paths = (int64_t)r->W * r->H * std::max(1, b->spp[j] / scale);
        need = std::max<int64_t>(need, paths);
        total += paths;
        if (r->stackNeed > b->jobs[deepest]->stackNeed) deepest = (int)j;
    }
    // the wavefront fills chunks with whole sample rows of several renders: size the pool by
    // all of them, and at least one row of the largest image
    if (b->wavefront) {
        int64_t row = 0;
        for (pt_renderer *r : b->jobs) row = std::max<int64_t>(row, (int64_t)r->W * r->H);
        b->jobs[0]->poolWanted = std::max<int64_t>(b->cfg.pool, row);
        need = std::max<int64_t>(std::min<int64_t>(total, b->cfg.pool), row);
    }
    ensurePool(b->jobs[0], need);
    int cap = 0;
    ensureSpill(b->jobs[deepest], b->jobs[0]->cus * b->cfg.trace, &cap);
    b->stamps.alloc((b->jobs.size() + 1) * sizeof(unsigned long long));
    b->order.resize(b->jobs.size());
    for (size_t j = 0; j < b->jobs.size(); j++) b->order[j] = (int)j;
    auto cost = [&](int j) {
        const pt_renderer *r = b->jobs[j];
        return (double)r->W * r->H * b->spp[j] * std::log2(2.0 + (double)r->triMat.size() + (double)r->sphMat.size());
    };
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return cost(x) < cost(y); });
    if (b->wavefront) {
        buildWave(b);
        applyCfg(b->wave, b->cfg);
        planWave(b, scale);
    }
}

// The merged scene of the batch wavefront: every render's nodes, primitives, normals,
// materials and emitters concatenated (references rebased), and the per-render table.
void buildWave(pt_batch *b) {
    if (b->wave && b->waveJobs == b->jobs.size()) return;
    delete b->wave;
    b->wave = nullptr;
    pt_renderer *w = new pt_renderer();
    b->wave = w;
    w->stream = b->stream;
    w->ownsStream = false;
    w->pp = b->pp;
    w->cus = b->jobs[0]->cus;
    int64_t nN = 0, nP = 0, nNrm = 0, nM = 0, nE = 0;
    for (pt_renderer *r : b->jobs) {
        nN += r->bvh.numNodes;
        nP += r->view.numPrims;
        nNrm += (int64_t)(r->normals.bytes / sizeof(float4));
        nM += (int64_t)r->mats.size();
        nE += r->view.numEmitters;
        w->stackNeed = std::max(w->stackNeed, r->stackNeed);
    }
    if (nP >= (1ll << PT_LEAF_COUNT_SHIFT) || nP * 48 >= (1ll << 32) || nN * (int64_t)PT_NODE_F4 * 16 >= (1ll << 32))
        throw HipFailure("batch: too many primitives for one wavefront");
    w->nodes.alloc(sizeof(float4) * PT_NODE_F4 * (size_t)std::max<int64_t>(nN, 1));
    w->prims.alloc(sizeof(float4) * 3 * (size_t)std::max<int64_t>(nP, 1));
    w->normals.alloc(sizeof(float4) * (size_t)std::max<int64_t>(nNrm, 1));
    w->materials.alloc(sizeof(PtMaterial) * (size_t)std::max<int64_t>(nM, 1));
    w->emitters.alloc(sizeof(float4) * 3 * (size_t)std::max<int64_t>(nE, 1));
    std::vector<BatchRender> table(b->jobs.size());
    int64_t oN = 0, oP = 0, oNrm = 0, oM = 0, oE = 0;
    for (size_t j = 0; j < b->jobs.size(); j++) {
        pt_renderer *r = b->jobs[j];
        const int nn = r->bvh.numNodes, np = r->view.numPrims, ne = r->view.numEmitters;
        const int64_t nrm = (int64_t)(r->normals.bytes / sizeof(float4)), nm = (int64_t)r->mats.size();
        float4 *dn = w->nodes.as<float4>() + (size_t)PT_NODE_F4 * oN, *dp = w->prims.as<float4>() + 3 * (size_t)oP;
        float4 *de = w->emitters.as<float4>() + 3 * (size_t)oE;
        PT_HIP_TRY(hipMemcpyAsync(dn, r->nodes.p, sizeof(float4) * PT_NODE_F4 * (size_t)nn, hipMemcpyDeviceToDevice, b->stream));
        PT_HIP_TRY(hipMemcpyAsync(dp, r->prims.p, sizeof(float4) * 3 * (size_t)np, hipMemcpyDeviceToDevice, b->stream));
        if (nrm > 0)
            PT_HIP_TRY(hipMemcpyAsync(w->normals.as<float4>() + oNrm, r->normals.p, sizeof(float4) * (size_t)nrm, hipMemcpyDeviceToDevice, b->stream));
        PT_HIP_TRY(hipMemcpyAsync(w->materials.as<PtMaterial>() + oM, r->materials.p, sizeof(PtMaterial) * (size_t)nm, hipMemcpyDeviceToDevice,
                                  b->stream));
        if (ne > 0) PT_HIP_TRY(hipMemcpyAsync(de, r->emitters.p, sizeof(float4) * 3 * (size_t)ne, hipMemcpyDeviceToDevice, b->stream));
        if (nn > 0) k_rebase_nodes<<<(nn + 255) / 256, 256, 0, b->stream>>>(dn, nn, (int)oN, (int)oP);
        if (np > 0) k_rebase_prims<<<(np + 255) / 256, 256, 0, b->stream>>>(dp, np, (int)oM, (int)oNrm);
        if (ne > 0) k_rebase_prims<<<(ne + 255) / 256, 256, 0, b->stream>>>(de, ne, (int)oM, (int)oNrm);
        PT_HIP_TRY(hipGetLastError());
        BatchRender &R = table[j];
        R.root = (int)oN;
        R.emitBegin = (int)oE;
        R.numEmitters = ne;
        R.maxDepth = r->maxDepth;
        R.W = r->W;
        R.H = r->H;
        R.npix = r->W * r->H;
        R.seed = r->seed;
        R.env = r->view.env;
        R.cam = r->cam;
        R.accum = r->accum.as<double>();
        oN += nn;
        oP += np;
        oNrm += nrm;
        oM += nm;
        oE += ne;
    }
    w->view.nodes = w->nodes.as<float4>();
    w->view.prims = w->prims.as<float4>();
    w->view.normals = w->normals.as<float4>();
    w->view.materials = w->materials.as<PtMaterial>();
    w->view.emitters = w->emitters.as<float4>();
    w->view.numPrims = (int)nP;
    w->view.numNodes = (int)nN;
    w->view.numEmitters = 0;   // per render (BatchRender)
    w->view.env = PtEnv{};
    b->renderTable.alloc(sizeof(BatchRender) * table.size());
    PT_HIP_TRY(hipMemcpyAsync(b->renderTable.p, table.data(), sizeof(BatchRender) * table.size(), hipMemcpyHostToDevice, b->stream));
    b->renderRays.alloc(sizeof(unsigned long long) * b->jobs.size());
    b->waveStats.alloc(S_COUNT * sizeof(unsigned long long));
    w->batchRenders = b->renderTable.as<BatchRender>();
    PT_HIP_TRY(hipStreamSynchronize(b->stream));
    b->waveJobs = b->jobs.size();
}

// Cuts the batch's sample rows (renders in batch order, each render's rows in order) into
// chunks of at most the pool's paths; a chunk's groups are runs of rows of one render.
void planWave(pt_batch *b, int scale) {
    // the configured capacity (at least one row of the largest image), within the allocated
    // pool (which does not shrink)
    int64_t row = 0;
    for (pt_renderer *r : b->jobs) row = std::max<int64_t>(row, (int64_t)r->W * r->H);
    const int64_t pool = std::min<int64_t>(b->pp->pool, std::max<int64_t>({4096, b->cfg.pool, row}));
    b->waveChunks.clear();
    b->waveGroups.clear();
    b->waveLead.clear();
    const pt_renderer *first = b->jobs[0];
    WaveChunk cur;
    int64_t paths = 0, acc = 0;
    auto close = [&]() {
        if (paths == 0) {
            // renders that added no rows (depth 0) complete with the chunk before, or first
            if (!b->waveChunks.empty()) {
                for (int k : cur.finishing) b->waveChunks.back().finishing.push_back(k);
            } else {
                b->waveLead = cur.finishing;
            }
            cur.finishing.clear();
            return;
        }
        ChunkParams &cp = cur.cp;
        cp.nPaths = (int)paths;
        cp.segLen = ptSegLen(paths);
        cp.nee = first->nee;
        cp.misSample = first->misSample;
        cp.numGroups = (int)b->waveGroups.size() - cur.firstGroup;
        cur.accThreads = (int)acc;
        b->waveChunks.push_back(cur);
        cur = WaveChunk();
        cur.firstGroup = (int)b->waveGroups.size();
        paths = 0;
        acc = 0;
    };
    cur.firstGroup = 0;
    for (size_t k = 0; k < b->order.size(); k++) {
        const int j = b->order[k];
        const pt_renderer *r = b->jobs[j];
        const int npix = r->W * r->H;
        // a render of depth 0 traces nothing (its image stays black, as on its own)
        const int spp = r->maxDepth > 0 ? std::max(1, b->spp[j] / scale) : 0;
        for (int s0 = 0; s0 < spp;) {
            const int64_t room = (pool - paths) / npix;
            if (room <= 0) {
                if (paths == 0) throw HipFailure("batch: a sample row does not fit the path pool");
                close();
                continue;
            }
            const int c = (int)std::min<int64_t>(spp - s0, room);
            BatchGroup g;
            g.pathBegin = (int)paths;
            g.accBegin = (int)acc;
            g.render = j;
            g.npix = npix;
            g.sampleBegin = s0;
            g.count = c;
            b->waveGroups.push_back(g);
            cur.cp.maxDepth = std::max(cur.cp.maxDepth, r->maxDepth);
            paths += (int64_t)c * npix;
            acc += npix;
            s0 += c;
        }
        cur.finishing.push_back((int)k);
    }
    close();
    auto consecutive = [](const std::vector<int> &v) {
        for (size_t i = 1; i < v.size(); i++)
            if (v[i] != v[i - 1] + 1) return false;
        return true;
    };
    bool ok = consecutive(b->waveLead);
    for (const WaveChunk &c : b->waveChunks) ok = ok && consecutive(c.finishing);
    if (!ok) throw HipFailure("batch: renders completing with a chunk are not consecutive");
    b->groupTable.alloc(sizeof(BatchGroup) * std::max<size_t>(b->waveGroups.size(), 1));
    PT_HIP_TRY(hipMemcpy(b->groupTable.p, b->waveGroups.data(), sizeof(BatchGroup) * b->waveGroups.size(), hipMemcpyHostToDevice));
    // per chunk: the group of each bucket's first path, with what a path lookup needs
    std::vector<BatchBucket> buckets;
    std::vector<size_t> bucketAt;
    for (const WaveChunk &c : b->waveChunks) {
        bucketAt.push_back(buckets.size());
        const int nb = ((c.cp.nPaths - 1) >> kBatchBucketShift) + 1;
        int g = 0;
        for (int k = 0; k < nb; k++) {
            const int path = k << kBatchBucketShift;
            while (g + 1 < c.cp.numGroups && b->waveGroups[c.firstGroup + g + 1].pathBegin <= path) g++;
            const BatchGroup &gr = b->waveGroups[c.firstGroup + g];
            BatchBucket bk;
            bk.group = g;
            bk.pathBegin = gr.pathBegin;
            bk.render = gr.render;
            bk.count = gr.count;
            bk.sampleBegin = gr.sampleBegin;
            bk.nextBegin = g + 1 < c.cp.numGroups ? b->waveGroups[c.firstGroup + g + 1].pathBegin : INT32_MAX;
            bk.seed = b->jobs[gr.render]->seed;
            buckets.push_back(bk);
        }
    }
    b->bucketTable.alloc(sizeof(BatchBucket) * std::max<size_t>(buckets.size(), 1));
    PT_HIP_TRY(hipMemcpy(b->bucketTable.p, buckets.data(), sizeof(BatchBucket) * buckets.size(), hipMemcpyHostToDevice));
    b->waveCamRays.assign(b->jobs.size(), 0);
    b->waveOwn.assign(b->jobs.size(), 0);
    for (const WaveChunk &c : b->waveChunks) {
        if (c.cp.numGroups == 1) b->waveOwn[b->waveGroups[c.firstGroup].render] = 1;
        else
            for (int k = 0; k < c.cp.numGroups; k++) {
                const BatchGroup &g = b->waveGroups[c.firstGroup + k];
                b->waveCamRays[g.render] += (int64_t)g.npix * g.count;
            }
    }
    for (size_t k = 0; k < b->waveChunks.size(); k++) {
        WaveChunk &c = b->waveChunks[k];
        c.cp.groups = b->groupTable.as<BatchGroup>() + c.firstGroup;
        c.cp.buckets = b->bucketTable.as<BatchBucket>() + bucketAt[k];
        c.cp.renders = b->renderTable.as<BatchRender>();
        c.cp.renderRays = b->renderRays.as<unsigned long long>();
    }
}

// One chunk of the batch wavefront: raygen from every group's camera, the bounces with the
// batch kernels, accumulation into each render's accumulator.  A chunk that holds sample rows
// of one render only (the large renders) runs that render's own kernels instead, exactly its
// chunk of the render-after-render schedule: the batch kernels' per-path render lookups buy
// nothing there.
void runWaveChunk(pt_batch *b, const WaveChunk &c) {
    pt_renderer *w = b->wave;
    const ChunkParams &cp = c.cp;
    PathPool &pp = *b->pp;
    if (cp.numGroups == 1) {
        const BatchGroup &g = b->waveGroups[c.firstGroup];
        pt_renderer *r = b->jobs[g.render];
        ChunkParams one;
        one.pixBegin = 0;
        one.npix = g.npix;
        one.samples = g.count;
        one.sampleBase = g.sampleBegin;
        one.nPaths = g.npix * g.count;
        one.segLen = ptSegLen(one.nPaths);
        one.maxDepth = r->maxDepth;
        one.nee = r->nee;
        one.misSample = r->misSample;
        one.W = r->W;
        one.H = r->H;
        one.seed = r->seed;
        one.pixelMajor = 1;
        runChunk(r, one);
        return;
    }
    int *ctr = pp.counters.as<int>();
    k_queue_reset<<<PT_QUEUE_SEGS + 1, 256, 0, b->stream>>>(ctr, PT_Q_RAYS, cp.nPaths, cp.segLen, 1);
    k_raygen_batch<<<(cp.nPaths + 255) / 256, 256, 0, b->stream>>>(cp, pp.rayD[0].as<float4>());
    PT_HIP_TRY(hipGetLastError());
    const int shadeGrid = (int)std::max<int64_t>(PT_QUEUE_SEGS / 4, std::min<int64_t>((cp.nPaths + 255) / 256, (int64_t)w->cus * w->shadeBlocksPerCU));
    int cur = 0;
    for (int bn = 0; bn < cp.maxDepth; bn++) {
        launchExtend(w, bn, cp.segLen, pp.rayO[cur].as<float4>(), pp.rayD[cur].as<float4>(), pp.hits.as<float4>(),
                     bn == 0 ? TRACE_CAMERA : TRACE_INCOHERENT, make_float4(0.f, 0.f, 0.f, 0.f), &cp);
        // w = 1 marks bounce 0 (camera vertices); batch kernels take each render's camera
        const float4 camFlag = make_float4(0.f, 0.f, 0.f, bn == 0 ? 1.f : 0.f);
        k_shade<false, true><<<shadeGrid, 256, 0, b->stream>>>(
            w->view, cp, bn, ctr, pp.rayO[cur].as<float4>(), pp.rayD[cur].as<float4>(), pp.st[cur].as<uint4>(), pp.beta[cur].as<float4>(),
            pp.hits.as<float4>(), pp.rayO[1 - cur].as<float4>(), pp.rayD[1 - cur].as<float4>(), pp.st[1 - cur].as<uint4>(),
            pp.beta[1 - cur].as<float4>(), pp.sO.as<float4>(), pp.sD.as<float4>(), pp.sL.as<float4>(), pp.Lpath.as<float4>(), GuideShadeArgs{},
            camFlag);
        PT_HIP_TRY(hipGetLastError());
        if (cp.nee) launchShadow(w, bn, cp.segLen);
        cur = 1 - cur;
    }
    k_accumulate_batch<<<(c.accThreads + 255) / 256, 256, 0, b->stream>>>(cp, c.accThreads, pp.Lpath.as<float4>(), ctr,
                                                                          b->waveStats.as<unsigned long long>());
    PT_HIP_TRY(hipGetLastError());
}

void enqueueWave(pt_batch *b, int scale) {
    unsigned long long *st = b->stamps.as<unsigned long long>();
    k_stamp<<<1, 64, 0, b->stream>>>(st);
    for (pt_renderer *r : b->jobs) zeroAccum(r);
    k_zero_u64<<<1, 512, 0, b->stream>>>(b->waveStats.as<unsigned long long>(), S_COUNT);
    const int nj = (int)b->jobs.size();
    k_zero_u64<<<(nj + 255) / 256, 256, 0, b->stream>>>(b->renderRays.as<unsigned long long>(), nj);
    // statistics of the chunks that run a render's own kernels
    for (int j = 0; j < nj; j++)
        if (b->waveOwn[j]) k_zero_u64<<<1, 512, 0, b->stream>>>(b->jobs[j]->stats.as<unsigned long long>(), S_COUNT);
    PT_HIP_TRY(hipGetLastError());
    // the renders completing with a chunk are consecutive in the batch order
    if (!b->waveLead.empty()) k_stamp_range<<<1, 64, 0, b->stream>>>(st + 1 + b->waveLead.front(), (int)b->waveLead.size());
    for (const WaveChunk &c : b->waveChunks) {
        runWaveChunk(b, c);
        if (!c.finishing.empty()) k_stamp_range<<<1, 64, 0, b->stream>>>(st + 1 + c.finishing.front(), (int)c.finishing.size());
    }
    for (size_t j = 0; j < b->jobs.size(); j++) b->jobs[j]->spp = std::max(1, b->spp[j] / scale);
    b->chunks = (int)b->waveChunks.size();
    PT_HIP_TRY(hipGetLastError());
}

void enqueueBatch(pt_batch *b, int scale) {
    if (b->wavefront) {
        enqueueWave(b, scale);
        return;
    }
    unsigned long long *st = b->stamps.as<unsigned long long>();
    k_stamp<<<1, 64, 0, b->stream>>>(st);
    b->chunks = 0;
    for (size_t k = 0; k < b->order.size(); k++) {
        pt_renderer *r = b->jobs[b->order[k]];
        const int spp = std::max(1, b->spp[b->order[k]] / scale);
        zeroAccum(r);
        k_zero_u64<<<1, 512, 0, b->stream>>>(r->stats.as<unsigned long long>(), S_COUNT);
        std::vector<ChunkParams> plan;
        planChunks(r, spp, plan);
        for (const ChunkParams &cp : plan) runChunk(r, cp);
        b->chunks += (int)plan.size();
        r->spp = spp;
        k_stamp<<<1, 64, 0, b->stream>>>(st + 1 + k);
        PT_HIP_TRY(hipGetLastError());
    }
}

// Runs the batch once (scale as in enqueueBatch); returns wall milliseconds of the launch.
double runBatch(pt_batch *b, int scale) {
    PT_HIP_TRY(hipStreamSynchronize(b->stream));
    double ms = 0.0;
    if (b->useGraph) {
        if (!b->exec || b->graphScale != scale || !sameCfg(b->graphCfg, b->cfg) || b->graphJobs != b->jobs.size()) {
            b->dropGraph();
            prepareBatch(b, scale);
            PT_HIP_TRY(hipStreamSynchronize(b->stream));
            PT_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueueBatch(b, scale);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(b->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            PT_HIP_TRY(hipStreamEndCapture(b->stream, &b->graph));
            PT_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            size_t n = 0;
            PT_HIP_TRY(hipGraphGetNodes(b->graph, nullptr, &n));
            b->graphNodes = n;
            b->graphScale = scale;
            b->graphCfg = b->cfg;
            b->graphJobs = b->jobs.size();
        }
        const auto t0 = std::chrono::steady_clock::now();
        PT_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
        PT_HIP_TRY(hipStreamSynchronize(b->stream));
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    } else {
        prepareBatch(b, scale);
        PT_HIP_TRY(hipStreamSynchronize(b->stream));
        const auto t0 = std::chrono::steady_clock::now();
        enqueueBatch(b, scale);
        PT_HIP_TRY(hipStreamSynchronize(b->stream));
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        b->graphNodes = 0;
    }
    std::vector<unsigned long long> st(b->jobs.size() + 1);
    PT_HIP_TRY(hipMemcpy(st.data(), b->stamps.p, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    b->latencyMs.assign(b->jobs.size(), 0.0);
    for (size_t k = 0; k < b->order.size(); k++) b->latencyMs[b->order[k]] = (double)(st[1 + k] - st[0]) / b->clockKHz;
    b->wallMs = ms;
    return ms;
}

int64_t jobRays(pt_renderer *r);

// Rays traced by render j of the batch in its last run: from the render's own statistics, or
// in the wavefront its paths' spawned rays plus its camera rays.
int64_t renderRays(pt_batch *b, int j) {
    if (!b->wavefront) return jobRays(b->jobs[j]);
    unsigned long long n = 0;
    PT_HIP_TRY(hipMemcpy(&n, b->renderRays.as<unsigned long long>() + j, sizeof n, hipMemcpyDeviceToHost));
    // spawned rays and camera rays of the chunks shared with other renders, and the rays of
    // the chunks that ran the render's own kernels
    if (j >= (int)b->waveOwn.size()) return (int64_t)n;   // not planned yet
    return (int64_t)n + b->waveCamRays[j] + (b->waveOwn[j] ? jobRays(b->jobs[j]) : 0);
}

int64_t batchRays(pt_batch *b) {
    int64_t n = 0;
    if (b->wavefront) {
        std::vector<unsigned long long> s(S_COUNT);
        PT_HIP_TRY(hipMemcpy(s.data(), b->waveStats.p, S_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (int k = 0; k < kMaxBounces; k++) n += (int64_t)(s[S_RAYS + k] + s[S_SHADOW + k]);
        for (size_t j = 0; j < b->waveOwn.size(); j++)
            if (b->waveOwn[j]) n += jobRays(b->jobs[j]);
    } else {
        for (pt_renderer *r : b->jobs) n += jobRays(r);
    }
    return n;
}

int64_t jobRays(pt_renderer *r) {
    std::vector<unsigned long long> s(S_COUNT);
    PT_HIP_TRY(hipMemcpy(s.data(), r->stats.p, S_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    int64_t n = 0;
    for (int k = 0; k < kMaxBounces; k++) n += (int64_t)(s[S_RAYS + k] + s[S_SHADOW + k]);
    return n;
}

}  // namespace

extern "C" {

int pt_batch_create(pt_batch **out) {
    if (!out) return fail(PT_ERR_INVALID, "pt_batch_create: out is NULL");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PT_ERR_DEVICE, "pt_batch_create: no HIP device");
    pt_batch *b = new pt_batch();
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->pp->counters.alloc(kNumCounters * sizeof(int));
    b->pp->overflow.alloc(sizeof(int));
    PT_HIP_TRY(hipMemsetAsync(b->pp->overflow.p, 0, sizeof(int), b->stream));
    int dev = 0, khz = 0;
    PT_HIP_TRY(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0) b->clockKHz = khz;
    PT_HIP_TRY(hipStreamSynchronize(b->stream));
    *out = b;
    return PT_OK;
    }
    catch (const HipFailure &e) {
        delete b;
        return fail(PT_ERR_DEVICE, "pt_batch_create: " + e.msg);
    }
}

int pt_batch_destroy(pt_batch *b) {
    if (!b) return PT_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->dropGraph();
    for (pt_renderer *r : b->jobs) pt_renderer_destroy(r);
    hipStream_t s = b->stream;
    delete b;
    if (s) (void)hipStreamDestroy(s);
    return PT_OK;
}

int pt_batch_add(pt_batch *b, const pt_scene *scene, int width, int height, int spp, int max_depth, int *index) {
    if (!b || !scene) return fail(PT_ERR_INVALID, "pt_batch_add: bad arguments");
    if (spp <= 0 || max_depth < 0 || max_depth > kMaxBounces) return fail(PT_ERR_INVALID, "pt_batch_add: spp must be positive and depth in [0, 64]");
    pt_renderer *r = nullptr;
    int rc = pt_renderer_create(width, height, &r);
    if (rc != PT_OK) return rc;
    rc = pt_renderer_set_scene(r, scene);
    if (rc != PT_OK) {
        pt_renderer_destroy(r);
        return rc;
    }
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(r->stream));
    PT_HIP_TRY(hipStreamDestroy(r->stream));
    r->stream = b->stream;
    r->ownsStream = false;
    r->pp = b->pp;
    r->maxDepth = max_depth;
    r->nee = 1;
    r->type = PT_RENDERER_OPTIX;
    b->jobs.push_back(r);
    b->spp.push_back(spp);
    b->dropGraph();
    if (index) *index = (int)b->jobs.size() - 1;
    return PT_OK;
    PT_API_END("pt_batch_add")
}

int pt_batch_set_option(pt_batch *b, int option, int64_t v) {
    if (!b) return fail(PT_ERR_INVALID, "pt_batch_set_option: batch is NULL");
    switch (option) {
    case PT_OPT_POOL_PATHS:
        if (v > (1ll << 27)) return fail(PT_ERR_INVALID, "pt_batch_set_option: pool larger than 2^27 paths");
        b->cfg.pool = v > 0 ? v : kDefaultPool;
        break;
    case PT_OPT_TRACE_BLOCKS_PER_CU: b->cfg.trace = v > 0 ? (int)std::min<int64_t>(v, 64) : kDefaultTraceBlocks; break;
    case PT_OPT_SHADE_BLOCKS_PER_CU: b->cfg.shade = v > 0 ? (int)std::min<int64_t>(v, 64) : 8; break;
    case PT_OPT_TRACE_LDS_STACK:
        if (v != 0 && v != 16 && v != 32) return fail(PT_ERR_INVALID, "pt_batch_set_option: LDS stack entries must be 16 or 32");
        b->cfg.lds = v == 0 ? 16 : (int)v;
        break;
    case PT_OPT_BATCH_GRAPH: b->useGraph = v != 0; break;
    case PT_OPT_BATCH_WAVEFRONT:
        b->wavefront = v != 0;
        b->dropGraph();
        break;
    default: return fail(PT_ERR_INVALID, "pt_batch_set_option: unknown option");
    }
    return PT_OK;
}

int pt_batch_get_option(pt_batch *b, int option, int64_t *v) {
    if (!b || !v) return fail(PT_ERR_INVALID, "pt_batch_get_option: bad arguments");
    switch (option) {
    case PT_OPT_POOL_PATHS: *v = b->cfg.pool; break;
    case PT_OPT_TRACE_BLOCKS_PER_CU: *v = b->cfg.trace; break;
    case PT_OPT_SHADE_BLOCKS_PER_CU: *v = b->cfg.shade; break;
    case PT_OPT_TRACE_LDS_STACK: *v = b->cfg.lds; break;
    case PT_OPT_BATCH_GRAPH: *v = b->useGraph ? 1 : 0; break;
    case PT_OPT_BATCH_WAVEFRONT: *v = b->wavefront ? 1 : 0; break;
    default: return fail(PT_ERR_INVALID, "pt_batch_get_option: unknown option");
    }
    return PT_OK;
}

int pt_batch_run(pt_batch *b) {
    if (!b) return fail(PT_ERR_INVALID, "pt_batch_run: batch is NULL");
    if (b->jobs.empty()) return fail(PT_ERR_STATE, "pt_batch_run: no renders");
    PT_API_BEGIN
    runBatch(b, 1);
    return PT_OK;
    PT_API_END("pt_batch_run")
}

int pt_batch_autotune(pt_batch *b, int probe_divisor, pt_tune_result *out) {
    if (!b || probe_divisor < 1) return fail(PT_ERR_INVALID, "pt_batch_autotune: bad arguments");
    if (b->jobs.empty()) return fail(PT_ERR_STATE, "pt_batch_autotune: no renders");
    PT_API_BEGIN
    int64_t maxPaths = 0, totalPaths = 0;
    for (size_t j = 0; j < b->jobs.size(); j++) {
        const int64_t paths = (int64_t)b->jobs[j]->W * b->jobs[j]->H * std::max(1, b->spp[j] / probe_divisor);
        maxPaths = std::max<int64_t>(maxPaths, paths);
        totalPaths += paths;
    }
    // the wavefront's chunks hold rows of several renders
    if (b->wavefront) maxPaths = totalPaths;
    auto probe = [&](const TuneCfg &c) {
        b->cfg = c;
        runBatch(b, probe_divisor);  // capture (or first enqueue) and warm-up
        const double ms = std::min(runBatch(b, probe_divisor), runBatch(b, probe_divisor));  // the faster of two
        return (double)batchRays(b) / (ms * 1e3);
    };
    int probes = 0;
    double rate = 0.0;
    b->cfg = coordinateDescent(b->cfg, poolCandidates(maxPaths), probe, &probes, &rate);
    b->dropGraph();
    if (out) {
        out->pool_paths = b->cfg.pool;
        out->trace_blocks_per_cu = b->cfg.trace;
        out->shade_blocks_per_cu = b->cfg.shade;
        out->lds_stack = b->cfg.lds;
        out->mrays_per_s = rate;
        out->probes = probes;
    }
    return PT_OK;
    PT_API_END("pt_batch_autotune")
}

int pt_batch_get_radiance(pt_batch *b, int index, float *out) {
    if (!b || !out || index < 0 || index >= (int)b->jobs.size()) return fail(PT_ERR_INVALID, "pt_batch_get_radiance: bad arguments");
    PT_API_BEGIN
    pt_renderer *r = b->jobs[index];
    PT_HIP_TRY(hipStreamSynchronize(b->stream));
    const size_t n = (size_t)r->W * r->H * 3;
    std::vector<double> acc(n);
    PT_HIP_TRY(hipMemcpy(acc.data(), r->accum.p, n * sizeof(double), hipMemcpyDeviceToHost));
    const double inv = r->spp > 0 ? 1.0 / (double)r->spp : 0.0;
    for (size_t i = 0; i < n; i++) out[i] = (float)(acc[i] * inv);
    return PT_OK;
    PT_API_END("pt_batch_get_radiance")
}

int pt_batch_get_render_stats(pt_batch *b, int index, pt_batch_render_stats *out) {
    if (!b || !out || index < 0 || index >= (int)b->jobs.size()) return fail(PT_ERR_INVALID, "pt_batch_get_render_stats: bad arguments");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(b->stream));
    pt_renderer *r = b->jobs[index];
    memset(out, 0, sizeof *out);
    out->width = r->W;
    out->height = r->H;
    out->spp = (int)r->spp;
    out->triangles = (int64_t)r->triMat.size();
    out->rays = renderRays(b, index);
    out->latency_ms = index < (int)b->latencyMs.size() ? b->latencyMs[index] : 0.0;
    return PT_OK;
    PT_API_END("pt_batch_get_render_stats")
}

int pt_batch_get_stats(pt_batch *b, pt_batch_stats *out) {
    if (!b || !out) return fail(PT_ERR_INVALID, "pt_batch_get_stats: bad arguments");
    PT_API_BEGIN
    PT_HIP_TRY(hipStreamSynchronize(b->stream));
    memset(out, 0, sizeof *out);
    out->renders = (int)b->jobs.size();
    out->chunks = b->chunks;
    out->graph_nodes = (int64_t)b->graphNodes;
    out->wall_ms = b->wallMs;
    out->rays = batchRays(b);
    int ov = 0;
    PT_HIP_TRY(hipMemcpy(&ov, b->pp->overflow.p, sizeof(int), hipMemcpyDeviceToHost));
    out->stack_overflows = ov;
    return PT_OK;
    PT_API_END("pt_batch_get_stats")
}

int pt_renderer_autotune(pt_renderer *r, int spp, pt_tune_result *out) {
    if (guard("pt_renderer_autotune", r) || spp <= 0) return fail(PT_ERR_INVALID, "pt_renderer_autotune: bad arguments");
    if (!r->hasScene) return fail(PT_ERR_STATE, "pt_renderer_autotune: no scene");
    if (r->inFlight) return fail(PT_ERR_STATE, "pt_renderer_autotune: a pass is still running");
    PT_API_BEGIN
    const bool prof = r->profile;
    r->profile = false;
    auto probe = [&](const TuneCfg &c) {
        applyCfg(r, c);
        if (r->pp->pool > r->poolWanted) r->pp->pool = 0;
        int rc = pt_renderer_render_async(r, spp);  // allocation and warm-up
        if (rc == PT_OK) rc = pt_renderer_synchronize(r);
        if (rc != PT_OK) throw HipFailure(pt_last_error());
        const int64_t before = jobRays(r);
        const auto t0 = std::chrono::steady_clock::now();
        rc = pt_renderer_render_async(r, spp);
        if (rc == PT_OK) rc = pt_renderer_synchronize(r);
        if (rc != PT_OK) throw HipFailure(pt_last_error());
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return (double)(jobRays(r) - before) / (ms * 1e3);
    };
    TuneCfg start{r->poolWanted, r->traceBlocksPerCU, r->shadeBlocksPerCU, r->ldsStack};
    int probes = 0;
    double rate = 0.0;
    const TuneCfg best = coordinateDescent(start, poolCandidates((int64_t)r->W * r->H * spp), probe, &probes, &rate);
    applyCfg(r, best);
    r->profile = prof;
    r->resetPending = true;  // the probe samples are not part of the image
    if (out) {
        out->pool_paths = best.pool;
        out->trace_blocks_per_cu = best.trace;
        out->shade_blocks_per_cu = best.shade;
        out->lds_stack = best.lds;
        out->mrays_per_s = rate;
        out->probes = probes;
    }
    return PT_OK;
    PT_API_END("pt_renderer_autotune")
}

}  // extern "C"
