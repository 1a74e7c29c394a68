// pt_bvh.hip -- GPU LBVH builder for gfx950.
//
// 1. 63-bit Morton codes of primitive centroids (21 bits per axis) + primitive AABBs
// 2. stable LSD radix sort of (code, index), 8 passes of 8 bits; ranking inside a tile with
//    wave64 ballots (one 64-bit mask per key bit)
// 3. Karras 2012 binary radix tree (one thread per internal node; equal codes ordered by
//    index)
// 4. bottom-up AABB refit (atomic arrival counter per node; the second child to arrive
//    merges) and depth
// 5. collapse into a 4-wide BVH of 128-byte nodes (on the host, over the refitted tree):
//    subtrees of <= PT_LEAF_MAX primitives become leaf ranges of the sorted primitive array,
//    and each 4-wide node opens the largest-area internal children of its binary node until
//    it has four, about halving the number of dependent node fetches per ray
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "pt_gpu.h"

namespace pt {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1fffffULL;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ Box primBox(const float4 *prims, int i) {
    const float4 a = prims[3 * i], b = prims[3 * i + 1], c = prims[3 * i + 2];
    Box bx;
    if (asInt(b.w) == PT_TAG_SPHERE) {
        const float r = b.x;
        bx.lo[0] = a.x - r; bx.lo[1] = a.y - r; bx.lo[2] = a.z - r;
        bx.hi[0] = a.x + r; bx.hi[1] = a.y + r; bx.hi[2] = a.z + r;
    } else {
        bx.lo[0] = fminf(a.x, fminf(b.x, c.x)); bx.hi[0] = fmaxf(a.x, fmaxf(b.x, c.x));
        bx.lo[1] = fminf(a.y, fminf(b.y, c.y)); bx.hi[1] = fmaxf(a.y, fmaxf(b.y, c.y));
        bx.lo[2] = fminf(a.z, fminf(b.z, c.z)); bx.hi[2] = fmaxf(a.z, fmaxf(b.z, c.z));
    }
    // widen by a few ulps so rounding in the slab test cannot cull a watertight hit
    for (int k = 0; k < 3; k++) {
        const float e = fmaxf(fabsf(bx.lo[k]), fabsf(bx.hi[k])) * 2.0e-6f + 1.0e-30f;
        bx.lo[k] -= e;
        bx.hi[k] += e;
    }
    return bx;
}

__global__ void k_morton(const float4 *prims, int n, float3 lo, float3 ext, uint64_t *keys, uint32_t *vals, float *boxes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Box b = primBox(prims, i);
    float c[3];
    const float e[3] = {ext.x, ext.y, ext.z}, l[3] = {lo.x, lo.y, lo.z};
    uint64_t key = 0;
    for (int k = 0; k < 3; k++) {
        c[k] = 0.5f * (b.lo[k] + b.hi[k]);
        float u = e[k] > 0.f ? (c[k] - l[k]) / e[k] : 0.5f;
        u = fminf(fmaxf(u, 0.f), 1.f);
        const uint32_t q = min((uint32_t)(u * 2097152.0f), 2097151u);
        key |= spread21(q) << (2 - k);
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
    for (int k = 0; k < 3; k++) {
        boxes[6 * i + k] = b.lo[k];
        boxes[6 * i + 3 + k] = b.hi[k];
    }
}

// --- radix sort -----------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint64_t *keys, int n, int shift, int numTiles, uint32_t *hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; it++) {
        const int idx = base + it * kSortThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * numTiles + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of m counters with one block of 1024 threads
__global__ __launch_bounds__(1024) void k_sort_scan(uint32_t *data, int m) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = (m + 1023) / 1024;
    const int b = t * per, e = min(b + per, m);
    uint32_t s = 0;
    for (int i = b; i < e; i++) s += data[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t > 0 ? part[t - 1] : 0u;
    for (int i = b; i < e; i++) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint64_t *keysIn, const uint32_t *valsIn, uint64_t *keysOut, uint32_t *valsOut,
                                                               const uint32_t *offsets, int n, int shift, int numTiles) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t base[256];
    __shared__ uint32_t waveCnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = offsets[tid * numTiles + blockIdx.x];
    running[tid] = 0;
    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int it = 0; it < kSortItems; it++) {
        for (int w = 0; w < 4; w++) waveCnt[w][tid] = 0;
        __syncthreads();
        const int idx = blockIdx.x * kSortTile + it * kSortThreads + tid;
        const bool valid = idx < n;
        const uint64_t key = valid ? keysIn[idx] : 0ull;
        const uint32_t val = valid ? valsIn[idx] : 0u;
        const uint32_t digit = (uint32_t)(key >> shift) & 255u;
        uint64_t m = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(m & ltMask);
        if (valid && rank == 0) waveCnt[wave][digit] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pre = running[digit];
            for (int w = 0; w < wave; w++) pre += waveCnt[w][digit];
            const uint32_t pos = base[digit] + pre + rank;
            keysOut[pos] = key;
            valsOut[pos] = val;
        }
        __syncthreads();
        running[tid] += waveCnt[0][tid] + waveCnt[1][tid] + waveCnt[2][tid] + waveCnt[3][tid];
        __syncthreads();
    }
}

__global__ void k_gather(const float4 *prims, const float *boxes, const uint32_t *order, int n, float4 *primsOut, float *boxesOut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = order[i];
    primsOut[3 * i] = prims[3 * s];
    primsOut[3 * i + 1] = prims[3 * s + 1];
    primsOut[3 * i + 2] = prims[3 * s + 2];
    for (int k = 0; k < 6; k++) boxesOut[6 * i + k] = boxes[6 * s + k];
}

// --- Karras 2012 hierarchy --------------------------------------------------------------
__device__ __forceinline__ int delta(const uint64_t *keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// child encoding during the build: >= 0 internal node, < 0 leaf ~primitive
__global__ void k_karras(const uint64_t *keys, int n, int2 *children, int2 *ranges, int *parentNode, int *parentLeaf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    int left, right;
    if (first == gamma) {
        left = ~gamma;
        parentLeaf[gamma] = i;
    } else {
        left = gamma;
        parentNode[gamma] = i;
    }
    if (last == gamma + 1) {
        right = ~(gamma + 1);
        parentLeaf[gamma + 1] = i;
    } else {
        right = gamma + 1;
        parentNode[gamma + 1] = i;
    }
    children[i] = make_int2(left, right);
    ranges[i] = make_int2(first, last);
}

__global__ void k_refit(const int2 *children, const int *parentNode, const int *parentLeaf, const float *leafBoxes, float *nodeBoxes,
                        int *arrivals, int n) {
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= n) return;
    int p = parentLeaf[leaf];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&arrivals[p], 1) == 0) break;  // the sibling subtree is not done yet
        __threadfence();
        const int2 c = children[p];
        // node boxes written by other workgroups are read with volatile (L1-bypassing) loads
        float x[6], y[6];
        if (c.x >= 0) {
            const volatile float *src = &nodeBoxes[6 * c.x];
            for (int k = 0; k < 6; k++) x[k] = src[k];
        } else {
            for (int k = 0; k < 6; k++) x[k] = leafBoxes[6 * (~c.x) + k];
        }
        if (c.y >= 0) {
            const volatile float *src = &nodeBoxes[6 * c.y];
            for (int k = 0; k < 6; k++) y[k] = src[k];
        } else {
            for (int k = 0; k < 6; k++) y[k] = leafBoxes[6 * (~c.y) + k];
        }
        volatile float *dst = &nodeBoxes[6 * p];
        for (int k = 0; k < 3; k++) {
            dst[k] = fminf(x[k], y[k]);
            dst[3 + k] = fmaxf(x[3 + k], y[3 + k]);
        }
        p = parentNode[p];
    }
}

// depth of each leaf = number of internal nodes on its root path
__global__ void k_depth(const int *parentNode, const int *parentLeaf, int n, int *maxDepth) {
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= n) return;
    int p = parentLeaf[leaf];
    int d = 0;
    while (p >= 0 && d < n) {
        d++;
        p = parentNode[p];
    }
    atomicMax(maxDepth, d);
}

// All primitives in one leaf below a single 4-wide root (n <= PT_LEAF_MAX).
__global__ void k_write_single(const float *leafBoxes, int n, float4 *nodes) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], leafBoxes[6 * i + k]);
            hi[k] = fmaxf(hi[k], leafBoxes[6 * i + 3 + k]);
        }
    for (int k = 0; k < 3; k++) {
        nodes[2 * k] = make_float4(lo[k], 1e30f, 1e30f, 1e30f);
        nodes[2 * k + 1] = make_float4(hi[k], -1e30f, -1e30f, -1e30f);
    }
    nodes[6] = make_float4(asFloat(~(((n - 1) << PT_LEAF_COUNT_SHIFT) | 0)), asFloat(PT_EMPTY_CHILD), asFloat(PT_EMPTY_CHILD),
                           asFloat(PT_EMPTY_CHILD));
    nodes[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_fill_int(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline int blocks(int64_t n, int t) { return (int)((n + t - 1) / t); }

// Collapses the binary tree into 4-wide nodes on the host.  Subtrees of <= PT_LEAF_MAX
// primitives become leaf ranges; which internal nodes become 4-wide nodes is chosen to
// minimise a surface-area cost (a 4-wide node costs kNodeCost x its area, a leaf kPrimCost x
// its area x its primitives): bottom up, every internal node gets the cheapest way to spend
// 1..4 child slots of its parent on its subtree (one slot: the node itself as a 4-wide node;
// more: its two children share them), then the nodes are written top down from the root.
// The root is node 0, and the live children of a node take consecutive indices (depth-first
// order).  Returns the node count; *depth receives the deepest 4-wide level.
constexpr float kNodeCost = 1.0f;
constexpr float kPrimCost = 0.6f;
int collapseWide(const std::vector<int2> &children, const std::vector<int2> &ranges, const std::vector<float> &nodeBoxes,
                 const std::vector<float> &leafBoxes, std::vector<float4> &out, int *depth) {
    auto live = [&](int c) { return c >= 0 && ranges[c].y - ranges[c].x + 1 > PT_LEAF_MAX; };
    auto box = [&](int c) { return c >= 0 ? &nodeBoxes[6 * (size_t)c] : &leafBoxes[6 * (size_t)(~c)]; };
    auto area = [&](int c) {
        const float *b = box(c);
        const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        return dx * dy + dy * dz + dz * dx;
    };
    auto leafRef = [&](int c) {
        if (c < 0) return ~(~c);
        const int2 r = ranges[c];
        return ~(((r.y - r.x) << PT_LEAF_COUNT_SHIFT) | r.x);
    };
    // bottom-up costs over the live internal nodes (children before parents)
    const size_t ni = children.size();
    std::vector<float> W(ni), S2(ni), S3(ni);   // as a 4-wide node; spread over 2 / 3 slots
    std::vector<uint8_t> split(3 * ni);         // slots of the left child when the children share 2..4
    std::vector<uint8_t> opened(ni);            // bit k: spreading over k + 2 slots opens the node
    auto cost = [&](int x, int slots) -> float {
        if (!live(x)) {
            const int cnt = x < 0 ? 1 : ranges[x].y - ranges[x].x + 1;
            return kPrimCost * area(x) * (float)cnt;
        }
        return slots == 1 ? W[x] : slots == 2 ? S2[x] : S3[x];
    };
    std::vector<int> order, todo;
    todo.push_back(0);
    while (!todo.empty()) {
        const int x = todo.back();
        todo.pop_back();
        order.push_back(x);
        if (live(children[x].x)) todo.push_back(children[x].x);
        if (live(children[x].y)) todo.push_back(children[x].y);
    }
    for (size_t k = order.size(); k-- > 0;) {
        const int x = order[k];
        const int l = children[x].x, r = children[x].y;
        float share[5] = {0.f, 0.f, 0.f, 0.f, 0.f};   // both children in j slots
        for (int j = 2; j <= 4; j++) {
            float best = INFINITY;
            int bi = 1;
            for (int i = 1; i < j; i++) {
                const float c = cost(l, std::min(i, 3)) + cost(r, std::min(j - i, 3));
                if (c < best) {
                    best = c;
                    bi = i;
                }
            }
            share[j] = best;
            split[3 * x + (j - 2)] = (uint8_t)bi;
        }
        W[x] = kNodeCost * area(x) + share[4];
        S2[x] = std::min(W[x], share[2]);
        S3[x] = std::min(W[x], share[3]);
        opened[x] = (uint8_t)((share[2] < W[x] ? 1 : 0) | (share[3] < W[x] ? 2 : 0));
    }
    // the children of a 4-wide node: its binary node's two children sharing 4 slots, opened
    // where the cost tables say so
    auto take = [&](auto &&self, int x, int slots, int *kids, int &nk) -> void {
        if (!live(x) || slots == 1 || !(opened[x] & (1 << (slots - 2)))) {
            kids[nk++] = x;
            return;
        }
        const int i = split[3 * x + (slots - 2)];
        self(self, children[x].x, i, kids, nk);
        self(self, children[x].y, slots - i, kids, nk);
    };
    out.clear();
    out.resize(PT_NODE_F4);
    int count = 1;
    *depth = 1;
    struct Item {
        int bin, idx, level;
    };
    std::vector<Item> stack;
    stack.push_back({0, 0, 1});
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        int kids[4] = {0, 0, 0, 0};
        int nk = 0;
        const int i4 = split[3 * it.bin + 2];
        take(take, children[it.bin].x, i4, kids, nk);
        take(take, children[it.bin].y, 4 - i4, kids, nk);
        float lo[3][4], hi[3][4];
        int ref[4];
        for (int k = 0; k < 4; k++) {
            if (k < nk) {
                const float *b = box(kids[k]);
                for (int a = 0; a < 3; a++) {
                    lo[a][k] = b[a];
                    hi[a][k] = b[3 + a];
                }
                if (live(kids[k])) {
                    ref[k] = count++;
                    out.resize((size_t)PT_NODE_F4 * count);
                    stack.push_back({kids[k], ref[k], it.level + 1});
                    *depth = std::max(*depth, it.level + 1);
                } else {
                    ref[k] = leafRef(kids[k]);
                }
            } else {
                // an inverted box: every slab test misses it
                for (int a = 0; a < 3; a++) {
                    lo[a][k] = 1e30f;
                    hi[a][k] = -1e30f;
                }
                ref[k] = PT_EMPTY_CHILD;
            }
        }
        float4 *o = &out[(size_t)PT_NODE_F4 * it.idx];
        for (int a = 0; a < 3; a++) {
            o[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
            o[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
        }
        o[6] = make_float4(asFloat(ref[0]), asFloat(ref[1]), asFloat(ref[2]), asFloat(ref[3]));
        o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    return count;
}

}  // namespace

int deviceCUs() {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    return prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
}

void buildBvh(const float4 *primsIn, int n, const float lo[3], const float hi[3], DevBuf &nodes, float4 *primsSorted, hipStream_t stream,
              BvhInfo *info) {
    info->numNodes = 0;
    info->depth = 0;
    info->binaryDepth = 0;
    if (n <= 0) return;
    DevBuf keys, keys2, vals, vals2, boxes, boxes2, hist;
    keys.alloc(sizeof(uint64_t) * n);
    keys2.alloc(sizeof(uint64_t) * n);
    vals.alloc(sizeof(uint32_t) * n);
    vals2.alloc(sizeof(uint32_t) * n);
    boxes.alloc(sizeof(float) * 6 * n);
    boxes2.alloc(sizeof(float) * 6 * n);
    const float3 flo = make_float3(lo[0], lo[1], lo[2]);
    const float3 ext = make_float3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
    k_morton<<<blocks(n, 256), 256, 0, stream>>>(primsIn, n, flo, ext, keys.as<uint64_t>(), vals.as<uint32_t>(), boxes.as<float>());
    PT_HIP_TRY(hipGetLastError());
    const int numTiles = blocks(n, kSortTile);
    hist.alloc(sizeof(uint32_t) * 256 * numTiles);
    uint64_t *kin = keys.as<uint64_t>(), *kout = keys2.as<uint64_t>();
    uint32_t *vin = vals.as<uint32_t>(), *vout = vals2.as<uint32_t>();
    for (int pass = 0; pass < 8; pass++) {
        const int shift = pass * 8;
        k_sort_hist<<<numTiles, kSortThreads, 0, stream>>>(kin, n, shift, numTiles, hist.as<uint32_t>());
        k_sort_scan<<<1, 1024, 0, stream>>>(hist.as<uint32_t>(), 256 * numTiles);
        k_sort_scatter<<<numTiles, kSortThreads, 0, stream>>>(kin, vin, kout, vout, hist.as<uint32_t>(), n, shift, numTiles);
        PT_HIP_TRY(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // after 8 passes the sorted data is back in keys/vals
    k_gather<<<blocks(n, 256), 256, 0, stream>>>(primsIn, boxes.as<float>(), vin, n, primsSorted, boxes2.as<float>());
    PT_HIP_TRY(hipGetLastError());
    if (n <= PT_LEAF_MAX) {
        nodes.alloc(sizeof(float4) * PT_NODE_F4);
        k_write_single<<<1, 64, 0, stream>>>(boxes2.as<float>(), n, nodes.as<float4>());
        PT_HIP_TRY(hipGetLastError());
        PT_HIP_TRY(hipStreamSynchronize(stream));
        info->numNodes = 1;
        info->depth = 1;
        info->binaryDepth = n > 1 ? 1 : 0;
        return;
    }
    const int ni = n - 1;
    DevBuf children, ranges, parentNode, parentLeaf, nodeBoxes, arrivals, counters;
    children.alloc(sizeof(int2) * ni);
    ranges.alloc(sizeof(int2) * ni);
    parentNode.alloc(sizeof(int) * ni);
    parentLeaf.alloc(sizeof(int) * n);
    nodeBoxes.alloc(sizeof(float) * 6 * ni);
    arrivals.alloc(sizeof(int) * ni);
    counters.alloc(sizeof(int) * 2);
    k_fill_int<<<blocks(ni, 256), 256, 0, stream>>>(parentNode.as<int>(), ni, -1);
    k_fill_int<<<blocks(ni, 256), 256, 0, stream>>>(arrivals.as<int>(), ni, 0);
    k_fill_int<<<1, 64, 0, stream>>>(counters.as<int>(), 2, 0);
    k_karras<<<blocks(ni, 256), 256, 0, stream>>>(kin, n, children.as<int2>(), ranges.as<int2>(), parentNode.as<int>(), parentLeaf.as<int>());
    k_refit<<<blocks(n, 256), 256, 0, stream>>>(children.as<int2>(), parentNode.as<int>(), parentLeaf.as<int>(), boxes2.as<float>(),
                                                 nodeBoxes.as<float>(), arrivals.as<int>(), n);
    k_depth<<<blocks(n, 256), 256, 0, stream>>>(parentNode.as<int>(), parentLeaf.as<int>(), n, counters.as<int>());
    PT_HIP_TRY(hipGetLastError());
    // the collapse runs on the host over the refitted binary tree
    std::vector<int2> hChildren(ni), hRanges(ni);
    std::vector<float> hNodeBoxes((size_t)6 * ni), hLeafBoxes((size_t)6 * n);
    int hc[2] = {0, 0};
    PT_HIP_TRY(hipMemcpyAsync(hChildren.data(), children.p, sizeof(int2) * ni, hipMemcpyDeviceToHost, stream));
    PT_HIP_TRY(hipMemcpyAsync(hRanges.data(), ranges.p, sizeof(int2) * ni, hipMemcpyDeviceToHost, stream));
    PT_HIP_TRY(hipMemcpyAsync(hNodeBoxes.data(), nodeBoxes.p, sizeof(float) * 6 * ni, hipMemcpyDeviceToHost, stream));
    PT_HIP_TRY(hipMemcpyAsync(hLeafBoxes.data(), boxes2.p, sizeof(float) * 6 * n, hipMemcpyDeviceToHost, stream));
    PT_HIP_TRY(hipMemcpyAsync(hc, counters.p, sizeof hc, hipMemcpyDeviceToHost, stream));
    PT_HIP_TRY(hipStreamSynchronize(stream));
    std::vector<float4> hNodes;
    int wideDepth = 1;
    const int count = collapseWide(hChildren, hRanges, hNodeBoxes, hLeafBoxes, hNodes, &wideDepth);
    nodes.alloc(sizeof(float4) * PT_NODE_F4 * (size_t)count);
    PT_HIP_TRY(hipMemcpyAsync(nodes.p, hNodes.data(), sizeof(float4) * hNodes.size(), hipMemcpyHostToDevice, stream));
    PT_HIP_TRY(hipStreamSynchronize(stream));
    info->numNodes = count;
    info->depth = wideDepth;
    info->binaryDepth = hc[0];
}

}  // namespace pt
