// pt_bvh.hip -- GPU BVH builder for gfx950.
//
// 1. 63-bit Morton codes of primitive centroids (21 bits per axis) + primitive AABBs
// 2. stable LSD radix sort of (code, index), 8 passes of 8 bits; ranking inside a tile with
//    wave64 ballots (one 64-bit mask per key bit)
// 3. a binary tree over the sorted primitives, one of
//    - LBVH: Karras 2012 binary radix tree (one thread per internal node; equal codes ordered
//      by index), then a bottom-up AABB refit (atomic arrival counter per node; the second
//      child to arrive merges);
//    - PLOC (Meister & Bittner 2018, parallel locally-ordered clustering): every cluster of the
//      Morton-ordered list finds the neighbour within PLOC_RADIUS positions whose merged box
//      has the least surface area; mutual nearest neighbours merge; the list is compacted with
//      a prefix sum; repeat until one cluster is left.  The primitives are then put in the
//      tree's depth-first leaf order so every subtree covers a contiguous range.
// 4. collapse into a 4-wide BVH of 128-byte nodes, on the GPU: subtrees of <= PT_LEAF_MAX
//    primitives become leaf ranges of the primitive array, and which internal nodes become
//    4-wide nodes is chosen to minimise a surface-area cost (bottom-up dynamic programme over
//    the binary tree with arrival counters; then the nodes are written level by level from
//    the root, the live children of a node at consecutive indices).
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

// --- prefix sums ------------------------------------------------------------------------
// Exclusive scan of m values in place, the total written to data[m]: tiles of kScanTile values
// scanned in LDS (one per block), the tile totals scanned by one block, then added back.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

template <class T> __device__ T blockExclusiveScan(T v, T *sh, T *total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const T x = t >= off ? sh[t - off] : (T)0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    const T incl = sh[t];
    *total = sh[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

template <class T> __global__ __launch_bounds__(kScanThreads) void k_scan_tiles(T *data, int m, T *tileSums) {
    __shared__ T sh[kScanThreads];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    T v[kScanItems];
    T sum = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < m ? data[base + k] : (T)0;
        sum += v[k];
    }
    T total;
    T run = blockExclusiveScan<T>(sum, sh, &total);
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < m) data[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tileSums[blockIdx.x] = total;
}

// one block: exclusive scan of the tile totals (any count), total to tileSums[tiles]
template <class T> __global__ __launch_bounds__(kScanThreads) void k_scan_totals(T *tileSums, int tiles) {
    __shared__ T sh[kScanThreads];
    T carry = 0;
    for (int b0 = 0; b0 < tiles; b0 += kScanThreads) {
        const int i = b0 + threadIdx.x;
        const T v = i < tiles ? tileSums[i] : (T)0;
        T total;
        const T ex = blockExclusiveScan<T>(v, sh, &total);
        if (i < tiles) tileSums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) tileSums[tiles] = carry;
}

template <class T> __global__ void k_scan_add(T *data, int m, const T *tileSums, int tiles) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) data[i] += tileSums[i / kScanTile];
    if (i == 0) data[m] = tileSums[tiles];
}

// data holds m + 1 entries; tmp at least tiles(m) + 1
template <class T> void exclusiveScan(T *data, int m, T *tmp, hipStream_t stream) {
    const int tiles = std::max(1, (m + kScanTile - 1) / kScanTile);
    k_scan_tiles<T><<<tiles, kScanThreads, 0, stream>>>(data, m, tmp);
    k_scan_totals<T><<<1, kScanThreads, 0, stream>>>(tmp, tiles);
    k_scan_add<T><<<std::max(1, (m + 255) / 256), 256, 0, stream>>>(data, m, tmp, tiles);
    PT_HIP_TRY(hipGetLastError());
}
inline int scanTmpEntries(int m) { return (m + kScanTile - 1) / kScanTile + 2; }

// --- binary tree shared by both builders --------------------------------------------------
// children: >= 0 internal node, < 0 ~leaf (leaf = position in the Morton order); cnt and
// first: primitives below an internal node and the position of the first in the final
// primitive order; leafPos: final position of a leaf (nullptr: the Morton position).
struct BinTree {
    int n = 0, root = 0;
    const int2 *children = nullptr;
    const float *nodeBoxes = nullptr, *leafBoxes = nullptr;
    const int *cnt = nullptr, *first = nullptr, *leafPos = nullptr;
    const int *parentNode = nullptr, *parentLeaf = nullptr;
};

__device__ __forceinline__ int childCount(const BinTree &t, int c) { return c < 0 ? 1 : t.cnt[c]; }

__device__ __forceinline__ float boxArea6(const float *b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
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

// LBVH: counts and first positions from the Karras ranges (leaves stay in Morton order)
__global__ void k_ranges_to_counts(const int2 *ranges, int ni, int *cnt, int *first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const int2 r = ranges[i];
    cnt[i] = r.y - r.x + 1;
    first[i] = r.x;
}

// --- PLOC --------------------------------------------------------------------------------
constexpr int PLOC_RADIUS = 16;
constexpr int kPlocThreads = 256;

__global__ void k_ploc_init(int n, const float *leafBoxes, int *clusters, float *cbox) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    clusters[i] = ~i;
    for (int k = 0; k < 6; k++) cbox[6 * i + k] = leafBoxes[6 * i + k];
}

// Nearest neighbour of every cluster among the PLOC_RADIUS positions on either side: least
// surface area of the merged box, ties to the lower position.
__global__ __launch_bounds__(kPlocThreads) void k_ploc_nn(const float *cbox, int m, int *nn) {
    __shared__ float sb[(kPlocThreads + 2 * PLOC_RADIUS) * 6];
    const int start = blockIdx.x * kPlocThreads;
    for (int k = threadIdx.x; k < (kPlocThreads + 2 * PLOC_RADIUS) * 6; k += kPlocThreads) {
        const int pos = start - PLOC_RADIUS + k / 6;
        sb[k] = pos >= 0 && pos < m ? cbox[(int64_t)pos * 6 + k % 6] : 0.f;
    }
    __syncthreads();
    const int i = start + threadIdx.x;
    if (i >= m) return;
    const float *bi = &sb[(threadIdx.x + PLOC_RADIUS) * 6];
    float best = INFINITY;
    int bj = i;
    const int j0 = max(0, i - PLOC_RADIUS), j1 = min(m - 1, i + PLOC_RADIUS);
    for (int j = j0; j <= j1; j++) {
        if (j == i) continue;
        const float *bj6 = &sb[(j - start + PLOC_RADIUS) * 6];
        float u[6];
        for (int k = 0; k < 3; k++) {
            u[k] = fminf(bi[k], bj6[k]);
            u[3 + k] = fmaxf(bi[3 + k], bj6[3 + k]);
        }
        const float d = boxArea6(u);
        if (d < best) {
            best = d;
            bj = j;
        }
    }
    nn[i] = bj;
}

// Merge pairs (2k, 2k + 1): the fallback for an iteration in which no two clusters chose each
// other (never observed; it keeps the loop finite).
__global__ void k_ploc_pairs(int m, int *nn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = i ^ 1;
    nn[i] = j < m ? j : i;
}

// keep flag (low 32 bits) and new-node flag (high 32 bits) per cluster position
__global__ void k_ploc_flags(const int *nn, int m, unsigned long long *flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = j != i && nn[j] == i;
    const unsigned long long keep = mutual && i > j ? 0ull : 1ull;
    const unsigned long long leader = mutual && i < j ? 1ull : 0ull;
    flags[i] = keep | leader << 32;
}

__global__ void k_ploc_merge(const int *clusters, const float *cbox, const int *nn, int m, const unsigned long long *pre, int nodeBase,
                             int *clustersOut, float *cboxOut, int2 *children, float *nodeBoxes, int *cnt, int *parentNode, int *parentLeaf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = j != i && nn[j] == i;
    if (mutual && i > j) return;
    const int pos = (int)(pre[i] & 0xffffffffull);
    const float *a = &cbox[6 * (size_t)i];
    if (!mutual) {
        clustersOut[pos] = clusters[i];
        for (int k = 0; k < 6; k++) cboxOut[6 * (size_t)pos + k] = a[k];
        return;
    }
    const int id = nodeBase + (int)(pre[i] >> 32);
    const int l = clusters[i], r = clusters[j];
    const float *b = &cbox[6 * (size_t)j];
    float u[6];
    for (int k = 0; k < 3; k++) {
        u[k] = fminf(a[k], b[k]);
        u[3 + k] = fmaxf(a[3 + k], b[3 + k]);
    }
    children[id] = make_int2(l, r);
    for (int k = 0; k < 6; k++) {
        nodeBoxes[6 * (size_t)id + k] = u[k];
        cboxOut[6 * (size_t)pos + k] = u[k];
    }
    cnt[id] = (l < 0 ? 1 : cnt[l]) + (r < 0 ? 1 : cnt[r]);
    if (l < 0) parentLeaf[~l] = id; else parentNode[l] = id;
    if (r < 0) parentLeaf[~r] = id; else parentNode[r] = id;
    clustersOut[pos] = id;
}

// depth-first leaf order: a leaf's position is the number of leaves in the left subtrees its
// root path passes
__global__ void k_leaf_positions(BinTree t, int *leafPos) {
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= t.n) return;
    int off = 0, cur = ~leaf, p = t.parentLeaf[leaf];
    // a root path has fewer than n nodes; the step bound only guards against a malformed tree
    for (int steps = 0; p >= 0 && steps < t.n; steps++) {
        const int2 c = t.children[p];
        if (c.y == cur) off += childCount(t, c.x);
        cur = p;
        p = t.parentNode[p];
    }
    leafPos[leaf] = off;
}

// first position of every internal node: written by its leftmost leaf
__global__ void k_node_first(BinTree t, int *first) {
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= t.n) return;
    const int o = t.leafPos[leaf];
    int cur = ~leaf, p = t.parentLeaf[leaf];
    for (int steps = 0; p >= 0 && t.children[p].x == cur && steps < t.n; steps++) {
        first[p] = o;
        cur = p;
        p = t.parentNode[p];
    }
}

__global__ void k_scatter_prims(const float4 *in, const int *leafPos, int n, float4 *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = leafPos[i];
    out[3 * (size_t)o] = in[3 * (size_t)i];
    out[3 * (size_t)o + 1] = in[3 * (size_t)i + 1];
    out[3 * (size_t)o + 2] = in[3 * (size_t)i + 2];
}

// --- 4-wide collapse -----------------------------------------------------------------------
// A 4-wide node costs kNodeCost x its area, a leaf kPrimCost x its area x its primitives.
// Bottom up, every live internal node (more than PT_LEAF_MAX primitives) gets the cheapest way
// to spend 1..3 child slots of its parent on its subtree: one slot is the node itself as a
// 4-wide node (W), two or three spread its subtree over its children (S2, S3), opening it.
constexpr float kNodeCost = 1.0f;
constexpr float kPrimCost = 0.6f;

struct CollapseTables {
    float *W, *S2, *S3;
    unsigned char *split;   // 3 per node: slots of the left child when the children share 2..4
    unsigned char *opened;  // bit k: spreading over k + 2 slots opens the node
};

__device__ __forceinline__ bool liveChild(const BinTree &t, int c) { return c >= 0 && t.cnt[c] > PT_LEAF_MAX; }
__device__ __forceinline__ const float *boxOf(const BinTree &t, int c) {
    return c >= 0 ? &t.nodeBoxes[6 * (size_t)c] : &t.leafBoxes[6 * (size_t)(~c)];
}

// cost of child c given `slots` of its parent (tables of live children are read volatile:
// other workgroups wrote them during this kernel)
__device__ float slotCost(const BinTree &t, const CollapseTables &tb, int c, int slots) {
    if (!liveChild(t, c)) return kPrimCost * boxArea6(boxOf(t, c)) * (float)childCount(t, c);
    const volatile float *src = slots == 1 ? tb.W : slots == 2 ? tb.S2 : tb.S3;
    return src[c];
}

__global__ void k_collapse_costs(BinTree t, CollapseTables tb, int *arrivals) {
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= t.n) return;
    int p = t.parentLeaf[leaf];
    for (int steps = 0; p >= 0 && steps < t.n; steps++) {
        __threadfence();
        if (atomicAdd(&arrivals[p], 1) == 0) break;  // the sibling subtree is not done yet
        __threadfence();
        if (t.cnt[p] > PT_LEAF_MAX) {
            const int2 c = t.children[p];
            float share[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
            for (int j = 2; j <= 4; j++) {
                float best = INFINITY;
                int bi = 1;
                for (int i = 1; i < j; i++) {
                    const float v = slotCost(t, tb, c.x, min(i, 3)) + slotCost(t, tb, c.y, min(j - i, 3));
                    if (v < best) {
                        best = v;
                        bi = i;
                    }
                }
                share[j] = best;
                tb.split[3 * (size_t)p + (j - 2)] = (unsigned char)bi;
            }
            const float w = kNodeCost * boxArea6(&t.nodeBoxes[6 * (size_t)p]) + share[4];
            volatile float *W = tb.W, *S2 = tb.S2, *S3 = tb.S3;
            W[p] = w;
            S2[p] = fminf(w, share[2]);
            S3[p] = fminf(w, share[3]);
            tb.opened[p] = (unsigned char)((share[2] < w ? 1 : 0) | (share[3] < w ? 2 : 0));
        }
        p = t.parentNode[p];
    }
}

// The children of the 4-wide node made from binary node x: its two children sharing four
// slots, opened where the cost tables say so, in left-to-right order.
__device__ int wideChildren(const BinTree &t, const CollapseTables &tb, int x, int *kids) {
    int stackNode[4], stackSlots[4], sp = 0, nk = 0;
    const int2 c = t.children[x];
    const int i4 = tb.split[3 * (size_t)x + 2];
    stackNode[sp] = c.y;
    stackSlots[sp++] = 4 - i4;
    stackNode[sp] = c.x;
    stackSlots[sp++] = i4;
    while (sp > 0) {
        sp--;
        const int y = stackNode[sp], s = stackSlots[sp];
        if (!liveChild(t, y) || s == 1 || !(tb.opened[y] & (1 << (s - 2)))) {
            kids[nk++] = y;
            continue;
        }
        const int i = tb.split[3 * (size_t)y + (s - 2)];
        const int2 cy = t.children[y];
        stackNode[sp] = cy.y;
        stackSlots[sp++] = s - i;
        stackNode[sp] = cy.x;
        stackSlots[sp++] = i;
    }
    return nk;
}

__global__ void k_level_count(BinTree t, CollapseTables tb, const int2 *frontier, int m, unsigned *live) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int kids[4];
    const int nk = wideChildren(t, tb, frontier[i].x, kids);
    unsigned c = 0;
    for (int k = 0; k < nk; k++) c += liveChild(t, kids[k]) ? 1u : 0u;
    live[i] = c;
}

// Writes the level's nodes; live children go to the next level at nodeBase + their rank.
__global__ void k_level_write(BinTree t, CollapseTables tb, const int2 *frontier, int m, const unsigned *rank, int nodeBase, int2 *next,
                              float4 *nodes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int2 f = frontier[i];
    int kids[4];
    const int nk = wideChildren(t, tb, f.x, kids);
    float lo[3][4], hi[3][4];
    int ref[4];
    int r = (int)rank[i];
    for (int k = 0; k < 4; k++) {
        if (k < nk) {
            const float *b = boxOf(t, kids[k]);
            for (int a = 0; a < 3; a++) {
                lo[a][k] = b[a];
                hi[a][k] = b[3 + a];
            }
            const int y = kids[k];
            if (liveChild(t, y)) {
                ref[k] = nodeBase + r;
                next[r] = make_int2(y, nodeBase + r);
                r++;
            } else {
                const int cnt = childCount(t, y);
                const int first = y < 0 ? (t.leafPos ? t.leafPos[~y] : ~y) : t.first[y];
                ref[k] = ~(((cnt - 1) << PT_LEAF_COUNT_SHIFT) | first);
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
    float4 *o = &nodes[(size_t)PT_NODE_F4 * f.y];
    for (int a = 0; a < 3; a++) {
        o[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
        o[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    o[6] = make_float4(asFloat(ref[0]), asFloat(ref[1]), asFloat(ref[2]), asFloat(ref[3]));
    o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Collapses the binary tree into 4-wide nodes (root = node 0); returns the node count and the
// number of levels.
int collapseWide(const BinTree &t, DevBuf &nodes, hipStream_t stream, int *levels) {
    const int ni = t.n - 1;
    DevBuf w, s2, s3, split, opened, arrivals, live, scanTmp, front[2], tmpNodes;
    w.alloc(sizeof(float) * ni);
    s2.alloc(sizeof(float) * ni);
    s3.alloc(sizeof(float) * ni);
    split.alloc((size_t)3 * ni);
    opened.alloc((size_t)ni);
    arrivals.alloc(sizeof(int) * ni);
    CollapseTables tb{w.as<float>(), s2.as<float>(), s3.as<float>(), split.as<unsigned char>(), opened.as<unsigned char>()};
    k_fill_int<<<blocks(ni, 256), 256, 0, stream>>>(arrivals.as<int>(), ni, 0);
    k_collapse_costs<<<blocks(t.n, 256), 256, 0, stream>>>(t, tb, arrivals.as<int>());
    PT_HIP_TRY(hipGetLastError());
    // every 4-wide node stands for a distinct live internal node: at most ni of them
    tmpNodes.alloc(sizeof(float4) * PT_NODE_F4 * (size_t)std::max(ni, 1));
    live.alloc(sizeof(unsigned) * ((size_t)ni + 1));
    scanTmp.alloc(sizeof(unsigned) * (size_t)scanTmpEntries(ni + 1));
    front[0].alloc(sizeof(int2) * (size_t)ni);
    front[1].alloc(sizeof(int2) * (size_t)ni);
    const int2 root = make_int2(t.root, 0);
    PT_HIP_TRY(hipMemcpyAsync(front[0].p, &root, sizeof root, hipMemcpyHostToDevice, stream));
    int m = 1, total = 1, level = 1, cur = 0;
    while (true) {
        k_level_count<<<blocks(m, 256), 256, 0, stream>>>(t, tb, front[cur].as<int2>(), m, live.as<unsigned>());
        exclusiveScan<unsigned>(live.as<unsigned>(), m, scanTmp.as<unsigned>(), stream);
        k_level_write<<<blocks(m, 256), 256, 0, stream>>>(t, tb, front[cur].as<int2>(), m, live.as<unsigned>(), total,
                                                          front[1 - cur].as<int2>(), tmpNodes.as<float4>());
        PT_HIP_TRY(hipGetLastError());
        unsigned nextM = 0;
        PT_HIP_TRY(hipMemcpyAsync(&nextM, live.as<unsigned>() + m, sizeof nextM, hipMemcpyDeviceToHost, stream));
        PT_HIP_TRY(hipStreamSynchronize(stream));
        if (nextM == 0) break;
        if (total + (int64_t)nextM > ni) throw HipFailure("BVH collapse: more 4-wide nodes than internal nodes");
        total += (int)nextM;
        m = (int)nextM;
        level++;
        cur = 1 - cur;
    }
    nodes.alloc(sizeof(float4) * PT_NODE_F4 * (size_t)total);
    PT_HIP_TRY(hipMemcpyAsync(nodes.p, tmpNodes.p, sizeof(float4) * PT_NODE_F4 * (size_t)total, hipMemcpyDeviceToDevice, stream));
    // temporaries are freed on return
    PT_HIP_TRY(hipStreamSynchronize(stream));
    *levels = level;
    return total;
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
              BvhInfo *info, int builder) {
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
    DevBuf children, cnt, first, parentNode, parentLeaf, nodeBoxes, counters, leafPos;
    children.alloc(sizeof(int2) * ni);
    cnt.alloc(sizeof(int) * ni);
    first.alloc(sizeof(int) * ni);
    parentNode.alloc(sizeof(int) * ni);
    parentLeaf.alloc(sizeof(int) * n);
    nodeBoxes.alloc(sizeof(float) * 6 * ni);
    counters.alloc(sizeof(int) * 2);
    k_fill_int<<<blocks(ni, 256), 256, 0, stream>>>(parentNode.as<int>(), ni, -1);
    k_fill_int<<<1, 64, 0, stream>>>(counters.as<int>(), 2, 0);
    BinTree t;
    t.n = n;
    t.children = children.as<int2>();
    t.nodeBoxes = nodeBoxes.as<float>();
    t.leafBoxes = boxes2.as<float>();
    t.cnt = cnt.as<int>();
    t.first = first.as<int>();
    t.parentNode = parentNode.as<int>();
    t.parentLeaf = parentLeaf.as<int>();
    if (builder == PT_BVH_PLOC) {
        // clusters start as the Morton-ordered leaves; each iteration merges mutual nearest
        // neighbours, whose new nodes take the next indices in position order
        DevBuf clus[2], cbox[2], nn, flags, scanTmp, primsTmp;
        for (int k = 0; k < 2; k++) {
            clus[k].alloc(sizeof(int) * n);
            cbox[k].alloc(sizeof(float) * 6 * n);
        }
        nn.alloc(sizeof(int) * n);
        flags.alloc(sizeof(unsigned long long) * ((size_t)n + 1));
        scanTmp.alloc(sizeof(unsigned long long) * (size_t)scanTmpEntries(n + 1));
        k_ploc_init<<<blocks(n, 256), 256, 0, stream>>>(n, boxes2.as<float>(), clus[0].as<int>(), cbox[0].as<float>());
        int m = n, nodeBase = 0, cur = 0;
        while (m > 1) {
            k_ploc_nn<<<blocks(m, kPlocThreads), kPlocThreads, 0, stream>>>(cbox[cur].as<float>(), m, nn.as<int>());
            for (int attempt = 0; attempt < 2; attempt++) {
                if (attempt == 1) k_ploc_pairs<<<blocks(m, 256), 256, 0, stream>>>(m, nn.as<int>());
                k_ploc_flags<<<blocks(m, 256), 256, 0, stream>>>(nn.as<int>(), m, flags.as<unsigned long long>());
                exclusiveScan<unsigned long long>(flags.as<unsigned long long>(), m, scanTmp.as<unsigned long long>(), stream);
                unsigned long long tot = 0;
                PT_HIP_TRY(hipMemcpyAsync(&tot, flags.as<unsigned long long>() + m, sizeof tot, hipMemcpyDeviceToHost, stream));
                PT_HIP_TRY(hipStreamSynchronize(stream));
                const int merged = (int)(tot >> 32);
                if (merged == 0 && attempt == 0) continue;  // no mutual pair: merge neighbours pairwise
                if (merged == 0 || nodeBase + merged > ni) throw HipFailure("PLOC: clustering made no progress");
                k_ploc_merge<<<blocks(m, 256), 256, 0, stream>>>(clus[cur].as<int>(), cbox[cur].as<float>(), nn.as<int>(), m,
                                                                 flags.as<unsigned long long>(), nodeBase, clus[1 - cur].as<int>(),
                                                                 cbox[1 - cur].as<float>(), children.as<int2>(), nodeBoxes.as<float>(),
                                                                 cnt.as<int>(), parentNode.as<int>(), parentLeaf.as<int>());
                PT_HIP_TRY(hipGetLastError());
                nodeBase += merged;
                m = (int)(tot & 0xffffffffull);
                cur = 1 - cur;
                break;
            }
        }
        if (nodeBase != ni) throw HipFailure("PLOC: the tree does not have n - 1 internal nodes");
        t.root = ni - 1;  // the last merge
        // depth-first leaf order, then every internal node's first position
        leafPos.alloc(sizeof(int) * n);
        k_leaf_positions<<<blocks(n, 256), 256, 0, stream>>>(t, leafPos.as<int>());
        t.leafPos = leafPos.as<int>();
        k_node_first<<<blocks(n, 256), 256, 0, stream>>>(t, first.as<int>());
        primsTmp.alloc(sizeof(float4) * 3 * (size_t)n);
        PT_HIP_TRY(hipMemcpyAsync(primsTmp.p, primsSorted, sizeof(float4) * 3 * (size_t)n, hipMemcpyDeviceToDevice, stream));
        k_scatter_prims<<<blocks(n, 256), 256, 0, stream>>>(primsTmp.as<float4>(), leafPos.as<int>(), n, primsSorted);
        PT_HIP_TRY(hipGetLastError());
        PT_HIP_TRY(hipStreamSynchronize(stream));  // primsTmp is freed on leaving this block
    } else {
        DevBuf ranges, arrivals;
        ranges.alloc(sizeof(int2) * ni);
        arrivals.alloc(sizeof(int) * ni);
        k_fill_int<<<blocks(ni, 256), 256, 0, stream>>>(arrivals.as<int>(), ni, 0);
        k_karras<<<blocks(ni, 256), 256, 0, stream>>>(kin, n, children.as<int2>(), ranges.as<int2>(), parentNode.as<int>(), parentLeaf.as<int>());
        k_refit<<<blocks(n, 256), 256, 0, stream>>>(children.as<int2>(), parentNode.as<int>(), parentLeaf.as<int>(), boxes2.as<float>(),
                                                     nodeBoxes.as<float>(), arrivals.as<int>(), n);
        k_ranges_to_counts<<<blocks(ni, 256), 256, 0, stream>>>(ranges.as<int2>(), ni, cnt.as<int>(), first.as<int>());
        PT_HIP_TRY(hipGetLastError());
        PT_HIP_TRY(hipStreamSynchronize(stream));
        t.root = 0;
    }
    k_depth<<<blocks(n, 256), 256, 0, stream>>>(parentNode.as<int>(), parentLeaf.as<int>(), n, counters.as<int>());
    PT_HIP_TRY(hipGetLastError());
    int levels = 1;
    const int count = collapseWide(t, nodes, stream, &levels);
    int hc[2] = {0, 0};
    PT_HIP_TRY(hipMemcpy(hc, counters.p, sizeof hc, hipMemcpyDeviceToHost));
    info->numNodes = count;
    info->depth = levels;
    info->binaryDepth = hc[0];
}

}  // namespace pt
