// pt_gpu.h -- device-resident builders shared by the renderer and the unit entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "pt_types.h"

namespace pt {

#define PT_HIP_TRY(expr)                                                                   \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            throw HipFailure(std::string(#expr) + ": " + hipGetErrorString(_e));           \
        }                                                                                  \
    } while (0)

struct HipFailure {
    std::string msg;
    explicit HipFailure(std::string m) : msg(std::move(m)) {}
};

// Owning device allocation (hipMalloc / hipFree).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n == bytes && p) return;
        release();
        if (n == 0) return;
        PT_HIP_TRY(hipMalloc(&p, n));
        bytes = n;
    }
    template <class T> T *as() const { return (T *)p; }
};

struct BvhInfo {
    int numNodes = 0;      // 4-wide nodes
    int depth = 0;         // levels of 4-wide nodes
    int binaryDepth = 0;   // levels of internal nodes of the binary radix tree
};

// Binary-tree builders of buildBvh.
enum { PT_BVH_LBVH = 0, PT_BVH_PLOC = 1 };

// Builds a binary BVH over `n` primitive records (3 float4 each, input order) on `stream`
// (builder: PT_BVH_LBVH or PT_BVH_PLOC) and collapses it into a 4-wide BVH on the GPU:
// allocates `nodes` (PT_NODE_F4 float4 per node) and writes the records in BVH order to
// `primsSorted`.  lo/hi bound all primitive centroids.
void buildBvh(const float4 *primsIn, int n, const float lo[3], const float hi[3], DevBuf &nodes,
              float4 *primsSorted, hipStream_t stream, BvhInfo *info, int builder);

// Environment tables in FP64 on the GPU: per-row and marginal alias tables packed as
// (prob as float, alias as int bits), texel pdf.  Optionally returns the FP64 tables.
void buildEnvTables(const float4 *rgba, int w, int h, int sinWeight, float2 *rows, float2 *marginal,
                    float *texelPdf, double *rowProbOut, int *rowAliasOut, double *margProbOut,
                    int *margAliasOut, hipStream_t stream);

// Copies the texel pdf into the w component of the environment texels (PtEnv::rgba).
void packEnvPdf(float4 *rgba, const float *texelPdf, int n, hipStream_t stream);

int deviceCUs();

}  // namespace pt
