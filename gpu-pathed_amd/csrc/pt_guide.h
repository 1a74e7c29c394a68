// pt_guide.h -- learned guiding distribution: a 5-layer MLP (one-blob input 8 x 8 = 64,
// four hidden layers of 64 ReLU units, 48 outputs of which 41 are used) that maps a path
// vertex to an 8-lobe von Mises-Fisher mixture over world directions plus the one-sample MIS
// fraction of the guide against BSDF sampling.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_device.h"
#include "pt_types.h"

namespace pt {

constexpr int GUIDE_RAW = 8;       // raw features: position (3), normal (2), wo (2), roughness (1)
constexpr int GUIDE_BINS = 8;      // one-blob bins per raw feature
constexpr int GUIDE_W = 64;        // width of every layer (input encoding, hidden, padded output)
constexpr int GUIDE_LAYERS = 5;
constexpr int GUIDE_OUT = 48;      // outputs computed (41 used)
constexpr int GUIDE_LOBES = 8;
constexpr int GUIDE_LAYER_PARAMS = GUIDE_W * GUIDE_W + GUIDE_W;          // W[out][in] then b[out]
constexpr int GUIDE_PARAMS = GUIDE_LAYERS * GUIDE_LAYER_PARAMS;           // 20800 (19305 live)
// output indices
constexpr int GO_LOGIT = 0;    // 8 mixture logits
constexpr int GO_MEAN = 8;     // 8 x 3 unnormalised lobe means
constexpr int GO_KAPPA = 32;   // 8 log concentrations
constexpr int GO_FRAC = 40;    // guide selection probability (logit)
constexpr float GUIDE_LOGK_MIN = -3.f, GUIDE_LOGK_MAX = 9.f;
constexpr float GUIDE_FRAC_MIN = 0.02f, GUIDE_FRAC_MAX = 0.98f;

// The last layer's outputs in the network kernels' order.  Output position op = 16 ob + 4 g
// + r is the C-fragment element r of output block ob in lane group g, so the four lanes that
// hold a row's outputs hold two lobes each: slot s = 4 ob + r of group g is the logit of lobe
// 2 g + s (s < 2), a mean coordinate of lobe 2 g + (s - 2) / 3 (s < 8), the log concentration
// of lobe 2 g + s - 8 (s < 10), and s = 10, 11 hold the selection logit (group 3, s = 10) and
// the seven computed outputs the loss does not use (41 .. 47).  Positions 48 .. 63 (block 3,
// never computed) map to themselves.  guideOutOf and guideOpOf are inverse permutations of
// 0 .. 63.
__host__ __device__ __forceinline__ int guideOutOf(int op) {
    if (op >= GUIDE_OUT) return op;
    const int g = (op >> 2) & 3, s = 4 * (op >> 4) + (op & 3);
    if (s < 2) return GO_LOGIT + 2 * g + s;
    if (s < 8) return GO_MEAN + 3 * (2 * g + (s - 2) / 3) + (s - 2) % 3;
    if (s < 10) return GO_KAPPA + 2 * g + (s - 8);
    if (g == 3) return s == 10 ? GO_FRAC : GUIDE_OUT - 1;
    return GO_FRAC + 1 + 2 * g + (s - 10);
}
__host__ __device__ __forceinline__ int guideOpOf(int o) {
    if (o >= GUIDE_OUT) return o;
    int g, s;
    if (o < GO_MEAN) {
        g = (o - GO_LOGIT) >> 1;
        s = (o - GO_LOGIT) & 1;
    } else if (o < GO_KAPPA) {
        const int k = (o - GO_MEAN) / 3;
        g = k >> 1;
        s = 2 + 3 * (k & 1) + (o - GO_MEAN) % 3;
    } else if (o < GO_FRAC) {
        g = (o - GO_KAPPA) >> 1;
        s = 8 + ((o - GO_KAPPA) & 1);
    } else if (o == GO_FRAC) {
        g = 3;
        s = 10;
    } else if (o == GUIDE_OUT - 1) {
        g = 3;
        s = 11;
    } else {
        g = (o - GO_FRAC - 1) >> 1;
        s = 10 + ((o - GO_FRAC - 1) & 1);
    }
    return 16 * (s >> 2) + 4 * g + (s & 3);
}

// Mixture record written by the inference kernel, 64 bytes (16 words) per path vertex:
// lobe k in words 2k (its unit mean, octahedral, two unorm16) and 2k + 1 (kappa as FP16 in
// the low half; the lobe's weight as FP16 in the high half for k < 7, the selection
// probability for k = 7, whose weight is 1 minus the others').  A vertex that is not guided
// has word 15 = 0 (selection probability 0).
constexpr int GUIDE_REC_F4 = 4;

__device__ __forceinline__ uint32_t guidePackHalf2(float lo, float hi) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)lo) | ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)hi) << 16);
}
__device__ __forceinline__ float guideHalfLo(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)); }
__device__ __forceinline__ float guideHalfHi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }
// Unit vector <-> octahedral map (Cigolle et al. 2014, "A Survey of Efficient Representations
// for Independent Unit Vectors"), 16 bits per coordinate.
__device__ __forceinline__ uint32_t guideOctEncode(f3 n) {
    const float il1 = __builtin_amdgcn_rcpf(fabsf(n.x) + fabsf(n.y) + fabsf(n.z));
    float u = n.x * il1, v = n.y * il1;
    if (n.z < 0.f) {
        const float ou = u;
        u = (1.f - fabsf(v)) * (ou >= 0.f ? 1.f : -1.f);
        v = (1.f - fabsf(ou)) * (v >= 0.f ? 1.f : -1.f);
    }
    const uint32_t qu = (uint32_t)(fminf(fmaxf(u * 0.5f + 0.5f, 0.f), 1.f) * 65535.f + 0.5f);
    const uint32_t qv = (uint32_t)(fminf(fmaxf(v * 0.5f + 0.5f, 0.f), 1.f) * 65535.f + 0.5f);
    return qu | (qv << 16);
}
__device__ __forceinline__ f3 guideOctDecode(uint32_t w) {
    const float u = (float)(w & 0xffffu) * (2.f / 65535.f) - 1.f, v = (float)(w >> 16) * (2.f / 65535.f) - 1.f;
    f3 n = mk(u, v, 1.f - fabsf(u) - fabsf(v));
    const float t = fmaxf(-n.z, 0.f);
    n.x += n.x >= 0.f ? -t : t;
    n.y += n.y >= 0.f ? -t : t;
    return n * __builtin_amdgcn_rsqf(dot(n, n));
}
// The lobe weights of a record, renormalised so that sampling and density use the same
// distribution whatever the FP16 rounding of the stored weights.
__device__ __forceinline__ void guideLobeWeights(const uint32_t *w, float *pi) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < GUIDE_LOBES - 1; k++) {
        pi[k] = guideHalfHi(w[2 * k + 1]);
        s += pi[k];
    }
    pi[GUIDE_LOBES - 1] = fmaxf(1.f - s, 0.f);
    const float inv = __builtin_amdgcn_rcpf(s + pi[GUIDE_LOBES - 1]);
#pragma unroll
    for (int k = 0; k < GUIDE_LOBES; k++) pi[k] *= inv;
}

// Training record, 4 float4: raw features (2 float4), (omega.xyz, bsdf pdf),
// (1 / luminance of the throughput at the vertex, path id bits, radiance luminance before the
// continuation's contribution, weight)
constexpr int GUIDE_TREC_F4 = 4;

// BF16 weights as MFMA A fragments of the transposed layer products (GUIDE_W * GUIDE_W bf16
// per layer, layout in pt_guide.hip): w for the forward pass (W), wt for the backward pass
// (W^T).
struct GuideNet {
    const __bf16 *w;
    const __bf16 *wt;
    const float *b;     // [5][64]
};

struct GuideSceneNorm {
    float lo[3], invExt[3];
};

// Raw guide features in [0, 1]: position in the scene box, shading normal and outgoing
// direction as (theta / pi, phi / 2 pi) in world space, roughness (Lambert 1, Beckmann
// alpha / 0.6).
__device__ __forceinline__ void guideFeatures(const GuideSceneNorm &nm, f3 p, f3 n, f3 wo, float rough, float *f) {
    f[0] = fminf(fmaxf((p.x - nm.lo[0]) * nm.invExt[0], 0.f), 1.f);
    f[1] = fminf(fmaxf((p.y - nm.lo[1]) * nm.invExt[1], 0.f), 1.f);
    f[2] = fminf(fmaxf((p.z - nm.lo[2]) * nm.invExt[2], 0.f), 1.f);
    f[3] = acosf(fminf(fmaxf(n.y, -1.f), 1.f)) * 0.31830988618f;
    f[4] = (atan2f(n.z, n.x) + 3.14159265358979f) * 0.15915494309f;
    f[5] = acosf(fminf(fmaxf(wo.y, -1.f), 1.f)) * 0.31830988618f;
    f[6] = (atan2f(wo.z, wo.x) + 3.14159265358979f) * 0.15915494309f;
    f[7] = rough;
}

// host-side entry points (pt_guide.hip)
void guideInitParams(float *params, uint64_t seed);
void guideUploadBf16(const float *dParams, __bf16 *dW, __bf16 *dWt, float *dB, hipStream_t s);
// Mixture per slot of a segmented ray queue (count of segment s at segCount[s * segStride],
// PT_QUEUE_SEGS segments of segLen slots); feat holds two float4 of features per slot.  With
// camO.w != 0 the queue holds camera rays starting at camO and rayO is not read.
void guideInfer(const PtSceneView &sc, GuideSceneNorm norm, GuideNet net, const int *segCount, int segStride, int segLen,
                const float4 *rayO, const float4 *rayD, float4 camO, const float4 *hits, float4 *feat, float4 *out, int cus, hipStream_t s);
void guideInferRaw(GuideNet net, const float *feats, int n, float *rawOut, hipStream_t s);
// The mixture records (GUIDE_REC_F4 float4 each) of n rows of raw features, every row guided.
void guideInferRecords(GuideNet net, const float *feats, int n, float4 *records, hipStream_t s);
// Scratch of one training step over `batch` records: FP32 gradient partials (blocks x
// GUIDE_PARAMS) and per-block loss sums of the training kernel, which runs
// guideTrainBlocks(batch, cus) blocks of GUIDE_TRAIN_ROWS records per iteration.
constexpr int GUIDE_TRAIN_ROWS = 128;
inline int guideTrainBlocks(int batch, int cus) {
    const int iters = (batch + GUIDE_TRAIN_ROWS - 1) / GUIDE_TRAIN_ROWS;
    return iters < 1 ? 1 : (iters < cus ? iters : cus);
}
struct GuideTrainBufs {
    float *partials;
    float *lossPart;
    int blocks;
};
// Gradients of the mean loss over `batch` records (rows (base + i) mod *count of the record
// sequence; weights scaled to mean one by *wsum when wsum is given) into `grad`, the mean loss
// into lossOut[0].  With prefix = nullptr the sequence is records[0 ..); otherwise record v of
// the sequence lies in segment s (prefix[s] <= v < prefix[s + 1], PT_QUEUE_SEGS segments) at
// records[s * recSeg + v - prefix[s]].
void guideGradients(GuideNet net, const float4 *records, const int *count, const int *prefix, int recSeg, int base, int batch,
                    const float *wsum, const GuideTrainBufs &tb, float *grad, float *lossOut, hipStream_t s);
void guideAdam(float *params, float *m, float *v, const float *grad, int step, float lr, __bf16 *dW, __bf16 *dWt, float *dB, hipStream_t s);
// C (32 x 32) = A (32 x 16) x B (16 x 32) with one v_mfma_f32_32x32x16_bf16 (lane-layout check)
void mfmaTile(const float *A, const float *B, float *C, hipStream_t s);
void mfmaTile16(const float *A, const float *B, float *C, hipStream_t s);

}  // namespace pt
