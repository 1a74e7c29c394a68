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

// Mixture record written by the inference kernel, 12 float4 per path vertex:
//   [0..7] (mu.xyz, kappa) per lobe, [8] pi 0..3, [9] pi 4..7, [10] (selection prob, eligible, 0, 0)
constexpr int GUIDE_REC_F4 = 12;

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
