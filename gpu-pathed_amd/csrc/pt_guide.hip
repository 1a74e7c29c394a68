// pt_guide.hip -- BF16 MFMA kernels of the guiding network (gfx950).
//
// Every layer is 64 x 64 (the 48-output last layer is zero-padded to 64), weights BF16
// ([out][in] for the forward pass, [in][out] for the backward pass), activations BF16,
// accumulation FP32 with v_mfma_f32_32x32x16_bf16: one wave multiplies a 32-row tile by a
// 64-wide layer as 2 output tiles x 4 K-steps.  Fragment layout of the 32x32x16 instruction
// used throughout (A: M x K, B: K x N, lane l = 0..63):
//   A: row (l & 31), 8 consecutive k starting at 8 * (l >> 5) of the 16-wide K step
//   B: col (l & 31), the same 8 k
//   C: 16 floats per lane, element r at row 8 * (r >> 2) + 4 * (l >> 5) + (r & 3), col (l & 31)
// (checked on the device by pt_unit_mfma_tile).  Operands are staged so that each lane's 8
// k-values are contiguous: activations row-major for the forward A operand, transposed copies
// for the weight-gradient products.
//
// Inference: blocks of 8 waves, each wave loops over 64-row tiles of the ray queue, encodes
// the rows' raw features one-blob straight into MFMA operands, runs the 5 layers with the
// activations in registers (v_mfma_f32_16x16x32_bf16 on the transposed product, weights
// permuted to match) and writes the mixture (8 vMF lobes + selection probability) per row.
// Training step: (1) one wave per block, persistent over 32-row tiles of the record batch:
// forward, KL loss gradient at the output, backward through the layers; every layer's input
// activations and output deltas are written transposed (BF16) to HBM; (2) the weight and bias
// gradients as one MFMA product per layer with K = the batch rows, split into slices with FP32
// partials; (3) the partials summed in slice order (deterministic) and Adam on the FP32
// master weights, refreshing the BF16 copies.
#include <hip/hip_runtime.h>
#include <math.h>

#include <random>
#include <vector>

#include "pt_device.h"
#include "pt_gpu.h"
#include "pt_guide.h"
#include "pt_internal.h"

namespace pt {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x16 __attribute__((__vector_size__(16 * sizeof(float))));

constexpr int ACT_STRIDE = 72;   // bf16 per row of a row-major 32 x 64 tile (padded)
constexpr int T_STRIDE = 40;     // bf16 per row of a transposed 64 x 32 tile (padded)

#define WAVE_SYNC()                                         \
    do {                                                    \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
        __builtin_amdgcn_wave_barrier();                    \
    } while (0)

__device__ __forceinline__ int cRow(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

__device__ __forceinline__ bf16x8 ld8(const __bf16 *p) { return *(const bf16x8 *)p; }

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }

// v_mfma_f32_16x16x32_bf16 (inference): A: row (l & 15), 8 consecutive k starting at
// 8 * (l >> 4) of the 32-wide K step; B: col (l & 15), the same 8 k; C: 4 floats per lane,
// element r at row 4 * (l >> 4) + r, col (l & 15) (checked on the device by
// pt_unit_mfma_tile16).  A 32 x 64 layer output is 2 x 4 such tiles of two K steps each:
// eight independent accumulation chains per wave instead of two.
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));
__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

__device__ __forceinline__ float luminance(float4 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

__device__ __forceinline__ float oneBlob(float s, int j) {
    const float x = s * (float)GUIDE_BINS - ((float)j + 0.5f);
    return __expf(-0.5f * x * x);
}

// forward through layer L for one 32-row tile: A from `act` (row-major), B from W[L]
__device__ __forceinline__ f32x16 layerTile(const __bf16 *act, const __bf16 *W, int L, int ct, int lane) {
    f32x16 acc = {};
    const __bf16 *arow = act + (lane & 31) * ACT_STRIDE + (lane >> 5) * 8;
    const __bf16 *wrow = W + (size_t)L * GUIDE_W * GUIDE_W + (ct * 32 + (lane & 31)) * GUIDE_W + (lane >> 5) * 8;
#pragma unroll
    for (int ks = 0; ks < GUIDE_W / 16; ks++) acc = mfma(ld8(arow + ks * 16), ld8(wrow + ks * 16), acc);
    return acc;
}

// mixture parameters from the 41 used raw outputs
struct Mixture {
    f3 mu[GUIDE_LOBES];
    float kappa[GUIDE_LOBES], pi[GUIDE_LOBES], mlen[GUIDE_LOBES];
    float frac;
};
__device__ __forceinline__ void mixtureFromRaw(const float *o, Mixture &mx) {
    float mxl = o[GO_LOGIT];
    for (int k = 1; k < GUIDE_LOBES; k++) mxl = fmaxf(mxl, o[GO_LOGIT + k]);
    float s = 0.f;
    for (int k = 0; k < GUIDE_LOBES; k++) {
        mx.pi[k] = __expf(o[GO_LOGIT + k] - mxl);
        s += mx.pi[k];
    }
    for (int k = 0; k < GUIDE_LOBES; k++) {
        mx.pi[k] /= s;
        const f3 m = mk(o[GO_MEAN + 3 * k], o[GO_MEAN + 3 * k + 1], o[GO_MEAN + 3 * k + 2]);
        const float l = sqrtf(dot(m, m));
        mx.mlen[k] = l;
        mx.mu[k] = l > 1e-12f ? m / l : mk(0.f, 1.f, 0.f);
        mx.kappa[k] = __expf(fminf(fmaxf(o[GO_KAPPA + k], GUIDE_LOGK_MIN), GUIDE_LOGK_MAX));
    }
    const float c = 1.f / (1.f + __expf(-o[GO_FRAC]));
    mx.frac = fminf(fmaxf(c, GUIDE_FRAC_MIN), GUIDE_FRAC_MAX);
}

// von Mises-Fisher density on S^2: kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1))
__device__ __forceinline__ float vmfPdf(f3 mu, float kappa, f3 w) {
    return kappa / (6.28318530718f * (1.f - __expf(-2.f * kappa))) * __expf(kappa * (dot(mu, w) - 1.f));
}

// --------------------------------------------------------------------------------------------
// inference
// --------------------------------------------------------------------------------------------
// Inference blocks: eight waves share the LDS-staged weight fragments (40 KiB) and biases;
// each wave owns a 64-row tile at a time and 12.25 KiB of LDS for its raw outputs.
constexpr int INFER_WAVES = 8;
constexpr int INFER_ROWS = 64;

// Raw features of the vertex at queue slot i (hit point normalised to the scene box, shading
// normal, outgoing direction, roughness); false where the vertex is not guided (a miss, a
// mirror or a glass surface).
__device__ __forceinline__ bool vertexFeatures(const PtSceneView &sc, const GuideSceneNorm &nm, int i, const float4 *rayO, const float4 *rayD,
                                               float4 camO, const float4 *hits, float *f) {
    const float4 h = hits[i];
    const int slot = asInt(h.w);
    if (slot == INT32_MIN) return false;
    // camera rays (camO.w != 0) start at the camera and have no stored origin
    const float4 ro = camO.w != 0.f ? camO : rayO[i], rd = rayD[i];
    const f3 o = xyz(ro), d = xyz(rd);
    const float4 p0 = sc.prims[3 * (size_t)slot], p1 = sc.prims[3 * (size_t)slot + 1];
    float4 p2 = sc.prims[3 * (size_t)slot + 2];
    // issues the third record's load with the first two instead of after the type tests
    asm volatile("" : "+v"(p2.x), "+v"(p2.y), "+v"(p2.z));
    const PtMaterial m = sc.materials[asInt(p0.w)];
    const int type = matType(m);
    if (type != PT_MAT_LAMBERTIAN && type != PT_MAT_MICROFACET) return false;
    f3 p, nn;
    const int tag = asInt(p1.w);
    if (tag == PT_TAG_SPHERE) {
        p = o + d * h.x;
        nn = normalize((p - xyz(p0)) / p1.x);
    } else {
        const f3 a = xyz(p0), b = xyz(p1), c = xyz(p2);
        p = a * (1.f - h.y - h.z) + b * h.y + c * h.z;
        if (tag >= 0) {
            const f3 na = xyz(sc.normals[3 * (size_t)tag]), nb = xyz(sc.normals[3 * (size_t)tag + 1]), nc = xyz(sc.normals[3 * (size_t)tag + 2]);
            nn = normalize(na * (1.f - h.y - h.z) + nb * h.y + nc * h.z);
        } else {
            nn = normalize(cross(b - a, c - a));
        }
    }
    const float rough = type == PT_MAT_LAMBERTIAN ? 1.f : fminf(m.b.w * (1.f / 0.6f), 1.f);
    guideFeatures(nm, p, nn, -d, rough, f);
    return true;
}

// Features of every filled slot of a segmented ray queue (count of segment s at
// count[s * segStride], segments segLen slots long): two float4 per slot, the eighth feature
// (roughness, > 0) set to -1 where the vertex is not guided.  A kernel of its own, at full
// occupancy, so that the dependent gathers (hit -> primitive -> material) overlap.
__global__ __launch_bounds__(256) void k_guide_features(PtSceneView sc, GuideSceneNorm nm, const int *count, int segStride, int segLen,
                                                        const float4 *rayO, const float4 *rayD, float4 camO, const float4 *hits, int maxRows,
                                                        float4 *feat) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < maxRows; i += gridDim.x * blockDim.x) {
        const int seg = i / segLen;
        if (i - seg * segLen >= count[seg * segStride]) continue;
        float f[GUIDE_RAW];
        for (int k = 0; k < GUIDE_RAW; k++) f[k] = 0.f;
        if (!vertexFeatures(sc, nm, i, rayO, rayD, camO, hits, f)) {
            for (int k = 0; k < GUIDE_RAW; k++) f[k] = 0.f;
            f[GUIDE_RAW - 1] = -1.f;
        }
        feat[2 * (size_t)i] = make_float4(f[0], f[1], f[2], f[3]);
        feat[2 * (size_t)i + 1] = make_float4(f[4], f[5], f[6], f[7]);
    }
}

// Inference computes every layer transposed, H^T = W A^T, on v_mfma_f32_16x16x32_bf16 with
// the weights as the A operand (M = 16 outputs) and the activations as the B operand (N = 16
// rows).  The C fragment of output block ob then leaves in lane l the outputs
// 16 ob + 4 (l >> 4) + r (r = 0..3) of row l & 15, and the eight of output blocks 2 ks and
// 2 ks + 1 are exactly the B fragment of the next layer's K step ks once the contraction index
// is permuted: K position 8 g + p (lane group g = l >> 4) stands for input feature kFeature(g,
// p, ks).  Activations therefore stay in registers from the input encoding to the last layer;
// the weights are stored permuted the same way (wFragIndex), one 16-byte fragment per lane per
// (layer, output block, K step).
__host__ __device__ __forceinline__ int kFeature(int g, int p, int ks) { return 32 * ks + (p < 4 ? 4 * g + p : 16 + 4 * g + (p - 4)); }
// position of W[L][out][in] in the fragment copy (the inverse of kFeature)
__host__ __device__ __forceinline__ int wFragIndex(int L, int out, int in) {
    const int ob = out >> 4, ks = in >> 5, q = in & 31;
    const int g = (q & 15) >> 2, p = (q & 3) + (q >= 16 ? 4 : 0);
    return (((L * 4 + ob) * 2 + ks) * 64 + 16 * g + (out & 15)) * 8 + p;
}

// Rows are the slots [0, maxRows) of a segmented queue (count of segment s at
// count[s * segStride], segments segLen slots long, a multiple of 64, so that a tile of 64 rows
// lies in one segment) whose features k_guide_features wrote to feat, or with count = nullptr
// all maxRows rows of rawFeats.
__global__ __launch_bounds__(64 * INFER_WAVES) void k_guide_infer(GuideNet net, const int *count, int segStride, int segLen, const float4 *feat,
                                                                  int maxRows, float4 *out, const float *rawFeats, float *rawOut) {
    __shared__ bf16x8 sWf[GUIDE_LAYERS * 8 * 64];
    __shared__ float4 sB[GUIDE_LAYERS * GUIDE_W / 4];
    __shared__ float sOut[INFER_WAVES][INFER_ROWS][GUIDE_OUT + 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, lr = lane & 15;
    for (int c = threadIdx.x; c < GUIDE_LAYERS * 8 * 64; c += blockDim.x) sWf[c] = ld8(net.w + GUIDE_WFRAG_OFF + 8 * (size_t)c);
    for (int c = threadIdx.x; c < GUIDE_LAYERS * GUIDE_W / 4; c += blockDim.x) sB[c] = ((const float4 *)net.b)[c];
    __syncthreads();
    float(*so)[GUIDE_OUT + 1] = sOut[wave];
    const int waveStride = gridDim.x * INFER_WAVES;
    for (int tile = blockIdx.x * INFER_WAVES + wave; tile * INFER_ROWS < maxRows; tile += waveStride) {
        const int rowBase = tile * INFER_ROWS;
        // rows [rowBase, n) of this tile hold records: the filled part of its queue segment
        int n = maxRows;
        if (count) {
            const int seg = rowBase / segLen;
            n = seg * segLen + min(count[seg * segStride], segLen);
            if (rowBase >= n) continue;
        }
        // this lane's row in the output stage
        const int mine = rowBase + lane;
        const bool elig = mine < n && (rawFeats != nullptr || feat[2 * (size_t)mine + 1].w >= 0.f);
        if (__ballot(elig) == 0) {
            // no guided vertex in the tile (misses, mirror and glass hits): only the flag that
            // makes shade sample the BSDF alone, no network evaluation
            if (mine < n && !rawOut) out[(size_t)mine * GUIDE_REC_F4 + 10] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        // layer 0's B fragments: one-blob encoding.  Input feature 8 k + j is bin j of raw
        // feature k; the lane's K positions need raw features 2 i + (g >> 1), i = 0..3, at bins
        // 4 (g & 1) + 0..3.
        bf16x8 act[4][2];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const int row = rowBase + rb * 16 + lr;
            float f[4] = {0.f, 0.f, 0.f, 0.f};
            if (row < n) {
                if (rawFeats) {
#pragma unroll
                    for (int i = 0; i < 4; i++) f[i] = rawFeats[(size_t)row * GUIDE_RAW + 2 * i + (g >> 1)];
                } else {
                    const float4 a = feat[2 * (size_t)row], b = feat[2 * (size_t)row + 1];
                    if (b.w >= 0.f) {
                        const bool hi = (g >> 1) != 0;
                        f[0] = hi ? a.y : a.x;
                        f[1] = hi ? a.w : a.z;
                        f[2] = hi ? b.y : b.x;
                        f[3] = hi ? b.w : b.z;
                    }
                }
            }
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
#pragma unroll
                for (int p = 0; p < 8; p++) {
                    const int kf = kFeature(g, p, ks);
                    act[rb][ks][p] = (__bf16)oneBlob(f[(kf >> 3) >> 1], kf & 7);
                }
        }
#pragma unroll
        for (int L = 0; L < GUIDE_LAYERS; L++) {
            const bool last = L == GUIDE_LAYERS - 1;
            f32x4 acc[4][4];
#pragma unroll
            for (int ob = 0; ob < 4; ob++) {
                if (last && ob * 16 >= GUIDE_OUT) continue;
                const float4 bias = sB[(L * GUIDE_W + ob * 16) / 4 + g];
                const bf16x8 w0 = sWf[((L * 4 + ob) * 2) * 64 + lane], w1 = sWf[((L * 4 + ob) * 2 + 1) * 64 + lane];
#pragma unroll
                for (int rb = 0; rb < 4; rb++) {
                    const f32x4 c = {bias.x, bias.y, bias.z, bias.w};
                    acc[rb][ob] = mfma16(w1, act[rb][1], mfma16(w0, act[rb][0], c));
                }
            }
            if (!last) {
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
#pragma unroll
                    for (int ks = 0; ks < 2; ks++)
#pragma unroll
                        for (int p = 0; p < 8; p++) act[rb][ks][p] = (__bf16)fmaxf(acc[rb][2 * ks + (p >> 2)][p & 3], 0.f);
            } else {
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
#pragma unroll
                    for (int ob = 0; ob * 16 < GUIDE_OUT; ob++)
#pragma unroll
                        for (int r = 0; r < 4; r++) so[rb * 16 + lr][ob * 16 + 4 * g + r] = acc[rb][ob][r];
            }
        }
        WAVE_SYNC();
        if (mine < n) {
            if (rawOut) {
                for (int k = 0; k < GUIDE_OUT; k++) rawOut[(size_t)mine * GUIDE_OUT + k] = so[lane][k];
            } else if (!elig) {
                // shade reads only the flag of a vertex that is not guided
                out[(size_t)mine * GUIDE_REC_F4 + 10] = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                Mixture mx;
                mixtureFromRaw(so[lane], mx);
                float4 *o = out + (size_t)mine * GUIDE_REC_F4;
                for (int k = 0; k < GUIDE_LOBES; k++) o[k] = f4(mx.mu[k], mx.kappa[k]);
                o[8] = make_float4(mx.pi[0], mx.pi[1], mx.pi[2], mx.pi[3]);
                o[9] = make_float4(mx.pi[4], mx.pi[5], mx.pi[6], mx.pi[7]);
                o[10] = make_float4(mx.frac, 1.f, 0.f, 0.f);
                o[11] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        WAVE_SYNC();
    }
}

// --------------------------------------------------------------------------------------------
// training: gradients
// --------------------------------------------------------------------------------------------
// d loss / d raw outputs for one record; loss = -w (log q + log(c q + (1 - c) p_bsdf))
// Stores the clamped gradient as one BF16 row (GUIDE_W entries, zero where the loss does not
// depend on the output) and returns the loss.
__device__ float lossGradient(const float *o, f3 w, float pb, float wt, __bf16 *g) {
    const bf16x8 zero = {};
#pragma unroll
    for (int k = 0; k < GUIDE_W; k += 8) *(bf16x8 *)(g + k) = zero;
    if (!(wt > 0.f)) return 0.f;
    // keep single outliers from dominating a step
    auto put = [&](int k, float v) { g[k] = (__bf16)fminf(fmaxf(v, -1e3f), 1e3f); };
    Mixture mx;
    mixtureFromRaw(o, mx);
    float fk[GUIDE_LOBES];
    float q = 0.f;
    for (int k = 0; k < GUIDE_LOBES; k++) {
        fk[k] = vmfPdf(mx.mu[k], mx.kappa[k], w);
        q += mx.pi[k] * fk[k];
    }
    q = fmaxf(q, 1e-12f);
    const float c = mx.frac;
    const float mix = fmaxf(c * q + (1.f - c) * pb, 1e-12f);
    const float loss = -wt * (__logf(q) + __logf(mix));
    const float dq = -wt * (1.f / q + c / mix);
    // selection probability (sigmoid, clamped)
    const float craw = 1.f / (1.f + __expf(-o[GO_FRAC]));
    if (craw > GUIDE_FRAC_MIN && craw < GUIDE_FRAC_MAX) put(GO_FRAC, -wt * (q - pb) / mix * craw * (1.f - craw));
    for (int k = 0; k < GUIDE_LOBES; k++) {
        // softmax logits
        put(GO_LOGIT + k, dq * mx.pi[k] * (fk[k] - q));
        const float df = dq * mx.pi[k] * fk[k];  // d loss / d log f_k
        const float kap = mx.kappa[k];
        const float cosw = dot(mx.mu[k], w);
        // d log f / d kappa = 1/kappa - 2 e^{-2k} / (1 - e^{-2k}) + (mu.w - 1); kappa = exp(raw)
        const float e2 = __expf(-2.f * kap);
        const float dlk = 1.f / kap - 2.f * e2 / (1.f - e2) + (cosw - 1.f);
        const float lk = o[GO_KAPPA + k];
        if (lk > GUIDE_LOGK_MIN && lk < GUIDE_LOGK_MAX) put(GO_KAPPA + k, df * dlk * kap);
        // d log f / d m = kappa (w - (mu.w) mu) / |m|
        if (mx.mlen[k] > 1e-12f) {
            const float s = df * kap / mx.mlen[k];
            put(GO_MEAN + 3 * k, s * (w.x - cosw * mx.mu[k].x));
            put(GO_MEAN + 3 * k + 1, s * (w.y - cosw * mx.mu[k].y));
            put(GO_MEAN + 3 * k + 2, s * (w.z - cosw * mx.mu[k].z));
        }
    }
    return loss;
}

// Training, phase 1: forward and backward of 32-row tiles of the record batch, one wave per
// block, persistent over tiles.  Each layer's input activations and its output deltas leave
// the kernel transposed ([layer][feature][row], BF16) for the weight-gradient product of phase
// 2.  In LDS a wave keeps one activation tile and one delta tile, each overwritten in place
// (both 32-column halves of a layer are computed before either is stored), and the ReLU
// masks of the backward pass stay in registers: the lane that computed an output element in
// the forward pass is the lane that masks its delta.  Rows past the batch get weight 0 and so
// zero deltas.
__global__ __launch_bounds__(64) void k_guide_fwdbwd(GuideNet net, const float4 *recs, const int *countPtr, const int *prefix, int recSeg,
                                                     const float *wsumPtr, int base, int batch, int rows, __bf16 *actT, __bf16 *deltaT,
                                                     float *lossOut) {
    __shared__ __bf16 sAct[32 * ACT_STRIDE];
    __shared__ __bf16 sD[32 * ACT_STRIDE];
    __shared__ float sOut[32][GUIDE_OUT + 1];
    __shared__ float sRec[32][8];
    const int lane = threadIdx.x;
    const int count = *countPtr;
    const float wsum = wsumPtr ? *wsumPtr : 0.f;
    // normalise the weights to mean one over the record buffer
    const float wscale = (wsumPtr == nullptr) ? 1.f : (wsum > 0.f ? (float)count / wsum : 0.f);
    // C-fragment rows of this lane: groups g = 0..3 of 4 consecutive rows from 8 g + 4 (lane >> 5)
    const int half4 = 4 * (lane >> 5);
    auto storeT = [&](__bf16 *dst, const float *v) {
        // 16 values of one column (C-fragment order) as four 8-byte runs of 4 consecutive rows
#pragma unroll
        for (int g = 0; g < 4; g++) {
            typedef __bf16 bf16x4 __attribute__((__vector_size__(4 * sizeof(__bf16))));
            bf16x4 q = {(__bf16)v[4 * g], (__bf16)v[4 * g + 1], (__bf16)v[4 * g + 2], (__bf16)v[4 * g + 3]};
            *(bf16x4 *)(dst + 8 * g + half4) = q;
        }
    };
    float lossAcc = 0.f;
    const int ntiles = rows / 32;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int rowBase = tile * 32;
        if (lane < 32) {
            const int r = rowBase + lane;
            float f[GUIDE_RAW], om[3] = {0.f, 0.f, 0.f}, pb = 0.f, wt = 0.f;
            for (int k = 0; k < GUIDE_RAW; k++) f[k] = 0.f;
            if (r < batch && count > 0) {
                const int v = (int)(((int64_t)base + r) % count);
                size_t idx = (size_t)v;
                if (prefix) {
                    // the last segment whose prefix does not exceed v holds record v
                    int lo = 0, hi = PT_QUEUE_SEGS;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (prefix[mid] <= v) lo = mid;
                        else hi = mid;
                    }
                    idx = (size_t)lo * recSeg + (v - prefix[lo]);
                }
                const float4 a = recs[idx * GUIDE_TREC_F4], b = recs[idx * GUIDE_TREC_F4 + 1];
                const float4 c = recs[idx * GUIDE_TREC_F4 + 2], d = recs[idx * GUIDE_TREC_F4 + 3];
                f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w;
                f[4] = b.x; f[5] = b.y; f[6] = b.z; f[7] = b.w;
                om[0] = c.x; om[1] = c.y; om[2] = c.z;
                pb = c.w;
                wt = fminf(d.w * wscale, 64.f);
            }
            for (int k = 0; k < GUIDE_RAW; k++) sOut[lane][k] = f[k];
            sRec[lane][0] = om[0]; sRec[lane][1] = om[1]; sRec[lane][2] = om[2];
            sRec[lane][3] = pb;
            sRec[lane][4] = wt;
        }
        WAVE_SYNC();
        {
            // one-blob input: lane = input column (raw feature lane >> 3, bin lane & 7); all 32
            // rows row-major into LDS and transposed to layer 0's activations
            const int k = lane / GUIDE_BINS, j = lane % GUIDE_BINS;
            __bf16 *dst = actT + (size_t)lane * rows + rowBase;
            for (int q = 0; q < 4; q++) {
                bf16x8 v;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int row = 8 * q + i;
                    v[i] = (__bf16)oneBlob(sOut[row][k], j);
                    sAct[row * ACT_STRIDE + lane] = v[i];
                }
                *(bf16x8 *)(dst + 8 * q) = v;
            }
        }
        WAVE_SYNC();
        // ReLU masks of the four hidden layers' outputs, 32 bits per layer (bit 16 ct + r for
        // this lane's C-fragment element r of column half ct): layers 0-1 in maskLo, 2-3 in
        // maskHi, so that rolled layer loops select them without indexing registers
        uint64_t maskLo = 0, maskHi = 0;
#pragma unroll 1
        for (int L = 0; L < GUIDE_LAYERS; L++) {
            const f32x16 acc0 = layerTile(sAct, net.w, L, 0, lane);
            const f32x16 acc1 = layerTile(sAct, net.w, L, 1, lane);
            WAVE_SYNC();
            uint32_t m = 0;
#pragma unroll
            for (int ct = 0; ct < 2; ct++) {
                const int col = ct * 32 + (lane & 31);
                const float bias = net.b[L * GUIDE_W + col];
                float h[16];
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int row = cRow(lane, r);
                    const float v = (ct ? acc1[r] : acc0[r]) + bias;
                    h[r] = fmaxf(v, 0.f);
                    if (L < GUIDE_LAYERS - 1) {
                        const __bf16 hb = (__bf16)h[r];
                        sAct[row * ACT_STRIDE + col] = hb;
                        m |= ((float)hb > 0.f ? 1u : 0u) << (16 * ct + r);
                    } else if (col < GUIDE_OUT) {
                        sOut[row][col] = v;
                    }
                }
                if (L < GUIDE_LAYERS - 1) storeT(actT + ((size_t)(L + 1) * GUIDE_W + col) * rows + rowBase, h);
            }
            if (L < GUIDE_LAYERS - 1) {
                const uint64_t mm = (uint64_t)m << (32 * (L & 1));
                if (L < 2) maskLo |= mm;
                else maskHi |= mm;
            }
            WAVE_SYNC();
        }
        if (lane < 32) {
            const f3 w = mk(sRec[lane][0], sRec[lane][1], sRec[lane][2]);
            lossAcc += lossGradient(sOut[lane], w, sRec[lane][3], sRec[lane][4], sD + lane * ACT_STRIDE);
        }
        WAVE_SYNC();
        {
            // output deltas transposed: lane = output column
            __bf16 *dst = deltaT + ((size_t)(GUIDE_LAYERS - 1) * GUIDE_W + lane) * rows + rowBase;
            for (int q = 0; q < 4; q++) {
                bf16x8 v;
#pragma unroll
                for (int i = 0; i < 8; i++) v[i] = sD[(8 * q + i) * ACT_STRIDE + lane];
                *(bf16x8 *)(dst + 8 * q) = v;
            }
        }
#pragma unroll 1
        for (int L = GUIDE_LAYERS - 1; L > 0; L--) {
            // delta_{L-1}[row][in] = sum_out delta_L[row][out] W_L[out][in], masked by relu'
            f32x16 acc[2];
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                f32x16 a = {};
                const __bf16 *arow = sD + (lane & 31) * ACT_STRIDE + (lane >> 5) * 8;
                const __bf16 *brow = net.wt + (size_t)L * GUIDE_W * GUIDE_W + (nt * 32 + (lane & 31)) * GUIDE_W + (lane >> 5) * 8;
#pragma unroll
                for (int ks = 0; ks < GUIDE_W / 16; ks++) a = mfma(ld8(arow + ks * 16), ld8(brow + ks * 16), a);
                acc[nt] = a;
            }
            WAVE_SYNC();
            const uint32_t m = (uint32_t)((L - 1 < 2 ? maskLo : maskHi) >> (32 * ((L - 1) & 1)));
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const int in = nt * 32 + (lane & 31);
                float dv[16];
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int row = cRow(lane, r);
                    dv[r] = (m >> (16 * nt + r)) & 1u ? acc[nt][r] : 0.f;
                    sD[row * ACT_STRIDE + in] = (__bf16)dv[r];
                }
                storeT(deltaT + ((size_t)(L - 1) * GUIDE_W + in) * rows + rowBase, dv);
            }
            WAVE_SYNC();
        }
    }
    // loss of this block: lanes 0..31 hold partial sums
    for (int off = 32; off >= 1; off >>= 1) lossAcc += __shfl_down(lossAcc, off);
    if (lane == 0 && lossOut) lossOut[blockIdx.x] = lossAcc;
}

// Training, phase 2: weight and bias gradients of one layer over one slice of
// GUIDE_WG_ROWS rows, dW[out][in] = sum_rows delta[row][out] a[row][in] with K = rows on the
// MFMA (four waves, one 32 x 32 output tile each) and db[out] = sum_rows delta[row][out]; the
// slice's partials are summed in slice order afterwards.
__global__ __launch_bounds__(256) void k_guide_wgrad(const __bf16 *actT, const __bf16 *deltaT, int rows, float *partials) {
    const int slice = blockIdx.x, L = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int mt = wave >> 1, nt = wave & 1;
    const int r0 = slice * GUIDE_WG_ROWS, r1 = min(r0 + GUIDE_WG_ROWS, rows);
    const __bf16 *dp = deltaT + ((size_t)L * GUIDE_W + mt * 32 + (lane & 31)) * rows + (lane >> 5) * 8;
    const __bf16 *ap = actT + ((size_t)L * GUIDE_W + nt * 32 + (lane & 31)) * rows + (lane >> 5) * 8;
    // two accumulators over alternate K steps (rows is a multiple of 32) keep two MFMA chains
    // and their loads in flight
    f32x16 acc = {}, acc2 = {};
    float bsum = 0.f;
#pragma unroll 2
    for (int k = r0; k < r1; k += 32) {
        const bf16x8 a = ld8(dp + k), b = ld8(ap + k);
        const bf16x8 a2 = ld8(dp + k + 16), b2 = ld8(ap + k + 16);
        acc = mfma(a, b, acc);
        acc2 = mfma(a2, b2, acc2);
        if (nt == 0) {
#pragma unroll
            for (int i = 0; i < 8; i++) bsum += (float)a[i] + (float)a2[i];
        }
    }
    acc += acc2;
    float *P = partials + (size_t)slice * GUIDE_PARAMS + (size_t)L * GUIDE_LAYER_PARAMS;
    const int in = nt * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; r++) P[(mt * 32 + cRow(lane, r)) * GUIDE_W + in] = acc[r];
    if (nt == 0) {
        // lanes l and l + 32 summed complementary rows of output column mt * 32 + (l & 31)
        bsum += __shfl_down(bsum, 32);
        if (lane < 32) P[GUIDE_W * GUIDE_W + mt * 32 + lane] = bsum;
    }
}

// --------------------------------------------------------------------------------------------
// Adam (FP32 master weights) and the BF16 copies
// --------------------------------------------------------------------------------------------
__device__ __forceinline__ void storeParam(int p, float v, __bf16 *w, __bf16 *wt, float *b) {
    const int L = p / GUIDE_LAYER_PARAMS, off = p - L * GUIDE_LAYER_PARAMS;
    if (off < GUIDE_W * GUIDE_W) {
        const int out = off / GUIDE_W, in = off % GUIDE_W;
        const __bf16 h = (__bf16)v;
        w[(size_t)L * GUIDE_W * GUIDE_W + out * GUIDE_W + in] = h;
        w[GUIDE_WFRAG_OFF + wFragIndex(L, out, in)] = h;
        wt[(size_t)L * GUIDE_W * GUIDE_W + in * GUIDE_W + out] = h;
    } else {
        b[L * GUIDE_W + (off - GUIDE_W * GUIDE_W)] = v;
    }
}
__device__ __forceinline__ bool liveParam(int p) {
    // rows >= 48 of the last layer are padding and stay zero
    const int L = p / GUIDE_LAYER_PARAMS, off = p - L * GUIDE_LAYER_PARAMS;
    if (L != GUIDE_LAYERS - 1) return true;
    const int out = off < GUIDE_W * GUIDE_W ? off / GUIDE_W : off - GUIDE_W * GUIDE_W;
    return out < GUIDE_OUT;
}

__global__ void k_guide_pack(const float *params, __bf16 *w, __bf16 *wt, float *b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= GUIDE_PARAMS) return;
    storeParam(p, liveParam(p) ? params[p] : 0.f, w, wt, b);
}

// Sum of the per-slice partial gradients: 64 parameters per block, four slices of the
// partials per parameter (k = slice, slice + 4, ...), the slices added in a fixed order, so
// the result does not depend on scheduling.  The last block sums the per-block losses the
// same way (256 fixed strided partial sums, then a fixed tree).
__global__ __launch_bounds__(256) void k_guide_reduce(const float *partials, int nblocks, const float *lossPart, int nloss, float invBatch,
                                                      float *grad, float *lossOut) {
    __shared__ float part[4][64];
    __shared__ float lsum[256];
    if (blockIdx.x == gridDim.x - 1) {
        if (!lossOut) return;
        float l = 0.f;
        for (int k = threadIdx.x; k < nloss; k += 256) l += lossPart[k];
        lsum[threadIdx.x] = l;
        __syncthreads();
        for (int off = 128; off >= 1; off >>= 1) {
            if ((int)threadIdx.x < off) lsum[threadIdx.x] += lsum[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) lossOut[0] = lsum[0] * invBatch;
        return;
    }
    const int p = blockIdx.x * 64 + (threadIdx.x & 63);
    const int slice = threadIdx.x >> 6;
    float g = 0.f;
    if (p < GUIDE_PARAMS)
        for (int k = slice; k < nblocks; k += 4) g += partials[(size_t)k * GUIDE_PARAMS + p];
    part[slice][threadIdx.x & 63] = g;
    __syncthreads();
    if (slice == 0 && p < GUIDE_PARAMS) {
        const int j = threadIdx.x;
        const float t = (part[0][j] + part[1][j]) + (part[2][j] + part[3][j]);
        grad[p] = liveParam(p) ? t * invBatch : 0.f;
    }
}

__global__ void k_guide_adam(float *params, float *m, float *v, const float *grad, float c1, float c2, float lr, __bf16 *w, __bf16 *wt,
                             float *b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= GUIDE_PARAMS) return;
    if (!liveParam(p)) return;
    const float g = grad[p];
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    const float val = params[p] - lr * (mm * c1) / (sqrtf(vv * c2) + 1e-8f);
    params[p] = val;
    storeParam(p, val, w, wt, b);
}

__global__ void k_mfma_tile(const float *A, const float *B, float *C) {
    // one wave: C (32 x 32) = A (32 x 16) * B (16 x 32), all row-major FP32 in and out
    const int lane = threadIdx.x;
    if (lane >= 64) return;
    bf16x8 a, b;
    for (int i = 0; i < 8; i++) {
        const int k = (lane >> 5) * 8 + i;
        a[i] = (__bf16)A[(lane & 31) * 16 + k];
        b[i] = (__bf16)B[k * 32 + (lane & 31)];
    }
    f32x16 acc = {};
    acc = mfma(a, b, acc);
    for (int r = 0; r < 16; r++) C[cRow(lane, r) * 32 + (lane & 31)] = acc[r];
}

__global__ void k_mfma_tile16(const float *A, const float *B, float *C) {
    // one wave: C (16 x 16) = A (16 x 32) * B (32 x 16), all row-major FP32 in and out
    const int lane = threadIdx.x;
    if (lane >= 64) return;
    bf16x8 a, b;
    for (int i = 0; i < 8; i++) {
        const int k = (lane >> 4) * 8 + i;
        a[i] = (__bf16)A[(lane & 15) * 32 + k];
        b[i] = (__bf16)B[k * 16 + (lane & 15)];
    }
    f32x4 acc = {};
    acc = mfma16(a, b, acc);
    for (int r = 0; r < 4; r++) C[(4 * (lane >> 4) + r) * 16 + (lane & 15)] = acc[r];
}

}  // namespace pt

// ============================================================================================
// host side
// ============================================================================================
namespace pt {

// He-uniform hidden layers; the last layer starts near zero with biases that give 8 lobes
// along the cube diagonals, kappa = 2, and a selection probability of 0.18.
void guideInitParams(float *params, uint64_t seed) {
    std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ULL + 12345);
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    for (int L = 0; L < GUIDE_LAYERS; L++) {
        float *W = params + (size_t)L * GUIDE_LAYER_PARAMS;
        float *b = W + GUIDE_W * GUIDE_W;
        const float lim = L < GUIDE_LAYERS - 1 ? sqrtf(6.f / GUIDE_W) : 1e-3f;
        for (int o = 0; o < GUIDE_W; o++)
            for (int i = 0; i < GUIDE_W; i++) W[o * GUIDE_W + i] = (L == GUIDE_LAYERS - 1 && o >= GUIDE_OUT) ? 0.f : lim * u(rng);
        for (int o = 0; o < GUIDE_W; o++) b[o] = 0.f;
        if (L == GUIDE_LAYERS - 1) {
            for (int k = 0; k < GUIDE_LOBES; k++) {
                const float s = 0.57735027f;
                b[GO_MEAN + 3 * k] = (k & 1) ? s : -s;
                b[GO_MEAN + 3 * k + 1] = (k & 2) ? s : -s;
                b[GO_MEAN + 3 * k + 2] = (k & 4) ? s : -s;
                b[GO_KAPPA + k] = logf(2.f);
            }
            b[GO_FRAC] = -1.5f;
        }
    }
}

void guideUploadBf16(const float *dParams, __bf16 *dW, __bf16 *dWt, float *dB, hipStream_t s) {
    k_guide_pack<<<(GUIDE_PARAMS + 255) / 256, 256, 0, s>>>(dParams, dW, dWt, dB);
    PT_HIP_TRY(hipGetLastError());
}

void guideInfer(const PtSceneView &sc, GuideSceneNorm norm, GuideNet net, const int *segCount, int segStride, int segLen,
                const float4 *rayO, const float4 *rayD, float4 camO, const float4 *hits, float4 *feat, float4 *out, int cus, hipStream_t s) {
    const int maxRows = PT_QUEUE_SEGS * segLen;
    const int fgrid = std::max(1, std::min((maxRows + 255) / 256, cus * 8));
    k_guide_features<<<fgrid, 256, 0, s>>>(sc, norm, segCount, segStride, segLen, rayO, rayD, camO, hits, maxRows, feat);
    const int grid = std::max(1, std::min((maxRows + INFER_ROWS * INFER_WAVES - 1) / (INFER_ROWS * INFER_WAVES), cus));
    k_guide_infer<<<grid, 64 * INFER_WAVES, 0, s>>>(net, segCount, segStride, segLen, feat, maxRows, out, nullptr, nullptr);
    PT_HIP_TRY(hipGetLastError());
}

void guideInferRaw(GuideNet net, const float *feats, int n, float *rawOut, hipStream_t s) {
    const int grid = std::max(1, (n + INFER_ROWS * INFER_WAVES - 1) / (INFER_ROWS * INFER_WAVES));
    k_guide_infer<<<grid, 64 * INFER_WAVES, 0, s>>>(net, nullptr, 0, 64, nullptr, n, nullptr, feats, rawOut);
    PT_HIP_TRY(hipGetLastError());
}

void guideGradients(GuideNet net, const float4 *records, const int *count, const int *prefix, int recSeg, int base, int batch,
                    const float *wsum, const GuideTrainBufs &tb, float *grad, float *lossOut, hipStream_t s) {
    const int rows = guideRows(batch), slices = guideSlices(batch);
    k_guide_fwdbwd<<<tb.fwdBlocks, 64, 0, s>>>(net, records, count, prefix, recSeg, wsum, base, batch, rows, tb.act, tb.delta, tb.lossPart);
    k_guide_wgrad<<<dim3(slices, GUIDE_LAYERS), 256, 0, s>>>(tb.act, tb.delta, rows, tb.partials);
    k_guide_reduce<<<(GUIDE_PARAMS + 63) / 64 + 1, 256, 0, s>>>(tb.partials, slices, tb.lossPart, tb.fwdBlocks, 1.f / (float)std::max(batch, 1),
                                                               grad, lossOut);
    PT_HIP_TRY(hipGetLastError());
}

void guideAdam(float *params, float *m, float *v, const float *grad, int step, float lr, __bf16 *dW, __bf16 *dWt, float *dB, hipStream_t s) {
    const float c1 = 1.f / (1.f - powf(0.9f, (float)step));
    const float c2 = 1.f / (1.f - powf(0.999f, (float)step));
    k_guide_adam<<<(GUIDE_PARAMS + 255) / 256, 256, 0, s>>>(params, m, v, grad, c1, c2, lr, dW, dWt, dB);
    PT_HIP_TRY(hipGetLastError());
}

void mfmaTile(const float *A, const float *B, float *C, hipStream_t s) {
    k_mfma_tile<<<1, 64, 0, s>>>(A, B, C);
    PT_HIP_TRY(hipGetLastError());
}

void mfmaTile16(const float *A, const float *B, float *C, hipStream_t s) {
    k_mfma_tile16<<<1, 64, 0, s>>>(A, B, C);
    PT_HIP_TRY(hipGetLastError());
}

}  // namespace pt
