// pt_guide.hip -- BF16 MFMA kernels of the guiding network (gfx950).
//
// Every layer is 64 x 64 (the 48-output last layer is zero-padded to 64), weights and
// activations BF16, accumulation FP32.  Every layer product runs transposed, H^T = W A^T, on
// v_mfma_f32_16x16x32_bf16 with the weights as the A operand, so that a layer's output
// fragments are the next layer's input fragments and activations stay in registers (the
// K-permuted weight layout is described at kFeature below).
//
// Inference: blocks of 8 waves, each wave loops over 64-row tiles of the ray queue, encodes
// the rows' raw features one-blob straight into MFMA operands, runs the 5 layers and writes
// the mixture (8 vMF lobes + selection probability) per row.
// Training step: (1) one kernel runs forward, loss gradient, backward and the weight and bias
// gradients over 128-row iterations per block, each block keeping its share of the gradient in
// registers and writing it once as an FP32 partial; (2) the partials summed in block order
// (deterministic) and Adam on the FP32 master weights, refreshing the BF16 fragments.
//
// The 32x32x16 instruction is kept for its lane-layout check (pt_unit_mfma_tile): A: row
// (l & 31), 8 consecutive k from 8 (l >> 5); B: col (l & 31), the same k; C: element r at row
// 8 (r >> 2) + 4 (l >> 5) + (r & 3), col (l & 31).
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


#define WAVE_SYNC()                                         \
    do {                                                    \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
        __builtin_amdgcn_wave_barrier();                    \
    } while (0)

__device__ __forceinline__ int cRow(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

__device__ __forceinline__ bf16x8 ld8(const __bf16 *p) { return *(const bf16x8 *)p; }

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }

// v_mfma_f32_16x16x32_bf16, the network's instruction: A: row (l & 15), 8 consecutive k
// starting at 8 * (l >> 4) of the 32-wide K step; B: col (l & 15), the same 8 k; C: 4 floats
// per lane, element r at row 4 * (l >> 4) + r, col (l & 15) (checked on the device by
// pt_unit_mfma_tile16).
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));
__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

__device__ __forceinline__ float luminance(float4 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// exp(-x^2 / 2) with x = 8 s - (j + 1/2), as exp2(-t^2) with t = x sqrt(log2(e) / 2): one fma,
// one multiply and v_exp_f32.
__device__ __forceinline__ float oneBlob(float s, int j) {
    constexpr float c = 0.8493218002880191f;   // sqrt(log2(e) / 2)
    const float t = fmaf(s, (float)GUIDE_BINS * c, -((float)j + 0.5f) * c);
    return __builtin_amdgcn_exp2f(-t * t);
}

// ReLU of two FP32 values as one packed BF16 pair: convert, then one packed 16-bit integer max
// with 0 (a BF16 value is negative, or -0, exactly when its bit pattern is a negative int16).
typedef short short2v __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint32_t reluPackBf16(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 h = {(__bf16)a, (__bf16)b};
    const short2v r = __builtin_elementwise_max(__builtin_bit_cast(short2v, h), (short2v){0, 0});
    return __builtin_bit_cast(uint32_t, r);
}

// Sums and maxima over the four lanes that hold one record's outputs (lanes l, l ^ 16,
// l ^ 32, l ^ 48); all four must be active.
__device__ __forceinline__ float sum4(float v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}
__device__ __forceinline__ float max4(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}

// A lobe's parameters from its raw outputs: unit mean (or +y for a degenerate output, with
// ilen = 0) and concentration exp(clamp(log kappa)).
__device__ __forceinline__ void lobeFromRaw(float mx, float my, float mz, float lk, f3 &mu, float &il, float &kappa) {
    const f3 m = mk(mx, my, mz);
    const float l2 = dot(m, m);
    const bool ok = l2 > 1e-24f;
    il = ok ? __builtin_amdgcn_rsqf(l2) : 0.f;
    mu = ok ? m * il : mk(0.f, 1.f, 0.f);
    kappa = __expf(fminf(fmaxf(lk, GUIDE_LOGK_MIN), GUIDE_LOGK_MAX));
}

// One record's mixture from its last-layer outputs, four lanes per record: lane group g holds
// slots v[s] = output guideOutOf(16 (s >> 2) + 4 g + (s & 3)), i.e. lobes 2 g and 2 g + 1 (and
// group 3 the selection logit), and returns its 16-byte quarter of the mixture record: the
// two lobes' octahedral means and (kappa, weight) pairs, the selection probability in place
// of lobe 7's weight.  Softmax weights, unit means, kappa = exp(clamp(log kappa, -3, 9)),
// selection probability clamp(sigmoid, 0.02, 0.98).
__device__ __forceinline__ uint4 mixtureQuarter(const float *v, int g) {
    const float mxl = max4(fmaxf(v[0], v[1]));
    const float e0 = __expf(v[0] - mxl), e1 = __expf(v[1] - mxl);
    const float is = frcp(sum4(e0 + e1));
    f3 mu0, mu1;
    float il, k0, k1;
    lobeFromRaw(v[2], v[3], v[4], v[8], mu0, il, k0);
    lobeFromRaw(v[5], v[6], v[7], v[9], mu1, il, k1);
    const float c = fminf(fmaxf(frcp(1.f + __expf(-v[10])), GUIDE_FRAC_MIN), GUIDE_FRAC_MAX);
    return make_uint4(guideOctEncode(mu0), guidePackHalf2(k0, e0 * is), guideOctEncode(mu1), guidePackHalf2(k1, g == 3 ? c : e1 * is));
}

// von Mises-Fisher density on S^2: kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1))
__device__ __forceinline__ float vmfPdf(f3 mu, float kappa, f3 w) {
    return kappa * frcp(6.28318530718f * (1.f - __expf(-2.f * kappa))) * __expf(kappa * (dot(mu, w) - 1.f));
}

// --------------------------------------------------------------------------------------------
// inference
// --------------------------------------------------------------------------------------------
// Inference blocks: four waves share the LDS-staged weight fragments (40 KiB) and biases;
// each wave owns a 64-row tile at a time, with its activations and outputs in registers.
// Three blocks per CU: three waves per SIMD (at most 168 registers per lane).
constexpr int INFER_WAVES = 4;
constexpr int INFER_BLOCKS_PER_CU = 3;
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
// count[s * segStride], segments segLen slots long): two float4 per slot (raw features 0, 2,
// 4, 6, then 1, 3, 5, 7), the eighth feature (roughness, > 0) set to -1 where the vertex is not
// guided.  A kernel of its own, at full
// occupancy, so that the dependent gathers (hit -> primitive -> material) overlap.
__global__ __launch_bounds__(256) void k_guide_features(PtSceneView sc, GuideSceneNorm nm, const int *count, int segStride, int segLen,
                                                        const float4 *rayO, const float4 *rayD, float4 camO, const float4 *hits, int maxRows,
                                                        float4 *feat) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < maxRows; i += gridDim.x * blockDim.x) {
        int off;
        const int seg = divmodEst(i, segLen, &off);
        if (off >= count[seg * segStride]) continue;
        float f[GUIDE_RAW];
        for (int k = 0; k < GUIDE_RAW; k++) f[k] = 0.f;
        if (!vertexFeatures(sc, nm, i, rayO, rayD, camO, hits, f)) {
            for (int k = 0; k < GUIDE_RAW; k++) f[k] = 0.f;
            f[GUIDE_RAW - 1] = -1.f;
        }
        // even raw features, then odd ones: the inference lanes that encode raw features
        // 2 i (2 i + 1) of a row read one float4
        feat[2 * (size_t)i] = make_float4(f[0], f[2], f[4], f[6]);
        feat[2 * (size_t)i + 1] = make_float4(f[1], f[3], f[5], f[7]);
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

// The biases in LDS, the last layer's in guideOutOf order.
__device__ __forceinline__ void stageBiases(const float *b, float *sb) {
    for (int c = threadIdx.x; c < GUIDE_LAYERS * GUIDE_W; c += blockDim.x) {
        const int L = c / GUIDE_W, o = c % GUIDE_W;
        sb[c] = b[L * GUIDE_W + (L == GUIDE_LAYERS - 1 ? guideOutOf(o) : o)];
    }
}

// Rows are the slots [0, maxRows) of a segmented queue (count of segment s at
// count[s * segStride], segments segLen slots long, a multiple of 64, so that a tile of 64 rows
// lies in one segment) whose features k_guide_features wrote to feat, or with count = nullptr
// all maxRows rows of rawFeats.
__global__ __launch_bounds__(64 * INFER_WAVES, INFER_BLOCKS_PER_CU) void k_guide_infer(GuideNet net, const int *count, int segStride, int segLen, const float4 *feat,
                                                                  int maxRows, float4 *out, const float *rawFeats, float *rawOut) {
    __shared__ bf16x8 sWf[GUIDE_LAYERS * 8 * 64];
    __shared__ float4 sB[GUIDE_LAYERS * GUIDE_W / 4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, lr = lane & 15;
    for (int c = threadIdx.x; c < GUIDE_LAYERS * 8 * 64; c += blockDim.x) sWf[c] = ld8(net.w + 8 * (size_t)c);
    stageBiases(net.b, (float *)sB);
    __syncthreads();
    const int waveStride = gridDim.x * INFER_WAVES;
    // A tile's inputs: for each of this lane's four rows rowBase + 16 rb + lr the four raw
    // features it encodes (2 i + hi, i = 0..3: one float4 of k_guide_features' interleaved
    // layout; gathered from the row-major rawFeats), and the segment's record count.  The next
    // tile's are loaded while the current one runs through the network.
    const int hi = g >> 1;
    auto loadTile = [&](int tile, float4 *fv, int &cnt) {
        const int rowBase = tile * INFER_ROWS;
        cnt = count ? count[(rowBase / segLen) * segStride] : maxRows;
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const int row = rowBase + rb * 16 + lr;
            fv[rb] = make_float4(0.f, 0.f, 0.f, -1.f);
            if (row < maxRows) {
                if (rawFeats) {
                    const float *r = rawFeats + (size_t)row * GUIDE_RAW + hi;
                    fv[rb] = make_float4(r[0], r[2], r[4], r[6]);
                } else {
                    fv[rb] = feat[2 * (size_t)row + hi];
                }
            }
        }
    };
    int tile = blockIdx.x * INFER_WAVES + wave;
    float4 fv[4];
    int cnt = 0;
    if (tile * INFER_ROWS < maxRows) loadTile(tile, fv, cnt);
    for (; tile * INFER_ROWS < maxRows; tile += waveStride) {
        float4 nv[4];
        int ncnt = 0;
        if ((tile + waveStride) * INFER_ROWS < maxRows) loadTile(tile + waveStride, nv, ncnt);
        const int rowBase = tile * INFER_ROWS;
        // rows [rowBase, n) of this tile hold records: the filled part of its queue segment
        const int n = count ? (rowBase / segLen) * segLen + min(cnt, segLen) : maxRows;
        // guided rows per 16-row block: bit 32 + lr of each ballot (the lanes with hi = 1 hold
        // raw feature 7, the roughness, -1 where the vertex is not guided)
        uint64_t bal[4];
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
            bal[rb] = __ballot(rowBase + rb * 16 + lr < n && (rawFeats != nullptr || (hi && fv[rb].w >= 0.f)));
        const int mine = rowBase + lane;
        if (rowBase < n && (bal[0] | bal[1] | bal[2] | bal[3]) == 0) {
            // no guided vertex in the tile (misses, mirror and glass hits): only the flag that
            // makes shade sample the BSDF alone, no network evaluation
            if (mine < n && !rawOut) out[(size_t)mine * GUIDE_REC_F4 + 3] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (rowBase < n) {
            // layer 0's B fragments: one-blob encoding.  Input feature 8 k + j is bin j of raw
            // feature k; the lane's K positions need raw features 2 i + (g >> 1), i = 0..3, at bins
            // 4 (g & 1) + 0..3.  Rows that are not guided enter as zeros.
            bf16x8 act[4][2];
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                const bool on = (bal[rb] >> (32 + lr)) & 1u;
                const float f[4] = {on ? fv[rb].x : 0.f, on ? fv[rb].y : 0.f, on ? fv[rb].z : 0.f, on ? fv[rb].w : 0.f};
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
                        for (int ks = 0; ks < 2; ks++) {
                            u32x4 q;
#pragma unroll
                            for (int h = 0; h < 4; h++)
                                q[h] = reluPackBf16(acc[rb][2 * ks + (h >> 1)][2 * (h & 1)], acc[rb][2 * ks + (h >> 1)][2 * (h & 1) + 1]);
                            act[rb][ks] = __builtin_bit_cast(bf16x8, q);
                        }
                } else {
                    // the last layer's outputs: the four lanes l, l ^ 16, l ^ 32, l ^ 48 hold a
                    // row's outputs in guideOutOf order (two lobes each)
#pragma unroll
                    for (int rb = 0; rb < 4; rb++) {
                        float v[12];
#pragma unroll
                        for (int k = 0; k < 12; k++) v[k] = acc[rb][k >> 2][k & 3];
                        const int row = rowBase + rb * 16 + lr;
                        if (rawOut) {
                            if (row < n) {
#pragma unroll
                                for (int k = 0; k < 12; k++) rawOut[(size_t)row * GUIDE_OUT + guideOutOf(16 * (k >> 2) + 4 * g + (k & 3))] = v[k];
                            }
                        } else {
                            const uint4 q = mixtureQuarter(v, g);
                            if (row < n) {
                                // a vertex that is not guided gets only the zero selection probability
                                if ((bal[rb] >> (32 + lr)) & 1u) ((uint4 *)out)[(size_t)row * GUIDE_REC_F4 + g] = q;
                                else if (g == 3) ((uint4 *)out)[(size_t)row * GUIDE_REC_F4 + 3] = make_uint4(0u, 0u, 0u, 0u);
                            }
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int rb = 0; rb < 4; rb++) fv[rb] = nv[rb];
        cnt = ncnt;
    }
}

// --------------------------------------------------------------------------------------------
// training: gradients
// --------------------------------------------------------------------------------------------
// d loss / d last-layer outputs for one record, loss = -w (log q + log(c q + (1 - c) p_bsdf)),
// four lanes per record in the layout of mixtureQuarter: d[s] is the gradient of output slot
// s (0 where the loss does not depend on it; single outliers clamped to 1e3).  The mixture's
// softmax maximum and sum and its density are reduced across the four lanes; the selection
// logit (group 3, slot 10) is read from group 3.  Returns the loss in group 0, 0 in the others.
__device__ float lossGradientQ(const float *v, f3 w, float pb, float wt, int g, float *d) {
#pragma unroll
    for (int k = 0; k < 12; k++) d[k] = 0.f;
    const float fracRaw = __shfl(v[10], (int)(threadIdx.x & 15) + 48);
    if (!(wt > 0.f)) return 0.f;
    auto clampG = [](float x) { return fminf(fmaxf(x, -1e3f), 1e3f); };
    const float mxl = max4(fmaxf(v[0], v[1]));
    const float e0 = __expf(v[0] - mxl), e1 = __expf(v[1] - mxl);
    const float is = frcp(sum4(e0 + e1));
    float pi[2], kap[2], il[2], fk[2];
    f3 mu[2];
    lobeFromRaw(v[2], v[3], v[4], v[8], mu[0], il[0], kap[0]);
    lobeFromRaw(v[5], v[6], v[7], v[9], mu[1], il[1], kap[1]);
    pi[0] = e0 * is;
    pi[1] = e1 * is;
    float qp = 0.f;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        fk[j] = vmfPdf(mu[j], kap[j], w);
        qp += pi[j] * fk[j];
    }
    const float qq = fmaxf(sum4(qp), 1e-12f);
    const float craw = frcp(1.f + __expf(-fracRaw));
    const float c = fminf(fmaxf(craw, GUIDE_FRAC_MIN), GUIDE_FRAC_MAX);
    const float mix = fmaxf(c * qq + (1.f - c) * pb, 1e-12f);
    const float imix = frcp(mix);
    const float dq = -wt * (frcp(qq) + c * imix);
    if (g == 3 && craw > GUIDE_FRAC_MIN && craw < GUIDE_FRAC_MAX) d[10] = clampG(-wt * (qq - pb) * imix * craw * (1.f - craw));
#pragma unroll
    for (int j = 0; j < 2; j++) {
        d[j] = clampG(dq * pi[j] * (fk[j] - qq));   // softmax logit
        const float df = dq * pi[j] * fk[j];       // d loss / d log f_j
        const float cosw = dot(mu[j], w);
        // d log f / d kappa = 1/kappa - 2 e^{-2k} / (1 - e^{-2k}) + (mu.w - 1); kappa = exp(raw)
        const float e2 = __expf(-2.f * kap[j]);
        const float dlk = frcp(kap[j]) - 2.f * e2 * frcp(1.f - e2) + (cosw - 1.f);
        const float lk = v[8 + j];
        if (lk > GUIDE_LOGK_MIN && lk < GUIDE_LOGK_MAX) d[8 + j] = clampG(df * dlk * kap[j]);
        // d log f / d m = kappa (w - (mu.w) mu) / |m|
        if (il[j] > 0.f) {
            const float sc = df * kap[j] * il[j];
            d[2 + 3 * j] = clampG(sc * (w.x - cosw * mu[j].x));
            d[3 + 3 * j] = clampG(sc * (w.y - cosw * mu[j].y));
            d[4 + 3 * j] = clampG(sc * (w.z - cosw * mu[j].z));
        }
    }
    return g == 0 ? -wt * (__logf(qq) + __logf(mix)) : 0.f;
}

// Training step in one kernel: blocks of eight waves, persistent over iterations of
// GUIDE_TRAIN_ROWS records (16 rows per wave).  Per iteration each wave runs the forward pass
// of its rows with the activations in registers (as inference), the loss gradient at the
// output (four lanes per record), and the backward pass (the transposed product W^T delta^T
// with permuted W^T fragments, ReLU masks kept as bits in the fragment order).  At every
// layer, from the last to the first, the waves write the layer's input activations and output
// deltas of their rows to LDS feature-major, and each wave adds the block's 128 rows to its
// share of the weight gradient (output block w >> 1, input blocks 2 (w & 1) and 2 (w & 1) + 1
// of every layer), dW = delta^T a with K = rows on the MFMA; the bias gradient sums the same
// delta fragments.  Each block keeps its gradient in registers across iterations and writes it
// once as a partial, summed in block order by k_guide_reduce: the result does not depend on
// scheduling.  Rows past the batch get weight 0 and so zero deltas.
constexpr int TRAIN_WAVES = 8;
constexpr int TRAIN_TILE = 16;                        // rows per wave per iteration
constexpr int FM_PITCH = GUIDE_TRAIN_ROWS + 8;        // bf16 per feature of a feature-major tile
static_assert(TRAIN_WAVES * TRAIN_TILE == GUIDE_TRAIN_ROWS, "an iteration is one tile per wave");
static_assert(TRAIN_WAVES * 8 == GUIDE_W, "a 16 x 32 block of the weight gradient per wave");

struct TrainGradLds {   // the block's weight-gradient stage
    __bf16 a[GUIDE_W][FM_PITCH];
    __bf16 d[GUIDE_W][FM_PITCH];
};

__global__ __launch_bounds__(64 * TRAIN_WAVES) void k_guide_train(GuideNet net, const float4 *recs, const int *countPtr, const int *prefix,
                                                                  int recSeg, const float *wsumPtr, int base, int batch, float *partials,
                                                                  float *lossOut) {
    __shared__ bf16x8 sWf[GUIDE_LAYERS * 8 * 64];
    __shared__ bf16x8 sWb[(GUIDE_LAYERS - 1) * 8 * 64];   // layers 1 .. 4 (layer 0 passes no delta back)
    __shared__ float4 sB[GUIDE_LAYERS * GUIDE_W / 4];
    __shared__ float4 sRec[TRAIN_WAVES][TRAIN_TILE][GUIDE_TREC_F4];
    __shared__ float sLoss[TRAIN_WAVES];
    __shared__ __align__(16) TrainGradLds gs;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, lr = lane & 15;
    const int gob = wave >> 1, gib = 2 * (wave & 1);   // this wave's block of the weight gradient
    for (int c = threadIdx.x; c < GUIDE_LAYERS * 8 * 64; c += blockDim.x) sWf[c] = ld8(net.w + 8 * (size_t)c);
    for (int c = threadIdx.x; c < (GUIDE_LAYERS - 1) * 8 * 64; c += blockDim.x) sWb[c] = ld8(net.wt + 8 * (size_t)(8 * 64 + c));
    stageBiases(net.b, (float *)sB);
    __syncthreads();
    const int count = *countPtr;
    const float wsum = wsumPtr ? *wsumPtr : 0.f;
    // normalise the weights to mean one over the record buffer
    const float wscale = (wsumPtr == nullptr) ? 1.f : (wsum > 0.f ? (float)count / wsum : 0.f);
    f32x4 dW[GUIDE_LAYERS][2];
    float db[GUIDE_LAYERS];
#pragma unroll
    for (int L = 0; L < GUIDE_LAYERS; L++) {
        db[L] = 0.f;
        dW[L][0] = dW[L][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float lossAcc = 0.f;
    const int iters = (batch + GUIDE_TRAIN_ROWS - 1) / GUIDE_TRAIN_ROWS;
    for (int it = blockIdx.x; it < iters; it += gridDim.x) {
        {
            // the tile's records: lane (g, lr) loads float4 g of row lr
            const int r = it * GUIDE_TRAIN_ROWS + wave * TRAIN_TILE + lr;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r < batch && count > 0) {
                const int vi = (int)(((int64_t)base + r) % count);
                size_t idx = (size_t)vi;
                if (prefix) {
                    // the last segment whose prefix does not exceed vi holds record vi
                    int lo = 0, hi = PT_QUEUE_SEGS;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (prefix[mid] <= vi) lo = mid;
                        else hi = mid;
                    }
                    idx = (size_t)lo * recSeg + (vi - prefix[lo]);
                }
                v = recs[idx * GUIDE_TREC_F4 + g];
                if (g == 3) v.w = fminf(v.w * wscale, 64.f);
            }
            sRec[wave][lr][g] = v;
        }
        WAVE_SYNC();
        // forward; act[L] = the input of layer L as B fragments [K step]
        bf16x8 act[GUIDE_LAYERS][2];
        {
            // one-blob encoding as in the inference kernel
            const float4 a = sRec[wave][lr][0], b = sRec[wave][lr][1];
            const bool hi = (g >> 1) != 0;
            const float f[4] = {hi ? a.y : a.x, hi ? a.w : a.z, hi ? b.y : b.x, hi ? b.w : b.z};
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
#pragma unroll
                for (int p = 0; p < 8; p++) {
                    const int kf = kFeature(g, p, ks);
                    act[0][ks][p] = (__bf16)oneBlob(f[(kf >> 3) >> 1], kf & 7);
                }
        }
        uint32_t mask[GUIDE_LAYERS - 1];   // ReLU' of hidden layer L's outputs, bit ks * 8 + p
        float v[12];                       // this lane's last-layer outputs (guideOutOf order)
#pragma unroll
        for (int L = 0; L < GUIDE_LAYERS; L++) {
            const bool last = L == GUIDE_LAYERS - 1;
            f32x4 acc[4];
#pragma unroll
            for (int ob = 0; ob < 4; ob++) {
                if (last && ob * 16 >= GUIDE_OUT) continue;
                const float4 bias = sB[(L * GUIDE_W + ob * 16) / 4 + g];
                const bf16x8 w0 = sWf[((L * 4 + ob) * 2) * 64 + lane], w1 = sWf[((L * 4 + ob) * 2 + 1) * 64 + lane];
                const f32x4 c = {bias.x, bias.y, bias.z, bias.w};
                acc[ob] = mfma16(w1, act[L][1], mfma16(w0, act[L][0], c));
            }
            if (!last) {
                uint32_t m = 0;
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    u32x4 q;
#pragma unroll
                    for (int h = 0; h < 4; h++) {
                        q[h] = reluPackBf16(acc[2 * ks + (h >> 1)][2 * (h & 1)], acc[2 * ks + (h >> 1)][2 * (h & 1) + 1]);
                        // a positive BF16 value is a positive int16
                        m |= ((q[h] & 0xffffu) != 0u ? 1u : 0u) << (ks * 8 + 2 * h);
                        m |= ((q[h] >> 16) != 0u ? 1u : 0u) << (ks * 8 + 2 * h + 1);
                    }
                    act[L + 1][ks] = __builtin_bit_cast(bf16x8, q);
                }
                mask[L] = m;
            } else {
#pragma unroll
                for (int k = 0; k < 12; k++) v[k] = acc[k >> 2][k & 3];
            }
        }
        // the output deltas, in registers in the C-fragment layout of the last layer, as the
        // backward pass's B fragments (output block 3 is not computed: zero)
        bf16x8 dl[2];
        {
            const float4 c = sRec[wave][lr][2];
            float d[12];
            lossAcc += lossGradientQ(v, mk(c.x, c.y, c.z), c.w, sRec[wave][lr][3].w, g, d);
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
#pragma unroll
                for (int p = 0; p < 8; p++) {
                    const int ob = 2 * ks + (p >> 2);
                    dl[ks][p] = (__bf16)(ob < 3 ? d[4 * ob + (p & 3)] : 0.f);
                }
        }
#pragma unroll
        for (int L = GUIDE_LAYERS - 1; L >= 0; L--) {
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
#pragma unroll
                for (int p = 0; p < 8; p++) {
                    const int kf = kFeature(g, p, ks), col = wave * TRAIN_TILE + lr;
                    gs.a[kf][col] = act[L][ks][p];
                    gs.d[kf][col] = dl[ks][p];
                }
            if (L > 0) {
                // delta of layer L - 1's outputs: W_L^T delta_L, masked by ReLU'
                f32x4 bacc[4];
#pragma unroll
                for (int ib = 0; ib < 4; ib++) {
                    const bf16x8 w0 = sWb[(((L - 1) * 4 + ib) * 2) * 64 + lane], w1 = sWb[(((L - 1) * 4 + ib) * 2 + 1) * 64 + lane];
                    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
                    bacc[ib] = mfma16(w1, dl[1], mfma16(w0, dl[0], z));
                }
                const uint32_t m = mask[L - 1];
#pragma unroll
                for (int ks = 0; ks < 2; ks++)
#pragma unroll
                    for (int p = 0; p < 8; p++) {
                        const float v = bacc[2 * ks + (p >> 2)][p & 3];
                        dl[ks][p] = (__bf16)((m >> (ks * 8 + p)) & 1u ? v : 0.f);
                    }
            }
            __syncthreads();
            // this wave's block of dW_L += delta_L^T a_L over the block's rows
#pragma unroll
            for (int kk = 0; kk < GUIDE_TRAIN_ROWS / 32; kk++) {
                const bf16x8 ad = *(const bf16x8 *)&gs.d[gob * 16 + lr][32 * kk + 8 * g];
                if (gib == 0) {
#pragma unroll
                    for (int i = 0; i < 8; i++) db[L] += (float)ad[i];
                }
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const bf16x8 ba = *(const bf16x8 *)&gs.a[(gib + j) * 16 + lr][32 * kk + 8 * g];
                    dW[L][j] = mfma16(ad, ba, dW[L][j]);
                }
            }
            __syncthreads();
        }
    }
    // the block's partial: out = 16 gob + 4 g + r, in = 16 (gib + j) + lr
    float *P = partials + (size_t)blockIdx.x * GUIDE_PARAMS;
#pragma unroll
    for (int L = 0; L < GUIDE_LAYERS; L++) {
        float *PL = P + (size_t)L * GUIDE_LAYER_PARAMS;
        // the last layer's rows back from the kernels' output order
        auto row = [&](int op) { return L == GUIDE_LAYERS - 1 ? guideOutOf(op) : op; };
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) PL[row(gob * 16 + 4 * g + r) * GUIDE_W + (gib + j) * 16 + lr] = dW[L][j][r];
        if (gib == 0) {
            // lanes lr, lr + 16, lr + 32, lr + 48 summed complementary rows of output 16 gob + lr
            float b = db[L];
            b += __shfl_xor(b, 16);
            b += __shfl_xor(b, 32);
            if (g == 0) PL[GUIDE_W * GUIDE_W + row(gob * 16 + lr)] = b;
        }
    }
    // loss of this block: quarter 0 of each wave holds the partial sums, added in wave order
    for (int off = 32; off >= 1; off >>= 1) lossAcc += __shfl_down(lossAcc, off);
    if (lane == 0) sLoss[wave] = lossAcc;
    __syncthreads();
    if (threadIdx.x == 0 && lossOut) {
        float t = 0.f;
        for (int w = 0; w < TRAIN_WAVES; w++) t += sLoss[w];
        lossOut[blockIdx.x] = t;
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
        // the last layer's rows in the kernels' output order
        const int o = L == GUIDE_LAYERS - 1 ? guideOpOf(out) : out;
        w[wFragIndex(L, o, in)] = h;    // forward: W[out][in], K = in
        wt[wFragIndex(L, in, o)] = h;   // backward: W^T[in][out], K = out
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

// Sum of the per-block partial gradients: 64 parameters per block, REDUCE_SLICES slices of the
// partials per parameter (k = slice, slice + REDUCE_SLICES, ...), the slices added in a fixed
// tree, so the result does not depend on scheduling.  The last block sums the per-block losses
// the same way (fixed strided partial sums, then a fixed tree).
constexpr int REDUCE_SLICES = 16;
__global__ __launch_bounds__(64 * REDUCE_SLICES) void k_guide_reduce(const float *partials, int nblocks, const float *lossPart, int nloss,
                                                                    float invBatch, float *grad, float *lossOut) {
    __shared__ float part[REDUCE_SLICES][64];
    __shared__ float lsum[64 * REDUCE_SLICES];
    const int t = threadIdx.x;
    if (blockIdx.x == gridDim.x - 1) {
        if (!lossOut) return;
        float l = 0.f;
        for (int k = t; k < nloss; k += 64 * REDUCE_SLICES) l += lossPart[k];
        lsum[t] = l;
        __syncthreads();
        for (int off = 32 * REDUCE_SLICES; off >= 1; off >>= 1) {
            if (t < off) lsum[t] += lsum[t + off];
            __syncthreads();
        }
        if (t == 0) lossOut[0] = lsum[0] * invBatch;
        return;
    }
    const int p = blockIdx.x * 64 + (t & 63);
    const int slice = t >> 6;
    float g = 0.f;
    if (p < GUIDE_PARAMS)
        for (int k = slice; k < nblocks; k += REDUCE_SLICES) g += partials[(size_t)k * GUIDE_PARAMS + p];
    part[slice][t & 63] = g;
    __syncthreads();
    for (int half = REDUCE_SLICES / 2; half >= 1; half >>= 1) {
        if (slice < half) part[slice][t & 63] += part[slice + half][t & 63];
        __syncthreads();
    }
    if (slice == 0 && p < GUIDE_PARAMS) grad[p] = liveParam(p) ? part[0][t] * invBatch : 0.f;
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
    const int grid = std::max(1, std::min((maxRows + INFER_ROWS * INFER_WAVES - 1) / (INFER_ROWS * INFER_WAVES), cus * INFER_BLOCKS_PER_CU));
    k_guide_infer<<<grid, 64 * INFER_WAVES, 0, s>>>(net, segCount, segStride, segLen, feat, maxRows, out, nullptr, nullptr);
    PT_HIP_TRY(hipGetLastError());
}

void guideInferRecords(GuideNet net, const float *feats, int n, float4 *records, hipStream_t s) {
    const int grid = std::max(1, (n + INFER_ROWS * INFER_WAVES - 1) / (INFER_ROWS * INFER_WAVES));
    k_guide_infer<<<grid, 64 * INFER_WAVES, 0, s>>>(net, nullptr, 0, 64, nullptr, n, records, feats, nullptr);
    PT_HIP_TRY(hipGetLastError());
}

void guideInferRaw(GuideNet net, const float *feats, int n, float *rawOut, hipStream_t s) {
    const int grid = std::max(1, (n + INFER_ROWS * INFER_WAVES - 1) / (INFER_ROWS * INFER_WAVES));
    k_guide_infer<<<grid, 64 * INFER_WAVES, 0, s>>>(net, nullptr, 0, 64, nullptr, n, nullptr, feats, rawOut);
    PT_HIP_TRY(hipGetLastError());
}

void guideGradients(GuideNet net, const float4 *records, const int *count, const int *prefix, int recSeg, int base, int batch,
                    const float *wsum, const GuideTrainBufs &tb, float *grad, float *lossOut, hipStream_t s) {
    k_guide_train<<<tb.blocks, 64 * TRAIN_WAVES, 0, s>>>(net, records, count, prefix, recSeg, wsum, base, batch, tb.partials, tb.lossPart);
    k_guide_reduce<<<(GUIDE_PARAMS + 63) / 64 + 1, 64 * REDUCE_SLICES, 0, s>>>(tb.partials, tb.blocks, tb.lossPart, tb.blocks,
                                                                              1.f / (float)std::max(batch, 1), grad, lossOut);
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
