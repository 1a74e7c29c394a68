// pt_unit.hip -- unit-level GPU entry points used by the parity tests: single ray/triangle
// tests, BSDF eval/pdf/sample, environment tables and sampling, PCG32 streams.
//
// Built without fused multiply-adds (this pragma precedes every include, and the file is
// also compiled with -ffp-contract=off) so that the triangle tests follow the reference's
// operation order exactly.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "pt_device.h"
#include "pt_gpu.h"
#include "pt_guide.h"
#include "pt_internal.h"

namespace pt {
namespace {

// triangle.cu:13-34 Moeller-Trumbore in the reference's operation order
__device__ bool mtHit(f3 p0, f3 p1, f3 p2, f3 o, f3 d, float tMin, float tMax, float *tt, float *uu, float *vv) {
    const f3 e1 = p1 - p0;
    const f3 e2 = p2 - p0;
    const f3 s1 = cross(d, e2);
    const float div = dot(s1, e1);
    if (div == 0.f) return false;
    const float inv = 1.f / div;
    const f3 s = o - p0;
    const float b1 = dot(s, s1) * inv;
    if (b1 < 0.f || b1 > 1.f) return false;
    const f3 s2 = cross(s, e1);
    const float b2 = dot(d, s2) * inv;
    if (b2 < 0.f || (b1 + b2) > 1.f) return false;
    const float t = dot(e2, s2) * inv;
    if (t < tMin || t > tMax) return false;
    *tt = t;
    *uu = b1;
    *vv = b2;
    return true;
}

__global__ void k_tri_test(const float *tris, const float *rays, int64_t n, int mode, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = tris + i * 9;
    const float *r = rays + i * 8;
    const f3 p0 = mk(p[0], p[1], p[2]), p1 = mk(p[3], p[4], p[5]), p2 = mk(p[6], p[7], p[8]);
    const f3 o = mk(r[0], r[1], r[2]), d = mk(r[4], r[5], r[6]);
    float t = 0.f, u = 0.f, v = 0.f;
    bool h;
    if (mode == 0) {
        h = mtHit(p0, p1, p2, o, d, r[3], r[7], &t, &u, &v);
    } else {
        const WtRay w = wtSetup(o, d);
        h = wtHit(w, p0, p1, p2, r[3], r[7], &t, &u, &v);
    }
    out[i * 4] = h ? 1.f : 0.f;
    out[i * 4 + 1] = t;
    out[i * 4 + 2] = u;
    out[i * 4 + 3] = v;
}

__device__ PtMaterial unpackMat(const float *m) {
    PtMaterial x;
    x.a = make_float4(m[1], m[2], m[3], m[0]);
    const int t = (int)m[0];
    x.b = make_float4(0.f, 0.f, 0.f, t == PT_MAT_GLASS ? m[4] : m[5]);
    return x;
}

__global__ void k_bsdf_eval(const float *mat, const float *wo, const float *wi, int64_t n, float *f, float *pdf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PtMaterial m = unpackMat(mat + i * 6);
    const f3 o = mk(wo[i * 3], wo[i * 3 + 1], wo[i * 3 + 2]), w = mk(wi[i * 3], wi[i * 3 + 1], wi[i * 3 + 2]);
    const f3 v = bsdfF(m, o, w);
    f[i * 3] = v.x;
    f[i * 3 + 1] = v.y;
    f[i * 3 + 2] = v.z;
    pdf[i] = bsdfPdf(m, o, w);
}

__global__ void k_bsdf_sample(const float *mat, const float *wo, const float *xi, int64_t n, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PtMaterial m = unpackMat(mat + i * 6);
    const BsdfSample s = bsdfSample(m, mk(wo[i * 3], wo[i * 3 + 1], wo[i * 3 + 2]), xi[i * 2], xi[i * 2 + 1]);
    float *r = out + i * 8;
    r[0] = s.wi.x; r[1] = s.wi.y; r[2] = s.wi.z;
    r[3] = s.pdf;
    r[4] = s.f.x; r[5] = s.f.y; r[6] = s.f.z;
    r[7] = s.delta ? 1.f : 0.f;
}

__global__ void k_env_sample(PtEnv env, const float *xi, int64_t n, float *dir, float *pdf, float *emit, const float *pdirs, int64_t m, float *pout) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float p;
        f3 e;
        const f3 w = envSample(env, xi[2 * i], xi[2 * i + 1], &p, &e);
        dir[3 * i] = w.x; dir[3 * i + 1] = w.y; dir[3 * i + 2] = w.z;
        pdf[i] = p;
        emit[3 * i] = e.x; emit[3 * i + 1] = e.y; emit[3 * i + 2] = e.z;
    }
    if (i < m) pout[i] = envPdf(env, mk(pdirs[3 * i], pdirs[3 * i + 1], pdirs[3 * i + 2]));
}

__global__ void k_path_uniforms(uint64_t seed, const uint32_t *pix, const uint32_t *smp, int64_t n, int k, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng r = seedRng(seed, pix[i], smp[i]);
    for (int j = 0; j < k; j++) out[i * k + j] = r.uniform();
}

__global__ void k_vmf_sample(const float *mu, const float *kappa, const float *xi, int64_t n, float *dir, float *pdf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 m = mk(mu[3 * i], mu[3 * i + 1], mu[3 * i + 2]);
    const f3 w = sampleVmf(m, kappa[i], xi[2 * i], xi[2 * i + 1]);
    dir[3 * i] = w.x;
    dir[3 * i + 1] = w.y;
    dir[3 * i + 2] = w.z;
    pdf[i] = vmfDensity(m, kappa[i], w);
}

// framebuffer.cu:35-40 bytes of a radiance buffer, through the renderer's device function
__global__ void k_tonemap_bytes(const float *radiance, int64_t n, uchar4 *rgba) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rgba[i] = make_uchar4(srgbByte(radiance[3 * i]), srgbByte(radiance[3 * i + 1]), srgbByte(radiance[3 * i + 2]), 255);
}

// Small RAII helper for the unit calls: device copies of host arrays on a private stream.
struct Stage {
    hipStream_t s = nullptr;
    std::vector<DevBuf *> bufs;
    Stage() { PT_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Stage() {
        if (s) (void)hipStreamSynchronize(s);
        for (DevBuf *b : bufs) delete b;
        if (s) (void)hipStreamDestroy(s);
    }
    template <class T> T *in(const T *h, size_t n) {
        DevBuf *b = new DevBuf();
        bufs.push_back(b);
        b->alloc(std::max<size_t>(n, 1) * sizeof(T));
        if (n) PT_HIP_TRY(hipMemcpyAsync(b->p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
        return b->as<T>();
    }
    template <class T> T *out(size_t n) {
        DevBuf *b = new DevBuf();
        bufs.push_back(b);
        b->alloc(std::max<size_t>(n, 1) * sizeof(T));
        return b->as<T>();
    }
    template <class T> void back(T *h, const T *d, size_t n) {
        if (n) PT_HIP_TRY(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void sync() { PT_HIP_TRY(hipStreamSynchronize(s)); }
};

inline unsigned nblk(int64_t n) { return (unsigned)((std::max<int64_t>(n, 1) + 255) / 256); }

int hasDevice(const char *fn) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(PT_ERR_DEVICE, std::string(fn) + ": no HIP device");
    return PT_OK;
}

}  // namespace
}  // namespace pt

using namespace pt;

#define PT_UNIT_BEGIN(fn) \
    if (hasDevice(fn)) return PT_ERR_DEVICE; \
    try {
#define PT_UNIT_END(fn)                                                       \
    }                                                                         \
    catch (const HipFailure &e) {                                             \
        return fail(PT_ERR_DEVICE, std::string(fn) + ": " + e.msg);           \
    }                                                                         \
    catch (const std::bad_alloc &) {                                          \
        return fail(PT_ERR_MEMORY, std::string(fn) + ": out of host memory"); \
    }

extern "C" {

int pt_unit_triangle_test(const float *tris, const float *rays, int64_t n, int mode, float *out) {
    if (n < 0 || (n > 0 && (!tris || !rays || !out)) || (mode != 0 && mode != 1)) return fail(PT_ERR_INVALID, "pt_unit_triangle_test: bad arguments");
    PT_UNIT_BEGIN("pt_unit_triangle_test")
    Stage st;
    const float *dt = st.in(tris, (size_t)n * 9), *dr = st.in(rays, (size_t)n * 8);
    float *dout = st.out<float>((size_t)n * 4);
    k_tri_test<<<nblk(n), 256, 0, st.s>>>(dt, dr, n, mode, dout);
    PT_HIP_TRY(hipGetLastError());
    st.back(out, dout, (size_t)n * 4);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_triangle_test")
}

int pt_unit_bsdf_eval(const float *mat, const float *wo, const float *wi, int64_t n, float *f, float *pdf) {
    if (n < 0 || (n > 0 && (!mat || !wo || !wi || !f || !pdf))) return fail(PT_ERR_INVALID, "pt_unit_bsdf_eval: bad arguments");
    PT_UNIT_BEGIN("pt_unit_bsdf_eval")
    Stage st;
    const float *dm = st.in(mat, (size_t)n * 6), *dwo = st.in(wo, (size_t)n * 3), *dwi = st.in(wi, (size_t)n * 3);
    float *df = st.out<float>((size_t)n * 3), *dp = st.out<float>((size_t)n);
    k_bsdf_eval<<<nblk(n), 256, 0, st.s>>>(dm, dwo, dwi, n, df, dp);
    PT_HIP_TRY(hipGetLastError());
    st.back(f, df, (size_t)n * 3);
    st.back(pdf, dp, (size_t)n);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_bsdf_eval")
}

int pt_unit_bsdf_sample(const float *mat, const float *wo, const float *xi, int64_t n, float *out) {
    if (n < 0 || (n > 0 && (!mat || !wo || !xi || !out))) return fail(PT_ERR_INVALID, "pt_unit_bsdf_sample: bad arguments");
    PT_UNIT_BEGIN("pt_unit_bsdf_sample")
    Stage st;
    const float *dm = st.in(mat, (size_t)n * 6), *dwo = st.in(wo, (size_t)n * 3), *dxi = st.in(xi, (size_t)n * 2);
    float *dout = st.out<float>((size_t)n * 8);
    k_bsdf_sample<<<nblk(n), 256, 0, st.s>>>(dm, dwo, dxi, n, dout);
    PT_HIP_TRY(hipGetLastError());
    st.back(out, dout, (size_t)n * 8);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_bsdf_sample")
}

int pt_unit_env_tables(const float *rgba, int w, int h, int sinw, double *rowProb, int *rowAlias, double *margProb, int *margAlias, float *texelPdf) {
    if (!rgba || w <= 0 || h <= 0) return fail(PT_ERR_INVALID, "pt_unit_env_tables: bad arguments");
    PT_UNIT_BEGIN("pt_unit_env_tables")
    Stage st;
    const size_t ne = (size_t)w * h;
    const float4 *drgba = (const float4 *)st.in(rgba, ne * 4);
    float2 *rows = st.out<float2>(ne), *marg = st.out<float2>((size_t)h);
    float *tp = st.out<float>(ne);
    buildEnvTables(drgba, w, h, sinw, rows, marg, tp, rowProb, rowAlias, margProb, margAlias, st.s);
    if (texelPdf) st.back(texelPdf, tp, ne);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_env_tables")
}

int pt_unit_env_sample(const float *rgba, int w, int h, int sinw, const float *xi, int64_t n, float *dir, float *pdf, float *emit,
                       const float *pdf_dirs, int64_t m, float *pdf_out) {
    if (!rgba || w <= 0 || h <= 0 || n < 0 || m < 0 || (n > 0 && (!xi || !dir || !pdf || !emit)) || (m > 0 && (!pdf_dirs || !pdf_out)))
        return fail(PT_ERR_INVALID, "pt_unit_env_sample: bad arguments");
    PT_UNIT_BEGIN("pt_unit_env_sample")
    Stage st;
    const size_t ne = (size_t)w * h;
    const float4 *drgba = (const float4 *)st.in(rgba, ne * 4);
    float2 *rows = st.out<float2>(ne), *marg = st.out<float2>((size_t)h);
    float *tp = st.out<float>(ne);
    buildEnvTables(drgba, w, h, sinw, rows, marg, tp, nullptr, nullptr, nullptr, nullptr, st.s);
    packEnvPdf((float4 *)drgba, tp, (int)ne, st.s);
    PtEnv env;
    env.rgba = drgba;
    env.rows = rows;
    env.marginal = marg;
    env.width = w;
    env.height = h;
    env.present = 1;
    const float *dxi = st.in(xi, (size_t)n * 2), *dpd = st.in(pdf_dirs, (size_t)m * 3);
    float *ddir = st.out<float>((size_t)n * 3), *dpdf = st.out<float>((size_t)n), *dem = st.out<float>((size_t)n * 3), *dpo = st.out<float>((size_t)m);
    k_env_sample<<<nblk(std::max(n, m)), 256, 0, st.s>>>(env, dxi, n, ddir, dpdf, dem, dpd, m, dpo);
    PT_HIP_TRY(hipGetLastError());
    st.back(dir, ddir, (size_t)n * 3);
    st.back(pdf, dpdf, (size_t)n);
    st.back(emit, dem, (size_t)n * 3);
    st.back(pdf_out, dpo, (size_t)m);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_env_sample")
}

int pt_unit_path_uniforms(uint64_t seed, const uint32_t *pixel, const uint32_t *sample, int64_t n, int k, float *out) {
    if (n < 0 || k < 0 || (n > 0 && (!pixel || !sample || !out))) return fail(PT_ERR_INVALID, "pt_unit_path_uniforms: bad arguments");
    PT_UNIT_BEGIN("pt_unit_path_uniforms")
    Stage st;
    const uint32_t *dp = st.in(pixel, (size_t)n), *ds = st.in(sample, (size_t)n);
    float *dout = st.out<float>((size_t)n * k);
    k_path_uniforms<<<nblk(n), 256, 0, st.s>>>(seed, dp, ds, n, k, dout);
    PT_HIP_TRY(hipGetLastError());
    st.back(out, dout, (size_t)n * k);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_path_uniforms")
}

int pt_unit_tonemap(const float *radiance, int64_t n, uint8_t *rgba) {
    if (n < 0 || (n > 0 && (!radiance || !rgba))) return fail(PT_ERR_INVALID, "pt_unit_tonemap: bad arguments");
    PT_UNIT_BEGIN("pt_unit_tonemap")
    Stage st;
    const float *dr = st.in(radiance, (size_t)n * 3);
    uchar4 *dout = st.out<uchar4>((size_t)n);
    k_tonemap_bytes<<<nblk(n), 256, 0, st.s>>>(dr, n, dout);
    PT_HIP_TRY(hipGetLastError());
    st.back((uchar4 *)rgba, dout, (size_t)n);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_tonemap")
}

int pt_unit_vmf_sample(const float *mu, const float *kappa, const float *xi, int64_t n, float *dir, float *pdf) {
    if (n < 0 || (n > 0 && (!mu || !kappa || !xi || !dir || !pdf))) return fail(PT_ERR_INVALID, "pt_unit_vmf_sample: bad arguments");
    for (int64_t i = 0; i < n; i++)
        if (!(kappa[i] > 0.f) || std::isinf(kappa[i])) return fail(PT_ERR_INVALID, "pt_unit_vmf_sample: kappa must be positive and finite");
    PT_UNIT_BEGIN("pt_unit_vmf_sample")
    Stage st;
    const float *dm = st.in(mu, (size_t)n * 3), *dk = st.in(kappa, (size_t)n), *dx = st.in(xi, (size_t)n * 2);
    float *dd = st.out<float>((size_t)n * 3), *dp = st.out<float>((size_t)n);
    k_vmf_sample<<<nblk(n), 256, 0, st.s>>>(dm, dk, dx, n, dd, dp);
    PT_HIP_TRY(hipGetLastError());
    st.back(dir, dd, (size_t)n * 3);
    st.back(pdf, dp, (size_t)n);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_vmf_sample")
}

namespace {
// BF16 copies of a parameter vector on the device
struct NetStage {
    DevBuf params, w, wt, b;
    GuideNet net;
    NetStage(const float *host, hipStream_t s) {
        params.alloc(GUIDE_PARAMS * sizeof(float));
        w.alloc(GUIDE_LAYERS * GUIDE_W * GUIDE_W * sizeof(__bf16));
        wt.alloc(GUIDE_LAYERS * GUIDE_W * GUIDE_W * sizeof(__bf16));
        b.alloc(GUIDE_LAYERS * GUIDE_W * sizeof(float));
        PT_HIP_TRY(hipMemcpyAsync(params.p, host, GUIDE_PARAMS * sizeof(float), hipMemcpyHostToDevice, s));
        guideUploadBf16(params.as<float>(), w.as<__bf16>(), wt.as<__bf16>(), b.as<float>(), s);
        net.w = w.as<__bf16>();
        net.wt = wt.as<__bf16>();
        net.b = b.as<float>();
    }
};
}  // namespace

int pt_unit_guide_forward(const float *params, const float *features, int64_t n, float *out) {
    if (!params || n < 0 || (n > 0 && (!features || !out)) || n > (1 << 24)) return fail(PT_ERR_INVALID, "pt_unit_guide_forward: bad arguments");
    PT_UNIT_BEGIN("pt_unit_guide_forward")
    Stage st;
    NetStage ns(params, st.s);
    const float *df = st.in(features, (size_t)n * GUIDE_RAW);
    float *dout = st.out<float>((size_t)n * GUIDE_OUT);
    if (n > 0) guideInferRaw(ns.net, df, (int)n, dout, st.s);
    st.back(out, dout, (size_t)n * GUIDE_OUT);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_guide_forward")
}

int pt_unit_guide_records(const float *params, const float *features, int64_t n, uint32_t *records) {
    if (!params || n < 0 || (n > 0 && (!features || !records)) || n > (1 << 24)) return fail(PT_ERR_INVALID, "pt_unit_guide_records: bad arguments");
    PT_UNIT_BEGIN("pt_unit_guide_records")
    Stage st;
    NetStage ns(params, st.s);
    const float *df = st.in(features, (size_t)n * GUIDE_RAW);
    float4 *dout = st.out<float4>((size_t)n * GUIDE_REC_F4);
    if (n > 0) guideInferRecords(ns.net, df, (int)n, dout, st.s);
    st.back((float4 *)records, dout, (size_t)n * GUIDE_REC_F4);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_guide_records")
}

int pt_unit_guide_gradients(const float *params, const float *records, int64_t n, float *grad, float *loss) {
    if (!params || !grad || n <= 0 || !records || n > (1 << 24)) return fail(PT_ERR_INVALID, "pt_unit_guide_gradients: bad arguments");
    PT_UNIT_BEGIN("pt_unit_guide_gradients")
    Stage st;
    NetStage ns(params, st.s);
    const float4 *dr = (const float4 *)st.in(records, (size_t)n * 16);
    const int cnt = (int)n;
    const int *dcount = st.in(&cnt, 1);
    GuideTrainBufs tb;
    tb.blocks = guideTrainBlocks(cnt, deviceCUs());
    tb.partials = st.out<float>((size_t)tb.blocks * GUIDE_PARAMS);
    tb.lossPart = st.out<float>((size_t)tb.blocks);
    float *dgrad = st.out<float>(GUIDE_PARAMS);
    float *dloss = st.out<float>(1);
    guideGradients(ns.net, dr, dcount, nullptr, 0, 0, cnt, nullptr, tb, dgrad, dloss, st.s);
    st.back(grad, dgrad, GUIDE_PARAMS);
    if (loss) st.back(loss, dloss, 1);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_guide_gradients")
}

int pt_unit_guide_adam(float *params, float *m, float *v, const float *grad, int step, float lr) {
    if (!params || !m || !v || !grad || step < 1) return fail(PT_ERR_INVALID, "pt_unit_guide_adam: bad arguments");
    PT_UNIT_BEGIN("pt_unit_guide_adam")
    Stage st;
    NetStage ns(params, st.s);
    float *dm = st.in(m, GUIDE_PARAMS), *dv = st.in(v, GUIDE_PARAMS);
    const float *dg = st.in(grad, GUIDE_PARAMS);
    guideAdam(ns.params.as<float>(), dm, dv, dg, step, lr, ns.w.as<__bf16>(), ns.wt.as<__bf16>(), ns.b.as<float>(), st.s);
    st.back(params, ns.params.as<float>(), GUIDE_PARAMS);
    st.back(m, dm, GUIDE_PARAMS);
    st.back(v, dv, GUIDE_PARAMS);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_guide_adam")
}

int pt_unit_mfma_tile(const float *a, const float *b, float *c) {
    if (!a || !b || !c) return fail(PT_ERR_INVALID, "pt_unit_mfma_tile: bad arguments");
    PT_UNIT_BEGIN("pt_unit_mfma_tile")
    Stage st;
    const float *da = st.in(a, 32 * 16), *db = st.in(b, 16 * 32);
    float *dc = st.out<float>(32 * 32);
    mfmaTile(da, db, dc, st.s);
    st.back(c, dc, 32 * 32);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_mfma_tile")
}

int pt_unit_mfma_tile16(const float *a, const float *b, float *c) {
    if (!a || !b || !c) return fail(PT_ERR_INVALID, "pt_unit_mfma_tile16: bad arguments");
    PT_UNIT_BEGIN("pt_unit_mfma_tile16")
    Stage st;
    const float *da = st.in(a, 16 * 32), *db = st.in(b, 32 * 16);
    float *dc = st.out<float>(16 * 16);
    mfmaTile16(da, db, dc, st.s);
    st.back(c, dc, 16 * 16);
    st.sync();
    return PT_OK;
    PT_UNIT_END("pt_unit_mfma_tile16")
}

}  // extern "C"
