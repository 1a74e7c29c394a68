// pt_device.h -- device functions of the wavefront path tracer (gfx950).
//
// Shading, sampling and intersection maths of one path vertex, written for the GPU: plain
// float registers, no virtual dispatch, material parameters packed in two float4s.  The
// semantics follow the reference renderer (file:line cited per function); where the
// product differs on purpose (merged BSDF/continuation ray, alias-table environment
// sampling, watertight triangles, normalised shading normals) the comment says so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_types.h"

namespace pt {

#define PT_INLINE __device__ __forceinline__

struct f3 {
    float x, y, z;
};
PT_INLINE f3 mk(float x, float y, float z) { return f3{x, y, z}; }
PT_INLINE f3 mk(float a) { return f3{a, a, a}; }
PT_INLINE f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_INLINE f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_INLINE f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
PT_INLINE f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
PT_INLINE f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
PT_INLINE f3 operator*(float s, f3 a) { return mk(a.x * s, a.y * s, a.z * s); }
PT_INLINE f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
PT_INLINE float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_INLINE f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x); }
// Shading arithmetic uses the hardware reciprocal, square root and reciprocal square root
// (v_rcp_f32, v_sqrt_f32, v_rsq_f32: within 1 ulp) instead of the correctly rounded quotient
// and root (about 10 and 16 instructions each, most of a shading vertex's instructions).  Two
// places keep IEEE arithmetic: the intersection tests (wtSetup, wtHit, sphereHit), whose
// decisions are compared bit for bit, and the Beckmann half vector and exponent (mfD), where
// exp(-tan^2 / alpha^2) at small alpha turns a last-bit difference into a relative error of
// 1e-4 and more.
PT_INLINE float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
PT_INLINE float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
PT_INLINE float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
PT_INLINE float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
constexpr float kInvPi = 0.318309886183790671538f;
PT_INLINE float length(f3 a) { return fsqrt(dot(a, a)); }
PT_INLINE f3 normalize(f3 a) { return a * frsq(dot(a, a)); }
PT_INLINE f3 normalizeIeee(f3 a) { return a / sqrtf(dot(a, a)); }

// n / d and n % d for 0 <= n < 2^31, d > 0, from a float estimate with the float reciprocal
// of d, corrected to the exact integer result (the estimate is off by at most a few units for
// quotients below 2^26): a per-lane integer division otherwise costs about 40 instructions.
PT_INLINE int divmodEst(int n, int d, int *r) {
    int q = (int)((float)n * __builtin_amdgcn_rcpf((float)d));
    int rr = n - q * d;
    while (rr < 0) {
        q--;
        rr += d;
    }
    while (rr >= d) {
        q++;
        rr -= d;
    }
    *r = rr;
    return q;
}
PT_INLINE bool isZero(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
PT_INLINE f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }
PT_INLINE float4 f4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
PT_INLINE float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
// vec3.h:57-59
PT_INLINE f3 reflect(f3 v, f3 n) { return n * (dot(v, n) * 2.f) - v; }

// ------------------------------------------------------------------------------------------
// PCG32 (O'Neill 2014).  One stream per (pixel, sample): inc = 2*splitmix64(stream)+1 with
// stream = sample << 32 | pixel; state seeded as pbrt-v3's RNG::SetSequence.  The state
// (8 bytes) lives in the path record; the increment is re-derived from the path id.
// ------------------------------------------------------------------------------------------
PT_INLINE uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
struct Rng {
    uint64_t state, inc;
    PT_INLINE uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    PT_INLINE float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};
PT_INLINE uint64_t streamInc(uint64_t seed, uint32_t pixel, uint32_t sample) {
    (void)seed;
    const uint64_t stream = ((uint64_t)sample << 32) | (uint64_t)pixel;
    return (splitmix64(stream) << 1u) | 1u;
}
PT_INLINE Rng seedRng(uint64_t seed, uint32_t pixel, uint32_t sample) {
    Rng r;
    r.inc = streamInc(seed, pixel, sample);
    r.state = 0u;
    r.next();
    r.state += 0x853c49e6748fea9bULL ^ seed;
    r.next();
    return r;
}

// ------------------------------------------------------------------------------------------
// Orthonormal frame around n (frame.h:7-18, pbrt's coordinateSystem)
// ------------------------------------------------------------------------------------------
struct Frame {
    f3 s, t, n;
};
PT_INLINE Frame makeFrame(f3 n) {
    Frame f;
    f.n = n;
    if (fabsf(n.x) > fabsf(n.y)) {
        const float inv = frsq(n.x * n.x + n.z * n.z);
        f.t = mk(n.z * inv, 0.f, -n.x * inv);
    } else {
        const float inv = frsq(n.y * n.y + n.z * n.z);
        f.t = mk(0.f, n.z * inv, -n.y * inv);
    }
    f.s = cross(f.t, n);
    return f;
}
PT_INLINE f3 toWorld(const Frame &f, f3 l) {
    return mk(f.s.x * l.x + f.t.x * l.y + f.n.x * l.z, f.s.y * l.x + f.t.y * l.y + f.n.y * l.z,
              f.s.z * l.x + f.t.z * l.y + f.n.z * l.z);
}
PT_INLINE f3 toLocal(const Frame &f, f3 w) { return mk(dot(f.s, w), dot(f.t, w), dot(f.n, w)); }

// ------------------------------------------------------------------------------------------
// BSDFs.  Material record: a = (albedo.rgb, type), b = (emit.rgb, ior or alpha).
// ------------------------------------------------------------------------------------------
PT_INLINE int matType(const PtMaterial &m) { return (int)m.a.w; }

struct BsdfSample {
    f3 wi;
    float pdf;
    f3 f;
    bool delta;
};

// sample.h:7-17 uniform hemisphere (z = xi1); the azimuth 2 pi xi2 through sincospi, whose
// argument reduction is exact (sincos of a radian angle reduces by a multi-word 2 pi)
PT_INLINE f3 uniformHemisphere(float xi1, float xi2) {
    const float z = xi1;
    const float r = fsqrt(fmaxf(0.f, 1.f - z * z));
    float sp, cp;
    sincospif(2.f * xi2, &sp, &cp);
    return mk(r * cp, r * sp, z);
}
// sample.h:26-37
PT_INLINE f3 uniformSphere(float xi1, float xi2) {
    const float z = xi1 * 2.f - 1.f;
    const float r = fsqrt(fmaxf(0.f, 1.f - z * z));
    float sp, cp;
    sincospif(2.f * xi2, &sp, &cp);
    return mk(r * cp, r * sp, z);
}

// fresnel.h:17-40
PT_INLINE float fresnelDielectric(float cosI, float etaI, float etaT) {
    const float sinT = fdiv(etaI, etaT) * fsqrt(fmaxf(0.f, 1.f - cosI * cosI));
    if (sinT > 1.f) return 1.f;
    const float cosT = fsqrt(fmaxf(0.f, 1.f - sinT * sinT));
    const float rPar = fdiv(etaT * cosI - etaI * cosT, etaT * cosI + etaI * cosT);
    const float rPer = fdiv(etaI * cosI - etaT * cosT, etaI * cosI + etaT * cosT);
    return 0.5f * (rPar * rPar + rPer * rPer);
}
// snell.h:10-43 (including its second eta swap for vectors below the normal)
PT_INLINE f3 snellRefract(f3 wi, float etaI, float etaT) {
    f3 n = mk(0.f, 0.f, 1.f);
    if (wi.z < 0.f) {
        n = mk(0.f, 0.f, -1.f);
        const float t = etaI; etaI = etaT; etaT = t;
    }
    const f3 perpI = wi - n * dot(wi, n);
    const f3 perpT = -perpI * fdiv(etaI, etaT);
    const float l = length(perpT);
    const float parLen = fsqrt(fmaxf(0.f, 1.f - l * l));
    return normalize(n * -parLen + perpT);
}

// microfacet.h:70-72 widened alpha
PT_INLINE float mfAlpha(float alpha, float absCos) { return (1.2f - 0.2f * sqrtf(absCos)) * alpha; }
// microfacet.h:37-51 Beckmann D with the widened alpha at |cos w|
PT_INLINE float mfD(float alpha, f3 w, f3 wh) {
    const float a = mfAlpha(alpha, fabsf(w.z));
    const float a2 = a * a;
    const float c2 = wh.z * wh.z;
    const float t2 = (1.f - c2) / c2;
    if (isinf(t2)) return 0.f;
    return expf(-t2 / a2) * frcp(3.14159265358979f * a2 * c2 * c2);
}
// microfacet.h:74-84 Walter 2007 rational G1
PT_INLINE float mfG1(float alpha, f3 v) {
    const float sinT = fsqrt(fmaxf(0.f, 1.f - v.z * v.z));
    const float a = frcp(alpha * fabsf(fdiv(sinT, v.z)));
    if (a >= 1.6f) return 1.f;
    return fdiv(3.535f * a + 2.181f * a * a, 1.f + 2.276f * a + 2.577f * a * a);
}
PT_INLINE float signOf(float x) { return x < 0.f ? -1.f : 1.f; }
// microfacet.h:98-122
PT_INLINE f3 microfacetF(float alpha, f3 wo, f3 wi) {
    if (wo.z < 0.f || wi.z < 0.f) return mk(0.f);
    const float cosO = fabsf(wo.z), cosI = fabsf(wi.z);
    if (cosO == 0.f || cosI == 0.f) return mk(0.f);
    const f3 sum = wo + wi;
    if (isZero(sum)) return mk(0.f);
    const f3 wh = normalizeIeee(sum);
    if (signOf(dot(wo, wh)) != signOf(wo.z) || signOf(dot(wi, wh)) != signOf(wi.z)) return mk(0.f);
    const float cosInc = fminf(1.f, fmaxf(dot(wi, wh), 0.f));
    const float F = fresnelDielectric(cosInc, 1.f, 1.5f);
    const float D = mfD(alpha, wi, wh);
    const float G = mfG1(alpha, wi) * mfG1(alpha, wo);
    return mk(fdiv(D * G * F, 4.f * cosI * cosO));
}
// microfacet.h:124-134 (D with the widened alpha at |cos wi|)
PT_INLINE float microfacetPdf(float alpha, f3 wo, f3 wi) {
    if (wo.z < 0.f || wi.z < 0.f) return 0.f;
    const f3 wh = normalizeIeee(wo + wi);
    return fdiv(mfD(alpha, wi, wh) * fabsf(wh.z), 4.f * fabsf(dot(wo, wh)));
}

// bsdf.h:20-39
PT_INLINE f3 bsdfF(const PtMaterial &m, f3 wo, f3 wi) {
    const int t = matType(m);
    if (t == PT_MAT_LAMBERTIAN) {  // lambertian.h:43-49
        if (wo.z < 0.f || wi.z < 0.f) return mk(0.f);
        return mk(m.a.x, m.a.y, m.a.z) * kInvPi;
    }
    if (t == PT_MAT_MICROFACET) return microfacetF(m.b.w, wo, wi);
    return mk(0.f);
}
// bsdf.h:41-61
PT_INLINE float bsdfPdf(const PtMaterial &m, f3 wo, f3 wi) {
    const int t = matType(m);
    if (t == PT_MAT_LAMBERTIAN) return wi.z < 0.f ? 0.f : 0.15915494309189535f;
    if (t == PT_MAT_MICROFACET) return microfacetPdf(m.b.w, wo, wi);
    return 0.f;
}
// Density of the directions sample() produces, evaluated at an arbitrary wi: for the
// Beckmann lobe this is the sample() pdf form (widened alpha at |cos wo|, half vector taken
// in the upper hemisphere), which differs from pdf() (alpha at |cos wi|, microfacet.h:124-134).
// Used where MIS weights must sum to one (guided vertices, consistent-MIS option).
PT_INLINE float bsdfSamplePdf(const PtMaterial &m, f3 wo, f3 wi) {
    const int t = matType(m);
    if (t == PT_MAT_LAMBERTIAN) return wi.z < 0.f ? 0.f : 0.15915494309189535f;
    if (t == PT_MAT_MICROFACET) {
        const f3 s = wo + wi;
        if (isZero(s)) return 0.f;
        f3 wh = normalizeIeee(s);
        if (wh.z < 0.f) wh = -wh;
        const float d = fabsf(dot(wi, wh));
        if (!(d > 0.f)) return 0.f;
        return fdiv(mfD(m.b.w, wo, wh) * fabsf(wh.z), 4.f * d);
    }
    return 0.f;
}

// von Mises-Fisher lobe around mu (Jakob 2012, "Numerically stable sampling of the von
// Mises-Fisher distribution on S^2") and its density kappa / (2 pi (1 - e^{-2 kappa}))
// exp(kappa (mu.w - 1)).
PT_INLINE float vmfDensity(f3 mu, float kappa, f3 w) {
    return kappa * frcp(6.28318530718f * (1.f - __expf(-2.f * kappa))) * __expf(kappa * (dot(mu, w) - 1.f));
}
PT_INLINE f3 sampleVmf(f3 mu, float kappa, float u1, float u2) {
    const float w = fminf(fmaxf(1.f + __logf(u1 + (1.f - u1) * __expf(-2.f * kappa)) * frcp(kappa), -1.f), 1.f);
    const float s = fsqrt(fmaxf(0.f, 1.f - w * w));
    float sp, cp;
    sincospif(2.f * u2, &sp, &cp);
    return toWorld(makeFrame(mu), mk(s * cp, s * sp, w));
}

// Uniforms consumed by sample() per material (lambertian.h:55-59, mirror.h:37-43,
// glass.h:35-37, microfacet.h:136-140)
PT_INLINE int bsdfDims(int t) { return t == PT_MAT_MIRROR ? 0 : (t == PT_MAT_GLASS ? 1 : 2); }

// bsdf.h:87-115 and the per-material sample() functions
PT_INLINE BsdfSample bsdfSample(const PtMaterial &m, f3 wo, float xi1, float xi2) {
    BsdfSample s;
    const int t = matType(m);
    if (t == PT_MAT_LAMBERTIAN) {  // lambertian.h:67-76
        s.wi = uniformHemisphere(xi1, xi2);
        s.pdf = 0.15915494309189535f;
        s.f = (wo.z < 0.f || s.wi.z < 0.f) ? mk(0.f) : mk(m.a.x, m.a.y, m.a.z) * kInvPi;
        s.delta = false;
    } else if (t == PT_MAT_MIRROR) {  // mirror.h:45-53
        s.wi = mk(-wo.x, -wo.y, wo.z);
        s.pdf = 1.f;
        s.f = mk(fmaxf(0.f, frcp(s.wi.z)));
        s.delta = true;
    } else if (t == PT_MAT_GLASS) {  // glass.h:44-109
        float etaI = 1.f, etaT = m.b.w;
        if (wo.z < 0.f) { const float tt = etaI; etaI = etaT; etaT = tt; }
        const float F = fresnelDielectric(fabsf(wo.z), etaI, etaT);
        s.delta = true;
        if (xi1 < F) {
            s.wi = mk(-wo.x, -wo.y, wo.z);
            const float c = fabsf(s.wi.z);
            s.pdf = F;
            s.f = c == 0.f ? mk(0.f) : mk(fdiv(F, c));
        } else {
            s.wi = snellRefract(wo, etaI, etaT);
            const float Tr = 1.f - F;
            const float c = fabsf(s.wi.z);
            const float eta = fdiv(etaI, etaT);
            s.pdf = Tr;
            s.f = c == 0.f ? mk(0.f) : mk(fdiv(Tr, c) * (eta * eta));
        }
    } else if (t == PT_MAT_MICROFACET) {  // microfacet.h:53-63, 148-158
        const float alpha = m.b.w;
        const float a = mfAlpha(alpha, fabsf(wo.z));
        const float sq = -(a * a) * logf(1.f - xi2);
        // theta = atan(sqrt(sq)): cos theta = 1 / sqrt(1 + sq), sin theta = sqrt(sq) cos theta
        const float ct = frsq(1.f + sq);
        const float st = fsqrt(sq) * ct;
        float sp, cp;
        sincospif(2.f * xi1, &sp, &cp);  // phi = 2 pi xi1
        const f3 wh = mk(st * cp, st * sp, ct);
        s.wi = reflect(wo, wh);
        s.pdf = fdiv(mfD(alpha, wo, wh) * fabsf(wh.z), 4.f * fabsf(dot(s.wi, wh)));
        s.f = microfacetF(alpha, wo, s.wi);
        s.delta = false;
    } else {
        s.wi = mk(0.f, 0.f, 1.f);
        s.pdf = 0.f;
        s.f = mk(0.f);
        s.delta = false;
    }
    return s;
}

// ------------------------------------------------------------------------------------------
// Environment (environment_light.h:46-83, distribution.h:157-197): nearest-texel radiance,
// solid-angle pdf = texel_pdf / sin(theta) with texel_pdf = pmf * W*H / (2 pi^2); sampling by
// two alias draws (row marginal, then column) with continuous offsets inside the texel.
// ------------------------------------------------------------------------------------------
PT_INLINE void envAngles(f3 w, float *phi, float *theta) {
    float p = atan2f(w.z, w.x);
    if (p < 0.f) p += 6.28318530717958647692f;
    *phi = p;
    *theta = acosf(fminf(1.f, fmaxf(w.y, -1.f)));
}
// environment_light.h:74-75: the quotient is formed in double, rounded to float, floored
PT_INLINE void envTexel(const PtEnv &e, float phi, float theta, int *x, int *y) {
    int xi = (int)floorf((float)((double)phi / 6.283185307179586 * (double)e.width));
    int yi = (int)floorf((float)((double)theta / 3.141592653589793 * (double)e.height));
    *x = min(max(xi, 0), e.width - 1);
    *y = min(max(yi, 0), e.height - 1);
}
PT_INLINE f3 envEmit(const PtEnv &e, f3 w) {
    if (!e.present) return mk(0.f);
    float phi, theta;
    envAngles(w, &phi, &theta);
    int x, y;
    envTexel(e, phi, theta, &x, &y);
    const float4 v = e.rgba[(size_t)y * e.width + x];
    return mk(v.x, v.y, v.z);
}
// Radiance of direction w and the sampling density of w (solid angle) from one texel read.
// sin(theta) from the xz components keeps its precision next to the poles, where y rounds to
// +-1; on a pole the (phi, theta) map degenerates: density 0 there, and a light sample that
// lands on one is dropped (pbrt-v3 InfiniteAreaLight::Pdf_Li).
PT_INLINE f3 envEmitPdf(const PtEnv &e, f3 w, float *pdf) {
    float phi, theta;
    envAngles(w, &phi, &theta);
    int x, y;
    envTexel(e, phi, theta, &x, &y);
    const float4 v = e.rgba[(size_t)y * e.width + x];
    const float st = fsqrt(w.x * w.x + w.z * w.z);
    *pdf = st > 0.f ? fdiv(v.w, st) : 0.f;
    return mk(v.x, v.y, v.z);
}
PT_INLINE float envPdf(const PtEnv &e, f3 w) {
    float pdf;
    envEmitPdf(e, w, &pdf);
    return pdf;
}
PT_INLINE int aliasDraw(const float2 *tab, int n, float xi, float *rem) {
    const float s = xi * (float)n;
    int k = (int)s;
    if (k > n - 1) k = n - 1;
    float f = s - (float)k;
    if (f >= 1.0f) f = 0x1.fffffep-1f;
    float2 e = tab[k];  // (prob, alias as int bits)
    // one 8-byte load: the alias is otherwise fetched after the comparison, a second round trip
    asm volatile("" : "+v"(e.y));
    if (f < e.x) {
        *rem = fdiv(f, e.x);
        return k;
    }
    *rem = fdiv(f - e.x, 1.0f - e.x);
    return asInt(e.y);
}
PT_INLINE f3 envSample(const PtEnv &e, float xi1, float xi2, float *pdf, f3 *emit) {
    float ry, rx;
    const int y = aliasDraw(e.marginal, e.height, xi1, &ry);
    const int x = aliasDraw(e.rows + (size_t)y * e.width, e.width, xi2, &rx);
    const float u = fminf(fdiv(x + rx, (float)e.width), 0x1.fffffep-1f);
    const float v = fminf(fdiv(y + ry, (float)e.height), 0x1.fffffep-1f);
    float st, ct, sp, cp;
    sincospif(v, &st, &ct);
    sincospif(2.f * u, &sp, &cp);
    const f3 w = mk(st * cp, ct, st * sp);
    *emit = envEmitPdf(e, w, pdf);
    return w;
}

// ------------------------------------------------------------------------------------------
// Primitives.  Record layout (3 float4, in BVH order): triangle p0 = (v0, mat), p1 = (v1, tag),
// p2 = (v2, original index); tag >= 0: index into the per-vertex normal array, -1: flat.
// Sphere: p0 = (centre, mat), p1 = (radius, 0, 0, -2), p2 = (0, 0, 0, ~sphere index).
// ------------------------------------------------------------------------------------------
PT_INLINE float triArea(f3 p0, f3 p1, f3 p2) { return fabsf(length(cross(p1 - p0, p2 - p0)) * 0.5f); }
PT_INLINE float sphereArea(float r) { return 12.566370614359172f * r * r; }

struct Ray {
    f3 o, d;
};

// Watertight ray/triangle intersection (Woop, Benthin, Wald 2013), two-sided, edges
// inclusive, FP64 fallback on exactly-zero edge functions.  Precomputed per ray.  Neither
// function lets the compiler fuse a product into an add: the edge functions of two triangles
// sharing an edge must be exact negatives of each other, which rounding each product on its
// own guarantees and a fused multiply-add does not.
struct WtRay {
    f3 o;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
PT_INLINE WtRay wtSetup(f3 o, f3 d) {
#pragma clang fp contract(off)
    WtRay r;
    r.o = o;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    const float dz = comp(d, kz);
    if (dz < 0.f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = comp(d, kx) / dz;
    r.Sy = comp(d, ky) / dz;
    r.Sz = 1.f / dz;
    return r;
}
PT_INLINE bool wtHit(const WtRay &r, f3 p0, f3 p1, f3 p2, float tMin, float tMax, float *tt, float *uu, float *vv) {
#pragma clang fp contract(off)
    const f3 A = p0 - r.o, B = p1 - r.o, C = p2 - r.o;
    const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.Sx * Akz, Ay = comp(A, r.ky) - r.Sy * Akz;
    const float Bx = comp(B, r.kx) - r.Sx * Bkz, By = comp(B, r.ky) - r.Sy * Bkz;
    const float Cx = comp(C, r.kx) - r.Sx * Ckz, Cy = comp(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float Tn = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
    const float rcp = 1.f / det;
    const float t = Tn * rcp;
    if (!(t >= tMin && t <= tMax)) return false;
    *tt = t;
    *uu = V * rcp;
    *vv = W * rcp;
    return true;
}
// sphere.cu:7-43 (both roots, strict interval)
PT_INLINE bool sphereHit(f3 c, float rad, f3 o, f3 d, float tMin, float tMax, float *tt) {
    const f3 oc = o - c;
    const float a = dot(d, d);
    const float b = dot(oc, d);
    const float cc = dot(oc, oc) - rad * rad;
    const float disc = b * b - a * cc;
    if (disc > 0.f) {
        const float sq = sqrtf(disc);
        float t = (-b - sq) / a;
        if (t < tMax && t > tMin) { *tt = t; return true; }
        t = (-b + sq) / a;
        if (t < tMax && t > tMin) { *tt = t; return true; }
    }
    return false;
}

PT_INLINE float balance(float p1, float p2) { return fdiv(p1, p1 + p2); }  // mis.h:5-8

// color.h:7-28 sRGB transfer and framebuffer.cu:35-40's byte conversion (clamp to [0, 1],
// times 255, truncated).  The power is evaluated in FP64 and rounded once to float (a
// correctly rounded powf), and no product is fused into an add, so the byte is a function of
// the radiance value alone, the same on every device and in the oracle.
PT_INLINE unsigned char srgbByte(float x) {
#pragma clang fp contract(off)
    const float s = x <= 0.0031308f ? x * 12.92f : 1.055f * (float)pow((double)x, (double)(1.f / 2.2f)) - 0.055f;
    return (unsigned char)(fmaxf(0.f, fminf(1.f, s)) * 255.f);
}

}  // namespace pt
