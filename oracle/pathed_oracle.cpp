// pathed_oracle.cpp -- CPU oracle for the progressive path-tracing pass.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++, the arithmetic of the
// reference renderer (chellmuth/gpu-pathed) on its hot path: camera rays, closest hit,
// the four BSDFs, next-event estimation with balance-heuristic MIS, the environment
// distribution and the running-mean framebuffer.  It is the checker that the HIP
// product is compared against; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  Nothing in the product links or calls it.
//
// Every function cites the reference file:line it follows.  The reference itself is
// CUDA (curand, cuda_runtime, OptiX) and cannot be compiled in this image, so this is a
// restatement, not a build of the reference.
//
// Parity pins (what checks each part; see tests/test_oracle_*.py):
//   distribution / 2-D / phi-theta pdf  -> reference known answers (test_distribution.cu)
//   PCG32                               -> pcg-c-basic published demo outputs
//   triangle / sphere hit               -> closed forms, brute force, invariants
//   BSDFs                               -> closed forms (Lambert a/pi, mirror f*cos/pdf=1,
//                                          Fresnel at normal incidence, TIR, Snell), chi-square
//                                          of sample() against its own pdf, Walter G1 constants
//   light sampling                      -> area->solid-angle closed form, choice pmf 1/n
//   integrator                          -> white-furnace closed form, NEE on/off agreement
//   alias table                         -> pmf equality with normalised weights
// Functions with no such check are marked "parity unpinned" at their definition.
//
// Templated on the scalar type: T=float is the restatement of the reference (FP32);
// T=double is the FP64 high-precision oracle.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

namespace orc {

static const double kPi = 3.14159265358979323846;

// ----------------------------------------------------------------------------------------
// Vector type (reference include/core/vec3.h:8-118)
// ----------------------------------------------------------------------------------------
template <class T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a) : x(a), y(a), z(a) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(const V3 &o) const { return V3(x * o.x, y * o.y, z * o.z); }
    V3 operator*(T t) const { return V3(x * t, y * t, z * t); }
    V3 operator/(T t) const { return V3(x / t, y / t, z / t); }
    V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3 &operator*=(const V3 &o) { x *= o.x; y *= o.y; z *= o.z; return *this; }
    bool isZero() const { return x == 0 && y == 0 && z == 0; }
    T length() const { return std::sqrt(x * x + y * y + z * z); }
};
template <class T> inline V3<T> operator*(T t, const V3<T> &v) { return V3<T>(t * v.x, t * v.y, t * v.z); }
template <class T> inline T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T absDot(const V3<T> &a, const V3<T> &b) { return std::fabs(dot(a, b)); }
// vec3.h:84-91
template <class T> inline V3<T> cross(const V3<T> &a, const V3<T> &b) {
    return V3<T>(a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x);
}
template <class T> inline V3<T> normalized(const V3<T> &v) { return v / v.length(); }
// vec3.h:57-59: reflect(n) = n * (v.n) * 2 - v
template <class T> inline V3<T> reflect(const V3<T> &v, const V3<T> &n) { return n * dot(v, n) * T(2) - v; }

// ----------------------------------------------------------------------------------------
// PCG32 (O'Neill 2014, pcg-c-basic pcg32_random_r / pcg32_srandom_r).  The reference uses
// cuRAND XORWOW (path_tracer.cu:26-36), which is not reproduced; parity with the reference
// is statistical.  Seeding follows pbrt-v3's RNG::SetSequence: state = 0, inc = 2*seq+1,
// step, state += initstate, step.
// ----------------------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state, inc;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    // [0, 1) with 24 bits, exactly representable in float
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// Stream of sample `sample` of pixel `pixel` under seed `seed`.  The product derives the
// same stream in its ray-generation kernel.
static inline Pcg32 pathRng(uint64_t seed, uint32_t pixel, uint32_t sample) {
    Pcg32 r;
    const uint64_t stream = ((uint64_t)sample << 32) | (uint64_t)pixel;
    r.seed(0x853c49e6748fea9bULL ^ seed, splitmix64(stream));
    return r;
}

// ----------------------------------------------------------------------------------------
// Frames (frame.h:7-59, tangent_frame.h:5-56, world_frame.h:5-22)
// ----------------------------------------------------------------------------------------
template <class T> struct Frame {
    V3<T> s, t, n;
    explicit Frame(const V3<T> &nn) : n(nn) {
        // frame.h:7-18 coordinateSystem
        if (std::fabs(n.x) > std::fabs(n.y)) {
            T invLen = T(1) / std::sqrt(n.x * n.x + n.z * n.z);
            t = V3<T>(n.z * invLen, T(0), -n.x * invLen);
        } else {
            T invLen = T(1) / std::sqrt(n.y * n.y + n.z * n.z);
            t = V3<T>(T(0), n.z * invLen, -n.y * invLen);
        }
        s = cross(t, n);
    }
    V3<T> toWorld(const V3<T> &l) const {
        return V3<T>(s.x * l.x + t.x * l.y + n.x * l.z, s.y * l.x + t.y * l.y + n.y * l.z,
                     s.z * l.x + t.z * l.y + n.z * l.z);
    }
    V3<T> toLocal(const V3<T> &w) const {
        return V3<T>(s.x * w.x + s.y * w.y + s.z * w.z, t.x * w.x + t.y * w.y + t.z * w.z,
                     n.x * w.x + n.y * w.y + n.z * w.z);
    }
};
template <class T> inline T absCosTheta(const V3<T> &v) { return std::fabs(v.z); }
template <class T> inline T cos2Theta(const V3<T> &v) { return v.z * v.z; }
template <class T> inline T sin2Theta(const V3<T> &v) { return T(1) - cos2Theta(v); }
template <class T> inline T tan2Theta(const V3<T> &v) { return sin2Theta(v) / cos2Theta(v); }
template <class T> inline T sinTheta(const V3<T> &v) { return std::sqrt(std::max(T(0), T(1) - cos2Theta(v))); }
template <class T> inline T tanTheta(const V3<T> &v) { return sinTheta(v) / v.z; }
template <class T> inline T signOf(T x) { return x < T(0) ? T(-1) : T(1); }  // util.h:16-20

// ----------------------------------------------------------------------------------------
// Materials (materials/types.h:5-11, params.h, lambertian.h, mirror.h, glass.h, microfacet.h)
// ----------------------------------------------------------------------------------------
enum MatType { MAT_LAMBERTIAN = 0, MAT_MIRROR = 1, MAT_GLASS = 2, MAT_MICROFACET = 3, MAT_NONE = 4 };

template <class T> struct Material {
    int type;
    V3<T> albedo, emit;
    T ior, alpha;
};

template <class T> struct BSDFSample {
    V3<T> wi;
    T pdf;
    V3<T> f;
    bool isDelta;
};

// sample.h:7-17
template <class T> inline V3<T> uniformHemisphere(T xi1, T xi2) {
    const T z = xi1;
    const T r = std::sqrt(std::max(T(0), T(1) - z * z));
    const T phi = (T)(2 * kPi * (double)xi2);
    return V3<T>(r * std::cos(phi), r * std::sin(phi), z);
}
// sample.h:19-24
template <class T> inline T uniformHemispherePdf(const V3<T> &v) { return v.z < T(0) ? T(0) : (T)(1.0 / (2.0 * kPi)); }
// sample.h:26-37
template <class T> inline V3<T> uniformSphere(T xi1, T xi2) {
    const T z = xi1 * T(2) - T(1);
    const T r = std::sqrt(std::max(T(0), T(1) - z * z));
    const T phi = (T)(2 * kPi * (double)xi2);
    return V3<T>(r * std::cos(phi), r * std::sin(phi), z);
}

// lambertian.h:43-49
template <class T> inline V3<T> lambertF(const Material<T> &m, const V3<T> &wo, const V3<T> &wi) {
    if (wo.z < T(0) || wi.z < T(0)) return V3<T>(T(0));
    return m.albedo / (T)kPi;
}

// snell.h:60-66, trig.h:10-14
template <class T> inline T transmittedSinTheta(T cosI, T etaI, T etaT) {
    return (etaI / etaT) * std::sqrt(std::max(T(0), T(1) - cosI * cosI));
}
// fresnel.h:17-40 (unpolarised dielectric reflectance)
template <class T> inline T fresnelDielectric(T cosI, T etaI, T etaT) {
    const T sinT = transmittedSinTheta(cosI, etaI, etaT);
    if (sinT > T(1)) return T(1);
    const T cosT = std::sqrt(std::max(T(0), T(1) - sinT * sinT));
    const T rPar = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
    const T rPer = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
    return T(0.5) * (rPar * rPar + rPer * rPer);
}
// snell.h:10-43.  Note the second eta swap when the incident vector lies below the
// normal: Glass::sample already swapped, so leaving rays refract with eta ratio
// 1/ior (reference quirk, kept for parity).
template <class T> inline V3<T> snellRefract(const V3<T> &wi, T etaI, T etaT) {
    V3<T> n(T(0), T(0), T(1));
    if (dot(wi, n) < T(0)) {
        n = n * T(-1);
        std::swap(etaI, etaT);
    }
    const V3<T> perpI = wi - n * dot(wi, n);
    const V3<T> perpT = -perpI * (etaI / etaT);
    const T l = perpT.length();
    const T parLen = std::sqrt(std::max(T(0), T(1) - l * l));
    const V3<T> parT = n * -parLen;
    return normalized(parT + perpT);
}

// microfacet.h:70-72
template <class T> inline T mfSampleAlpha(T alpha, T absCos) { return (T(1.2) - T(0.2) * std::sqrt(absCos)) * alpha; }
// microfacet.h:37-51: D evaluated with the widened alpha at |cos wi|
template <class T> inline T mfD(T alpha, const V3<T> &wi, const V3<T> &wh) {
    const T a = mfSampleAlpha(alpha, absCosTheta(wi));
    const T a2 = a * a;
    const T c2 = cos2Theta(wh);
    const T c4 = c2 * c2;
    const T t2 = tan2Theta(wh);
    if (std::isinf(t2)) return T(0);
    return std::exp(-t2 / a2) / ((T)kPi * a2 * c4);
}
// microfacet.h:74-84 (Walter et al. 2007 rational G1 approximation)
template <class T> inline T mfG1(T alpha, const V3<T> &v) {
    const T a = T(1) / (alpha * std::fabs(tanTheta(v)));
    if (a >= T(1.6)) return T(1);
    return (T(3.535) * a + T(2.181) * a * a) / (T(1) + T(2.276) * a + T(2.577) * a * a);
}
// microfacet.h:27-35
template <class T> inline T mfG(T alpha, const V3<T> &wi, const V3<T> &wo, const V3<T> &wh) {
    if (signOf(dot(wo, wh)) != signOf(wo.z)) return T(0);
    if (signOf(dot(wi, wh)) != signOf(wi.z)) return T(0);
    return mfG1(alpha, wi) * mfG1(alpha, wo);
}
// microfacet.h:98-122
template <class T> inline V3<T> microfacetF(T alpha, const V3<T> &wo, const V3<T> &wi) {
    if (wo.z < T(0) || wi.z < T(0)) return V3<T>(T(0));
    const T cosO = absCosTheta(wo), cosI = absCosTheta(wi);
    const V3<T> wh = normalized(wo + wi);
    if (cosO == T(0) || cosI == T(0)) return V3<T>(T(0));
    if (wh.isZero()) return V3<T>(T(0));
    const T cosInc = std::min(T(1), std::max(dot(wi, wh), T(0)));
    const T F = fresnelDielectric(cosInc, T(1), T(1.5));
    const T D = mfD(alpha, wi, wh);
    const T G = mfG(alpha, wi, wo, wh);
    return V3<T>(T(1)) * D * G * F / (T(4) * cosI * cosO);
}
// microfacet.h:124-134
template <class T> inline T microfacetPdf(T alpha, const V3<T> &wo, const V3<T> &wi) {
    if (wo.z < T(0) || wi.z < T(0)) return T(0);
    const V3<T> wh = normalized(wo + wi);
    return mfD(alpha, wi, wh) * std::fabs(wh.z) / (T(4) * absDot(wo, wh));
}

// BSDF dispatch: bsdf.h:20-61 (f, pdf) and bsdf.h:87-115 (sample)
template <class T> inline V3<T> bsdfF(const Material<T> &m, const V3<T> &wo, const V3<T> &wi) {
    switch (m.type) {
    case MAT_LAMBERTIAN: return lambertF(m, wo, wi);
    case MAT_MICROFACET: return microfacetF(m.alpha, wo, wi);
    default: return V3<T>(T(0));  // mirror.h:25-27, glass.h:27-29
    }
}
template <class T> inline T bsdfPdf(const Material<T> &m, const V3<T> &wo, const V3<T> &wi) {
    switch (m.type) {
    case MAT_LAMBERTIAN: return uniformHemispherePdf(wi);  // lambertian.h:51-53
    case MAT_MICROFACET: return microfacetPdf(m.alpha, wo, wi);
    default: return T(0);
    }
}
// Density of the directions bsdfSample() draws, at an arbitrary wi.  For the Beckmann lobe
// sample() draws the half vector from the alpha widened at |cos wo| in the upper hemisphere
// and reflects wo about it (microfacet.h:53-63, 148-158), so p(wi) = D(wh) |cos wh| / (4 |wi.wh|)
// with wh = +-normalize(wo + wi) taken upward; this differs from pdf() (microfacet.h:124-134),
// which widens alpha at |cos wi|.  The consistent-MIS option weights light samples with it.
template <class T> inline T bsdfSamplePdf(const Material<T> &m, const V3<T> &wo, const V3<T> &wi) {
    switch (m.type) {
    case MAT_LAMBERTIAN: return uniformHemispherePdf(wi);
    case MAT_MICROFACET: {
        const V3<T> s = wo + wi;
        if (s.isZero()) return T(0);
        V3<T> wh = normalized(s);
        if (wh.z < T(0)) wh = -wh;
        const T d = absDot(wi, wh);
        if (!(d > T(0))) return T(0);
        return mfD(m.alpha, wo, wh) * std::fabs(wh.z) / (T(4) * d);
    }
    default: return T(0);
    }
}
// Number of uniforms each material's sample() consumes (lambertian.h:55-59, mirror.h:37-43,
// glass.h:35-37, microfacet.h:136-140).
inline int bsdfSampleDims(int type) {
    switch (type) {
    case MAT_LAMBERTIAN: return 2;
    case MAT_MIRROR: return 0;
    case MAT_GLASS: return 1;
    case MAT_MICROFACET: return 2;
    default: return 0;
    }
}
template <class T> inline BSDFSample<T> bsdfSample(const Material<T> &m, const V3<T> &wo, T xi1, T xi2) {
    BSDFSample<T> s;
    switch (m.type) {
    case MAT_LAMBERTIAN: {  // lambertian.h:67-76
        s.wi = uniformHemisphere(xi1, xi2);
        s.pdf = uniformHemispherePdf(s.wi);
        s.f = lambertF(m, wo, s.wi);
        s.isDelta = false;
        return s;
    }
    case MAT_MIRROR: {  // mirror.h:45-53
        s.wi = reflect(wo, V3<T>(T(0), T(0), T(1)));
        s.pdf = T(1);
        s.f = V3<T>(std::max(T(0), T(1) / s.wi.z));
        s.isDelta = true;
        return s;
    }
    case MAT_GLASS: {  // glass.h:44-109
        T etaI = T(1), etaT = m.ior;
        if (wo.z < T(0)) std::swap(etaI, etaT);
        V3<T> wi = snellRefract(wo, etaI, etaT);
        const T F = fresnelDielectric(absCosTheta(wo), etaI, etaT);
        if (xi1 < F) {
            wi = reflect(wo, V3<T>(T(0), T(0), T(1)));
            const T c = absCosTheta(wi);
            s.wi = wi;
            s.pdf = F;
            s.f = c == T(0) ? V3<T>(T(0)) : V3<T>(F / c);
            s.isDelta = true;
            return s;
        }
        const T Tr = T(1) - F;
        const T c = absCosTheta(wi);
        const V3<T> thr = c == T(0) ? V3<T>(T(0)) : V3<T>(Tr / c);
        const T corr = (etaI / etaT) * (etaI / etaT);
        s.wi = wi;
        s.pdf = Tr;
        s.f = thr * corr;
        s.isDelta = true;
        return s;
    }
    case MAT_MICROFACET: {  // microfacet.h:53-63, 148-158
        const T phi = (T)((double)xi1 * kPi * 2.0);
        const T a = mfSampleAlpha(m.alpha, absCosTheta(wo));
        const T a2 = a * a;
        const T sq = -a2 * std::log(T(1) - xi2);
        const T theta = std::atan(std::sqrt(sq));
        const T ct = std::cos(theta), st = std::sin(theta);
        const V3<T> wh(st * std::cos(phi), st * std::sin(phi), ct);  // spherical_coordinates.h:35-42
        const V3<T> wi = reflect(wo, wh);
        s.wi = wi;
        s.pdf = mfD(m.alpha, wo, wh) * std::fabs(wh.z) / (T(4) * absDot(wi, wh));
        s.f = microfacetF(m.alpha, wo, wi);
        s.isDelta = false;
        return s;
    }
    default:
        s.wi = V3<T>(T(0));
        s.pdf = T(0);
        s.f = V3<T>(T(0));
        s.isDelta = false;
        return s;
    }
}

// ----------------------------------------------------------------------------------------
// 1-D / 2-D piecewise-constant distributions (math/distribution.h:11-201, distribution.cu)
// ----------------------------------------------------------------------------------------
struct Distribution1D {
    std::vector<float> cmf;
    float totalMass = 0.f;
};
// distribution.cu:9-56
static Distribution1D buildDistribution(const float *values, int n) {
    Distribution1D d;
    float total = 0.f;
    for (int i = 0; i < n; i++) total += values[i];
    d.totalMass = total;
    d.cmf.resize(n);
    int lastIndex = -1;
    for (int i = 0; i < n; i++) {
        const float mass = values[i] / total;
        if (mass > 0.f) lastIndex = i;
        d.cmf[i] = mass;
        if (i > 0) d.cmf[i] += d.cmf[i - 1];
    }
    if (total > 0.f)
        for (int i = lastIndex; i < n; i++) d.cmf[i] = 1.f;
    return d;
}
// distribution.h:48-55
static float distPmf(const Distribution1D &d, int i) {
    if (d.totalMass == 0.f) return 0.f;
    const float c = d.cmf[i];
    if (i == 0) return c;
    return c - d.cmf[i - 1];
}
// distribution.h:19-46 (linear scan; xi == 1 handled; empty distribution -> -1 where the
// reference asserts)
static int distSample(const Distribution1D &d, float xi, float *pmf) {
    if (d.totalMass == 0.f) { *pmf = 0.f; return -1; }
    const int n = (int)d.cmf.size();
    for (int i = 0; i < n; i++)
        if (xi < d.cmf[i]) { *pmf = distPmf(d, i); return i; }
    if (xi == 1.f)
        for (int i = 0; i < n; i++)
            if (d.cmf[i] == 1.f) { *pmf = distPmf(d, i); return i; }
    *pmf = 0.f;
    return -1;
}
struct Distribution2D {
    Distribution1D y;
    std::vector<Distribution1D> x;
    int w = 0, h = 0;
};
// distribution.cu:58-90
static Distribution2D buildDistribution2D(const float *data, int w, int h) {
    Distribution2D d;
    d.w = w; d.h = h;
    std::vector<float> rowSums(h);
    d.x.resize(h);
    for (int r = 0; r < h; r++) {
        d.x[r] = buildDistribution(data + (size_t)r * w, w);
        float t = 0.f;
        for (int c = 0; c < w; c++) t += data[(size_t)r * w + c];
        rowSums[r] = t;
    }
    d.y = buildDistribution(rowSums.data(), h);
    return d;
}
// distribution.h:113-120
static float dist2Pmf(const Distribution2D &d, int x, int y) {
    if (d.y.totalMass == 0.f) return 0.f;
    return distPmf(d.x[y], x) * distPmf(d.y, y);
}
// distribution.h:103-111
static void dist2Sample(const Distribution2D &d, float xi0, float xi1, int *x, int *y, float *pmf) {
    float py, px;
    *y = distSample(d.y, xi0, &py);
    if (*y < 0) { *x = -1; *pmf = 0.f; return; }
    *x = distSample(d.x[*y], xi1, &px);
    *pmf = px * py;
}

// spherical_coordinates.h:8-19 (y-up)
template <class T> inline void cartesianToSpherical(const V3<T> &c, T *phi, T *theta) {
    *phi = std::atan2(c.z, c.x);
    if (*phi < T(0)) *phi += (T)(2 * kPi);
    *theta = std::acos(std::min(T(1), std::max(c.y, T(-1))));
}

// distribution.h:186-197 (pdf w.r.t. the (phi,theta) unit square, canonical coordinates)
static float phiThetaPdfCanonical(const Distribution2D &d, float xc, float yc) {
    const int xs = (int)std::min(std::floor(xc * d.w), (float)(d.w - 1));
    const int ys = (int)std::min(std::floor(yc * d.h), (float)(d.h - 1));
    return dist2Pmf(d, xs, ys) * d.w * d.h;
}

// ----------------------------------------------------------------------------------------
// Alias tables (the product's environment sampler).  Per row: an alias table over the
// row's texels; marginal: an alias table over row sums.  Built in FP64 with the sweep
// construction (Huebschle-Schneider & Sanders 2019, "SWEEP"): the first heavy item absorbs
// the deficits of lights in index order and turns light when its residual drops to <= 1.
// The product builds the same tables on the GPU with the same operation order; the pmf of
// a table must equal the normalised weights (checked in tests).
// ----------------------------------------------------------------------------------------
static void buildAliasRow(const double *w, int n, double *prob, int *alias, double *sumOut) {
    double sum = 0.0;
    for (int i = 0; i < n; i++) sum += w[i];
    *sumOut = sum;
    for (int i = 0; i < n; i++) { prob[i] = 1.0; alias[i] = i; }
    if (!(sum > 0.0)) return;
    const double scale = (double)n / sum;
    // q_i = w_i * n / sum; lights have q < 1, heavies q >= 1 (classification never changes)
    auto q = [&](int k) { return w[k] * scale; };
    int i = 0;
    while (i < n && !(q(i) < 1.0)) i++;  // first light
    int j = 0;
    while (j < n && q(j) < 1.0) j++;     // first heavy
    if (j >= n) return;
    double wj = q(j);
    while (true) {
        if (wj >= 1.0) {
            if (i >= n) break;
            // light i keeps its own mass q_i and takes the rest of its bucket from heavy j
            prob[i] = q(i);
            alias[i] = j;
            wj = wj + (q(i) - 1.0);
            i++;
            while (i < n && !(q(i) < 1.0)) i++;
        } else {
            // heavy j turned light with residual wj: its bucket aliases to the next heavy
            int jn = j + 1;
            while (jn < n && q(jn) < 1.0) jn++;
            if (jn >= n) break;
            prob[j] = wj;
            alias[j] = jn;
            wj = q(jn) + (wj - 1.0);
            j = jn;
        }
    }
    // buckets never reached (heavies after the sweep ends, rounding leftovers) keep prob 1
}

struct EnvTables {
    int w = 0, h = 0;
    std::vector<float> rgba;
    // reference (CMF, texel centres)
    Distribution2D cdf;
    // product (alias, continuous in texel)
    std::vector<double> rowProb, margProb;
    std::vector<int> rowAlias, margAlias;
    std::vector<float> texelPdf;  // pmf * w*h / (2 pi^2)
    bool present = false;
};

static void buildEnv(EnvTables &e, const float *rgba, int w, int h, bool sinWeight) {
    e.w = w; e.h = h;
    e.rgba.assign(rgba, rgba + (size_t)w * h * 4);
    e.present = true;
    // environment_light.cu:44-50: weight = r + g + b (no sin theta)
    std::vector<float> inten((size_t)w * h);
    for (size_t i = 0; i < (size_t)w * h; i++) inten[i] = rgba[i * 4 + 0] + rgba[i * 4 + 1] + rgba[i * 4 + 2];
    e.cdf = buildDistribution2D(inten.data(), w, h);
    std::vector<double> wd((size_t)w * h);
    for (int y = 0; y < h; y++) {
        const double st = sinWeight ? std::sin((y + 0.5) / h * kPi) : 1.0;
        for (int x = 0; x < w; x++) {
            const size_t i = (size_t)y * w + x;
            wd[i] = ((double)rgba[i * 4 + 0] + (double)rgba[i * 4 + 1] + (double)rgba[i * 4 + 2]) * st;
        }
    }
    e.rowProb.resize((size_t)w * h);
    e.rowAlias.resize((size_t)w * h);
    std::vector<double> rowSum(h);
    for (int y = 0; y < h; y++)
        buildAliasRow(&wd[(size_t)y * w], w, &e.rowProb[(size_t)y * w], &e.rowAlias[(size_t)y * w], &rowSum[y]);
    e.margProb.resize(h);
    e.margAlias.resize(h);
    double total;
    buildAliasRow(rowSum.data(), h, e.margProb.data(), e.margAlias.data(), &total);
    e.texelPdf.resize((size_t)w * h);
    const double norm = total > 0.0 ? (double)w * h / (2.0 * kPi * kPi) / total : 0.0;
    for (size_t i = 0; i < (size_t)w * h; i++) e.texelPdf[i] = (float)(wd[i] * norm);
}

// environment_light.h:74-75 texel index: the quotient is formed in double and rounded to
// float before floorf; clamped to the map (the reference reads one past the last row at
// theta == pi).
static inline void envTexel(const EnvTables &e, double phi, double theta, int *x, int *y) {
    int xi = (int)std::floor((float)(phi / (2 * kPi) * (double)e.w));
    int yi = (int)std::floor((float)(theta / kPi * (double)e.h));
    *x = std::min(std::max(xi, 0), e.w - 1);
    *y = std::min(std::max(yi, 0), e.h - 1);
}

// environment_light.h:68-83 nearest texel; x/y clamped (the reference reads one past the
// last row at theta == pi)
template <class T> inline V3<T> envEmit(const EnvTables &e, const V3<T> &wi) {
    if (!e.present) return V3<T>(T(0));
    T phi, theta;
    cartesianToSpherical(wi, &phi, &theta);
    int x, y;
    envTexel(e, (double)phi, (double)theta, &x, &y);
    const size_t o = ((size_t)y * e.w + x) * 4;
    return V3<T>((T)e.rgba[o], (T)e.rgba[o + 1], (T)e.rgba[o + 2]);
}

// distribution.h:176-184 solid-angle pdf; refMode uses the FP32 CMF pmf, product mode the
// FP64-built texel pdf table
template <class T> inline T envPdf(const EnvTables &e, const V3<T> &wi, bool refMode) {
    T phi, theta;
    cartesianToSpherical(wi, &phi, &theta);
    if (refMode) {
        const float xc = (float)(phi / (T)(2 * kPi));
        const float yc = (float)(theta / (T)kPi);
        return (T)phiThetaPdfCanonical(e.cdf, xc, yc) / (std::sin(theta) * (T)kPi * (T)kPi * T(2));
    }
    int x, y;
    envTexel(e, (double)phi, (double)theta, &x, &y);
    // sin(theta) from the xz components keeps its precision next to the poles, where y
    // rounds to +-1; on a pole the (phi, theta) map degenerates: density 0 there, and a
    // light sample that lands on one is dropped (pbrt-v3 InfiniteAreaLight::Pdf_Li)
    const T st = std::sqrt(wi.x * wi.x + wi.z * wi.z);
    return st > T(0) ? (T)e.texelPdf[(size_t)y * e.w + x] / st : T(0);
}

// Alias draw: bucket = floor(xi*n); the remainder decides prob vs alias and is rescaled to a
// fresh uniform for the in-texel offset.
static inline int aliasDraw(const double *prob, const int *alias, int n, float xi, float *rem) {
    const float s = xi * (float)n;
    int k = (int)s;
    if (k > n - 1) k = n - 1;
    float f = s - (float)k;
    if (f >= 1.0f) f = 0x1.fffffep-1f;
    const float p = (float)prob[k];
    if (f < p) {
        *rem = f / p;
        return k;
    }
    *rem = (f - p) / (1.0f - p);
    return alias[k];
}

// environment_light.h:50-66 and distribution.h:157-174 (reference: texel centre) or the
// product sampler (alias + continuous in-texel offset).
template <class T> inline V3<T> envSample(const EnvTables &e, float xi1, float xi2, bool refMode, T *pdf) {
    if (refMode) {
        int x, y;
        float pmf;
        dist2Sample(e.cdf, xi1, xi2, &x, &y, &pmf);
        if (x < 0) { *pdf = T(0); return V3<T>(T(0), T(1), T(0)); }
        const float phiC = (x + 0.5f) / e.w;
        const float thetaC = (y + 0.5f) / e.h;
        const T phi = (T)(phiC * kPi * 2.0);
        const T theta = (T)(thetaC * kPi);
        const V3<T> wi(std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi));
        *pdf = envPdf(e, wi, true);
        return wi;
    }
    float ry, rx;
    const int y = aliasDraw(e.margProb.data(), e.margAlias.data(), e.h, xi1, &ry);
    const int x = aliasDraw(&e.rowProb[(size_t)y * e.w], &e.rowAlias[(size_t)y * e.w], e.w, xi2, &rx);
    const float u = std::min((x + rx) / (float)e.w, 0x1.fffffep-1f);
    const float v = std::min((y + ry) / (float)e.h, 0x1.fffffep-1f);
    const T phi = (T)u * (T)(2 * kPi);
    const T theta = (T)v * (T)kPi;
    const V3<T> wi(std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi));
    *pdf = envPdf(e, wi, false);
    return wi;
}

// ----------------------------------------------------------------------------------------
// Geometry (primitives/triangle.h, triangle.cu:7-48, sphere.h, sphere.cu:7-43)
// ----------------------------------------------------------------------------------------
template <class T> struct Tri {
    V3<T> p0, p1, p2, n0, n1, n2;
    int mat;
};
template <class T> struct Sph {
    V3<T> c;
    T r;
    int mat;
};
template <class T> struct Hit {
    T t, u, v;
    int prim;  // >= 0 triangle index, < 0: sphere ~index
};

// triangle.cu:13-34 Moeller-Trumbore in the reference's operation order.  Edges inclusive.
template <class T> inline bool hitMT(const Tri<T> &tr, const V3<T> &o, const V3<T> &d, T tMin, T tMax, T *tt, T *uu, T *vv) {
    const V3<T> e1 = tr.p1 - tr.p0;
    const V3<T> e2 = tr.p2 - tr.p0;
    const V3<T> s1 = cross(d, e2);
    const T div = dot(s1, e1);
    if (div == T(0)) return false;
    const T inv = T(1) / div;
    const V3<T> s = o - tr.p0;
    const T b1 = dot(s, s1) * inv;
    if (b1 < T(0) || b1 > T(1)) return false;
    const V3<T> s2 = cross(s, e1);
    const T b2 = dot(d, s2) * inv;
    if (b2 < T(0) || (b1 + b2) > T(1)) return false;
    const T t = dot(e2, s2) * inv;
    if (t < tMin || t > tMax) return false;
    *tt = t; *uu = b1; *vv = b2;
    return true;
}

// Watertight ray/triangle test (Woop, Benthin, Wald 2013, JCGT 2(1)), two-sided, edges
// inclusive, FP64 fallback when an edge function is exactly zero.  u/v are the weights of
// p1/p2 as in the MT test.
template <class T> inline bool hitWatertight(const Tri<T> &tr, const V3<T> &o, const V3<T> &d, T tMin, T tMax, T *tt, T *uu, T *vv) {
    const T ad[3] = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    int kz = 0;
    if (ad[1] > ad[kz]) kz = 1;
    if (ad[2] > ad[kz]) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < T(0)) std::swap(kx, ky);
    const T Sx = d[kx] / d[kz], Sy = d[ky] / d[kz], Sz = T(1) / d[kz];
    const V3<T> A = tr.p0 - o, B = tr.p1 - o, C = tr.p2 - o;
    const T Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    const T Bx = B[kx] - Sx * B[kz], By = B[ky] - Sy * B[kz];
    const T Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    T U = Cx * By - Cy * Bx;
    T V = Ax * Cy - Ay * Cx;
    T W = Bx * Ay - By * Ax;
    if (U == T(0) || V == T(0) || W == T(0)) {
        U = (T)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (T)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (T)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < T(0) || V < T(0) || W < T(0)) && (U > T(0) || V > T(0) || W > T(0))) return false;
    const T det = U + V + W;
    if (det == T(0)) return false;
    const T Az = Sz * A[kz], Bz = Sz * B[kz], Cz = Sz * C[kz];
    const T Tn = U * Az + V * Bz + W * Cz;
    const T rcp = T(1) / det;
    const T t = Tn * rcp;
    if (!(t >= tMin && t <= tMax)) return false;
    *tt = t; *uu = V * rcp; *vv = W * rcp;
    return true;
}

// sphere.cu:7-43 both roots, strict (tMin, tMax)
template <class T> inline bool hitSphere(const Sph<T> &s, const V3<T> &o, const V3<T> &d, T tMin, T tMax, T *tt) {
    const V3<T> oc = o - s.c;
    const T a = dot(d, d);
    const T b = dot(oc, d);
    const T c = dot(oc, oc) - s.r * s.r;
    const T disc = b * b - a * c;
    if (disc > T(0)) {
        T tmp = (-b - std::sqrt(disc)) / a;
        if (tmp < tMax && tmp > tMin) { *tt = tmp; return true; }
        tmp = (-b + std::sqrt(disc)) / a;
        if (tmp < tMax && tmp > tMin) { *tt = tmp; return true; }
    }
    return false;
}

// ----------------------------------------------------------------------------------------
// Scene, BVH (CPU baseline only: median split; the reference brute-forces, world.cu:5-41)
// ----------------------------------------------------------------------------------------
template <class T> struct Camera {
    T m[3][4];
    T vfov;
    int W, H;
};

// transform.cu:7-33 lookAt
template <class T> static Camera<T> makeCamera(const float o[3], const float t[3], const float u[3], float vfov, int flip, int W, int H) {
    Camera<T> c;
    const V3<T> src(o[0], o[1], o[2]), tgt(t[0], t[1], t[2]), up(u[0], u[1], u[2]);
    const V3<T> dir = normalized(src - tgt);
    const V3<T> xa = normalized(cross(normalized(up), dir));
    const V3<T> ya = normalized(cross(dir, xa));
    const T sign = flip ? T(-1) : T(1);
    const V3<T> cols[4] = {xa * sign, ya, dir, src};
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 4; k++) c.m[r][k] = cols[k][r];
    c.vfov = vfov; c.W = W; c.H = H;
    return c;
}
// camera.h:40-61 (row 0 = bottom)
template <class T> static void cameraRay(const Camera<T> &c, int row, int col, T sx, T sy, V3<T> *o, V3<T> *d) {
    const T top = std::tan(c.vfov / T(2));
    const T height = top * T(2);
    const T aspect = T(1) * c.W / c.H;
    const T width = height * aspect;
    const T right = width / T(2);
    const T xc = (col + sx) / c.W;
    const T yc = (row + sy) / c.H;
    const T y = yc * height - top;
    const T x = xc * width - right;
    const V3<T> dl = normalized(V3<T>(x, y, T(-1)));
    *o = V3<T>(c.m[0][3], c.m[1][3], c.m[2][3]);
    *d = V3<T>(c.m[0][0] * dl.x + c.m[0][1] * dl.y + c.m[0][2] * dl.z,
               c.m[1][0] * dl.x + c.m[1][1] * dl.y + c.m[1][2] * dl.z,
               c.m[2][0] * dl.x + c.m[2][1] * dl.y + c.m[2][2] * dl.z);
}

struct BvhNode {
    float lo[3], hi[3];
    int left, right;  // internal: children; leaf: left = first prim, right = -count
};

template <class T> struct Scene {
    std::vector<Tri<T>> tris;
    std::vector<Sph<T>> sphs;
    std::vector<Material<T>> mats;
    std::vector<int> lights;  // >= 0 triangle, < 0 sphere ~index (scene_data.cu:92-115 order)
    EnvTables env;
    Camera<T> cam;
    // BVH over primitive refs (same encoding as Hit::prim)
    std::vector<BvhNode> nodes;
    std::vector<int> refs;
    bool useBvh = false;
};

template <class T> static void primBounds(const Scene<T> &s, int p, float lo[3], float hi[3]) {
    if (p >= 0) {
        const Tri<T> &t = s.tris[p];
        for (int k = 0; k < 3; k++) {
            lo[k] = (float)std::min(t.p0[k], std::min(t.p1[k], t.p2[k]));
            hi[k] = (float)std::max(t.p0[k], std::max(t.p1[k], t.p2[k]));
        }
    } else {
        const Sph<T> &sp = s.sphs[~p];
        for (int k = 0; k < 3; k++) {
            lo[k] = (float)(sp.c[k] - sp.r);
            hi[k] = (float)(sp.c[k] + sp.r);
        }
    }
    for (int k = 0; k < 3; k++) {  // conservative padding
        lo[k] = std::nextafter(lo[k], -INFINITY);
        hi[k] = std::nextafter(hi[k], INFINITY);
    }
}

template <class T> static void buildBvh(Scene<T> &s) {
    const int n = (int)s.tris.size() + (int)s.sphs.size();
    s.refs.resize(n);
    for (int i = 0; i < (int)s.tris.size(); i++) s.refs[i] = i;
    for (int i = 0; i < (int)s.sphs.size(); i++) s.refs[s.tris.size() + i] = ~i;
    std::vector<float> blo((size_t)n * 3), bhi((size_t)n * 3), cen((size_t)n * 3);
    for (int i = 0; i < n; i++) {
        primBounds(s, s.refs[i], &blo[i * 3], &bhi[i * 3]);
        for (int k = 0; k < 3; k++) cen[i * 3 + k] = 0.5f * (blo[i * 3 + k] + bhi[i * 3 + k]);
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    s.nodes.clear();
    s.nodes.reserve(2 * n + 1);
    struct Item { int node, b, e; };
    std::vector<Item> st;
    s.nodes.push_back(BvhNode());
    st.push_back({0, 0, n});
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        BvhNode nd;
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; k++) { nd.lo[k] = INFINITY; nd.hi[k] = -INFINITY; }
        for (int i = it.b; i < it.e; i++) {
            const int p = idx[i];
            for (int k = 0; k < 3; k++) {
                nd.lo[k] = std::min(nd.lo[k], blo[p * 3 + k]);
                nd.hi[k] = std::max(nd.hi[k], bhi[p * 3 + k]);
                clo[k] = std::min(clo[k], cen[p * 3 + k]);
                chi[k] = std::max(chi[k], cen[p * 3 + k]);
            }
        }
        const int cnt = it.e - it.b;
        int axis = 0;
        for (int k = 1; k < 3; k++)
            if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
        if (cnt <= 4 || chi[axis] - clo[axis] <= 0.f) {
            nd.left = it.b;
            nd.right = -cnt;
            s.nodes[it.node] = nd;
            continue;
        }
        const int mid = (it.b + it.e) / 2;
        std::nth_element(idx.begin() + it.b, idx.begin() + mid, idx.begin() + it.e,
                         [&](int a, int b) { return cen[a * 3 + axis] < cen[b * 3 + axis]; });
        const int l = (int)s.nodes.size();
        s.nodes.push_back(BvhNode());
        const int r = (int)s.nodes.size();
        s.nodes.push_back(BvhNode());
        nd.left = l;
        nd.right = r;
        s.nodes[it.node] = nd;
        st.push_back({l, it.b, mid});
        st.push_back({r, mid, it.e});
    }
    std::vector<int> refs2(n);
    for (int i = 0; i < n; i++) refs2[i] = s.refs[idx[i]];
    s.refs.swap(refs2);
    s.useBvh = true;
}

template <class T> static inline bool slab(const BvhNode &nd, const V3<T> &o, const V3<T> &inv, T tMin, T tMax) {
    T t0 = tMin, t1 = tMax;
    for (int k = 0; k < 3; k++) {
        T a = ((T)nd.lo[k] - o[k]) * inv[k];
        T b = ((T)nd.hi[k] - o[k]) * inv[k];
        if (a > b) std::swap(a, b);
        if (!(a <= t1 && b >= t0)) { if (a != a || b != b) continue; return false; }
        t0 = std::max(t0, a);
        t1 = std::min(t1, b);
    }
    return t0 <= t1;
}

// world.cu:5-41 closest hit (triangles then spheres; a later equal-t hit replaces an earlier
// one).  triTest: 0 = Moeller-Trumbore (reference), 1 = watertight (product).
template <class T> static bool testPrim(const Scene<T> &s, int p, const V3<T> &o, const V3<T> &d, T tMin, T tMax, int triTest, Hit<T> *h) {
    T t, u = 0, v = 0;
    if (p >= 0) {
        const bool ok = triTest ? hitWatertight(s.tris[p], o, d, tMin, tMax, &t, &u, &v)
                                : hitMT(s.tris[p], o, d, tMin, tMax, &t, &u, &v);
        if (!ok) return false;
    } else {
        if (!hitSphere(s.sphs[~p], o, d, tMin, tMax, &t)) return false;
    }
    h->t = t; h->u = u; h->v = v; h->prim = p;
    return true;
}

template <class T> static bool worldHit(const Scene<T> &s, const V3<T> &o, const V3<T> &d, T tMin, T tMax, int triTest, bool anyHit, Hit<T> *rec) {
    bool isHit = false;
    T tNear = tMax;
    if (!s.useBvh) {
        for (int i = 0; i < (int)s.tris.size(); i++) {
            Hit<T> h;
            if (testPrim(s, i, o, d, tMin, tNear, triTest, &h)) {
                isHit = true; tNear = h.t; *rec = h;
                if (anyHit) return true;
            }
        }
        for (int i = 0; i < (int)s.sphs.size(); i++) {
            Hit<T> h;
            if (testPrim(s, ~i, o, d, tMin, tNear, triTest, &h)) {
                isHit = true; tNear = h.t; *rec = h;
                if (anyHit) return true;
            }
        }
        return isHit;
    }
    const V3<T> inv(T(1) / d.x, T(1) / d.y, T(1) / d.z);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const BvhNode &nd = s.nodes[stack[--sp]];
        if (!slab(nd, o, inv, tMin, tNear)) continue;
        if (nd.right < 0) {
            for (int i = nd.left; i < nd.left - nd.right; i++) {
                Hit<T> h;
                if (testPrim(s, s.refs[i], o, d, tMin, tNear, triTest, &h)) {
                    isHit = true; tNear = h.t; *rec = h;
                    if (anyHit) return true;
                }
            }
        } else {
            if (sp + 2 > 128) return isHit;
            stack[sp++] = nd.right;
            stack[sp++] = nd.left;
        }
    }
    return isHit;
}

// HitRecord construction: triangle.cu:36-46 / sphere.cu:17-26.
template <class T> struct Record {
    T t;
    V3<T> point, normal, wo;
    int mat, prim;
};
template <class T> static Record<T> makeRecord(const Scene<T> &s, const Hit<T> &h, const V3<T> &o, const V3<T> &d, bool normalizeShading) {
    Record<T> r;
    r.t = h.t;
    r.prim = h.prim;
    if (h.prim >= 0) {
        const Tri<T> &tr = s.tris[h.prim];
        const T b1 = h.u, b2 = h.v;
        r.point = tr.p0 * (T(1) - b1 - b2) + tr.p1 * b1 + tr.p2 * b2;
        r.normal = tr.n0 * (T(1) - b1 - b2) + tr.n1 * b1 + tr.n2 * b2;  // triangle.h:52-56
        if (normalizeShading) r.normal = normalized(r.normal);
        r.mat = tr.mat;
    } else {
        const Sph<T> &sp = s.sphs[~h.prim];
        r.point = o + d * h.t;
        r.normal = normalized((r.point - sp.c) / sp.r);
        r.mat = sp.mat;
    }
    const Frame<T> f(r.normal);
    r.wo = normalized(f.toLocal(-d));
    return r;
}

// lambertian.cu:5-11 front-facing emission; only Lambertians emit (world.h:118-131)
template <class T> static V3<T> recordEmit(const Scene<T> &s, const Record<T> &r) {
    const Material<T> &m = s.mats[r.mat];
    if (m.type != MAT_LAMBERTIAN) return V3<T>(T(0));
    if (r.wo.z >= T(0)) return m.emit;
    return V3<T>(T(0));
}

template <class T> static T triArea(const Tri<T> &t) { return std::fabs(cross(t.p1 - t.p0, t.p2 - t.p0).length() / T(2)); }
template <class T> static T sphArea(const Sph<T> &s) { return (T)(4.0 * kPi) * s.r * s.r; }

// sampler.h:89-125 + measure.h:8-23 (max(0, cos) at the light)
template <class T> static T pdfSceneLights(const Scene<T> &s, const V3<T> &ref, const Record<T> &lr) {
    const T pdfArea = lr.prim >= 0 ? T(1) / triArea(s.tris[lr.prim]) : T(1) / sphArea(s.sphs[~lr.prim]);
    const V3<T> dir = ref - lr.point;
    const V3<T> wo = normalized(dir);
    const T dist = dir.length();
    const T proj = std::max(T(0), dot(lr.normal, wo));
    const T pdfSA = pdfArea * (dist * dist) / proj;
    int choice = (int)s.lights.size() + (s.env.present ? 1 : 0);
    return pdfSA / (T)choice;
}
// sampler.h:127-140
template <class T> static T pdfEnvLight(const Scene<T> &s, const V3<T> &wi, bool refEnv) {
    if (!s.env.present) return T(0);
    return envPdf(s.env, wi, refEnv) * (T(1) / (T)(s.lights.size() + 1));
}

template <class T> struct LightSample {
    V3<T> wi, normal, emitted;
    T distance, pdf;
};
// sampler.h:142-183, 20-87; triangle.h:66-90; sphere.h:30-44; surface_sample.h:13-24
template <class T> static LightSample<T> sampleDirectLights(const Scene<T> &s, const V3<T> &p, float xi1, float xi2, float xi3, bool refEnv) {
    LightSample<T> ls;
    const int nl = (int)s.lights.size();
    const int choiceSize = nl + (s.env.present ? 1 : 0);
    if (choiceSize == 0) {
        // no emitter and no environment: the reference would index lightIndices[-1]
        // (sampler.h:155-163); return a sample that the one-sidedness test rejects
        ls.wi = V3<T>(T(0), T(0), T(1));
        ls.normal = ls.wi;
        ls.distance = T(0);
        ls.pdf = T(0);
        ls.emitted = V3<T>(T(0));
        return ls;
    }
    const int choice = (int)std::min((float)(choiceSize - 1), std::floor(xi1 * (float)choiceSize));
    const T choicePdf = T(1) / (T)choiceSize;
    if (choice < nl) {
        const int L = s.lights[choice];
        V3<T> pt, n;
        T pdfA;
        int mat;
        if (L >= 0) {
            const Tri<T> &tr = s.tris[L];
            const T r1 = xi2, r2 = xi3;
            const T a = T(1) - std::sqrt(r1);
            const T b = std::sqrt(r1) * (T(1) - r2);
            const T c = T(1) - a - b;
            pt = tr.p0 * a + tr.p1 * b + tr.p2 * c;
            n = normalized(cross(tr.p1 - tr.p0, tr.p2 - tr.p0));
            pdfA = T(1) / triArea(tr);
            mat = tr.mat;
        } else {
            const Sph<T> &sp = s.sphs[~L];
            const V3<T> pl = uniformSphere((T)xi2, (T)xi3);
            pt = pl * sp.r + sp.c;
            n = pl;
            pdfA = T(1) / sphArea(sp);
            mat = sp.mat;
        }
        const V3<T> dir = pt - p;
        const T dist = dir.length();
        // surface_sample.h:13-24
        const V3<T> lwo = -normalized(dir);
        const T proj = std::fabs(dot(n, lwo));
        const T sa = proj == T(0) ? T(0) : pdfA * (dist * dist) / proj;
        ls.wi = normalized(dir);
        ls.normal = n;
        ls.distance = dist;
        ls.pdf = sa * choicePdf;
        ls.emitted = s.mats[mat].emit;
        return ls;
    }
    T pdf;
    const V3<T> wi = envSample(s.env, xi2, xi3, refEnv, &pdf);
    ls.wi = normalized(wi);
    ls.normal = -ls.wi;
    ls.distance = (T)1e16;
    ls.pdf = pdf * choicePdf;
    ls.emitted = envEmit(s.env, wi);
    return ls;
}

template <class T> static T balance(T p1, T p2) { return p1 / (p1 + p2); }  // mis.h:5-8

struct Opts {
    int maxDepth;
    int nee;
    int mode;      // 0 = reference (re-traced BSDF ray), 1 = merged (product wavefront)
    int triTest;   // 0 = MT, 1 = watertight
    int refEnv;    // 1 = CMF texel-centre env sampling (reference), 0 = alias (product)
    int misSample; // 1 = light-sample MIS weight uses bsdfSamplePdf (weights sum to one)
};

struct RayCount { int64_t rays = 0; };

// path_tracer.cu:38-184 (NEE + MIS) and :186-233 (naive).  mode 1 merges the BSDF re-trace
// of directSampleBSDF with the continuation ray (one trace at tMin 1e-3) and ends paths
// whose throughput became zero; this is the product's wavefront schedule.
template <class T> static V3<T> Li(const Scene<T> &s, V3<T> o, V3<T> d, Pcg32 &rng, const Opts &op, RayCount &rc) {
    if (op.maxDepth == 0) return V3<T>(T(0));
    const bool norm = op.mode == 1;
    V3<T> beta(T(1)), result(T(0));
    Hit<T> h;
    rc.rays++;
    if (!worldHit(s, o, d, T(0), (T)FLT_MAX, op.triTest, false, &h)) return envEmit(s.env, d);
    Record<T> rec = makeRecord(s, h, o, d, norm);
    {
        const V3<T> e = recordEmit(s, rec);
        if (!e.isZero()) result += e * beta;
    }
    for (int path = 1; path < op.maxDepth; path++) {
        const Frame<T> fr(rec.normal);
        const Material<T> &m = s.mats[rec.mat];
        float x1 = 0.f, x2 = 0.f;
        const int dims = bsdfSampleDims(m.type);
        if (dims >= 1) x1 = rng.uniform();
        if (dims >= 2) x2 = rng.uniform();
        const BSDFSample<T> bs = bsdfSample(m, rec.wo, (T)x1, (T)x2);
        if (op.nee) {
            V3<T> direct(T(0));
            // directSampleLights (path_tracer.cu:83-126)
            if (!bs.isDelta) {
                const float l1 = rng.uniform(), l2 = rng.uniform(), l3 = rng.uniform();
                const LightSample<T> ls = sampleDirectLights(s, rec.point, l1, l2, l3, op.refEnv != 0);
                if (dot(ls.normal, ls.wi) < T(0) && ls.pdf > T(0)) {
                    Hit<T> oh;
                    rc.rays++;
                    if (!worldHit(s, rec.point, ls.wi, (T)1e-4, ls.distance - (T)2e-4, op.triTest, true, &oh)) {
                        const V3<T> wiL = fr.toLocal(ls.wi);
                        const T pB = op.misSample ? bsdfSamplePdf(m, rec.wo, wiL) : bsdfPdf(m, rec.wo, wiL);
                        const T w = balance(ls.pdf, pB);
                        direct += V3<T>(T(1)) * ls.emitted * w * bsdfF(m, rec.wo, wiL) * absDot(rec.normal, ls.wi) / ls.pdf;
                    }
                }
            }
            if (op.mode == 0) {
                // directSampleBSDF (path_tracer.cu:38-81): separate trace at tMin 1e-4
                if (!bs.f.isZero()) {
                    const V3<T> bd = fr.toWorld(bs.wi);
                    Hit<T> bh;
                    rc.rays++;
                    if (worldHit(s, rec.point, bd, (T)1e-4, (T)FLT_MAX, op.triTest, false, &bh)) {
                        const Record<T> br = makeRecord(s, bh, rec.point, bd, norm);
                        const V3<T> e = recordEmit(s, br);
                        if (!e.isZero()) {
                            const T lp = pdfSceneLights(s, rec.point, br);
                            const T w = bs.isDelta ? T(1) : balance(bs.pdf, lp);
                            direct += e * w * bs.f * absCosTheta(bs.wi) / bs.pdf;
                        }
                    } else {
                        const T lp = pdfEnvLight(s, bd, op.refEnv != 0);
                        const T w = bs.isDelta ? T(1) : balance(bs.pdf, lp);
                        direct += envEmit(s.env, bd) * w * bs.f * absCosTheta(bs.wi) / bs.pdf;
                    }
                }
                result += beta * direct;
                beta *= bs.f * absCosTheta(bs.wi) / bs.pdf;
                const V3<T> nd = fr.toWorld(bs.wi);
                rc.rays++;
                if (!worldHit(s, rec.point, nd, (T)1e-3, (T)FLT_MAX, op.triTest, false, &h)) return result;
                rec = makeRecord(s, h, rec.point, nd, norm);
            } else {
                result += beta * direct;
                if (bs.f.isZero() || !(bs.pdf > T(0))) return result;
                const V3<T> nbeta = beta * (bs.f * absCosTheta(bs.wi) / bs.pdf);
                const V3<T> nd = fr.toWorld(bs.wi);
                rc.rays++;
                if (!worldHit(s, rec.point, nd, (T)1e-3, (T)FLT_MAX, op.triTest, false, &h)) {
                    const T lp = pdfEnvLight(s, nd, op.refEnv != 0);
                    const T w = bs.isDelta ? T(1) : balance(bs.pdf, lp);
                    result += nbeta * envEmit(s.env, nd) * w;
                    return result;
                }
                const Record<T> nr = makeRecord(s, h, rec.point, nd, norm);
                const V3<T> e = recordEmit(s, nr);
                if (!e.isZero()) {
                    const T lp = pdfSceneLights(s, rec.point, nr);
                    const T w = bs.isDelta ? T(1) : balance(bs.pdf, lp);
                    result += nbeta * e * w;
                }
                beta = nbeta;
                rec = nr;
            }
        } else {
            // calculateLiNaive (path_tracer.cu:186-233)
            if (op.mode == 1 && (bs.f.isZero() || !(bs.pdf > T(0)))) return result;
            beta *= V3<T>(T(1)) * bs.f * absCosTheta(bs.wi) / bs.pdf;
            const V3<T> nd = fr.toWorld(bs.wi);
            rc.rays++;
            if (!worldHit(s, rec.point, nd, (T)1e-3, (T)FLT_MAX, op.triTest, false, &h))
                return result + envEmit(s.env, nd) * beta;
            rec = makeRecord(s, h, rec.point, nd, norm);
            const V3<T> e = recordEmit(s, rec);
            if (!e.isZero()) result += e * beta;
        }
    }
    return result;
}

}  // namespace orc

// ========================================================================================
// C interface (ctypes)
// ========================================================================================
using namespace orc;

extern "C" {

typedef struct {
    int64_t n_tris;
    const float *tri_pos;   // 9 per triangle
    const float *tri_nrm;   // 9 per triangle, or NULL for geometric normals
    const int *tri_mat;
    int64_t n_spheres;
    const float *sph;       // cx cy cz r
    const int *sph_mat;
    int64_t n_mats;
    const int *mat_type;
    const float *mat_albedo;  // 3 per material
    const float *mat_emit;    // 3 per material
    const float *mat_ior;
    const float *mat_alpha;
    int env_w, env_h;
    const float *env_rgba;  // NULL: no environment
    int env_sin_weight;
    float cam_origin[3], cam_target[3], cam_up[3];
    float cam_vfov;
    int cam_flip;
} or_scene_desc;

typedef struct {
    int width, height;
    int max_depth, nee, mode, tri_test, ref_env;
    int use_bvh;
    int precision;  // 32 or 64
    int threads;
    uint64_t seed;
    int64_t sample_start;
    int spp;
    int row_begin, row_end;  // rows [begin, end) (end <= 0: all)
    int pixel_stride;        // render every k-th pixel of each row (>= 1)
    int mis_sample;          // 1: consistent MIS (light weight from the sampling density)
} or_render_params;

}  // extern "C"

struct OrHandle {
    Scene<float> sf;
    Scene<double> sd;
};

template <class T> static void loadScene(Scene<T> &s, const or_scene_desc *d, int W, int H) {
    s.tris.resize((size_t)d->n_tris);
    for (int64_t i = 0; i < d->n_tris; i++) {
        Tri<T> &t = s.tris[i];
        const float *p = d->tri_pos + i * 9;
        t.p0 = V3<T>(p[0], p[1], p[2]);
        t.p1 = V3<T>(p[3], p[4], p[5]);
        t.p2 = V3<T>(p[6], p[7], p[8]);
        if (d->tri_nrm) {
            const float *n = d->tri_nrm + i * 9;
            t.n0 = V3<T>(n[0], n[1], n[2]);
            t.n1 = V3<T>(n[3], n[4], n[5]);
            t.n2 = V3<T>(n[6], n[7], n[8]);
        } else {
            // obj_parser.cpp:185-194: flat normal from the winding
            const V3<T> g = normalized(cross(t.p1 - t.p0, t.p2 - t.p0));
            t.n0 = t.n1 = t.n2 = g;
        }
        t.mat = d->tri_mat[i];
    }
    s.sphs.resize((size_t)d->n_spheres);
    for (int64_t i = 0; i < d->n_spheres; i++) {
        s.sphs[i].c = V3<T>(d->sph[i * 4], d->sph[i * 4 + 1], d->sph[i * 4 + 2]);
        s.sphs[i].r = d->sph[i * 4 + 3];
        s.sphs[i].mat = d->sph_mat[i];
    }
    s.mats.resize((size_t)d->n_mats);
    for (int64_t i = 0; i < d->n_mats; i++) {
        Material<T> &m = s.mats[i];
        m.type = d->mat_type[i];
        m.albedo = V3<T>(d->mat_albedo[i * 3], d->mat_albedo[i * 3 + 1], d->mat_albedo[i * 3 + 2]);
        m.emit = V3<T>(d->mat_emit[i * 3], d->mat_emit[i * 3 + 1], d->mat_emit[i * 3 + 2]);
        m.ior = d->mat_ior[i];
        m.alpha = d->mat_alpha[i];
    }
    // scene_data.cu:92-115: emitters = primitives whose material has non-zero emission
    // (only Lambertian parameters carry emission, params.h:37-70)
    s.lights.clear();
    for (int64_t i = 0; i < d->n_tris; i++) {
        const Material<T> &m = s.mats[s.tris[i].mat];
        if (m.type == MAT_LAMBERTIAN && !m.emit.isZero()) s.lights.push_back((int)i);
    }
    for (int64_t i = 0; i < d->n_spheres; i++) {
        const Material<T> &m = s.mats[s.sphs[i].mat];
        if (m.type == MAT_LAMBERTIAN && !m.emit.isZero()) s.lights.push_back(~(int)i);
    }
    s.env = EnvTables();
    if (d->env_rgba && d->env_w > 0 && d->env_h > 0) buildEnv(s.env, d->env_rgba, d->env_w, d->env_h, d->env_sin_weight != 0);
    s.cam = makeCamera<T>(d->cam_origin, d->cam_target, d->cam_up, d->cam_vfov, d->cam_flip, W, H);
}

extern "C" void *or_scene_create(const or_scene_desc *d, int width, int height, int build_bvh) {
    OrHandle *h = new OrHandle();
    loadScene(h->sf, d, width, height);
    loadScene(h->sd, d, width, height);
    if (build_bvh) {
        buildBvh(h->sf);
        buildBvh(h->sd);
    }
    return h;
}
extern "C" void or_scene_destroy(void *h) { delete (OrHandle *)h; }

template <class T> static void renderImpl(const Scene<T> &s, const or_render_params *p, double *sum, double *sumsq, int64_t *rays) {
    Opts op{p->max_depth, p->nee, p->mode, p->tri_test, p->ref_env, p->mis_sample};
    const int W = p->width, H = p->height;
    const int rb = p->row_begin, re = p->row_end > 0 ? p->row_end : H;
    const int stride = p->pixel_stride > 0 ? p->pixel_stride : 1;
    int nthreads = p->threads > 0 ? p->threads : (int)std::thread::hardware_concurrency();
    if (nthreads < 1) nthreads = 1;
    std::atomic<int> nextRow(rb);
    std::atomic<int64_t> rayTotal(0);
    auto work = [&]() {
        RayCount rc;
        for (;;) {
            const int row = nextRow.fetch_add(1);
            if (row >= re) break;
            for (int col = 0; col < W; col += stride) {
                const uint32_t pix = (uint32_t)(row * W + col);
                double acc[3] = {0, 0, 0}, acc2[3] = {0, 0, 0};
                for (int k = 0; k < p->spp; k++) {
                    Pcg32 rng = pathRng(p->seed, pix, (uint32_t)(p->sample_start + k));
                    const float sx = rng.uniform(), sy = rng.uniform();
                    V3<T> o, d;
                    cameraRay(s.cam, row, col, (T)sx, (T)sy, &o, &d);
                    const V3<T> L = Li(s, o, d, rng, op, rc);
                    for (int c = 0; c < 3; c++) {
                        acc[c] += (double)L[c];
                        acc2[c] += (double)L[c] * (double)L[c];
                    }
                }
                for (int c = 0; c < 3; c++) {
                    sum[(size_t)pix * 3 + c] += acc[c];
                    if (sumsq) sumsq[(size_t)pix * 3 + c] += acc2[c];
                }
            }
        }
        rayTotal += rc.rays;
    };
    std::vector<std::thread> th;
    for (int i = 0; i < nthreads; i++) th.emplace_back(work);
    for (auto &t : th) t.join();
    if (rays) *rays = rayTotal.load();
}

extern "C" {

// Accumulates per-pixel FP64 sums (and sums of squares if non-NULL) of the radiance of
// samples [sample_start, sample_start + spp) into sum[3*W*H].
int or_render(void *hv, const or_render_params *p, double *sum, double *sumsq, int64_t *rays) {
    OrHandle *h = (OrHandle *)hv;
    if (p->precision == 64) renderImpl(h->sd, p, sum, sumsq, rays);
    else renderImpl(h->sf, p, sum, sumsq, rays);
    return 0;
}

// Radiance of single paths (pixel, sample) in FP32/FP64 -> out[3n]
int or_trace_paths(void *hv, const or_render_params *p, const int *pix, const int *smp, int64_t n, double *out) {
    OrHandle *h = (OrHandle *)hv;
    Opts op{p->max_depth, p->nee, p->mode, p->tri_test, p->ref_env, p->mis_sample};
    RayCount rc;
    for (int64_t i = 0; i < n; i++) {
        const int row = pix[i] / p->width, col = pix[i] % p->width;
        Pcg32 rng = pathRng(p->seed, (uint32_t)pix[i], (uint32_t)smp[i]);
        const float sx = rng.uniform(), sy = rng.uniform();
        if (p->precision == 64) {
            V3<double> o, d;
            cameraRay(h->sd.cam, row, col, (double)sx, (double)sy, &o, &d);
            const V3<double> L = Li(h->sd, o, d, rng, op, rc);
            out[i * 3] = L.x; out[i * 3 + 1] = L.y; out[i * 3 + 2] = L.z;
        } else {
            V3<float> o, d;
            cameraRay(h->sf.cam, row, col, sx, sy, &o, &d);
            const V3<float> L = Li(h->sf, o, d, rng, op, rc);
            out[i * 3] = L.x; out[i * 3 + 1] = L.y; out[i * 3 + 2] = L.z;
        }
    }
    return 0;
}

// Per-pixel FP64 sums (and sums of squares) of samples [sample_start, sample_start + spp) for
// a list of pixels: sum[3n], sumsq[3n].  Work is split into blocks of up to 1024 samples that
// the threads take in turn; block partials are reduced in a fixed order, so the result does
// not depend on the thread count.
int or_render_pixels(void *hv, const or_render_params *p, const int *pix, int64_t n, double *sum, double *sumsq, int64_t *rays) {
    OrHandle *h = (OrHandle *)hv;
    const Opts op{p->max_depth, p->nee, p->mode, p->tri_test, p->ref_env, p->mis_sample};
    const int blk = 1024;
    const int64_t nb = (p->spp + blk - 1) / blk;
    const int64_t items = n * nb;
    std::vector<double> part((size_t)items * 6, 0.0);
    int nthreads = p->threads > 0 ? p->threads : (int)std::thread::hardware_concurrency();
    if (nthreads < 1) nthreads = 1;
    std::atomic<int64_t> next(0), rayTotal(0);
    auto work = [&]() {
        RayCount rc;
        for (;;) {
            const int64_t it = next.fetch_add(1);
            if (it >= items) break;
            const int64_t i = it / nb, b = it % nb;
            const int row = pix[i] / p->width, col = pix[i] % p->width;
            const int k0 = (int)(b * blk), k1 = std::min(p->spp, k0 + blk);
            double *acc = &part[(size_t)it * 6];
            for (int k = k0; k < k1; k++) {
                Pcg32 rng = pathRng(p->seed, (uint32_t)pix[i], (uint32_t)(p->sample_start + k));
                const float sx = rng.uniform(), sy = rng.uniform();
                double L[3];
                if (p->precision == 64) {
                    V3<double> o, d;
                    cameraRay(h->sd.cam, row, col, (double)sx, (double)sy, &o, &d);
                    const V3<double> v = Li(h->sd, o, d, rng, op, rc);
                    L[0] = v.x; L[1] = v.y; L[2] = v.z;
                } else {
                    V3<float> o, d;
                    cameraRay(h->sf.cam, row, col, sx, sy, &o, &d);
                    const V3<float> v = Li(h->sf, o, d, rng, op, rc);
                    L[0] = v.x; L[1] = v.y; L[2] = v.z;
                }
                for (int c = 0; c < 3; c++) {
                    acc[c] += L[c];
                    acc[3 + c] += L[c] * L[c];
                }
            }
        }
        rayTotal += rc.rays;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
    for (int64_t i = 0; i < n; i++) {
        for (int c = 0; c < 3; c++) {
            double a = 0.0, q = 0.0;
            for (int64_t b = 0; b < nb; b++) {
                a += part[(size_t)(i * nb + b) * 6 + c];
                q += part[(size_t)(i * nb + b) * 6 + 3 + c];
            }
            sum[i * 3 + c] = a;
            if (sumsq) sumsq[i * 3 + c] = q;
        }
    }
    if (rays) *rays = rayTotal.load();
    return 0;
}

// Closest / any hit for n rays (8 floats each: o.xyz tmin d.xyz tmax) on the FP32 scene.
// out: 4 per ray: t, u, v, prim (as float bits of int; -2^31 for a miss)
int or_intersect(void *hv, const float *rays, int64_t n, int tri_test, int any_hit, float *out) {
    OrHandle *h = (OrHandle *)hv;
    for (int64_t i = 0; i < n; i++) {
        const float *r = rays + i * 8;
        Hit<float> hit;
        const bool ok = worldHit(h->sf, V3<float>(r[0], r[1], r[2]), V3<float>(r[4], r[5], r[6]), r[3], r[7], tri_test, any_hit != 0, &hit);
        int prim = ok ? hit.prim : INT32_MIN;
        out[i * 4] = ok ? hit.t : -1.f;
        out[i * 4 + 1] = ok ? hit.u : 0.f;
        out[i * 4 + 2] = ok ? hit.v : 0.f;
        std::memcpy(&out[i * 4 + 3], &prim, 4);
    }
    return 0;
}

// Single triangle tests in float (for bit-level parity): tris[9n], rays[8n] -> out[4n]
int or_tri_test(const float *tris, const float *rays, int64_t n, int tri_test, float *out) {
    for (int64_t i = 0; i < n; i++) {
        Tri<float> t;
        const float *p = tris + i * 9;
        t.p0 = V3<float>(p[0], p[1], p[2]);
        t.p1 = V3<float>(p[3], p[4], p[5]);
        t.p2 = V3<float>(p[6], p[7], p[8]);
        const float *r = rays + i * 8;
        float tt = 0, u = 0, v = 0;
        const bool ok = tri_test ? hitWatertight(t, V3<float>(r[0], r[1], r[2]), V3<float>(r[4], r[5], r[6]), r[3], r[7], &tt, &u, &v)
                                 : hitMT(t, V3<float>(r[0], r[1], r[2]), V3<float>(r[4], r[5], r[6]), r[3], r[7], &tt, &u, &v);
        out[i * 4] = ok ? 1.f : 0.f;
        out[i * 4 + 1] = tt;
        out[i * 4 + 2] = u;
        out[i * 4 + 3] = v;
    }
    return 0;
}

// BSDF queries in FP32 (precision 32) or FP64 (64).  mat: type, albedo[3], ior, alpha per entry
// (6 floats).  wo/wi: 3 floats each.  out_f[3n], out_pdf[n]: pdf() for precision 32/64, the
// sampling density bsdfSamplePdf for precision 33/65.
int or_bsdf_eval(const float *mat, const float *wo, const float *wi, int64_t n, int precision, double *out_f, double *out_pdf) {
    for (int64_t i = 0; i < n; i++) {
        const float *m = mat + i * 6;
        if (precision == 64 || precision == 65) {
            Material<double> M{(int)m[0], V3<double>(m[1], m[2], m[3]), V3<double>(0.0), (double)m[4], (double)m[5]};
            const V3<double> o(wo[i * 3], wo[i * 3 + 1], wo[i * 3 + 2]), w(wi[i * 3], wi[i * 3 + 1], wi[i * 3 + 2]);
            const V3<double> f = bsdfF(M, o, w);
            out_f[i * 3] = f.x; out_f[i * 3 + 1] = f.y; out_f[i * 3 + 2] = f.z;
            out_pdf[i] = precision == 65 ? bsdfSamplePdf(M, o, w) : bsdfPdf(M, o, w);
        } else {
            Material<float> M{(int)m[0], V3<float>(m[1], m[2], m[3]), V3<float>(0.f), m[4], m[5]};
            const V3<float> o(wo[i * 3], wo[i * 3 + 1], wo[i * 3 + 2]), w(wi[i * 3], wi[i * 3 + 1], wi[i * 3 + 2]);
            const V3<float> f = bsdfF(M, o, w);
            out_f[i * 3] = f.x; out_f[i * 3 + 1] = f.y; out_f[i * 3 + 2] = f.z;
            out_pdf[i] = precision == 33 ? bsdfSamplePdf(M, o, w) : bsdfPdf(M, o, w);
        }
    }
    return 0;
}

// BSDF sampling in FP32/FP64: xi[2n] -> out: wi[3], pdf, f[3], isDelta (8 per entry)
int or_bsdf_sample(const float *mat, const float *wo, const float *xi, int64_t n, int precision, double *out) {
    for (int64_t i = 0; i < n; i++) {
        const float *m = mat + i * 6;
        double *r = out + i * 8;
        if (precision == 64) {
            Material<double> M{(int)m[0], V3<double>(m[1], m[2], m[3]), V3<double>(0.0), (double)m[4], (double)m[5]};
            const BSDFSample<double> s = bsdfSample(M, V3<double>(wo[i * 3], wo[i * 3 + 1], wo[i * 3 + 2]), (double)xi[i * 2], (double)xi[i * 2 + 1]);
            r[0] = s.wi.x; r[1] = s.wi.y; r[2] = s.wi.z; r[3] = s.pdf; r[4] = s.f.x; r[5] = s.f.y; r[6] = s.f.z; r[7] = s.isDelta;
        } else {
            Material<float> M{(int)m[0], V3<float>(m[1], m[2], m[3]), V3<float>(0.f), m[4], m[5]};
            const BSDFSample<float> s = bsdfSample(M, V3<float>(wo[i * 3], wo[i * 3 + 1], wo[i * 3 + 2]), xi[i * 2], xi[i * 2 + 1]);
            r[0] = s.wi.x; r[1] = s.wi.y; r[2] = s.wi.z; r[3] = s.pdf; r[4] = s.f.x; r[5] = s.f.y; r[6] = s.f.z; r[7] = s.isDelta;
        }
    }
    return 0;
}

// Fresnel reflectance (fresnel.h:17-40) in FP64
double or_fresnel(double cosI, double etaI, double etaT) { return fresnelDielectric(cosI, etaI, etaT); }

// 1-D distribution (reference CMF): pmf[n], total mass, samples for xis
int or_dist1d(const float *values, int n, float *pmf_out, float *total_out, const float *xis, int nxi, int *idx_out, float *xpmf_out) {
    Distribution1D d = buildDistribution(values, n);
    for (int i = 0; i < n; i++) pmf_out[i] = distPmf(d, i);
    *total_out = d.totalMass;
    for (int i = 0; i < nxi; i++) idx_out[i] = distSample(d, xis[i], &xpmf_out[i]);
    return 0;
}
// 2-D distribution: pmf[w*h]; samples from xi pairs -> (x, y, pmf)
int or_dist2d(const float *values, int w, int h, float *pmf_out, const float *xis, int nxi, int *xy_out, float *xpmf_out) {
    Distribution2D d = buildDistribution2D(values, w, h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) pmf_out[y * w + x] = dist2Pmf(d, x, y);
    for (int i = 0; i < nxi; i++) dist2Sample(d, xis[2 * i], xis[2 * i + 1], &xy_out[2 * i], &xy_out[2 * i + 1], &xpmf_out[i]);
    return 0;
}
// phi-theta pdf (canonical) on a grid of nx*ny query points (canonical centres)
int or_phitheta_pdf_canonical(const float *values, int w, int h, int nx, int ny, float *out) {
    Distribution2D d = buildDistribution2D(values, w, h);
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) out[y * nx + x] = phiThetaPdfCanonical(d, (x + 0.5f) / nx, (y + 0.5f) / ny);
    return 0;
}
// environment pdf for directions (FP32; ref=1 CMF, ref=0 alias tables)
int or_env_pdf(const float *rgba, int w, int h, int sin_weight, const float *dirs, int64_t n, int ref, float *out) {
    EnvTables e;
    buildEnv(e, rgba, w, h, sin_weight != 0);
    for (int64_t i = 0; i < n; i++) out[i] = envPdf(e, V3<float>(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]), ref != 0);
    return 0;
}
// environment samples: xi[2n] -> dir[3n], pdf[n], emit[3n]
int or_env_sample(const float *rgba, int w, int h, int sin_weight, const float *xi, int64_t n, int ref, float *dir, float *pdf, float *emit) {
    EnvTables e;
    buildEnv(e, rgba, w, h, sin_weight != 0);
    for (int64_t i = 0; i < n; i++) {
        float p;
        const V3<float> d = envSample(e, xi[2 * i], xi[2 * i + 1], ref != 0, &p);
        dir[i * 3] = d.x; dir[i * 3 + 1] = d.y; dir[i * 3 + 2] = d.z;
        pdf[i] = p;
        const V3<float> L = envEmit(e, d);
        emit[i * 3] = L.x; emit[i * 3 + 1] = L.y; emit[i * 3 + 2] = L.z;
    }
    return 0;
}
// alias tables: row prob/alias [w*h], marginal prob/alias [h], texel pdf [w*h]
int or_env_alias(const float *rgba, int w, int h, int sin_weight, double *row_prob, int *row_alias, double *marg_prob, int *marg_alias, float *texel_pdf) {
    EnvTables e;
    buildEnv(e, rgba, w, h, sin_weight != 0);
    std::memcpy(row_prob, e.rowProb.data(), sizeof(double) * w * h);
    std::memcpy(row_alias, e.rowAlias.data(), sizeof(int) * w * h);
    std::memcpy(marg_prob, e.margProb.data(), sizeof(double) * h);
    std::memcpy(marg_alias, e.margAlias.data(), sizeof(int) * h);
    std::memcpy(texel_pdf, e.texelPdf.data(), sizeof(float) * w * h);
    return 0;
}
// generic alias build over n weights (FP64)
int or_alias_build(const double *w, int n, double *prob, int *alias, double *sum) {
    buildAliasRow(w, n, prob, alias, sum);
    return 0;
}
// PCG32 outputs for (initstate, initseq)
int or_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t *out) {
    Pcg32 r;
    r.seed(initstate, initseq);
    for (int i = 0; i < n; i++) out[i] = r.next();
    return 0;
}
// first n uniforms of the path stream (seed, pixel, sample)
int or_path_uniforms(uint64_t seed, uint32_t pixel, uint32_t sample, int n, float *out) {
    Pcg32 r = pathRng(seed, pixel, sample);
    for (int i = 0; i < n; i++) out[i] = r.uniform();
    return 0;
}
// camera rays (FP32) for (row, col, sx, sy) -> o[3], d[3]
int or_camera_rays(const float o[3], const float t[3], const float u[3], float vfov, int flip, int W, int H, const float *q, int64_t n, float *out) {
    Camera<float> c = makeCamera<float>(o, t, u, vfov, flip, W, H);
    for (int64_t i = 0; i < n; i++) {
        V3<float> ro, rd;
        cameraRay(c, (int)q[i * 4], (int)q[i * 4 + 1], q[i * 4 + 2], q[i * 4 + 3], &ro, &rd);
        out[i * 6] = ro.x; out[i * 6 + 1] = ro.y; out[i * 6 + 2] = ro.z;
        out[i * 6 + 3] = rd.x; out[i * 6 + 4] = rd.y; out[i * 6 + 5] = rd.z;
    }
    return 0;
}
// framebuffer.cu:7-41 running mean in FP32 + color.h:7-28 sRGB -> uchar
int or_framebuffer_update(const float *pass, float *mean, int64_t npix, int passSamples, int previousSamples, uint8_t *rgba) {
    for (int64_t i = 0; i < npix; i++) {
        for (int c = 0; c < 3; c++) {
            const float p = pass[i * 3 + c];
            float v;
            if (previousSamples > 0) {
                const int total = previousSamples + passSamples;
                v = mean[i * 3 + c] * previousSamples / total + p * (1.f * passSamples / total);
            } else {
                v = p;
            }
            mean[i * 3 + c] = v;
            // color.h:7-28 with the float power correctly rounded (evaluated in FP64, rounded
            // once), so the bytes are a function of the radiance alone
            const float s = v <= 0.0031308f ? v * 12.92f : 1.055f * (float)std::pow((double)v, (double)(1.f / 2.2f)) - 0.055f;
            rgba[i * 4 + c] = (uint8_t)(std::max(0.f, std::min(1.f, s)) * 255);
        }
        rgba[i * 4 + 3] = 255;
    }
    return 0;
}

}  // extern "C"
