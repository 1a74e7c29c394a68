// pt_scene.cpp -- host scene description, procedural workloads and the scene part of the
// C ABI.  Host-only C++ (no device code).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <unordered_map>

#include "pt_internal.h"

namespace pt {

static thread_local std::string g_lastError;

void setError(const std::string &msg) { g_lastError = msg; }
int fail(int code, const std::string &msg) {
    g_lastError = msg;
    return code;
}

bool isEmissive(const HostMaterial &m) {
    // params.h:37-70: only Lambertian parameters carry emission
    return m.type == PT_LAMBERTIAN && (m.emit[0] != 0.f || m.emit[1] != 0.f || m.emit[2] != 0.f);
}

std::vector<int> emitterList(const pt_scene &s) {
    std::vector<int> out;
    const int64_t nt = (int64_t)s.triMat.size();
    for (int64_t i = 0; i < nt; i++)
        if (isEmissive(s.mats[s.triMat[i]])) out.push_back((int)i);
    for (size_t i = 0; i < s.sphMat.size(); i++)
        if (isEmissive(s.mats[s.sphMat[i]])) out.push_back(~(int)i);
    return out;
}

static void norm3(float v[3]) {
    const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l; v[1] /= l; v[2] /= l;
}
static void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = -(a[0] * b[2] - a[2] * b[0]);
    o[2] = a[0] * b[1] - a[1] * b[0];
}

bool lookAt(const float o[3], const float t[3], const float u[3], int flip, float m[3][4]) {
    float dir[3] = {o[0] - t[0], o[1] - t[1], o[2] - t[2]};
    norm3(dir);
    float up[3] = {u[0], u[1], u[2]};
    norm3(up);
    if (dir[0] == u[0] && dir[1] == u[1] && dir[2] == u[2]) return false;  // transform.cu:15-18
    float xa[3], ya[3];
    cross3(up, dir, xa);
    const float lx = sqrtf(xa[0] * xa[0] + xa[1] * xa[1] + xa[2] * xa[2]);
    if (!(lx > 0.f)) return false;
    norm3(xa);
    cross3(dir, xa, ya);
    norm3(ya);
    const float sign = flip ? -1.f : 1.f;
    for (int r = 0; r < 3; r++) {
        m[r][0] = sign * xa[r];
        m[r][1] = ya[r];
        m[r][2] = dir[r];
        m[r][3] = o[r];
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// Procedural workloads
// ---------------------------------------------------------------------------------------
static int addMat(pt_scene &s, int type, float r, float g, float b, float er, float eg, float eb, float ior, float alpha) {
    HostMaterial m;
    m.type = type;
    m.albedo[0] = r; m.albedo[1] = g; m.albedo[2] = b;
    m.emit[0] = er; m.emit[1] = eg; m.emit[2] = eb;
    m.ior = ior;
    m.alpha = alpha;
    s.mats.push_back(m);
    return (int)s.mats.size() - 1;
}
static void addTri(pt_scene &s, const float a[3], const float b[3], const float c[3], int mat) {
    for (int k = 0; k < 3; k++) s.pos.push_back(a[k]);
    for (int k = 0; k < 3; k++) s.pos.push_back(b[k]);
    for (int k = 0; k < 3; k++) s.pos.push_back(c[k]);
    s.triMat.push_back(mat);
}
// quad a b c d (counter-clockwise seen from the side the normal points to)
static void addQuad(pt_scene &s, const float a[3], const float b[3], const float c[3], const float d[3], int mat) {
    addTri(s, a, b, c, mat);
    addTri(s, a, c, d, mat);
}
static void addSphere(pt_scene &s, float x, float y, float z, float r, int mat) {
    s.sph.push_back(x); s.sph.push_back(y); s.sph.push_back(z); s.sph.push_back(r);
    s.sphMat.push_back(mat);
}

// Cornell box: x in [-1, 1], y in [0, 2], z in [-1, 1], open towards +z, walls facing in.
static void presetCornell(pt_scene &s) {
    const int grey = addMat(s, PT_LAMBERTIAN, 0.75f, 0.75f, 0.75f, 0, 0, 0, 1.f, 0.1f);
    const int red = addMat(s, PT_LAMBERTIAN, 0.63f, 0.065f, 0.05f, 0, 0, 0, 1.f, 0.1f);
    const int green = addMat(s, PT_LAMBERTIAN, 0.14f, 0.45f, 0.09f, 0, 0, 0, 1.f, 0.1f);
    const int light = addMat(s, PT_LAMBERTIAN, 0.78f, 0.78f, 0.78f, 17.f, 12.f, 4.f, 1.f, 0.1f);
    const int diffuse = addMat(s, PT_LAMBERTIAN, 0.75f, 0.75f, 0.75f, 0, 0, 0, 1.f, 0.1f);
    const int glass = addMat(s, PT_GLASS, 0, 0, 0, 0, 0, 0, 1.5f, 0.1f);
    const float A[3] = {-1, 0, 1}, B[3] = {1, 0, 1}, C[3] = {1, 0, -1}, D[3] = {-1, 0, -1};
    const float E[3] = {-1, 2, 1}, F[3] = {1, 2, 1}, G[3] = {1, 2, -1}, H[3] = {-1, 2, -1};
    addQuad(s, A, B, C, D, grey);   // floor, normal +y
    addQuad(s, E, H, G, F, grey);   // ceiling, normal -y
    addQuad(s, D, C, G, H, grey);   // back wall, normal +z
    addQuad(s, A, D, H, E, red);    // left wall, normal +x
    addQuad(s, B, F, G, C, green);  // right wall, normal -x
    // ceiling light (CornellBox-Original.obj extents), normal -y
    const float L0[3] = {-0.24f, 1.98f, 0.16f}, L1[3] = {-0.24f, 1.98f, -0.22f};
    const float L2[3] = {0.23f, 1.98f, -0.22f}, L3[3] = {0.23f, 1.98f, 0.16f};
    addQuad(s, L0, L1, L2, L3, light);
    addSphere(s, -0.45f, 0.3f, -0.35f, 0.3f, diffuse);
    addSphere(s, 0.45f, 0.3f, 0.3f, 0.3f, glass);
    s.cam = HostCamera();  // scene.cu:176-184
}

// Seeded lattice value noise in [0, 1): hash of the integer lattice, trilinear with a
// smoothstep fade.
static inline float latticeValue(int x, int y, int z, uint64_t seed) {
    uint64_t h = seed ^ 0x9e3779b97f4a7c15ULL;
    h ^= (uint64_t)(uint32_t)x * 0xbf58476d1ce4e5b9ULL;
    h = (h ^ (h >> 29)) * 0x94d049bb133111ebULL;
    h ^= (uint64_t)(uint32_t)y * 0xd6e8feb86659fd93ULL;
    h = (h ^ (h >> 32)) * 0xbf58476d1ce4e5b9ULL;
    h ^= (uint64_t)(uint32_t)z * 0x9e3779b97f4a7c15ULL;
    h = (h ^ (h >> 31)) * 0x94d049bb133111ebULL;
    h ^= h >> 29;
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}
static float valueNoise(float x, float y, float z, uint64_t seed) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float tx = x - fx, ty = y - fy, tz = z - fz;
    tx = tx * tx * (3 - 2 * tx);
    ty = ty * ty * (3 - 2 * ty);
    tz = tz * tz * (3 - 2 * tz);
    float c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) c[a][b][d] = latticeValue(ix + a, iy + b, iz + d, seed);
    const float x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    const float x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    const float x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    const float x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    const float y0 = x00 + (x10 - x00) * ty;
    const float y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}
// 3 octaves, base frequency 4, lacunarity 2, gain 0.5; result in [-1, 1]
static float fbm3(const float p[3], uint64_t seed) {
    float sum = 0.f, amp = 1.f, freq = 4.f, norm = 0.f;
    for (int o = 0; o < 3; o++) {
        sum += amp * valueNoise(p[0] * freq, p[1] * freq, p[2] * freq, seed + 1000003ULL * o);
        norm += amp;
        amp *= 0.5f;
        freq *= 2.f;
    }
    return 2.f * (sum / norm) - 1.f;
}

// Synthetic HDR environment: per-texel log-normal(mu = 0, sigma = 1) white luminance plus a
// Gaussian sun lobe of peak 1e4 and 0.5 degree full width at half maximum.
static void syntheticEnvironment(pt_scene &s, int W, int H, uint64_t seed) {
    s.envW = W;
    s.envH = H;
    s.env.assign((size_t)W * H * 4, 0.f);
    std::mt19937_64 rng(seed ^ 0xE17E17E1ULL);
    std::lognormal_distribution<double> ln(0.0, 1.0);
    const double sunPhi = 60.0 / 180.0 * M_PI, sunElev = 35.0 / 180.0 * M_PI;
    const double sd[3] = {cos(sunElev) * cos(sunPhi), sin(sunElev), cos(sunElev) * sin(sunPhi)};
    const double sigma = (0.5 / 2.354820045) / 180.0 * M_PI;
    for (int y = 0; y < H; y++) {
        const double theta = (y + 0.5) / H * M_PI;
        for (int x = 0; x < W; x++) {
            const double phi = (x + 0.5) / W * 2 * M_PI;
            const double d[3] = {sin(theta) * cos(phi), cos(theta), sin(theta) * sin(phi)};
            const double c = std::min(1.0, std::max(-1.0, d[0] * sd[0] + d[1] * sd[1] + d[2] * sd[2]));
            const double ang = acos(c);
            const double L = ln(rng) + 1e4 * exp(-ang * ang / (2 * sigma * sigma));
            float *t = &s.env[((size_t)y * W + x) * 4];
            t[0] = t[1] = t[2] = (float)L;
            t[3] = 1.f;
        }
    }
}

static void finishIcosphere(pt_scene &s, std::vector<float> &V, const std::vector<int> &F, uint64_t seed);

static void icosahedron(std::vector<float> &V, std::vector<int> &F) {
    const float t = (1.f + sqrtf(5.f)) / 2.f;
    V = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
         0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    F = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
         11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
         3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (size_t i = 0; i < V.size(); i += 3) norm3(&V[i]);
}

// Geodesic displaced sphere of about `target` triangles: each icosahedron face split into
// n^2 triangles (n = round(sqrt(target / 20))), points projected to the unit sphere, then
// displaced and given materials as the subdivided icosphere.  A point's position is summed
// over the face corners in global-index order, so points on a shared edge come out bitwise
// equal from both faces and the mesh stays closed.
static void presetGeodesic(pt_scene &s, uint64_t seed, int64_t target) {
    std::vector<float> C;
    std::vector<int> CF;
    icosahedron(C, CF);
    const int n = std::max(1, (int)llround(sqrt((double)target / 20.0)));
    std::vector<float> V;
    std::vector<int> F;
    V.reserve((size_t)20 * (n + 1) * (n + 2) / 2 * 3);
    F.reserve((size_t)20 * n * n * 3);
    for (int f = 0; f < 20; f++) {
        const int ci[3] = {CF[3 * f], CF[3 * f + 1], CF[3 * f + 2]};
        int order[3] = {0, 1, 2};
        std::sort(order, order + 3, [&](int a, int b) { return ci[a] < ci[b]; });
        const int base = (int)(V.size() / 3);
        // point (i, j): weights (n - i - j, i, j) on corners (A, B, C) of the face
        for (int i = 0; i <= n; i++)
            for (int j = 0; j <= n - i; j++) {
                const int w[3] = {n - i - j, i, j};
                float p[3] = {0.f, 0.f, 0.f};
                for (int k = 0; k < 3; k++) {
                    const int c = order[k];
                    const float wf = (float)w[c] / (float)n;
                    for (int a = 0; a < 3; a++) p[a] += wf * C[3 * ci[c] + a];
                }
                norm3(p);
                V.insert(V.end(), p, p + 3);
            }
        auto idx = [&](int i, int j) { return base + i * (n + 1) - i * (i - 1) / 2 + j; };
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n - i; j++) {
                const int t0[3] = {idx(i, j), idx(i + 1, j), idx(i, j + 1)};
                F.insert(F.end(), t0, t0 + 3);
                if (i + j < n - 1) {
                    const int t1[3] = {idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)};
                    F.insert(F.end(), t1, t1 + 3);
                }
            }
    }
    finishIcosphere(s, V, F, seed);
}

// Displaced icosphere: level-`level` subdivision of the icosahedron (20 * 4^level faces),
// vertices on the unit sphere pushed out radially by 0.05 * fbm3; per-face materials:
// 30% Lambert (albedo U[0.2, 0.8]^3), 70% Beckmann alpha U[0.05, 0.5].
static void presetIcosphere(pt_scene &s, uint64_t seed, int level) {
    std::vector<float> V;
    std::vector<int> F;
    icosahedron(V, F);
    for (int l = 0; l < level; l++) {
        std::unordered_map<uint64_t, int> mid;
        mid.reserve(F.size());
        std::vector<int> F2;
        F2.reserve(F.size() * 4);
        auto midpoint = [&](int a, int b) {
            const uint64_t key = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            float m[3] = {(V[a * 3] + V[b * 3]) / 2, (V[a * 3 + 1] + V[b * 3 + 1]) / 2, (V[a * 3 + 2] + V[b * 3 + 2]) / 2};
            norm3(m);
            const int id = (int)(V.size() / 3);
            V.insert(V.end(), m, m + 3);
            mid.emplace(key, id);
            return id;
        };
        for (size_t f = 0; f < F.size(); f += 3) {
            const int a = F[f], b = F[f + 1], c = F[f + 2];
            const int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const int nf[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            F2.insert(F2.end(), nf, nf + 12);
        }
        F.swap(F2);
    }
    finishIcosphere(s, V, F, seed);
}

// Displacement, per-face materials, environment and camera shared by the icosphere workloads.
static void finishIcosphere(pt_scene &s, std::vector<float> &V, const std::vector<int> &F, uint64_t seed) {
    const size_t nv = V.size() / 3;
    for (size_t i = 0; i < nv; i++) {
        float *p = &V[i * 3];
        const float r = 1.f + 0.05f * fbm3(p, seed);
        p[0] *= r; p[1] *= r; p[2] *= r;
    }
    const size_t nf = F.size() / 3;
    s.mats.reserve(s.mats.size() + nf);
    std::mt19937_64 rng(seed * 0x2545F4914F6CDD1DULL + 17);
    std::uniform_real_distribution<float> u01(0.f, 1.f);
    s.pos.reserve(nf * 9);
    s.triMat.reserve(nf);
    for (size_t f = 0; f < nf; f++) {
        int id;
        if (u01(rng) < 0.3f) {
            const float r = 0.2f + 0.6f * u01(rng), g = 0.2f + 0.6f * u01(rng), b = 0.2f + 0.6f * u01(rng);
            id = addMat(s, PT_LAMBERTIAN, r, g, b, 0, 0, 0, 1.f, 0.1f);
        } else {
            id = addMat(s, PT_MICROFACET, 0, 0, 0, 0, 0, 0, 1.f, 0.05f + 0.45f * u01(rng));
        }
        addTri(s, &V[F[f * 3] * 3], &V[F[f * 3 + 1] * 3], &V[F[f * 3 + 2] * 3], id);
    }
    syntheticEnvironment(s, 1024, 512, seed);
    s.cam = HostCamera();
    const float o[3] = {0.f, 0.35f, 3.2f}, tg[3] = {0.f, 0.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
    memcpy(s.cam.origin, o, sizeof o);
    memcpy(s.cam.target, tg, sizeof tg);
    memcpy(s.cam.up, up, sizeof up);
    s.cam.vfov = 40.f / 180.f * 3.14159265358979f;
    s.cam.flip = 0;
}

// Tessellated ellipsoid (UV sphere scaled by radii), outward winding.
static void addEllipsoid(pt_scene &s, const float c[3], const float r[3], int nu, int nv, int mat) {
    auto P = [&](int i, int j, float out[3]) {
        const float th = (float)M_PI * j / nv, ph = 2.f * (float)M_PI * i / nu;
        out[0] = c[0] + r[0] * sinf(th) * cosf(ph);
        out[1] = c[1] + r[1] * cosf(th);
        out[2] = c[2] + r[2] * sinf(th) * sinf(ph);
    };
    for (int j = 0; j < nv; j++)
        for (int i = 0; i < nu; i++) {
            float a[3], b[3], cc[3], d[3];
            P(i, j, a); P(i + 1, j, b); P(i + 1, j + 1, cc); P(i, j + 1, d);
            if (j != 0) addTri(s, a, b, d, mat);
            if (j != nv - 1) addTri(s, b, cc, d, mat);
        }
}
static void addBox(pt_scene &s, const float lo[3], const float hi[3], int mat) {
    const float v[8][3] = {{lo[0], lo[1], lo[2]}, {hi[0], lo[1], lo[2]}, {hi[0], hi[1], lo[2]}, {lo[0], hi[1], lo[2]},
                           {lo[0], lo[1], hi[2]}, {hi[0], lo[1], hi[2]}, {hi[0], hi[1], hi[2]}, {lo[0], hi[1], hi[2]}};
    const int q[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 4, 7, 3}, {1, 2, 6, 5}};
    for (int f = 0; f < 6; f++) addQuad(s, v[q[f][0]], v[q[f][1]], v[q[f][2]], v[q[f][3]], mat);
}

// Closed room x in [-3, 3], y in [0, 3], z in [-3, 3] with a 0.3-wide slot along z in the
// ceiling; an emitter of radiance 1e3 above the slot; clutter of boxes and tessellated
// ellipsoids (Beckmann alpha U[0.02, 0.3]) up to ~target triangles; a glass sphere r = 0.5.
static void presetRoom(pt_scene &s, uint64_t seed, int64_t target) {
    std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ULL + 3);
    std::uniform_real_distribution<float> u01(0.f, 1.f);
    const int wall = addMat(s, PT_LAMBERTIAN, 0.7f, 0.7f, 0.7f, 0, 0, 0, 1.f, 0.1f);
    const int light = addMat(s, PT_LAMBERTIAN, 0.f, 0.f, 0.f, 1000.f, 1000.f, 1000.f, 1.f, 0.1f);
    const int glass = addMat(s, PT_GLASS, 0, 0, 0, 0, 0, 0, 1.5f, 0.1f);
    const float X = 3.f, Y = 3.f, Z = 3.f, hs = 0.15f;
    // floor (normal +y), back/front/left/right walls facing in
    { const float a[3] = {-X, 0, Z}, b[3] = {X, 0, Z}, c[3] = {X, 0, -Z}, d[3] = {-X, 0, -Z}; addQuad(s, a, b, c, d, wall); }
    { const float a[3] = {-X, 0, -Z}, b[3] = {X, 0, -Z}, c[3] = {X, Y, -Z}, d[3] = {-X, Y, -Z}; addQuad(s, a, b, c, d, wall); }
    { const float a[3] = {X, 0, Z}, b[3] = {-X, 0, Z}, c[3] = {-X, Y, Z}, d[3] = {X, Y, Z}; addQuad(s, a, b, c, d, wall); }
    { const float a[3] = {-X, 0, Z}, b[3] = {-X, 0, -Z}, c[3] = {-X, Y, -Z}, d[3] = {-X, Y, Z}; addQuad(s, a, b, c, d, wall); }
    { const float a[3] = {X, 0, -Z}, b[3] = {X, 0, Z}, c[3] = {X, Y, Z}, d[3] = {X, Y, -Z}; addQuad(s, a, b, c, d, wall); }
    // ceiling in two parts around the slot x in [-hs, hs], normal -y (inside) ...
    { const float a[3] = {-X, Y, Z}, b[3] = {-X, Y, -Z}, c[3] = {-hs, Y, -Z}, d[3] = {-hs, Y, Z}; addQuad(s, a, b, c, d, wall); }
    { const float a[3] = {hs, Y, Z}, b[3] = {hs, Y, -Z}, c[3] = {X, Y, -Z}, d[3] = {X, Y, Z}; addQuad(s, a, b, c, d, wall); }
    // ... and its outer face (normal +y), so the roof is opaque from both sides
    { const float a[3] = {-X, Y + 0.01f, Z}, b[3] = {-hs, Y + 0.01f, Z}, c[3] = {-hs, Y + 0.01f, -Z}, d[3] = {-X, Y + 0.01f, -Z}; addQuad(s, a, b, c, d, wall); }
    { const float a[3] = {hs, Y + 0.01f, Z}, b[3] = {X, Y + 0.01f, Z}, c[3] = {X, Y + 0.01f, -Z}, d[3] = {hs, Y + 0.01f, -Z}; addQuad(s, a, b, c, d, wall); }
    // emitter strip above the slot facing down (x in [-0.6, 0.6], z in [-1, 1], y = 3.6)
    { const float a[3] = {-0.6f, 3.6f, 1.f}, b[3] = {-0.6f, 3.6f, -1.f}, c[3] = {0.6f, 3.6f, -1.f}, d[3] = {0.6f, 3.6f, 1.f}; addQuad(s, a, b, c, d, light); }
    addSphere(s, 0.8f, 0.5f, 0.6f, 0.5f, glass);
    // clutter, kept clear of the camera (-2.4, 1.4, 2.6) and of the glass sphere
    int64_t budget = target - (int64_t)s.triMat.size();
    int obj = 0;
    while (budget > 0) {
        const float alpha = 0.02f + 0.28f * u01(rng);
        const float cx = -2.6f + 5.2f * u01(rng), cz = -2.6f + 5.2f * u01(rng);
        const float dcx = cx + 2.4f, dcz = cz - 2.6f, dsx = cx - 0.8f, dsz = cz - 0.6f;
        if (dcx * dcx + dcz * dcz < 1.5f * 1.5f || dsx * dsx + dsz * dsz < 0.9f * 0.9f) continue;
        const int mat = addMat(s, PT_MICROFACET, 0, 0, 0, 0, 0, 0, 1.f, alpha);
        if ((obj % 4) == 3) {
            const float sx = 0.1f + 0.3f * u01(rng), sy = 0.1f + 0.6f * u01(rng), sz = 0.1f + 0.3f * u01(rng);
            const float lo[3] = {cx - sx, 0.f, cz - sz}, hi[3] = {cx + sx, 2 * sy, cz + sz};
            addBox(s, lo, hi, mat);
            budget -= 12;
        } else {
            const float r[3] = {0.08f + 0.25f * u01(rng), 0.08f + 0.35f * u01(rng), 0.08f + 0.25f * u01(rng)};
            const float c[3] = {cx, r[1] + 1.2f * u01(rng), cz};
            int nu = 256, nv = 128;
            const int64_t per = (int64_t)nu * nv * 2 - 2 * nu;
            if (per > budget) {
                nv = std::max(4, (int)sqrt((double)std::max<int64_t>(budget, 64) / 4.0));
                nu = 2 * nv;
            }
            addEllipsoid(s, c, r, nu, nv, mat);
            budget -= (int64_t)nu * nv * 2 - 2 * nu;
        }
        obj++;
    }
    syntheticEnvironment(s, 2048, 1024, seed);
    s.cam = HostCamera();
    const float o[3] = {-2.4f, 1.4f, 2.6f}, tg[3] = {0.5f, 1.1f, -1.0f}, up[3] = {0.f, 1.f, 0.f};
    memcpy(s.cam.origin, o, sizeof o);
    memcpy(s.cam.target, tg, sizeof tg);
    memcpy(s.cam.up, up, sizeof up);
    s.cam.vfov = 55.f / 180.f * 3.14159265358979f;
}

}  // namespace pt

using namespace pt;

extern "C" {

const char *pt_last_error(void) { return g_lastError.c_str(); }
int pt_abi_version(void) { return PT_ABI_VERSION; }

int pt_scene_create(pt_scene **out) {
    if (!out) return fail(PT_ERR_INVALID, "pt_scene_create: out is NULL");
    try {
        *out = new pt_scene();
    } catch (...) {
        return fail(PT_ERR_MEMORY, "pt_scene_create: out of memory");
    }
    return PT_OK;
}

int pt_scene_destroy(pt_scene *s) {
    delete s;
    return PT_OK;
}

int pt_scene_add_material(pt_scene *s, int type, const float albedo[3], const float emit[3], float ior, float alpha, int *out_id) {
    if (!s) return fail(PT_ERR_INVALID, "pt_scene_add_material: scene is NULL");
    if (type < PT_LAMBERTIAN || type > PT_MICROFACET) return fail(PT_ERR_INVALID, "pt_scene_add_material: unknown material type");
    const float z[3] = {0.f, 0.f, 0.f};
    const float *a = albedo ? albedo : z, *e = emit ? emit : z;
    const int id = addMat(*s, type, a[0], a[1], a[2], e[0], e[1], e[2], ior, alpha);
    if (out_id) *out_id = id;
    return PT_OK;
}

int pt_scene_add_triangles(pt_scene *s, const float *positions, const float *normals, const int *material_ids, int64_t n) {
    if (!s || (n > 0 && (!positions || !material_ids)) || n < 0) return fail(PT_ERR_INVALID, "pt_scene_add_triangles: bad arguments");
    for (int64_t i = 0; i < n; i++)
        if (material_ids[i] < 0 || material_ids[i] >= (int)s->mats.size())
            return fail(PT_ERR_INVALID, "pt_scene_add_triangles: material id out of range");
    const size_t before = s->triMat.size();
    if (normals && !s->hasNormals) {
        // materialise flat normals of the triangles added so far
        s->nrm.resize(before * 9);
        for (size_t i = 0; i < before; i++) {
            const float *p = &s->pos[i * 9];
            float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]}, e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]}, g[3];
            cross3(e1, e2, g);
            norm3(g);
            for (int k = 0; k < 3; k++) memcpy(&s->nrm[i * 9 + k * 3], g, sizeof g);
        }
        s->hasNormals = true;
    }
    s->pos.insert(s->pos.end(), positions, positions + n * 9);
    s->triMat.insert(s->triMat.end(), material_ids, material_ids + n);
    if (s->hasNormals) {
        if (normals) {
            s->nrm.insert(s->nrm.end(), normals, normals + n * 9);
        } else {
            for (int64_t i = 0; i < n; i++) {
                const float *p = positions + i * 9;
                float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]}, e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]}, g[3];
                cross3(e1, e2, g);
                norm3(g);
                for (int k = 0; k < 3; k++) s->nrm.insert(s->nrm.end(), g, g + 3);
            }
        }
    }
    return PT_OK;
}

int pt_scene_add_spheres(pt_scene *s, const float *centers, const float *radii, const int *material_ids, int64_t n) {
    if (!s || n < 0 || (n > 0 && (!centers || !radii || !material_ids))) return fail(PT_ERR_INVALID, "pt_scene_add_spheres: bad arguments");
    for (int64_t i = 0; i < n; i++) {
        if (material_ids[i] < 0 || material_ids[i] >= (int)s->mats.size())
            return fail(PT_ERR_INVALID, "pt_scene_add_spheres: material id out of range");
        if (!(radii[i] > 0.f)) return fail(PT_ERR_INVALID, "pt_scene_add_spheres: radius must be positive");
    }
    for (int64_t i = 0; i < n; i++) addSphere(*s, centers[i * 3], centers[i * 3 + 1], centers[i * 3 + 2], radii[i], material_ids[i]);
    return PT_OK;
}

int pt_scene_set_environment(pt_scene *s, const float *rgba, int w, int h, int sin_theta_weight) {
    if (!s) return fail(PT_ERR_INVALID, "pt_scene_set_environment: scene is NULL");
    if (!rgba || w <= 0 || h <= 0) {
        s->env.clear();
        s->envW = s->envH = 0;
        return PT_OK;
    }
    double total = 0.0;
    for (size_t i = 0; i < (size_t)w * h; i++) {
        for (int c = 0; c < 3; c++) {
            const float v = rgba[i * 4 + c];
            if (!(v >= 0.f) || isinf(v)) return fail(PT_ERR_NUMERIC, "pt_scene_set_environment: texels must be finite and non-negative");
            total += v;
        }
    }
    if (!(total > 0.0)) return fail(PT_ERR_NUMERIC, "pt_scene_set_environment: environment has zero total radiance");
    s->env.assign(rgba, rgba + (size_t)w * h * 4);
    s->envW = w;
    s->envH = h;
    s->envSinWeight = sin_theta_weight ? 1 : 0;
    return PT_OK;
}

int pt_scene_set_camera(pt_scene *s, const float origin[3], const float target[3], const float up[3], float vfov, int flip) {
    if (!s || !origin || !target || !up) return fail(PT_ERR_INVALID, "pt_scene_set_camera: bad arguments");
    float m[3][4];
    if (!lookAt(origin, target, up, flip, m)) return fail(PT_ERR_NUMERIC, "pt_scene_set_camera: view direction equals the up vector");
    if (!(vfov > 0.f && vfov < 3.14159f)) return fail(PT_ERR_INVALID, "pt_scene_set_camera: vfov out of (0, pi)");
    memcpy(s->cam.origin, origin, 12);
    memcpy(s->cam.target, target, 12);
    memcpy(s->cam.up, up, 12);
    s->cam.vfov = vfov;
    s->cam.flip = flip ? 1 : 0;
    return PT_OK;
}

int pt_scene_preset(pt_scene *s, int preset, uint64_t seed, int64_t param) {
    if (!s) return fail(PT_ERR_INVALID, "pt_scene_preset: scene is NULL");
    *s = pt_scene();
    try {
        switch (preset) {
        case 0: presetCornell(*s); break;
        case 1: {
            const int level = param > 0 ? (int)param : 8;
            if (level > 10) return fail(PT_ERR_INVALID, "pt_scene_preset: icosphere level must be <= 10");
            presetIcosphere(*s, seed, level);
            break;
        }
        case 2: presetRoom(*s, seed, param > 0 ? param : 4000000); break;
        case 4:
            if (param > 50000000) return fail(PT_ERR_INVALID, "pt_scene_preset: geodesic sphere limited to 5e7 triangles");
            presetGeodesic(*s, seed, param > 0 ? param : 1000000);
            break;
        case 3: {
            const int w = param > 0 ? (int)param : 1024;
            syntheticEnvironment(*s, w, std::max(1, w / 2), seed);
            break;
        }
        default: return fail(PT_ERR_INVALID, "pt_scene_preset: unknown preset");
        }
    } catch (const std::bad_alloc &) {
        return fail(PT_ERR_MEMORY, "pt_scene_preset: out of memory");
    }
    return PT_OK;
}

int pt_scene_get_info(const pt_scene *s, pt_scene_info *info) {
    if (!s || !info) return fail(PT_ERR_INVALID, "pt_scene_get_info: bad arguments");
    memset(info, 0, sizeof *info);
    info->n_triangles = (int64_t)s->triMat.size();
    info->n_spheres = (int64_t)s->sphMat.size();
    info->n_materials = (int64_t)s->mats.size();
    info->n_emitters = (int64_t)emitterList(*s).size();
    info->env_width = s->envW;
    info->env_height = s->envH;
    info->has_normals = s->hasNormals;
    memcpy(info->camera_origin, s->cam.origin, 12);
    memcpy(info->camera_target, s->cam.target, 12);
    memcpy(info->camera_up, s->cam.up, 12);
    info->camera_vfov = s->cam.vfov;
    info->camera_flip = s->cam.flip;
    for (int k = 0; k < 3; k++) { info->bounds_lo[k] = INFINITY; info->bounds_hi[k] = -INFINITY; }
    for (size_t i = 0; i < s->pos.size(); i += 3)
        for (int k = 0; k < 3; k++) {
            info->bounds_lo[k] = std::min(info->bounds_lo[k], s->pos[i + k]);
            info->bounds_hi[k] = std::max(info->bounds_hi[k], s->pos[i + k]);
        }
    for (size_t i = 0; i < s->sphMat.size(); i++)
        for (int k = 0; k < 3; k++) {
            info->bounds_lo[k] = std::min(info->bounds_lo[k], s->sph[i * 4 + k] - s->sph[i * 4 + 3]);
            info->bounds_hi[k] = std::max(info->bounds_hi[k], s->sph[i * 4 + k] + s->sph[i * 4 + 3]);
        }
    return PT_OK;
}

int pt_scene_export(const pt_scene *s, float *positions, float *normals, int *tri_mat, float *spheres, int *sph_mat, int *mat_type,
                    float *mat_albedo, float *mat_emit, float *mat_ior, float *mat_alpha, float *env_rgba) {
    if (!s) return fail(PT_ERR_INVALID, "pt_scene_export: scene is NULL");
    const size_t nt = s->triMat.size();
    if (positions) memcpy(positions, s->pos.data(), nt * 9 * sizeof(float));
    if (normals) {
        if (s->hasNormals) {
            memcpy(normals, s->nrm.data(), nt * 9 * sizeof(float));
        } else {
            for (size_t i = 0; i < nt; i++) {
                const float *p = &s->pos[i * 9];
                float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]}, e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]}, g[3];
                cross3(e1, e2, g);
                norm3(g);
                for (int k = 0; k < 3; k++) memcpy(&normals[i * 9 + k * 3], g, sizeof g);
            }
        }
    }
    if (tri_mat) memcpy(tri_mat, s->triMat.data(), nt * sizeof(int));
    if (spheres) memcpy(spheres, s->sph.data(), s->sph.size() * sizeof(float));
    if (sph_mat) memcpy(sph_mat, s->sphMat.data(), s->sphMat.size() * sizeof(int));
    for (size_t i = 0; i < s->mats.size(); i++) {
        const HostMaterial &m = s->mats[i];
        if (mat_type) mat_type[i] = m.type;
        if (mat_albedo) memcpy(&mat_albedo[i * 3], m.albedo, 12);
        if (mat_emit) memcpy(&mat_emit[i * 3], m.emit, 12);
        if (mat_ior) mat_ior[i] = m.ior;
        if (mat_alpha) mat_alpha[i] = m.alpha;
    }
    if (env_rgba && !s->env.empty()) memcpy(env_rgba, s->env.data(), s->env.size() * sizeof(float));
    return PT_OK;
}

}  // extern "C"
