// pt_io.cpp -- scene and image files: Wavefront OBJ + MTL and binary PLY meshes, OpenEXR
// images (read: scanline NONE / ZIPS / ZIP with HALF or FLOAT channels; write: NONE, HALF or
// FLOAT), and the reference's four built-in scenes assembled from its scene files.
//
// Behaviour follows the reference's loaders (src/parsers/obj_parser.cpp:20-269,
// src/parsers/mtl_parser.cpp:15-109, src/parsers/ply_parser.cpp:9-112,
// src/lights/environment_light.cu:14-56, src/io/image.cpp:10-78, src/scene.cu:14-176,
// src/scene_data.cu:9-90); the OpenEXR layout follows the published OpenEXR file format
// description (magic 20000630, attribute header, offset table, scanline blocks; the ZIP
// codec is zlib over a byte-interleaved, delta-predicted block).  Errors are status codes.
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "pt_internal.h"

namespace pt {
namespace {

// ------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------
std::string ltrim(const std::string &s) {
    const size_t i = s.find_first_not_of(" \t");
    return i == std::string::npos ? std::string() : s.substr(i);
}

std::vector<std::string> tokens(const std::string &line) {
    std::vector<std::string> out;
    std::string rest = ltrim(line);
    while (!rest.empty()) {
        const size_t e = rest.find_first_of(" \t");
        if (e == std::string::npos) {
            out.push_back(rest);
            break;
        }
        out.push_back(rest.substr(0, e));
        rest = ltrim(rest.substr(e));
    }
    return out;
}

std::string dirOf(const std::string &path) {
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

bool readFile(const std::string &path, std::vector<unsigned char> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (n < 0) {
        fclose(f);
        return false;
    }
    out.resize((size_t)n);
    const size_t got = n > 0 ? fread(out.data(), 1, (size_t)n, f) : 0;
    fclose(f);
    return got == (size_t)n;
}

void flatNormal(const float *p, float n[3]) {
    const float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]};
    const float e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
    // parsers/types.h Vertex::cross and normalized()
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = -(e1[0] * e2[2] - e1[2] * e2[0]);
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; k++) n[k] /= l;
}

// ------------------------------------------------------------------------------------------
// MTL: newmtl / Kd / Ke; materials ordered by name (the reference keeps them in a std::map,
// mtl_parser.cpp:21-27), each a Lambertian with albedo Kd and emission Ke
// ------------------------------------------------------------------------------------------
struct Mtl {
    float kd[3] = {0.f, 0.f, 0.f};
    float ke[3] = {0.f, 0.f, 0.f};
};

bool parseMtl(const std::string &path, std::map<std::string, Mtl> &out, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::string line, current;
    while (std::getline(f, line)) {
        const std::vector<std::string> t = tokens(line);
        if (t.empty() || t[0][0] == '#') continue;
        if (t[0] == "newmtl") {
            if (t.size() < 2) {
                err = "newmtl without a name in " + path;
                return false;
            }
            current = t[1];
            out[current] = Mtl();
        } else if (t[0] == "Kd" || t[0] == "Ke") {
            if (t.size() < 4) {
                err = t[0] + " needs three values in " + path;
                return false;
            }
            Mtl &m = out[current];
            float *dst = t[0] == "Kd" ? m.kd : m.ke;
            for (int k = 0; k < 3; k++) dst[k] = strtof(t[1 + k].c_str(), nullptr);
        }
        // Ka, Ks, Ni, Ns, illum and anything else: ignored (mtl_parser.cpp:47-57)
    }
    return true;
}

// ------------------------------------------------------------------------------------------
// OBJ
// ------------------------------------------------------------------------------------------
struct ObjFace {
    float p[9];
    float n[9];
    bool hasNormals;
    int mtl;
};

// OBJ index: 1-based, or negative = relative to the end (obj_parser.cpp:220-228)
bool fixIndex(long idx, size_t count, long *out) {
    const long i = idx < 0 ? idx + (long)count : idx - 1;
    if (i < 0 || i >= (long)count) return false;
    *out = i;
    return true;
}

// A face vertex "v", "v/t/n", "v//n" or "v/t": vertex index and normal index (0 = none).
bool parseCorner(const std::string &s, long *v, long *vn, int *form) {
    const size_t a = s.find('/');
    char *end = nullptr;
    *v = strtol(s.c_str(), &end, 10);
    if (end == s.c_str()) return false;
    *vn = 0;
    if (a == std::string::npos) {
        *form = 0;
        return *end == '\0';
    }
    const size_t b = s.find('/', a + 1);
    if (b == std::string::npos) {  // v/t
        *form = 3;
        return true;
    }
    *form = (b == a + 1) ? 2 : 1;  // v//n or v/t/n
    const char *ns = s.c_str() + b + 1;
    *vn = strtol(ns, &end, 10);
    return end != ns;
}

bool parseObj(const std::string &path, std::vector<ObjFace> &faces, std::vector<Mtl> &mtls, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::vector<float> V, N;
    std::map<std::string, int> lookup;
    int current = 0;
    std::string line;
    int lineNo = 0;
    while (std::getline(f, line)) {
        lineNo++;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        const size_t sp = line.find_first_of(" \t");
        if (sp == std::string::npos) continue;
        const std::string cmd = line.substr(0, sp);
        if (cmd[0] == '#') continue;
        const std::string rest = ltrim(line.substr(sp + 1));
        if (cmd == "v" || cmd == "vn") {
            const char *c = rest.c_str();
            char *end = nullptr;
            float xyz[3];
            for (int k = 0; k < 3; k++) {
                xyz[k] = strtof(c, &end);
                if (end == c) {
                    err = path + ":" + std::to_string(lineNo) + ": bad " + cmd;
                    return false;
                }
                c = end;
            }
            std::vector<float> &dst = cmd == "v" ? V : N;
            dst.insert(dst.end(), xyz, xyz + 3);
        } else if (cmd == "f") {
            const std::vector<std::string> t = tokens(rest);
            long vi[4], ni[4];
            int form[4];
            const size_t nc = t.size();
            // the reference reads quads of plain indices and triangles of v/t/n or v//n
            // corners (obj_parser.cpp:95-193); plain-index triangles are read as well here
            if (nc < 3 || nc > 4) {
                err = path + ":" + std::to_string(lineNo) + ": faces must have 3 or 4 corners";
                return false;
            }
            bool ok = true;
            for (size_t k = 0; k < nc && ok; k++) ok = parseCorner(t[k], &vi[k], &ni[k], &form[k]);
            if (!ok) {
                err = path + ":" + std::to_string(lineNo) + ": bad face corner";
                return false;
            }
            for (size_t k = 1; k < nc; k++)
                if (form[k] != form[0]) {
                    err = path + ":" + std::to_string(lineNo) + ": mixed face corner forms";
                    return false;
                }
            const bool withNormals = form[0] == 1 || form[0] == 2;
            if (nc == 4 && withNormals) {
                err = path + ":" + std::to_string(lineNo) + ": quads with normals are not supported";
                return false;
            }
            long v[4], n[4];
            for (size_t k = 0; k < nc; k++) {
                if (!fixIndex(vi[k], V.size() / 3, &v[k]) || (withNormals && !fixIndex(ni[k], N.size() / 3, &n[k]))) {
                    err = path + ":" + std::to_string(lineNo) + ": face index out of range";
                    return false;
                }
            }
            const int tris[2][3] = {{0, 1, 2}, {0, 2, 3}};
            for (int tI = 0; tI < (nc == 4 ? 2 : 1); tI++) {
                ObjFace face;
                for (int c = 0; c < 3; c++) {
                    const int k = tris[tI][c];
                    memcpy(&face.p[3 * c], &V[3 * v[k]], 12);
                    if (withNormals) memcpy(&face.n[3 * c], &N[3 * n[k]], 12);
                }
                face.hasNormals = withNormals;
                if (!withNormals) {
                    float g[3];
                    flatNormal(face.p, g);
                    for (int c = 0; c < 3; c++) memcpy(&face.n[3 * c], g, 12);
                }
                face.mtl = current;
                faces.push_back(face);
            }
        } else if (cmd == "mtllib") {
            std::map<std::string, Mtl> m;
            if (!parseMtl(dirOf(path) + "/" + rest, m, err)) return false;
            // the last library wins; ids follow name order (obj_parser.cpp:250-263)
            mtls.clear();
            lookup.clear();
            int idx = 0;
            for (const auto &kv : m) {
                mtls.push_back(kv.second);
                lookup[kv.first] = idx++;
            }
        } else if (cmd == "usemtl") {
            // an unknown name selects material 0, as std::map::operator[] does in the reference
            const auto it = lookup.find(rest);
            current = it == lookup.end() ? 0 : it->second;
        }
    }
    return true;
}

// ------------------------------------------------------------------------------------------
// PLY (binary little endian: float x, y, z vertices; uchar count + int indices triangles)
// ------------------------------------------------------------------------------------------
bool parsePly(const std::string &path, std::vector<float> &verts, std::vector<float> &tris, std::string &err) {
    std::vector<unsigned char> d;
    if (!readFile(path, d)) {
        err = "cannot open " + path;
        return false;
    }
    size_t pos = 0;
    auto nextLine = [&](std::string &l) {
        const size_t e = std::find(d.begin() + pos, d.end(), '\n') - d.begin();
        if (e >= d.size()) return false;
        l.assign((const char *)&d[pos], e - pos);
        if (!l.empty() && l.back() == '\r') l.pop_back();
        pos = e + 1;
        return true;
    };
    std::string l;
    long nv = -1, nf = -1;
    if (!nextLine(l) || l != "ply") {
        err = path + ": not a PLY file";
        return false;
    }
    if (!nextLine(l) || l != "format binary_little_endian 1.0") {
        err = path + ": only binary_little_endian 1.0 PLY files are supported";
        return false;
    }
    const char *expectVertex[3] = {"property float x", "property float y", "property float z"};
    if (!nextLine(l) || sscanf(l.c_str(), "element vertex %ld", &nv) != 1) {
        err = path + ": expected 'element vertex N'";
        return false;
    }
    for (int k = 0; k < 3; k++)
        if (!nextLine(l) || l != expectVertex[k]) {
            err = path + ": expected '" + expectVertex[k] + "'";
            return false;
        }
    if (!nextLine(l) || sscanf(l.c_str(), "element face %ld", &nf) != 1) {
        err = path + ": expected 'element face N'";
        return false;
    }
    if (!nextLine(l) || (l != "property list uint8 int vertex_indices" && l != "property list uchar int vertex_indices")) {
        err = path + ": expected 'property list uchar int vertex_indices'";
        return false;
    }
    if (!nextLine(l) || l != "end_header") {
        err = path + ": expected 'end_header'";
        return false;
    }
    if (nv < 0 || nf < 0 || d.size() - pos < (size_t)nv * 12) {
        err = path + ": truncated vertex data";
        return false;
    }
    verts.resize((size_t)nv * 3);
    memcpy(verts.data(), &d[pos], (size_t)nv * 12);
    pos += (size_t)nv * 12;
    tris.reserve((size_t)nf * 9);
    for (long i = 0; i < nf; i++) {
        if (pos + 13 > d.size() || d[pos] != 3) {
            err = path + ": faces must be triangles";
            return false;
        }
        int32_t idx[3];
        memcpy(idx, &d[pos + 1], 12);
        pos += 13;
        for (int k = 0; k < 3; k++) {
            if (idx[k] < 0 || idx[k] >= nv) {
                err = path + ": face index out of range";
                return false;
            }
            tris.insert(tris.end(), &verts[3 * (size_t)idx[k]], &verts[3 * (size_t)idx[k]] + 3);
        }
    }
    return true;
}

// ------------------------------------------------------------------------------------------
// OpenEXR
// ------------------------------------------------------------------------------------------
constexpr uint32_t kExrMagic = 20000630u;
enum { EXR_UINT = 0, EXR_HALF = 1, EXR_FLOAT = 2 };
enum { EXR_NONE = 0, EXR_RLE = 1, EXR_ZIPS = 2, EXR_ZIP = 3, EXR_PIZ = 4 };

// IEEE half <-> float, round to nearest even
float halfToFloat(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31;
    const uint32_t e = (h >> 10) & 0x1f;
    uint32_t m = h & 0x3ff;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: normalise
            int ee = -1;
            do {
                ee++;
                m <<= 1;
            } while (!(m & 0x400));
            bits = s | (uint32_t)(127 - 15 - ee) << 23 | (m & 0x3ff) << 13;
        }
    } else if (e == 31) {
        bits = s | 0x7f800000u | m << 13;
    } else {
        bits = s | (e - 15 + 127) << 23 | m << 13;
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

uint16_t floatToHalf(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint16_t s = (uint16_t)((x >> 16) & 0x8000);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return s | 0x7c00 | (ax > 0x7f800000u ? 0x200 : 0);  // inf / nan
    if (ax >= 0x477ff000u) return s | 0x7c00;                                   // overflow
    if (ax < 0x38800000u) {                                                     // subnormal half
        if (ax < 0x33000000u) return s;                                         // below half min / 2
        // |f| / 2^-24 = M * 2^(e - 126) with M the 24-bit significand: round M >> (126 - e)
        const uint32_t e = ax >> 23;
        const uint32_t m = (ax & 0x7fffff) | 0x800000;
        const int shift = 126 - (int)e;
        const uint32_t q = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1);
        const uint32_t half = 1u << (shift - 1);
        uint32_t r = q;
        if (rem > half || (rem == half && (q & 1))) r++;
        return s | (uint16_t)r;
    }
    uint32_t r = ax - 0x38000000u;  // rebias exponent 127 -> 15
    const uint32_t rem = r & 0x1fff;
    r >>= 13;
    if (rem > 0x1000 || (rem == 0x1000 && (r & 1))) r++;
    return s | (uint16_t)r;
}

struct ExrChannel {
    std::string name;
    int type;
};

struct ExrHeader {
    std::vector<ExrChannel> channels;
    int compression = -1;
    int xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    int lineOrder = 0;
    size_t end = 0;  // offset of the offset table
};

bool parseExrHeader(const std::vector<unsigned char> &d, ExrHeader &h, std::string &err) {
    if (d.size() < 8) {
        err = "not an OpenEXR file";
        return false;
    }
    uint32_t magic, version;
    memcpy(&magic, &d[0], 4);
    memcpy(&version, &d[4], 4);
    if (magic != kExrMagic) {
        err = "not an OpenEXR file";
        return false;
    }
    if ((version & 0xff) != 2 || (version & 0x1a00)) {  // tiled, deep or multi-part
        err = "only single-part scanline OpenEXR files (version 2) are supported";
        return false;
    }
    size_t p = 8;
    auto cstr = [&](std::string &out) {
        const size_t e = std::find(d.begin() + p, d.end(), '\0') - d.begin();
        if (e >= d.size()) return false;
        out.assign((const char *)&d[p], e - p);
        p = e + 1;
        return true;
    };
    while (true) {
        std::string name, type;
        if (!cstr(name)) {
            err = "truncated header";
            return false;
        }
        if (name.empty()) break;
        if (!cstr(type) || p + 4 > d.size()) {
            err = "truncated header";
            return false;
        }
        uint32_t size;
        memcpy(&size, &d[p], 4);
        p += 4;
        if (p + size > d.size()) {
            err = "truncated header";
            return false;
        }
        const unsigned char *v = &d[p];
        if (name == "channels" && type == "chlist") {
            size_t q = 0;
            while (q < size && v[q] != 0) {
                const size_t e = std::find(v + q, v + size, '\0') - v;
                if (e + 17 > size) {
                    err = "bad channel list";
                    return false;
                }
                ExrChannel c;
                c.name.assign((const char *)v + q, e - q);
                int32_t pt, xs, ys;
                memcpy(&pt, v + e + 1, 4);
                memcpy(&xs, v + e + 9, 4);
                memcpy(&ys, v + e + 13, 4);
                if (xs != 1 || ys != 1) {
                    err = "subsampled channels are not supported";
                    return false;
                }
                c.type = pt;
                h.channels.push_back(c);
                q = e + 17;
            }
        } else if (name == "compression" && size >= 1) {
            h.compression = v[0];
        } else if (name == "dataWindow" && size >= 16) {
            int32_t b[4];
            memcpy(b, v, 16);
            h.xmin = b[0];
            h.ymin = b[1];
            h.xmax = b[2];
            h.ymax = b[3];
        } else if (name == "lineOrder" && size >= 1) {
            h.lineOrder = v[0];
        } else if (name == "tiles") {
            err = "tiled OpenEXR files are not supported";
            return false;
        }
        p += size;
    }
    h.end = p;
    if (h.channels.empty() || h.xmax < h.xmin || h.ymax < h.ymin) {
        err = "missing channels or data window";
        return false;
    }
    // widths and heights in 64-bit: xmax - xmin + 1 overflows int32 for extreme windows
    const int64_t w = (int64_t)h.xmax - h.xmin + 1, hh = (int64_t)h.ymax - h.ymin + 1;
    if (w > (1 << 16) || hh > (1 << 16) || w * hh > (1ll << 28)) {
        err = "data window too large";
        return false;
    }
    return true;
}

// Inverse of the ZIP / ZIPS block transform: zlib, then the delta predictor (each byte is
// the previous plus the stored difference minus 128), then the two half-buffers interleaved.
bool unzipBlock(const unsigned char *src, size_t n, std::vector<unsigned char> &out, size_t expect) {
    std::vector<unsigned char> t(expect);
    uLongf got = (uLongf)expect;
    if (uncompress(t.data(), &got, src, (uLong)n) != Z_OK || got != expect) return false;
    for (size_t i = 1; i < expect; i++) t[i] = (unsigned char)(t[i - 1] + t[i] - 128);
    out.resize(expect);
    const size_t halfN = (expect + 1) / 2;
    for (size_t i = 0, j = 0; j < expect; i++) {
        out[j++] = t[i];
        if (j < expect) out[j++] = t[halfN + i];
    }
    return true;
}

// ---------------------------------------------------------------------------------------------
// PIZ blocks (OpenEXR's wavelet + Huffman codec, 32 scanlines per block).  A block holds, in this
// order: the range [lo, hi] of a bitmap of the 16-bit values present, those bitmap bytes, the
// length of the Huffman stream, and the stream.  Decoding reverses the encoder's three steps:
// Huffman decoding to 16-bit words, the 2D Haar-like wavelet per channel plane, and the
// value-compaction table built from the bitmap.  The bit layouts and arithmetic below are the
// format's; this is an independent implementation of them.
// ---------------------------------------------------------------------------------------------
namespace piz {

constexpr int kRange = 1 << 16;             // 16-bit values
constexpr int kBitmapBytes = kRange / 8;
constexpr int kMaxCodeLen = 58;
constexpr int kShortZeroRun = 59;           // 6-bit length codes 59..62: 2..5 zero lengths
constexpr int kLongZeroRun = 63;            // 63 + 8 bits: 6..261 zero lengths
constexpr int kShortestLongRun = 2 + kLongZeroRun - kShortZeroRun;

// MSB-first bit reader over [p, end) with at most `bits` bits in total.
struct Bits {
    const unsigned char *p = nullptr, *end = nullptr;
    int64_t left = 0;   // bits still allowed
    uint32_t cur = 0;
    int have = 0;       // bits of cur not yet consumed
    bool get(int n, uint64_t &v) {
        v = 0;
        for (int i = 0; i < n; i++) {
            if (left <= 0) return false;
            if (have == 0) {
                if (p >= end) return false;
                cur = *p++;
                have = 8;
            }
            have--;
            left--;
            v = (v << 1) | ((cur >> have) & 1u);
        }
        return true;
    }
};

uint32_t readU32(const unsigned char *q) { return (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 | (uint32_t)q[3] << 24; }

// Huffman stream: 20-byte header (lowest symbol, highest symbol = the run-length pseudo-symbol,
// table bytes, stream bits, reserved), the run-length-coded table of code lengths, then the
// codes.  Codes are canonical: per length from the longest down, the first code of length l
// is (first code of length l + 1 + count of length l + 1) / 2, and symbols of one length take
// consecutive codes in increasing order.  The pseudo-symbol repeats the previous word
// (8-bit count).
bool hufDecode(const unsigned char *in, size_t n, uint16_t *out, size_t nOut, std::string &err) {
    if (n == 0) {
        if (nOut != 0) err = "PIZ: empty Huffman stream";
        return nOut == 0;
    }
    if (n < 20) {
        err = "PIZ: truncated Huffman header";
        return false;
    }
    const uint32_t im = readU32(in), iM = readU32(in + 4), tableBytes = readU32(in + 8), nBits = readU32(in + 12);
    if (im > iM || iM >= (uint32_t)kRange + 1 || tableBytes > n - 20) {
        err = "PIZ: bad Huffman table header";
        return false;
    }
    std::vector<uint8_t> len((size_t)kRange + 1, 0);
    Bits tb;
    tb.p = in + 20;
    tb.end = in + 20 + tableBytes;
    tb.left = (int64_t)tableBytes * 8;
    for (uint32_t sym = im; sym <= iM; sym++) {
        uint64_t l = 0;
        if (!tb.get(6, l)) {
            err = "PIZ: truncated code-length table";
            return false;
        }
        if (l == kLongZeroRun || l >= kShortZeroRun) {
            uint64_t run = 0;
            if (l == kLongZeroRun) {
                if (!tb.get(8, run)) {
                    err = "PIZ: truncated code-length table";
                    return false;
                }
                run += kShortestLongRun;
            } else {
                run = l - kShortZeroRun + 2;
            }
            if (sym + run - 1 > iM) {
                err = "PIZ: code-length run past the last symbol";
                return false;
            }
            sym += (uint32_t)run - 1;  // lengths stay 0
            continue;
        }
        len[sym] = (uint8_t)l;
    }
    // canonical codes
    uint64_t count[kMaxCodeLen + 2] = {0}, first[kMaxCodeLen + 2] = {0};
    for (uint32_t sym = im; sym <= iM; sym++) count[len[sym]]++;
    uint64_t c = 0;
    for (int l = kMaxCodeLen; l > 0; l--) {
        first[l] = c;
        c = (c + count[l]) >> 1;
    }
    std::vector<std::vector<uint32_t>> byLen(kMaxCodeLen + 1);
    for (uint32_t sym = im; sym <= iM; sym++)
        if (len[sym] > 0) byLen[len[sym]].push_back(sym);
    // decode
    Bits bs;
    bs.p = in + 20 + tableBytes;
    bs.end = in + n;
    bs.left = nBits;
    if ((uint64_t)nBits > (uint64_t)(bs.end - bs.p) * 8) {
        err = "PIZ: Huffman stream shorter than its bit count";
        return false;
    }
    size_t o = 0;
    while (bs.left > 0 && o < nOut) {
        uint64_t code = 0;
        int l = 0;
        uint32_t sym = 0;
        bool found = false;
        while (!found) {
            uint64_t bit = 0;
            if (l == kMaxCodeLen || !bs.get(1, bit)) break;
            code = (code << 1) | bit;
            l++;
            if (count[l] && code >= first[l] && code - first[l] < count[l]) {
                sym = byLen[l][(size_t)(code - first[l])];
                found = true;
            }
        }
        if (!found) {
            // the last byte's padding bits may not form a code
            if (bs.left <= 0) break;
            err = "PIZ: invalid Huffman code";
            return false;
        }
        if (sym == iM) {
            uint64_t rep = 0;
            if (!bs.get(8, rep) || o == 0 || o + rep > nOut) {
                err = "PIZ: bad run of repeated words";
                return false;
            }
            const uint16_t prev = out[o - 1];
            for (uint64_t k = 0; k < rep; k++) out[o++] = prev;
        } else {
            out[o++] = (uint16_t)sym;
        }
    }
    if (o != nOut) {
        err = "PIZ: Huffman stream holds " + std::to_string(o) + " words, the block needs " + std::to_string(nOut);
        return false;
    }
    return true;
}

// One inverse lifting step of the wavelet: 14-bit variant (values < 2^14, signed arithmetic on
// 16-bit words) and 16-bit variant (modular arithmetic).
inline void unlift14(uint16_t l, uint16_t h, uint16_t &a, uint16_t &b) {
    const int hs = (int16_t)h;
    const int as = (int16_t)l + (hs & 1) + (hs >> 1);
    a = (uint16_t)(int16_t)as;
    b = (uint16_t)(int16_t)(as - hs);
}
inline void unlift16(uint16_t l, uint16_t h, uint16_t &a, uint16_t &b) {
    const int m = l, d = h;
    const int bb = (m - (d >> 1)) & 0xffff;
    const int aa = (d + bb - 0x8000) & 0xffff;
    a = (uint16_t)aa;
    b = (uint16_t)bb;
}

// Inverse 2D wavelet of an nx x ny plane (element (x, y) at v[x * ox + y * oy]), coarsest level
// first.  maxValue < 2^14 selects the 14-bit lifting.
void unwavelet(uint16_t *v, int nx, int ox, int ny, int oy, uint16_t maxValue) {
    auto step = maxValue < (1 << 14) ? unlift14 : unlift16;
    const int n = std::min(nx, ny);
    int p = 1;
    while (p <= n) p <<= 1;
    p >>= 1;
    int p2 = p;
    p >>= 1;
    while (p >= 1) {
        const int oy1 = oy * p, ox1 = ox * p;
        int y = 0;
        for (; y + p2 <= ny; y += p2) {
            uint16_t *row = v + (size_t)y * oy;
            int x = 0;
            for (; x + p2 <= nx; x += p2) {
                uint16_t *q00 = row + (size_t)x * ox, *q01 = q00 + ox1, *q10 = q00 + oy1, *q11 = q10 + ox1;
                uint16_t i00, i01, i10, i11;
                step(*q00, *q10, i00, i10);
                step(*q01, *q11, i01, i11);
                step(i00, i01, *q00, *q01);
                step(i10, i11, *q10, *q11);
            }
            if (nx & p) {
                uint16_t *q00 = row + (size_t)x * ox, *q10 = q00 + oy1;
                uint16_t a;
                step(*q00, *q10, a, *q10);
                *q00 = a;
            }
        }
        if (ny & p) {
            uint16_t *row = v + (size_t)y * oy;
            for (int x = 0; x + p2 <= nx; x += p2) {
                uint16_t *q00 = row + (size_t)x * ox, *q01 = q00 + ox1;
                uint16_t a;
                step(*q00, *q01, a, *q01);
                *q00 = a;
            }
        }
        p2 = p;
        p >>= 1;
    }
}

// One PIZ block of `lines` scanlines, `w` pixels wide, channels of `words` 16-bit words per
// pixel (1 half, 2 float/uint), decoded into `out` in scanline order (per line, each channel's
// w pixels).
bool decodeBlock(const unsigned char *src, size_t n, int w, int lines, const std::vector<int> &words, std::vector<unsigned char> &out,
                 std::string &err) {
    if (n < 4) {
        err = "PIZ: truncated block";
        return false;
    }
    const unsigned lo = src[0] | src[1] << 8, hi = src[2] | src[3] << 8;
    size_t at = 4;
    std::vector<uint8_t> bitmap(kBitmapBytes, 0);
    if (hi >= (unsigned)kBitmapBytes) {
        err = "PIZ: bad bitmap range";
        return false;
    }
    if (lo <= hi) {
        if (n - at < hi - lo + 1) {
            err = "PIZ: truncated bitmap";
            return false;
        }
        memcpy(&bitmap[lo], src + at, hi - lo + 1);
        at += hi - lo + 1;
    }
    // the values present, in increasing order (0 always); the stored words index this list
    std::vector<uint16_t> lut(kRange, 0);
    int k = 0;
    for (int v = 0; v < kRange; v++)
        if (v == 0 || (bitmap[v >> 3] & (1 << (v & 7)))) lut[k++] = (uint16_t)v;
    const uint16_t maxValue = (uint16_t)(k - 1);
    if (n - at < 4) {
        err = "PIZ: truncated block";
        return false;
    }
    const uint32_t hufBytes = readU32(src + at);
    at += 4;
    if (hufBytes > n - at) {
        err = "PIZ: Huffman stream past the block";
        return false;
    }
    size_t total = 0;
    for (int wd : words) total += (size_t)w * lines * wd;
    std::vector<uint16_t> buf(total);
    if (!hufDecode(src + at, hufBytes, buf.data(), total, err)) return false;
    // each channel's plane(s): w x lines pixels of `wd` interleaved words
    size_t start = 0;
    for (int wd : words) {
        for (int j = 0; j < wd; j++) unwavelet(buf.data() + start + j, w, wd, lines, w * wd, maxValue);
        start += (size_t)w * lines * wd;
    }
    for (uint16_t &v : buf) v = lut[v];
    // channel planes -> scanlines
    out.resize(total * 2);
    unsigned char *o = out.data();
    std::vector<size_t> chanAt(words.size());
    size_t acc = 0;
    for (size_t c = 0; c < words.size(); c++) {
        chanAt[c] = acc;
        acc += (size_t)w * lines * words[c];
    }
    for (int y = 0; y < lines; y++)
        for (size_t c = 0; c < words.size(); c++) {
            const size_t cnt = (size_t)w * words[c];
            memcpy(o, buf.data() + chanAt[c] + (size_t)y * cnt, cnt * 2);
            o += cnt * 2;
        }
    return true;
}

}  // namespace piz

bool readExr(const std::string &path, int *W, int *H, std::vector<float> *rgba, std::string &err) {
    std::vector<unsigned char> d;
    if (!readFile(path, d)) {
        err = "cannot open " + path;
        return false;
    }
    ExrHeader h;
    if (!parseExrHeader(d, h, err)) {
        err = path + ": " + err;
        return false;
    }
    const int w = h.xmax - h.xmin + 1, hh = h.ymax - h.ymin + 1;
    *W = w;
    *H = hh;
    if (!rgba) return true;
    int linesPerBlock;
    switch (h.compression) {
    case EXR_NONE: linesPerBlock = 1; break;
    case EXR_ZIPS: linesPerBlock = 1; break;
    case EXR_ZIP: linesPerBlock = 16; break;
    case EXR_PIZ: linesPerBlock = 32; break;
    default:
        err = path + ": unsupported OpenEXR compression " + std::to_string(h.compression) + " (NONE, ZIPS, ZIP and PIZ are read)";
        return false;
    }
    // channels are stored sorted by name; pick R, G, B, A
    int bytesPerPixel = 0;
    std::vector<int> offset(h.channels.size());
    for (size_t c = 0; c < h.channels.size(); c++) {
        offset[c] = bytesPerPixel;
        const int t = h.channels[c].type;
        if (t != EXR_HALF && t != EXR_FLOAT && t != EXR_UINT) {
            err = path + ": unknown channel type";
            return false;
        }
        bytesPerPixel += t == EXR_HALF ? 2 : 4;
    }
    const size_t lineBytes = (size_t)bytesPerPixel * w;
    const int blocks = (hh + linesPerBlock - 1) / linesPerBlock;
    if (h.end + (size_t)blocks * 8 > d.size()) {
        err = path + ": truncated offset table";
        return false;
    }
    rgba->assign((size_t)w * hh * 4, 0.f);
    for (size_t i = 0; i < (size_t)w * hh; i++) (*rgba)[4 * i + 3] = 1.f;
    std::vector<unsigned char> block;
    for (int b = 0; b < blocks; b++) {
        uint64_t off;
        memcpy(&off, &d[h.end + 8 * (size_t)b], 8);
        if (off > d.size() || d.size() - off < 8) {
            err = path + ": bad block offset";
            return false;
        }
        int32_t y0, size;
        memcpy(&y0, &d[off], 4);
        memcpy(&size, &d[off + 4], 4);
        const int64_t first64 = (int64_t)y0 - h.ymin;
        if (first64 < 0 || first64 >= hh || size < 0 || d.size() - off - 8 < (uint64_t)size) {
            err = path + ": bad block";
            return false;
        }
        const int first = (int)first64;
        const int lines = std::min(linesPerBlock, hh - first);
        if (lines <= 0) {
            err = path + ": bad block";
            return false;
        }
        const size_t expect = lineBytes * lines;
        const unsigned char *src = &d[off + 8];
        if ((size_t)size == expect) {
            block.assign(src, src + expect);  // stored uncompressed (always for NONE)
        } else if (h.compression == EXR_PIZ) {
            std::vector<int> words;
            for (const ExrChannel &c : h.channels) words.push_back(c.type == EXR_HALF ? 1 : 2);
            if (!piz::decodeBlock(src, (size_t)size, w, lines, words, block, err)) {
                err = path + ": " + err;
                return false;
            }
        } else if (!unzipBlock(src, (size_t)size, block, expect)) {
            err = path + ": corrupt ZIP block";
            return false;
        }
        for (int l = 0; l < lines; l++) {
            const unsigned char *row = &block[lineBytes * l];
            for (size_t c = 0; c < h.channels.size(); c++) {
                const std::string &nm = h.channels[c].name;
                const int dst = nm == "R" ? 0 : nm == "G" ? 1 : nm == "B" ? 2 : nm == "A" ? 3 : -1;
                if (dst < 0) continue;
                const int t = h.channels[c].type;
                const unsigned char *base = row + (size_t)offset[c] * w;
                for (int x = 0; x < w; x++) {
                    float v;
                    if (t == EXR_HALF) {
                        uint16_t hv;
                        memcpy(&hv, base + 2 * (size_t)x, 2);
                        v = halfToFloat(hv);
                    } else if (t == EXR_FLOAT) {
                        memcpy(&v, base + 4 * (size_t)x, 4);
                    } else {
                        uint32_t u;
                        memcpy(&u, base + 4 * (size_t)x, 4);
                        v = (float)u;
                    }
                    (*rgba)[4 * ((size_t)(first + l) * w + x) + dst] = v;
                }
            }
        }
    }
    return true;
}

void putAttr(std::vector<unsigned char> &o, const char *name, const char *type, const void *v, uint32_t size) {
    o.insert(o.end(), name, name + strlen(name) + 1);
    o.insert(o.end(), type, type + strlen(type) + 1);
    const unsigned char *s = (const unsigned char *)&size;
    o.insert(o.end(), s, s + 4);
    o.insert(o.end(), (const unsigned char *)v, (const unsigned char *)v + size);
}

// RGB image, rows top to bottom, channels B, G, R (the order image.cpp:40-53 writes),
// uncompressed scanlines.
bool writeExr(const std::string &path, const float *rgbTopDown, int w, int h, bool half, std::string &err) {
    std::vector<unsigned char> o;
    const uint32_t magic = kExrMagic, version = 2;
    o.insert(o.end(), (const unsigned char *)&magic, (const unsigned char *)&magic + 4);
    o.insert(o.end(), (const unsigned char *)&version, (const unsigned char *)&version + 4);
    std::vector<unsigned char> ch;
    const char *names[3] = {"B", "G", "R"};
    for (const char *n : names) {
        ch.insert(ch.end(), n, n + 2);
        const int32_t pt = half ? EXR_HALF : EXR_FLOAT;
        const unsigned char pl[4] = {0, 0, 0, 0};
        const int32_t one = 1;
        ch.insert(ch.end(), (const unsigned char *)&pt, (const unsigned char *)&pt + 4);
        ch.insert(ch.end(), pl, pl + 4);
        ch.insert(ch.end(), (const unsigned char *)&one, (const unsigned char *)&one + 4);
        ch.insert(ch.end(), (const unsigned char *)&one, (const unsigned char *)&one + 4);
    }
    ch.push_back(0);
    putAttr(o, "channels", "chlist", ch.data(), (uint32_t)ch.size());
    const unsigned char comp = EXR_NONE;
    putAttr(o, "compression", "compression", &comp, 1);
    const int32_t box[4] = {0, 0, w - 1, h - 1};
    putAttr(o, "dataWindow", "box2i", box, 16);
    putAttr(o, "displayWindow", "box2i", box, 16);
    const unsigned char lo = 0;
    putAttr(o, "lineOrder", "lineOrder", &lo, 1);
    const float par = 1.f, sww = 1.f;
    const float swc[2] = {0.f, 0.f};
    putAttr(o, "pixelAspectRatio", "float", &par, 4);
    putAttr(o, "screenWindowCenter", "v2f", swc, 8);
    putAttr(o, "screenWindowWidth", "float", &sww, 4);
    o.push_back(0);
    const size_t tableAt = o.size();
    o.resize(o.size() + 8 * (size_t)h);
    const int bpc = half ? 2 : 4;
    const int32_t lineBytes = 3 * bpc * w;
    for (int y = 0; y < h; y++) {
        const uint64_t off = o.size();
        memcpy(&o[tableAt + 8 * (size_t)y], &off, 8);
        const int32_t yy = y;
        o.insert(o.end(), (const unsigned char *)&yy, (const unsigned char *)&yy + 4);
        o.insert(o.end(), (const unsigned char *)&lineBytes, (const unsigned char *)&lineBytes + 4);
        for (int c = 2; c >= 0; c--) {  // B, G, R
            for (int x = 0; x < w; x++) {
                const float v = rgbTopDown[3 * ((size_t)y * w + x) + c];
                if (half) {
                    const uint16_t hv = floatToHalf(v);
                    o.insert(o.end(), (const unsigned char *)&hv, (const unsigned char *)&hv + 2);
                } else {
                    o.insert(o.end(), (const unsigned char *)&v, (const unsigned char *)&v + 4);
                }
            }
        }
    }
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    const size_t put = fwrite(o.data(), 1, o.size(), f);
    const int closed = fclose(f);
    if (put != o.size() || closed != 0) {
        err = "short write to " + path;
        return false;
    }
    return true;
}

int addMaterial(pt_scene *s, int type, const float a[3], const float e[3], float ior, float alpha) {
    int id = -1;
    const int rc = pt_scene_add_material(s, type, a, e, ior, alpha, &id);
    return rc == PT_OK ? id : rc;
}

int loadObjInto(pt_scene *s, const std::string &path, int defaultMaterial, int *firstMaterial) {
    std::vector<ObjFace> faces;
    std::vector<Mtl> mtls;
    std::string err;
    if (!parseObj(path, faces, mtls, err)) return fail(PT_ERR_IO, "pt_scene_load_obj: " + err);
    // scene_data.cu:20-52: the file's materials become Lambertians appended to the scene;
    // without any, every face takes the default material
    const int offset = (int)s->mats.size();
    if (firstMaterial) *firstMaterial = mtls.empty() ? -1 : offset;
    if (mtls.empty() && (defaultMaterial < 0 || defaultMaterial >= (int)s->mats.size()))
        return fail(PT_ERR_INVALID, "pt_scene_load_obj: the file has no materials and the default material id is out of range");
    for (const Mtl &m : mtls) {
        const int id = addMaterial(s, PT_LAMBERTIAN, m.kd, m.ke, 1.f, 0.1f);
        if (id < 0) return id;
    }
    bool anyNormals = false;
    for (const ObjFace &f : faces) anyNormals = anyNormals || f.hasNormals;
    std::vector<float> pos, nrm;
    std::vector<int> mat;
    pos.reserve(faces.size() * 9);
    for (const ObjFace &f : faces) {
        pos.insert(pos.end(), f.p, f.p + 9);
        if (anyNormals) nrm.insert(nrm.end(), f.n, f.n + 9);
        mat.push_back(mtls.empty() ? defaultMaterial : offset + std::min(f.mtl, (int)mtls.size() - 1));
    }
    return pt_scene_add_triangles(s, pos.data(), anyNormals ? nrm.data() : nullptr, mat.data(), (int64_t)mat.size());
}

int loadPlyInto(pt_scene *s, const std::string &path, int material) {
    std::vector<float> verts, tris;
    std::string err;
    if (!parsePly(path, verts, tris, err)) return fail(PT_ERR_IO, "pt_scene_load_ply: " + err);
    std::vector<int> mat(tris.size() / 9, material);
    return pt_scene_add_triangles(s, tris.data(), nullptr, mat.data(), (int64_t)mat.size());
}

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" {

int pt_scene_load_obj(pt_scene *s, const char *path, int default_material, int *first_material) {
    if (!s || !path) return fail(PT_ERR_INVALID, "pt_scene_load_obj: bad arguments");
    try {
        return loadObjInto(s, path, default_material, first_material);
    } catch (const std::bad_alloc &) {
        return fail(PT_ERR_MEMORY, "pt_scene_load_obj: out of memory");
    }
}

int pt_scene_load_ply(pt_scene *s, const char *path, int material) {
    if (!s || !path) return fail(PT_ERR_INVALID, "pt_scene_load_ply: bad arguments");
    if (material < 0 || material >= (int)s->mats.size()) return fail(PT_ERR_INVALID, "pt_scene_load_ply: material id out of range");
    try {
        return loadPlyInto(s, path, material);
    } catch (const std::bad_alloc &) {
        return fail(PT_ERR_MEMORY, "pt_scene_load_ply: out of memory");
    }
}

int pt_exr_info(const char *path, int *width, int *height) {
    if (!path || !width || !height) return fail(PT_ERR_INVALID, "pt_exr_info: bad arguments");
    std::string err;
    if (!readExr(path, width, height, nullptr, err)) return fail(PT_ERR_IO, "pt_exr_info: " + err);
    return PT_OK;
}

int pt_exr_read_rgba(const char *path, float *rgba, int64_t capacity) {
    if (!path || !rgba) return fail(PT_ERR_INVALID, "pt_exr_read_rgba: bad arguments");
    try {
        int w = 0, h = 0;
        std::vector<float> px;
        std::string err;
        if (!readExr(path, &w, &h, &px, err)) return fail(PT_ERR_IO, "pt_exr_read_rgba: " + err);
        if ((int64_t)px.size() > capacity) return fail(PT_ERR_INVALID, "pt_exr_read_rgba: buffer too small");
        memcpy(rgba, px.data(), px.size() * sizeof(float));
        return PT_OK;
    } catch (const std::bad_alloc &) {
        return fail(PT_ERR_MEMORY, "pt_exr_read_rgba: out of memory");
    }
}

int pt_exr_write_rgb(const char *path, const float *radiance, int width, int height, int half) {
    if (!path || !radiance || width <= 0 || height <= 0) return fail(PT_ERR_INVALID, "pt_exr_write_rgb: bad arguments");
    try {
        // radiance rows run bottom to top (row 0 = bottom); EXR scanlines top to bottom
        std::vector<float> top((size_t)width * height * 3);
        for (int row = 0; row < height; row++)
            memcpy(&top[(size_t)(height - 1 - row) * width * 3], &radiance[(size_t)row * width * 3], (size_t)width * 12);
        std::string err;
        if (!writeExr(path, top.data(), width, height, half != 0, err)) return fail(PT_ERR_IO, "pt_exr_write_rgb: " + err);
        return PT_OK;
    } catch (const std::bad_alloc &) {
        return fail(PT_ERR_MEMORY, "pt_exr_write_rgb: out of memory");
    }
}

int pt_scene_load_exr_environment(pt_scene *s, const char *path, int sin_theta_weight) {
    if (!s || !path) return fail(PT_ERR_INVALID, "pt_scene_load_exr_environment: bad arguments");
    try {
        int w = 0, h = 0;
        std::vector<float> px;
        std::string err;
        if (!readExr(path, &w, &h, &px, err)) return fail(PT_ERR_IO, "pt_scene_load_exr_environment: " + err);
        return pt_scene_set_environment(s, px.data(), w, h, sin_theta_weight);
    } catch (const std::bad_alloc &) {
        return fail(PT_ERR_MEMORY, "pt_scene_load_exr_environment: out of memory");
    }
}

int pt_scene_reference(pt_scene *s, int index, const char *scenes_dir) {
    if (!s || !scenes_dir) return fail(PT_ERR_INVALID, "pt_scene_reference: bad arguments");
    if (!s->pos.empty() || !s->sph.empty() || !s->mats.empty()) return fail(PT_ERR_STATE, "pt_scene_reference: scene is not empty");
    const std::string dir(scenes_dir);
    const float z[3] = {0.f, 0.f, 0.f}, red[3] = {1.f, 0.f, 0.f}, one[3] = {1.f, 1.f, 1.f};
    int rc = PT_OK;
    // cameras of scene.cu:142-176
    const float o012[3] = {0.f, 1.f, 6.8f}, t012[3] = {0.f, 1.f, 0.f}, u012[3] = {0.f, 1.f, 0.f};
    const float fov012 = 19.5f / 180.f * 3.14159265358979f;
    switch (index) {
    case 0: {  // scene.cu:18-32
        addMaterial(s, PT_LAMBERTIAN, red, z, 1.f, 0.1f);
        rc = pt_scene_load_obj(s, (dir + "/cornell-box/CornellBox-Original.obj").c_str(), 0, nullptr);
        if (rc == PT_OK) rc = pt_scene_set_camera(s, o012, t012, u012, fov012, 0);
        break;
    }
    case 1: {  // scene.cu:33-72
        addMaterial(s, PT_LAMBERTIAN, red, z, 1.f, 0.1f);
        addMaterial(s, PT_MIRROR, z, z, 1.f, 0.1f);
        const int glass = addMaterial(s, PT_GLASS, z, z, 1.4f, 0.1f);
        rc = pt_scene_load_obj(s, (dir + "/cornell-glossy/CornellBox-Glossy.obj").c_str(), 0, nullptr);
        if (rc == PT_OK) rc = pt_scene_load_obj(s, (dir + "/cornell-glossy/box.obj").c_str(), 1, nullptr);
        if (rc == PT_OK) {
            const float c[3] = {0.45f, 0.451f, -0.4f}, r = 0.45f;
            rc = pt_scene_add_spheres(s, c, &r, &glass, 1);
        }
        if (rc == PT_OK) rc = pt_scene_set_camera(s, o012, t012, u012, fov012, 0);
        break;
    }
    case 2: {  // scene.cu:73-94
        addMaterial(s, PT_LAMBERTIAN, one, z, 1.f, 0.1f);
        rc = pt_scene_load_obj(s, (dir + "/tests/plane.obj").c_str(), 0, nullptr);
        if (rc == PT_OK) rc = pt_scene_load_exr_environment(s, (dir + "/assets/1_pixel_test.exr").c_str(), 0);
        if (rc == PT_OK) rc = pt_scene_set_camera(s, o012, t012, u012, fov012, 0);
        break;
    }
    case 3: {  // scene.cu:95-139 (the Veach MIS plates)
        const float grey[3] = {0.4f, 0.4f, 0.4f};
        const int base = addMaterial(s, PT_LAMBERTIAN, grey, z, 1.f, 0.1f);
        const float alphas[4] = {0.005f, 0.02f, 0.05f, 0.1f};
        for (float a : alphas) addMaterial(s, PT_MICROFACET, z, z, 1.f, a);
        const char *files[5] = {"floor.ply", "plate1.ply", "plate2.ply", "plate3.ply", "plate4.ply"};
        for (int k = 0; k < 5 && rc == PT_OK; k++) rc = pt_scene_load_ply(s, (dir + "/mis/" + files[k]).c_str(), base + k);
        if (rc != PT_OK) break;
        const float emit[5] = {800.f, 100.f, 901.803f, 11.1111f, 1.234567f};
        const float centers[15] = {10.f, 10.f, 4.f, -1.25f, 0.f, 0.f, -3.75f, 0.f, 0.f, 1.25f, 0.f, 0.f, 3.75f, 0.f, 0.f};
        const float radii[5] = {0.5f, 0.1f, 0.03333f, 0.3f, 0.9f};
        int ids[5];
        for (int k = 0; k < 5; k++) {
            const float e[3] = {emit[k], emit[k], emit[k]};
            ids[k] = addMaterial(s, PT_LAMBERTIAN, z, e, 1.f, 0.1f);
        }
        rc = pt_scene_add_spheres(s, centers, radii, ids, 5);
        if (rc == PT_OK) rc = pt_scene_load_exr_environment(s, (dir + "/assets/20060807_wells6_hd.exr").c_str(), 0);
        const float o[3] = {0.f, 2.f, 15.f}, t[3] = {0.f, 1.69521f, 14.0476f}, u[3] = {0.f, 0.952421f, -0.304787f};
        if (rc == PT_OK) rc = pt_scene_set_camera(s, o, t, u, 28.0000262073138f / 180.f * 3.14159265358979f, 1);
        break;
    }
    default: return fail(PT_ERR_INVALID, "pt_scene_reference: scene index must be 0..3");
    }
    if (rc != PT_OK) {
        // leave the scene empty on failure
        s->pos.clear();
        s->nrm.clear();
        s->hasNormals = false;
        s->triMat.clear();
        s->sph.clear();
        s->sphMat.clear();
        s->mats.clear();
        s->env.clear();
        s->envW = s->envH = 0;
    }
    return rc;
}

}  // extern "C"
