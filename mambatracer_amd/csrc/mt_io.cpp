// Host-side readers for the two input formats of the reference's scene and
// environment: Wavefront OBJ meshes (OBJParser.readString, called at
// src/cl/ui/fx/main/TracerAPI.java:321-322) and Radiance RGBE images
// (HDRBitmapReader, src/cl/ui/fx/main/UserInterfaceFXMLController.java:
// 176-179).  Both readers live in un-vendored jars (coordinate.parser.obj,
// bitmap.reader); these are restatements of the published formats, writing
// the reference's device layouts (points/normals float4 with w = 0, uvs
// float2, faces int32[10] = {v0,v1,v2, t0,t1,t2, n0,n1,n2, material}, -1
// for an absent uv or normal; images float4 RGBA with a = 1).  No HIP.
#include "../../include/mambart.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

void mt_set_error(const char* fmt, ...);

namespace {

struct Cursor {
    const char* p;
    const char* end;
    bool eol() const { return p >= end || *p == '\n' || *p == '\r'; }
    void skip_blank()
    {
        while (p < end && (*p == ' ' || *p == '\t')) p++;
    }
    void skip_line()
    {
        while (p < end && *p != '\n') p++;
        if (p < end) p++;
    }
    // next whitespace-delimited token on this line
    bool token(const char*& b, const char*& e)
    {
        skip_blank();
        if (eol()) return false;
        b = p;
        while (p < end && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') p++;
        e = p;
        return true;
    }
};

bool parse_float(const char* b, const char* e, float& out)
{
    std::string s(b, e);
    char* stop = nullptr;
    out = std::strtof(s.c_str(), &stop);
    return stop == s.c_str() + s.size();
}

bool parse_int(const char* b, const char* e, long& out)
{
    if (b == e) return false;
    std::string s(b, e);
    char* stop = nullptr;
    out = std::strtol(s.c_str(), &stop, 10);
    return stop == s.c_str() + s.size();
}

// OBJ index (1-based, or negative = relative to the end) -> 0-based, -1 absent
bool resolve(long v, long count, int32_t& out)
{
    if (v > 0 && v <= count) {
        out = (int32_t)(v - 1);
        return true;
    }
    if (v < 0 && -v <= count) {
        out = (int32_t)(count + v);
        return true;
    }
    return false;
}

} // namespace

extern "C" int32_t mt_obj_parse(const char* text, uint64_t len, mt_obj_counts* counts, float* points,
                                float* normals, float* uvs, int32_t* faces)
{
    if (!text || !counts) {
        mt_set_error("mt_obj_parse: NULL argument");
        return MT_ERR_ARG;
    }
    const bool fill = points != nullptr;
    if (fill && !faces) {
        mt_set_error("mt_obj_parse: faces is NULL");
        return MT_ERR_ARG;
    }
    const mt_obj_counts cap = *counts;
    long nv = 0, nn = 0, nt = 0, nf = 0;
    // Material index of a face = its group's rank among the groups that
    // receive faces: `o`, `g` and `usemtl` each start a new group (several
    // in a row start one), faces before any of them form group 0.
    int32_t group = -1;         // index of the current group, -1 until its first face
    int32_t groups_used = 0;
    Cursor c{text, text + len};
    long line = 0;
    while (c.p < c.end) {
        line++;
        const char *b, *e;
        if (!c.token(b, e) || *b == '#') {
            c.skip_line();
            continue;
        }
        const std::string key(b, e);
        if (key == "v" || key == "vn" || key == "vt") {
            float x[3] = {0.0f, 0.0f, 0.0f};
            const int need = key == "vt" ? 2 : 3;
            for (int k = 0; k < need; k++) {
                if (!c.token(b, e) || !parse_float(b, e, x[k])) {
                    mt_set_error("mt_obj_parse: line %ld: bad '%s' record", line, key.c_str());
                    return MT_ERR_ARG;
                }
            }
            if (key == "v") {
                if (fill && nv < cap.num_points) {
                    float* q = points + 4 * (size_t)nv;
                    q[0] = x[0]; q[1] = x[1]; q[2] = x[2]; q[3] = 0.0f;
                }
                nv++;
            } else if (key == "vn") {
                if (fill && normals && nn < cap.num_normals) {
                    float* q = normals + 4 * (size_t)nn;
                    q[0] = x[0]; q[1] = x[1]; q[2] = x[2]; q[3] = 0.0f;
                }
                nn++;
            } else {
                if (fill && uvs && nt < cap.num_uvs) {
                    uvs[2 * (size_t)nt] = x[0];
                    uvs[2 * (size_t)nt + 1] = x[1];
                }
                nt++;
            }
        } else if (key == "o" || key == "g" || key == "usemtl") {
            group = -1;
        } else if (key == "f") {
            if (group < 0) group = groups_used++;
            int32_t vi[3], ti[3], ni[3];
            int corners = 0;
            while (c.token(b, e)) {
                // v, v/t, v/t/n or v//n
                const char* s1 = static_cast<const char*>(std::memchr(b, '/', (size_t)(e - b)));
                const char* s2 = s1 ? static_cast<const char*>(std::memchr(s1 + 1, '/', (size_t)(e - s1 - 1))) : nullptr;
                long v, t = 0, n = 0;
                int32_t cv, ct = -1, cn = -1;
                bool ok = parse_int(b, s1 ? s1 : e, v) && resolve(v, nv, cv);
                if (ok && s1) {
                    const char* te = s2 ? s2 : e;
                    if (te > s1 + 1) ok = parse_int(s1 + 1, te, t) && resolve(t, nt, ct);
                }
                if (ok && s2 && e > s2 + 1) ok = parse_int(s2 + 1, e, n) && resolve(n, nn, cn);
                if (!ok) {
                    mt_set_error("mt_obj_parse: line %ld: bad face corner '%s'", line, std::string(b, e).c_str());
                    return MT_ERR_ARG;
                }
                if (corners < 2) {
                    vi[corners] = cv; ti[corners] = ct; ni[corners] = cn;
                } else { // fan triangulation (first corner, previous, this)
                    vi[2] = cv; ti[2] = ct; ni[2] = cn;
                    if (fill && nf < cap.num_tris) {
                        int32_t* F = faces + 10 * (size_t)nf;
                        for (int k = 0; k < 3; k++) {
                            F[k] = vi[k];
                            F[3 + k] = ti[k];
                            F[6 + k] = ni[k];
                        }
                        F[9] = group;
                    }
                    nf++;
                    vi[1] = vi[2]; ti[1] = ti[2]; ni[1] = ni[2];
                }
                corners++;
            }
            if (corners < 3) {
                mt_set_error("mt_obj_parse: line %ld: face with %d corners", line, corners);
                return MT_ERR_ARG;
            }
        }
        c.skip_line();
    }
    if (fill && (nv > cap.num_points || nf > cap.num_tris || (normals && nn > cap.num_normals) ||
                 (uvs && nt > cap.num_uvs))) {
        mt_set_error("mt_obj_parse: output arrays smaller than the counts");
        return MT_ERR_CAPACITY;
    }
    if (nv > INT32_MAX || nf > INT32_MAX) {
        mt_set_error("mt_obj_parse: more than 2^31 records");
        return MT_ERR_CAPACITY;
    }
    counts->num_points = (int32_t)nv;
    counts->num_normals = (int32_t)nn;
    counts->num_uvs = (int32_t)nt;
    counts->num_tris = (int32_t)nf;
    counts->num_groups = groups_used;
    return MT_OK;
}

// ---------------------------------------------------------------- RGBE
namespace {

inline void rgbe_to_float(const uint8_t* c, float* out)
{ // Ward's RGBE: mantissas scaled by 2^(e - 136); e == 0 is black
    if (c[3] == 0) {
        out[0] = out[1] = out[2] = 0.0f;
    } else {
        const float f = std::ldexp(1.0f, (int)c[3] - (128 + 8));
        out[0] = (float)c[0] * f;
        out[1] = (float)c[1] * f;
        out[2] = (float)c[2] * f;
    }
    out[3] = 1.0f;
}

} // namespace

extern "C" int32_t mt_hdr_decode(const uint8_t* data, uint64_t len, int32_t* width, int32_t* height, float* rgba)
{
    if (!data || !width || !height) {
        mt_set_error("mt_hdr_decode: NULL argument");
        return MT_ERR_ARG;
    }
    const uint8_t* p = data;
    const uint8_t* end = data + len;
    auto next_line = [&](std::string& s) -> bool {
        s.clear();
        while (p < end && *p != '\n') s.push_back((char)*p++);
        if (p >= end) return false;
        p++;
        return true;
    };
    std::string s;
    if (!next_line(s) || s.compare(0, 2, "#?") != 0) {
        mt_set_error("mt_hdr_decode: not a Radiance file (no #? signature)");
        return MT_ERR_ARG;
    }
    for (;;) { // header ends with an empty line
        if (!next_line(s)) {
            mt_set_error("mt_hdr_decode: truncated header");
            return MT_ERR_ARG;
        }
        if (s.empty()) break;
        if (s.compare(0, 7, "FORMAT=") == 0 && s != "FORMAT=32-bit_rle_rgbe") {
            mt_set_error("mt_hdr_decode: unsupported %s", s.c_str());
            return MT_ERR_ARG;
        }
    }
    if (!next_line(s)) {
        mt_set_error("mt_hdr_decode: missing resolution line");
        return MT_ERR_ARG;
    }
    int h = 0, w = 0;
    char ys = 0, xs = 0, yc = 0, xc = 0;
    if (std::sscanf(s.c_str(), "%c%c %d %c%c %d", &ys, &yc, &h, &xs, &xc, &w) != 6 || yc != 'Y' || xc != 'X' ||
        ys != '-' || xs != '+' || w <= 0 || h <= 0) {
        mt_set_error("mt_hdr_decode: unsupported resolution line '%s' (need -Y h +X w)", s.c_str());
        return MT_ERR_ARG;
    }
    if ((int64_t)w * h > ((int64_t)1 << 30)) { // float4 pixels: 16 GiB, beyond any environment map
        mt_set_error("mt_hdr_decode: %d x %d pixels exceed the 2^30 limit", w, h);
        return MT_ERR_ARG;
    }
    *width = w;
    *height = h;
    if (!rgba) return MT_OK;
    std::vector<uint8_t> line((size_t)w * 4);
    for (int y = 0; y < h; y++) {
        if (end - p >= 4 && w >= 8 && w < 32768 && p[0] == 2 && p[1] == 2 && ((p[2] << 8) | p[3]) == w) {
            // adaptive RLE: four channel planes, each of runs and literals
            p += 4;
            for (int ch = 0; ch < 4; ch++) {
                int x = 0;
                while (x < w) {
                    if (p >= end) goto truncated;
                    int n = *p++;
                    if (n > 128) {
                        n -= 128;
                        if (n > w - x || p >= end) goto corrupt;
                        const uint8_t v = *p++;
                        for (int k = 0; k < n; k++) line[4 * (size_t)(x + k) + ch] = v;
                    } else {
                        if (n == 0 || n > w - x || end - p < n) goto corrupt;
                        for (int k = 0; k < n; k++) line[4 * (size_t)(x + k) + ch] = *p++;
                    }
                    x += n;
                }
            }
        } else {
            // flat pixels, with the old run form (1,1,1,n): repeat the last pixel
            int x = 0, shift = 0;
            while (x < w) {
                if (end - p < 4) goto truncated;
                if (p[0] == 1 && p[1] == 1 && p[2] == 1) {
                    // a run count takes at most 3 consecutive records (24 bits);
                    // a longer chain cannot describe a line of < 2^31 pixels
                    if (x == 0 || shift > 16) goto corrupt;
                    const long n = (long)p[3] << shift;
                    if (n > w - x) goto corrupt;
                    for (long k = 0; k < n; k++) std::memcpy(&line[4 * (size_t)(x + k)], &line[4 * (size_t)(x - 1)], 4);
                    x += (int)n;
                    shift += 8;
                } else {
                    std::memcpy(&line[4 * (size_t)x], p, 4);
                    x++;
                    shift = 0;
                }
                p += 4;
            }
        }
        for (int x = 0; x < w; x++) rgbe_to_float(&line[4 * (size_t)x], rgba + 4 * ((size_t)y * w + x));
    }
    return MT_OK;
truncated:
    mt_set_error("mt_hdr_decode: truncated pixel data");
    return MT_ERR_ARG;
corrupt:
    mt_set_error("mt_hdr_decode: corrupt run-length data");
    return MT_ERR_ARG;
}
