// Host-side code of the library (OBJ / RGBE readers, BVH builder, workload
// generators) driven over well-formed and malformed inputs; built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitize.py,
// so any out-of-bounds access, overflow or undefined shift aborts the run.
// Exit status 0: every call returned (an error code on bad input is fine).
#include "../../include/mambart.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

static int obj(const std::string& text)
{
    mt_obj_counts c;
    std::memset(&c, 0, sizeof(c));
    int32_t rc = mt_obj_parse(text.data(), text.size(), &c, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    std::vector<float> p(4 * (size_t)c.num_points + 4), n(4 * (size_t)c.num_normals + 4), t(2 * (size_t)c.num_uvs + 2);
    std::vector<int32_t> f(10 * (size_t)c.num_tris + 10);
    return mt_obj_parse(text.data(), text.size(), &c, p.data(), c.num_normals ? n.data() : nullptr,
                        c.num_uvs ? t.data() : nullptr, f.data());
}

static int hdr(const std::vector<uint8_t>& data)
{
    int32_t w = 0, h = 0;
    int32_t rc = mt_hdr_decode(data.data(), data.size(), &w, &h, nullptr);
    if (rc) return rc;
    std::vector<float> out(4 * (size_t)w * (size_t)h);
    return mt_hdr_decode(data.data(), data.size(), &w, &h, out.data());
}

static std::vector<uint8_t> bytes(const std::string& s) { return std::vector<uint8_t>(s.begin(), s.end()); }

int main()
{
    std::mt19937 rng(7);
    // ---- OBJ
    const char* objs[] = {
        "", "v 1 2 3\n", "v 1 2\nf 1 2 3\n", "f 1 2 3\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf -1 -2 -3\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 99999999\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\nf 1/1/1 2/1/1 3/1/1\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//1 2//1 3//1\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/ 2/ 3/\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 3 4\ng a\nusemtl m\nf 4 3 2\n",
        "v nan inf -inf\nv 1e40 0 0\nv 0 1 0\nf 1 2 3", "v 0 0 0\nf 1 1\n", "f\n", "vn\nvt\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf -4 -5 -6\n",
        "v 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\nf 1 2 3\r\n", "# comment only\n\n\n",
    };
    for (const char* t : objs) obj(t);
    for (int k = 0; k < 300; k++) { // random bytes and random token soup
        std::string s;
        const int n = (int)(rng() % 400);
        const char* toks[] = {"v ", "vt ", "vn ", "f ", "g x\n", "o y\n", "usemtl z\n", "/", "//", "-", "1", "2",
                              "3", "0", "1e9", ".5", " ", "\n", "#", "nan"};
        for (int i = 0; i < n; i++) {
            if (k % 3 == 0) s.push_back((char)(rng() & 0xff));
            else s += toks[rng() % 20];
        }
        obj(s);
    }
    std::string big(1 << 20, 'v');
    obj(big);

    // ---- RGBE
    std::string head = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n";
    hdr(bytes(""));
    hdr(bytes("#?RADIANCE\n"));
    hdr(bytes(head));
    hdr(bytes(head + "-Y 2 +X 2\n"));
    hdr(bytes(head + "-Y 0 +X 2\n"));
    hdr(bytes(head + "-Y 99999999 +X 99999999\n"));
    hdr(bytes(head + "+Y 2 +X 2\n\x01\x02\x03\x04"));
    {   // flat pixels, then old-style runs chained far past 24 bits
        std::vector<uint8_t> d = bytes(head + "-Y 1 +X 64\n");
        for (int i = 0; i < 4; i++) d.insert(d.end(), {128, 64, 32, 130});
        for (int i = 0; i < 12; i++) d.insert(d.end(), {1, 1, 1, 255});
        hdr(d);
    }
    {   // new-style run-length lines with bad counts and truncation
        std::vector<uint8_t> d = bytes(head + "-Y 2 +X 16\n");
        d.insert(d.end(), {2, 2, 0, 16});
        d.insert(d.end(), {200, 5});      // run past the line
        hdr(d);
        std::vector<uint8_t> e = bytes(head + "-Y 2 +X 16\n");
        e.insert(e.end(), {2, 2, 0, 16, 0}); // zero-length literal
        hdr(e);
        std::vector<uint8_t> g = bytes(head + "-Y 2 +X 16\n");
        g.insert(g.end(), {2, 2, 0, 16, 16, 1, 2});  // truncated literal
        hdr(g);
    }
    for (int k = 0; k < 300; k++) {  // valid header, random pixel bytes
        std::vector<uint8_t> d = bytes(head + "-Y 3 +X " + std::to_string(8 + k % 40) + "\n");
        const int n = (int)(rng() % 600);
        for (int i = 0; i < n; i++) d.push_back((uint8_t)(k % 2 ? rng() & 0xff : (rng() % 3 ? 1 : rng() & 0xff)));
        hdr(d);
    }
    {   // a well-formed new-style image
        const int w = 32, h = 4;
        std::vector<uint8_t> d = bytes(head + "-Y 4 +X 32\n");
        for (int y = 0; y < h; y++) {
            d.insert(d.end(), {2, 2, (uint8_t)(w >> 8), (uint8_t)(w & 255)});
            for (int ch = 0; ch < 4; ch++) {
                d.push_back(128 + 16);
                d.push_back((uint8_t)(100 + ch));
                d.push_back(16);
                for (int i = 0; i < 16; i++) d.push_back((uint8_t)(i + ch));
            }
        }
        if (hdr(d) != MT_OK) {
            std::fprintf(stderr, "well-formed RGBE image rejected\n");
            return 1;
        }
    }

    // ---- BVH builder and generators
    for (int n : {1, 2, 3, 17, 1000, 20000}) {
        std::vector<float> p(4 * 3 * (size_t)n);
        std::vector<int32_t> f(10 * (size_t)n);
        if (mt_gen_soup(5 + n, n, -1.f, 1.f, 0.01f, 0.2f, 0, 0, p.data(), f.data())) return 1;
        std::vector<int32_t> nodes(7 * (size_t)(2 * n - 1));
        std::vector<float> bounds(8 * (size_t)(2 * n - 1));
        int32_t depth = 0;
        if (mt_bvh_build(n, p.data(), f.data(), nodes.data(), bounds.data(), &depth, 4)) return 1;
    }
    {   // coincident triangles (median-split fallback)
        const int n = 513;
        std::vector<float> p(4 * 3 * (size_t)n, 0.0f);
        std::vector<int32_t> f(10 * (size_t)n, -1);
        for (int i = 0; i < n; i++) {
            p[12 * i + 4] = 1.0f;
            p[12 * i + 9] = 1.0f;
            f[10 * i] = 3 * i; f[10 * i + 1] = 3 * i + 1; f[10 * i + 2] = 3 * i + 2; f[10 * i + 9] = 0;
        }
        std::vector<int32_t> nodes(7 * (size_t)(2 * n - 1));
        std::vector<float> bounds(8 * (size_t)(2 * n - 1));
        int32_t depth = 0;
        if (mt_bvh_build(n, p.data(), f.data(), nodes.data(), bounds.data(), &depth, 2)) return 1;
    }
    mt_bvh_build(0, nullptr, nullptr, nullptr, nullptr, nullptr, 1);   // rejected
    {
        std::vector<float> p(12 * 4), rn(12 * 4);
        std::vector<int32_t> f(12 * 10);
        mt_gen_room(0, p.data(), f.data());
        const float pos[3] = {0, 0, -0.95f}, look[3] = {0, 0, 1}, up[3] = {0, 1, 0};
        std::vector<float> o(4 * 17 * 9), d(4 * 17 * 9);
        if (mt_gen_camera_rays(pos, look, up, 60.0f, 17, 9, o.data(), d.data())) return 1;
        const float lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};
        std::vector<float> ro(4 * 1000), rd(4 * 1000);
        if (mt_gen_random_rays(3, 12345, 1000, lo, hi, ro.data(), rd.data(), 3)) return 1;
        const int count = 3, tris = count * 4096 + 2;
        std::vector<float> sp(4 * (size_t)tris * 3), sn(4 * (size_t)tris * 3);
        std::vector<int32_t> sf(10 * (size_t)tris);
        if (mt_gen_spheres(2, count, -2.5f, 0.5f, -5.f, 5.f, -5.f, 50.f, sp.data(), sn.data(), sf.data())) return 1;
    }
    char msg[256];
    mt_last_error(msg, sizeof(msg));
    std::printf("host check done (last error: %s)\n", msg);
    return 0;
}
