// Host-side scene preparation: the BVH builder and the seeded workload
// generators.  No HIP calls; usable without a GPU.
//
// mt_bvh_build restates CNormalBVH.build/subdivide/getMid/calculateBounds
// (src/cl/scene/CNormalBVH.java:45-148).  The reference allocates nodes
// depth-first with one shared counter (nodesPtr, :110-111): a node processed
// when the counter reads `base` gets children base and base+1, its left
// subtree of nl primitives then allocates 2*nl-2 nodes, so its right subtree
// starts at base + 2*nl.  Computing those bases up front makes subtrees
// independent, so they are built on several threads while producing the
// reference's single-threaded numbering exactly.  The partition step and the
// split-axis choice come from the un-vendored OpenCLTracing jar
// (CNormalBVH.java:118,132): a two-pointer partition (centroid < split goes
// left) and pbrt's largest-diagonal axis are used.
#include "../../include/mambart.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

void mt_set_error(const char* fmt, ...);

namespace {

struct Box3 {
    float mn[3], mx[3];
    void init()
    { // CBound.java:22-34
        for (int i = 0; i < 3; i++) { mn[i] = INFINITY; mx[i] = -INFINITY; }
    }
    void include(const float* p)
    { // CBound.include, CBound.java:54-70
        for (int i = 0; i < 3; i++) {
            if (p[i] < mn[i]) mn[i] = p[i];
            if (p[i] > mx[i]) mx[i] = p[i];
        }
    }
};

struct BvhBuilder {
    int32_t* nodes;
    float* bounds;
    std::vector<int32_t> objects;
    std::vector<Box3> prim_box;
    std::vector<float> centroid;
    std::atomic<int32_t> max_depth{0};
    std::atomic<int> spare_threads{0};

    static int axis_of(const Box3& b)
    {
        float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
        if (dx > dy && dx > dz) return 0;
        if (dy > dz) return 1;
        return 2;
    }

    int32_t partition(int32_t start, int32_t end, int dim, float split)
    {
        int32_t l = start, r = end - 1;
        for (;;) {
            while (l <= r && centroid[3 * (size_t)objects[l] + dim] < split) l++;
            while (l <= r && centroid[3 * (size_t)objects[r] + dim] >= split) r--;
            if (l > r) break;
            std::swap(objects[l], objects[r]);
            l++;
            r--;
        }
        return l;
    }

    void build(int32_t node, int32_t start, int32_t end, int32_t base, int32_t depth)
    {
        for (;;) {
            Box3 bb, bc;
            bb.init();
            bc.init();
            for (int32_t p = start; p < end; p++) {
                const Box3& pb = prim_box[objects[p]];
                bb.include(pb.mn);
                bb.include(pb.mx);
                bc.include(&centroid[3 * (size_t)objects[p]]);
            }
            int32_t* pn = nodes + 7 * (size_t)node;
            float* pbd = bounds + 8 * (size_t)node;
            pn[0] = node;
            pbd[0] = bb.mn[0]; pbd[1] = bb.mn[1]; pbd[2] = bb.mn[2]; pbd[3] = 0.0f;
            pbd[4] = bb.mx[0]; pbd[5] = bb.mx[1]; pbd[6] = bb.mx[2]; pbd[7] = 0.0f;
            int32_t md = max_depth.load(std::memory_order_relaxed);
            while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {}
            if (end - start < 2) {
                pn[6] = objects[start];
                pn[5] = 1;
                return;
            }
            int32_t li = base, ri = base + 1;
            int32_t* ln = nodes + 7 * (size_t)li;
            int32_t* rn = nodes + 7 * (size_t)ri;
            std::memset(ln, 0, 7 * sizeof(int32_t));
            std::memset(rn, 0, 7 * sizeof(int32_t));
            ln[4] = node;
            rn[4] = node;
            pn[2] = li;
            pn[3] = ri;
            ln[1] = ri;
            rn[1] = li;
            int dim = axis_of(bc);
            float split = 0.5f * (bc.mn[dim] + bc.mx[dim]);
            int32_t mid = partition(start, end, dim, split);
            if (mid == start || mid == end) mid = start + (end - start) / 2;
            int32_t nl = mid - start;
            int32_t left_base = base + 2;
            int32_t right_base = base + 2 * nl;
            bool forked = false;
            std::thread th;
            if (end - start > 32768 && spare_threads.fetch_sub(1) > 0) {
                th = std::thread([=] { build(li, start, mid, left_base, depth + 1); });
                forked = true;
            } else if (end - start > 32768) {
                spare_threads.fetch_add(1);
            }
            if (!forked) build(li, start, mid, left_base, depth + 1);
            // continue with the right subtree in this thread (tail iteration)
            node = ri;
            start = mid;
            base = right_base;
            depth = depth + 1;
            if (forked) {
                build(node, start, end, base, depth);
                th.join();
                spare_threads.fetch_add(1);
                return;
            }
        }
    }
};

// splitmix64 (Steele, Lea, Flood 2014): the generators' only random source.
inline uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream)
    {
        s = seed * 0xD1B54A32D192ED03ull + stream * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        splitmix64(s);
    }
    double uniform() { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
    void unit_vector(double* v)
    { // uniform on S^2: z = 1 - 2u, phi = 2 pi w
        double z = 1.0 - 2.0 * uniform();
        double phi = 2.0 * M_PI * uniform();
        double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        v[0] = r * std::cos(phi);
        v[1] = r * std::sin(phi);
        v[2] = z;
    }
    double normal()
    { // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

template <class F>
void parallel_for(int64_t n, int nthreads, F f)
{
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    if (nthreads < 1) nthreads = 1;
    if (n < 4096 || nthreads == 1) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    int64_t chunk = (n + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; k++) {
        int64_t b = k * chunk, e = std::min<int64_t>(n, (k + 1) * chunk);
        if (b >= e) break;
        th.emplace_back([=] { f(b, e); });
    }
    for (auto& t : th) t.join();
}

} // namespace

extern "C" int32_t mt_bvh_build(int32_t ntris, const float* points, const int32_t* faces,
                                int32_t* nodes, float* bounds, int32_t* max_depth,
                                int32_t nthreads)
{
    if (ntris < 1 || !points || !faces || !nodes || !bounds) {
        mt_set_error("mt_bvh_build: invalid arguments (ntris=%d)", ntris);
        return MT_ERR_ARG;
    }
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    BvhBuilder B;
    B.nodes = nodes;
    B.bounds = bounds;
    try {
        B.objects.resize(ntris);
        B.prim_box.resize(ntris);
        B.centroid.resize(3 * (size_t)ntris);
    } catch (...) {
        mt_set_error("mt_bvh_build: out of host memory");
        return MT_ERR_ALLOC;
    }
    parallel_for(ntris, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) {
            B.objects[i] = (int32_t)i;
            Box3& bx = B.prim_box[i];
            bx.init(); // CMesh.getBound, CMesh.java:64-71
            for (int k = 0; k < 3; k++) bx.include(points + 4 * (size_t)faces[10 * i + k]);
            for (int k = 0; k < 3; k++) B.centroid[3 * i + k] = 0.5f * (bx.mn[k] + bx.mx[k]);
        }
    });
    std::memset(nodes, 0, 7 * sizeof(int32_t));
    B.spare_threads = nthreads - 1;
    B.build(0, 0, ntris, 1, 0);
    if (max_depth) *max_depth = B.max_depth.load();
    return MT_OK;
}

extern "C" int32_t mt_gen_soup(uint64_t seed, int32_t n, float c_lo, float c_hi,
                               float e_lo, float e_hi, int32_t first_vertex, int32_t mat,
                               float* points, int32_t* faces)
{
    if (n < 0 || !points || !faces) {
        mt_set_error("mt_gen_soup: invalid arguments");
        return MT_ERR_ARG;
    }
    parallel_for(n, 0, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) {
            Rng r(seed, (uint64_t)i);
            double c[3], u1[3], u2[3];
            for (int k = 0; k < 3; k++) c[k] = r.range(c_lo, c_hi);
            r.unit_vector(u1);
            double l1 = r.range(e_lo, e_hi);
            r.unit_vector(u2);
            double l2 = r.range(e_lo, e_hi);
            double p0[3], p1[3], p2[3];
            for (int k = 0; k < 3; k++) {
                double e1 = l1 * u1[k], e2 = l2 * u2[k];
                p0[k] = c[k] - (e1 + e2) / 3.0;
                p1[k] = p0[k] + e1;
                p2[k] = p0[k] + e2;
            }
            float* P = points + 12 * i;
            for (int k = 0; k < 3; k++) {
                P[k] = (float)p0[k];
                P[4 + k] = (float)p1[k];
                P[8 + k] = (float)p2[k];
            }
            P[3] = P[7] = P[11] = 0.0f;
            int32_t* F = faces + 10 * i;
            int32_t v = first_vertex + 3 * (int32_t)i;
            F[0] = v; F[1] = v + 1; F[2] = v + 2;
            for (int k = 3; k < 9; k++) F[k] = -1;
            F[9] = mat;
        }
    });
    return MT_OK;
}

extern "C" int32_t mt_gen_room(int32_t mat, float* points, int32_t* faces)
{
    if (!points || !faces) return MT_ERR_ARG;
    for (int i = 0; i < 8; i++) {
        points[4 * i + 0] = (i & 1) ? 1.0f : -1.0f;
        points[4 * i + 1] = (i & 2) ? 1.0f : -1.0f;
        points[4 * i + 2] = (i & 4) ? 1.0f : -1.0f;
        points[4 * i + 3] = 0.0f;
    }
    // two triangles per face, wound so that cross(e1, e2) points inward
    static const int tri[12][3] = {
        {0, 1, 2}, {1, 3, 2},   // z = -1 face, normal +z
        {4, 6, 5}, {5, 6, 7},   // z = +1 face, normal -z
        {0, 4, 1}, {1, 4, 5},   // y = -1 face, normal +y
        {2, 3, 6}, {3, 7, 6},   // y = +1 face, normal -y
        {0, 2, 4}, {2, 6, 4},   // x = -1 face, normal +x
        {1, 5, 3}, {3, 5, 7},   // x = +1 face, normal -x
    };
    for (int t = 0; t < 12; t++) {
        int32_t* F = faces + 10 * t;
        F[0] = tri[t][0]; F[1] = tri[t][1]; F[2] = tri[t][2];
        for (int k = 3; k < 9; k++) F[k] = -1;
        F[9] = mat;
    }
    return MT_OK;
}

extern "C" int32_t mt_gen_camera_rays(const float* pos, const float* look, const float* up,
                                      float vfov_deg, int32_t w, int32_t h,
                                      float* o_tmin, float* d_tmax)
{
    if (!pos || !look || !up || w <= 0 || h <= 0 || !o_tmin || !d_tmax) return MT_ERR_ARG;
    double f[3] = {(double)look[0] - pos[0], (double)look[1] - pos[1], (double)look[2] - pos[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& x : r) x /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    double th = std::tan(0.5 * vfov_deg * M_PI / 180.0);
    double aspect = (double)w / (double)h;
    for (int32_t py = 0; py < h; py++)
        for (int32_t px = 0; px < w; px++) {
            double sx = (2.0 * (px + 0.5) / w - 1.0) * th * aspect;
            double sy = (1.0 - 2.0 * (py + 0.5) / h) * th;
            double d[3];
            for (int k = 0; k < 3; k++) d[k] = f[k] + sx * r[k] + sy * u[k];
            double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            size_t i = (size_t)py * w + px;
            for (int k = 0; k < 3; k++) {
                o_tmin[4 * i + k] = pos[k];
                d_tmax[4 * i + k] = (float)(d[k] / dl);
            }
            o_tmin[4 * i + 3] = 0.001f;
            d_tmax[4 * i + 3] = INFINITY;
        }
    return MT_OK;
}

extern "C" int32_t mt_gen_random_rays(uint64_t seed, int64_t first, int64_t n,
                                      const float* lo, const float* hi,
                                      float* o_tmin, float* d_tmax, int32_t nthreads)
{
    if (n < 0 || !lo || !hi || !o_tmin || !d_tmax) return MT_ERR_ARG;
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t j = b; j < e; j++) {
            Rng r(seed, (uint64_t)(first + j));
            double d[3];
            for (int k = 0; k < 3; k++) o_tmin[4 * j + k] = (float)r.range(lo[k], hi[k]);
            r.unit_vector(d);
            for (int k = 0; k < 3; k++) d_tmax[4 * j + k] = (float)d[k];
            o_tmin[4 * j + 3] = 0.001f;
            d_tmax[4 * j + 3] = INFINITY;
        }
    });
    return MT_OK;
}

extern "C" int32_t mt_gen_spheres(uint64_t seed, int32_t count, float mu, float sigma,
                                  float c_lo, float c_hi, float ground_y, float g,
                                  float* points, float* normals, int32_t* faces)
{
    if (count < 0 || !points || !normals || !faces) return MT_ERR_ARG;
    const int SL = 64, RG = 33, VPS = SL * RG, TPS = 2 * SL * (RG - 1);
    parallel_for(count, 0, [&](int64_t b, int64_t e) {
        for (int64_t s = b; s < e; s++) {
            Rng r(seed, (uint64_t)s);
            double c[3];
            for (int k = 0; k < 3; k++) c[k] = r.range(c_lo, c_hi);
            double rad = std::exp((double)mu + (double)sigma * r.normal());
            int32_t v0 = (int32_t)s * VPS;
            for (int j = 0; j < RG; j++) {
                double th = M_PI * j / (RG - 1);
                for (int i = 0; i < SL; i++) {
                    double ph = 2.0 * M_PI * i / SL;
                    double l[3] = {std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph)};
                    size_t v = (size_t)v0 + (size_t)j * SL + i;
                    for (int k = 0; k < 3; k++) {
                        points[4 * v + k] = (float)(c[k] + rad * l[k]);
                        normals[4 * v + k] = (float)l[k];
                    }
                    points[4 * v + 3] = 0.0f;
                    normals[4 * v + 3] = 0.0f;
                }
            }
            int32_t* F = faces + 10 * (size_t)s * TPS;
            int t = 0;
            for (int j = 0; j < RG - 1; j++)
                for (int i = 0; i < SL; i++) {
                    int32_t a = v0 + j * SL + i, bb = v0 + j * SL + (i + 1) % SL;
                    int32_t cc = v0 + (j + 1) * SL + i, d = v0 + (j + 1) * SL + (i + 1) % SL;
                    int32_t q[2][3] = {{a, cc, bb}, {bb, cc, d}};
                    for (int h = 0; h < 2; h++, t++) {
                        int32_t* G = F + 10 * t;
                        G[0] = q[h][0]; G[1] = q[h][1]; G[2] = q[h][2];
                        G[3] = G[4] = G[5] = -1;
                        G[6] = q[h][0]; G[7] = q[h][1]; G[8] = q[h][2];
                        G[9] = (int32_t)s;
                    }
                }
        }
    });
    size_t gv = (size_t)count * VPS;
    const float gx[4] = {-g, g, g, -g}, gz[4] = {-g, -g, g, g};
    for (int k = 0; k < 4; k++) {
        points[4 * (gv + k) + 0] = gx[k];
        points[4 * (gv + k) + 1] = ground_y;
        points[4 * (gv + k) + 2] = gz[k];
        points[4 * (gv + k) + 3] = 0.0f;
        normals[4 * (gv + k) + 0] = 0.0f;
        normals[4 * (gv + k) + 1] = 1.0f;
        normals[4 * (gv + k) + 2] = 0.0f;
        normals[4 * (gv + k) + 3] = 0.0f;
    }
    int32_t* F = faces + 10 * (size_t)count * TPS;
    const int32_t q[2][3] = {{0, 2, 1}, {0, 3, 2}};
    for (int h = 0; h < 2; h++) {
        int32_t* G = F + 10 * h;
        for (int k = 0; k < 3; k++) G[k] = (int32_t)gv + q[h][k];
        for (int k = 3; k < 9; k++) G[k] = -1;
        G[9] = count;
    }
    return MT_OK;
}
