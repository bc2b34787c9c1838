/*
 * ORACLE (test infrastructure only) — stage A: BVH build, stackless
 * closest-hit / any-hit traversal, hit shading, shadow-ray setup.
 *
 * Parity status
 *   orc_bvh_build      restates src/cl/scene/CNormalBVH.java:45-148.  The
 *                      partition step (CNormalBVH.java:132) and
 *                      maximumExtentAxis (:118) live in the un-vendored
 *                      OpenCLTracing jar: the two-pointer partition and the
 *                      pbrt largest-diagonal axis used here are our choice.
 *                      Node ORDER is therefore parity unpinned; the tree is
 *                      pinned only by invariants (2N-1 nodes, bound == index,
 *                      parent/sibling links, leaves a permutation, boxes
 *                      nested) checked in tests/test_oracle_bvh.py.
 *   orc_trace_closest  restates NormalBVH.cl:14-92 + Primitive.cl:3-29 +
 *                      Util.cl:552-555,577-600,707-732; pinned against brute
 *                      force over all triangles (orc_trace_brute) and the
 *                      analytic room-box distances.
 *   orc_trace_any      restates NormalBVH.cl:235-317; pinned by
 *                      any-hit == (closest hit exists with t < tMax).
 *   orc_shade_hits     restates NormalBVH.cl:124-220 + Primitive.cl:31-71;
 *                      pinned by barycentric reconstruction (p = sum b_i p_i)
 *                      and unit-length/orientation checks.
 *
 * Deviation (documented): the reference's bitstack is a 64-bit long
 * (NormalBVH.cl:20) and silently corrupts traversal past depth 64.  The
 * oracle carries 128 bits (identical behaviour for depth <= 64) and reports
 * ORC_ERR_DEPTH for trees deeper than 128.
 */
#include "orc.h"
#include "orc_math.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* BVH builder: CNormalBVH.java:45-148                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    float mn[3], mx[3];
} obox;

static void obox_init(obox* b)
{ /* CBound.java:22-34 */
    for (int i = 0; i < 3; i++) { b->mn[i] = INFINITY; b->mx[i] = -INFINITY; }
}

static void obox_include_pt(obox* b, const float* p)
{ /* CBound.java:54-70 */
    for (int i = 0; i < 3; i++) {
        if (p[i] < b->mn[i]) b->mn[i] = p[i];
        if (p[i] > b->mx[i]) b->mx[i] = p[i];
    }
}

typedef struct {
    const float* points;
    const int32_t* faces;
    int32_t* nodes;
    float* bounds;
    int32_t* objects;
    obox* prim_box;     /* CMesh.getBound(primID), CMesh.java:64-71 */
    float* centroid;    /* CBound.getCenter(), CBound.java:72-79 */
    int32_t nodes_ptr;
    int32_t max_depth;
} obuild;

static int max_extent_axis(const obox* b)
{ /* un-vendored AbstractBound.maximumExtentAxis: pbrt largest diagonal */
    float dx = b->mx[0] - b->mn[0], dy = b->mx[1] - b->mn[1], dz = b->mx[2] - b->mn[2];
    if (dx > dy && dx > dz) return 0;
    if (dy > dz) return 1;
    return 2;
}

static int32_t partition_objs(obuild* B, int32_t start, int32_t end, int dim, float split)
{ /* un-vendored AbstractAccelerator.partition: two-pointer, left = centroid < split */
    int32_t l = start, r = end - 1;
    for (;;) {
        while (l <= r && B->centroid[3 * (size_t)B->objects[l] + dim] < split) l++;
        while (l <= r && B->centroid[3 * (size_t)B->objects[r] + dim] >= split) r--;
        if (l > r) break;
        int32_t tmp = B->objects[l]; B->objects[l] = B->objects[r]; B->objects[r] = tmp;
        l++; r--;
    }
    return l;
}

static void subdivide(obuild* B, int32_t parent, int32_t start, int32_t end, int32_t depth)
{ /* CNormalBVH.java:81-124 */
    obox bb, bc;
    obox_init(&bb); obox_init(&bc);
    for (int32_t p = start; p < end; p++) { /* calculateBounds, :141-148 */
        const obox* pb = &B->prim_box[B->objects[p]];
        obox_include_pt(&bb, pb->mn);
        obox_include_pt(&bb, pb->mx);
        obox_include_pt(&bc, &B->centroid[3 * (size_t)B->objects[p]]);
    }
    int32_t* pn = B->nodes + 7 * (size_t)parent;
    float* pbd = B->bounds + 8 * (size_t)parent;
    pn[0] = parent; /* bound == own index, :88 */
    pbd[0] = bb.mn[0]; pbd[1] = bb.mn[1]; pbd[2] = bb.mn[2]; pbd[3] = 0.0f;
    pbd[4] = bb.mx[0]; pbd[5] = bb.mx[1]; pbd[6] = bb.mx[2]; pbd[7] = 0.0f;
    if (depth > B->max_depth) B->max_depth = depth;
    if (end - start < 2) { /* leaf, :92-98 */
        pn[6] = B->objects[start];
        pn[5] = 1;
        return;
    }
    int32_t li = B->nodes_ptr++;
    int32_t ri = B->nodes_ptr++;
    int32_t* ln = B->nodes + 7 * (size_t)li;
    int32_t* rn = B->nodes + 7 * (size_t)ri;
    memset(ln, 0, 7 * sizeof(int32_t));
    memset(rn, 0, 7 * sizeof(int32_t));
    ln[4] = parent; rn[4] = parent;     /* parent links, :107-108 */
    pn[2] = li; pn[3] = ri;             /* :110-111 */
    ln[1] = ri; rn[1] = li;             /* siblings, :113-114 */
    int dim = max_extent_axis(&bc);      /* :118 */
    float split = 0.5f * (bc.mn[dim] + bc.mx[dim]); /* getCenter(dim), :129 */
    int32_t mid = partition_objs(B, start, end, dim, split);
    if (mid == start || mid == end) mid = start + (end - start) / 2; /* :135-136 */
    subdivide(B, li, start, mid, depth + 1);
    subdivide(B, ri, mid, end, depth + 1);
}

int orc_bvh_build(int32_t ntris, const float* points, const int32_t* faces,
                  int32_t* nodes, float* bounds, int32_t* max_depth)
{
    if (ntris < 1 || !points || !faces || !nodes || !bounds) return ORC_ERR_ARG;
    obuild B;
    memset(&B, 0, sizeof(B));
    B.points = points; B.faces = faces; B.nodes = nodes; B.bounds = bounds;
    B.objects = (int32_t*)malloc(sizeof(int32_t) * (size_t)ntris);
    B.prim_box = (obox*)malloc(sizeof(obox) * (size_t)ntris);
    B.centroid = (float*)malloc(sizeof(float) * 3 * (size_t)ntris);
    if (!B.objects || !B.prim_box || !B.centroid) {
        free(B.objects); free(B.prim_box); free(B.centroid);
        return ORC_ERR_ARG;
    }
    for (int32_t i = 0; i < ntris; i++) {
        B.objects[i] = i;
        obox* b = &B.prim_box[i];
        obox_init(b);
        for (int k = 0; k < 3; k++) obox_include_pt(b, points + 4 * (size_t)faces[10 * (size_t)i + k]);
        for (int k = 0; k < 3; k++) B.centroid[3 * (size_t)i + k] = 0.5f * (b->mn[k] + b->mx[k]);
    }
    memset(nodes, 0, sizeof(int32_t) * 7); /* root = new CNode(), :62-64 */
    B.nodes_ptr = 1;
    subdivide(&B, 0, 0, ntris, 0);
    if (max_depth) *max_depth = B.max_depth;
    free(B.objects); free(B.prim_box); free(B.centroid);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Traversal                                                           */
/* ------------------------------------------------------------------ */

typedef struct {
    f4 o, d, inv_d;
    int sign[3];
    float tmin, tmax;
} oray;

static void oray_init(oray* r, const float* o_tmin, const float* d_tmax)
{ /* Util.cl:577-600: inv_d = 1/d, sign = inv_d < 0 */
    r->o = f4_make(o_tmin[0], o_tmin[1], o_tmin[2], 0.0f);
    r->d = f4_make(d_tmax[0], d_tmax[1], d_tmax[2], 0.0f);
    r->inv_d = f4_make(1.0f / r->d.x, 1.0f / r->d.y, 1.0f / r->d.z, 0.0f);
    r->sign[0] = r->inv_d.x < 0.0f;
    r->sign[1] = r->inv_d.y < 0.0f;
    r->sign[2] = r->inv_d.z < 0.0f;
    r->tmin = o_tmin[3];
    r->tmax = d_tmax[3];
}

static int box_hit(const oray* r, const float* b, float* t_near)
{ /* intersectBoundT, Util.cl:707-732 (getExtent :661-667) */
    const float* ext0 = b;      /* minimum */
    const float* ext1 = b + 4;  /* maximum */
    float tmin  = ((r->sign[0] ? ext1 : ext0)[0] - r->o.x) * r->inv_d.x;
    float tmax  = ((r->sign[0] ? ext0 : ext1)[0] - r->o.x) * r->inv_d.x;
    float tymin = ((r->sign[1] ? ext1 : ext0)[1] - r->o.y) * r->inv_d.y;
    float tymax = ((r->sign[1] ? ext0 : ext1)[1] - r->o.y) * r->inv_d.y;
    if ((tmin > tymax) || (tymin > tmax)) return 0;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = ((r->sign[2] ? ext1 : ext0)[2] - r->o.z) * r->inv_d.z;
    float tzmax = ((r->sign[2] ? ext0 : ext1)[2] - r->o.z) * r->inv_d.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return 0;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    tmax *= 1.00000024f;
    *t_near = tmin;
    return (tmin < r->tmax) && (tmax > r->tmin);
}

static float tri_hit(const oray* r, f4 p1, f4 p2, f4 p3)
{ /* fastTriangleIntersection, Primitive.cl:3-29 */
    f4 e1 = f4_sub(p2, p1);
    f4 e2 = f4_sub(p3, p1);
    f4 h = cross3(r->d, e2);
    float a = dot3(e1, h);
    if ((double)a > -0.0000001 && (double)a < 0.0000001) return r->tmax; /* double compare, :13 */
    float f = 1.0f / a;
    f4 s = f4_sub(r->o, p1);
    float b1 = f * dot3(s, h);
    f4 q = cross3(s, e1);
    float b2 = f * dot3(r->d, q);
    float t = f * dot3(e2, q);
    if (b1 < 0.0f || b1 > 1.0f || b2 < 0.0f || b1 + b2 > 1.0f || t < 0.0f || t > r->tmax)
        return r->tmax;
    return t;
}

typedef struct {
    const float* points;
    const int32_t* faces;
    const int32_t* nodes;
    const float* bounds;
} oscene;

static inline f4 vtx(const oscene* S, int32_t prim, int k)
{ /* getP1/2/3, Util.cl:229-247 */
    return f4_load(S->points + 4 * (size_t)S->faces[10 * (size_t)prim + k]);
}

/* intersectMesh, NormalBVH.cl:14-92.  stats = {interior steps, leaf visits, pops}.
 * Returns 1 hit, 0 miss, -1 depth overflow. */
static int intersect_mesh(const oscene* S, oray* r, int32_t* child, int bailout, int32_t* stats)
{
    int hit = 0;
    int32_t node_id = 0;
    uint64_t bs_lo = 0, bs_hi = 0; /* 128-bit bitstack */
    int32_t parent_id = 0, sibling_id = 0;
    int32_t depth = 0;
    for (;;) {
        while (!S->nodes[7 * (size_t)node_id + 5]) {
            const int32_t* n = S->nodes + 7 * (size_t)node_id;
            parent_id = n[4];
            sibling_id = n[1];
            const int32_t* ln = S->nodes + 7 * (size_t)n[2];
            const int32_t* rn = S->nodes + 7 * (size_t)n[3];
            float lt, rt;
            int lh = box_hit(r, S->bounds + 8 * (size_t)ln[0], &lt);
            int rh = box_hit(r, S->bounds + 8 * (size_t)rn[0], &rt);
            stats[0]++;
            if (!lh && !rh) break;
            if (++depth > 128) return -1;
            bs_hi = (bs_hi << 1) | (bs_lo >> 63);
            bs_lo <<= 1;
            if (lh && rh) {
                node_id = (rt < lt) ? n[3] : n[2];
                bs_lo |= 1;
            } else {
                node_id = lh ? n[2] : n[3];
            }
        }
        if (S->nodes[7 * (size_t)node_id + 5]) {
            const int32_t* n = S->nodes + 7 * (size_t)node_id;
            stats[1]++;
            int32_t c = n[6];
            float t = tri_hit(r, vtx(S, c, 0), vtx(S, c, 1), vtx(S, c, 2));
            if ((r->tmin < t) && (t < r->tmax)) { /* isInside, Util.cl:552-555 */
                if (bailout) return 1;
                hit = 1;
                r->tmax = t;
                *child = c;
            }
            parent_id = n[4];
            sibling_id = n[1];
        }
        while ((bs_lo & 1) == 0) {
            if (bs_lo == 0 && bs_hi == 0) return hit;
            node_id = parent_id;
            const int32_t* n = S->nodes + 7 * (size_t)node_id;
            parent_id = n[4];
            sibling_id = n[1];
            bs_lo = (bs_lo >> 1) | (bs_hi << 63);
            bs_hi >>= 1;
            depth--;
            stats[2]++;
        }
        node_id = sibling_id;
        bs_lo ^= 1;
    }
}

typedef struct {
    const oscene* S;
    int64_t begin, end;
    const float* o_tmin;
    const float* d_tmax;
    int32_t* prim;
    float* t;
    int32_t* occluded;
    int32_t* ray_stats;
    int64_t tot[3];
    int any;
    int err;
} ojob;

static void* trace_worker(void* arg)
{
    ojob* J = (ojob*)arg;
    for (int64_t i = J->begin; i < J->end; i++) {
        oray r;
        oray_init(&r, J->o_tmin + 4 * i, J->d_tmax + 4 * i);
        int32_t st[3] = {0, 0, 0};
        int32_t child = -1;
        int h = intersect_mesh(J->S, &r, &child, J->any, st);
        if (h < 0) { J->err = ORC_ERR_DEPTH; h = 0; }
        if (J->any) {
            J->occluded[i] = h;
        } else {
            J->prim[i] = h ? child : -1;
            J->t[i] = r.tmax; /* ray->tMax is mutated in place, NormalBVH.cl:71 */
        }
        if (J->ray_stats) {
            J->ray_stats[3 * i + 0] = st[0];
            J->ray_stats[3 * i + 1] = st[1];
            J->ray_stats[3 * i + 2] = st[2];
        }
        J->tot[0] += st[0]; J->tot[1] += st[1]; J->tot[2] += st[2];
    }
    return NULL;
}

static int run_trace(const float* points, const int32_t* faces, const int32_t* nodes,
                     const float* bounds, int64_t nrays, const float* o_tmin,
                     const float* d_tmax, int32_t* prim, float* t, int32_t* occluded,
                     int32_t* ray_stats, int64_t* total_stats, int32_t nthreads, int any)
{
    if (nrays < 0 || !points || !faces || !nodes || !bounds || !o_tmin || !d_tmax) return ORC_ERR_ARG;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    oscene S = {points, faces, nodes, bounds};
    ojob jobs[256];
    pthread_t th[256];
    int64_t chunk = (nrays + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; k++) {
        ojob* J = &jobs[k];
        memset(J, 0, sizeof(*J));
        J->S = &S;
        J->begin = k * chunk < nrays ? k * chunk : nrays;
        J->end = (k + 1) * chunk < nrays ? (k + 1) * chunk : nrays;
        J->o_tmin = o_tmin; J->d_tmax = d_tmax;
        J->prim = prim; J->t = t; J->occluded = occluded; J->ray_stats = ray_stats;
        J->any = any;
    }
    if (nthreads == 1) {
        trace_worker(&jobs[0]);
    } else {
        for (int k = 0; k < nthreads; k++) pthread_create(&th[k], NULL, trace_worker, &jobs[k]);
        for (int k = 0; k < nthreads; k++) pthread_join(th[k], NULL);
    }
    int err = ORC_OK;
    int64_t tot[3] = {0, 0, 0};
    for (int k = 0; k < nthreads; k++) {
        if (jobs[k].err) err = jobs[k].err;
        for (int q = 0; q < 3; q++) tot[q] += jobs[k].tot[q];
    }
    if (total_stats) { total_stats[0] = tot[0]; total_stats[1] = tot[1]; total_stats[2] = tot[2]; }
    return err;
}

int orc_trace_closest(const float* points, const int32_t* faces,
                      const int32_t* nodes, const float* bounds,
                      int64_t nrays, const float* o_tmin, const float* d_tmax,
                      int32_t* prim, float* t, int32_t* ray_stats,
                      int64_t* total_stats, int32_t nthreads)
{ /* IntersectPrimitives' traversal part, NormalBVH.cl:94-123 */
    if (!prim || !t) return ORC_ERR_ARG;
    return run_trace(points, faces, nodes, bounds, nrays, o_tmin, d_tmax, prim, t, NULL,
                     ray_stats, total_stats, nthreads, 0);
}

int orc_trace_any(const float* points, const int32_t* faces,
                  const int32_t* nodes, const float* bounds,
                  int64_t nrays, const float* o_tmin, const float* d_tmax,
                  int32_t* occluded, int32_t* ray_stats,
                  int64_t* total_stats, int32_t nthreads)
{ /* intersectOcclusion / testOcclusion, NormalBVH.cl:235-348 */
    if (!occluded) return ORC_ERR_ARG;
    return run_trace(points, faces, nodes, bounds, nrays, o_tmin, d_tmax, NULL, NULL, occluded,
                     ray_stats, total_stats, nthreads, 1);
}

int orc_trace_brute(const float* points, const int32_t* faces, int32_t ntris,
                    int64_t nrays, const float* o_tmin, const float* d_tmax,
                    int32_t* prim, float* t)
{ /* Pinning check: the same triangle test applied to every triangle in index order. */
    oscene S = {points, faces, NULL, NULL};
    for (int64_t i = 0; i < nrays; i++) {
        oray r;
        oray_init(&r, o_tmin + 4 * i, d_tmax + 4 * i);
        int32_t best = -1;
        for (int32_t c = 0; c < ntris; c++) {
            float tt = tri_hit(&r, vtx(&S, c, 0), vtx(&S, c, 1), vtx(&S, c, 2));
            if ((r.tmin < tt) && (tt < r.tmax)) { r.tmax = tt; best = c; }
        }
        prim[i] = best;
        t[i] = r.tmax;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Hit shading: IntersectPrimitives, NormalBVH.cl:124-220              */
/* ------------------------------------------------------------------ */

static void barycentrics(f4 p, f4 p1, f4 p2, f4 p3, float* b1, float* b2)
{ /* triangleBarycentrics, Primitive.cl:31-53 */
    f4 e1 = f4_sub(p2, p1), e2 = f4_sub(p3, p1), e = f4_sub(p, p1);
    float d00 = dot3(e1, e1), d01 = dot3(e1, e2), d11 = dot3(e2, e2);
    float d20 = dot3(e, e1), d21 = dot3(e, e2);
    float denom = (d00 * d11 - d01 * d01);
    if (denom == 0.0f) { *b1 = 0.0f; *b2 = 0.0f; return; }
    float inv = 1.0f / denom;
    *b1 = (d11 * d20 - d01 * d21) * inv;
    *b2 = (d00 * d21 - d01 * d20) * inv;
}

static void uv_barycentrics(f4 p, f4 p1, f4 p2, f4 p3, const float* uv1, const float* uv2,
                            const float* uv3, float* out)
{ /* triangleBarycentricsFromUVMesh, Primitive.cl:56-71 */
    f4 f1 = f4_sub(p1, p), f2 = f4_sub(p2, p), f3 = f4_sub(p3, p);
    float a = length3(cross3(f4_sub(p1, p2), f4_sub(p1, p3)));
    float a1 = length3(cross3(f2, f3)) / a;
    float a2 = length3(cross3(f3, f1)) / a;
    float a3 = length3(cross3(f1, f2)) / a;
    out[0] = (uv1[0] * a1 + uv2[0] * a2) + uv3[0] * a3;
    out[1] = (uv1[1] * a1 + uv2[1] * a2) + uv3[1] * a3;
}

int orc_shade_hits(const float* points, const float* normals, const float* uvs,
                   const int32_t* faces, int64_t nrays, const float* o_tmin,
                   const float* d_tmax, const int32_t* prim, const float* t,
                   float* p_out, float* n_out, float* ng_out, float* uv_out, int32_t* mat_out)
{
    for (int64_t i = 0; i < nrays; i++) {
        int32_t c = prim[i];
        if (c < 0) {
            mat_out[i] = -1; /* MISS: hit = 0, mat = -1, NormalBVH.cl:224-225 */
            for (int k = 0; k < 4; k++) { p_out[4 * i + k] = 0; n_out[4 * i + k] = 0; ng_out[4 * i + k] = 0; }
            uv_out[2 * i] = 0; uv_out[2 * i + 1] = 0;
            continue;
        }
        const int32_t* F = faces + 10 * (size_t)c;
        f4 d = f4_make(d_tmax[4 * i], d_tmax[4 * i + 1], d_tmax[4 * i + 2], 0.0f);
        f4 o = f4_make(o_tmin[4 * i], o_tmin[4 * i + 1], o_tmin[4 * i + 2], 0.0f);
        f4 p1 = f4_load(points + 4 * (size_t)F[0]);
        f4 p2 = f4_load(points + 4 * (size_t)F[1]);
        f4 p3 = f4_load(points + 4 * (size_t)F[2]);
        float tt = t[i];
        /* getPoint, Util.cl:558-565 (.w left 0 here) */
        f4 p = f4_make(o.x + tt * d.x, o.y + tt * d.y, o.z + tt * d.z, 0.0f);
        float texuv[2] = {0, 0};
        int has_uv = F[3] > -1; /* hasUV, Util.cl:297-301 */
        if (has_uv && uvs)
            uv_barycentrics(p, p1, p2, p3, uvs + 2 * (size_t)F[3], uvs + 2 * (size_t)F[4],
                            uvs + 2 * (size_t)F[5], texuv);
        float u, v;
        barycentrics(p, p1, p2, p3, &u, &v);
        f4 n;
        if (F[6] > -1 && normals) { /* hasNormals, Util.cl:270-274 */
            f4 n1 = f4_load(normals + 4 * (size_t)F[6]);
            f4 n2 = f4_load(normals + 4 * (size_t)F[7]);
            f4 n3 = f4_load(normals + 4 * (size_t)F[8]);
            n1.w = n2.w = n3.w = 0.0f;
            float w0 = (1.0f - u) - v;
            n = f4_add(f4_add(f4_scale(n1, w0), f4_scale(n2, u)), f4_scale(n3, v));
            if (dot3(n, d) > 0.00001f) {
                n = f4_neg(n); n1 = f4_neg(n1); n2 = f4_neg(n2); n3 = f4_neg(n3);
            }
            n = normalize3(n);
            /* Hanika shadow-terminator offset, NormalBVH.cl:175-193 */
            f4 tu = f4_sub(p, p1), tv = f4_sub(p, p2), tw = f4_sub(p, p3);
            f4 nn1 = normalize3(n1), nn2 = normalize3(n2), nn3 = normalize3(n3);
            float du = fminf(0.0f, dot3(tu, nn1));
            float dv = fminf(0.0f, dot3(tv, nn2));
            float dw = fminf(0.0f, dot3(tw, nn3));
            tu = f4_sub(tu, f4_scale(nn1, du));
            tv = f4_sub(tv, f4_scale(nn2, dv));
            tw = f4_sub(tw, f4_scale(nn3, dw));
            tu.w = tv.w = tw.w = 0.0f;
            p = f4_add(f4_add(f4_add(p, f4_scale(tu, w0)), f4_scale(tv, u)), f4_scale(tw, v));
        } else {
            n = normalize3(cross3(f4_sub(p2, p1), f4_sub(p3, p1))); /* getNormal, Util.cl:349-357 */
            if (dot3(n, d) > 0.00001f) n = f4_neg(n);
        }
        f4 ng = normalize3(cross3(f4_sub(p2, p1), f4_sub(p3, p1)));
        if (dot3(ng, d) > 0.00001f) ng = f4_neg(ng);
        p_out[4 * i] = p.x; p_out[4 * i + 1] = p.y; p_out[4 * i + 2] = p.z; p_out[4 * i + 3] = 0.0f;
        n_out[4 * i] = n.x; n_out[4 * i + 1] = n.y; n_out[4 * i + 2] = n.z; n_out[4 * i + 3] = 0.0f;
        ng_out[4 * i] = ng.x; ng_out[4 * i + 1] = ng.y; ng_out[4 * i + 2] = ng.z; ng_out[4 * i + 3] = 0.0f;
        if (has_uv && uvs) { uv_out[2 * i] = texuv[0]; uv_out[2 * i + 1] = texuv[1]; }
        else { uv_out[2 * i] = u; uv_out[2 * i + 1] = v; }
        mat_out[i] = F[9] & 0xFFFF; /* getMaterial, Util.cl:11-14 */
    }
    return ORC_OK;
}

int orc_shadow_rays(int64_t n, const float* o_tmin, const float* d_tmax,
                    const int32_t* prim, const float* t, const float* light,
                    float* s_o_tmin, float* s_d_tmax, int32_t* valid)
{ /* shadow ray toward a point: illuminateAreaLight's direction/distance
     (Light.cl:93-97) and initGlobalRayT(p, dir, dist - 2*EPS_RAY) (Render.cl:453,
     Util.cl:590-600).  The origin is getPoint(ray, t) (Util.cl:558-565). */
    for (int64_t i = 0; i < n; i++) {
        int ok = prim[i] >= 0;
        valid[i] = ok;
        float tt = ok ? t[i] : 0.0f;
        float px = o_tmin[4 * i] + tt * d_tmax[4 * i];
        float py = o_tmin[4 * i + 1] + tt * d_tmax[4 * i + 1];
        float pz = o_tmin[4 * i + 2] + tt * d_tmax[4 * i + 2];
        float vx = light[0] - px, vy = light[1] - py, vz = light[2] - pz;
        float dist2 = (vx * vx + vy * vy) + vz * vz;
        float dist = sqrtf(dist2);
        s_o_tmin[4 * i] = px; s_o_tmin[4 * i + 1] = py; s_o_tmin[4 * i + 2] = pz;
        s_o_tmin[4 * i + 3] = 0.0001f;
        s_d_tmax[4 * i] = vx / dist; s_d_tmax[4 * i + 1] = vy / dist; s_d_tmax[4 * i + 2] = vz / dist;
        s_d_tmax[4 * i + 3] = dist - 2 * 1e-3f;
    }
    return ORC_OK;
}
