// Multi-job wavefront engine: many render jobs share one set of wavefront
// queues (continuous batching).  The per-path work is shade_path
// (mt_path.h), the same code mt_render runs, and path RNG is keyed by
// (job seed, path id, bounce, purpose), so a job's image is bit-identical to
// its plain mt_render however it is mixed with other jobs.
//
// Device layout.
//   job table   JobRec[max_jobs]: scene view, camera, sizes, seed and the
//               job's per-path radiance buffer (a slice of the arena);
//               MtSceneRef[max_jobs]: the packed traversal layout per job;
//               npaths/issued int32[max_jobs] counters; done counters
//               int32[DONE_SEGS][max_jobs] (work-group b counts into row
//               b % DONE_SEGS) summed per job into done_tot[max_jobs].
//   queue       two copies (ping-pong) of four SoA streams per slot:
//               tag int2 (job << 6 | bounce, path id), ray o/tmin and
//               d/tmax float4, throughput float4 (w = last BSDF pdf);
//               plus per-iteration scratch: hit prim/t, alive flag,
//               shadow ray o/d, contribution and slot tag.
// One iteration (all device-side, counts never leave the device):
//   1. k_refill (one wave): free slots = capacity - live; hand them to the
//      active jobs in admission order (wave prefix sum over the jobs'
//      remaining paths), advancing each job's issued counter;
//   2. k_generate: camera rays for the new slots (binary search of the
//      slot's job in the hand-out table);
//   3. closest hit over the queue, each ray against its job's scene;
//   4. k_shade_jobs: shade_path per slot with its job's scene view; ended
//      paths count into their job's done counter (wave-aggregated atomics);
//      shadow rays go to a segmented queue (MtSegQueue) with their slot's
//      tag;
//   5. k_shadow_jobs: any hit over the shadow queue, each unoccluded ray
//      adding its light to its path's radiance;
//   6. stable compaction of live slots (4 streams) into the other copy;
//   7. k_done_sum folds the done rows of the active jobs.
// The host then reads the done totals (4 B per job slot), resolves the
// images of finished jobs (k_resolve, fixed sample order) and admits waiting
// jobs into the freed slots and arena space.
#include <algorithm>
#include <deque>
#include <map>
#include <vector>

#include "mt_path.h"
#include "mt_traverse.h"

namespace {

constexpr int BT = 256;
constexpr int BOUNCE_BITS = 6; // max_bounces <= 63
// Rows of done counters: a job's ended paths are counted by many waves of
// one launch, and one counter per job would serialise their atomics.
constexpr int DONE_SEGS = 64;
// Shadow-queue segments of the multi-job engine (its iterations are smaller
// than a whole render's bounces, so fewer, fuller segments).
constexpr int JOB_SEGS = 256;
static_assert(JOB_SEGS % 64 == 0, "k_refill clears the segment counters one wave-width at a time");

struct JobRec {
    SceneView V;
    Camera C;
    float4* L;         // per-path radiance, pid-indexed
    int npaths, max_bounces, rr_start, flags;
    uint32_t seed;
};

struct Queue {
    int2* tag;
    float4 *ro, *rd, *T;
};

// One wave of DONE_SEGS lanes.
__global__ void k_set_job(JobRec* jobs, MtSceneRef* refs, int* npaths, int* issued, int* done, int* done_tot,
                          int* cancel, int max_jobs, int slot, JobRec rec, MtSceneRef ref)
{
    if (threadIdx.x < DONE_SEGS) done[(size_t)threadIdx.x * max_jobs + slot] = 0;
    if (threadIdx.x == 0) {
        jobs[slot] = rec;
        refs[slot] = ref;
        npaths[slot] = rec.npaths;
        issued[slot] = 0;
        done_tot[slot] = 0;
        cancel[slot] = 0;
    }
}

// Cancels the job in `slot`: it is issued no more paths (its path count
// drops to what it has been issued) and its paths in the queue end at their
// next shading (k_shade_jobs), so it finishes once they have drained.
__global__ void k_cancel_job(int* npaths, const int* issued, int* cancel, int slot)
{
    if (threadIdx.x == 0) {
        npaths[slot] = issued[slot];
        cancel[slot] = 1;
    }
}

// done_tot[j] = sum of job j's DONE_SEGS rows, for the active jobs, and
// fin[j] = 1 once every path the job will ever be issued has ended.
__global__ __launch_bounds__(BT) void k_done_sum(const int* __restrict__ active, const int* __restrict__ nact,
                                                 const int* __restrict__ done, int max_jobs,
                                                 const int* __restrict__ npaths, int* __restrict__ done_tot,
                                                 int* __restrict__ fin)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nact) return;
    const int j = active[i];
    int s = 0;
    for (int r = 0; r < DONE_SEGS; r++) s += done[(size_t)r * max_jobs + j];
    done_tot[j] = s;
    fin[j] = s >= npaths[j] ? 1 : 0;
}

__global__ void k_zero_counts(int* cnt, unsigned long long* tot)
{
    if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
    if (threadIdx.x < 2) tot[threadIdx.x] = 0ull;
}

// One wave.  cnt[0] = live slots after the last compaction; writes
// cnt[1] = live + new.  gen_start[i] = first new-slot offset of active job
// i, gen_pid0[i] = its first new path id.
__global__ __launch_bounds__(64) void k_refill(const int* __restrict__ active, const int* __restrict__ nact,
                                               const int* __restrict__ npaths, int* __restrict__ issued,
                                               int capacity, int* __restrict__ cnt, int* __restrict__ gen_start,
                                               int* __restrict__ gen_pid0, int* __restrict__ gen_job,
                                               int* __restrict__ s_seg, unsigned long long* __restrict__ tot)
{
    const int lane = threadIdx.x;
    const int n_active = *nact;
    for (int p = lane; p < JOB_SEGS; p += 64) s_seg[p] = 0; // shadow queue of this iteration
    const int live = cnt[0];
    const int64_t avail = (int64_t)capacity - live;
    int64_t carry = 0;
    for (int base = 0; base < n_active; base += 64) {
        const int i = base + lane;
        int j = -1, iss = 0;
        int64_t rem = 0;
        if (i < n_active) {
            j = active[i];
            iss = issued[j];
            rem = (int64_t)npaths[j] - iss;
        }
        int64_t incl = rem;
        for (int o = 1; o < 64; o <<= 1) {
            int64_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const int64_t excl = incl - rem;
        const int64_t a = avail - carry > 0 ? avail - carry : 0;
        int64_t give = a - excl;
        give = give < 0 ? 0 : (give > rem ? rem : give);
        if (i < n_active) {
            gen_start[i] = (int)(carry + (excl < a ? excl : a));
            gen_pid0[i] = iss;
            gen_job[i] = j;
            issued[j] = iss + (int)give;
        }
        const int64_t chunk = __shfl(incl, 63);
        carry += chunk < a ? chunk : a;
    }
    if (lane == 0) {
        cnt[1] = live + (int)carry;
        tot[0] += (unsigned long long)(live + carry);
    }
}

__global__ __launch_bounds__(BT) void k_generate(const JobRec* __restrict__ jobs, const int* __restrict__ cnt,
                                                 const int* __restrict__ gen_start,
                                                 const int* __restrict__ gen_pid0, const int* __restrict__ gen_job,
                                                 const int* __restrict__ nact, Queue Q)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int live = cnt[0], total = cnt[1];
    if (slot < live || slot >= total) return;
    const int off = slot - live;
    // last i with gen_start[i] <= off (entries are non-decreasing; a job
    // given no slots shares its start with the next job, which wins)
    int lo = 0, hi = *nact - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (gen_start[mid] <= off) lo = mid;
        else hi = mid - 1;
    }
    const int j = gen_job[lo];
    const int pid = gen_pid0[lo] + (off - gen_start[lo]);
    const JobRec& J = jobs[j];
    float4 o, d;
    camera_ray(J.C, pid, J.seed, o, d);
    Q.ro[slot] = o;
    Q.rd[slot] = d;
    Q.T[slot] = make_float4(1.0f, 1.0f, 1.0f, 1.0f); // InitPathData, Initialize.cl:25-48
    Q.tag[slot] = make_int2(j << BOUNCE_BITS, pid);
    J.L[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

struct ShadeJobsArgs {
    const JobRec* jobs;
    const int* cancel;       // per job slot: 1 = cancelled, its paths end unshaded
    int* done;               // [DONE_SEGS][max_jobs]
    int max_jobs;
    const int* cnt;
    Queue Q;
    const int* prim;
    const float* tin;
    int* alive;
    float4 *s_o, *s_d, *s_c; // segmented shadow queue: rays and contributions
    int2* s_tag;             // the entry's slot tag (job word, path id)
    int* s_seg;              // JOB_SEGS segment lengths
    int seg_cap;
};

__global__ __launch_bounds__(BT) void k_shade_jobs(ShadeJobsArgs A)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int count = A.cnt[1];
    if ((int)(blockIdx.x * blockDim.x) >= count) return; // whole work-group idle
    const int lane = __lane_id();
    bool ended = false, has_shadow = false;
    int job = -1;
    float4 so, sd, sc;
    int2 stag;
    if (slot < count) {
        const int2 tg = A.Q.tag[slot];
        job = tg.x >> BOUNCE_BITS;
        const int bounce = tg.x & ((1 << BOUNCE_BITS) - 1), pid = tg.y;
        const JobRec& J = A.jobs[job];
        PathOut P;
        P.alive = false;
        P.has_shadow = false;
        P.Ladd = mk3(0.0f, 0.0f, 0.0f);
        if (!A.cancel[job])
            shade_path(J.V, J.seed, J.rr_start, J.flags, pid, bounce, A.Q.ro[slot], A.Q.rd[slot], A.Q.T[slot],
                       A.prim[slot], A.tin[slot], P);
        const bool alive = P.alive && bounce + 1 < J.max_bounces;
        if (alive) {
            A.Q.T[slot] = P.T;
            A.Q.ro[slot] = P.ro;
            A.Q.rd[slot] = P.rd;
            A.Q.tag[slot] = make_int2(tg.x + 1, pid);
        }
        if (P.Ladd.x != 0.0f || P.Ladd.y != 0.0f || P.Ladd.z != 0.0f) {
            float4 L = J.L[pid];
            J.L[pid] = make_float4(L.x + P.Ladd.x, L.y + P.Ladd.y, L.z + P.Ladd.z, L.w);
        }
        A.alive[slot] = alive ? 1 : 0;
        ended = !alive;
        has_shadow = P.has_shadow;
        so = P.so;
        sd = P.sd;
        sc = P.contrib;
        stag = make_int2(tg.x, pid);
    }
    const int seg = blockIdx.x % JOB_SEGS;
    const int k = seg * A.seg_cap + wave_append(has_shadow, A.s_seg + seg);
    if (has_shadow) {
        A.s_o[k] = so;
        A.s_d[k] = sd;
        A.s_c[k] = sc;
        A.s_tag[k] = stag;
    }
    // ended paths per job: one atomic per (wave, job) into this work-group's row
    int* done_row = A.done + (size_t)(blockIdx.x % DONE_SEGS) * A.max_jobs;
    uint64_t em = __ballot(ended);
    while (em) {
        const int leader = __builtin_ctzll(em);
        const int lj = __shfl(job, leader);
        const uint64_t same = __ballot(ended && job == lj);
        if (lane == leader) atomicAdd(&done_row[lj], (int)__popcll(same));
        em &= ~same;
    }
}

// Any hit over the segmented shadow queue, each ray against its job's scene,
// with the direct-light add fused in: an unoccluded entry adds its
// contribution to its path's radiance in its job's buffer.  Work-group 0
// adds the iteration's shadow rays to the running total.
__global__ __launch_bounds__(BT) void k_shadow_jobs(const JobRec* __restrict__ jobs,
                                                    const MtSceneRef* __restrict__ refs,
                                                    const int* __restrict__ s_seg, int seg_cap,
                                                    const int2* __restrict__ s_tag, const float4* __restrict__ s_o,
                                                    const float4* __restrict__ s_d,
                                                    const float4* __restrict__ contrib,
                                                    unsigned long long* __restrict__ tot)
{
    const int k = seg_entry<JOB_SEGS>(seg_cap); // chunk-major: live chunks dispatch first
    if (blockIdx.x == 0) {
        int c = 0;
        for (int p = threadIdx.x; p < JOB_SEGS; p += BT) c += s_seg[p];
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (__lane_id() == 0 && c) atomicAdd(&tot[1], (unsigned long long)c);
    }
    if (!seg_live<JOB_SEGS>(s_seg, seg_cap, k)) return;
    const int2 tg = s_tag[k];
    const int job = tg.x >> BOUNCE_BITS;
    const MtSceneRef R = refs[job];
    RayT r;
    ray_init(r, s_o[k], s_d[k]);
    int st[3] = {0, 0, 0};
    int child = -1;
    if (trace_packed<true, false, true>(R.pn, R.pt, R.root, r, child, st)) return;
    const float4 c = contrib[k];
    float4* L = jobs[job].L;
    const float4 v = L[tg.y];
    L[tg.y] = make_float4(v.x + c.x, v.y + c.y, v.z + c.z, v.w);
}

} // namespace

struct mt_batch {
    mt_context* ctx;
    int capacity, max_jobs;
    uint64_t arena_bytes;
    char* mem;   // fixed state (one allocation)
    char* arena; // per-job radiance buffers
    JobRec* jobs;
    MtSceneRef* refs;
    int *npaths, *issued, *done, *done_tot, *cancel, *fin;
    int *active, *gen_start, *gen_pid0, *gen_job;
    int* cnt;                 // [0] live after compaction, [1] live + new
    int* s_seg;               // shadow queue segment lengths
    int seg_cap;              // shadow queue segment capacity
    unsigned long long* tot;  // [0] extension rays, [1] shadow rays
    Queue q[2];
    int cur;
    int *prim, *alive;
    float* t;
    float4 *s_o, *s_d, *s_c;
    int2* s_tag;
    int* h_active;            // pinned
    int* h_done;              // pinned: fin flags of the job slots
    unsigned long long* h_tot; // pinned
    struct Job {
        mt_job_desc d;
        int slot;
        int npaths;
        uint64_t off, len;
        bool cancelled;
    };
    std::deque<Job> pending;
    std::vector<Job> active_jobs; // admission order
    std::vector<int> free_slots;
    std::map<uint64_t, uint64_t> free_blocks; // arena offset -> length
    std::deque<int64_t> finished;             // tags not yet reported
    bool active_dirty;
    int policy;
    int iters_per_step;       // wavefront iterations between host round trips
    int* nact;                // device: number of active jobs (read by the iteration kernels)
    int trace_mode;           // closest hit: MT_BATCH_TRACE_SPLIT (default), MT_TRACE_PACKED_STACK or _POOL
    // The iterations of a step as one hipGraph per start parity of the queue
    // ping-pong: every size they use lives in device memory, so a capture
    // stays valid until the iterations per step change.
    bool use_graphs;
    hipGraph_t graph[2];
    hipGraphExec_t gexec[2];
    int g_iters[2];
    void* gscratch;           // the captured compactions' look-back workspace
    size_t gscratch_bytes;
    mt_batch_stats st;
};

namespace {

int32_t validate_job(const mt_job_desc* j)
{
    MT_ARG_CHECK(j && j->scene && j->materials && j->image, "mt_batch_submit: NULL scene/materials/image");
    const mt_render_desc& r = j->render;
    MT_ARG_CHECK(r.width > 0 && r.height > 0 && r.spp > 0, "mt_batch_submit: empty image");
    MT_ARG_CHECK(r.max_bounces >= 1 && r.max_bounces <= 63, "mt_batch_submit: max_bounces %d not in [1,63]",
                 r.max_bounces);
    MT_ARG_CHECK((int64_t)r.width * r.height * r.spp <= (1ll << 30), "mt_batch_submit: more than 2^30 paths");
    MT_ARG_CHECK(r.mode == MT_TRACE_PACKED || r.mode == MT_TRACE_PACKED_STACK,
                 "mt_batch_submit: traversal mode %d not supported by the batch engine", r.mode);
    MT_ARG_CHECK((r.flags & ~MT_RENDER_PICKED_LIGHT) == 0, "mt_batch_submit: flags 0x%x not supported", r.flags);
    MT_ARG_CHECK(!j->light_faces == !j->d_num_lights, "mt_batch_submit: light_faces and d_num_lights go together");
    return MT_OK;
}

uint64_t job_bytes(const mt_render_desc& r)
{
    return (((uint64_t)r.width * r.height * r.spp * 16u) + 255u) & ~(uint64_t)255u;
}

bool arena_take(mt_batch* b, uint64_t len, uint64_t& off)
{
    for (auto it = b->free_blocks.begin(); it != b->free_blocks.end(); ++it) {
        if (it->second >= len) {
            off = it->first;
            uint64_t rest = it->second - len;
            b->free_blocks.erase(it);
            if (rest) b->free_blocks[off + len] = rest;
            return true;
        }
    }
    return false;
}

void arena_give(mt_batch* b, uint64_t off, uint64_t len)
{
    auto it = b->free_blocks.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != b->free_blocks.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        b->free_blocks.erase(nx);
    }
    if (it != b->free_blocks.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            b->free_blocks.erase(it);
        }
    }
}

int32_t admit(mt_batch* b)
{
    while (!b->pending.empty() && !b->free_slots.empty()) {
        mt_batch::Job& J = b->pending.front();
        uint64_t off;
        if (!arena_take(b, J.len, off)) break; // strict FIFO: a large job waits for space
        J.off = off;
        J.slot = b->free_slots.back();
        b->free_slots.pop_back();
        const mt_job_desc& d = J.d;
        JobRec rec;
        int32_t rc = camera_from_desc(&d.render, rec.C);
        if (rc) return rc;
        rec.V.points = reinterpret_cast<const float4*>(d.scene->desc.points);
        rec.V.normals = reinterpret_cast<const float4*>(d.scene->desc.normals);
        rec.V.faces = d.scene->desc.faces;
        rec.V.mats = reinterpret_cast<const MtMaterial*>(d.materials);
        rec.V.light_faces = d.light_faces;
        rec.V.num_lights = d.d_num_lights;
        rec.V.env = env_view(d.env);
        rec.V.G = LgView{nullptr, nullptr, nullptr, 0.0f, 0.0f, 0.0f, nullptr, 0}; // full-map environment sampling
        rec.L = reinterpret_cast<float4*>(b->arena + off);
        rec.npaths = J.npaths;
        rec.max_bounces = d.render.max_bounces;
        rec.rr_start = d.render.rr_start;
        rec.flags = d.render.flags;
        rec.seed = d.render.seed;
        MtSceneRef ref{d.scene->pnodes, d.scene->ptris, d.scene->root_ref, 0};
        k_set_job<<<1, DONE_SEGS, 0, b->ctx->stream>>>(b->jobs, b->refs, b->npaths, b->issued, b->done, b->done_tot,
                                                       b->cancel, b->max_jobs, J.slot, rec, ref);
        MT_LAUNCH_CHECK();
        b->st.arena_used += J.len;
        if (b->st.arena_used > b->st.arena_peak) b->st.arena_peak = b->st.arena_used;
        b->st.jobs_admitted++;
        b->active_jobs.push_back(J);
        b->pending.pop_front();
        b->active_dirty = true;
    }
    return MT_OK;
}

__global__ void k_set_nact(int* nact, int n) { *nact = n; }

// The active jobs' slots in hand-out order, and their count, to the device
// (stream-ordered; before the iterations that read them).
int32_t sync_active(mt_batch* b)
{
    hipStream_t s = b->ctx->stream;
    const int n_active = (int)b->active_jobs.size();
    if (b->active_dirty) {
        // hand-out order of free queue slots: admission order, or fewest
        // paths first (ties by admission order)
        std::vector<int> order(n_active);
        for (int i = 0; i < n_active; i++) order[i] = i;
        if (b->policy == MT_BATCH_SMALLEST_FIRST)
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
                return b->active_jobs[x].npaths < b->active_jobs[y].npaths;
            });
        for (int i = 0; i < n_active; i++) b->h_active[i] = b->active_jobs[order[i]].slot;
        MT_HIP_CHECK(hipMemcpyAsync(b->active, b->h_active, sizeof(int) * (size_t)std::max(n_active, 1),
                                    hipMemcpyHostToDevice, s));
        k_set_nact<<<1, 1, 0, s>>>(b->nact, n_active);
        MT_LAUNCH_CHECK();
        b->active_dirty = false;
    }
    return MT_OK;
}

// One wavefront iteration: launches only, every size from device memory
// (capturable).
int32_t iterate(mt_batch* b)
{
    hipStream_t s = b->ctx->stream;
    const int Qc = b->capacity;
    const int grid = mt_blocks(Qc, BT);
    Queue& q = b->q[b->cur];
    Queue& nq = b->q[1 - b->cur];
    k_refill<<<1, 64, 0, s>>>(b->active, b->nact, b->npaths, b->issued, Qc, b->cnt, b->gen_start, b->gen_pid0,
                              b->gen_job, b->s_seg, b->tot);
    k_generate<<<grid, BT, 0, s>>>(b->jobs, b->cnt, b->gen_start, b->gen_pid0, b->gen_job, b->nact, q);
    MT_LAUNCH_CHECK();
    int32_t rc = mt_trace_jobs(b->ctx, b->refs, &q.tag->x, BOUNCE_BITS, q.ro, q.rd, b->cnt + 1, Qc, b->prim,
                               b->t, b->trace_mode, b->cnt); // cnt[0]: survivors, cnt[1]: survivors + new
    if (rc) return rc;
    ShadeJobsArgs A;
    A.jobs = b->jobs;
    A.cancel = b->cancel;
    A.done = b->done;
    A.max_jobs = b->max_jobs;
    A.cnt = b->cnt;
    A.Q = q;
    A.prim = b->prim;
    A.tin = b->t;
    A.alive = b->alive;
    A.s_o = b->s_o;
    A.s_d = b->s_d;
    A.s_c = b->s_c;
    A.s_tag = b->s_tag;
    A.s_seg = b->s_seg;
    A.seg_cap = b->seg_cap;
    k_shade_jobs<<<grid, BT, 0, s>>>(A);
    MT_LAUNCH_CHECK();
    k_shadow_jobs<<<mt_blocks((int64_t)JOB_SEGS * b->seg_cap, BT), BT, 0, s>>>(
        b->jobs, b->refs, b->s_seg, b->seg_cap, b->s_tag, b->s_o, b->s_d, b->s_c, b->tot);
    MT_LAUNCH_CHECK();
    mt_payload P[4] = {{q.tag, nq.tag, 8}, {q.ro, nq.ro, 16}, {q.rd, nq.rd, 16}, {q.T, nq.T, 16}};
    rc = mt_compact(b->ctx, b->alive, 0, b->cnt + 1, Qc, nullptr, b->cnt, P, 4);
    if (rc) return rc;
    k_done_sum<<<mt_blocks(b->max_jobs, BT), BT, 0, s>>>(b->active, b->nact, b->done, b->max_jobs, b->npaths,
                                                      b->done_tot, b->fin);
    MT_LAUNCH_CHECK();
    b->cur = 1 - b->cur;
    return MT_OK;
}

void drop_graphs(mt_batch* b)
{
    for (int p = 0; p < 2; p++) {
        if (b->gexec[p]) (void)hipGraphExecDestroy(b->gexec[p]);
        if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
        b->gexec[p] = nullptr;
        b->graph[p] = nullptr;
        b->g_iters[p] = 0;
    }
}

// Records iters_per_step iterations starting at parity p into graph[p].
// Captured on a stream of its own (the context's may be the null stream)
// with a look-back workspace of the batch's own (the context's may be
// regrown and freed later, while the graph keeps addresses).
int32_t capture_iterations(mt_batch* b, int p)
{
    mt_context* ctx = b->ctx;
    if (b->gexec[p]) (void)hipGraphExecDestroy(b->gexec[p]);
    if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
    b->gexec[p] = nullptr;
    b->graph[p] = nullptr;
    b->g_iters[p] = 0;
    void* ctx_scratch = ctx->scratch;
    const size_t ctx_scratch_bytes = ctx->scratch_bytes;
    const hipStream_t ctx_stream = ctx->stream;
    int32_t rc = MT_OK;
    if (!b->gscratch) {
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        rc = mt_scan_reserve(ctx, b->capacity); // allocated and cleared on the context stream
        b->gscratch = ctx->scratch;
        b->gscratch_bytes = ctx->scratch_bytes;
        ctx->scratch = ctx_scratch;
        ctx->scratch_bytes = ctx_scratch_bytes;
        if (rc) return rc;
    }
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        mt_set_error("mt_batch_step: hipStreamCreate failed");
        return MT_ERR_HIP;
    }
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        mt_set_error("mt_batch_step: hipStreamBeginCapture failed: %s", hipGetErrorString(e));
        return MT_ERR_HIP;
    }
    ctx->stream = cs;
    ctx->scratch = b->gscratch;
    ctx->scratch_bytes = b->gscratch_bytes;
    for (int it = 0; it < b->iters_per_step && rc == MT_OK; it++) rc = iterate(b);
    ctx->stream = ctx_stream;
    ctx->scratch = ctx_scratch;
    ctx->scratch_bytes = ctx_scratch_bytes;
    b->cur = p; // nothing ran: the queues are where they were
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc == MT_OK && e != hipSuccess) {
        mt_set_error("mt_batch_step: hipStreamEndCapture failed: %s", hipGetErrorString(e));
        rc = MT_ERR_HIP;
    }
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    b->graph[p] = g;
    e = hipGraphInstantiate(&b->gexec[p], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        b->gexec[p] = nullptr;
        mt_set_error("mt_batch_step: hipGraphInstantiate failed: %s", hipGetErrorString(e));
        return MT_ERR_HIP;
    }
    b->g_iters[p] = b->iters_per_step;
    return MT_OK;
}

// The step's iterations: one launch of the captured graph, or the kernels.
int32_t run_iterations(mt_batch* b)
{
    int32_t rc = sync_active(b);
    if (rc) return rc;
    if (b->use_graphs) {
        const int p = b->cur;
        if (!b->gexec[p] || b->g_iters[p] != b->iters_per_step) {
            rc = capture_iterations(b, p);
            if (rc) return rc;
        }
        MT_HIP_CHECK(hipGraphLaunch(b->gexec[p], b->ctx->stream));
        b->cur = p ^ (b->iters_per_step & 1);
    } else {
        for (int it = 0; it < b->iters_per_step; it++) {
            rc = iterate(b);
            if (rc) return rc;
        }
    }
    b->st.iterations += b->iters_per_step;
    return MT_OK;
}

} // namespace

extern "C" int32_t mt_batch_job_bytes(const mt_render_desc* desc, uint64_t* bytes)
{
    MT_ARG_CHECK(desc && bytes, "mt_batch_job_bytes: NULL argument");
    MT_ARG_CHECK(desc->width > 0 && desc->height > 0 && desc->spp > 0, "mt_batch_job_bytes: empty image");
    *bytes = job_bytes(*desc);
    return MT_OK;
}

extern "C" int32_t mt_batch_create(mt_context* ctx, int32_t queue_capacity, uint64_t arena_bytes, int32_t max_jobs,
                                   mt_batch** out)
{
    MT_ARG_CHECK(ctx && out, "mt_batch_create: NULL argument");
    MT_ARG_CHECK(queue_capacity >= 1 && queue_capacity <= (1 << 30), "mt_batch_create: queue_capacity %d",
                 queue_capacity);
    MT_ARG_CHECK(max_jobs >= 1 && max_jobs <= (1 << 20), "mt_batch_create: max_jobs %d", max_jobs);
    MT_ARG_CHECK(arena_bytes >= 256, "mt_batch_create: arena of %llu bytes", (unsigned long long)arena_bytes);
    *out = nullptr;
    mt_batch* b = new (std::nothrow) mt_batch();
    if (!b) {
        mt_set_error("mt_batch_create: host allocation failed");
        return MT_ERR_ALLOC;
    }
    b->ctx = ctx;
    b->capacity = queue_capacity;
    b->max_jobs = max_jobs;
    b->arena_bytes = arena_bytes & ~(uint64_t)255u;
    const size_t n = (size_t)queue_capacity, J = (size_t)max_jobs;
    size_t bytes = 0;
    auto take = [&](size_t sz) { size_t off = bytes; bytes += (sz + 255) & ~(size_t)255; return off; };
    size_t oJobs = take(sizeof(JobRec) * J), oRefs = take(sizeof(MtSceneRef) * J);
    size_t oNp = take(4 * J), oIss = take(4 * J), oDone = take(4 * J * DONE_SEGS), oDtot = take(4 * J);
    size_t oCan = take(4 * J), oFin = take(4 * J);
    size_t oAct = take(4 * J), oGs = take(4 * J), oGp = take(4 * J), oGj = take(4 * J);
    size_t oCnt = take(16), oTot = take(16), oSeg = take(4 * JOB_SEGS), oNact = take(16);
    b->seg_cap = shadow_seg_cap<JOB_SEGS>(queue_capacity, BT);
    const size_t sq = (size_t)JOB_SEGS * b->seg_cap;
    size_t oQ[2][4];
    for (int k = 0; k < 2; k++) {
        oQ[k][0] = take(8 * n);
        for (int m = 1; m < 4; m++) oQ[k][m] = take(16 * n);
    }
    size_t oPrim = take(4 * n), oT = take(4 * n), oAlive = take(4 * n);
    size_t oSo = take(16 * sq), oSd = take(16 * sq), oSc = take(16 * sq), oStag = take(8 * sq);
    if (hipMalloc(&b->mem, bytes) != hipSuccess) {
        delete b;
        mt_set_error("mt_batch_create: cannot allocate %.2f GB of queue state", bytes / 1e9);
        return MT_ERR_ALLOC;
    }
    if (hipMalloc(&b->arena, b->arena_bytes) != hipSuccess) {
        (void)hipFree(b->mem);
        delete b;
        mt_set_error("mt_batch_create: cannot allocate an arena of %.2f GB", arena_bytes / 1e9);
        return MT_ERR_ALLOC;
    }
    if (hipHostMalloc(&b->h_active, 4 * J) != hipSuccess || hipHostMalloc(&b->h_done, 4 * J) != hipSuccess ||
        hipHostMalloc(&b->h_tot, 16) != hipSuccess) {
        if (b->h_active) (void)hipHostFree(b->h_active);
        if (b->h_done) (void)hipHostFree(b->h_done);
        (void)hipFree(b->arena);
        (void)hipFree(b->mem);
        delete b;
        mt_set_error("mt_batch_create: pinned host allocation failed");
        return MT_ERR_ALLOC;
    }
    char* m = b->mem;
    b->jobs = reinterpret_cast<JobRec*>(m + oJobs);
    b->refs = reinterpret_cast<MtSceneRef*>(m + oRefs);
    b->npaths = reinterpret_cast<int*>(m + oNp);
    b->issued = reinterpret_cast<int*>(m + oIss);
    b->done = reinterpret_cast<int*>(m + oDone);
    b->done_tot = reinterpret_cast<int*>(m + oDtot);
    b->cancel = reinterpret_cast<int*>(m + oCan);
    b->fin = reinterpret_cast<int*>(m + oFin);
    b->s_seg = reinterpret_cast<int*>(m + oSeg);
    b->active = reinterpret_cast<int*>(m + oAct);
    b->gen_start = reinterpret_cast<int*>(m + oGs);
    b->gen_pid0 = reinterpret_cast<int*>(m + oGp);
    b->gen_job = reinterpret_cast<int*>(m + oGj);
    b->cnt = reinterpret_cast<int*>(m + oCnt);
    b->nact = reinterpret_cast<int*>(m + oNact);
    b->tot = reinterpret_cast<unsigned long long*>(m + oTot);
    for (int k = 0; k < 2; k++) {
        b->q[k].tag = reinterpret_cast<int2*>(m + oQ[k][0]);
        b->q[k].ro = reinterpret_cast<float4*>(m + oQ[k][1]);
        b->q[k].rd = reinterpret_cast<float4*>(m + oQ[k][2]);
        b->q[k].T = reinterpret_cast<float4*>(m + oQ[k][3]);
    }
    b->prim = reinterpret_cast<int*>(m + oPrim);
    b->t = reinterpret_cast<float*>(m + oT);
    b->alive = reinterpret_cast<int*>(m + oAlive);
    b->s_o = reinterpret_cast<float4*>(m + oSo);
    b->s_d = reinterpret_cast<float4*>(m + oSd);
    b->s_c = reinterpret_cast<float4*>(m + oSc);
    b->s_tag = reinterpret_cast<int2*>(m + oStag);
    b->cur = 0;
    b->active_dirty = true;
    b->st = mt_batch_stats{};
    b->policy = MT_BATCH_OLDEST_FIRST;
    b->iters_per_step = 1;
    b->use_graphs = true;
    b->trace_mode = MT_BATCH_TRACE_SPLIT; // the serving autotuner's choice on both multi-job workloads
    for (int p = 0; p < 2; p++) {
        b->graph[p] = nullptr;
        b->gexec[p] = nullptr;
        b->g_iters[p] = 0;
    }
    b->gscratch = nullptr;
    b->gscratch_bytes = 0;
    for (int k = max_jobs - 1; k >= 0; k--) b->free_slots.push_back(k);
    b->free_blocks[0] = b->arena_bytes;
    k_zero_counts<<<1, 64, 0, ctx->stream>>>(b->cnt, b->tot);
    if (hipGetLastError() != hipSuccess) {
        mt_batch_destroy(b);
        mt_set_error("mt_batch_create: launch failed");
        return MT_ERR_HIP;
    }
    *out = b;
    return MT_OK;
}

extern "C" int32_t mt_batch_destroy(mt_batch* b)
{
    if (!b) return MT_OK;
    MtFirstError fe;
    fe(hipStreamSynchronize(b->ctx->stream));
    fe(hipHostFree(b->h_active));
    fe(hipHostFree(b->h_done));
    fe(hipHostFree(b->h_tot));
    drop_graphs(b);
    if (b->gscratch) fe(hipFree(b->gscratch));
    fe(hipFree(b->arena));
    fe(hipFree(b->mem));
    delete b;
    return fe.status("mt_batch_destroy");
}

extern "C" int32_t mt_batch_submit(mt_batch* b, const mt_job_desc* job)
{
    MT_ARG_CHECK(b, "mt_batch_submit: NULL batch");
    int32_t rc = validate_job(job);
    if (rc) return rc;
    mt_batch::Job J;
    J.d = *job;
    J.slot = -1;
    J.npaths = job->render.width * job->render.height * job->render.spp;
    J.len = job_bytes(job->render);
    J.off = 0;
    J.cancelled = false;
    if (J.len > b->arena_bytes) {
        mt_set_error("mt_batch_submit: job needs %llu arena bytes, the arena holds %llu",
                     (unsigned long long)J.len, (unsigned long long)b->arena_bytes);
        return MT_ERR_CAPACITY;
    }
    b->pending.push_back(J);
    b->st.jobs_submitted++;
    return MT_OK;
}

extern "C" int32_t mt_batch_step(mt_batch* b, int32_t max_done, int64_t* done_tags, int32_t* n_done, int32_t* busy)
{
    MT_ARG_CHECK(b && n_done && busy && max_done >= 0 && (max_done == 0 || done_tags),
                 "mt_batch_step: bad arguments");
    hipStream_t s = b->ctx->stream;
    int32_t rc = admit(b);
    if (rc) return rc;
    if (!b->active_jobs.empty()) {
        // several iterations back to back: finished jobs are detected on the
        // device (k_done_sum's flags) and read back once per step
        rc = run_iterations(b);
        if (rc) return rc;
        MT_HIP_CHECK(hipMemcpyAsync(b->h_done, b->fin, 4 * (size_t)b->max_jobs, hipMemcpyDeviceToHost, s));
        MT_HIP_CHECK(hipMemcpyAsync(b->h_tot, b->tot, 16, hipMemcpyDeviceToHost, s));
        MT_HIP_CHECK(hipStreamSynchronize(s));
        b->st.extension_rays = (int64_t)b->h_tot[0];
        b->st.shadow_rays = (int64_t)b->h_tot[1];
        bool any = false;
        for (size_t i = 0; i < b->active_jobs.size();) {
            mt_batch::Job& J = b->active_jobs[i];
            if (!b->h_done[J.slot]) {
                i++;
                continue;
            }
            const mt_render_desc& r = J.d.render;
            const int npix = r.width * r.height;
            if (!J.cancelled) {
                k_resolve<<<mt_blocks(npix, resolve_pixels_per_group(r.spp)), BT, 0, s>>>(
                    reinterpret_cast<const float4*>(b->arena + J.off), r.width, r.height, r.spp,
                    reinterpret_cast<float4*>(J.d.image));
                MT_LAUNCH_CHECK();
                b->st.jobs_completed++;
            }
            // stream order keeps the freed slot and arena range untouched
            // until this resolve has read them
            arena_give(b, J.off, J.len);
            b->st.arena_used -= J.len;
            b->free_slots.push_back(J.slot);
            b->finished.push_back(J.d.tag);
            b->active_jobs.erase(b->active_jobs.begin() + (ptrdiff_t)i);
            b->active_dirty = true;
            any = true;
        }
        if (any) MT_HIP_CHECK(hipStreamSynchronize(s));
        rc = mt_scan_check(b->ctx);
        if (rc == MT_OK && b->gscratch) { // the captured compactions' workspace
            mt_context* c = b->ctx;
            void* sv = c->scratch;
            const size_t svb = c->scratch_bytes;
            c->scratch = b->gscratch;
            c->scratch_bytes = b->gscratch_bytes;
            rc = mt_scan_check(c);
            c->scratch = sv;
            c->scratch_bytes = svb;
        }
        if (rc) return rc;
    }
    int nd = 0;
    while (nd < max_done && !b->finished.empty()) {
        done_tags[nd++] = b->finished.front();
        b->finished.pop_front();
    }
    *n_done = nd;
    *busy = (!b->active_jobs.empty() || !b->pending.empty() || !b->finished.empty()) ? 1 : 0;
    return MT_OK;
}

extern "C" int32_t mt_batch_cancel(mt_batch* b, int64_t tag, int32_t* found)
{
    MT_ARG_CHECK(b && found, "mt_batch_cancel: NULL argument");
    *found = 0;
    for (auto it = b->pending.begin(); it != b->pending.end(); ++it) {
        if (it->d.tag != tag) continue;
        b->pending.erase(it);
        b->finished.push_back(tag); // never admitted: nothing to drain
        b->st.jobs_cancelled++;
        *found = 1;
        return MT_OK;
    }
    for (mt_batch::Job& J : b->active_jobs) {
        if (J.d.tag != tag || J.cancelled) continue;
        k_cancel_job<<<1, 64, 0, b->ctx->stream>>>(b->npaths, b->issued, b->cancel, J.slot);
        MT_LAUNCH_CHECK();
        J.cancelled = true;
        b->st.jobs_cancelled++;
        *found = 1;
        return MT_OK;
    }
    return MT_OK;
}

extern "C" int32_t mt_batch_set_iterations(mt_batch* b, int32_t iterations)
{
    MT_ARG_CHECK(b, "mt_batch_set_iterations: NULL batch");
    MT_ARG_CHECK(iterations >= 1 && iterations <= 1024, "mt_batch_set_iterations: %d not in [1,1024]", iterations);
    b->iters_per_step = iterations;
    return MT_OK;
}

extern "C" int32_t mt_batch_set_trace_mode(mt_batch* b, int32_t mode)
{
    MT_ARG_CHECK(b, "mt_batch_set_trace_mode: NULL batch");
    MT_ARG_CHECK(mode == MT_TRACE_PACKED || mode == MT_TRACE_PACKED_STACK || mode == MT_TRACE_PACKED_POOL ||
                     mode == MT_BATCH_TRACE_SPLIT,
                 "mt_batch_set_trace_mode: mode %d not supported by the batch engine", mode);
    MT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream)); // recorded graphs hold the old kernel
    b->trace_mode = mode == MT_TRACE_PACKED ? MT_BATCH_TRACE_SPLIT : mode;
    drop_graphs(b);
    return MT_OK;
}

extern "C" int32_t mt_batch_set_graphs(mt_batch* b, int32_t on)
{
    MT_ARG_CHECK(b, "mt_batch_set_graphs: NULL batch");
    MT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream)); // a graph being dropped may still run
    b->use_graphs = on != 0;
    if (!b->use_graphs) drop_graphs(b);
    return MT_OK;
}

extern "C" int32_t mt_batch_set_policy(mt_batch* b, int32_t policy)
{
    MT_ARG_CHECK(b, "mt_batch_set_policy: NULL batch");
    MT_ARG_CHECK(policy == MT_BATCH_OLDEST_FIRST || policy == MT_BATCH_SMALLEST_FIRST,
                 "mt_batch_set_policy: unknown policy %d", policy);
    b->policy = policy;
    b->active_dirty = true;
    return MT_OK;
}

extern "C" int32_t mt_batch_get_stats(mt_batch* b, mt_batch_stats* out)
{
    MT_ARG_CHECK(b && out, "mt_batch_get_stats: NULL argument");
    b->st.active_jobs = (int32_t)b->active_jobs.size();
    b->st.pending_jobs = (int32_t)b->pending.size();
    *out = b->st;
    return MT_OK;
}
