// Pooled trace: camera requests of several models and LODs traced together.
//   nglod_model_table_bytes / nglod_model_table_fill   device table of models
//   nglod_trace_pool_workspace_size / nglod_trace_pool  one launch per pass
// The reference has no serving layer; the nearest analogue is the batch loop
// of Renderer.render (sdf-net/lib/renderer.py:118-123), which traces one
// model's rays at a time.
#include <stdlib.h>

#include <algorithm>

#include "api_common.hpp"
#include "trace_common.hpp"

using namespace nglod;

namespace {

size_t pool_flag_bytes(int64_t n) { return ((size_t)n + 255) / 256 * 256; }

}  // namespace

extern "C" {

// The table holds one kernel argument block per (model, lod): entry
// model * NGLOD_MAX_LODS + lod is make_model_args(model, lod, NGLOD_MIXED),
// the cumulative grid of that lod as level 0 (unused lods stay zero).
size_t nglod_model_table_bytes(int32_t n) { return n > 0 ? (size_t)n * NGLOD_MAX_LODS * sizeof(ModelArgs) : 0; }

int nglod_model_table_fill(const nglod_model_t* models, int32_t n, void* table, size_t table_bytes, void* stream) {
    if (!models || !table || n < 1) return fail(NGLOD_ERR_ARG, "models / table NULL or n < 1");
    if (table_bytes < nglod_model_table_bytes(n)) return fail(NGLOD_ERR_CAPACITY, "table smaller than n models");
    const size_t entries = (size_t)n * NGLOD_MAX_LODS;
    ModelArgs* host = (ModelArgs*)calloc(entries, sizeof(ModelArgs));
    if (!host) return fail(NGLOD_ERR_ARG, "host allocation failed");
    for (int i = 0; i < n; ++i) {
        if (models[i].num_lods < 1 || models[i].num_lods > NGLOD_MAX_LODS) {
            free(host);
            return fail(NGLOD_ERR_ARG, "num_lods out of range");
        }
        for (int l = 0; l < models[i].num_lods; ++l) {
            const int rc = make_model_args(&models[i], l, NGLOD_MIXED, &host[(size_t)i * NGLOD_MAX_LODS + l]);
            if (rc) {
                free(host);
                return rc;
            }
        }
    }
    hipStream_t s = (hipStream_t)stream;
    const bool ok = hipMemcpyAsync(table, host, entries * sizeof(ModelArgs), hipMemcpyHostToDevice, s) == hipSuccess &&
                    hipStreamSynchronize(s) == hipSuccess;
    free(host);
    return ok ? NGLOD_OK : fail(NGLOD_ERR_HIP, "model table copy failed");
}

size_t nglod_trace_pool_workspace_size(int64_t pool_rays) {
    if (pool_rays < 0) pool_rays = 0;
    return (size_t)pool_rays * sizeof(Handoff) + pool_flag_bytes(pool_rays);
}

int nglod_trace_pool(const nglod_model_t* models, const void* model_table, int32_t num_models,
                     const nglod_pool_request_t* requests, int32_t num_requests, const nglod_trace_params_t* params,
                     const float* ray_o, const float* ray_d, float* x, float* t, uint8_t* hit, float* normal,
                     int64_t pool_rays, void* workspace, size_t workspace_bytes, void* counters,
                     uint64_t* evals_out, void* stream) {
    if (!models || !model_table || num_models < 1) return fail(NGLOD_ERR_ARG, "models / model table missing");
    if (!requests || num_requests < 0 || num_requests > kPoolMaxRequests)
        return fail(NGLOD_ERR_ARG, "num_requests must be 0..64");
    if (!params || params->num_steps < 1) return fail(NGLOD_ERR_ARG, "params NULL or num_steps < 1");
    if (params->refine_steps < 1 || params->refine_steps > 4) return fail(NGLOD_ERR_ARG, "refine_steps must be 1..4");
    if (!(params->near_dist >= 0.0)) return fail(NGLOD_ERR_ARG, "near_dist must be >= 0");
    if (params->return_dist > 0.0) return fail(NGLOD_ERR_UNSUPPORTED, "the pooled trace runs one round: return_dist 0");
    if (pool_rays < 0 || pool_rays > (int64_t)0x7fff0000) return fail(NGLOD_ERR_ARG, "pool_rays out of range");
    PoolArgs p;
    memset(&p, 0, sizeof(p));
    p.models = (const ModelArgs*)model_table;
    int pos_inv = -1;
    struct Span {
        int64_t off, end;
    } spans[kPoolMaxRequests];
    int64_t total = 0;
    int nreq = 0;
    for (int i = 0; i < num_requests; ++i) {
        const nglod_pool_request_t& q = requests[i];
        if (q.model < 0 || q.model >= num_models) return fail(NGLOD_ERR_ARG, "request: model index out of range");
        const nglod_model_t& m = models[q.model];
        if (q.lod < 0 || q.lod >= m.num_lods) return fail(NGLOD_ERR_ARG, "request: lod out of range");
        if (q.offset < 0 || q.count < 0 || q.offset + q.count > pool_rays)
            return fail(NGLOD_ERR_ARG, "request: rays outside the pool");
        const int pi = m.pos_invariant ? 1 : 0;
        if (pos_inv >= 0 && pi != pos_inv) return fail(NGLOD_ERR_UNSUPPORTED, "requests mix pos_invariant models");
        pos_inv = pi;
        if (q.count == 0) continue;
        spans[nreq] = {q.offset, q.offset + q.count};
        PoolReq& r = p.req[nreq++];
        r.model = q.model * NGLOD_MAX_LODS + q.lod;  // the (model, lod) table entry
        r.lod = gather_lod(NGLOD_MIXED, q.lod);
        r.dec = decoder_of(&m, q.lod);
        r.vbeg = total;
        r.off = q.offset;
        total += q.count;
    }
    std::sort(spans, spans + nreq, [](const Span& a, const Span& b) { return a.off < b.off; });
    for (int i = 1; i < nreq; ++i)
        if (spans[i].off < spans[i - 1].end) return fail(NGLOD_ERR_ARG, "requests' pool ranges overlap");
    if (nreq == 0) return NGLOD_OK;
    if (!ray_o || !ray_d || !x || !t || !hit || !normal) return fail(NGLOD_ERR_ARG, "NULL pointer");
    if (!workspace || workspace_bytes < nglod_trace_pool_workspace_size(pool_rays) || !aligned16(workspace))
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small or not 16-byte aligned");
    if (!counters || ((uintptr_t)counters & 3u)) return fail(NGLOD_ERR_ARG, "counters NULL or not 4-byte aligned");
    p.total = total;
    p.nreq = nreq;
    TraceArgs a;
    memset(&a, 0, sizeof(a));
    a.num_steps = params->num_steps;
    a.normals = params->normals ? 1 : 0;
    a.min_dis = (float)params->min_dis;
    a.min_dis3 = (float)(params->min_dis * 3.0);
    a.step = (float)params->step_size;
    a.far_clip = (float)params->far_clip;
    a.ray_o = ray_o;
    a.ray_d = ray_d;
    a.x = x;
    a.t = t;
    a.hit = hit;
    a.normal = normal;
    static_assert(3 * kQueues * kQueueStride * sizeof(uint32_t) <= NGLOD_POOL_COUNTER_BYTES,
                  "the pool's claim counters fit NGLOD_POOL_COUNTER_BYTES");
    a.queue = (uint32_t*)counters;
    a.queue_stride = kQueueStride;
    a.handoff_queue = a.queue + kQueues * kQueueStride;
    a.ret_queue = a.queue + 2 * kQueues * kQueueStride;
    a.handoff = (Handoff*)workspace;
    a.hflag = (uint8_t*)workspace + (size_t)pool_rays * sizeof(Handoff);
    a.ret = nullptr;
    a.rflag = nullptr;
    a.return_dist = 0.0f;
    a.defer_normals = 0;  // the pooled passes set their own
    a.solo_lanes = solo_lanes_setting();
    a.coarse = CoarsePlan{0, 0};
    a.replay_steps = params->refine_steps;
    a.near_dist = (float)params->near_dist;
    a.evals = (unsigned long long*)evals_out;
    a.ray_steps = nullptr;
    a.nrays = pool_rays;  // bound of the record indices (pool slots)
    hipStream_t s = (hipStream_t)stream;
    PassTimer timer(params->pass_ms, kPassSlots, s);
    const int rc = launch_trace_pool(a, p, pos_inv == 1, s, timer);
    return rc ? rc : timer.finish();
}

}  // extern "C"
