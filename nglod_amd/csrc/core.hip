// gfx950 kernels and C ABI for the dense NGLOD query path:
//   nglod_aabb          box entry                      (sol_nglod_kernel.cu:77-188)
//   nglod_grid_sample   one FeatureVolume          (OctreeSDF.py:46-57)
//   nglod_sdf_fwd       fused multi-LOD gather+decoder (OctreeSDF.py:94-155)
//   nglod_sdf_fwd_all   every level's prediction       (OctreeSDF.py:148-153)
//   nglod_normals_fd    central-difference normals     (diffutils.py:61-70)
//   nglod_trace_dense   persistent sphere tracer       (SphereTracer.py:41-132)
#include <stdio.h>
#include <stdlib.h>

#include <mutex>

#include "api_common.hpp"
#include "fp32_wave.hpp"
#include "trace_common.hpp"

namespace nglod {

static thread_local char g_last_error[512] = "";
void set_last_error(const char* msg) {
    strncpy(g_last_error, msg, sizeof(g_last_error) - 1);
    g_last_error[sizeof(g_last_error) - 1] = 0;
}

int device_cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

// ---------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t* p, int n, uint32_t v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void aabb_kernel(const float* __restrict__ o, const float* __restrict__ d,
                                                   int64_t n, float* __restrict__ x, float* __restrict__ t,
                                                   uint8_t* __restrict__ hit) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BoxEntry e = box_entry(o[3 * i], o[3 * i + 1], o[3 * i + 2], d[3 * i], d[3 * i + 1], d[3 * i + 2]);
    x[3 * i] = e.x;
    x[3 * i + 1] = e.y;
    x[3 * i + 2] = e.z;
    t[i] = e.t;
    hit[i] = e.hit ? 1 : 0;
}

// One feature volume: out[i][c] = trilinear sample (FeatureVolume.forward,
// sdf-net/lib/models/OctreeSDF.py:46-51).
__global__ __launch_bounds__(256) void grid_sample_kernel(const float* __restrict__ grid, int r,
                                                          const float* __restrict__ x, int64_t n,
                                                          float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int S = r + 1;
    const AxisW ax = axis_weights(x[3 * i], r);
    const AxisW ay = axis_weights(x[3 * i + 1], r);
    const AxisW az = axis_weights(x[3 * i + 2], r);
    const float* base = grid + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC;
    const size_t sx = kC, sy = (size_t)S * kC, sz = (size_t)S * S * kC;
    const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
    const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
    float f[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) f[c] = 0.0f;
    corner_fma(f, base, wxy00 * az.w0);
    corner_fma(f, base + sx, wxy10 * az.w0);
    corner_fma(f, base + sy, wxy01 * az.w0);
    corner_fma(f, base + sy + sx, wxy11 * az.w0);
    corner_fma(f, base + sz, wxy00 * az.w1);
    corner_fma(f, base + sz + sx, wxy10 * az.w1);
    corner_fma(f, base + sz + sy, wxy01 * az.w1);
    corner_fma(f, base + sz + sy + sx, wxy11 * az.w1);
    float4* o4 = reinterpret_cast<float4*>(out + i * kC);
#pragma unroll
    for (int v = 0; v < kC / 4; ++v) o4[v] = make_float4(f[4 * v], f[4 * v + 1], f[4 * v + 2], f[4 * v + 3]);
}

// Batch query; with `order`, thread i evaluates point order[i] (a spatially
// sorted permutation: neighbouring lanes gather neighbouring corners).
template <bool POSINV>
__global__ __launch_bounds__(256) void sdf_fp32_kernel(ModelArgs m, int lod, int dec, const float* __restrict__ x,
                                                       int64_t n, float* __restrict__ out,
                                                       const uint32_t* __restrict__ order) {
    __shared__ Fp32Lds lds;
    stage_fp32_lds(m, dec, &lds);
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < n;
    if (__ballot(valid) == 0) return;  // whole wave past the end (no barrier follows)
    const int64_t i = valid ? (order ? (int64_t)order[t] : t) : 0;
    const float px = valid ? x[3 * i] : 0.0f;
    const float py = valid ? x[3 * i + 1] : 0.0f;
    const float pz = valid ? x[3 * i + 2] : 0.0f;
    const float d = sdf_fp32_wave<POSINV>(m, lod, &lds, m.b1[dec][0], lane_id(), valid, px, py, pz);
    if (valid) out[i] = d;
}

int launch_sdf_fp32(const ModelArgs& m, int lod, int dec, const float* x, int64_t n, float* out,
                    const uint32_t* order, hipStream_t s) {
    const int64_t blocks = (n + 255) / 256;
    if (m.pos_inv)
        hipLaunchKernelGGL(sdf_fp32_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, s, m, lod, dec, x, n, out,
                           order);
    else
        hipLaunchKernelGGL(sdf_fp32_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, s, m, lod, dec, x, n, out,
                           order);
    return check_launch("sdf_fp32_kernel");
}

// Every level's prediction (the training forward): the running FP32 feature
// sum per lane, and after each level the FP32 MFMA decoder of that level on
// the wave's 64 queries (decode_fp32_wave, the trace's decoder), each level's
// decoder staged in LDS in turn.  Every thread of the block runs every level
// (no early exit: the staging barriers are block-wide).
template <bool POSINV>
__global__ __launch_bounds__(256) void sdf_all_fp32_kernel(ModelArgs m, int lod, const float* __restrict__ x,
                                                           int64_t n, float* __restrict__ out) {
    __shared__ Fp32Lds lds;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const uint32_t lane = lane_id();
    const float px = valid ? x[3 * i] : 0.0f, py = valid ? x[3 * i + 1] : 0.0f, pz = valid ? x[3 * i + 2] : 0.0f;
    float f[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) f[c] = 0.0f;
    int staged = -1;
    for (int l = 0; l <= lod; ++l) {
        const int dec = m.num_dec == 1 ? 0 : l;
        if (dec != staged) {  // block-uniform
            __syncthreads();
            stage_fp32_lds(m, dec, &lds);
            __syncthreads();
            staged = dec;
        }
        if (valid) {
            const int r = m.res[l];
            const int S = r + 1;
            const AxisW ax = axis_weights(px, r);
            const AxisW ay = axis_weights(py, r);
            const AxisW az = axis_weights(pz, r);
            const float* base = m.grid[l] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC;
            const size_t sx = kC, sy = (size_t)S * kC, sz = (size_t)S * S * kC;
            const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
            const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
            corner_fma(f, base, wxy00 * az.w0);
            corner_fma(f, base + sx, wxy10 * az.w0);
            corner_fma(f, base + sy, wxy01 * az.w0);
            corner_fma(f, base + sy + sx, wxy11 * az.w0);
            corner_fma(f, base + sz, wxy00 * az.w1);
            corner_fma(f, base + sz + sx, wxy10 * az.w1);
            corner_fma(f, base + sz + sy, wxy01 * az.w1);
            corner_fma(f, base + sz + sy + sx, wxy11 * az.w1);
        }
        const float d = decode_fp32_wave<POSINV>(f, &lds, m.b1[dec][0], lane, valid, px, py, pz);
        if (valid) out[(int64_t)l * n + i] = d;
    }
}

template <bool POSINV>
__global__ __launch_bounds__(256) void normals_fp32_kernel(ModelArgs m, int lod, int dec, const float* __restrict__ x,
                                                           int64_t n, float* __restrict__ nrm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    float g[3];
#pragma unroll 1
    for (int a = 0; a < 3; ++a) {
        float pp[3] = {p[0], p[1], p[2]}, pm[3] = {p[0], p[1], p[2]};
        pp[a] = p[a] + kFdH;
        pm[a] = p[a] - kFdH;
        const float fp = sdf_fp32<POSINV>(m, lod, dec, pp[0], pp[1], pp[2]);
        const float fm = sdf_fp32<POSINV>(m, lod, dec, pm[0], pm[1], pm[2]);
        g[a] = (fp - fm) / kFdTwoH;
    }
    normalize_grad(g[0], g[1], g[2], nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
}

template <bool POSINV>
__global__ __launch_bounds__(kMarchBlock) void trace_dense_fp32_kernel(TraceArgs a) {
    __shared__ Fp32Lds lds;
    __shared__ CoarseLds32 coarse;
    __shared__ __align__(16) float solo_sv[kMarchBlock / 64][kH];
    __shared__ SoloLevels solo_lv;
    stage_fp32_lds(a.m, a.dec, &lds);
    stage_coarse32(a.m, a.coarse, &coarse);
    stage_solo_levels(a.m, &solo_lv);
    __syncthreads();
    const float b1 = a.m.b1[a.dec][0];
    const uint32_t lane = lane_id();
    float* sv = solo_sv[threadIdx.x >> 6];
    RayFeed feed;
    feed.init(a.nrays, a.queue_stride);
    RayState s;
    s.ray = -1;
    s.phase = PH_IDLE;
    uint32_t nevals = 0;
    bool tail = false;
    for (;;) {
        feed.refill<0>(s, a, lane);
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        if (feed.empty() && __popcll(act) <= a.solo_lanes) {
            tail = true;
            break;
        }
        const bool busy = s.phase != PH_IDLE;
        const float d = sdf_fp32_wave<POSINV>(a.m, a.lod, &lds, b1, lane, busy, s.qx, s.qy, s.qz, a.coarse, &coarse);
        if (busy) {
            ++nevals;
            s.consume(a, d);
        }
    }
    // the wave's last few rays, one query at a time (fp32_solo_step)
    while (tail) {
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        const float d = fp32_solo_step<POSINV>(a, &lds, &coarse, a.coarse, &solo_lv, sv, b1, lane, act, s.qx, s.qy,
                                               s.qz);
        if (s.phase != PH_IDLE) {
            ++nevals;
            s.consume(a, d);
        }
    }
    if (a.evals) {
        unsigned long long tot = nevals;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0) atomicAdd(a.evals, tot);
    }
}

template <typename K>
static int persistent_grid(K kernel, int64_t n, int block) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
    const int64_t want = (n + block - 1) / block;
    const int64_t cap = (int64_t)device_cu_count() * per_cu;
    return (int)(want < cap ? (want < 1 ? 1 : want) : cap);
}


}  // namespace nglod

using namespace nglod;

extern "C" {

int nglod_abi_version(void) { return NGLOD_ABI_VERSION; }

const char* nglod_last_error(void) { return g_last_error; }

int nglod_aabb(const float* ray_o, const float* ray_d, int64_t n, float* x, float* t, uint8_t* hit, void* stream) {
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n == 0) return NGLOD_OK;
    if (!ray_o || !ray_d || !x || !t || !hit) return fail(NGLOD_ERR_ARG, "NULL pointer");
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(aabb_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, ray_o, ray_d, n, x, t, hit);
    return check_launch("aabb_kernel");
}

int nglod_grid_sample(const float* grid, int32_t res, int32_t feature_dim, const float* x, int64_t n, float* out,
                      void* stream) {
    if (feature_dim != kC) return fail(NGLOD_ERR_UNSUPPORTED, "kernels are built for feature_dim 32");
    if (res < 1 || res > 1024) return fail(NGLOD_ERR_ARG, "res out of range");
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n == 0) return NGLOD_OK;
    if (!grid || !x || !out || !aligned16(grid) || !aligned16(out)) return fail(NGLOD_ERR_ARG, "NULL or misaligned pointer");
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(grid_sample_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, grid, res, x, n, out);
    return check_launch("grid_sample_kernel");
}

int nglod_sdf_fwd(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, int32_t form, float* out,
                  void* stream) {
    ModelArgs m;
    int rc = make_model_args(model, lod, form, &m);
    if (rc) return rc;
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n == 0) return NGLOD_OK;
    if (!x || !out) return fail(NGLOD_ERR_ARG, "NULL pointer");
    const int dec = decoder_of(model, lod);
    if (form == NGLOD_MIXED) return launch_sdf_mixed(m, 0, dec, x, n, out, nullptr, (hipStream_t)stream);
    return launch_sdf_fp32(m, lod, dec, x, n, out, nullptr, (hipStream_t)stream);
}

int nglod_sdf_fwd_all(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, float* out, void* stream) {
    ModelArgs m;
    int rc = make_model_args(model, lod, NGLOD_FP32, &m);
    if (rc) return rc;
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n == 0) return NGLOD_OK;
    if (!x || !out) return fail(NGLOD_ERR_ARG, "NULL pointer");
    const int64_t blocks = (n + 255) / 256;
    if (m.pos_inv)
        hipLaunchKernelGGL(sdf_all_fp32_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, m, lod, x, n, out);
    else
        hipLaunchKernelGGL(sdf_all_fp32_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, m, lod, x, n, out);
    return check_launch("sdf_all_fp32_kernel");
}

int nglod_normals_fd(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, int32_t form, float* normal,
                     void* stream) {
    // Normals are FP32 in both forms (finite differences of h = 1/192 need
    // more than BF16 resolution); the mixed form takes them from the FP32
    // cumulative grid, as its trace's FP32 pass does (that pass decodes with
    // exact BF16 splits on the matrix cores: FP32-level, not the same bits).
    ModelArgs m;
    int rc = make_model_args(model, lod, form, &m);
    if (rc) return rc;
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n == 0) return NGLOD_OK;
    if (!x || !normal) return fail(NGLOD_ERR_ARG, "NULL pointer");
    const int dec = decoder_of(model, lod);
    const int64_t blocks = (n + 255) / 256;
    if (m.pos_inv)
        hipLaunchKernelGGL(normals_fp32_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, m,
                           gather_lod(form, lod), dec, x, n, normal);
    else
        hipLaunchKernelGGL(normals_fp32_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, m,
                           gather_lod(form, lod), dec, x, n, normal);
    return check_launch("normals_fp32_kernel");
}

// Workspace: kWsCounters bytes of counters (the ray queues, then the claim
// counters of pass 2 and of a pass 1 over returned rays, kQueues each, one
// 128-byte line per counter); for the mixed form then hand-off records [n], returned records [n], hand-off flags
// [n] and return flags [n] (each flag array padded to 256 bytes).
static constexpr int kCounterWords = 3 * kQueues * kQueueStride;
static constexpr size_t kWsCounters = (4 * kCounterWords + 255) / 256 * 256;

static size_t flag_bytes(int64_t n) { return ((size_t)n + 255) / 256 * 256; }


size_t nglod_trace_dense_workspace_size(int64_t n, int32_t form) {
    if (n < 0) n = 0;
    return kWsCounters + (form == NGLOD_MIXED ? (size_t)n * 2 * sizeof(Handoff) + 2 * flag_bytes(n) : 0);
}

}  // extern "C"

// Argument checks shared by the trace entry points (before any workspace).
static int check_trace_args(const nglod_model_t* model, const float* ray_o, const float* ray_d, int64_t n,
                            int32_t lod, const nglod_trace_params_t* params, int32_t form, float* x, float* t,
                            uint8_t* hit, float* normal, ModelArgs* m) {
    int rc = make_model_args(model, lod, form, m);
    if (rc) return rc;
    if (!params) return fail(NGLOD_ERR_ARG, "params is NULL");
    if (params->num_steps < 1) return fail(NGLOD_ERR_ARG, "num_steps must be >= 1");
    if (form == NGLOD_MIXED && (params->refine_steps < 1 || params->refine_steps > 4))
        return fail(NGLOD_ERR_ARG, "refine_steps must be 1..4");
    if (form == NGLOD_MIXED && !(params->near_dist >= 0.0)) return fail(NGLOD_ERR_ARG, "near_dist must be >= 0");
    if (n < 0 || n > (int64_t)0x7fff0000) return fail(NGLOD_ERR_ARG, "n out of range");
    if (n > 0 && (!ray_o || !ray_d || !x || !t || !hit || !normal)) return fail(NGLOD_ERR_ARG, "NULL pointer");
    return NGLOD_OK;
}

// Scratch of one trace call: kWsCounters bytes of claim counters and, for
// the mixed form, the hand-off records and flags [n] and (when near misses
// may go back to mixed precision) the hand-back records and flags [n].
struct TraceScratch {
    uint32_t* counters;
    Handoff* handoff;
    Handoff* ret;
    uint8_t* hflag;
    uint8_t* rflag;
};

// The trace on checked arguments (n > 0) and laid-out scratch.
static int trace_dense_run(const nglod_model_t* model, const ModelArgs& m, const float* ray_o, const float* ray_d,
                           int64_t n, int32_t lod, const nglod_trace_params_t* params, int32_t form, float* x,
                           float* t, uint8_t* hit, float* normal, const TraceScratch& w, uint64_t* evals_out,
                           int32_t* ray_steps, hipStream_t s) {
    int rc = NGLOD_OK;
    TraceArgs a;
    a.m = m;
    a.lod = gather_lod(form, lod);
    a.dec = decoder_of(model, lod);
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
    a.queue = w.counters;
    a.queue_stride = kQueueStride;
    a.handoff_queue = a.queue + kQueues * kQueueStride;
    a.ret_queue = a.queue + 2 * kQueues * kQueueStride;
    a.handoff = w.handoff;
    a.ret = w.ret;
    a.hflag = w.hflag;
    a.rflag = w.rflag;
    a.return_dist = (float)params->return_dist;
    // In-lane normal taps; the mixed passes set their own.  A separate
    // normals pass measured 3-4% slower for the FP32 form.
    a.defer_normals = 0;
    a.solo_lanes = solo_lanes_setting();
    a.coarse = coarse_plan(m, a.lod);
    a.replay_steps = params->refine_steps < 1 ? 1 : params->refine_steps;
    a.near_dist = (float)params->near_dist;
    a.evals = (unsigned long long*)evals_out;
    a.ray_steps = ray_steps;
    a.nrays = n;
    PassTimer timer(params->pass_ms, kPassSlots, s);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((kCounterWords + 63) / 64), dim3(64), 0, s, a.queue, kCounterWords, 0u);
    rc = check_launch("fill_u32_kernel");
    if (rc) return rc;
    timer.mark(kPassOther);
    if (form == NGLOD_MIXED) {
        rc = launch_trace_mixed(a, n, s, timer);
        return rc ? rc : timer.finish();
    }
    if (m.pos_inv) {
        const int grid = persistent_grid(trace_dense_fp32_kernel<true>, n, kMarchBlock);
        hipLaunchKernelGGL(trace_dense_fp32_kernel<true>, dim3(grid), dim3(kMarchBlock), 0, s, a);
    } else {
        const int grid = persistent_grid(trace_dense_fp32_kernel<false>, n, kMarchBlock);
        hipLaunchKernelGGL(trace_dense_fp32_kernel<false>, dim3(grid), dim3(kMarchBlock), 0, s, a);
    }
    rc = check_launch("trace_dense_fp32_kernel");
    if (rc) return rc;
    timer.mark(kPassMarch);
    return timer.finish();
}

extern "C" int nglod_trace_dense(const nglod_model_t* model, const float* ray_o, const float* ray_d, int64_t n,
                                 int32_t lod, const nglod_trace_params_t* params, int32_t form, float* x, float* t,
                                 uint8_t* hit, float* normal, void* workspace, size_t workspace_bytes,
                                 uint64_t* evals_out, int32_t* ray_steps, void* stream) {
    ModelArgs m;
    int rc = check_trace_args(model, ray_o, ray_d, n, lod, params, form, x, t, hit, normal, &m);
    if (rc) return rc;
    if (n == 0) return NGLOD_OK;
    if (!workspace || workspace_bytes < nglod_trace_dense_workspace_size(n, form))
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    if (!aligned16(workspace)) return fail(NGLOD_ERR_ARG, "workspace not 16-byte aligned");
    TraceScratch w{(uint32_t*)workspace, nullptr, nullptr, nullptr, nullptr};
    if (form == NGLOD_MIXED) {
        char* p = (char*)workspace + kWsCounters;
        w.handoff = (Handoff*)p;
        w.ret = (Handoff*)(p + (size_t)n * sizeof(Handoff));
        w.hflag = (uint8_t*)(p + 2 * (size_t)n * sizeof(Handoff));
        w.rflag = w.hflag + flag_bytes(n);
    }
    return trace_dense_run(model, m, ray_o, ray_d, n, lod, params, form, x, t, hit, normal, w, evals_out, ray_steps,
                           (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// nglod_trace_dense_noalloc: the same trace with its scratch in a file-scope
// device array instead of a caller workspace, for callers that cannot
// allocate (graph capture, tight serving loops).  Layout: claim counters,
// then hand-off records [kNoAllocMaxMixedRays], then hand-off flags.  The
// mixed form's hand-back records are not provided, so it requires
// return_dist = 0 (one round).  Every call on a device shares that device's
// array: calls that overlap on one device race.
static constexpr int64_t kNoAllocMaxMixedRays = NGLOD_NOALLOC_MAX_MIXED_RAYS;
static constexpr size_t kNoAllocHandoffBytes = (size_t)kNoAllocMaxMixedRays * sizeof(Handoff);
static constexpr size_t kNoAllocBytes = kWsCounters + kNoAllocHandoffBytes + (size_t)kNoAllocMaxMixedRays;
static_assert(kWsCounters % 256 == 0 && kNoAllocHandoffBytes % 256 == 0, "scratch sections stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on each device, looked up once per device with
// that device current (each device holds its own copy of the array).
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[64] = {nullptr};

static int scratch_on_current_device(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return fail(NGLOD_ERR_HIP, "nglod_trace_dense_noalloc: no current device (or device index >= 64)");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            char buf[256];
            snprintf(buf, sizeof(buf), "nglod_trace_dense_noalloc: hipGetSymbolAddress(scratch): %s",
                     hipGetErrorString(e));
            return fail(NGLOD_ERR_HIP, buf);
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return NGLOD_OK;
}

extern "C" int nglod_trace_dense_noalloc(const nglod_model_t* model, const float* ray_o, const float* ray_d,
                                         int64_t n, int32_t lod, const nglod_trace_params_t* params, int32_t form,
                                         float* x, float* t, uint8_t* hit, float* normal, uint64_t* evals_out,
                                         int32_t* ray_steps, void* stream) {
    ModelArgs m;
    int rc = check_trace_args(model, ray_o, ray_d, n, lod, params, form, x, t, hit, normal, &m);
    if (rc) return rc;
    if (form == NGLOD_MIXED && n > kNoAllocMaxMixedRays)
        return fail(NGLOD_ERR_CAPACITY, "nglod_trace_dense_noalloc: the mixed form takes at most "
                                        "NGLOD_NOALLOC_MAX_MIXED_RAYS rays per call");
    if (form == NGLOD_MIXED && params->return_dist > 0.0)
        return fail(NGLOD_ERR_UNSUPPORTED, "nglod_trace_dense_noalloc: the mixed form needs return_dist = 0");
    if (n == 0) return NGLOD_OK;
    void* base = nullptr;
    rc = scratch_on_current_device(&base);
    if (rc) return rc;
    unsigned char* p = (unsigned char*)base;
    TraceScratch w{(uint32_t*)p, nullptr, nullptr, nullptr, nullptr};
    if (form == NGLOD_MIXED) {
        w.handoff = (Handoff*)(p + kWsCounters);
        w.hflag = p + kWsCounters + kNoAllocHandoffBytes;
    }
    return trace_dense_run(model, m, ray_o, ray_d, n, lod, params, form, x, t, hit, normal, w, evals_out, ray_steps,
                           (hipStream_t)stream);
}
