// Mixed-precision form: FP16 grids, BF16 MFMA first decoder layer with FP32
// accumulation, and FP32 re-evaluation near the surface.
//   nglod_pack_mixed        FP32 model -> FP16 grids + BF16 W0 fragments
//   launch_sdf_mixed        nglod_sdf_fwd(form = NGLOD_MIXED)
//   launch_trace_mixed      nglod_trace_dense(form = NGLOD_MIXED):
//     pass 1  persistent mixed march.  When a ray stops on a surface test
//             (|d| <= min_dis or the oscillation test) it rewinds its last
//             near-surface evaluations (at most replay_steps, back to the
//             last step of at least near_dist) and is appended to a hand-off list
//             (ballot + one atomic per wave); finished hits go to the
//             normal list.
//     pass 2  persistent FP32 kernel over the hand-off list: re-evaluates the
//             rewound steps and goes on in FP32 to its FP32 stop; a ray whose
//             FP32 step grows past return_dist (the mixed stop was a near
//             miss) is handed back and pass 1 resumes it.  Passes 1 and 2 run
//             kMixedRounds times (once when return_dist = 0); the last
//             pass 2 keeps every ray.
//     pass 3  FP32 normals of every finished hit, one hit per lane.
#include "api_common.hpp"
#include "fp32_wave.hpp"
#include "mixed_device.hpp"
#include "trace_common.hpp"

namespace nglod {

// Passes 1 and 2 read every level from global memory: with LDS copies of the
// coarse levels (FP16 or FP32 layout) they measured slower (pass 1, by 10-20%)
// or the same (pass 2) on the 1280x720 LOD-4 frame.
constexpr int kMixedBlock = 256;
constexpr int kPass2Block = 256;

// ---------------------------------------------------------------------------
__global__ void pack_grid_kernel(const float* __restrict__ src, __half* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = __float2half_rn(src[i]);
}

// Fragment (h, kb, lane, e) = W0[32h + lane%32][off + 16kb + 8*(lane/32) + e].
__global__ void pack_w0_kernel(const float* __restrict__ w0, int din, int off, uint16_t* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0 .. 4*2*64*8-1
    if (i >= 4 * 2 * 64 * 8) return;
    const int e = i & 7;
    const int lane = (i >> 3) & 63;
    const int kb = (i >> 9) & 1;
    const int h = i >> 10;
    const int row = 32 * h + (lane & 31);
    const int col = off + 16 * kb + 8 * (lane >> 5) + e;
    const __bf16 b = (__bf16)w0[row * din + col];
    dst[i] = __builtin_bit_cast(uint16_t, b);
}

// ---------------------------------------------------------------------------
template <bool POSINV>
__global__ __launch_bounds__(256) void sdf_mixed_kernel(ModelArgs m, int lod, int dec, const float* __restrict__ x,
                                                        int64_t n, float* __restrict__ out) {
    __shared__ MixedLds lds;
    stage_mixed_lds(m, dec, &lds);
    __syncthreads();
    const uint32_t lane = lane_id();
    const float b1 = m.b1[dec][0];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const float px = valid ? x[3 * i] : 0.0f;
    const float py = valid ? x[3 * i + 1] : 0.0f;
    const float pz = valid ? x[3 * i + 2] : 0.0f;
    if (__ballot(valid) == 0) return;  // whole wave past the end (no barrier follows)
    const float d = sdf_mixed_wave<POSINV>(m, lod, &lds, b1, lane, valid, px, py, pz);
    if (valid) out[i] = d;
}

// ---------------------------------------------------------------------------
// Wave-aggregated append: lanes with `want` get consecutive slots of a list
// whose fill counter is `count` (one atomic per wave).
__device__ __forceinline__ uint32_t append_slot(uint32_t* count, uint32_t lane, bool want, bool& any) {
    const uint64_t mask = __ballot(want);
    any = mask != 0;
    if (!any) return 0;
    const int cnt = __popcll(mask);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(count, (uint32_t)cnt);
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    return base + mask_rank(mask);
}

// Moves lanes that need FP32 work (`refine`) to the hand-off list and
// finished hits (phase PH_NQ) to the normal list.
__device__ __forceinline__ void flush_lists(RayState& s, const TraceArgs& a, uint32_t lane, bool refine) {
    bool any;
    const uint32_t hs = append_slot(a.handoff_count, lane, refine, any);
    if (refine) s.hand_off(a, hs);
    const bool nq = s.phase == PH_NQ;
    const uint32_t ns = append_slot(a.ncount, lane, nq, any);
    if (nq) s.queue_normal(a, ns);
}

template <bool POSINV>
__global__ __launch_bounds__(kMixedBlock) void trace_mixed_pass1_kernel(TraceArgs a) {
    __shared__ MixedLds lds;
    stage_mixed_lds(a.m, a.dec, &lds);
    __syncthreads();
    const uint32_t lane = lane_id();
    const float b1 = a.m.b1[a.dec][0];
    int q = blockIdx.x % kQueues;
    int tried = 0;
    bool drained = false;
    RayState s;
    s.ray = -1;
    s.phase = PH_IDLE;
    uint32_t nevals = 0;
    float ht0 = 0, ht1 = 0, ht2 = 0, ht3 = 0, hd0 = 0, hd1 = 0, hd2 = 0, hd3 = 0;
    int hist_n = 0;
    bool hist_init = false;
    bool replay = false;
    const uint32_t ret_total = a.from_returns ? *a.ret_count : 0u;
    for (;;) {
        for (;;) {
            const bool need = (s.phase == PH_IDLE) && !drained;
            const uint64_t mask = __ballot(need);
            if (mask == 0) break;
            const int cnt = __popcll(mask);
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if (a.from_returns) {
                if ((int)lane == leader) base = atomicAdd(a.ret_queue, (uint32_t)cnt);
                base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
                if (need) {
                    const uint32_t idx = base + mask_rank(mask);
                    if (idx < ret_total) {
                        s.resume(a, a.ret[idx]);
                        hist_n = 0;
                        hist_init = false;
                    }
                }
                if (base + (uint32_t)cnt >= ret_total) drained = true;
            } else {
                if ((int)lane == leader) base = atomicAdd(&a.queue[q], (uint32_t)cnt);
                base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
                const int64_t qsize = a.qend[q] - a.qbeg[q];
                if (need) {
                    const int64_t idx = (int64_t)base + (int64_t)mask_rank(mask);
                    if (idx < qsize) s.start(a, a.qbeg[q] + idx);
                }
                if ((int64_t)base + cnt >= qsize) {
                    q = (q + 1) % kQueues;
                    if (++tried == kQueues) drained = true;
                }
            }
            // origins inside the box finish at once as hits
            flush_lists(s, a, lane, false);
        }
        if (__ballot(s.phase != PH_IDLE) == 0) break;
        const bool busy = s.phase != PH_IDLE;
        const float d = sdf_mixed_wave<POSINV>(a.m, a.lod, &lds, b1, lane, busy, s.qx, s.qy, s.qz);
        if (busy) {
            ++nevals;
            // pre-evaluation state of the last kReplay evaluations (shift register)
            if (s.phase == PH_INIT) {
                hist_n = 0;
                hist_init = true;
            }
            ht3 = ht2; ht2 = ht1; ht1 = ht0; ht0 = s.t;
            hd3 = hd2; hd2 = hd1; hd1 = hd0; hd0 = s.D;
            ++hist_n;
            s.consume(a, d);
            if (s.phase == PH_REPLAY) {
                // Rewind j evaluations (at most replay_steps, back to the
                // box-entry evaluation if the ray made fewer steps): the
                // stopping one and those before it back to (not including)
                // the last step of at least near_dist.  A step that long was
                // taken away from the surface, where the mixed-precision
                // error cannot carry the ray past it.
                const int jmax = hist_n < a.replay_steps ? hist_n : a.replay_steps;
                int j = 1;
                if (a.near_dist <= 0.0f) {
                    j = jmax;
                } else {
                    if (j < jmax && fabsf(hd0) < a.near_dist) ++j;
                    if (j == 2 && j < jmax && fabsf(hd1) < a.near_dist) ++j;
                    if (j == 3 && j < jmax && fabsf(hd2) < a.near_dist) ++j;
                }
                const float tj = j == 1 ? ht0 : j == 2 ? ht1 : j == 3 ? ht2 : ht3;
                const float dj = j == 1 ? hd0 : j == 2 ? hd1 : j == 3 ? hd2 : hd3;
                if (hist_init && j == hist_n) {
                    const BoxEntry e = box_entry(s.ox, s.oy, s.oz, s.dx, s.dy, s.dz);
                    s.phase = PH_INIT;
                    s.x0 = e.x;
                    s.x1 = e.y;
                    s.x2 = e.z;
                    s.t = e.t;
                    s.it = 0;
                    s.steps = 0;
                } else {
                    s.phase = PH_MARCH;
                    s.t = tj;
                    s.D = dj;
                    s.it -= j;
                    s.steps -= j;
                    s.x0 = s.ox + s.dx * tj;
                    s.x1 = s.oy + s.dy * tj;
                    s.x2 = s.oz + s.dz * tj;
                }
                s.qx = s.x0;
                s.qy = s.x1;
                s.qz = s.x2;
                replay = true;
            }
        }
        flush_lists(s, a, lane, replay);
        replay = false;
    }
    if (a.evals) {
        unsigned long long tot = nevals;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0) atomicAdd(a.evals + 1, tot);
    }
}

template <bool POSINV>
__global__ __launch_bounds__(kPass2Block) void trace_mixed_pass2_kernel(TraceArgs a) {
    __shared__ Fp32Lds lds;
    stage_fp32_lds(a.m, a.dec, &lds);
    __syncthreads();
    const float b1 = a.m.b1[a.dec][0];
    const uint32_t lane = lane_id();
    const uint32_t total = *a.handoff_count;
    bool drained = false;
    RayState s;
    s.ray = -1;
    s.phase = PH_IDLE;
    uint32_t nevals = 0;
    for (;;) {
        {
            const bool need = (s.phase == PH_IDLE) && !drained;
            const uint64_t mask = __ballot(need);
            if (mask != 0) {
                const int cnt = __popcll(mask);
                const int leader = __ffsll((unsigned long long)mask) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(a.handoff_queue, (uint32_t)cnt);
                base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
                if (need) {
                    const uint32_t idx = base + mask_rank(mask);
                    if (idx < total) s.resume(a, a.handoff[idx]);
                }
                if (base + (uint32_t)cnt >= total) drained = true;
            }
        }
        if (__ballot(s.phase != PH_IDLE) == 0) break;
        const bool busy = s.phase != PH_IDLE;
        const float d = sdf_fp32_wave<POSINV>(a.m, a.lod, &lds, b1, lane, busy, s.qx, s.qy, s.qz);
        if (busy) {
            ++nevals;
            s.consume(a, d);
        }
        bool any;
        // hand back only a ray that moves away from the surface (step growing)
        const bool back = a.allow_return && s.phase == PH_MARCH && fabsf(s.D) > a.return_dist &&
                          fabsf(s.D) > fabsf(s.Dp);
        const uint32_t rs = append_slot(a.ret_count, lane, back, any);
        if (back) s.hand_off_to(a.ret, rs);
        const bool nq = s.phase == PH_NQ;
        const uint32_t ns = append_slot(a.ncount, lane, nq, any);
        if (nq) s.queue_normal(a, ns);
    }
    if (a.evals) {
        unsigned long long tot = nevals;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0) atomicAdd(a.evals, tot);
    }
}

// Pass 3: FP32 central-difference normals of the finished hits, one hit per
// lane (six evaluations each, uniform work), persistent over the list.
template <bool POSINV>
__global__ __launch_bounds__(kMarchBlock) void normals_list_kernel(TraceArgs a) {
    const uint32_t total = *a.ncount;
    if ((int64_t)blockIdx.x * blockDim.x >= (int64_t)total) return;  // block-uniform, before the barrier
    __shared__ Fp32Lds lds;
    __shared__ CoarseLds32 coarse;
    stage_fp32_lds(a.m, a.dec, &lds);
    stage_coarse32(a.m, a.coarse, &coarse);
    __syncthreads();
    const float b1 = a.m.b1[a.dec][0];
    const uint32_t lane = lane_id();
    uint32_t nevals = 0;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < (int64_t)total; i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        const bool valid = i < total;
        if (__ballot(valid) == 0) break;  // whole wave past the end (no barrier follows)
        NormalItem it = {0, 0.0f, 0.0f, 0.0f};
        if (valid) it = a.nlist[i];
        const float p[3] = {it.x0, it.x1, it.x2};
        float g[3];
#pragma unroll 1
        for (int ax = 0; ax < 3; ++ax) {
            float pp[3] = {p[0] + 0.0f, p[1] + 0.0f, p[2] + 0.0f}, pm[3] = {p[0] - 0.0f, p[1] - 0.0f, p[2] - 0.0f};
            pp[ax] = p[ax] + kFdH;
            pm[ax] = p[ax] - kFdH;
            const float fp =
                sdf_fp32_wave<POSINV>(a.m, a.lod, &lds, b1, lane, valid, pp[0], pp[1], pp[2], a.coarse, &coarse);
            const float fm =
                sdf_fp32_wave<POSINV>(a.m, a.lod, &lds, b1, lane, valid, pm[0], pm[1], pm[2], a.coarse, &coarse);
            g[ax] = (fp - fm) / kFdTwoH;
        }
        if (valid) {
            nevals += 6;
            normalize_grad(g[0], g[1], g[2], a.normal[3 * it.ray], a.normal[3 * it.ray + 1],
                           a.normal[3 * it.ray + 2]);
        }
    }
    if (a.evals) {
        unsigned long long tot = nevals;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if ((threadIdx.x & 63) == 0 && tot) atomicAdd(a.evals, tot);
    }
}

template <typename K>
static int grid_for(K kernel, int64_t n, int block) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
    const int64_t want = (n + block - 1) / block;
    const int64_t cap = (int64_t)device_cu_count() * per_cu;
    return (int)(want < cap ? (want < 1 ? 1 : want) : cap);
}

int launch_sdf_mixed(const ModelArgs& m, int lod, int dec, const float* x, int64_t n, float* out, hipStream_t s) {
    const int64_t blocks = (n + 255) / 256;
    if (m.pos_inv)
        hipLaunchKernelGGL(sdf_mixed_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, s, m, lod, dec, x, n, out);
    else
        hipLaunchKernelGGL(sdf_mixed_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, s, m, lod, dec, x, n, out);
    return check_launch("sdf_mixed_kernel");
}

// Counter slots (uint32 in the workspace): 0..7 ray queues, 8 handoff_count,
// 9 handoff_queue, 10 ncount, 11 ret_count, 12 ret_queue.
__global__ void reset_counters_kernel(uint32_t* c, int which) {
    if (threadIdx.x != 0) return;
    if (which == 0) {  // before a pass 1 that reads the return list
        c[8] = 0;
        c[9] = 0;
        c[12] = 0;
    } else {  // before a pass 2
        c[11] = 0;
    }
}

// Rounds of (pass 1, pass 2) when near misses may go back to mixed precision
// (return_dist > 0); with return_dist = 0 one round, pass 2 keeps every ray.
constexpr int kMixedRounds = 3;

template <bool POSINV>
static int trace_mixed_rounds(TraceArgs a, int64_t n, hipStream_t s, PassTimer& timer) {
    const int g1 = grid_for(trace_mixed_pass1_kernel<POSINV>, n, kMixedBlock);
    const int g2 = grid_for(trace_mixed_pass2_kernel<POSINV>, n, kPass2Block);
    const int g3 = grid_for(normals_list_kernel<POSINV>, n, kMarchBlock);
    const int rounds = a.return_dist > 0.0f ? kMixedRounds : 1;
    for (int r = 0; r < rounds; ++r) {
        a.from_returns = r > 0;
        if (r > 0) {
            hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, a.queue, 0);
            int rc = check_launch("reset_counters_kernel");
            if (rc) return rc;
            timer.mark(kPassOther);
        }
        a.replay = 1;
        hipLaunchKernelGGL(trace_mixed_pass1_kernel<POSINV>, dim3(g1), dim3(kMixedBlock), 0, s, a);
        a.replay = 0;
        int rc = check_launch("trace_mixed_pass1_kernel");
        if (rc) return rc;
        timer.mark(kPassMarch);
        hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, a.queue, 1);
        rc = check_launch("reset_counters_kernel");
        if (rc) return rc;
        timer.mark(kPassOther);
        a.allow_return = r + 1 < rounds;
        hipLaunchKernelGGL(trace_mixed_pass2_kernel<POSINV>, dim3(g2), dim3(kPass2Block), 0, s, a);
        rc = check_launch("trace_mixed_pass2_kernel");
        if (rc) return rc;
        timer.mark(kPassFp32);
    }
    hipLaunchKernelGGL(normals_list_kernel<POSINV>, dim3(g3), dim3(kMarchBlock), 0, s, a);
    int rc = check_launch("normals_list_kernel");
    if (rc) return rc;
    timer.mark(kPassNormals);
    return NGLOD_OK;
}

int launch_trace_mixed(const TraceArgs& a, int64_t n, hipStream_t s, PassTimer& timer) {
    return a.m.pos_inv ? trace_mixed_rounds<true>(a, n, s, timer) : trace_mixed_rounds<false>(a, n, s, timer);
}

}  // namespace nglod

using namespace nglod;

extern "C" int nglod_pack_mixed(const nglod_model_t* model, void* stream) {
    if (!model) return fail(NGLOD_ERR_ARG, "model is NULL");
    if (model->feature_dim != kC || model->hidden_dim != kH)
        return fail(NGLOD_ERR_UNSUPPORTED, "kernels are built for feature_dim 32, hidden_dim 128");
    if (model->num_lods < 1 || model->num_lods > NGLOD_MAX_LODS) return fail(NGLOD_ERR_ARG, "num_lods out of range");
    if (model->num_decoders != 1 && model->num_decoders != model->num_lods)
        return fail(NGLOD_ERR_ARG, "num_decoders must be 1 or num_lods");
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < model->num_lods; ++i) {
        const int r = model->res[i];
        if (r < 1 || r > 1024) return fail(NGLOD_ERR_ARG, "grid resolution out of range");
        if (!model->grid[i] || !model->grid_h[i] || !aligned16(model->grid_h[i]))
            return fail(NGLOD_ERR_ARG, "grid / grid_h pointer NULL or misaligned");
        const int64_t cnt = (int64_t)(r + 1) * (r + 1) * (r + 1) * kC;
        hipLaunchKernelGGL(pack_grid_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, model->grid[i],
                           (__half*)model->grid_h[i], cnt);
        int rc = check_launch("pack_grid_kernel");
        if (rc) return rc;
    }
    const int din = model->pos_invariant ? kC : kC + 3;
    const int off = model->pos_invariant ? 0 : 3;
    for (int i = 0; i < model->num_decoders; ++i) {
        if (!model->w0[i] || !model->w0_bf16[i] || !aligned16(model->w0_bf16[i]))
            return fail(NGLOD_ERR_ARG, "w0 / w0_bf16 pointer NULL or misaligned");
        hipLaunchKernelGGL(pack_w0_kernel, dim3(16), dim3(256), 0, s, model->w0[i], din, off,
                           (uint16_t*)model->w0_bf16[i]);
        int rc = check_launch("pack_w0_kernel");
        if (rc) return rc;
    }
    return NGLOD_OK;
}
