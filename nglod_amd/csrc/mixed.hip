// Mixed-precision form: FP16 grids, BF16 MFMA first decoder layer with FP32
// accumulation, and FP32 re-evaluation near the surface.
//   nglod_pack_mixed        FP32 model -> cumulative FP32 / FP16 grids + BF16 W0 fragments
//   launch_trace_mixed      nglod_trace_dense(form = NGLOD_MIXED):
//     pass 1  persistent mixed march.  When a ray stops on a surface test
//             (|d| <= min_dis or the oscillation test) it rewinds its last
//             near-surface evaluations (at most replay_steps, back to the
//             last step of at least near_dist) and writes its hand-off
//             record at handoff[ray] (flag hflag[ray]); a ray it finishes as
//             a hit is handed over the same way, in phase PH_N0.
//     pass 2  persistent FP32 kernel over the hand-off records, claimed in
//             ray order: re-evaluates the rewound steps and goes on in FP32
//             to its FP32 stop; a ray whose FP32 step grows past return_dist
//             (the mixed stop was a near miss) is handed back (ret[ray],
//             rflag[ray]) and pass 1 resumes it.  Passes 1 and 2 run
//             kMixedRounds times (once when return_dist = 0); the last pass 2
//             keeps every ray.  Pass 2 takes the six FP32 normal taps of
//             every hit, in-lane after its continuation or from a PH_N0
//             record (there is no separate normals pass).
// No pass appends to a shared list: the records are indexed by ray, so the
// only atomics are the claim counters (one per wave per refill).  A single
// append counter hit by every wave serialised the march kernels measurably.
#include "api_common.hpp"
#include "fp32_wave.hpp"
#include "mixed_device.hpp"
#include "split_wave.hpp"
#include "trace_common.hpp"

namespace nglod {

// Pass 1 reads its grid from global memory: with LDS copies of coarse levels
// (when it gathered every level) it measured 10-20% slower on the 1280x720
// LOD-4 frame.  Pass 2 gathers one FP32 cumulative grid per query and keeps
// only the decoder in LDS (134 VGPRs); 512-thread blocks measured faster than
// 256 for it (0.51 against 0.59 ms, before it took the normal taps), and more
// waves per SIMD did not help.
constexpr int kMixedBlock = 256;
#ifndef NGLOD_MIXED_FP32_BLOCK
#define NGLOD_MIXED_FP32_BLOCK 512
#endif
#ifndef NGLOD_MIXED_FP32_WAVES
#define NGLOD_MIXED_FP32_WAVES 2
#endif
constexpr int kPass2Block = NGLOD_MIXED_FP32_BLOCK;
constexpr int kMixedFp32Waves = NGLOD_MIXED_FP32_WAVES;
// Pass 1 claims rays 32 ids at a time (the FP32 march claims per idle lane):
// measured faster for the mixed march, slightly slower for the FP32 one.  On
// the 1280x720 LOD-4 frame with the claim counters one line apart: 16 ids
// 0.725 ms, 32 0.705, 64 0.728, 128 0.858.
constexpr int kPass1Chunk = 32;

// ---------------------------------------------------------------------------
// Cumulative grid of level li: at corner (cx, cy, cz) of grid li, the sum over
// levels l <= li of grid l's trilinear sample there (OctreeSDF.py:102-110
// sums the per-level samples of a query the same way).  With R = res[li] and
// r = res[l], q = R / r is an integer (checked on the host), the corner sits
// at c / q in grid l's corner units: lower corner c / q (kept <= r - 1, the
// border convention of axis_weights, weights (0, 1) at c = R) and weights
// (q - c % q) / q, (c % q) / q.  Summed in FP64, levels in order and each
// level's 8 corners in x, y, z order, weight ((wx * wy) * wz); rounded once to
// FP32, and that FP32 value rounded to FP16.  One thread per (corner, channel).
struct CumLevels {
    const float* grid[NGLOD_MAX_LODS];
    int res[NGLOD_MAX_LODS];
};

__device__ __forceinline__ void cum_axis(int c, int q, int r, int& i0, double& w0, double& w1) {
    i0 = c / q;
    int rem = c - i0 * q;
    if (i0 > r - 1) {
        i0 = r - 1;
        rem = q;
    }
    w0 = (double)(q - rem) / (double)q;
    w1 = (double)rem / (double)q;
}

__global__ void cumulative_grid_kernel(CumLevels lv, int li, float* __restrict__ out32, __half* __restrict__ out16) {
    const int R = lv.res[li];
    const int S = R + 1;
    const int64_t n = (int64_t)S * S * S * kC;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % kC);
    const int64_t corner = i / kC;
    const int cx = (int)(corner % S), cy = (int)((corner / S) % S), cz = (int)(corner / ((int64_t)S * S));
    double acc = 0.0;
    for (int l = 0; l <= li; ++l) {
        const int r = lv.res[l], q = R / r, s = r + 1;
        int ix, iy, iz;
        double wx[2], wy[2], wz[2];
        cum_axis(cx, q, r, ix, wx[0], wx[1]);
        cum_axis(cy, q, r, iy, wy[0], wy[1]);
        cum_axis(cz, q, r, iz, wz[0], wz[1]);
        const float* g = lv.grid[l] + ((size_t)((iz * s + iy) * s + ix)) * kC + c;
        const size_t sx = kC, sy = (size_t)s * kC, sz = (size_t)s * s * kC;
        for (int k = 0; k < 8; ++k) {
            const int bx = k & 1, by = (k >> 1) & 1, bz = k >> 2;
            const double w = (wx[bx] * wy[by]) * wz[bz];
            acc = acc + w * (double)g[bx * sx + by * sy + bz * sz];
        }
    }
    const float f = (float)acc;
    out32[i] = f;
    out16[i] = __float2half_rn(f);
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

// Pre-evaluation (t, D) of a ray's last four mixed-precision evaluations (a
// shift register), and the rewind at a mixed surface stop.
struct MarchHistory {
    float t0, t1, t2, t3, d0, d1, d2, d3;
    int n;      // evaluations recorded since the ray started (or came back)
    bool init;  // the first recorded evaluation was the box-entry one

    __device__ __forceinline__ void clear() {
        n = 0;
        init = false;
        t0 = t1 = t2 = t3 = d0 = d1 = d2 = d3 = 0.0f;  // finite: rewind weighs every entry
    }

    __device__ __forceinline__ void push(const RayState& s) {
        // The box-entry evaluation has no step yet (a fresh ray's D is unset);
        // record 0 there: rewind weighs every entry, so all must be finite.
        const bool entry = s.phase == PH_INIT;
        if (entry) {
            n = 0;
            init = true;
        }
        t3 = t2; t2 = t1; t1 = t0; t0 = s.t;
        d3 = d2; d2 = d1; d1 = d0; d0 = entry ? 0.0f : s.D;
        ++n;
    }

    // Rewind j evaluations (at most replay_steps, back to the box-entry
    // evaluation if the ray made fewer steps): the stopping one and those
    // before it back to (not including) the last step of at least near_dist.
    // A step that long was taken away from the surface, where the
    // mixed-precision error cannot carry the ray past it.  Leaves the ray with
    // a pending evaluation (phase INIT or MARCH) for the FP32 form.
    __device__ __forceinline__ void rewind(RayState& s, const TraceArgs& a) const {
        const int jmax = n < a.replay_steps ? n : a.replay_steps;
        int j = 1;
        if (a.near_dist <= 0.0f) {
            j = jmax;
        } else {
            if (j < jmax && fabsf(d0) < a.near_dist) ++j;
            if (j == 2 && j < jmax && fabsf(d1) < a.near_dist) ++j;
            if (j == 3 && j < jmax && fabsf(d2) < a.near_dist) ++j;
        }
        // Entry j as a sum with weights 1 and 0 (exact: one product by 1, the
        // rest add zeros; every entry is finite).  A ?: chain over the fields
        // was compiled into an indexed load, which kept the whole history in
        // scratch memory and stored it there on every step.
        const float w1 = j == 1 ? 1.0f : 0.0f, w2 = j == 2 ? 1.0f : 0.0f;
        const float w3 = j == 3 ? 1.0f : 0.0f, w4 = j == 4 ? 1.0f : 0.0f;
        const float tj = ((t0 * w1 + t1 * w2) + t2 * w3) + t3 * w4;
        const float dj = ((d0 * w1 + d1 * w2) + d2 * w3) + d3 * w4;
        if (init && j == n) {
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
    }
};

__device__ __forceinline__ void add_evals(const TraceArgs& a, int slot, uint32_t nevals, uint32_t lane) {
    if (!a.evals) return;
    unsigned long long tot = nevals;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (lane == 0 && tot) atomicAdd(a.evals + slot, tot);
}

// RESUME: a later round, over the rays pass 2 handed back (ret, rflag)
// instead of the frame's rays.  A separate instantiation: compiled into the
// first-round kernel, the resume path slowed its march by half even unused.
// Up to this many busy lanes, a drained wave of the FP32 passes and of the
// pooled march gathers cooperatively (16 lanes per query).
constexpr int kCoopLanes = 4;
// Pass 1's tail: up to 8 busy lanes, 16 lanes per query up to 4 of them
// (gather_fp16_coop) and 8 lanes per query above (gather_fp16_coop8); on
// unless TraceArgs::solo_lanes is 0.  8 against 4 measured pass 1 0.68 ms
// against 0.69 ms on the 1280x720 frame.
constexpr int kPass1CoopLanes = 8;

template <bool POSINV, bool RESUME>
__global__ __launch_bounds__(kMixedBlock) void trace_mixed_pass1_kernel(TraceArgs a) {
    __shared__ MixedLds lds;
    __shared__ MixedLevels levels;
    stage_mixed_lds(a.m, a.dec, &lds);
    stage_mixed_levels(a.m, &levels);
    __syncthreads();
    const uint32_t lane = lane_id();
    const float b1 = a.m.b1[a.dec][0];
    RayFeed feed;
    feed.init(a.nrays, a.queue_stride);
    RayState s;
    s.ray = -1;
    s.phase = PH_IDLE;
    uint32_t nevals = 0;
    MarchHistory hist;
    hist.clear();
    const int coop = a.solo_lanes == 0 ? 0 : kPass1CoopLanes;
    bool tail = false;
    for (;;) {
        if (RESUME) {
            const int64_t id = feed.next_flagged<true>(a.ret_queue, a.rflag, lane, s.phase == PH_IDLE);
            if (id >= 0) {
                s.resume(a, a.ret[id]);
                hist.clear();
            }
        } else {
            feed.refill<kPass1Chunk>(s, a, lane);
        }
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        // nothing left to claim and few rays: the tail loop below
        if ((RESUME ? feed.drained && feed.fmask == 0 : feed.empty()) && __popcll(act) <= coop) {
            tail = true;
            break;
        }
        const bool busy = s.phase != PH_IDLE;
        __half2 acc[kC / 2];
        gather_fp16_cum(a.m, busy, s.qx, s.qy, s.qz, acc);
        const float d = decode_mixed_wave<POSINV>(acc, &lds, b1, lane, busy, s.qx, s.qy, s.qz);
        if (busy) {
            ++nevals;
            hist.push(s);
            s.consume(a, d, true);
            if (s.phase == PH_REPLAY) {
                hist.rewind(s, a);
                s.hand_off_to(a.handoff, a.hflag);
            }
        }
    }
    // The wave's last few rays: the same steps with the gather spread over the
    // idle lanes (same bits).  A separate loop, so that its registers do not
    // weigh on the main loop's.
    while (tail) {
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        const bool busy = s.phase != PH_IDLE;
        __half2 acc[kC / 2];
        if (__popcll(act) <= 4)
            gather_fp16_coop(&levels, lane, act, s.qx, s.qy, s.qz, acc);
        else
            gather_fp16_coop8(&levels, lane, act, s.qx, s.qy, s.qz, acc);
        const float d = decode_mixed_wave_few<POSINV>(acc, &lds, b1, lane, busy, s.qx, s.qy, s.qz);
        if (busy) {
            ++nevals;
            hist.push(s);
            s.consume(a, d, true);
            if (s.phase == PH_REPLAY) {
                hist.rewind(s, a);
                s.hand_off_to(a.handoff, a.hflag);
            }
        }
    }
    add_evals(a, 1, nevals, lane);
}

// RETURNS: near misses may be handed back to a later pass 1 (not in the last
// round); a separate instantiation, as for pass 1.
template <bool POSINV, bool RETURNS>
__global__ __launch_bounds__(kPass2Block, kMixedFp32Waves) void trace_mixed_pass2_kernel(TraceArgs a) {
    __shared__ SplitLds lds;
    __shared__ SoloLevels solo_lv;
    stage_split_lds(a.m, a.dec, &lds);
    stage_solo_levels(a.m, &solo_lv);
    __syncthreads();
    const float b1 = a.m.b1[a.dec][0];
    const uint32_t lane = lane_id();
    RayFeed feed;
    feed.init(a.nrays, a.queue_stride);
    RayState s;
    s.ray = -1;
    s.phase = PH_IDLE;
    uint32_t nevals = 0;
    const int coop = a.solo_lanes < kCoopLanes ? a.solo_lanes : kCoopLanes;
    bool tail = false;
    for (;;) {
        const int64_t id = feed.next_flagged<true>(a.handoff_queue, a.hflag, lane, s.phase == PH_IDLE);
        if (id >= 0) s.resume(a, a.handoff[id]);
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        if (feed.drained && feed.fmask == 0 && __popcll(act) <= coop) {
            tail = true;
            break;
        }
        const bool busy = s.phase != PH_IDLE;
        const float d = sdf_split_wave<POSINV>(a.m, &lds, b1, lane, busy, s.qx, s.qy, s.qz);
        if (busy) {
            ++nevals;
            s.consume(a, d);
            // hand back only a ray that moves away from the surface (step growing)
            if (RETURNS && s.phase == PH_MARCH && fabsf(s.D) > a.return_dist && fabsf(s.D) > fabsf(s.Dp))
                s.hand_off_to(a.ret, a.rflag);
        }
    }
    // The wave's last few rays: the same steps with the gather spread over the
    // idle lanes (same bits).  A separate loop, so that its registers do not
    // weigh on the main loop's.
    while (tail) {
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        const bool busy = s.phase != PH_IDLE;
        float f[kC];
        gather_fp32_coop(&solo_lv, lane, act, s.qx, s.qy, s.qz, f);
        const float d = decode_split_wave_few<POSINV>(f, &lds, b1, lane, busy, s.qx, s.qy, s.qz);
        if (busy) {
            ++nevals;
            s.consume(a, d);
            if (RETURNS && s.phase == PH_MARCH && fabsf(s.D) > a.return_dist && fabsf(s.D) > fabsf(s.Dp))
                s.hand_off_to(a.ret, a.rflag);
        }
    }
    add_evals(a, 0, nevals, lane);
}


// ---------------------------------------------------------------------------
// Pooled trace: the rays of several requests (each its own model and LOD) in
// one launch per pass.  The rays live in one pool of per-ray arrays; request
// r owns pool slots [off_r, off_r + count_r) and virtual ids
// [vbeg_r, vbeg_r + count_r) of the launch's id space, which the ray feed
// hands out (XCD ranges and chunk claims as for one frame).  A lane keeps its
// ray's request; each wave trip evaluates the busy lanes request by request
// (usually one group; two where a wave's rays straddle a request boundary),
// with a decoder slot per wave in LDS that is restaged when the group's
// request differs from the one it holds.  Records passed between passes are
// indexed by pool slot.  Per-ray arithmetic is that of the single-request
// kernels, so a request's output does not depend on the other requests.

// Index of the request that owns virtual id v (requests sorted by vbeg).
__device__ __forceinline__ int pool_request_of(const PoolArgs& p, int64_t v) {
    int lo = 0, hi = p.nreq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.req[mid].vbeg <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int64_t pool_slot_of(const PoolArgs& p, int r, int64_t v) {
    return p.req[r].off + (v - p.req[r].vbeg);
}

// Orders a wave's LDS writes before its own later LDS reads (one wave, no
// block barrier: the waves of a block run different requests).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One model's level table in the device model table, read through the
// constant address space: the grid pointers and resolutions come in with
// scalar loads, as a single-request kernel reads them from its arguments.
struct ScalarInts {
    const int* p;
    __device__ __forceinline__ int operator[](int i) const { return ((const NGLOD_CONST int*)p)[i]; }
};
template <typename T>
struct ScalarPtrs {
    const T* const* p;
    __device__ __forceinline__ const T* operator[](int i) const { return ((const T* const NGLOD_CONST*)p)[i]; }
};
struct PoolLevelsH {  // FP16 grids (gather_fp16_cum)
    ScalarInts res;
    ScalarPtrs<uint16_t> grid_h;
};
struct PoolLevels32 {  // FP32 grids (gather_fp32)
    ScalarInts res;
    ScalarPtrs<float> grid;
};
__device__ __forceinline__ PoolLevelsH pool_levels_h(const PoolArgs& p, int g) {
    const ModelArgs* m = p.models + p.req[g].model;
    return PoolLevelsH{ScalarInts{m->res}, ScalarPtrs<uint16_t>{m->grid_h}};
}
__device__ __forceinline__ PoolLevels32 pool_levels32(const PoolArgs& p, int g) {
    const ModelArgs* m = p.models + p.req[g].model;
    return PoolLevels32{ScalarInts{m->res}, ScalarPtrs<float>{m->grid}};
}

// A wave's LDS slot in the pooled mixed march: the decoder, the level
// table of the request it belongs to (for the cooperative tail gather, which
// reads MixedLevels as vector values), and b1.
struct PoolMixedSlot {
    MixedLds dec;
    MixedLevels lv;
    float b1;
};

// A wave's LDS slot in the pooled FP32 passes: decoder, level table, b1 and
// the one-query-at-a-time tail's hidden vector.
struct PoolFp32Slot {
    Fp32Lds dec;
    SoloLevels lv;
    float b1;
};

// Fills a wave's mixed slot for request e (the wave's 64 lanes).  Out of
// line: restaging is rare, and its code inlined into the march loop would
// weigh on the loop's registers and instruction cache.
__device__ __noinline__ void stage_pool_mixed(const ModelArgs* mp, int dec, PoolMixedSlot* s, uint32_t lane) {
    const ModelArgs& m = *mp;
    const int din = m.pos_inv ? kC : kC + 3;
    const uint4* frag = reinterpret_cast<const uint4*>(m.w0_bf16[dec]);
    for (int j = (int)lane; j < 4 * 2 * 64; j += 64) s->dec.wfrag[j] = frag[j];
    for (int j = (int)lane; j < kH; j += 64) {
        s->dec.b0[j] = m.b0[dec][j];
        s->dec.w1[j] = m.w1[dec][j];
        const bool xyz = !m.pos_inv;
        s->dec.wxyz[j][0] = xyz ? m.w0[dec][j * din + 0] : 0.0f;
        s->dec.wxyz[j][1] = xyz ? m.w0[dec][j * din + 1] : 0.0f;
        s->dec.wxyz[j][2] = xyz ? m.w0[dec][j * din + 2] : 0.0f;
        s->dec.wxyz[j][3] = 0.0f;
    }
    if (lane < (uint32_t)NGLOD_MAX_LODS) {
        s->lv.grid_h[lane] = reinterpret_cast<const __half*>(m.grid_h[lane]);
        s->lv.res[lane] = m.res[lane];
    }
    if (lane == 0) s->b1 = m.b1[dec][0];
}

// Fills a wave's FP32 slot for request e (the wave's 64 lanes); out of line.
__device__ __noinline__ void stage_pool_fp32(const ModelArgs* mp, int dec, PoolFp32Slot* s, uint32_t lane) {
    const ModelArgs& m = *mp;
    const float* W0 = m.w0[dec];
    const int din = m.pos_inv ? kC : kC + 3;
    const int off = m.pos_inv ? 0 : 3;
    for (int j = (int)lane; j < 4 * 4 * 64; j += 64) {
        const int l = j & 63;
        const int kq = (j >> 6) & 3;
        const int h = j >> 8;
        const int row = 32 * h + (l & 31);
        const int col = off + 8 * kq + (l >> 5);
        float4 v;
        v.x = W0[row * din + col + 0];
        v.y = W0[row * din + col + 2];
        v.z = W0[row * din + col + 4];
        v.w = W0[row * din + col + 6];
        s->dec.afrag[j] = v;
    }
    for (int j = (int)lane; j < kH; j += 64) {
        s->dec.b0[j] = m.b0[dec][j];
        s->dec.w1[j] = m.w1[dec][j];
        const bool xyz = !m.pos_inv;
        s->dec.wxyz[j][0] = xyz ? W0[j * din + 0] : 0.0f;
        s->dec.wxyz[j][1] = xyz ? W0[j * din + 1] : 0.0f;
        s->dec.wxyz[j][2] = xyz ? W0[j * din + 2] : 0.0f;
        s->dec.wxyz[j][3] = 0.0f;
    }
    if (lane < (uint32_t)NGLOD_MAX_LODS) {
        s->lv.grid[lane] = m.grid[lane];
        s->lv.res[lane] = m.res[lane];
    }
    if (lane == 0) s->b1 = m.b1[dec][0];
}

// Makes the wave's slot hold request g (restaging it if it holds another).
template <typename SLOT, typename STAGE>
__device__ __forceinline__ void hold_request(const PoolArgs& p, int g, int& held, SLOT* slot, uint32_t lane,
                                             STAGE stage) {
    if (g == held) return;
    wave_lds_sync();
    stage(p.models + p.req[g].model, p.req[g].dec, slot, lane);
    wave_lds_sync();
    held = g;
}

constexpr int kPoolFp32Block = 256;

// Three waves per SIMD, as the single-request march (its registers allow
// three; the group loop would otherwise push this kernel to two).
template <bool POSINV>
__global__ __launch_bounds__(kMixedBlock, 3) void pool_pass1_kernel(TraceArgs a, PoolArgs p) {
    __shared__ PoolMixedSlot slots[kMixedBlock / 64];
    PoolMixedSlot* sl = &slots[threadIdx.x >> 6];
    const uint32_t lane = lane_id();
    int held = -1;  // request whose decoder this wave's slot holds (wave-uniform)
    RayFeed feed;
    feed.init(p.total, a.queue_stride);
    RayState s;
    s.ray = -1;
    s.phase = PH_IDLE;
    int rq = 0;
    uint32_t nevals = 0;
    MarchHistory hist;
    hist.clear();
    const int coop = a.solo_lanes < kCoopLanes ? a.solo_lanes : kCoopLanes;
    bool tail = false;
    for (;;) {
        // the claim loop only when a lane needs a ray (as RayFeed::refill)
        if (__ballot(s.phase == PH_IDLE && !feed.empty()) != 0) {
            for (;;) {
                const bool need = s.phase == PH_IDLE;
                if (__ballot(need && !feed.empty()) == 0) break;
                const int64_t v = feed.next<kPass1Chunk>(a.queue, nullptr, lane, need);
                if (v >= 0) {
                    rq = pool_request_of(p, v);
                    s.start(a, pool_slot_of(p, rq, v));
                }
            }
        }
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        if (feed.empty() && __popcll(act) <= coop) {
            tail = true;
            break;
        }
        const bool busy = s.phase != PH_IDLE;
        float d = 0.0f;
        uint64_t todo = act;
        while (todo) {
            const int g = __builtin_amdgcn_readlane(rq, __ffsll((unsigned long long)todo) - 1);
            const bool in = busy && rq == g;
            todo &= ~__ballot(in);
            hold_request(p, g, held, sl, lane, stage_pool_mixed);
            const float dg = sdf_mixed_wave<POSINV>(pool_levels_h(p, g), &sl->dec, sl->b1, lane, in,
                                                    s.qx, s.qy, s.qz);
            if (in) d = dg;
        }
        if (busy) {
            ++nevals;
            hist.push(s);
            s.consume(a, d, true);
            if (s.phase == PH_REPLAY) {
                hist.rewind(s, a);
                s.hand_off_to(a.handoff, a.hflag);
            }
        }
    }
    // the wave's last few rays: the gather spread over the idle lanes, group by group
    while (tail) {
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        const bool busy = s.phase != PH_IDLE;
        float d = 0.0f;
        uint64_t todo = act;
        while (todo) {
            const int g = __builtin_amdgcn_readlane(rq, __ffsll((unsigned long long)todo) - 1);
            const bool in = busy && rq == g;
            const uint64_t gm = __ballot(in);
            todo &= ~gm;
            hold_request(p, g, held, sl, lane, stage_pool_mixed);
            __half2 acc[kC / 2];
            gather_fp16_coop(&sl->lv, lane, gm, s.qx, s.qy, s.qz, acc);
            const float dg = decode_mixed_wave_few<POSINV>(acc, &sl->dec, sl->b1, lane, in, s.qx, s.qy, s.qz);
            if (in) d = dg;
        }
        if (busy) {
            ++nevals;
            hist.push(s);
            s.consume(a, d, true);
            if (s.phase == PH_REPLAY) {
                hist.rewind(s, a);
                s.hand_off_to(a.handoff, a.hflag);
            }
        }
    }
    add_evals(a, 1, nevals, lane);
}

// The next flagged record for each lane with `need`: the wave claims 64
// virtual ids at a time, each lane loads the flag of its id's pool slot, and
// the flagged ids are handed out in order from the ballot mask (as
// RayFeed::next_flagged, through the slot map).  Returns the pool slot (-1:
// none left) and the request in r.
template <bool SPLIT>
__device__ __forceinline__ int64_t pool_next_flagged(RayFeed& f, const PoolArgs& p, uint32_t* counters,
                                                     const uint8_t* flag, uint32_t lane, bool need, int& r) {
    int64_t slot = -1;
    for (;;) {
        const bool want = need && slot < 0;
        const uint64_t mask = __ballot(want);
        if (mask == 0) break;
        if (f.fmask == 0) {
            if (f.drained) break;
            const int leader = __ffsll((unsigned long long)mask) - 1;
            if (!f.claim<SPLIT>(counters, lane, leader, 64)) continue;
            const int64_t c = f.cb + (int64_t)lane;
            bool fl = false;
            if (c < f.ce) {
                const int rr = pool_request_of(p, c);
                fl = flag[pool_slot_of(p, rr, c)] != 0;
            }
            f.fmask = __ballot(fl);
            if (f.fmask == 0) continue;
        }
        const int avail = __popcll(f.fmask);
        const int cnt = __popcll(mask);
        if (want) {
            const int k = (int)mask_rank(mask);
            if (k < avail) {
                const int64_t v = f.cb + nth_set_bit(f.fmask, k);
                r = pool_request_of(p, v);
                slot = pool_slot_of(p, r, v);
            }
        }
        if (cnt >= avail)
            f.fmask = 0;
        else
            f.fmask &= ~((1ull << nth_set_bit(f.fmask, cnt)) - 1ull);
    }
    return slot;
}

// Pass 2 over the pool: FP32 continuation of the handed-off rays, claimed in
// virtual-id order (flags indexed by pool slot), with the one-query-at-a-time
// tail of the single-request pass.
template <bool POSINV>
__global__ __launch_bounds__(kPoolFp32Block) void pool_pass2_kernel(TraceArgs a, PoolArgs p) {
    __shared__ PoolFp32Slot slots[kPoolFp32Block / 64];
    PoolFp32Slot* sl = &slots[threadIdx.x >> 6];
    const uint32_t lane = lane_id();
    int held = -1;
    RayFeed feed;
    feed.init(p.total, a.queue_stride);
    RayState s;
    s.ray = -1;
    s.phase = PH_IDLE;
    int rq = 0;
    uint32_t nevals = 0;
    const int coop = a.solo_lanes < kCoopLanes ? a.solo_lanes : kCoopLanes;
    bool tail = false;
    for (;;) {
        int r = 0;
        const int64_t slot = pool_next_flagged<true>(feed, p, a.handoff_queue, a.hflag, lane, s.phase == PH_IDLE, r);
        if (slot >= 0) {
            rq = r;
            s.resume(a, a.handoff[slot]);
        }
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        if (feed.drained && feed.fmask == 0 && __popcll(act) <= coop) {
            tail = true;
            break;
        }
        const bool busy = s.phase != PH_IDLE;
        float d = 0.0f;
        uint64_t todo = act;
        while (todo) {
            const int g = __builtin_amdgcn_readlane(rq, __ffsll((unsigned long long)todo) - 1);
            const bool in = busy && rq == g;
            todo &= ~__ballot(in);
            hold_request(p, g, held, sl, lane, stage_pool_fp32);
            const float dg = sdf_split_wave<POSINV>(pool_levels32(p, g), &sl->dec, sl->b1, lane, in,
                                                    s.qx, s.qy, s.qz);
            if (in) d = dg;
        }
        if (busy) {
            ++nevals;
            s.consume(a, d);
        }
    }
    // the wave's last few rays, request by request: the gather spread over
    // the idle lanes (gather_fp32_coop), the same decoder (same bits)
    while (tail) {
        const uint64_t act = __ballot(s.phase != PH_IDLE);
        if (act == 0) break;
        const bool busy = s.phase != PH_IDLE;
        float d = 0.0f;
        uint64_t todo = act;
        while (todo) {
            const int g = __builtin_amdgcn_readlane(rq, __ffsll((unsigned long long)todo) - 1);
            const bool in = busy && rq == g;
            const uint64_t gm = __ballot(in);
            todo &= ~gm;
            hold_request(p, g, held, sl, lane, stage_pool_fp32);
            float f[kC];
            gather_fp32_coop(&sl->lv, lane, gm, s.qx, s.qy, s.qz, f);
            const float dg = decode_split_wave_few<POSINV>(f, &sl->dec, sl->b1, lane, in, s.qx, s.qy, s.qz);
            if (in) d = dg;
        }
        if (s.phase != PH_IDLE) {
            ++nevals;
            s.consume(a, d);
        }
    }
    add_evals(a, 0, nevals, lane);
}


// Clears the hand-off flags of the launch's slots (one thread per virtual id)
// and the ncounters claim counters (`stride` words apart).
__global__ void pool_clear_kernel(PoolArgs p, uint8_t* hflag, uint32_t* counters, int ncounters, int stride) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < ncounters) counters[v * stride] = 0u;
    if (v < p.total) {
        const int r = pool_request_of(p, v);
        hflag[pool_slot_of(p, r, v)] = 0;
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


__global__ void clear_words_kernel(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Clears `words` uint32 of flags and the kQueues claim counters of a pass
// (`stride` words apart).
__global__ void clear_round_kernel(uint32_t* flags, int words, uint32_t* counters, int stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) flags[i] = 0u;
    if (i < kQueues) counters[i * stride] = 0u;
}

// Rounds of (pass 1, pass 2) when near misses may go back to mixed precision
// (return_dist > 0); with return_dist = 0 one round, pass 2 keeps every ray.
constexpr int kMixedRounds = 3;

template <bool POSINV>
// The mixed form's normals: pass 1 hands every ray it finishes as a hit to
// pass 2 as a PH_N0 record (defer_normals = 1), and pass 2 takes the six FP32
// taps in-lane after the continuation (defer_normals = 0), so its drain
// fills with tap work and there is no third launch.  The taps are the FP32
// evaluations a separate normals pass would make (same bits).
static int trace_mixed_rounds(TraceArgs a, int64_t n, hipStream_t s, PassTimer& timer) {
    const int g1 = grid_for(trace_mixed_pass1_kernel<POSINV, false>, n, kMixedBlock);
    const int g1r = grid_for(trace_mixed_pass1_kernel<POSINV, true>, n, kMixedBlock);
    const int g2 = grid_for(trace_mixed_pass2_kernel<POSINV, false>, n, kPass2Block);
    const int words = (int)((n + 3) / 4);
    const dim3 cg((unsigned)((words + 255) / 256));
    const int rounds = a.return_dist > 0.0f ? kMixedRounds : 1;
    TraceArgs a1 = a, a2 = a;  // pass 1 / pass 2 arguments
    a1.defer_normals = 1;
    a2.defer_normals = 0;
    for (int r = 0; r < rounds; ++r) {
        const bool allow_return = r + 1 < rounds;
        hipLaunchKernelGGL(clear_round_kernel, cg, dim3(256), 0, s, (uint32_t*)a.hflag, words, a.handoff_queue, a.queue_stride);
        int rc = check_launch("clear_round_kernel");
        if (rc) return rc;
        timer.mark(kPassOther);
        if (r == 0)
            hipLaunchKernelGGL((trace_mixed_pass1_kernel<POSINV, false>), dim3(g1), dim3(kMixedBlock), 0, s, a1);
        else
            hipLaunchKernelGGL((trace_mixed_pass1_kernel<POSINV, true>), dim3(g1r), dim3(kMixedBlock), 0, s, a1);
        rc = check_launch("trace_mixed_pass1_kernel");
        if (rc) return rc;
        timer.mark(kPassMarch);
        if (allow_return) {
            hipLaunchKernelGGL(clear_round_kernel, cg, dim3(256), 0, s, (uint32_t*)a.rflag, words, a.ret_queue, a.queue_stride);
            rc = check_launch("clear_round_kernel");
            if (rc) return rc;
            timer.mark(kPassOther);
        }
        if (allow_return)
            hipLaunchKernelGGL((trace_mixed_pass2_kernel<POSINV, true>), dim3(g2), dim3(kPass2Block), 0, s, a2);
        else
            hipLaunchKernelGGL((trace_mixed_pass2_kernel<POSINV, false>), dim3(g2), dim3(kPass2Block), 0, s, a2);
        rc = check_launch("trace_mixed_pass2_kernel");
        if (rc) return rc;
        timer.mark(kPassFp32);
    }
    return NGLOD_OK;
}

int launch_trace_mixed(const TraceArgs& a, int64_t n, hipStream_t s, PassTimer& timer) {
    // the mixed kernels gather one level, the cumulative grid of the LOD
    if (a.lod != 0 || a.m.num_lods != 1) return fail(NGLOD_ERR_ARG, "mixed form: one cumulative level (gather lod 0)");
    return a.m.pos_inv ? trace_mixed_rounds<true>(a, n, s, timer) : trace_mixed_rounds<false>(a, n, s, timer);
}

template <bool POSINV>
static int trace_pool_passes(const TraceArgs& a, const PoolArgs& p, hipStream_t s, PassTimer& timer) {
    const int64_t n = p.total;
    TraceArgs a1 = a, a2 = a;  // normals as in the dense trace: taken in pass 2
    a1.defer_normals = 1;
    a2.defer_normals = 0;
    hipLaunchKernelGGL(pool_clear_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, a.hflag, a.queue,
                       3 * kQueues, a.queue_stride);
    int rc = check_launch("pool_clear_kernel");
    if (rc) return rc;
    timer.mark(kPassOther);
    hipLaunchKernelGGL(pool_pass1_kernel<POSINV>, dim3(grid_for(pool_pass1_kernel<POSINV>, n, kMixedBlock)),
                       dim3(kMixedBlock), 0, s, a1, p);
    rc = check_launch("pool_pass1_kernel");
    if (rc) return rc;
    timer.mark(kPassMarch);
    hipLaunchKernelGGL(pool_pass2_kernel<POSINV>, dim3(grid_for(pool_pass2_kernel<POSINV>, n, kPoolFp32Block)),
                       dim3(kPoolFp32Block), 0, s, a2, p);
    rc = check_launch("pool_pass2_kernel");
    if (rc) return rc;
    timer.mark(kPassFp32);
    return NGLOD_OK;
}

int launch_trace_pool(const TraceArgs& a, const PoolArgs& p, bool pos_inv, hipStream_t s, PassTimer& timer) {
    return pos_inv ? trace_pool_passes<true>(a, p, s, timer) : trace_pool_passes<false>(a, p, s, timer);
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
    CumLevels lv;
    memset(&lv, 0, sizeof(lv));
    for (int i = 0; i < model->num_lods; ++i) {
        const int r = model->res[i];
        if (r < 1 || r > 1024) return fail(NGLOD_ERR_ARG, "grid resolution out of range");
        for (int l = 0; l < i; ++l)
            if (r % model->res[l] != 0)
                return fail(NGLOD_ERR_UNSUPPORTED, "the mixed form needs each grid resolution to divide the later ones");
        if (!model->grid[i] || !aligned16(model->grid[i]) || !model->grid_cum[i] || !aligned16(model->grid_cum[i]) ||
            !model->grid_h[i] || !aligned16(model->grid_h[i]))
            return fail(NGLOD_ERR_ARG, "grid / grid_cum / grid_h pointer NULL or misaligned");
        lv.grid[i] = model->grid[i];
        lv.res[i] = r;
    }
    for (int i = 0; i < model->num_lods; ++i) {
        const int r = model->res[i];
        const int64_t cnt = (int64_t)(r + 1) * (r + 1) * (r + 1) * kC;
        hipLaunchKernelGGL(cumulative_grid_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, lv, i,
                           (float*)model->grid_cum[i], (__half*)model->grid_h[i]);
        int rc = check_launch("cumulative_grid_kernel");
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
