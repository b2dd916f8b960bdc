// Sphere-tracer state machine shared by the FP32 and mixed-precision
// persistent trace kernels.
#pragma once

#include <stdlib.h>

#include "coarse_lds.hpp"
#include "device_common.hpp"
#include "fp32_wave.hpp"

namespace nglod {

// Central differences with h = 1/192 and normalize(eps=1e-5), op by op as
// diffutils.py:61-70 + SphereTracer.py:129.
constexpr float kFdH = (float)(1.0 / (64.0 * 3.0));
constexpr float kFdTwoH = (float)((1.0 / (64.0 * 3.0)) * 2.0);

__device__ __forceinline__ void normalize_grad(float g0, float g1, float g2, float& n0, float& n1, float& n2) {
    float nn = sqrtf(g0 * g0 + g1 * g1 + g2 * g2);
    nn = nn < 1e-5f ? 1e-5f : nn;
    n0 = g0 / nn;
    n1 = g1 / nn;
    n2 = g2 / nn;
}

// ---------------------------------------------------------------------------
// Persistent sphere tracer.
//
// Each lane owns one ray at a time and runs SphereTracer.forward's per-ray
// recurrence as a small state machine; every loop trip of the wave is one SDF
// evaluation for every busy lane.  Finished lanes are refilled from a ray
// queue (one atomic per wave per refill, ballot + mbcnt to hand out ids).
// The frame is split into kQueues contiguous ray ranges, one per XCD
// (workgroups are dispatched to XCDs round-robin, so blockIdx % 8 picks the
// XCD-local range and its L2 sees one band of the image); a wave that drains
// its own range moves on to the next ones.
// Threads per block of the FP32 march: one block of 8
// waves per CU (the 128 KB of decoder and coarse-level tables in LDS leave room
// for one), sharing those tables.  Measured on the 1280x720 LOD-4 frame: 512
// threads 4.66 ms, 768 4.99, 384 5.10, 1024 5.36, 256 5.68 (more waves per CU
// evict each other's cells from the L1).
constexpr int kMarchBlock = 512;
constexpr int kQueues = 8;  // a multiple of 8 (one share per XCD)
static_assert(kQueues % 8 == 0 && kQueues <= 256, "kQueues: a multiple of 8, at most 256");
// Stride (in words) of the single-request traces' claim counters: one
// 128-byte line per counter.  Packed into one line, the counters of all eight
// XCDs' ranges made every wave's claim atomic contend for that line.
constexpr int kQueueStride = 32;

// PH_REPLAY: the mixed-precision march stopped on a surface test; the kernel
// rewinds the last steps and hands them to the FP32 pass.
enum : int { PH_IDLE = -1, PH_INIT = 0, PH_MARCH = 1, PH_N0 = 2, PH_N5 = 7, PH_REPLAY = 9 };

// State of a ray handed from the mixed-precision march to the FP32 pass:
// either a pending evaluation at x (phase INIT or MARCH, re-evaluated in
// FP32) or the normal taps of a finished hit (phase N0).  48 bytes.
struct Handoff {
    int32_t ray;
    int32_t phase;
    int32_t it;
    int32_t steps;
    float t, D, Dp;
    float x0, x1, x2;
    int32_t hitflag;
    int32_t pad;
};

struct TraceArgs {
    ModelArgs m;
    int lod, dec, num_steps, normals;
    float min_dis, min_dis3, step, far_clip;
    const float* ray_o;
    const float* ray_d;
    float* x;
    float* t;
    uint8_t* hit;
    float* normal;
    uint32_t* queue;
    unsigned long long* evals;  // [0] FP32-form evaluations, [1] mixed-form evaluations
    int32_t* ray_steps;
    CoarsePlan coarse; // leading levels the march kernels stage in LDS
    // Mixed form.  Records passed between the passes are indexed by ray
    // (record[ray], flag[ray] = 1 when written): no shared append counter, and
    // a pass that reads them claims consecutive ray ids.
    int replay_steps;  // at most this many evaluations rewound and re-evaluated in FP32
    float near_dist;   // ... back to the last step of at least this length (0: all replay_steps)
    struct Handoff* handoff;  // rays handed from the mixed march to the FP32 pass
    uint8_t* hflag;
    uint32_t* handoff_queue;  // [kQueues * queue_stride] claim counters of the FP32 pass
    struct Handoff* ret;      // rays handed back from the FP32 pass to the mixed march
    uint8_t* rflag;
    uint32_t* ret_queue;      // [kQueues * queue_stride] claim counters of a pass 1 that resumes them
    int queue_stride;         // words between the claim counters of the kQueues ranges
    float return_dist;        // FP32 step length above which a ray goes back to mixed
    int defer_normals;        // 0: a finished hit takes its normal taps in-lane; 1: it is handed to the FP32
                              // pass (hand-off record in phase PH_N0), which takes the taps
    int solo_lanes;           // FP32 passes: a wave with no rays left to claim and at most this many busy
                              // lanes evaluates them one at a time with the whole wave (sdf_fp32_solo);
                              // the mixed passes' cooperative tails use min(this, 4), pass 1's 8 unless 0
    int64_t nrays;  // the frame is split into kQueues ranges [nrays*q/kQueues, nrays*(q+1)/kQueues)
};

// Default of TraceArgs::solo_lanes; NGLOD_TRACE_SOLO_LANES (0..64) overrides it.
constexpr int kSoloLanesDefault = 4;

// TraceArgs::solo_lanes: NGLOD_TRACE_SOLO_LANES (0 = off .. 64) or the default.
inline int solo_lanes_setting() {
    const char* e = getenv("NGLOD_TRACE_SOLO_LANES");
    if (!e || !*e) return kSoloLanesDefault;
    const int v = atoi(e);
    return v < 0 ? 0 : (v > 64 ? 64 : v);
}

struct RayState {
    int64_t ray;
    int phase;
    int it;
    int steps;  // in-loop evaluations of this ray
    bool hitflag;
    float ox, oy, oz, dx, dy, dz;
    float x0, x1, x2;
    float t, D, Dp;
    float qx, qy, qz;  // next query point
    float fplus, g0, g1, g2;

    __device__ __forceinline__ void retire(const TraceArgs& a, bool hit, float n0, float n1, float n2) {
        a.x[3 * ray] = x0;
        a.x[3 * ray + 1] = x1;
        a.x[3 * ray + 2] = x2;
        a.t[ray] = t;
        a.hit[ray] = hit ? 1 : 0;
        a.normal[3 * ray] = n0;
        a.normal[3 * ray + 1] = n1;
        a.normal[3 * ray + 2] = n2;
        if (a.ray_steps) a.ray_steps[ray] = steps;
        phase = PH_IDLE;
        ray = -1;
    }

    // After the march: hit &= all(|x| <= 1) (SphereTracer.py:119), then
    // start the six normal taps (in-lane, or in the FP32 pass the hit is
    // handed to) or write the ray out.
    __device__ __forceinline__ void finish(const TraceArgs& a) {
        const bool outside = (fabsf(x0) > 1.0f) || (fabsf(x1) > 1.0f) || (fabsf(x2) > 1.0f);
        const bool hit = hitflag && !outside;
        if (hit && a.normals && a.defer_normals) {
            phase = PH_N0;
            hand_off_to(a.handoff, a.hflag);
        } else if (hit && a.normals) {
            phase = PH_N0;
            qx = x0 + kFdH;
            qy = x1 + 0.0f;
            qz = x2 + 0.0f;
        } else {
            retire(a, hit, 0.0f, 0.0f, 0.0f);
        }
    }

    // Start of loop iteration `it` (SphereTracer.py:84-102).  With `replay`
    // (the mixed-precision march) a surface stop becomes PH_REPLAY instead of
    // finishing the ray.
    __device__ __forceinline__ void check(const TraceArgs& a, bool replay) {
        if (it >= a.num_steps) {
            finish(a);
            return;
        }
        hitflag = fabsf(t) < a.far_clip;
        const bool cond = (fabsf(D) > a.min_dis) && (fabsf((D + Dp) * 0.5f) > a.min_dis3) && hitflag;
        if (!cond) {
            if (replay && hitflag) {
                phase = PH_REPLAY;
                return;
            }
            finish(a);
            return;
        }
        x0 = ox + dx * t;
        x1 = oy + dy * t;
        x2 = oz + dz * t;
        qx = x0;
        qy = x1;
        qz = x2;
        phase = PH_MARCH;
    }

    __device__ __forceinline__ void start(const TraceArgs& a, int64_t r) {
        ray = r;
        ox = a.ray_o[3 * r];
        oy = a.ray_o[3 * r + 1];
        oz = a.ray_o[3 * r + 2];
        dx = a.ray_d[3 * r];
        dy = a.ray_d[3 * r + 1];
        dz = a.ray_d[3 * r + 2];
        const BoxEntry e = box_entry(ox, oy, oz, dx, dy, dz);
        x0 = e.x;
        x1 = e.y;
        x2 = e.z;
        t = e.t;
        it = 0;
        steps = 0;
        hitflag = false;
        if (e.hit) {
            phase = PH_INIT;
            qx = x0;
            qy = x1;
            qz = x2;
        } else {
            // cond is false from the start: the first loop trip only sets
            // hit = |t| < far and stops (SphereTracer.py:84, :93-99).
            hitflag = (a.num_steps >= 1) && (fabsf(t) < a.far_clip);
            finish(a);
        }
    }

    // Pass 2 of the mixed form: continue a ray from its hand-off record.  A
    // record whose ray index is out of range (never the case for a written
    // record) leaves the lane idle instead of indexing outside the arrays.
    __device__ __forceinline__ void resume(const TraceArgs& a, const Handoff& h) {
        if (h.ray < 0 || (int64_t)h.ray >= a.nrays) {
            phase = PH_IDLE;
            ray = -1;
            return;
        }
        ray = h.ray;
        ox = a.ray_o[3 * ray];
        oy = a.ray_o[3 * ray + 1];
        oz = a.ray_o[3 * ray + 2];
        dx = a.ray_d[3 * ray];
        dy = a.ray_d[3 * ray + 1];
        dz = a.ray_d[3 * ray + 2];
        phase = h.phase;
        it = h.it;
        steps = h.steps;
        t = h.t;
        D = h.D;
        Dp = h.Dp;
        x0 = h.x0;
        x1 = h.x1;
        x2 = h.x2;
        hitflag = h.hitflag != 0;
        if (phase == PH_N0) {
            qx = x0 + kFdH;
            qy = x1 + 0.0f;
            qz = x2 + 0.0f;
        } else {
            qx = x0;
            qy = x1;
            qz = x2;
        }
    }

    // Write this ray's record at list[ray], set flag[ray] and free the lane.
    __device__ __forceinline__ void hand_off_to(Handoff* list, uint8_t* flag) {
        Handoff h;
        h.ray = (int32_t)ray;
        h.phase = phase;
        h.it = it;
        h.steps = steps;
        h.t = t;
        h.D = D;
        h.Dp = Dp;
        h.x0 = x0;
        h.x1 = x1;
        h.x2 = x2;
        h.hitflag = hitflag ? 1 : 0;
        h.pad = 0;
        list[ray] = h;
        flag[ray] = 1;
        phase = PH_IDLE;
        ray = -1;
    }

    __device__ __forceinline__ void consume(const TraceArgs& a, float d, bool replay = false) {
        if (phase == PH_INIT) {
            D = d;  // d = net(x) is not scaled by step_size (SphereTracer.py:63)
            Dp = D;
            it = 0;
            check(a, replay);
        } else if (phase == PH_MARCH) {
            Dp = D;
            D = d * a.step;
            t = t + D;
            ++it;
            ++steps;
            check(a, replay);
        } else if (phase < PH_N0 || phase > PH_N5) {
            // not an evaluating phase: end the ray rather than loop on it
            finish(a);
        } else {
            const int k = phase - PH_N0;  // 0..5: axis k/2, + tap then - tap
            const int axis = k >> 1;
            if ((k & 1) == 0) {
                fplus = d;
                qx = axis == 0 ? x0 - kFdH : x0 - 0.0f;
                qy = axis == 1 ? x1 - kFdH : x1 - 0.0f;
                qz = axis == 2 ? x2 - kFdH : x2 - 0.0f;
                ++phase;
            } else {
                const float g = (fplus - d) / kFdTwoH;
                if (axis == 0) g0 = g;
                else if (axis == 1) g1 = g;
                else g2 = g;
                if (axis == 2) {
                    float n0, n1, n2;
                    normalize_grad(g0, g1, g2, n0, n1, n2);
                    retire(a, true, n0, n1, n2);
                } else {
                    const int na = axis + 1;
                    qx = na == 0 ? x0 + kFdH : x0 + 0.0f;
                    qy = na == 1 ? x1 + kFdH : x1 + 0.0f;
                    qz = na == 2 ? x2 + kFdH : x2 + 0.0f;
                    ++phase;
                }
            }
        }
    }
};

// Every lane claims the next id in [0, total) whose flag is set (flag2, when
// not NULL: or whose flag2 is set), from one counter (one atomic per wave per round, as many ids as lanes still
// looking); -1 once the ids run out (`drained`, wave-uniform).
__device__ __forceinline__ int64_t claim_hits(uint32_t* counter, const uint8_t* flag, uint32_t total, uint32_t lane,
                                              bool& drained, const uint8_t* flag2 = nullptr) {
    int64_t id = -1;
    for (;;) {
        const bool want = id < 0 && !drained;
        const uint64_t mask = __ballot(want);
        if (mask == 0) break;
        const int cnt = __popcll(mask);
        const int leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)cnt);
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        if (want) {
            const uint32_t idx = base + mask_rank(mask);
            if (idx < total && (flag[idx] || (flag2 && flag2[idx]))) id = idx;
        }
        if (base + (uint32_t)cnt >= total) drained = true;
    }
    return id;
}

// The tail of a persistent FP32 pass: a wave with nothing left to claim and
// at most TraceArgs::solo_lanes busy lanes leaves its main loop (its busy
// count can only fall from there) and evaluates the remaining queries one at
// a time with the whole wave (sdf_fp32_solo), the same bits as the wave-wide
// path.  A separate loop after the main one, so that its registers do not
// weigh on the main loop's (inlined into the main loop, the unused solo code
// made the pass 25% slower).
template <bool POSINV>
__device__ __forceinline__ float fp32_solo_step(const TraceArgs& a, const Fp32Lds* lds, const CoarseLds32* coarse,
                                                CoarsePlan cp, const SoloLevels* lv, float* sv, float b1,
                                                uint32_t lane, uint64_t act, float qx, float qy, float qz) {
    float d = 0.0f;
    uint64_t mm = act;
    while (mm) {
        const int r = __ffsll((unsigned long long)mm) - 1;
        mm &= mm - 1ull;
        // broadcast through a lane shuffle: the address arithmetic of the
        // solo gather then stays in vector registers
        const float px = __shfl(qx, r), py = __shfl(qy, r), pz = __shfl(qz, r);
        const float dr = sdf_fp32_solo<POSINV>(lv, a.lod, lds, b1, lane, sv, px, py, pz, cp, coarse);
        if ((int)lane == r) d = dr;
    }
    return d;
}

// Hands ray ids to the lanes of one wave.  The frame is split into kQueues
// contiguous ranges with one claim counter each (counters[0..kQueues-1]);
// blockIdx % kQueues picks the wave's own range first (one per XCD), and a
// wave that drains it moves on to the next.  A wave claims CHUNK consecutive
// ids per atomic (CHUNK = 0: as many as it has lanes asking) and hands them
// out in lane order, so its rays stay neighbours in the frame.  With a flag
// array, ids whose flag is 0 are skipped (the passes over ray-indexed records).

struct RayFeed {
    int64_t nrays;
    int q0, q, tried;
    int stride;      // words between the ranges' claim counters
    bool drained;    // no ids left to claim (wave-uniform)
    int64_t cb, ce;  // the wave's unused ids of its current chunk (wave-uniform)
    uint64_t fmask;  // next_flagged: flagged ids cb + bit not yet handed out (wave-uniform)

    __device__ __forceinline__ void init(int64_t n, int counter_stride = 1) {
        nrays = n;
        stride = counter_stride;
        q0 = q = blockIdx.x % kQueues;
        tried = 0;
        drained = false;
        cb = ce = 0;
        fmask = 0;
    }

    __device__ __forceinline__ bool empty() const { return drained && cb >= ce; }

    // The next range to claim from: first the other ranges of this XCD
    // (q0 + 8j), then those of the next XCDs; each range once.
    __device__ __forceinline__ void advance() {
        constexpr int m = kQueues / 8;
        ++tried;
        q = (q0 + 8 * (tried % m) + tried / m) % kQueues;
    }

    // Claims the next chunk of `take` ids into [cb, ce); false if this claim
    // found the range empty (the caller tries again unless drained).
    __device__ __forceinline__ int64_t range_begin(int qi) const { return nrays * qi / kQueues; }

    template <bool SPLIT>
    __device__ __forceinline__ bool claim(uint32_t* counters, uint32_t lane, int leader, int take) {
        const int qi = SPLIT ? q : 0;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&counters[qi * stride], (uint32_t)take);
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        const int64_t lo = SPLIT ? range_begin(qi) : 0;
        const int64_t qsize = (SPLIT ? range_begin(qi + 1) : nrays) - lo;
        const bool got = (int64_t)base < qsize;
        if (got) {
            cb = lo + base;
            ce = lo + ((int64_t)base + take < qsize ? (int64_t)base + take : qsize);
        }
        if ((int64_t)base + take >= qsize) {
            advance();
            if (!SPLIT || tried == kQueues) drained = true;
        }
        return got;
    }

    // The next id whose flag is set, for each lane with `need` (-1: none
    // left).  The wave claims 64 ids at a time, loads their 64 flags in one
    // coalesced load (a flag per lane) and hands the flagged ids out in order
    // from the ballot mask.
    template <bool SPLIT>
    __device__ __forceinline__ int64_t next_flagged(uint32_t* counters, const uint8_t* flag, uint32_t lane,
                                                    bool need) {
        int64_t id = -1;
        for (;;) {
            const bool want = need && id < 0;
            const uint64_t mask = __ballot(want);
            if (mask == 0) break;
            if (fmask == 0) {
                if (drained) break;
                const int leader = __ffsll((unsigned long long)mask) - 1;
                if (!claim<SPLIT>(counters, lane, leader, 64)) continue;
                const int64_t c = cb + (int64_t)lane;
                fmask = __ballot(c < ce && flag[c] != 0);
                if (fmask == 0) continue;
            }
            const int avail = __popcll(fmask);
            const int cnt = __popcll(mask);
            if (want) {
                const int r = (int)mask_rank(mask);
                if (r < avail) id = cb + nth_set_bit(fmask, r);
            }
            if (cnt >= avail)
                fmask = 0;
            else
                fmask &= ~((1ull << nth_set_bit(fmask, cnt)) - 1ull);
        }
        return id;
    }

    // The next id for each lane with `need` (-1: none left).  SPLIT = false:
    // the whole frame is one range with counter counters[0].
    template <int CHUNK, bool SPLIT = true>
    __device__ __forceinline__ int64_t next(uint32_t* counters, const uint8_t* flag, uint32_t lane, bool need) {
        int64_t id = -1;
        for (;;) {
            const bool want = need && id < 0 && !empty();
            const uint64_t mask = __ballot(want);
            if (mask == 0) break;
            const int cnt = __popcll(mask);
            if (cb >= ce) {
                const int take = CHUNK > 0 ? CHUNK : cnt;
                const int leader = __ffsll((unsigned long long)mask) - 1;
                uint32_t base = 0;
                const int qi = SPLIT ? q : 0;
                if ((int)lane == leader) base = atomicAdd(&counters[qi * stride], (uint32_t)take);
                base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
                const int64_t lo = SPLIT ? range_begin(qi) : 0;
                const int64_t qsize = (SPLIT ? range_begin(qi + 1) : nrays) - lo;
                if ((int64_t)base < qsize) {
                    cb = lo + base;
                    ce = lo + ((int64_t)base + take < qsize ? (int64_t)base + take : qsize);
                }
                if ((int64_t)base + take >= qsize) {
                    advance();
                    if (!SPLIT || tried == kQueues) drained = true;
                }
                if (cb >= ce) continue;
            }
            if (want) {
                const int64_t c = cb + (int64_t)mask_rank(mask);
                if (c < ce && (!flag || flag[c])) id = c;
            }
            cb = cb + cnt < ce ? cb + cnt : ce;
        }
        return id;
    }

    // Start a ray in every idle lane, until each lane is busy or the rays run
    // out.  A ray that misses the box, or starts inside it, finishes at once
    // and its lane takes the next one.
    template <int CHUNK>
    __device__ __forceinline__ void refill(RayState& s, const TraceArgs& a, uint32_t lane) {
        if (__ballot(s.phase == PH_IDLE && !empty()) == 0) return;
        for (;;) {
            const bool need = s.phase == PH_IDLE;
            if (__ballot(need && !empty()) == 0) break;
            const int64_t id = next<CHUNK>(a.queue, nullptr, lane, need);
            if (id >= 0) s.start(a, id);
        }
    }
};

}  // namespace nglod
