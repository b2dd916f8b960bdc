// Sphere-tracer state machine shared by the FP32 and mixed-precision
// persistent trace kernels.
#pragma once

#include "coarse_lds.hpp"
#include "device_common.hpp"

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
// Threads per block of the march kernels: one block of 12 waves per CU, all
// sharing the decoder and coarse-level tables in LDS.
constexpr int kMarchBlock = 768;
constexpr int kQueues = 8;

// PH_NQ: a finished hit whose normal is computed by a separate pass (the ray's
// x, t, hit are already written; it waits to be appended to the normal list).
// PH_REPLAY: the mixed-precision march stopped on a surface test; the kernel
// rewinds the last steps and hands them to the FP32 pass.
enum : int { PH_IDLE = -1, PH_INIT = 0, PH_MARCH = 1, PH_N0 = 2, PH_N5 = 7, PH_NQ = 8, PH_REPLAY = 9 };

// A finished hit waiting for its normal (mixed form).  16 bytes.
struct NormalItem {
    int32_t ray;
    float x0, x1, x2;
};

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
    // Mixed form: rays handed from the mixed march to the FP32 pass.
    int replay;        // 1 in the mixed march: a surface stop becomes PH_REPLAY
    CoarsePlan coarse; // leading levels the march kernels stage in LDS
    int replay_steps;  // at most this many evaluations rewound and re-evaluated in FP32
    float near_dist;   // ... back to the last step of at least this length (0: all replay_steps)
    struct Handoff* handoff;
    uint32_t* handoff_count;
    uint32_t* handoff_queue;
    struct Handoff* ret;     // rays handed back from FP32 to the mixed march
    uint32_t* ret_count;
    uint32_t* ret_queue;
    int from_returns;        // pass 1 refills from `ret` instead of the ray queues
    int allow_return;        // pass 2 may hand rays back (not in the last round)
    float return_dist;       // FP32 step length above which a ray goes back to mixed
    int defer_normals;  // 1: finished hits go to the normal list instead of in-lane taps
    NormalItem* nlist;
    uint32_t* ncount;
    int64_t qbeg[kQueues];
    int64_t qend[kQueues];
};

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
    // either start the six normal taps or write the ray out.
    __device__ __forceinline__ void finish(const TraceArgs& a) {
        const bool outside = (fabsf(x0) > 1.0f) || (fabsf(x1) > 1.0f) || (fabsf(x2) > 1.0f);
        const bool hit = hitflag && !outside;
        if (hit && a.normals && a.defer_normals) {
            a.x[3 * ray] = x0;
            a.x[3 * ray + 1] = x1;
            a.x[3 * ray + 2] = x2;
            a.t[ray] = t;
            a.hit[ray] = 1;
            if (a.ray_steps) a.ray_steps[ray] = steps;
            phase = PH_NQ;
        } else if (hit && a.normals) {
            phase = PH_N0;
            qx = x0 + kFdH;
            qy = x1 + 0.0f;
            qz = x2 + 0.0f;
        } else {
            retire(a, hit, 0.0f, 0.0f, 0.0f);
        }
    }

    // Start of loop iteration `it` (SphereTracer.py:84-102).
    __device__ __forceinline__ void check(const TraceArgs& a) {
        if (it >= a.num_steps) {
            finish(a);
            return;
        }
        hitflag = fabsf(t) < a.far_clip;
        const bool cond = (fabsf(D) > a.min_dis) && (fabsf((D + Dp) * 0.5f) > a.min_dis3) && hitflag;
        if (!cond) {
            if (a.replay && hitflag) {
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
        if (h.ray < 0 || (int64_t)h.ray >= a.qend[kQueues - 1]) {
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

    __device__ __forceinline__ void queue_normal(const TraceArgs& a, uint32_t slot) {
        NormalItem it_;
        it_.ray = (int32_t)ray;
        it_.x0 = x0;
        it_.x1 = x1;
        it_.x2 = x2;
        a.nlist[slot] = it_;
        phase = PH_IDLE;
        ray = -1;
    }

    __device__ __forceinline__ void hand_off(const TraceArgs& a, uint32_t slot) { hand_off_to(a.handoff, slot); }

    __device__ __forceinline__ void hand_off_to(Handoff* list, uint32_t slot) {
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
        list[slot] = h;
        phase = PH_IDLE;
        ray = -1;
    }

    __device__ __forceinline__ void consume(const TraceArgs& a, float d) {
        if (phase == PH_INIT) {
            D = d;  // d = net(x) is not scaled by step_size (SphereTracer.py:63)
            Dp = D;
            it = 0;
            check(a);
        } else if (phase == PH_MARCH) {
            Dp = D;
            D = d * a.step;
            t = t + D;
            ++it;
            ++steps;
            check(a);
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

}  // namespace nglod
