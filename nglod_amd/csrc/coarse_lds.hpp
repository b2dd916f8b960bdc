// The leading coarse levels of the FP32 feature grids, staged once per
// workgroup in LDS, so that the gather of those levels costs LDS reads instead
// of vector-memory (TA/TD) work.  At the standard resolutions LODs 0 and 1
// (r = 4, 8) hold 125 + 729 = 854 corners: 109 KB.  Used by the FP32 march and
// the AO shading pass, where it measured faster; the mixed form's passes
// gather one cumulative grid from global memory.
//
// Layout: corner `idx` (level 0 first, then level 1) is 8 16-byte chunks;
// logical chunk k is stored at chunk k ^ (idx & 7), which spreads the lanes of
// a ds_read_b128 that read different corners over all 32 banks (a corner is
// 128 bytes, so without it every lane reading chunk k of any corner would hit
// the same 4 banks).  Lanes reading the same corner read the same address and
// are served by broadcast.
#pragma once

#include "device_common.hpp"

namespace nglod {

constexpr int kCoarseCorners = 854;

struct CoarseLds32 {
    uint4 c[kCoarseCorners * 8];
};

__device__ __forceinline__ int swz32(int idx, int k) { return idx * 8 + (k ^ (idx & 7)); }

// Number of leading levels (at most 2, none beyond lod) whose corners fit the
// LDS budget, and the first corner of each.
struct CoarsePlan {
    int levels;
    int off1;  // first corner of level 1
};

inline CoarsePlan coarse_plan(const ModelArgs& m, int lod) {
    CoarsePlan p{0, 0};
    int cum = 0;
    for (int l = 0; l < 2 && l <= lod; ++l) {
        const int s = m.res[l] + 1;
        const int n = s * s * s;
        if (cum + n > kCoarseCorners) break;
        if (l == 1) p.off1 = cum;
        cum += n;
        p.levels = l + 1;
    }
    return p;
}

// Stage the planned levels.  Every thread of the block must call this,
// followed by __syncthreads().
__device__ __forceinline__ void stage_coarse32(const ModelArgs& m, CoarsePlan p, CoarseLds32* s) {
    if (p.levels == 0) return;
    const int s0 = m.res[0] + 1;
    const int n0 = s0 * s0 * s0;
    int total = n0;
    if (p.levels > 1) {
        const int s1 = m.res[1] + 1;
        total += s1 * s1 * s1;
    }
    for (int j = threadIdx.x; j < total * 8; j += blockDim.x) {
        const int idx = j >> 3, k = j & 7;
        const int l = idx < n0 ? 0 : 1;
        const int local = l == 0 ? idx : idx - p.off1;
        s->c[swz32(idx, k)] = reinterpret_cast<const uint4*>(m.grid[l] + (size_t)local * kC)[k];
    }
}

__device__ __forceinline__ void corner_fma_lds(float (&f)[kC], const CoarseLds32* s, int idx, float w) {
#pragma unroll
    for (int k = 0; k < kC / 4; ++k) {
        const float4 g = __builtin_bit_cast(float4, s->c[swz32(idx, k)]);
        f[4 * k + 0] = __builtin_fmaf(g.x, w, f[4 * k + 0]);
        f[4 * k + 1] = __builtin_fmaf(g.y, w, f[4 * k + 1]);
        f[4 * k + 2] = __builtin_fmaf(g.z, w, f[4 * k + 2]);
        f[4 * k + 3] = __builtin_fmaf(g.w, w, f[4 * k + 3]);
    }
}

// gather_fp32's accumulation onto f (levels 0..lod) with levels < p.levels
// read from LDS: the same corners, weights and FMA order, so the result is
// bit-identical.
__device__ __forceinline__ void gather_fp32_coarse(const ModelArgs& m, int lod, CoarsePlan p, const CoarseLds32* s,
                                                   float px, float py, float pz, float (&f)[kC]) {
    for (int l = 0; l <= lod; ++l) {
        const int r = m.res[l];
        const int S = r + 1;
        const AxisW ax = axis_weights(px, r);
        const AxisW ay = axis_weights(py, r);
        const AxisW az = axis_weights(pz, r);
        const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
        const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
        if (l < p.levels) {
            const int i0 = (l == 0 ? 0 : p.off1) + (az.i0 * S + ay.i0) * S + ax.i0;
            const int sy = S, sz = S * S;
            corner_fma_lds(f, s, i0, wxy00 * az.w0);
            corner_fma_lds(f, s, i0 + 1, wxy10 * az.w0);
            corner_fma_lds(f, s, i0 + sy, wxy01 * az.w0);
            corner_fma_lds(f, s, i0 + sy + 1, wxy11 * az.w0);
            corner_fma_lds(f, s, i0 + sz, wxy00 * az.w1);
            corner_fma_lds(f, s, i0 + sz + 1, wxy10 * az.w1);
            corner_fma_lds(f, s, i0 + sz + sy, wxy01 * az.w1);
            corner_fma_lds(f, s, i0 + sz + sy + 1, wxy11 * az.w1);
        } else {
            // global levels: the corner loads in two batches of four corners
            // ahead of their FMAs (as gather_fp32_add; same order)
            const float* base = m.grid[l] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC;
            const size_t sx = kC, sy = (size_t)S * kC, sz = (size_t)S * S * kC;
            const float* cp[8] = {base, base + sx, base + sy, base + sy + sx,
                                  base + sz, base + sz + sx, base + sz + sy, base + sz + sy + sx};
            const float w[8] = {wxy00 * az.w0, wxy10 * az.w0, wxy01 * az.w0, wxy11 * az.w0,
                                wxy00 * az.w1, wxy10 * az.w1, wxy01 * az.w1, wxy11 * az.w1};
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                float4 g[4][kC / 4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int v = 0; v < kC / 4; ++v) g[c][v] = reinterpret_cast<const float4*>(cp[4 * b + c])[v];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
#pragma unroll
                    for (int v = 0; v < kC / 4; ++v) {
                        f[4 * v + 0] = __builtin_fmaf(g[c][v].x, w[4 * b + c], f[4 * v + 0]);
                        f[4 * v + 1] = __builtin_fmaf(g[c][v].y, w[4 * b + c], f[4 * v + 1]);
                        f[4 * v + 2] = __builtin_fmaf(g[c][v].z, w[4 * b + c], f[4 * v + 2]);
                        f[4 * v + 3] = __builtin_fmaf(g[c][v].w, w[4 * b + c], f[4 * v + 3]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
}

}  // namespace nglod
