// Device-side building blocks shared by the gfx950 kernels of libnglod_hip.so:
// model argument block, box entry, FP32 trilinear gather and decoder.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nglod_hip.h"

namespace nglod {

constexpr int kC = 32;   // feature_dim the kernels are built for
constexpr int kH = 128;  // hidden_dim the kernels are built for

// Kernel-argument copy of nglod_model_t (passed by value, lives in SGPRs /
// the kernarg segment).
struct ModelArgs {
    int num_lods;
    int pos_inv;
    int num_dec;
    int res[NGLOD_MAX_LODS];
    const float* grid[NGLOD_MAX_LODS];
    const float* w0[NGLOD_MAX_LODS];
    const float* b0[NGLOD_MAX_LODS];
    const float* w1[NGLOD_MAX_LODS];
    const float* b1[NGLOD_MAX_LODS];
    const uint16_t* grid_h[NGLOD_MAX_LODS];
    const uint16_t* w0_bf16[NGLOD_MAX_LODS];
};

#define NGLOD_CONST __attribute__((address_space(4)))

// Wave-uniform read-only load through the constant address space, so the
// compiler issues scalar (SMEM) loads and feeds VALU ops an SGPR operand.
__device__ __forceinline__ float ldc(const float* p, int i) {
    return ((const NGLOD_CONST float*)p)[i];
}

// ---------------------------------------------------------------------------
// Box entry, sol_nglod.aabb arithmetic (sdf-net/lib/extensions/sol_nglod/
// sol_nglod_kernel.cu:84-156): reciprocal in double, explicit fmaf exit tests,
// x/y/z face priority, entries at or beyond 500 dropped, origins inside the
// unit box not entered.  Everything else rounds op by op (-ffp-contract=off).
struct BoxEntry {
    float x, y, z, t;
    bool hit;
};

__device__ __forceinline__ BoxEntry box_entry(float ox, float oy, float oz,
                                              float dx, float dy, float dz) {
    BoxEntry e{ox, oy, oz, 0.0f, false};
    const float ivx = (float)(1.0 / (double)dx);
    const float ivy = (float)(1.0 / (double)dy);
    const float ivz = (float)(1.0 / (double)dz);
    const float s0 = __builtin_signbit(dx) ? 1.0f : -1.0f;
    const float s1 = __builtin_signbit(dy) ? 1.0f : -1.0f;
    const float s2 = __builtin_signbit(dz) ? 1.0f : -1.0f;
    const float cmax = fmaxf(fmaxf(fabsf(ox), fabsf(oy)), fabsf(oz));
    const float r = 1.0f;
    const float winding = cmax < r ? -1.0f : 1.0f;
    if (winding < 0.0f) return e;
    const float d0 = (r * winding * s0 - ox) * ivx;
    const float d1 = (r * winding * s1 - oy) * ivy;
    const float d2 = (r * winding * s2 - oz) * ivz;
    const float ltxy = __builtin_fmaf(dy, d0, oy);
    const float ltxz = __builtin_fmaf(dz, d0, oz);
    const float ltyx = __builtin_fmaf(dx, d1, ox);
    const float ltyz = __builtin_fmaf(dz, d1, oz);
    const float ltzx = __builtin_fmaf(dx, d2, ox);
    const float ltzy = __builtin_fmaf(dy, d2, oy);
    const bool t0 = (d0 >= 0.0f) && (fabsf(ltxy) < r) && (fabsf(ltxz) < r);
    const bool t1 = (d1 >= 0.0f) && (fabsf(ltyx) < r) && (fabsf(ltyz) < r);
    const bool t2 = (d2 >= 0.0f) && (fabsf(ltzx) < r) && (fabsf(ltzy) < r);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (t0) sx = s0;
    else if (t1) sy = s1;
    else if (t2) sz = s2;
    float d = 0.0f;
    if (sx != 0.0f) d = d0;
    else if (sy != 0.0f) d = d1;
    else if (sz != 0.0f) d = d2;
    const bool h = (sx != 0.0f) | (sy != 0.0f) | (sz != 0.0f);
    if (h && d < 500.0f) {
        e.hit = true;
        e.t = d;
        e.x = __builtin_fmaf(dx, d, ox);
        e.y = __builtin_fmaf(dy, d, oy);
        e.z = __builtin_fmaf(dz, d, oz);
    }
    return e;
}

// ---------------------------------------------------------------------------
// Trilinear coordinates of one axis for a grid of resolution r (r+1 corners):
// u = ((p + 1) / 2) * r clipped to [0, r] (grid_sample align_corners=True,
// padding_mode='border', sdf-net/lib/models/OctreeSDF.py:50-51).  The lower
// corner is kept <= r-1 so all 8 corners are in bounds; at u == r that gives
// weights (0, 1) on corners (r-1, r), the same value as the reference's
// (1, 0) on (r, r+1-skipped).  Any float input (NaN, Inf) lands in [0, r].
struct AxisW {
    int i0;
    float w0, w1;
};

__device__ __forceinline__ AxisW axis_weights(float p, int r) {
    const float rf = (float)r;
    float u = ((p + 1.0f) * 0.5f) * rf;
    u = fminf(fmaxf(u, 0.0f), rf);
    int i = (int)floorf(u);
    i = i > r - 1 ? r - 1 : i;
    AxisW a;
    a.i0 = i;
    a.w1 = u - (float)i;
    a.w0 = (float)(i + 1) - u;
    return a;
}

// Accumulates w * grid[corner][0..31] into f (FP32 corner-major grid).
__device__ __forceinline__ void corner_fma(float (&f)[kC], const float* __restrict__ p, float w) {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int v = 0; v < kC / 4; ++v) {
        const float4 g = q[v];
        f[4 * v + 0] = __builtin_fmaf(g.x, w, f[4 * v + 0]);
        f[4 * v + 1] = __builtin_fmaf(g.y, w, f[4 * v + 1]);
        f[4 * v + 2] = __builtin_fmaf(g.z, w, f[4 * v + 2]);
        f[4 * v + 3] = __builtin_fmaf(g.w, w, f[4 * v + 3]);
    }
}

// Sum over levels 0..lod of the trilinear samples (OctreeSDF.py:102-110),
// FP32, features held in registers.
// LV: any table with res[] and grid[] per level (ModelArgs, or a copy of
// them in LDS).
template <typename LV>
__device__ __forceinline__ void gather_fp32(const LV& m, int lod, float px, float py,
                                            float pz, float (&f)[kC]) {
#pragma unroll
    for (int c = 0; c < kC; ++c) f[c] = 0.0f;
    for (int l = 0; l <= lod; ++l) {
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
}

// gather_fp32's accumulation of levels 0..lod onto f (NL levels when NL > 0
// fixes the count at compile time) with the corner loads issued in two
// batches of four corners (32 16-byte loads each) ahead of their FMAs, so
// that a wave waits about two memory latencies per level instead of one per
// handful of loads (a wave-wide load completes with its slowest lane, and on
// an L1 miss that is an L2 round trip).  Same corners, weights and FMA order
// as gather_fp32: the same bits.  For kernels with registers to spare (128
// VGPRs hold a batch).
template <int NL = 0, typename LV>
__device__ __forceinline__ void gather_fp32_add(const LV& m, int lod, float px, float py, float pz,
                                                float (&f)[kC]) {
    const int nl = NL > 0 ? NL : lod + 1;
    for (int l = 0; l < nl; ++l) {
        const int r = m.res[l];
        const int S = r + 1;
        const AxisW ax = axis_weights(px, r);
        const AxisW ay = axis_weights(py, r);
        const AxisW az = axis_weights(pz, r);
        const float* base = m.grid[l] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC;
        const size_t sx = kC, sy = (size_t)S * kC, sz = (size_t)S * S * kC;
        const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
        const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
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

// The mixed form's FP32 gather: one level, the cumulative grid of the LOD (a
// mixed-form ModelArgs holds only that level, gathered as lod 0), zero on
// inactive lanes, the accumulators cleared once for both cases.
template <typename LV>
__device__ __forceinline__ void gather_fp32_cum(const LV& m, bool active, float px, float py, float pz,
                                                float (&f)[kC]) {
#pragma unroll
    for (int c = 0; c < kC; ++c) f[c] = 0.0f;
    if (active) gather_fp32_add<1>(m, 0, px, py, pz, f);
}

// relu(e W0^T + b0) W1^T + b1 with e = [xyz, f] (POSINV=false) or f
// (OctreeSDF.py:79-88, :113-115).  Weights are wave-uniform: scalar loads.
template <bool POSINV>
__device__ __forceinline__ float decode_fp32(const ModelArgs& m, int dec, float px, float py,
                                             float pz, const float (&f)[kC]) {
    constexpr int D = POSINV ? kC : kC + 3;
    constexpr int OFF = POSINV ? 0 : 3;
    const float* W0 = m.w0[dec];
    const float* B0 = m.b0[dec];
    const float* W1 = m.w1[dec];
    float out = 0.0f;
#pragma unroll 2
    for (int j = 0; j < kH; ++j) {
        float acc = 0.0f;
        if (!POSINV) {
            acc = __builtin_fmaf(ldc(W0, j * D + 0), px, acc);
            acc = __builtin_fmaf(ldc(W0, j * D + 1), py, acc);
            acc = __builtin_fmaf(ldc(W0, j * D + 2), pz, acc);
        }
#pragma unroll
        for (int k = 0; k < kC; ++k) acc = __builtin_fmaf(ldc(W0, j * D + OFF + k), f[k], acc);
        const float h = fmaxf(acc + ldc(B0, j), 0.0f);
        out = __builtin_fmaf(ldc(W1, j), h, out);
    }
    return out + ldc(m.b1[dec], 0);
}

template <bool POSINV>
__device__ __forceinline__ float sdf_fp32(const ModelArgs& m, int lod, int dec, float px,
                                          float py, float pz) {
    float f[kC];
    gather_fp32(m, lod, px, py, pz, f);
    return decode_fp32<POSINV>(m, dec, px, py, pz, f);
}

// Wave-wide MFMA decoders take the 64 lanes' queries as two 32-column
// tiles.  The coordinates of the queries in this lane's column of tile 0
// (lanes 0..31's queries) and tile 1 (lanes 32..63's): one v_permlane32_swap
// each.
__device__ __forceinline__ void tile_coords(float px, float py, float pz, float& qx0, float& qy0, float& qz0,
                                            float& qx1, float& qy1, float& qz1) {
    const auto sx = __builtin_amdgcn_permlane32_swap(__float_as_uint(px), __float_as_uint(px), false, false);
    const auto sy = __builtin_amdgcn_permlane32_swap(__float_as_uint(py), __float_as_uint(py), false, false);
    const auto sz = __builtin_amdgcn_permlane32_swap(__float_as_uint(pz), __float_as_uint(pz), false, false);
    qx0 = __uint_as_float(sx[0]);
    qx1 = __uint_as_float(sx[1]);
    qy0 = __uint_as_float(sy[0]);
    qy1 = __uint_as_float(sy[1]);
    qz0 = __uint_as_float(sz[0]);
    qz1 = __uint_as_float(sz[1]);
}

// A lane's query result from the two half-sums of its tile: part0 (tile 0)
// on lanes 0..31 and part1 (tile 1) on lanes 32..63 hold one half each; the
// swap brings the other half, and the sum is the same whichever operand
// comes first.
__device__ __forceinline__ float halves_sum(float part0, float part1) {
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(part0), __float_as_uint(part1), false, false);
    return __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of set bits of `mask` below this lane.
// Position of the r-th (0-based) set bit of m; r < popcount(m).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const int c = __popcll((m >> pos) & ((1ull << w) - 1ull));
        if (r >= c) {
            r -= c;
            pos += w;
        }
    }
    return pos;
}

__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

}  // namespace nglod
