// Mixed-precision SDF evaluation for one wavefront (64 queries, one per lane):
// FP16 corner-major grids gathered with packed FP16 FMAs, then the first
// decoder layer as BF16 MFMA 32x32x16 tiles with FP32 accumulation, the
// second layer (128 -> 1) and the ReLU in FP32 on the MFMA output registers.
//
// MFMA operand layout (v_mfma_f32_32x32x16_bf16): lane l holds A[m = l%32][k in
// half l/32] and B[k in half l/32][n = l%32]; A and B use the same (half,
// element) -> k map, so any consistent channel order is valid.  Output: lane l
// holds C[m][n = l%32] for m = 8*(i/4) + 4*(l/32) + i%4, i = 0..15.
//
// A = W0 (hidden x 32 feature channels), B = features^T (32 x 32 queries).
// The 64 queries of a wave form two 32-query tiles; a lane's own features
// feed tile (l/32) through one lane-half swap.
#pragma once

#include <hip/hip_fp16.h>

#include "device_common.hpp"

namespace nglod {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

// Per-workgroup constants of the mixed decoder, staged once in LDS (11 KB):
// the W0 A fragments ([h tile 0..3][k block 0..1][lane], 8 bf16 each), b0, w1
// and the FP32 xyz columns of W0.
struct MixedLds {
    uint4 wfrag[4 * 2 * 64];
    float b0[kH];
    float w1[kH];
    float wxyz[kH][4];  // W0 columns 0..2 (xyz input), FP32; [3] unused
};

// Stage decoder `dec` into LDS.  Every thread of the block must call this,
// followed by __syncthreads().
__device__ __forceinline__ void stage_mixed_lds(const ModelArgs& m, int dec, MixedLds* s) {
    const int din = m.pos_inv ? kC : kC + 3;
    const uint4* frag = reinterpret_cast<const uint4*>(m.w0_bf16[dec]);
    for (int j = threadIdx.x; j < 4 * 2 * 64; j += blockDim.x) s->wfrag[j] = frag[j];
    for (int j = threadIdx.x; j < kH; j += blockDim.x) {
        s->b0[j] = m.b0[dec][j];
        s->w1[j] = m.w1[dec][j];
        if (!m.pos_inv) {
            s->wxyz[j][0] = m.w0[dec][j * din + 0];
            s->wxyz[j][1] = m.w0[dec][j * din + 1];
            s->wxyz[j][2] = m.w0[dec][j * din + 2];
        } else {
            s->wxyz[j][0] = s->wxyz[j][1] = s->wxyz[j][2] = 0.0f;
        }
        s->wxyz[j][3] = 0.0f;
    }
}

// The FP16 gather: adds to acc, for levels 0..lod (NL levels when NL > 0
// fixes the count at compile time; LV: any table with res[] and grid_h[] per
// level), the trilinear sample of the FP16 corner-major grid: per level the 8
// corners in the order 000, 100, 010, 110, 001, 101, 011, 111 (x fastest),
// each an __hfma2 of its 16 channel pairs with the FP16-rounded FP32 corner
// weight.  The corner loads are issued in two batches of four corners (16
// 16-byte loads each) ahead of their FMAs: about two memory latencies per
// level instead of one per handful of loads.
template <int NL = 0, typename LV>
__device__ __forceinline__ void gather_fp16_add(const LV& m, int lod, float px, float py, float pz,
                                                __half2 (&acc)[kC / 2]) {
    const int nl = NL > 0 ? NL : lod + 1;
    for (int l = 0; l < nl; ++l) {
        const int r = m.res[l];
        const int S = r + 1;
        const AxisW ax = axis_weights(px, r);
        const AxisW ay = axis_weights(py, r);
        const AxisW az = axis_weights(pz, r);
        const __half* base = reinterpret_cast<const __half*>(m.grid_h[l]) +
                             ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC;
        const size_t sx = kC, sy = (size_t)S * kC, sz = (size_t)S * S * kC;
        const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
        const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
        const __half* cp[8] = {base, base + sx, base + sy, base + sy + sx,
                               base + sz, base + sz + sx, base + sz + sy, base + sz + sy + sx};
        const float w[8] = {wxy00 * az.w0, wxy10 * az.w0, wxy01 * az.w0, wxy11 * az.w0,
                            wxy00 * az.w1, wxy10 * az.w1, wxy01 * az.w1, wxy11 * az.w1};
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            uint4 g[4][kC / 8];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v < kC / 8; ++v) g[c][v] = reinterpret_cast<const uint4*>(cp[4 * b + c])[v];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const __half2 wh = __float2half2_rn(w[4 * b + c]);
#pragma unroll
                for (int v = 0; v < kC / 8; ++v) {
                    acc[4 * v + 0] = __hfma2(__builtin_bit_cast(__half2, g[c][v].x), wh, acc[4 * v + 0]);
                    acc[4 * v + 1] = __hfma2(__builtin_bit_cast(__half2, g[c][v].y), wh, acc[4 * v + 1]);
                    acc[4 * v + 2] = __hfma2(__builtin_bit_cast(__half2, g[c][v].z), wh, acc[4 * v + 2]);
                    acc[4 * v + 3] = __hfma2(__builtin_bit_cast(__half2, g[c][v].w), wh, acc[4 * v + 3]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// The mixed form's FP16 gather: a mixed-form ModelArgs (make_model_args)
// holds one level, the cumulative grid of the LOD, which every mixed launch
// gathers as lod 0, so the level count is fixed at one; zero on inactive
// lanes.  The accumulators are cleared once for both cases (a clear on each
// side of the busy-lane branch costs both on every trip).
template <typename LV>
__device__ __forceinline__ void gather_fp16_cum(const LV& m, bool active, float px, float py, float pz,
                                                __half2 (&acc)[kC / 2]) {
#pragma unroll
    for (int c = 0; c < kC / 2; ++c) acc[c] = __float2half2_rn(0.0f);
    if (active) gather_fp16_add<1>(m, 0, px, py, pz, acc);
}

__device__ __forceinline__ uint32_t pack_bf16(__half2 h) {
    const float2 f = __half22float2(h);
    const bf16x2v v = {(__bf16)f.x, (__bf16)f.y};
    return __builtin_bit_cast(uint32_t, v);
}

// b0 of this lane's 16 rows of hidden tile h (rows 32h + 8g + 4 half + 0..3,
// g = 0..3): the first MFMA of the tile accumulates onto it, so the bias
// costs no vector instruction.  LDS: any staged decoder with b0[], w1[] and
// wxyz[][4] (MixedLds, SplitLds).
template <typename LDS>
__device__ __forceinline__ f32x16 bias_tile(int h, uint32_t half, const LDS* s) {
    f32x16 c;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const float4 b = *reinterpret_cast<const float4*>(&s->b0[32 * h + 8 * g + 4 * (int)half]);
        c[4 * g + 0] = b.x;
        c[4 * g + 1] = b.y;
        c[4 * g + 2] = b.z;
        c[4 * g + 3] = b.w;
    }
    return c;
}

// ReLU as an integer max on the float's bits: the sign bit makes every
// negative value (and -0) a negative integer, so max(bits, 0) keeps positive
// values and gives +0 otherwise.  One v_max_i32; fmaxf(v, 0) on an MFMA
// output costs a canonicalising v_max_f32 before the max itself.
__device__ __forceinline__ float relu_bits(float v) {
    const int b = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, b > 0 ? b : 0);
}

// Second layer over one 32x32 output tile held in registers (the first
// layer's output with b0 already in it): sum of
// w1[row] * relu(C + wxyz[row] . xyz) over this lane's 16 rows.
template <typename LDS>
__device__ __forceinline__ float layer2_partial(const f32x16& c, int h, uint32_t half, const LDS* s, bool xyz,
                                                float px, float py, float pz) {
    float acc = 0.0f;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int row0 = 32 * h + 8 * g + 4 * (int)half;
        const float4 w = *reinterpret_cast<const float4*>(&s->w1[row0]);
        float v0 = c[4 * g + 0], v1 = c[4 * g + 1], v2 = c[4 * g + 2], v3 = c[4 * g + 3];
        if (xyz) {
            const float4 q0 = *reinterpret_cast<const float4*>(s->wxyz[row0 + 0]);
            const float4 q1 = *reinterpret_cast<const float4*>(s->wxyz[row0 + 1]);
            const float4 q2 = *reinterpret_cast<const float4*>(s->wxyz[row0 + 2]);
            const float4 q3 = *reinterpret_cast<const float4*>(s->wxyz[row0 + 3]);
            v0 = __builtin_fmaf(q0.x, px, __builtin_fmaf(q0.y, py, __builtin_fmaf(q0.z, pz, v0)));
            v1 = __builtin_fmaf(q1.x, px, __builtin_fmaf(q1.y, py, __builtin_fmaf(q1.z, pz, v1)));
            v2 = __builtin_fmaf(q2.x, px, __builtin_fmaf(q2.y, py, __builtin_fmaf(q2.z, pz, v2)));
            v3 = __builtin_fmaf(q3.x, px, __builtin_fmaf(q3.y, py, __builtin_fmaf(q3.z, pz, v3)));
        }
        acc = __builtin_fmaf(w.x, relu_bits(v0), acc);
        acc = __builtin_fmaf(w.y, relu_bits(v1), acc);
        acc = __builtin_fmaf(w.z, relu_bits(v2), acc);
        acc = __builtin_fmaf(w.w, relu_bits(v3), acc);
    }
    return acc;
}

// SDF of the 64 lanes' query points.  Every lane of the wave must call this
// (no divergent branch around it): the MFMAs and lane swaps are wave-wide.
// Lanes without a query pass active = false: they load nothing, and a
// 32-query tile with no active lane skips its MFMAs.  Each output column
// depends only on its own query.
// The decoder on the 64 lanes' FP16 feature sums acc (zero on inactive
// lanes); wave-wide, as sdf_mixed_wave.
template <bool POSINV>
__device__ __forceinline__ float decode_mixed_wave(const __half2 (&acc)[kC / 2], const MixedLds* s, float b1,
                                                   uint32_t lane, bool active, float px, float py, float pz) {
    const uint64_t am = __ballot(active);
    const bool use0 = (uint32_t)am != 0u, use1 = (uint32_t)(am >> 32) != 0u;
    const uint32_t half = lane >> 5;
    // own channels 16kb + 0..7 (Y) and 16kb + 8..15 (X) as bf16; one
    // v_permlane32_swap per word gives tile 0 (own Y | X of lane l-32) and
    // tile 1 (Y of lane l+32 | own X)
    bf16x8 B0[2], B1[2];  // [kb] of tile 0 and tile 1
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        uint32_t t0[4], t1[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const auto sw = __builtin_amdgcn_permlane32_swap(pack_bf16(acc[8 * kb + j]),
                                                             pack_bf16(acc[8 * kb + 4 + j]), false, false);
            t0[j] = sw[0];
            t1[j] = sw[1];
        }
        B0[kb] = __builtin_bit_cast(bf16x8, make_uint4(t0[0], t0[1], t0[2], t0[3]));
        B1[kb] = __builtin_bit_cast(bf16x8, make_uint4(t1[0], t1[1], t1[2], t1[3]));
    }
    // query coordinates of each tile's column (tile t, column l%32 = lane l%32 + 32t)
    float qx0, qy0, qz0, qx1, qy1, qz1;
    tile_coords(px, py, pz, qx0, qy0, qz0, qx1, qy1, qz1);
    float part0 = 0.0f, part1 = 0.0f;
#pragma unroll 1
    for (int h = 0; h < 4; ++h) {
        const bf16x8 a0 = __builtin_bit_cast(bf16x8, s->wfrag[(h * 2 + 0) * 64 + lane]);
        const bf16x8 a1 = __builtin_bit_cast(bf16x8, s->wfrag[(h * 2 + 1) * 64 + lane]);
        if (use0) {
            f32x16 c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, B0[0], bias_tile(h, half, s), 0, 0, 0);
            c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, B0[1], c0, 0, 0, 0);
            part0 += layer2_partial(c0, h, half, s, !POSINV, qx0, qy0, qz0);
        }
        if (use1) {
            f32x16 c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, B1[0], bias_tile(h, half, s), 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, B1[1], c1, 0, 0, 0);
            part1 += layer2_partial(c1, h, half, s, !POSINV, qx1, qy1, qz1);
        }
    }
    return halves_sum(part0, part1) + b1;
}

// decode_mixed_wave for a wave with few busy lanes, where the decode's
// latency rather than its issue rate sets the trip time: a used tile's four
// hidden tiles are independent, so their LDS reads, MFMA pairs and
// second-layer chains are issued together instead of one hidden tile after
// another.  Per output the same instructions on the same operands, and the
// partial sums added in the same order: the same bits as decode_mixed_wave.
template <bool POSINV>
__device__ __forceinline__ float decode_mixed_wave_few(const __half2 (&acc)[kC / 2], const MixedLds* s, float b1,
                                                       uint32_t lane, bool active, float px, float py, float pz) {
    const uint64_t am = __ballot(active);
    const uint32_t half = lane >> 5;
    bf16x8 B[2][2];  // [tile][kb]
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        uint32_t t0[4], t1[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const auto sw = __builtin_amdgcn_permlane32_swap(pack_bf16(acc[8 * kb + j]),
                                                             pack_bf16(acc[8 * kb + 4 + j]), false, false);
            t0[j] = sw[0];
            t1[j] = sw[1];
        }
        B[0][kb] = __builtin_bit_cast(bf16x8, make_uint4(t0[0], t0[1], t0[2], t0[3]));
        B[1][kb] = __builtin_bit_cast(bf16x8, make_uint4(t1[0], t1[1], t1[2], t1[3]));
    }
    float qx0, qy0, qz0, qx1, qy1, qz1;
    tile_coords(px, py, pz, qx0, qy0, qz0, qx1, qy1, qz1);
    float part[2] = {0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        if ((t == 0 ? (uint32_t)am : (uint32_t)(am >> 32)) == 0u) continue;
        float p[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const bf16x8 a0 = __builtin_bit_cast(bf16x8, s->wfrag[(h * 2 + 0) * 64 + lane]);
            const bf16x8 a1 = __builtin_bit_cast(bf16x8, s->wfrag[(h * 2 + 1) * 64 + lane]);
            f32x16 c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, B[t][0], bias_tile(h, half, s), 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, B[t][1], c, 0, 0, 0);
            p[h] = layer2_partial(c, h, half, s, !POSINV, t == 0 ? qx0 : qx1, t == 0 ? qy0 : qy1,
                                  t == 0 ? qz0 : qz1);
        }
#pragma unroll
        for (int h = 0; h < 4; ++h) part[t] += p[h];
    }
    return halves_sum(part[0], part[1]) + b1;
}

template <bool POSINV, typename LV>
__device__ __forceinline__ float sdf_mixed_wave(const LV& m, const MixedLds* s, float b1, uint32_t lane,
                                                bool active, float px, float py, float pz) {
    __half2 acc[kC / 2];
    gather_fp16_cum(m, active, px, py, pz, acc);
    return decode_mixed_wave<POSINV>(acc, s, b1, lane, active, px, py, pz);
}

// FP16 grid pointers and resolutions, copied to LDS so that the cooperative
// gather reads them as vector values (the mixed march uses nearly all of its
// scalar registers).
struct MixedLevels {
    const __half* grid_h[NGLOD_MAX_LODS];
    int res[NGLOD_MAX_LODS];
};

// Every thread of the block calls this before the block's __syncthreads().
__device__ __forceinline__ void stage_mixed_levels(const ModelArgs& m, MixedLevels* lv) {
#pragma unroll
    for (int l = 0; l < NGLOD_MAX_LODS; ++l) {
        if ((int)threadIdx.x == l) {
            lv->grid_h[l] = reinterpret_cast<const __half*>(m.grid_h[l]);
            lv->res[l] = m.res[l];
        }
    }
}

// gather_fp16_cum for the busy lanes of a wave with at most 4 of them (act,
// a wave-uniform mask), done by 16 lanes per query: lane 16 i + j computes
// channel pair j of the i-th busy lane's query with the 8 corner loads of the
// cumulative grid (level 0) issued first, then the same __hfma2 chain in the
// same corner order; each busy lane then collects its 16 pairs.
// Bit-identical to gather_fp16_cum.  Every lane of the wave must call this.
__device__ __forceinline__ void gather_fp16_coop(const MixedLevels* lv, uint32_t lane, uint64_t act, float qx,
                                                 float qy, float qz, __half2 (&acc)[kC / 2]) {
    constexpr int lod = 0;
    const int i = (int)(lane >> 4), j = (int)(lane & 15);
    const int k = __popcll(act);
    const bool helper = i < k;
    const int src = helper ? nth_set_bit(act, i) : 0;
    const float px = __shfl(qx, src), py = __shfl(qy, src), pz = __shfl(qz, src);
    uint32_t g[lod + 1][8];
    if (helper) {
#pragma unroll
        for (int l = 0; l <= lod; ++l) {
            const int r = lv->res[l];
            const int S = r + 1;
            const AxisW ax = axis_weights(px, r);
            const AxisW ay = axis_weights(py, r);
            const AxisW az = axis_weights(pz, r);
            const uint32_t* base = reinterpret_cast<const uint32_t*>(
                lv->grid_h[l] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC) + j;
            const size_t sx = kC / 2, sy = (size_t)S * kC / 2, sz = (size_t)S * S * kC / 2;
            g[l][0] = base[0];
            g[l][1] = base[sx];
            g[l][2] = base[sy];
            g[l][3] = base[sy + sx];
            g[l][4] = base[sz];
            g[l][5] = base[sz + sx];
            g[l][6] = base[sz + sy];
            g[l][7] = base[sz + sy + sx];
        }
    }
    __half2 v = __float2half2_rn(0.0f);
    if (helper) {
#pragma unroll
        for (int l = 0; l <= lod; ++l) {
            const int r = lv->res[l];
            const AxisW ax = axis_weights(px, r);
            const AxisW ay = axis_weights(py, r);
            const AxisW az = axis_weights(pz, r);
            const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
            const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
            v = __hfma2(__builtin_bit_cast(__half2, g[l][0]), __float2half2_rn(wxy00 * az.w0), v);
            v = __hfma2(__builtin_bit_cast(__half2, g[l][1]), __float2half2_rn(wxy10 * az.w0), v);
            v = __hfma2(__builtin_bit_cast(__half2, g[l][2]), __float2half2_rn(wxy01 * az.w0), v);
            v = __hfma2(__builtin_bit_cast(__half2, g[l][3]), __float2half2_rn(wxy11 * az.w0), v);
            v = __hfma2(__builtin_bit_cast(__half2, g[l][4]), __float2half2_rn(wxy00 * az.w1), v);
            v = __hfma2(__builtin_bit_cast(__half2, g[l][5]), __float2half2_rn(wxy10 * az.w1), v);
            v = __hfma2(__builtin_bit_cast(__half2, g[l][6]), __float2half2_rn(wxy01 * az.w1), v);
            v = __hfma2(__builtin_bit_cast(__half2, g[l][7]), __float2half2_rn(wxy11 * az.w1), v);
        }
    }
    const bool mine = (act >> lane) & 1ull;
    const int q = mine ? (int)mask_rank(act) : 0;
    const uint32_t vb = __builtin_bit_cast(uint32_t, v);
#pragma unroll
    for (int c = 0; c < kC / 2; ++c) {
        const uint32_t w = (uint32_t)__shfl((int)vb, 16 * q + c);
        acc[c] = mine ? __builtin_bit_cast(__half2, w) : __float2half2_rn(0.0f);
    }
}

// gather_fp16_coop for up to 8 busy lanes: 8 lanes per query, each two
// channel pairs (lanes 8 i + j: pairs 2 j and 2 j + 1).  Same per-pair
// __hfma2 chain, the same bits.
__device__ __forceinline__ void gather_fp16_coop8(const MixedLevels* lv, uint32_t lane, uint64_t act, float qx,
                                                  float qy, float qz, __half2 (&acc)[kC / 2]) {
    const int i = (int)(lane >> 3), j = (int)(lane & 7);
    const int k = __popcll(act);
    const bool helper = i < k;
    const int src = helper ? nth_set_bit(act, i) : 0;
    const float px = __shfl(qx, src), py = __shfl(qy, src), pz = __shfl(qz, src);
    uint2 g[8];
    __half2 v0 = __float2half2_rn(0.0f), v1 = __float2half2_rn(0.0f);
    if (helper) {
        const int r = lv->res[0];
        const int S = r + 1;
        const AxisW ax = axis_weights(px, r);
        const AxisW ay = axis_weights(py, r);
        const AxisW az = axis_weights(pz, r);
        const uint2* base = reinterpret_cast<const uint2*>(
            lv->grid_h[0] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC) + j;
        const size_t sx = kC / 4, sy = (size_t)S * kC / 4, sz = (size_t)S * S * kC / 4;
        g[0] = base[0];
        g[1] = base[sx];
        g[2] = base[sy];
        g[3] = base[sy + sx];
        g[4] = base[sz];
        g[5] = base[sz + sx];
        g[6] = base[sz + sy];
        g[7] = base[sz + sy + sx];
        const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
        const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
        const float w[8] = {wxy00 * az.w0, wxy10 * az.w0, wxy01 * az.w0, wxy11 * az.w0,
                            wxy00 * az.w1, wxy10 * az.w1, wxy01 * az.w1, wxy11 * az.w1};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const __half2 wh = __float2half2_rn(w[c]);
            v0 = __hfma2(__builtin_bit_cast(__half2, g[c].x), wh, v0);
            v1 = __hfma2(__builtin_bit_cast(__half2, g[c].y), wh, v1);
        }
    }
    const bool mine = (act >> lane) & 1ull;
    const int q = mine ? (int)mask_rank(act) : 0;
    const uint32_t b0 = __builtin_bit_cast(uint32_t, v0), b1 = __builtin_bit_cast(uint32_t, v1);
#pragma unroll
    for (int c = 0; c < kC / 2; ++c) {
        const uint32_t w = (uint32_t)__shfl((int)((c & 1) ? b1 : b0), 8 * q + (c >> 1));
        acc[c] = mine ? __builtin_bit_cast(__half2, w) : __float2half2_rn(0.0f);
    }
}

}  // namespace nglod
