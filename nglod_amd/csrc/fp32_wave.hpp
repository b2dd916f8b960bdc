// FP32 SDF evaluation for one wavefront (64 queries, one per lane): the FP32
// trilinear gather per lane, then the first decoder layer as FP32 MFMA
// 32x32x2 tiles (v_mfma_f32_32x32x2_f32: exact FP32 products, FP32
// accumulation), then bias, ReLU and the 128 -> 1 layer in FP32 on the MFMA
// output registers.  Twice the FP32 rate of the vector FMA decoder, and it
// runs beside the vector ALU work of the other waves.
//
// MFMA operand layout (32x32x2, one block): lane l holds A[m = l%32][k = l/32]
// and B[k = l/32][n = l%32]; output: lane l holds C[m][n = l%32] for
// m = 8*(i/4) + 4*(l/32) + i%4, i = 0..15.  A = W0 (hidden x feature
// channels), B = features^T (channel pair x 32 queries).  The 64 queries of
// a wave form two 32-query tiles: tile 0 is lanes 0..31, tile 1 lanes 32..63;
// one v_permlane32_swap per channel pair turns a lane's own two channels into
// both tiles' B operands.
#pragma once

#include "coarse_lds.hpp"
#include "device_common.hpp"

namespace nglod {

typedef float f32x16_t __attribute__((ext_vector_type(16)));

// Per-workgroup constants of the FP32 MFMA decoder, staged once in LDS
// (18.5 KB): A fragments [h tile 0..3][channel-pair quad 0..3][lane] (4
// channel pairs each), b0, w1 and the xyz columns of W0.
struct Fp32Lds {
    float4 afrag[4 * 4 * 64];
    float b0[kH];
    float w1[kH];
    float wxyz[kH][4];  // W0 columns 0..2 (xyz input); [3] unused
};

// Stage a decoder (W0 [128][din], din = 32 or 35 with xyz first; b0; w1) into
// LDS.  Every thread of the block must call this, followed by __syncthreads().
__device__ __forceinline__ void stage_fp32_lds_raw(const float* W0, int din, const float* b0, const float* w1,
                                                   Fp32Lds* s) {
    const bool pos_inv = din == kC;
    const int off = pos_inv ? 0 : 3;
    for (int j = threadIdx.x; j < 4 * 4 * 64; j += blockDim.x) {
        const int lane = j & 63;
        const int kq = (j >> 6) & 3;
        const int h = j >> 8;
        const int row = 32 * h + (lane & 31);
        const int col = off + 8 * kq + (lane >> 5);  // channel 2*(4kq+e) + lane/32
        float4 v;
        v.x = W0[row * din + col + 0];
        v.y = W0[row * din + col + 2];
        v.z = W0[row * din + col + 4];
        v.w = W0[row * din + col + 6];
        s->afrag[j] = v;
    }
    for (int j = threadIdx.x; j < kH; j += blockDim.x) {
        s->b0[j] = b0[j];
        s->w1[j] = w1[j];
        if (!pos_inv) {
            s->wxyz[j][0] = W0[j * din + 0];
            s->wxyz[j][1] = W0[j * din + 1];
            s->wxyz[j][2] = W0[j * din + 2];
        } else {
            s->wxyz[j][0] = s->wxyz[j][1] = s->wxyz[j][2] = 0.0f;
        }
        s->wxyz[j][3] = 0.0f;
    }
}

// Stage decoder `dec` of the model.
__device__ __forceinline__ void stage_fp32_lds(const ModelArgs& m, int dec, Fp32Lds* s) {
    stage_fp32_lds_raw(m.w0[dec], m.pos_inv ? kC : kC + 3, m.b0[dec], m.w1[dec], s);
}

// Second layer over one 32x32 output tile held in registers: sum of
// w1[row] * relu(C + b0[row] + wxyz[row] . xyz) over this lane's 16 rows.
template <typename LDS>
__device__ __forceinline__ float layer2_tile(const f32x16_t& c, int h, uint32_t half, const LDS* s, bool xyz,
                                             float px, float py, float pz) {
    float acc = 0.0f;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int row0 = 32 * h + 8 * g + 4 * (int)half;
        const float4 b = *reinterpret_cast<const float4*>(&s->b0[row0]);
        const float4 w = *reinterpret_cast<const float4*>(&s->w1[row0]);
        float v0 = c[4 * g + 0] + b.x, v1 = c[4 * g + 1] + b.y, v2 = c[4 * g + 2] + b.z, v3 = c[4 * g + 3] + b.w;
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
        acc = __builtin_fmaf(w.x, fmaxf(v0, 0.0f), acc);
        acc = __builtin_fmaf(w.y, fmaxf(v1, 0.0f), acc);
        acc = __builtin_fmaf(w.z, fmaxf(v2, 0.0f), acc);
        acc = __builtin_fmaf(w.w, fmaxf(v3, 0.0f), acc);
    }
    return acc;
}

// SDF of the 64 lanes' query points in FP32.  Every lane of the wave must
// call this (no divergent branch around it): the MFMAs and lane swaps are
// wave-wide.  Lanes without a query pass active = false: they load nothing,
// and a 32-query tile with no active lane skips its MFMAs.  Each output
// column depends only on its own query, so a lane's result is the same
// whatever the other lanes hold.
// The decoder on the 64 lanes' feature vectors f (zero on inactive lanes).
// Wave-wide, as sdf_fp32_wave.
template <bool POSINV>
__device__ __forceinline__ float decode_fp32_wave(const float (&f)[kC], const Fp32Lds* s, float b1, uint32_t lane,
                                                  bool active, float px, float py, float pz) {
    const uint64_t am = __ballot(active);
    const bool use0 = (uint32_t)am != 0u, use1 = (uint32_t)(am >> 32) != 0u;
    const uint32_t half = lane >> 5;
    const bool lo = half == 0;
    // B operands of both tiles for channel pair kk: tile 0 takes (own channel
    // 2kk | channel 2kk+1 of lane l-32), tile 1 (channel 2kk of lane l+32 |
    // own channel 2kk+1).
    float B0[kC / 2], B1[kC / 2];
#pragma unroll
    for (int kk = 0; kk < kC / 2; ++kk) {
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(f[2 * kk]), __float_as_uint(f[2 * kk + 1]),
                                                         false, false);
        B0[kk] = __uint_as_float(sw[0]);
        B1[kk] = __uint_as_float(sw[1]);
    }
    const float qx0 = __shfl(px, lane & 31), qy0 = __shfl(py, lane & 31), qz0 = __shfl(pz, lane & 31);
    const float qx1 = __shfl(px, (lane & 31) + 32), qy1 = __shfl(py, (lane & 31) + 32),
                qz1 = __shfl(pz, (lane & 31) + 32);
    float part0 = 0.0f, part1 = 0.0f;
#pragma unroll 1
    for (int h = 0; h < 4; ++h) {
        // a tile with no active query (a draining wave) skips its MFMAs
        if (use0) {
            f32x16_t c = {};
#pragma unroll
            for (int kq = 0; kq < 4; ++kq) {
                const float4 a = s->afrag[(h * 4 + kq) * 64 + lane];
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, B0[4 * kq + 0], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, B0[4 * kq + 1], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, B0[4 * kq + 2], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, B0[4 * kq + 3], c, 0, 0, 0);
            }
            part0 += layer2_tile(c, h, half, s, !POSINV, qx0, qy0, qz0);
        }
        if (use1) {
            f32x16_t c = {};
#pragma unroll
            for (int kq = 0; kq < 4; ++kq) {
                const float4 a = s->afrag[(h * 4 + kq) * 64 + lane];
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, B1[4 * kq + 0], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, B1[4 * kq + 1], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, B1[4 * kq + 2], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, B1[4 * kq + 3], c, 0, 0, 0);
            }
            part1 += layer2_tile(c, h, half, s, !POSINV, qx1, qy1, qz1);
        }
    }
    part0 += __shfl_xor(part0, 32);
    part1 += __shfl_xor(part1, 32);
    return (lo ? part0 : part1) + b1;
}

template <bool POSINV>
__device__ __forceinline__ float sdf_fp32_wave(const ModelArgs& m, int lod, const Fp32Lds* s, float b1,
                                               uint32_t lane, bool active, float px, float py, float pz,
                                               CoarsePlan cp = CoarsePlan{0, 0}, const CoarseLds32* cs = nullptr) {
    float f[kC];
    if (active) {
        if (cp.levels > 0)
            gather_fp32_coarse(m, lod, cp, cs, px, py, pz, f);
        else
            gather_fp32(m, lod, px, py, pz, f);
    } else {
#pragma unroll
        for (int c = 0; c < kC; ++c) f[c] = 0.0f;
    }
    return decode_fp32_wave<POSINV>(f, s, b1, lane, active, px, py, pz);
}

}  // namespace nglod
