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
    float qx0, qy0, qz0, qx1, qy1, qz1;
    tile_coords(px, py, pz, qx0, qy0, qz0, qx1, qy1, qz1);
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
    return halves_sum(part0, part1) + b1;
}

// The SDF of ONE query (wave-uniform px, py, pz) computed by the whole wave,
// bit-identical to sdf_fp32_wave's result for that query: for a wave with
// one or two rays left, where a wave-wide step costs the latency of several
// dependent rounds of gathers plus 64 MFMAs whatever the lane count.
//   gather: lane c (and c + 32) owns channel c; every level's 8 corner values
//     are loaded first (one round of memory latency), then summed with the
//     same weights in the same level / corner order as gather_fp32(_coarse);
//   layer 1: lane L computes hidden rows L and L + 64 as the FMA chain
//     z = fma(W0[R][2kk+1], f[2kk+1], fma(W0[R][2kk], f[2kk], z)), kk = 0..15,
//     which is how v_mfma_f32_32x32x2_f32 accumulates (exact products, one
//     rounding per product, k = 0 first; checked bit for bit on the device);
//   layer 2: the ReLU outputs go through `sv` (this wave's 128 floats of LDS)
//     and lanes 0 / 32 sum rows in layer2_tile's order for half 0 / half 1.
// The grid pointers and resolutions come from an LDS copy (SoloLevels): read
// as vector values they add no scalar registers to the persistent kernels,
// which already use nearly all of theirs.
// Every lane of the wave must call this with the same arguments.
struct SoloLevels {
    const float* grid[NGLOD_MAX_LODS];
    int res[NGLOD_MAX_LODS];
};

// Every thread of the block calls this before the block's __syncthreads().
__device__ __forceinline__ void stage_solo_levels(const ModelArgs& m, SoloLevels* lv) {
#pragma unroll
    for (int l = 0; l < NGLOD_MAX_LODS; ++l) {
        if ((int)threadIdx.x == l) {
            lv->grid[l] = m.grid[l];
            lv->res[l] = m.res[l];
        }
    }
}

template <bool POSINV>
__device__ __forceinline__ float sdf_fp32_solo(const SoloLevels* lv, int lod, const Fp32Lds* s, float b1,
                                               uint32_t lane, float* sv, float px, float py, float pz, CoarsePlan cp,
                                               const CoarseLds32* cs) {
    const int c = (int)(lane & 31);
    float g[NGLOD_MAX_LODS][8];
#pragma unroll
    for (int l = 0; l < NGLOD_MAX_LODS; ++l) {
        if (l > lod) break;
        const int r = lv->res[l];
        const int S = r + 1;
        const AxisW ax = axis_weights(px, r);
        const AxisW ay = axis_weights(py, r);
        const AxisW az = axis_weights(pz, r);
        if (l < cp.levels) {
            const float* cf = reinterpret_cast<const float*>(cs->c);
            const int i0 = (l == 0 ? 0 : cp.off1) + (az.i0 * S + ay.i0) * S + ax.i0;
            const int sy = S, sz = S * S;
            const int k4 = c >> 2, j = c & 3;
            g[l][0] = cf[swz32(i0, k4) * 4 + j];
            g[l][1] = cf[swz32(i0 + 1, k4) * 4 + j];
            g[l][2] = cf[swz32(i0 + sy, k4) * 4 + j];
            g[l][3] = cf[swz32(i0 + sy + 1, k4) * 4 + j];
            g[l][4] = cf[swz32(i0 + sz, k4) * 4 + j];
            g[l][5] = cf[swz32(i0 + sz + 1, k4) * 4 + j];
            g[l][6] = cf[swz32(i0 + sz + sy, k4) * 4 + j];
            g[l][7] = cf[swz32(i0 + sz + sy + 1, k4) * 4 + j];
        } else {
            const float* base = lv->grid[l] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC + c;
            const size_t sx = kC, sy = (size_t)S * kC, sz = (size_t)S * S * kC;
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
    float f = 0.0f;
#pragma unroll
    for (int l = 0; l < NGLOD_MAX_LODS; ++l) {
        if (l > lod) break;
        const int r = lv->res[l];
        const AxisW ax = axis_weights(px, r);
        const AxisW ay = axis_weights(py, r);
        const AxisW az = axis_weights(pz, r);
        const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
        const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
        f = __builtin_fmaf(g[l][0], wxy00 * az.w0, f);
        f = __builtin_fmaf(g[l][1], wxy10 * az.w0, f);
        f = __builtin_fmaf(g[l][2], wxy01 * az.w0, f);
        f = __builtin_fmaf(g[l][3], wxy11 * az.w0, f);
        f = __builtin_fmaf(g[l][4], wxy00 * az.w1, f);
        f = __builtin_fmaf(g[l][5], wxy10 * az.w1, f);
        f = __builtin_fmaf(g[l][6], wxy01 * az.w1, f);
        f = __builtin_fmaf(g[l][7], wxy11 * az.w1, f);
    }
    // layer 1: rows R0 = lane (tiles 0, 1) and R1 = lane + 64 (tiles 2, 3)
    const int m0 = (int)(lane & 31), h0 = (int)(lane >> 5), h1 = h0 + 2;
    float z0 = 0.0f, z1 = 0.0f;
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) {
        const float4 a0 = s->afrag[(h0 * 4 + kq) * 64 + m0], a0k = s->afrag[(h0 * 4 + kq) * 64 + 32 + m0];
        const float4 a1 = s->afrag[(h1 * 4 + kq) * 64 + m0], a1k = s->afrag[(h1 * 4 + kq) * 64 + 32 + m0];
        const float ea0[4] = {a0.x, a0.y, a0.z, a0.w}, ea0k[4] = {a0k.x, a0k.y, a0k.z, a0k.w};
        const float ea1[4] = {a1.x, a1.y, a1.z, a1.w}, ea1k[4] = {a1k.x, a1k.y, a1k.z, a1k.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ch = 8 * kq + 2 * e;
            const float f0 = __shfl(f, ch), f1 = __shfl(f, ch + 1);
            z0 = __builtin_fmaf(ea0k[e], f1, __builtin_fmaf(ea0[e], f0, z0));
            z1 = __builtin_fmaf(ea1k[e], f1, __builtin_fmaf(ea1[e], f0, z1));
        }
    }
    auto relu_row = [&](float z, int R) {
        float v = z + s->b0[R];
        if (!POSINV)
            v = __builtin_fmaf(s->wxyz[R][0], px, __builtin_fmaf(s->wxyz[R][1], py, __builtin_fmaf(s->wxyz[R][2], pz, v)));
        return fmaxf(v, 0.0f);
    };
    sv[lane] = relu_row(z0, (int)lane);
    sv[lane + 64] = relu_row(z1, (int)lane + 64);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // layer 2 in layer2_tile's order: lanes 0..31 sum half 0, lanes 32..63 half 1
    const int half = (int)(lane >> 5);
    float part = 0.0f;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        float acc = 0.0f;
#pragma unroll
        for (int gq = 0; gq < 4; ++gq) {
            const int row0 = 32 * h + 8 * gq + 4 * half;
            const float4 w = *reinterpret_cast<const float4*>(&s->w1[row0]);
            const float4 v = *reinterpret_cast<const float4*>(&sv[row0]);
            acc = __builtin_fmaf(w.x, v.x, acc);
            acc = __builtin_fmaf(w.y, v.y, acc);
            acc = __builtin_fmaf(w.z, v.z, acc);
            acc = __builtin_fmaf(w.w, v.w, acc);
        }
        part += acc;
    }
    const float p0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(part), 0));
    const float p1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(part), 32));
    return (p0 + p1) + b1;
}

template <bool POSINV>
__device__ __forceinline__ float sdf_fp32_wave(const ModelArgs& m, int lod, const Fp32Lds* s, float b1,
                                               uint32_t lane, bool active, float px, float py, float pz,
                                               CoarsePlan cp = CoarsePlan{0, 0}, const CoarseLds32* cs = nullptr) {
    // cleared once for both cases (zero on inactive lanes)
    float f[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) f[c] = 0.0f;
    if (active) {
        if (cp.levels > 0)
            gather_fp32_coarse(m, lod, cp, cs, px, py, pz, f);
        else
            gather_fp32_add(m, lod, px, py, pz, f);
    }
    return decode_fp32_wave<POSINV>(f, s, b1, lane, active, px, py, pz);
}

}  // namespace nglod
