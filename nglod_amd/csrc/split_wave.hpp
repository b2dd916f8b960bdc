// FP32-accurate SDF decoder for one wavefront on the BF16 matrix cores: the
// first decoder layer as products of three-term BF16 splits.
//
// Every FP32 value x is split exactly into x = xh + xm + xl with xh = bf16(x),
// xm = bf16(x - xh) and xl = x - xh - xm (a BF16 holds 8 significant bits, so
// the last residual, at most 8 bits wide, is exact).  Of the nine cross
// products of W0 = Wh + Wm + Wl and f = fh + fm + fl the six of order 2^-16
// and above are kept:
//   Wl fh + Wh fl + Wm fm,   Wm fh + Wh fm,   Wh fh
// (smallest first); the three dropped terms are at most 2^-24 |W0 f| each, an
// FP32 rounding.  Each product of two BF16 values is exact in FP32 and the
// MFMA accumulates in FP32, so the result carries FP32-level rounding, like
// the FP32 MFMA decoder, at 12 v_mfma_f32_32x32x16_bf16 per 32x32 tile (32
// cycles each) instead of 16 v_mfma_f32_32x32x2_f32 (64 cycles each).
//
// Operand layout (v_mfma_f32_32x32x16_bf16): lane l holds A[m = l%32][8
// k values of half l/32] and B[the same k][n = l%32]; output as in
// mixed_device.hpp (lane l holds C[m][n = l%32] for m = 8*(i/4) + 4*(l/32) +
// i%4).  Channel map: in block kb (0, 1) element e of half hf is channel
// 2*(8kb + e) + hf, so one v_permlane32_swap per channel pair builds both
// 32-query tiles' operands from a lane's own features, as in fp32_wave.hpp.
#pragma once

#include "fp32_wave.hpp"
#include "mixed_device.hpp"

namespace nglod {

// Per-workgroup constants of the split decoder (26.5 KB): A fragments
// [part h/m/l][h tile 0..3][k block 0..1][lane] (8 bf16 each), b0, w1 and the
// FP32 xyz columns of W0.
struct SplitLds {
    uint4 wfrag[3][4 * 2 * 64];
    float b0[kH];
    float w1[kH];
    float wxyz[kH][4];  // W0 columns 0..2 (xyz input); [3] unused
};

// x = h + m + l exactly (finite x; the grids and weights are far from the
// BF16 underflow range).
__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
    h = (__bf16)x;
    const float r = x - (float)h;
    m = (__bf16)r;
    l = (__bf16)(r - (float)m);
}

__device__ __forceinline__ uint32_t bf16_pair(__bf16 lo, __bf16 hi) {
    const bf16x2v v = {lo, hi};
    return __builtin_bit_cast(uint32_t, v);
}

// Stage decoder (W0 [128][din], din = 32 or 35 with xyz first; b0; w1) into
// LDS, splitting W0.  Every thread of the block must call this, followed by
// __syncthreads().
__device__ __forceinline__ void stage_split_lds_raw(const float* W0, int din, const float* b0, const float* w1,
                                                    SplitLds* s) {
    const bool pos_inv = din == kC;
    const int off = pos_inv ? 0 : 3;
    for (int j = threadIdx.x; j < 4 * 2 * 64; j += blockDim.x) {
        const int lane = j & 63;
        const int kb = (j >> 6) & 1;
        const int h = j >> 7;
        const int row = 32 * h + (lane & 31);
        uint32_t ph[4], pm[4], pl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            __bf16 h0, m0, l0, h1, m1, l1;
            // elements e = 2q, 2q + 1: channels 2*(8kb + e) + lane/32
            split3(W0[row * din + off + 2 * (8 * kb + 2 * q) + (lane >> 5)], h0, m0, l0);
            split3(W0[row * din + off + 2 * (8 * kb + 2 * q + 1) + (lane >> 5)], h1, m1, l1);
            ph[q] = bf16_pair(h0, h1);
            pm[q] = bf16_pair(m0, m1);
            pl[q] = bf16_pair(l0, l1);
        }
        s->wfrag[0][j] = make_uint4(ph[0], ph[1], ph[2], ph[3]);
        s->wfrag[1][j] = make_uint4(pm[0], pm[1], pm[2], pm[3]);
        s->wfrag[2][j] = make_uint4(pl[0], pl[1], pl[2], pl[3]);
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

__device__ __forceinline__ void stage_split_lds(const ModelArgs& m, int dec, SplitLds* s) {
    stage_split_lds_raw(m.w0[dec], m.pos_inv ? kC : kC + 3, m.b0[dec], m.w1[dec], s);
}

// The three BF16 parts of this lane's A fragments of hidden tile h, k block
// kb (8 values each), split from the FP32 fragments the FP32 MFMA decoder
// stages (Fp32Lds::afrag): element e of half l/32 is channel 2*(8kb + e) +
// l/32, which are the float4 fragments of k quads 2kb and 2kb + 1.
__device__ __forceinline__ void split_afrag(const Fp32Lds* s, int h, int kb, uint32_t lane, bf16x8& ah, bf16x8& am,
                                            bf16x8& al) {
    const float4 q0 = s->afrag[(h * 4 + 2 * kb) * 64 + lane];
    const float4 q1 = s->afrag[(h * 4 + 2 * kb + 1) * 64 + lane];
    const float v[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    uint32_t ph[4], pm[4], pl[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        __bf16 h0, m0, l0, h1, m1, l1;
        split3(v[2 * q], h0, m0, l0);
        split3(v[2 * q + 1], h1, m1, l1);
        ph[q] = bf16_pair(h0, h1);
        pm[q] = bf16_pair(m0, m1);
        pl[q] = bf16_pair(l0, l1);
    }
    ah = __builtin_bit_cast(bf16x8, make_uint4(ph[0], ph[1], ph[2], ph[3]));
    am = __builtin_bit_cast(bf16x8, make_uint4(pm[0], pm[1], pm[2], pm[3]));
    al = __builtin_bit_cast(bf16x8, make_uint4(pl[0], pl[1], pl[2], pl[3]));
}

// The same parts staged in SplitLds.
__device__ __forceinline__ void split_afrag(const SplitLds* s, int h, int kb, uint32_t lane, bf16x8& ah, bf16x8& am,
                                            bf16x8& al) {
    ah = __builtin_bit_cast(bf16x8, s->wfrag[0][(h * 2 + kb) * 64 + lane]);
    am = __builtin_bit_cast(bf16x8, s->wfrag[1][(h * 2 + kb) * 64 + lane]);
    al = __builtin_bit_cast(bf16x8, s->wfrag[2][(h * 2 + kb) * 64 + lane]);
}

// One 32x32 tile of the first layer, b0 included: the six split products,
// smallest first, each over the two k blocks.
__device__ __forceinline__ f32x16 split_tile(const bf16x8 (&ah)[2], const bf16x8 (&am)[2], const bf16x8 (&al)[2],
                                             const bf16x8 (&bh)[2], const bf16x8 (&bm)[2], const bf16x8 (&bl)[2],
                                             f32x16 c) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[kb], bh[kb], c, 0, 0, 0);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[kb], bl[kb], c, 0, 0, 0);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[kb], bm[kb], c, 0, 0, 0);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[kb], bh[kb], c, 0, 0, 0);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[kb], bm[kb], c, 0, 0, 0);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[kb], bh[kb], c, 0, 0, 0);
    return c;
}

// The decoder on the 64 lanes' FP32 feature vectors f (zero on inactive
// lanes).  Every lane of the wave must call this (no divergent branch around
// it): the MFMAs and lane swaps are wave-wide.  A 32-query tile with no
// active lane skips its MFMAs; each output column depends only on its own
// query, so a lane's result is the same whatever the other lanes hold.
template <bool POSINV, typename LDS>
__device__ __forceinline__ float decode_split_wave(const float (&f)[kC], const LDS* s, float b1, uint32_t lane,
                                                   bool active, float px, float py, float pz) {
    const uint64_t am_ = __ballot(active);
    const bool use0 = (uint32_t)am_ != 0u, use1 = (uint32_t)(am_ >> 32) != 0u;
    const uint32_t half = lane >> 5;
    // [tile][kb] B operands of the three parts
    bf16x8 Bh[2][2], Bm[2][2], Bl[2][2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        uint32_t th[2][4], tm[2][4], tl[2][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            __bf16 h[2][2], m[2][2], l[2][2];  // [tile][element parity]
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int kk = 8 * kb + 2 * q + p;  // channel pair
                const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(f[2 * kk]),
                                                                 __float_as_uint(f[2 * kk + 1]), false, false);
                split3(__uint_as_float(sw[0]), h[0][p], m[0][p], l[0][p]);
                split3(__uint_as_float(sw[1]), h[1][p], m[1][p], l[1][p]);
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                th[t][q] = bf16_pair(h[t][0], h[t][1]);
                tm[t][q] = bf16_pair(m[t][0], m[t][1]);
                tl[t][q] = bf16_pair(l[t][0], l[t][1]);
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            Bh[t][kb] = __builtin_bit_cast(bf16x8, make_uint4(th[t][0], th[t][1], th[t][2], th[t][3]));
            Bm[t][kb] = __builtin_bit_cast(bf16x8, make_uint4(tm[t][0], tm[t][1], tm[t][2], tm[t][3]));
            Bl[t][kb] = __builtin_bit_cast(bf16x8, make_uint4(tl[t][0], tl[t][1], tl[t][2], tl[t][3]));
        }
    }
    float qx0, qy0, qz0, qx1, qy1, qz1;
    tile_coords(px, py, pz, qx0, qy0, qz0, qx1, qy1, qz1);
    float part0 = 0.0f, part1 = 0.0f;
#pragma unroll 1
    for (int h = 0; h < 4; ++h) {
        bf16x8 ah[2], am[2], al[2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) split_afrag(s, h, kb, lane, ah[kb], am[kb], al[kb]);
        if (use0) {
            const f32x16 c = split_tile(ah, am, al, Bh[0], Bm[0], Bl[0], bias_tile(h, half, s));
            part0 += layer2_partial(c, h, half, s, !POSINV, qx0, qy0, qz0);
        }
        if (use1) {
            const f32x16 c = split_tile(ah, am, al, Bh[1], Bm[1], Bl[1], bias_tile(h, half, s));
            part1 += layer2_partial(c, h, half, s, !POSINV, qx1, qy1, qz1);
        }
    }
    return halves_sum(part0, part1) + b1;
}

// decode_split_wave for a wave with few busy lanes (latency-bound): a used
// tile's four hidden tiles are independent and issued together instead of
// one after another.  Per output the same MFMA chain on the same operands and
// the partial sums added in the same order: the same bits as
// decode_split_wave.
template <bool POSINV, typename LDS>
__device__ __forceinline__ float decode_split_wave_few(const float (&f)[kC], const LDS* s, float b1, uint32_t lane,
                                                       bool active, float px, float py, float pz) {
    const uint64_t am_ = __ballot(active);
    const uint32_t half = lane >> 5;
    float qx0, qy0, qz0, qx1, qy1, qz1;
    tile_coords(px, py, pz, qx0, qy0, qz0, qx1, qy1, qz1);
    float part[2] = {0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        if ((t == 0 ? (uint32_t)am_ : (uint32_t)(am_ >> 32)) == 0u) continue;
        bf16x8 Bh[2], Bm[2], Bl[2];  // [kb] of tile t
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            uint32_t th[4], tm[4], tl[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                __bf16 h[2], m[2], l[2];
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const int kk = 8 * kb + 2 * q + p;
                    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(f[2 * kk]),
                                                                     __float_as_uint(f[2 * kk + 1]), false, false);
                    split3(__uint_as_float(sw[t]), h[p], m[p], l[p]);
                }
                th[q] = bf16_pair(h[0], h[1]);
                tm[q] = bf16_pair(m[0], m[1]);
                tl[q] = bf16_pair(l[0], l[1]);
            }
            Bh[kb] = __builtin_bit_cast(bf16x8, make_uint4(th[0], th[1], th[2], th[3]));
            Bm[kb] = __builtin_bit_cast(bf16x8, make_uint4(tm[0], tm[1], tm[2], tm[3]));
            Bl[kb] = __builtin_bit_cast(bf16x8, make_uint4(tl[0], tl[1], tl[2], tl[3]));
        }
        float p[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            bf16x8 ah[2], am[2], al[2];
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) split_afrag(s, h, kb, lane, ah[kb], am[kb], al[kb]);
            const f32x16 c = split_tile(ah, am, al, Bh, Bm, Bl, bias_tile(h, half, s));
            p[h] = layer2_partial(c, h, half, s, !POSINV, t == 0 ? qx0 : qx1, t == 0 ? qy0 : qy1,
                                  t == 0 ? qz0 : qz1);
        }
#pragma unroll
        for (int h = 0; h < 4; ++h) part[t] += p[h];
    }
    return halves_sum(part[0], part[1]) + b1;
}

// Gather of the cumulative grid (gather_fp32_cum, any level table LV whose
// level 0 is that grid) and decode of the 64 lanes' queries.  Wave-wide.
template <bool POSINV, typename LV, typename LDS>
__device__ __forceinline__ float sdf_split_wave(const LV& lv, const LDS* s, float b1, uint32_t lane, bool active,
                                                float px, float py, float pz) {
    float f[kC];
    gather_fp32_cum(lv, active, px, py, pz, f);
    return decode_split_wave<POSINV, LDS>(f, s, b1, lane, active, px, py, pz);
}

// gather_fp32_cum for the busy lanes of a wave with at most 4 of them (act,
// a wave-uniform mask), done by 16 lanes per query: lane 16 i + j computes
// channels 2j and 2j + 1 of the i-th busy lane's query with the 8 corner
// loads of the cumulative grid (level 0) issued first, then the same FMAs in
// the same corner order; each busy lane then collects its 32 channels.
// Bit-identical to gather_fp32_cum.  Every lane of the wave must call this.
template <typename LV>
__device__ __forceinline__ void gather_fp32_coop(const LV* lv, uint32_t lane, uint64_t act, float qx, float qy,
                                                 float qz, float (&f)[kC]) {
    constexpr int lod = 0;
    const int i = (int)(lane >> 4), j = (int)(lane & 15);
    const int k = __popcll(act);
    const bool helper = i < k;
    const int src = helper ? nth_set_bit(act, i) : 0;
    const float px = __shfl(qx, src), py = __shfl(qy, src), pz = __shfl(qz, src);
    float2 g[lod + 1][8];
    if (helper) {
#pragma unroll
        for (int l = 0; l <= lod; ++l) {
            const int r = lv->res[l];
            const int S = r + 1;
            const AxisW ax = axis_weights(px, r);
            const AxisW ay = axis_weights(py, r);
            const AxisW az = axis_weights(pz, r);
            const float2* base = reinterpret_cast<const float2*>(
                lv->grid[l] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC) + j;
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
    float v0 = 0.0f, v1 = 0.0f;
    if (helper) {
#pragma unroll
        for (int l = 0; l <= lod; ++l) {
            const int r = lv->res[l];
            const AxisW ax = axis_weights(px, r);
            const AxisW ay = axis_weights(py, r);
            const AxisW az = axis_weights(pz, r);
            const float wxy00 = ax.w0 * ay.w0, wxy10 = ax.w1 * ay.w0;
            const float wxy01 = ax.w0 * ay.w1, wxy11 = ax.w1 * ay.w1;
            const float w[8] = {wxy00 * az.w0, wxy10 * az.w0, wxy01 * az.w0, wxy11 * az.w0,
                                wxy00 * az.w1, wxy10 * az.w1, wxy01 * az.w1, wxy11 * az.w1};
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                v0 = __builtin_fmaf(g[l][c].x, w[c], v0);
                v1 = __builtin_fmaf(g[l][c].y, w[c], v1);
            }
        }
    }
    const bool mine = (act >> lane) & 1ull;
    const int q = mine ? (int)mask_rank(act) : 0;
#pragma unroll
    for (int c = 0; c < kC / 2; ++c) {
        const float a = __shfl(v0, 16 * q + c), b = __shfl(v1, 16 * q + c);
        f[2 * c] = mine ? a : 0.0f;
        f[2 * c + 1] = mine ? b : 0.0f;
    }
}

}  // namespace nglod
