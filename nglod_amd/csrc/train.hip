// Hand-written backward of the fused query (OctreeSDF.sdf, every level up
// to lod) for training, deterministic by construction (no float atomics):
//
//   1. per-query kernel: recomputes the trilinear weights and running
//      features, and for each level with an incoming gradient the decoder
//      pre-activations Z = W0 e + b0; writes E (decoder input), Z and
//      ds = W0^T dz (feature part), dz = g * W1 * [Z > 0].
//   2. decoder gradients: each block reduces a fixed chunk of queries in
//      query order into partial dW0 (FP32 MFMA), db0, dW1, db1; a second
//      kernel sums fixed groups of partials in chunk order and a third the
//      group sums in group order (and over levels, in level order, for a
//      joint decoder).
//   3. grid gradients, per level j: every (query, corner) contribution
//      w_c * dg_j with dg_j = sum_{i >= j} ds_i gets the corner index as a
//      key; a stable radix sort groups equal corners with queries in
//      ascending order; run-length encoding finds each corner's run; runs
//      are cut into fixed pieces summed in order by one thread each, and the
//      piece sums of a run are added in order (coarse levels have runs of
//      thousands of contributions, so a run is not left to one thread).
//
// Replaces autograd through F.grid_sample and nn.Linear in
// Trainer.step_geometry (sdf-net/lib/trainer.py:317-339).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "api_common.hpp"
#include "fp32_wave.hpp"

namespace nglod {

constexpr int kChunk = 256;   // queries per partial of the decoder-gradient reduction
constexpr int kRunPiece = 64;  // sorted contributions per partial of a corner's gradient

struct BwdArgs {
    ModelArgs m;
    int lod;
    int din;       // decoder input width (32 or 35)
    int off;       // 0 or 3: offset of the features in the decoder input
    int mask;      // bit i: level i has an incoming gradient
    int64_t n;
    const float* x;
    const float* g;  // [lod+1][n]
    float* E;        // [lod+1][n][din]
    float* Z;        // [lod+1][n][128]
    float* ds;       // [lod+1][n][32]
};

// Per-block copy of one level's decoder for the two MFMA products of the
// query kernel: fwd = W0 fragments for Z^T = W0 e (as the trace decoders,
// Fp32Lds), bwd[h][g][j][lane] = W0[32h + 8g + 4 (lane / 32) + j][off + lane % 32]
// for ds^T = W0^T dz, whose k-steps pair hidden rows 8g + j and 8g + 4 + j:
// exactly the rows lanes n and n + 32 hold for query n in the first product's
// output, so each lane feeds its own registers as the second product's B
// operand.
struct BwdLds {
    Fp32Lds fwd;
    float bwd[4 * 4 * 4 * 64];
};

__device__ __forceinline__ void stage_bwd_lds(const ModelArgs& m, int dec, int din, int off, BwdLds* s) {
    stage_fp32_lds_raw(m.w0[dec], din, m.b0[dec], m.w1[dec], &s->fwd);
    const float* W0 = m.w0[dec];
    for (int t = threadIdx.x; t < 4 * 4 * 4 * 64; t += blockDim.x) {
        const int lane = t & 63, j = (t >> 6) & 3, g = (t >> 8) & 3, h = t >> 10;
        const int row = 32 * h + 8 * g + 4 * (lane >> 5) + j;
        s->bwd[t] = W0[row * din + off + (lane & 31)];
    }
}

// One wave per 64 queries, every level up to lod: the running trilinear
// features (stored as E), then for a level with an incoming gradient
// Z^T = W0 e + b0 as FP32 MFMA 32x32x2 tiles (stored as Z), dz = g * W1 * [Z > 0]
// in the output registers, and ds^T = W0^T dz as a second chain of MFMA tiles
// (stored as ds).  A 32-query tile with no query in range skips its MFMAs.
// Every thread of the block runs every level (the decoder tables are
// restaged between levels when each level has its own decoder).
__global__ __launch_bounds__(256) void bwd_query_kernel(BwdArgs a) {
    __shared__ BwdLds lds;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = q < a.n;
    const uint32_t lane = lane_id();
    const uint32_t half = lane >> 5;
    const bool lo = half == 0;
    const uint64_t am = __ballot(valid);
    const bool use0 = (uint32_t)am != 0u, use1 = (uint32_t)(am >> 32) != 0u;
    const int64_t wave0 = q - (int64_t)lane;  // query of lane 0
    const int64_t q0 = wave0 + (lane & 31), q1 = q0 + 32;  // this lane's column of tile 0 / tile 1
    const float px = valid ? a.x[3 * q] : 0.0f, py = valid ? a.x[3 * q + 1] : 0.0f, pz = valid ? a.x[3 * q + 2] : 0.0f;
    const float qx0 = __shfl(px, lane & 31), qy0 = __shfl(py, lane & 31), qz0 = __shfl(pz, lane & 31);
    const float qx1 = __shfl(px, (lane & 31) + 32), qy1 = __shfl(py, (lane & 31) + 32),
                qz1 = __shfl(pz, (lane & 31) + 32);
    float f[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) f[c] = 0.0f;
    int staged = -1;
    for (int l = 0; l <= a.lod; ++l) {
        if (valid) {
            const int r = a.m.res[l];
            const int S = r + 1;
            const AxisW ax = axis_weights(px, r);
            const AxisW ay = axis_weights(py, r);
            const AxisW az = axis_weights(pz, r);
            const float* base = a.m.grid[l] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC;
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
            float* E = a.E + ((size_t)l * a.n + q) * a.din;
            if (a.off) {
                E[0] = px;
                E[1] = py;
                E[2] = pz;
#pragma unroll
                for (int c = 0; c < kC; ++c) E[a.off + c] = f[c];
            } else {
                float4* E4 = reinterpret_cast<float4*>(E);  // rows of 32 floats: 16-byte aligned
#pragma unroll
                for (int c = 0; c < kC / 4; ++c) E4[c] = make_float4(f[4 * c], f[4 * c + 1], f[4 * c + 2], f[4 * c + 3]);
            }
        }
        if (!((a.mask >> l) & 1)) {
            if (valid) {
                float4* ds4 = reinterpret_cast<float4*>(a.ds + ((size_t)l * a.n + q) * kC);
#pragma unroll
                for (int c = 0; c < kC / 4; ++c) ds4[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            continue;
        }
        const int dec = a.m.num_dec == 1 ? 0 : l;
        if (dec != staged) {  // block-uniform
            __syncthreads();
            stage_bwd_lds(a.m, dec, a.din, a.off, &lds);
            __syncthreads();
            staged = dec;
        }
        // B operands of both query tiles (as decode_fp32_wave)
        float B0[kC / 2], B1[kC / 2];
#pragma unroll
        for (int kk = 0; kk < kC / 2; ++kk) {
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(f[2 * kk]), __float_as_uint(f[2 * kk + 1]),
                                                             false, false);
            B0[kk] = __uint_as_float(sw[0]);
            B1[kk] = __uint_as_float(sw[1]);
        }
        const float gq = valid ? a.g[(size_t)l * a.n + q] : 0.0f;
        const float g0 = __shfl(gq, lane & 31), g1 = __shfl(gq, (lane & 31) + 32);
        f32x16_t d0 = {}, d1 = {};  // ds^T tiles: channel x query
#pragma unroll 1
        for (int h = 0; h < 4; ++h) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (t == 0 ? !use0 : !use1) continue;
                f32x16_t c = {};
#pragma unroll
                for (int kq = 0; kq < 4; ++kq) {
                    const float4 w = lds.fwd.afrag[(h * 4 + kq) * 64 + lane];
                    const float* B = t == 0 ? B0 : B1;
                    c = __builtin_amdgcn_mfma_f32_32x32x2f32(w.x, B[4 * kq + 0], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x2f32(w.y, B[4 * kq + 1], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x2f32(w.z, B[4 * kq + 2], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x2f32(w.w, B[4 * kq + 3], c, 0, 0, 0);
                }
                const int64_t qt = t == 0 ? q0 : q1;
                const bool in = qt < a.n;
                const float tx = t == 0 ? qx0 : qx1, ty = t == 0 ? qy0 : qy1, tz = t == 0 ? qz0 : qz1;
                const float gt = t == 0 ? g0 : g1;
                float* Zq = a.Z + ((size_t)l * a.n + (in ? qt : 0)) * kH;
                f32x16_t dz;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int row0 = 32 * h + 8 * g + 4 * (int)half;
                    float z[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float v = c[4 * g + j] + lds.fwd.b0[row0 + j];
                        if (a.off)
                            v = __builtin_fmaf(lds.fwd.wxyz[row0 + j][0], tx,
                                               __builtin_fmaf(lds.fwd.wxyz[row0 + j][1], ty,
                                                              __builtin_fmaf(lds.fwd.wxyz[row0 + j][2], tz, v)));
                        z[j] = v;
                        dz[4 * g + j] = v > 0.0f ? gt * lds.fwd.w1[row0 + j] : 0.0f;
                    }
                    if (in) *reinterpret_cast<float4*>(Zq + row0) = make_float4(z[0], z[1], z[2], z[3]);
                }
                // ds^T += W0^T dz over this tile's 32 hidden rows (k-steps pair rows 8g + j, 8g + 4 + j)
                f32x16_t& d = t == 0 ? d0 : d1;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        d = __builtin_amdgcn_mfma_f32_32x32x2f32(lds.bwd[((h * 4 + g) * 4 + j) * 64 + lane],
                                                                 dz[4 * g + j], d, 0, 0, 0);
                }
            }
        }
        // ds rows: lane n / n + 32 hold channels 8 i + 4 half + (0..3) of query n of each tile
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int64_t qt = t == 0 ? q0 : q1;
            if (qt >= a.n) continue;
            const f32x16_t& d = t == 0 ? d0 : d1;
            float* dsq = a.ds + ((size_t)l * a.n + qt) * kC;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                *reinterpret_cast<float4*>(dsq + 8 * i + 4 * half) =
                    make_float4(d[4 * i + 0], d[4 * i + 1], d[4 * i + 2], d[4 * i + 3]);
        }
    }
}

typedef float f32x16_bwd __attribute__((ext_vector_type(16)));

// Partial decoder gradients of one level over one chunk of kChunk queries.
// Output layout per (level, chunk): dW0 [128][din], db0 [128], dW1 [128],
// db1 [1].  dW0 = dZ^T E over the chunk as FP32 MFMA 32x32x2 products
// (v_mfma_f32_32x32x2_f32: lane l holds A[m = l%32][k = l/32] and
// B[k = l/32][n = l%32]): wave w owns hidden rows 32w..32w+31, and each step
// takes two queries, A[m][k] = dz of query 2s+k at row m, B[k][n] = E of that
// query at column n.  The products are exact and accumulate in FP32 in query
// order, so a partial is the same every run.  db0, dW1 and db1 accumulate in
// the lanes beside it (each half-wave over its query of every pair, the two
// halves added at the end).
__global__ __launch_bounds__(256) void bwd_decoder_partial_kernel(BwdArgs a, int level, float* __restrict__ part,
                                                                  int64_t nchunks) {
    const int64_t chunk = blockIdx.x;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int col = lane & 31;
    const int half = lane >> 5;
    const int j = 32 * wave + col;
    const int dec = a.m.num_dec == 1 ? 0 : level;
    const float w1 = a.m.w1[dec][j];
    const bool two = a.din > 32;  // the xyz columns put E over 32 columns: a second tile
    const int64_t q0 = chunk * kChunk;
    const int64_t q1 = q0 + kChunk < a.n ? q0 + kChunk : a.n;
    const float* Zl = a.Z + (size_t)level * a.n * kH;
    const float* El = a.E + (size_t)level * a.n * a.din;
    const float* gl = a.g + (size_t)level * a.n;
    f32x16_bwd c0 = {}, c1 = {};
    float acc_b0 = 0.0f, acc_w1 = 0.0f, acc_b1 = 0.0f;
    for (int64_t qb = q0; qb < q1; qb += 2) {
        const int64_t q = qb + half;
        float dz = 0.0f, e0 = 0.0f, e1 = 0.0f;
        if (q < q1) {
            const float gq = gl[q];
            const float z = Zl[q * kH + j];
            dz = z > 0.0f ? gq * w1 : 0.0f;
            acc_b0 += dz;
            acc_w1 = __builtin_fmaf(gq, z > 0.0f ? z : 0.0f, acc_w1);
            acc_b1 += gq;
            e0 = El[q * a.din + col];
            if (two && 32 + col < a.din) e1 = El[q * a.din + 32 + col];
        }
        c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(dz, e0, c0, 0, 0, 0);
        if (two) c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(dz, e1, c1, 0, 0, 0);
    }
    const int outs = kH * a.din + 2 * kH + 1;
    float* P = part + ((size_t)level * nchunks + chunk) * outs;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = 32 * wave + 8 * (i >> 2) + 4 * half + (i & 3);
        P[row * a.din + col] = c0[i];
        if (two && 32 + col < a.din) P[row * a.din + 32 + col] = c1[i];
    }
    acc_b0 += __shfl_xor(acc_b0, 32);
    acc_w1 += __shfl_xor(acc_w1, 32);
    acc_b1 += __shfl_xor(acc_b1, 32);
    if (half == 0) {
        P[kH * a.din + j] = acc_b0;
        P[kH * a.din + kH + j] = acc_w1;
    }
    if (threadIdx.x == 0) P[kH * a.din + 2 * kH] = acc_b1;
}

// First stage of the partials' sum: group g of level l (grid z) adds chunks
// g*kGroup .. g*kGroup+kGroup-1 in order for every output.
constexpr int kGroup = 32;

__global__ void bwd_decoder_group_kernel(BwdArgs a, const float* __restrict__ part, int64_t nchunks,
                                         float* __restrict__ gpart, int64_t ngroups) {
    const int outs = kH * a.din + 2 * kH + 1;
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.z;
    const int64_t g = blockIdx.y;
    if (o >= outs || !((a.mask >> l) & 1)) return;
    const int64_t c0 = g * kGroup;
    const int64_t c1 = c0 + kGroup < nchunks ? c0 + kGroup : nchunks;
    const float* P = part + (size_t)l * nchunks * outs + o;
    float s = 0.0f;
    for (int64_t c = c0; c < c1; ++c) s += P[c * outs];
    gpart[((size_t)l * ngroups + g) * outs + o] = s;
}

// Sums the group sums of decoder `dec` over its levels (ascending) and groups
// (ascending) and writes dW0, db0, dW1, db1.
__global__ void bwd_decoder_final_kernel(BwdArgs a, int dec, const float* __restrict__ gpart, int64_t ngroups,
                                         float* __restrict__ dw0, float* __restrict__ db0, float* __restrict__ dw1,
                                         float* __restrict__ db1) {
    const int outs = kH * a.din + 2 * kH + 1;
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= outs) return;
    float s = 0.0f;
    for (int l = 0; l <= a.lod; ++l) {
        const int d = a.m.num_dec == 1 ? 0 : l;
        if (d != dec || !((a.mask >> l) & 1)) continue;
        const float* P = gpart + (size_t)l * ngroups * outs + o;
        for (int64_t g = 0; g < ngroups; ++g) s += P[g * outs];
    }
    if (o < kH * a.din) dw0[o] = s;
    else if (o < kH * a.din + kH) db0[o - kH * a.din] = s;
    else if (o < kH * a.din + 2 * kH) dw1[o - kH * a.din - kH] = s;
    else db1[0] = s;
}

// Keys (corner index at `level`) and values (query * 8 + corner slot).
__global__ void bwd_corner_keys_kernel(BwdArgs a, int level, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n) return;
    const int r = a.m.res[level];
    const int S = r + 1;
    const AxisW ax = axis_weights(a.x[3 * q], r);
    const AxisW ay = axis_weights(a.x[3 * q + 1], r);
    const AxisW az = axis_weights(a.x[3 * q + 2], r);
    const uint32_t base = (uint32_t)((az.i0 * S + ay.i0) * S + ax.i0);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t off = (uint32_t)((c & 1) + ((c >> 1) & 1) * S + ((c >> 2) & 1) * S * S);
        keys[q * 8 + c] = base + off;
        vals[q * 8 + c] = (uint32_t)(q * 8 + c);
    }
}

// ds_l(q) -> dg_l(q) = sum_{i >= l} ds_i(q) in place, added from the finest
// level down (((ds_lod + ds_lod-1) + ...) + ds_l): one thread per query and
// group of 4 channels.
__global__ void bwd_suffix_kernel(BwdArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n * (kC / 4)) return;
    const int64_t q = t / (kC / 4);
    const int c4 = (int)(t % (kC / 4));
    float4 acc = reinterpret_cast<const float4*>(a.ds + ((size_t)a.lod * a.n + q) * kC)[c4];
    for (int l = a.lod - 1; l >= 0; --l) {
        float4* d = reinterpret_cast<float4*>(a.ds + ((size_t)l * a.n + q) * kC) + c4;
        const float4 v = *d;
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
        *d = acc;
    }
}

// Contribution of sorted entry e: w_c(x_q) * dg_j(q), dg_j = sum_{l >= j} ds_l
// (ds already holds these suffix sums, bwd_suffix_kernel).
__device__ __forceinline__ void corner_contrib(const BwdArgs& a, int level, uint32_t v, float (&acc)[kC]) {
    const int r = a.m.res[level];
    const int64_t q = v >> 3;
    const int c = v & 7;
    const AxisW ax = axis_weights(a.x[3 * q], r);
    const AxisW ay = axis_weights(a.x[3 * q + 1], r);
    const AxisW az = axis_weights(a.x[3 * q + 2], r);
    const float wx = (c & 1) ? ax.w1 : ax.w0;
    const float wy = ((c >> 1) & 1) ? ay.w1 : ay.w0;
    const float wz = ((c >> 2) & 1) ? az.w1 : az.w0;
    const float w = (wx * wy) * wz;
    const float4* d4 = reinterpret_cast<const float4*>(a.ds + ((size_t)level * a.n + q) * kC);
#pragma unroll
    for (int k = 0; k < kC / 4; ++k) {
        const float4 d = d4[k];
        acc[4 * k] = __builtin_fmaf(w, d.x, acc[4 * k]);
        acc[4 * k + 1] = __builtin_fmaf(w, d.y, acc[4 * k + 1]);
        acc[4 * k + 2] = __builtin_fmaf(w, d.z, acc[4 * k + 2]);
        acc[4 * k + 3] = __builtin_fmaf(w, d.w, acc[4 * k + 3]);
    }
}

// Pieces: run r (a run of equal corner keys, start rstart[r], length cnt[r])
// is cut into ceil(cnt / kRunPiece) pieces; piece_start[r] is the exclusive
// scan of those piece counts.  One thread per piece sums its entries in order.
__global__ void run_pieces_kernel(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ nruns,
                                  uint32_t* __restrict__ pieces) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= *nruns) return;
    pieces[r] = (cnt[r] + kRunPiece - 1) / kRunPiece;
}

__global__ void bwd_piece_sum_kernel(BwdArgs a, int level, const uint32_t* __restrict__ vals,
                                     const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ cnt,
                                     const uint32_t* __restrict__ piece_start, const uint32_t* __restrict__ nruns,
                                     uint32_t max_pieces, float* __restrict__ partial) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t R = *nruns;
    if (R == 0 || p >= max_pieces) return;
    const uint32_t total = piece_start[R - 1] + (cnt[R - 1] + kRunPiece - 1) / kRunPiece;
    if (p >= total) return;
    // run of this piece: last r with piece_start[r] <= p
    uint32_t lo = 0, hi = R - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (piece_start[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t k = p - piece_start[lo];
    const uint32_t b = rstart[lo] + k * kRunPiece;
    const uint32_t e_end = min(rstart[lo] + cnt[lo], b + kRunPiece);
    float acc[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) acc[c] = 0.0f;
    for (uint32_t e = b; e < e_end; ++e) corner_contrib(a, level, vals[e], acc);
    float4* o4 = reinterpret_cast<float4*>(partial + (size_t)p * kC);
#pragma unroll
    for (int c = 0; c < kC / 4; ++c) o4[c] = make_float4(acc[4 * c], acc[4 * c + 1], acc[4 * c + 2], acc[4 * c + 3]);
}

// One thread per (run, channel): sums the run's piece partials in order.
__global__ void bwd_run_final_kernel(const uint32_t* __restrict__ ukeys, const uint32_t* __restrict__ cnt,
                                     const uint32_t* __restrict__ piece_start, const uint32_t* __restrict__ nruns,
                                     uint32_t max_runs, const float* __restrict__ partial, float* __restrict__ grad) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t r = (uint32_t)(t / kC);
    const int c = (int)(t % kC);
    if (r >= max_runs || r >= *nruns) return;
    const uint32_t np = (cnt[r] + kRunPiece - 1) / kRunPiece;
    float s = 0.0f;
    for (uint32_t i = 0; i < np; ++i) s += partial[(size_t)(piece_start[r] + i) * kC + c];
    grad[(size_t)ukeys[r] * kC + c] = s;
}

// The same sums for levels whose runs hold many pieces (the coarse grids:
// 2^19 queries put ~500 pieces in each of LOD 0's 125 corners): one block of
// kRunGroups x kC threads per run; thread (g, c) sums pieces g, g + kRunGroups,
// ... of channel c in order (a row of kC partials is one coalesced load for
// the block's 32 channels), then thread (0, c) adds the kRunGroups group sums
// in group order.  A fixed order, so the result is the same every run.
constexpr int kRunGroups = 8;

__global__ __launch_bounds__(kRunGroups * kC) void bwd_run_final_block_kernel(
    const uint32_t* __restrict__ ukeys, const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ piece_start,
    const uint32_t* __restrict__ nruns, const float* __restrict__ partial, float* __restrict__ grad) {
    __shared__ float gs[kRunGroups][kC];
    const uint32_t r = blockIdx.x;
    if (r >= *nruns) return;  // block-uniform
    const int c = threadIdx.x % kC;
    const int g = threadIdx.x / kC;
    const uint32_t np = (cnt[r] + kRunPiece - 1) / kRunPiece;
    const size_t p0 = piece_start[r];
    float s = 0.0f;
    for (uint32_t i = g; i < np; i += kRunGroups) s += partial[(p0 + i) * kC + c];
    gs[g][c] = s;
    __syncthreads();
    if (g == 0) {
        float t = gs[0][c];
#pragma unroll
        for (int k = 1; k < kRunGroups; ++k) t += gs[k][c];
        grad[(size_t)ukeys[r] * kC + c] = t;
    }
}

__global__ void fill_f32_kernel(float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

struct BwdLayout {
    size_t E, Z, ds, part, gpart, keys_in, keys_out, vals_in, vals_out, sort_tmp, total;
    size_t ukeys, counts, nruns, rstart, pstart, pcount, piece_part, rle_tmp;
    size_t sort_bytes, rle_bytes;
    int64_t nchunks, ngroups;
    uint32_t max_runs, max_pieces;
};

static int key_bits(int r) {
    const uint64_t corners = (uint64_t)(r + 1) * (r + 1) * (r + 1);
    int b = 1;
    while ((1ull << b) < corners) ++b;
    return b;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

static int bwd_layout(int64_t n, int lod, int din, const int32_t* res, BwdLayout* L) {
    const size_t levels = (size_t)lod + 1;
    L->nchunks = (n + kChunk - 1) / kChunk;
    const size_t outs = (size_t)kH * din + 2 * kH + 1;
    size_t off = 0;
    L->E = off;
    off = align256(off + levels * n * din * sizeof(float));
    L->Z = off;
    off = align256(off + levels * n * kH * sizeof(float));
    L->ds = off;
    off = align256(off + levels * n * kC * sizeof(float));
    L->part = off;
    off = align256(off + levels * L->nchunks * outs * sizeof(float));
    L->ngroups = (L->nchunks + kGroup - 1) / kGroup;
    L->gpart = off;
    off = align256(off + levels * L->ngroups * outs * sizeof(float));
    const size_t m = (size_t)n * 8;
    L->keys_in = off;
    off = align256(off + m * 4);
    L->keys_out = off;
    off = align256(off + m * 4);
    L->vals_in = off;
    off = align256(off + m * 4);
    L->vals_out = off;
    off = align256(off + m * 4);
    size_t tmp = 0;
    for (int j = 0; j <= lod; ++j) {
        size_t t = 0;
        if (rocprim::radix_sort_pairs((void*)nullptr, t, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                      (uint32_t*)nullptr, m, 0, key_bits(res[j]), (hipStream_t)0) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "radix sort size query failed");
        tmp = t > tmp ? t : tmp;
    }
    L->sort_tmp = off;
    L->sort_bytes = tmp;
    off = align256(off + tmp);
    // run-length encoding of the sorted keys and the run pieces
    uint64_t max_corners = 0;
    for (int j = 0; j <= lod; ++j) {
        const uint64_t c = (uint64_t)(res[j] + 1) * (res[j] + 1) * (res[j] + 1);
        max_corners = c > max_corners ? c : max_corners;
    }
    L->max_runs = (uint32_t)(max_corners < m ? max_corners : m);
    L->max_pieces = (uint32_t)(m / kRunPiece + L->max_runs);
    L->ukeys = off;
    off = align256(off + (size_t)L->max_runs * 4);
    L->counts = off;
    off = align256(off + (size_t)L->max_runs * 4);
    L->rstart = off;
    off = align256(off + (size_t)L->max_runs * 4);
    L->pcount = off;
    off = align256(off + (size_t)L->max_runs * 4);
    L->pstart = off;
    off = align256(off + (size_t)L->max_runs * 4);
    L->nruns = off;
    off = align256(off + 16);
    L->piece_part = off;
    off = align256(off + (size_t)L->max_pieces * kC * sizeof(float));
    size_t t1 = 0, t2 = 0, t3 = 0;
    if (rocprim::run_length_encode((void*)nullptr, t1, (uint32_t*)nullptr, m, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, (hipStream_t)0) != hipSuccess ||
        rocprim::exclusive_scan((void*)nullptr, t2, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u,
                                (size_t)L->max_runs, rocprim::plus<uint32_t>(), (hipStream_t)0) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "run-length / scan size query failed");
    t3 = t1 > t2 ? t1 : t2;
    L->rle_tmp = off;
    L->rle_bytes = t3;
    off = align256(off + t3);
    L->total = off;
    return NGLOD_OK;
}

}  // namespace nglod

using namespace nglod;

extern "C" {

size_t nglod_sdf_bwd_workspace_size(const nglod_model_t* model, int64_t n, int32_t lod) {
    if (!model || n < 0 || lod < 0) return 0;
    if (lod >= model->num_lods || lod >= NGLOD_MAX_LODS) return 0;
    BwdLayout L;
    const int din = model->pos_invariant ? kC : kC + 3;
    if (bwd_layout(n, lod, din, model->res, &L)) return 0;
    return L.total;
}

int nglod_sdf_bwd(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, const float* grad_out,
                  int32_t level_mask, const nglod_model_grads_t* grads, void* workspace, size_t workspace_bytes,
                  void* stream) {
    ModelArgs m;
    int rc = make_model_args(model, lod, NGLOD_FP32, &m);
    if (rc) return rc;
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n > ((int64_t)1 << 28)) return fail(NGLOD_ERR_ARG, "n too large for 32-bit sort values");
    if (!grads) return fail(NGLOD_ERR_ARG, "grads is NULL");
    const int din = m.pos_inv ? kC : kC + 3;
    hipStream_t s = (hipStream_t)stream;
    level_mask &= (1 << (lod + 1)) - 1;
    // zero every gradient this call owns
    for (int l = 0; l <= lod; ++l) {
        if (!grads->grid[l] || !aligned16(grads->grid[l])) return fail(NGLOD_ERR_ARG, "grid grad NULL or misaligned");
        const int64_t cnt = (int64_t)(m.res[l] + 1) * (m.res[l] + 1) * (m.res[l] + 1) * kC;
        hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, grads->grid[l], cnt);
    }
    for (int d = 0; d < m.num_dec; ++d) {
        if (!grads->w0[d] || !grads->b0[d] || !grads->w1[d] || !grads->b1[d])
            return fail(NGLOD_ERR_ARG, "decoder grad NULL");
        hipLaunchKernelGGL(fill_f32_kernel, dim3((kH * din + 255) / 256), dim3(256), 0, s, grads->w0[d], (int64_t)kH * din);
        hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(256), 0, s, grads->b0[d], (int64_t)kH);
        hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(256), 0, s, grads->w1[d], (int64_t)kH);
        hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(64), 0, s, grads->b1[d], (int64_t)1);
    }
    rc = check_launch("fill_f32_kernel");
    if (rc) return rc;
    if (n == 0 || level_mask == 0) return NGLOD_OK;
    if (!x || !grad_out) return fail(NGLOD_ERR_ARG, "NULL pointer");
    BwdLayout L;
    rc = bwd_layout(n, lod, din, model->res, &L);
    if (rc) return rc;
    if (!workspace || workspace_bytes < L.total || !aligned16(workspace))
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small or misaligned");
    char* ws = (char*)workspace;
    BwdArgs a;
    a.m = m;
    a.lod = lod;
    a.din = din;
    a.off = m.pos_inv ? 0 : 3;
    a.mask = level_mask;
    a.n = n;
    a.x = x;
    a.g = grad_out;
    a.E = (float*)(ws + L.E);
    a.Z = (float*)(ws + L.Z);
    a.ds = (float*)(ws + L.ds);
    float* part = (float*)(ws + L.part);
    const unsigned qblocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(bwd_query_kernel, dim3(qblocks), dim3(256), 0, s, a);
    rc = check_launch("bwd_query_kernel");
    if (rc) return rc;
    // decoder gradients
    for (int l = 0; l <= lod; ++l) {
        if (!((level_mask >> l) & 1)) continue;
        hipLaunchKernelGGL(bwd_decoder_partial_kernel, dim3((unsigned)L.nchunks), dim3(256), 0, s, a, l, part, L.nchunks);
    }
    rc = check_launch("bwd_decoder_partial_kernel");
    if (rc) return rc;
    const int outs = kH * din + 2 * kH + 1;
    float* gpart = (float*)(ws + L.gpart);
    hipLaunchKernelGGL(bwd_decoder_group_kernel, dim3((outs + 255) / 256, (unsigned)L.ngroups, (unsigned)(lod + 1)),
                       dim3(256), 0, s, a, part, L.nchunks, gpart, L.ngroups);
    rc = check_launch("bwd_decoder_group_kernel");
    if (rc) return rc;
    for (int d = 0; d < m.num_dec; ++d) {
        bool used = false;
        for (int l = 0; l <= lod; ++l) used |= ((m.num_dec == 1 ? 0 : l) == d) && ((level_mask >> l) & 1);
        if (!used) continue;
        hipLaunchKernelGGL(bwd_decoder_final_kernel, dim3((outs + 255) / 256), dim3(256), 0, s, a, d, gpart,
                           L.ngroups, grads->w0[d], grads->b0[d], grads->w1[d], grads->b1[d]);
    }
    rc = check_launch("bwd_decoder_final_kernel");
    if (rc) return rc;
    if (lod > 0) {
        hipLaunchKernelGGL(bwd_suffix_kernel, dim3((unsigned)((n * (kC / 4) + 255) / 256)), dim3(256), 0, s, a);
        rc = check_launch("bwd_suffix_kernel");
        if (rc) return rc;
    }
    // grid gradients: levels that receive gradient from some level i >= j
    uint32_t* kin = (uint32_t*)(ws + L.keys_in);
    uint32_t* kout = (uint32_t*)(ws + L.keys_out);
    uint32_t* vin = (uint32_t*)(ws + L.vals_in);
    uint32_t* vout = (uint32_t*)(ws + L.vals_out);
    const int64_t mm = n * 8;
    for (int j = 0; j <= lod; ++j) {
        if ((level_mask >> j) == 0) continue;  // no level >= j carries gradient
        hipLaunchKernelGGL(bwd_corner_keys_kernel, dim3(qblocks), dim3(256), 0, s, a, j, kin, vin);
        rc = check_launch("bwd_corner_keys_kernel");
        if (rc) return rc;
        size_t tmp = 0;
        if (rocprim::radix_sort_pairs((void*)nullptr, tmp, kin, kout, vin, vout, (size_t)mm, 0, key_bits(m.res[j]), s) !=
                hipSuccess || tmp > L.sort_bytes)
            return fail(NGLOD_ERR_HIP, "radix sort temp size");
        tmp = L.sort_bytes;
        if (rocprim::radix_sort_pairs((void*)(ws + L.sort_tmp), tmp, kin, kout, vin, vout, (size_t)mm, 0,
                                      key_bits(m.res[j]), s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "radix sort failed");
        // runs of equal corners -> pieces of kRunPiece entries -> in-order sums
        uint32_t* ukeys = (uint32_t*)(ws + L.ukeys);
        uint32_t* cnts = (uint32_t*)(ws + L.counts);
        uint32_t* rstart = (uint32_t*)(ws + L.rstart);
        uint32_t* pcount = (uint32_t*)(ws + L.pcount);
        uint32_t* pstart = (uint32_t*)(ws + L.pstart);
        uint32_t* nruns = (uint32_t*)(ws + L.nruns);
        float* ppart = (float*)(ws + L.piece_part);
        size_t t = L.rle_bytes;
        if (rocprim::run_length_encode((void*)(ws + L.rle_tmp), t, kout, (size_t)mm, ukeys, cnts, nruns, s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "run-length encode failed");
        t = L.rle_bytes;
        if (rocprim::exclusive_scan((void*)(ws + L.rle_tmp), t, cnts, rstart, 0u, (size_t)L.max_runs,
                                    rocprim::plus<uint32_t>(), s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "scan failed");
        hipLaunchKernelGGL(run_pieces_kernel, dim3((L.max_runs + 255) / 256), dim3(256), 0, s, cnts, nruns, pcount);
        rc = check_launch("run_pieces_kernel");
        if (rc) return rc;
        t = L.rle_bytes;
        if (rocprim::exclusive_scan((void*)(ws + L.rle_tmp), t, pcount, pstart, 0u, (size_t)L.max_runs,
                                    rocprim::plus<uint32_t>(), s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "scan failed");
        hipLaunchKernelGGL(bwd_piece_sum_kernel, dim3((L.max_pieces + 255) / 256), dim3(256), 0, s, a, j, vout, rstart,
                           cnts, pstart, nruns, L.max_pieces, ppart);
        rc = check_launch("bwd_piece_sum_kernel");
        if (rc) return rc;
        // levels averaging 8 or more pieces per corner: a block per run
        const int64_t corners = (int64_t)(m.res[j] + 1) * (m.res[j] + 1) * (m.res[j] + 1);
        if (mm >= corners * 8 * kRunPiece) {
            const uint32_t runs = (uint32_t)(corners < (int64_t)L.max_runs ? corners : (int64_t)L.max_runs);
            hipLaunchKernelGGL(bwd_run_final_block_kernel, dim3(runs), dim3(kRunGroups * kC), 0, s, ukeys, cnts, pstart,
                               nruns, ppart, grads->grid[j]);
            rc = check_launch("bwd_run_final_block_kernel");
        } else {
            const uint64_t fin = (uint64_t)L.max_runs * kC;
            hipLaunchKernelGGL(bwd_run_final_kernel, dim3((unsigned)((fin + 255) / 256)), dim3(256), 0, s, ukeys, cnts,
                               pstart, nruns, L.max_runs, ppart, grads->grid[j]);
            rc = check_launch("bwd_run_final_kernel");
        }
        if (rc) return rc;
    }
    return NGLOD_OK;
}

}  // extern "C"
