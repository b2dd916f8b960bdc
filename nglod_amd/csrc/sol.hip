// Sparse ("SOL") model and its sphere march.
//   nglod_sol_corners    unique corners of the occupied voxels per level
//                        (geoutils.unique_corners, sdf-net/lib/geoutils.py:60-77)
//   nglod_sol_features   corner features from the dense volumes (SOL_NGLOD.py:64-70)
//   nglod_sol_trinkets   per voxel: 8 corner indices + parent (index_trinket.cuh:31-98),
//                        by binary search instead of the reference's O(n*m) scan
//   nglod_trace_sparse   SDF::sphereTrace + getNormal (sol-renderer/SDF.cu:218-472)
//
// The march is one persistent kernel: each lane owns a ray (refilled from a
// queue when it retires) and runs the reference's per-ray recurrence: enter
// the first voxel of its nugget list that contains or lies ahead of the
// current point, evaluate the sparse SDF there (features summed over the
// voxel's ancestors, decoder in registers), step, re-enter, at most 50 times;
// then six taps for the normal.  The (n, 35) feature matrix the reference
// writes to HBM and re-reads in addmm never exists here.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "api_common.hpp"
#include "fp32_wave.hpp"
#include "trace_common.hpp"

namespace nglod {

__host__ __device__ __forceinline__ uint64_t morton_u64(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t m = 0;
    for (int i = 0; i < 16; ++i) {
        m |= (uint64_t)((z >> i) & 1u) << (3 * i);
        m |= (uint64_t)((y >> i) & 1u) << (3 * i + 1);
        m |= (uint64_t)((x >> i) & 1u) << (3 * i + 2);
    }
    return m;
}

struct SolArgs {
    int target, lod, din;
    const int16_t* points;
    int64_t psum_target;  // global index of the first point of level `target`
    const uint32_t* tv;   // [P][8]
    const int32_t* parent;
    const int32_t* cc;    // [Nc][3]
    const float* cf;      // [Nc][32]
    const float* w0;
    const float* b0;
    const float* w1;
    const float* b1;
    // rays
    const float* o;
    const float* d;
    int64_t n;
    const int32_t* offs;  // [n+1] CSR over nuggets
    const int32_t* nug;   // [M][2]
    float* x;
    float* t;
    uint8_t* hit;
    float* normal;
    int32_t* iters;
    uint32_t* queue;
    unsigned long long* evals;
    float r;  // voxel half size at `target`
};

__device__ __forceinline__ float sol_ray_aabb(float qx, float qy, float qz, float dx, float dy, float dz, float ivx,
                                              float ivy, float ivz, float sx_, float sy_, float sz_, float cx, float cy,
                                              float cz, float r) {
    const float ox = qx - cx, oy = qy - cy, oz = qz - cz;
    const float cmax = fmaxf(fmaxf(fabsf(ox), fabsf(oy)), fabsf(oz));
    float winding = cmax < r ? -1.0f : 1.0f;
    winding *= r;
    if (winding < 0) return winding;
    const float d0 = __builtin_fmaf(winding, sx_, -ox) * ivx;
    const float d1 = __builtin_fmaf(winding, sy_, -oy) * ivy;
    const float d2 = __builtin_fmaf(winding, sz_, -oz) * ivz;
    const float ltxy = __builtin_fmaf(dy, d0, oy), ltxz = __builtin_fmaf(dz, d0, oz);
    const float ltyx = __builtin_fmaf(dx, d1, ox), ltyz = __builtin_fmaf(dz, d1, oz);
    const float ltzx = __builtin_fmaf(dx, d2, ox), ltzy = __builtin_fmaf(dy, d2, oy);
    const bool t0 = (d0 >= 0.0f) && (fabsf(ltxy) < r) && (fabsf(ltxz) < r);
    const bool t1 = (d1 >= 0.0f) && (fabsf(ltyx) < r) && (fabsf(ltyz) < r);
    const bool t2 = (d2 >= 0.0f) && (fabsf(ltzx) < r) && (fabsf(ltzy) < r);
    float ssx = 0.0f, ssy = 0.0f, ssz = 0.0f;
    if (t0) ssx = sx_;
    else if (t1) ssy = sy_;
    else if (t2) ssz = sz_;
    float dd = 0.0f;
    if (ssx != 0.0f) dd = d0;
    else if (ssy != 0.0f) dd = d1;
    else if (ssz != 0.0f) dd = d2;
    return dd;
}

// Sparse features at p, voxel cidx (global point index) and its ancestors,
// finest level first, in the reference's term order (no contraction).
__device__ __forceinline__ void sol_features(const SolArgs& a, float px, float py, float pz, int64_t cidx,
                                             float (&f)[kC]) {
#pragma unroll
    for (int k = 0; k < kC; ++k) f[k] = 0.0f;
    const float nx = __builtin_fmaf(px, 0.5f, 0.5f), ny = __builtin_fmaf(py, 0.5f, 0.5f), nz = __builtin_fmaf(pz, 0.5f, 0.5f);
    for (int lv = a.lod; lv >= 0; --lv) {
        if (lv != a.lod) cidx = a.parent[cidx];
        const float res = (float)(4u << lv);
        const float iz = nx * res, iy = ny * res, ix = nz * res;
        const uint4 va = reinterpret_cast<const uint4*>(a.tv + 8 * cidx)[0];
        const uint4 vb = reinterpret_cast<const uint4*>(a.tv + 8 * cidx)[1];
        const uint32_t v000 = va.x, v100 = va.y, v010 = va.z, v110 = va.w;
        const uint32_t v001 = vb.x, v101 = vb.y, v011 = vb.z, v111 = vb.w;
#define CCX(q) ((float)a.cc[3 * (q)])
#define CCY(q) ((float)a.cc[3 * (q) + 1])
#define CCZ(q) ((float)a.cc[3 * (q) + 2])
        const float w0 = ((CCZ(v111) - ix) * (CCY(v111) - iy)) * (CCX(v111) - iz);
        const float w1 = ((CCZ(v110) - ix) * (CCY(v110) - iy)) * (iz - CCX(v110));
        const float w2 = ((CCZ(v101) - ix) * (iy - CCY(v101))) * (CCX(v101) - iz);
        const float w3 = ((CCZ(v100) - ix) * (iy - CCY(v100))) * (iz - CCX(v100));
        const float w4 = ((ix - CCZ(v011)) * (CCY(v011) - iy)) * (CCX(v011) - iz);
        const float w5 = ((ix - CCZ(v010)) * (CCY(v010) - iy)) * (iz - CCX(v010));
        const float w6 = ((ix - CCZ(v001)) * (iy - CCY(v001))) * (CCX(v001) - iz);
        const float w7 = ((ix - CCZ(v000)) * (iy - CCY(v000))) * (iz - CCX(v000));
#undef CCX
#undef CCY
#undef CCZ
        const float* c0 = a.cf + (size_t)v000 * kC;
        const float* c1 = a.cf + (size_t)v001 * kC;
        const float* c2 = a.cf + (size_t)v010 * kC;
        const float* c3 = a.cf + (size_t)v011 * kC;
        const float* c4 = a.cf + (size_t)v100 * kC;
        const float* c5 = a.cf + (size_t)v101 * kC;
        const float* c6 = a.cf + (size_t)v110 * kC;
        const float* c7 = a.cf + (size_t)v111 * kC;
#pragma unroll
        for (int k = 0; k < kC; ++k) {
            // The reference rounds ((f*wa)*wb)*wc per term; here the weight
            // product is formed once per corner and fused into the sum.
            f[k] = __builtin_fmaf(c0[k], w0, f[k]);
            f[k] = __builtin_fmaf(c1[k], w1, f[k]);
            f[k] = __builtin_fmaf(c2[k], w2, f[k]);
            f[k] = __builtin_fmaf(c3[k], w3, f[k]);
            f[k] = __builtin_fmaf(c4[k], w4, f[k]);
            f[k] = __builtin_fmaf(c5[k], w5, f[k]);
            f[k] = __builtin_fmaf(c6[k], w6, f[k]);
            f[k] = __builtin_fmaf(c7[k], w7, f[k]);
        }
    }
}

__device__ __forceinline__ float sol_decode(const SolArgs& a, float px, float py, float pz, const float (&f)[kC]) {
    const int off = a.din - kC;
    float out = 0.0f;
    for (int j = 0; j < kH; ++j) {
        float acc = 0.0f;
        if (off) {
            acc = __builtin_fmaf(ldc(a.w0, j * a.din + 0), px, acc);
            acc = __builtin_fmaf(ldc(a.w0, j * a.din + 1), py, acc);
            acc = __builtin_fmaf(ldc(a.w0, j * a.din + 2), pz, acc);
        }
#pragma unroll
        for (int k = 0; k < kC; ++k) acc = __builtin_fmaf(ldc(a.w0, j * a.din + off + k), f[k], acc);
        const float h = fmaxf(acc + ldc(a.b0, j), 0.0f);
        out = __builtin_fmaf(ldc(a.w1, j), h, out);
    }
    return out + ldc(a.b1, 0);
}

__device__ __forceinline__ float sol_eval(const SolArgs& a, float px, float py, float pz, int64_t pidx_local) {
    float f[kC];
    sol_features(a, px, py, pz, pidx_local + a.psum_target, f);
    return sol_decode(a, px, py, pz, f);
}

enum : int { SP_IDLE = -1, SP_EVAL = 0, SP_N0 = 1 };

struct SolRay {
    int64_t ray;
    int phase;
    int it;
    bool hitf;
    int64_t pidx;
    int32_t nb, ne;
    float ox, oy, oz, dx, dy, dz, ivx, ivy, ivz, sgx, sgy, sgz;
    float x0, x1, x2, t, dprev;
    float qx, qy, qz, fplus, g0, g1, g2;

    // First nugget voxel that contains (x0,x1,x2) or lies ahead; false: miss.
    __device__ __forceinline__ bool walk(const SolArgs& a) {
        for (int32_t i = nb; i < ne; ++i) {
            const int32_t pl = a.nug[2 * i + 1];
            const int16_t* p = a.points + 4 * (a.psum_target + pl);
            const float cx = __builtin_fmaf(a.r, __builtin_fmaf(2.0f, (float)p[0], 1.0f), -1.0f);
            const float cy = __builtin_fmaf(a.r, __builtin_fmaf(2.0f, (float)p[1], 1.0f), -1.0f);
            const float cz = __builtin_fmaf(a.r, __builtin_fmaf(2.0f, (float)p[2], 1.0f), -1.0f);
            const float dd = sol_ray_aabb(x0, x1, x2, dx, dy, dz, ivx, ivy, ivz, sgx, sgy, sgz, cx, cy, cz, a.r);
            if (dd != 0.0f) {
                pidx = pl;
                if (dd > 0.0f) {
                    t += dd;
                    x0 = __builtin_fmaf(dx, t, ox);
                    x1 = __builtin_fmaf(dy, t, oy);
                    x2 = __builtin_fmaf(dz, t, oz);
                }
                return true;
            }
        }
        t = 100.0f;
        x0 = __builtin_fmaf(dx, t, ox);
        x1 = __builtin_fmaf(dy, t, oy);
        x2 = __builtin_fmaf(dz, t, oz);
        return false;
    }

    __device__ __forceinline__ void retire(const SolArgs& a, float n0, float n1, float n2) {
        a.x[3 * ray] = x0;
        a.x[3 * ray + 1] = x1;
        a.x[3 * ray + 2] = x2;
        a.t[ray] = t;
        a.hit[ray] = hitf ? 1 : 0;
        a.normal[3 * ray] = n0;
        a.normal[3 * ray + 1] = n1;
        a.normal[3 * ray + 2] = n2;
        if (a.iters) a.iters[ray] = it;
        phase = SP_IDLE;
        ray = -1;
    }

    __device__ __forceinline__ void finish(const SolArgs& a) {
        if (hitf) {
            phase = SP_N0;
            qx = x0 + 0.001f;
            qy = x1 + 0.0f;
            qz = x2 + 0.0f;
        } else {
            retire(a, 1.0f, 1.0f, 1.0f);
        }
    }

    __device__ __forceinline__ void start(const SolArgs& a, int64_t r) {
        ray = r;
        ox = a.o[3 * r];
        oy = a.o[3 * r + 1];
        oz = a.o[3 * r + 2];
        dx = a.d[3 * r];
        dy = a.d[3 * r + 1];
        dz = a.d[3 * r + 2];
        ivx = 1.0f / dx;
        ivy = 1.0f / dy;
        ivz = 1.0f / dz;
        sgx = __builtin_signbit(dx) ? 1.0f : -1.0f;
        sgy = __builtin_signbit(dy) ? 1.0f : -1.0f;
        sgz = __builtin_signbit(dz) ? 1.0f : -1.0f;
        x0 = ox;
        x1 = oy;
        x2 = oz;
        t = 0.0f;
        dprev = 0.0f;
        it = 0;
        hitf = false;
        pidx = -1;
        nb = a.offs[r];
        ne = a.offs[r + 1];
        const bool cond = ne > nb && walk(a);
        if (cond) {
            phase = SP_EVAL;
            qx = x0;
            qy = x1;
            qz = x2;
        } else {
            finish(a);
        }
    }

    __device__ __forceinline__ void consume(const SolArgs& a, float d) {
        if (phase == SP_EVAL) {
            ++it;
            t = t + d;
            hitf = ((double)fabsf(d) < 0.0003) || (((double)fabsf(d + dprev) * 0.5) < 0.0015);
            bool cond = ((double)t < 5.0) && !hitf;
            x0 = __builtin_fmaf(dx, t, ox);
            x1 = __builtin_fmaf(dy, t, oy);
            x2 = __builtin_fmaf(dz, t, oz);
            dprev = d;
            if (cond) cond = walk(a);
            if (cond && it < 50) {
                qx = x0;
                qy = x1;
                qz = x2;
            } else {
                finish(a);
            }
            return;
        }
        const int k = phase - SP_N0;  // 0..5
        const int axis = k >> 1;
        if ((k & 1) == 0) {
            fplus = d;
            qx = axis == 0 ? x0 - 0.001f : x0 - 0.0f;
            qy = axis == 1 ? x1 - 0.001f : x1 - 0.0f;
            qz = axis == 2 ? x2 - 0.001f : x2 - 0.0f;
            ++phase;
            return;
        }
        const float g = fplus - d;
        if (axis == 0) g0 = g;
        else if (axis == 1) g1 = g;
        else g2 = g;
        if (axis == 2) {
            const float len = sqrtf(g0 * g0 + g1 * g1 + g2 * g2);
            retire(a, g0 / len, g1 / len, g2 / len);
        } else {
            const int na = axis + 1;
            qx = na == 0 ? x0 + 0.001f : x0 + 0.0f;
            qy = na == 1 ? x1 + 0.001f : x1 + 0.0f;
            qz = na == 2 ? x2 + 0.001f : x2 + 0.0f;
            ++phase;
        }
    }
};

// Persistent sparse march: one ray per lane (SolRay), refilled from kQueues
// ray ranges (RayFeed); the features come from the sparse corner table, the
// decoder is the FP32 MFMA one of the dense form (fp32_wave.hpp), staged in
// LDS once per block.  512-thread blocks, as the dense FP32 march.
// 512 and 256 threads measured the same (3.23 / 3.25 ms on the 1280x720 LOD-4
// frame), 768 slower (3.62 ms).
constexpr int kSparseBlock = 512;

template <bool POSINV>
__global__ __launch_bounds__(kSparseBlock) void trace_sparse_kernel(SolArgs a) {
    __shared__ Fp32Lds lds;
    stage_fp32_lds_raw(a.w0, a.din, a.b0, a.w1, &lds);
    __syncthreads();
    const float b1 = a.b1[0];
    const uint32_t lane = lane_id();
    RayFeed feed;
    feed.init(a.n, kQueueStride);
    SolRay s;
    s.ray = -1;
    s.phase = SP_IDLE;
    uint32_t nevals = 0;
    for (;;) {
        // a ray without nuggets finishes at once and its lane takes the next
        for (;;) {
            const bool need = s.phase == SP_IDLE;
            if (__ballot(need && !feed.empty()) == 0) break;
            const int64_t id = feed.next<0>(a.queue, nullptr, lane, need);
            if (id >= 0) s.start(a, id);
        }
        if (__ballot(s.phase != SP_IDLE) == 0) break;
        const bool busy = s.phase != SP_IDLE;
        float f[kC];
        if (busy) {
            sol_features(a, s.qx, s.qy, s.qz, s.pidx + a.psum_target, f);
        } else {
#pragma unroll
            for (int k = 0; k < kC; ++k) f[k] = 0.0f;
        }
        const float d = decode_fp32_wave<POSINV>(f, &lds, b1, lane, busy, s.qx, s.qy, s.qz);
        if (busy) {
            ++nevals;
            s.consume(a, d);
        }
    }
    if (a.evals) {
        unsigned long long tot = nevals;
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
        if (lane == 0 && tot) atomicAdd(a.evals, tot);
    }
}

__global__ __launch_bounds__(256) void sol_sdf_kernel(SolArgs a, const float* __restrict__ x,
                                                      const int32_t* __restrict__ pidx, int64_t n,
                                                      float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = sol_eval(a, x[3 * i], x[3 * i + 1], x[3 * i + 2], pidx[i]);
}

// Leaf code at `level` of the cell holding each point:
// floor((x + 1) / 2 * 2^level) per axis (renderutils.py:67-70), clamped to
// the grid so that a point on the +1 face lands in the last cell.
__global__ void point_codes_kernel(const float* __restrict__ x, const uint8_t* __restrict__ mask, int64_t n, int level,
                                   uint32_t* __restrict__ codes) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float res = (float)(1 << level);
    uint32_t c[3];
    for (int k = 0; k < 3; ++k) {
        const float v = floorf(((x[3 * i + k] + 1.0f) / 2.0f) * res);
        const float cl = fminf(fmaxf(v, 0.0f), res - 1.0f);
        c[k] = (uint32_t)cl;
    }
    const bool keep = mask ? mask[i] != 0 : true;
    codes[i] = keep ? (uint32_t)morton_u64(c[0], c[1], c[2]) : 0xFFFFFFFFu;
}

// Zeroes kQueues counters of a workspace, `stride` words apart.
__global__ void clear_counters_kernel(uint32_t* q, int stride) {
    if (threadIdx.x < kQueues) q[threadIdx.x * stride] = 0;
}

// The sparse march's workspace: its claim counters, one 128-byte line each.
constexpr size_t kSparseWorkspace = (size_t)kQueues * kQueueStride * sizeof(uint32_t);

// ---------------------------------------------------------------------------
// sparse structure builders
__global__ void corner_keys_kernel(const int16_t* __restrict__ pts, int64_t n, uint32_t R1, uint32_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = (uint16_t)pts[4 * i], y = (uint16_t)pts[4 * i + 1], z = (uint16_t)pts[4 * i + 2];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t cx = x + ((j & 4) >> 2), cy = y + ((j & 2) >> 1), cz = z + (j & 1);
        keys[8 * i + j] = (cx * R1 + cy) * R1 + cz;
    }
}

__global__ void decode_corners_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ nu, uint32_t R1,
                                      int32_t* __restrict__ cc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nu) return;
    const uint32_t k = keys[i];
    cc[3 * i] = (int32_t)(k / (R1 * R1));
    cc[3 * i + 1] = (int32_t)((k / R1) % R1);
    cc[3 * i + 2] = (int32_t)(k % R1);
}

__global__ void gather_cf_kernel(const float* __restrict__ grid, int r, const int32_t* __restrict__ cc, int64_t n,
                                 float* __restrict__ cf) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = t / kC;
    const int c = (int)(t % kC);
    if (i >= n) return;
    const int S = r + 1;
    const int32_t x = cc[3 * i], y = cc[3 * i + 1], z = cc[3 * i + 2];
    cf[i * kC + c] = grid[(((int64_t)z * S + y) * S + x) * kC + c];
}

__device__ __forceinline__ int cmp3i(const int32_t* a, uint32_t x, uint32_t y, uint32_t z) {
    if ((uint32_t)a[0] != x) return (uint32_t)a[0] < x ? -1 : 1;
    if ((uint32_t)a[1] != y) return (uint32_t)a[1] < y ? -1 : 1;
    if ((uint32_t)a[2] != z) return (uint32_t)a[2] < z ? -1 : 1;
    return 0;
}

__global__ void trinket_kernel(const int16_t* __restrict__ pts, int64_t beg, int64_t n, const int32_t* __restrict__ cc,
                               int64_t cc_off, int64_t cc_cnt, int64_t pbeg, int64_t pcnt, int has_parent,
                               uint32_t* __restrict__ tv, int32_t* __restrict__ parent, uint32_t* __restrict__ missing) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t g = beg + k;
    const uint32_t x = (uint16_t)pts[4 * g], y = (uint16_t)pts[4 * g + 1], z = (uint16_t)pts[4 * g + 2];
    for (int j = 0; j < 8; ++j) {
        const uint32_t cx = x + ((j & 4) >> 2), cy = y + ((j & 2) >> 1), cz = z + (j & 1);
        int64_t lo = 0, hi = cc_cnt - 1;
        uint32_t found = 0xFFFFFFFFu;
        while (lo <= hi) {
            const int64_t mid = (lo + hi) >> 1;
            const int c = cmp3i(cc + 3 * (cc_off + mid), cx, cy, cz);
            if (c == 0) {
                found = (uint32_t)(cc_off + mid);
                break;
            }
            if (c < 0) lo = mid + 1;
            else hi = mid - 1;
        }
        tv[8 * g + j] = found;
        if (found == 0xFFFFFFFFu) atomicAdd(missing, 1u);
    }
    int32_t par = -1;
    if (has_parent) {
        const uint64_t key = morton_u64(x >> 1, y >> 1, z >> 1);
        int64_t lo = 0, hi = pcnt - 1;
        while (lo <= hi) {
            const int64_t mid = (lo + hi) >> 1;
            const int16_t* pp = pts + 4 * (pbeg + mid);
            const uint64_t mk = morton_u64((uint16_t)pp[0], (uint16_t)pp[1], (uint16_t)pp[2]);
            if (mk == key) {
                par = (int32_t)(pbeg + mid);
                break;
            }
            if (mk < key) lo = mid + 1;
            else hi = mid - 1;
        }
    }
    parent[g] = par;
    if (has_parent && par < 0) atomicAdd(missing, 1u);
}

static size_t a256s(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace nglod

using namespace nglod;

extern "C" {

int nglod_sol_corners(const int16_t* points, const int32_t* pyramid, int32_t level, int32_t target, int32_t* cc,
                      int64_t cap, int64_t* counts, void* workspace, size_t workspace_bytes, size_t* workspace_needed,
                      void* stream) {
    if (!pyramid || !counts || !workspace_needed) return fail(NGLOD_ERR_ARG, "NULL out-parameter");
    if (target < 2 || target > level || target > 10) return fail(NGLOD_ERR_ARG, "target level out of range");
    hipStream_t s = (hipStream_t)stream;
    const int32_t* pcnt = pyramid;
    const int32_t* psum = pyramid + (level + 2);
    int64_t maxn = 0;
    for (int l = 2; l <= target; ++l) maxn = pcnt[l] > maxn ? pcnt[l] : maxn;
    const size_t m = (size_t)maxn * 8;
    size_t sort_tmp = 0, uniq_tmp = 0;
    if (rocprim::radix_sort_keys((void*)nullptr, sort_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, m, 0, 32, s) !=
            hipSuccess ||
        rocprim::unique((void*)nullptr, uniq_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, m,
                        rocprim::equal_to<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "size query");
    const size_t o_a = 0, o_b = a256s(m * 4), o_n = o_b + a256s(m * 4), o_t = o_n + 256;
    const size_t need = o_t + a256s(sort_tmp > uniq_tmp ? sort_tmp : uniq_tmp);
    if (!workspace || workspace_bytes < need) {
        *workspace_needed = need;
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    }
    char* ws = (char*)workspace;
    uint32_t* A = (uint32_t*)(ws + o_a);
    uint32_t* B = (uint32_t*)(ws + o_b);
    uint32_t* nu = (uint32_t*)(ws + o_n);
    int64_t off = 0;
    for (int l = 2; l <= target; ++l) {
        const int64_t n = pcnt[l];
        const uint32_t R1 = (1u << l) + 1u;
        if (n == 0) {
            counts[l - 2] = 0;
            continue;
        }
        hipLaunchKernelGGL(corner_keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, points + 4 * (int64_t)psum[l],
                           n, R1, A);
        int rc = check_launch("corner_keys_kernel");
        if (rc) return rc;
        size_t t = sort_tmp;
        if (rocprim::radix_sort_keys((void*)(ws + o_t), t, A, B, (size_t)n * 8, 0, 32, s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "sort failed");
        t = uniq_tmp;
        if (rocprim::unique((void*)(ws + o_t), t, B, A, nu, (size_t)n * 8, rocprim::equal_to<uint32_t>(), s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "unique failed");
        uint32_t u = 0;
        if (hipMemcpyAsync(&u, nu, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "read failed");
        counts[l - 2] = u;
        if (off + (int64_t)u <= cap && cc) {
            hipLaunchKernelGGL(decode_corners_kernel, dim3((u + 255) / 256), dim3(256), 0, s, A, nu, R1, cc + 3 * off);
            rc = check_launch("decode_corners_kernel");
            if (rc) return rc;
        }
        off += u;
    }
    if (off > cap || !cc) return fail(NGLOD_ERR_CAPACITY, "corner capacity too small");
    return hipStreamSynchronize(s) == hipSuccess ? NGLOD_OK : fail(NGLOD_ERR_HIP, "sync failed");
}

int nglod_sol_features(const nglod_model_t* model, int32_t target, const int32_t* cc, const int64_t* counts, float* cf,
                       void* stream) {
    if (!model || !cc || !counts || !cf) return fail(NGLOD_ERR_ARG, "NULL pointer");
    if (target < 2 || target - 2 >= model->num_lods) return fail(NGLOD_ERR_ARG, "target level out of range");
    if (model->feature_dim != kC) return fail(NGLOD_ERR_UNSUPPORTED, "feature_dim must be 32");
    hipStream_t s = (hipStream_t)stream;
    int64_t off = 0;
    for (int l = 2; l <= target; ++l) {
        const int lod = l - 2;
        if (model->res[lod] != (1 << l)) return fail(NGLOD_ERR_UNSUPPORTED, "LOD i must have resolution 4*2^i");
        const int64_t n = counts[lod];
        const int64_t tot = n * kC;
        if (n > 0)
            hipLaunchKernelGGL(gather_cf_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, model->grid[lod],
                               model->res[lod], cc + 3 * off, n, cf + off * kC);
        off += n;
    }
    return check_launch("gather_cf_kernel");
}

int nglod_sol_trinkets(const int16_t* points, const int32_t* pyramid, int32_t level, int32_t target, const int32_t* cc,
                       const int64_t* counts, uint32_t* trinket_v, int32_t* parent, void* workspace,
                       size_t workspace_bytes, void* stream) {
    if (!points || !pyramid || !cc || !counts || !trinket_v || !parent) return fail(NGLOD_ERR_ARG, "NULL pointer");
    if (target < 2 || target > level) return fail(NGLOD_ERR_ARG, "target level out of range");
    if (!workspace || workspace_bytes < 256) return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    const int32_t* pcnt = pyramid;
    const int32_t* psum = pyramid + (level + 2);
    uint32_t* missing = (uint32_t*)workspace;
    hipLaunchKernelGGL(clear_counters_kernel, dim3(1), dim3(64), 0, s, missing, 1);
    int rc = check_launch("clear_counters_kernel");
    if (rc) return rc;
    int64_t off = 0;
    for (int l = 2; l <= target; ++l) {
        const int64_t n = pcnt[l];
        if (n > 0)
            hipLaunchKernelGGL(trinket_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, points, (int64_t)psum[l], n,
                               cc, off, counts[l - 2], (int64_t)psum[l - 1], (int64_t)pcnt[l - 1], l > 2 ? 1 : 0,
                               trinket_v, parent, missing);
        rc = check_launch("trinket_kernel");
        if (rc) return rc;
        off += counts[l - 2];
    }
    uint32_t miss = 0;
    if (hipMemcpyAsync(&miss, missing, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "read failed");
    if (miss) return fail(NGLOD_ERR_ARG, "a cell corner or parent is missing from the corner list / octree");
    return NGLOD_OK;
}

static int sol_args(const nglod_sol_t* sol, SolArgs& a) {
    if (!sol) return fail(NGLOD_ERR_ARG, "sol is NULL");
    if (sol->target < 2 || sol->target > 12 || sol->lod != sol->target - 2)
        return fail(NGLOD_ERR_ARG, "target must be lod + 2");
    if (sol->din != kC && sol->din != kC + 3) return fail(NGLOD_ERR_UNSUPPORTED, "decoder input must be 32 or 35");
    if (!sol->points || !sol->trinket_v || !sol->parent || !sol->cc || !sol->cf || !sol->w0 || !sol->b0 || !sol->w1 ||
        !sol->b1)
        return fail(NGLOD_ERR_ARG, "NULL model pointer");
    if (((uintptr_t)sol->trinket_v & 15) != 0) return fail(NGLOD_ERR_ARG, "trinkets not 16-byte aligned");
    if (sol->psum_target < 0) return fail(NGLOD_ERR_ARG, "psum_target < 0");
    memset(&a, 0, sizeof(a));
    a.target = sol->target;
    a.lod = sol->lod;
    a.din = sol->din;
    a.points = sol->points;
    a.psum_target = sol->psum_target;
    a.tv = sol->trinket_v;
    a.parent = sol->parent;
    a.cc = sol->cc;
    a.cf = sol->cf;
    a.w0 = sol->w0;
    a.b0 = sol->b0;
    a.w1 = sol->w1;
    a.b1 = sol->b1;
    a.r = (float)(1.0 / (double)(1 << sol->target));
    return NGLOD_OK;
}

int nglod_sol_sdf(const nglod_sol_t* sol, const float* x, const int32_t* pidx, int64_t n, float* out, void* stream) {
    SolArgs a;
    int rc = sol_args(sol, a);
    if (rc) return rc;
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n == 0) return NGLOD_OK;
    if (!x || !pidx || !out) return fail(NGLOD_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(sol_sdf_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a, x, pidx,
                       n, out);
    return check_launch("sol_sdf_kernel");
}

size_t nglod_trace_sparse_workspace_size(int64_t n) {
    (void)n;
    return kSparseWorkspace;
}

int nglod_trace_sparse(const nglod_sol_t* sol, const float* ray_o, const float* ray_d, int64_t n,
                       const int32_t* ray_offsets, const int32_t* nuggets, float* x, float* t, uint8_t* hit,
                       float* normal, int32_t* iters, void* workspace, size_t workspace_bytes, uint64_t* evals_out,
                       void* stream) {
    SolArgs a;
    int rc = sol_args(sol, a);
    if (rc) return rc;
    if (n < 0 || n > (int64_t)0x7fff0000) return fail(NGLOD_ERR_ARG, "n out of range");
    if (n == 0) return NGLOD_OK;
    if (!ray_o || !ray_d || !ray_offsets || !x || !t || !hit || !normal) return fail(NGLOD_ERR_ARG, "NULL pointer");
    if (!workspace || workspace_bytes < kSparseWorkspace) return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    a.o = ray_o;
    a.d = ray_d;
    a.n = n;
    a.offs = ray_offsets;
    a.nug = nuggets;
    a.x = x;
    a.t = t;
    a.hit = hit;
    a.normal = normal;
    a.iters = iters;
    a.queue = (uint32_t*)workspace;
    a.evals = (unsigned long long*)evals_out;
    hipLaunchKernelGGL(clear_counters_kernel, dim3(1), dim3(64), 0, s, a.queue, kQueueStride);
    rc = check_launch("clear_counters_kernel");
    if (rc) return rc;
    const bool posinv = a.din == kC;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, posinv ? trace_sparse_kernel<true> : trace_sparse_kernel<false>, kSparseBlock, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    const int64_t want = (n + kSparseBlock - 1) / kSparseBlock;
    const int64_t cap = (int64_t)device_cu_count() * per_cu;
    const int grid = (int)(want < cap ? want : cap);
    if (posinv)
        hipLaunchKernelGGL(trace_sparse_kernel<true>, dim3(grid), dim3(kSparseBlock), 0, s, a);
    else
        hipLaunchKernelGGL(trace_sparse_kernel<false>, dim3(grid), dim3(kSparseBlock), 0, s, a);
    return check_launch("trace_sparse_kernel");
}

int nglod_spc_codes_from_points(const float* x, const uint8_t* mask, int64_t n, int32_t level, uint32_t* codes,
                                int64_t cap, int64_t* count, void* workspace, size_t workspace_bytes,
                                size_t* workspace_needed, void* stream) {
    if (!count || !workspace_needed) return fail(NGLOD_ERR_ARG, "NULL out-parameter");
    if (level < 1 || level > 10) return fail(NGLOD_ERR_ARG, "level must be 1..10");
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    *count = 0;
    if (n == 0) return NGLOD_OK;
    if (!x) return fail(NGLOD_ERR_ARG, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    size_t sort_tmp = 0, uniq_tmp = 0;
    if (rocprim::radix_sort_keys((void*)nullptr, sort_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)n, 0, 32,
                                 s) != hipSuccess ||
        rocprim::unique((void*)nullptr, uniq_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                        (size_t)n, rocprim::equal_to<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "size query");
    const size_t m = (size_t)n;
    const size_t o_a = 0, o_b = a256s(m * 4), o_n = o_b + a256s(m * 4), o_t = o_n + 256;
    const size_t need = o_t + a256s(sort_tmp > uniq_tmp ? sort_tmp : uniq_tmp);
    if (!workspace || workspace_bytes < need) {
        *workspace_needed = need;
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    }
    char* ws = (char*)workspace;
    uint32_t* A = (uint32_t*)(ws + o_a);
    uint32_t* B = (uint32_t*)(ws + o_b);
    uint32_t* nu = (uint32_t*)(ws + o_n);
    hipLaunchKernelGGL(point_codes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, mask, n, level, A);
    int rc = check_launch("point_codes_kernel");
    if (rc) return rc;
    size_t tb = sort_tmp;
    if (rocprim::radix_sort_keys((void*)(ws + o_t), tb, A, B, m, 0, 32, s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "sort failed");
    tb = uniq_tmp;
    if (rocprim::unique((void*)(ws + o_t), tb, B, A, nu, m, rocprim::equal_to<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "unique failed");
    uint32_t u = 0;
    if (hipMemcpyAsync(&u, nu, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "read failed");
    uint32_t last = 0;
    if (u > 0 && (hipMemcpyAsync(&last, A + (u - 1), 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
                  hipStreamSynchronize(s) != hipSuccess))
        return fail(NGLOD_ERR_HIP, "read failed");
    if (u > 0 && last == 0xFFFFFFFFu) --u;  // masked points sort last
    *count = u;
    if ((int64_t)u > cap || (!codes && u > 0)) return fail(NGLOD_ERR_CAPACITY, "code capacity too small");
    if (u > 0 && hipMemcpyAsync(codes, A, (size_t)u * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "copy failed");
    return hipStreamSynchronize(s) == hipSuccess ? NGLOD_OK : fail(NGLOD_ERR_HIP, "sync failed");
}

}  // extern "C"
