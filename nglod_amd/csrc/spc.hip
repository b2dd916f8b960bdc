// Sparse octree (SPC) on the GPU: occupancy from a mesh, octree bytes from
// Morton codes, points from bytes, and ray-octree traversal to the
// (ray, voxel) "nugget" list.
//
//   nglod_spc_occupancy    level-L cells a triangle soup overlaps (separating
//                          axes, FP64), sorted unique Morton codes
//   nglod_spc_from_codes   SPC3D.points_to_nodes (sdf-net/lib/spc3d.py:224-255)
//   nglod_spc_points       SPC::SetGeometry (sol-renderer/include/spc/spc/SPC.cu:148-238)
//   nglod_spc_raytrace     spc_raytrace (spc_raytrace_cuda.cpp:141-199,
//                          spc_raytrace_cuda_kernel.cu:194-265)
//   nglod_mark_unique_rays d_MarkUniqueRays (sol-renderer/sdfRenderer.cu:108-120)
//
// Traversal: the reference expands the whole (ray, cell) frontier one level at
// a time (scan + host sync per level) and keeps, per ray, the children of each
// hit cell in the Order table's front-to-back order.  The resulting list is
// the depth-first order of each ray's hit cells with children visited in
// Order, so here every lane runs that depth-first walk for one ray with an
// explicit stack in LDS: a counting pass, an exclusive scan of the per-ray
// counts, and an emitting pass that writes each ray's nuggets at its offset.
// No host sync inside the traversal; the face tests use the reference's FP32
// operation order (no contraction), so the list is bit-identical.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "api_common.hpp"

namespace nglod {

constexpr int kSpcMaxLevel = 12;
constexpr int kStack = 1 + 7 * kSpcMaxLevel + 3;  // deepest DFS stack
constexpr int kDfsBlock = 64;

__constant__ uint8_t kOrder[8][8] = {
    {0, 1, 2, 4, 3, 5, 6, 7}, {1, 0, 3, 5, 2, 4, 7, 6}, {2, 0, 3, 6, 1, 4, 7, 5}, {3, 1, 2, 7, 0, 5, 6, 4},
    {4, 0, 5, 6, 1, 2, 7, 3}, {5, 1, 4, 7, 0, 3, 6, 2}, {6, 2, 4, 7, 0, 3, 5, 1}, {7, 3, 5, 6, 1, 2, 4, 0}};

__host__ __device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t m = 0;
    for (int i = 0; i < 10; ++i) {
        m |= ((z >> i) & 1u) << (3 * i);
        m |= ((y >> i) & 1u) << (3 * i + 1);
        m |= ((x >> i) & 1u) << (3 * i + 2);
    }
    return m;
}

// ---------------------------------------------------------------------------
// occupancy
__device__ __forceinline__ void sub3d(const double* a, const double* b, double* r) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
__device__ __forceinline__ void crs3d(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3d(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Separating-axis triangle / closed-box test (Akenine-Moller 2001).
__device__ bool tri_box_overlap(const double* c, double h, const double (&t)[3][3]) {
    double v[3][3], e[3][3];
    sub3d(t[0], c, v[0]);
    sub3d(t[1], c, v[1]);
    sub3d(t[2], c, v[2]);
    sub3d(v[1], v[0], e[0]);
    sub3d(v[2], v[1], e[1]);
    sub3d(v[0], v[2], e[2]);
    for (int i = 0; i < 3; ++i) {
        for (int a = 0; a < 3; ++a) {
            double ax[3], u[3] = {0.0, 0.0, 0.0};
            u[a] = 1.0;
            crs3d(u, e[i], ax);
            const double p0 = dot3d(ax, v[0]), p1 = dot3d(ax, v[1]), p2 = dot3d(ax, v[2]);
            const double mn = fmin(p0, fmin(p1, p2)), mx = fmax(p0, fmax(p1, p2));
            const double r = h * (fabs(ax[0]) + fabs(ax[1]) + fabs(ax[2]));
            if (mn > r || mx < -r) return false;
        }
    }
    for (int a = 0; a < 3; ++a) {
        const double mn = fmin(v[0][a], fmin(v[1][a], v[2][a])), mx = fmax(v[0][a], fmax(v[1][a], v[2][a]));
        if (mn > h || mx < -h) return false;
    }
    double n[3];
    crs3d(e[0], e[1], n);
    const double d = dot3d(n, v[0]);
    const double r = h * (fabs(n[0]) + fabs(n[1]) + fabs(n[2]));
    return !(d > r || d < -r);
}

template <bool EMIT>
__global__ void occupancy_kernel(const float* __restrict__ tris, int64_t nf, int level, uint32_t* __restrict__ count,
                                 const uint32_t* __restrict__ offset, uint32_t* __restrict__ codes) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int R = 1 << level;
    const double h = 1.0 / R;
    double t[3][3];
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) t[k][a] = tris[9 * f + 3 * k + a];
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const double mn = fmin(t[0][a], fmin(t[1][a], t[2][a])), mx = fmax(t[0][a], fmax(t[1][a], t[2][a]));
        lo[a] = (int)floor((mn + 1.0) * 0.5 * R) - 1;
        hi[a] = (int)floor((mx + 1.0) * 0.5 * R) + 1;
        lo[a] = lo[a] < 0 ? 0 : lo[a];
        hi[a] = hi[a] > R - 1 ? R - 1 : hi[a];
    }
    uint32_t w = EMIT ? offset[f] : 0u;
    for (int x = lo[0]; x <= hi[0]; ++x)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int z = lo[2]; z <= hi[2]; ++z) {
                const double c[3] = {-1.0 + (2.0 * x + 1.0) * h, -1.0 + (2.0 * y + 1.0) * h, -1.0 + (2.0 * z + 1.0) * h};
                if (!tri_box_overlap(c, h, t)) continue;
                if (EMIT) codes[w] = morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
                ++w;
            }
    if (!EMIT) count[f] = w;
}

// ---------------------------------------------------------------------------
// octree bytes from sorted unique codes: one level up
__global__ void parent_bytes_kernel(const uint32_t* __restrict__ codes, const uint32_t* __restrict__ run_start,
                                    const uint32_t* __restrict__ run_len, const uint32_t* __restrict__ nruns,
                                    uint32_t* __restrict__ parents, uint8_t* __restrict__ bytes) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= *nruns) return;
    uint32_t b = 0;
    const uint32_t s = run_start[r];
    for (uint32_t i = 0; i < run_len[r]; ++i) b |= 1u << (codes[s + i] & 7u);
    parents[r] = codes[s] >> 3;
    bytes[r] = (uint8_t)b;
}

__global__ void shift3_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] >> 3;
}

// ---------------------------------------------------------------------------
// points from bytes
__global__ void popc_kernel(const uint8_t* __restrict__ bytes, int64_t n, uint32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)__builtin_popcount(bytes[i]);
}

// Children codes of the nodes [beg, end) (global indices) into codes[].
__global__ void child_codes_kernel(const uint8_t* __restrict__ bytes, const uint32_t* __restrict__ S, int64_t beg,
                                   int64_t end, uint64_t* __restrict__ codes) {
    const int64_t k = beg + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    const uint8_t b = bytes[k];
    uint32_t w = 1 + S[k];
    const uint64_t m = codes[k];
    for (int i = 0; i < 8; ++i)
        if (b & (1u << i)) codes[w++] = 8ull * m + (uint64_t)i;
}

__global__ void codes_to_points_kernel(const uint64_t* __restrict__ codes, int64_t n, int16_t* __restrict__ pts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t m = codes[i];
    uint32_t x = 0, y = 0, z = 0;
    for (int b = 0; b < 16; ++b) {
        z |= (uint32_t)((m >> (3 * b)) & 1u) << b;
        y |= (uint32_t)((m >> (3 * b + 1)) & 1u) << b;
        x |= (uint32_t)((m >> (3 * b + 2)) & 1u) << b;
    }
    pts[4 * i] = (int16_t)x;
    pts[4 * i + 1] = (int16_t)y;
    pts[4 * i + 2] = (int16_t)z;
    pts[4 * i + 3] = 0;
}

// ---------------------------------------------------------------------------
// traversal
struct SpcArgs {
    const uint8_t* oct;
    const uint32_t* S;  // exclusive popcount sums over the bytes
    const int16_t* pts;
    int32_t psum[kSpcMaxLevel + 2];
    int target;
    const float* o;
    const float* d;
    int64_t n;
    const uint32_t* offs;
    uint32_t* counts;
    int32_t* nug;
};

// d_FaceEval (spc_raytrace_cuda_kernel.cu:85-105), written order, no contraction.
__device__ __forceinline__ bool face_eval(uint16_t i, uint16_t j, float a, float b, float c) {
    float r0 = a * (float)i + b * (float)j + c;
    float r1 = r0 + a;
    float r2 = r0 + b;
    float r3 = r1 + b;
    float mn = 1.0f, mx = -1.0f;
    mn = r0 < mn ? r0 : mn;
    mx = r0 > mx ? r0 : mx;
    mn = r1 < mn ? r1 : mn;
    mx = r1 > mx ? r1 : mx;
    mn = r2 < mn ? r2 : mn;
    mx = r2 > mx ? r2 : mx;
    mn = r3 < mn ? r3 : mn;
    mx = r3 > mx ? r3 : mx;
    return mn <= 0.0f && mx >= 0.0f;
}

template <bool EMIT>
__global__ __launch_bounds__(kDfsBlock) void spc_dfs_kernel(SpcArgs a) {
    __shared__ uint32_t stack_lds[kDfsBlock][kStack];
    const int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= a.n) return;
    uint32_t* st = stack_lds[threadIdx.x];
    const float ogx = a.o[3 * ray], ogy = a.o[3 * ray + 1], ogz = a.o[3 * ray + 2];
    // d_Decide's per-ray values (:116-122)
    float ox = 0.5f * ogx, oy = 0.5f * ogy, oz = 0.5f * ogz;
    const float dx = 0.5f * a.d[3 * ray], dy = 0.5f * a.d[3 * ray + 1], dz = 0.5f * a.d[3 * ray + 2];
    ox += 0.5f;
    oy += 0.5f;
    oz += 0.5f;
    const float cx = oy * dz - dy * oz, cy = oz * dx - dz * ox, cz = ox * dy - dx * oy;
    // d_Subdivide's origin term (:160-162)
    const double hx = (double)(0.5f * ogx + 0.5f), hy = (double)(0.5f * ogy + 0.5f), hz = (double)(0.5f * ogz + 0.5f);
    uint32_t w = EMIT ? a.offs[ray] : 0u;
    int sp = 0;
    st[sp++] = 0u;  // level 0, node 0
    while (sp > 0) {
        const uint32_t e = st[--sp];
        const int l = (int)(e >> 28);
        const uint32_t k = e & 0x0FFFFFFFu;
        const int16_t* p = a.pts + 4 * ((int64_t)a.psum[l] + k);
        const uint16_t px = (uint16_t)p[0], py = (uint16_t)p[1], pz = (uint16_t)p[2];
        const float s1 = (float)(1.0 / (double)(float)(1u << l));
        const float s2 = s1 * s1;
        const bool hit = face_eval(py, pz, -s2 * dz, s2 * dy, s1 * cx) && face_eval(px, pz, s2 * dz, -s2 * dx, s1 * cy) &&
                         face_eval(px, py, -s2 * dy, s2 * dx, s1 * cz);
        if (!hit) continue;
        if (l == a.target) {
            if (EMIT) {
                a.nug[2 * w] = (int32_t)ray;
                a.nug[2 * w + 1] = (int32_t)k;
            }
            ++w;
            continue;
        }
        const uint8_t ob = a.oct[a.psum[l] + k];
        const uint32_t s = a.S[a.psum[l] + k];
        const float x = (float)(hx - (double)s1 * ((double)(float)p[0] + 0.5));
        const float y = (float)(hy - (double)s1 * ((double)(float)p[1] + 0.5));
        const float z = (float)(hz - (double)s1 * ((double)(float)p[2] + 0.5));
        uint32_t code = 0;
        if (x > 0) code = 4;
        if (y > 0) code += 2;
        if (z > 0) code += 1;
        for (int i = 7; i >= 0; --i) {  // pushed in reverse: popped front-to-back
            const uint32_t j = kOrder[code][i];
            if (ob & (1u << j)) {
                const uint32_t cnt = (uint32_t)__builtin_popcount(ob & ((2u << j) - 1u));
                const uint32_t child = s + cnt - (uint32_t)a.psum[l + 1];
                st[sp++] = ((uint32_t)(l + 1) << 28) | child;
            }
        }
    }
    if (!EMIT) a.counts[ray] = w;
}

// ---------------------------------------------------------------------------
// Wave-cooperative traversal, one pass (targets up to kBfsMaxTarget).
//
// A wave takes a tile of 64 consecutive rays (ticket order) and walks the
// octree level by level in LDS: the list of level-l cells its rays hit,
// sorted by ray and, per ray, in depth-first (front-to-back) order, becomes
// the level-(l+1) list by giving each parent eight lanes (one per slot of
// the Order row of its octant) that test the child in that slot; the hits
// are compacted with a ballot and mbcnt, so the new list keeps the order
// (parents in list order, children in Order).  At the target level the list
// is the tile's nuggets.  Its position in the output comes from a
// decoupled look-back over the tiles' counts (tile i publishes its count,
// then adds its predecessors' until one has published an inclusive sum), so
// the list is written once, ray-major, with no host read-back.  A tile whose
// list outgrows the LDS buffers falls back to the per-ray depth-first walk
// (count, look-back, emit) with the same buffers as its stacks.  The face
// tests and child ranks are those of the depth-first walk: same bits.
constexpr int kBfsCap = 2048;        // entries per LDS list (ray-local << 26 | node)
constexpr int kBfsMaxTarget = 9;     // a depth-first stack of 7 * 9 + 1 <= 64 entries per lane fits the lists
constexpr int kBfsStack = 2 * kBfsCap / 64;
constexpr uint32_t kNodeMask = (1u << 26) - 1u;
constexpr uint64_t kTileAggregate = 1ull << 32;
constexpr uint64_t kTileInclusive = 2ull << 32;
constexpr uint32_t kLookbackSpinLimit = 1u << 22;
static_assert(7 * kBfsMaxTarget + 1 <= kBfsStack, "fallback stack fits the LDS lists");

struct SpcBfsArgs {
    SpcArgs s;
    uint32_t* ticket;      // tile ticket counter (zeroed before the launch)
    uint64_t* tile_state;  // [ntiles] (flag << 32 | count), zeroed before the launch
    uint32_t* status;      // [0] total nuggets, [1] flags: 1 = capacity exceeded, 2 = look-back gave up
    uint32_t* offs;        // [n + 1] CSR offsets of each ray's nuggets
    uint32_t ntiles;
    int64_t cap;           // nugget capacity (entries beyond it are not written)
};

// Per-ray values of the face tests and octant codes (d_Decide :116-122,
// d_Subdivide :160-162), as the depth-first walk computes them.
struct RayConsts {
    float dx, dy, dz, cx, cy, cz;
    double hx, hy, hz;
};

__device__ __forceinline__ RayConsts ray_consts(const float* o, const float* d, int64_t ray, bool valid) {
    RayConsts c;
    const float ogx = valid ? o[3 * ray] : 0.0f, ogy = valid ? o[3 * ray + 1] : 0.0f, ogz = valid ? o[3 * ray + 2] : 0.0f;
    float ox = 0.5f * ogx, oy = 0.5f * ogy, oz = 0.5f * ogz;
    c.dx = 0.5f * (valid ? d[3 * ray] : 0.0f);
    c.dy = 0.5f * (valid ? d[3 * ray + 1] : 0.0f);
    c.dz = 0.5f * (valid ? d[3 * ray + 2] : 0.0f);
    ox += 0.5f;
    oy += 0.5f;
    oz += 0.5f;
    c.cx = oy * c.dz - c.dy * oz;
    c.cy = oz * c.dx - c.dz * ox;
    c.cz = ox * c.dy - c.dx * oy;
    c.hx = (double)(0.5f * ogx + 0.5f);
    c.hy = (double)(0.5f * ogy + 0.5f);
    c.hz = (double)(0.5f * ogz + 0.5f);
    return c;
}

__device__ __forceinline__ double shfl_d(double v, int src) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(b >> 32), src);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ bool cell_hit(const RayConsts& c, int l, uint16_t px, uint16_t py, uint16_t pz) {
    const float s1 = (float)(1.0 / (double)(float)(1u << l));
    const float s2 = s1 * s1;
    return face_eval(py, pz, -s2 * c.dz, s2 * c.dy, s1 * c.cx) && face_eval(px, pz, s2 * c.dz, -s2 * c.dx, s1 * c.cy) &&
           face_eval(px, py, -s2 * c.dy, s2 * c.dx, s1 * c.cz);
}

// Octant of the ray origin relative to the centre of cell p at level l (the Order row).
__device__ __forceinline__ uint32_t octant_code(const RayConsts& c, int l, const int16_t* p) {
    const float s1 = (float)(1.0 / (double)(float)(1u << l));
    const float x = (float)(c.hx - (double)s1 * ((double)(float)p[0] + 0.5));
    const float y = (float)(c.hy - (double)s1 * ((double)(float)p[1] + 0.5));
    const float z = (float)(c.hz - (double)s1 * ((double)(float)p[2] + 0.5));
    uint32_t code = 0;
    if (x > 0) code = 4;
    if (y > 0) code += 2;
    if (z > 0) code += 1;
    return code;
}

// One ray's depth-first walk (the fallback), stack `st` of kBfsStack
// entries; EMIT writes its nuggets from position w.
template <bool EMIT>
__device__ uint32_t dfs_one(const SpcArgs& a, const RayConsts& c, int64_t ray, uint32_t* st, uint32_t w, int64_t cap) {
    int sp = 0;
    st[sp++] = 0u;
    uint32_t cnt = 0;
    while (sp > 0) {
        const uint32_t e = st[--sp];
        const int l = (int)(e >> 28);
        const uint32_t k = e & 0x0FFFFFFFu;
        const int16_t* p = a.pts + 4 * ((int64_t)a.psum[l] + k);
        if (!cell_hit(c, l, (uint16_t)p[0], (uint16_t)p[1], (uint16_t)p[2])) continue;
        if (l == a.target) {
            if (EMIT && (int64_t)(w + cnt) < cap) {
                a.nug[2 * (int64_t)(w + cnt)] = (int32_t)ray;
                a.nug[2 * (int64_t)(w + cnt) + 1] = (int32_t)k;
            }
            ++cnt;
            continue;
        }
        const uint8_t ob = a.oct[a.psum[l] + k];
        const uint32_t s = a.S[a.psum[l] + k];
        const uint32_t code = octant_code(c, l, p);
        for (int i = 7; i >= 0; --i) {  // pushed in reverse: popped front-to-back
            const uint32_t j = kOrder[code][i];
            if (ob & (1u << j)) {
                const uint32_t rank = (uint32_t)__builtin_popcount(ob & ((2u << j) - 1u));
                st[sp++] = ((uint32_t)(l + 1) << 28) | (s + rank - (uint32_t)a.psum[l + 1]);
            }
        }
    }
    return cnt;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
    return v;
}

__device__ __forceinline__ uint32_t wave_excl_scan_u32(uint32_t v, uint32_t lane) {
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    return x - v;
}

// Decoupled look-back: publishes the tile's count, then sums its
// predecessors' until one has published an inclusive sum; returns the
// tile's exclusive prefix.  Predecessors took their tickets earlier and
// publish their count before they wait on anyone, so the wait ends; the
// spin limit only guards against a broken invariant (status flag 2).  A
// state word carries its flag and its value together and no other data is
// published with it, so the loads and stores are relaxed (device-scope
// atomics, which bypass the CU's L1): an acquire load would also invalidate
// the CU's L1 on every poll, under the waves that are still walking the
// octree.
__device__ uint32_t tile_lookback(const SpcBfsArgs& b, uint32_t tile, uint32_t count, uint32_t lane) {
    if (tile == 0) {
        if (lane == 0) __hip_atomic_store(&b.tile_state[0], kTileInclusive | count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0u;
    }
    if (lane == 0) __hip_atomic_store(&b.tile_state[tile], kTileAggregate | count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = 0;
    int64_t top = (int64_t)tile - 1;
    uint32_t spins = 0;
    for (;;) {
        const int64_t k = top - (int64_t)lane;
        const uint64_t v = k >= 0 ? __hip_atomic_load(&b.tile_state[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : kTileInclusive;
        const uint32_t flag = (uint32_t)(v >> 32);
        const uint64_t inc = __ballot(flag == 2u);
        const uint64_t unready = __ballot(flag == 0u);
        const int kinc = inc ? __ffsll((unsigned long long)inc) - 1 : 64;
        const uint64_t upto = kinc >= 63 ? ~0ull : ((2ull << kinc) - 1ull);  // lanes 0..kinc
        if (unready & upto) {
            if (++spins > kLookbackSpinLimit) {
                if (lane == 0) atomicOr(&b.status[1], 2u);
                return 0u;
            }
            __builtin_amdgcn_s_sleep(8);
            continue;
        }
        excl += wave_sum_u32(((uint64_t)1 << lane) & upto ? (uint32_t)v : 0u);
        if (kinc < 64) break;
        top -= 64;
    }
    if (lane == 0) __hip_atomic_store(&b.tile_state[tile], kTileInclusive | (excl + count), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}

__global__ __launch_bounds__(64) void spc_bfs_kernel(SpcBfsArgs b) {
    __shared__ uint32_t lists[2][kBfsCap];
    const SpcArgs& a = b.s;
    const uint32_t lane = threadIdx.x;
    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(b.ticket, 1u);
        tile = (uint32_t)__builtin_amdgcn_readfirstlane((int)tile);
        if (tile >= b.ntiles) break;
        const int64_t r0 = (int64_t)tile * 64;
        const int64_t ray = r0 + lane;
        const bool valid = ray < a.n;
        const RayConsts mine = ray_consts(a.o, a.d, ray, valid);
        // level 0: the root cell (point (0, 0, 0))
        const bool root = valid && cell_hit(mine, 0, (uint16_t)a.pts[0], (uint16_t)a.pts[1], (uint16_t)a.pts[2]);
        uint64_t m0 = __ballot(root);
        if (root) lists[0][mask_rank(m0)] = lane << 26;
        __syncthreads();
        uint32_t count = (uint32_t)__popcll(m0);
        int cur = 0;
        bool overflow = false;
        for (int l = 0; l < a.target; ++l) {
            const uint32_t parents = count;
            uint32_t next = 0;
            const int32_t base_l = a.psum[l], base_n = a.psum[l + 1];
            for (uint32_t j0 = 0; j0 < parents; j0 += 8) {
                const uint32_t pi = j0 + (lane >> 3);
                const uint32_t slot = lane & 7u;
                const bool act = pi < parents;
                const uint32_t e = act ? lists[cur][pi] : 0u;
                const int rl = (int)(e >> 26);
                const uint32_t k = e & kNodeMask;
                RayConsts c;
                c.dx = __shfl(mine.dx, rl);
                c.dy = __shfl(mine.dy, rl);
                c.dz = __shfl(mine.dz, rl);
                c.cx = __shfl(mine.cx, rl);
                c.cy = __shfl(mine.cy, rl);
                c.cz = __shfl(mine.cz, rl);
                c.hx = shfl_d(mine.hx, rl);
                c.hy = shfl_d(mine.hy, rl);
                c.hz = shfl_d(mine.hz, rl);
                bool h = false;
                uint32_t child = 0;
                if (act) {
                    const int16_t* p = a.pts + 4 * ((int64_t)base_l + k);
                    const uint8_t ob = a.oct[base_l + k];
                    const uint32_t jj = kOrder[octant_code(c, l, p)][slot];
                    if (ob & (1u << jj)) {
                        const uint32_t rank = (uint32_t)__builtin_popcount(ob & ((2u << jj) - 1u));
                        child = a.S[base_l + k] + rank - (uint32_t)base_n;
                        const int16_t* q = a.pts + 4 * ((int64_t)base_n + child);
                        h = cell_hit(c, l + 1, (uint16_t)q[0], (uint16_t)q[1], (uint16_t)q[2]);
                    }
                }
                const uint64_t hm = __ballot(h);
                const uint32_t pos = next + mask_rank(hm);
                if (h && pos < (uint32_t)kBfsCap) lists[cur ^ 1][pos] = ((uint32_t)rl << 26) | child;
                next += (uint32_t)__popcll(hm);
            }
            __syncthreads();  // one wave per block: orders the list writes before the next level's reads
            if (next > (uint32_t)kBfsCap) {
                overflow = true;
                break;
            }
            count = next;
            cur ^= 1;
        }
        __syncthreads();
        if (!overflow) {
            const uint32_t base = tile_lookback(b, tile, count, lane);
            // this ray's first entry: lower bound of lane << 26 in the sorted list
            uint32_t lo = 0, hi = count;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if ((lists[cur][mid] >> 26) < lane) lo = mid + 1;
                else hi = mid;
            }
            if (valid) b.offs[ray] = base + lo;
            for (uint32_t j = lane; j < count; j += 64) {
                const uint32_t e = lists[cur][j];
                const int64_t pos = (int64_t)base + j;
                if (pos < b.cap) {
                    int2 v;
                    v.x = (int32_t)(r0 + (e >> 26));
                    v.y = (int32_t)(e & kNodeMask);
                    reinterpret_cast<int2*>(a.nug)[pos] = v;
                }
            }
            if (tile == b.ntiles - 1 && lane == 0) {
                const uint32_t total = base + count;
                b.offs[a.n] = total;
                b.status[0] = total;
                if ((int64_t)total > b.cap) atomicOr(&b.status[1], 1u);
            }
        } else {
            // the per-ray depth-first walk, the two lists as 64 stacks
            uint32_t* st = &lists[0][0] + lane * kBfsStack;
            const uint32_t mycount = valid ? dfs_one<false>(a, mine, ray, st, 0u, 0) : 0u;
            const uint32_t total = wave_sum_u32(mycount);
            const uint32_t base = tile_lookback(b, tile, total, lane);
            const uint32_t w = base + wave_excl_scan_u32(mycount, lane);
            if (valid) {
                b.offs[ray] = w;
                dfs_one<true>(a, mine, ray, st, w, b.cap);
            }
            if (tile == b.ntiles - 1 && lane == 0) {
                const uint32_t all = base + total;
                b.offs[a.n] = all;
                b.status[0] = all;
                if ((int64_t)all > b.cap) atomicOr(&b.status[1], 1u);
            }
        }
        __syncthreads();  // the next tile rewrites the lists
    }
}

__global__ void spc_bfs_clear_kernel(uint32_t* ticket, uint64_t* state, uint32_t ntiles, uint32_t* status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntiles) state[i] = 0ull;
    if (i == 0) {
        if (ticket) *ticket = 0u;
        status[0] = 0u;
        status[1] = 0u;
    }
}

__global__ void mark_unique_kernel(const int32_t* __restrict__ nug, int64_t m, int32_t* __restrict__ info) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    info[i] = (i == 0) ? 1 : (nug[2 * (i - 1)] == nug[2 * i] ? 0 : 1);
}

static size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }

static int sync_read_u32(const uint32_t* dev, uint32_t* host, hipStream_t s) {
    if (hipMemcpyAsync(host, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "device read failed");
    return NGLOD_OK;
}

__global__ void spc_zero_u32_kernel(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Depth-first traversal in two passes (count, scan, emit) with the count
// read back in between: targets deeper than the single pass supports.
static int spc_raytrace_dfs(const SpcArgs& base_args, int64_t n, int32_t* nuggets, int64_t cap, int64_t* count,
                            uint32_t* ray_offsets, void* workspace, size_t workspace_bytes, size_t* workspace_needed,
                            hipStream_t s) {
    size_t scan_tmp = 0;
    if (rocprim::exclusive_scan((void*)nullptr, scan_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u, (size_t)n + 1,
                                rocprim::plus<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "scan size query");
    const size_t o_cnt = 0, o_tmp = a256((size_t)(n + 1) * 4);
    const size_t need = o_tmp + a256(scan_tmp);
    if (!workspace || workspace_bytes < need) {
        *workspace_needed = need;
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    }
    SpcArgs a = base_args;
    a.offs = ray_offsets;
    a.counts = (uint32_t*)((char*)workspace + o_cnt);
    a.nug = nuggets;
    hipLaunchKernelGGL(spc_zero_u32_kernel, dim3(1), dim3(64), 0, s, a.counts + n, 1);
    int rc = check_launch("spc_zero_u32_kernel");
    if (rc) return rc;
    const unsigned blocks = (unsigned)((n + kDfsBlock - 1) / kDfsBlock);
    hipLaunchKernelGGL(spc_dfs_kernel<false>, dim3(blocks), dim3(kDfsBlock), 0, s, a);
    rc = check_launch("spc_dfs_kernel");
    if (rc) return rc;
    size_t t = scan_tmp;
    if (rocprim::exclusive_scan((void*)((char*)workspace + o_tmp), t, a.counts, ray_offsets, 0u, (size_t)n + 1,
                                rocprim::plus<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "scan failed");
    uint32_t total = 0;
    rc = sync_read_u32(ray_offsets + n, &total, s);
    if (rc) return rc;
    *count = total;
    if ((int64_t)total > cap || (total > 0 && !nuggets)) return fail(NGLOD_ERR_CAPACITY, "nugget capacity too small");
    if (total == 0) return NGLOD_OK;
    hipLaunchKernelGGL(spc_dfs_kernel<true>, dim3(blocks), dim3(kDfsBlock), 0, s, a);
    return check_launch("spc_dfs_kernel");
}

static int spc_args(const uint8_t* octree, const uint32_t* prefix, const int16_t* points, const int32_t* pyramid,
                    int32_t level, const float* ray_o, const float* ray_d, int64_t n, int32_t target, SpcArgs* a) {
    if (!pyramid) return fail(NGLOD_ERR_ARG, "pyramid is NULL");
    if (level < 1 || level > kSpcMaxLevel) return fail(NGLOD_ERR_UNSUPPORTED, "octree level out of range");
    if (target < 0 || target > level) return fail(NGLOD_ERR_ARG, "target level out of range");
    if (n < 0 || n > (int64_t)0x7fffffff) return fail(NGLOD_ERR_ARG, "ray count out of range");
    if (pyramid[level + 1] != 0) return fail(NGLOD_ERR_ARG, "bad pyramid");
    if (n > 0 && (!octree || !prefix || !points || !ray_o || !ray_d)) return fail(NGLOD_ERR_ARG, "NULL pointer");
    memset(a, 0, sizeof(*a));
    a->oct = octree;
    a->S = prefix;
    a->pts = points;
    for (int l = 0; l <= level + 1 && l < kSpcMaxLevel + 2; ++l) a->psum[l] = pyramid[(level + 2) + l];
    a->target = target;
    a->o = ray_o;
    a->d = ray_d;
    a->n = n;
    return NGLOD_OK;
}

static uint32_t bfs_tiles(int64_t n) { return (uint32_t)((n + 63) / 64); }

// Workspace of the single pass: the ticket counter, then the tiles' states.
static size_t bfs_workspace_size(int64_t n) { return 256 + a256((size_t)bfs_tiles(n) * 8); }

static int spc_raytrace_bfs(const SpcArgs& a, int32_t* nuggets, int64_t cap, uint32_t* ray_offsets, uint32_t* status,
                            void* workspace, hipStream_t s) {
    SpcBfsArgs b;
    b.s = a;
    b.s.nug = nuggets;
    b.ticket = (uint32_t*)workspace;
    b.tile_state = (uint64_t*)((char*)workspace + 256);
    b.status = status;
    b.offs = ray_offsets;
    b.ntiles = bfs_tiles(a.n);
    b.cap = nuggets ? cap : 0;
    hipLaunchKernelGGL(spc_bfs_clear_kernel, dim3((b.ntiles + 255) / 256), dim3(256), 0, s, b.ticket, b.tile_state,
                       b.ntiles, b.status);
    int rc = check_launch("spc_bfs_clear_kernel");
    if (rc) return rc;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, spc_bfs_kernel, 64, 0) != hipSuccess || per_cu <= 0)
        per_cu = 8;
    const int64_t cap_blocks = (int64_t)device_cu_count() * per_cu;
    const unsigned grid = (unsigned)((int64_t)b.ntiles < cap_blocks ? (int64_t)b.ntiles : cap_blocks);
    hipLaunchKernelGGL(spc_bfs_kernel, dim3(grid), dim3(64), 0, s, b);
    return check_launch("spc_bfs_kernel");
}

}  // namespace nglod

using namespace nglod;

extern "C" {

int nglod_spc_occupancy(const float* tris, int64_t nf, int32_t level, uint32_t* codes, int64_t cap, int64_t* count,
                        void* workspace, size_t workspace_bytes, size_t* workspace_needed, void* stream) {
    if (!count || !workspace_needed) return fail(NGLOD_ERR_ARG, "NULL out-parameter");
    if (level < 1 || level > 10) return fail(NGLOD_ERR_UNSUPPORTED, "occupancy level must be 1..10");
    if (nf <= 0 || !tris) return fail(NGLOD_ERR_ARG, "empty mesh");
    hipStream_t s = (hipStream_t)stream;
    // phase 1: candidate counts per triangle (needs nf*8 + scan temp)
    size_t scan_tmp = 0;
    if (rocprim::exclusive_scan((void*)nullptr, scan_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u, (size_t)nf + 1,
                                rocprim::plus<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "scan size query");
    const size_t o_cnt = 0, o_off = a256((size_t)(nf + 1) * 4), o_tmp = o_off + a256((size_t)(nf + 1) * 4);
    size_t need = o_tmp + a256(scan_tmp);
    if (!workspace || workspace_bytes < need) {
        *workspace_needed = need + (size_t)(1 << 20);
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small (phase 1)");
    }
    char* ws = (char*)workspace;
    uint32_t* cnt = (uint32_t*)(ws + o_cnt);
    uint32_t* off = (uint32_t*)(ws + o_off);
    hipMemsetAsync(cnt + nf, 0, 4, s);
    hipLaunchKernelGGL(occupancy_kernel<false>, dim3((unsigned)((nf + 127) / 128)), dim3(128), 0, s, tris, nf, level, cnt,
                       (const uint32_t*)nullptr, (uint32_t*)nullptr);
    int rc = check_launch("occupancy_kernel");
    if (rc) return rc;
    size_t t = scan_tmp;
    if (rocprim::exclusive_scan((void*)(ws + o_tmp), t, cnt, off, 0u, (size_t)nf + 1, rocprim::plus<uint32_t>(), s) !=
        hipSuccess)
        return fail(NGLOD_ERR_HIP, "scan failed");
    uint32_t total = 0;
    rc = sync_read_u32(off + nf, &total, s);
    if (rc) return rc;
    // phase 2: emit, sort, unique
    size_t sort_tmp = 0, uniq_tmp = 0;
    if (rocprim::radix_sort_keys((void*)nullptr, sort_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)total, 0,
                                 3 * level, s) != hipSuccess ||
        rocprim::unique((void*)nullptr, uniq_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                        (size_t)total, rocprim::equal_to<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "sort size query");
    const size_t o_a = need, o_b = o_a + a256((size_t)total * 4 + 4), o_n = o_b + a256((size_t)total * 4 + 4),
                 o_t2 = o_n + 256;
    need = o_t2 + a256(sort_tmp > uniq_tmp ? sort_tmp : uniq_tmp);
    if (workspace_bytes < need) {
        *workspace_needed = need;
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small (phase 2)");
    }
    uint32_t* A = (uint32_t*)(ws + o_a);
    uint32_t* B = (uint32_t*)(ws + o_b);
    uint32_t* nu = (uint32_t*)(ws + o_n);
    hipLaunchKernelGGL(occupancy_kernel<true>, dim3((unsigned)((nf + 127) / 128)), dim3(128), 0, s, tris, nf, level, cnt,
                       off, A);
    rc = check_launch("occupancy_kernel");
    if (rc) return rc;
    t = sort_tmp;
    if (total > 0 && rocprim::radix_sort_keys((void*)(ws + o_t2), t, A, B, (size_t)total, 0, 3 * level, s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "sort failed");
    t = uniq_tmp;
    hipMemsetAsync(nu, 0, 4, s);
    if (total > 0 && rocprim::unique((void*)(ws + o_t2), t, B, A, nu, (size_t)total, rocprim::equal_to<uint32_t>(), s) !=
                         hipSuccess)
        return fail(NGLOD_ERR_HIP, "unique failed");
    uint32_t u = 0;
    rc = sync_read_u32(nu, &u, s);
    if (rc) return rc;
    *count = u;
    if ((int64_t)u > cap || (u > 0 && !codes)) return fail(NGLOD_ERR_CAPACITY, "codes capacity too small");
    if (u > 0 && hipMemcpyAsync(codes, A, (size_t)u * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "copy failed");
    return NGLOD_OK;
}

int nglod_spc_from_codes(const uint32_t* codes, int64_t n, int32_t level, uint8_t* octree, int64_t cap, int64_t* osize,
                         int32_t* level_counts, void* workspace, size_t workspace_bytes, size_t* workspace_needed,
                         void* stream) {
    if (!osize || !level_counts || !workspace_needed) return fail(NGLOD_ERR_ARG, "NULL out-parameter");
    if (level < 1 || level > 10 || n <= 0 || !codes) return fail(NGLOD_ERR_ARG, "bad level or empty codes");
    hipStream_t s = (hipStream_t)stream;
    size_t rle_tmp = 0, scan_tmp = 0;
    if (rocprim::run_length_encode((void*)nullptr, rle_tmp, (uint32_t*)nullptr, (size_t)n, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, (uint32_t*)nullptr, s) != hipSuccess ||
        rocprim::exclusive_scan((void*)nullptr, scan_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u, (size_t)n,
                                rocprim::plus<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "size query");
    const size_t nb = a256((size_t)n * 4);
    // cur, par, keys, ukeys, lens, starts, bytes per level (level x n), tmp
    const size_t o_cur = 0, o_par = nb, o_key = 2 * nb, o_uk = 3 * nb, o_len = 4 * nb, o_st = 5 * nb, o_nr = 6 * nb,
                 o_by = o_nr + 256, o_tmp = o_by + a256((size_t)level * n);
    const size_t need = o_tmp + a256(rle_tmp > scan_tmp ? rle_tmp : scan_tmp);
    if (!workspace || workspace_bytes < need) {
        *workspace_needed = need;
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    }
    char* ws = (char*)workspace;
    uint32_t* cur = (uint32_t*)(ws + o_cur);
    uint32_t* par = (uint32_t*)(ws + o_par);
    uint32_t* key = (uint32_t*)(ws + o_key);
    uint32_t* uk = (uint32_t*)(ws + o_uk);
    uint32_t* len = (uint32_t*)(ws + o_len);
    uint32_t* st = (uint32_t*)(ws + o_st);
    uint32_t* nr = (uint32_t*)(ws + o_nr);
    uint8_t* by = (uint8_t*)(ws + o_by);
    if (hipMemcpyAsync(cur, codes, (size_t)n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "copy failed");
    int64_t m = n;
    int64_t counts[16];
    counts[level] = n;
    for (int l = level; l > 0; --l) {
        hipLaunchKernelGGL(shift3_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, cur, key, m);
        size_t t = rle_tmp;
        if (rocprim::run_length_encode((void*)(ws + o_tmp), t, key, (size_t)m, uk, len, nr, s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "run-length encode failed");
        t = scan_tmp;
        if (rocprim::exclusive_scan((void*)(ws + o_tmp), t, len, st, 0u, (size_t)m, rocprim::plus<uint32_t>(), s) !=
            hipSuccess)
            return fail(NGLOD_ERR_HIP, "scan failed");
        hipLaunchKernelGGL(parent_bytes_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, cur, st, len, nr, par,
                           by + (size_t)(l - 1) * n);
        int rc = check_launch("parent_bytes_kernel");
        if (rc) return rc;
        uint32_t r = 0;
        rc = sync_read_u32(nr, &r, s);
        if (rc) return rc;
        counts[l - 1] = r;
        uint32_t* tmp = cur;
        cur = par;
        par = tmp;
        m = r;
    }
    int64_t total = 0;
    for (int l = 0; l < level; ++l) total += counts[l];
    *osize = total;
    for (int l = 0; l <= level; ++l) level_counts[l] = (int32_t)counts[l];
    if (total > cap || !octree) return fail(NGLOD_ERR_CAPACITY, "octree capacity too small");
    int64_t o = 0;
    for (int l = 0; l < level; ++l) {
        if (hipMemcpyAsync(octree + o, by + (size_t)l * n, (size_t)counts[l], hipMemcpyDeviceToDevice, s) != hipSuccess)
            return fail(NGLOD_ERR_HIP, "copy failed");
        o += counts[l];
    }
    return hipStreamSynchronize(s) == hipSuccess ? NGLOD_OK : fail(NGLOD_ERR_HIP, "sync failed");
}

int nglod_spc_points(const uint8_t* octree, int64_t osize, int16_t* points, int64_t pcap, int32_t* pyramid,
                     int32_t pyramid_cap, int32_t* level_out, int64_t* psize_out, uint32_t* prefix, void* workspace,
                     size_t workspace_bytes, size_t* workspace_needed, void* stream) {
    if (!pyramid || !level_out || !psize_out || !workspace_needed || !prefix)
        return fail(NGLOD_ERR_ARG, "NULL out-parameter");
    if (osize <= 0 || !octree) return fail(NGLOD_ERR_ARG, "empty octree");
    hipStream_t s = (hipStream_t)stream;
    size_t scan_tmp = 0;
    if (rocprim::exclusive_scan((void*)nullptr, scan_tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u, (size_t)osize + 1,
                                rocprim::plus<uint32_t>(), s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "scan size query");
    const size_t o_pc = 0, o_tmp = a256((size_t)(osize + 1) * 4);
    size_t need = o_tmp + a256(scan_tmp);
    if (!workspace || workspace_bytes < need) {
        *workspace_needed = need + a256((size_t)(8 * osize + 1) * 8);
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small (phase 1)");
    }
    char* ws = (char*)workspace;
    uint32_t* pc = (uint32_t*)(ws + o_pc);
    hipLaunchKernelGGL(popc_kernel, dim3((unsigned)((osize + 255) / 256)), dim3(256), 0, s, octree, osize, pc);
    hipMemsetAsync(pc + osize, 0, 4, s);
    size_t t = scan_tmp;
    if (rocprim::exclusive_scan((void*)(ws + o_tmp), t, pc, prefix, 0u, (size_t)osize + 1, rocprim::plus<uint32_t>(), s) !=
        hipSuccess)
        return fail(NGLOD_ERR_HIP, "scan failed");
    uint32_t tot = 0;
    int rc = sync_read_u32(prefix + osize, &tot, s);
    if (rc) return rc;
    const int64_t psize = (int64_t)tot + 1;
    *psize_out = psize;
    const size_t o_codes = o_tmp + a256(scan_tmp);
    need = o_codes + a256((size_t)psize * 8);
    if (workspace_bytes < need) {
        *workspace_needed = need;
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small (phase 2)");
    }
    if (psize > pcap || !points) return fail(NGLOD_ERR_CAPACITY, "points capacity too small");
    uint64_t* codes = (uint64_t*)(ws + o_codes);
    hipMemsetAsync(codes, 0, 8, s);
    int64_t counts[kSpcMaxLevel + 2];
    int L = 0;
    counts[0] = 1;
    int64_t beg = 0, end = 1;
    while (beg < osize) {
        if (L + 1 > kSpcMaxLevel) return fail(NGLOD_ERR_UNSUPPORTED, "octree deeper than supported");
        const int64_t e = end < osize ? end : osize;
        hipLaunchKernelGGL(child_codes_kernel, dim3((unsigned)((e - beg + 255) / 256)), dim3(256), 0, s, octree, prefix, beg,
                           e, codes);
        rc = check_launch("child_codes_kernel");
        if (rc) return rc;
        uint32_t nxt = 0;
        rc = sync_read_u32(prefix + e, &nxt, s);
        if (rc) return rc;
        ++L;
        counts[L] = (int64_t)nxt + 1 - end;
        beg = end;
        end = (int64_t)nxt + 1;
    }
    *level_out = L;
    if (pyramid_cap < 2 * (L + 2)) return fail(NGLOD_ERR_CAPACITY, "pyramid capacity too small");
    int64_t sum = 0;
    for (int l = 0; l <= L; ++l) {
        pyramid[l] = (int32_t)counts[l];
        pyramid[(L + 2) + l] = (int32_t)sum;
        sum += counts[l];
    }
    pyramid[L + 1] = 0;
    pyramid[(L + 2) + L + 1] = (int32_t)sum;
    hipLaunchKernelGGL(codes_to_points_kernel, dim3((unsigned)((psize + 255) / 256)), dim3(256), 0, s, codes, psize,
                       points);
    rc = check_launch("codes_to_points_kernel");
    if (rc) return rc;
    return hipStreamSynchronize(s) == hipSuccess ? NGLOD_OK : fail(NGLOD_ERR_HIP, "sync failed");
}

size_t nglod_spc_raytrace_async_workspace_size(int64_t n) { return bfs_workspace_size(n < 0 ? 0 : n); }

int nglod_spc_raytrace_async(const uint8_t* octree, const uint32_t* prefix, const int16_t* points,
                             const int32_t* pyramid, int32_t level, const float* ray_o, const float* ray_d, int64_t n,
                             int32_t target, int32_t* nuggets, int64_t cap, uint32_t* ray_offsets, uint32_t* status,
                             void* workspace, size_t workspace_bytes, void* stream) {
    SpcArgs a;
    int rc = spc_args(octree, prefix, points, pyramid, level, ray_o, ray_d, n, target, &a);
    if (rc) return rc;
    if (target > kBfsMaxTarget) return fail(NGLOD_ERR_UNSUPPORTED, "single-pass traversal: target level above 9");
    if (cap < 0 || (cap > 0 && !nuggets)) return fail(NGLOD_ERR_ARG, "nugget capacity without a buffer");
    if (!ray_offsets || !status) return fail(NGLOD_ERR_ARG, "NULL pointer");
    if (nuggets && ((uintptr_t)nuggets & 7u)) return fail(NGLOD_ERR_ARG, "nuggets not 8-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) {
        hipLaunchKernelGGL(spc_bfs_clear_kernel, dim3(1), dim3(64), 0, s, (uint32_t*)nullptr, (uint64_t*)nullptr, 0u,
                           status);
        return check_launch("spc_bfs_clear_kernel");
    }
    if (!workspace || workspace_bytes < bfs_workspace_size(n) || ((uintptr_t)workspace & 7u))
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small or not 8-byte aligned");
    return spc_raytrace_bfs(a, nuggets, cap, ray_offsets, status, workspace, s);
}

int nglod_spc_raytrace(const uint8_t* octree, const uint32_t* prefix, const int16_t* points, const int32_t* pyramid,
                       int32_t level, const float* ray_o, const float* ray_d, int64_t n, int32_t target,
                       int32_t* nuggets, int64_t cap, int64_t* count, uint32_t* ray_offsets, void* workspace,
                       size_t workspace_bytes, size_t* workspace_needed, void* stream) {
    if (!count || !workspace_needed) return fail(NGLOD_ERR_ARG, "NULL out-parameter");
    SpcArgs a;
    int rc = spc_args(octree, prefix, points, pyramid, level, ray_o, ray_d, n, target, &a);
    if (rc) return rc;
    *count = 0;
    if (n == 0) return NGLOD_OK;
    if (!ray_offsets) return fail(NGLOD_ERR_ARG, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    if (target > kBfsMaxTarget)
        return spc_raytrace_dfs(a, n, nuggets, cap, count, ray_offsets, workspace, workspace_bytes, workspace_needed,
                                s);
    if (nuggets && ((uintptr_t)nuggets & 7u)) return fail(NGLOD_ERR_ARG, "nuggets not 8-byte aligned");
    // the single pass, with its two status words after its workspace
    const size_t need = bfs_workspace_size(n) + 256;
    if (!workspace || workspace_bytes < need || ((uintptr_t)workspace & 7u)) {
        *workspace_needed = need;
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    }
    uint32_t* status = (uint32_t*)((char*)workspace + bfs_workspace_size(n));
    rc = spc_raytrace_bfs(a, nuggets, nuggets ? cap : 0, ray_offsets, status, workspace, s);
    if (rc) return rc;
    uint32_t st[2] = {0, 0};
    if (hipMemcpyAsync(st, status, sizeof(st), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(NGLOD_ERR_HIP, "device read failed");
    if (st[1] & 2u) return fail(NGLOD_ERR_HIP, "traversal: a tile's look-back did not complete");
    *count = st[0];
    if ((int64_t)st[0] > cap || (st[0] > 0 && !nuggets)) return fail(NGLOD_ERR_CAPACITY, "nugget capacity too small");
    return NGLOD_OK;
}

int nglod_mark_unique_rays(const int32_t* nuggets, int64_t m, int32_t* info, void* stream) {
    if (m < 0) return fail(NGLOD_ERR_ARG, "m < 0");
    if (m == 0) return NGLOD_OK;
    if (!nuggets || !info) return fail(NGLOD_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(mark_unique_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, (hipStream_t)stream, nuggets,
                       m, info);
    return check_launch("mark_unique_kernel");
}

}  // extern "C"
