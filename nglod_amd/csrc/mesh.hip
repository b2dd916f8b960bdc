// Training targets and samples for the mesh the model is fitted to.
//   nglod_mesh2sdf_f64     signed point-to-mesh distance in FP64
//                          (mesh2sdf.mesh2sdf_gpu, mesh2sdf_kernel.cu:307-616, 895-927)
//   nglod_sample_uniform   points uniform in [-1,1]^3      (torchgp/sample_uniform.py)
//   nglod_sample_surface   area-weighted surface points (+ Gaussian offset)
//                          (torchgp/sample_surface.py, sample_near_surface.py)
//
// mesh2sdf: one thread per point, triangles streamed through LDS in tiles.
// Everything that depends only on the triangle (edges, normal, the 13
// stab-direction cross products and determinants) is computed once per
// triangle by a prep kernel, so the per-pair work is the edge/face distance
// and 13 barycentric tests.  The 13 x 2 stab flags of a point live in one
// register (the reference keeps them in a [64, N, 13, 2] byte buffer).
#include "api_common.hpp"

namespace nglod {

__constant__ double kStab[13][3] = {
    {1.0f, 0.0f, 0.0f},
    {0.0f, 1.0f, 0.0f},
    {0.0f, 0.0f, 1.0f},
    {0.0f, 0.707106781f, 0.707106781f},
    {0.707106781f, 0.0f, 0.707106781f},
    {0.707106781f, 0.707106781f, 0.0f},
    {0.0f, 0.707106781f, -0.707106781f},
    {0.707106781f, 0.0f, -0.707106781f},
    {0.707106781f, -0.707106781f, 0.0f},
    {0.577350269f, 0.577350269f, 0.577350269f},
    {-0.577350269f, 0.577350269f, 0.577350269f},
    {0.577350269f, -0.577350269f, 0.577350269f},
    {0.577350269f, 0.577350269f, -0.577350269f},
};

// Per-triangle FP64 data (80 doubles).
struct TriPrep {
    double a[3], b[3], c[3];
    double v10[3], v21[3], v02[3];
    double nor[3];
    double c10[3], c21[3], c02[3];
    double i10, i21, i02, inor;  // reciprocal squared lengths of v10, v21, v02, nor
    double degenerate;           // 1 if nor == 0
    double pvec[13][3];
    double idet[13];             // 0 where |det| < 1e-8 (no intersection)
};

// The FP64 distance and stabbing arithmetic lets the compiler fuse a multiply
// and an add (one rounding instead of two): the library is otherwise built
// without contraction so that FP32 results follow the reference op by op, but
// here the bar is the distance to 1e-12 relative (1e-15 absolute) and exact
// signs, and the fused form issues about a quarter fewer FP64 instructions.
#define NGLOD_FP64_CONTRACT _Pragma("clang fp contract(fast)")

__device__ __forceinline__ double dot3(const double* a, const double* b) {
    NGLOD_FP64_CONTRACT
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const double* a, const double* b, double* r) {
    NGLOD_FP64_CONTRACT
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

__global__ void tri_prep_kernel(const float* __restrict__ mesh, int64_t nf, TriPrep* __restrict__ out) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    TriPrep T;
    for (int k = 0; k < 3; ++k) {
        T.a[k] = mesh[9 * f + k];
        T.b[k] = mesh[9 * f + 3 + k];
        T.c[k] = mesh[9 * f + 6 + k];
    }
    for (int k = 0; k < 3; ++k) {
        T.v10[k] = T.b[k] - T.a[k];
        T.v21[k] = T.c[k] - T.b[k];
        T.v02[k] = T.a[k] - T.c[k];
    }
    cross3(T.v10, T.v02, T.nor);
    cross3(T.v10, T.nor, T.c10);
    cross3(T.v21, T.nor, T.c21);
    cross3(T.v02, T.nor, T.c02);
    T.i10 = 1.0 / dot3(T.v10, T.v10);
    T.i21 = 1.0 / dot3(T.v21, T.v21);
    T.i02 = 1.0 / dot3(T.v02, T.v02);
    T.inor = 1.0 / dot3(T.nor, T.nor);
    T.degenerate = (T.nor[0] == 0.0 && T.nor[1] == 0.0 && T.nor[2] == 0.0) ? 1.0 : 0.0;
    const double e2[3] = {-T.v02[0], -T.v02[1], -T.v02[2]};
    for (int i = 0; i < 13; ++i) {
        const double dir[3] = {kStab[i][0], kStab[i][1], kStab[i][2]};
        cross3(dir, e2, T.pvec[i]);
        const double det = dot3(T.v10, T.pvec[i]);
        T.idet[i] = (det > -1e-8 && det < 1e-8) ? 0.0 : 1.0 / det;
    }
    out[f] = T;
}

__device__ __forceinline__ double clamp01(double x) { return fmax(0.0, fmin(x, 1.0)); }
__device__ __forceinline__ double d2axmb(const double* a, double x, const double* b) {
    NGLOD_FP64_CONTRACT
    const double t0 = a[0] * x - b[0], t1 = a[1] * x - b[1], t2 = a[2] * x - b[2];
    return t0 * t0 + t1 * t1 + t2 * t2;
}

constexpr int kTriTile = 24;  // triangles per LDS tile (24 x 640 B = 15 KB)

__global__ __launch_bounds__(256) void mesh2sdf_kernel(const float* __restrict__ pts, int64_t n,
                                                       const TriPrep* __restrict__ tris, int64_t nf,
                                                       float* __restrict__ dist, double* __restrict__ dist64) {
    NGLOD_FP64_CONTRACT
    __shared__ TriPrep tile[kTriTile];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const double p[3] = {valid ? (double)pts[3 * i] : 0.0, valid ? (double)pts[3 * i + 1] : 0.0,
                         valid ? (double)pts[3 * i + 2] : 0.0};
    double best = INFINITY;
    uint32_t flags = 0;
    for (int64_t f0 = 0; f0 < nf; f0 += kTriTile) {
        const int cnt = (int)((nf - f0) < kTriTile ? (nf - f0) : kTriTile);
        __syncthreads();
        // cooperative copy of the tile (doubles)
        const double* src = reinterpret_cast<const double*>(tris + f0);
        double* dst = reinterpret_cast<double*>(tile);
        const int words = cnt * (int)(sizeof(TriPrep) / sizeof(double));
        for (int w = threadIdx.x; w < words; w += blockDim.x) dst[w] = src[w];
        __syncthreads();
        for (int k = 0; k < cnt; ++k) {
            const TriPrep& T = tile[k];
            const double p0[3] = {p[0] - T.a[0], p[1] - T.a[1], p[2] - T.a[2]};
            if (T.degenerate == 0.0) {
                const double p1[3] = {p[0] - T.b[0], p[1] - T.b[1], p[2] - T.b[2]};
                const double p2[3] = {p[0] - T.c[0], p[1] - T.c[1], p[2] - T.c[2]};
                const double s = copysign(1.0, dot3(T.c10, p0)) + copysign(1.0, dot3(T.c21, p1)) +
                                 copysign(1.0, dot3(T.c02, p2));
                double d2;
                if (s < 2.0) {
                    const double e1 = d2axmb(T.v10, clamp01(dot3(T.v10, p0) * T.i10), p0);
                    const double e2 = d2axmb(T.v21, clamp01(dot3(T.v21, p1) * T.i21), p1);
                    const double e3 = d2axmb(T.v02, clamp01(dot3(T.v02, p2) * T.i02), p2);
                    d2 = fmin(e1, fmin(e2, e3));
                } else {
                    const double dn = dot3(T.nor, p0);
                    d2 = dn * dn * T.inor;
                }
                best = fmin(best, d2 < 0.0 ? 0.0 : d2);
            }
            double qvec[3];
            cross3(p0, T.v10, qvec);
            const double e2v[3] = {-T.v02[0], -T.v02[1], -T.v02[2]};
            const double tnum = dot3(e2v, qvec);
#pragma unroll
            for (int d = 0; d < 13; ++d) {
                const double inv = T.idet[d];
                if (inv == 0.0) continue;
                const double u = dot3(p0, T.pvec[d]) * inv;
                if (u < 0.0 || u > 1.0) continue;
                const double dir[3] = {kStab[d][0], kStab[d][1], kStab[d][2]};
                const double v = dot3(dir, qvec) * inv;
                if (v < 0.0 || u + v > 1.0) continue;
                const double t = tnum * inv;
                flags |= (t >= 0.0) ? (1u << d) : (1u << (13 + d));
            }
        }
    }
    if (!valid) return;
    const uint32_t all = (1u << 13) - 1u;
    const bool inside = ((flags & all) == all) && (((flags >> 13) & all) == all);
    double d = sqrt(best < 0.0 ? 0.0 : best);
    d = inside ? -d : d;
    dist[i] = (float)d;
    if (dist64) dist64[i] = d;
}

// ---------------------------------------------------------------------------
// Counter-based RNG: SplitMix64 (Steele, Lea, Flood 2014) of (sample, draw).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double rnd01(uint64_t seed, uint64_t i, uint32_t k) {
    const uint64_t z = splitmix64(seed ^ splitmix64(i * 8ull + k));
    return (double)(z >> 40) * (1.0 / 16777216.0);
}

__global__ void sample_uniform_kernel(uint64_t seed, int64_t n, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; ++k) out[3 * i + k] = (float)(rnd01(seed, (uint64_t)i, (uint32_t)k) * 2.0 - 1.0);
}

// Directions uniform on S^2 (geoutils.sample_unif_sphere, sdf-net/lib/geoutils.py:31-40):
// z = 1 - 2u, phi = 2*pi*v, (sqrt(1 - z^2) cos phi, sqrt(1 - z^2) sin phi, z).
__global__ void sample_directions_kernel(uint64_t seed, int64_t n, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double z = 1.0 - 2.0 * rnd01(seed, (uint64_t)i, 0);
    const double r = sqrt(1.0 - z * z);
    const double phi = 2.0 * M_PI * rnd01(seed, (uint64_t)i, 1);
    out[3 * i] = (float)(r * cos(phi));
    out[3 * i + 1] = (float)(r * sin(phi));
    out[3 * i + 2] = (float)z;
}

__global__ void sample_surface_kernel(const float* __restrict__ mesh, const double* __restrict__ cdf, int64_t nf,
                                      uint64_t seed, int64_t n, double sigma, float* __restrict__ out,
                                      int32_t* __restrict__ face) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r0 = rnd01(seed, (uint64_t)i, 0);
    int64_t lo = 0, hi = nf - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cdf[mid] > r0) hi = mid;
        else lo = mid + 1;
    }
    if (face) face[i] = (int32_t)lo;
    const float* t = mesh + 9 * lo;
    const double u = sqrt(rnd01(seed, (uint64_t)i, 1));
    const double v = rnd01(seed, (uint64_t)i, 2);
    double g[3] = {0.0, 0.0, 0.0};
    if (sigma > 0.0) {  // Box-Muller
        const double u1 = 1.0 - rnd01(seed, (uint64_t)i, 3), u2 = rnd01(seed, (uint64_t)i, 4);
        const double u3 = 1.0 - rnd01(seed, (uint64_t)i, 5), u4 = rnd01(seed, (uint64_t)i, 6);
        const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
        g[0] = r1 * cos(2.0 * M_PI * u2);
        g[1] = r1 * sin(2.0 * M_PI * u2);
        g[2] = r2 * cos(2.0 * M_PI * u4);
    }
    for (int k = 0; k < 3; ++k) {
        const double p = (1.0 - u) * t[k] + (u * (1.0 - v)) * t[3 + k] + (u * v) * t[6 + k];
        out[3 * i + k] = (float)(p + sigma * g[k]);
    }
}

}  // namespace nglod

using namespace nglod;

extern "C" {

size_t nglod_mesh2sdf_workspace_size(int64_t nf) { return nf > 0 ? (size_t)nf * sizeof(TriPrep) : 0; }

int nglod_mesh2sdf_f64(const float* points, int64_t n, const float* mesh, int64_t nf, float* dist, double* dist64,
                       void* workspace, size_t workspace_bytes, void* stream) {
    if (n < 0 || nf < 0) return fail(NGLOD_ERR_ARG, "negative size");
    if (n == 0) return NGLOD_OK;
    if (nf == 0) return fail(NGLOD_ERR_ARG, "mesh has no triangles");
    if (!points || !mesh || !dist) return fail(NGLOD_ERR_ARG, "NULL pointer");
    if (!workspace || workspace_bytes < nglod_mesh2sdf_workspace_size(nf) || ((uintptr_t)workspace & 7))
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small or misaligned");
    hipStream_t s = (hipStream_t)stream;
    TriPrep* T = (TriPrep*)workspace;
    hipLaunchKernelGGL(tri_prep_kernel, dim3((unsigned)((nf + 127) / 128)), dim3(128), 0, s, mesh, nf, T);
    int rc = check_launch("tri_prep_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(mesh2sdf_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, points, n, T, nf, dist,
                       dist64);
    return check_launch("mesh2sdf_kernel");
}

int nglod_sample_uniform(int64_t n, uint64_t seed, float* out, void* stream) {
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n == 0) return NGLOD_OK;
    if (!out) return fail(NGLOD_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(sample_uniform_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       seed, n, out);
    return check_launch("sample_uniform_kernel");
}

int nglod_sample_directions(int64_t n, uint64_t seed, float* out, void* stream) {
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n == 0) return NGLOD_OK;
    if (!out) return fail(NGLOD_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(sample_directions_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       seed, n, out);
    return check_launch("sample_directions_kernel");
}

int nglod_sample_surface(const float* mesh, const double* cdf, int64_t nf, int64_t n, uint64_t seed, double sigma,
                         float* out, int32_t* face, void* stream) {
    if (n < 0 || nf < 0) return fail(NGLOD_ERR_ARG, "negative size");
    if (n == 0) return NGLOD_OK;
    if (nf == 0 || !mesh || !cdf || !out) return fail(NGLOD_ERR_ARG, "NULL pointer or empty mesh");
    hipLaunchKernelGGL(sample_surface_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       mesh, cdf, nf, seed, n, sigma, out, face);
    return check_launch("sample_surface_kernel");
}

}  // extern "C"
