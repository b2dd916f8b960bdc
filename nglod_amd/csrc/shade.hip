// Post-trace shading of a rendered frame, the consumers of the trace outputs:
//   nglod_shadow_rays   ground-plane hits + shadow-ray set-up (renderer.py:131-160)
//   nglod_ao            ambient occlusion, 40 SDF taps per hit (renderer.py:184-209)
//   nglod_shadow_map    shadow mask -> clamp -> Gaussian blur   (renderer.py:301-304)
//   nglod_matcap_shade  matcap lookup + segmentation + shadow/AO (renderer.py:279-307,
//                       geoutils.py:237-275)
// Every expression keeps the reference's operation order (the library is built
// with -ffp-contract=off, so a*b+c stays two roundings as in PyTorch); the
// matcap bilinear lookup and the blur run in FP64 as SciPy does.
#include <math.h>

#include "api_common.hpp"
#include "fp32_wave.hpp"
#include "split_wave.hpp"
#include "trace_common.hpp"

namespace nglod {

constexpr int kAoTaps = 40;

// ---------------------------------------------------------------------------
// Ground plane y = ground_y and shadow rays (renderer.py:131-160).  Per ray:
// plane_t = (o.y - ground_y) / (-d.y); the plane is hit when 0 < plane_t < 500
// and plane_t < t; such a ray leaves the surface-hit set and takes the plane's
// point, depth and normal (0, 1, 0).  Shadow ray: origin x + 0.1 n, direction
// normalize((noise + light) - origin); light_hit = dir . n > 0.
__global__ __launch_bounds__(256) void shadow_rays_kernel(const float* __restrict__ o, const float* __restrict__ d,
                                                          int64_t n, float ground_y, float lx, float ly, float lz,
                                                          const float* __restrict__ noise, float* __restrict__ x,
                                                          float* __restrict__ t, uint8_t* __restrict__ hit,
                                                          float* __restrict__ normal, uint8_t* __restrict__ plane,
                                                          float* __restrict__ so, float* __restrict__ sd,
                                                          uint8_t* __restrict__ light_hit) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float rate = -d[3 * i + 1];
    const float delta = o[3 * i + 1] - ground_y;
    const float pt = delta / rate;
    const bool ph = (pt > 0.0f) && (pt < 500.0f) && (pt < t[i]);
    float X0 = x[3 * i], X1 = x[3 * i + 1], X2 = x[3 * i + 2];
    float N0 = normal[3 * i], N1 = normal[3 * i + 1], N2 = normal[3 * i + 2];
    if (ph) {
        hit[i] = 0;
        t[i] = pt;
        X0 = o[3 * i] + d[3 * i] * pt;
        X1 = o[3 * i + 1] + d[3 * i + 1] * pt;
        X2 = o[3 * i + 2] + d[3 * i + 2] * pt;
        N0 = 0.0f;
        N1 = 1.0f;
        N2 = 0.0f;
        x[3 * i] = X0;
        x[3 * i + 1] = X1;
        x[3 * i + 2] = X2;
        normal[3 * i] = N0;
        normal[3 * i + 1] = N1;
        normal[3 * i + 2] = N2;
    }
    plane[i] = ph ? 1 : 0;
    const float s0 = X0 + 0.1f * N0, s1 = X1 + 0.1f * N1, s2 = X2 + 0.1f * N2;
    const float e0 = noise ? noise[3 * i] : 0.0f, e1 = noise ? noise[3 * i + 1] : 0.0f,
                e2 = noise ? noise[3 * i + 2] : 0.0f;
    const float v0 = (e0 + lx) - s0, v1 = (e1 + ly) - s1, v2 = (e2 + lz) - s2;
    float nn = sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
    nn = nn < 1e-12f ? 1e-12f : nn;
    const float r0 = v0 / nn, r1 = v1 / nn, r2 = v2 / nn;
    so[3 * i] = s0;
    so[3 * i + 1] = s1;
    so[3 * i + 2] = s2;
    sd[3 * i] = r0;
    sd[3 * i + 1] = r1;
    sd[3 * i + 2] = r2;
    light_hit[i] = (r0 * N0 + r1 * N1 + r2 * N2) > 0.0f ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Ambient occlusion (renderer.py:184-209): for tap i = 0..39 at distance
// dist[i] = 0.1 * 0.25 * ((i + 1) / 41)^1.6 along the normal, r = sdf(x + n *
// dist[i]) (a plane hit takes min(r, dist[i])), acc += 3.5 * relu(dist[i] - r -
// 0.0015); ao = clamp(1 - acc, 0.1, 1)^2, and 1 where neither hit is set.
// Persistent: lanes claim the marked rays in frame order (claim_hits), each
// runs its 40 taps through the FP32 MFMA decoder, LODs 0-1 from LDS.
struct AoArgs {
    ModelArgs m;
    int lod, dec;
    CoarsePlan coarse;
    const float* x;
    const float* normal;
    const uint8_t* hit;
    const uint8_t* plane;  // may be NULL
    int64_t n;
    float* ao;
    uint32_t* counter;
    float dist[kAoTaps];
};

__global__ __launch_bounds__(256) void ao_fill_kernel(const uint8_t* __restrict__ hit, const uint8_t* __restrict__ plane,
                                                      int64_t n, float* __restrict__ ao, uint32_t* counter) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) counter[0] = 0u;
    if (i >= n) return;
    if (!hit[i] && !(plane && plane[i])) ao[i] = 1.0f;
}

template <bool POSINV>
__global__ __launch_bounds__(kMarchBlock) void ao_kernel(AoArgs a) {
    __shared__ Fp32Lds lds;
    __shared__ CoarseLds32 coarse;
    stage_fp32_lds(a.m, a.dec, &lds);
    stage_coarse32(a.m, a.coarse, &coarse);
    __syncthreads();
    const float b1 = a.m.b1[a.dec][0];
    const uint32_t lane = lane_id();
    bool drained = false;
    for (;;) {
        const int64_t id = claim_hits(a.counter, a.hit, (uint32_t)a.n, lane, drained, a.plane);
        const bool valid = id >= 0;
        if (__ballot(valid) == 0) break;
        float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f, n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
        bool on_plane = false;
        if (valid) {
            p0 = a.x[3 * id];
            p1 = a.x[3 * id + 1];
            p2 = a.x[3 * id + 2];
            n0 = a.normal[3 * id];
            n1 = a.normal[3 * id + 1];
            n2 = a.normal[3 * id + 2];
            on_plane = !a.hit[id];
        }
        float acc = 0.0f;
#pragma unroll 1
        for (int k = 0; k < kAoTaps; ++k) {
            const float dk = a.dist[k];
            const float q0 = p0 + n0 * dk, q1 = p1 + n1 * dk, q2 = p2 + n2 * dk;
            float r = sdf_fp32_wave<POSINV>(a.m, a.lod, &lds, b1, lane, valid, q0, q1, q2, a.coarse, &coarse);
            if (on_plane) r = fminf(r, dk);
            const float u = (dk - r) - 0.0015f;
            acc = acc + 3.5f * (u > 0.0f ? u : 0.0f);
        }
        if (valid) {
            float v = 1.0f - acc;
            v = v < 0.1f ? 0.1f : (v > 1.0f ? 1.0f : v);
            a.ao[id] = v * v;
        }
    }
}

// The mixed form's AO: the same loop on the FP32 cumulative grid (one level,
// no coarse levels to stage) with the decoder of the mixed trace's FP32 pass
// (exact three-term BF16 splits on the matrix cores, FP32-level results).
template <bool POSINV>
__global__ __launch_bounds__(kMarchBlock) void ao_split_kernel(AoArgs a) {
    __shared__ SplitLds lds;
    stage_split_lds(a.m, a.dec, &lds);
    __syncthreads();
    const float b1 = a.m.b1[a.dec][0];
    const uint32_t lane = lane_id();
    bool drained = false;
    for (;;) {
        const int64_t id = claim_hits(a.counter, a.hit, (uint32_t)a.n, lane, drained, a.plane);
        const bool valid = id >= 0;
        if (__ballot(valid) == 0) break;
        float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f, n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
        bool on_plane = false;
        if (valid) {
            p0 = a.x[3 * id];
            p1 = a.x[3 * id + 1];
            p2 = a.x[3 * id + 2];
            n0 = a.normal[3 * id];
            n1 = a.normal[3 * id + 1];
            n2 = a.normal[3 * id + 2];
            on_plane = !a.hit[id];
        }
        float acc = 0.0f;
#pragma unroll 1
        for (int k = 0; k < kAoTaps; ++k) {
            const float dk = a.dist[k];
            const float q0 = p0 + n0 * dk, q1 = p1 + n1 * dk, q2 = p2 + n2 * dk;
            float r = sdf_split_wave<POSINV>(a.m, &lds, b1, lane, valid, q0, q1, q2);
            if (on_plane) r = fminf(r, dk);
            const float u = (dk - r) - 0.0015f;
            acc = acc + 3.5f * (u > 0.0f ? u : 0.0f);
        }
        if (valid) {
            float v = 1.0f - acc;
            v = v < 0.1f ? 0.1f : (v > 1.0f ? 1.0f : v);
            a.ao[id] = v * v;
        }
    }
}

// ---------------------------------------------------------------------------
// Shadow map [width][height]: clamp(1 - shadow + 0.9, 0, 1) with shadow =
// shadow_hit & light_hit, then scipy.ndimage.gaussian_filter(sigma = 2,
// mode 'reflect', truncate 4): a 17-tap kernel exp(-x^2 / 8) normalised, applied
// along axis 0 then axis 1, each pass accumulating in FP64 as
// c * w[0] + sum_{j = 8..1} (left_j + right_j) * w[j] and rounding to FP32.
constexpr int kBlurRadius = 8;

struct BlurWeights {
    double w[kBlurRadius + 1];  // w[j]: weight at distance j
};

__device__ __forceinline__ int reflect_index(int i, int len) {
    // half-sample symmetric extension: (d c b a | a b c d | d c b a)
    for (;;) {
        if (i < 0)
            i = -i - 1;
        else if (i >= len)
            i = 2 * len - i - 1;
        else
            return i;
    }
}

__global__ __launch_bounds__(256) void shadow_mask_kernel(const uint8_t* __restrict__ shadow_hit,
                                                          const uint8_t* __restrict__ light_hit, int64_t n,
                                                          float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = (shadow_hit[i] && light_hit[i]) ? 1.0f : 0.0f;
    float v = (1.0f - s) + 0.9f;
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    out[i] = v;
}

// One 1-D pass over an [n0][n1] array along `axis` (0: stride n1, 1: stride 1).
__global__ __launch_bounds__(256) void blur_pass_kernel(const float* __restrict__ in, float* __restrict__ out, int n0,
                                                        int n1, int axis, BlurWeights bw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n0 * n1) return;
    const int r = (int)(i / n1), c = (int)(i % n1);
    const int len = axis == 0 ? n0 : n1;
    const int pos = axis == 0 ? r : c;
    auto at = [&](int k) -> double {
        const int j = reflect_index(k, len);
        return (double)(axis == 0 ? in[(int64_t)j * n1 + c] : in[(int64_t)r * n1 + j]);
    };
    double acc = at(pos) * bw.w[0];
#pragma unroll
    for (int j = kBlurRadius; j >= 1; --j) acc += (at(pos - j) + at(pos + j)) * bw.w[j];
    out[i] = (float)acc;
}

// ---------------------------------------------------------------------------
// Matcap shading (renderer.py:279-307): view' = view @ mm^T (when given);
// spherical_envmap (geoutils.py:253-275) gives vN in [0,1]^2; rgb =
// RegularGridInterpolator over linspace(0, 1, tw) x linspace(0, 1, th) of the
// texture [tw][th][channels] (the image transposed to x-major), linear, in FP64,
// then FP32 / 255.  Pixels without a surface hit take rgb = normal = 1; rgb is
// then scaled by the shadow map and the AO when given.
__device__ __forceinline__ int grid_interval(double u, int npts, double step, double& frac) {
    // numpy.linspace(0, 1, npts): g_i = i * step, g_{npts-1} = 1; SciPy's
    // interval i has g_i <= u < g_{i+1} (u = 1 takes the last interval).
    auto g = [&](int k) -> double { return k == npts - 1 ? 1.0 : (double)k * step; };
    int i = (int)(u * (double)(npts - 1));
    i = i < 0 ? 0 : (i > npts - 2 ? npts - 2 : i);
    while (i > 0 && g(i) > u) --i;
    while (i < npts - 2 && g(i + 1) <= u) ++i;
    frac = (u - g(i)) / (g(i + 1) - g(i));
    return i;
}

__global__ __launch_bounds__(256) void matcap_kernel(const float* __restrict__ view, float* __restrict__ normal,
                                                     const uint8_t* __restrict__ hit, int64_t n, int use_mm, float m00,
                                                     float m01, float m02, float m10, float m11, float m12, float m20,
                                                     float m21, float m22, const float* __restrict__ tex, int tw,
                                                     int th, int channels, const float* __restrict__ shadow_map,
                                                     const float* __restrict__ ao, float* __restrict__ rgb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float d0 = view[3 * i], d1 = view[3 * i + 1], d2 = view[3 * i + 2];
    if (use_mm) {
        // torch.mm(view, mm.transpose(1, 0)): row i of the result is view . mm[j, :]
        const float e0 = d0 * m00 + d1 * m01 + d2 * m02;
        const float e1 = d0 * m10 + d1 * m11 + d2 * m12;
        const float e2 = d0 * m20 + d1 * m21 + d2 * m22;
        d0 = e0;
        d1 = e1;
        d2 = e2;
    }
    const float N0 = normal[3 * i], N1 = normal[3 * i + 1], N2 = normal[3 * i + 2];
    const float s0 = d0, s1 = d1, s2 = d2 * -1.0f;
    const float dot = N0 * s0 + N1 * s1 + N2 * s2;
    float r0 = s0 - 2.0f * dot * N0, r1 = s1 - 2.0f * dot * N1, r2 = s2 - 2.0f * dot * N2;
    r2 = r2 - 1.0f;
    const float m = 2.0f * sqrtf(r0 * r0 + r1 * r1 + r2 * r2);
    float u = 1.0f - (r0 / m + 0.5f), v = 1.0f - (r1 / m + 0.5f);
    u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    if (u != u) u = 0.0f;
    if (v != v) v = 0.0f;
    double fu, fv;
    const int iu = grid_interval((double)u, tw, 1.0 / (double)(tw - 1), fu);
    const int iv = grid_interval((double)v, th, 1.0 / (double)(th - 1), fv);
    const bool h = hit[i] != 0;
    float out[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double t00 = tex[((int64_t)iu * th + iv) * channels + c];
        const double t01 = tex[((int64_t)iu * th + iv + 1) * channels + c];
        const double t10 = tex[((int64_t)(iu + 1) * th + iv) * channels + c];
        const double t11 = tex[((int64_t)(iu + 1) * th + iv + 1) * channels + c];
        double val = 0.0;
        val = val + t00 * ((1.0 * (1.0 - fu)) * (1.0 - fv));
        val = val + t01 * ((1.0 * (1.0 - fu)) * fv);
        val = val + t10 * ((1.0 * fu) * (1.0 - fv));
        val = val + t11 * ((1.0 * fu) * fv);
        float px = (float)val / 255.0f;
        if (!h) px = 1.0f;
        if (shadow_map) px = px * shadow_map[i];
        if (ao) px = px * ao[i];
        out[c] = px;
    }
    rgb[3 * i] = out[0];
    rgb[3 * i + 1] = out[1];
    rgb[3 * i + 2] = out[2];
    if (!h) {
        normal[3 * i] = 1.0f;
        normal[3 * i + 1] = 1.0f;
        normal[3 * i + 2] = 1.0f;
    }
}

}  // namespace nglod

using namespace nglod;

static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

extern "C" int nglod_shadow_rays(const float* ray_o, const float* ray_d, int64_t n, float ground_y, const float* light,
                                 const float* noise, float* x, float* t, uint8_t* hit, float* normal,
                                 uint8_t* plane_hit, float* shadow_o, float* shadow_d, uint8_t* light_hit,
                                 void* stream) {
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (!light) return fail(NGLOD_ERR_ARG, "light is NULL");
    if (n == 0) return NGLOD_OK;
    if (!ray_o || !ray_d || !x || !t || !hit || !normal || !plane_hit || !shadow_o || !shadow_d || !light_hit)
        return fail(NGLOD_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(shadow_rays_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, ray_o, ray_d, n,
                       ground_y, light[0], light[1], light[2], noise, x, t, hit, normal, plane_hit, shadow_o,
                       shadow_d, light_hit);
    return check_launch("shadow_rays_kernel");
}

extern "C" size_t nglod_ao_workspace_size(void) { return 256; }

extern "C" int nglod_ao(const nglod_model_t* model, const float* x, const float* normal, const uint8_t* hit,
                        const uint8_t* plane_hit, int64_t n, int32_t lod, int32_t form, float* ao, void* workspace,
                        size_t workspace_bytes, void* stream) {
    // AO taps are FP32-level in both forms, as the normals are; the mixed
    // form takes them from the FP32 cumulative grid with its trace's FP32-pass
    // decoder.
    AoArgs a;
    int rc = make_model_args(model, lod, form, &a.m);
    if (rc) return rc;
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (n > (int64_t)0xFFFFFFFFll - 1024) return fail(NGLOD_ERR_UNSUPPORTED, "n >= 2^32");
    if (n == 0) return NGLOD_OK;
    if (!x || !normal || !hit || !ao) return fail(NGLOD_ERR_ARG, "NULL pointer");
    if (!workspace || workspace_bytes < nglod_ao_workspace_size())
        return fail(NGLOD_ERR_WORKSPACE, "workspace smaller than nglod_ao_workspace_size()");
    a.lod = gather_lod(form, lod);
    a.dec = decoder_of(model, lod);
    // the mixed AO kernel gathers one level, the cumulative grid of the LOD
    if (form == NGLOD_MIXED && (a.lod != 0 || a.m.num_lods != 1))
        return fail(NGLOD_ERR_ARG, "mixed form: one cumulative level (gather lod 0)");
    a.coarse = coarse_plan(a.m, a.lod);
    a.x = x;
    a.normal = normal;
    a.hit = hit;
    a.plane = plane_hit;
    a.n = n;
    a.ao = ao;
    a.counter = (uint32_t*)workspace;
    for (int k = 0; k < kAoTaps; ++k)
        a.dist[k] = (float)(0.1 * 0.25 * pow((double)(k + 1) / (double)(kAoTaps + 1), 1.6));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(ao_fill_kernel, dim3(blocks_for(n)), dim3(256), 0, s, hit, plane_hit, n, ao, a.counter);
    rc = check_launch("ao_fill_kernel");
    if (rc) return rc;
    int per_cu = 0;
    auto kern = form == NGLOD_MIXED ? (a.m.pos_inv ? ao_split_kernel<true> : ao_split_kernel<false>)
                                    : (a.m.pos_inv ? ao_kernel<true> : ao_kernel<false>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kMarchBlock, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const int64_t want = (n + kMarchBlock - 1) / kMarchBlock;
    const int64_t cap = (int64_t)device_cu_count() * per_cu;
    const unsigned grid = (unsigned)(want < cap ? want : cap);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kMarchBlock), 0, s, a);
    return check_launch("ao_kernel");
}

// SciPy's _gaussian_kernel1d(sigma = 2, radius 8): phi = exp(-0.5 / sigma^2 *
// x^2) over x = -8..8, divided by its sum, the sum taken in NumPy's pairwise
// order for 17 values (8 running partial sums over x[0..15], combined as
// ((p0+p1)+(p2+p3))+((p4+p5)+(p6+p7)), then + x[16]).
static BlurWeights blur_weights() {
    double phi[2 * kBlurRadius + 1];
    const double sigma = 2.0;
    const double c = -0.5 / (sigma * sigma);
    for (int k = 0; k <= 2 * kBlurRadius; ++k) {
        const double xx = (double)(k - kBlurRadius);
        phi[k] = exp(c * (xx * xx));
    }
    double p[8];
    for (int j = 0; j < 8; ++j) p[j] = phi[j];
    for (int k = 8; k < 16; ++k) p[k - 8] += phi[k];
    double sum = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    sum += phi[16];
    BlurWeights bw;
    for (int j = 0; j <= kBlurRadius; ++j) bw.w[j] = phi[kBlurRadius - j] / sum;
    return bw;
}

extern "C" int nglod_shadow_map(const uint8_t* shadow_hit, const uint8_t* light_hit, int32_t width, int32_t height,
                                float* map, float* tmp, void* stream) {
    if (width < 0 || height < 0) return fail(NGLOD_ERR_ARG, "negative size");
    const int64_t n = (int64_t)width * height;
    if (n == 0) return NGLOD_OK;
    if (!shadow_hit || !light_hit || !map || !tmp) return fail(NGLOD_ERR_ARG, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    const BlurWeights bw = blur_weights();
    hipLaunchKernelGGL(shadow_mask_kernel, dim3(blocks_for(n)), dim3(256), 0, s, shadow_hit, light_hit, n, map);
    int rc = check_launch("shadow_mask_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(blur_pass_kernel, dim3(blocks_for(n)), dim3(256), 0, s, map, tmp, width, height, 0, bw);
    rc = check_launch("blur_pass_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(blur_pass_kernel, dim3(blocks_for(n)), dim3(256), 0, s, tmp, map, width, height, 1, bw);
    return check_launch("blur_pass_kernel");
}

extern "C" int nglod_matcap_shade(const float* view, float* normal, const uint8_t* hit, int64_t n, const float* mm,
                                  const float* texture, int32_t tw, int32_t th, int32_t channels,
                                  const float* shadow_map, const float* ao, float* rgb, void* stream) {
    if (n < 0) return fail(NGLOD_ERR_ARG, "n < 0");
    if (tw < 2 || th < 2 || channels < 3) return fail(NGLOD_ERR_ARG, "texture must be at least 2x2 with 3 channels");
    if (n == 0) return NGLOD_OK;
    if (!view || !normal || !hit || !texture || !rgb) return fail(NGLOD_ERR_ARG, "NULL pointer");
    float m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (mm)
        for (int k = 0; k < 9; ++k) m[k] = mm[k];
    hipLaunchKernelGGL(matcap_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, view, normal, hit, n,
                       mm ? 1 : 0, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], texture, tw, th, channels,
                       shadow_map, ao, rgb);
    return check_launch("matcap_kernel");
}
