// Gradients of the SDF with respect to the query position and the
// minimum-tracking sphere trace:
//   nglod_gradient     diffutils.gradient (sdf-net/lib/diffutils.py:29-84):
//                      'finitediff', 'tetrahedron' and 'autodiff', optionally
//                      masked and normalised as the tracers use it
//                      (SphereTracer.py:128-130, :214-215)
//   nglod_trace_min    SphereTracer.get_min (sdf-net/lib/tracer/SphereTracer.py:134-218)
// FP32, one query per lane, the decoder on the vector ALU (decode_fp32): these
// are not the frame's hot path.
#include "api_common.hpp"
#include "trace_common.hpp"

namespace nglod {

// Tetrahedron taps (diffutils.py:38-62): f at x + h k_j for the four corners
// k_j of a tetrahedron, gradient sum_j k_j f(x + h k_j) / (4 h), h = 1/192.
constexpr float kTetraFourH = (float)((1.0 / (64.0 * 3.0)) * 4.0);

// One query's SDF at `lod` and its gradient with respect to the position, as
// torch autograd of OctreeSDF.sdf (OctreeSDF.py:94-155) gives it:
//  - F.grid_sample's backward with align_corners=True and padding_mode=
//    'border': d u / d p = r / 2 per axis, and 0 where the unclamped u is <= 0
//    or >= r (the border counts as outside, so the clamp passes no gradient);
//    within the cell [floor(u), floor(u) + 1] the derivative of the trilinear
//    weights, (-1, +1) on the axis times the other two axes' weights;
//  - ReLU's backward: hidden unit j contributes where its pre-activation is > 0;
//  - a position-dependent decoder's xyz columns of W0 add their own term.
// The SDF value is computed exactly as decode_fp32 computes it.
template <bool POSINV>
__device__ __forceinline__ float sdf_and_grad_fp32(const ModelArgs& m, int lod, int dec, float px, float py,
                                                   float pz, float& gx, float& gy, float& gz) {
    constexpr int D = POSINV ? kC : kC + 3;
    constexpr int OFF = POSINV ? 0 : 3;
    float f[kC];
    gather_fp32(m, lod, px, py, pz, f);
    const float* W0 = m.w0[dec];
    const float* B0 = m.b0[dec];
    const float* W1 = m.w1[dec];
    float v[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) v[c] = 0.0f;
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
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
        const float pre = acc + ldc(B0, j);
        const float w = ldc(W1, j);
        out = __builtin_fmaf(w, fmaxf(pre, 0.0f), out);
        if (pre > 0.0f) {
#pragma unroll
            for (int k = 0; k < kC; ++k) v[k] = __builtin_fmaf(w, ldc(W0, j * D + OFF + k), v[k]);
            if (!POSINV) {
                vx = __builtin_fmaf(w, ldc(W0, j * D + 0), vx);
                vy = __builtin_fmaf(w, ldc(W0, j * D + 1), vy);
                vz = __builtin_fmaf(w, ldc(W0, j * D + 2), vz);
            }
        }
    }
    gx = vx;
    gy = vy;
    gz = vz;
    for (int l = 0; l <= lod; ++l) {
        const int r = m.res[l];
        const int S = r + 1;
        const float rf = (float)r;
        const AxisW ax = axis_weights(px, r);
        const AxisW ay = axis_weights(py, r);
        const AxisW az = axis_weights(pz, r);
        const float ux = ((px + 1.0f) * 0.5f) * rf, uy = ((py + 1.0f) * 0.5f) * rf, uz = ((pz + 1.0f) * 0.5f) * rf;
        const float sxg = (ux > 0.0f && ux < rf) ? rf * 0.5f : 0.0f;
        const float syg = (uy > 0.0f && uy < rf) ? rf * 0.5f : 0.0f;
        const float szg = (uz > 0.0f && uz < rf) ? rf * 0.5f : 0.0f;
        const float* base = m.grid[l] + ((size_t)((az.i0 * S + ay.i0) * S + ax.i0)) * kC;
        const size_t stride[3] = {(size_t)kC, (size_t)S * kC, (size_t)S * S * kC};
        float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int bx = k & 1, by = (k >> 1) & 1, bz = k >> 2;
            const float4* q = reinterpret_cast<const float4*>(base + bx * stride[0] + by * stride[1] + bz * stride[2]);
            float pk = 0.0f;  // corner k's features projected on d sdf / d feature
#pragma unroll
            for (int c4 = 0; c4 < kC / 4; ++c4) {
                const float4 gq = q[c4];
                pk = __builtin_fmaf(gq.x, v[4 * c4 + 0], pk);
                pk = __builtin_fmaf(gq.y, v[4 * c4 + 1], pk);
                pk = __builtin_fmaf(gq.z, v[4 * c4 + 2], pk);
                pk = __builtin_fmaf(gq.w, v[4 * c4 + 3], pk);
            }
            const float wx = bx ? ax.w1 : ax.w0, wy = by ? ay.w1 : ay.w0, wz = bz ? az.w1 : az.w0;
            dx = __builtin_fmaf(bx ? pk : -pk, wy * wz, dx);
            dy = __builtin_fmaf(by ? pk : -pk, wx * wz, dy);
            dz = __builtin_fmaf(bz ? pk : -pk, wx * wy, dz);
        }
        gx = __builtin_fmaf(sxg, dx, gx);
        gy = __builtin_fmaf(syg, dy, gy);
        gz = __builtin_fmaf(szg, dz, gz);
    }
    return out + ldc(m.b1[dec], 0);
}

template <bool POSINV, int METHOD>
__global__ __launch_bounds__(256) void gradient_fp32_kernel(ModelArgs m, int lod, int dec, const float* __restrict__ x,
                                                            const uint8_t* __restrict__ mask, int64_t n, int normalize,
                                                            float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    if (!mask || mask[i]) {
        const float p0 = x[3 * i], p1 = x[3 * i + 1], p2 = x[3 * i + 2];
        if (METHOD == NGLOD_GRAD_FINITEDIFF) {
            // diffutils.py:61-70: (f(x + h e_a) - f(x - h e_a)) / (2h)
            float g[3];
#pragma unroll 1
            for (int a = 0; a < 3; ++a) {
                const float fp = sdf_fp32<POSINV>(m, lod, dec, a == 0 ? p0 + kFdH : p0 + 0.0f,
                                                  a == 1 ? p1 + kFdH : p1 + 0.0f, a == 2 ? p2 + kFdH : p2 + 0.0f);
                const float fm = sdf_fp32<POSINV>(m, lod, dec, a == 0 ? p0 - kFdH : p0 - 0.0f,
                                                  a == 1 ? p1 - kFdH : p1 - 0.0f, a == 2 ? p2 - kFdH : p2 - 0.0f);
                g[a] = (fp - fm) / kFdTwoH;
            }
            g0 = g[0];
            g1 = g[1];
            g2 = g[2];
        } else if (METHOD == NGLOD_GRAD_TETRAHEDRON) {
            // k0 = (1,-1,-1), k1 = (-1,-1,1), k2 = (-1,1,-1), k3 = (1,1,1)
            const float f0 = sdf_fp32<POSINV>(m, lod, dec, p0 + kFdH, p1 + (-kFdH), p2 + (-kFdH));
            const float f1 = sdf_fp32<POSINV>(m, lod, dec, p0 + (-kFdH), p1 + (-kFdH), p2 + kFdH);
            const float f2 = sdf_fp32<POSINV>(m, lod, dec, p0 + (-kFdH), p1 + kFdH, p2 + (-kFdH));
            const float f3 = sdf_fp32<POSINV>(m, lod, dec, p0 + kFdH, p1 + kFdH, p2 + kFdH);
            g0 = (((f0 + (-f1)) + (-f2)) + f3) / kTetraFourH;
            g1 = ((((-f0) + (-f1)) + f2) + f3) / kTetraFourH;
            g2 = ((((-f0) + f1) + (-f2)) + f3) / kTetraFourH;
        } else {
            sdf_and_grad_fp32<POSINV>(m, lod, dec, p0, p1, p2, g0, g1, g2);
        }
        if (normalize) normalize_grad(g0, g1, g2, g0, g1, g2);
    }
    out[3 * i] = g0;
    out[3 * i + 1] = g1;
    out[3 * i + 2] = g2;
}

// SDF values and autodiff gradients together (out_sdf [n], out_grad [n][3]).
template <bool POSINV>
__global__ __launch_bounds__(256) void sdf_grad_fp32_kernel(ModelArgs m, int lod, int dec, const float* __restrict__ x,
                                                            int64_t n, float* __restrict__ out_sdf,
                                                            float* __restrict__ out_grad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g0, g1, g2;
    const float y = sdf_and_grad_fp32<POSINV>(m, lod, dec, x[3 * i], x[3 * i + 1], x[3 * i + 2], g0, g1, g2);
    if (out_sdf) out_sdf[i] = y;
    out_grad[3 * i] = g0;
    out_grad[3 * i + 1] = g1;
    out_grad[3 * i + 2] = g2;
}

// ---------------------------------------------------------------------------
// SphereTracer.get_min, per ray (SphereTracer.py:134-218).  Differences from
// forward: every ray marches from its box entry (or from its origin, t = 0,
// when there is none: cond does not start from the box test), and the loop
// keeps the smallest distance seen, mind, with minx.  In the loop body the
// comparison d < mind runs after x has moved to o + d t, so minx is the point
// reached by the step of length d, not the point where d was evaluated.  The
// batch loop stops at the first iteration at which no ray is still marching;
// a ray that stops at iteration k < that one still runs the comparison once
// more at k (with x unchanged), a ray that stops at the batch's last
// iteration does not.  The first kernel marches each ray and records k and
// whether that last comparison would move minx; the batch's last iteration is
// the largest k (atomicMax); the second kernel applies the comparison.
struct TraceMinArgs {
    const float* ray_o;
    const float* ray_d;
    int64_t n;
    int num_steps;
    float min_dis, min_dis3, step, far_clip;
    float* x;
    float* t;
    uint8_t* hit;
    float* min_x;
    uint32_t* kinfo;  // [n]: k | (pending comparison << 31)
    uint32_t* kmax;   // [1]: the batch's last iteration
};

__global__ void trace_min_clear_kernel(uint32_t* kmax) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *kmax = 0u;
}

template <bool POSINV>
__global__ __launch_bounds__(256) void trace_min_kernel(ModelArgs m, int lod, int dec, TraceMinArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float ox = a.ray_o[3 * i], oy = a.ray_o[3 * i + 1], oz = a.ray_o[3 * i + 2];
    const float dx = a.ray_d[3 * i], dy = a.ray_d[3 * i + 1], dz = a.ray_d[3 * i + 2];
    const BoxEntry e = box_entry(ox, oy, oz, dx, dy, dz);
    float x0 = e.x, x1 = e.y, x2 = e.z, t = e.t;
    float D = sdf_fp32<POSINV>(m, lod, dec, x0, x1, x2);
    float Dp = D, mind = D;
    float m0 = x0, m1 = x1, m2 = x2;
    bool hitflag = false;
    int k = 0;
    bool pending = false;
    for (; k < a.num_steps; ++k) {
        hitflag = fabsf(t) < a.far_clip;
        const bool cond = (fabsf(D) > a.min_dis) && (fabsf((D + Dp) * 0.5f) > a.min_dis3) && hitflag;
        if (!cond) {
            pending = D < mind;
            break;
        }
        x0 = ox + dx * t;
        x1 = oy + dy * t;
        x2 = oz + dz * t;
        if (D < mind) {
            mind = D;
            m0 = x0;
            m1 = x1;
            m2 = x2;
        }
        Dp = D;
        D = sdf_fp32<POSINV>(m, lod, dec, x0, x1, x2) * a.step;
        t = t + D;
    }
    const bool outside = (fabsf(x0) > 1.0f) || (fabsf(x1) > 1.0f) || (fabsf(x2) > 1.0f);
    a.x[3 * i] = x0;
    a.x[3 * i + 1] = x1;
    a.x[3 * i + 2] = x2;
    a.t[i] = t;
    a.hit[i] = (hitflag && !outside) ? 1 : 0;
    a.min_x[3 * i] = m0;
    a.min_x[3 * i + 1] = m1;
    a.min_x[3 * i + 2] = m2;
    a.kinfo[i] = (uint32_t)k | (pending ? 0x80000000u : 0u);
    atomicMax(a.kmax, (uint32_t)k);
}

__global__ __launch_bounds__(256) void trace_min_finish_kernel(TraceMinArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t info = a.kinfo[i];
    const uint32_t k = info & 0x7fffffffu;
    if ((info & 0x80000000u) && k < *a.kmax) {
        a.min_x[3 * i] = a.x[3 * i];
        a.min_x[3 * i + 1] = a.x[3 * i + 1];
        a.min_x[3 * i + 2] = a.x[3 * i + 2];
    }
}

template <bool POSINV>
static int launch_gradient(const ModelArgs& m, int lod, int dec, const float* x, const uint8_t* mask, int64_t n,
                           int method, int normalize, float* out, hipStream_t s) {
    const dim3 grid((unsigned)((n + 255) / 256));
    if (method == NGLOD_GRAD_FINITEDIFF)
        hipLaunchKernelGGL((gradient_fp32_kernel<POSINV, NGLOD_GRAD_FINITEDIFF>), grid, dim3(256), 0, s, m, lod, dec,
                           x, mask, n, normalize, out);
    else if (method == NGLOD_GRAD_TETRAHEDRON)
        hipLaunchKernelGGL((gradient_fp32_kernel<POSINV, NGLOD_GRAD_TETRAHEDRON>), grid, dim3(256), 0, s, m, lod, dec,
                           x, mask, n, normalize, out);
    else
        hipLaunchKernelGGL((gradient_fp32_kernel<POSINV, NGLOD_GRAD_AUTODIFF>), grid, dim3(256), 0, s, m, lod, dec, x,
                           mask, n, normalize, out);
    return check_launch("gradient_fp32_kernel");
}

}  // namespace nglod

using namespace nglod;

extern "C" {

int nglod_gradient(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, int32_t form, int32_t method,
                   const uint8_t* mask, int32_t normalize, float* grad, void* stream) {
    if (method != NGLOD_GRAD_FINITEDIFF && method != NGLOD_GRAD_TETRAHEDRON && method != NGLOD_GRAD_AUTODIFF)
        return fail(NGLOD_ERR_ARG, "unknown gradient method");
    ModelArgs m;
    int rc = make_model_args(model, lod, form, &m);
    if (rc) return rc;
    if (n < 0 || n > (int64_t)0x7fff0000) return fail(NGLOD_ERR_ARG, "n out of range");
    if (n == 0) return NGLOD_OK;
    if (!x || !grad) return fail(NGLOD_ERR_ARG, "NULL pointer");
    const int dec = decoder_of(model, lod);
    const int gl = gather_lod(form, lod);
    hipStream_t s = (hipStream_t)stream;
    return m.pos_inv ? launch_gradient<true>(m, gl, dec, x, mask, n, method, normalize ? 1 : 0, grad, s)
                     : launch_gradient<false>(m, gl, dec, x, mask, n, method, normalize ? 1 : 0, grad, s);
}

int nglod_sdf_grad(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, int32_t form, float* sdf,
                   float* grad, void* stream) {
    ModelArgs m;
    int rc = make_model_args(model, lod, form, &m);
    if (rc) return rc;
    if (n < 0 || n > (int64_t)0x7fff0000) return fail(NGLOD_ERR_ARG, "n out of range");
    if (n == 0) return NGLOD_OK;
    if (!x || !grad) return fail(NGLOD_ERR_ARG, "NULL pointer");
    const int dec = decoder_of(model, lod);
    const int gl = gather_lod(form, lod);
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (m.pos_inv)
        hipLaunchKernelGGL(sdf_grad_fp32_kernel<true>, grid, dim3(256), 0, s, m, gl, dec, x, n, sdf, grad);
    else
        hipLaunchKernelGGL(sdf_grad_fp32_kernel<false>, grid, dim3(256), 0, s, m, gl, dec, x, n, sdf, grad);
    return check_launch("sdf_grad_fp32_kernel");
}

size_t nglod_trace_min_workspace_size(int64_t n) {
    if (n < 0) n = 0;
    return 256 + ((size_t)n * 4 + 255) / 256 * 256;
}

int nglod_trace_min(const nglod_model_t* model, const float* ray_o, const float* ray_d, int64_t n, int32_t lod,
                    const nglod_trace_params_t* params, float* x, float* t, uint8_t* hit, float* min_x,
                    void* workspace, size_t workspace_bytes, void* stream) {
    ModelArgs m;
    int rc = make_model_args(model, lod, NGLOD_FP32, &m);
    if (rc) return rc;
    if (!params) return fail(NGLOD_ERR_ARG, "params is NULL");
    if (params->num_steps < 1) return fail(NGLOD_ERR_ARG, "num_steps must be >= 1");
    if (n < 0 || n > (int64_t)0x7fff0000) return fail(NGLOD_ERR_ARG, "n out of range");
    if (n == 0) return NGLOD_OK;
    if (!ray_o || !ray_d || !x || !t || !hit || !min_x) return fail(NGLOD_ERR_ARG, "NULL pointer");
    if (!workspace || workspace_bytes < nglod_trace_min_workspace_size(n) || ((uintptr_t)workspace & 3u))
        return fail(NGLOD_ERR_WORKSPACE, "workspace too small or not 4-byte aligned");
    TraceMinArgs a;
    a.ray_o = ray_o;
    a.ray_d = ray_d;
    a.n = n;
    a.num_steps = params->num_steps;
    a.min_dis = (float)params->min_dis;
    a.min_dis3 = (float)(params->min_dis * 3.0);
    a.step = (float)params->step_size;
    a.far_clip = (float)params->far_clip;
    a.x = x;
    a.t = t;
    a.hit = hit;
    a.min_x = min_x;
    a.kmax = (uint32_t*)workspace;
    a.kinfo = (uint32_t*)((char*)workspace + 256);
    const int dec = decoder_of(model, lod);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(trace_min_clear_kernel, dim3(1), dim3(64), 0, s, a.kmax);
    rc = check_launch("trace_min_clear_kernel");
    if (rc) return rc;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (m.pos_inv)
        hipLaunchKernelGGL(trace_min_kernel<true>, grid, dim3(256), 0, s, m, lod, dec, a);
    else
        hipLaunchKernelGGL(trace_min_kernel<false>, grid, dim3(256), 0, s, m, lod, dec, a);
    rc = check_launch("trace_min_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(trace_min_finish_kernel, grid, dim3(256), 0, s, a);
    return check_launch("trace_min_finish_kernel");
}

}  // extern "C"
