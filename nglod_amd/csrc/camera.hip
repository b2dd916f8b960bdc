// Pinhole / orthographic camera rays in one launch (geoutils.look_at and
// normalized_grid, sdf-net/lib/geoutils.py:140-206, without the jitter).
// Layout as the reference's 'ij' meshgrid: ray index = ix * height + iy,
// ix along the window's x (left to right), iy along y (top to bottom).
// The camera basis is formed on the host in FP64 and rounded once; each ray
// then follows the reference's FP32 expression order.
#include <math.h>

#include "api_common.hpp"

namespace nglod {

struct CameraArgs {
    float ox, oy, oz;     // camera origin
    float vx, vy, vz;     // view (unit)
    float rx, ry, rz;     // right (unit)
    float ux, uy, uz;     // up (unit)
    float tanh;           // tan(fov / 2)
    float aspect;         // width / height
    int32_t width, height;
    int32_t ortho;
};

// torch.linspace(start, end, steps)[i]: the first half counts up from
// start, the second half down from end.
__device__ __forceinline__ float linspace_at(float start, float end, int steps, int i) {
    if (steps == 1) return start;
    const float step = (end - start) / (float)(steps - 1);
    const int half = steps / 2;
    return i < half ? start + step * (float)i : end - step * (float)(steps - i - 1);
}

__global__ void camera_rays_kernel(CameraArgs c, float* __restrict__ ro, float* __restrict__ rd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)c.width * c.height;
    if (i >= n) return;
    const int ix = (int)(i / c.height);
    const int iy = (int)(i % c.height);
    const float cx = linspace_at(-1.0f, 1.0f, c.width, ix) * c.aspect;
    const float cy = linspace_at(1.0f, -1.0f, c.height, iy);
    if (c.ortho) {
        // ray_o = right*cx*tanh + up*cy*tanh + origin + view; ray_d = view
        ro[3 * i] = c.rx * cx * c.tanh + c.ux * cy * c.tanh + c.ox + c.vx;
        ro[3 * i + 1] = c.ry * cx * c.tanh + c.uy * cy * c.tanh + c.oy + c.vy;
        ro[3 * i + 2] = c.rz * cx * c.tanh + c.uz * cy * c.tanh + c.oz + c.vz;
        rd[3 * i] = c.vx;
        rd[3 * i + 1] = c.vy;
        rd[3 * i + 2] = c.vz;
        return;
    }
    const float px = c.rx * cx * c.tanh + c.ux * cy * c.tanh + c.ox + c.vx;
    const float py = c.ry * cx * c.tanh + c.uy * cy * c.tanh + c.oy + c.vy;
    const float pz = c.rz * cx * c.tanh + c.uz * cy * c.tanh + c.oz + c.vz;
    const float dx = px - c.ox, dy = py - c.oy, dz = pz - c.oz;
    const float len = fmaxf(sqrtf(dx * dx + dy * dy + dz * dz), 1e-12f);
    ro[3 * i] = c.ox;
    ro[3 * i + 1] = c.oy;
    ro[3 * i + 2] = c.oz;
    rd[3 * i] = dx / len;
    rd[3 * i + 1] = dy / len;
    rd[3 * i + 2] = dz / len;
}

static void normalize3(double* v) {
    const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double s = 1.0 / (l > 1e-12 ? l : 1e-12);
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
}

static void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace nglod

using namespace nglod;

extern "C" int nglod_camera_rays(const float* eye, const float* at, float fov_deg, int32_t width, int32_t height,
                                 int32_t ortho, float* ray_o, float* ray_d, void* stream) {
    if (!eye || !at) return fail(NGLOD_ERR_ARG, "NULL camera");
    if (width < 1 || height < 1 || (int64_t)width * height > (int64_t)1 << 30)
        return fail(NGLOD_ERR_ARG, "resolution out of range");
    if (!ray_o || !ray_d) return fail(NGLOD_ERR_ARG, "NULL output");
    double o[3] = {eye[0], eye[1], eye[2]};
    double v[3] = {(double)at[0] - o[0], (double)at[1] - o[1], (double)at[2] - o[2]};
    if (v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0) return fail(NGLOD_ERR_ARG, "eye equals target");
    normalize3(v);
    const double yaxis[3] = {0.0, 1.0, 0.0};
    double r[3], u[3];
    cross3(v, yaxis, r);
    normalize3(r);
    cross3(r, v, u);
    normalize3(u);
    CameraArgs c;
    c.ox = eye[0];
    c.oy = eye[1];
    c.oz = eye[2];
    c.vx = (float)v[0];
    c.vy = (float)v[1];
    c.vz = (float)v[2];
    c.rx = (float)r[0];
    c.ry = (float)r[1];
    c.rz = (float)r[2];
    c.ux = (float)u[0];
    c.uy = (float)u[1];
    c.uz = (float)u[2];
    c.tanh = (float)tan(fov_deg * M_PI / 360.0);
    c.aspect = (float)((double)width / (double)height);
    c.width = width;
    c.height = height;
    c.ortho = ortho ? 1 : 0;
    const int64_t n = (int64_t)width * height;
    hipLaunchKernelGGL(camera_rays_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, c,
                       ray_o, ray_d);
    return check_launch("camera_rays_kernel");
}
