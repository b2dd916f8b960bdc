// Batch query in the mixed form: launch_sdf_mixed (nglod_sdf_fwd and
// nglod_sdf_fwd_ordered with form = NGLOD_MIXED), one point per lane, the
// FP16 gather of the cumulative grid and the BF16 MFMA decoder of the trace.
// A file of its own: it is built without the VGPR form of MFMA that the
// trace kernels use, which measured slower for this kernel.
#include "api_common.hpp"
#include "mixed_device.hpp"

namespace nglod {

template <bool POSINV>
__global__ __launch_bounds__(256) void sdf_mixed_kernel(ModelArgs m, int dec, const float* __restrict__ x,
                                                        int64_t n, float* __restrict__ out,
                                                        const uint32_t* __restrict__ order) {
    __shared__ MixedLds lds;
    stage_mixed_lds(m, dec, &lds);
    __syncthreads();
    const uint32_t lane = lane_id();
    const float b1 = m.b1[dec][0];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < n;
    const int64_t i = valid ? (order ? (int64_t)order[t] : t) : 0;
    const float px = valid ? x[3 * i] : 0.0f;
    const float py = valid ? x[3 * i + 1] : 0.0f;
    const float pz = valid ? x[3 * i + 2] : 0.0f;
    if (__ballot(valid) == 0) return;  // whole wave past the end (no barrier follows)
    __half2 acc[kC / 2];
    gather_fp16_cum(m, valid, px, py, pz, acc);
    const float d = decode_mixed_wave<POSINV>(acc, &lds, b1, lane, valid, px, py, pz);
    if (valid) out[i] = d;
}

int launch_sdf_mixed(const ModelArgs& m, int lod, int dec, const float* x, int64_t n, float* out,
                     const uint32_t* order, hipStream_t s) {
    // a mixed-form ModelArgs holds one level, the cumulative grid of the LOD
    if (lod != 0 || m.num_lods != 1) return fail(NGLOD_ERR_ARG, "mixed form: one cumulative level (gather lod 0)");
    const int64_t blocks = (n + 255) / 256;
    if (m.pos_inv)
        hipLaunchKernelGGL(sdf_mixed_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, s, m, dec, x, n, out,
                           order);
    else
        hipLaunchKernelGGL(sdf_mixed_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, s, m, dec, x, n, out,
                           order);
    return check_launch("sdf_mixed_kernel");
}

}  // namespace nglod
