// Calibration: the vector-L1 -> VGPR ceiling of the CUs for 16-byte loads.
//   nglod_vl1d_probe
// Every lane reads 16-byte words of a small buffer that stays in each CU's
// vector L1 (16 KB against its 32 KB), fully coalesced (a wave instruction
// reads 64 consecutive words, 1 KB), with no dependency between the loads;
// the XOR/sum of the words keeps the loads live.  Bytes read per call =
// blocks * 256 * iters * 16, the rate the gather kernels' vL1D-bound
// roofline is compared with.
#include "api_common.hpp"

namespace nglod {

__global__ __launch_bounds__(256) void vl1d_probe_kernel(const uint4* __restrict__ buf, uint32_t mask, int iters,
                                                         uint4* __restrict__ out) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t idx = gid & mask;
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll 8
    for (int k = 0; k < iters; ++k) {
        const uint4 v = buf[idx];
        acc.x ^= v.x;
        acc.y += v.y;
        acc.z ^= v.z;
        acc.w += v.w;
        idx = (idx + 64u) & mask;
    }
    out[gid] = acc;
}

}  // namespace nglod

using namespace nglod;

extern "C" int nglod_vl1d_probe(const void* buf, int64_t buf_bytes, int32_t iters, int32_t blocks, void* out,
                                int64_t out_bytes, void* stream) {
    if (!buf || !out || !aligned16(buf) || !aligned16(out)) return fail(NGLOD_ERR_ARG, "NULL or misaligned pointer");
    // a power-of-two number of 16-byte words, at least one wave's worth
    if (buf_bytes < 1024 || buf_bytes > (1 << 20) || (buf_bytes & (buf_bytes - 1)) != 0)
        return fail(NGLOD_ERR_ARG, "buf_bytes must be a power of two in [1 KB, 1 MB]");
    if (iters < 1 || blocks < 1 || blocks > (1 << 20)) return fail(NGLOD_ERR_ARG, "iters or blocks out of range");
    if (out_bytes < (int64_t)blocks * 256 * 16) return fail(NGLOD_ERR_CAPACITY, "out needs blocks * 4096 bytes");
    const uint32_t mask = (uint32_t)(buf_bytes / 16) - 1u;
    hipLaunchKernelGGL(vl1d_probe_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)buf, mask, iters, (uint4*)out);
    return check_launch("vl1d_probe_kernel");
}
