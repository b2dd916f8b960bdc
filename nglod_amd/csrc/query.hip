// Batch SDF queries in spatial order: nglod_sdf_fwd_ordered.
//
// Uniformly scattered query points put unrelated grid cells in neighbouring
// lanes, so a wave's corner loads touch as many cache lines as it has lanes.
// Sorting the points by the Morton code of their position (10 bits per axis)
// first makes a wave's 64 points a compact cluster: its loads share lines in
// L1 and L2.  Keys are 30 bits, sorted with their indices by rocPRIM's radix
// sort; the batch kernels then evaluate point order[t] in thread t and write
// out[order[t]].
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "api_common.hpp"

namespace nglod {

// Spreads the low 10 bits of v to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t quantize10(float p) {
    const float q = (p + 1.0f) * 512.0f;  // [-1, 1] -> [0, 1024]
    return q <= 0.0f ? 0u : (q >= 1023.0f ? 1023u : (uint32_t)q);  // NaN -> 0
}

__global__ void morton_keys_kernel(const float* __restrict__ x, int64_t n, uint32_t* __restrict__ keys,
                                   uint32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t qx = quantize10(x[3 * i]), qy = quantize10(x[3 * i + 1]), qz = quantize10(x[3 * i + 2]);
    keys[i] = (spread3(qz) << 2) | (spread3(qy) << 1) | spread3(qx);
    idx[i] = (uint32_t)i;
}

static inline size_t up256(size_t b) { return (b + 255) / 256 * 256; }

static size_t sort_temp_bytes(int64_t n) {
    size_t t = 0;
    if (rocprim::radix_sort_pairs((void*)nullptr, t, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (uint32_t*)nullptr, (size_t)n, 0, 30, (hipStream_t)0) != hipSuccess)
        return 0;
    return t;
}

}  // namespace nglod

using namespace nglod;

extern "C" {

size_t nglod_sdf_fwd_ordered_workspace_size(int64_t n) {
    if (n < 1) n = 1;
    return 4 * up256((size_t)n * 4) + up256(sort_temp_bytes(n));
}

int nglod_sdf_fwd_ordered(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, int32_t form, float* out,
                          void* workspace, size_t workspace_bytes, void* stream) {
    ModelArgs m;
    int rc = make_model_args(model, lod, form, &m);
    if (rc) return rc;
    if (n < 0 || n > (int64_t)0x7fff0000) return fail(NGLOD_ERR_ARG, "n out of range");
    if (n == 0) return NGLOD_OK;
    if (!x || !out) return fail(NGLOD_ERR_ARG, "NULL pointer");
    const size_t seg = up256((size_t)n * 4);
    size_t tmp = sort_temp_bytes(n);
    if (tmp == 0) return fail(NGLOD_ERR_HIP, "sort size query failed");
    if (!workspace || workspace_bytes < 4 * seg + up256(tmp)) return fail(NGLOD_ERR_WORKSPACE, "workspace too small");
    if (!aligned16(workspace)) return fail(NGLOD_ERR_ARG, "workspace not 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    char* w = (char*)workspace;
    uint32_t* k_in = (uint32_t*)w;
    uint32_t* k_out = (uint32_t*)(w + seg);
    uint32_t* i_in = (uint32_t*)(w + 2 * seg);
    uint32_t* i_out = (uint32_t*)(w + 3 * seg);
    hipLaunchKernelGGL(morton_keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, k_in, i_in);
    rc = check_launch("morton_keys_kernel");
    if (rc) return rc;
    if (rocprim::radix_sort_pairs((void*)(w + 4 * seg), tmp, k_in, k_out, i_in, i_out, (size_t)n, 0, 30, s) !=
        hipSuccess)
        return fail(NGLOD_ERR_HIP, "sort failed");
    const int dec = decoder_of(model, lod);
    if (form == NGLOD_MIXED) return launch_sdf_mixed(m, 0, dec, x, n, out, i_out, s);
    return launch_sdf_fp32(m, lod, dec, x, n, out, i_out, s);
}

}  // extern "C"
