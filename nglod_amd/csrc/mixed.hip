// Mixed-precision form (FP16 grids, BF16 MFMA decoder): placeholder launchers.
#include "api_common.hpp"
#include "trace_common.hpp"

namespace nglod {

int launch_sdf_mixed(const ModelArgs&, int, int, const float*, int64_t, float*, hipStream_t) {
    return fail(NGLOD_ERR_UNSUPPORTED, "mixed form not built yet");
}
int launch_normals_mixed(const ModelArgs&, int, int, const float*, int64_t, float*, hipStream_t) {
    return fail(NGLOD_ERR_UNSUPPORTED, "mixed form not built yet");
}
int launch_trace_mixed(const TraceArgs&, int, hipStream_t) {
    return fail(NGLOD_ERR_UNSUPPORTED, "mixed form not built yet");
}

}  // namespace nglod

extern "C" int nglod_pack_mixed(const nglod_model_t*, void*) {
    return nglod::fail(NGLOD_ERR_UNSUPPORTED, "mixed form not built yet");
}
