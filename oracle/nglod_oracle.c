/*
 * CPU oracle for the NGLOD neural-SDF query and sphere-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.
 * The product path (nglod_amd + libnglod_hip.so) never calls it.
 *
 * What it restates (reference files, read as text):
 *   orc_gather_*    FeatureVolume.forward / F.grid_sample   sdf-net/lib/models/OctreeSDF.py:46-57
 *   orc_decode_*    per-LOD decoder                          OctreeSDF.py:79-88
 *   orc_sdf_*       OctreeSDF.sdf(x, lod, return_lst)        OctreeSDF.py:94-155
 *   orc_aabb_*      sol_nglod.aabb                           sdf-net/lib/extensions/sol_nglod/sol_nglod_kernel.cu:77-188
 *   orc_trace_*     SphereTracer.forward                     sdf-net/lib/tracer/SphereTracer.py:41-132
 *   orc_normals_*   gradient(method='finitediff')            sdf-net/lib/diffutils.py:61-70
 *
 * Two precisions from one body (oracle_impl.h): _f32 follows the reference's
 * FP32 operation order with no contraction (build with -ffp-contract=off);
 * _f64 is the FP64 oracle.  Weights and grids are FP32 in both.
 *
 * Pinning (tests/test_oracle_*.py): the gather against torch's own
 * grid_sample (the library call the reference makes) in FP64 and FP32, plus
 * affine-field, grid-vertex and border known answers; the decoder against
 * torch.nn.functional.linear; aabb against closed-form box entries; the
 * tracer against a plane SDF whose hit depth and normal are known exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#define ORC_MAX_LODS 16
#define ORC_MAX_C 256

typedef struct {
    int32_t num_lods;
    int32_t fdim;
    int32_t hidden;
    int32_t pos_invariant;
    int32_t num_decoders;
    int32_t res[ORC_MAX_LODS];
    const float* grid[ORC_MAX_LODS];
    const float* w0[ORC_MAX_LODS];
    const float* b0[ORC_MAX_LODS];
    const float* w1[ORC_MAX_LODS];
    const float* b1[ORC_MAX_LODS];
} orc_model;

typedef struct {
    int32_t num_steps;
    double min_dis;
    double step_size;
    double far_clip;
} orc_trace_params;

#define REAL float
#define SFX f32
#define FMA fmaf
#define FABS fabsf
#define FMAX fmaxf
#include "oracle_impl.h"
#undef REAL
#undef SFX
#undef FMA
#undef FABS
#undef FMAX

#define REAL double
#define SFX f64
#define FMA fma
#define FABS fabs
#define FMAX fmax
#include "oracle_impl.h"
#undef REAL
#undef SFX
#undef FMA
#undef FABS
#undef FMAX

int orc_abi_version(void) { return 1; }
