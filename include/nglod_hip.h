/*
 * libnglod_hip.so — C ABI of the MI355X (gfx950) NGLOD neural-SDF query and
 * sphere-tracing path.
 *
 * Every entry point takes device pointers the caller allocated, element
 * counts, and a HIP stream handle (hipStream_t passed as void*; NULL = the
 * null stream).  Calls are asynchronous on that stream unless stated.  The
 * library never frees caller memory.  Return value: NGLOD_OK (0) or a
 * negative status; no C++ exception crosses this boundary.
 *
 * Each declaration names the reference interface it replaces (file:line,
 * paths relative to the nv-tlabs/nglod repository root).
 */
#ifndef NGLOD_HIP_H
#define NGLOD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NGLOD_ABI_VERSION 1
#define NGLOD_MAX_LODS 8

/* status codes */
#define NGLOD_OK 0
#define NGLOD_ERR_ARG (-1)         /* null pointer, negative size, bad lod */
#define NGLOD_ERR_UNSUPPORTED (-2) /* shape the kernels are not built for */
#define NGLOD_ERR_CAPACITY (-4)    /* output buffer too small; see *needed */
#define NGLOD_ERR_HIP (-5)         /* HIP runtime error (launch, device) */
#define NGLOD_ERR_WORKSPACE (-6)   /* workspace smaller than *_workspace_size */

/* arithmetic forms */
#define NGLOD_FP32 0  /* FP32 grids, FP32 decoder */
#define NGLOD_MIXED 1 /* FP16 cumulative grids, BF16 MFMA decoder with FP32 accumulation */

/* Dense multi-LOD model (OctreeSDF, sdf-net/lib/models/OctreeSDF.py:59-88).
 * Device pointers.  grid[i] is corner-major: element
 * ((z*(r+1) + y)*(r+1) + x)*feature_dim + c  is  fm[0, c, z, y, x] of the
 * reference FeatureVolume (OctreeSDF.py:43) — i.e. torch channels_last_3d.
 * Decoder i: w0[i] [hidden][din], b0[i] [hidden], w1[i] [hidden], b1[i] [1],
 * din = feature_dim + (pos_invariant ? 0 : 3), xyz first (OctreeSDF.py:114-115).
 * grid_cum[i] / grid_h[i] / w0_bf16[i] are only read by NGLOD_MIXED and are
 * written by nglod_pack_mixed.  grid_cum[i] (same layout and size as grid[i])
 * holds, at each corner of grid i, the sum over levels l <= i of grid l's
 * trilinear sample at that corner's position, so that one trilinear sample of
 * grid_cum[i] is the level sum OctreeSDF.sdf accumulates up to lod i
 * (OctreeSDF.py:102-110); grid_h[i] is its FP16 rounding. */
typedef struct nglod_model {
    int32_t num_lods;      /* feature volumes in use, 1..NGLOD_MAX_LODS */
    int32_t feature_dim;   /* C; kernels are built for 32 */
    int32_t hidden_dim;    /* H; kernels are built for 128 */
    int32_t pos_invariant; /* 1: decoder input is the feature only */
    int32_t num_decoders;  /* 1 (joint_decoder) or num_lods */
    int32_t res[NGLOD_MAX_LODS];
    const float* grid[NGLOD_MAX_LODS];
    const float* w0[NGLOD_MAX_LODS];
    const float* b0[NGLOD_MAX_LODS];
    const float* w1[NGLOD_MAX_LODS];
    const float* b1[NGLOD_MAX_LODS];
    const uint16_t* grid_h[NGLOD_MAX_LODS];  /* FP16 rounding of grid_cum[i] */
    const uint16_t* w0_bf16[NGLOD_MAX_LODS]; /* BF16 W0 in MFMA fragment order */
    const float* grid_cum[NGLOD_MAX_LODS];   /* cumulative grid of level i, FP32 */
} nglod_model_t;

/* Sphere-tracer parameters (BaseTracer.__init__, sdf-net/lib/tracer/BaseTracer.py:29-45;
 * defaults sdf-net/lib/options.py:209-221). */
typedef struct nglod_trace_params {
    int32_t num_steps;  /* >= 1 */
    double min_dis;     /* hit threshold; oscillation threshold is 3*min_dis */
    double step_size;   /* scales every in-loop SDF step */
    double far_clip;    /* camera_clamp[1] */
    int32_t normals;    /* 1: central-difference normals at hit points */
    int32_t refine_steps; /* NGLOD_MIXED only (1..4): when the mixed-precision
                             march stops on a surface test, its last
                             refine_steps evaluations are re-evaluated in FP32
                             and the march goes on in FP32 to its FP32 stop */
    double return_dist;   /* NGLOD_MIXED only: if > 0, an FP32 step that grows
                             past this (the ray leaves the surface: a near
                             miss) returns the ray to mixed precision, at most
                             twice; 0: the FP32 pass keeps every ray */
    double near_dist;     /* NGLOD_MIXED only: of the last refine_steps
                             evaluations, the stopping one and every earlier
                             one back to the first step of at least near_dist
                             (a step taken far from the surface, exact enough
                             in mixed precision) are re-evaluated; 0 redoes
                             all refine_steps */
    float* pass_ms;       /* optional host float[4] (NULL: off).  When set the
                             call records HIP events between its launches,
                             waits for the last one (the call becomes
                             synchronous) and writes milliseconds per pass:
                             [0] the march kernel (FP32 form) or the mixed
                             march, [1] the FP32 continuation (mixed form;
                             it also takes the normals of the form's hits),
                             [2] 0 (no separate normals pass), [3] counter
                             resets.  For measurement, not for serving. */
} nglod_trace_params_t;

int nglod_abi_version(void);

/* Last HIP error string recorded by this thread (static storage). */
const char* nglod_last_error(void);

/* Ray vs [-1,1]^3 box entry.  Replaces sol_nglod.aabb
 * (sdf-net/lib/extensions/sol_nglod/sol_nglod_kernel.cu:161-188, bound at :190-192,
 * called from sdf-net/lib/tracer/SphereTracer.py:53).
 * ray_o, ray_d: [n][3] fp32.  Writes x [n][3] (entry point, or ray_o when no
 * entry), t [n] (0 when no entry), hit [n] (0/1).  Bit-exact with the reference
 * kernel's arithmetic (double reciprocal, explicit fmaf). */
int nglod_aabb(const float* ray_o, const float* ray_d, int64_t n,
               float* x, float* t, uint8_t* hit, void* stream);

/* One feature volume sampled trilinearly: FeatureVolume.forward
 * (sdf-net/lib/models/OctreeSDF.py:46-57; F.grid_sample with
 * align_corners=True, padding_mode='border').  grid: corner-major
 * [(res+1)^3][feature_dim] fp32; x [n][3]; out [n][feature_dim]. */
int nglod_grid_sample(const float* grid, int32_t res, int32_t feature_dim, const float* x, int64_t n,
                      float* out, void* stream);

/* Fused multi-LOD gather + decoder: out[i] = OctreeSDF.sdf(x[i], lod)
 * (sdf-net/lib/models/OctreeSDF.py:94-155; net(x) via BaseLOD.forward,
 * sdf-net/lib/models/BaseLOD.py:37-41).  Features never leave registers.
 * x [n][3] fp32, out [n] fp32.  form = NGLOD_FP32 or NGLOD_MIXED. */
int nglod_sdf_fwd(const nglod_model_t* model, const float* x, int64_t n, int32_t lod,
                  int32_t form, float* out, void* stream);

/* Workspace bytes nglod_sdf_fwd_ordered needs for n points. */
size_t nglod_sdf_fwd_ordered_workspace_size(int64_t n);

/* nglod_sdf_fwd (the same reference interface, OctreeSDF.sdf) for large,
 * spatially scattered query sets: the points are first sorted by the Morton
 * code of their position (10 bits per axis over [-1, 1]) and evaluated in that
 * order, so the lanes of a wave gather neighbouring grid corners; out is in
 * input order and bit-identical to nglod_sdf_fwd's (each point's arithmetic
 * does not depend on the other points of its wave).  n < 2^31.  workspace:
 * device memory of nglod_sdf_fwd_ordered_workspace_size(n) bytes, 16-byte
 * aligned, owned by this call while it runs. */
int nglod_sdf_fwd_ordered(const nglod_model_t* model, const float* x, int64_t n, int32_t lod,
                          int32_t form, float* out, void* workspace, size_t workspace_bytes, void* stream);

/* Every level's prediction, OctreeSDF.sdf(x, return_lst=True)[0..lod]
 * (OctreeSDF.py:148-153).  out is [lod+1][n] fp32.  FP32 form only. */
int nglod_sdf_fwd_all(const nglod_model_t* model, const float* x, int64_t n, int32_t lod,
                      float* out, void* stream);

/* Workspace bytes nglod_trace_dense needs for n rays in `form` (256 bytes of
 * claim counters; for NGLOD_MIXED also the ray-indexed hand-off and hand-back
 * records, 48 B each, and their flags, 1 B each, per ray). */
size_t nglod_trace_dense_workspace_size(int64_t n, int32_t form);

/* Persistent sphere tracer with SphereTracer semantics
 * (sdf-net/lib/tracer/SphereTracer.py:41-132 incl. box entry :53, FD normals
 * :128-130 via sdf-net/lib/diffutils.py:61-70).  One kernel: box entry, march,
 * retirement and refill of finished rays, normals.  Outputs x [n][3] (the
 * position of the last evaluation, which lags t by one step as in the
 * reference), t [n], hit [n], normal [n][3] (zeros for misses).  Per-ray
 * arithmetic does not depend on which other rays are in the launch.
 * NGLOD_MIXED: a mixed-precision persistent march, then an FP32 persistent
 * pass over the rays handed off at their surface stop (see refine_steps),
 * then FP32 normals of the hits.
 * evals_out (device, may be NULL): uint64[2], SDF evaluations in FP32 form
 * ([0]) and in mixed form ([1]), accumulated (the caller zeroes it).
 * ray_steps (device, may be NULL): per ray, the in-loop SDF evaluations
 * (the reference's loop trips in which the ray was active), int32 [n].
 * Environment: NGLOD_TRACE_SOLO_LANES (0..64, default 4): in the FP32
 * passes, a wave with no rays left to claim and at most this many busy lanes
 * evaluates their queries one at a time with the whole wave (the mixed FP32
 * pass: spreads their gather over the idle lanes, at most 4 lanes); the mixed
 * pass 1 spreads its gather at up to 8 busy lanes unless the value is 0.  Same
 * bits for any value; 0 turns it off. */
int nglod_trace_dense(const nglod_model_t* model, const float* ray_o, const float* ray_d,
                      int64_t n, int32_t lod, const nglod_trace_params_t* params, int32_t form,
                      float* x, float* t, uint8_t* hit, float* normal,
                      void* workspace, size_t workspace_bytes, uint64_t* evals_out,
                      int32_t* ray_steps, void* stream);

/* Most rays one nglod_trace_dense_noalloc call takes in NGLOD_MIXED (2^20: a
 * 1280x720 frame; the FP32 form has no such limit). */
#define NGLOD_NOALLOC_MAX_MIXED_RAYS (1 << 20)

/* No counterpart in the reference.  For callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): nglod_trace_dense without a workspace argument; the same
 * kernels and the same outputs, bit for bit.  Its scratch (claim counters,
 * and for NGLOD_MIXED the hand-off records and flags) is a file-scope device
 * array of 3 KiB + 49 B * NGLOD_NOALLOC_MAX_MIXED_RAYS (about 49 MiB) that each
 * device holds from the library's load until the process ends.  The call
 * allocates nothing.  Limits: NGLOD_MIXED takes n <= NGLOD_NOALLOC_MAX_MIXED_RAYS
 * (larger n: NGLOD_ERR_CAPACITY; rays are independent, so a larger frame can
 * be traced in chunks with the same results) and needs params->return_dist = 0
 * (else NGLOD_ERR_UNSUPPORTED).  Every call on a device shares that device's
 * array and clears what must start at zero: calls that overlap on one device
 * (concurrent streams or threads) race. */
int nglod_trace_dense_noalloc(const nglod_model_t* model, const float* ray_o, const float* ray_d,
                              int64_t n, int32_t lod, const nglod_trace_params_t* params, int32_t form,
                              float* x, float* t, uint8_t* hit, float* normal, uint64_t* evals_out,
                              int32_t* ray_steps, void* stream);

/* ---- pooled trace (serving) -------------------------------------------- */

/* No counterpart in the reference (its nearest analogue is the batch loop of
 * Renderer.render, sdf-net/lib/renderer.py:118-123).  A device table of
 * models for nglod_trace_pool: nglod_model_table_fill validates the models
 * (packed for NGLOD_MIXED, all LODs) and copies one descriptor per model and
 * LOD to `table` (device, nglod_model_table_bytes(n) = n * NGLOD_MAX_LODS
 * descriptors).  Synchronous; refill after a model's buffers move. */
size_t nglod_model_table_bytes(int32_t n);
int nglod_model_table_fill(const nglod_model_t* models, int32_t n, void* table, size_t table_bytes, void* stream);

/* One request of a pooled trace: its rays are pool slots [offset, offset+count)
 * of the per-ray arrays; model indexes the model table. */
typedef struct nglod_pool_request {
    int32_t model;
    int32_t lod;
    int64_t offset;
    int64_t count;
} nglod_pool_request_t;

/* Workspace bytes of nglod_trace_pool for a pool of pool_rays ray slots
 * (hand-off records and flags indexed by slot: 49 B per slot).  Calls on
 * disjoint slots may share it, also while they overlap. */
size_t nglod_trace_pool_workspace_size(int64_t pool_rays);

/* Device bytes of one nglod_trace_pool call's claim counters (one 128-byte
 * line per counter: counters sharing a line slowed every claim); calls that
 * may overlap (on different streams) need their own. */
#define NGLOD_POOL_COUNTER_BYTES 4096

/* nglod_trace_dense (NGLOD_MIXED, SphereTracer semantics) for up to 64
 * requests at once, each with its own model (an index into model_table;
 * `models` is the same host array, for validation) and LOD: one launch of the
 * persistent mixed march over all their rays and one FP32 continuation pass
 * (which also takes the hits' normals).  ray_o, ray_d, x, t, hit, normal are the pool's per-slot
 * arrays ([pool_rays][3] or [pool_rays]); each request's outputs land in its
 * own slots and are bit-identical to a nglod_trace_dense call on those rays
 * alone (the FP32 continuation pass here runs without the single-request
 * kernels' one-query-at-a-time tail, which does not change results).  The
 * requests' slot ranges must not overlap, their models must agree on
 * pos_invariant, and params->return_dist must be 0.  workspace: the pool's
 * records (nglod_trace_pool_workspace_size(pool_rays)); counters:
 * NGLOD_POOL_COUNTER_BYTES of device memory for this call. */
int nglod_trace_pool(const nglod_model_t* models, const void* model_table, int32_t num_models,
                     const nglod_pool_request_t* requests, int32_t num_requests, const nglod_trace_params_t* params,
                     const float* ray_o, const float* ray_d, float* x, float* t, uint8_t* hit, float* normal,
                     int64_t pool_rays, void* workspace, size_t workspace_bytes, void* counters,
                     uint64_t* evals_out, void* stream);

/* Central-difference normals h = 1/192 then normalize(eps=1e-5)
 * (sdf-net/lib/diffutils.py:61-70, SphereTracer.py:128-130) at points x [n][3].
 * Computed in FP32 for either form. */
int nglod_normals_fd(const nglod_model_t* model, const float* x, int64_t n, int32_t lod,
                     int32_t form, float* normal, void* stream);

/* gradient methods of diffutils.gradient (sdf-net/lib/diffutils.py:29-84) */
#define NGLOD_GRAD_FINITEDIFF 0  /* central differences, h = 1/192 (:61-70) */
#define NGLOD_GRAD_TETRAHEDRON 1 /* four tetrahedron taps, h = 1/192 (:38-60) */
#define NGLOD_GRAD_AUTODIFF 2    /* d sdf / d x as torch autograd gives it (:32-37) */

/* Gradient of OctreeSDF.sdf(x, lod) with respect to x: diffutils.gradient(x,
 * net, method) (sdf-net/lib/diffutils.py:29-84).  x [n][3], grad [n][3], FP32
 * arithmetic in both forms (NGLOD_MIXED: on the FP32 cumulative grid of lod).
 * NGLOD_GRAD_AUTODIFF is the analytic derivative torch autograd computes:
 * grid_sample's backward with align_corners=True, padding_mode='border' (no
 * gradient through the border clamp: zero on and beyond the box faces along
 * that axis), ReLU's (units with pre-activation <= 0 pass nothing) and the
 * decoder's xyz columns for a position-dependent model.  mask: optional
 * uint8 [n]; where 0 the output is 0 and nothing is evaluated (the tracers'
 * normal[hit] = ..., SphereTracer.py:128-130).  normalize: 1 applies
 * F.normalize(p=2, eps=1e-5) as SphereTracer.forward does (SphereTracer.py:129);
 * 0 returns the raw gradient as diffutils.gradient and get_min do. */
int nglod_gradient(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, int32_t form,
                   int32_t method, const uint8_t* mask, int32_t normalize, float* grad, void* stream);

/* OctreeSDF.sdf(x, lod) and its autodiff gradient in one pass (sdf [n], may
 * be NULL; grad [n][3]): the position gradient of the reference's autograd
 * graph (OctreeSDF.py:94-155), used by the training backward when the query
 * positions require grad. */
int nglod_sdf_grad(const nglod_model_t* model, const float* x, int64_t n, int32_t lod, int32_t form, float* sdf,
                   float* grad, void* stream);

/* SphereTracer.get_min (sdf-net/lib/tracer/SphereTracer.py:134-218), FP32:
 * the march of every ray from its box entry (from its origin, t = 0, when it
 * has none), keeping the point min_x [n][3] that the reference's loop
 * associates with the smallest distance seen (the comparison runs after x has
 * moved by that distance).  x [n][3], t [n], hit [n] as the reference leaves
 * them (hit: |t| < far_clip at the last loop iteration it ran, inside the
 * box).  Normals are not computed here (nglod_gradient with mask = hit,
 * normalize = 0 gives get_min's).  The batch semantics of the reference's
 * loop (it ends when no ray marches) are kept: the result for a ray depends
 * on when the batch's last ray stops, as in the reference.  params: num_steps
 * (>= 1), min_dis, step_size, far_clip; the rest ignored.  workspace:
 * nglod_trace_min_workspace_size(n) bytes, 4-byte aligned. */
size_t nglod_trace_min_workspace_size(int64_t n);
int nglod_trace_min(const nglod_model_t* model, const float* ray_o, const float* ray_d, int64_t n, int32_t lod,
                    const nglod_trace_params_t* params, float* x, float* t, uint8_t* hit, float* min_x,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Builds what NGLOD_MIXED reads from an FP32 model: the cumulative grids
 * grid_cum[i] (float, (res[i]+1)^3*feature_dim) and their FP16 rounding
 * grid_h[i] (uint16, same count), and the feature columns of W0 in the BF16
 * MFMA fragment layout, w0_bf16[i] (hidden*feature_dim uint16; the xyz
 * columns of a position-dependent decoder stay FP32).  A cumulative corner
 * value is summed in FP64 (levels in order, each level's 8 corners in x, y, z
 * order) and rounded once to FP32; grid_h is the FP16 rounding of that FP32
 * value.  Every level's resolution must divide the resolutions of the levels
 * after it (the reference's 2^(i+base_lod) do), so that each cell of grid i
 * lies inside one cell of every coarser grid: then one trilinear sample of
 * grid_cum[i] equals the sum of the per-level samples in exact arithmetic;
 * otherwise NGLOD_ERR_UNSUPPORTED.  Replaces the half-precision export of
 * SOL_NGLOD (sdf-net/lib/models/SOL_NGLOD.py:33-46), kept on device. */
int nglod_pack_mixed(const nglod_model_t* model, void* stream);

/* ---- sparse octree (SPC) --------------------------------------------------
 * Setup calls below are synchronous (they read sizes back) and follow one
 * workspace protocol: if workspace_bytes is too small they return
 * NGLOD_ERR_WORKSPACE and set *workspace_needed; call again with at least that
 * much.  Output capacities too small return NGLOD_ERR_CAPACITY with the
 * required count already written to the count out-parameter. */

/* Level-`level` cells (1..10) that triangles [nf][3][3] overlap (closed-box
 * separating-axis test in FP64): sorted unique 30-bit Morton codes (x in the
 * high bit of each triple, SPC.h:65-80).  The benchmark octree's occupancy;
 * the reference samples it from traced surface points instead
 * (sdf-net/lib/renderutils.py:57-73). */
int nglod_spc_occupancy(const float* tris, int64_t nf, int32_t level, uint32_t* codes, int64_t cap,
                        int64_t* count, void* workspace, size_t workspace_bytes, size_t* workspace_needed,
                        void* stream);

/* Sorted unique leaf codes at `level` -> octree occupancy bytes, level 0
 * first (SPC3D.points_to_nodes, sdf-net/lib/spc3d.py:224-255).  level_counts
 * (host, level+1 entries): nodes per level. */
int nglod_spc_from_codes(const uint32_t* codes, int64_t n, int32_t level, uint8_t* octree, int64_t cap,
                         int64_t* osize, int32_t* level_counts, void* workspace, size_t workspace_bytes,
                         size_t* workspace_needed, void* stream);

/* Octree bytes -> points [P][4] int16 (x, y, z, 0) and the pyramid (host,
 * [2][L+2]: row 0 points per level, row 1 exclusive sums) as SPC::SetGeometry
 * (sol-renderer/include/spc/spc/SPC.cu:148-238).  prefix (device, osize+1):
 * exclusive sums of the bytes' popcounts, needed by nglod_spc_raytrace. */
int nglod_spc_points(const uint8_t* octree, int64_t osize, int16_t* points, int64_t pcap, int32_t* pyramid,
                     int32_t pyramid_cap, int32_t* level_out, int64_t* psize_out, uint32_t* prefix,
                     void* workspace, size_t workspace_bytes, size_t* workspace_needed, void* stream);

/* Ray-octree traversal to level `target`: nuggets [M][2] int32 (ray, point
 * index within level `target`), rays in ascending order, each ray's cells
 * front-to-back as the reference orders them.  Replaces spc_raytrace
 * (sol-renderer/include/spc/spc/spc_raytrace_cuda.cpp:141-199).  ray_offsets
 * (device, n+1): CSR offsets of each ray's nuggets.  Bit-identical list.
 * For target <= 9 this is nglod_spc_raytrace_async's single launch followed
 * by one read-back of the count; deeper targets run a per-ray depth-first
 * walk twice (count, then emit) with the count read back in between. */
int nglod_spc_raytrace(const uint8_t* octree, const uint32_t* prefix, const int16_t* points,
                       const int32_t* pyramid, int32_t level, const float* ray_o, const float* ray_d, int64_t n,
                       int32_t target, int32_t* nuggets, int64_t cap, int64_t* count, uint32_t* ray_offsets,
                       void* workspace, size_t workspace_bytes, size_t* workspace_needed, void* stream);

/* Workspace bytes nglod_spc_raytrace_async needs for n rays (a ticket
 * counter and 8 bytes per tile of 64 rays). */
size_t nglod_spc_raytrace_async_workspace_size(int64_t n);

/* The same traversal and nugget list as nglod_spc_raytrace (spc_raytrace,
 * sol-renderer/include/spc/spc/spc_raytrace_cuda.cpp:141-199) in one
 * launch with no host read-back, for target <= 9 (else
 * NGLOD_ERR_UNSUPPORTED).  Waves take tiles of 64 rays and walk the octree
 * level by level in LDS with ballot/prefix-sum compaction; tiles find their
 * output position by a decoupled look-back over the tiles' counts.  Writes
 * ray_offsets [n+1] (complete even when the list is cut) and status (device,
 * uint32[2]): [0] the nugget count, [1] flags: 1 = more than cap nuggets
 * (only the first cap written; call again with a larger buffer), 2 = a tile's
 * look-back gave up (an internal error; the outputs are not valid).
 * nuggets may be NULL with cap = 0 (count and offsets only). */
int nglod_spc_raytrace_async(const uint8_t* octree, const uint32_t* prefix, const int16_t* points,
                             const int32_t* pyramid, int32_t level, const float* ray_o, const float* ray_d, int64_t n,
                             int32_t target, int32_t* nuggets, int64_t cap, uint32_t* ray_offsets, uint32_t* status,
                             void* workspace, size_t workspace_bytes, void* stream);

/* info[i] = 1 where nugget i starts a new ray (d_MarkUniqueRays,
 * sol-renderer/sdfRenderer.cu:108-120). */
int nglod_mark_unique_rays(const int32_t* nuggets, int64_t m, int32_t* info, void* stream);

/* ---- training: backward of the fused query ------------------------------ */

/* Parameter gradients, device pointers laid out like nglod_model_t:
 * grid[i] corner-major [(res[i]+1)^3][feature_dim] (= fm.grad in
 * channels_last_3d), w0[i] [hidden][din], b0[i] [hidden], w1[i] [hidden],
 * b1[i] [1]. */
typedef struct nglod_model_grads {
    float* grid[NGLOD_MAX_LODS];
    float* w0[NGLOD_MAX_LODS];
    float* b0[NGLOD_MAX_LODS];
    float* w1[NGLOD_MAX_LODS];
    float* b1[NGLOD_MAX_LODS];
} nglod_model_grads_t;

/* Workspace bytes nglod_sdf_bwd needs for n queries up to `lod`. */
size_t nglod_sdf_bwd_workspace_size(const nglod_model_t* model, int64_t n, int32_t lod);

/* Backward of OctreeSDF.sdf for levels 0..lod (the graph autograd builds
 * through F.grid_sample and nn.Linear in Trainer.step_geometry,
 * sdf-net/lib/trainer.py:317-339).  grad_out [lod+1][n]: dL/dpred_i; only
 * levels whose bit is set in level_mask are read.  Writes (does not
 * accumulate) grads->grid[0..lod] and the decoders' gradients.  Interpolation
 * weights and decoder activations are recomputed, not stored by the forward.
 * Deterministic: fixed-order reductions, grid gradients by stable radix sort
 * of (corner, query) keys and an in-order segmented sum; no float atomics.
 * Position gradients are not computed.  n <= 2^28. */
int nglod_sdf_bwd(const nglod_model_t* model, const float* x, int64_t n, int32_t lod,
                  const float* grad_out, int32_t level_mask, const nglod_model_grads_t* grads,
                  void* workspace, size_t workspace_bytes, void* stream);

/* ---- training targets and samples ------------------------------------- */

/* Workspace bytes nglod_mesh2sdf_f64 needs for nf triangles (per-triangle
 * FP64 table). */
size_t nglod_mesh2sdf_workspace_size(int64_t nf);

/* Signed distance from points [n][3] to a closed triangle mesh [nf][3][3],
 * computed in FP64.  Replaces mesh2sdf.mesh2sdf_gpu
 * (sdf-net/lib/extensions/mesh2sdf_cuda/mesh2sdf_kernel.cu:895-927, bound at
 * :1009, called from sdf-net/lib/torchgp/compute_sdf.py:38): unsigned distance
 * by the edge/face rule of :423-445, negative iff all 13 stab directions
 * (:341-353) have a hit on both sides of the point (:479-529, :584-602).
 * dist [n] fp32; dist64 [n] fp64 (may be NULL). */
int nglod_mesh2sdf_f64(const float* points, int64_t n, const float* mesh, int64_t nf, float* dist,
                       double* dist64, void* workspace, size_t workspace_bytes, void* stream);

/* n points uniform in [-1,1]^3 (sdf-net/lib/torchgp/sample_uniform.py:25-31),
 * from a counter-based SplitMix64 stream: sample i depends on (seed, i) only. */
int nglod_sample_uniform(int64_t n, uint64_t seed, float* out, void* stream);

/* n area-weighted surface samples of mesh [nf][3][3] with barycentrics
 * u = sqrt(r1), v = r2 (sdf-net/lib/torchgp/sample_surface.py:27-52), plus
 * sigma * N(0,1) per axis when sigma > 0 (sample_near_surface.py:27-45).
 * cdf [nf] fp64: inclusive normalized cumulative face areas
 * (area_weighted_distribution.py:26-45).  face [n] (may be NULL): chosen face. */
int nglod_sample_surface(const float* mesh, const double* cdf, int64_t nf, int64_t n, uint64_t seed,
                         double sigma, float* out, int32_t* face, void* stream);

/* n directions uniform on the unit sphere: z = 1 - 2u, phi = 2*pi*v
 * (geoutils.sample_unif_sphere, sdf-net/lib/geoutils.py:31-40), same
 * counter-based stream as nglod_sample_uniform. */
int nglod_sample_directions(int64_t n, uint64_t seed, float* out, void* stream);

/* Camera rays of a look-at camera, width*height rays, ray index =
 * ix * height + iy (geoutils.look_at + normalized_grid,
 * sdf-net/lib/geoutils.py:140-206, jitter off; world up (0, 1, 0)).  eye, at:
 * host float[3]; ortho = 0 for the perspective camera.  ray_o, ray_d [N][3]. */
int nglod_camera_rays(const float* eye, const float* at, float fov_deg, int32_t width, int32_t height,
                      int32_t ortho, float* ray_o, float* ray_d, void* stream);

/* ---- sparse ("SOL") model and its march ---------------------------------- */

/* Sparse model: features live only on the corners of occupied octree cells.
 * Levels 2..target of the octree carry LODs 0..lod (target = lod + 2,
 * sdfRenderer.cu:220).  Replaces the state SDF::loadWeights and
 * SDF::initTrinkets build (sol-renderer/SDF.cu:65-216).  All pointers are
 * device pointers the caller owns. */
typedef struct nglod_sol {
    int32_t target;              /* octree level of the march cells (2..12) */
    int32_t lod;                 /* = target - 2; levels lod..0 are summed */
    int32_t din;                 /* decoder input: 35 (xyz first) or 32 */
    int32_t reserved;
    int64_t psum_target;         /* index of the first level-`target` point */
    const int16_t* points;       /* [P][4] octree points, all levels */
    const uint32_t* trinket_v;   /* [P][8] corner indices, 16-byte aligned */
    const int32_t* parent;       /* [P] parent point index, -1 at level 2 */
    const int32_t* cc;           /* [Nc][3] corner coordinates, per level
                                    blocks sorted by (x, y, z) */
    const float* cf;             /* [Nc][32] corner features */
    const float* w0;             /* [128][din] decoder `lod` */
    const float* b0;             /* [128] */
    const float* w1;             /* [128] */
    const float* b1;             /* [1] */
} nglod_sol_t;

/* Unique corners of the octree cells at levels 2..target, one block per
 * level sorted by (x, y, z) as torch.unique(dim=0) orders them
 * (geoutils.unique_corners, sdf-net/lib/geoutils.py:60-77).  pyramid: host
 * [2][level+2] as nglod_spc_points returns it.  counts (host, target-1
 * entries): corners per level; cc [cap][3].  Setup call (synchronous,
 * workspace protocol above); cap too small returns NGLOD_ERR_CAPACITY with
 * counts filled. */
int nglod_sol_corners(const int16_t* points, const int32_t* pyramid, int32_t level, int32_t target, int32_t* cc,
                      int64_t cap, int64_t* counts, void* workspace, size_t workspace_bytes,
                      size_t* workspace_needed, void* stream);

/* cf[c] = dense corner-major grid of LOD (level - 2) at cc[c]
 * (feat[0, :, z, y, x] in unique_corners, geoutils.py:76), FP32.  Requires
 * res[i] = 4 * 2^i. */
int nglod_sol_features(const nglod_model_t* model, int32_t target, const int32_t* cc, const int64_t* counts,
                       float* cf, void* stream);

/* Per point of levels 2..target: the 8 corner indices (corner j at
 * x + (j&4)>>2, y + (j&2)>>1, z + (j&1)) and the parent point
 * (index_trinket_kernel, solr/sdf/index_trinket.cuh:31-98; binary searches
 * instead of its O(n*m) scan).  trinket_v [P][8], parent [P]; entries of
 * other levels are left untouched.  Synchronous; workspace >= 256 bytes.
 * Returns NGLOD_ERR_ARG when a corner is not in cc or a parent is not in
 * the octree (the reference would read out of bounds). */
int nglod_sol_trinkets(const int16_t* points, const int32_t* pyramid, int32_t level, int32_t target,
                       const int32_t* cc, const int64_t* counts, uint32_t* trinket_v, int32_t* parent,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Sparse SDF at x [n][3] inside level-`target` cells pidx [n] (index within
 * the level): sparse_grid_sample_kernel (solr/sdf/sparse_grid_sample.cuh:31-109)
 * then the decoder (SDF.cu:420-421), fused.  out [n]. */
int nglod_sol_sdf(const nglod_sol_t* sol, const float* x, const int32_t* pidx, int64_t n, float* out,
                  void* stream);

/* Workspace bytes of nglod_trace_sparse (its claim counters). */
size_t nglod_trace_sparse_workspace_size(int64_t n);

/* Sparse sphere march with normals (SDF::sphereTrace + SDF::getNormal,
 * sol-renderer/SDF.cu:218-472, with ray_aabb_kernel and step_kernel): entry
 * into the first nugget cell holding or ahead of the point, at most 50
 * steps, hit when |d| < 3e-4 or |d + dprev|/2 < 1.5e-3, stop at t >= 5 or
 * when no cell is left (t = 100); normals by central differences (1e-3) in
 * the same cell, normalised; misses get normal (1, 1, 1).  ray_offsets [n+1]
 * and nuggets [M][2] as nglod_spc_raytrace returns them at `target`.
 * Outputs x [n][3], t [n], hit [n], normal [n][3], iters [n] (may be NULL).
 * evals_out (device, may be NULL): SDF evaluations are added to it. */
int nglod_trace_sparse(const nglod_sol_t* sol, const float* ray_o, const float* ray_d, int64_t n,
                       const int32_t* ray_offsets, const int32_t* nuggets, float* x, float* t, uint8_t* hit,
                       float* normal, int32_t* iters, void* workspace, size_t workspace_bytes,
                       uint64_t* evals_out, void* stream);

/* Sorted unique leaf codes at `level` of the cells holding points x [n][3]
 * (where mask[i] != 0; mask may be NULL): the occupancy voxel_sparsify
 * derives from traced surface points (sdf-net/lib/renderutils.py:57-73),
 * cell = floor((x + 1) / 2 * 2^level) clamped to the grid.  Setup call
 * (synchronous, workspace protocol above). */
int nglod_spc_codes_from_points(const float* x, const uint8_t* mask, int64_t n, int32_t level, uint32_t* codes,
                                int64_t cap, int64_t* count, void* workspace, size_t workspace_bytes,
                                size_t* workspace_needed, void* stream);

/* ---- post-trace shading (Renderer's shadow, AO and matcap options) ------ */

/* Ground plane and shadow rays (sdf-net/lib/renderer.py:131-160).  Per ray,
 * plane_t = (ray_o.y - ground_y) / (-ray_d.y); the plane is hit when
 * 0 < plane_t < 500 and plane_t < t.  A plane hit clears hit[i], sets
 * t = plane_t, x = o + d * plane_t, normal = (0, 1, 0) (x, t, hit, normal are
 * updated in place) and plane_hit[i] = 1.  Every ray gets a shadow ray:
 * shadow_o = x + 0.1 * normal, shadow_d = normalize((noise + light) - shadow_o)
 * (noise [n][3] may be NULL: zero; the reference draws it from N(0, 0.01)),
 * light_hit = shadow_d . normal > 0.  light: host float[3] (the reference's
 * light is (-1.5, 4.5, -1.5)).  The caller traces (shadow_o, shadow_d) with
 * nglod_trace_dense; shadow = that trace's hit & light_hit. */
int nglod_shadow_rays(const float* ray_o, const float* ray_d, int64_t n, float ground_y, const float* light,
                      const float* noise, float* x, float* t, uint8_t* hit, float* normal, uint8_t* plane_hit,
                      float* shadow_o, float* shadow_d, uint8_t* light_hit, void* stream);

/* Workspace bytes of nglod_ao (one claim counter; reset by every call). */
size_t nglod_ao_workspace_size(void);

/* Ambient occlusion (sdf-net/lib/renderer.py:184-209): per ray with hit[i]
 * (or plane_hit[i]; plane_hit may be NULL; the two sets are disjoint as
 * nglod_shadow_rays leaves them), 40 SDF taps at x + normal * d_k,
 * d_k = 0.025 * ((k + 1) / 41)^1.6; r = sdf (a plane hit: min(sdf, d_k));
 * acc += 3.5 * relu(d_k - r - 0.0015); ao = clamp(1 - acc, 0.1, 1)^2; rays
 * with neither flag get ao = 1.  The taps are FP32 in both forms (as the
 * normals); ao [n]. */
int nglod_ao(const nglod_model_t* model, const float* x, const float* normal, const uint8_t* hit,
             const uint8_t* plane_hit, int64_t n, int32_t lod, int32_t form, float* ao, void* workspace,
             size_t workspace_bytes, void* stream);

/* Shadow map of a width x height frame (ray index = ix * height + iy)
 * (sdf-net/lib/renderer.py:301-304): clamp(1 - (shadow_hit & light_hit) + 0.9,
 * 0, 1), then scipy.ndimage.gaussian_filter(sigma = 2) over the [width][height]
 * array (mode 'reflect', radius 8, FP64 accumulation, FP32 after each axis).
 * map, tmp: float [width * height]. */
int nglod_shadow_map(const uint8_t* shadow_hit, const uint8_t* light_hit, int32_t width, int32_t height, float* map,
                     float* tmp, void* stream);

/* Matcap shading (sdf-net/lib/renderer.py:279-307 with
 * geoutils.spherical_envmap :253-275 and matcap_sampler :237-250).  view [n][3]
 * (the rays' directions), mm: host float[3][3] (may be NULL: identity; the
 * view is rotated by mm^T).  texture [tw][th][channels] float: the matcap
 * image transposed to x-major, values 0..255, sampled bilinearly on
 * linspace(0, 1, tw) x linspace(0, 1, th) in FP64, then / 255.  Pixels with
 * hit[i] = 0 get rgb = 1 and normal = 1 (normal is updated in place).  rgb is
 * then multiplied by shadow_map [n] and by ao [n] (each may be NULL).
 * rgb [n][3]. */
int nglod_matcap_shade(const float* view, float* normal, const uint8_t* hit, int64_t n, const float* mm,
                       const float* texture, int32_t tw, int32_t th, int32_t channels, const float* shadow_map,
                       const float* ao, float* rgb, void* stream);

/* ---- calibration ------------------------------------------------------ */

/* No counterpart in the reference: a measurement aid for the roofline of the
 * gather kernels.  blocks x 256 lanes each read `iters` 16-byte words of buf
 * (buf_bytes: a power of two in [1 KB, 1 MB], small enough to stay in each
 * CU's vector L1), a wave instruction reading 1 KB of consecutive words, and
 * write one 16-byte word each to out (out_bytes >= blocks * 4096).  Bytes
 * read = blocks * 256 * iters * 16; timing the call gives the vector L1 ->
 * register ceiling of 16-byte loads. */
int nglod_vl1d_probe(const void* buf, int64_t buf_bytes, int32_t iters, int32_t blocks, void* out,
                     int64_t out_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NGLOD_HIP_H */
