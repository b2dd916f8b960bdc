// Host-side helpers shared by the C-ABI translation units of libnglod_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/nglod_hip.h"
#include "device_common.hpp"

namespace nglod {

void set_last_error(const char* msg);

// Returns NGLOD_ERR_HIP and records the message if the last launch failed.
inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
        set_last_error(buf);
        return NGLOD_ERR_HIP;
    }
    return NGLOD_OK;
}

inline int fail(int code, const char* msg) {
    set_last_error(msg);
    return code;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// Validates the model for `lod` / `form` and fills the kernel argument copy.
// NGLOD_FP32: every level 0..lod, gathered level by level.  NGLOD_MIXED: one
// level, the cumulative grid of `lod` (grid_cum[lod] FP32, grid_h[lod] FP16)
// presented as level 0 with grid lod's resolution; the kernels then gather
// with lod gather_lod(form, lod) = 0 and decode with decoder_of(model, lod).
inline int make_model_args(const nglod_model_t* m, int lod, int form, ModelArgs* a) {
    if (!m) return fail(NGLOD_ERR_ARG, "model is NULL");
    if (m->num_lods < 1 || m->num_lods > NGLOD_MAX_LODS) return fail(NGLOD_ERR_ARG, "num_lods out of range");
    if (lod < 0 || lod >= m->num_lods) return fail(NGLOD_ERR_ARG, "lod out of range");
    if (m->feature_dim != kC || m->hidden_dim != kH)
        return fail(NGLOD_ERR_UNSUPPORTED, "kernels are built for feature_dim 32, hidden_dim 128");
    if (m->num_decoders != 1 && m->num_decoders != m->num_lods)
        return fail(NGLOD_ERR_ARG, "num_decoders must be 1 or num_lods");
    if (form != NGLOD_FP32 && form != NGLOD_MIXED) return fail(NGLOD_ERR_ARG, "unknown form");
    memset(a, 0, sizeof(*a));
    a->pos_inv = m->pos_invariant ? 1 : 0;
    a->num_dec = m->num_decoders;
    for (int i = 0; i < m->num_lods; ++i)
        if (m->res[i] < 1 || m->res[i] > 1024) return fail(NGLOD_ERR_ARG, "grid resolution out of range");
    if (form == NGLOD_FP32) {
        a->num_lods = m->num_lods;
        for (int i = 0; i < m->num_lods; ++i) {
            a->res[i] = m->res[i];
            a->grid[i] = m->grid[i];
            if (i <= lod && (!m->grid[i] || !aligned16(m->grid[i])))
                return fail(NGLOD_ERR_ARG, "grid pointer NULL or not 16-byte aligned");
        }
    } else {
        if (!m->grid_cum[lod] || !aligned16(m->grid_cum[lod]) || !m->grid_h[lod] || !aligned16(m->grid_h[lod]))
            return fail(NGLOD_ERR_ARG, "cumulative grid (grid_cum / grid_h) NULL or not 16-byte aligned: "
                                       "nglod_pack_mixed builds them");
        a->num_lods = 1;
        a->res[0] = m->res[lod];
        a->grid[0] = m->grid_cum[lod];
        a->grid_h[0] = m->grid_h[lod];
    }
    for (int i = 0; i < m->num_decoders; ++i) {
        a->w0[i] = m->w0[i];
        a->b0[i] = m->b0[i];
        a->w1[i] = m->w1[i];
        a->b1[i] = m->b1[i];
        a->w0_bf16[i] = m->w0_bf16[i];
        if (!m->w0[i] || !m->b0[i] || !m->w1[i] || !m->b1[i]) return fail(NGLOD_ERR_ARG, "decoder pointer NULL");
    }
    if (form == NGLOD_MIXED) {
        const int dec = m->num_decoders == 1 ? 0 : lod;
        if (!m->w0_bf16[dec] || !aligned16(m->w0_bf16[dec])) return fail(NGLOD_ERR_ARG, "bf16 W0 not packed");
    }
    return NGLOD_OK;
}

// The lod the kernels gather to for arguments built by make_model_args.
inline int gather_lod(int form, int lod) { return form == NGLOD_MIXED ? 0 : lod; }

inline int decoder_of(const nglod_model_t* m, int lod) { return m->num_decoders == 1 ? 0 : lod; }

int device_cu_count();

// Optional host-side timing of a launch sequence: mark(slot) records an event
// and charges the time since the previous mark to out[slot].  Inactive (no
// events, no sync) when out is NULL; finish() synchronizes on the last event.
class PassTimer {
  public:
    static constexpr int kMax = 32;
    PassTimer(float* out, int nslots, hipStream_t s) : out_(out), nslots_(nslots), s_(s) {
        if (!out_) return;
        for (int i = 0; i < nslots_; ++i) out_[i] = 0.0f;
        for (int i = 0; i < kMax; ++i) ev_[i] = nullptr;
        mark(-1);
    }
    ~PassTimer() {
        for (int i = 0; i < n_; ++i) (void)hipEventDestroy(ev_[i]);
    }
    void mark(int slot) {
        if (!out_ || n_ >= kMax) return;
        if (hipEventCreate(&ev_[n_]) != hipSuccess) return;
        if (hipEventRecord(ev_[n_], s_) != hipSuccess) {
            (void)hipEventDestroy(ev_[n_]);
            return;
        }
        slot_[n_++] = slot;
    }
    int finish() {
        if (!out_ || n_ == 0) return NGLOD_OK;
        if (hipEventSynchronize(ev_[n_ - 1]) != hipSuccess) return fail(NGLOD_ERR_HIP, "timing sync failed");
        for (int i = 1; i < n_; ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev_[i - 1], ev_[i]) != hipSuccess) return fail(NGLOD_ERR_HIP, "timing failed");
            if (slot_[i] >= 0 && slot_[i] < nslots_) out_[slot_[i]] += ms;
        }
        return NGLOD_OK;
    }

  private:
    float* out_;
    int nslots_;
    hipStream_t s_;
    hipEvent_t ev_[kMax];
    int slot_[kMax];
    int n_ = 0;
};

// Slots of nglod_trace_params_t::pass_ms.  Slot 2 (a separate normals pass)
// stays 0: the mixed form's FP32 pass takes the normal taps.
enum : int { kPassMarch = 0, kPassFp32 = 1, kPassUnused = 2, kPassOther = 3, kPassSlots = 4 };

struct TraceArgs;
// Mixed-precision launchers (mixed.hip).
// Batch queries; order (optional) is the evaluation order, a permutation of [0, n).
int launch_sdf_fp32(const ModelArgs& m, int lod, int dec, const float* x, int64_t n, float* out,
                    const uint32_t* order, hipStream_t s);
int launch_sdf_mixed(const ModelArgs& m, int lod, int dec, const float* x, int64_t n, float* out,
                     const uint32_t* order, hipStream_t s);
int launch_trace_mixed(const TraceArgs& a, int64_t n, hipStream_t s, PassTimer& timer);

// Pooled trace (several requests, one launch per pass): the request table is
// a kernel argument, the models a device table of ModelArgs.
constexpr int kPoolMaxRequests = 64;
struct PoolReq {
    int32_t model;  // index into PoolArgs::models
    int32_t lod;
    int32_t dec;    // decoder of that lod
    int32_t pad;
    int64_t vbeg;   // first virtual id (requests in ascending vbeg)
    int64_t off;    // first pool slot
};
struct PoolArgs {
    const ModelArgs* models;
    int64_t total;  // virtual ids of the launch
    int32_t nreq;
    int32_t pad;
    PoolReq req[kPoolMaxRequests];
};
int launch_trace_pool(const TraceArgs& a, const PoolArgs& p, bool pos_inv, hipStream_t s, PassTimer& timer);

}  // namespace nglod
