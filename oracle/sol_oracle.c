#define _DEFAULT_SOURCE
/*
 * CPU oracle for the sparse ("SOL") march.  TEST INFRASTRUCTURE ONLY.
 *
 *   orc_sol_trinkets   what index_trinket_kernel computes
 *                      (sol-renderer/include/solr/solr/sdf/index_trinket.cuh:31-98):
 *                      per octree point at levels 2..target, the 8 corner
 *                      indices (corner j at x + (j&4)>>2, y + (j&2)>>1,
 *                      z + (j&1)) into the concatenated corner list, and the
 *                      parent point (-1 at level 2)
 *   orc_sol_trace      SDF::sphereTrace (sol-renderer/SDF.cu:297-472) with
 *                      ray_aabb_kernel (solr/gfx/ray_aabb.cuh:42-192),
 *                      sparse_grid_sample_kernel (solr/sdf/sparse_grid_sample.cuh:31-109),
 *                      the addmm decoder (SDF.cu:420-421), step_kernel
 *                      (solr/sdf/step.cuh:31-86) and SDF::getNormal
 *                      (SDF.cu:218-295).  FP32 in the written order, no
 *                      contraction; explicit fmaf where the reference has it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

uint64_t orc_morton(uint32_t x, uint32_t y, uint32_t z);

typedef struct {
    int32_t target;          /* octree level of the march voxels (= lod + 2) */
    int32_t lod;             /* finest LOD; levels target..2 are summed */
    int32_t din;             /* decoder input width: 35 (xyz + 32) or 32 */
    const int16_t* points;   /* [P][4] */
    const int32_t* psum;     /* exclusive point sums per octree level */
    const uint32_t* tv;      /* [P][8] trinket corner indices */
    const int32_t* parent;   /* [P] */
    const int32_t* cc;       /* [Nc][3] corner coordinates */
    const float* cf;         /* [Nc][32] corner features */
    const float* w0;         /* [128][din] */
    const float* b0;         /* [128] */
    const float* w1;         /* [128] */
    const float* b1;         /* [1] */
} orc_sol;

static int cmp3(const int32_t* a, const int32_t* b) {
    for (int k = 0; k < 3; ++k) {
        if (a[k] != b[k]) return a[k] < b[k] ? -1 : 1;
    }
    return 0;
}

/* cc: per level l (2..target) a lexicographically sorted block of corners,
 * block l starting at cc_off[l - 2] with cc_cnt[l - 2] entries. */
void orc_sol_trinkets(const int16_t* points, const int32_t* pyr_cnt, const int32_t* psum, int target, const int32_t* cc,
                      const int64_t* cc_off, const int64_t* cc_cnt, uint32_t* tv, int32_t* parent) {
    for (int l = 2; l <= target; ++l) {
        for (int64_t k = 0; k < pyr_cnt[l]; ++k) {
            const int64_t g = psum[l] + k;
            const int16_t* p = points + 4 * g;
            for (int j = 0; j < 8; ++j) {
                const int32_t c[3] = {p[0] + ((j & 4) >> 2), p[1] + ((j & 2) >> 1), p[2] + (j & 1)};
                int64_t lo = 0, hi = cc_cnt[l - 2] - 1, found = -1;
                while (lo <= hi) {
                    const int64_t mid = (lo + hi) / 2;
                    const int r = cmp3(cc + 3 * (cc_off[l - 2] + mid), c);
                    if (r == 0) { found = mid; break; }
                    if (r < 0) lo = mid + 1; else hi = mid - 1;
                }
                tv[8 * g + j] = found < 0 ? 0xFFFFFFFFu : (uint32_t)(cc_off[l - 2] + found);
            }
            parent[g] = -1;
            if (l > 2) {
                /* points of a level are in Morton order: binary search the parent */
                const uint64_t key = orc_morton((uint32_t)(p[0] / 2), (uint32_t)(p[1] / 2), (uint32_t)(p[2] / 2));
                int64_t lo = 0, hi = pyr_cnt[l - 1] - 1;
                while (lo <= hi) {
                    const int64_t mid = (lo + hi) / 2;
                    const int16_t* pp = points + 4 * (psum[l - 1] + mid);
                    const uint64_t mk = orc_morton((uint32_t)pp[0], (uint32_t)pp[1], (uint32_t)pp[2]);
                    if (mk == key) { parent[g] = (int32_t)(psum[l - 1] + mid); break; }
                    if (mk < key) lo = mid + 1; else hi = mid - 1;
                }
            }
        }
    }
}

/* Slab test against a cube of half-size r around vc (ray_aabb.cuh:42-89). */
static float ray_aabb(const float* q, const float* dir, const float* invdir, const float* sgn, const float* vc, float r) {
    const float o[3] = {q[0] - vc[0], q[1] - vc[1], q[2] - vc[2]};
    const float cmax = fmaxf(fmaxf(fabsf(o[0]), fabsf(o[1])), fabsf(o[2]));
    float winding = cmax < r ? -1.0f : 1.0f;
    winding *= r;
    if (winding < 0) return winding;
    const float d0 = fmaf(winding, sgn[0], -o[0]) * invdir[0];
    const float d1 = fmaf(winding, sgn[1], -o[1]) * invdir[1];
    const float d2 = fmaf(winding, sgn[2], -o[2]) * invdir[2];
    const float ltxy = fmaf(dir[1], d0, o[1]), ltxz = fmaf(dir[2], d0, o[2]);
    const float ltyx = fmaf(dir[0], d1, o[0]), ltyz = fmaf(dir[2], d1, o[2]);
    const float ltzx = fmaf(dir[0], d2, o[0]), ltzy = fmaf(dir[1], d2, o[1]);
    const int t0 = (d0 >= 0.0f) && (fabsf(ltxy) < r) && (fabsf(ltxz) < r);
    const int t1 = (d1 >= 0.0f) && (fabsf(ltyx) < r) && (fabsf(ltyz) < r);
    const int t2 = (d2 >= 0.0f) && (fabsf(ltzx) < r) && (fabsf(ltzy) < r);
    float sx = 0, sy = 0, sz = 0;
    if (t0) sx = sgn[0];
    else if (t1) sy = sgn[1];
    else if (t2) sz = sgn[2];
    float d = 0.0f;
    if (sx != 0.0f) d = d0;
    else if (sy != 0.0f) d = d1;
    else if (sz != 0.0f) d = d2;
    if (d != 0.0f) return d;
    return 0.0f;
}

/* Sum over levels lod..0 of the trilinear samples at x, starting at the
 * target-level voxel cidx (global point index), following parents. */
static void sol_sample(const orc_sol* s, const float* x, int64_t cidx, float* f) {
    for (int k = 0; k < 32; ++k) f[k] = 0.0f;
    const float nx = fmaf(x[0], 0.5f, 0.5f), ny = fmaf(x[1], 0.5f, 0.5f), nz = fmaf(x[2], 0.5f, 0.5f);
    for (int lv = s->lod; lv >= 0; --lv) {
        if (lv != s->lod) cidx = s->parent[cidx];
        const float res = (float)(4u << lv);
        const float iz = nx * res, iy = ny * res, ix = nz * res;
        const uint32_t* v = s->tv + 8 * cidx;
        const uint32_t v000 = v[0], v100 = v[1], v010 = v[2], v110 = v[3], v001 = v[4], v101 = v[5], v011 = v[6], v111 = v[7];
        const int32_t* cc = s->cc;
#define CCX(q) ((float)cc[3 * (q)])
#define CCY(q) ((float)cc[3 * (q) + 1])
#define CCZ(q) ((float)cc[3 * (q) + 2])
        for (int k = 0; k < 32; ++k) {
            f[k] = f[k] + s->cf[32 * v000 + k] * (CCZ(v111) - ix) * (CCY(v111) - iy) * (CCX(v111) - iz);
            f[k] = f[k] + s->cf[32 * v001 + k] * (CCZ(v110) - ix) * (CCY(v110) - iy) * (iz - CCX(v110));
            f[k] = f[k] + s->cf[32 * v010 + k] * (CCZ(v101) - ix) * (iy - CCY(v101)) * (CCX(v101) - iz);
            f[k] = f[k] + s->cf[32 * v011 + k] * (CCZ(v100) - ix) * (iy - CCY(v100)) * (iz - CCX(v100));
            f[k] = f[k] + s->cf[32 * v100 + k] * (ix - CCZ(v011)) * (CCY(v011) - iy) * (CCX(v011) - iz);
            f[k] = f[k] + s->cf[32 * v101 + k] * (ix - CCZ(v010)) * (CCY(v010) - iy) * (iz - CCX(v010));
            f[k] = f[k] + s->cf[32 * v110 + k] * (ix - CCZ(v001)) * (iy - CCY(v001)) * (CCX(v001) - iz);
            f[k] = f[k] + s->cf[32 * v111 + k] * (ix - CCZ(v000)) * (iy - CCY(v000)) * (iz - CCX(v000));
        }
#undef CCX
#undef CCY
#undef CCZ
    }
}

static float sol_decode(const orc_sol* s, const float* x, const float* f) {
    const int off = s->din - 32;
    float out = 0.0f;
    for (int j = 0; j < 128; ++j) {
        float acc = 0.0f;
        const float* row = s->w0 + (int64_t)j * s->din;
        for (int k = 0; k < off; ++k) acc += row[k] * x[k];
        for (int k = 0; k < 32; ++k) acc += row[off + k] * f[k];
        float h = acc + s->b0[j];
        h = h < 0.0f ? 0.0f : h;
        out += s->w1[j] * h;
    }
    return out + s->b1[0];
}

static float sol_eval(const orc_sol* s, const float* x, int64_t pidx_local) {
    float f[32];
    sol_sample(s, x, pidx_local + s->psum[s->target], f);
    return sol_decode(s, x, f);
}

/* Walk the ray's nuggets from the start; the first voxel containing q or
 * ahead of it sets pidx and advances t (ray_aabb_kernel, ray_aabb.cuh:104-192). */
static int voxel_walk(const orc_sol* s, const int32_t* nug, int64_t beg, int64_t end, const float* o, const float* dir,
                      const float* inv, const float* sgn, float r, float* x, float* t, int64_t* pidx) {
    for (int64_t i = beg; i < end; ++i) {
        const int64_t pl = nug[2 * i + 1];
        const int16_t* p = s->points + 4 * (s->psum[s->target] + pl);
        const float vc[3] = {fmaf(r, fmaf(2.0f, (float)p[0], 1.0f), -1.0f), fmaf(r, fmaf(2.0f, (float)p[1], 1.0f), -1.0f),
                             fmaf(r, fmaf(2.0f, (float)p[2], 1.0f), -1.0f)};
        const float d = ray_aabb(x, dir, inv, sgn, vc, r);
        if (d != 0.0f) {
            *pidx = pl;
            if (d > 0.0f) {
                *t += d;
                for (int k = 0; k < 3; ++k) x[k] = fmaf(dir[k], *t, o[k]);
            }
            return 1;
        }
    }
    *t = 100.0f;
    for (int k = 0; k < 3; ++k) x[k] = fmaf(dir[k], *t, o[k]);
    return 0;
}

void orc_sol_trace(const orc_sol* s, const float* ray_o, const float* ray_d, int64_t n, const int32_t* offs,
                   const int32_t* nug, float* xo, float* to, uint8_t* hito, float* nrmo, int32_t* iters) {
    const float r = (float)(1.0 / (double)(1 << s->target));
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float* o = ray_o + 3 * i;
        const float* dir = ray_d + 3 * i;
        const float inv[3] = {1.0f / dir[0], 1.0f / dir[1], 1.0f / dir[2]};
        const float sgn[3] = {signbit(dir[0]) ? 1.0f : -1.0f, signbit(dir[1]) ? 1.0f : -1.0f, signbit(dir[2]) ? 1.0f : -1.0f};
        float x[3] = {o[0], o[1], o[2]};
        float t = 0.0f, dprev = 0.0f;
        int cond = 0, hit = 0, it = 0;
        int64_t pidx = -1;
        const int64_t beg = offs[i], end = offs[i + 1];
        if (end > beg) cond = voxel_walk(s, nug, beg, end, o, dir, inv, sgn, r, x, &t, &pidx);
        for (it = 0; it < 50 && cond; ++it) {
            const float d = sol_eval(s, x, pidx);
            t = t + d;
            hit = ((double)fabsf(d) < 0.0003) || (((double)fabsf(d + dprev) * 0.5) < 0.0015);
            cond = ((double)t < 5.0) && !hit;
            for (int k = 0; k < 3; ++k) x[k] = fmaf(dir[k], t, o[k]);
            dprev = d;
            if (cond) cond = voxel_walk(s, nug, beg, end, o, dir, inv, sgn, r, x, &t, &pidx);
        }
        float nm[3] = {1.0f, 1.0f, 1.0f};
        if (hit) {
            for (int a = 0; a < 3; ++a) {
                float xf[3] = {x[0] + 0.0f, x[1] + 0.0f, x[2] + 0.0f}, xb[3] = {x[0] - 0.0f, x[1] - 0.0f, x[2] - 0.0f};
                xf[a] = x[a] + 0.001f;
                xb[a] = x[a] - 0.001f;
                nm[a] = sol_eval(s, xf, pidx) - sol_eval(s, xb, pidx);
            }
            const float len = sqrtf(nm[0] * nm[0] + nm[1] * nm[1] + nm[2] * nm[2]);
            nm[0] = nm[0] / len;
            nm[1] = nm[1] / len;
            nm[2] = nm[2] / len;
        }
        for (int k = 0; k < 3; ++k) {
            xo[3 * i + k] = x[k];
            nrmo[3 * i + k] = nm[k];
        }
        to[i] = t;
        hito[i] = (uint8_t)hit;
        if (iters) iters[i] = it;
    }
}

/* Single SDF evaluation (for tests). */
void orc_sol_eval(const orc_sol* s, const float* x, const int64_t* pidx_local, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = sol_eval(s, x + 3 * i, pidx_local[i]);
}
