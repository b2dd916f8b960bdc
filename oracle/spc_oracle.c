#define _DEFAULT_SOURCE
/*
 * CPU oracle for the sparse octree (SPC) path.  TEST INFRASTRUCTURE ONLY.
 *
 *   orc_morton / orc_point     ToMorton / ToPoint, x in the high bit of each
 *                              triple (sol-renderer/include/spc/spc/SPC.h:65-97,
 *                              sdf-net/lib/spc3d.py:24-40)
 *   orc_tri_box_overlap        Akenine-Moller 2001 separating-axis test (FP64),
 *                              the occupancy rule of the benchmark octree
 *   orc_occupancy              sorted unique Morton codes of the level-L cells
 *                              a triangle soup overlaps
 *   orc_octree_from_codes      SPC3D.points_to_nodes (spc3d.py:224-255):
 *                              sorted leaf codes -> occupancy bytes, level 0 first
 *   orc_octree_points          SPC::SetGeometry (SPC.cu:148-238): bytes ->
 *                              points (int16 x4) and pyramid [2][L+2]
 *   orc_spc_raytrace           spc_raytrace_cuda (spc_raytrace_cuda_kernel.cu:194-265),
 *                              level-synchronous as written: Decide (:108-138,
 *                              face test :85-105), inclusive scan, Subdivide with
 *                              the Order table (:39-47, :141-180) / Compactify
 *                              (:183-190).  FP32 arithmetic in the written order,
 *                              no contraction.
 *   orc_mark_unique_rays       d_MarkUniqueRays (sol-renderer/sdfRenderer.cu:108-120)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

uint64_t orc_morton(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t m = 0;
    for (int i = 0; i < 16; ++i) {
        m |= (uint64_t)((z >> i) & 1u) << (3 * i);
        m |= (uint64_t)((y >> i) & 1u) << (3 * i + 1);
        m |= (uint64_t)((x >> i) & 1u) << (3 * i + 2);
    }
    return m;
}

void orc_point(uint64_t m, int16_t* p) {
    uint32_t x = 0, y = 0, z = 0;
    for (int i = 0; i < 16; ++i) {
        z |= (uint32_t)((m >> (3 * i)) & 1u) << i;
        y |= (uint32_t)((m >> (3 * i + 1)) & 1u) << i;
        x |= (uint32_t)((m >> (3 * i + 2)) & 1u) << i;
    }
    p[0] = (int16_t)x;
    p[1] = (int16_t)y;
    p[2] = (int16_t)z;
    p[3] = 0;
}

/* ---- triangle / box overlap (separating axes) --------------------------- */
static void sub3(const double* a, const double* b, double* r) { r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2]; }
static void crs(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static double dt(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Does the triangle overlap the closed box (center c, half size h)? */
int orc_tri_box_overlap(const double* c, double h, const double* t0, const double* t1, const double* t2) {
    double v[3][3], e[3][3];
    sub3(t0, c, v[0]);
    sub3(t1, c, v[1]);
    sub3(t2, c, v[2]);
    sub3(v[1], v[0], e[0]);
    sub3(v[2], v[1], e[1]);
    sub3(v[0], v[2], e[2]);
    /* 9 axes: box axis x triangle edge */
    for (int i = 0; i < 3; ++i) {
        for (int a = 0; a < 3; ++a) {
            double ax[3] = {0, 0, 0}, u[3] = {0, 0, 0};
            u[a] = 1.0;
            crs(u, e[i], ax);
            const double p0 = dt(ax, v[0]), p1 = dt(ax, v[1]), p2 = dt(ax, v[2]);
            const double mn = fmin(p0, fmin(p1, p2)), mx = fmax(p0, fmax(p1, p2));
            const double r = h * (fabs(ax[0]) + fabs(ax[1]) + fabs(ax[2]));
            if (mn > r || mx < -r) return 0;
        }
    }
    /* 3 box face normals */
    for (int a = 0; a < 3; ++a) {
        const double mn = fmin(v[0][a], fmin(v[1][a], v[2][a])), mx = fmax(v[0][a], fmax(v[1][a], v[2][a]));
        if (mn > h || mx < -h) return 0;
    }
    /* triangle plane */
    double n[3];
    crs(e[0], e[1], n);
    const double d = dt(n, v[0]);
    const double r = h * (fabs(n[0]) + fabs(n[1]) + fabs(n[2]));
    if (d > r || d < -r) return 0;
    return 1;
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/* Occupied level-L cells of tris [nf][9] (float): sorted unique Morton codes.
 * Returns the count (written up to cap). */
int64_t orc_occupancy(const float* tris, int64_t nf, int level, uint64_t* out, int64_t cap) {
    const int R = 1 << level;
    const double h = 1.0 / R;
    int64_t cnt = 0, alloc = 1 << 16;
    uint64_t* buf = (uint64_t*)malloc(alloc * sizeof(uint64_t));
    for (int64_t f = 0; f < nf; ++f) {
        double t[3][3];
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) t[k][a] = tris[9 * f + 3 * k + a];
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            const double mn = fmin(t[0][a], fmin(t[1][a], t[2][a])), mx = fmax(t[0][a], fmax(t[1][a], t[2][a]));
            lo[a] = (int)floor((mn + 1.0) * 0.5 * R) - 1;
            hi[a] = (int)floor((mx + 1.0) * 0.5 * R) + 1;
            lo[a] = lo[a] < 0 ? 0 : lo[a];
            hi[a] = hi[a] > R - 1 ? R - 1 : hi[a];
        }
        for (int x = lo[0]; x <= hi[0]; ++x)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int z = lo[2]; z <= hi[2]; ++z) {
                    const double c[3] = {-1.0 + (2.0 * x + 1.0) * h, -1.0 + (2.0 * y + 1.0) * h, -1.0 + (2.0 * z + 1.0) * h};
                    if (!orc_tri_box_overlap(c, h, t[0], t[1], t[2])) continue;
                    if (cnt == alloc) {
                        alloc *= 2;
                        buf = (uint64_t*)realloc(buf, alloc * sizeof(uint64_t));
                    }
                    buf[cnt++] = orc_morton((uint32_t)x, (uint32_t)y, (uint32_t)z);
                }
    }
    qsort(buf, (size_t)cnt, sizeof(uint64_t), cmp_u64);
    int64_t u = 0;
    for (int64_t i = 0; i < cnt; ++i)
        if (i == 0 || buf[i] != buf[i - 1]) {
            if (u < cap) out[u] = buf[i];
            ++u;
        }
    free(buf);
    return u;
}

/* Leaf codes (sorted, unique) at `level` -> octree bytes, level 0 first.
 * pyramid_counts[0..level] = nodes per level.  Returns osize. */
int64_t orc_octree_from_codes(const uint64_t* codes, int64_t n, int level, uint8_t* bytes, int64_t cap,
                              int64_t* pyramid_counts) {
    uint8_t** lv_bytes = (uint8_t**)calloc((size_t)level, sizeof(uint8_t*));
    int64_t* lv_n = (int64_t*)calloc((size_t)level + 1, sizeof(int64_t));
    uint64_t* cur = (uint64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    memcpy(cur, codes, (size_t)n * sizeof(uint64_t));
    int64_t m = n;
    lv_n[level] = n;
    for (int l = level; l > 0; --l) {
        uint64_t* par = (uint64_t*)malloc((size_t)(m > 0 ? m : 1) * sizeof(uint64_t));
        uint8_t* by = (uint8_t*)malloc((size_t)(m > 0 ? m : 1));
        int64_t k = -1;
        for (int64_t t = 0; t < m; ++t) {
            const uint64_t p = cur[t] >> 3;
            if (t == 0 || p != (cur[t - 1] >> 3)) {
                ++k;
                par[k] = p;
                by[k] = 0;
            }
            by[k] |= (uint8_t)(1u << (cur[t] & 7u));
        }
        lv_bytes[l - 1] = by;
        lv_n[l - 1] = k + 1;
        free(cur);
        cur = par;
        m = k + 1;
    }
    free(cur);
    int64_t osize = 0;
    for (int l = 0; l < level; ++l) {
        for (int64_t i = 0; i < lv_n[l]; ++i) {
            if (osize < cap) bytes[osize] = lv_bytes[l][i];
            ++osize;
        }
        free(lv_bytes[l]);
    }
    for (int l = 0; l <= level; ++l) pyramid_counts[l] = lv_n[l];
    free(lv_bytes);
    free(lv_n);
    return osize;
}

/* Bytes -> points [P][4] and pyramid [2][L+2] (row 0 counts, row 1 exclusive
 * sums, last column 0 / total).  Returns L (levels below the root). */
int orc_octree_points(const uint8_t* bytes, int64_t osize, int16_t* points, int64_t pcap, int32_t* pyramid, int pmax) {
    int64_t psize = 1;
    for (int64_t i = 0; i < osize; ++i) psize += __builtin_popcount(bytes[i]);
    uint64_t* m = (uint64_t*)malloc((size_t)psize * sizeof(uint64_t));
    m[0] = 0;
    int64_t counts[64];
    int L = 0;
    counts[0] = 1;
    int64_t lbeg = 0, lend = 1, next = 1;
    while (lbeg < osize) {
        for (int64_t k = lbeg; k < lend && k < osize; ++k)
            for (int i = 0; i < 8; ++i)
                if (bytes[k] & (1u << i)) m[next++] = 8 * m[k] + (uint64_t)i;
        ++L;
        counts[L] = next - lend;
        lbeg = lend;
        lend = next;
    }
    for (int64_t i = 0; i < psize && i < pcap; ++i) orc_point(m[i], points + 4 * i);
    /* pyramid [2][L+2] */
    if (L + 2 <= pmax) {
        int64_t s = 0;
        for (int l = 0; l <= L; ++l) {
            pyramid[l] = (int32_t)counts[l];
            pyramid[(L + 2) + l] = (int32_t)s;
            s += counts[l];
        }
        pyramid[L + 1] = 0;
        pyramid[(L + 2) + L + 1] = (int32_t)s;
    }
    free(m);
    return L;
}

static const uint32_t ORDER[8][8] = {
    {0, 1, 2, 4, 3, 5, 6, 7}, {1, 0, 3, 5, 2, 4, 7, 6}, {2, 0, 3, 6, 1, 4, 7, 5}, {3, 1, 2, 7, 0, 5, 6, 4},
    {4, 0, 5, 6, 1, 2, 7, 3}, {5, 1, 4, 7, 0, 3, 6, 2}, {6, 2, 4, 7, 0, 3, 5, 1}, {7, 3, 5, 6, 1, 2, 4, 0}};

static int face_eval(uint16_t i, uint16_t j, float a, float b, float c) {
    float r[4];
    r[0] = a * (float)i + b * (float)j + c;
    r[1] = r[0] + a;
    r[2] = r[0] + b;
    r[3] = r[1] + b;
    float mn = 1, mx = -1;
    for (int k = 0; k < 4; ++k) {
        if (r[k] < mn) mn = r[k];
        if (r[k] > mx) mx = r[k];
    }
    return mn <= 0.0f && mx >= 0.0f;
}

static int decide(const int16_t* p, const float* org, const float* dir, uint32_t level) {
    float o[3] = {0.5f * org[0], 0.5f * org[1], 0.5f * org[2]};
    const float d[3] = {0.5f * dir[0], 0.5f * dir[1], 0.5f * dir[2]};
    o[0] += 0.5f;
    o[1] += 0.5f;
    o[2] += 0.5f;
    const float oxd[3] = {o[1] * d[2] - d[1] * o[2], o[2] * d[0] - d[2] * o[0], o[0] * d[1] - d[0] * o[1]};
    const float s1 = (float)(1.0 / (double)(float)(1u << level));
    const float s2 = s1 * s1;
    return face_eval((uint16_t)p[1], (uint16_t)p[2], -s2 * d[2], s2 * d[1], s1 * oxd[0]) &&
           face_eval((uint16_t)p[0], (uint16_t)p[2], s2 * d[2], -s2 * d[0], s1 * oxd[1]) &&
           face_eval((uint16_t)p[0], (uint16_t)p[1], -s2 * d[1], s2 * d[0], s1 * oxd[2]);
}

/* Nuggets [M][2] int32 (ray, local point index at target) of rays o, d [n][3].
 * points/pyramid as produced by orc_octree_points.  Returns M (written up to cap). */
int64_t orc_spc_raytrace(const uint8_t* bytes, int64_t osize, const int16_t* points, const int32_t* pyramid, int L,
                         const float* org, const float* dir, int64_t n, int target, int32_t* nuggets, int64_t cap) {
    const int32_t* psum = pyramid + (L + 2);
    int64_t* S = (int64_t*)malloc((size_t)(osize + 1) * sizeof(int64_t));
    S[0] = 0;
    for (int64_t i = 0; i < osize; ++i) S[i + 1] = S[i] + __builtin_popcount(bytes[i]);
    int64_t num = n;
    int32_t* cur = (int32_t*)malloc((size_t)(num > 0 ? num : 1) * 2 * sizeof(int32_t));
    for (int64_t i = 0; i < num; ++i) {
        cur[2 * i] = (int32_t)i;
        cur[2 * i + 1] = 0;
    }
    for (int l = 0; l <= target && num > 0; ++l) {
        const int16_t* proot = points + 4 * (int64_t)psum[l];
        int64_t outn = 0;
        for (int64_t t = 0; t < num; ++t) {
            const int32_t ridx = cur[2 * t], pidx = cur[2 * t + 1];
            if (!decide(proot + 4 * (int64_t)pidx, org + 3 * ridx, dir + 3 * ridx, (uint32_t)l)) continue;
            outn += (l < target) ? __builtin_popcount(bytes[psum[l] + pidx]) : 1;
        }
        int32_t* nxt = (int32_t*)malloc((size_t)(outn > 0 ? outn : 1) * 2 * sizeof(int32_t));
        int64_t w = 0;
        for (int64_t t = 0; t < num; ++t) {
            const int32_t ridx = cur[2 * t], pidx = cur[2 * t + 1];
            const int16_t* p = proot + 4 * (int64_t)pidx;
            if (!decide(p, org + 3 * ridx, dir + 3 * ridx, (uint32_t)l)) continue;
            if (l == target) {
                nxt[2 * w] = ridx;
                nxt[2 * w + 1] = pidx;
                ++w;
                continue;
            }
            const uint8_t ob = bytes[psum[l] + pidx];
            const int64_t s = S[psum[l] + pidx];
            const float scale = (float)(1.0 / (double)(float)(1u << l));
            const float* og = org + 3 * ridx;
            const float x = (float)((double)(0.5f * og[0] + 0.5f) - (double)scale * ((double)(float)p[0] + 0.5));
            const float y = (float)((double)(0.5f * og[1] + 0.5f) - (double)scale * ((double)(float)p[1] + 0.5));
            const float z = (float)((double)(0.5f * og[2] + 0.5f) - (double)scale * ((double)(float)p[2] + 0.5));
            uint32_t code = 0;
            if (x > 0) code = 4;
            if (y > 0) code += 2;
            if (z > 0) code += 1;
            for (int i = 0; i < 8; ++i) {
                const uint32_t j = ORDER[code][i];
                if (ob & (1u << j)) {
                    const int cnt = __builtin_popcount(ob & ((2u << j) - 1u));
                    nxt[2 * w] = ridx;
                    nxt[2 * w + 1] = (int32_t)(s + cnt - psum[l + 1]);
                    ++w;
                }
            }
        }
        free(cur);
        cur = nxt;
        num = w;
    }
    for (int64_t i = 0; i < num && i < cap; ++i) {
        nuggets[2 * i] = cur[2 * i];
        nuggets[2 * i + 1] = cur[2 * i + 1];
    }
    free(cur);
    free(S);
    return num;
}

void orc_mark_unique_rays(const int32_t* nuggets, int64_t m, int32_t* info) {
    for (int64_t i = 0; i < m; ++i) info[i] = (i == 0 || nuggets[2 * (i - 1)] != nuggets[2 * i]) ? 1 : 0;
}
