#define _DEFAULT_SOURCE
/*
 * CPU oracle for the training-target path (TEST INFRASTRUCTURE ONLY).
 *
 *   orc_mesh2sdf_f64   signed point-to-mesh distance, FP64 restatement of
 *                      kernel_mesh2sdf_quad + kernel_quad_aggr
 *                      (sdf-net/lib/extensions/mesh2sdf_cuda/mesh2sdf_kernel.cu:307-616):
 *                      unsigned distance from the edge/face rule (:423-445),
 *                      sign from 13 ray-stab directions (:341-353, :479-529):
 *                      inside iff every direction has a hit at t >= 0 and a
 *                      hit at t < 0 (:584-602).
 *   orc_point_tri_dist2_f64  an independent closed form (Ericson, Real-Time
 *                      Collision Detection, 5.1.5, closest point on triangle),
 *                      used by the tests to pin the edge/face rule.
 *   orc_winding_f64    generalized winding number (sum of signed solid angles,
 *                      Van Oosterom & Strackee 1983), used by the tests to pin
 *                      the ray-stab sign on closed meshes.
 *   orc_splitmix64 / orc_sample_*  the samplers' counter-based RNG and
 *                      formulas (sdf-net/lib/torchgp/sample_surface.py:36-52,
 *                      sample_near_surface.py:42-45, sample_uniform.py:29-31).
 */
#include <math.h>
#include <stdint.h>

static const double STAB[13][3] = {
    {1.0f, 0.0f, 0.0f},
    {0.0f, 1.0f, 0.0f},
    {0.0f, 0.0f, 1.0f},
    {0.0f, 0.707106781f, 0.707106781f},
    {0.707106781f, 0.0f, 0.707106781f},
    {0.707106781f, 0.707106781f, 0.0f},
    {0.0f, 0.707106781f, -0.707106781f},
    {0.707106781f, 0.0f, -0.707106781f},
    {0.707106781f, -0.707106781f, 0.0f},
    {0.577350269f, 0.577350269f, 0.577350269f},
    {-0.577350269f, 0.577350269f, 0.577350269f},
    {0.577350269f, -0.577350269f, 0.577350269f},
    {0.577350269f, 0.577350269f, -0.577350269f},
};

static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double sgn1(double x) { return copysign(1.0, x); }
static inline double clamp01(double x) { return fmax(0.0, fmin(x, 1.0)); }
static inline double d2axmb(const double* a, double x, const double* b) {
    const double t0 = a[0] * x - b[0], t1 = a[1] * x - b[1], t2 = a[2] * x - b[2];
    return t0 * t0 + t1 * t1 + t2 * t2;
}

/* dist2 and stab flags of one point against one triangle (tri: 9 doubles). */
static void point_triangle(const double* p, const double* tri, double* dist2, uint32_t* flags) {
    const double* a = tri;
    const double* b = tri + 3;
    const double* c = tri + 6;
    double v10[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double v21[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
    double v02[3] = {a[0] - c[0], a[1] - c[1], a[2] - c[2]};
    double nor[3], c10[3], c21[3], c02[3];
    cross3(v10, v02, nor);
    cross3(v10, nor, c10);
    cross3(v21, nor, c21);
    cross3(v02, nor, c02);
    double p0[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double p1[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double p2[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    *dist2 = INFINITY;
    if (nor[0] != 0.0 || nor[1] != 0.0 || nor[2] != 0.0) {
        const double s = sgn1(dot3(c10, p0)) + sgn1(dot3(c21, p1)) + sgn1(dot3(c02, p2));
        double d2;
        if (s < 2.0) {
            const double e1 = d2axmb(v10, clamp01(dot3(v10, p0) / dot3(v10, v10)), p0);
            const double e2 = d2axmb(v21, clamp01(dot3(v21, p1) / dot3(v21, v21)), p1);
            const double e3 = d2axmb(v02, clamp01(dot3(v02, p2) / dot3(v02, v02)), p2);
            d2 = fmin(e1, fmin(e2, e3));
        } else {
            const double dn = dot3(nor, p0);
            d2 = dn * dn / dot3(nor, nor);
        }
        *dist2 = d2 < 0.0 ? 0.0 : d2;
    }
    const double e2v[3] = {-v02[0], -v02[1], -v02[2]};
    for (int i = 0; i < 13; ++i) {
        const double* dir = STAB[i];
        double pvec[3], qvec[3];
        cross3(dir, e2v, pvec);
        const double det = dot3(v10, pvec);
        if (det > -1e-8 && det < 1e-8) continue;
        const double inv = 1.0 / det;
        const double u = dot3(p0, pvec) * inv;
        if (u < 0.0 || u > 1.0) continue;
        cross3(p0, v10, qvec);
        const double v = dot3(dir, qvec) * inv;
        if (v < 0.0 || u + v > 1.0) continue;
        const double t = dot3(e2v, qvec) * inv;
        *flags |= (t >= 0.0) ? (1u << i) : (1u << (13 + i));
    }
}

void orc_mesh2sdf_f64(const double* pts, int64_t n, const double* mesh, int64_t nf, double* out) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double best = INFINITY;
        uint32_t flags = 0;
        for (int64_t f = 0; f < nf; ++f) {
            double d2;
            point_triangle(pts + 3 * i, mesh + 9 * f, &d2, &flags);
            best = fmin(best, d2);
        }
        if (best < 0.0) best = 0.0;
        double d = sqrt(best);
        const uint32_t all = (1u << 13) - 1u;
        const int inside = ((flags & all) == all) && (((flags >> 13) & all) == all);
        out[i] = inside ? -d : d;
    }
}

/* Closest point on triangle abc to p (Ericson 5.1.5), squared distance. */
double orc_point_tri_dist2_f64(const double* p, const double* a, const double* b, const double* c) {
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double q[3];
    const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { q[0] = a[0]; q[1] = a[1]; q[2] = a[2]; goto done; }
    {
        double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
        const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
        if (d3 >= 0.0 && d4 <= d3) { q[0] = b[0]; q[1] = b[1]; q[2] = b[2]; goto done; }
        const double vc = d1 * d4 - d3 * d2;
        if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
            const double v = d1 / (d1 - d3);
            for (int k = 0; k < 3; ++k) q[k] = a[k] + v * ab[k];
            goto done;
        }
        double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
        const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
        if (d6 >= 0.0 && d5 <= d6) { q[0] = c[0]; q[1] = c[1]; q[2] = c[2]; goto done; }
        const double vb = d5 * d2 - d1 * d6;
        if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
            const double w = d2 / (d2 - d6);
            for (int k = 0; k < 3; ++k) q[k] = a[k] + w * ac[k];
            goto done;
        }
        const double va = d3 * d6 - d5 * d4;
        if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
            const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            for (int k = 0; k < 3; ++k) q[k] = b[k] + w * (c[k] - b[k]);
            goto done;
        }
        const double den = 1.0 / (va + vb + vc);
        const double v = vb * den, w = vc * den;
        for (int k = 0; k < 3; ++k) q[k] = a[k] + ab[k] * v + ac[k] * w;
    }
done:;
    double r[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    return dot3(r, r);
}

/* Generalized winding number of a closed triangle mesh at p. */
double orc_winding_f64(const double* p, const double* mesh, int64_t nf) {
    double total = 0.0;
    for (int64_t f = 0; f < nf; ++f) {
        const double* t = mesh + 9 * f;
        double a[3], b[3], c[3];
        for (int k = 0; k < 3; ++k) { a[k] = t[k] - p[k]; b[k] = t[3 + k] - p[k]; c[k] = t[6 + k] - p[k]; }
        const double la = sqrt(dot3(a, a)), lb = sqrt(dot3(b, b)), lc = sqrt(dot3(c, c));
        double bc[3];
        cross3(b, c, bc);
        const double num = dot3(a, bc);
        const double den = la * lb * lc + dot3(a, b) * lc + dot3(b, c) * la + dot3(c, a) * lb;
        total += 2.0 * atan2(num, den);
    }
    return total / (4.0 * M_PI);
}

/* SplitMix64 (Steele, Lea, Flood 2014; constants as published by Vigna). */
uint64_t orc_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform float in [0, 1) from draw k of sample i of stream `seed`. */
static double rnd01(uint64_t seed, uint64_t i, uint32_t k) {
    const uint64_t z = orc_splitmix64(seed ^ orc_splitmix64(i * 8ull + k));
    return (double)(z >> 40) * (1.0 / 16777216.0);
}

void orc_sample_uniform(uint64_t seed, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) out[3 * i + k] = (float)(rnd01(seed, (uint64_t)i, (uint32_t)k) * 2.0 - 1.0);
}

/* Directions uniform on S^2 (sdf-net/lib/geoutils.py:31-40), one stream of
 * the same counter RNG: draws 0 (z) and 1 (phi). */
void orc_sample_directions(uint64_t seed, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        const double z = 1.0 - 2.0 * rnd01(seed, (uint64_t)i, 0);
        const double r = sqrt(1.0 - z * z);
        const double phi = 2.0 * M_PI * rnd01(seed, (uint64_t)i, 1);
        out[3 * i] = (float)(r * cos(phi));
        out[3 * i + 1] = (float)(r * sin(phi));
        out[3 * i + 2] = (float)z;
    }
}

/* Area-weighted surface sample + optional Gaussian offset of std sigma.
 * cdf: inclusive cumulative normalized areas [nf]; face = first f with
 * cdf[f] > r0.  Returns face indices too. */
void orc_sample_surface(const float* mesh, const double* cdf, int64_t nf, uint64_t seed, int64_t n, double sigma,
                        float* out, int32_t* face) {
    for (int64_t i = 0; i < n; ++i) {
        const double r0 = rnd01(seed, (uint64_t)i, 0);
        int64_t lo = 0, hi = nf - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (cdf[mid] > r0) hi = mid; else lo = mid + 1;
        }
        face[i] = (int32_t)lo;
        const float* t = mesh + 9 * lo;
        const double u = sqrt(rnd01(seed, (uint64_t)i, 1));
        const double v = rnd01(seed, (uint64_t)i, 2);
        double g[3] = {0, 0, 0};
        if (sigma > 0.0) {
            const double u1 = 1.0 - rnd01(seed, (uint64_t)i, 3), u2 = rnd01(seed, (uint64_t)i, 4);
            const double u3 = 1.0 - rnd01(seed, (uint64_t)i, 5), u4 = rnd01(seed, (uint64_t)i, 6);
            const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
            g[0] = r1 * cos(2.0 * M_PI * u2);
            g[1] = r1 * sin(2.0 * M_PI * u2);
            g[2] = r2 * cos(2.0 * M_PI * u4);
        }
        for (int k = 0; k < 3; ++k) {
            const double p = (1.0 - u) * t[k] + (u * (1.0 - v)) * t[3 + k] + (u * v) * t[6 + k];
            out[3 * i + k] = (float)(p + sigma * g[k]);
        }
    }
}
