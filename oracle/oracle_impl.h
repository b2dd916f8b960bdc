/*
 * Precision-generic body of the CPU oracle.  nglod_oracle.c includes this
 * file twice: once with REAL=float / SFX=f32 (the reference's FP32 op order,
 * compiled with -ffp-contract=off so every + and * rounds on its own, as the
 * reference's PyTorch CPU kernels do) and once with REAL=double / SFX=f64
 * (the FP64 oracle the precision form is measured against).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker, never the product path.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* Trilinear sample of one feature volume at p, accumulated into out[C] with
 * weight order and corner order of ATen's 3-D grid sampler, which the
 * reference calls with align_corners=True, padding_mode='border'
 * (sdf-net/lib/models/OctreeSDF.py:50-51):
 *   u = ((x + 1) / 2) * r, clipped to [0, r]            (unnormalize + border)
 *   corners tnw, tne, tsw, tse, bnw, bne, bsw, bse       (x fastest, then y, z)
 *   weight = wx * wy * wz, corners outside [0, r] skipped.
 * Grid layout here is corner-major: grid[((z*(r+1) + y)*(r+1) + x)*C + c],
 * which is fm[0, c, z, y, x] of the reference (geoutils.py:76 indexing). */
static void FN(orc_gather)(const orc_model* m, int li, const REAL p[3], REAL* out)
{
    const int r = m->res[li];
    const int S = r + 1;
    const int C = m->fdim;
    const float* g = m->grid[li];
    REAL w[3][2];
    int64_t i0[3];
    for (int a = 0; a < 3; ++a) {
        REAL u = ((p[a] + (REAL)1) / (REAL)2) * (REAL)r;
        /* border clip: min(r, max(u, 0)) with std::max / std::min semantics */
        u = (u < (REAL)0) ? (REAL)0 : u;
        u = ((REAL)r < u) ? (REAL)r : u;
        REAL f = (REAL)floor((double)u);
        i0[a] = (int64_t)f;
        w[a][0] = (REAL)(i0[a] + 1) - u;
        w[a][1] = u - (REAL)i0[a];
    }
    for (int c = 0; c < C; ++c) out[c] = (REAL)0;
    for (int corner = 0; corner < 8; ++corner) {
        const int bx = corner & 1, by = (corner >> 1) & 1, bz = (corner >> 2) & 1;
        const int64_t ix = i0[0] + bx, iy = i0[1] + by, iz = i0[2] + bz;
        if (ix < 0 || iy < 0 || iz < 0 || ix > r || iy > r || iz > r) continue;
        const REAL wt = (w[0][bx] * w[1][by]) * w[2][bz];
        const float* v = g + ((iz * S + iy) * S + ix) * (int64_t)C;
        for (int c = 0; c < C; ++c) out[c] += (REAL)v[c] * wt;
    }
}

/* One decoder: relu(e . W0^T + b0) . W1^T + b1 (OctreeSDF.py:82-87). */
static REAL FN(orc_decode)(const orc_model* m, int di, const REAL* e, int din)
{
    const float* w0 = m->w0[di];
    const float* b0 = m->b0[di];
    const float* w1 = m->w1[di];
    REAL out = (REAL)0;
    for (int j = 0; j < m->hidden; ++j) {
        REAL acc = (REAL)0;
        const float* row = w0 + (int64_t)j * din;
        for (int k = 0; k < din; ++k) acc += (REAL)row[k] * e[k];
        REAL h = acc + (REAL)b0[j];
        if (h < (REAL)0) h = (REAL)0;
        out += (REAL)w1[j] * h;
    }
    return out + (REAL)m->b1[di][0];
}

/* OctreeSDF.sdf(x, lod) for one point (OctreeSDF.py:94-155): every LOD up to
 * `lod` is gathered and summed into the running feature (:105-110); the
 * decoder of level `lod` (or decoder 0 when joint) maps [xyz, s] or s to the
 * distance (:113-124, :142-146).  If all_out is non-null the decoder of every
 * level 0..lod is evaluated and written there (return_lst=True, :152-153). */
static REAL FN(orc_sdf_point)(const orc_model* m, const REAL p[3], int lod, REAL* all_out)
{
    REAL s[ORC_MAX_C + 3];
    REAL g[ORC_MAX_C];
    const int C = m->fdim;
    const int off = m->pos_invariant ? 0 : 3;
    REAL d = (REAL)0;
    if (!m->pos_invariant) { s[0] = p[0]; s[1] = p[1]; s[2] = p[2]; }
    for (int i = 0; i <= lod; ++i) {
        FN(orc_gather)(m, i, p, g);
        if (i == 0) for (int c = 0; c < C; ++c) s[off + c] = g[c];
        else        for (int c = 0; c < C; ++c) s[off + c] = g[c] + s[off + c];
        if (all_out || i == lod) {
            const int di = (m->num_decoders == 1) ? 0 : i;
            d = FN(orc_decode)(m, di, s, C + off);
            if (all_out) all_out[i] = d;
        }
    }
    return d;
}

void FN(orc_sdf)(const orc_model* m, const REAL* x, int64_t n, int lod, REAL* out)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = FN(orc_sdf_point)(m, x + 3 * i, lod, NULL);
}

/* out is [n, lod+1] row-major: every level's prediction (return_lst=True). */
void FN(orc_sdf_all)(const orc_model* m, const REAL* x, int64_t n, int lod, REAL* out)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) (void)FN(orc_sdf_point)(m, x + 3 * i, lod, out + i * (lod + 1));
}

/* Ray vs [-1,1]^3 (sol_nglod_kernel.cu:77-159): Majercik 2018 slab form.
 * invdir is 1/d computed in double and rounded (:94); origins inside the box
 * are skipped (:109-112); the first passing face in x,y,z order wins
 * (:133-140); hits farther than 500 are dropped (:89,:143). */
static void FN(orc_aabb_ray)(const REAL o[3], const REAL dir[3], REAL x[3], REAL* t, uint8_t* hit)
{
    x[0] = o[0]; x[1] = o[1]; x[2] = o[2];
    *t = (REAL)0;
    *hit = 0;
    REAL inv[3], sgn[3];
    for (int a = 0; a < 3; ++a) {
        inv[a] = (REAL)(1.0 / (double)dir[a]);
        sgn[a] = signbit(dir[a]) ? (REAL)1 : (REAL)-1;
    }
    REAL cmax = FMAX(FMAX(FABS(o[0]), FABS(o[1])), FABS(o[2]));
    const REAL r = (REAL)1;
    const REAL winding = cmax < r ? (REAL)-1 : (REAL)1;
    if (winding < (REAL)0) return;
    REAL d0 = (r * winding * sgn[0] - o[0]) * inv[0];
    REAL d1 = (r * winding * sgn[1] - o[1]) * inv[1];
    REAL d2 = (r * winding * sgn[2] - o[2]) * inv[2];
    REAL ltxy = FMA(dir[1], d0, o[1]);
    REAL ltxz = FMA(dir[2], d0, o[2]);
    REAL ltyx = FMA(dir[0], d1, o[0]);
    REAL ltyz = FMA(dir[2], d1, o[2]);
    REAL ltzx = FMA(dir[0], d2, o[0]);
    REAL ltzy = FMA(dir[1], d2, o[1]);
    int test0 = (d0 >= (REAL)0) && (FABS(ltxy) < r) && (FABS(ltxz) < r);
    int test1 = (d1 >= (REAL)0) && (FABS(ltyx) < r) && (FABS(ltyz) < r);
    int test2 = (d2 >= (REAL)0) && (FABS(ltzx) < r) && (FABS(ltzy) < r);
    REAL sx = 0, sy = 0, sz = 0;
    if (test0) sx = sgn[0];
    else if (test1) sy = sgn[1];
    else if (test2) sz = sgn[2];
    REAL d = 0;
    if (sx != (REAL)0) d = d0;
    else if (sy != (REAL)0) d = d1;
    else if (sz != (REAL)0) d = d2;
    int h = (sx != (REAL)0) | (sy != (REAL)0) | (sz != (REAL)0);
    if (h && d < (REAL)500) {
        *hit = 1;
        x[0] = FMA(dir[0], d, o[0]);
        x[1] = FMA(dir[1], d, o[1]);
        x[2] = FMA(dir[2], d, o[2]);
        *t = d;
    }
}

void FN(orc_aabb)(const REAL* o, const REAL* d, REAL* x, REAL* t, uint8_t* hit, int64_t n)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) FN(orc_aabb_ray)(o + 3 * i, d + 3 * i, x + 3 * i, t + i, hit + i);
}

/* Central-difference gradient, h = 1/(64*3) (diffutils.py:61-70), then
 * F.normalize(p=2, eps=1e-5) (SphereTracer.py:128-130). */
static void FN(orc_fd_normal)(const orc_model* m, const REAL x[3], int lod, REAL nrm[3])
{
    const REAL h = (REAL)(1.0 / (64.0 * 3.0));
    const REAL twoh = (REAL)((1.0 / (64.0 * 3.0)) * 2.0);
    REAL g[3];
    for (int a = 0; a < 3; ++a) {
        REAL xp[3] = {x[0], x[1], x[2]}, xm[3] = {x[0], x[1], x[2]};
        xp[a] = x[a] + h;
        xm[a] = x[a] - h;
        REAL fp = FN(orc_sdf_point)(m, xp, lod, NULL);
        REAL fm = FN(orc_sdf_point)(m, xm, lod, NULL);
        g[a] = (fp - fm) / twoh;
    }
    REAL nn = (REAL)sqrt((double)(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]));
    if (nn < (REAL)1e-5) nn = (REAL)1e-5;
    nrm[0] = g[0] / nn; nrm[1] = g[1] / nn; nrm[2] = g[2] / nn;
}

/* SphereTracer.forward for one ray (SphereTracer.py:41-132), per-ray form of
 * the batched loop: a ray whose `cond` drops stays frozen, so its state does
 * not depend on the other rays.  `hit` is |t| < far taken at the start of the
 * last iteration the ray saw (:84), x lags t by one step (:102, :109, :114),
 * then hit &= all(|x| <= 1) (:119) and FD normals at x[hit] (:128-130).
 * margin (optional) is the smallest distance of any decision input to its
 * threshold, so tests can set apart threshold-ambiguous rays. evals counts
 * SDF evaluations (init + march + 6 for normals). */
static void FN(orc_trace_ray)(const orc_model* m, const REAL o[3], const REAL dir[3], int lod,
                              const orc_trace_params* P,
                              REAL x[3], REAL* tout, uint8_t* hitout, REAL nrm[3],
                              int32_t* evals, REAL* margin)
{
    const REAL min_dis = (REAL)P->min_dis;
    const REAL min_dis3 = (REAL)(P->min_dis * 3.0);
    const REAL far_ = (REAL)P->far_clip;
    const REAL step = (REAL)P->step_size;
    REAL t;
    uint8_t cond;
    int32_t ne = 0;
    REAL mg = (REAL)1e30;
    FN(orc_aabb_ray)(o, dir, x, &t, &cond);
    REAL D = 0, Dp = 0;
    if (cond) { D = FN(orc_sdf_point)(m, x, lod, NULL); ne++; }
    Dp = D;
    int hitflag = 0;
    for (int it = 0; it < P->num_steps; ++it) {
        hitflag = FABS(t) < far_;
        if (cond) {
            REAL a1 = FABS(D), a2 = FABS((D + Dp) / (REAL)2), a3 = FABS(t);
            REAL m1 = FABS(a1 - min_dis), m2 = FABS(a2 - min_dis3), m3 = FABS(a3 - far_);
            if (m1 < mg) mg = m1;
            if (m2 < mg) mg = m2;
            if (m3 < mg) mg = m3;
        }
        cond = cond && (FABS(D) > min_dis) && (FABS((D + Dp) / (REAL)2) > min_dis3) && hitflag;
        if (!cond) break;
        x[0] = o[0] + dir[0] * t;
        x[1] = o[1] + dir[1] * t;
        x[2] = o[2] + dir[2] * t;
        Dp = D;
        D = FN(orc_sdf_point)(m, x, lod, NULL) * step;
        ne++;
        t = t + D;
    }
    int inside = !(FABS(x[0]) > (REAL)1 || FABS(x[1]) > (REAL)1 || FABS(x[2]) > (REAL)1);
    for (int a = 0; a < 3; ++a) {
        REAL mb = FABS(FABS(x[a]) - (REAL)1);
        if (mb < mg) mg = mb;
    }
    uint8_t hit = (uint8_t)(hitflag && inside);
    *tout = t;
    *hitout = hit;
    nrm[0] = nrm[1] = nrm[2] = (REAL)0;
    if (hit) { FN(orc_fd_normal)(m, x, lod, nrm); ne += 6; }
    if (evals) *evals = ne;
    if (margin) *margin = mg;
}

void FN(orc_trace)(const orc_model* m, const REAL* o, const REAL* d, int64_t n, int lod,
                   const orc_trace_params* P, REAL* x, REAL* t, uint8_t* hit, REAL* nrm,
                   int32_t* evals, REAL* margin)
{
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i)
        FN(orc_trace_ray)(m, o + 3 * i, d + 3 * i, lod, P, x + 3 * i, t + i, hit + i, nrm + 3 * i,
                          evals ? evals + i : NULL, margin ? margin + i : NULL);
}

/* FD normals for given points (diffutils.py:61-70 + F.normalize). */
void FN(orc_normals)(const orc_model* m, const REAL* x, int64_t n, int lod, REAL* nrm)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) FN(orc_fd_normal)(m, x + 3 * i, lod, nrm + 3 * i);
}

#undef FN
#undef CAT
#undef CAT2
