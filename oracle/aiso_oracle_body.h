/*
 * aiso_oracle_body.h -- precision-generic body of the CPU oracle.
 * Included twice by aiso_oracle.c with REAL/SFX set to (float,_f) and
 * (double,_d).  TEST INFRASTRUCTURE ONLY (see aiso_oracle.h).
 *
 * Kernel order, indexing and rounding follow the reference CUDA kernels one
 * for one; each function cites the kernel or driver it restates.  Loops over
 * interior points are OpenMP-parallel over ix (each output element is written
 * by exactly one iteration, as in the reference's one-thread-per-point launch),
 * so results do not depend on the thread count.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

#define FAT ORC_FAT

/* Laplacian coefficients, time-interpolation filter and cosine sponge table.
 * F/src/BornShotsGpuFunctions.cu:135-147 (coefficients), :151-167 (filter),
 * :169-182 (sponge); identical in nonlinearShotsGpuFunctions.cu:123-170.
 * filt has 2*(sub+1) entries, cd has minPad entries. */
void FN(orc_tables)(const orc_grid* g, REAL* zc, REAL* xc, REAL* filt, REAL* cd)
{
    static const double C[6] = {-2.927222222, 1.666666667, -0.238095238,
                                0.039682539, -0.004960317, 0.000317460};
    REAL dz = (REAL)g->dz, dx = (REAL)g->dx;
    for (int k = 0; k < 6; ++k) {
        zc[k] = (REAL)(C[k] / (double)(dz * dz));
        xc[k] = (REAL)(C[k] / (double)(dx * dx));
    }
    int sub = g->sub, nts = g->nts, ntw = (nts - 1) * sub + 1, h = sub + 1;
    double s = 1.0 / sqrt((double)((REAL)ntw / (REAL)nts));
    for (int i = 0; i < h; ++i) {
        REAL a = (REAL)(1.0 - 1.0 * i / sub);
        REAL b = (REAL)(1.0 - (double)a);
        filt[i] = (REAL)((double)a * s);
        filt[i + h] = (REAL)((double)b * s);
    }
    REAL alpha = (REAL)g->alphaCos;
    for (int j = 0; j < g->minPad; ++j) {
        REAL arg = (REAL)(M_PI / (1.0 * g->minPad) * 1.0 * (g->minPad - j));
        cd[j] = (REAL)((double)alpha + (1.0 - (double)alpha) * cos((double)arg));
    }
}

/* v^2*dtw^2 and the reflectivity scale 2/v^3.  F/src/fdParam.cpp:23,122-139. */
void FN(orc_vel_scales)(const orc_grid* g, const REAL* vel, REAL* v2dt2, REAL* refl)
{
    REAL dtw = (REAL)g->dts / (REAL)g->sub;
    long n = (long)g->nz * g->nx;
    for (long i = 0; i < n; ++i) {
        REAL v = vel[i];
        v2dt2[i] = v * v * dtw * dtw;
        refl[i] = (REAL)(2.0 / (double)(v * v * v));
    }
}

/* The timing builds (ORC_SIMD) vectorize the inner loop: every loop body
 * below writes per point only, and vector lanes round each operation as the
 * scalar code does, so the bits do not change. */
#ifdef ORC_SIMD
#define ORC_INNER _Pragma("omp simd")
#else
#define ORC_INNER
#endif
#define INTERIOR_BEGIN(g)                                              \
    _Pragma("omp parallel for schedule(static)")                       \
    for (int ix = FAT; ix < (g)->nx - FAT; ++ix) {                     \
        ORC_INNER                                                      \
        for (int iz = FAT; iz < (g)->nz - FAT; ++iz) {                 \
            long i = (long)ix * (g)->nz + iz;
#define INTERIOR_END }}

/* setFsConditionFwdGpu, F/src/kernelsGpu.cu:555-567: on the current slice,
 * before each forward step with a free surface, the surface row (iz = FAT)
 * is zeroed and the FAT rows above it take the opposite of their mirror
 * images below it. */
static void FN(set_fs)(const orc_grid* g, REAL* c)
{
    const long nz = g->nz;
    for (int ix = FAT; ix < g->nx - FAT; ++ix) {
        REAL* col = c + (long)ix * nz;
        col[FAT] = 0;
        for (int k = 0; k < FAT; ++k) col[k] = -col[2 * FAT - k];
    }
}

/* stepFwdGpu, F/src/kernelsGpu.cu:513-552.  In place when o == n.  With a
 * free surface the current slice is first conditioned by set_fs, in place,
 * as the reference's drivers do before every forward step
 * (e.g. F/src/BornShotsGpuFunctions.cu:579-583). */
void FN(orc_step_fwd)(const orc_grid* g, const REAL* zc, const REAL* xc,
                      REAL* o, REAL* c, REAL* n, const REAL* v2dt2)
{
    const long nz = g->nz;
    if (g->freeSurface) FN(set_fs)(g, c);
    INTERIOR_BEGIN(g)
        REAL lap = zc[0] * c[i]
                 + zc[1] * (c[i - 1] + c[i + 1])
                 + zc[2] * (c[i - 2] + c[i + 2])
                 + zc[3] * (c[i - 3] + c[i + 3])
                 + zc[4] * (c[i - 4] + c[i + 4])
                 + zc[5] * (c[i - 5] + c[i + 5])
                 + xc[0] * c[i]
                 + xc[1] * (c[i + nz] + c[i - nz])
                 + xc[2] * (c[i + 2 * nz] + c[i - 2 * nz])
                 + xc[3] * (c[i + 3 * nz] + c[i - 3 * nz])
                 + xc[4] * (c[i + 4 * nz] + c[i - 4 * nz])
                 + xc[5] * (c[i + 5 * nz] + c[i - 5 * nz]);
        n[i] = v2dt2[i] * lap + c[i] + c[i] - o[i];
    INTERIOR_END
}

/* stepAdjGpu, F/src/kernelsGpu.cu:574-610: the Laplacian acts on v2dt2*c
 * (formed on the halo too).  In place when o == n.  tmp holds nz*nx values.
 * With a free surface, stepAdjFsGpu (:613-700): the surface row is left as
 * it is and the four rows below it drop the z terms that the forward's
 * mirror folds onto them. */
void FN(orc_step_adj)(const orc_grid* g, const REAL* zc, const REAL* xc,
                      REAL* o, const REAL* c, REAL* n, const REAL* v2dt2, REAL* tmp)
{
    const long nz = g->nz, nzx = (long)g->nz * g->nx;
    const int fs = g->freeSurface;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nzx; ++i) tmp[i] = c[i] * v2dt2[i];
    const REAL* s = tmp;
    INTERIOR_BEGIN(g)
        if (fs && iz == FAT) continue;
        REAL lap = zc[0] * s[i]
                 + zc[1] * (s[i - 1] + s[i + 1])
                 + zc[2] * (s[i - 2] + s[i + 2])
                 + zc[3] * (s[i - 3] + s[i + 3])
                 + zc[4] * (s[i - 4] + s[i + 4])
                 + zc[5] * (s[i - 5] + s[i + 5]);
        if (fs) {
            if (iz == FAT + 1) lap = lap - zc[2] * s[i] - zc[3] * s[i + 1] - zc[4] * s[i + 2] - zc[5] * s[i + 3];
            else if (iz == FAT + 2) lap = lap - zc[3] * s[i - 1] - zc[4] * s[i] - zc[5] * s[i + 1];
            else if (iz == FAT + 3) lap = lap - zc[4] * s[i - 2] - zc[5] * s[i - 1];
            else if (iz == FAT + 4) lap = lap - zc[5] * s[i - 3];
        }
        lap = lap
                 + xc[0] * s[i]
                 + xc[1] * (s[i + nz] + s[i - nz])
                 + xc[2] * (s[i + 2 * nz] + s[i - 2 * nz])
                 + xc[3] * (s[i + 3 * nz] + s[i - 3 * nz])
                 + xc[4] * (s[i + 4 * nz] + s[i - 4 * nz])
                 + xc[5] * (s[i + 5 * nz] + s[i - 5 * nz]);
        o[i] = lap + c[i] + c[i] - n[i];
    INTERIOR_END
}

/* dampCosineEdge, F/src/kernelsGpu.cu:202-219: both slices are damped.
 * With a free surface (dampCosineEdgeFs, :221-239) there is no sponge at
 * the top. */
void FN(orc_damp)(const orc_grid* g, const REAL* cd, REAL* p1, REAL* p2)
{
    const int nz = g->nz, nx = g->nx, minPad = g->minPad, fs = g->freeSurface;
    INTERIOR_BEGIN(g)
        int a = fs ? ix - FAT : iz - FAT, b = ix - FAT, c = nz - iz - 1 - FAT, d = nx - ix - 1 - FAT;
        int m1 = a < b ? a : b, m2 = c < d ? c : d;
        int dist = m1 < m2 ? m1 : m2;
        if (dist < minPad) {
            REAL damp = cd[dist];
            p1[i] *= damp;
            p2[i] *= damp;
        }
    INTERIOR_END
}

/* injectSource, F/src/kernelsGpu.cu:88-91. */
static void FN(inject_source)(int ntw, const REAL* sig, REAL* p, int itw,
                              const int* pos, int nPos)
{
    for (int s = 0; s < nPos; ++s) p[pos[s]] += sig[(long)s * ntw + itw];
}

/* recordInterpData, F/src/kernelsGpu.cu:119-127. */
static void FN(record_interp)(int nts, int h, const REAL* f, const REAL* p, REAL* d,
                              int its, int it2, const int* pos, int nPos)
{
    for (int r = 0; r < nPos; ++r) {
        d[(long)nts * r + its] += p[pos[r]] * f[it2];
        d[(long)nts * r + its + 1] += p[pos[r]] * f[h + it2];
    }
}

/* interpInjectData, F/src/kernelsGpu.cu:94-99. */
static void FN(interp_inject)(int nts, int h, const REAL* f, const REAL* d, REAL* p,
                              int its, int it2, const int* pos, int nPos)
{
    for (int r = 0; r < nPos; ++r)
        p[pos[r]] += d[(long)nts * r + its] * f[it2 + 1]
                   + d[(long)nts * r + its + 1] * f[h + it2 + 1];
}

/* recordSource, F/src/kernelsGpu.cu:113-116. */
static void FN(record_source)(int ntw, const REAL* p, REAL* sig, int itw,
                              const int* pos, int nPos)
{
    for (int s = 0; s < nPos; ++s) sig[(long)ntw * s + itw] += p[pos[s]];
}

/* interpWavefield (F/src/kernelsGpu.cu:132-141) and
 * extractInterpAdjointWavefield (:166-173): same arithmetic into two slices. */
static void FN(capture2)(const orc_grid* g, const REAL* f, int h, const REAL* p,
                         REAL* left, REAL* right, int it2)
{
    INTERIOR_BEGIN(g)
        left[i] += p[i] * f[it2];
        right[i] += p[i] * f[h + it2];
    INTERIOR_END
}

/* injectSecondarySource, F/src/kernelsGpu.cu:102-107. */
static void FN(inject_secondary)(const orc_grid* g, const REAL* f, int h,
                                 const REAL* ssL, const REAL* ssR, REAL* p, int k)
{
    INTERIOR_BEGIN(g)
        p[i] += ssL[i] * f[k] + ssR[i] * f[h + k];
    INTERIOR_END
}

static void FN(swap)(REAL** a, REAL** b) { REAL* t = *a; *a = *b; *b = t; }

static REAL* FN(zalloc)(long n) { return (REAL*)calloc((size_t)n, sizeof(REAL)); }

/* propShotsFwdGpu, F/src/nonlinearShotsGpuFunctions.cu:227-309.
 * sig: nSrc x ntw (already scaled and time-interpolated); dreg: nRec x nts.
 * With wfld (nts x nz x nx, or NULL): propShotsFwdGpuWavefield (:310-390),
 * which also captures the coarse wavefield (interpWavefield) after the
 * damping of every fine step. */
void FN(orc_nl_fwd_wfld)(const orc_grid* g, const REAL* v2dt2, const REAL* sig,
                         const int* srcPos, int nSrc, const int* recPos, int nRec, REAL* dreg,
                         REAL* wfld)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const int nts = g->nts, sub = g->sub, ntw = (nts - 1) * sub + 1, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    memset(dreg, 0, sizeof(REAL) * (size_t)nRec * nts);
    for (int its = 0; its < nts - 1; ++its) {
        for (int it2 = 1; it2 < sub + 1; ++it2) {
            int itw = its * sub + it2;
            FN(orc_step_fwd)(g, zc, xc, p0, p1, p0, v2dt2);
            FN(inject_source)(ntw, sig, p0, itw - 1, srcPos, nSrc);
            FN(orc_damp)(g, cd, p0, p1);
            if (wfld) FN(capture2)(g, f, h, p0, wfld + (long)its * nzx, wfld + (long)(its + 1) * nzx, it2);
            FN(record_interp)(nts, h, f, p0, dreg, its, it2, recPos, nRec);
            FN(swap)(&p0, &p1);
        }
    }
    free(p0); free(p1);
}

void FN(orc_nl_fwd)(const orc_grid* g, const REAL* v2dt2, const REAL* sig,
                    const int* srcPos, int nSrc, const int* recPos, int nRec, REAL* dreg)
{
    FN(orc_nl_fwd_wfld)(g, v2dt2, sig, srcPos, nSrc, recPos, nRec, dreg, NULL);
}

/* propShotsAdjGpu, F/src/nonlinearShotsGpuFunctions.cu:579-652.
 * sig (output): nSrc x ntw, overwritten.  With wfld (nts x nz x nx, zeroed
 * here, or NULL): propShotsAdjGpuWavefield (:653-740), which also captures
 * the adjoint field at the coarse rate (interpWavefield with the backward
 * fine index it2) after the damping. */
void FN(orc_nl_adj_wfld)(const orc_grid* g, const REAL* v2dt2, REAL* sig,
                         const int* srcPos, int nSrc, const int* recPos, int nRec,
                         const REAL* dreg, REAL* wfld)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const int nts = g->nts, sub = g->sub, ntw = (nts - 1) * sub + 1, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    REAL* tmp = FN(zalloc)(nzx);
    memset(sig, 0, sizeof(REAL) * (size_t)nSrc * ntw);
    if (wfld) memset(wfld, 0, sizeof(REAL) * (size_t)nzx * nts);
    for (int its = nts - 2; its > -1; --its) {
        for (int it2 = sub - 1; it2 > -1; --it2) {
            int itw = its * sub + it2;
            FN(orc_step_adj)(g, zc, xc, p0, p1, p0, v2dt2, tmp);
            FN(interp_inject)(nts, h, f, dreg, p0, its, it2, recPos, nRec);
            FN(orc_damp)(g, cd, p0, p1);
            if (wfld) FN(capture2)(g, f, h, p0, wfld + (long)its * nzx, wfld + (long)(its + 1) * nzx, it2);
            FN(record_source)(ntw, p0, sig, itw, srcPos, nSrc);
            FN(swap)(&p0, &p1);
        }
    }
    free(p0); free(p1); free(tmp);
}

void FN(orc_nl_adj)(const orc_grid* g, const REAL* v2dt2, REAL* sig,
                    const int* srcPos, int nSrc, const int* recPos, int nRec,
                    const REAL* dreg)
{
    FN(orc_nl_adj_wfld)(g, v2dt2, sig, srcPos, nSrc, recPos, nRec, dreg, NULL);
}

/* 10th-order Laplacian of c at interior point i, in the reference's order. */
static REAL FN(lap_at)(const REAL* zc, const REAL* xc, const REAL* c, long i, long nz)
{
    return zc[0] * c[i]
         + zc[1] * (c[i - 1] + c[i + 1])
         + zc[2] * (c[i - 2] + c[i + 2])
         + zc[3] * (c[i - 3] + c[i + 3])
         + zc[4] * (c[i - 4] + c[i + 4])
         + zc[5] * (c[i - 5] + c[i + 5])
         + xc[0] * c[i]
         + xc[1] * (c[i + nz] + c[i - nz])
         + xc[2] * (c[i + 2 * nz] + c[i - 2 * nz])
         + xc[3] * (c[i + 3 * nz] + c[i - 3 * nz])
         + xc[4] * (c[i + 4 * nz] + c[i - 4 * nz])
         + xc[5] * (c[i + 5 * nz] + c[i - 5 * nz]);
}

/* Adjoint-state gradient of J(v2dt2) = <r, d(v2dt2)> for the nonlinear
 * operator of one shot, where d is the data orc_nl_fwd records and dres = dJ/dd
 * on the regular receiver grid.  The forward recursion is
 *   p_t = D (v2dt2 * L p_{t-1} + 2 p_{t-1} - D p_{t-2} + s_t)
 * (step, injection before the sponge D, the old slice damped twice), whose
 * transpose is the adjoint propagation of orc_nl_adj driven by dres; its
 * damped slices nu_t give
 *   dJ/dv2dt2 += sum_t (L p_{t-1}) * nu_t          (grad, accumulated, t descending)
 * and sigAdj (nSrc x ntw) the adjoint recorded at the source points, which
 * pairs with the unscaled source for the v2dt2 factor of the injection.
 * Not a reference function: the reference's FWI uses the Born adjoint as its
 * gradient (PY/fwiMain.py:123-142). */
void FN(orc_nl_vgrad)(const orc_grid* g, const REAL* v2dt2, const REAL* sig,
                      const int* srcPos, int nSrc, const int* recPos, int nRec,
                      const REAL* dres, REAL* grad, REAL* sigAdj)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const int nts = g->nts, sub = g->sub, ntw = (nts - 1) * sub + 1, h = sub + 1, T = ntw - 1;
    const long nz = g->nz, nzx = (long)g->nz * g->nx;
    REAL* P = FN(zalloc)((long)T * nzx);      /* P[t] = p_t, t = 0 .. T-1 (p_0 = 0) */
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    for (int its = 0; its < nts - 1; ++its) {
        for (int it2 = 1; it2 < sub + 1; ++it2) {
            int itw = its * sub + it2;
            FN(orc_step_fwd)(g, zc, xc, p0, p1, p0, v2dt2);
            FN(inject_source)(ntw, sig, p0, itw - 1, srcPos, nSrc);
            FN(orc_damp)(g, cd, p0, p1);
            if (itw < T) memcpy(P + (long)itw * nzx, p0, sizeof(REAL) * (size_t)nzx);
            FN(swap)(&p0, &p1);
        }
    }
    memset(p0, 0, sizeof(REAL) * (size_t)nzx);
    memset(p1, 0, sizeof(REAL) * (size_t)nzx);
    REAL* tmp = FN(zalloc)(nzx);
    memset(sigAdj, 0, sizeof(REAL) * (size_t)nSrc * ntw);
    for (int its = nts - 2; its > -1; --its) {
        for (int it2 = sub - 1; it2 > -1; --it2) {
            int itw = its * sub + it2;                /* produces nu_t, t = itw + 1 */
            FN(orc_step_adj)(g, zc, xc, p0, p1, p0, v2dt2, tmp);
            FN(interp_inject)(nts, h, f, dres, p0, its, it2, recPos, nRec);
            FN(orc_damp)(g, cd, p0, p1);
            FN(record_source)(ntw, p0, sigAdj, itw, srcPos, nSrc);
            const REAL* pf = P + (long)itw * nzx;     /* p_{t-1} */
            INTERIOR_BEGIN(g)
                grad[i] += FN(lap_at)(zc, xc, pf, i, nz) * p0[i];
            INTERIOR_END
            FN(swap)(&p0, &p1);
        }
    }
    free(P); free(p0); free(p1); free(tmp);
}

/* Source-wavefield pass shared by every Born-type driver, e.g.
 * F/src/BornShotsGpuFunctions.cu:281-321: W (nts slices) is zeroed, then
 * step -> injectSource -> damp -> interpWavefield -> swap. */
static void FN(src_pass)(const orc_grid* g, const REAL* zc, const REAL* xc,
                         const REAL* f, const REAL* cd, const REAL* v2dt2,
                         const REAL* sig, const int* srcPos, int nSrc, REAL* W)
{
    const int nts = g->nts, sub = g->sub, ntw = (nts - 1) * sub + 1, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    memset(W, 0, sizeof(REAL) * (size_t)nzx * nts);
    for (int its = 0; its < nts - 1; ++its) {
        for (int it2 = 1; it2 < sub + 1; ++it2) {
            int itw = its * sub + it2;
            FN(orc_step_fwd)(g, zc, xc, p0, p1, p0, v2dt2);
            FN(inject_source)(ntw, sig, p0, itw - 1, srcPos, nSrc);
            FN(orc_damp)(g, cd, p0, p1);
            FN(capture2)(g, f, h, p0, W + (long)its * nzx, W + (long)(its + 1) * nzx, it2);
            FN(swap)(&p0, &p1);
        }
    }
    free(p0); free(p1);
}

/* Scattered pass shared by Born and extended Born forward, e.g.
 * F/src/BornShotsGpuFunctions.cu:344-376.  make_ss(k, ss) fills the secondary
 * source for coarse sample k into a zeroed slice. */
typedef void (*FN(ss_fn))(const orc_grid*, const void*, int, REAL*);

static void FN(scat_pass)(const orc_grid* g, const REAL* zc, const REAL* xc,
                          const REAL* f, const REAL* cd, const REAL* v2dt2,
                          const int* recPos, int nRec, REAL* dreg,
                          FN(ss_fn) make_ss, const void* ctx, int clearData, REAL* wsec)
{
    const int nts = g->nts, sub = g->sub, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    REAL* ssL = FN(zalloc)(nzx);
    REAL* ssR = FN(zalloc)(nzx);
    if (clearData) memset(dreg, 0, sizeof(REAL) * (size_t)nRec * nts);
    make_ss(g, ctx, 0, ssL);
    for (int its = 0; its < nts - 1; ++its) {
        make_ss(g, ctx, its + 1, ssR);
        for (int it2 = 1; it2 < sub + 1; ++it2) {
            FN(orc_step_fwd)(g, zc, xc, p0, p1, p0, v2dt2);
            FN(inject_secondary)(g, f, h, ssL, ssR, p0, it2 - 1);
            FN(orc_damp)(g, cd, p0, p1);
            /* BornShotsFwdGpuWavefield (F/src/BornShotsGpuFunctions.cu:485-486):
             * the scattered field at the coarse rate (interpWavefield) */
            if (wsec) FN(capture2)(g, f, h, p0, wsec + (long)its * nzx, wsec + (long)(its + 1) * nzx, it2);
            FN(record_interp)(nts, h, f, p0, dreg, its, it2, recPos, nRec);
            FN(swap)(&p0, &p1);
        }
        FN(swap)(&ssL, &ssR);
        memset(ssR, 0, sizeof(REAL) * (size_t)nzx);
    }
    free(p0); free(p1); free(ssL); free(ssR);
}

/* Receiver pass shared by Born and extended Born adjoint, e.g.
 * F/src/BornShotsGpuFunctions.cu:893-942.  image(k, ssR) applies the imaging
 * condition for coarse sample k. */
typedef void (*FN(img_fn))(const orc_grid*, void*, int, REAL*);

/* recordScaleWavefield (F/src/kernelsGpu.cu:145-153): the receiver
 * wavefield sample, scaled like the image, added into a zeroed slice. */
static void FN(record_scale)(const orc_grid* g, REAL* out, const REAL* ss, const REAL* refl,
                             const REAL* v2dt2)
{
    INTERIOR_BEGIN(g)
        out[i] += ss[i] * refl[i] * v2dt2[i];
    INTERIOR_END
}

static void FN(rec_pass)(const orc_grid* g, const REAL* zc, const REAL* xc,
                         const REAL* f, const REAL* cd, const REAL* v2dt2,
                         const int* recPos, int nRec, const REAL* dreg,
                         FN(img_fn) image, void* ctx, REAL* wsec, const REAL* refl)
{
    const int nts = g->nts, sub = g->sub, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    REAL* ssL = FN(zalloc)(nzx);
    REAL* ssR = FN(zalloc)(nzx);
    REAL* tmp = FN(zalloc)(nzx);
    for (int its = nts - 2; its > -1; --its) {
        for (int it2 = sub - 1; it2 > -1; --it2) {
            FN(orc_step_adj)(g, zc, xc, p0, p1, p0, v2dt2, tmp);
            FN(interp_inject)(nts, h, f, dreg, p0, its, it2, recPos, nRec);
            FN(orc_damp)(g, cd, p0, p1);
            FN(capture2)(g, f, h, p0, ssL, ssR, it2);
            FN(swap)(&p0, &p1);
        }
        image(g, ctx, its + 1, ssR);
        /* BornShotsAdjGpuWavefield (F/src/BornShotsGpuFunctions.cu:1055-1056) */
        if (wsec) FN(record_scale)(g, wsec + (long)(its + 1) * nzx, ssR, refl, v2dt2);
        FN(swap)(&ssR, &ssL);
        memset(ssL, 0, sizeof(REAL) * (size_t)nzx);
    }
    if (wsec) FN(record_scale)(g, wsec, ssR, refl, v2dt2);
    image(g, ctx, 0, ssR);
    free(p0); free(p1); free(ssL); free(ssR); free(tmp);
}

typedef struct { const REAL* m; const REAL* W; const REAL* v2dt2; const REAL* refl; } FN(born_ctx);

/* imagingFwdGpu, F/src/kernelsGpu.cu:309-316: ss = m * W[k]. */
static void FN(born_ss)(const orc_grid* g, const void* vctx, int k, REAL* ss)
{
    const FN(born_ctx)* c = (const FN(born_ctx)*)vctx;
    const long nzx = (long)g->nz * g->nx;
    const REAL* Wk = c->W + (long)k * nzx;
    INTERIOR_BEGIN(g)
        ss[i] = c->m[i] * Wk[i];
    INTERIOR_END
}

/* imagingAdjGpu, F/src/kernelsGpu.cu:318-325: m += W[k] * ss. */
static void FN(born_img)(const orc_grid* g, void* vctx, int k, REAL* ss)
{
    FN(born_ctx)* c = (FN(born_ctx)*)vctx;
    const long nzx = (long)g->nz * g->nx;
    const REAL* Wk = c->W + (long)k * nzx;
    REAL* m = (REAL*)c->m;
    INTERIOR_BEGIN(g)
        m[i] += Wk[i] * ss[i];
    INTERIOR_END
}

/* scaleReflectivity, F/src/kernelsGpu.cu:246-252. */
static void FN(scale_refl)(const orc_grid* g, REAL* m, const REAL* refl, const REAL* v2dt2)
{
    INTERIOR_BEGIN(g)
        m[i] *= v2dt2[i] * refl[i];
    INTERIOR_END
}

/* BornShotsFwdGpu, F/src/BornShotsGpuFunctions.cu:260-388.  model: nz*nx
 * (read only, copied before scaling); W: nts*nz*nx scratch. */
void FN(orc_born_fwd_wfld)(const orc_grid* g, const REAL* v2dt2, const REAL* refl,
                           const REAL* sig, const int* srcPos, int nSrc,
                           const int* recPos, int nRec, const REAL* model, REAL* dreg, REAL* W,
                           REAL* Wsec)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    REAL* m = FN(zalloc)(nzx);
    memcpy(m, model, sizeof(REAL) * (size_t)nzx);
    FN(scale_refl)(g, m, refl, v2dt2);
    FN(born_ctx) ctx = {m, W, v2dt2, refl};
    if (Wsec) memset(Wsec, 0, sizeof(REAL) * (size_t)nzx * g->nts);
    FN(scat_pass)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, FN(born_ss), &ctx, 1, Wsec);
    free(m);
}

/* W (output) is the coarse source wavefield; with Wsec (or NULL) the
 * scattered wavefield of BornShotsFwdGpuWavefield (:390-520). */
void FN(orc_born_fwd)(const orc_grid* g, const REAL* v2dt2, const REAL* refl,
                      const REAL* sig, const int* srcPos, int nSrc,
                      const int* recPos, int nRec, const REAL* model, REAL* dreg, REAL* W)
{
    FN(orc_born_fwd_wfld)(g, v2dt2, refl, sig, srcPos, nSrc, recPos, nRec, model, dreg, W, NULL);
}

/* BornShotsAdjGpu, F/src/BornShotsGpuFunctions.cu:829-957.  model (output)
 * is overwritten with this shot's image. */
void FN(orc_born_adj_wfld)(const orc_grid* g, const REAL* v2dt2, const REAL* refl,
                           const REAL* sig, const int* srcPos, int nSrc,
                           const int* recPos, int nRec, REAL* model, const REAL* dreg, REAL* W,
                           REAL* Wsec)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    memset(model, 0, sizeof(REAL) * (size_t)nzx);
    FN(born_ctx) ctx = {model, W, v2dt2, refl};
    if (Wsec) memset(Wsec, 0, sizeof(REAL) * (size_t)nzx * g->nts);
    FN(rec_pass)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, FN(born_img), &ctx, Wsec, refl);
    FN(scale_refl)(g, model, refl, v2dt2);
}

/* W (output) is the coarse source wavefield; with Wsec (or NULL) the scaled
 * receiver wavefield of BornShotsAdjGpuWavefield (:959-1080). */
void FN(orc_born_adj)(const orc_grid* g, const REAL* v2dt2, const REAL* refl,
                      const REAL* sig, const int* srcPos, int nSrc,
                      const int* recPos, int nRec, REAL* model, const REAL* dreg, REAL* W)
{
    FN(orc_born_adj_wfld)(g, v2dt2, refl, sig, srcPos, nSrc, recPos, nRec, model, dreg, W, NULL);
}

/* imagingOffsetFwdGpu (F/src/kernelsGpu.cu:381-396) followed by
 * scaleSecondarySourceFd (:270-275). */
static void FN(ext_ss)(const orc_grid* g, const void* vctx, int k, REAL* ss)
{
    const FN(born_ctx)* c = (const FN(born_ctx)*)vctx;
    const int nz = g->nz, nx = g->nx, hExt = (g->nExt - 1) / 2;
    const long nzx = (long)nz * nx;
    INTERIOR_BEGIN(g)
        int lo = -hExt, t = ix + 1 + FAT + hExt - nx;
        if (t > lo) lo = t;
        int hi = ix - hExt - FAT;
        if (hExt < hi) hi = hExt;
        for (int e = lo; e < hi + 1; ++e) {
            long iModel = nzx * (e + hExt) + (long)nz * (ix - e) + iz;
            long iSrc = nzx * k + (long)(ix - 2 * e) * nz + iz;
            ss[i] += c->m[iModel] * c->W[iSrc];
        }
    INTERIOR_END
    INTERIOR_BEGIN(g)
        ss[i] *= c->v2dt2[i];
    INTERIOR_END
}

/* scaleSecondarySourceFd (F/src/kernelsGpu.cu:270-275) followed by
 * imagingOffsetAdjGpu (:398-411). */
static void FN(ext_img)(const orc_grid* g, void* vctx, int k, REAL* ss)
{
    FN(born_ctx)* c = (FN(born_ctx)*)vctx;
    const int nz = g->nz, nx = g->nx, nExt = g->nExt, hExt = (nExt - 1) / 2;
    const long nzx = (long)nz * nx;
    REAL* m = (REAL*)c->m;
    INTERIOR_BEGIN(g)
        ss[i] *= c->v2dt2[i];
    INTERIOR_END
#pragma omp parallel for schedule(static)
    for (int ix = FAT + hExt; ix < nx - FAT - hExt; ++ix)
        for (int e = 0; e < nExt; ++e)
            for (int iz = FAT; iz < nz - FAT; ++iz) {
                int sh = e - hExt;
                long iModel = nzx * e + (long)nz * ix + iz;
                long iSrc = nzx * k + (long)nz * (ix - sh) + iz;
                long iRec = (long)nz * (ix + sh) + iz;
                m[iModel] += ss[iRec] * c->W[iSrc];
            }
}

/* scaleReflectivityLinExt, F/src/kernelsGpu.cu:256-267. */
static void FN(scale_refl_ext)(const orc_grid* g, REAL* m, const REAL* refl)
{
    const long nzx = (long)g->nz * g->nx;
    for (int e = 0; e < g->nExt; ++e) {
        REAL* me = m + nzx * e;
        INTERIOR_BEGIN(g)
            me[i] *= refl[i];
        INTERIOR_END
    }
}

/* BornOffsetShotsFwdGpu, F/src/BornExtShotsGpuFunctions.cu:582-724.
 * model: nExt*nz*nx (read only). */
void FN(orc_bornext_fwd)(const orc_grid* g, const REAL* v2dt2, const REAL* refl,
                         const REAL* sig, const int* srcPos, int nSrc,
                         const int* recPos, int nRec, const REAL* model, REAL* dreg, REAL* W)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx, nm = nzx * g->nExt;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    REAL* m = FN(zalloc)(nm);
    memcpy(m, model, sizeof(REAL) * (size_t)nm);
    FN(scale_refl_ext)(g, m, refl);
    FN(born_ctx) ctx = {m, W, v2dt2, refl};
    FN(scat_pass)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, FN(ext_ss), &ctx, 1, NULL);
    free(m);
}

/* BornOffsetShotsAdjGpu, F/src/BornExtShotsGpuFunctions.cu:1848-1989.
 * model (output, nExt*nz*nx) is overwritten with this shot's image. */
void FN(orc_bornext_adj)(const orc_grid* g, const REAL* v2dt2, const REAL* refl,
                         const REAL* sig, const int* srcPos, int nSrc,
                         const int* recPos, int nRec, REAL* model, const REAL* dreg, REAL* W)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    memset(model, 0, sizeof(REAL) * (size_t)nzx * g->nExt);
    FN(born_ctx) ctx = {model, W, v2dt2, refl};
    FN(rec_pass)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, FN(ext_img), &ctx, NULL, NULL);
    FN(scale_refl_ext)(g, model, refl);
}

/* ---------------------------------------------------------------------------
 * Time-lag extended Born (BornExtShotsGpu, extension=time).  Model slice e
 * is the image at time lag e - hExt (coarse samples); it scatters the source
 * wavefield 2 (e - hExt) coarse samples earlier.
 */

/* Slices e in [lo, hi) whose source sample k - 2 (e - hExt) lies in
 * [0, nts) (F/src/BornExtShotsGpuFunctions.cu:362-366, 374-378).  The
 * reference's (k + 1 - nts) / 2 truncates toward zero, which for the
 * non-positive numerator is the rounding up the bound needs. */
static void FN(time_lags)(const orc_grid* g, int k, int* lo, int* hi)
{
    const int hExt = (g->nExt - 1) / 2;
    int a = (k + 1 - g->nts) / 2;
    if (a < -hExt) a = -hExt;
    int b = k / 2;
    if (b > hExt) b = hExt;
    *lo = a + hExt;
    *hi = b + hExt + 1;
}

/* imagingTimeFwdGpu, F/src/kernelsGpu.cu:337-350: ss += m[e] * W[k - 2 (e - hExt)]. */
static void FN(time_ss)(const orc_grid* g, const void* vctx, int k, REAL* ss)
{
    const FN(born_ctx)* c = (const FN(born_ctx)*)vctx;
    const int hExt = (g->nExt - 1) / 2;
    const long nzx = (long)g->nz * g->nx;
    int lo, hi;
    FN(time_lags)(g, k, &lo, &hi);
    INTERIOR_BEGIN(g)
        for (int e = lo; e < hi; ++e)
            ss[i] += c->m[nzx * e + i] * c->W[nzx * (k - 2 * (e - hExt)) + i];
    INTERIOR_END
}

/* imagingTimeAdjGpu, F/src/kernelsGpu.cu:352-364: m[e] += ss * W[k - 2 (e - hExt)]. */
static void FN(time_img)(const orc_grid* g, void* vctx, int k, REAL* ss)
{
    FN(born_ctx)* c = (FN(born_ctx)*)vctx;
    const int hExt = (g->nExt - 1) / 2;
    const long nzx = (long)g->nz * g->nx;
    REAL* m = (REAL*)c->m;
    int lo, hi;
    FN(time_lags)(g, k, &lo, &hi);
    INTERIOR_BEGIN(g)
        for (int e = lo; e < hi; ++e)
            m[nzx * e + i] += ss[i] * c->W[nzx * (k - 2 * (e - hExt)) + i];
    INTERIOR_END
}

/* Slices e in [lo, hi) whose sample k + 2 (e - hExt) lies in [0, nts): the
 * bounds of the adjoint-side time-lag kernels
 * (F/src/tomoExtShotsGpuFunctions.cu:2630-2633, 2638-2641). */
static void FN(time_lags_rev)(const orc_grid* g, int k, int* lo, int* hi)
{
    const int hExt = (g->nExt - 1) / 2;
    int a = (-k) / 2;
    if (a < -hExt) a = -hExt;
    int b = (g->nts - 1 - k) / 2;
    if (b > hExt) b = hExt;
    *lo = a + hExt;
    *hi = b + hExt + 1;
}

/* imagingTimeTomoAdjGpu (F/src/kernelsGpu.cu:445-458): secondary source of
 * the backward leg from the receiver wavefield R at coarse sample k,
 * ss += r[e] * R[k + 2 (e - hExt)] into a zeroed slice. */
static void FN(time_adj_ss)(const orc_grid* g, const REAL* R, const REAL* r,
                            const REAL* v2dt2, int k, REAL* ss)
{
    (void)v2dt2;
    const int hExt = (g->nExt - 1) / 2;
    const long nzx = (long)g->nz * g->nx;
    int lo, hi;
    FN(time_lags_rev)(g, k, &lo, &hi);
    INTERIOR_BEGIN(g)
        for (int e = lo; e < hi; ++e)
            ss[i] += r[nzx * e + i] * R[nzx * (k + 2 * (e - hExt)) + i];
    INTERIOR_END
}

/* imagingWemvaTimeAdjGpu (F/src/kernelsGpu.cu:366-378):
 * img[e] += scat * R[k + 2 (e - hExt)]. */
static void FN(wemva_time_img)(const orc_grid* g, REAL* img, const REAL* scat, const REAL* R, int k)
{
    const int hExt = (g->nExt - 1) / 2;
    const long nzx = (long)g->nz * g->nx;
    int lo, hi;
    FN(time_lags_rev)(g, k, &lo, &hi);
    INTERIOR_BEGIN(g)
        for (int e = lo; e < hi; ++e)
            img[nzx * e + i] += scat[i] * R[nzx * (k + 2 * (e - hExt)) + i];
    INTERIOR_END
}

/* scaleReflectivityExt (F/src/kernelsGpu.cu:277-288): m *= 2/v^3 * v^2 dtw^2;
 * sloth != 0: scaleReflectivitySlothExt (:290-302), m *= -v^2 dtw^2.  The
 * free-surface routines always take the first (BornTimeShotsFwdFsGpu and
 * BornTimeShotsAdjFsGpu have no sloth argument,
 * F/src/BornExtShotsGpuFunctions.cu:885, 2152). */
static void FN(scale_refl_time)(const orc_grid* g, REAL* m, const REAL* refl, const REAL* v2dt2, int sloth)
{
    const long nzx = (long)g->nz * g->nx;
    if (g->freeSurface) sloth = 0;
    for (int e = 0; e < g->nExt; ++e) {
        REAL* me = m + nzx * e;
        if (sloth) {
            INTERIOR_BEGIN(g)
                me[i] *= -v2dt2[i];
            INTERIOR_END
        } else {
            INTERIOR_BEGIN(g)
                me[i] *= refl[i] * v2dt2[i];
            INTERIOR_END
        }
    }
}

/* BornTimeShotsFwdGpu, F/src/BornExtShotsGpuFunctions.cu:269-418 (free
 * surface: BornTimeShotsFwdFsGpu, :885-1036).  model: nExt*nz*nx (read only). */
void FN(orc_borntime_fwd)(const orc_grid* g, const REAL* v2dt2, const REAL* refl, int sloth,
                          const REAL* sig, const int* srcPos, int nSrc,
                          const int* recPos, int nRec, const REAL* model, REAL* dreg, REAL* W)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx, nm = nzx * g->nExt;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    REAL* m = FN(zalloc)(nm);
    memcpy(m, model, sizeof(REAL) * (size_t)nm);
    FN(scale_refl_time)(g, m, refl, v2dt2, sloth);
    FN(born_ctx) ctx = {m, W, v2dt2, refl};
    FN(scat_pass)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, FN(time_ss), &ctx, 1, NULL);
    free(m);
}

/* BornTimeShotsAdjGpu, F/src/BornExtShotsGpuFunctions.cu:1525-1678 (free
 * surface: BornTimeShotsAdjFsGpu, :2152-2304).  model (output, nExt*nz*nx)
 * is overwritten with this shot's image. */
void FN(orc_borntime_adj)(const orc_grid* g, const REAL* v2dt2, const REAL* refl, int sloth,
                          const REAL* sig, const int* srcPos, int nSrc,
                          const int* recPos, int nRec, REAL* model, const REAL* dreg, REAL* W)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    memset(model, 0, sizeof(REAL) * (size_t)nzx * g->nExt);
    FN(born_ctx) ctx = {model, W, v2dt2, refl};
    FN(rec_pass)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, FN(time_img), &ctx, NULL, NULL);
    FN(scale_refl_time)(g, model, refl, v2dt2, sloth);
}

/* ---------------------------------------------------------------------------
 * Offset-extended tomographic operator (tomoExtShotsGpu, extension=offset).
 */

/* computeTomoSrcWfldDt2 (F/src/tomoExtShotsGpuFunctions.cu:1251-1360) with
 * srcWfldSecondTimeDerivative (F/src/kernelsGpu.cu:188-197): the coarse
 * source wavefield W from src_pass is replaced in place by
 * cSide*(W[k-1] + W[k+1]) + cCenter*W[k], zero beyond both ends, where
 * cSide = 1/dts^2 and cCenter = -2/dts^2 are rounded to REAL
 * (F/src/tomoExtShotsGpuFunctions.cu:138-139). */
static void FN(dt2_coarse)(const orc_grid* g, REAL* W)
{
    const int nts = g->nts;
    const long nzx = (long)g->nz * g->nx;
    const REAL dts = (REAL)g->dts;
    const REAL d2 = dts * dts;
    const REAL cs = (REAL)(1.0 / (double)d2), cc = (REAL)(-2.0 / (double)d2);
    INTERIOR_BEGIN(g)
        REAL prev = 0, cur = W[i];
        for (int k = 0; k < nts; ++k) {
            const REAL next = k + 1 < nts ? W[(long)(k + 1) * nzx + i] : (REAL)0;
            W[(long)k * nzx + i] = cs * (prev + next) + cc * cur;
            prev = cur;
            cur = next;
        }
    INTERIOR_END
}

/* Scattered wavefield captured at the coarse rate: the first half of
 * computeTomoLeg1OffsetFwd and computeTomoLeg2OffsetFwd
 * (F/src/tomoExtShotsGpuFunctions.cu:2005-2062, 2227-2261).  P (nts slices)
 * is zeroed, then per fine step: step -> injectSecondarySource -> damp ->
 * interpWavefield. */
static void FN(cap_pass)(const orc_grid* g, const REAL* zc, const REAL* xc,
                         const REAL* f, const REAL* cd, const REAL* v2dt2,
                         FN(ss_fn) make_ss, const void* ctx, REAL* P)
{
    const int nts = g->nts, sub = g->sub, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    REAL* ssL = FN(zalloc)(nzx);
    REAL* ssR = FN(zalloc)(nzx);
    memset(P, 0, sizeof(REAL) * (size_t)nzx * nts);
    make_ss(g, ctx, 0, ssL);
    for (int its = 0; its < nts - 1; ++its) {
        make_ss(g, ctx, its + 1, ssR);
        for (int it2 = 1; it2 < sub + 1; ++it2) {
            FN(orc_step_fwd)(g, zc, xc, p0, p1, p0, v2dt2);
            FN(inject_secondary)(g, f, h, ssL, ssR, p0, it2 - 1);
            FN(orc_damp)(g, cd, p0, p1);
            FN(capture2)(g, f, h, p0, P + (long)its * nzx, P + (long)(its + 1) * nzx, it2);
            FN(swap)(&p0, &p1);
        }
        FN(swap)(&ssL, &ssR);
        memset(ssR, 0, sizeof(REAL) * (size_t)nzx);
    }
    free(p0); free(p1); free(ssL); free(ssR);
}

/* computeTomoRecWfld (F/src/tomoExtShotsGpuFunctions.cu:1481-1515): the
 * receiver wavefield R (nts slices) of the adjoint propagation of the data,
 * captured unscaled at the coarse rate. */
static void FN(rec_wfld)(const orc_grid* g, const REAL* zc, const REAL* xc,
                         const REAL* f, const REAL* cd, const REAL* v2dt2,
                         const int* recPos, int nRec, const REAL* dreg, REAL* R)
{
    const int nts = g->nts, sub = g->sub, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    REAL* tmp = FN(zalloc)(nzx);
    memset(R, 0, sizeof(REAL) * (size_t)nzx * nts);
    for (int its = nts - 2; its > -1; --its) {
        for (int it2 = sub - 1; it2 > -1; --it2) {
            FN(orc_step_adj)(g, zc, xc, p0, p1, p0, v2dt2, tmp);
            FN(interp_inject)(nts, h, f, dreg, p0, its, it2, recPos, nRec);
            FN(orc_damp)(g, cd, p0, p1);
            FN(capture2)(g, f, h, p0, R + (long)its * nzx, R + (long)(its + 1) * nzx, it2);
            FN(swap)(&p0, &p1);
        }
    }
    free(p0); free(p1); free(tmp);
}

typedef void (*FN(adj_ss_fn))(const orc_grid*, const REAL*, const REAL*, const REAL*, int, REAL*);

/* imagingOffsetTomoAdjGpu (F/src/kernelsGpu.cu:460-475): secondary source of
 * leg 2 of the adjoint from the receiver wavefield R at coarse sample k,
 * ss[x] += r[e][x+e] * R[k][x+2e] * v2dt2[x+2e] into a zeroed slice. */
static void FN(tomo_adj_ss)(const orc_grid* g, const REAL* R, const REAL* r,
                            const REAL* v2dt2, int k, REAL* ss)
{
    const int nz = g->nz, nx = g->nx, hExt = (g->nExt - 1) / 2;
    const long nzx = (long)nz * nx;
    INTERIOR_BEGIN(g)
        int lo = -hExt, t = FAT + hExt - ix;
        if (t > lo) lo = t;
        int hi = nx - 1 - FAT - ix - hExt;
        if (hExt < hi) hi = hExt;
        for (int e = lo; e < hi + 1; ++e) {
            long iModel = nzx * (e + hExt) + (long)nz * (ix + e) + iz;
            long iRec = (long)(ix + 2 * e) * nz + iz;
            ss[i] += r[iModel] * R[nzx * k + iRec] * v2dt2[iRec];
        }
    INTERIOR_END
}

/* imagingAdjGpu (F/src/kernelsGpu.cu:318-325): m += X[k] * slice. */
static void FN(img_acc)(const orc_grid* g, REAL* m, const REAL* X, int k, const REAL* slice)
{
    const long nzx = (long)g->nz * g->nx;
    const REAL* Xk = X + (long)k * nzx;
    INTERIOR_BEGIN(g)
        m[i] += Xk[i] * slice[i];
    INTERIOR_END
}

/* The scaled extended reflectivity r of the tomographic and WEMVA
 * operators: scaleReflectivityLinExt for offsets, scaleReflectivityExt (or
 * with sloth scaleReflectivitySlothExt) for time lags. */
static void FN(scale_ext_any)(const orc_grid* g, REAL* r, const REAL* refl, const REAL* v2dt2,
                              int timeLag, int sloth)
{
    if (timeLag) FN(scale_refl_time)(g, r, refl, v2dt2, sloth);
    else FN(scale_refl_ext)(g, r, refl);
}

/* tomoOffsetShotsFwdGpu (F/src/tomoExtShotsGpuFunctions.cu:467-572) with
 * computeTomoLeg1OffsetFwd (:2005-2113) and computeTomoLeg2OffsetFwd
 * (:2227-2317); timeLag != 0: tomoTimeShotsFwdGpu (:355-466) with
 * computeTomoLeg1TimeFwd (:1554-1657) and computeTomoLeg2TimeFwd
 * (:1782-1882), the same legs with the time-lag scattering (free surface:
 * tomoTimeShotsFwdFsGpu, :576-683, no sloth).  model: nz*nx velocity
 * perturbation; extRefl: nExt*nz*nx extended reflectivity (both read only);
 * W, P: nts*nz*nx scratch.
 *   leg 1: s -> extended reflectivity -> P -> model -> data
 *   leg 2: s -> model -> P -> extended reflectivity -> data */
void FN(orc_tomo_fwd)(const orc_grid* g, const REAL* v2dt2, const REAL* refl,
                      const REAL* sig, const int* srcPos, int nSrc,
                      const int* recPos, int nRec, const REAL* model, const REAL* extRefl,
                      int leg1, int leg2, int timeLag, int sloth, REAL* dreg, REAL* W, REAL* P)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx, nr = nzx * g->nExt;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    FN(dt2_coarse)(g, W);
    REAL* m = FN(zalloc)(nzx);
    memcpy(m, model, sizeof(REAL) * (size_t)nzx);
    FN(scale_refl)(g, m, refl, v2dt2);
    REAL* r = FN(zalloc)(nr);
    memcpy(r, extRefl, sizeof(REAL) * (size_t)nr);
    FN(scale_ext_any)(g, r, refl, v2dt2, timeLag, sloth);
    memset(dreg, 0, sizeof(REAL) * (size_t)nRec * g->nts);
    FN(ss_fn) extSS = timeLag ? FN(time_ss) : FN(ext_ss);
    FN(born_ctx) rW = {r, W, v2dt2, refl}, mW = {m, W, v2dt2, refl};
    FN(born_ctx) rP = {r, P, v2dt2, refl}, mP = {m, P, v2dt2, refl};
    if (leg1) {
        FN(cap_pass)(g, zc, xc, f, cd, v2dt2, extSS, &rW, P);
        FN(scat_pass)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, FN(born_ss), &mP, 0, NULL);
    }
    if (leg2) {
        FN(cap_pass)(g, zc, xc, f, cd, v2dt2, FN(born_ss), &mW, P);
        FN(scat_pass)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, extSS, &rP, 0, NULL);
    }
    free(m); free(r);
}

/* Forward in time from the extended scattering of W (ext_ss with the
 * scaled extended reflectivity r), each captured coarse sample k imaged as
 * model += X[k] * sample: leg 1 of the tomographic adjoint
 * (computeTomoLeg1OffsetAdj, F/src/tomoExtShotsGpuFunctions.cu:2788-2860)
 * and leg 2 of the WEMVA adjoint (computeWemvaLeg2OffsetAdj,
 * F/src/wemvaExtShotsGpuFunctions.cu:1502-1571). */
static void FN(adj_leg_fwd)(const orc_grid* g, const REAL* zc, const REAL* xc, const REAL* f,
                            const REAL* cd, const REAL* v2dt2, const REAL* r, const REAL* W,
                            const REAL* X, REAL* model, FN(ss_fn) make_ss)
{
    const int nts = g->nts, sub = g->sub, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    REAL* ssL = FN(zalloc)(nzx);
    REAL* ssR = FN(zalloc)(nzx);
    REAL* scL = FN(zalloc)(nzx);
    REAL* scR = FN(zalloc)(nzx);
    FN(born_ctx) rW = {r, W, v2dt2, NULL};
    make_ss(g, &rW, 0, ssL);
    for (int its = 0; its < nts - 1; ++its) {
        make_ss(g, &rW, its + 1, ssR);
        for (int it2 = 1; it2 < sub + 1; ++it2) {
            FN(orc_step_fwd)(g, zc, xc, p0, p1, p0, v2dt2);
            FN(inject_secondary)(g, f, h, ssL, ssR, p0, it2 - 1);
            FN(orc_damp)(g, cd, p0, p1);
            FN(capture2)(g, f, h, p0, scL, scR, it2);
            FN(swap)(&p0, &p1);
        }
        FN(img_acc)(g, model, X, its, scL);
        FN(swap)(&ssL, &ssR);
        memset(ssR, 0, sizeof(REAL) * (size_t)nzx);
        FN(swap)(&scL, &scR);
        memset(scR, 0, sizeof(REAL) * (size_t)nzx);
    }
    FN(img_acc)(g, model, X, nts - 1, scL);
    free(p0); free(p1); free(ssL); free(ssR); free(scL); free(scR);
}

/* Backward in time from the adjoint extended scattering of the receiver
 * wavefield R (tomo_adj_ss with r), each captured coarse sample k imaged as
 * model += W[k] * sample: leg 2 of the tomographic adjoint
 * (computeTomoLeg2OffsetAdj, F/src/tomoExtShotsGpuFunctions.cu:2949-3020)
 * and leg 1 of the WEMVA adjoint (computeWemvaLeg1OffsetAdj,
 * F/src/wemvaExtShotsGpuFunctions.cu:1428-1500). */
static void FN(adj_leg_bwd)(const orc_grid* g, const REAL* zc, const REAL* xc, const REAL* f,
                            const REAL* cd, const REAL* v2dt2, const REAL* r, const REAL* W,
                            const REAL* R, REAL* model, FN(adj_ss_fn) make_ss)
{
    const int nts = g->nts, sub = g->sub, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    REAL* ssL = FN(zalloc)(nzx);
    REAL* ssR = FN(zalloc)(nzx);
    REAL* scL = FN(zalloc)(nzx);
    REAL* scR = FN(zalloc)(nzx);
    REAL* tmp = FN(zalloc)(nzx);
    make_ss(g, R, r, v2dt2, nts - 1, ssR);
    for (int its = nts - 2; its > -1; --its) {
        make_ss(g, R, r, v2dt2, its, ssL);
        for (int it2 = sub - 1; it2 > -1; --it2) {
            FN(orc_step_adj)(g, zc, xc, p0, p1, p0, v2dt2, tmp);
            FN(inject_secondary)(g, f, h, ssL, ssR, p0, it2 + 1);
            FN(orc_damp)(g, cd, p0, p1);
            FN(capture2)(g, f, h, p0, scL, scR, it2);
            FN(swap)(&p0, &p1);
        }
        FN(img_acc)(g, model, W, its + 1, scR);
        FN(swap)(&ssR, &ssL);
        memset(ssL, 0, sizeof(REAL) * (size_t)nzx);
        FN(swap)(&scR, &scL);
        memset(scL, 0, sizeof(REAL) * (size_t)nzx);
    }
    FN(img_acc)(g, model, W, 0, scR);
    free(p0); free(p1); free(ssL); free(ssR); free(scL); free(scR); free(tmp);
}

/* tomoOffsetShotsAdjGpu (F/src/tomoExtShotsGpuFunctions.cu:912-1020) with
 * computeTomoLeg1OffsetAdj (:2788-2860) and computeTomoLeg2OffsetAdj
 * (:2949-3020); timeLag != 0: tomoTimeShotsAdjGpu (:797-911) with
 * computeTomoLeg1TimeAdj (:2442-2528) and computeTomoLeg2TimeAdj
 * (:2619-2702).  model (output, nz*nx) is overwritten with this shot's
 * contribution; W holds the source wavefield's second time derivative and P
 * the receiver wavefield. */
void FN(orc_tomo_adj)(const orc_grid* g, const REAL* v2dt2, const REAL* refl,
                      const REAL* sig, const int* srcPos, int nSrc,
                      const int* recPos, int nRec, REAL* model, const REAL* dreg,
                      const REAL* extRefl, int leg1, int leg2, int timeLag, int sloth, REAL* W, REAL* P)
{
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx, nr = nzx * g->nExt;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    FN(dt2_coarse)(g, W);
    FN(rec_wfld)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dreg, P);
    REAL* r = FN(zalloc)(nr);
    memcpy(r, extRefl, sizeof(REAL) * (size_t)nr);
    FN(scale_ext_any)(g, r, refl, v2dt2, timeLag, sloth);
    memset(model, 0, sizeof(REAL) * (size_t)nzx);
    if (leg1) FN(adj_leg_fwd)(g, zc, xc, f, cd, v2dt2, r, W, P, model, timeLag ? FN(time_ss) : FN(ext_ss));
    if (leg2) FN(adj_leg_bwd)(g, zc, xc, f, cd, v2dt2, r, W, P, model, timeLag ? FN(time_adj_ss) : FN(tomo_adj_ss));
    FN(scale_refl)(g, model, refl, v2dt2);
    free(r);
}

/* ---------------------------------------------------------------------------
 * Offset-extended WEMVA operator (wemvaExtShotsGpu, extension=offset):
 * velocity perturbation <-> extended image, about the receiver wavefield of
 * the observed data dobs.
 */

/* imagingOffsetWemvaScaleFwdGpu (F/src/kernelsGpu.cu:429-443):
 * img[e][x] += scat[x-h] * R[k][x+h] * v2dt2[x+h], h = e - hExt. */
static void FN(wemva_img_fwd)(const orc_grid* g, REAL* img, const REAL* scat, const REAL* R, int k,
                              const REAL* v2dt2)
{
    const int nz = g->nz, nx = g->nx, nExt = g->nExt, hExt = (nExt - 1) / 2;
    const long nzx = (long)nz * nx;
#pragma omp parallel for schedule(static)
    for (int ix = FAT + hExt; ix < nx - FAT - hExt; ++ix)
        for (int e = 0; e < nExt; ++e)
            for (int iz = FAT; iz < nz - FAT; ++iz) {
                int sh = e - hExt;
                long iRec = (long)(ix + sh) * nz + iz;
                img[nzx * e + (long)ix * nz + iz] += scat[(long)(ix - sh) * nz + iz] * R[nzx * k + iRec] * v2dt2[iRec];
            }
}

/* wemvaOffsetShotsFwdGpu (F/src/wemvaExtShotsGpuFunctions.cu:458-574) with
 * computeWemvaLeg1OffsetFwd (:1123-1188) and computeWemvaLeg2OffsetFwd
 * (:1190-1264); timeLag != 0: wemvaTimeShotsFwdGpu (:344-457) with
 * computeWemvaLeg1TimeFwd (:961-1036) and computeWemvaLeg2TimeFwd
 * (:1038-1121).  The reference's WEMVA has no free-surface routines and
 * propagates with the absorbing top whatever the freeSurface key
 * (F/src/wemvaExtGpu.cpp:131-135).  model: nz*nx velocity perturbation
 * (read only); dobs: the shot's observed data on the regular receiver grid;
 * image (output, nExt*nz*nx) is overwritten with this shot's extended image. */
void FN(orc_wemva_fwd)(const orc_grid* gIn, const REAL* v2dt2, const REAL* refl,
                       const REAL* sig, const int* srcPos, int nSrc,
                       const int* recPos, int nRec, const REAL* model, const REAL* dobs,
                       int leg1, int leg2, int timeLag, REAL* image, REAL* W, REAL* P)
{
    orc_grid g0 = *gIn;
    g0.freeSurface = 0;
    const orc_grid* g = &g0;
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const int nts = g->nts, sub = g->sub, h = sub + 1;
    const long nzx = (long)g->nz * g->nx;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    FN(dt2_coarse)(g, W);
    FN(rec_wfld)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dobs, P);
    REAL* m = FN(zalloc)(nzx);
    memcpy(m, model, sizeof(REAL) * (size_t)nzx);
    FN(scale_refl)(g, m, refl, v2dt2);
    memset(image, 0, sizeof(REAL) * (size_t)nzx * g->nExt);
    REAL* p0 = FN(zalloc)(nzx);
    REAL* p1 = FN(zalloc)(nzx);
    REAL* ssL = FN(zalloc)(nzx);
    REAL* ssR = FN(zalloc)(nzx);
    REAL* scL = FN(zalloc)(nzx);
    REAL* scR = FN(zalloc)(nzx);
    REAL* tmp = FN(zalloc)(nzx);
    if (leg1) {   /* forward from m * W'', imaged against R */
        FN(born_ctx) mW = {m, W, v2dt2, refl};
        FN(born_ss)(g, &mW, 0, ssL);
        for (int its = 0; its < nts - 1; ++its) {
            FN(born_ss)(g, &mW, its + 1, ssR);
            for (int it2 = 1; it2 < sub + 1; ++it2) {
                FN(orc_step_fwd)(g, zc, xc, p0, p1, p0, v2dt2);
                FN(inject_secondary)(g, f, h, ssL, ssR, p0, it2 - 1);
                FN(orc_damp)(g, cd, p0, p1);
                FN(capture2)(g, f, h, p0, scL, scR, it2);
                FN(swap)(&p0, &p1);
            }
            if (timeLag) FN(wemva_time_img)(g, image, scL, P, its);
            else FN(wemva_img_fwd)(g, image, scL, P, its, v2dt2);
            FN(swap)(&ssL, &ssR);
            FN(swap)(&scL, &scR);
            memset(scR, 0, sizeof(REAL) * (size_t)nzx);
        }
        if (timeLag) FN(wemva_time_img)(g, image, scL, P, nts - 1);
        else FN(wemva_img_fwd)(g, image, scL, P, nts - 1, v2dt2);
    }
    if (leg2) {   /* backward from m * R, imaged against W'' */
        memset(p0, 0, sizeof(REAL) * (size_t)nzx);
        memset(p1, 0, sizeof(REAL) * (size_t)nzx);
        memset(ssL, 0, sizeof(REAL) * (size_t)nzx);
        memset(ssR, 0, sizeof(REAL) * (size_t)nzx);
        memset(scL, 0, sizeof(REAL) * (size_t)nzx);
        memset(scR, 0, sizeof(REAL) * (size_t)nzx);
        FN(born_ctx) mR = {m, P, v2dt2, refl};
        FN(born_ctx) iW = {image, W, v2dt2, refl};
        FN(born_ss)(g, &mR, nts - 1, ssR);
        for (int its = nts - 2; its > -1; --its) {
            FN(born_ss)(g, &mR, its, ssL);
            for (int it2 = sub - 1; it2 > -1; --it2) {
                FN(orc_step_adj)(g, zc, xc, p0, p1, p0, v2dt2, tmp);
                FN(inject_secondary)(g, f, h, ssL, ssR, p0, it2 + 1);
                FN(orc_damp)(g, cd, p0, p1);
                FN(capture2)(g, f, h, p0, scL, scR, it2);
                FN(swap)(&p0, &p1);
            }
            /* offsets: scR *= v2dt2, then the offset imaging */
            (timeLag ? FN(time_img) : FN(ext_img))(g, &iW, its + 1, scR);
            FN(swap)(&ssR, &ssL);
            FN(swap)(&scR, &scL);
            memset(scL, 0, sizeof(REAL) * (size_t)nzx);
        }
        (timeLag ? FN(time_img) : FN(ext_img))(g, &iW, 0, scR);
    }
    FN(scale_ext_any)(g, image, refl, v2dt2, timeLag, 0);
    free(m); free(p0); free(p1); free(ssL); free(ssR); free(scL); free(scR); free(tmp);
}

/* wemvaOffsetShotsAdjGpu (F/src/wemvaExtShotsGpuFunctions.cu:689-808);
 * timeLag != 0: wemvaTimeShotsAdjGpu (:575-688) with computeWemvaLeg1TimeAdj
 * (:1270-1350) and computeWemvaLeg2TimeAdj (:1352-1426); absorbing top
 * whatever the freeSurface key, as in the forward.  image (read only,
 * nExt*nz*nx) -> model (output, nz*nx), overwritten with this shot's
 * contribution.  Leg 1 is the backward pass, leg 2 the forward one, in that
 * order. */
void FN(orc_wemva_adj)(const orc_grid* gIn, const REAL* v2dt2, const REAL* refl,
                       const REAL* sig, const int* srcPos, int nSrc,
                       const int* recPos, int nRec, REAL* model, const REAL* image, const REAL* dobs,
                       int leg1, int leg2, int timeLag, REAL* W, REAL* P)
{
    orc_grid g0 = *gIn;
    g0.freeSurface = 0;
    const orc_grid* g = &g0;
    REAL zc[6], xc[6], f[2 * 101], cd[512];
    FN(orc_tables)(g, zc, xc, f, cd);
    const long nzx = (long)g->nz * g->nx, nr = nzx * g->nExt;
    FN(src_pass)(g, zc, xc, f, cd, v2dt2, sig, srcPos, nSrc, W);
    FN(dt2_coarse)(g, W);
    FN(rec_wfld)(g, zc, xc, f, cd, v2dt2, recPos, nRec, dobs, P);
    REAL* r = FN(zalloc)(nr);
    memcpy(r, image, sizeof(REAL) * (size_t)nr);
    FN(scale_ext_any)(g, r, refl, v2dt2, timeLag, 0);
    memset(model, 0, sizeof(REAL) * (size_t)nzx);
    if (leg1) FN(adj_leg_bwd)(g, zc, xc, f, cd, v2dt2, r, W, P, model, timeLag ? FN(time_adj_ss) : FN(tomo_adj_ss));
    if (leg2) FN(adj_leg_fwd)(g, zc, xc, f, cd, v2dt2, r, W, P, model, timeLag ? FN(time_ss) : FN(ext_ss));
    FN(scale_refl)(g, model, refl, v2dt2);
    free(r);
}

/* Test hook for dt2_coarse: W (nts*nz*nx) in place. */
void FN(orc_dt2_coarse)(const orc_grid* g, REAL* W) { FN(dt2_coarse)(g, W); }

/* interpTimeLinTbb::forward, F/src/interpTimeLinTbb.cpp:9-62.
 * m: nDev x nts (coarse), d: nDev x ntw (fine). */
void FN(orc_interp_fwd)(int nts, double dtsIn, double otsIn, int sub, int nDev,
                        int add, const REAL* m, REAL* d)
{
    const int ntw = (nts - 1) * sub + 1;
    const REAL dts = (REAL)dtsIn, ots = (REAL)otsIn, dtw = dts / sub, otw = ots;
    REAL scale = (REAL)ntw / (REAL)nts;
    scale = (REAL)(1.0 / sqrt((double)scale));
    if (!add) memset(d, 0, sizeof(REAL) * (size_t)nDev * ntw);
    for (int dv = 0; dv < nDev; ++dv) {
        const REAL* md = m + (long)dv * nts;
        REAL* dd = d + (long)dv * ntw;
        for (int itw = 0; itw < ntw - 1; ++itw) {
            REAL tw = otw + itw * dtw;
            REAL w = (tw - ots) / dts;
            int k = (int)w;
            w = w - k;
            double v = ((double)md[k] * (1.0 - (double)w) + (double)(md[k + 1] * w)) * (double)scale;
            dd[itw] = (REAL)((double)dd[itw] + v);
        }
        dd[ntw - 1] += scale * md[nts - 1];
    }
}

/* interpTimeLinTbb::adjoint, F/src/interpTimeLinTbb.cpp:64-102. */
void FN(orc_interp_adj)(int nts, double dtsIn, double otsIn, int sub, int nDev,
                        int add, REAL* m, const REAL* d)
{
    const int ntw = (nts - 1) * sub + 1;
    const REAL dts = (REAL)dtsIn, ots = (REAL)otsIn, dtw = dts / sub, otw = ots;
    REAL scale = (REAL)ntw / (REAL)nts;
    scale = (REAL)(1.0 / sqrt((double)scale));
    if (!add) memset(m, 0, sizeof(REAL) * (size_t)nDev * nts);
    for (int dv = 0; dv < nDev; ++dv) {
        REAL* md = m + (long)dv * nts;
        const REAL* dd = d + (long)dv * ntw;
        for (int itw = 0; itw < ntw - 1; ++itw) {
            REAL tw = otw + itw * dtw;
            REAL w = (tw - ots) / dts;
            int k = (int)w;
            w = w - k;
            md[k] = (REAL)((double)md[k] + (double)(scale * dd[itw]) * (1.0 - (double)w));
            md[k + 1] += scale * dd[itw] * w;
        }
        md[nts - 1] += scale * dd[ntw - 1];
    }
}

/* secondTimeDerivative::forward, F/src/secondTimeDerivative.cpp:9-28. */
void FN(orc_d2t_fwd)(int nt, double dt, int nDev, int add, const REAL* m, REAL* d)
{
    REAL dtf = (REAL)dt;
    REAL dt2 = (REAL)(1.0 / (double)(dtf * dtf));
    if (!add) memset(d, 0, sizeof(REAL) * (size_t)nDev * nt);
    for (int dv = 0; dv < nDev; ++dv) {
        const REAL* x = m + (long)dv * nt;
        REAL* y = d + (long)dv * nt;
        y[0] = (REAL)((double)y[0] + (double)dt2 * ((double)x[1] - 2.0 * (double)x[0]));
        for (int it = 1; it < nt - 1; ++it)
            y[it] = (REAL)((double)y[it] + (double)dt2 * ((double)x[it + 1] - 2.0 * (double)x[it] + (double)x[it - 1]));
        y[nt - 1] = (REAL)((double)y[nt - 1] + (double)dt2 * ((double)x[nt - 2] - 2.0 * (double)x[nt - 1]));
    }
}

/* secondTimeDerivative::adjoint, F/src/secondTimeDerivative.cpp:30-38
 * (the operator is symmetric). */
void FN(orc_d2t_adj)(int nt, double dt, int nDev, int add, REAL* m, const REAL* d)
{
    FN(orc_d2t_fwd)(nt, dt, nDev, add, d, m);
}

/* seismicOperator2D::scaleSeismicSource, F/include/seismicOperator2D.cpp:73-86:
 * sig[p][it] *= dtw*dtw*v[pos]*v[pos]. */
void FN(orc_scale_source)(const orc_grid* g, const REAL* vel, const int* pos, int nPos,
                          int nt, REAL* sig)
{
    REAL dtw = (REAL)g->dts / (REAL)g->sub;
    for (int p = 0; p < nPos; ++p) {
        REAL scale = dtw * dtw * vel[pos[p]] * vel[pos[p]];
        for (int it = 0; it < nt; ++it) sig[(long)p * nt + it] = sig[(long)p * nt + it] * scale;
    }
}

/* deviceGpu constructor #1 (irregular coordinates, bilinear corners),
 * F/src/deviceGpu.cpp:8-84; bounds check :263-274.  Returns 0 or -1 when a
 * device is out of bounds. */
int FN(orc_dev_coords)(int nDev, const REAL* zCoord, const REAL* xCoord,
                       int nzVel, int nxVel, double ozIn, double dzIn, double oxIn,
                       double dxIn, int dipole, double zShiftIn, double xShiftIn,
                       int* gridIdx, REAL* weight)
{
    const REAL oz = (REAL)ozIn, dz = (REAL)dzIn, ox = (REAL)oxIn, dx = (REAL)dxIn;
    const REAL zShift = (REAL)zShiftIn, xShift = (REAL)xShiftIn;
    const int nb = dipole == 1 ? 8 : 4;
    REAL zMax = oz + (nzVel - 1) * dz, xMax = ox + (nxVel - 1) * dx;
    for (int i = 0; i < nDev; ++i)
        if (zCoord[i] >= zMax || xCoord[i] >= xMax) return -1;
    for (int i = 0; i < nDev; ++i) {
        int i1 = i * nb;
        REAL wz = (zCoord[i] - oz) / dz, wx = (xCoord[i] - ox) / dx;
        int zReg = (int)wz;
        wz = wz - zReg;
        wz = (REAL)(1.0 - (double)wz);
        int xReg = (int)wx;
        wx = wx - xReg;
        wx = (REAL)(1.0 - (double)wx);
        gridIdx[i1] = xReg * nzVel + zReg;
        weight[i1] = wz * wx;
        gridIdx[i1 + 1] = gridIdx[i1] + 1;
        weight[i1 + 1] = (REAL)((1.0 - (double)wz) * (double)wx);
        gridIdx[i1 + 2] = gridIdx[i1] + nzVel;
        weight[i1 + 2] = (REAL)((double)wz * (1.0 - (double)wx));
        gridIdx[i1 + 3] = gridIdx[i1] + nzVel + 1;
        weight[i1 + 3] = (REAL)((1.0 - (double)wz) * (1.0 - (double)wx));
        if (dipole == 1) {
            REAL wzD = (zCoord[i] + zShift - oz) / dz, wxD = (xCoord[i] + xShift - ox) / dx;
            int zRD = (int)wzD;
            wzD = wzD - zRD;
            wzD = (REAL)(1.0 - (double)wzD);
            int xRD = (int)wxD;
            wxD = wxD - xRD;
            wxD = (REAL)(1.0 - (double)wxD);
            gridIdx[i1 + 4] = xRD * nzVel + zRD;
            weight[i1 + 4] = (REAL)((-1.0) * (double)wzD * (double)wxD);
            gridIdx[i1 + 5] = gridIdx[i1 + 4] + 1;
            weight[i1 + 5] = (REAL)((-1.0) * (1.0 - (double)wzD) * (double)wxD);
            gridIdx[i1 + 6] = gridIdx[i1 + 4] + nzVel;
            weight[i1 + 6] = (REAL)((-1.0) * (double)wzD * (1.0 - (double)wxD));
            gridIdx[i1 + 7] = gridIdx[i1 + 4] + nzVel + 1;
            weight[i1 + 7] = (REAL)((-1.0) * (1.0 - (double)wzD) * (1.0 - (double)wxD));
        }
    }
    return 0;
}

/* deviceGpu::forward, F/src/deviceGpu.cpp:226-243 (regular -> irregular).
 * cornerReg[i*nbCorner+c] is the unique-point index of that corner. */
void FN(orc_dev_fwd)(int nDev, int nbCorner, const int* cornerReg, const REAL* weight,
                     int nt, int add, const REAL* mReg, REAL* dIrreg)
{
    if (!add) memset(dIrreg, 0, sizeof(REAL) * (size_t)nDev * nt);
    for (int i = 0; i < nDev; ++i)
        for (int c = 0; c < nbCorner; ++c) {
            int i1 = i * nbCorner + c, i2 = cornerReg[i1];
            for (int it = 0; it < nt; ++it)
                dIrreg[(long)i * nt + it] += weight[i1] * mReg[(long)i2 * nt + it];
        }
}

/* deviceGpu::adjoint, F/src/deviceGpu.cpp:245-261 (irregular -> regular). */
void FN(orc_dev_adj)(int nDev, int nbCorner, const int* cornerReg, const REAL* weight,
                     int nt, int add, REAL* mReg, const REAL* dIrreg)
{
    int nReg = 0;
    for (int k = 0; k < nDev * nbCorner; ++k)
        if (cornerReg[k] + 1 > nReg) nReg = cornerReg[k] + 1;
    if (!add) memset(mReg, 0, sizeof(REAL) * (size_t)nReg * nt);
    for (int i = 0; i < nDev; ++i)
        for (int c = 0; c < nbCorner; ++c) {
            int i1 = i * nbCorner + c, i2 = cornerReg[i1];
            for (int it = 0; it < nt; ++it)
                mReg[(long)i2 * nt + it] += weight[i1] * dIrreg[(long)i * nt + it];
        }
}

#undef FAT
#undef INTERIOR_BEGIN
#undef INTERIOR_END
#undef FN
#undef CAT
#undef CAT2
