/*
 * ctr_oracle.c - FP64 CPU restatement of CT-Research's projector pair,
 * FDK / fan-beam FBP and SIRT.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * kernels: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  The product library never calls it.
 *
 * Each function cites the reference file:line it restates.  Two families:
 *   *_ref    line-by-line restatements of the reference CUDA kernels,
 *            including their arithmetic quirks (selected by `quirks`);
 *   *_exact  the mathematically intended operator: exact parametric Siddon
 *            (merged plane crossings, voxel from the segment midpoint, as
 *            Projection_computation_matlab/Siddon3D.m:21-142 and
 *            GenSysMatCircular3D.m:72-93), and its exact transpose.
 * Inputs are float (as on the device); outputs are double.
 *
 * Pinning (tests/test_oracle.py): the exact Siddon against per-voxel
 * box clipping in long double and against closed-form chord lengths of the
 * volume box and of ellipsoids; the ref restatement against the exact
 * Siddon; the adjoint against <Ax,y> = <x,A'y>; filters against their
 * closed-form taps and DFT; FBP/FDK against analytic phantoms.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.141592653589793

typedef struct {
    int nt, ns, nz;
    double size[3];
    double D;
    int nb, np, nx;
    const float *beta, *p, *xi;
} or_cone;

typedef struct {
    int nt, ns;
    double size[2];
    double D;
    int nb, np;
    const double *beta, *p;
} or_fan;

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------
 * Exact Siddon: merged plane crossings, voxel of each segment from its
 * midpoint (Siddon 1985; Siddon3D.m:131-142 merges the t/s/z crossing
 * parameters with sortrows, GenSysMatCircular3D.m:72-93 assigns each
 * segment to a voxel).  Works for 2-D (n[2] == 1 and d[2] == 0) too.
 * Returns the number of segments written to idx/len.
 * ---------------------------------------------------------------------- */
static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static int siddon_exact(const double S[3], const double d[3], const int n[3], const double r[3], int dims,
                        long *idx, double *len, double *work)
{
    double lo = -INFINITY, hi = INFINITY;
    for (int a = 0; a < dims; a++) {
        double bmax = n[a] * r[a];
        if (d[a] == 0.0) {
            /* a ray inside a boundary plane belongs to the voxel above it (the
             * reference's floor-of-midpoint rule) */
            if (!(S[a] >= 0.0 && S[a] < bmax)) return 0;
            continue;
        }
        double t1 = (0.0 - S[a]) / d[a], t2 = (bmax - S[a]) / d[a];
        lo = dmax(lo, dmin(t1, t2));
        hi = dmin(hi, dmax(t1, t2));
    }
    if (!(hi > lo)) return 0;
    int m = 0;
    work[m++] = lo;
    for (int a = 0; a < dims; a++) {
        if (d[a] == 0.0) continue;
        for (int k = 1; k < n[a]; k++) {
            double t = (k * r[a] - S[a]) / d[a];
            if (t > lo && t < hi) work[m++] = t;
        }
    }
    work[m++] = hi;
    qsort(work, m, sizeof(double), cmp_double);
    int nseg = 0;
    for (int q = 0; q + 1 < m; q++) {
        double t0 = work[q], t1 = work[q + 1];
        if (!(t1 > t0)) continue;
        double tm = 0.5 * (t0 + t1);
        long id = 0, stride = 1;
        int ok = 1;
        for (int a = 0; a < dims; a++) {
            long i = (long)floor((S[a] + tm * d[a]) / r[a]);
            if (i < 0 || i >= n[a]) { ok = 0; break; }
            id += i * stride;
            stride *= n[a];
        }
        if (!ok) continue;
        idx[nseg] = id;
        len[nseg] = t1 - t0;
        nseg++;
    }
    return nseg;
}

/* Ray of view b, column ip, row ix: source and unit direction (mm).
 * Geometry of ProjectionCone_3D.cu:26-42: S = (Ct - D sin b, Cs + D cos b, Cz),
 * direction proportional to (p cos b + D sin b, p sin b - D cos b, xi). */
static void cone_ray(const or_cone *g, int b, int ip, int ix, double S[3], double d[3])
{
    double beta = (double)g->beta[b], P = (double)g->p[ip], X = (double)g->xi[ix];
    double sb = sin(beta), cb = cos(beta);
    S[0] = g->size[0] / 2 - g->D * sb;
    S[1] = g->size[1] / 2 + g->D * cb;
    S[2] = g->size[2] / 2;
    d[0] = P * cb + g->D * sb;
    d[1] = P * sb - g->D * cb;
    d[2] = X;
    double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] /= nrm; d[1] /= nrm; d[2] /= nrm;
}

static void fan_ray(const or_fan *g, int b, int ip, double S[3], double d[3])
{
    double beta = g->beta[b], P = g->p[ip];
    double sb = sin(beta), cb = cos(beta);
    S[0] = g->size[0] / 2 - g->D * sb;
    S[1] = g->size[1] / 2 + g->D * cb;
    S[2] = 0.5;
    d[0] = P * cb + g->D * sb;
    d[1] = P * sb - g->D * cb;
    d[2] = 0.0;
    double nrm = sqrt(d[0] * d[0] + d[1] * d[1]);
    d[0] /= nrm; d[1] /= nrm;
}

static int max_segments(const int n[3]) { return n[0] + n[1] + n[2] + 8; }

/* Exact Siddon forward projection (cone). */
void or_cone_fp_exact(const or_cone *g, const float *vol, double *proj)
{
    const int n[3] = {g->nt, g->ns, g->nz};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, g->size[2] / g->nz};
    const long nray = (long)g->nb * g->nx * g->np;
#pragma omp parallel
    {
        int cap = max_segments(n);
        long *idx = malloc(sizeof(long) * cap);
        double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
#pragma omp for schedule(dynamic, 64)
        for (long ray = 0; ray < nray; ray++) {
            int ip = (int)(ray % g->np), ix = (int)((ray / g->np) % g->nx), b = (int)(ray / ((long)g->np * g->nx));
            double S[3], d[3];
            cone_ray(g, b, ip, ix, S, d);
            int m = siddon_exact(S, d, n, r, 3, idx, len, work);
            double acc = 0.0;
            for (int q = 0; q < m; q++) acc += len[q] * (double)vol[idx[q]];
            proj[ray] = acc;
        }
        free(idx); free(len); free(work);
    }
}

/* Total chord of each ray through the volume box: the row sums A*1 of the
 * Siddon system matrix (SIRT3D.m:81 sum(SysMatrix,2)). */
void or_cone_rowsum(const or_cone *g, double *out)
{
    const int n[3] = {g->nt, g->ns, g->nz};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, g->size[2] / g->nz};
    const long nray = (long)g->nb * g->nx * g->np;
#pragma omp parallel
    {
        int cap = max_segments(n);
        long *idx = malloc(sizeof(long) * cap);
        double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
#pragma omp for schedule(dynamic, 64)
        for (long ray = 0; ray < nray; ray++) {
            int ip = (int)(ray % g->np), ix = (int)((ray / g->np) % g->nx), b = (int)(ray / ((long)g->np * g->nx));
            double S[3], d[3];
            cone_ray(g, b, ip, ix, S, d);
            int m = siddon_exact(S, d, n, r, 3, idx, len, work);
            double acc = 0.0;
            for (int q = 0; q < m; q++) acc += len[q];
            out[ray] = acc;
        }
        free(idx); free(len); free(work);
    }
}

/* Exact transpose of or_cone_fp_exact (ray-driven scatter). */
void or_cone_adjoint(const or_cone *g, const float *proj, double *vol)
{
    const int n[3] = {g->nt, g->ns, g->nz};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, g->size[2] / g->nz};
    const long nray = (long)g->nb * g->nx * g->np;
    const long nvox = (long)g->nt * g->ns * g->nz;
    memset(vol, 0, sizeof(double) * nvox);
#pragma omp parallel
    {
        int cap = max_segments(n);
        long *idx = malloc(sizeof(long) * cap);
        double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
        double *priv = calloc(nvox, sizeof(double));
#pragma omp for schedule(dynamic, 64)
        for (long ray = 0; ray < nray; ray++) {
            double y = (double)proj[ray];
            if (y == 0.0) continue;
            int ip = (int)(ray % g->np), ix = (int)((ray / g->np) % g->nx), b = (int)(ray / ((long)g->np * g->nx));
            double S[3], d[3];
            cone_ray(g, b, ip, ix, S, d);
            int m = siddon_exact(S, d, n, r, 3, idx, len, work);
            for (int q = 0; q < m; q++) priv[idx[q]] += len[q] * y;
        }
#pragma omp critical
        for (long v = 0; v < nvox; v++) vol[v] += priv[v];
        free(priv); free(idx); free(len); free(work);
    }
}

/* Long-double brute force: clip the ray against every voxel box.  Used only
 * to pin siddon_exact on tiny volumes. */
double or_cone_ray_brute(const or_cone *g, const float *vol, int b, int ip, int ix)
{
    double Sd[3], dd[3];
    cone_ray(g, b, ip, ix, Sd, dd);
    long double S[3] = {Sd[0], Sd[1], Sd[2]}, d[3] = {dd[0], dd[1], dd[2]};
    long double r[3] = {(long double)g->size[0] / g->nt, (long double)g->size[1] / g->ns,
                        (long double)g->size[2] / g->nz};
    long double acc = 0.0L;
    for (int k = 0; k < g->nz; k++)
        for (int j = 0; j < g->ns; j++)
            for (int i = 0; i < g->nt; i++) {
                long double blo[3] = {i * r[0], j * r[1], k * r[2]};
                long double bhi[3] = {(i + 1) * r[0], (j + 1) * r[1], (k + 1) * r[2]};
                long double lo = -INFINITY, hi = INFINITY;
                int miss = 0;
                for (int a = 0; a < 3; a++) {
                    if (d[a] == 0.0L) {
                        if (!(S[a] >= blo[a] && S[a] < bhi[a])) miss = 1;
                        continue;
                    }
                    long double t1 = (blo[a] - S[a]) / d[a], t2 = (bhi[a] - S[a]) / d[a];
                    long double a1 = t1 < t2 ? t1 : t2, a2 = t1 < t2 ? t2 : t1;
                    if (a1 > lo) lo = a1;
                    if (a2 < hi) hi = a2;
                }
                if (miss || !(hi > lo)) continue;
                acc += (hi - lo) * (long double)vol[((long)k * g->ns + j) * g->nt + i];
            }
    return (double)acc;
}

/* ------------------------------------------------------------------------
 * Restatement of the ProjectionCone kernel
 * (GPU_accelerated_Operator/ProjectionCone_3D/ProjectionCone_3D.cu:8-379)
 * in FP64: 1e-16 slope clamps (:78-84, ProjectionCone_3D.h:29-30), 6-face
 * entry/exit (:110-218), start at the face point nearer the source
 * (:220-233), nearest-next-plane marching with EPS integer tests and the
 * Nt+Ns+Nz-2 step cap (:270-375), voxel = floor of segment midpoint.
 * ---------------------------------------------------------------------- */
static double clamp16(double x)
{
    if (x > 1e16) x = 1e16;
    if (x < -1e16) x = -1e16;
    return x;
}

static double ref_cone_ray(const or_cone *g, const float *vol, double beta_f, double P, double X)
{
    const double EPS = 1e-15;
    const int tl = g->nt, sl = g->ns, zl = g->nz;
    const double r1 = g->size[0] / tl, r2 = g->size[1] / sl, r3 = g->size[2] / zl;
    const double Ct = g->size[0] / 2, Cs = g->size[1] / 2, Cz = g->size[2] / 2, D = g->D;
    double st = Ct - D * sin(beta_f), ss = Cs + D * cos(beta_f), sz = Cz;
    double gam = atan(X / D), Dsh = D / cos(gam), th = atan(P / Dsh), Smax = 2 * D;
    double te = Ct + Smax * sin(th) * cos(beta_f) - (D - Smax * cos(th) * cos(gam)) * sin(beta_f);
    double se = Cs + Smax * sin(th) * sin(beta_f) + (D - Smax * cos(th) * cos(gam)) * cos(beta_f);
    double ze = Cz + Smax * cos(th) * sin(gam);
    double dt = te - st, ds = se - ss, dz = ze - sz;
    double T2S = ds / (dt == 0 ? dt + EPS : dt), T2Z = dz / (dt == 0 ? dt + EPS : dt);
    double S2T = dt / (ds == 0 ? ds + EPS : ds), S2Z = dz / (ds == 0 ? ds + EPS : ds);
    double Z2T = dt / (dz == 0 ? dz + EPS : dz), Z2S = ds / (dz == 0 ? dz + EPS : dz);
    T2S = clamp16(T2S); T2Z = clamp16(T2Z); S2T = clamp16(S2T);
    S2Z = clamp16(S2Z); Z2T = clamp16(Z2T); Z2S = clamp16(Z2S);
    int tsg = te >= st ? 1 : -1, ssg = se >= ss ? 1 : -1, zsg = ze >= sz ? 1 : -1;
    double thi = tl * r1, shi = sl * r2, zhi = zl * r3;
    /* the six face crossings, in the reference's search order */
    double fp[6][3] = {
        {0, ss + (0 - st) * T2S, sz + (0 - st) * T2Z},
        {thi, ss + (thi - st) * T2S, sz + (thi - st) * T2Z},
        {st + (0 - ss) * S2T, 0, sz + (0 - ss) * S2Z},
        {st + (shi - ss) * S2T, shi, sz + (shi - ss) * S2Z},
        {st + (0 - sz) * Z2T, ss + (0 - sz) * Z2S, 0},
        {st + (zhi - sz) * Z2T, ss + (zhi - sz) * Z2S, zhi},
    };
    int ok[6];
    for (int f = 0; f < 6; f++)
        ok[f] = fp[f][0] >= 0 && fp[f][0] <= thi && fp[f][1] >= 0 && fp[f][1] <= shi && fp[f][2] >= 0 &&
                fp[f][2] <= zhi;
    int f1 = -1, f2 = -1;
    for (int f = 0; f < 6 && f1 < 0; f++)
        if (ok[f]) f1 = f;
    if (f1 < 0) return 0.0;
    for (int f = f1 + 1; f < 6 && f2 < 0; f++)
        if (ok[f]) f2 = f;
    double P1[3] = {fp[f1][0], fp[f1][1], fp[f1][2]};
    double P2[3] = {0, 0, 0};
    if (f2 >= 0) { P2[0] = fp[f2][0]; P2[1] = fp[f2][1]; P2[2] = fp[f2][2]; }
#define DSQ(a, b, c, x, y, z) (((a) - (x)) * ((a) - (x)) + ((b) - (y)) * ((b) - (y)) + ((c) - (z)) * ((c) - (z)))
    double T, Sv, Z;
    if (DSQ(P1[0], P1[1], P1[2], st, ss, sz) >= DSQ(P2[0], P2[1], P2[2], st, ss, sz)) {
        T = P2[0]; Sv = P2[1]; Z = P2[2];
    } else {
        T = P1[0]; Sv = P1[1]; Z = P1[2];
    }
    double Ray = 0;
    double TC = T / r1, SC = Sv / r2, ZC = Z / r3;
    int cap = tl + sl + zl - 2;
    for (int it = 0; it < cap; it++) {
        short GT, GS, GZ;
        if (TC - (double)((short)TC) < EPS) GT = (short)(TC + tsg);
        else GT = (short)(floor(TC) + (1 + tsg) / 2);
        double GT_s = (ss + (GT * r1 - st) * T2S) / r2, GT_z = (sz + (GT * r1 - st) * T2Z) / r3;
        if (SC - (double)((short)SC) < EPS) GS = (short)(SC + ssg);
        else GS = (short)(floor(SC) + (1 + ssg) / 2);
        double GS_t = (st + (GS * r2 - ss) * S2T) / r1, GS_z = (sz + (GS * r2 - ss) * S2Z) / r3;
        if (ZC - (double)((short)ZC) < EPS) GZ = (short)(ZC + zsg);
        else GZ = (short)(floor(ZC) + (1 + zsg) / 2);
        double GZ_t = (st + (GZ * r3 - sz) * Z2T) / r1, GZ_s = (ss + (GZ * r3 - sz) * Z2S) / r2;
        double dT = DSQ((double)GT, GT_s, GT_z, TC, SC, ZC);
        double dS = DSQ(GS_t, (double)GS, GS_z, TC, SC, ZC);
        double dZ = DSQ(GZ_t, GZ_s, (double)GZ, TC, SC, ZC);
        double nT, nS, nZ;
        if (dT <= dS) {
            if (dZ <= dT) { nT = GZ_t; nS = GZ_s; nZ = GZ; }
            else { nT = GT; nS = GT_s; nZ = GT_z; }
        } else {
            if (dZ <= dS) { nT = GZ_t; nS = GZ_s; nZ = GZ; }
            else { nT = GS_t; nS = GS; nZ = GS_z; }
        }
        double w = sqrt(DSQ(nT * r1, nS * r2, nZ * r3, TC * r1, SC * r2, ZC * r3));
        short it_ = (short)floor((nT + TC) / 2), is_ = (short)floor((nS + SC) / 2), iz_ = (short)floor((nZ + ZC) / 2);
        TC = nT; SC = nS; ZC = nZ;
        if (it_ >= 0 && it_ < tl && is_ >= 0 && is_ < sl && iz_ >= 0 && iz_ < zl)
            Ray += w * (double)vol[((long)iz_ * sl + is_) * tl + it_];
        else
            break;
    }
#undef DSQ
    return Ray;
}

void or_cone_fp_ref(const or_cone *g, const float *vol, double *proj)
{
    const long nray = (long)g->nb * g->nx * g->np;
#pragma omp parallel for schedule(dynamic, 64)
    for (long ray = 0; ray < nray; ray++) {
        int ip = (int)(ray % g->np), ix = (int)((ray / g->np) % g->nx), b = (int)(ray / ((long)g->np * g->nx));
        proj[ray] = ref_cone_ray(g, vol, (double)g->beta[b], (double)g->p[ip], (double)g->xi[ix]);
    }
}

/* ------------------------------------------------------------------------
 * Detector sampling helpers shared by the voxel-driven backprojectors:
 * the reference derives the pitch from the first two samples and the
 * acceptance window from the end samples
 * (Backprojection3D.cu:304-311, FDK.cu:140-147).
 * ---------------------------------------------------------------------- */
typedef struct { double first, pitch, step, lo, hi; int n; const float *pos; } or_axis;

/* Interpolation grid of the voxel-driven backprojectors.
 *   0 (reference): index floor(|x - x0| / |x1 - x0|) with the pitch of the
 *     first two float samples, weights from the sample positions
 *     (FDK.cu:102-112).  For a pitch that is not exact in float32 this pitch
 *     drifts across the detector and near sample boundaries far from x0 the
 *     index is one too low, where the |.| weights then over-weight.
 *   1 (uniform): linear interpolation on the uniform grid through the end
 *     samples x0 + k (x_{n-1} - x0)/(n-1), as the HIP kernels compute.
 * The two agree when the positions are equally spaced in float32. */
static int g_interp_mode = 0;

void or_set_interp_mode(int mode) { g_interp_mode = mode; }

static or_axis make_axis(const float *pos, int n)
{
    or_axis a;
    a.first = pos[0];
    a.pitch = (double)fabsf(pos[1] - pos[0]);
    a.step = ((double)pos[n - 1] - (double)pos[0]) / (n - 1);
    if (g_interp_mode == 1) a.pitch = fabs(a.step);
    a.lo = (double)(pos[0] < pos[n - 1] ? pos[0] : pos[n - 1]);
    a.hi = (double)(pos[0] > pos[n - 1] ? pos[0] : pos[n - 1]);
    a.n = n;
    a.pos = pos;
    return a;
}

/* Linear interpolation coordinates of x on an axis (FDK.cu:102-112): lower
 * sample index and the two distance weights (each divided by the pitch). */
static int axis_lerp(const or_axis *a, double x, int *i1, double *w1, double *w2)
{
    if (!(x >= a->lo && x < a->hi)) return 0;
    if (g_interp_mode == 1) {
        double t = (x - a->first) / a->step;
        int i = (int)floor(t);
        if (i > a->n - 2) i = a->n - 2;
        if (i < 0) i = 0;
        double f = t - i;
        *i1 = i;
        *w2 = f * a->pitch;
        *w1 = (1.0 - f) * a->pitch;
        return 1;
    }
    int i = (int)floor(fabs(x - a->first) / a->pitch);
    if (i > a->n - 2) i = a->n - 2; /* the reference would read past the end here */
    *i1 = i;
    *w2 = fabs(x - (double)a->pos[i]);      /* distance to the lower sample */
    *w1 = fabs((double)a->pos[i + 1] - x);  /* distance to the upper sample */
    return 1;
}

/* Restatement of the BackProjection3D kernel
 * (GPU_accelerated_Operator/BackprojectionCone_3D/Backprojection3D.cu:8-291).
 * quirks != 0 keeps the `return` out of the z loop when no low face is hit
 * (:277-281); otherwise that voxel is skipped and the loop continues. */
void or_cone_bp_ref(const or_cone *g, const float *proj, double *vol, int quirks)
{
    const double EPS = 1e-15;
    const int tl = g->nt, sl = g->ns, zl = g->nz;
    const double rt = g->size[0] / tl, rs = g->size[1] / sl, rz = g->size[2] / zl;
    const double Ct = g->size[0] / 2, Cs = g->size[1] / 2, Cz = g->size[2] / 2, D = g->D;
    or_axis ap = make_axis(g->p, g->np), ax = make_axis(g->xi, g->nx);
    memset(vol, 0, sizeof(double) * (long)tl * sl * zl);
#pragma omp parallel for collapse(2) schedule(static)
    for (int S = 0; S < sl; S++)
        for (int T = 0; T < tl; T++) {
            for (int b = 0; b < g->nb; b++) {
                double beta = (double)g->beta[b];
                double it = (T + 0.5) * rt - Ct, is = (S + 0.5) * rs - Cs;
                double dt = it * cos(beta) + is * sin(beta), dsv = -it * sin(beta) + is * cos(beta);
                double LR = D / (D - dsv), P = dt * LR;
                double st = Ct - D * sin(beta), ss = Cs + D * cos(beta), sz = Cz;
                double te = it + Ct, se = is + Cs;
                const float *R = proj + (long)b * g->np * g->nx;
                for (int Z = 0; Z < zl; Z++) {
                    double iz = (Z + 0.5) * rz - Cz, X = iz * LR;
                    int ip, ix;
                    double p1, p2, x1, x2;
                    if (!axis_lerp(&ap, P, &ip, &p2, &p1)) continue;
                    if (!axis_lerp(&ax, X, &ix, &x2, &x1)) continue;
                    /* p2/x2 = distance to upper sample (weight of lower), p1/x1 = distance to lower */
                    double val = (x2 * p2 * R[(long)ix * g->np + ip] + x1 * p2 * R[(long)(ix + 1) * g->np + ip] +
                                  x2 * p1 * R[(long)ix * g->np + ip + 1] + x1 * p1 * R[(long)(ix + 1) * g->np + ip + 1]) /
                                 (ap.pitch * ax.pitch);
                    double ze = iz + Cz;
                    double tlo = T * rt, thi = (T + 1) * rt, slo = S * rs, shi = (S + 1) * rs, zlo = Z * rz,
                           zhi = (Z + 1) * rz;
                    double tlo_s = ss + (tlo - st) * (se - ss) / (te - st + EPS);
                    double tlo_z = sz + (tlo - st) * (ze - sz) / (te - st + EPS);
                    double slo_t = st + (slo - ss) * (te - st) / (se - ss + EPS);
                    double slo_z = sz + (slo - ss) * (ze - sz) / (se - ss + EPS);
                    double zlo_t = st + (zlo - sz) * (te - st) / (ze - sz + EPS);
                    double zlo_s = ss + (zlo - sz) * (se - ss) / (ze - sz + EPS);
                    double T1, S1, Z1;
                    if (tlo_s >= slo && tlo_s <= shi && tlo_z >= zlo && tlo_z <= zhi) {
                        T1 = tlo; S1 = tlo_s; Z1 = tlo_z;
                    } else if (slo_t >= tlo && slo_t <= thi && slo_z >= zlo && slo_z <= zhi) {
                        T1 = slo_t; S1 = slo; Z1 = slo_z;
                    } else if (zlo_t >= tlo && zlo_t <= thi && zlo_s >= slo && zlo_s <= shi) {
                        T1 = zlo_t; S1 = zlo_s; Z1 = zlo;
                    } else {
                        if (quirks) break;
                        continue;
                    }
                    double chord = 2 * sqrt((T1 - te) * (T1 - te) + (S1 - se) * (S1 - se) + (Z1 - ze) * (Z1 - ze));
                    vol[((long)Z * sl + S) * tl + T] += val * chord;
                }
            }
        }
}

/* ------------------------------------------------------------------------
 * Ramp filters.  Half Shepp-Logan taps of FDK_GPU/FDK.cu:6-19 (GFunction),
 * half Ram-Lak taps of FBP_type_analytic_algorithm/FanProjection2.m:249-261
 * (commented variant), equiangular Ram-Lak of FanProjection.m:241-254
 * halved.  taps[n] for n = 0..L-1 (symmetric).
 * ---------------------------------------------------------------------- */
void or_filter_taps(int filter, int equiangular, double d, int L, double *taps)
{
    for (int n = 0; n < L; n++) {
        double h;
        if (filter == 0) { /* SL half: stored as float in the reference (dev_G is float*) */
            h = -1.0 / ((d * OR_PI) * (d * OR_PI)) / (4.0 * n * n - 1.0);
            h = (double)(float)h;
        } else if (n == 0) {
            h = 1.0 / (8.0 * d * d);
        } else if (n % 2 == 0) {
            h = 0.0;
        } else if (equiangular) {
            double s = sin(n * d);
            h = -1.0 / (2.0 * OR_PI * OR_PI * s * s);
        } else {
            h = -1.0 / (2.0 * OR_PI * OR_PI * (double)n * n * d * d);
        }
        taps[n] = h;
    }
}

/* Row convolution out[p] = d * sum_g w[g] taps[|p-g|] over the detector row
 * (FDK.cu:38-47: Rcov += PInt * sum_g R[g] G[LP-1+p-g]). */
static void conv_row(const double *w, int L, const double *taps, double d, double *out)
{
    for (int p = 0; p < L; p++) {
        double acc = 0.0;
        for (int q = 0; q < L; q++) acc += w[q] * taps[p > q ? p - q : q - p];
        out[p] = d * acc;
    }
}

/* FDK pre-weight and filter (FDK.cu:21-48), race-free: every sample is
 * pre-weighted (rounded to float, as the in-place float store does) before
 * any convolution reads it. */
void or_cone_filter(const or_cone *g, int filter, const float *proj, double *out)
{
    const double D = g->D;
    const double d = (double)fabsf(g->p[1] - g->p[0]);
    double *taps = malloc(sizeof(double) * g->np);
    or_filter_taps(filter, 0, (double)(float)d, g->np, taps);
    const long rows = (long)g->nb * g->nx;
#pragma omp parallel
    {
        double *w = malloc(sizeof(double) * g->np), *o = malloc(sizeof(double) * g->np);
#pragma omp for schedule(static)
        for (long row = 0; row < rows; row++) {
            int ix = (int)(row % g->nx);
            const float *R = proj + row * g->np;
            double X = g->xi[ix];
            for (int q = 0; q < g->np; q++) {
                double P = g->p[q];
                w[q] = (double)(float)((double)R[q] * (D / sqrt(D * D + P * P + X * X)));
            }
            conv_row(w, g->np, taps, d, o);
            memcpy(out + row * g->np, o, sizeof(double) * g->np);
        }
        free(w); free(o);
    }
    free(taps);
}

/* Restatement of the FDK BackProjection kernel (FDK_GPU/FDK.cu:50-126):
 * bilinear / (dp dxi) x mag^2 x dbeta, Weight == 1 (:114).  With quirks != 0
 * the sample of the last in-detector z is re-added for every z above the
 * detector (Display_pBeta is not reset, :100-123). */
void or_fdk_bp(const or_cone *g, const float *filt, double *vol, int quirks)
{
    const int tl = g->nt, sl = g->ns, zl = g->nz;
    const double rt = g->size[0] / tl, rs = g->size[1] / sl, rz = g->size[2] / zl;
    const double Ct = g->size[0] / 2, Cs = g->size[1] / 2, Cz = g->size[2] / 2, D = g->D;
    or_axis ap = make_axis(g->p, g->np), ax = make_axis(g->xi, g->nx);
    const double dbeta = (double)fabsf(g->beta[1] - g->beta[0]);
    memset(vol, 0, sizeof(double) * (long)tl * sl * zl);
#pragma omp parallel for collapse(2) schedule(static)
    for (int S = 0; S < sl; S++)
        for (int T = 0; T < tl; T++) {
            for (int b = 0; b < g->nb; b++) {
                double beta = (double)g->beta[b];
                double it = (T + 0.5) * rt - Ct, is = (S + 0.5) * rs - Cs;
                double dt = it * cos(beta) + is * sin(beta), dsv = -it * sin(beta) + is * cos(beta);
                double LR = D / (D - dsv), P = dt * LR;
                const float *R = filt + (long)b * g->np * g->nx;
                double carried = 0.0;
                for (int Z = 0; Z < zl; Z++) {
                    double iz = (Z + 0.5) * rz - Cz, X = iz * LR;
                    int ip, ix;
                    double p1, p2, x1, x2, val;
                    if (axis_lerp(&ap, P, &ip, &p2, &p1) && axis_lerp(&ax, X, &ix, &x2, &x1)) {
                        val = (x2 * p2 * R[(long)ix * g->np + ip] + x1 * p2 * R[(long)(ix + 1) * g->np + ip] +
                               x2 * p1 * R[(long)ix * g->np + ip + 1] + x1 * p1 * R[(long)(ix + 1) * g->np + ip + 1]) /
                              (ap.pitch * ax.pitch) * LR * LR * dbeta;
                        carried = val;
                    } else {
                        val = quirks ? carried : 0.0;
                    }
                    vol[((long)Z * sl + S) * tl + T] += val;
                }
            }
        }
}

/* ------------------------------------------------------------------------
 * Fan beam.
 * ---------------------------------------------------------------------- */
void or_fan_fp_exact(const or_fan *g, const float *img, double *proj)
{
    const int n[3] = {g->nt, g->ns, 1};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, 1.0};
    const long nray = (long)g->nb * g->np;
#pragma omp parallel
    {
        int cap = max_segments(n);
        long *idx = malloc(sizeof(long) * cap);
        double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
#pragma omp for schedule(dynamic, 16)
        for (long ray = 0; ray < nray; ray++) {
            double S[3], d[3];
            fan_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
            int m = siddon_exact(S, d, n, r, 2, idx, len, work);
            double acc = 0.0;
            for (int q = 0; q < m; q++) acc += len[q] * (double)img[idx[q]];
            proj[ray] = acc;
        }
        free(idx); free(len); free(work);
    }
}

void or_fan_rowsum(const or_fan *g, double *out)
{
    const int n[3] = {g->nt, g->ns, 1};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, 1.0};
    const long nray = (long)g->nb * g->np;
    long *idx = malloc(sizeof(long) * max_segments(n));
    double *len = malloc(sizeof(double) * max_segments(n)), *work = malloc(sizeof(double) * (max_segments(n) + 4));
    for (long ray = 0; ray < nray; ray++) {
        double S[3], d[3];
        fan_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
        int m = siddon_exact(S, d, n, r, 2, idx, len, work);
        double acc = 0.0;
        for (int q = 0; q < m; q++) acc += len[q];
        out[ray] = acc;
    }
    free(idx); free(len); free(work);
}

void or_fan_adjoint(const or_fan *g, const float *proj, double *img)
{
    const int n[3] = {g->nt, g->ns, 1};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, 1.0};
    const long nray = (long)g->nb * g->np;
    memset(img, 0, sizeof(double) * (long)g->nt * g->ns);
    long *idx = malloc(sizeof(long) * max_segments(n));
    double *len = malloc(sizeof(double) * max_segments(n)), *work = malloc(sizeof(double) * (max_segments(n) + 4));
    for (long ray = 0; ray < nray; ray++) {
        double y = (double)proj[ray];
        if (y == 0.0) continue;
        double S[3], d[3];
        fan_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
        int m = siddon_exact(S, d, n, r, 2, idx, len, work);
        for (int q = 0; q < m; q++) img[idx[q]] += len[q] * y;
    }
    free(idx); free(len); free(work);
}

/* Restatement of the ProjectionFan kernel (ProjectionFan_2D.cu:8-220):
 * `round`-based integer tests (:165,175) and the Nt+Ns-1 step cap (:162). */
static double ref_fan_ray(const or_fan *g, const float *img, double beta, double P)
{
    const double EPS = 1e-15;
    const int tl = g->nt, sl = g->ns;
    const double r1 = g->size[0] / tl, r2 = g->size[1] / sl;
    const double Ct = g->size[0] / 2, Cs = g->size[1] / 2, D = g->D;
    double st = Ct + D * sin(-beta), ss = Cs + D * cos(-beta);
    double th = atan(P / D), Smax = 2 * D;
    double te = Ct + Smax * sin(th) * cos(-beta) + (D - Smax * cos(th)) * sin(-beta);
    double se = Cs - Smax * sin(th) * sin(-beta) + (D - Smax * cos(th)) * cos(-beta);
    double dt = te - st, ds = se - ss;
    double T2S = clamp16(ds / (dt == 0 ? dt + EPS : dt)), S2T = clamp16(dt / (ds == 0 ? ds + EPS : ds));
    int tsg = te >= st ? 1 : -1, ssg = se >= ss ? 1 : -1;
    double thi = tl * r1, shi = sl * r2;
    double fp[4][2] = {{0, ss + (0 - st) * T2S}, {thi, ss + (thi - st) * T2S},
                       {st + (0 - ss) * S2T, 0}, {st + (shi - ss) * S2T, shi}};
    int ok[4];
    for (int f = 0; f < 4; f++) ok[f] = fp[f][0] >= 0 && fp[f][0] <= thi && fp[f][1] >= 0 && fp[f][1] <= shi;
    int f1 = -1, f2 = -1;
    for (int f = 0; f < 4 && f1 < 0; f++)
        if (ok[f]) f1 = f;
    if (f1 < 0) return 0.0;
    for (int f = f1 + 1; f < 4 && f2 < 0; f++)
        if (ok[f]) f2 = f;
    double P1[2] = {fp[f1][0], fp[f1][1]}, P2[2] = {0, 0};
    if (f2 >= 0) { P2[0] = fp[f2][0]; P2[1] = fp[f2][1]; }
#define DSQ2(a, b, x, y) (((a) - (x)) * ((a) - (x)) + ((b) - (y)) * ((b) - (y)))
    double T, Sv;
    if (DSQ2(P1[0], P1[1], st, ss) >= DSQ2(P2[0], P2[1], st, ss)) { T = P2[0]; Sv = P2[1]; }
    else { T = P1[0]; Sv = P1[1]; }
    double Ray = 0, TC = T / r1, SC = Sv / r2;
    for (int it = 0; it < tl + sl - 1; it++) {
        short GT, GS;
        if (fabs(TC - round(TC)) < EPS) GT = (short)(round(TC) + tsg);
        else GT = (short)(floor(TC) + (1 + tsg) / 2);
        double GT_s = (ss + (GT * r1 - st) * T2S) / r2;
        if (fabs(SC - round(SC)) < EPS) GS = (short)(round(SC) + ssg);
        else GS = (short)(floor(SC) + (1 + ssg) / 2);
        double GS_t = (st + (GS * r2 - ss) * S2T) / r1;
        double nT, nS;
        if (DSQ2((double)GT, GT_s, TC, SC) <= DSQ2(GS_t, (double)GS, TC, SC)) { nT = GT; nS = GT_s; }
        else { nT = GS_t; nS = GS; }
        double w = sqrt(DSQ2(nT * r1, nS * r2, TC * r1, SC * r2));
        short it_ = (short)floor((nT + TC) / 2), is_ = (short)floor((nS + SC) / 2);
        TC = nT; SC = nS;
        if (it_ >= 0 && it_ < tl && is_ >= 0 && is_ < sl) Ray += w * (double)img[(long)is_ * tl + it_];
        else break;
    }
#undef DSQ2
    return Ray;
}

void or_fan_fp_ref(const or_fan *g, const float *img, double *proj)
{
    const long nray = (long)g->nb * g->np;
#pragma omp parallel for schedule(dynamic, 16)
    for (long ray = 0; ray < nray; ray++)
        proj[ray] = ref_fan_ray(g, img, g->beta[ray / g->np], g->p[ray % g->np]);
}

/* Fan-geometry detector axis from double positions rounded to float, as the
 * BackprojectionFan2D / FBPfan gateways receive single arrays
 * (BackprojectionFan2D.cpp:8-11). */
typedef struct { double first, pitch, step, lo, hi; int n; float *pos; } or_faxis;

static or_faxis make_faxis(const double *p, int n)
{
    or_faxis a;
    a.pos = malloc(sizeof(float) * n);
    for (int i = 0; i < n; i++) a.pos[i] = (float)p[i];
    a.first = a.pos[0];
    a.pitch = n > 1 ? (double)fabsf(a.pos[1] - a.pos[0]) : 0.0;
    a.step = n > 1 ? ((double)a.pos[n - 1] - (double)a.pos[0]) / (n - 1) : 0.0;
    a.lo = (double)(a.pos[0] < a.pos[n - 1] ? a.pos[0] : a.pos[n - 1]);
    a.hi = (double)(a.pos[0] > a.pos[n - 1] ? a.pos[0] : a.pos[n - 1]);
    a.n = n;
    return a;
}

static int faxis_lerp(const or_faxis *a, double x, int *i1, double *wl, double *wu)
{
    if (!(x >= a->lo && x < a->hi)) return 0;
    if (g_interp_mode == 1) {
        double t = (x - a->first) / a->step;
        int i = (int)floor(t);
        if (i > a->n - 2) i = a->n - 2;
        if (i < 0) i = 0;
        double f = t - i;
        *i1 = i;
        *wl = (1.0 - f) * a->pitch;
        *wu = f * a->pitch;
        return 1;
    }
    int i = (int)floor(fabs(x - a->first) / a->pitch);
    if (i > a->n - 2) i = a->n - 2;
    *i1 = i;
    *wl = fabs((double)a->pos[i + 1] - x); /* weight of the lower sample */
    *wu = fabs(x - (double)a->pos[i]);     /* weight of the upper sample */
    return 1;
}

/* Restatement of the fan BackProjection kernel (BackprojectionFan2D.cu:7-116).
 * quirks != 0 keeps (a) the entry-face test against 0 instead of the pixel's
 * low edge (:95-101) and (b) the chord measured to the ray's far end point
 * (:59-65,109); otherwise the chord is the pixel chord of the line through
 * the pixel centre. */
void or_fan_bp_ref(const or_fan *g, const float *proj, double *img, int quirks)
{
    const double EPS = 1e-15;
    const int tl = g->nt, sl = g->ns;
    const double rt = g->size[0] / tl, rs = g->size[1] / sl;
    const double Ct = g->size[0] / 2, Cs = g->size[1] / 2, D = g->D;
    or_faxis ap = make_faxis(g->p, g->np);
    memset(img, 0, sizeof(double) * (long)tl * sl);
#pragma omp parallel for collapse(2) schedule(static)
    for (int S = 0; S < sl; S++)
        for (int T = 0; T < tl; T++) {
            for (int b = 0; b < g->nb; b++) {
                double beta = (double)(float)g->beta[b];
                double it = (T + 0.5) * rt - Ct, is = (S + 0.5) * rs - Cs;
                double P = it * cos(beta) + is * sin(beta), Sv = -it * sin(beta) + is * cos(beta);
                double U = (D - Sv) / D, Pd = P / U;
                double val = 0.0;
                int ip;
                double wl, wu;
                if (faxis_lerp(&ap, Pd, &ip, &wl, &wu))
                    val = (wl * proj[(long)b * g->np + ip] + wu * proj[(long)b * g->np + ip + 1]) / ap.pitch;
                double st = Ct - D * sin(beta), ss = Cs + D * cos(beta);
                double th = atan(Pd / D), Smax = 2 * D;
                double te = Ct + Smax * sin(th) * cos(beta) - (D - Smax * cos(th)) * sin(beta);
                double se = Cs + Smax * sin(th) * sin(beta) + (D - Smax * cos(th)) * cos(beta);
                double T2S = clamp16((se - ss) / ((te - st) == 0 ? (te - st) + EPS : (te - st)));
                double S2T = clamp16((te - st) / ((se - ss) == 0 ? (se - ss) + EPS : (se - ss)));
                double tlo = T * rt, thi = (T + 1) * rt, slo = S * rs, shi = (S + 1) * rs;
                double tlo_s = ss + (tlo - st) * T2S, slo_t = st + (slo - ss) * S2T;
                double sb = quirks ? 0.0 : slo, tb = quirks ? 0.0 : tlo;
                double T1, S1;
                if (tlo_s >= sb && tlo_s <= shi) { T1 = tlo; S1 = tlo_s; }
                else if (slo_t >= tb && slo_t <= thi) { T1 = slo_t; S1 = slo; }
                else continue;
                /* te/se is the far end point of the ray (2D from the source).  The
                 * reference measures the "chord" from the entry point to that far
                 * point (:109); the intended chord is twice the distance from the
                 * entry point to the pixel centre, as in Backprojection3D.cu:283. */
                double ct = quirks ? te : it + Ct, cs = quirks ? se : is + Cs;
                double chord = 2 * sqrt((T1 - ct) * (T1 - ct) + (S1 - cs) * (S1 - cs));
                img[(long)S * tl + T] += val * chord;
            }
        }
    free(ap.pos);
}

/* Fan pre-weight + filter.
 *   flat (FBPfan.cu:21-43): w = R * D / sqrt(D^2 + p^2), out = dp * (w * h)
 *   equiangular (FanProjection.m:236-239,286-293): w = R * D cos(gamma),
 *   out = dgamma * (w * h) with h in gamma. */
void or_fan_filter(const or_fan *g, int equiangular, int filter, const float *proj, double *out)
{
    const double D = g->D;
    or_faxis ap = make_faxis(g->p, g->np);
    double d;
    double *gam = malloc(sizeof(double) * g->np);
    for (int q = 0; q < g->np; q++) gam[q] = atan(g->p[q] / D);
    d = equiangular ? fabs(gam[1] - gam[0]) : ap.pitch;
    double *taps = malloc(sizeof(double) * g->np);
    or_filter_taps(filter, equiangular, equiangular ? d : (double)(float)d, g->np, taps);
    double *w = malloc(sizeof(double) * g->np), *o = malloc(sizeof(double) * g->np);
    for (int b = 0; b < g->nb; b++) {
        const float *R = proj + (long)b * g->np;
        for (int q = 0; q < g->np; q++) {
            if (equiangular) w[q] = (double)R[q] * D * cos(gam[q]);
            else w[q] = (double)(float)((double)R[q] * (D / sqrt(D * D + (double)ap.pos[q] * ap.pos[q])));
        }
        conv_row(w, g->np, taps, d, o);
        memcpy(out + (long)b * g->np, o, sizeof(double) * g->np);
    }
    free(w); free(o); free(taps); free(gam); free(ap.pos);
}

/* Fan FBP backprojection.
 *   flat (FBPfan.cu:45-96): sum_b lerp(filtered)(p/U) / U^2 * dbeta
 *   equiangular (FanProjection.m:295-324): sum_b lerp in gamma / L^2 * dbeta. */
void or_fan_fbp_bp(const or_fan *g, int equiangular, const float *filt, double *img)
{
    const int tl = g->nt, sl = g->ns;
    const double rt = g->size[0] / tl, rs = g->size[1] / sl;
    const double Ct = g->size[0] / 2, Cs = g->size[1] / 2, D = g->D;
    or_faxis ap = make_faxis(g->p, g->np);
    const double dbeta = fabs((double)(float)g->beta[1] - (double)(float)g->beta[0]);
    double *gam = malloc(sizeof(double) * g->np);
    for (int q = 0; q < g->np; q++) gam[q] = atan(g->p[q] / D);
    double g0 = gam[0], dg = fabs(gam[1] - gam[0]);
    double glo = dmin(gam[0], gam[g->np - 1]), ghi = dmax(gam[0], gam[g->np - 1]);
    memset(img, 0, sizeof(double) * (long)tl * sl);
#pragma omp parallel for collapse(2) schedule(static)
    for (int S = 0; S < sl; S++)
        for (int T = 0; T < tl; T++) {
            double acc = 0.0;
            for (int b = 0; b < g->nb; b++) {
                double beta = equiangular ? g->beta[b] : (double)(float)g->beta[b];
                double it = (T + 0.5) * rt - Ct, is = (S + 0.5) * rs - Cs;
                double u = it * cos(beta) + is * sin(beta), w = -it * sin(beta) + is * cos(beta);
                const float *R = filt + (long)b * g->np;
                if (!equiangular) {
                    double U = (D - w) / D, Pd = u / U;
                    int ip;
                    double wl, wu;
                    if (faxis_lerp(&ap, Pd, &ip, &wl, &wu))
                        acc += (wl * R[ip] + wu * R[ip + 1]) / (U * U * ap.pitch) * dbeta;
                } else {
                    double gb = atan2(u, fabs(w - D)), L2 = u * u + (w - D) * (w - D);
                    if (gb >= glo && gb < ghi) {
                        if (g_interp_mode == 1) {
                            double t = (gb - g0) / ((gam[g->np - 1] - g0) / (g->np - 1));
                            int i = (int)floor(t);
                            if (i > g->np - 2) i = g->np - 2;
                            if (i < 0) i = 0;
                            double f = t - i;
                            acc += ((1.0 - f) * R[i] + f * R[i + 1]) / L2;
                        } else {
                            int i = (int)floor(fabs(gb - g0) / dg);
                            if (i > g->np - 2) i = g->np - 2;
                            acc += (fabs(gam[i + 1] - gb) * R[i] + fabs(gb - gam[i]) * R[i + 1]) / (L2 * dg);
                        }
                    }
                }
            }
            img[(long)S * tl + T] = equiangular ? acc * fabs(g->beta[1] - g->beta[0]) : acc;
        }
    free(gam); free(ap.pos);
}

/* ------------------------------------------------------------------------
 * SIRT (Iterative_algorithm/SIRT3D.m:67-92 and SIRT.m:126-154), matrix
 * free with the exact Siddon pair:
 *   x <- clamp(x + lambda * (A'((b - A x) ./ (A 1 + 1e-9))) ./ (A' 1 + 1e-9), lo, hi)
 * residual_l2[k] = ||b - A x_k|| / ||b||; residual_l1[k] = sum|b - A x_k| / (sum(b) * Nvox)
 * ---------------------------------------------------------------------- */
void or_sirt3d(const or_cone *g, const float *b, double *x, int n_iter, double lambda, double lo, double hi,
               double *res_l2, double *res_l1)
{
    const long nray = (long)g->nb * g->nx * g->np, nvox = (long)g->nt * g->ns * g->nz;
    double *rows = malloc(sizeof(double) * nray), *cols = malloc(sizeof(double) * nvox);
    double *ax = malloc(sizeof(double) * nray), *bp = malloc(sizeof(double) * nvox);
    float *rf = malloc(sizeof(float) * nray), *xf = malloc(sizeof(float) * nvox);
    or_cone_rowsum(g, rows);
    for (long i = 0; i < nray; i++) rf[i] = 1.0f;
    /* column sums from the adjoint of ones, computed in double */
    {
        const int n[3] = {g->nt, g->ns, g->nz};
        const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, g->size[2] / g->nz};
        memset(cols, 0, sizeof(double) * nvox);
        int cap = max_segments(n);
        long *idx = malloc(sizeof(long) * cap);
        double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
        for (long ray = 0; ray < nray; ray++) {
            int ip = (int)(ray % g->np), ix = (int)((ray / g->np) % g->nx), bb = (int)(ray / ((long)g->np * g->nx));
            double S[3], d[3];
            cone_ray(g, bb, ip, ix, S, d);
            int m = siddon_exact(S, d, n, r, 3, idx, len, work);
            for (int q = 0; q < m; q++) cols[idx[q]] += len[q];
        }
        free(idx); free(len); free(work);
    }
    double bn2 = 0, bsum = 0;
    for (long i = 0; i < nray; i++) { bn2 += (double)b[i] * b[i]; bsum += b[i]; }
    for (int k = 0; k <= n_iter; k++) {
        for (long v = 0; v < nvox; v++) xf[v] = (float)x[v];
        /* A x in double: use the exact FP on the double iterate via float copy
         * would lose precision; project the double iterate directly */
        {
            const int n[3] = {g->nt, g->ns, g->nz};
            const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, g->size[2] / g->nz};
#pragma omp parallel
            {
                int cap = max_segments(n);
                long *idx = malloc(sizeof(long) * cap);
                double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
#pragma omp for schedule(dynamic, 64)
                for (long ray = 0; ray < nray; ray++) {
                    int ip = (int)(ray % g->np), ix = (int)((ray / g->np) % g->nx),
                        bb = (int)(ray / ((long)g->np * g->nx));
                    double S[3], d[3];
                    cone_ray(g, bb, ip, ix, S, d);
                    int m = siddon_exact(S, d, n, r, 3, idx, len, work);
                    double acc = 0.0;
                    for (int q = 0; q < m; q++) acc += len[q] * x[idx[q]];
                    ax[ray] = acc;
                }
                free(idx); free(len); free(work);
            }
        }
        double r2 = 0, r1 = 0;
        for (long i = 0; i < nray; i++) {
            double e = (double)b[i] - ax[i];
            r2 += e * e;
            r1 += fabs(e);
        }
        if (res_l2) res_l2[k] = sqrt(r2) / (sqrt(bn2) > 0 ? sqrt(bn2) : 1.0);
        if (res_l1) res_l1[k] = r1 / ((bsum != 0 ? bsum : 1.0) * (double)nvox);
        if (k == n_iter) break;
        /* weighted residual, then the exact adjoint in double */
        {
            const int n[3] = {g->nt, g->ns, g->nz};
            const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, g->size[2] / g->nz};
            memset(bp, 0, sizeof(double) * nvox);
#pragma omp parallel
            {
                int cap = max_segments(n);
                long *idx = malloc(sizeof(long) * cap);
                double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
                double *priv = calloc(nvox, sizeof(double));
#pragma omp for schedule(dynamic, 64)
                for (long ray = 0; ray < nray; ray++) {
                    double e = ((double)b[ray] - ax[ray]) / (rows[ray] + 1e-9);
                    if (e == 0.0) continue;
                    int ip = (int)(ray % g->np), ix = (int)((ray / g->np) % g->nx),
                        bb = (int)(ray / ((long)g->np * g->nx));
                    double S[3], d[3];
                    cone_ray(g, bb, ip, ix, S, d);
                    int m = siddon_exact(S, d, n, r, 3, idx, len, work);
                    for (int q = 0; q < m; q++) priv[idx[q]] += len[q] * e;
                }
#pragma omp critical
                for (long v = 0; v < nvox; v++) bp[v] += priv[v];
                free(priv); free(idx); free(len); free(work);
            }
        }
        for (long v = 0; v < nvox; v++) {
            double nx = x[v] + lambda * bp[v] / (cols[v] + 1e-9);
            if (nx < lo) nx = lo;
            if (nx > hi) nx = hi;
            x[v] = nx;
        }
    }
    free(rows); free(cols); free(ax); free(bp); free(rf); free(xf);
}

void or_sirt2d(const or_fan *g, const float *b, double *x, int n_iter, double lambda, double lo, double hi,
               double *res_l2, double *res_l1)
{
    const long nray = (long)g->nb * g->np, nvox = (long)g->nt * g->ns;
    const int n[3] = {g->nt, g->ns, 1};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, 1.0};
    double *rows = malloc(sizeof(double) * nray), *cols = calloc(nvox, sizeof(double));
    double *ax = malloc(sizeof(double) * nray), *bp = malloc(sizeof(double) * nvox);
    int cap = max_segments(n);
    long *idx = malloc(sizeof(long) * cap);
    double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
    for (long ray = 0; ray < nray; ray++) {
        double S[3], d[3];
        fan_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
        int m = siddon_exact(S, d, n, r, 2, idx, len, work);
        double acc = 0;
        for (int q = 0; q < m; q++) { acc += len[q]; cols[idx[q]] += len[q]; }
        rows[ray] = acc;
    }
    double bn2 = 0, bsum = 0;
    for (long i = 0; i < nray; i++) { bn2 += (double)b[i] * b[i]; bsum += b[i]; }
    for (int k = 0; k <= n_iter; k++) {
        double r2 = 0, r1 = 0;
        for (long ray = 0; ray < nray; ray++) {
            double S[3], d[3];
            fan_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
            int m = siddon_exact(S, d, n, r, 2, idx, len, work);
            double acc = 0;
            for (int q = 0; q < m; q++) acc += len[q] * x[idx[q]];
            ax[ray] = acc;
            double e = (double)b[ray] - acc;
            r2 += e * e;
            r1 += fabs(e);
        }
        if (res_l2) res_l2[k] = sqrt(r2) / (sqrt(bn2) > 0 ? sqrt(bn2) : 1.0);
        if (res_l1) res_l1[k] = r1 / ((bsum != 0 ? bsum : 1.0) * (double)nvox);
        if (k == n_iter) break;
        memset(bp, 0, sizeof(double) * nvox);
        for (long ray = 0; ray < nray; ray++) {
            double e = ((double)b[ray] - ax[ray]) / (rows[ray] + 1e-9);
            if (e == 0.0) continue;
            double S[3], d[3];
            fan_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
            int m = siddon_exact(S, d, n, r, 2, idx, len, work);
            for (int q = 0; q < m; q++) bp[idx[q]] += len[q] * e;
        }
        for (long v = 0; v < nvox; v++) {
            double nx = x[v] + lambda * bp[v] / (cols[v] + 1e-9);
            if (nx < lo) nx = lo;
            if (nx > hi) nx = hi;
            x[v] = nx;
        }
    }
    free(rows); free(cols); free(ax); free(bp); free(idx); free(len); free(work);
}

/* ------------------------------------------------------------------------
 * Parallel beam (GPU_accelerated_Operator/ProjectionParallel_2D and
 * FBP_parallel_GPU/FBPparallel.cu).  Image [s][t], t fastest.  Ray (theta, t)
 * is the line (x - Ct) cos(theta) + (y - Cs) sin(theta) = t traversed along
 * (-sin theta, cos theta) (ProjectionParallel_2D.cu:22-28: start and end
 * points Center + t (cos, sin) -/+ Smax (-sin, cos)).
 * ---------------------------------------------------------------------- */
typedef struct {
    int nt, ns;
    double size[2];
    int nb, np;
    const double *theta, *t;
} or_par;

static void par_ray(const or_par *g, int b, int it, double S[3], double d[3])
{
    double th = g->theta[b], T = g->t[it];
    double c = cos(th), s = sin(th);
    S[0] = g->size[0] / 2 + T * c;
    S[1] = g->size[1] / 2 + T * s;
    S[2] = 0.5;
    d[0] = -s;
    d[1] = c;
    d[2] = 0.0;
}

/* Exact Siddon line integrals (the operator ProjectionParallel_2D.cu:9-220
 * computes with near-integer tolerance ERR = 1e-5 and a height+width-1 cap). */
void or_par_fp_exact(const or_par *g, const float *img, double *proj)
{
    const int n[3] = {g->nt, g->ns, 1};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, 1.0};
    const long nray = (long)g->nb * g->np;
#pragma omp parallel
    {
        int cap = max_segments(n);
        long *idx = malloc(sizeof(long) * cap);
        double *len = malloc(sizeof(double) * cap), *work = malloc(sizeof(double) * (cap + 4));
#pragma omp for schedule(dynamic, 16)
        for (long ray = 0; ray < nray; ray++) {
            double S[3], d[3];
            par_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
            int m = siddon_exact(S, d, n, r, 2, idx, len, work);
            double acc = 0.0;
            for (int q = 0; q < m; q++) acc += len[q] * (double)img[idx[q]];
            proj[ray] = acc;
        }
        free(idx); free(len); free(work);
    }
}

/* Transpose of or_par_fp_exact: every ray scatters its Siddon lengths. */
void or_par_adjoint(const or_par *g, const float *proj, double *img)
{
    const int n[3] = {g->nt, g->ns, 1};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, 1.0};
    const long nray = (long)g->nb * g->np;
    memset(img, 0, sizeof(double) * (long)g->nt * g->ns);
    long *idx = malloc(sizeof(long) * max_segments(n));
    double *len = malloc(sizeof(double) * max_segments(n)), *work = malloc(sizeof(double) * (max_segments(n) + 4));
    for (long ray = 0; ray < nray; ray++) {
        double y = (double)proj[ray];
        if (y == 0.0) continue;
        double S[3], d[3];
        par_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
        int m = siddon_exact(S, d, n, r, 2, idx, len, work);
        for (int q = 0; q < m; q++) img[idx[q]] += len[q] * y;
    }
    free(idx); free(len); free(work);
}

/* The exact Siddon segments of one ray (its system-matrix row), in visiting
 * order; returns their number (at most nt + ns + nz + 8). */
int or_cone_segments(const or_cone *g, long ray, long *idx, double *len)
{
    const int n[3] = {g->nt, g->ns, g->nz};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, g->size[2] / g->nz};
    double S[3], d[3], *work = malloc(sizeof(double) * (max_segments(n) + 4));
    cone_ray(g, (int)(ray / ((long)g->nx * g->np)), (int)(ray % g->np), (int)((ray / g->np) % g->nx), S, d);
    int m = siddon_exact(S, d, n, r, 3, idx, len, work);
    free(work);
    return m;
}

int or_fan_segments(const or_fan *g, long ray, long *idx, double *len)
{
    const int n[3] = {g->nt, g->ns, 1};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, 1.0};
    double S[3], d[3], *work = malloc(sizeof(double) * (max_segments(n) + 4));
    fan_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
    int m = siddon_exact(S, d, n, r, 2, idx, len, work);
    free(work);
    return m;
}

int or_par_segments(const or_par *g, long ray, long *idx, double *len)
{
    const int n[3] = {g->nt, g->ns, 1};
    const double r[3] = {g->size[0] / g->nt, g->size[1] / g->ns, 1.0};
    double S[3], d[3], *work = malloc(sizeof(double) * (max_segments(n) + 4));
    par_ray(g, (int)(ray / g->np), (int)(ray % g->np), S, d);
    int m = siddon_exact(S, d, n, r, 2, idx, len, work);
    free(work);
    return m;
}

/* Long-double brute force of one parallel ray: clip against every pixel box. */
double or_par_ray_brute(const or_par *g, const float *img, int b, int it)
{
    double Sd[3], dd[3];
    par_ray(g, b, it, Sd, dd);
    long double S[2] = {Sd[0], Sd[1]}, d[2] = {dd[0], dd[1]};
    long double r[2] = {(long double)g->size[0] / g->nt, (long double)g->size[1] / g->ns};
    long double acc = 0.0L;
    for (int j = 0; j < g->ns; j++)
        for (int i = 0; i < g->nt; i++) {
            long double blo[2] = {i * r[0], j * r[1]}, bhi[2] = {(i + 1) * r[0], (j + 1) * r[1]};
            long double lo = -INFINITY, hi = INFINITY;
            int miss = 0;
            for (int a = 0; a < 2; a++) {
                if (d[a] == 0.0L) {
                    if (!(S[a] >= blo[a] && S[a] < bhi[a])) miss = 1;
                    continue;
                }
                long double t1 = (blo[a] - S[a]) / d[a], t2 = (bhi[a] - S[a]) / d[a];
                long double a1 = t1 < t2 ? t1 : t2, a2 = t1 < t2 ? t2 : t1;
                if (a1 > lo) lo = a1;
                if (a2 < hi) hi = a2;
            }
            if (miss || !(hi > lo)) continue;
            acc += (hi - lo) * (long double)img[(long)j * g->nt + i];
        }
    return (double)acc;
}

/* Parallel-beam filter (FBPparallel.cu:6-39): no pre-weight,
 * Rcov[t] = dt * sum_g R[g] G[LT-1+t-g] with the float half-SL taps of
 * GFunction (or half Ram-Lak); t positions arrive as single precision. */
void or_par_filter(const or_par *g, int filter, const float *proj, double *out)
{
    or_faxis ap = make_faxis(g->t, g->np);
    const double d = ap.pitch;
    double *taps = malloc(sizeof(double) * g->np);
    or_filter_taps(filter, 0, (double)(float)d, g->np, taps);
    double *w = malloc(sizeof(double) * g->np), *o = malloc(sizeof(double) * g->np);
    for (int b = 0; b < g->nb; b++) {
        const float *R = proj + (long)b * g->np;
        for (int q = 0; q < g->np; q++) w[q] = (double)R[q];
        conv_row(w, g->np, taps, d, o);
        memcpy(out + (long)b * g->np, o, sizeof(double) * g->np);
    }
    free(w); free(o); free(taps); free(ap.pos);
}

/* Parallel-beam FBP backprojection (FBPparallel.cu:41-89): for each pixel
 * and view, t = x cos(theta) + y sin(theta) (pixel centre relative to the
 * centre), linear interpolation on the detector / dt, times dtheta; theta
 * and t arrive as single precision (FBPparallel.cpp:7-11). */
void or_par_fbp_bp(const or_par *g, const float *filt, double *img)
{
    const int tl = g->nt, sl = g->ns;
    const double rt = g->size[0] / tl, rs = g->size[1] / sl;
    const double Ct = g->size[0] / 2, Cs = g->size[1] / 2;
    or_faxis ap = make_faxis(g->t, g->np);
    const double dtheta = fabs((double)(float)g->theta[1] - (double)(float)g->theta[0]);
    memset(img, 0, sizeof(double) * (long)tl * sl);
#pragma omp parallel for collapse(2) schedule(static)
    for (int S = 0; S < sl; S++)
        for (int T = 0; T < tl; T++) {
            double acc = 0.0;
            const double it = (T + 0.5) * rt - Ct, is = (S + 0.5) * rs - Cs;
            for (int b = 0; b < g->nb; b++) {
                const double th = (double)(float)g->theta[b];
                const double t = it * cos(th) + is * sin(th);
                const float *R = filt + (long)b * g->np;
                int ip;
                double wl, wu;
                if (faxis_lerp(&ap, t, &ip, &wl, &wu)) acc += (wl * R[ip] + wu * R[ip + 1]) / ap.pitch * dtheta;
            }
            img[(long)S * tl + T] = acc;
        }
    free(ap.pos);
}

/* ----------------------------------------------------------------------
 * Curved (equiangular, cylindrical) cone-beam FDK with Parker short-scan
 * weights: My_toolbox/Implement_UnitedImaging_Data/ImplementUI.cu:6-153
 * (FDKUI), with ParkFunction.m:6-16.  Projections [beta][xi][gamma], gamma
 * fastest; gamma are channel angles (rad), xi are row positions (mm) on the
 * cylinder of radius Dsd; the source sits at (-D sin b, D cos b, 0).
 * ---------------------------------------------------------------------- */
typedef struct {
    int nt, ns, nz;
    double size[3];
    double D, Dsd;
    int nb, ng, nx;
    const double *beta, *gamma, *xi;
    double start, fan;
} or_curved;

/* ParkFunction.m:6-16 (fan = full fan angle).  The sine terms use beta
 * itself, not beta - start, as the reference does.  A channel at exactly
 * +-fan/2 can enter a ramp region of zero width through rounding, where the
 * reference divides by zero (NaN); its weight is 0 here. */
double or_parker(double beta, double start, double gama, double fan)
{
    if (beta >= start && beta < start + fan - 2 * gama) {
        const double den = fan / 2 - gama;
        if (!(den > 0.0)) return 0.0;
        double v = sin(OR_PI / 4 * beta / den);
        return v * v;
    }
    if (beta >= start + fan - 2 * gama && beta <= start + OR_PI - 2 * gama) return 1.0;
    if (beta > start + OR_PI - 2 * gama && beta <= start + OR_PI + fan) {
        const double den = fan / 2 + gama;
        if (!(den > 0.0)) return 0.0;
        double v = sin(OR_PI / 4 * (OR_PI + fan - beta) / den);
        return v * v;
    }
    return 0.0;
}

/* Row scale of ImplementUI.cu:14,45: P = Dsd / sqrt(Dsd^2 + xi^2). */
static double curved_row_scale(const or_curved *g, int x)
{
    return g->Dsd / sqrt(g->Dsd * g->Dsd + g->xi[x] * g->xi[x]);
}

/* Equiangular Shepp-Logan taps of row x (ImplementUI.cu:6-22), full scale:
 * G[0] = 2/(pi dgP)^2, G[n] = -2 n^2 / (pi^2 (4n^2-1) sin^2(n dgP)), dgP = dgamma * P. */
void or_curved_taps(const or_curved *g, int x, double *taps)
{
    const double P = curved_row_scale(g, x), dg = fabs(g->gamma[1] - g->gamma[0]) * P;
    taps[0] = 2.0 / ((OR_PI * dg) * (OR_PI * dg));
    for (int n = 1; n < g->ng; n++) {
        double sn = sin(n * dg);
        taps[n] = -2.0 * n * n / (OR_PI * OR_PI) / (4.0 * n * n - 1.0) / (sn * sn);
    }
}

/* Pre-weight and filter (ImplementUI.cu:24-78): per row xi, with P the row
 * scale, R * (D/P) cos(gamma P) * Parker(beta P, start P, gamma P, fan P),
 * then Rcov[c] = dgamma P * sum_g R'[g] G[|c-g|]. */
void or_curved_filter(const or_curved *g, const float *proj, double *out)
{
    const int ng = g->ng;
    const double dg0 = fabs(g->gamma[1] - g->gamma[0]);
#pragma omp parallel
    {
        double *taps = malloc(sizeof(double) * ng), *w = malloc(sizeof(double) * ng);
#pragma omp for collapse(2) schedule(static)
        for (int b = 0; b < g->nb; b++)
            for (int x = 0; x < g->nx; x++) {
                const double P = curved_row_scale(g, x);
                or_curved_taps(g, x, taps);
                const float *R = proj + ((long)b * g->nx + x) * ng;
                for (int c = 0; c < ng; c++) {
                    const double gp = g->gamma[c] * P;
                    const double W = or_parker(g->beta[b] * P, g->start * P, gp, g->fan * P);
                    w[c] = (double)R[c] * (g->D / P) * cos(gp) * W;
                }
                double *o = out + ((long)b * g->nx + x) * ng;
                for (int c = 0; c < ng; c++) {
                    double acc = 0.0;
                    for (int q = 0; q < ng; q++) acc += w[q] * taps[abs(c - q)];
                    o[c] = dg0 * P * acc;
                }
            }
        free(taps); free(w);
    }
}

/* Linear-interpolation coordinates on a double axis (interp modes as make_axis). */
typedef struct { double first, pitch, step, lo, hi; int n; const double *pos; } or_daxis;

static or_daxis make_daxis(const double *pos, int n)
{
    or_daxis a;
    a.first = pos[0];
    a.pitch = fabs(pos[1] - pos[0]);
    a.step = (pos[n - 1] - pos[0]) / (n - 1);
    if (g_interp_mode == 1) a.pitch = fabs(a.step);
    a.lo = pos[0] < pos[n - 1] ? pos[0] : pos[n - 1];
    a.hi = pos[0] > pos[n - 1] ? pos[0] : pos[n - 1];
    a.n = n;
    a.pos = pos;
    return a;
}

/* weights returned as fractions (sum 1) */
static int daxis_lerp(const or_daxis *a, double x, int *i1, double *w1, double *w2)
{
    if (!(x >= a->lo && x < a->hi)) return 0;
    int i;
    if (g_interp_mode == 1) {
        double t = (x - a->first) / a->step;
        i = (int)floor(t);
        if (i > a->n - 2) i = a->n - 2;
        if (i < 0) i = 0;
        *w2 = t - i;
        *w1 = 1.0 - *w2;
    } else {
        i = (int)floor(fabs(x - a->first) / a->pitch);
        if (i > a->n - 2) i = a->n - 2;
        *w2 = fabs(x - a->pos[i]) / a->pitch;
        *w1 = fabs(a->pos[i + 1] - x) / a->pitch;
    }
    *i1 = i;
    return 1;
}

/* Backprojection (ImplementUI.cu:80-153): voxel centre (i + o) r - C with
 * o = +0.5, or the reference's -0.5 when quirks (ImplementUI.cu:110: the
 * 1-based MATLAB offset of ImplementUIData.m:160 kept for 0-based indices);
 * gamma = atan(u / (D - s)), xi = Dsd tan(asin(z / |S - v|)), value =
 * bilinear sample * dbeta / L2 / (|S - v| / rho). */
void or_curved_bp(const or_curved *g, const float *filt, double *vol, int quirks)
{
    const double rt = g->size[0] / g->nt, rs = g->size[1] / g->ns, rz = g->size[2] / g->nz;
    const double Ct = g->size[0] / 2, Cs = g->size[1] / 2, Cz = g->size[2] / 2;
    const double o = quirks ? -0.5 : 0.5;
    const double dbeta = fabs(g->beta[1] - g->beta[0]);
    or_daxis ag = make_daxis(g->gamma, g->ng), ax = make_daxis(g->xi, g->nx);
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < g->nz; k++)
        for (int j = 0; j < g->ns; j++)
            for (int i = 0; i < g->nt; i++) {
                const double x = (i + o) * rt - Ct, y = (j + o) * rs - Cs, z = (k + o) * rz - Cz;
                double acc = 0.0;
                for (int b = 0; b < g->nb; b++) {
                    const double sb = sin(g->beta[b]), cb = cos(g->beta[b]);
                    const double xs = -g->D * sb, ys = g->D * cb;
                    const double L2 = (x - xs) * (x - xs) + (y - ys) * (y - ys) + z * z;
                    const double u = x * cb + y * sb, s = -x * sb + y * cb;
                    const double xi = g->Dsd * tan(asin(z / sqrt(L2)));
                    const double ga = atan(u / (g->D - s));
                    const double prop = sqrt(L2) / sqrt(L2 - z * z);
                    int ig, ix;
                    double g1, g2, x1, x2;
                    if (!daxis_lerp(&ag, ga, &ig, &g1, &g2) || !daxis_lerp(&ax, xi, &ix, &x1, &x2)) continue;
                    const float *F = filt + (long)b * g->nx * g->ng;
                    const double v = x1 * g1 * F[(long)ix * g->ng + ig] + x2 * g1 * F[(long)(ix + 1) * g->ng + ig] +
                                     x1 * g2 * F[(long)ix * g->ng + ig + 1] + x2 * g2 * F[(long)(ix + 1) * g->ng + ig + 1];
                    acc += v / L2 * (dbeta / prop);
                }
                vol[((long)k * g->ns + j) * g->nt + i] = acc;
            }
}

/* ----------------------------------------------------------------------
 * Error slice of the iterative FDK correction
 * (My_novel_algorithm/IterativeFDK/IterativeFDKerrCompute.cu:12-75): for voxel
 * slice t of a volume [z][s][t] and every view, row U of the integration
 * domain and detector sample P, the difference between the voxel reached at
 * (P (D-U)/D, U) in the rotated frame (its z moved along the ray) and the
 * voxel of the same (t, s) in slice z, weighted by the ramp tap at the
 * voxel's projected offset.  Output [beta][z][s].  The float/double types of
 * the reference's intermediates are kept, because each voxel index is a
 * floor of them; cos/sin of the view angles are float(cos(double beta)).
 * ---------------------------------------------------------------------- */
typedef struct {
    int nt, ns, nz;
    double size[3];
    double D;
    int nb, nu, np, ng, t;
    const float *beta, *u;
    const double *p, *g;
} or_fdkerr;

void or_fdk_error_slice(const or_fdkerr *a, const float *pic, double *out)
{
    const double r0 = a->size[0] / a->nt, r1 = a->size[1] / a->ns, r2 = a->size[2] / a->nz;
    const double Ct = a->size[0] / 2, Cs = a->size[1] / 2, Cz = a->size[2] / 2;
    const double D = a->D;
    const float dbeta = fabsf(a->beta[1] - a->beta[0]), dU = fabsf(a->u[1] - a->u[0]);
    const double PInt = fabs(a->p[1] - a->p[0]);
    const long plane = (long)a->nt * a->ns;
#pragma omp parallel for collapse(3) schedule(static)
    for (int b = 0; b < a->nb; b++)
        for (int z = 0; z < a->nz; z++)
            for (int s = 0; s < a->ns; s++) {
                const float cb = (float)cos((double)a->beta[b]), sb = (float)sin((double)a->beta[b]);
                const float image_t = (float)((a->t + 0.5) * r0 - Ct);
                const float image_s = (float)((s + 0.5) * r1 - Cs);
                const float image_z = (float)((z + 0.5) * r2 - Cz);
                const float ts = image_t * sb, tc = image_t * cb, ss = image_s * sb, sc = image_s * cb;
                const float dect_t = tc + ss, dect_s = sc - ts, dect_z = image_z;
                const float LR = (float)(D / (D - (double)dect_s));
                const float dtl = dect_t * LR;
                double acc = 0.0;
                for (int iu = 0; iu < a->nu; iu++) {
                    const float U = a->u[iu];
                    const float num = (dect_s - U) * dect_z;
                    const float Zd = (float)((double)dect_z + (double)num / (D - (double)dect_s));
                    for (int ip = 0; ip < a->np; ip++) {
                        const double P = a->p[ip];
                        const double ff = floor(((double)dtl - P) / PInt) + a->np;
                        if (!(ff >= 0 && ff < a->ng)) continue;
                        const float T = (float)(P * (D - (double)U) / D);
                        const float tcb = T * cb, usb = U * sb, tsb = T * sb, ucb = U * cb;
                        const float X = tcb - usb, Y = tsb + ucb;
                        const double fx = floor(((double)X + Ct) / r0), fy = floor(((double)Y + Cs) / r1);
                        const double fz = floor(((double)Zd + Cz) / r2);
                        if (!(fx >= 0 && fx < a->nt && fy >= 0 && fy < a->ns && fz >= 0 && fz < a->nz)) continue;
                        const long xy = (long)fy * a->nt + (long)fx;
                        const float diff = pic[(long)fz * plane + xy] - pic[(long)z * plane + xy];
                        acc += 0.5 * ((double)LR * LR) * diff * sqrt(1 + (P / D) * (P / D)) * a->g[(long)ff] * D /
                               sqrt(D * D + P * P) * dU * PInt * dbeta;
                    }
                }
                out[((long)b * a->nz + z) * a->ns + s] = acc;
            }
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Thread count of the OpenMP loops above (the CPU baseline's 1-thread figure). */
void or_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
