/*
 * mrf_oracle_mt.c -- multi-threaded restatement of the oracle's DILU-BiCGStab
 * (orc_solve with solver BiCGStab), for the all-cores CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY (see mrf_oracle.h): bench.py's cpu_baseline leg
 * times it beside the single-core oracle; tests/ check it against orc_solve.
 *
 * Same algorithm and per-row arithmetic order as mrf_oracle.c, run on all
 * host cores with OpenMP:
 *   - A*x row by row from a CSR of each region's LDU in face order (lower
 *     neighbours, then upper neighbours), then the interface terms;
 *   - the DILU sweeps level by level (a cell's level is one more than its
 *     dependencies'): within a level every row is independent and keeps the
 *     face order of its own subtractions, so the sweeps are bit-identical to
 *     the sequential face loops;
 *   - vector updates per cell;
 *   - global sums compensated per thread and merged in thread order (equal
 *     to the sequential compensated sum up to the rounding of its last bit).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#include "mrf_oracle.h"

typedef struct {
    int n;
    int32_t *off, *col;      /* off-diagonal entries of each row in face order */
    double *val;
    int32_t *loOff, *loCol;  /* forward DILU: lower neighbours, ascending face */
    double *loVal;
    int32_t *upOff, *upCol;  /* backward DILU: upper neighbours, descending face */
    double *upVal;
    int nLevF, nLevB;
    int32_t *levFOff, *levF, *levBOff, *levB;   /* cells grouped by level */
} mt_region;

static int32_t *levels(int n, const int32_t *off, const int32_t *col, int forward, int *nLev, int32_t **cells)
{
    int32_t *lev = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    int maxL = 0;
    for (int k = 0; k < n; ++k) {
        int c = forward ? k : n - 1 - k, L = 0;
        for (int q = off[c]; q < off[c + 1]; ++q)
            if (lev[col[q]] + 1 > L) L = lev[col[q]] + 1;
        lev[c] = L;
        if (L > maxL) maxL = L;
    }
    *nLev = maxL + 1;
    int32_t *lo = (int32_t *)calloc((size_t)maxL + 2, sizeof(int32_t));
    for (int c = 0; c < n; ++c) lo[lev[c] + 1]++;
    for (int l = 0; l <= maxL; ++l) lo[l + 1] += lo[l];
    *cells = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxL + 1));
    memcpy(pos, lo, sizeof(int32_t) * (size_t)(maxL + 1));
    for (int c = 0; c < n; ++c) (*cells)[pos[lev[c]]++] = c;
    free(pos);
    free(lev);
    return lo;
}

static void mt_build(const orc_ldu *L, mt_region *R)
{
    const int n = L->nCells, F = L->nFaces;
    const double *lo = L->symmetric ? L->upper : L->lower;
    R->n = n;
    int32_t *nl = (int32_t *)calloc((size_t)n, sizeof(int32_t)), *nu = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    for (int f = 0; f < F; ++f) {
        nl[L->uAddr[f]]++;
        nu[L->lAddr[f]]++;
    }
    R->off = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    R->loOff = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    R->upOff = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    for (int c = 0; c < n; ++c) {
        R->off[c + 1] = R->off[c] + nl[c] + nu[c];
        R->loOff[c + 1] = R->loOff[c] + nl[c];
        R->upOff[c + 1] = R->upOff[c] + nu[c];
    }
    R->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * F + 1));
    R->val = (double *)malloc(sizeof(double) * (size_t)(2 * F + 1));
    R->loCol = (int32_t *)malloc(sizeof(int32_t) * (size_t)(F + 1));
    R->loVal = (double *)malloc(sizeof(double) * (size_t)(F + 1));
    R->upCol = (int32_t *)malloc(sizeof(int32_t) * (size_t)(F + 1));
    R->upVal = (double *)malloc(sizeof(double) * (size_t)(F + 1));
    /* lower entries of row u: faces (l, u) ascending; upper entries of row l:
       faces (l, u) ascending after the lower ones (A*x order), descending
       for the backward sweep */
    int32_t *fl = (int32_t *)malloc(sizeof(int32_t) * (size_t)n), *fu = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int c = 0; c < n; ++c) {
        fl[c] = 0;
        fu[c] = 0;
    }
    for (int f = 0; f < F; ++f) {
        int l = L->lAddr[f], u = L->uAddr[f];
        int q = R->off[u] + fl[u];
        R->col[q] = l;
        R->val[q] = lo[f];
        R->loCol[R->loOff[u] + fl[u]] = l;
        R->loVal[R->loOff[u] + fl[u]] = lo[f];
        fl[u]++;
    }
    for (int f = 0; f < F; ++f) {
        int l = L->lAddr[f], u = L->uAddr[f];
        int q = R->off[l] + nl[l] + fu[l];
        R->col[q] = u;
        R->val[q] = L->upper[f];
        fu[l]++;
        int b = R->upOff[l] + (nu[l] - fu[l]);   /* descending face order */
        R->upCol[b] = u;
        R->upVal[b] = L->upper[f];
    }
    free(fl);
    free(fu);
    free(nl);
    free(nu);
    R->levFOff = levels(n, R->loOff, R->loCol, 1, &R->nLevF, &R->levF);
    R->levBOff = levels(n, R->upOff, R->upCol, 0, &R->nLevB, &R->levB);
}

static void mt_free(mt_region *R)
{
    free(R->off); free(R->col); free(R->val);
    free(R->loOff); free(R->loCol); free(R->loVal);
    free(R->upOff); free(R->upCol); free(R->upVal);
    free(R->levFOff); free(R->levF); free(R->levBOff); free(R->levB);
}

typedef struct {
    double hi, lo;
} csum_t;

static inline void cs_add(csum_t *a, double x)
{
    double s = a->hi + x;
    double bb = s - a->hi;
    a->lo += (a->hi - (s - bb)) + (x - bb);
    a->hi = s;
}

/* sum of f(g, c) over all cells, compensated per thread, merged in thread order */
#define MT_SUM(RESULT, EXPR)                                                          \
    do {                                                                              \
        int nt_ = omp_get_max_threads();                                              \
        csum_t *part_ = (csum_t *)calloc((size_t)nt_, sizeof(csum_t));                \
        for (int g = 0; g < nR; ++g) {                                                \
            const int n_ = M[g].n;                                                    \
            _Pragma("omp parallel")                                                   \
            {                                                                         \
                csum_t s_ = {0.0, 0.0};                                               \
                _Pragma("omp for schedule(static)")                                   \
                for (int c = 0; c < n_; ++c) cs_add(&s_, (EXPR));                     \
                csum_t *p_ = &part_[omp_get_thread_num()];                            \
                cs_add(p_, s_.hi);                                                    \
                cs_add(p_, s_.lo);                                                    \
            }                                                                         \
        }                                                                             \
        csum_t t_ = {0.0, 0.0};                                                       \
        for (int i_ = 0; i_ < nt_; ++i_) {                                            \
            cs_add(&t_, part_[i_].hi);                                                \
            cs_add(&t_, part_[i_].lo);                                                \
        }                                                                             \
        free(part_);                                                                  \
        RESULT = t_.hi + t_.lo;                                                       \
    } while (0)

static void mt_amul(const orc_system *sys, const mt_region *M, double *const *x, double **y)
{
    for (int g = 0; g < sys->nRegions; ++g) {
        const orc_ldu *L = &sys->regions[g];
        const mt_region *R = &M[g];
        const double *xr = x[g];
        double *yr = y[g];
#pragma omp parallel for schedule(static)
        for (int c = 0; c < R->n; ++c) {
            double s = L->diag[c] * xr[c];
            for (int q = R->off[c]; q < R->off[c + 1]; ++q) s += R->val[q] * xr[R->col[q]];
            yr[c] = s;
        }
    }
    for (int i = 0; i < sys->nInterfaces; ++i) {
        const orc_interface *I = &sys->interfaces[i];
        for (int side = 0; side < 2; ++side) {
            const int n = side ? I->nB : I->nA;
            const int32_t *off = side ? I->offB : I->offA, *idx = side ? I->idxB : I->idxA;
            const int32_t *fc = side ? I->fcB : I->fcA, *fcN = side ? I->fcA : I->fcB;
            const double *w = side ? I->wB : I->wA, *coeff = side ? I->coeffB : I->coeffA;
            const double *xN = x[side ? I->regionA : I->regionB];
            double *yO = y[side ? I->regionB : I->regionA];
            for (int f = 0; f < n; ++f) {
                double pnf = 0.0;
                for (int p = off[f]; p < off[f + 1]; ++p) pnf += w[p] * xN[fcN[idx[p]]];
                yO[fc[f]] -= coeff[f] * pnf;
            }
        }
    }
}

static void mt_precondition(const orc_system *sys, const mt_region *M, double *const *rD, double **w,
                            double *const *r)
{
    for (int g = 0; g < sys->nRegions; ++g) {
        const mt_region *R = &M[g];
        const double *d = rD[g], *b = r[g];
        double *x = w[g];
        /* one team for both sweeps; the implicit barrier of each level's
           work-sharing loop orders the levels */
#pragma omp parallel
        {
            for (int l = 0; l < R->nLevF; ++l) {
#pragma omp for schedule(static)
                for (int k = R->levFOff[l]; k < R->levFOff[l + 1]; ++k) {
                    int c = R->levF[k];
                    double v = b[c] * d[c];
                    for (int q = R->loOff[c]; q < R->loOff[c + 1]; ++q) v -= d[c] * R->loVal[q] * x[R->loCol[q]];
                    x[c] = v;
                }
            }
            for (int l = 0; l < R->nLevB; ++l) {
#pragma omp for schedule(static)
                for (int k = R->levBOff[l]; k < R->levBOff[l + 1]; ++k) {
                    int c = R->levB[k];
                    double v = x[c];
                    for (int q = R->upOff[c]; q < R->upOff[c + 1]; ++q) v -= d[c] * R->upVal[q] * x[R->upCol[q]];
                    x[c] = v;
                }
            }
        }
    }
}

static int mt_stop(orc_perf *p, double tol, double relTol, int minIter, int maxIter)
{
    if (!isfinite(p->finalResidual)) {   /* diverged, as the sequential solve */
        p->converged = 0;
        p->singular = 2;
        return 1;
    }
    if (p->nIterations < minIter) return 0;
    int conv = p->finalResidual < tol || (relTol > 1e-20 && p->finalResidual < relTol * p->initialResidual);
    p->converged = conv;
    return p->nIterations >= maxIter || conv;
}

static double **vec_new(const orc_system *sys)
{
    double **v = (double **)malloc(sizeof(double *) * (size_t)sys->nRegions);
    for (int g = 0; g < sys->nRegions; ++g) v[g] = (double *)calloc((size_t)sys->regions[g].nCells, sizeof(double));
    return v;
}

static void vec_free(const orc_system *sys, double **v)
{
    for (int g = 0; g < sys->nRegions; ++g) free(v[g]);
    free(v);
}

int orc_solve_mt(orc_system *sys, double tol, double relTol, int minIter, int maxIter, orc_perf *perf)
{
    const int nR = sys->nRegions;
    mt_region *M = (mt_region *)calloc((size_t)nR, sizeof(mt_region));
    for (int g = 0; g < nR; ++g) mt_build(&sys->regions[g], &M[g]);
    double **rD = vec_new(sys);
    orc_dilu_factor(sys, rD);
    double **x = (double **)malloc(sizeof(double *) * (size_t)nR), **b = (double **)malloc(sizeof(double *) * (size_t)nR);
    for (int g = 0; g < nR; ++g) {
        x[g] = sys->regions[g].psi;
        b[g] = sys->regions[g].source;
    }
    double **p = vec_new(sys), **r = vec_new(sys), **tmp = vec_new(sys), **ph = vec_new(sys), **v = vec_new(sys);
    double **s = vec_new(sys), **sh = vec_new(sys), **t = vec_new(sys), **rw = vec_new(sys);
    mt_amul(sys, M, x, p);
    for (int g = 0; g < nR; ++g)
#pragma omp parallel for schedule(static)
        for (int c = 0; c < M[g].n; ++c) r[g][c] = b[g][c] - p[g][c];
    /* normFactor: gAverage(x), sumA*xRef, sum(|Ax - tmp| + |b - tmp|) + small */
    double sx, nf, sr;
    MT_SUM(sx, x[g][c]);
    long nC = 0;
    for (int g = 0; g < nR; ++g) nC += M[g].n;
    const double xRef = sx / (double)nC;
    orc_sumA(sys, tmp);
    MT_SUM(nf, fabs(p[g][c] - tmp[g][c] * xRef) + fabs(b[g][c] - tmp[g][c] * xRef));
    nf += 1e-20;
    MT_SUM(sr, fabs(r[g][c]));
    perf->initialResidual = sr / nf;
    perf->finalResidual = perf->initialResidual;
    perf->nIterations = 0;
    perf->converged = 0;
    perf->singular = 0;
    if (!mt_stop(perf, tol, relTol, minIter, maxIter)) {
        const double great = 1e20;
        double rho = great, rhoOld, alpha = 0.0, omega = great, beta;
        for (int g = 0; g < nR; ++g) {
            memset(p[g], 0, sizeof(double) * (size_t)M[g].n);
            memcpy(rw[g], r[g], sizeof(double) * (size_t)M[g].n);
        }
        do {
            rhoOld = rho;
            MT_SUM(rho, rw[g][c] * r[g][c]);
            beta = rho / rhoOld * (alpha / omega);
            for (int g = 0; g < nR; ++g)
#pragma omp parallel for schedule(static)
                for (int c = 0; c < M[g].n; ++c) p[g][c] = r[g][c] + beta * p[g][c] - beta * omega * v[g][c];
            mt_precondition(sys, M, rD, ph, p);
            mt_amul(sys, M, ph, v);
            double rwv;
            MT_SUM(rwv, rw[g][c] * v[g][c]);
            if (rho == 0.0 || rwv == 0.0) {
                perf->singular = 1;
                break;
            }
            alpha = rho / rwv;
            for (int g = 0; g < nR; ++g)
#pragma omp parallel for schedule(static)
                for (int c = 0; c < M[g].n; ++c) s[g][c] = r[g][c] - alpha * v[g][c];
            mt_precondition(sys, M, rD, sh, s);
            mt_amul(sys, M, sh, t);
            double ts, tt;
            MT_SUM(ts, t[g][c] * s[g][c]);
            MT_SUM(tt, t[g][c] * t[g][c]);
            /* s vanished: take the step with omega = 0, as the sequential
               solve does */
            const int sZero = tt == 0.0;
            omega = sZero ? 0.0 : ts / tt;
            for (int g = 0; g < nR; ++g)
#pragma omp parallel for schedule(static)
                for (int c = 0; c < M[g].n; ++c) {
                    x[g][c] = x[g][c] + alpha * ph[g][c] + omega * sh[g][c];
                    r[g][c] = s[g][c] - omega * t[g][c];
                }
            MT_SUM(sr, fabs(r[g][c]));
            perf->finalResidual = sr / nf;
            perf->nIterations++;
            if (mt_stop(perf, tol, relTol, minIter, maxIter)) break;
            if (sZero) {
                perf->singular = 1;
                break;
            }
        } while (1);
    }
    vec_free(sys, p); vec_free(sys, r); vec_free(sys, tmp); vec_free(sys, ph); vec_free(sys, v);
    vec_free(sys, s); vec_free(sys, sh); vec_free(sys, t); vec_free(sys, rw); vec_free(sys, rD);
    for (int g = 0; g < nR; ++g) mt_free(&M[g]);
    free(M);
    free(x);
    free(b);
    return 0;
}

int orc_threads(void) { return omp_get_max_threads(); }
