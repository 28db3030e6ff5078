/*
 * mrf_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the monolithic
 * interface-coupled temperature solve of multiRegionFoam.
 *
 * Plain C99, FP64, int32 addressing, single-threaded: the precision and index
 * width of the reference build (compile_commands.json:9-10, -DWM_DP,
 * -DWM_LABEL_SIZE=32).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library.
 *
 * The arithmetic of the hot path lives in foam-extend 4.1, which the
 * reference links but does not vendor (src/multiRegionSystem/Make/options:
 * 6,21-29).  Every function below therefore cites the reference call site it
 * stands behind and restates the published foam-extend algorithm.
 */
#include "mrf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Deterministic case randomness                                              */
/* ------------------------------------------------------------------------- */

/* splitmix64 finaliser on (seed, index); uniform double in [0,1) from the top
   53 bits.  Shared definition with the product's case generator so both
   produce identical seeded fields. */
double orc_u01(uint64_t seed, uint64_t index)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (index + 1ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------- */
/* Region geometry                                                            */
/* ------------------------------------------------------------------------- */

typedef struct {
    const orc_region_spec *spec;
    int n[3];
    int N, F;
    double d[3];       /* cell size per axis */
    double area[3];    /* face area normal to axis */
    double *k;         /* per cell */
    double *U;         /* per cell, 3 components (fluid only) */
} region_geom;

static void tangential_axes(int a, int *b, int *c)
{
    if (a == 0) { *b = 1; *c = 2; }
    else if (a == 1) { *b = 0; *c = 2; }
    else { *b = 0; *c = 1; }
}

/* Cell of boundary face `f` on side `s`; faces of a side are ordered with the
   first tangential axis fastest. */
static int side_face_cell(const region_geom *g, int s, int f)
{
    int a = s / 2, hi = s % 2, b, c, idx[3];
    tangential_axes(a, &b, &c);
    idx[a] = hi ? g->n[a] - 1 : 0;
    idx[b] = f % g->n[b];
    idx[c] = f / g->n[b];
    return idx[0] + g->n[0] * (idx[1] + g->n[1] * idx[2]);
}

static int side_nfaces(const region_geom *g, int s)
{
    int a = s / 2, b, c;
    tangential_axes(a, &b, &c);
    if (g->spec->patch_type[s] == ORC_EMPTY) return 0;
    return g->n[b] * g->n[c];
}

/* Parabolic profile factor: prod over profile axes of 4*eta*(1-eta), with
   eta the cell-centre (or face-centre) relative coordinate (i+0.5)/n. */
static double profile_factor(const orc_region_spec *sp, const int idx[3])
{
    double f = 1.0;
    for (int a = 0; a < 3; ++a) {
        if (sp->profile_mask & (1 << a)) {
            double eta = ((double)idx[a] + 0.5) / (double)sp->n[a];
            f *= 4.0 * eta * (1.0 - eta);
        }
    }
    return f;
}

static void region_geom_init(region_geom *g, const orc_region_spec *sp)
{
    g->spec = sp;
    for (int a = 0; a < 3; ++a) {
        g->n[a] = sp->n[a];
        g->d[a] = sp->length[a] / (double)sp->n[a];
    }
    g->area[0] = g->d[1] * g->d[2];
    g->area[1] = g->d[0] * g->d[2];
    g->area[2] = g->d[0] * g->d[1];
    g->N = g->n[0] * g->n[1] * g->n[2];
    g->F = (g->n[0] - 1) * g->n[1] * g->n[2]
         + g->n[0] * (g->n[1] - 1) * g->n[2]
         + g->n[0] * g->n[1] * (g->n[2] - 1);

    /* conductivity: conductTemperature.C:139-141 / transportTemperature.C:
       96-102 read k per cell; here uniform or block-random log-uniform */
    g->k = (double *)malloc(sizeof(double) * (size_t)g->N);
    for (int kk = 0; kk < g->n[2]; ++kk)
    for (int j = 0; j < g->n[1]; ++j)
    for (int i = 0; i < g->n[0]; ++i) {
        int c = i + g->n[0] * (j + g->n[1] * kk);
        if (sp->k_block > 0) {
            int B = sp->k_block;
            int nbx = (g->n[0] + B - 1) / B, nby = (g->n[1] + B - 1) / B;
            uint64_t bidx = (uint64_t)(i / B)
                          + (uint64_t)nbx * ((uint64_t)(j / B) + (uint64_t)nby * (uint64_t)(kk / B));
            double u = orc_u01(sp->k_seed, bidx);
            double lo = log(sp->k_min), hi = log(sp->k_max);
            g->k[c] = exp(lo + u * (hi - lo));
        } else {
            g->k[c] = sp->k_uniform;
        }
    }

    /* frozen velocity (transportTemperature.C:81-94: phi = interpolate(U)
       & Sf): cell velocities carry the parabolic profile along flow_axis */
    g->U = NULL;
    if (sp->transport) {
        g->U = (double *)calloc((size_t)g->N * 3, sizeof(double));
        for (int kk = 0; kk < g->n[2]; ++kk)
        for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
            int idx[3] = { i, j, kk };
            int c = i + g->n[0] * (j + g->n[1] * kk);
            g->U[3 * c + sp->flow_axis] = sp->Umax * profile_factor(sp, idx);
        }
    }
}

/* Divergence-free velocity noise: the face flux is the discrete curl of a
   seeded edge vector potential A (uniform in +-noise*Umax*h/4, zero on
   boundary edges), so every cell's net noise flux cancels and boundary faces
   carry none.  Returns the noise flux of the internal face normal to axis a
   whose owner cell is (i,j,k). */
static double edge_pot(const region_geom *g, int comp, int I, int J, int K)
{
    const orc_region_spec *sp = g->spec;
    int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    uint64_t id;
    /* edges along `comp`; the other two node indices must be interior */
    if (comp == 0) {
        if (J <= 0 || J >= ny || K <= 0 || K >= nz) return 0.0;
        id = (uint64_t)I + (uint64_t)nx * ((uint64_t)J + (uint64_t)(ny + 1) * (uint64_t)K);
    } else if (comp == 1) {
        if (I <= 0 || I >= nx || K <= 0 || K >= nz) return 0.0;
        id = (uint64_t)I + (uint64_t)(nx + 1) * ((uint64_t)J + (uint64_t)ny * (uint64_t)K);
    } else {
        if (I <= 0 || I >= nx || J <= 0 || J >= ny) return 0.0;
        id = (uint64_t)I + (uint64_t)(nx + 1) * ((uint64_t)J + (uint64_t)(ny + 1) * (uint64_t)K);
    }
    double xi = orc_u01(sp->vel_seed, 3ULL * id + (uint64_t)comp);
    return 0.25 * sp->noise * sp->Umax * g->d[comp] * (2.0 * xi - 1.0);
}

static double noise_flux(const region_geom *g, int a, int i, int j, int k)
{
    const double *d = g->d;
    if (g->spec->noise <= 0.0) return 0.0;
    if (a == 0) {      /* x-face at node plane I = i+1, covering (j,k) */
        int I = i + 1;
        return (edge_pot(g, 2, I, j + 1, k) - edge_pot(g, 2, I, j, k)) * d[2]
             - (edge_pot(g, 1, I, j, k + 1) - edge_pot(g, 1, I, j, k)) * d[1];
    } else if (a == 1) { /* y-face at node plane J = j+1, covering (i,k) */
        int J = j + 1;
        return (edge_pot(g, 0, i, J, k + 1) - edge_pot(g, 0, i, J, k)) * d[0]
             - (edge_pot(g, 2, i + 1, J, k) - edge_pot(g, 2, i, J, k)) * d[2];
    } else {           /* z-face at node plane K = k+1, covering (i,j) */
        int K = k + 1;
        return (edge_pot(g, 1, i + 1, j, K) - edge_pot(g, 1, i, j, K)) * d[1]
             - (edge_pot(g, 0, i, j + 1, K) - edge_pot(g, 0, i, j, K)) * d[0];
    }
}

static void region_geom_free(region_geom *g)
{
    free(g->U);
    /* g->k ownership moves to orc_ldu */
}

/* ------------------------------------------------------------------------- */
/* GGI weights for axis-aligned rectangular faces                             */
/* ------------------------------------------------------------------------- */

typedef struct { int j; double ov; } ov1d;
typedef struct { int n; ov1d *items; } ovlist;

/* 1-D interval overlaps between side A intervals [oA+i dA, oA+(i+1) dA] and
   side B intervals; slivers below 1e-10 of the narrower width are dropped
   (GGI non-overlapping face tolerance, ggiInterfaceToInterfaceMapping.C:
   71-72 passes SMALL). */
static ovlist *overlaps_1d(int nA, double oA, double dA, int nB, double oB, double dB)
{
    ovlist *L = (ovlist *)calloc((size_t)nA, sizeof(ovlist));
    for (int i = 0; i < nA; ++i) {
        double lo = oA + (double)i * dA, hi = oA + (double)(i + 1) * dA;
        int j0 = (int)floor((lo - oB) / dB) - 1, j1 = (int)floor((hi - oB) / dB) + 1;
        if (j0 < 0) j0 = 0;
        if (j1 > nB - 1) j1 = nB - 1;
        L[i].items = (ov1d *)malloc(sizeof(ov1d) * (size_t)(j1 - j0 + 2 > 1 ? j1 - j0 + 2 : 1));
        L[i].n = 0;
        for (int j = j0; j <= j1; ++j) {
            double blo = oB + (double)j * dB, bhi = oB + (double)(j + 1) * dB;
            double ov = fmin(hi, bhi) - fmax(lo, blo);
            double wmin = fmin(hi - lo, bhi - blo);
            if (ov > 1e-10 * wmin) {
                L[i].items[L[i].n].j = j;
                L[i].items[L[i].n].ov = ov;
                L[i].n++;
            }
        }
    }
    return L;
}

static void free_ovlist(ovlist *L, int n)
{
    for (int i = 0; i < n; ++i) free(L[i].items);
    free(L);
}

/* Weights for faces of side (gA,sA) gathering from faces of side (gB,sB):
   w_ij = overlap area(i,j) / sum_j overlap area(i,j) (GGI with rescaled
   weighting factors, ggiInterfaceToInterfaceMapping.C:62-77).  Pairs of a
   face are ordered by ascending B face index. */
static void ggi_weights(const region_geom *gA, int sA, const region_geom *gB, int sB,
                        int32_t **off, int32_t **idx, double **w)
{
    int a = sA / 2, b, c;
    tangential_axes(a, &b, &c);
    const orc_region_spec *pA = gA->spec, *pB = gB->spec;
    int nAb = gA->n[b], nAc = gA->n[c], nBb = gB->n[b];
    ovlist *Lb = overlaps_1d(nAb, pA->origin[b], gA->d[b], gB->n[b], pB->origin[b], gB->d[b]);
    ovlist *Lc = overlaps_1d(nAc, pA->origin[c], gA->d[c], gB->n[c], pB->origin[c], gB->d[c]);
    int nf = nAb * nAc;
    size_t cap = 0;
    for (int f = 0; f < nf; ++f) cap += (size_t)Lb[f % nAb].n * (size_t)Lc[f / nAb].n;
    *off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nf + 1));
    *idx = (int32_t *)malloc(sizeof(int32_t) * (cap ? cap : 1));
    *w = (double *)malloc(sizeof(double) * (cap ? cap : 1));
    size_t p = 0;
    double faceArea = gA->area[a];
    for (int f = 0; f < nf; ++f) {
        const ovlist *lb = &Lb[f % nAb], *lc = &Lc[f / nAb];
        (*off)[f] = (int32_t)p;
        size_t p0 = p;
        double sum = 0.0;
        for (int q = 0; q < lc->n; ++q)
            for (int r = 0; r < lb->n; ++r) {
                double ar = lb->items[r].ov * lc->items[q].ov;
                (*idx)[p] = lb->items[r].j + nBb * lc->items[q].j;
                (*w)[p] = ar;
                sum += ar;
                ++p;
            }
        if (sum > 1e-12 * faceArea) {
            for (size_t q = p0; q < p; ++q) (*w)[q] = (*w)[q] / sum;
        } else {
            p = p0;   /* uncovered face: no contributions */
        }
    }
    (*off)[nf] = (int32_t)p;
    free_ovlist(Lb, nAb);
    free_ovlist(Lc, nAc);
}

/* Interface conductance of one side's faces.  Restates
   monolithicThermalDiffusivityFvPatchScalarField.C:167-277 with radiation
   off (Qr = 0): kOwn = fOwn/(1-w)/|d|, kNei = fNei/w/|d|,
   k = kOwn*kNei/(deltaCoeffs*(kOwn+kNei)) clamped to [0.01,100]*kHarm, and
   the laplacian's coupled-patch coefficient gammaMagSf*deltaCoeffs
   (fvm::laplacian internal/boundary coeffs of a coupled patch). */
static void iface_coeffs(const region_geom *gO, int sO, const region_geom *gN, int sN,
                         const int32_t *off, const int32_t *idx, const double *w,
                         int32_t *fcO, double *coeff, double *magSfO)
{
    int a = sO / 2;
    int nf = side_nfaces(gO, sO);
    double dOwn = 0.5 * gO->d[a], dNei = 0.5 * gN->d[a];
    double magDelta = dOwn + dNei;
    double wP = dNei / magDelta;
    double delta = 1.0 / magDelta;
    for (int f = 0; f < nf; ++f) {
        int cO = side_face_cell(gO, sO, f);
        fcO[f] = cO;
        magSfO[f] = gO->area[a];
        double fOwn = gO->k[cO];
        double fNei = 0.0;
        for (int p = off[f]; p < off[f + 1]; ++p)
            fNei += w[p] * gN->k[side_face_cell(gN, sN, idx[p])];
        double kOwn = fOwn / (1.0 - wP) / magDelta;
        double kNei = fNei / wP / magDelta;
        double kf = kOwn * kNei;
        kf /= delta * (kOwn + kNei);
        double kHarm = 1.0 / (wP / fOwn + (1.0 - wP) / fNei);
        kf = fmax(fmin(kf, 100.0 * kHarm), 0.01 * kHarm);
        double gammaMagSf = kf * gO->area[a];
        coeff[f] = gammaMagSf * delta;
    }
}

/* ------------------------------------------------------------------------- */
/* FV assembly                                                                */
/* ------------------------------------------------------------------------- */

/* Per region: conductTemperature.C:137-142 (ddt(rho cv,T) == laplacian(k,T))
   and transportTemperature.C:131-140 (rho cp (ddt(T) + div(phi,T)) ==
   laplacian(k,T)), steady ddt, laplacianSchemes Gauss linear (uniform
   orthogonal hex: no correction), divSchemes Gauss upwind.  Coefficients
   are combined in the order fvMatrix arithmetic does: A == B is A - B,
   negSumDiag accumulates in face order, then addBoundaryDiag /
   addBoundarySource(couples=false) in patch order as the coupled solve does
   before calling the coupled solver (multiRegionSystem.C:150-153). */
static void assemble_region(const region_geom *g, orc_ldu *L,
                            const double *ifaceCoeffPerSide[6], const int32_t *ifaceFc[6])
{
    const orc_region_spec *sp = g->spec;
    int N = g->N, F = g->F;
    int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    L->nCells = N;
    L->nFaces = F;
    L->lAddr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(F ? F : 1));
    L->uAddr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(F ? F : 1));
    L->lower = (double *)malloc(sizeof(double) * (size_t)(F ? F : 1));
    L->upper = (double *)malloc(sizeof(double) * (size_t)(F ? F : 1));
    L->diag = (double *)calloc((size_t)N, sizeof(double));
    L->source = (double *)calloc((size_t)N, sizeof(double));
    L->psi = (double *)malloc(sizeof(double) * (size_t)N);
    L->symmetric = sp->transport ? 0 : 1;
    L->k = g->k;
    L->cellSize = (double *)malloc(sizeof(double) * 3);
    for (int a = 0; a < 3; ++a) L->cellSize[a] = g->d[a];
    for (int c = 0; c < N; ++c) L->psi[c] = sp->T0;

    double *diagL = (double *)calloc((size_t)N, sizeof(double));
    double *diagC = (double *)calloc((size_t)N, sizeof(double));

    /* internal faces, upper-triangular order: owner ascending, then
       neighbour c+1, c+nx, c+nx*ny */
    int f = 0;
    for (int c = 0; c < N; ++c) {
        int i = c % nx, j = (c / nx) % ny, kk = c / (nx * ny);
        int nb[3] = { i + 1 < nx ? c + 1 : -1,
                      j + 1 < ny ? c + nx : -1,
                      kk + 1 < nz ? c + nx * ny : -1 };
        for (int a = 0; a < 3; ++a) {
            if (nb[a] < 0) continue;
            int u = nb[a];
            L->lAddr[f] = c;
            L->uAddr[f] = u;
            /* laplacian: upper = deltaCoeffs * (interpolate(k) * magSf) */
            double kf = 0.5 * g->k[c] + 0.5 * g->k[u];
            double gammaMagSf = kf * g->area[a];
            double aL = (1.0 / g->d[a]) * gammaMagSf;
            diagL[c] -= aL;
            diagL[u] -= aL;
            if (sp->transport) {
                double phi = (0.5 * g->U[3 * c + a] + 0.5 * g->U[3 * u + a]) * g->area[a]
                           + noise_flux(g, a, i, j, kk);
                double wup = phi >= 0.0 ? 1.0 : 0.0;
                double lowerC = -wup * phi;
                double upperC = lowerC + phi;
                diagC[c] -= lowerC;
                diagC[u] -= upperC;
                L->upper[f] = sp->rhoCp * upperC - aL;
                L->lower[f] = sp->rhoCp * lowerC - aL;
            } else {
                L->upper[f] = 0.0 - aL;
                L->lower[f] = L->upper[f];
            }
            ++f;
        }
    }
    for (int c = 0; c < N; ++c)
        L->diag[c] = sp->transport ? sp->rhoCp * diagC[c] - diagL[c] : 0.0 - diagL[c];

    /* boundary patches in patch order */
    for (int s = 0; s < 6; ++s) {
        int type = sp->patch_type[s];
        if (type == ORC_EMPTY) continue;
        int a = s / 2, hi = s % 2, b, cc;
        tangential_axes(a, &b, &cc);
        int nf = side_nfaces(g, s);
        double magSf = g->area[a];
        double deltab = 1.0 / (0.5 * g->d[a]);
        for (int fb = 0; fb < nf; ++fb) {
            int cell = side_face_cell(g, s, fb);
            double ic, bc;
            if (type == ORC_INTERFACE) {
                /* coupled patch: internalCoeffs == boundaryCoeffs ==
                   gammaMagSf*deltaCoeffs; boundaryCoeffs go to the
                   interface (Amul), not to the source */
                if (ifaceFc[s][fb] != cell) abort();
                L->diag[cell] += ifaceCoeffPerSide[s][fb];
                continue;
            }
            double kb = g->k[cell];
            double gammaMagSf = kb * magSf;
            double icL, bcL;
            double Tb = sp->patch_value[s];
            if (type == ORC_FIXED_VALUE) {
                icL = gammaMagSf * (-deltab);
                bcL = -gammaMagSf * (deltab * Tb);
            } else {
                icL = gammaMagSf * 0.0;
                bcL = -gammaMagSf * 0.0;
            }
            if (sp->transport) {
                double phi = 0.0;
                int kind = sp->patch_flux[s];
                if (kind == ORC_INLET) {
                    /* fixedValue U: the undisturbed profile at the face */
                    int idx[3];
                    idx[a] = hi ? g->n[a] - 1 : 0;
                    idx[b] = fb % g->n[b];
                    idx[cc] = fb / g->n[b];
                    double Ub = a == sp->flow_axis ? sp->Umax * profile_factor(sp, idx) : 0.0;
                    phi = (hi ? 1.0 : -1.0) * Ub * magSf;
                } else if (kind == ORC_OUTLET) {
                    /* zeroGradient U: the face cell's velocity */
                    phi = (hi ? 1.0 : -1.0) * g->U[3 * cell + a] * magSf;
                }
                double icC, bcC;
                if (type == ORC_FIXED_VALUE) {
                    icC = phi * 0.0;
                    bcC = -phi * Tb;
                } else {
                    icC = phi * 1.0;
                    bcC = -phi * 0.0;
                }
                ic = sp->rhoCp * icC - icL;
                bc = sp->rhoCp * bcC - bcL;
            } else {
                ic = 0.0 - icL;
                bc = 0.0 - bcL;
            }
            L->diag[cell] += ic;
            L->source[cell] += bc;
        }
    }
    free(diagL);
    free(diagC);
}

int orc_build(int nRegions, const orc_region_spec *regions,
              int nInterfaces, const orc_interface_spec *ifs, orc_system *out)
{
    memset(out, 0, sizeof(*out));
    region_geom *G = (region_geom *)calloc((size_t)nRegions, sizeof(region_geom));
    for (int r = 0; r < nRegions; ++r) region_geom_init(&G[r], &regions[r]);

    out->nRegions = nRegions;
    out->nInterfaces = nInterfaces;
    out->regions = (orc_ldu *)calloc((size_t)nRegions, sizeof(orc_ldu));
    out->interfaces = (orc_interface *)calloc((size_t)(nInterfaces ? nInterfaces : 1), sizeof(orc_interface));

    const double **coeffSide = (const double **)calloc((size_t)nRegions * 6, sizeof(double *));
    const int32_t **fcSide = (const int32_t **)calloc((size_t)nRegions * 6, sizeof(int32_t *));

    for (int i = 0; i < nInterfaces; ++i) {
        const orc_interface_spec *s = &ifs[i];
        const region_geom *gA = &G[s->regionA], *gB = &G[s->regionB];
        if (s->sideA / 2 != s->sideB / 2) { free(G); return -1; }
        if (regions[s->regionA].patch_type[s->sideA] != ORC_INTERFACE ||
            regions[s->regionB].patch_type[s->sideB] != ORC_INTERFACE) { free(G); return -2; }
        orc_interface *I = &out->interfaces[i];
        I->regionA = s->regionA;
        I->regionB = s->regionB;
        I->nA = side_nfaces(gA, s->sideA);
        I->nB = side_nfaces(gB, s->sideB);
        ggi_weights(gA, s->sideA, gB, s->sideB, &I->offA, &I->idxA, &I->wA);
        ggi_weights(gB, s->sideB, gA, s->sideA, &I->offB, &I->idxB, &I->wB);
        I->fcA = (int32_t *)malloc(sizeof(int32_t) * (size_t)I->nA);
        I->fcB = (int32_t *)malloc(sizeof(int32_t) * (size_t)I->nB);
        I->coeffA = (double *)malloc(sizeof(double) * (size_t)I->nA);
        I->coeffB = (double *)malloc(sizeof(double) * (size_t)I->nB);
        I->magSfA = (double *)malloc(sizeof(double) * (size_t)I->nA);
        I->magSfB = (double *)malloc(sizeof(double) * (size_t)I->nB);
        iface_coeffs(gA, s->sideA, gB, s->sideB, I->offA, I->idxA, I->wA, I->fcA, I->coeffA, I->magSfA);
        iface_coeffs(gB, s->sideB, gA, s->sideA, I->offB, I->idxB, I->wB, I->fcB, I->coeffB, I->magSfB);
        coeffSide[6 * s->regionA + s->sideA] = I->coeffA;
        fcSide[6 * s->regionA + s->sideA] = I->fcA;
        coeffSide[6 * s->regionB + s->sideB] = I->coeffB;
        fcSide[6 * s->regionB + s->sideB] = I->fcB;
    }
    for (int r = 0; r < nRegions; ++r) {
        for (int s = 0; s < 6; ++s)
            if (regions[r].patch_type[s] == ORC_INTERFACE && !coeffSide[6 * r + s]) {
                free(G); return -3;   /* interface patch without a partner */
            }
        assemble_region(&G[r], &out->regions[r], coeffSide + 6 * r, fcSide + 6 * r);
        region_geom_free(&G[r]);
    }
    free(coeffSide);
    free(fcSide);
    free(G);
    return 0;
}

void orc_free(orc_system *sys)
{
    for (int r = 0; r < sys->nRegions; ++r) {
        orc_ldu *L = &sys->regions[r];
        free(L->lAddr); free(L->uAddr); free(L->diag); free(L->lower);
        free(L->upper); free(L->source); free(L->psi); free(L->k); free(L->cellSize);
    }
    for (int i = 0; i < sys->nInterfaces; ++i) {
        orc_interface *I = &sys->interfaces[i];
        free(I->fcA); free(I->fcB); free(I->coeffA); free(I->coeffB);
        free(I->offA); free(I->idxA); free(I->wA);
        free(I->offB); free(I->idxB); free(I->wB);
        free(I->magSfA); free(I->magSfB);
    }
    free(sys->regions);
    free(sys->interfaces);
    memset(sys, 0, sizeof(*sys));
}

/* ------------------------------------------------------------------------- */
/* Coupled LDU operations                                                     */
/* ------------------------------------------------------------------------- */

/* GGI interpolation onto one side's faces: buf_i = sum_p w_p * x[fc[idx_p]]
   (regionCouplePatch().interpolate, monolithicCouplingFvPatchField.C:
   401-404). */
static void ggi_gather(int n, const int32_t *off, const int32_t *idx, const double *w,
                       const int32_t *fcOther, const double *xOther, double *buf)
{
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int p = off[i]; p < off[i + 1]; ++p) s += w[p] * xOther[fcOther[idx[p]]];
        buf[i] = s;
    }
}

/* coupledLduMatrix::Amul: result = 0, init interfaces, per-region
   lduMatrix::AmulCore face loop, then update interfaces
   (monolithicCouplingFvPatchField.C:381-414 init: GGI gather of the other
   side; :419-464 update: result[fc] -= coeffs*pnf). */
void orc_amul(const orc_system *sys, double *const *x, double **result)
{
    int nI = sys->nInterfaces;
    double **bufA = (double **)malloc(sizeof(double *) * (size_t)(nI ? nI : 1));
    double **bufB = (double **)malloc(sizeof(double *) * (size_t)(nI ? nI : 1));
    for (int i = 0; i < nI; ++i) {
        const orc_interface *I = &sys->interfaces[i];
        bufA[i] = (double *)malloc(sizeof(double) * (size_t)(I->nA ? I->nA : 1));
        bufB[i] = (double *)malloc(sizeof(double) * (size_t)(I->nB ? I->nB : 1));
        ggi_gather(I->nA, I->offA, I->idxA, I->wA, I->fcB, x[I->regionB], bufA[i]);
        ggi_gather(I->nB, I->offB, I->idxB, I->wB, I->fcA, x[I->regionA], bufB[i]);
    }
    for (int r = 0; r < sys->nRegions; ++r) {
        const orc_ldu *L = &sys->regions[r];
        double *y = result[r];
        const double *xr = x[r];
        for (int c = 0; c < L->nCells; ++c) y[c] = L->diag[c] * xr[c];
        for (int f = 0; f < L->nFaces; ++f) {
            y[L->uAddr[f]] += L->lower[f] * xr[L->lAddr[f]];
            y[L->lAddr[f]] += L->upper[f] * xr[L->uAddr[f]];
        }
    }
    for (int i = 0; i < nI; ++i) {
        const orc_interface *I = &sys->interfaces[i];
        double *yA = result[I->regionA], *yB = result[I->regionB];
        for (int f = 0; f < I->nA; ++f) yA[I->fcA[f]] -= I->coeffA[f] * bufA[i][f];
        for (int f = 0; f < I->nB; ++f) yB[I->fcB[f]] -= I->coeffB[f] * bufB[i][f];
        free(bufA[i]);
        free(bufB[i]);
    }
    free(bufA);
    free(bufB);
}

/* lduMatrix::sumA over the coupled system: diag + off-diagonal row sums,
   minus each interface face's boundary coefficient. */
void orc_sumA(const orc_system *sys, double **sumA)
{
    for (int r = 0; r < sys->nRegions; ++r) {
        const orc_ldu *L = &sys->regions[r];
        double *s = sumA[r];
        for (int c = 0; c < L->nCells; ++c) s[c] = L->diag[c];
        for (int f = 0; f < L->nFaces; ++f) {
            s[L->uAddr[f]] += L->lower[f];
            s[L->lAddr[f]] += L->upper[f];
        }
    }
    for (int i = 0; i < sys->nInterfaces; ++i) {
        const orc_interface *I = &sys->interfaces[i];
        for (int f = 0; f < I->nA; ++f) sumA[I->regionA][I->fcA[f]] -= I->coeffA[f];
        for (int f = 0; f < I->nB; ++f) sumA[I->regionB][I->fcB[f]] -= I->coeffB[f];
    }
}

/* Compensated summation (TwoSum error terms accumulated beside the sum).
   A global sum (gSum / gSumProd / gSumMag) is then the sum of its terms to
   about 2^-100 relative, so its rounding to double does not depend on the
   order in which the terms are added: a sequential restatement and a
   parallel reduction of the same terms agree.  Plain sequential summation
   differs from either by rounding only. */
typedef struct {
    double hi, lo;
} csum;

static inline void csum_add(csum *a, double x)
{
    double s = a->hi + x;
    double bb = s - a->hi;
    a->lo += (a->hi - (s - bb)) + (x - bb);
    a->hi = s;
}

static inline double csum_val(csum a) { return a.hi + a.lo; }

/* lduMatrix::solver::normFactor: xRef = gAverage(x); tmp = sumA*xRef;
   sum(|Ax - tmp| + |b - tmp|) + small (1e-20). */
double orc_norm_factor(const orc_system *sys, double *const *x, double *const *b,
                       double *const *Ax, double **tmp)
{
    csum sx = {0.0, 0.0};
    long n = 0;
    for (int r = 0; r < sys->nRegions; ++r) {
        for (int c = 0; c < sys->regions[r].nCells; ++c) csum_add(&sx, x[r][c]);
        n += sys->regions[r].nCells;
    }
    double xRef = csum_val(sx) / (double)n;
    orc_sumA(sys, tmp);
    csum nf = {0.0, 0.0};
    for (int r = 0; r < sys->nRegions; ++r)
        for (int c = 0; c < sys->regions[r].nCells; ++c) {
            double t = tmp[r][c] * xRef;
            csum_add(&nf, fabs(Ax[r][c] - t) + fabs(b[r][c] - t));
        }
    return csum_val(nf) + 1e-20;
}

/* coupled Cholesky preconditioner, calcPreconDiag: per region, in face
   order, d[u] -= upper*lower/d[l] (upper*upper when symmetric), then
   invert.  Interface coefficients are not part of the factor (block-Jacobi
   over regions), as the coupled preconditioner works on each region's
   lduMatrix. */
void orc_dilu_factor(const orc_system *sys, double **rD)
{
    for (int r = 0; r < sys->nRegions; ++r) {
        const orc_ldu *L = &sys->regions[r];
        double *d = rD[r];
        for (int c = 0; c < L->nCells; ++c) d[c] = L->diag[c];
        for (int f = 0; f < L->nFaces; ++f) {
            double lo = L->symmetric ? L->upper[f] : L->lower[f];
            d[L->uAddr[f]] -= L->upper[f] * lo / d[L->lAddr[f]];
        }
        for (int c = 0; c < L->nCells; ++c) d[c] = 1.0 / d[c];
    }
}

/* coupled Cholesky preconditioner, precondition: w = rD*r; forward face
   loop w[u] -= rD[u]*lower*w[l]; reverse face loop w[l] -= rD[l]*upper*w[u]. */
void orc_dilu_precondition(const orc_system *sys, double *const *rD,
                           double **w, double *const *r)
{
    for (int g = 0; g < sys->nRegions; ++g) {
        const orc_ldu *L = &sys->regions[g];
        const double *d = rD[g];
        double *x = w[g];
        const double *b = r[g];
        const double *lo = L->symmetric ? L->upper : L->lower;
        for (int c = 0; c < L->nCells; ++c) x[c] = b[c] * d[c];
        for (int f = 0; f < L->nFaces; ++f)
            x[L->uAddr[f]] -= d[L->uAddr[f]] * lo[f] * x[L->lAddr[f]];
        for (int f = L->nFaces - 1; f >= 0; --f)
            x[L->lAddr[f]] -= d[L->lAddr[f]] * L->upper[f] * x[L->uAddr[f]];
    }
}

void orc_interface_flux(const orc_system *sys, int i, double *const *psi,
                        double *QA, double *QB, double *srcA, double *srcB)
{
    const orc_interface *I = &sys->interfaces[i];
    const orc_ldu *LA = &sys->regions[I->regionA], *LB = &sys->regions[I->regionB];
    for (int side = 0; side < 2; ++side) {
        int n = side ? I->nB : I->nA;
        const int32_t *off = side ? I->offB : I->offA, *idx = side ? I->idxB : I->idxA;
        const double *w = side ? I->wB : I->wA;
        const int32_t *fc = side ? I->fcB : I->fcA, *fcN = side ? I->fcA : I->fcB;
        const double *coeff = side ? I->coeffB : I->coeffA;
        const double *magSf = side ? I->magSfB : I->magSfA;
        const orc_ldu *LO = side ? LB : LA, *LN = side ? LA : LB;
        const double *xO = psi[side ? I->regionB : I->regionA];
        const double *xN = psi[side ? I->regionA : I->regionB];
        /* normal axis: the one whose cell size enters the coefficient; the
           interface is planar, so take the axis with faces of area magSf */
        int a = 0;
        for (int t = 0; t < 3; ++t) {
            double ar = 1.0;
            for (int u = 0; u < 3; ++u) if (u != t) ar *= LO->cellSize[u];
            if (n > 0 && fabs(ar - magSf[0]) <= 1e-12 * magSf[0]) a = t;
        }
        double dOwn = 0.5 * LO->cellSize[a], dNei = 0.5 * LN->cellSize[a];
        double magDelta = dOwn + dNei, wP = dNei / magDelta;
        csum Q = {0.0, 0.0};
        double S = 0.0;
        for (int f = 0; f < n; ++f) {
            double Tn = 0.0, kn = 0.0;
            for (int p = off[f]; p < off[f + 1]; ++p) {
                Tn += w[p] * xN[fcN[idx[p]]];
                kn += w[p] * LN->k[fcN[idx[p]]];
            }
            double To = xO[fc[f]];
            csum_add(&Q, coeff[f] * (Tn - To));
            /* monolithicTemperatureFvPatchScalarField.C:204-224 source with
               Tw from monolithicThermalDiffusivity...C:381 */
            double kOwn = LO->k[fc[f]] / (1.0 - wP) / magDelta;
            double kNei = kn / wP / magDelta;
            if (off[f + 1] > off[f]) {
                double Tw = (kOwn * To + kNei * Tn) / (kOwn + kNei);
                double kd = coeff[f] / magSf[f];
                S += (kOwn * (Tw - To) - kd * (Tn - To)) * magSf[f];
            }
        }
        if (side) { *QB = csum_val(Q); *srcB = S; } else { *QA = csum_val(Q); *srcA = S; }
    }
}

/* Coupled patch values of one side of interface i from the field psi, as
   correctBoundaryConditions leaves them on the monolithicTemperature
   patches (box-built systems: k and cell sizes known).  tn: the patch
   neighbour field, the GGI interpolate of the other side's face-cell values
   (monolithicCouplingFvPatchField.C:208-236).  tw: the patch value
   calcTemperature gives with radiation off
   (monolithicThermalDiffusivityFvPatchScalarField.C:300-380, called from
   monolithicTemperatureFvPatchScalarField.C:147-202):
   Tw = (kOwn*TcOwn + kNei*TcNei)/(kOwn + kNei), kOwn = forig/(1-w)/|d|,
   kNei = interpolate(forig_nei)/w/|d|. */
void orc_interface_patch_values(const orc_system *sys, int i, int side, double *const *psi,
                                double *tn, double *tw)
{
    const orc_interface *I = &sys->interfaces[i];
    int n = side ? I->nB : I->nA;
    const int32_t *off = side ? I->offB : I->offA, *idx = side ? I->idxB : I->idxA;
    const double *w = side ? I->wB : I->wA;
    const int32_t *fc = side ? I->fcB : I->fcA, *fcN = side ? I->fcA : I->fcB;
    const double *magSf = side ? I->magSfB : I->magSfA;
    const orc_ldu *LO = &sys->regions[side ? I->regionB : I->regionA];
    const orc_ldu *LN = &sys->regions[side ? I->regionA : I->regionB];
    const double *xO = psi[side ? I->regionB : I->regionA];
    const double *xN = psi[side ? I->regionA : I->regionB];
    /* the normal axis, as in orc_interface_flux */
    int a = 0;
    for (int t = 0; t < 3; ++t) {
        double ar = 1.0;
        for (int u = 0; u < 3; ++u) if (u != t) ar *= LO->cellSize[u];
        if (n > 0 && fabs(ar - magSf[0]) <= 1e-12 * magSf[0]) a = t;
    }
    double dOwn = 0.5 * LO->cellSize[a], dNei = 0.5 * LN->cellSize[a];
    double magDelta = dOwn + dNei, wP = dNei / magDelta;
    for (int f = 0; f < n; ++f) {
        double Tn = 0.0, kn = 0.0;
        for (int p = off[f]; p < off[f + 1]; ++p) {
            Tn += w[p] * xN[fcN[idx[p]]];
            kn += w[p] * LN->k[fcN[idx[p]]];
        }
        double kOwn = LO->k[fc[f]] / (1.0 - wP) / magDelta;
        double kNei = kn / wP / magDelta;
        tn[f] = Tn;
        tw[f] = (kOwn * xO[fc[f]] + kNei * Tn) / (kOwn + kNei);
    }
}

/* ------------------------------------------------------------------------- */
/* Krylov solvers                                                             */
/* ------------------------------------------------------------------------- */

typedef struct {
    int nR;
    int *n;
    double **v;
} ffield;

static ffield ff_new(const orc_system *sys)
{
    ffield f;
    f.nR = sys->nRegions;
    f.n = (int *)malloc(sizeof(int) * (size_t)f.nR);
    f.v = (double **)malloc(sizeof(double *) * (size_t)f.nR);
    for (int r = 0; r < f.nR; ++r) {
        f.n[r] = sys->regions[r].nCells;
        f.v[r] = (double *)calloc((size_t)f.n[r], sizeof(double));
    }
    return f;
}

static void ff_free(ffield *f)
{
    for (int r = 0; r < f->nR; ++r) free(f->v[r]);
    free(f->v);
    free(f->n);
}

static double ff_sumprod(const ffield *a, const ffield *b)
{
    csum s = {0.0, 0.0};
    for (int r = 0; r < a->nR; ++r)
        for (int c = 0; c < a->n[r]; ++c) csum_add(&s, a->v[r][c] * b->v[r][c]);
    return csum_val(s);
}

static double ff_summag(const ffield *a)
{
    csum s = {0.0, 0.0};
    for (int r = 0; r < a->nR; ++r)
        for (int c = 0; c < a->n[r]; ++c) csum_add(&s, fabs(a->v[r][c]));
    return csum_val(s);
}

/* lduSolver::stop + SolverPerformance::checkConvergence.  A residual that
   is not finite (NaN / Inf in the system, the field or the Krylov scalars)
   stops the solve at once, flagged diverged (singular = 2); the reference
   has no such test and would iterate to maxIter. */
static int solver_stop(orc_perf *p, double tol, double relTol, int minIter, int maxIter)
{
    if (!isfinite(p->finalResidual)) {
        p->converged = 0;
        p->singular = 2;
        return 1;
    }
    if (p->nIterations < minIter) return 0;
    int conv = p->finalResidual < tol
            || (relTol > 1e-20 && p->finalResidual < relTol * p->initialResidual);
    p->converged = conv;
    return p->nIterations >= maxIter || conv;
}

/* Right-preconditioned van der Vorst BiCGStab, selected by `solver
   BiCGStab;` with `preconditioner { preconditioner Cholesky; }`
   (flowOverHeatedPlate/system/fluid/fvSolution:43-57), called from
   coupledFvMatrix::solve at multiRegionSystem.C:150-153. */
static void bicgstab(orc_system *sys, const ffield *x, const ffield *b, double **rD,
                     double tol, double relTol, int minIter, int maxIter, orc_perf *perf)
{
    ffield p = ff_new(sys), r = ff_new(sys), tmp = ff_new(sys);
    orc_amul(sys, x->v, p.v);
    for (int g = 0; g < r.nR; ++g)
        for (int c = 0; c < r.n[g]; ++c) r.v[g][c] = b->v[g][c] - p.v[g][c];
    double nf = orc_norm_factor(sys, x->v, b->v, p.v, tmp.v);
    perf->initialResidual = ff_summag(&r) / nf;
    perf->finalResidual = perf->initialResidual;
    perf->nIterations = 0;
    perf->converged = 0;
    perf->singular = 0;
    if (!solver_stop(perf, tol, relTol, minIter, maxIter)) {
        const double great = 1e20;
        double rho = great, rhoOld, alpha = 0.0, omega = great, beta;
        ffield ph = ff_new(sys), v = ff_new(sys), s = ff_new(sys), sh = ff_new(sys),
               t = ff_new(sys), rw = ff_new(sys);
        for (int g = 0; g < r.nR; ++g) {
            memset(p.v[g], 0, sizeof(double) * (size_t)p.n[g]);
            memcpy(rw.v[g], r.v[g], sizeof(double) * (size_t)r.n[g]);
        }
        do {
            rhoOld = rho;
            rho = ff_sumprod(&rw, &r);
            beta = rho / rhoOld * (alpha / omega);
            for (int g = 0; g < r.nR; ++g)
                for (int c = 0; c < r.n[g]; ++c)
                    p.v[g][c] = r.v[g][c] + beta * p.v[g][c] - beta * omega * v.v[g][c];
            orc_dilu_precondition(sys, rD, ph.v, p.v);
            orc_amul(sys, ph.v, v.v);
            double rwv = ff_sumprod(&rw, &v);
            if (rho == 0.0 || rwv == 0.0) { perf->singular = 1; break; }
            alpha = rho / rwv;
            for (int g = 0; g < r.nR; ++g)
                for (int c = 0; c < r.n[g]; ++c) s.v[g][c] = r.v[g][c] - alpha * v.v[g][c];
            orc_dilu_precondition(sys, rD, sh.v, s.v);
            orc_amul(sys, sh.v, t.v);
            double ts = ff_sumprod(&t, &s), tt = ff_sumprod(&t, &t);
            /* t.t == 0: t = A M^-1 s vanished, so s did and x + alpha*ph
               solves the system.  The step is taken with omega = 0
               (x += alpha*ph, r = s) and the stop test decides; if it does
               not stop, the solve cannot continue (singular). */
            int sZero = tt == 0.0;
            omega = sZero ? 0.0 : ts / tt;
            for (int g = 0; g < r.nR; ++g)
                for (int c = 0; c < r.n[g]; ++c) {
                    x->v[g][c] = x->v[g][c] + alpha * ph.v[g][c] + omega * sh.v[g][c];
                    r.v[g][c] = s.v[g][c] - omega * t.v[g][c];
                }
            perf->finalResidual = ff_summag(&r) / nf;
            perf->nIterations++;
            if (solver_stop(perf, tol, relTol, minIter, maxIter)) break;
            if (sZero) { perf->singular = 1; break; }
        } while (1);
        ff_free(&ph); ff_free(&v); ff_free(&s); ff_free(&sh); ff_free(&t); ff_free(&rw);
    }
    ff_free(&p); ff_free(&r); ff_free(&tmp);
}

/* Preconditioned conjugate gradients (lduMatrix PCG restated), for
   symmetric block systems. */
static void pcg(orc_system *sys, const ffield *x, const ffield *b, double **rD,
                double tol, double relTol, int minIter, int maxIter, orc_perf *perf)
{
    ffield wA = ff_new(sys), rA = ff_new(sys), pA = ff_new(sys), tmp = ff_new(sys);
    orc_amul(sys, x->v, wA.v);
    for (int g = 0; g < rA.nR; ++g)
        for (int c = 0; c < rA.n[g]; ++c) rA.v[g][c] = b->v[g][c] - wA.v[g][c];
    double nf = orc_norm_factor(sys, x->v, b->v, wA.v, tmp.v);
    perf->initialResidual = ff_summag(&rA) / nf;
    perf->finalResidual = perf->initialResidual;
    perf->nIterations = 0;
    perf->converged = 0;
    perf->singular = 0;
    if (!solver_stop(perf, tol, relTol, minIter, maxIter)) {
        double rho = 1e20, rhoOld;
        do {
            rhoOld = rho;
            orc_dilu_precondition(sys, rD, wA.v, rA.v);
            rho = ff_sumprod(&wA, &rA);
            if (perf->nIterations == 0) {
                for (int g = 0; g < pA.nR; ++g)
                    memcpy(pA.v[g], wA.v[g], sizeof(double) * (size_t)pA.n[g]);
            } else {
                double beta = rho / rhoOld;
                for (int g = 0; g < pA.nR; ++g)
                    for (int c = 0; c < pA.n[g]; ++c) pA.v[g][c] = wA.v[g][c] + beta * pA.v[g][c];
            }
            orc_amul(sys, pA.v, wA.v);
            double wApA = ff_sumprod(&wA, &pA);
            if (fabs(wApA) / nf < 1e-300 || wApA == 0.0) { perf->singular = 1; break; }
            double alpha = rho / wApA;
            for (int g = 0; g < pA.nR; ++g)
                for (int c = 0; c < pA.n[g]; ++c) {
                    x->v[g][c] += alpha * pA.v[g][c];
                    rA.v[g][c] -= alpha * wA.v[g][c];
                }
            perf->finalResidual = ff_summag(&rA) / nf;
            perf->nIterations++;
        } while (!solver_stop(perf, tol, relTol, minIter, maxIter));
    }
    ff_free(&wA); ff_free(&rA); ff_free(&pA); ff_free(&tmp);
}

int orc_solve(orc_system *sys, int solver, double tol, double relTol,
              int minIter, int maxIter, orc_perf *perf)
{
    ffield x, b;
    x.nR = b.nR = sys->nRegions;
    x.n = (int *)malloc(sizeof(int) * (size_t)x.nR);
    b.n = (int *)malloc(sizeof(int) * (size_t)b.nR);
    x.v = (double **)malloc(sizeof(double *) * (size_t)x.nR);
    b.v = (double **)malloc(sizeof(double *) * (size_t)b.nR);
    double **rD = (double **)malloc(sizeof(double *) * (size_t)x.nR);
    for (int r = 0; r < x.nR; ++r) {
        x.n[r] = b.n[r] = sys->regions[r].nCells;
        x.v[r] = sys->regions[r].psi;
        b.v[r] = sys->regions[r].source;
        rD[r] = (double *)malloc(sizeof(double) * (size_t)x.n[r]);
    }
    orc_dilu_factor(sys, rD);
    if (solver == ORC_PCG) pcg(sys, &x, &b, rD, tol, relTol, minIter, maxIter, perf);
    else bicgstab(sys, &x, &b, rD, tol, relTol, minIter, maxIter, perf);
    for (int r = 0; r < x.nR; ++r) free(rD[r]);
    free(rD); free(x.n); free(b.n); free(x.v); free(b.v);
    return 0;
}
