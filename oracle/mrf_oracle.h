/*
 * mrf_oracle.h -- CPU oracle for the monolithic interface-coupled solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference algorithm used by multiRegionFoam's
 * monolithic coupled temperature solve.  It exists to check the HIP
 * product path; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and never as the thing measured or shipped.
 *
 * Coverage (each function cites the reference file:line it follows; the
 * foam-extend 4.1 arithmetic behind those call sites is not vendored in the
 * reference, so it is restated from the published algorithm):
 *   - structured hex box regions, OpenFOAM upper-triangular face order
 *   - fvm::laplacian (Gauss linear), fvm::div (Gauss upwind), steady ddt,
 *     fixedValue / zeroGradient / empty boundaries
 *   - monolithic interface conductance (series resistance, harmonic clamp)
 *   - GGI weights for axis-aligned rectangular faces (exact overlap areas,
 *     rescaled to sum 1), directMap as the conformal special case
 *   - coupled Amul (LDU face loop + interface update), sumA, normFactor
 *   - Cholesky/DILU factorisation and sweeps (block-Jacobi over regions)
 *   - van der Vorst BiCGStab and PCG with lduSolver stop/convergence rules
 *
 * Parity pinning (tests/test_oracle_*.py): closed-form conductances,
 * two-slab 1-D exact solutions, dense brute-force A*x and DILU factor
 * M=(D+L)D^-1(D+U), direct sparse solves, GGI weight sums and overlap
 * conservation.  Functions that only upstream foam-extend could pin exactly
 * (iteration counts of its BiCGStab) are "parity unpinned" w.r.t. the
 * upstream binary and pinned only against the textbook algorithm.
 */
#ifndef MRF_ORACLE_H
#define MRF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Patch sides of a box region: xmin xmax ymin ymax zmin zmax */
enum { ORC_XMIN = 0, ORC_XMAX, ORC_YMIN, ORC_YMAX, ORC_ZMIN, ORC_ZMAX };

/* Temperature boundary types */
enum { ORC_ZERO_GRADIENT = 0, ORC_FIXED_VALUE = 1, ORC_EMPTY = 2, ORC_INTERFACE = 3 };

/* Flux (velocity) boundary kinds */
enum { ORC_WALL = 0, ORC_INLET = 1, ORC_OUTLET = 2 };

typedef struct {
    int n[3];             /* cells per axis */
    double origin[3];     /* box origin */
    double length[3];     /* box size */
    int transport;        /* 1: transportTemperature (fluid), 0: conductTemperature (solid) */
    double rhoCp;         /* rho*cp for the convective term (fluid) */
    /* conductivity: k_block > 0 selects block-random log-uniform k */
    double k_uniform;
    int k_block;
    double k_min, k_max;
    uint64_t k_seed;
    /* velocity: Umax * prod_{axes in profile_mask} 4 eta (1-eta) along flow_axis;
       internal face fluxes add a divergence-free noise flux, the discrete curl
       of an edge potential uniform in +-noise*Umax*h/4 (zero on boundary edges) */
    int flow_axis;
    int profile_mask;     /* bit a set -> parabolic across axis a */
    double Umax;
    double noise;
    uint64_t vel_seed;
    int patch_type[6];    /* ORC_ZERO_GRADIENT ... */
    double patch_value[6];/* fixedValue temperature */
    int patch_flux[6];    /* ORC_WALL (phi=0) / ORC_INLET (profile U) /
                             ORC_OUTLET (face-cell U); phi = +-U_n |Sf|,
                             + on max sides, - on min sides */
    double T0;            /* uniform initial temperature */
} orc_region_spec;

typedef struct {
    int regionA, sideA, regionB, sideB;
} orc_interface_spec;

/* LDU matrix of one region, in the form the coupled solver receives it:
   diag includes every patch's internalCoeffs (addBoundaryDiag), source
   includes non-coupled patches' boundaryCoeffs (addBoundarySource with
   couples=false). */
typedef struct {
    int nCells, nFaces;
    int32_t *lAddr, *uAddr;
    double *diag, *lower, *upper, *source, *psi;
    int symmetric;        /* 1 if lower == upper (pure diffusion) */
    /* per-cell properties kept for tests */
    double *k;
    double *cellSize;     /* 3 entries: dx dy dz */
} orc_ldu;

/* One monolithic interface between region A and region B.
   Side A face i gathers from B faces: pnfA_i = sum_p wA[p] * psiB[fcB[idxA[p]]],
   p in [offA[i], offA[i+1]).  Same for B with the roles swapped.
   Amul adds result_A[fcA[i]] -= coeffA[i] * pnfA_i. */
typedef struct {
    int regionA, regionB;
    int nA, nB;
    int32_t *fcA, *fcB;
    double *coeffA, *coeffB;
    int32_t *offA, *idxA;
    double *wA;
    int32_t *offB, *idxB;
    double *wB;
    double *magSfA, *magSfB;   /* face areas (for flux checks) */
} orc_interface;

typedef struct {
    int nRegions, nInterfaces;
    orc_ldu *regions;
    orc_interface *interfaces;
} orc_system;

typedef struct {
    double initialResidual, finalResidual;
    int nIterations, converged, singular;
} orc_perf;

enum { ORC_BICGSTAB = 0, ORC_PCG = 1 };

/* Build and assemble a case.  Returns 0 on success. */
int orc_build(int nRegions, const orc_region_spec *regions,
              int nInterfaces, const orc_interface_spec *ifs,
              orc_system *out);
void orc_free(orc_system *sys);

/* Building blocks (exposed for tests) */
void orc_amul(const orc_system *sys, double *const *x, double **result);
void orc_sumA(const orc_system *sys, double **sumA);
double orc_norm_factor(const orc_system *sys, double *const *x, double *const *b,
                       double *const *Ax, double **tmp);
void orc_dilu_factor(const orc_system *sys, double **rD);
void orc_dilu_precondition(const orc_system *sys, double *const *rD,
                           double **w, double *const *r);
/* Interface checks: heat flow through side A and side B (W), evaluated with
   the face conductances and GGI-interpolated neighbour temperatures.  Also
   the monolithicTemperature explicit source sum (zero with radiation off). */
void orc_interface_flux(const orc_system *sys, int i, double *const *psi,
                        double *QA, double *QB, double *srcA, double *srcB);

/* Patch values of one side (0 = A, 1 = B) of interface i: the neighbour
   field tn and the monolithicTemperature wall value tw, per face. */
void orc_interface_patch_values(const orc_system *sys, int i, int side, double *const *psi,
                                double *tn, double *tw);

/* Solve A psi = source for all regions; psi is in/out (sys->regions[r].psi). */
int orc_solve(orc_system *sys, int solver, double tol, double relTol,
              int minIter, int maxIter, orc_perf *perf);

/* The same BiCGStab solve on all host cores (mrf_oracle_mt.c): row-parallel
   A*x, level-scheduled DILU sweeps (bit-identical to the sequential ones),
   compensated sums merged across threads.  For the all-cores CPU baseline. */
int orc_solve_mt(orc_system *sys, double tol, double relTol, int minIter, int maxIter, orc_perf *perf);
/* threads orc_solve_mt uses (OMP_NUM_THREADS or the host's cores) */
int orc_threads(void);

/* Deterministic hash RNG used by the case generator (splitmix64). */
double orc_u01(uint64_t seed, uint64_t index);

#ifdef __cplusplus
}
#endif
#endif
