/*
 * mrfg.h -- C ABI of the MI355X-native monolithic interface-coupled solver
 * (libmrfg.so, built from multiregionfoam_amd/csrc for gfx950).
 *
 * The boundary is the one multiRegionFoam crosses when it solves a
 * monolithically coupled field:
 *
 *   multiRegionSystem::assembleAndSolveCoupledMatrix  (src/multiRegionSystem/
 *   multiRegionSystem.C:61-191) collects each region's assembled fvMatrix
 *   into a coupledFvMatrix and calls coupledEqns.solve(dict) with region 0's
 *   "<field>coupled" solver dictionary (multiRegionSystem.C:150-153).
 *
 * foam-extend's coupledFvMatrix<scalar>::solve then hands the coupled
 * solver, per region, the LDU matrix with addBoundaryDiag applied, the
 * source with addBoundarySource(couples=false) applied, and the interface
 * boundary coefficients; the monolithic interfaces contribute to A*x through
 * monolithicCouplingFvPatchField::initInterfaceMatrixUpdate /
 * updateInterfaceMatrix (monolithicCouplingFvPatchField.C:381-464):
 *     result[fc[i]] -= coeffs[i] * sum_p w[p] * psi_other[fc_other[idx[p]]].
 * The entry points below take exactly that data, as plain arrays.
 *
 * Conventions: FP64 values, int32 addressing (the reference's -DWM_DP,
 * -DWM_LABEL_SIZE=32, compile_commands.json:9-10); faces in OpenFOAM
 * upper-triangular order (l_addr[f] < u_addr[f], sorted by owner then
 * neighbour).  Every function returns MRFG_OK (0) or a negative code and
 * never throws; non-convergence is not an error (mrfg_perf.converged = 0).
 * A context is not thread-safe; several contexts may coexist.
 */
#ifndef MRFG_H
#define MRFG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRFG_OK 0
#define MRFG_ERR_ARG (-1)      /* bad argument or inconsistent addressing */
#define MRFG_ERR_DEVICE (-2)   /* HIP runtime error (no device, out of memory, ...) */
#define MRFG_ERR_DICT (-3)     /* dictionary syntax or unknown keyword/solver */
#define MRFG_ERR_STATE (-4)    /* call out of sequence (e.g. solve before commit) */

typedef struct mrfg_ctx mrfg_ctx;
typedef struct mrfg_case mrfg_case;

/* lduSolverPerformance of one coupled solve */
typedef struct {
    double initial_residual;   /* sum|b-Ax0| / normFactor */
    double final_residual;     /* sum|b-Ax| / normFactor (FP64 in every mode) */
    int n_iterations;          /* Krylov iterations (mixed: FP32 inner iterations plus the FP64
                                  continuation of a system whose refinement stalled) */
    int converged;
    int singular;              /* 0; 1 = breakdown (rho, (rw,v) or (w,p) zero, or s = 0 without
                                  meeting the tolerance); 2 = diverged: a residual became NaN or Inf
                                  (the system is stopped at once and releases its batch) */
    int n_outer;               /* mixed: FP64 refinement steps; fp64: 0 */
    double seconds;            /* wall time of the solve on the device */
    double bytes_moved;        /* algorithmic bytes of the solve */
} mrfg_perf;

enum { MRFG_BICGSTAB = 0, MRFG_PCG = 1 };
enum { MRFG_FP64 = 0, MRFG_MIXED = 1 };

/* Solver controls: the keys of an fvSolution solver entry
   (flowOverHeatedPlate/system/fluid/fvSolution:43-57) plus the GPU keys
   precision / refinement / checkInterval. */
typedef struct {
    int solver;            /* MRFG_BICGSTAB (BiCGStab) or MRFG_PCG (PCG/CG) */
    int precision;         /* MRFG_FP64 or MRFG_MIXED */
    double tolerance;
    double rel_tol;
    int min_iter;
    int max_iter;
    double inner_rel_tol;  /* mixed: FP32 inner solve relative tolerance */
    int inner_max_iter;
    int max_outer;
    int check_interval;    /* iterations between host convergence checks */
    int launch_mode;       /* MRFG_LAUNCH_GRAPH (0): each check interval's iterations replayed as one
                              captured HIP graph; MRFG_LAUNCH_KERNELS (1): kernels launched one by
                              one (fvSolution key graphs on|off) */
    int tuning;            /* MRFG_TUNE_AUTO (0): the first solve after a commit times the SpMV in
                              every launch configuration (blocks of 4, 8 or 16 slices of 64 rows;
                              ELL chunk 2, 4, 8 or 16) on the batch and keeps the fastest;
                              MRFG_TUNE_FIXED (1): 4 slices, chunk 8 (fvSolution key
                              autotune on|off).  Results do not depend on the choice. */
    int inner_preconditioner; /* mixed: the FP32 inner solve's preconditioner (fvSolution key
                              refinement { innerPreconditioner tileDILU|DILU; }).
                              MRFG_INNER_TILE_DILU (0, default): each structured-box tile's DILU
                              alone (couplings across tile edges left out: the sweeps wait for no
                              other tile); MRFG_INNER_DILU (1): the full DILU in FP32.  The FP64
                              refinement sets the result's accuracy either way. */
} mrfg_controls;

enum { MRFG_LAUNCH_GRAPH = 0, MRFG_LAUNCH_KERNELS = 1 };
enum { MRFG_TUNE_AUTO = 0, MRFG_TUNE_FIXED = 1 };
enum { MRFG_INNER_TILE_DILU = 0, MRFG_INNER_DILU = 1 };

/* ---- context --------------------------------------------------------- */

/* Replaces the run-time selected coupledLduSolver object built per call by
   coupledFvMatrix::solve (multiRegionSystem.C:150-153); one context holds
   one or more coupled systems resident on `device`. */
int mrfg_create(mrfg_ctx **ctx, int device);
void mrfg_destroy(mrfg_ctx *ctx);
const char *mrfg_last_error(const mrfg_ctx *ctx);
const char *mrfg_version(void);

/* Start a coupled system (one coupledFvMatrix: multiRegionSystem.C:85). */
int mrfg_begin_system(mrfg_ctx *ctx, int n_regions, int n_interfaces, int *sys_id);

/* One region's fvMatrix as the coupled solver receives it
   (coupledEqns.set(nReg, ...) at multiRegionSystem.C:143-145): diag with
   all patch internalCoeffs, source with non-coupled boundaryCoeffs, psi0 the
   current field.  lower == NULL means a symmetric matrix. */
int mrfg_set_region(mrfg_ctx *ctx, int sys, int region, int n_cells, int n_faces,
                    const int32_t *l_addr, const int32_t *u_addr,
                    const double *diag, const double *lower, const double *upper,
                    const double *source, const double *psi0);

/* The ddt term of a transient region, for mrfg_time_steps: per cell the
   Euler ddt coefficient a = rho*c*V/dt (the diagonal fvm::ddt adds,
   EulerDdtScheme) and the region's diagonal and source without it
   (ddt_coeff == NULL makes the region steady again).  The region's
   diag / source set by mrfg_set_region are the first step's system. */
int mrfg_set_region_ddt(mrfg_ctx *ctx, int sys, int region, const double *ddt_coeff,
                        const double *diag_base, const double *source_base);

/* One monolithic interface (a coupledPatchPair of regionInterfaceProperties;
   monolithicCouplingFvPatchField.C:381-464).  Side A faces gather B face
   cells with weights w_a over pairs [off_a[i], off_a[i+1]) (idx_a = B face
   index), and vice versa; coeff_* are the coupled patches' boundaryCoeffs. */
int mrfg_set_interface(mrfg_ctx *ctx, int sys, int iface, int region_a, int region_b,
                       int n_faces_a, const int32_t *fc_a, const double *coeff_a,
                       const int32_t *off_a, const int32_t *idx_a, const double *w_a,
                       int n_faces_b, const int32_t *fc_b, const double *coeff_b,
                       const int32_t *off_b, const int32_t *idx_b, const double *w_b);

/* The two conductances of the monolithicTemperature wall value per face of
   each side of interface `iface` (calcTemperature,
   monolithicThermalDiffusivityFvPatchScalarField.C:300-380, radiation off):
   k_own = forig / (1 - w) / |delta| of the face (monolithicBase::korig,
   monolithicBase.C:97-159) and k_nei = the GGI-interpolated other side's
   forig / w / |delta|; positive.  Needed by mrfg_interface_patch_values for
   systems defined from LDU arrays (cases built or read by this library
   carry them).  May be called before or after mrfg_commit. */
int mrfg_set_interface_wall(mrfg_ctx *ctx, int sys, int iface, const double *k_own_a, const double *k_nei_a,
                            const double *k_own_b, const double *k_nei_b);

/* DILU schedule selection for the next mrfg_commit: MRFG_SCHEDULE_AUTO
   (structured hex regions as line tiles, others by levels) or
   MRFG_SCHEDULE_LEVELS (levels everywhere).  Both compute the same factor
   and sweeps; the choice only changes the parallel execution order. */
enum { MRFG_SCHEDULE_AUTO = 0, MRFG_SCHEDULE_LEVELS = 1 };
int mrfg_set_schedule(mrfg_ctx *ctx, int schedule);

/* Build the device layout of every system defined so far and upload it. */
int mrfg_commit(mrfg_ctx *ctx);

/* Value refresh of a committed system defined from LDU arrays: the next
   time step's assembled values on the same mesh.  multiRegionFoam's time
   loop (applications/solvers/multiRegionFoam/multiRegionFoam.C:94-117)
   re-assembles every region's fvMatrix each step (regionType::setCoupledEqns,
   multiRegionSystem.C:115) and hands it to coupledFvMatrix::solve
   (multiRegionSystem.C:143-153); the addressing and the GGI weights stay
   those of the mesh.  These calls replace the values only -- region: diag,
   lower (NULL for a region committed symmetric), upper, source, in the
   arrays' committed order; interface: the boundary coefficients of side a
   and / or side b (a NULL side keeps its values) -- and the device layout,
   DILU schedule and maps built by mrfg_commit are kept: the values are
   staged on the device and put in place (SpMV values, DILU records,
   diagonal, source, row sums, interface coefficients) before the next solve
   or kernel call.  Commit once per mesh, refresh per time step.  Not for
   systems read from disk (their values come from mrfg_set_field /
   mrfg_assemble) or batches with ddt terms (MRFG_ERR_ARG). */
int mrfg_update_region(mrfg_ctx *ctx, int sys, int region, const double *diag, const double *lower,
                       const double *upper, const double *source);
int mrfg_update_interface(mrfg_ctx *ctx, int sys, int iface, const double *coeff_a, const double *coeff_b);

/* coupledFvMatrix<scalar>::solve(dict) for every committed system
   (multiRegionSystem.C:150-153).  perf receives one entry per system. */
int mrfg_solve(mrfg_ctx *ctx, const mrfg_controls *controls, mrfg_perf *perf);

/* Same, with the controls read from an fvSolution text: `entry` names the
   solver dictionary under `solvers` (e.g. "Tcoupled"). */
int mrfg_solve_dict(mrfg_ctx *ctx, const char *fv_solution, const char *entry, mrfg_perf *perf);

/* n_steps time steps of the committed systems: the time loop of
   multiRegionFoam around assembleAndSolveCoupledMatrix
   (multiRegionSystem.C:61-191) for a fixed flow, where each step's
   setCoupledEqns changes only the ddt terms.  They are rebuilt on the
   device from the current and previous fields (no matrix upload per step):
   Euler, or "backward" (constant time step, coefficients 3/2, 2, 1/2; the
   first step of a run, after mrfg_commit or mrfg_reset_psi, is Euler, as
   backwardDdtScheme's first step).  Each step solves from the current field.
   perf receives n_steps x n_systems entries (step-major).  Needs at least
   one transient region (mrfg_set_region_ddt or a case read with a ddt
   scheme); precision fp64. */
int mrfg_time_steps(mrfg_ctx *ctx, const mrfg_controls *controls, const char *ddt_scheme, int n_steps,
                    mrfg_perf *perf);

/* Result of the partitioned coupling of one case (dnaControl). */
typedef struct {
    int iterations;          /* DNA iterations run ("DNA: converged in N iterations") */
    int converged;           /* residual control met (else maxCoupleIter reached) */
    int region_solves;       /* region solves run */
    int solver_iterations;   /* Krylov iterations over all region solves */
    double jump_residual;    /* last normalised temperature-jump residual (max over pairs) */
    double flux_residual;    /* last normalised heat-flux residual (max over pairs) */
    double seconds;          /* wall time of the whole batch's coupling */
} mrfg_dna_perf;

/* Partitioned (Dirichlet-Neumann) coupling of every committed system, each
   a case read with regionCoupledTemperatureJump / regionCoupledHeatFlux
   patches: the loop of multiRegionSystem::solve for partitioned fields
   (multiRegionSystem.C:571-610) under dnaControl::loop (dnaControl.C:
   993-1061), steady.  Per DNA iteration every region in order gets its
   partitioned patches updated on the device -- the Dirichlet side the other
   side's patch values times K, relaxed (fixed or Aitken); the Neumann side
   the gradient that matches the other side's heat flux, relaxed -- and the
   region's systems of all cases still iterating are solved in one batched
   solve (controls, fp64).  Each case stops when its jump and flux residuals
   are within maxJumpRes / maxFluxRes (after the first iteration) or after
   maxCoupleIter iterations, and runs from the committed initial state.  perf
   receives one entry per system; the coupled fields become the systems'
   fields (mrfg_get_psi). */
int mrfg_dna_solve(mrfg_ctx *ctx, const mrfg_controls *controls, mrfg_dna_perf *perf);
/* n_steps time steps of the committed partitioned cases (regions read with
   a ddt scheme), each a DNA loop as in mrfg_dna_solve: the ddt terms are
   rebuilt on the device from the current and previous fields (Euler, or
   "backward" with a first Euler step, as mrfg_time_steps), the interface
   cells take the step's source plus their partitioned terms, and the
   acceleration models start a new time value (correction count reset;
   IQN-ILS keeps columns within the step, couplingReuse 0) while keeping
   their previous field and Aitken factor (accelerationModel.C,
   aitkenRelaxation.C:83-110, IQNILS.C:70-140).  Runs from the committed
   initial state; perf receives n_steps x n_systems entries (step-major). */
int mrfg_dna_time_steps(mrfg_ctx *ctx, const mrfg_controls *controls, const char *ddt_scheme, int n_steps,
                        mrfg_dna_perf *perf);
/* Per DNA iteration of the last mrfg_dna_solve (or of the last step of
   mrfg_dna_time_steps) for system `sys`: the jump
   and flux residuals and the relaxation factor of its first Dirichlet side
   (at most max_iter values each; *n receives the number of iterations). */
int mrfg_dna_history(mrfg_ctx *ctx, int sys, int max_iter, double *jump, double *flux, double *omega, int *n);

/* Parse a solver entry of an fvSolution text (lduSolver::New selection by
   name, keys solver / preconditioner / tolerance / relTol / minIter /
   maxIter).  err receives the message on MRFG_ERR_DICT. */
int mrfg_parse_controls(const char *fv_solution, const char *entry, mrfg_controls *out,
                        char *err, int err_len);

/* Reset every system's field to the psi0 given at definition (a device
   copy; repeated solves from the same initial field). */
int mrfg_reset_psi(mrfg_ctx *ctx);

/* Field in / out in the region's own cell order (the psi that
   coupledFvMatrix::solve writes back into each region's T). */
int mrfg_get_psi(mrfg_ctx *ctx, int sys, int region, double *psi);
int mrfg_set_psi(mrfg_ctx *ctx, int sys, int region, const double *psi);

/* FP64 interface heat-flux conservation check on the device: the heat flow
   into each side of interface `iface` from the system's current field,
   Q = sum_f coeff_f * (T_neighbour,f - T_own,f), with T_neighbour the GGI
   interpolate of the other side (monolithicCouplingFvPatchField.C:381-464
   coefficients; the interface flux of monolithicTemperature...C:204-224).
   A conserving coupled solution gives q_a + q_b = 0. */
int mrfg_interface_flux(mrfg_ctx *ctx, int sys, int iface, double *q_a, double *q_b);

/* The coupled patch values of side `side` (0 = a, 1 = b) of interface
   `iface` from the system's current field, on the device in FP64, as
   coupledFvMatrix::solve's correctBoundaryConditions leaves them on the
   monolithicTemperature patches: t_neighbour = the patch neighbour field,
   the GGI interpolate of the other side's face-cell values
   (monolithicCouplingFvPatchField::patchNeighbourField,
   monolithicCouplingFvPatchField.C:208-236), and t_wall = the patch value,
   (k_own T_own + k_nei t_neighbour) / (k_own + k_nei)
   (monolithicTemperatureFvPatchScalarField::initEvaluate / updateCoeffs ->
   calcTemperature, monolithicTemperatureFvPatchScalarField.C:147-202).  One value per face
   of that side.  MRFG_ERR_ARG if the interface has no wall conductances. */
int mrfg_interface_patch_values(mrfg_ctx *ctx, int sys, int iface, int side, double *t_neighbour, double *t_wall);

/* The FP64 interface heat-flux check of mrfg_interface_flux without any
   device allocation: no counterpart in the reference; for callers that
   cannot allocate device memory (code under HIP graph capture, tight
   serving or time-stepping loops, harnesses that forbid hipMalloc).  Same
   result as mrfg_interface_flux bit for bit.  The reduction partials and
   the two results live in a process-lifetime device array of 256 KiB per
   GPU (taken at mrfg_create), so each side of the interface may have at
   most 1,048,576 faces (MRFG_ERR_ARG beyond); every call on one GPU shares
   that array, so calls that overlap in time on one GPU (other threads or
   contexts) race and must be serialised by the caller.  The interface's
   face lists are resident from mrfg_commit. */
int mrfg_interface_flux_noalloc(mrfg_ctx *ctx, int sys, int iface, double *q_a, double *q_b);

/* ---- finite-volume assembly on the device (cases read from disk) ------ */

/* Replace one field of region `region` of committed system `sys`, a case
   read by mrfg_case_read_foam (the fields setCoupledEqns reads):
   "k"     conductivity, n = n_cells values (positive);
   "U"     velocity, n = 3 * n_cells (x y z per cell; transportTemperature
           regions only; phi = linearInterpolate(U) & Sf is formed on the
           device, transportTemperature.C:86-94);
   "T"     temperature, n = n_cells: the old-time field of the ddt term and
           the system's initial guess for the region;
   "rhoCp" rho*cp (rho*cv for conductTemperature), n = 1.
   The field takes effect at the next mrfg_assemble. */
int mrfg_set_field(mrfg_ctx *ctx, int sys, int region, const char *name, const double *values,
                   long long n);

/* Assemble, on the device, every committed system read from disk from its
   current fields, and put the result in place of its committed matrix
   (SpMV values, DILU records, diagonal, source, interface coefficients,
   ddt terms; no matrix crosses the bus).  This is the assembly half of
   assembleAndSolveCoupledMatrix (multiRegionSystem.C:115 rg.setCoupledEqns()
   per region, then coupledEqns.set at :143-145): the region equations of
   transportTemperature.C:129-150 (upwind convection, Gauss-linear
   Laplacian, fixedValue / zeroGradient / empty / monolithicTemperature
   patches), the interface conductivity of
   monolithicThermalDiffusivityFvPatchScalarField.C:168 (calcThermalDiffusivity,
   radiation off) and the Euler ddt term.  The result equals the host
   reader's assembly of the same fields bit for bit. */
int mrfg_assemble(mrfg_ctx *ctx);

/* The device-assembled LDU matrix of one region in its own cell and face
   order (diag with patch internalCoeffs, source with non-coupled
   boundaryCoeffs; lower == upper for a symmetric region), and one side's
   interface coefficients (side 0 = a, 1 = b).  After mrfg_assemble. */
int mrfg_get_region_matrix(mrfg_ctx *ctx, int sys, int region, double *diag, double *lower,
                           double *upper, double *source);
int mrfg_get_interface_coeff(mrfg_ctx *ctx, int sys, int iface, int side, double *coeff);

/* Bytes the two passes of mrfg_assemble read and write, each array counted
   once (the algorithmic traffic of mrfg_time_kernel 7 and 8). */
int mrfg_assembly_bytes(mrfg_ctx *ctx, double *assemble_bytes, double *place_bytes);

/* ---- case builder ------------------------------------------------------ */

/* Assemble a case dictionary (regions of transportTemperature /
   conductTemperature type and monolithicCoupledPatches) into its coupled
   system: the work of regionType::setCoupledEqns
   (conductTemperature.C:135-152, transportTemperature.C:129-150) and of the
   monolithic patch fields (monolithicThermalDiffusivity...C:167-277,
   GGI weights of ggiInterfaceToInterfaceMapping.C:62-77). */
int mrfg_case_build(const char *dict_text, mrfg_case **out, char *err, int err_len);
void mrfg_case_free(mrfg_case *c);
int mrfg_case_sizes(const mrfg_case *c, int *n_regions, int *n_interfaces);
int mrfg_case_region_sizes(const mrfg_case *c, int region, int *n_cells, int *n_faces, int *symmetric);
int mrfg_case_region(const mrfg_case *c, int region, int32_t *l_addr, int32_t *u_addr,
                     double *diag, double *lower, double *upper, double *source, double *psi0);
/* The ddt split of a transient region of a built case (a case read with an
   Euler / backward ddt): *transient = 1 and, per cell, the Euler ddt
   coefficient a = rho*c*V/dt and the diagonal and source without the ddt
   term (diag = diag_base + a, source = source_base + a*psi0); *transient = 0
   for a steady region (arrays untouched).  Null arrays are skipped. */
int mrfg_case_region_ddt(const mrfg_case *c, int region, int *transient, double *ddt_coeff,
                         double *diag_base, double *source_base);
/* Partitioned coupling of a case read from disk (Dirichlet-Neumann through
   heatTransferInterface pairs: regionInterfacePropertiesPartitioned,
   0/<region>/T with regionCoupledTemperatureJump / regionCoupledHeatFlux
   patches, constant/multiRegionProperties DNA { T { maxCoupleIter;
   residualControl { maxJumpRes; maxFluxRes; } } }, dnaControl.C:40-200).
   n_pairs = 0 for a monolithic case; a residual limit < 0 means none.
   Side 0 is the Dirichlet side, 1 the Neumann side: its face cells, source
   term per unit boundary value (temperature or gradient; added to
   source[fc[i]] after every other boundary contribution), face conductivity,
   patch deltaCoeffs, initial patch values, the GGI gather of the other
   side's faces, acceleration (0 fixed, 1 aitken, 2 IQN-ILS) and relaxation factor; K
   is the pair's temperature jump coefficient.  source_base is a region's
   source without the partitioned faces. */
int mrfg_case_dna_controls(const mrfg_case *c, int *n_pairs, int *max_couple_iter,
                           double *max_jump_res, double *max_flux_res);
int mrfg_case_dna_side_sizes(const mrfg_case *c, int pair, int side, int *region, int *n_faces,
                             int *n_weights, int *acc_type, double *relax, double *K);
int mrfg_case_dna_side(const mrfg_case *c, int pair, int side, int32_t *fc, double *coeff,
                       double *kb, double *dc, double *value0, int32_t *off, int32_t *idx,
                       double *w);
int mrfg_case_region_source_base(const mrfg_case *c, int region, double *source_base);
int mrfg_case_interface_sizes(const mrfg_case *c, int iface, int *region_a, int *region_b,
                              int *n_faces_a, int *n_faces_b, int *n_pairs_a, int *n_pairs_b);
int mrfg_case_interface(const mrfg_case *c, int iface,
                        int32_t *fc_a, double *coeff_a, int32_t *off_a, int32_t *idx_a, double *w_a,
                        int32_t *fc_b, double *coeff_b, int32_t *off_b, int32_t *idx_b, double *w_b);
/* The wall conductances (see mrfg_set_interface_wall) of a built case's
   interface, n_faces_a / n_faces_b values each; MRFG_ERR_STATE if the case
   has none. */
int mrfg_case_interface_wall(const mrfg_case *c, int iface, double *k_own_a, double *k_nei_a, double *k_own_b,
                             double *k_nei_b);
/* Read an OpenFOAM / foam-extend multi-region case directory and assemble
   it like mrfg_case_build: constant/multiRegionProperties (region types),
   constant/<region>/polyMesh (ASCII points, faces, owner, neighbour,
   boundary), 0/<region>/{T,k,U}, constant/<region>/transportProperties,
   constant/regionInterfaceProperties (monolithicCoupledPatches),
   system/controlDict (deltaT) and system/<region>/fvSchemes (ddt), the
   layout of tutorials/conjugateHeatTransfer/flowOverHeatedPlate.  `options`
   is dictionary text (may be empty): fieldDir, interfaceProperties, ddt,
   deltaT.  The equations are those of setCoupledEqns
   (conductTemperature.C:135-152, transportTemperature.C:129-150) with
   phi = interpolate(U) & Sf (transportTemperature.C:86-94); a `backward`
   ddt is taken at its first step (Euler).  Meshes must be orthogonal to
   1 degree; the two patches of an interface must lie in one plane (any
   orientation) and have convex faces, whose overlap areas in that plane,
   rescaled per face, are the GGI weights (ggiInterfaceToInterfaceMapping.C:62-77). */
int mrfg_case_read_foam(const char *case_dir, const char *options, mrfg_case **out, char *err, int err_len);
/* Add a built case as a new system of the context (begin_system +
   set_region + set_interface). */
int mrfg_add_case(mrfg_ctx *ctx, const mrfg_case *c, int *sys_id);

/* ---- kernel-level entry points (tests, profiling) ---------------------- */

/* y = A x over all regions of one system (lduMatrix::Amul plus the
   interface updates); x, y in concatenated region order. */
int mrfg_spmv(mrfg_ctx *ctx, int sys, const double *x, double *y);
/* Reciprocal DILU diagonal (coupled Cholesky calcPreconDiag). */
int mrfg_dilu_factor(mrfg_ctx *ctx, int sys, double *rD);
/* w = M^-1 r with the DILU preconditioner (coupled Cholesky precondition). */
int mrfg_precondition(mrfg_ctx *ctx, int sys, const double *r, double *w);
/* The same in FP32: the single-precision records of the mixed-precision
   solver, with rD rounded from the FP64 factor (no counterpart in the
   reference, which is FP64 only). */
int mrfg_precondition_fp32(mrfg_ctx *ctx, int sys, const float *r, float *w);
/* The mixed solver's FP32 inner preconditioner, MRFG_INNER_TILE_DILU (each
   structured-box tile's DILU alone, couplings across tile edges left out;
   rD of that tile-local factor) or MRFG_INNER_DILU (as mrfg_precondition_fp32).
   No counterpart in the reference. */
int mrfg_precondition_inner(mrfg_ctx *ctx, int sys, int inner_preconditioner, const float *r, float *w);
/* Per cell of system sys (concatenated region order): the index of the
   structured-box tile holding it in the DILU schedule, or -1 for a cell
   scheduled on general waves (the tiles of the tile-local preconditioner). */
int mrfg_dilu_tiles(mrfg_ctx *ctx, int sys, int *tile_of_cell);
/* Average device time (ms) of one launch over `reps` launches, from HIP
   events on the context's stream: which 0 = SpMV, 1 = DILU forward sweep,
   2 = DILU backward sweep, 3 = DILU factorisation, each launched back to
   back; 4 = the first forward sweep, 5 = the first backward sweep and 6 =
   the SpMV with the alpha reduction of each of `reps` BiCGStab iterations
   (run from the initial guess with tolerance 0; psi is restored after);
   7 = the finite-volume assembly passes (faces, interface sides, cells) and
   8 = the placement pass of mrfg_assemble, back to back (systems read from
   disk; the matrices are left as one mrfg_assemble leaves them). */
int mrfg_time_kernel(mrfg_ctx *ctx, int which, int reps, double *ms);
/* Trace one DILU forward/backward pair (tracing facility; no counterpart
   in the reference): out receives, per wave in dispatch order, [start ns,
   end ns, slow-path waits, steps], forward sweep then backward sweep
   (8*min(n_waves, max_waves) values).  n_waves receives the wave count. */
int mrfg_trace_dilu(mrfg_ctx *ctx, long long *out, int max_waves, int *n_waves);
/* The launch configuration the autotuner chose for the committed batch
   (profiling facility; no counterpart in the reference): threads per block
   of the batch kernels, the SpMV's ELL chunk, and the SpMV time (ms) with
   it and with the fixed default (256 threads, chunk 8).  *tuned = 0 before
   the first autotuned solve. */
int mrfg_tuning(mrfg_ctx *ctx, int *tuned, int *block_threads, int *ell_chunk, double *spmv_ms,
                double *spmv_ms_default);
/* Force a launch configuration (block_threads 256, 512 or 1024; ell_chunk
   2, 4, 8 or 16) for the following solves of the committed batch, as if the
   autotuner had chosen it (testing and profiling facility). */
int mrfg_set_tuning(mrfg_ctx *ctx, int block_threads, int ell_chunk);
/* info[0..5] = cells, internal faces, GGI pairs, rows incl. padding,
   64-row slices, DILU waves. */
int mrfg_layout_info(mrfg_ctx *ctx, int sys, long long *info);
/* Canonical algorithmic bytes at value size 8 (fp64) or 4 (fp32). */
int mrfg_bytes(mrfg_ctx *ctx, int sys, int value_size, double *b_iter, double *b_spmv, double *b_dilu);
/* Average time (ms) of a plain device-to-device copy of `bytes` bytes over
   `reps` launches (each reads and writes `bytes`): the reference kernel for
   scaling HBM traffic counters (profiling facility; no counterpart in the
   reference). */
int mrfg_copy_probe(mrfg_ctx *ctx, long long bytes, int reps, double *ms);

/* ---- serving (continuous batching; no counterpart in the reference) ------

   A server admits coupled solves as they arrive.  n_prep host threads
   (0: the host core share -- OMP_NUM_THREADS when set -- less n_workers)
   assemble each request and build its device layout, in arrival order.
   n_workers device slots, each a resident batch on its own HIP stream,
   take the oldest prepared requests whose device footprint fits the HBM
   budget left by the requests in flight (up to max_batch per slot) and
   solve them with one set of launches; slots run concurrently on the
   device, so a request arriving while others run starts as soon as a slot
   and memory are free.  Each request is retired at the convergence check
   that finds it finished, while the rest of its batch continues.  A result
   does not depend on the batch, slot or time it ran in.  Every request is
   solved with the controls of solver entry `entry` of `fv_solution`;
   keep_fields = 1 keeps each result's field for mrfg_server_get_psi.
   Results may be polled while requests run (MRFG_ERR_STATE until done). */
typedef struct mrfg_server mrfg_server;
int mrfg_server_create(mrfg_server **srv, int device, long long hbm_budget_bytes, int n_workers, int max_batch,
                       int n_prep, const char *fv_solution, const char *entry, int keep_fields);
void mrfg_server_destroy(mrfg_server *srv);
/* Queue one request (a case dictionary, as mrfg_case_build reads). */
int mrfg_server_submit(mrfg_server *srv, const char *case_dict, int *request_id);
/* Block until every submitted request has finished or failed. */
int mrfg_server_wait_all(mrfg_server *srv);
/* Result of a finished request: perf, times[3] = submit, start, done (s
   since the server started), and the number of systems of its batch.
   MRFG_ERR_STATE while it runs; MRFG_ERR_DEVICE if it failed
   (mrfg_server_last_error). */
int mrfg_server_result(mrfg_server *srv, int request_id, mrfg_perf *perf, double *times, int *batch_size);
int mrfg_server_get_psi(mrfg_server *srv, int request_id, int region, double *psi);
/* Where a finished request's time went: times[5] = submitted, prepared
   (assembled and laid out on the host), admitted to a device slot, solve
   started (its batch's layout uploaded), done -- seconds since the server
   started. */
int mrfg_server_times(mrfg_server *srv, int request_id, double *times);
const char *mrfg_server_last_error(const mrfg_server *srv);

#ifdef __cplusplus
}
#endif
#endif
