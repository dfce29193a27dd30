/*
 * jetgpu.h -- C ABI of the MI355X (gfx950) pressure-projection solver.
 *
 * This is the boundary a host build of Jet (doyubkim/fluid-engine-dev) binds
 * to replace its CPU linear solvers for grid pressure projection.  Plain C
 * types only: sizes, host pointers, device pointers, opaque handles.  Every
 * function returns a status code (JG_OK == 0); jg_last_error() describes the
 * last failure of a solver.
 *
 * Data model (identical to Jet's, so no conversion is needed at the
 * boundary):
 *   - cell arrays are x-fastest, index = i + nx*(j + ny*k)
 *     (Jet Array3, include/jet/detail/array3-inl.h:142-151);
 *   - a matrix is an array of FdmMatrixRow3 {center, right, up, front}
 *     doubles, 32 bytes per cell (include/jet/fdm_linear_system3.h:18-30):
 *     the 7-point symmetric stencil stored as its upper couplings;
 *   - a multigrid system is a list of levels, level 0 finest, each level
 *     exactly half the previous one (FdmMgLinearSystem3,
 *     include/jet/fdm_mg_linear_system3.h:23-55).
 *
 * Threading: one jg_solver is used by one host thread at a time (like Jet's
 * solver objects, which keep scratch vectors as members).
 */
#ifndef JETGPU_H_
#define JETGPU_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
enum {
    JG_OK = 0,
    JG_ERR_INVALID_ARG = 1,  /* Jet: JET_THROW_INVALID_ARG_IF (size mismatch) */
    JG_ERR_OUT_OF_MEMORY = 2,
    JG_ERR_HIP = 3,          /* HIP runtime error (message in jg_last_error) */
    JG_ERR_NONFINITE = 4,    /* NaN/Inf met in a reduction */
    JG_ERR_NO_DEVICE = 5,
    JG_ERR_UNSUPPORTED = 6   /* e.g. Jet's no-op entry points */
};

enum { JG_FP64 = 0, JG_MIXED = 1 };

/* Stop rules.  JG_STOP_JET is Jet's pcg rule: stop when r.z <= tol^2
 * (include/jet/detail/cg-inl.h:48).  JG_STOP_RELATIVE stops when
 * ||b - A x||_2 <= rel_tol * ||b||_2 (FP64 residual). */
enum { JG_STOP_JET = 0, JG_STOP_RELATIVE = 1 };

/* V-cycle smoothers: red-black Gauss-Seidel with SOR factor omega
 * (FdmGaussSeidelSolver3::relaxRedBlack, src/jet/fdm_gauss_seidel_solver3.cpp:
 * 150-221), or weighted Jacobi x += omega ((b - s) / c - x)
 * (FdmJacobiSolver3::relax, src/jet/fdm_jacobi_solver3.cpp:95-112, weighted). */
enum { JG_SMOOTH_RBGS = 0, JG_SMOOTH_JACOBI = 1 };

typedef struct jg_solver jg_solver;

/* MgParameters + FdmMgSolver3 knobs (include/jet/mg.h:58-86,
 * include/jet/fdm_mg_solver3.h:24-39).  Defaults: see jg_mg_params_default. */
typedef struct {
    size_t max_levels;
    unsigned restrict_iter;  /* pre-smoothing sweeps per level           */
    unsigned correct_iter;   /* post-smoothing sweeps on levels > 0      */
    unsigned coarsest_iter;  /* sweeps on the coarsest level             */
    unsigned final_iter;     /* post-smoothing sweeps on level 0         */
    double max_tol;          /* MGPCG tolerance (Jet: sqrt(r.z) <= max_tol) */
    double sor;              /* Jet's sorFactor                          */
    int red_black;           /* Jet's useRedBlackOrdering                */
} jg_mg_params;

/* GPU-only options (no Jet counterpart).  Zero-initialised = defaults. */
typedef struct {
    int precision;           /* JG_FP64 (default) or JG_MIXED            */
    int stop_rule;           /* JG_STOP_JET (default) or JG_STOP_RELATIVE */
    double rel_tol;          /* for JG_STOP_RELATIVE (default 1e-8)      */
    int smoother;            /* JG_SMOOTH_RBGS (default) or JG_SMOOTH_JACOBI */
    double omega;            /* RBGS: relaxation in (0,2), 0 -> Jet's sor;
                                Jacobi: weight in (0,1], 0 -> 6/7 */
    unsigned max_refine;     /* mixed: outer FP64 refinement steps (default 8) */
    double inner_rel_tol;    /* mixed: FP32 inner reduction per step (default 1e-4) */
    unsigned check_every;    /* host convergence poll interval in iterations (default 1) */
    int sweep_pairs;         /* smoothing: two sweeps per launch where the tuner
                                measured it faster (0, default), wherever two sweeps
                                follow each other (1), never (-1); the result is
                                bit-identical either way */
} jg_gpu_options;

typedef struct {
    unsigned iterations;     /* PCG iterations (all inner iterations in mixed mode) */
    unsigned refinements;    /* mixed: outer refinement steps taken */
    double last_residual;    /* Jet's lastResidual(): sqrt(|r.z|) at exit */
    double rel_residual;     /* ||b - A x||_2 / ||b||_2 in FP64 at exit */
    int converged;           /* Jet's bool return value */
    double solve_ms;         /* device time of the solve (HIP events) */
} jg_solve_info;

/* ----------------------------------------------------------- lifetime */
void jg_mg_params_default(jg_mg_params* p);
void jg_gpu_options_default(jg_gpu_options* o);

/* Creates a solver bound to HIP device `device` with its own stream. */
int jg_solver_create(int device, jg_solver** out);
void jg_solver_destroy(jg_solver* s);
const char* jg_last_error(const jg_solver* s);
/* Device stream used by the solver (a hipStream_t). */
void* jg_solver_stream(jg_solver* s);

/* --------------------------------------------- device-resident system */
/* Declares the level sizes and allocates device storage (cached: a call
 * with the same sizes allocates nothing).  Level l+1 must be exactly half of
 * level l in every direction -- except in a 2-D hierarchy, one whose finest
 * level is one plane thick (nz[0] == 1) and which has coarser levels: its
 * levels all have nz == 1 and halve in x and y (FdmMgLinearSystem2,
 * include/jet/fdm_mg_linear_system2.h:25-60; rows are FdmMatrixRow2 with
 * front = 0), and jg_solver_mgpcg / jg_mg_vcycle then use FdmMgUtils2's
 * transfers (FdmMgpcgSolver2, src/jet/fdm_mgpcg_solver2.cpp:20-92). */
int jg_solver_set_levels(jg_solver* s, size_t nlev, const size_t* nx,
                         const size_t* ny, const size_t* nz);
/* Device pointers of the resident system: matrix rows of level l (32-byte
 * FdmMatrixRow3), the finest right-hand side b and solution x. */
double* jg_solver_matrix(jg_solver* s, size_t level);
double* jg_solver_rhs(jg_solver* s);
double* jg_solver_solution(jg_solver* s);
/* Host <-> device copies of the resident system (b may be NULL). */
int jg_solver_upload(jg_solver* s, size_t level, const double* rows,
                     const double* b);
int jg_solver_download_x(jg_solver* s, double* x);
/* Marks the matrices as changed after they were written in place on the
 * device (refreshes the FP32 shadow copies used by JG_MIXED). */
int jg_solver_matrices_changed(jg_solver* s);

/* FdmCgSolver3::solve(FdmLinearSystem3*) on the resident level 0
 * (src/jet/fdm_cg_solver3.cpp:20-47): x <- 0, Jet's cg<FdmBlas3> control
 * flow, Jet's bool in info->converged. */
int jg_solver_cg(jg_solver* s, unsigned max_iter, double tol,
                 jg_solve_info* info);

/* FdmMgpcgSolver3::solve(FdmMgLinearSystem3*) on the resident hierarchy
 * (src/jet/fdm_mgpcg_solver3.cpp:54-80): x <- 0, PCG on level 0
 * preconditioned by one V-cycle over all levels.  The GPU V-cycle is
 * symmetric (zero initial guess, red-black pre-sweeps, black-red
 * post-sweeps), unlike Jet's warm-started lexicographic SOR. */
int jg_solver_mgpcg(jg_solver* s, unsigned max_iter, const jg_mg_params* p,
                    const jg_gpu_options* o, jg_solve_info* info);

/* z = M r: one application of jg_solver_mgpcg's preconditioner -- the
 * symmetric V-cycle over the solver's hierarchy -- to device vectors r and z
 * (finest-level size, r != z), in opt->precision (JG_MIXED: FP32 V-cycle,
 * r and z converted).  Replaces the mgVCycle call FdmMgpcgSolver3's
 * preconditioner makes (include/jet/mg.h:101, include/jet/detail/mg-inl.h:
 * 17-113, called at src/jet/fdm_mgpcg_solver3.cpp:24-29). */
int jg_mg_vcycle(jg_solver* s, const jg_mg_params* p, const jg_gpu_options* o,
                 const double* r, double* z);

/* FdmMgSolver3::solve(FdmMgLinearSystem3*) on the resident hierarchy
 * (src/jet/fdm_mg_solver3.cpp:69-74): one V-cycle from the resident x, which
 * is NOT zeroed (Jet's mgVCycle starts from the caller's x): x += M (b - A x)
 * with jg_mg_vcycle's M.  info->last_residual is ||b - A x||_2 afterwards
 * (mgVCycle's lastResidualNorm, include/jet/detail/mg-inl.h:62-66) and
 * info->converged is Jet's result, last_residual < p->max_tol;
 * info->iterations is 1. */
int jg_solver_mg(jg_solver* s, const jg_mg_params* p, const jg_gpu_options* o,
                 jg_solve_info* info);

/* ---------------------------------------------------- batched solves */
/* Many independent systems of uneven sizes solved together.  Systems of
 * one hierarchy (finest size and level count) form a group held back to
 * back in device memory; every launch of the V-cycle and of the PCG
 * iteration covers a whole group, each block taking its system from its grid
 * index and exiting at once once that system has stopped (its own
 * convergence flag).  jg_batch_system(b, i) is system i's own handle (fill
 * it with jg_solver_upload / jg_sp_build / jg_fr_build, read it with
 * jg_solver_download_x; it keeps the hierarchy it was created with), and
 * jg_batch_mgpcg runs the MGPCG of every system as one captured hipGraph (one
 * branch per group) replayed until every system has stopped.  No
 * counterpart in Jet, which solves one system per FdmLinearSystemSolver3::
 * solve call (include/jet/fdm_linear_system_solver3.h:17-31); per system the
 * semantics are jg_solver_mgpcg's (3-D hierarchies, red-black smoothing). */
typedef struct jg_batch jg_batch;
/* nsys systems; system i has nlev[i] levels halving from nx[i] x ny[i] x nz[i]. */
int jg_batch_create(int device, size_t nsys, const size_t* nx, const size_t* ny,
                    const size_t* nz, const size_t* nlev, jg_batch** out);
void jg_batch_destroy(jg_batch* b);
const char* jg_batch_last_error(const jg_batch* b);
size_t jg_batch_size(const jg_batch* b);
jg_solver* jg_batch_system(jg_batch* b, size_t i);
/* infos: one per system (may be NULL); total_ms: device time of the batch. */
int jg_batch_mgpcg(jg_batch* b, unsigned max_iter, const jg_mg_params* p,
                   const jg_gpu_options* o, jg_solve_info* infos, double* total_ms);

/* ------------------------------------------------ host-memory drop-ins */
/* FdmCgSolver3::solveCompressed (src/jet/fdm_cg_solver3.cpp:49-74): Jet's cg
 * (include/jet/detail/cg-inl.h:15-94) on a compressed system
 * (FdmCompressedLinearSystem3, include/jet/fdm_linear_system3.h: rows for the
 * fluid cells only), with FdmCompressedBlas3's mvm and residual
 * (src/jet/fdm_linear_system3.cpp:160-199) on the device.  Host arrays
 * exactly as MatrixCsrD stores them: `values` and `col_idx` (nnz entries,
 * each row's in ascending column order), `row_ptr` (rows + 1, from 0 to
 * nnz); b and x have `rows` entries.  x is overwritten (Jet zeroes it).
 * Rejects (JG_ERR_INVALID_ARG, before any device work) row pointers that do
 * not run from 0 to nnz or decrease, columns >= rows, and sizes >= 2^31. */
int jg_csr_cg_solve_host(jg_solver* s, size_t rows, size_t nnz, const double* values,
                         const size_t* row_ptr, const size_t* col_idx, const double* b,
                         double* x, unsigned max_iter, double tol, jg_solve_info* info);
/* FdmCompressedBlas3::mvm and ::residual (src/jet/fdm_linear_system3.cpp:
 * 160-199) on the device for a host MatrixCsrD as above: result = A v,
 * result = b - A x (rows entries each); bitwise Jet's. */
int jg_csr_mvm_host(jg_solver* s, size_t rows, size_t nnz, const double* values,
                    const size_t* row_ptr, const size_t* col_idx, const double* v,
                    double* result);
int jg_csr_residual_host(jg_solver* s, size_t rows, size_t nnz, const double* values,
                         const size_t* row_ptr, const size_t* col_idx, const double* x,
                         const double* b, double* result);

/* One-call forms of the two solves on host arrays: upload, solve, download
 * x.  rows[l] / b are host arrays in Jet's layout. */
int jg_cg_solve_host(jg_solver* s, size_t nx, size_t ny, size_t nz,
                     const double* rows, const double* b, double* x,
                     unsigned max_iter, double tol, jg_solve_info* info);
int jg_mgpcg_solve_host(jg_solver* s, size_t nlev, const size_t* nx,
                        const size_t* ny, const size_t* nz,
                        const double* const* rows, const double* b, double* x,
                        unsigned max_iter, const jg_mg_params* p,
                        const jg_gpu_options* o, jg_solve_info* info);
/* jg_solver_mg on host arrays: x is read (the V-cycle's starting value) and
 * overwritten. */
int jg_mg_solve_host(jg_solver* s, size_t nlev, const size_t* nx,
                     const size_t* ny, const size_t* nz,
                     const double* const* rows, const double* b, double* x,
                     const jg_mg_params* p, const jg_gpu_options* o,
                     jg_solve_info* info);

/* The flat-entry MGPCG: FdmLinearSystemSolver3::solve(FdmLinearSystem3*)
 * (include/jet/fdm_linear_system_solver3.h:23) with multigrid, which Jet
 * itself does not offer (FdmMgSolver3::solve(FdmLinearSystem3*) is final and
 * does nothing, include/jet/fdm_mg_solver3.h:42), so that callers holding a
 * single-level system -- GridBackwardEulerDiffusionSolver3 (src/jet/
 * grid_backward_euler_diffusion_solver3.cpp:42,71,85,99) -- get multigrid.
 * The hierarchy is sized by Jet's FdmMgUtils3::resizeArrayWithFinest rule
 * with p->max_levels (x and y only for a one-plane system) and its coarse
 * matrices are built on the device from A (jg_solver_coarsen); then
 * jg_solver_mgpcg.  A and b are host arrays in Jet's layout; x is
 * overwritten. */
int jg_mgpcg_flat_solve_host(jg_solver* s, size_t nx, size_t ny, size_t nz,
                             const double* rows, const double* b, double* x,
                             unsigned max_iter, const jg_mg_params* p,
                             const jg_gpu_options* o, jg_solve_info* info);
/* Coarse matrices of every level below the finest from the finest level's
 * resident A (uploaded or built on the device).  A coarse cell aggregates
 * the 2x2x2 fine cells it covers (2x2 on one-plane levels); with A = M + L
 * (M the row sums, L the couplings) its row is that of kappa (P'AP + P'MP),
 * kappa = 1/16 (1/8), P the piecewise-constant injection: a uniform stencil
 * keeps its mass and quarters its couplings (the operator at twice the
 * spacing).  No counterpart in Jet, whose multigrid callers build each level
 * from markers; FP64, bitwise equal to the oracle's or_mg_coarsen. */
int jg_solver_coarsen(jg_solver* s);

/* ------------------------------------ FdmBlas3 / FdmMgUtils3 on device */
/* All pointers are device pointers; these run on the solver's stream and
 * return after the work is complete. */
/* FdmBlas3::mvm -- src/jet/fdm_linear_system3.cpp:79-96 */
int jg_blas_mvm(jg_solver* s, size_t nx, size_t ny, size_t nz,
                const double* rows, const double* x, double* y);
/* FdmBlas3::residual -- src/jet/fdm_linear_system3.cpp:98-116 */
int jg_blas_residual(jg_solver* s, size_t nx, size_t ny, size_t nz,
                     const double* rows, const double* x, const double* b,
                     double* r);
/* FdmBlas3::dot -- src/jet/fdm_linear_system3.cpp:49-65 (FP64) */
int jg_blas_dot(jg_solver* s, size_t n, const double* a, const double* b,
                double* result);
/* FdmBlas3::axpy -- src/jet/fdm_linear_system3.cpp:67-77: out = a*x + y */
int jg_blas_axpy(jg_solver* s, size_t n, double a, const double* x,
                 const double* y, double* out);
/* FdmGaussSeidelSolver3::relaxRedBlack -- src/jet/fdm_gauss_seidel_solver3.cpp:150-221 */
int jg_relax_red_black(jg_solver* s, size_t nx, size_t ny, size_t nz,
                       const double* rows, const double* b, double sor,
                       double* x);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): jg_relax_red_black with the same arguments and the same
 * result, bit for bit, but the copy of x that the sweep reads lives in a
 * static per-GPU array of 2^21 doubles (16 MiB) instead of a temporary, and
 * nothing is allocated or freed.  Grids of more than 2^21 cells return
 * JG_ERR_INVALID_ARG.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (two streams, two threads) race. */
int jg_relax_red_black_noalloc(jg_solver* s, size_t nx, size_t ny, size_t nz,
                               const double* rows, const double* b, double sor,
                               double* x);
/* FdmJacobiSolver3::relax -- src/jet/fdm_jacobi_solver3.cpp:95-112 */
int jg_relax_jacobi(jg_solver* s, size_t nx, size_t ny, size_t nz,
                    const double* rows, const double* b, const double* x,
                    double* xtemp);
/* FdmMgUtils3::restrict -- src/jet/fdm_mg_linear_system3.cpp:43-98
 * (coarse size cnx*cny*cnz, fine size twice that) */
int jg_mg_restrict(jg_solver* s, size_t cnx, size_t cny, size_t cnz,
                   const double* fine, double* coarse);
/* FdmMgUtils3::correct -- src/jet/fdm_mg_linear_system3.cpp:100-179
 * (fine += prolongation of coarse; fine size fnx*fny*fnz) */
int jg_mg_correct(jg_solver* s, size_t fnx, size_t fny, size_t fnz,
                  const double* coarse, double* fine);
/* FdmMgUtils2::restrict -- src/jet/fdm_mg_linear_system2.cpp:44-84
 * (coarse size cnx*cny, fine size 2cnx * 2cny) */
int jg_mg_restrict2(jg_solver* s, size_t cnx, size_t cny, const double* fine, double* coarse);
/* FdmMgUtils2::correct -- src/jet/fdm_mg_linear_system2.cpp:86-143
 * (fine += prolongation of coarse; fine size fnx*fny, both even) */
int jg_mg_correct2(jg_solver* s, size_t fnx, size_t fny, const double* coarse, double* fine);

/* ------------------------------------------------- kernel measurement */
/* Times `reps` back-to-back launches of one level-0 kernel of the solve on
 * the resident system with HIP events on the solver's stream (the solution
 * and work vectors are overwritten; the matrices and the right-hand side are
 * kept, so a later solve is unchanged).  Returns the average launch duration
 * and the kernel's algorithmic HBM bytes per launch (compulsory reads and
 * writes of the cells it touches, 4 coefficients per matrix row). */
enum {
    JG_KERNEL_RBGS = 0,       /* one colour of a red-black sweep              */
    JG_KERNEL_DIR_SPMV = 1,   /* d = z + beta d, q = A d, partial d.q          */
    JG_KERNEL_RESIDUAL = 2,   /* r = b - A x                                   */
    JG_KERNEL_UPDATE_XR = 3,  /* x += alpha d, r -= alpha q                    */
    JG_KERNEL_DOT = 4,        /* partial r.z                                   */
    JG_KERNEL_COPY = 5,       /* plain copy of one vector (bandwidth reference) */
    JG_KERNEL_SWEEP = 6,      /* fused red-black sweep, both colours           */
    JG_KERNEL_SWEEP2 = 7,     /* two fused sweeps in one pass (JG_MIXED only);
                                 bytes = two sweeps' algorithmic bytes          */
    JG_KERNEL_PRE0 = 8,       /* the first level-0 pre-smoothing launch of a
                                 V(2,2) cycle as the MGPCG solve issues it (its
                                 sweep plan and tuning, omega 1.3); bytes = its
                                 sweeps plus the residual it writes           */
    JG_KERNEL_POST0 = 9       /* the first level-0 post-smoothing launch (the
                                 coarse correction added as x is loaded)      */
};
int jg_solver_time_kernel(jg_solver* s, int kernel, int precision,
                          unsigned reps, double* avg_ms, double* bytes);

/* ------------------------------------- pressure-system assembly on device */
/* GridSinglePhasePressureSolver3::buildMarkers + buildSystem
 * (src/jet/grid_single_phase_pressure_solver3.cpp:248-399) on the solver's
 * resident hierarchy (jg_solver_set_levels first).  Device inputs: face
 * velocities u ((nx+1)*ny*nz), v (nx*(ny+1)*nz), w (nx*ny*(nz+1)) and the
 * boundary / fluid SDFs sampled at the cell centres of the finest level.
 * Markers on every level (coarse: 4^3 majority vote), A on every level
 * (re-discretised at spacing 2^l h), b = div u on level 0.  The coarse
 * right-hand sides Jet also builds are not built: its V-cycle overwrites them
 * by restriction before any use (include/jet/detail/mg-inl.h:29-31). */
int jg_sp_build(jg_solver* s, const double* u, const double* v, const double* w,
                const double* boundary_sdf, const double* fluid_sdf, double hx,
                double hy, double hz);
/* applyPressureGradient (src/jet/grid_single_phase_pressure_solver3.cpp:401-431)
 * with the resident solution x and level-0 markers; device arrays; the
 * output may alias the input. */
int jg_sp_apply_gradient(jg_solver* s, const double* u, const double* v,
                         const double* w, double* uo, double* vo, double* wo,
                         double hx, double hy, double hz);
/* GridSinglePhasePressureSolver2::buildMarkers + buildSystem
 * (src/jet/grid_single_phase_pressure_solver2.cpp:218-354) on a resident 2-D
 * hierarchy (one-plane levels, jg_solver_set_levels with nz = 1).  Device
 * inputs: u ((nx+1)*ny), v (nx*(ny+1)), the boundary / fluid SDFs at the
 * finest cell centres.  Markers on every level (coarse: 4^2 majority vote),
 * A on every level (spacing 2^l h, front = 0), b = div u on level 0. */
int jg_sp_build2(jg_solver* s, const double* u, const double* v, const double* boundary_sdf,
                 const double* fluid_sdf, double hx, double hy);
/* applyPressureGradient (src/jet/grid_single_phase_pressure_solver2.cpp:356-378)
 * with the resident solution and level-0 markers of a 2-D hierarchy. */
int jg_sp_apply_gradient2(jg_solver* s, const double* u, const double* v, double* uo,
                          double* vo, double hx, double hy);
/* Device pointer of the markers of a level (int8: 0 fluid, 1 air, 2 boundary). */
const signed char* jg_solver_markers(jg_solver* s, size_t level);

/* GridBackwardEulerDiffusionSolver3's system (buildMarkers, buildMatrix,
 * buildVectors; src/jet/grid_backward_euler_diffusion_solver3.cpp:175-330) on
 * the device for one data grid of the solver's finest-level size: a scalar
 * grid's data, or one component of a collocated or face-centred vector grid
 * (its u, v or w array, with the level sized to that array).  f is the field,
 * boundary_sdf / fluid_sdf the SDFs at its data points; dirichlet != 0 selects
 * Jet's Dirichlet boundary type (Neumann otherwise).  A (level 0) and b become
 * the resident system: jg_solver_cg then gives the diffused field in the
 * solution vector, as Jet's solve (:25-50) with a CG linear solver. */
int jg_diffusion_build(jg_solver* s, const double* f, const double* boundary_sdf,
                       const double* fluid_sdf, double dt, double mu, double hx, double hy,
                       double hz, int dirichlet);

/* The compressed form of jg_sp_build: buildMarkers and the compressed
 * buildSingleSystem of GridSinglePhasePressureSolver3
 * (src/jet/grid_single_phase_pressure_solver3.cpp:81-171, 248-269) on the
 * device, for the finest level: one row per fluid cell in x-fastest order,
 * its entries in ascending column order (as MatrixCsr::addRow keeps them),
 * b = div u.  The system is the solver's resident compressed system (solve
 * it with jg_solver_csr_cg); *rows and *nnz receive its size. */
int jg_sp_build_compressed(jg_solver* s, const double* u, const double* v, const double* w,
                           const double* boundary_sdf, const double* fluid_sdf, double hx,
                           double hy, double hz, size_t* rows, size_t* nnz);
/* The compressed 2-D single-phase build (src/jet/grid_single_phase_pressure_solver2.cpp:
 * 69-130, rows in MatrixCsr::addRow's sorted order) of a one-plane finest level:
 * u ((nx+1)*ny), v (nx*(ny+1)) and the SDFs as for jg_sp_build2; otherwise as
 * jg_sp_build_compressed. */
int jg_sp_build_compressed2(jg_solver* s, const double* u, const double* v,
                            const double* boundary_sdf, const double* fluid_sdf, double hx,
                            double hy, size_t* rows_out, size_t* nnz_out);
/* FdmCgSolver3::solveCompressed (src/jet/fdm_cg_solver3.cpp:49-74) on the
 * resident compressed system (jg_sp_build_compressed or the last
 * jg_csr_*_host call). */
int jg_solver_csr_cg(jg_solver* s, unsigned max_iter, double tol, jg_solve_info* info);
/* GridSinglePhasePressureSolver3::decompressSolution (:330-341): the
 * compressed solution into the solver's finest-level solution vector
 * (jg_solver_solution), 0 at non-fluid cells -- after which
 * jg_sp_apply_gradient projects as for a flat solve. */
int jg_csr_decompress(jg_solver* s);
/* Device pointers of the resident compressed system: jg_csr_array 0 values,
 * 1 b, 2 x (double); jg_csr_index 0 row pointers, 1 column indices (int32). */
const double* jg_csr_array(jg_solver* s, int which);
const int* jg_csr_index(jg_solver* s, int which);

/* GridFractionalSinglePhasePressureSolver3::buildWeights + buildSystem
 * (src/jet/grid_fractional_single_phase_pressure_solver3.cpp:520-713) on the
 * solver's resident hierarchy.  Device inputs: face velocities u, v, w; the
 * collider SDF at the (nx+1)*(ny+1)*(nz+1) grid vertices of the finest level
 * (the corners buildWeights samples, :553-617); the fluid SDF at the finest
 * cell centres (stored as float, as Jet does, :548-551); collider velocity
 * components at the u / v / w face centres (bu, bv, bw; any may be NULL for
 * a resting collider).  Face weights and fluid SDF on every level (coarse:
 * float restriction, :29-115, :619-635), A on every level (spacing 2^l h),
 * b on level 0; coarse right-hand sides are not built (as jg_sp_build). */
int jg_fr_build(jg_solver* s, const double* u, const double* v, const double* w,
                const double* boundary_sdf_vertex, const double* fluid_sdf,
                const double* bu, const double* bv, const double* bw, double hx,
                double hy, double hz);
/* applyPressureGradient, variational form
 * (src/jet/grid_fractional_single_phase_pressure_solver3.cpp:715-765) with the
 * resident solution x and the level-0 weights / fluid SDF of jg_fr_build;
 * device arrays; the output may alias the input. */
int jg_fr_apply_gradient(jg_solver* s, const double* u, const double* v,
                         const double* w, double* uo, double* vo, double* wo,
                         double hx, double hy, double hz);
/* GridFractionalSinglePhasePressureSolver2::buildWeights + buildSystem
 * (src/jet/grid_fractional_single_phase_pressure_solver2.cpp:414-563) on the
 * solver's resident 2-D hierarchy (one-plane levels, or one nx*ny*1 level).
 * Device inputs: face velocities u ((nx+1)*ny) and v (nx*(ny+1)); the collider
 * SDF at the cell centres of the grid extended by one ghost cell on every
 * side ((nx+2)*(ny+2), cell (i,j) at (i+1) + (nx+2)(j+1): the points a 2-D
 * face weight samples, :440-480); the fluid SDF at the cell centres; the
 * collider velocity x at the u faces and y at the v faces (bu, bv; NULL for a
 * resting collider).  Face weights and fluid SDF on every level (float
 * restriction, :29-93), A on every level (spacing 2^l h), b on level 0. */
int jg_fr_build2(jg_solver* s, const double* u, const double* v,
                 const double* boundary_sdf_ghost, const double* fluid_sdf,
                 const double* bu, const double* bv, double hx, double hy);
/* applyPressureGradient, 2-D variational form
 * (src/jet/grid_fractional_single_phase_pressure_solver2.cpp:565-600) with the
 * resident solution x and the level-0 weights of jg_fr_build2; the output may
 * alias the input. */
int jg_fr_apply_gradient2(jg_solver* s, const double* u, const double* v,
                          double* uo, double* vo, double hx, double hy);
/* Device pointers of a level's float arrays built by jg_fr_build or jg_fr_build2: axis 0/1/2
 * = u/v/w face weights, 3 = fluid SDF (NULL before the first build). */
const float* jg_fr_array(jg_solver* s, size_t level, int which);

/* ------------------------------- velocity boundary conditions on device */
/* A rigid-body collider: RigidBodyCollider3's linear and angular velocity
 * (include/jet/rigid_body_collider3.h), its surface's transform translation
 * (the centre of the rotation, src/jet/rigid_body_collider3.cpp:25-28) and
 * Collider3's friction coefficient. */
typedef struct jg_rigid_collider {
    double linear_velocity[3];
    double angular_velocity[3];
    double translation[3];
    double friction;
} jg_rigid_collider;

/* Closed domain walls: Jet's kDirection* bits (include/jet/constants.h:312-335). */
enum {
    JG_DIR_LEFT = 1, JG_DIR_RIGHT = 2, JG_DIR_DOWN = 4, JG_DIR_UP = 8,
    JG_DIR_BACK = 16, JG_DIR_FRONT = 32, JG_DIR_ALL = 63
};

/* GridFractionalBoundaryConditionSolver3::constrainVelocity
 * (src/jet/grid_fractional_boundary_condition_solver3.cpp:26-232): faces fully
 * inside the collider take its velocity, fluid velocity is extrapolated
 * extrapolation_depth cells into the collider, velocities inside it lose
 * their normal component relative to the collider (friction scales the
 * tangential one), and closed domain walls get zero normal velocity.
 * blocked != 0: GridBlockedBoundaryConditionSolver3::constrainVelocity
 * (src/jet/grid_blocked_boundary_condition_solver3.cpp:23-69), which then also
 * sets every face between a collider cell and a fluid cell to the collider
 * velocity.  The grid is the solver's finest level (nx*ny*nz cells) with
 * spacing h and origin o; u, v, w are device face arrays updated in place;
 * collider_sdf is the collider's signed distance at the cell centres (Jet's
 * colliderSdf grid; DBL_MAX everywhere when there is no collider); collider
 * NULL means a resting collider (zero velocity and friction).  Jet's caller
 * passes extrapolation_depth = ceil(max CFL) (src/jet/grid_fluid_solver3.cpp:
 * 346-353).  Scratch: 8 B + 2 B per face, allocated on first use. */
int jg_constrain_velocity(jg_solver* s, int blocked, const double h[3], const double o[3],
                          double* u, double* v, double* w, const double* collider_sdf,
                          const jg_rigid_collider* collider, unsigned extrapolation_depth,
                          int closed_walls);

/* --------------------------------------------------- synthetic scenes */
/* Deterministic synthetic inputs (counter-based SplitMix64, jg_synth.h):
 *   JG_SCENE_CLOSED_SPHERE : solid sphere r = n/4 at the centre, all fluid,
 *                            solid walls (all-Neumann, singular);
 *   JG_SCENE_DAM_BREAK     : fluid in x < n/2, y < 0.4 n, air elsewhere,
 *                            v += -9.8/60;
 *   JG_SCENE_SPHERE_POOL   : 8 solid spheres, radii U(8,32)*n/256, fluid
 *                            below a plane placed so the fluid fraction is
 *                            `fluid_fraction`;
 *   JG_SCENE_BATCH_MEMBER  : 1-4 spheres, radii U(n/16, n/6), fluid below
 *                            y = U(0.3, 0.8) n;
 *   JG_SCENE_TORUS_POOL    : torus (R = 150/512 n, r = 50/512 n) around the
 *                            y axis through the centre, fluid below 0.35 n.
 * Velocities are U(-1,1) with domain-wall faces and faces touching a solid
 * cell set to zero (the blocked boundary condition). */
enum {
    JG_SCENE_CLOSED_SPHERE = 1,
    JG_SCENE_DAM_BREAK = 2,
    JG_SCENE_SPHERE_POOL = 3,
    JG_SCENE_BATCH_MEMBER = 4,
    JG_SCENE_TORUS_POOL = 5
};

typedef struct {
    int kind;
    size_t n;                 /* cubic grid edge (cells)           */
    double h;                 /* grid spacing (0 -> 1)             */
    unsigned long long seed;
    double fluid_fraction;    /* JG_SCENE_SPHERE_POOL (0 -> 0.6)   */
} jg_scene_desc;

/* Host generator (no GPU needed).  Any output pointer may be NULL.
 * bsdf_vertex is the boundary SDF at the (n+1)^3 grid vertices. */
int jg_scene_generate_host(const jg_scene_desc* d, double* u, double* v,
                           double* w, double* bsdf_cc, double* fsdf_cc,
                           double* bsdf_vertex);
/* Same inputs generated on the device (device pointers; any may be NULL). */
int jg_scene_generate_device(jg_solver* s, const jg_scene_desc* d, double* u,
                             double* v, double* w, double* bsdf_cc,
                             double* fsdf_cc, double* bsdf_vertex);

#ifdef __cplusplus
}
#endif

#endif /* JETGPU_H_ */
