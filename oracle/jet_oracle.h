/*
 * jet_oracle.h -- CPU restatement of Jet's pressure-projection solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * solver is compared against.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load it; the product library never
 * links it and never falls back to it.
 *
 * Every function restates the reference code cited beside it (paths are
 * relative to the doyubkim/fluid-engine-dev source tree).  Storage follows
 * Jet's Array3 layout: x fastest, index = i + nx*(j + ny*k)
 * (include/jet/detail/array3-inl.h:142-151).  A matrix row is Jet's
 * FdmMatrixRow3 {center, right, up, front} (include/jet/fdm_linear_system3.h:18-30).
 *
 * Pinning (how each group is checked independently of this code):
 *   - BLAS (mvm/residual/dot/axpy):   bitwise against oracle/_ref (Jet's own
 *     FdmBlas3 compiled from its sources) and against a dense brute-force
 *     matrix-vector product on <= 6^3 grids.
 *   - CG / MG / MGPCG / GS / Jacobi:  iteration-for-iteration against
 *     oracle/_ref on seeded systems; Jet's 2x2 CG known answer (x=(1/11,7/11));
 *     the closed form x = c - j of Jet's test helper system.
 *   - restrict / correct:             against oracle/_ref and closed forms
 *     (constants preserved, linear fields exact away from clamped edges).
 *   - single-phase / fractional system build and gradient: Jet's unit-test
 *     known answers (p(j+1)-p(j) = -1, p = 2-j, p = 1.5-j, velocities -> 0),
 *     the post-projection identity div(u') = b - A p, and symmetry of A.
 *     (Jet's pressure solvers cannot be linked here: they pull in
 *     TriangleMesh3, whose source needs the absent tinyobj submodule.)
 *   - fractionInside / fractionInsideSdf: closed-form areas of the square cut
 *     by a straight interface (every sign case), brute-force area counting.
 */
#ifndef JET_ORACLE_H_
#define JET_ORACLE_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_LEVELS 16

typedef struct {
    double center, right, up, front;
} or_row;

/* ---------------------------------------------------------------- FdmBlas3 */
void or_set(double s, double* v, size_t n);
void or_copy(const double* src, double* dst, size_t n);
double or_dot(const double* a, const double* b, size_t nx, size_t ny, size_t nz);
void or_axpy(double a, const double* x, const double* y, double* result,
             size_t nx, size_t ny, size_t nz);
void or_mvm(const or_row* A, const double* v, double* result, size_t nx,
            size_t ny, size_t nz);
void or_residual(const or_row* A, const double* x, const double* b,
                 double* result, size_t nx, size_t ny, size_t nz);
double or_l2norm(const double* v, size_t nx, size_t ny, size_t nz);
double or_linfnorm(const double* v, size_t nx, size_t ny, size_t nz);

/* ------------------------------------------------------------------ CG/PCG */
typedef void (*or_precond_fn)(void* ctx, const double* b, double* x);

void or_pcg(const or_row* A, const double* b, size_t nx, size_t ny, size_t nz,
            unsigned max_iter, double tol, or_precond_fn M, void* mctx,
            double* x, double* r, double* d, double* q, double* s,
            unsigned* last_iter, double* last_residual);

/* FdmCgSolver3::solve: zeroes x and scratch, runs cg, returns Jet's bool. */
int or_cg_solve(const or_row* A, const double* b, double* x, size_t nx,
                size_t ny, size_t nz, unsigned max_iter, double tol,
                unsigned* last_iter, double* last_residual);

/* -------------------------------------------------------------- Smoothers */
void or_gs_relax(const or_row* A, const double* b, double sor, double* x,
                 size_t nx, size_t ny, size_t nz);
void or_gs_relax_rb(const or_row* A, const double* b, double sor, double* x,
                    size_t nx, size_t ny, size_t nz);
void or_jacobi_relax(const or_row* A, const double* b, const double* x,
                     double* xtemp, size_t nx, size_t ny, size_t nz);

/* ------------------------------------------------------------- Multigrid */
typedef struct {
    size_t max_levels;
    unsigned restrict_iter, correct_iter, coarsest_iter, final_iter;
    double max_tol;
    double sor;
    int red_black;
} or_mg_params;

typedef struct {
    size_t nlev;
    size_t nx[OR_MAX_LEVELS], ny[OR_MAX_LEVELS], nz[OR_MAX_LEVELS];
    or_row* A[OR_MAX_LEVELS];
    double* x[OR_MAX_LEVELS];
    double* b[OR_MAX_LEVELS];
} or_mg_system;

size_t or_levels_with_finest(size_t nx, size_t ny, size_t nz, size_t max_levels,
                             size_t* lnx, size_t* lny, size_t* lnz);
void or_mg_restrict(const double* finer, double* coarser, size_t cnx,
                    size_t cny, size_t cnz);
void or_mg_correct(const double* coarser, double* finer, size_t fnx,
                   size_t fny, size_t fnz);
/* 2-D transfers and level sizing (FdmMgUtils2); a hierarchy with nz[0] == 1
 * and more than one level is a 2-D one and takes them in every V-cycle. */
size_t or_levels_with_finest2(size_t nx, size_t ny, size_t max_levels,
                              size_t* lnx, size_t* lny);
void or_mg_restrict2(const double* finer, double* coarser, size_t cnx, size_t cny);
void or_mg_correct2(const double* coarser, double* finer, size_t fnx, size_t fny);
int or_mg_is_2d(const or_mg_system* s);
/* Coarse operator of the flat-entry MGPCG from the next finer level's rows
 * (2x2x2 aggregation; 2x2 when fnz == 1); coarse is (fnx/2, fny/2, fnz/2 or 1). */
void or_mg_coarsen(const double* fine, size_t fnx, size_t fny, size_t fnz, double* coarse);
double or_mg_vcycle(const or_mg_system* sys, const or_mg_params* p,
                    double** x, double** b, double** buffer);
int or_mg_solve(or_mg_system* sys, const or_mg_params* p, double* last_norm);
int or_mgpcg_solve(or_mg_system* sys, const or_mg_params* p, unsigned max_iter,
                   unsigned* last_iter, double* last_residual);
/* Clean-mode FP64 reference solve (not a Jet code path): PCG with a symmetric
 * red-black V-cycle built from the restated operators; stops on
 * ||r||/||b|| <= rel_tol and reports the true relative residual. */
int or_mgpcg_clean_solve(or_mg_system* sys, const or_mg_params* p,
                         unsigned max_iter, double rel_tol,
                         unsigned* last_iter, double* true_rel_residual);
/* One application z = M r of that symmetric V-cycle (the preconditioner
 * alone): restrict_iter forward sweeps, correct_iter / final_iter backward
 * sweeps, coarsest_iter sweeps on the coarsest level split into forward and
 * backward halves. */
void or_mg_clean_vcycle(const or_mg_system* sys, const or_mg_params* p,
                        const double* r, double* z);
/* The same with weighted Jacobi smoothing instead of red-black
 * (x += omega ((b - s) / c - x), FdmJacobiSolver3::relax for omega = 1,
 * src/jet/fdm_jacobi_solver3.cpp:95-112; zero-centre rows set to 0):
 * restrict_iter sweeps before and correct_iter / final_iter sweeps after the
 * coarse correction, coarsest_iter sweeps on the coarsest level. */
void or_mg_clean_vcycle_jacobi(const or_mg_system* sys, const or_mg_params* p,
                               double omega, const double* r, double* z);
int or_mgpcg_clean_solve_jacobi(or_mg_system* sys, const or_mg_params* p, double omega,
                                unsigned max_iter, double rel_tol,
                                unsigned* last_iter, double* true_rel_residual);

/* ------------------------------------------------ Single-phase pressure */
enum { OR_FLUID = 0, OR_AIR = 1, OR_BOUNDARY = 2 };

void or_sp_markers(const double* boundary_sdf, const double* fluid_sdf,
                   char* markers, size_t n);
void or_sp_coarsen_markers(const char* finer, char* coarser, size_t cnx,
                           size_t cny, size_t cnz);
double or_divergence(const double* u, const double* v, const double* w,
                     size_t nx, size_t ny, size_t nz, const double h[3],
                     size_t i, size_t j, size_t k);
void or_sp_build_system(const char* markers, const double* u, const double* v,
                        const double* w, size_t nx, size_t ny, size_t nz,
                        const double h[3], or_row* A, double* b);
void or_face_fill_coarse(const double* fu, const double* fv, const double* fw,
                         size_t fnx, size_t fny, size_t fnz, const double fh[3],
                         const double origin[3], double* cu, double* cv,
                         double* cw, size_t cnx, size_t cny, size_t cnz,
                         const double ch[3]);
/* 2-D single phase: markers as or_sp_markers; the rest below. */
void or_sp_coarsen_markers2(const char* finer, char* coarser, size_t cnx, size_t cny);
void or_sp_build_system2(const char* markers, const double* u, const double* v, size_t nx,
                         size_t ny, const double h[2], or_row* A, double* b);
size_t or_sp_build_compressed2(const char* markers, const double* u, const double* v, size_t nx,
                               size_t ny, const double h[2], double* val, size_t* rp, size_t* ci,
                               double* b, size_t* nnz);
void or_sp_apply_gradient2(const char* markers, const double* p, const double* u,
                           const double* v, double* uo, double* vo, size_t nx, size_t ny,
                           const double h[2]);
void or_sp_apply_gradient(const char* markers, const double* p,
                          const double* u, const double* v, const double* w,
                          double* uo, double* vo, double* wo, size_t nx,
                          size_t ny, size_t nz, const double h[3]);

/* --------------------------------------------------- Fractional pressure */
double or_fraction_inside_sdf(double phi0, double phi1);
double or_fraction_inside(double bl, double br, double tl, double tr);
void or_fr_face_weights(const double* vertex_sdf, size_t nx, size_t ny,
                        size_t nz, float* uw, float* vw, float* ww);
void or_fr_restrict_float(const float* finer, size_t fnx, size_t fny,
                          size_t fnz, float* coarser, size_t cnx, size_t cny,
                          size_t cnz);
void or_fr_build_system(const float* fluid_sdf, const float* uw,
                        const float* vw, const float* ww, const double* u,
                        const double* v, const double* w, const double* bu,
                        const double* bv, const double* bw, size_t nx,
                        size_t ny, size_t nz, const double h[3], or_row* A,
                        double* b);
void or_fr_apply_gradient(const float* fluid_sdf, const float* uw,
                          const float* vw, const float* ww, const double* p,
                          const double* u, const double* v, const double* w,
                          double* uo, double* vo, double* wo, size_t nx,
                          size_t ny, size_t nz, const double h[3]);

/* 2-D (GridFractionalSinglePhasePressureSolver2): bsdf at the cell centres
 * of the grid extended by one ghost cell per side ((nx+2)*(ny+2)); uw is
 * (nx+1)*ny, vw nx*(ny+1); rows are {center, right, up, 0}. */
void or_fr_face_weights2(const double* bsdf, size_t nx, size_t ny, float* uw,
                         float* vw);
void or_fr_restrict_float2(const float* finer, size_t fnx, size_t fny,
                           float* coarser, size_t cnx, size_t cny);
void or_fr_build_system2(const float* fsdf, const float* uw, const float* vw,
                         const double* u, const double* v, const double* bu,
                         const double* bv, size_t nx, size_t ny, const double h[2],
                         or_row* A, double* b);
void or_fr_apply_gradient2(const float* fsdf, const float* uw, const float* vw,
                           const double* p, const double* u, const double* v,
                           double* uo, double* vo, size_t nx, size_t ny,
                           const double h[2]);

/* ---------------------------------------- boundary conditions (velocity) */
/* Rigid-body collider: RigidBodyCollider3 (include/jet/rigid_body_collider3.h)
 * with its surface's transform translation and Collider3's friction. */
typedef struct {
    double linear_velocity[3];
    double angular_velocity[3];
    double translation[3];
    double friction;
} or_rigid_collider;

/* Jet's kDirection* bits (include/jet/constants.h:312-335). */
enum { OR_DIR_LEFT = 1, OR_DIR_RIGHT = 2, OR_DIR_DOWN = 4, OR_DIR_UP = 8, OR_DIR_BACK = 16,
       OR_DIR_FRONT = 32 };

/* GridFractionalBoundaryConditionSolver3::constrainVelocity and, with
 * blocked != 0, GridBlockedBoundaryConditionSolver3::constrainVelocity, on a
 * face-centred grid of nx*ny*nz cells (u: (nx+1)*ny*nz, v: nx*(ny+1)*nz,
 * w: nx*ny*(nz+1), updated in place) with spacing h and origin o.  sdf: the
 * collider SDF at the cell centres (DBL_MAX everywhere without a collider);
 * collider NULL: zero collider velocity. */
void or_constrain_velocity(int blocked, size_t nx, size_t ny, size_t nz, const double h[3],
                           const double o[3], double* u, double* v, double* w,
                           const double* sdf, const or_rigid_collider* collider,
                           unsigned depth, int closed);
/* Building blocks, exposed for their own checks: LinearArraySampler3 of a
 * cell-centred grid, ScalarGrid3::gradient, extrapolateToRegion (3-D),
 * projectAndApplyFriction, RigidBodyCollider3::velocityAt. */
double or_sample_cell(const double* data, size_t nx, size_t ny, size_t nz, const double h[3],
                      const double o[3], const double p[3]);
void or_gradient_cell(const double* data, size_t nx, size_t ny, size_t nz, const double h[3],
                      const double o[3], const double p[3], double g[3]);
void or_extrapolate_to_region(const double* input, const char* valid, size_t sx, size_t sy,
                              size_t sz, unsigned iterations, double* output);
void or_project_and_apply_friction(const double vel[3], const double normal[3], double friction,
                                   double out[3]);
void or_collider_velocity(const or_rigid_collider* c, const double p[3], double out[3]);

/* ------------------------------------------- compressed (CSR) systems */
/* MatrixCsrD as Jet stores it: values and column indices (nnz), row
 * pointers (rows + 1), each row's entries in ascending column order. */
void or_csr_mvm(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                const double* v, double* result);
void or_csr_residual(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                     const double* x, const double* b, double* result);
/* FdmCgSolver3::solveCompressed: x zeroed, Jet's cg with FdmCompressedBlas3
 * (serial dot products); returns Jet's bool. */
int or_csr_cg_solve(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                    const double* b, double* x, unsigned max_iter, double tol,
                    unsigned* last_iter, double* last_residual);
/* Compressed buildSingleSystem of GridSinglePhasePressureSolver3: rows for
 * the fluid cells in x-fastest order.  Returns the row count; with NULL
 * outputs it only counts rows and, in *nnz, non-zeros. */
size_t or_sp_build_compressed(const char* markers, const double* u, const double* v,
                              const double* w, size_t nx, size_t ny, size_t nz, const double h[3],
                              double* val, size_t* rp, size_t* ci, double* b, size_t* nnz);

/* ------------------------------------------- backward Euler diffusion */
/* GridBackwardEulerDiffusionSolver3's buildMarkers / buildMatrix /
 * buildVectors for one data grid of sx*sy*sz points: f the field, bsdf and
 * fsdf the boundary / fluid SDFs at the data points, c = dt*mu/h^2 per axis,
 * dirichlet != 0 for Jet's Dirichlet boundary type.  A: Jet rows, b. */
void or_diffusion_build(const double* f, const double* bsdf, const double* fsdf, size_t sx,
                        size_t sy, size_t sz, const double c[3], int dirichlet, or_row* A,
                        double* b);

#ifdef __cplusplus
}
#endif

#endif /* JET_ORACLE_H_ */
