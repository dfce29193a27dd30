/*
 * jet_oracle.c -- CPU restatement of Jet's pressure-projection solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by the tests, the smoke
 * check and the CPU-baseline leg of the benchmark, never by the product.
 *
 * Written from the semantics of the cited reference files.  Arithmetic is
 * done in the same operand order as Jet so that, compiled with
 * -ffp-contract=off, results are bitwise equal to Jet's own build.
 * Loops Jet runs with parallelFor over k are OpenMP-parallel over k here;
 * loops Jet runs serially (dot, set, lexicographic Gauss-Seidel) are serial.
 */
#include "jet_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, k, nx, ny) ((i) + (nx) * ((j) + (ny) * (k)))

static const double kEps = DBL_EPSILON; /* include/jet/constants.h:74 */

/* ======================================================= FdmBlas3 ====== */

/* Array3::set(value) -- include/jet/detail/array3-inl.h:49-54 (serial). */
void or_set(double s, double* v, size_t n) {
    for (size_t i = 0; i < n; ++i) v[i] = s;
}

/* Array3::set(other) -- include/jet/detail/array3-inl.h:56-61 (serial). */
void or_copy(const double* src, double* dst, size_t n) {
    if (src != dst) memmove(dst, src, n * sizeof(double));
}

/* FdmBlas3::dot -- src/jet/fdm_linear_system3.cpp:49-65.  Serial k,j,i. */
double or_dot(const double* a, const double* b, size_t nx, size_t ny,
              size_t nz) {
    double result = 0.0;
    const size_t n = nx * ny * nz;
    for (size_t idx = 0; idx < n; ++idx) result += a[idx] * b[idx];
    return result;
}

/* FdmBlas3::axpy -- src/jet/fdm_linear_system3.cpp:67-77. */
void or_axpy(double a, const double* x, const double* y, double* result,
             size_t nx, size_t ny, size_t nz) {
    const size_t plane = nx * ny;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t p = k * plane; p < (k + 1) * plane; ++p)
            result[p] = a * x[p] + y[p];
}

/* FdmBlas3::mvm -- src/jet/fdm_linear_system3.cpp:79-96.  The -x/-y/-z
 * couplings are the neighbour's right/up/front (upper-triangle storage). */
void or_mvm(const or_row* m, const double* v, double* result, size_t nx,
            size_t ny, size_t nz) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                result[c] =
                    m[c].center * v[c] +
                    ((i > 0) ? m[c - 1].right * v[c - 1] : 0.0) +
                    ((i + 1 < nx) ? m[c].right * v[c + 1] : 0.0) +
                    ((j > 0) ? m[c - nx].up * v[c - nx] : 0.0) +
                    ((j + 1 < ny) ? m[c].up * v[c + nx] : 0.0) +
                    ((k > 0) ? m[c - nx * ny].front * v[c - nx * ny] : 0.0) +
                    ((k + 1 < nz) ? m[c].front * v[c + nx * ny] : 0.0);
            }
}

/* FdmBlas3::residual -- src/jet/fdm_linear_system3.cpp:98-116. */
void or_residual(const or_row* a, const double* x, const double* b,
                 double* result, size_t nx, size_t ny, size_t nz) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                result[c] =
                    b[c] - a[c].center * x[c] -
                    ((i > 0) ? a[c - 1].right * x[c - 1] : 0.0) -
                    ((i + 1 < nx) ? a[c].right * x[c + 1] : 0.0) -
                    ((j > 0) ? a[c - nx].up * x[c - nx] : 0.0) -
                    ((j + 1 < ny) ? a[c].up * x[c + nx] : 0.0) -
                    ((k > 0) ? a[c - nx * ny].front * x[c - nx * ny] : 0.0) -
                    ((k + 1 < nz) ? a[c].front * x[c + nx * ny] : 0.0);
            }
}

/* FdmBlas3::l2Norm -- src/jet/fdm_linear_system3.cpp:118. */
double or_l2norm(const double* v, size_t nx, size_t ny, size_t nz) {
    return sqrt(or_dot(v, v, nx, ny, nz));
}

/* FdmBlas3::lInfNorm -- src/jet/fdm_linear_system3.cpp:120-134 (absmax,
 * include/jet/detail/math_utils-inl.h:65-67). */
double or_linfnorm(const double* v, size_t nx, size_t ny, size_t nz) {
    double result = 0.0;
    const size_t n = nx * ny * nz;
    for (size_t idx = 0; idx < n; ++idx)
        result = (result * result > v[idx] * v[idx]) ? result : v[idx];
    return fabs(result);
}

/* ======================================================== CG / PCG ===== */

/* pcg<BlasType, PrecondType> -- include/jet/detail/cg-inl.h:15-94.
 * Stops when sigma = r.z <= tol^2; recomputes the true residual every 50
 * iterations and on the iteration after sigma grew. */
void or_pcg(const or_row* A, const double* b, size_t nx, size_t ny, size_t nz,
            unsigned max_iter, double tol, or_precond_fn M, void* mctx,
            double* x, double* r, double* d, double* q, double* s,
            unsigned* last_iter, double* last_residual) {
    const size_t n = nx * ny * nz;
    or_set(0.0, r, n);
    or_set(0.0, d, n);
    or_set(0.0, q, n);
    or_set(0.0, s, n);

    or_residual(A, x, b, r, nx, ny, nz);
    M(mctx, r, d);
    double sigma_new = or_dot(r, d, nx, ny, nz);

    unsigned iter = 0;
    int trigger = 0;
    while (sigma_new > tol * tol && iter < max_iter) {
        or_mvm(A, d, q, nx, ny, nz);
        double alpha = sigma_new / or_dot(d, q, nx, ny, nz);
        or_axpy(alpha, d, x, x, nx, ny, nz);
        if (trigger || (iter % 50 == 0 && iter > 0)) {
            or_residual(A, x, b, r, nx, ny, nz);
            trigger = 0;
        } else {
            or_axpy(-alpha, q, r, r, nx, ny, nz);
        }
        M(mctx, r, s);
        double sigma_old = sigma_new;
        sigma_new = or_dot(r, s, nx, ny, nz);
        if (sigma_new > sigma_old) trigger = 1;
        double beta = sigma_new / sigma_old;
        or_axpy(beta, d, s, d, nx, ny, nz);
        ++iter;
    }
    *last_iter = iter;
    *last_residual = sqrt(fabs(sigma_new));
}

typedef struct {
    size_t n;
} null_precond_ctx;

/* NullCgPreconditioner::solve -- include/jet/cg.h:20-29 (copy). */
static void null_precond(void* ctx, const double* b, double* x) {
    or_copy(b, x, ((null_precond_ctx*)ctx)->n);
}

/* FdmCgSolver3::solve -- src/jet/fdm_cg_solver3.cpp:20-47. */
int or_cg_solve(const or_row* A, const double* b, double* x, size_t nx,
                size_t ny, size_t nz, unsigned max_iter, double tol,
                unsigned* last_iter, double* last_residual) {
    const size_t n = nx * ny * nz;
    double* buf = (double*)calloc(4 * (n ? n : 1), sizeof(double));
    if (!buf) return -1;
    or_set(0.0, x, n);
    null_precond_ctx ctx = {n};
    or_pcg(A, b, nx, ny, nz, max_iter, tol, null_precond, &ctx, x, buf,
           buf + n, buf + 2 * n, buf + 3 * n, last_iter, last_residual);
    free(buf);
    return (*last_residual <= tol) || (*last_iter < max_iter);
}

/* ====================================================== Smoothers ====== */

#define OR_OFFDIAG_SUM(A, x, c, i, j, k, nx, ny, nz)                          \
    (((i > 0) ? A[c - 1].right * x[c - 1] : 0.0) +                          \
     ((i + 1 < nx) ? A[c].right * x[c + 1] : 0.0) +                         \
     ((j > 0) ? A[c - nx].up * x[c - nx] : 0.0) +                           \
     ((j + 1 < ny) ? A[c].up * x[c + nx] : 0.0) +                           \
     ((k > 0) ? A[c - nx * ny].front * x[c - nx * ny] : 0.0) +              \
     ((k + 1 < nz) ? A[c].front * x[c + nx * ny] : 0.0))

/* FdmGaussSeidelSolver3::relax -- src/jet/fdm_gauss_seidel_solver3.cpp:103-120.
 * Lexicographic (k,j,i), in place, SOR factor sor.  Serial. */
void or_gs_relax(const or_row* A, const double* b, double sor, double* x,
                 size_t nx, size_t ny, size_t nz) {
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                double r = OR_OFFDIAG_SUM(A, x, c, i, j, k, nx, ny, nz);
                x[c] = (1.0 - sor) * x[c] + sor * (b[c] - r) / A[c].center;
            }
}

/* zero_rows: a row with a zero centre is set to 0 instead of divided by
 * (the clean-mode smoother below; Jet's relaxRedBlack divides). */
static void rb_pass_rows(const or_row* A, const double* b, double sor, double* x,
                         size_t nx, size_t ny, size_t nz, size_t color, int zero_rows) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = (j + k + color) % 2; i < nx; i += 2) {
                const size_t c = IX(i, j, k, nx, ny);
                double r = OR_OFFDIAG_SUM(A, x, c, i, j, k, nx, ny, nz);
                if (zero_rows && A[c].center == 0.0)
                    x[c] = 0.0;
                else
                    x[c] = (1.0 - sor) * x[c] + sor * (b[c] - r) / A[c].center;
            }
}

static void rb_pass(const or_row* A, const double* b, double sor, double* x,
                    size_t nx, size_t ny, size_t nz, size_t color) {
    rb_pass_rows(A, b, sor, x, nx, ny, nz, color, 0);
}

/* FdmGaussSeidelSolver3::relaxRedBlack --
 * src/jet/fdm_gauss_seidel_solver3.cpp:150-221.  Red = (i+j+k) even first
 * (:163), then black = (i+j+k) odd (:196).  Each colour only reads the other
 * colour, so the visiting order within a colour does not change results. */
void or_gs_relax_rb(const or_row* A, const double* b, double sor, double* x,
                    size_t nx, size_t ny, size_t nz) {
    rb_pass(A, b, sor, x, nx, ny, nz, 0);
    rb_pass(A, b, sor, x, nx, ny, nz, 1);
}

/* FdmJacobiSolver3::relax -- src/jet/fdm_jacobi_solver3.cpp:95-112. */
void or_jacobi_relax(const or_row* A, const double* b, const double* x,
                     double* xtemp, size_t nx, size_t ny, size_t nz) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                double r = OR_OFFDIAG_SUM(A, x, c, i, j, k, nx, ny, nz);
                xtemp[c] = (b[c] - r) / A[c].center;
            }
}

/* ====================================================== Multigrid ====== */

/* FdmMgUtils3::resizeArrayWithFinest / WithCoarsest --
 * include/jet/detail/fdm_mg_linear_system3-inl.h:14-50.  Halve while all of
 * nx, ny, nz are even, up to max_levels levels (at least one). */
size_t or_levels_with_finest(size_t nx, size_t ny, size_t nz, size_t max_levels,
                             size_t* lnx, size_t* lny, size_t* lnz) {
    size_t rx = nx, ry = ny, rz = nz;
    size_t i = 1;
    for (; i < max_levels; ++i) {
        if (rx % 2 == 0 && ry % 2 == 0 && rz % 2 == 0) {
            rx >>= 1;
            ry >>= 1;
            rz >>= 1;
        } else {
            break;
        }
    }
    size_t nlev = i < 1 ? 1 : i;
    if (nlev > OR_MAX_LEVELS) nlev = OR_MAX_LEVELS;
    for (size_t l = 0; l < nlev; ++l) {
        lnx[nlev - 1 - l] = rx << l;
        lny[nlev - 1 - l] = ry << l;
        lnz[nlev - 1 - l] = rz << l;
    }
    return nlev;
}

/* FdmMgUtils3::restrict -- src/jet/fdm_mg_linear_system3.cpp:43-98.
 * 1-D weights {1/8,3/8,3/8,1/8} over fine {2i-1,2i,2i+1,2i+2}, clamped to 2i
 * at the low end and to 2i+1 at the high end.  No 4x scaling. */
void or_mg_restrict(const double* finer, double* coarser, size_t cnx,
                    size_t cny, size_t cnz) {
    static const double kernel[4] = {0.125, 0.375, 0.375, 0.125};
    const size_t fnx = 2 * cnx, fny = 2 * cny;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < cnz; ++k) {
        size_t kI[4] = {(k > 0) ? 2 * k - 1 : 2 * k, 2 * k, 2 * k + 1,
                        (k + 1 < cnz) ? 2 * k + 2 : 2 * k + 1};
        for (size_t j = 0; j < cny; ++j) {
            size_t jI[4] = {(j > 0) ? 2 * j - 1 : 2 * j, 2 * j, 2 * j + 1,
                            (j + 1 < cny) ? 2 * j + 2 : 2 * j + 1};
            for (size_t i = 0; i < cnx; ++i) {
                size_t iI[4] = {(i > 0) ? 2 * i - 1 : 2 * i, 2 * i, 2 * i + 1,
                                (i + 1 < cnx) ? 2 * i + 2 : 2 * i + 1};
                double sum = 0.0;
                for (int z = 0; z < 4; ++z)
                    for (int y = 0; y < 4; ++y)
                        for (int x = 0; x < 4; ++x) {
                            double w = kernel[x] * kernel[y] * kernel[z];
                            sum += w * finer[IX(iI[x], jI[y], kI[z], fnx, fny)];
                        }
                coarser[IX(i, j, k, cnx, cny)] = sum;
            }
        }
    }
}

/* FdmMgUtils3::correct -- src/jet/fdm_mg_linear_system3.cpp:100-179.
 * Trilinear (1/4, 3/4) prolongation added to the finer grid.  Jet bounds the
 * k upper clamp with n.y (:159); restated (no effect on cubes). */
void or_mg_correct(const double* coarser, double* finer, size_t fnx,
                   size_t fny, size_t fnz) {
    const size_t cnx = fnx / 2, cny = fny / 2;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < fnz; ++k)
        for (size_t j = 0; j < fny; ++j)
            for (size_t i = 0; i < fnx; ++i) {
                size_t iI[2], jI[2], kI[2];
                double iW[2], jW[2], kW[2];
                const size_t ci = i / 2, cj = j / 2, ck = k / 2;
                if (i % 2 == 0) {
                    iI[0] = (i > 1) ? ci - 1 : ci;
                    iI[1] = ci;
                    iW[0] = 0.25;
                    iW[1] = 0.75;
                } else {
                    iI[0] = ci;
                    iI[1] = (i + 1 < fnx) ? ci + 1 : ci;
                    iW[0] = 0.75;
                    iW[1] = 0.25;
                }
                if (j % 2 == 0) {
                    jI[0] = (j > 1) ? cj - 1 : cj;
                    jI[1] = cj;
                    jW[0] = 0.25;
                    jW[1] = 0.75;
                } else {
                    jI[0] = cj;
                    jI[1] = (j + 1 < fny) ? cj + 1 : cj;
                    jW[0] = 0.75;
                    jW[1] = 0.25;
                }
                if (k % 2 == 0) {
                    kI[0] = (k > 1) ? ck - 1 : ck;
                    kI[1] = ck;
                    kW[0] = 0.25;
                    kW[1] = 0.75;
                } else {
                    kI[0] = ck;
                    /* Jet tests k + 1 < n.y.  Where n.y > n.z that reads past
                     * the coarse array (undefined behaviour, nothing to
                     * restate), so the bound is min(n.y, n.z): identical to
                     * Jet wherever Jet stays in bounds. */
                    kI[1] = (k + 1 < (fny < fnz ? fny : fnz)) ? ck + 1 : ck;
                    kW[0] = 0.75;
                    kW[1] = 0.25;
                }
                const size_t f = IX(i, j, k, fnx, fny);
                for (int z = 0; z < 2; ++z)
                    for (int y = 0; y < 2; ++y)
                        for (int x = 0; x < 2; ++x) {
                            double w = iW[x] * jW[y] * kW[z] *
                                       coarser[IX(iI[x], jI[y], kI[z], cnx, cny)];
                            finer[f] += w;
                        }
            }
}

/* FdmMgUtils2::resizeArrayWithFinest -- include/jet/detail/
 * fdm_mg_linear_system2-inl.h:34-48: halve while nx and ny are even. */
size_t or_levels_with_finest2(size_t nx, size_t ny, size_t max_levels,
                              size_t* lnx, size_t* lny) {
    size_t rx = nx, ry = ny;
    size_t i = 1;
    for (; i < max_levels; ++i) {
        if (rx % 2 == 0 && ry % 2 == 0) {
            rx >>= 1;
            ry >>= 1;
        } else {
            break;
        }
    }
    size_t nlev = i > OR_MAX_LEVELS ? OR_MAX_LEVELS : i;
    for (size_t l = 0; l < nlev; ++l) {
        lnx[nlev - 1 - l] = rx << l;
        lny[nlev - 1 - l] = ry << l;
    }
    return nlev;
}

/* FdmMgUtils2::restrict -- src/jet/fdm_mg_linear_system2.cpp:44-84: the
 * same 1-D weights and clamps as in 3-D; the 16 taps are summed y-major,
 * each weight kernel[x] * kernel[y] formed first. */
void or_mg_restrict2(const double* finer, double* coarser, size_t cnx, size_t cny) {
    static const double kernel[4] = {0.125, 0.375, 0.375, 0.125};
    const size_t fnx = 2 * cnx;
    for (size_t j = 0; j < cny; ++j) {
        const size_t jI[4] = {(j > 0) ? 2 * j - 1 : 2 * j, 2 * j, 2 * j + 1,
                              (j + 1 < cny) ? 2 * j + 2 : 2 * j + 1};
        for (size_t i = 0; i < cnx; ++i) {
            const size_t iI[4] = {(i > 0) ? 2 * i - 1 : 2 * i, 2 * i, 2 * i + 1,
                                  (i + 1 < cnx) ? 2 * i + 2 : 2 * i + 1};
            double sum = 0.0;
            for (int y = 0; y < 4; ++y)
                for (int x = 0; x < 4; ++x) {
                    const double w = kernel[x] * kernel[y];
                    sum += w * finer[iI[x] + fnx * jI[y]];
                }
            coarser[i + cnx * j] = sum;
        }
    }
}

/* FdmMgUtils2::correct -- src/jet/fdm_mg_linear_system2.cpp:86-143: bilinear
 * (1/4, 3/4) prolongation added to the finer grid, four separate adds of
 * (wi * wj) * coarse, y-major. */
void or_mg_correct2(const double* coarser, double* finer, size_t fnx, size_t fny) {
    const size_t cnx = fnx / 2;
    for (size_t j = 0; j < fny; ++j)
        for (size_t i = 0; i < fnx; ++i) {
            size_t iI[2], jI[2];
            double iW[2], jW[2];
            const size_t ci = i / 2, cj = j / 2;
            if (i % 2 == 0) {
                iI[0] = (i > 1) ? ci - 1 : ci;
                iI[1] = ci;
                iW[0] = 0.25;
                iW[1] = 0.75;
            } else {
                iI[0] = ci;
                iI[1] = (i + 1 < fnx) ? ci + 1 : ci;
                iW[0] = 0.75;
                iW[1] = 0.25;
            }
            if (j % 2 == 0) {
                jI[0] = (j > 1) ? cj - 1 : cj;
                jI[1] = cj;
                jW[0] = 0.25;
                jW[1] = 0.75;
            } else {
                jI[0] = cj;
                jI[1] = (j + 1 < fny) ? cj + 1 : cj;
                jW[0] = 0.75;
                jW[1] = 0.25;
            }
            for (int y = 0; y < 2; ++y)
                for (int x = 0; x < 2; ++x) {
                    const double w = iW[x] * jW[y] * coarser[iI[x] + cnx * jI[y]];
                    finer[i + fnx * j] += w;
                }
        }
}

/* A hierarchy whose finest level is one plane thick and which has coarser
 * levels is a 2-D one (FdmMgLinearSystem2, levels halving in x and y): a 3-D
 * hierarchy cannot halve an odd nz.  Its rows are 2-D rows with front = 0,
 * so the 3-D relax, residual and products give the 2-D values (the absent z
 * terms add +0); only the transfers differ. */
/* Coarse operator of the flat-entry MGPCG.  No counterpart in Jet: its
 * FdmMgSolver3::solve(FdmLinearSystem3*) is final and does nothing
 * (include/jet/fdm_mg_solver3.h:42), and its multigrid callers build every
 * level from markers.  A coarse cell C aggregates the 2x2x2 fine cells it
 * covers (2x2 on one-plane levels, fnz == 1).  Split the fine matrix as
 * A = M + L, M = diag(row sums), L the coupling part with zero row sums; then
 *   A_c = (1/8) P'MP + (1/16) P'LP   (one plane: (1/4) P'MP + (1/8) P'LP)
 * with P the piecewise-constant injection.  A uniform stencil (mass m,
 * couplings -c) coarsens to mass m and couplings -c/4: the operator
 * rediscretised at twice the spacing, the scale that FdmMgUtils3's restriction
 * (a weighted average) and correction (an interpolation) expect.  In rows,
 * with kappa = 1/16 (1/8):
 *   right_C  = kappa * (the fine +x couplings leaving C), 0 on the last coarse
 *              column; up_C and front_C likewise;
 *   center_C = kappa * ((2 * sum of centers + 4 * sum of couplings inside C)
 *              + sum of couplings leaving C through its faces),
 * a coupling to a cell outside the grid ignored (as mvm ignores it).  Sums
 * run in the fixed order below; the GPU kernel uses the same one. */
void or_mg_coarsen(const double* fine, size_t fnx, size_t fny, size_t fnz, double* coarse) {
    const int plane = fnz == 1;
    const size_t cnx = fnx / 2, cny = fny / 2, cnz = plane ? 1 : fnz / 2;
    const size_t dz = plane ? 1 : 2;
    const double kappa = plane ? 0.125 : 0.0625;
#define FR(i, j, k, f) fine[4 * IX(i, j, k, fnx, fny) + (f)]
#pragma omp parallel for schedule(static)
    for (size_t K = 0; K < cnz; ++K)
        for (size_t J = 0; J < cny; ++J)
            for (size_t I = 0; I < cnx; ++I) {
                const size_t i0 = 2 * I, j0 = 2 * J, k0 = dz * K;
                double sc = 0.0, si = 0.0;
                double px = 0.0, mx = 0.0, py = 0.0, my = 0.0, pz = 0.0, mz = 0.0;
                for (size_t c = 0; c < dz; ++c)
                    for (size_t b = 0; b < 2; ++b)
                        for (size_t a = 0; a < 2; ++a) sc += FR(i0 + a, j0 + b, k0 + c, 0);
                for (size_t c = 0; c < dz; ++c)
                    for (size_t b = 0; b < 2; ++b) si += FR(i0, j0 + b, k0 + c, 1);
                for (size_t c = 0; c < dz; ++c)
                    for (size_t a = 0; a < 2; ++a) si += FR(i0 + a, j0, k0 + c, 2);
                if (!plane)
                    for (size_t b = 0; b < 2; ++b)
                        for (size_t a = 0; a < 2; ++a) si += FR(i0 + a, j0 + b, k0, 3);
                for (size_t c = 0; c < dz; ++c)
                    for (size_t b = 0; b < 2; ++b) {
                        if (I + 1 < cnx) px += FR(i0 + 1, j0 + b, k0 + c, 1);
                        if (I > 0) mx += FR(i0 - 1, j0 + b, k0 + c, 1);
                    }
                for (size_t c = 0; c < dz; ++c)
                    for (size_t a = 0; a < 2; ++a) {
                        if (J + 1 < cny) py += FR(i0 + a, j0 + 1, k0 + c, 2);
                        if (J > 0) my += FR(i0 + a, j0 - 1, k0 + c, 2);
                    }
                if (!plane)
                    for (size_t b = 0; b < 2; ++b)
                        for (size_t a = 0; a < 2; ++a) {
                            if (K + 1 < cnz) pz += FR(i0 + a, j0 + b, k0 + 1, 3);
                            if (K > 0) mz += FR(i0 + a, j0 + b, k0 - 1, 3);
                        }
                const double se = ((((px + mx) + py) + my) + pz) + mz;
                double* row = coarse + 4 * IX(I, J, K, cnx, cny);
                row[0] = kappa * ((2.0 * sc + 4.0 * si) + se);
                row[1] = kappa * px;
                row[2] = kappa * py;
                row[3] = kappa * pz;
            }
#undef FR
}


int or_mg_is_2d(const or_mg_system* s) { return s->nlev > 1 && s->nz[0] == 1; }

static void mg_restrict_level(const or_mg_system* s, size_t l, const double* fine,
                              double* coarse) {
    if (or_mg_is_2d(s))
        or_mg_restrict2(fine, coarse, s->nx[l + 1], s->ny[l + 1]);
    else
        or_mg_restrict(fine, coarse, s->nx[l + 1], s->ny[l + 1], s->nz[l + 1]);
}

static void mg_correct_level(const or_mg_system* s, size_t l, const double* coarse,
                             double* fine) {
    if (or_mg_is_2d(s))
        or_mg_correct2(coarse, fine, s->nx[l], s->ny[l]);
    else
        or_mg_correct(coarse, fine, s->nx[l], s->ny[l], s->nz[l]);
}

static void mg_relax(const or_row* A, const double* b, unsigned iters,
                     double* x, size_t nx, size_t ny, size_t nz,
                     const or_mg_params* p) {
    /* FdmMgSolver3 relax closures -- src/jet/fdm_mg_solver3.cpp:26-50. */
    for (unsigned it = 0; it < iters; ++it) {
        if (p->red_black)
            or_gs_relax_rb(A, b, p->sor, x, nx, ny, nz);
        else
            or_gs_relax(A, b, p->sor, x, nx, ny, nz);
    }
}

/* internal::mgVCycle -- include/jet/detail/mg-inl.h:16-68. */
static double vcycle_level(const or_mg_system* s, const or_mg_params* p,
                           size_t l, double** x, double** b, double** buf) {
    const size_t nx = s->nx[l], ny = s->ny[l], nz = s->nz[l];
    mg_relax(s->A[l], b[l], p->restrict_iter, x[l], nx, ny, nz, p);
    if (l < s->nlev - 1) {
        or_residual(s->A[l], x[l], b[l], buf[l], nx, ny, nz);
        mg_restrict_level(s, l, buf[l], b[l + 1]);
        or_set(0.0, x[l + 1], s->nx[l + 1] * s->ny[l + 1] * s->nz[l + 1]);
        vcycle_level(s, p, l + 1, x, b, buf);
        mg_correct_level(s, l, x[l + 1], x[l]);
        mg_relax(s->A[l], b[l], l > 0 ? p->correct_iter : p->final_iter, x[l],
                 nx, ny, nz, p);
    } else {
        mg_relax(s->A[l], b[l], p->coarsest_iter, x[l], nx, ny, nz, p);
        or_residual(s->A[l], x[l], b[l], buf[l], nx, ny, nz);
    }
    or_residual(s->A[l], x[l], b[l], buf[l], nx, ny, nz);
    return or_l2norm(buf[l], nx, ny, nz);
}

double or_mg_vcycle(const or_mg_system* sys, const or_mg_params* p,
                    double** x, double** b, double** buffer) {
    return vcycle_level(sys, p, 0, x, b, buffer);
}

static size_t level_n(const or_mg_system* s, size_t l) {
    return s->nx[l] * s->ny[l] * s->nz[l];
}

static double** alloc_levels(const or_mg_system* s) {
    double** v = (double**)calloc(s->nlev, sizeof(double*));
    for (size_t l = 0; l < s->nlev; ++l)
        v[l] = (double*)calloc(level_n(s, l) ? level_n(s, l) : 1,
                               sizeof(double));
    return v;
}

static void free_levels(const or_mg_system* s, double** v) {
    for (size_t l = 0; l < s->nlev; ++l) free(v[l]);
    free(v);
}

/* FdmMgSolver3::solve(FdmMgLinearSystem3*) -- src/jet/fdm_mg_solver3.cpp:69-74:
 * buffer = copy of x, one V-cycle, converged if norm < maxTolerance. */
int or_mg_solve(or_mg_system* sys, const or_mg_params* p, double* last_norm) {
    double** buf = alloc_levels(sys);
    for (size_t l = 0; l < sys->nlev; ++l)
        or_copy(sys->x[l], buf[l], level_n(sys, l));
    double norm = or_mg_vcycle(sys, p, sys->x, sys->b, buf);
    free_levels(sys, buf);
    if (last_norm) *last_norm = norm;
    return norm < p->max_tol;
}

typedef struct {
    or_mg_system* sys;
    const or_mg_params* p;
    double** mx;
    double** mb;
    double** mbuf;
} mgpcg_precond_ctx;

/* FdmMgpcgSolver3::Preconditioner::solve -- src/jet/fdm_mgpcg_solver3.cpp:21-36.
 * Jet deep-copies the x hierarchy three times; the coarse copies are always
 * overwritten before use (restriction writes b[l+1], x[l+1] is zeroed, the
 * residual writes the buffer), so only the top-level warm start -- the output
 * vector's current contents -- carries state.  That warm start is kept. */
static void mgpcg_precond(void* vctx, const double* b, double* x) {
    mgpcg_precond_ctx* c = (mgpcg_precond_ctx*)vctx;
    or_mg_system* s = c->sys;
    for (size_t l = 1; l < s->nlev; ++l) {
        or_copy(s->x[l], c->mx[l], level_n(s, l));
        or_copy(s->x[l], c->mb[l], level_n(s, l));
        or_copy(s->x[l], c->mbuf[l], level_n(s, l));
    }
    or_copy(s->x[0], c->mbuf[0], level_n(s, 0));
    or_copy(x, c->mx[0], level_n(s, 0));
    or_copy(b, c->mb[0], level_n(s, 0));
    or_mg_vcycle(s, c->p, c->mx, c->mb, c->mbuf);
    or_copy(c->mx[0], x, level_n(s, 0));
}

/* FdmMgpcgSolver3::solve -- src/jet/fdm_mgpcg_solver3.cpp:54-80. */
int or_mgpcg_solve(or_mg_system* sys, const or_mg_params* p, unsigned max_iter,
                   unsigned* last_iter, double* last_residual) {
    const size_t nx = sys->nx[0], ny = sys->ny[0], nz = sys->nz[0];
    const size_t n = nx * ny * nz;
    double* scratch = (double*)calloc(4 * (n ? n : 1), sizeof(double));
    or_set(0.0, sys->x[0], n);
    mgpcg_precond_ctx c = {sys, p, alloc_levels(sys), alloc_levels(sys),
                           alloc_levels(sys)};
    or_pcg(sys->A[0], sys->b[0], nx, ny, nz, max_iter, p->max_tol,
           mgpcg_precond, &c, sys->x[0], scratch, scratch + n, scratch + 2 * n,
           scratch + 3 * n, last_iter, last_residual);
    free_levels(sys, c.mx);
    free_levels(sys, c.mb);
    free_levels(sys, c.mbuf);
    free(scratch);
    return (*last_residual <= p->max_tol) || (*last_iter < max_iter);
}

/* ============================================ Clean-mode reference solve ==
 * Not a Jet code path: a tightly converging FP64 reference solve of the same
 * (A, b), used to produce the reference pressure p_ref that GPU results are
 * compared with.  It is composed only of the restated Jet operators above
 * (relaxRedBlack, restrict, correct, residual, dot) arranged as a symmetric
 * V-cycle (zero initial guess, red-black forward pre-smoothing, black-red
 * backward post-smoothing, symmetric coarsest sweeps) inside plain PCG.  Jet's
 * restriction equals 1/8 of the transpose of its prolongation, clamps
 * included, so the preconditioner is symmetric.  A zero-centre row (a fluid
 * cell with only solid neighbours: an all-zero row with b = 0 in Jet's
 * single-phase build) is set to 0 by the smoother instead of divided by, so
 * the preconditioner stays finite and symmetric.  Its answer is pinned by the
 * true residual ||b - A x|| (the bitwise-pinned residual above), which the
 * caller checks. */

static void rb_sweep_ordered(const or_row* A, const double* b, double sor,
                             double* x, size_t nx, size_t ny, size_t nz,
                             int backward) {
    rb_pass_rows(A, b, sor, x, nx, ny, nz, backward ? 1 : 0, 1);
    rb_pass_rows(A, b, sor, x, nx, ny, nz, backward ? 0 : 1, 1);
}

static void clean_vcycle(const or_mg_system* s, const or_mg_params* p,
                         size_t l, double** x, double** b, double** buf) {
    const size_t nx = s->nx[l], ny = s->ny[l], nz = s->nz[l];
    const size_t n = nx * ny * nz;
    or_set(0.0, x[l], n);
    if (l + 1 < s->nlev) {
        for (unsigned it = 0; it < p->restrict_iter; ++it)
            rb_sweep_ordered(s->A[l], b[l], p->sor, x[l], nx, ny, nz, 0);
        or_residual(s->A[l], x[l], b[l], buf[l], nx, ny, nz);
        mg_restrict_level(s, l, buf[l], b[l + 1]);
        clean_vcycle(s, p, l + 1, x, b, buf);
        mg_correct_level(s, l, x[l + 1], x[l]);
        unsigned post = l > 0 ? p->correct_iter : p->final_iter;
        for (unsigned it = 0; it < post; ++it)
            rb_sweep_ordered(s->A[l], b[l], p->sor, x[l], nx, ny, nz, 1);
    } else {
        /* coarsest_iter sweeps in total (Jet's count), split into forward
         * then backward halves to keep the cycle symmetric */
        const unsigned half = p->coarsest_iter > 1 ? (p->coarsest_iter + 1) / 2 : 1;
        for (unsigned it = 0; it < half; ++it)
            rb_sweep_ordered(s->A[l], b[l], p->sor, x[l], nx, ny, nz, 0);
        for (unsigned it = 0; it < half; ++it)
            rb_sweep_ordered(s->A[l], b[l], p->sor, x[l], nx, ny, nz, 1);
    }
}

/* Weighted Jacobi sweep of the clean V-cycle: xo = x + omega ((b - s) / c - x)
 * with s summed in FdmJacobiSolver3::relax's order, and for omega = 1 that
 * relax itself, (b - s) / c (src/jet/fdm_jacobi_solver3.cpp:95-112).  A
 * zero-centre row is set to 0, as in the clean red-black sweep. */
static void jacobi_sweep_clean(const or_row* A, const double* b, double omega,
                               const double* x, double* xo, size_t nx,
                               size_t ny, size_t nz) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                if (A[c].center == 0.0) {
                    xo[c] = 0.0;
                    continue;
                }
                const double s = OR_OFFDIAG_SUM(A, x, c, i, j, k, nx, ny, nz);
                xo[c] = omega == 1.0 ? (b[c] - s) / A[c].center
                                     : x[c] + omega * ((b[c] - s) / A[c].center - x[c]);
            }
}

/* `count` Jacobi sweeps of x, through tmp; the result ends in x. */
static void jacobi_sweeps(const or_row* A, const double* b, double omega, double* x,
                          double* tmp, size_t nx, size_t ny, size_t nz,
                          unsigned count) {
    for (unsigned it = 0; it < count; ++it) {
        jacobi_sweep_clean(A, b, omega, x, tmp, nx, ny, nz);
        or_copy(tmp, x, nx * ny * nz);
    }
}

/* The clean V-cycle with weighted Jacobi smoothing: zero initial guess,
 * restrict_iter sweeps, residual and Jet's restriction, recursion, Jet's
 * correction, correct_iter (final_iter on level 0) sweeps; coarsest_iter
 * sweeps on the coarsest level.  With as many sweeps after as before the
 * coarse correction the cycle is symmetric, and positive definite for
 * 0 < omega <= 1 on diagonally dominant rows. */
static void clean_vcycle_jacobi(const or_mg_system* s, const or_mg_params* p,
                                double omega, size_t l, double** x, double** b,
                                double** buf) {
    const size_t nx = s->nx[l], ny = s->ny[l], nz = s->nz[l];
    or_set(0.0, x[l], nx * ny * nz);
    if (l + 1 < s->nlev) {
        jacobi_sweeps(s->A[l], b[l], omega, x[l], buf[l], nx, ny, nz, p->restrict_iter);
        or_residual(s->A[l], x[l], b[l], buf[l], nx, ny, nz);
        mg_restrict_level(s, l, buf[l], b[l + 1]);
        clean_vcycle_jacobi(s, p, omega, l + 1, x, b, buf);
        mg_correct_level(s, l, x[l + 1], x[l]);
        const unsigned post = l > 0 ? p->correct_iter : p->final_iter;
        jacobi_sweeps(s->A[l], b[l], omega, x[l], buf[l], nx, ny, nz, post);
    } else {
        jacobi_sweeps(s->A[l], b[l], omega, x[l], buf[l], nx, ny, nz, p->coarsest_iter);
    }
}

void or_mg_clean_vcycle_jacobi(const or_mg_system* sys, const or_mg_params* p,
                               double omega, const double* r, double* z) {
    double** mx = alloc_levels(sys);
    double** mb = alloc_levels(sys);
    double** mbuf = alloc_levels(sys);
    const size_t n = sys->nx[0] * sys->ny[0] * sys->nz[0];
    or_copy(r, mb[0], n);
    clean_vcycle_jacobi(sys, p, omega, 0, mx, mb, mbuf);
    or_copy(mx[0], z, n);
    free_levels(sys, mx);
    free_levels(sys, mb);
    free_levels(sys, mbuf);
}

void or_mg_clean_vcycle(const or_mg_system* sys, const or_mg_params* p,
                        const double* r, double* z) {
    double** mx = alloc_levels(sys);
    double** mb = alloc_levels(sys);
    double** mbuf = alloc_levels(sys);
    const size_t n = sys->nx[0] * sys->ny[0] * sys->nz[0];
    or_copy(r, mb[0], n);
    clean_vcycle(sys, p, 0, mx, mb, mbuf);
    or_copy(mx[0], z, n);
    free_levels(sys, mx);
    free_levels(sys, mb);
    free_levels(sys, mbuf);
}

static int mgpcg_clean(or_mg_system* sys, const or_mg_params* p, double jacobi_omega,
                       unsigned max_iter, double rel_tol, unsigned* last_iter,
                       double* true_rel_residual);

int or_mgpcg_clean_solve(or_mg_system* sys, const or_mg_params* p,
                         unsigned max_iter, double rel_tol,
                         unsigned* last_iter, double* true_rel_residual) {
    return mgpcg_clean(sys, p, 0.0, max_iter, rel_tol, last_iter, true_rel_residual);
}

int or_mgpcg_clean_solve_jacobi(or_mg_system* sys, const or_mg_params* p, double omega,
                                unsigned max_iter, double rel_tol,
                                unsigned* last_iter, double* true_rel_residual) {
    return mgpcg_clean(sys, p, omega, max_iter, rel_tol, last_iter, true_rel_residual);
}

/* jacobi_omega > 0: the weighted-Jacobi V-cycle, else the red-black one. */
static int mgpcg_clean(or_mg_system* sys, const or_mg_params* p, double jacobi_omega,
                       unsigned max_iter, double rel_tol, unsigned* last_iter,
                       double* true_rel_residual) {
    const size_t nx = sys->nx[0], ny = sys->ny[0], nz = sys->nz[0];
    const size_t n = nx * ny * nz;
    double* r = (double*)calloc(n, sizeof(double));
    double* d = (double*)calloc(n, sizeof(double));
    double* q = (double*)calloc(n, sizeof(double));
    double** mx = alloc_levels(sys);
    double** mb = alloc_levels(sys);
    double** mbuf = alloc_levels(sys);
    double* x = sys->x[0];
    or_set(0.0, x, n);
    or_copy(sys->b[0], r, n);
    const double bnorm = or_l2norm(sys->b[0], nx, ny, nz);
    double rnorm = bnorm;
    unsigned iter = 0;
    double sigma = 0.0;
    while (iter < max_iter && rnorm > rel_tol * bnorm && bnorm > 0.0) {
        or_copy(r, mb[0], n);
        if (jacobi_omega > 0.0)
            clean_vcycle_jacobi(sys, p, jacobi_omega, 0, mx, mb, mbuf);
        else
            clean_vcycle(sys, p, 0, mx, mb, mbuf); /* z = mx[0] */
        double sigma_new = or_dot(r, mx[0], nx, ny, nz);
        if (iter == 0)
            or_copy(mx[0], d, n);
        else
            or_axpy(sigma_new / sigma, d, mx[0], d, nx, ny, nz);
        sigma = sigma_new;
        or_mvm(sys->A[0], d, q, nx, ny, nz);
        double alpha = sigma / or_dot(d, q, nx, ny, nz);
        or_axpy(alpha, d, x, x, nx, ny, nz);
        or_axpy(-alpha, q, r, r, nx, ny, nz);
        rnorm = or_l2norm(r, nx, ny, nz);
        ++iter;
    }
    or_residual(sys->A[0], x, sys->b[0], r, nx, ny, nz);
    *last_iter = iter;
    *true_rel_residual =
        bnorm > 0.0 ? or_l2norm(r, nx, ny, nz) / bnorm : 0.0;
    free(r);
    free(d);
    free(q);
    free_levels(sys, mx);
    free_levels(sys, mb);
    free_levels(sys, mbuf);
    return *true_rel_residual <= rel_tol;
}

/* ============================================ Single-phase pressure ===== */

/* GridSinglePhasePressureSolver3::buildMarkers top level --
 * src/jet/grid_single_phase_pressure_solver3.cpp:258-269.  SDFs are sampled
 * at cell centres (src/jet/grid3.cpp:32-38); isInsideSdf(phi) = phi < 0
 * (include/jet/detail/level_set_utils-inl.h:23-26). */
void or_sp_markers(const double* boundary_sdf, const double* fluid_sdf,
                   char* markers, size_t n) {
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; ++c) {
        if (boundary_sdf[c] < 0)
            markers[c] = OR_BOUNDARY;
        else if (fluid_sdf[c] < 0)
            markers[c] = OR_FLUID;
        else
            markers[c] = OR_AIR;
    }
}

/* argmax3 -- include/jet/detail/math_utils-inl.h:107-113 (ties go to the
 * later argument). */
static int argmax3i(int x, int y, int z) {
    if (x > y) return (x > z) ? 0 : 2;
    return (y > z) ? 1 : 2;
}

/* Coarse markers: 4^3 clamped neighbourhood majority vote --
 * src/jet/grid_single_phase_pressure_solver3.cpp:271-327. */
void or_sp_coarsen_markers(const char* finer, char* coarser, size_t cnx,
                           size_t cny, size_t cnz) {
    const size_t fnx = 2 * cnx, fny = 2 * cny;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < cnz; ++k) {
        size_t kI[4] = {(k > 0) ? 2 * k - 1 : 2 * k, 2 * k, 2 * k + 1,
                        (k + 1 < cnz) ? 2 * k + 2 : 2 * k + 1};
        for (size_t j = 0; j < cny; ++j) {
            size_t jI[4] = {(j > 0) ? 2 * j - 1 : 2 * j, 2 * j, 2 * j + 1,
                            (j + 1 < cny) ? 2 * j + 2 : 2 * j + 1};
            for (size_t i = 0; i < cnx; ++i) {
                size_t iI[4] = {(i > 0) ? 2 * i - 1 : 2 * i, 2 * i, 2 * i + 1,
                                (i + 1 < cnx) ? 2 * i + 2 : 2 * i + 1};
                int cnt[3] = {0, 0, 0};
                for (int z = 0; z < 4; ++z)
                    for (int y = 0; y < 4; ++y)
                        for (int x = 0; x < 4; ++x) {
                            char f = finer[IX(iI[x], jI[y], kI[z], fnx, fny)];
                            if (f == OR_BOUNDARY)
                                ++cnt[OR_BOUNDARY];
                            else if (f == OR_FLUID)
                                ++cnt[OR_FLUID];
                            else
                                ++cnt[OR_AIR];
                        }
                coarser[IX(i, j, k, cnx, cny)] =
                    (char)argmax3i(cnt[0], cnt[1], cnt[2]);
            }
        }
    }
}

/* ----- 2-D single phase (GridSinglePhasePressureSolver2).  Cells are
 * nx*ny, x fastest; u is (nx+1)*ny and v is nx*(ny+1). */

/* Coarse markers: 4^2 clamped neighbourhood majority vote --
 * src/jet/grid_single_phase_pressure_solver2.cpp:238-283. */
void or_sp_coarsen_markers2(const char* finer, char* coarser, size_t cnx, size_t cny) {
    const size_t fnx = 2 * cnx;
    for (size_t j = 0; j < cny; ++j) {
        const size_t jI[4] = {(j > 0) ? 2 * j - 1 : 2 * j, 2 * j, 2 * j + 1,
                              (j + 1 < cny) ? 2 * j + 2 : 2 * j + 1};
        for (size_t i = 0; i < cnx; ++i) {
            const size_t iI[4] = {(i > 0) ? 2 * i - 1 : 2 * i, 2 * i, 2 * i + 1,
                                  (i + 1 < cnx) ? 2 * i + 2 : 2 * i + 1};
            int cnt[3] = {0, 0, 0};
            for (int y = 0; y < 4; ++y)
                for (int x = 0; x < 4; ++x) {
                    const char f = finer[iI[x] + fnx * jI[y]];
                    if (f == OR_BOUNDARY)
                        ++cnt[OR_BOUNDARY];
                    else if (f == OR_FLUID)
                        ++cnt[OR_FLUID];
                    else
                        ++cnt[OR_AIR];
                }
            coarser[i + cnx * j] = (char)argmax3i(cnt[0], cnt[1], cnt[2]);
        }
    }
}

/* buildSingleSystem (2-D) -- src/jet/grid_single_phase_pressure_solver2.cpp:
 * 25-67, with b = FaceCenteredGrid2::divergenceAtCellCenter
 * (src/jet/face_centered_grid2.cpp:107-118).  Rows are written as the
 * 4-coefficient rows of a one-plane system (front = 0); b is written only
 * when u is given (coarse levels need A alone). */
void or_sp_build_system2(const char* markers, const double* u, const double* v, size_t nx,
                         size_t ny, const double h[2], or_row* A, double* b) {
    const double ihx = 1.0 / h[0], ihy = 1.0 / h[1];
    const double ihx2 = ihx * ihx, ihy2 = ihy * ihy;
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            const size_t c = i + nx * j;
            or_row row = {0.0, 0.0, 0.0, 0.0};
            double bc = 0.0;
            if (markers[c] == OR_FLUID) {
                if (u) {
                    const double lu = u[i + (nx + 1) * j], ru = u[i + 1 + (nx + 1) * j];
                    const double bv = v[i + nx * j], tv = v[i + nx * (j + 1)];
                    bc = (ru - lu) / h[0] + (tv - bv) / h[1];
                }
                if (i + 1 < nx && markers[c + 1] != OR_BOUNDARY) {
                    row.center += ihx2;
                    if (markers[c + 1] == OR_FLUID) row.right -= ihx2;
                }
                if (i > 0 && markers[c - 1] != OR_BOUNDARY) row.center += ihx2;
                if (j + 1 < ny && markers[c + nx] != OR_BOUNDARY) {
                    row.center += ihy2;
                    if (markers[c + nx] == OR_FLUID) row.up -= ihy2;
                }
                if (j > 0 && markers[c - nx] != OR_BOUNDARY) row.center += ihy2;
            } else {
                row.center = 1.0;
            }
            A[c] = row;
            if (b) b[c] = bc;
        }
}

/* Compressed buildSingleSystem (2-D) -- src/jet/grid_single_phase_pressure_solver2.cpp:
 * 69-130: fluid cells numbered in x-fastest order; a row lists the centre
 * (its coefficient summed over non-boundary neighbours in the order +x, -x,
 * +y, -y) and the fluid neighbours, then MatrixCsr::addRow sorts the entries
 * by column (include/jet/detail/matrix_csr-inl.h:253-275), which for this
 * numbering is -y, -x, centre, +x, +y.  b = divergenceAtCellCenter.  With
 * val == NULL only the row and entry counts are returned. */
size_t or_sp_build_compressed2(const char* markers, const double* u, const double* v, size_t nx,
                               size_t ny, const double h[2], double* val, size_t* rp, size_t* ci,
                               double* b, size_t* nnz) {
    const double ihx = 1.0 / h[0], ihy = 1.0 / h[1];
    const double ihx2 = ihx * ihx, ihy2 = ihy * ihy;
    const size_t n = nx * ny;
    size_t* index = (size_t*)malloc((n ? n : 1) * sizeof(size_t));
    size_t rows = 0, e = 0;
    for (size_t c = 0; c < n; ++c)
        if (markers[c] == OR_FLUID) index[c] = rows++;
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            const size_t c = i + nx * j;
            if (markers[c] != OR_FLUID) continue;
            double centre = 0.0;
            int has[4] = {0, 0, 0, 0}; /* -y, -x, +x, +y fluid neighbours */
            if (i + 1 < nx && markers[c + 1] != OR_BOUNDARY) {
                centre += ihx2;
                has[2] = markers[c + 1] == OR_FLUID;
            }
            if (i > 0 && markers[c - 1] != OR_BOUNDARY) {
                centre += ihx2;
                has[1] = markers[c - 1] == OR_FLUID;
            }
            if (j + 1 < ny && markers[c + nx] != OR_BOUNDARY) {
                centre += ihy2;
                has[3] = markers[c + nx] == OR_FLUID;
            }
            if (j > 0 && markers[c - nx] != OR_BOUNDARY) {
                centre += ihy2;
                has[0] = markers[c - nx] == OR_FLUID;
            }
            const size_t row = index[c];
            if (val) {
                rp[row] = e;
                const size_t nb[4] = {c - nx, c - 1, c + 1, c + nx};
                const double cf[4] = {-ihy2, -ihx2, -ihx2, -ihy2};
                for (int q = 0; q < 4; ++q) {
                    if (q == 2) {
                        val[e] = centre;
                        ci[e++] = row;
                    }
                    if (!has[q]) continue;
                    val[e] = cf[q];
                    ci[e++] = index[nb[q]];
                }
                const double lu = u[i + (nx + 1) * j], ru = u[i + 1 + (nx + 1) * j];
                const double bv = v[i + nx * j], tv = v[i + nx * (j + 1)];
                b[row] = (ru - lu) / h[0] + (tv - bv) / h[1];
            } else {
                e += 1 + has[0] + has[1] + has[2] + has[3];
            }
        }
    if (val) rp[rows] = e;
    *nnz = e;
    free(index);
    return rows;
}

/* applyPressureGradient (2-D) -- src/jet/grid_single_phase_pressure_solver2.cpp:
 * 356-378: +side faces of fluid cells whose neighbour is not Boundary. */
void or_sp_apply_gradient2(const char* markers, const double* p, const double* u,
                           const double* v, double* uo, double* vo, size_t nx, size_t ny,
                           const double h[2]) {
    const double ihx = 1.0 / h[0], ihy = 1.0 / h[1];
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            const size_t c = i + nx * j;
            if (markers[c] != OR_FLUID) continue;
            if (i + 1 < nx && markers[c + 1] != OR_BOUNDARY) {
                const size_t f = i + 1 + (nx + 1) * j;
                uo[f] = u[f] + ihx * (p[c + 1] - p[c]);
            }
            if (j + 1 < ny && markers[c + nx] != OR_BOUNDARY) {
                const size_t f = i + nx * (j + 1);
                vo[f] = v[f] + ihy * (p[c + nx] - p[c]);
            }
        }
}

/* FaceCenteredGrid3::divergenceAtCellCenter -- src/jet/face_centered_grid3.cpp:137-152.
 * u is (nx+1)*ny*nz, v is nx*(ny+1)*nz, w is nx*ny*(nz+1), x fastest. */
double or_divergence(const double* u, const double* v, const double* w,
                     size_t nx, size_t ny, size_t nz, const double h[3],
                     size_t i, size_t j, size_t k) {
    (void)nz;
    double leftU = u[IX(i, j, k, nx + 1, ny)];
    double rightU = u[IX(i + 1, j, k, nx + 1, ny)];
    double bottomV = v[IX(i, j, k, nx, ny + 1)];
    double topV = v[IX(i, j + 1, k, nx, ny + 1)];
    double backW = w[IX(i, j, k, nx, ny)];
    double frontW = w[IX(i, j, k + 1, nx, ny)];
    return (rightU - leftU) / h[0] + (topV - bottomV) / h[1] +
           (frontW - backW) / h[2];
}

/* buildSingleSystem (flat) -- src/jet/grid_single_phase_pressure_solver3.cpp:25-79. */
void or_sp_build_system(const char* markers, const double* u, const double* v,
                        const double* w, size_t nx, size_t ny, size_t nz,
                        const double h[3], or_row* A, double* b) {
    double invH[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
    double invHSqr[3] = {invH[0] * invH[0], invH[1] * invH[1],
                         invH[2] * invH[2]};
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                or_row row = {0.0, 0.0, 0.0, 0.0};
                b[c] = 0.0;
                if (markers[c] == OR_FLUID) {
                    b[c] = or_divergence(u, v, w, nx, ny, nz, h, i, j, k);
                    if (i + 1 < nx && markers[c + 1] != OR_BOUNDARY) {
                        row.center += invHSqr[0];
                        if (markers[c + 1] == OR_FLUID) row.right -= invHSqr[0];
                    }
                    if (i > 0 && markers[c - 1] != OR_BOUNDARY)
                        row.center += invHSqr[0];
                    if (j + 1 < ny && markers[c + nx] != OR_BOUNDARY) {
                        row.center += invHSqr[1];
                        if (markers[c + nx] == OR_FLUID) row.up -= invHSqr[1];
                    }
                    if (j > 0 && markers[c - nx] != OR_BOUNDARY)
                        row.center += invHSqr[1];
                    if (k + 1 < nz && markers[c + nx * ny] != OR_BOUNDARY) {
                        row.center += invHSqr[2];
                        if (markers[c + nx * ny] == OR_FLUID)
                            row.front -= invHSqr[2];
                    }
                    if (k > 0 && markers[c - nx * ny] != OR_BOUNDARY)
                        row.center += invHSqr[2];
                } else {
                    row.center = 1.0;
                }
                A[c] = row;
            }
}

/* getBarycentric -- include/jet/detail/math_utils-inl.h:147-175 with
 * iLow = 0, iHigh = size - 1. */
static void barycentric(double x, long iHigh, long* i, double* f) {
    double s = floor(x);
    *i = (long)s;
    if (0 == iHigh) {
        *i = 0;
        *f = 0;
    } else if (*i < 0) {
        *i = 0;
        *f = 0;
    } else if (*i > iHigh - 1) {
        *i = iHigh - 1;
        *f = 1;
    } else {
        *f = x - s;
    }
}

static double lerp(double a, double b, double f) {
    return (1 - f) * a + f * b; /* include/jet/detail/math_utils-inl.h:177-180 */
}

/* LinearArraySampler3::operator() -- include/jet/detail/array_samplers3-inl.h:116-156. */
static double linear_sample(const double* data, long sx, long sy, long sz,
                            const double h[3], const double o[3],
                            const double p[3]) {
    double nxp = (p[0] - o[0]) / h[0];
    double nyp = (p[1] - o[1]) / h[1];
    double nzp = (p[2] - o[2]) / h[2];
    long i, j, k;
    double fx, fy, fz;
    barycentric(nxp, sx - 1, &i, &fx);
    barycentric(nyp, sy - 1, &j, &fy);
    barycentric(nzp, sz - 1, &k, &fz);
    long ip1 = (i + 1 < sx - 1) ? i + 1 : sx - 1;
    long jp1 = (j + 1 < sy - 1) ? j + 1 : sy - 1;
    long kp1 = (k + 1 < sz - 1) ? k + 1 : sz - 1;
#define D(a, b, c) data[(a) + sx * ((b) + sy * (c))]
    double b0 = lerp(lerp(D(i, j, k), D(ip1, j, k), fx),
                     lerp(D(i, jp1, k), D(ip1, jp1, k), fx), fy);
    double b1 = lerp(lerp(D(i, j, kp1), D(ip1, j, kp1), fx),
                     lerp(D(i, jp1, kp1), D(ip1, jp1, kp1), fx), fy);
#undef D
    return lerp(b0, b1, fz);
}

/* Coarse-level velocity for multigrid system building:
 * coarser.fill(finer->sampler()) -- src/jet/grid_single_phase_pressure_solver3.cpp:380-392,
 * FaceCenteredGrid3::fill / uPosition / resetSampler --
 * src/jet/face_centered_grid3.cpp:275-299,309-331,429-467. */
void or_face_fill_coarse(const double* fu, const double* fv, const double* fw,
                         size_t fnx, size_t fny, size_t fnz, const double fh[3],
                         const double origin[3], double* cu, double* cv,
                         double* cw, size_t cnx, size_t cny, size_t cnz,
                         const double ch[3]) {
    const double fou[3] = {origin[0] + 0.5 * 0.0, origin[1] + 0.5 * fh[1],
                           origin[2] + 0.5 * fh[2]};
    const double fov[3] = {origin[0] + 0.5 * fh[0], origin[1] + 0.5 * 0.0,
                           origin[2] + 0.5 * fh[2]};
    const double fow[3] = {origin[0] + 0.5 * fh[0], origin[1] + 0.5 * fh[1],
                           origin[2] + 0.5 * 0.0};
    const double cou[3] = {origin[0] + 0.5 * 0.0, origin[1] + 0.5 * ch[1],
                           origin[2] + 0.5 * ch[2]};
    const double cov[3] = {origin[0] + 0.5 * ch[0], origin[1] + 0.5 * 0.0,
                           origin[2] + 0.5 * ch[2]};
    const double cow[3] = {origin[0] + 0.5 * ch[0], origin[1] + 0.5 * ch[1],
                           origin[2] + 0.5 * 0.0};
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < cnz; ++k)
        for (size_t j = 0; j < cny; ++j)
            for (size_t i = 0; i <= cnx; ++i) {
                double p[3] = {cou[0] + ch[0] * (double)i,
                               cou[1] + ch[1] * (double)j,
                               cou[2] + ch[2] * (double)k};
                cu[IX(i, j, k, cnx + 1, cny)] = linear_sample(
                    fu, (long)fnx + 1, (long)fny, (long)fnz, fh, fou, p);
            }
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < cnz; ++k)
        for (size_t j = 0; j <= cny; ++j)
            for (size_t i = 0; i < cnx; ++i) {
                double p[3] = {cov[0] + ch[0] * (double)i,
                               cov[1] + ch[1] * (double)j,
                               cov[2] + ch[2] * (double)k};
                cv[IX(i, j, k, cnx, cny + 1)] = linear_sample(
                    fv, (long)fnx, (long)fny + 1, (long)fnz, fh, fov, p);
            }
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k <= cnz; ++k)
        for (size_t j = 0; j < cny; ++j)
            for (size_t i = 0; i < cnx; ++i) {
                double p[3] = {cow[0] + ch[0] * (double)i,
                               cow[1] + ch[1] * (double)j,
                               cow[2] + ch[2] * (double)k};
                cw[IX(i, j, k, cnx, cny)] = linear_sample(
                    fw, (long)fnx, (long)fny, (long)fnz + 1, fh, fow, p);
            }
}

/* applyPressureGradient -- src/jet/grid_single_phase_pressure_solver3.cpp:401-431.
 * Only the +side faces of fluid cells are updated.  Output may alias input. */
void or_sp_apply_gradient(const char* markers, const double* p,
                          const double* u, const double* v, const double* w,
                          double* uo, double* vo, double* wo, size_t nx,
                          size_t ny, size_t nz, const double h[3]) {
    double invH[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                if (markers[c] != OR_FLUID) continue;
                if (i + 1 < nx && markers[c + 1] != OR_BOUNDARY) {
                    size_t f = IX(i + 1, j, k, nx + 1, ny);
                    uo[f] = u[f] + invH[0] * (p[c + 1] - p[c]);
                }
                if (j + 1 < ny && markers[c + nx] != OR_BOUNDARY) {
                    size_t f = IX(i, j + 1, k, nx, ny + 1);
                    vo[f] = v[f] + invH[1] * (p[c + nx] - p[c]);
                }
                if (k + 1 < nz && markers[c + nx * ny] != OR_BOUNDARY) {
                    size_t f = IX(i, j, k + 1, nx, ny);
                    wo[f] = w[f] + invH[2] * (p[c + nx * ny] - p[c]);
                }
            }
}

/* ============================================== Fractional pressure ===== */

/* fractionInsideSdf -- include/jet/detail/level_set_utils-inl.h:51-62. */
double or_fraction_inside_sdf(double phi0, double phi1) {
    if (phi0 < 0 && phi1 < 0) return 1;
    if (phi0 < 0 && !(phi1 < 0)) return phi0 / (phi0 - phi1);
    if (!(phi0 < 0) && phi1 < 0) return phi1 / (phi1 - phi0);
    return 0;
}

static void cycle4(double* a) {
    double t = a[0];
    a[0] = a[1];
    a[1] = a[2];
    a[2] = a[3];
    a[3] = t;
}

/* fractionInside (Batty's square-corner area) --
 * include/jet/detail/level_set_utils-inl.h:71-147. */
double or_fraction_inside(double bl, double br, double tl, double tr) {
    int inside = (bl < 0 ? 1 : 0) + (tl < 0 ? 1 : 0) + (br < 0 ? 1 : 0) +
                 (tr < 0 ? 1 : 0);
    double list[4] = {bl, br, tr, tl};
    if (inside == 4) return 1;
    if (inside == 3) {
        while (list[0] < 0) cycle4(list);
        double side0 = 1 - or_fraction_inside_sdf(list[0], list[3]);
        double side1 = 1 - or_fraction_inside_sdf(list[0], list[1]);
        return 1 - 0.5 * side0 * side1;
    }
    if (inside == 2) {
        while (list[0] >= 0 || !(list[1] < 0 || list[2] < 0)) cycle4(list);
        if (list[1] < 0) {
            double side_left = or_fraction_inside_sdf(list[0], list[3]);
            double side_right = or_fraction_inside_sdf(list[1], list[2]);
            return 0.5 * (side_left + side_right);
        }
        double middle = 0.25 * (list[0] + list[1] + list[2] + list[3]);
        if (middle < 0) {
            double area = 0;
            double side1 = 1 - or_fraction_inside_sdf(list[0], list[3]);
            double side3 = 1 - or_fraction_inside_sdf(list[2], list[3]);
            area += 0.5 * side1 * side3;
            double side2 = 1 - or_fraction_inside_sdf(list[2], list[1]);
            double side0 = 1 - or_fraction_inside_sdf(list[0], list[1]);
            area += 0.5 * side0 * side2;
            return 1 - area;
        } else {
            double area = 0;
            double side0 = or_fraction_inside_sdf(list[0], list[1]);
            double side1 = or_fraction_inside_sdf(list[0], list[3]);
            area += 0.5 * side0 * side1;
            double side2 = or_fraction_inside_sdf(list[2], list[1]);
            double side3 = or_fraction_inside_sdf(list[2], list[3]);
            area += 0.5 * side2 * side3;
            return area;
        }
    }
    if (inside == 1) {
        while (list[0] >= 0) cycle4(list);
        double side0 = or_fraction_inside_sdf(list[0], list[3]);
        double side1 = or_fraction_inside_sdf(list[0], list[1]);
        return 0.5 * side0 * side1;
    }
    return 0;
}

static float face_weight(double p0, double p1, double p2, double p3) {
    double frac = or_fraction_inside(p0, p1, p2, p3);
    double weight = 1.0 - frac;
    if (weight < 0.0) weight = 0.0;
    if (weight > 1.0) weight = 1.0;
    if (weight < 0.01 && weight > 0.0) weight = 0.01; /* kMinWeight */
    return (float)weight;
}

/* buildWeights top level -- src/jet/grid_fractional_single_phase_pressure_solver3.cpp:553-617.
 * The four corner samples of every face are grid vertices; vertex_sdf holds
 * the boundary SDF at vertex (i,j,k) = origin + h*(i,j,k), an
 * (nx+1)*(ny+1)*(nz+1) array. */
void or_fr_face_weights(const double* vs, size_t nx, size_t ny, size_t nz,
                        float* uw, float* vw, float* ww) {
    const size_t vx = nx + 1, vy = ny + 1;
#define V(a, b, c) vs[(a) + vx * ((b) + vy * (c))]
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i <= nx; ++i)
                uw[IX(i, j, k, nx + 1, ny)] =
                    face_weight(V(i, j, k), V(i, j + 1, k), V(i, j, k + 1),
                                V(i, j + 1, k + 1));
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j <= ny; ++j)
            for (size_t i = 0; i < nx; ++i)
                vw[IX(i, j, k, nx, ny + 1)] =
                    face_weight(V(i, j, k), V(i, j, k + 1), V(i + 1, j, k),
                                V(i + 1, j, k + 1));
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k <= nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i)
                ww[IX(i, j, k, nx, ny)] =
                    face_weight(V(i, j, k), V(i, j + 1, k), V(i + 1, j, k),
                                V(i + 1, j + 1, k));
#undef V
}

/* Float restriction of fluid SDF / face weights -- anonymous restrict() in
 * src/jet/grid_fractional_single_phase_pressure_solver3.cpp:29-115.  A
 * direction whose fine size is not twice the coarse size (the staggered
 * direction of a face array) uses injection {0,1,0}. */
void or_fr_restrict_float(const float* finer, size_t fnx, size_t fny,
                          size_t fnz, float* coarser, size_t cnx, size_t cny,
                          size_t cnz) {
    static const float centered[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    static const float staggered[4] = {0.f, 1.f, 0.f, 0.f};
    int ks[3] = {fnx != 2 * cnx ? 3 : 4, fny != 2 * cny ? 3 : 4,
                 fnz != 2 * cnz ? 3 : 4};
    const float* kern[3] = {ks[0] == 3 ? staggered : centered,
                            ks[1] == 3 ? staggered : centered,
                            ks[2] == 3 ? staggered : centered};
    size_t cn[3] = {cnx, cny, cnz};
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < cnz; ++k) {
        size_t idx[3][4];
        size_t c3[3];
        c3[2] = k;
        for (size_t j = 0; j < cny; ++j) {
            c3[1] = j;
            for (size_t i = 0; i < cnx; ++i) {
                c3[0] = i;
                for (int d = 0; d < 3; ++d) {
                    size_t c = c3[d];
                    if (ks[d] == 3) {
                        idx[d][0] = (c > 0) ? 2 * c - 1 : 2 * c;
                        idx[d][1] = 2 * c;
                        idx[d][2] = (c + 1 < cn[d]) ? 2 * c + 1 : 2 * c;
                    } else {
                        idx[d][0] = (c > 0) ? 2 * c - 1 : 2 * c;
                        idx[d][1] = 2 * c;
                        idx[d][2] = 2 * c + 1;
                        idx[d][3] = (c + 1 < cn[d]) ? 2 * c + 2 : 2 * c + 1;
                    }
                }
                float sum = 0.0f;
                for (int z = 0; z < ks[2]; ++z)
                    for (int y = 0; y < ks[1]; ++y)
                        for (int x = 0; x < ks[0]; ++x) {
                            float w = kern[0][x] * kern[1][y] * kern[2][z];
                            sum += w * finer[IX(idx[0][x], idx[1][y], idx[2][z],
                                                fnx, fny)];
                        }
                coarser[IX(i, j, k, cnx, cny)] = sum;
            }
        }
    }
}

/* Fractional buildSingleSystem (flat) --
 * src/jet/grid_fractional_single_phase_pressure_solver3.cpp:117-267.
 * bu/bv/bw: boundary-velocity x/y/z components at u/v/w face centres (NULL
 * means a zero boundary velocity). */
void or_fr_build_system(const float* fsdf, const float* uw, const float* vw,
                        const float* ww, const double* u, const double* v,
                        const double* w, const double* bu, const double* bv,
                        const double* bw, size_t nx, size_t ny, size_t nz,
                        const double h[3], or_row* A, double* b) {
    double invH[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
    double invHSqr[3] = {invH[0] * invH[0], invH[1] * invH[1],
                         invH[2] * invH[2]};
    const size_t ux = nx + 1, vy = ny + 1;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                or_row row = {0.0, 0.0, 0.0, 0.0};
                double bc = 0.0;
                double centerPhi = fsdf[c];
                if (centerPhi < 0) {
                    double term;
                    const size_t uR = IX(i + 1, j, k, ux, ny);
                    const size_t uL = IX(i, j, k, ux, ny);
                    const size_t vU = IX(i, j + 1, k, nx, vy);
                    const size_t vD = IX(i, j, k, nx, vy);
                    const size_t wF = IX(i, j, k + 1, nx, ny);
                    const size_t wB = IX(i, j, k, nx, ny);
                    if (i + 1 < nx) {
                        term = uw[uR] * invHSqr[0];
                        double rightPhi = fsdf[c + 1];
                        if (rightPhi < 0) {
                            row.center += term;
                            row.right -= term;
                        } else {
                            double theta =
                                or_fraction_inside_sdf(centerPhi, rightPhi);
                            theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                            row.center += term / theta;
                        }
                        bc += uw[uR] * u[uR] * invH[0];
                    } else {
                        bc += u[uR] * invH[0];
                    }
                    if (i > 0) {
                        term = uw[uL] * invHSqr[0];
                        double leftPhi = fsdf[c - 1];
                        if (leftPhi < 0) {
                            row.center += term;
                        } else {
                            double theta =
                                or_fraction_inside_sdf(centerPhi, leftPhi);
                            theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                            row.center += term / theta;
                        }
                        bc -= uw[uL] * u[uL] * invH[0];
                    } else {
                        bc -= u[uL] * invH[0];
                    }
                    if (j + 1 < ny) {
                        term = vw[vU] * invHSqr[1];
                        double upPhi = fsdf[c + nx];
                        if (upPhi < 0) {
                            row.center += term;
                            row.up -= term;
                        } else {
                            double theta =
                                or_fraction_inside_sdf(centerPhi, upPhi);
                            theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                            row.center += term / theta;
                        }
                        bc += vw[vU] * v[vU] * invH[1];
                    } else {
                        bc += v[vU] * invH[1];
                    }
                    if (j > 0) {
                        term = vw[vD] * invHSqr[1];
                        double downPhi = fsdf[c - nx];
                        if (downPhi < 0) {
                            row.center += term;
                        } else {
                            double theta =
                                or_fraction_inside_sdf(centerPhi, downPhi);
                            theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                            row.center += term / theta;
                        }
                        bc -= vw[vD] * v[vD] * invH[1];
                    } else {
                        bc -= v[vD] * invH[1];
                    }
                    if (k + 1 < nz) {
                        term = ww[wF] * invHSqr[2];
                        double frontPhi = fsdf[c + nx * ny];
                        if (frontPhi < 0) {
                            row.center += term;
                            row.front -= term;
                        } else {
                            double theta =
                                or_fraction_inside_sdf(centerPhi, frontPhi);
                            theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                            row.center += term / theta;
                        }
                        bc += ww[wF] * w[wF] * invH[2];
                    } else {
                        bc += w[wF] * invH[2];
                    }
                    if (k > 0) {
                        term = ww[wB] * invHSqr[2];
                        double backPhi = fsdf[c - nx * ny];
                        if (backPhi < 0) {
                            row.center += term;
                        } else {
                            double theta =
                                or_fraction_inside_sdf(centerPhi, backPhi);
                            theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                            row.center += term / theta;
                        }
                        bc -= ww[wB] * w[wB] * invH[2];
                    } else {
                        bc -= w[wB] * invH[2];
                    }
                    double buR = bu ? bu[uR] : 0.0, buL = bu ? bu[uL] : 0.0;
                    double bvU = bv ? bv[vU] : 0.0, bvD = bv ? bv[vD] : 0.0;
                    double bwF = bw ? bw[wF] : 0.0, bwB = bw ? bw[wB] : 0.0;
                    double contrib = (1.0 - uw[uR]) * buR * invH[0] -
                                     (1.0 - uw[uL]) * buL * invH[0] +
                                     (1.0 - vw[vU]) * bvU * invH[1] -
                                     (1.0 - vw[vD]) * bvD * invH[1] +
                                     (1.0 - ww[wF]) * bwF * invH[2] -
                                     (1.0 - ww[wB]) * bwB * invH[2];
                    bc += contrib;
                    if (row.center < kEps) {
                        row.center = 1.0;
                        bc = 0.0;
                    }
                } else {
                    row.center = 1.0;
                }
                A[c] = row;
                b[c] = bc;
            }
}

/* Fractional applyPressureGradient --
 * src/jet/grid_fractional_single_phase_pressure_solver3.cpp:715-765.
 * A face is updated when its weight is positive and either side is fluid;
 * the gradient is divided by the ghost-fluid fraction theta (>= 0.01). */
void or_fr_apply_gradient(const float* fsdf, const float* uw, const float* vw,
                          const float* ww, const double* p, const double* u,
                          const double* v, const double* w, double* uo,
                          double* vo, double* wo, size_t nx, size_t ny,
                          size_t nz, const double h[3]) {
    double invH[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                double centerPhi = fsdf[c];
                if (i + 1 < nx) {
                    size_t f = IX(i + 1, j, k, nx + 1, ny);
                    if (uw[f] > 0.0 && (centerPhi < 0 || fsdf[c + 1] < 0)) {
                        double theta = or_fraction_inside_sdf(centerPhi, fsdf[c + 1]);
                        theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                        uo[f] = u[f] + invH[0] / theta * (p[c + 1] - p[c]);
                    }
                }
                if (j + 1 < ny) {
                    size_t f = IX(i, j + 1, k, nx, ny + 1);
                    if (vw[f] > 0.0 && (centerPhi < 0 || fsdf[c + nx] < 0)) {
                        double theta = or_fraction_inside_sdf(centerPhi, fsdf[c + nx]);
                        theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                        vo[f] = v[f] + invH[1] / theta * (p[c + nx] - p[c]);
                    }
                }
                if (k + 1 < nz) {
                    size_t f = IX(i, j, k + 1, nx, ny);
                    if (ww[f] > 0.0 && (centerPhi < 0 || fsdf[c + nx * ny] < 0)) {
                        double theta =
                            or_fraction_inside_sdf(centerPhi, fsdf[c + nx * ny]);
                        theta = (theta < 0.01) ? 0.01 : theta; /* std::max */
                        wo[f] = w[f] + invH[2] / theta * (p[c + nx * ny] - p[c]);
                    }
                }
            }
}

/* ====================================== Fractional pressure, 2-D ======= */

/* buildWeights top level, 2-D -- src/jet/grid_fractional_single_phase_pressure_solver2.cpp:
 * 440-480.  Unlike the 3-D solver (fractionInside over a face's corners), a
 * 2-D face weight is 1 - fractionInsideSdf of the two boundary-SDF samples half
 * a cell before and after the face along its normal, which are the centres
 * of the two cells the face separates.  bsdf holds the boundary SDF at the
 * cell centres of the grid extended by one ghost cell on every side:
 * cell (i, j), -1 <= i <= nx, -1 <= j <= ny, at bsdf[(i + 1) + (nx + 2)(j + 1)]. */
void or_fr_face_weights2(const double* bsdf, size_t nx, size_t ny, float* uw,
                         float* vw) {
    const size_t gx = nx + 2;
#define G(a, b) bsdf[(size_t)((a) + 1) + gx * (size_t)((b) + 1)]
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i <= nx; ++i) {
            double frac = or_fraction_inside_sdf(G((long)i - 1, (long)j), G((long)i, (long)j));
            double weight = 1.0 - frac;
            if (weight < 0.0) weight = 0.0;
            if (weight > 1.0) weight = 1.0;
            if (weight < 0.01 && weight > 0.0) weight = 0.01; /* kMinWeight */
            uw[i + (nx + 1) * j] = (float)weight;
        }
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j <= ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            double frac = or_fraction_inside_sdf(G((long)i, (long)j - 1), G((long)i, (long)j));
            double weight = 1.0 - frac;
            if (weight < 0.0) weight = 0.0;
            if (weight > 1.0) weight = 1.0;
            if (weight < 0.01 && weight > 0.0) weight = 0.01;
            vw[i + nx * j] = (float)weight;
        }
#undef G
}

/* Float restriction, 2-D -- the anonymous restrict() in
 * src/jet/grid_fractional_single_phase_pressure_solver2.cpp:29-93: per axis
 * the centred kernel over fine 2c-1 .. 2c+2 (clamped), or injection
 * {0, 1, 0} over 2c-1 .. 2c+1 where the fine size is not twice the coarse
 * size; y outer, x inner, w = kx * ky. */
void or_fr_restrict_float2(const float* finer, size_t fnx, size_t fny,
                           float* coarser, size_t cnx, size_t cny) {
    static const float centered[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    static const float staggered[4] = {0.f, 1.f, 0.f, 0.f};
    const int kx = fnx != 2 * cnx ? 3 : 4, ky = fny != 2 * cny ? 3 : 4;
    const float* kernx = kx == 3 ? staggered : centered;
    const float* kerny = ky == 3 ? staggered : centered;
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < cny; ++j) {
        size_t jj[4] = {0, 0, 0, 0};
        if (ky == 3) {
            jj[0] = (j > 0) ? 2 * j - 1 : 2 * j;
            jj[1] = 2 * j;
            jj[2] = (j + 1 < cny) ? 2 * j + 1 : 2 * j;
        } else {
            jj[0] = (j > 0) ? 2 * j - 1 : 2 * j;
            jj[1] = 2 * j;
            jj[2] = 2 * j + 1;
            jj[3] = (j + 1 < cny) ? 2 * j + 2 : 2 * j + 1;
        }
        for (size_t i = 0; i < cnx; ++i) {
            size_t ii[4] = {0, 0, 0, 0};
            if (kx == 3) {
                ii[0] = (i > 0) ? 2 * i - 1 : 2 * i;
                ii[1] = 2 * i;
                ii[2] = (i + 1 < cnx) ? 2 * i + 1 : 2 * i;
            } else {
                ii[0] = (i > 0) ? 2 * i - 1 : 2 * i;
                ii[1] = 2 * i;
                ii[2] = 2 * i + 1;
                ii[3] = (i + 1 < cnx) ? 2 * i + 2 : 2 * i + 1;
            }
            float sum = 0.0f;
            for (int y = 0; y < ky; ++y)
                for (int x = 0; x < kx; ++x) {
                    float w = kernx[x] * kerny[y];
                    sum += w * finer[ii[x] + fnx * jj[y]];
                }
            coarser[i + cnx * j] = sum;
        }
    }
}

/* buildSingleSystem, 2-D -- src/jet/grid_fractional_single_phase_pressure_solver2.cpp:95-203.
 * Rows are written as 4 doubles {center, right, up, 0} (a one-plane level).
 * bu / bv: collider velocity x at the u faces and y at the v faces (NULL:
 * zero); the collider term sum (1 - w) u_collider / h is Jet's
 * boundaryContribution. */
void or_fr_build_system2(const float* fsdf, const float* uw, const float* vw,
                         const double* u, const double* v, const double* bu,
                         const double* bv, size_t nx, size_t ny, const double h[2],
                         or_row* A, double* b) {
    const double invH[2] = {1.0 / h[0], 1.0 / h[1]};
    const double invHSqr[2] = {invH[0] * invH[0], invH[1] * invH[1]};
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            const size_t c = i + nx * j;
            or_row row = {0.0, 0.0, 0.0, 0.0};
            double bij = 0.0;
            const double centerPhi = fsdf[c];
            const size_t uR = (i + 1) + (nx + 1) * j, uL = uR - 1;
            const size_t vU = i + nx * (j + 1), vD = i + nx * j;
            if (centerPhi < 0) {
                double term;
                if (i + 1 < nx) {
                    term = uw[uR] * invHSqr[0];
                    const double rightPhi = fsdf[c + 1];
                    if (rightPhi < 0) {
                        row.center += term;
                        row.right -= term;
                    } else {
                        double theta = or_fraction_inside_sdf(centerPhi, rightPhi);
                        theta = theta < 0.01 ? 0.01 : theta; /* std::max */
                        row.center += term / theta;
                    }
                    bij += uw[uR] * u[uR] * invH[0];
                } else {
                    bij += u[uR] * invH[0];
                }
                if (i > 0) {
                    term = uw[uL] * invHSqr[0];
                    const double leftPhi = fsdf[c - 1];
                    if (leftPhi < 0) {
                        row.center += term;
                    } else {
                        double theta = or_fraction_inside_sdf(centerPhi, leftPhi);
                        theta = theta < 0.01 ? 0.01 : theta;
                        row.center += term / theta;
                    }
                    bij -= uw[uL] * u[uL] * invH[0];
                } else {
                    bij -= u[uL] * invH[0];
                }
                if (j + 1 < ny) {
                    term = vw[vU] * invHSqr[1];
                    const double upPhi = fsdf[c + nx];
                    if (upPhi < 0) {
                        row.center += term;
                        row.up -= term;
                    } else {
                        double theta = or_fraction_inside_sdf(centerPhi, upPhi);
                        theta = theta < 0.01 ? 0.01 : theta;
                        row.center += term / theta;
                    }
                    bij += vw[vU] * v[vU] * invH[1];
                } else {
                    bij += v[vU] * invH[1];
                }
                if (j > 0) {
                    term = vw[vD] * invHSqr[1];
                    const double downPhi = fsdf[c - nx];
                    if (downPhi < 0) {
                        row.center += term;
                    } else {
                        double theta = or_fraction_inside_sdf(centerPhi, downPhi);
                        theta = theta < 0.01 ? 0.01 : theta;
                        row.center += term / theta;
                    }
                    bij -= vw[vD] * v[vD] * invH[1];
                } else {
                    bij -= v[vD] * invH[1];
                }
                const double buR = bu ? bu[uR] : 0.0, buL = bu ? bu[uL] : 0.0;
                const double bvU = bv ? bv[vU] : 0.0, bvD = bv ? bv[vD] : 0.0;
                const double contribution = (1.0 - uw[uR]) * buR * invH[0] -
                                            (1.0 - uw[uL]) * buL * invH[0] +
                                            (1.0 - vw[vU]) * bvU * invH[1] -
                                            (1.0 - vw[vD]) * bvD * invH[1];
                bij += contribution;
                if (row.center < 2.220446049250313080847e-16) { /* kEpsilonD */
                    row.center = 1.0;
                    bij = 0.0;
                }
            } else {
                row.center = 1.0;
            }
            A[c] = row;
            if (b) b[c] = bij;
        }
}

/* applyPressureGradient, 2-D -- src/jet/grid_fractional_single_phase_pressure_solver2.cpp:565-600. */
void or_fr_apply_gradient2(const float* fsdf, const float* uw, const float* vw,
                           const double* p, const double* u, const double* v,
                           double* uo, double* vo, size_t nx, size_t ny,
                           const double h[2]) {
    const double invH[2] = {1.0 / h[0], 1.0 / h[1]};
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            const size_t c = i + nx * j;
            const double centerPhi = fsdf[c];
            if (i + 1 < nx) {
                const size_t f = (i + 1) + (nx + 1) * j;
                if (uw[f] > 0.0 && (centerPhi < 0 || fsdf[c + 1] < 0)) {
                    const double rightPhi = fsdf[c + 1];
                    double theta = or_fraction_inside_sdf(centerPhi, rightPhi);
                    theta = theta < 0.01 ? 0.01 : theta;
                    uo[f] = u[f] + invH[0] / theta * (p[c + 1] - p[c]);
                }
            }
            if (j + 1 < ny) {
                const size_t f = i + nx * (j + 1);
                if (vw[f] > 0.0 && (centerPhi < 0 || fsdf[c + nx] < 0)) {
                    const double upPhi = fsdf[c + nx];
                    double theta = or_fraction_inside_sdf(centerPhi, upPhi);
                    theta = theta < 0.01 ? 0.01 : theta;
                    vo[f] = v[f] + invH[1] / theta * (p[c + nx] - p[c]);
                }
            }
        }
}

/* ============================================ Boundary conditions ======= */

/* RigidBodyCollider3::velocityAt -- src/jet/rigid_body_collider3.cpp:25-28:
 * linearVelocity + angularVelocity x (p - translation), with Vector3::cross
 * (include/jet/detail/vector3-inl.h:125-127).  NULL: zero velocity. */
void or_collider_velocity(const or_rigid_collider* c, const double p[3], double out[3]) {
    if (!c) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double* a = c->angular_velocity;
    double r[3] = {p[0] - c->translation[0], p[1] - c->translation[1], p[2] - c->translation[2]};
    out[0] = c->linear_velocity[0] + (a[1] * r[2] - r[1] * a[2]);
    out[1] = c->linear_velocity[1] + (a[2] * r[0] - r[2] * a[0]);
    out[2] = c->linear_velocity[2] + (a[0] * r[1] - r[0] * a[1]);
}

/* projectAndApplyFriction -- src/jet/physics_helpers.h:34-45, with
 * Vector3::projected / dot / length (include/jet/detail/vector3-inl.h:
 * 120-122, 276-282, 302-305). */
void or_project_and_apply_friction(const double vel[3], const double n[3], double friction,
                                   double out[3]) {
    double d = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
    out[0] = vel[0] - n[0] * d;
    out[1] = vel[1] - n[1] * d;
    out[2] = vel[2] - n[2] * d;
    double l2 = out[0] * out[0] + out[1] * out[1] + out[2] * out[2];
    if (l2 > 0) {
        double nd = -(vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2]);
        double veln = nd < 0.0 ? 0.0 : nd; /* std::max(nd, 0.0) */
        double sc = 1.0 - friction * veln / sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
        if (sc < 0.0) sc = 0.0; /* std::max(sc, 0.0) */
        out[0] *= sc;
        out[1] *= sc;
        out[2] *= sc;
    }
}

/* CellCenteredScalarGrid3 sample: LinearArraySampler3::operator() over the
 * data at origin o + h/2 (src/jet/cell_centered_scalar_grid3.cpp:60-62,
 * src/jet/scalar_grid3.cpp:83,214-218). */
static void cell_origin(const double h[3], const double o[3], double d[3]) {
    d[0] = o[0] + 0.5 * h[0];
    d[1] = o[1] + 0.5 * h[1];
    d[2] = o[2] + 0.5 * h[2];
}

double or_sample_cell(const double* data, size_t nx, size_t ny, size_t nz, const double h[3],
                      const double o[3], const double p[3]) {
    double d[3];
    cell_origin(h, o, d);
    return linear_sample(data, (long)nx, (long)ny, (long)nz, h, d, p);
}

/* ScalarGrid3::gradient -- src/jet/scalar_grid3.cpp:89-102: the weights of
 * LinearArraySampler3::getCoordinatesAndWeights (include/jet/detail/
 * array_samplers3-inl.h:152-194, which scales by 1/h instead of dividing)
 * applied to gradient3 at the 8 data points (src/jet/fdm_utils.cpp:49-67). */
void or_gradient_cell(const double* data, size_t nx, size_t ny, size_t nz, const double h[3],
                      const double o[3], const double p[3], double g[3]) {
    double d[3], inv[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
    cell_origin(h, o, d);
    long sz3[3] = {(long)nx, (long)ny, (long)nz};
    long lo[3], hi[3];
    double f[3];
    for (int a = 0; a < 3; ++a) {
        barycentric((p[a] - d[a]) * inv[a], sz3[a] - 1, &lo[a], &f[a]);
        hi[a] = lo[a] + 1 < sz3[a] - 1 ? lo[a] + 1 : sz3[a] - 1;
    }
    const double fx = f[0], fy = f[1], fz = f[2];
    const double wt[8] = {(1 - fx) * (1 - fy) * (1 - fz), fx * (1 - fy) * (1 - fz),
                          (1 - fx) * fy * (1 - fz),       fx * fy * (1 - fz),
                          (1 - fx) * (1 - fy) * fz,       fx * (1 - fy) * fz,
                          (1 - fx) * fy * fz,             fx * fy * fz};
    g[0] = g[1] = g[2] = 0.0;
    for (int q = 0; q < 8; ++q) {
        const long i = (q & 1) ? hi[0] : lo[0];
        const long j = (q & 2) ? hi[1] : lo[1];
        const long k = (q & 4) ? hi[2] : lo[2];
#define D(a, b, c) data[(a) + (long)nx * ((b) + (long)ny * (c))]
        const double left = D(i > 0 ? i - 1 : i, j, k);
        const double right = D(i + 1 < (long)nx ? i + 1 : i, j, k);
        const double down = D(i, j > 0 ? j - 1 : j, k);
        const double up = D(i, j + 1 < (long)ny ? j + 1 : j, k);
        const double back = D(i, j, k > 0 ? k - 1 : k);
        const double front = D(i, j, k + 1 < (long)nz ? k + 1 : k);
#undef D
        g[0] += wt[q] * (0.5 * (right - left) / h[0]);
        g[1] += wt[q] * (0.5 * (up - down) / h[1]);
        g[2] += wt[q] * (0.5 * (front - back) / h[2]);
    }
}

/* extrapolateToRegion, 3-D -- include/jet/detail/array_utils-inl.h:172-241.
 * Each pass writes only cells that were invalid and reads only cells that
 * were valid at its start, so the passes are order-independent. */
void or_extrapolate_to_region(const double* input, const char* valid, size_t sx, size_t sy,
                              size_t sz, unsigned iterations, double* output) {
    const size_t n = sx * sy * sz;
    char* v0 = (char*)malloc(n);
    char* v1 = (char*)calloc(n, 1);
    for (size_t c = 0; c < n; ++c) {
        v0[c] = valid[c];
        output[c] = input[c];
    }
    for (unsigned it = 0; it < iterations; ++it) {
        for (size_t k = 0; k < sz; ++k)
            for (size_t j = 0; j < sy; ++j)
                for (size_t i = 0; i < sx; ++i) {
                    const size_t c = IX(i, j, k, sx, sy);
                    if (v0[c]) {
                        v1[c] = 1;
                        continue;
                    }
                    double sum = 0.0;
                    unsigned count = 0;
                    if (i + 1 < sx && v0[c + 1]) { sum += output[c + 1]; ++count; }
                    if (i > 0 && v0[c - 1]) { sum += output[c - 1]; ++count; }
                    if (j + 1 < sy && v0[c + sx]) { sum += output[c + sx]; ++count; }
                    if (j > 0 && v0[c - sx]) { sum += output[c - sx]; ++count; }
                    if (k + 1 < sz && v0[c + sx * sy]) { sum += output[c + sx * sy]; ++count; }
                    if (k > 0 && v0[c - sx * sy]) { sum += output[c - sx * sy]; ++count; }
                    if (count > 0) {
                        output[c] = sum / (double)count;
                        v1[c] = 1;
                    }
                }
        char* t = v0;
        v0 = v1;
        v1 = t;
    }
    free(v0);
    free(v1);
}

/* Face grid a = 0 / 1 / 2 (u / v / w): sizes and data origin
 * o + h/2 in the two other axes (src/jet/face_centered_grid3.cpp:216-238,
 * 437-447). */
static void face_dims(int a, size_t nx, size_t ny, size_t nz, size_t s[3]) {
    s[0] = nx + (a == 0);
    s[1] = ny + (a == 1);
    s[2] = nz + (a == 2);
}
static void face_origin(int a, const double h[3], const double o[3], double d[3]) {
    for (int b = 0; b < 3; ++b) d[b] = o[b] + 0.5 * (b == a ? 0.0 : h[b]);
}
static void face_pos(const double d[3], const double h[3], size_t i, size_t j, size_t k,
                     double p[3]) {
    p[0] = d[0] + h[0] * (double)i;
    p[1] = d[1] + h[1] * (double)j;
    p[2] = d[2] + h[2] * (double)k;
}

/* clamp -- include/jet/detail/math_utils-inl.h */
static double clamp01(double x) {
    if (x < 0.0) return 0.0;
    if (x > 1.0) return 1.0;
    return x;
}

/* GridFractionalBoundaryConditionSolver3::constrainVelocity --
 * src/jet/grid_fractional_boundary_condition_solver3.cpp:26-232, then for
 * blocked GridBlockedBoundaryConditionSolver3::constrainVelocity's marker
 * pass -- src/jet/grid_blocked_boundary_condition_solver3.cpp:23-69 with the
 * marker of :86-94 (collider where the cell-centre SDF is negative). */
void or_constrain_velocity(int blocked, size_t nx, size_t ny, size_t nz, const double h[3],
                           const double o[3], double* u, double* v, double* w,
                           const double* sdf, const or_rigid_collider* collider,
                           unsigned depth, int closed) {
    double* comp[3] = {u, v, w};
    size_t dims[3][3];
    double org[3][3];
    for (int a = 0; a < 3; ++a) {
        face_dims(a, nx, ny, nz, dims[a]);
        face_origin(a, h, o, org[a]);
    }
    /* collider velocity at the face; phi at the two cell centres it joins */
    for (int a = 0; a < 3; ++a) {
        const size_t* s = dims[a];
        const size_t n = s[0] * s[1] * s[2];
        char* marker = (char*)malloc(n);
        for (size_t k = 0; k < s[2]; ++k)
            for (size_t j = 0; j < s[1]; ++j)
                for (size_t i = 0; i < s[0]; ++i) {
                    double p[3], q0[3], q1[3], cv[3];
                    face_pos(org[a], h, i, j, k, p);
                    for (int b = 0; b < 3; ++b) {
                        const double off = b == a ? 0.5 * h[a] : 0.0;
                        q0[b] = p[b] - off;
                        q1[b] = p[b] + off;
                    }
                    const double phi0 = or_sample_cell(sdf, nx, ny, nz, h, o, q0);
                    const double phi1 = or_sample_cell(sdf, nx, ny, nz, h, o, q1);
                    double frac = or_fraction_inside_sdf(phi0, phi1);
                    frac = 1.0 - clamp01(frac);
                    const size_t c = IX(i, j, k, s[0], s[1]);
                    if (frac > 0.0) {
                        marker[c] = 1;
                    } else {
                        or_collider_velocity(collider, p, cv);
                        comp[a][c] = cv[a];
                        marker[c] = 0;
                    }
                }
        /* free slip: extrapolate fluid velocity into the collider */
        double* tmp = (double*)malloc(n * sizeof(double));
        or_extrapolate_to_region(comp[a], marker, s[0], s[1], s[2], depth, tmp);
        memcpy(comp[a], tmp, n * sizeof(double));
        free(tmp);
        free(marker);
    }
    /* no flux: project onto the collider surface (reads all three
     * extrapolated components; results land after all three) */
    double* out[3];
    for (int a = 0; a < 3; ++a) {
        const size_t* s = dims[a];
        out[a] = (double*)malloc(s[0] * s[1] * s[2] * sizeof(double));
        for (size_t k = 0; k < s[2]; ++k)
            for (size_t j = 0; j < s[1]; ++j)
                for (size_t i = 0; i < s[0]; ++i) {
                    const size_t c = IX(i, j, k, s[0], s[1]);
                    double p[3];
                    face_pos(org[a], h, i, j, k, p);
                    if (!(or_sample_cell(sdf, nx, ny, nz, h, o, p) < 0)) {
                        out[a][c] = comp[a][c];
                        continue;
                    }
                    double cv[3], vel[3], g[3];
                    or_collider_velocity(collider, p, cv);
                    for (int b = 0; b < 3; ++b)
                        vel[b] = linear_sample(comp[b], (long)dims[b][0], (long)dims[b][1],
                                               (long)dims[b][2], h, org[b], p);
                    or_gradient_cell(sdf, nx, ny, nz, h, o, p, g);
                    if (g[0] * g[0] + g[1] * g[1] + g[2] * g[2] > 0.0) {
                        const double l = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                        const double nrm[3] = {g[0] / l, g[1] / l, g[2] / l};
                        const double velr[3] = {vel[0] - cv[0], vel[1] - cv[1], vel[2] - cv[2]};
                        double velt[3];
                        or_project_and_apply_friction(velr, nrm, collider ? collider->friction : 0.0,
                                                      velt);
                        out[a][c] = velt[a] + cv[a];
                    } else {
                        out[a][c] = cv[a];
                    }
                }
    }
    for (int a = 0; a < 3; ++a) {
        memcpy(comp[a], out[a], dims[a][0] * dims[a][1] * dims[a][2] * sizeof(double));
        free(out[a]);
    }
    /* no flux through closed domain walls */
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j) {
            if (closed & OR_DIR_LEFT) u[IX(0, j, k, nx + 1, ny)] = 0;
            if (closed & OR_DIR_RIGHT) u[IX(nx, j, k, nx + 1, ny)] = 0;
        }
    for (size_t k = 0; k < nz; ++k)
        for (size_t i = 0; i < nx; ++i) {
            if (closed & OR_DIR_DOWN) v[IX(i, 0, k, nx, ny + 1)] = 0;
            if (closed & OR_DIR_UP) v[IX(i, ny, k, nx, ny + 1)] = 0;
        }
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            if (closed & OR_DIR_BACK) w[IX(i, j, 0, nx, ny)] = 0;
            if (closed & OR_DIR_FRONT) w[IX(i, j, nz, nx, ny)] = 0;
        }
    if (!blocked) return;
    /* blocked: faces between a collider cell and a fluid cell take the
     * collider velocity */
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
#define COLL(a, b, c) (sdf[IX(a, b, c, nx, ny)] < 0)
                if (!COLL(i, j, k)) continue;
                double p[3], cv[3];
                if (i > 0 && !COLL(i - 1, j, k)) {
                    face_pos(org[0], h, i, j, k, p);
                    or_collider_velocity(collider, p, cv);
                    u[IX(i, j, k, nx + 1, ny)] = cv[0];
                }
                if (i + 1 < nx && !COLL(i + 1, j, k)) {
                    face_pos(org[0], h, i + 1, j, k, p);
                    or_collider_velocity(collider, p, cv);
                    u[IX(i + 1, j, k, nx + 1, ny)] = cv[0];
                }
                if (j > 0 && !COLL(i, j - 1, k)) {
                    face_pos(org[1], h, i, j, k, p);
                    or_collider_velocity(collider, p, cv);
                    v[IX(i, j, k, nx, ny + 1)] = cv[1];
                }
                if (j + 1 < ny && !COLL(i, j + 1, k)) {
                    face_pos(org[1], h, i, j + 1, k, p);
                    or_collider_velocity(collider, p, cv);
                    v[IX(i, j + 1, k, nx, ny + 1)] = cv[1];
                }
                if (k > 0 && !COLL(i, j, k - 1)) {
                    face_pos(org[2], h, i, j, k, p);
                    or_collider_velocity(collider, p, cv);
                    w[IX(i, j, k, nx, ny)] = cv[2];
                }
                if (k + 1 < nz && !COLL(i, j, k + 1)) {
                    face_pos(org[2], h, i, j, k + 1, p);
                    or_collider_velocity(collider, p, cv);
                    w[IX(i, j, k + 1, nx, ny)] = cv[2];
                }
#undef COLL
            }
}

/* ============================================= Compressed systems ====== */

/* FdmCompressedBlas3::mvm -- src/jet/fdm_linear_system3.cpp:160-179. */
void or_csr_mvm(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                const double* v, double* result) {
    for (size_t i = 0; i < rows; ++i) {
        double sum = 0.0;
        for (size_t jj = rp[i]; jj < rp[i + 1]; ++jj) sum += val[jj] * v[ci[jj]];
        result[i] = sum;
    }
}

/* FdmCompressedBlas3::residual -- src/jet/fdm_linear_system3.cpp:181-199. */
void or_csr_residual(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                     const double* x, const double* b, double* result) {
    for (size_t i = 0; i < rows; ++i) {
        double sum = 0.0;
        for (size_t jj = rp[i]; jj < rp[i + 1]; ++jj) sum += val[jj] * x[ci[jj]];
        result[i] = b[i] - sum;
    }
}

static double vdot(const double* a, const double* b, size_t n) {
    double s = 0.0; /* VectorN::dot (include/jet/detail/vector_n-inl.h:361-372), one chunk */
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void vaxpy(double a, const double* x, const double* y, double* r, size_t n) {
    for (size_t i = 0; i < n; ++i) r[i] = a * x[i] + y[i]; /* fdm_linear_system3.cpp:155-158 */
}

/* FdmCgSolver3::solveCompressed -- src/jet/fdm_cg_solver3.cpp:49-74 -- with
 * cg / pcg of include/jet/detail/cg-inl.h:15-94 (NullCgPreconditioner: s = r). */
int or_csr_cg_solve(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                    const double* b, double* x, unsigned max_iter, double tol,
                    unsigned* last_iter, double* last_residual) {
    const size_t n = rows;
    double* buf = (double*)calloc(4 * (n ? n : 1), sizeof(double));
    if (!buf) return -1;
    double *r = buf, *d = buf + n, *q = buf + 2 * n, *s = buf + 3 * n;
    for (size_t i = 0; i < n; ++i) x[i] = 0.0;
    or_csr_residual(val, rp, ci, n, x, b, r);
    memcpy(d, r, n * sizeof(double));
    double sigma_new = vdot(r, d, n);
    unsigned iter = 0;
    int trigger = 0;
    while (sigma_new > tol * tol && iter < max_iter) {
        or_csr_mvm(val, rp, ci, n, d, q);
        const double alpha = sigma_new / vdot(d, q, n);
        vaxpy(alpha, d, x, x, n);
        if (trigger || (iter % 50 == 0 && iter > 0)) {
            or_csr_residual(val, rp, ci, n, x, b, r);
            trigger = 0;
        } else {
            vaxpy(-alpha, q, r, r, n);
        }
        memcpy(s, r, n * sizeof(double));
        const double sigma_old = sigma_new;
        sigma_new = vdot(r, s, n);
        if (sigma_new > sigma_old) trigger = 1;
        const double beta = sigma_new / sigma_old;
        vaxpy(beta, d, s, d, n);
        ++iter;
    }
    *last_iter = iter;
    *last_residual = sqrt(fabs(sigma_new));
    free(buf);
    return (*last_residual <= tol) || (*last_iter < max_iter);
}

/* buildSingleSystem, compressed -- src/jet/grid_single_phase_pressure_solver3.cpp:81-171:
 * fluid cells numbered in x-fastest order; a row holds the centre (the sum of
 * 1/h^2 over non-boundary neighbours, in the order +x, -x, +y, -y, +z, -z)
 * and -1/h^2 for each fluid neighbour, sorted by column as MatrixCsr::addRow
 * does (include/jet/detail/matrix_csr-inl.h:253-275) -- for x-fastest
 * numbering that is -z, -y, -x, centre, +x, +y, +z. */
size_t or_sp_build_compressed(const char* m, const double* u, const double* v, const double* w,
                              size_t nx, size_t ny, size_t nz, const double h[3], double* val,
                              size_t* rp, size_t* ci, double* b, size_t* nnz_out) {
    const double ih[3] = {1.0 / h[0], 1.0 / h[1], 1.0 / h[2]};
    const double ih2[3] = {ih[0] * ih[0], ih[1] * ih[1], ih[2] * ih[2]};
    const size_t n = nx * ny * nz;
    size_t* idx = (size_t*)malloc((n ? n : 1) * sizeof(size_t));
    size_t rows = 0, nnz = 0;
    for (size_t c = 0; c < n; ++c)
        if (m[c] == 0 /* fluid */) idx[c] = rows++;
    if (rp) rp[0] = 0;
    size_t row = 0;
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t c = IX(i, j, k, nx, ny);
                if (m[c] != 0) continue;
                double centre = 0.0;
                /* neighbour (cell, axis, exists), in Jet's order */
                const long nb[6] = {i + 1 < nx ? (long)(c + 1) : -1, i > 0 ? (long)(c - 1) : -1,
                                    j + 1 < ny ? (long)(c + nx) : -1, j > 0 ? (long)(c - nx) : -1,
                                    k + 1 < nz ? (long)(c + nx * ny) : -1,
                                    k > 0 ? (long)(c - nx * ny) : -1};
                int fluid[6] = {0, 0, 0, 0, 0, 0};
                for (int q = 0; q < 6; ++q) {
                    if (nb[q] < 0 || m[nb[q]] == 2 /* boundary */) continue;
                    centre += ih2[q / 2];
                    fluid[q] = m[nb[q]] == 0;
                }
                /* ascending column order: -z, -y, -x, centre, +x, +y, +z */
                const int order[6] = {5, 3, 1, 0, 2, 4};
                size_t e = nnz;
                for (int o = 0; o < 3; ++o) {
                    const int q = order[o];
                    if (!fluid[q]) continue;
                    if (val) {
                        val[e] = -ih2[q / 2];
                        ci[e] = idx[nb[q]];
                    }
                    ++e;
                }
                if (val) {
                    val[e] = centre;
                    ci[e] = idx[c];
                }
                ++e;
                for (int o = 3; o < 6; ++o) {
                    const int q = order[o];
                    if (!fluid[q]) continue;
                    if (val) {
                        val[e] = -ih2[q / 2];
                        ci[e] = idx[nb[q]];
                    }
                    ++e;
                }
                nnz = e;
                if (b) b[row] = or_divergence(u, v, w, nx, ny, nz, h, i, j, k);
                ++row;
                if (rp) rp[row] = nnz;
            }
    free(idx);
    if (nnz_out) *nnz_out = nnz;
    return rows;
}

/* ========================================= Backward Euler diffusion ====== */

/* GridBackwardEulerDiffusionSolver3::buildMarkers / buildMatrix /
 * buildVectors -- src/jet/grid_backward_euler_diffusion_solver3.cpp:175-330
 * (markers 0 fluid, 1 air, 2 boundary, :16-18). */
void or_diffusion_build(const double* f, const double* bsdf, const double* fsdf, size_t sx,
                        size_t sy, size_t sz, const double c[3], int dirichlet, or_row* A,
                        double* b) {
    const size_t n = sx * sy * sz;
    char* m = (char*)malloc(n ? n : 1);
    for (size_t q = 0; q < n; ++q) m[q] = bsdf[q] < 0 ? 2 : (fsdf[q] < 0 ? 0 : 1);
    const size_t ox = 1, oy = sx, oz = sx * sy;
    for (size_t k = 0; k < sz; ++k)
        for (size_t j = 0; j < sy; ++j)
            for (size_t i = 0; i < sx; ++i) {
                const size_t q = IX(i, j, k, sx, sy);
                or_row row = {1.0, 0.0, 0.0, 0.0};
                if (m[q] == 0) {
                    /* a neighbour counts on the diagonal if it is fluid, or
                     * (Dirichlet) anything but air; fluid ones couple */
#define COUNTS(qq) ((dirichlet && m[qq] != 1) || m[qq] == 0)
                    if (i + 1 < sx) {
                        if (COUNTS(q + ox)) row.center += c[0];
                        if (m[q + ox] == 0) row.right -= c[0];
                    }
                    if (i > 0 && COUNTS(q - ox)) row.center += c[0];
                    if (j + 1 < sy) {
                        if (COUNTS(q + oy)) row.center += c[1];
                        if (m[q + oy] == 0) row.up -= c[1];
                    }
                    if (j > 0 && COUNTS(q - oy)) row.center += c[1];
                    if (k + 1 < sz) {
                        if (COUNTS(q + oz)) row.center += c[2];
                        if (m[q + oz] == 0) row.front -= c[2];
                    }
                    if (k > 0 && COUNTS(q - oz)) row.center += c[2];
#undef COUNTS
                }
                A[q] = row;
                double bv = f[q];
                if (dirichlet && m[q] == 0) {
                    if (i + 1 < sx && m[q + ox] == 2) bv += c[0] * f[q + ox];
                    if (i > 0 && m[q - ox] == 2) bv += c[0] * f[q - ox];
                    if (j + 1 < sy && m[q + oy] == 2) bv += c[1] * f[q + oy];
                    if (j > 0 && m[q - oy] == 2) bv += c[1] * f[q - oy];
                    if (k + 1 < sz && m[q + oz] == 2) bv += c[2] * f[q + oz];
                    if (k > 0 && m[q - oz] == 2) bv += c[2] * f[q - oz];
                }
                b[q] = bv;
            }
    free(m);
}
