// ref_shim.cpp -- C entry points over Jet's own compiled linear solvers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target "ref") against
// the reference sources where they lie, into oracle/_ref/libjetref.so.  It is
// used to pin the C restatement in jet_oracle.c (bitwise for BLAS and
// transfer operators, iteration-for-iteration for the solvers) and, on the
// benchmark host, as the "reference" CPU baseline.  Nothing here is product
// code and nothing is copied from the reference: this file only calls Jet's
// public API (include/jet/*.h).
#include <jet/fdm_cg_solver3.h>
#include <jet/fdm_gauss_seidel_solver3.h>
#include <jet/fdm_jacobi_solver3.h>
#include <jet/fdm_linear_system3.h>
#include <jet/fdm_mg_linear_system3.h>
#include <jet/fdm_mg_solver3.h>
#include <jet/fdm_mgpcg_solver3.h>
#include <jet/level_set_utils.h>
#include <jet/logging.h>

#include <cstring>

using namespace jet;

namespace {

struct Row {
    double center, right, up, front;
};

void toMatrix(const Row* src, size_t nx, size_t ny, size_t nz, FdmMatrix3* m) {
    m->resize(Size3(nx, ny, nz));
    m->parallelForEachIndex([&](size_t i, size_t j, size_t k) {
        const Row& r = src[i + nx * (j + ny * k)];
        FdmMatrixRow3& d = (*m)(i, j, k);
        d.center = r.center;
        d.right = r.right;
        d.up = r.up;
        d.front = r.front;
    });
}

void toVector(const double* src, size_t nx, size_t ny, size_t nz,
              FdmVector3* v) {
    v->resize(Size3(nx, ny, nz));
    std::memcpy(v->data(), src, sizeof(double) * nx * ny * nz);
}

void fromVector(const FdmVector3& v, double* dst) {
    std::memcpy(dst, v.data(), sizeof(double) * v.size().x * v.size().y *
                                   v.size().z);
}

struct MgIn {
    size_t nlev;
    const size_t* nx;
    const size_t* ny;
    const size_t* nz;
};

void toMgSystem(size_t nlev, const size_t* nx, const size_t* ny,
                const size_t* nz, const Row* const* A, const double* const* x,
                const double* const* b, FdmMgLinearSystem3* s) {
    s->A.levels.resize(nlev);
    s->x.levels.resize(nlev);
    s->b.levels.resize(nlev);
    for (size_t l = 0; l < nlev; ++l) {
        toMatrix(A[l], nx[l], ny[l], nz[l], &s->A.levels[l]);
        toVector(x[l], nx[l], ny[l], nz[l], &s->x.levels[l]);
        toVector(b[l], nx[l], ny[l], nz[l], &s->b.levels[l]);
    }
}

}  // namespace

extern "C" {

void ref_quiet(void) { Logging::mute(); }

void ref_mvm(const Row* A, const double* v, double* out, size_t nx, size_t ny,
             size_t nz) {
    FdmMatrix3 m;
    FdmVector3 x, y;
    toMatrix(A, nx, ny, nz, &m);
    toVector(v, nx, ny, nz, &x);
    y.resize(Size3(nx, ny, nz));
    FdmBlas3::mvm(m, x, &y);
    fromVector(y, out);
}

void ref_residual(const Row* A, const double* x, const double* b, double* out,
                  size_t nx, size_t ny, size_t nz) {
    FdmMatrix3 m;
    FdmVector3 xv, bv, r;
    toMatrix(A, nx, ny, nz, &m);
    toVector(x, nx, ny, nz, &xv);
    toVector(b, nx, ny, nz, &bv);
    r.resize(Size3(nx, ny, nz));
    FdmBlas3::residual(m, xv, bv, &r);
    fromVector(r, out);
}

double ref_dot(const double* a, const double* b, size_t nx, size_t ny,
               size_t nz) {
    FdmVector3 av, bv;
    toVector(a, nx, ny, nz, &av);
    toVector(b, nx, ny, nz, &bv);
    return FdmBlas3::dot(av, bv);
}

double ref_linfnorm(const double* a, size_t nx, size_t ny, size_t nz) {
    FdmVector3 av;
    toVector(a, nx, ny, nz, &av);
    return FdmBlas3::lInfNorm(av);
}

void ref_axpy(double a, const double* x, const double* y, double* out,
              size_t nx, size_t ny, size_t nz) {
    FdmVector3 xv, yv, r;
    toVector(x, nx, ny, nz, &xv);
    toVector(y, nx, ny, nz, &yv);
    r.resize(Size3(nx, ny, nz));
    FdmBlas3::axpy(a, xv, yv, &r);
    fromVector(r, out);
}

int ref_cg_solve(const Row* A, const double* b, double* x, size_t nx,
                 size_t ny, size_t nz, unsigned max_iter, double tol,
                 unsigned* iters, double* residual) {
    FdmLinearSystem3 sys;
    toMatrix(A, nx, ny, nz, &sys.A);
    toVector(b, nx, ny, nz, &sys.b);
    sys.x.resize(Size3(nx, ny, nz));
    FdmCgSolver3 solver(max_iter, tol);
    bool ok = solver.solve(&sys);
    fromVector(sys.x, x);
    *iters = solver.lastNumberOfIterations();
    *residual = solver.lastResidual();
    return ok ? 1 : 0;
}

void ref_gs_relax(const Row* A, const double* b, double sor, double* x,
                  size_t nx, size_t ny, size_t nz, int red_black) {
    FdmMatrix3 m;
    FdmVector3 bv, xv;
    toMatrix(A, nx, ny, nz, &m);
    toVector(b, nx, ny, nz, &bv);
    toVector(x, nx, ny, nz, &xv);
    if (red_black)
        FdmGaussSeidelSolver3::relaxRedBlack(m, bv, sor, &xv);
    else
        FdmGaussSeidelSolver3::relax(m, bv, sor, &xv);
    fromVector(xv, x);
}

void ref_jacobi_relax(const Row* A, const double* b, const double* x,
                      double* xtemp, size_t nx, size_t ny, size_t nz) {
    FdmMatrix3 m;
    FdmVector3 bv, xv, tv;
    toMatrix(A, nx, ny, nz, &m);
    toVector(b, nx, ny, nz, &bv);
    toVector(x, nx, ny, nz, &xv);
    tv.resize(Size3(nx, ny, nz));
    FdmJacobiSolver3::relax(m, bv, &xv, &tv);
    fromVector(tv, xtemp);
}

void ref_mg_restrict(const double* finer, double* coarser, size_t cnx,
                     size_t cny, size_t cnz) {
    FdmVector3 f, c;
    toVector(finer, 2 * cnx, 2 * cny, 2 * cnz, &f);
    c.resize(Size3(cnx, cny, cnz));
    FdmMgUtils3::restrict(f, &c);
    fromVector(c, coarser);
}

void ref_mg_correct(const double* coarser, double* finer, size_t fnx,
                    size_t fny, size_t fnz) {
    FdmVector3 f, c;
    toVector(coarser, fnx / 2, fny / 2, fnz / 2, &c);
    toVector(finer, fnx, fny, fnz, &f);
    FdmMgUtils3::correct(c, &f);
    fromVector(f, finer);
}

size_t ref_levels_with_finest(size_t nx, size_t ny, size_t nz,
                              size_t max_levels, size_t* lnx, size_t* lny,
                              size_t* lnz) {
    std::vector<Array3<double>> levels;
    FdmMgUtils3::resizeArrayWithFinest(Size3(nx, ny, nz), max_levels, &levels);
    for (size_t l = 0; l < levels.size(); ++l) {
        lnx[l] = levels[l].size().x;
        lny[l] = levels[l].size().y;
        lnz[l] = levels[l].size().z;
    }
    return levels.size();
}

// FdmMgSolver3::solve(FdmMgLinearSystem3*): one V-cycle.  x levels are
// updated in place; returns Jet's bool and the last residual norm measured
// by the caller-side residual (as FdmMgSolver3 does not expose it).
int ref_mg_solve(size_t nlev, const size_t* nx, const size_t* ny,
                 const size_t* nz, const Row* const* A, double* const* x,
                 const double* const* b, size_t max_levels,
                 unsigned restrict_iter, unsigned correct_iter,
                 unsigned coarsest_iter, unsigned final_iter, double max_tol,
                 double sor, int red_black) {
    FdmMgLinearSystem3 s;
    toMgSystem(nlev, nx, ny, nz, A, (const double* const*)x, b, &s);
    FdmMgSolver3 solver(max_levels, restrict_iter, correct_iter, coarsest_iter,
                        final_iter, max_tol, sor, red_black != 0);
    bool ok = solver.solve(&s);
    for (size_t l = 0; l < nlev; ++l) fromVector(s.x.levels[l], x[l]);
    return ok ? 1 : 0;
}

int ref_mgpcg_solve(size_t nlev, const size_t* nx, const size_t* ny,
                    const size_t* nz, const Row* const* A, double* const* x,
                    const double* const* b, unsigned max_cg_iter,
                    size_t max_levels, unsigned restrict_iter,
                    unsigned correct_iter, unsigned coarsest_iter,
                    unsigned final_iter, double max_tol, double sor,
                    int red_black, unsigned* iters, double* residual) {
    FdmMgLinearSystem3 s;
    toMgSystem(nlev, nx, ny, nz, A, (const double* const*)x, b, &s);
    FdmMgpcgSolver3 solver(max_cg_iter, max_levels, restrict_iter,
                           correct_iter, coarsest_iter, final_iter, max_tol,
                           sor, red_black != 0);
    bool ok = solver.solve(&s);
    fromVector(s.x.levels[0], x[0]);
    *iters = solver.lastNumberOfIterations();
    *residual = solver.lastResidual();
    return ok ? 1 : 0;
}

double ref_fraction_inside_sdf(double a, double b) {
    return fractionInsideSdf(a, b);
}

double ref_fraction_inside(double bl, double br, double tl, double tr) {
    return fractionInside(bl, br, tl, tr);
}

}  // extern "C"
