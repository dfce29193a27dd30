// ref_shim.cpp -- C entry points over Jet's own compiled linear solvers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target "ref") against
// the reference sources where they lie, into oracle/_ref/libjetref.so.  It is
// used to pin the C restatement in jet_oracle.c (bitwise for BLAS and
// transfer operators, iteration-for-iteration for the solvers) and, on the
// benchmark host, as the "reference" CPU baseline.  Nothing here is product
// code and nothing is copied from the reference: this file only calls Jet's
// public API (include/jet/*.h).
#include <jet/array_utils.h>
#include <jet/cell_centered_scalar_grid3.h>
#include <jet/face_centered_grid3.h>
#include <jet/custom_scalar_field3.h>
#include <jet/fdm_cg_solver3.h>
#include <jet/grid_backward_euler_diffusion_solver3.h>
#include <jet/rigid_body_collider3.h>
#include <jet/sphere3.h>
#include <physics_helpers.h>
#include <jet/fdm_gauss_seidel_solver3.h>
#include <jet/fdm_jacobi_solver3.h>
#include <jet/fdm_linear_system3.h>
#include <jet/fdm_mg_linear_system3.h>
#include <jet/fdm_mg_solver3.h>
#include <jet/fdm_mgpcg_solver3.h>
#include <jet/fdm_mgpcg_solver2.h>
#include <jet/custom_scalar_field2.h>
#include <jet/custom_vector_field2.h>
#include <jet/face_centered_grid2.h>
#include <jet/grid_fractional_single_phase_pressure_solver2.h>
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

// ------------------------------------------------- 2-D multigrid
// FdmMgUtils2 and FdmMgpcgSolver2 on (ny, nx) arrays; rows arrive as the
// 4-coefficient rows of a one-plane system and their front is dropped.

static void toVector2(const double* src, size_t nx, size_t ny, FdmVector2* v) {
    v->resize(Size2(nx, ny));
    v->forEachIndex([&](size_t i, size_t j) { (*v)(i, j) = src[i + nx * j]; });
}

static void fromVector2(const FdmVector2& v, double* dst) {
    const size_t nx = v.size().x;
    v.forEachIndex([&](size_t i, size_t j) { dst[i + nx * j] = v(i, j); });
}

void ref_mg_restrict2(const double* finer, double* coarser, size_t cnx, size_t cny) {
    FdmVector2 f, c;
    toVector2(finer, 2 * cnx, 2 * cny, &f);
    c.resize(Size2(cnx, cny));
    FdmMgUtils2::restrict(f, &c);
    fromVector2(c, coarser);
}

void ref_mg_correct2(const double* coarser, double* finer, size_t fnx, size_t fny) {
    FdmVector2 f, c;
    toVector2(coarser, fnx / 2, fny / 2, &c);
    toVector2(finer, fnx, fny, &f);
    FdmMgUtils2::correct(c, &f);
    fromVector2(f, finer);
}

size_t ref_levels_with_finest2(size_t nx, size_t ny, size_t max_levels, size_t* lnx,
                               size_t* lny) {
    std::vector<Array2<double>> levels;
    FdmMgUtils2::resizeArrayWithFinest(Size2(nx, ny), max_levels, &levels);
    for (size_t l = 0; l < levels.size(); ++l) {
        lnx[l] = levels[l].size().x;
        lny[l] = levels[l].size().y;
    }
    return levels.size();
}

int ref_mgpcg_solve2(size_t nlev, const size_t* nx, const size_t* ny, const Row* const* A,
                     double* const* x, const double* const* b, unsigned max_cg_iter,
                     size_t max_levels, unsigned restrict_iter, unsigned correct_iter,
                     unsigned coarsest_iter, unsigned final_iter, double max_tol, double sor,
                     int red_black, unsigned* iters, double* residual) {
    FdmMgLinearSystem2 s;
    s.resizeWithCoarsest(Size2(nx[nlev - 1], ny[nlev - 1]), nlev);
    for (size_t l = 0; l < nlev; ++l) {
        FdmMatrix2& m = s.A.levels[l];
        m.forEachIndex([&](size_t i, size_t j) {
            const Row& r = A[l][i + nx[l] * j];
            m(i, j).center = r.center;
            m(i, j).right = r.right;
            m(i, j).up = r.up;
        });
        toVector2(x[l], nx[l], ny[l], &s.x.levels[l]);
        toVector2(b[l], nx[l], ny[l], &s.b.levels[l]);
    }
    FdmMgpcgSolver2 solver(max_cg_iter, max_levels, restrict_iter, correct_iter, coarsest_iter,
                           final_iter, max_tol, sor, red_black != 0);
    bool ok = solver.solve(&s);
    fromVector2(s.x.levels[0], x[0]);
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

// ------------------------------------------------ compressed systems
// FdmCompressedBlas3 and FdmCgSolver3::solveCompressed on a MatrixCsrD built
// row by row with MatrixCsr::addRow from CSR arrays.

static void toCsr(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                  MatrixCsrD* m) {
    m->clear();
    for (size_t i = 0; i < rows; ++i) {
        std::vector<double> nz(val + rp[i], val + rp[i + 1]);
        std::vector<size_t> cols(ci + rp[i], ci + rp[i + 1]);
        m->addRow(nz, cols);
    }
}

static void toVectorN(const double* src, size_t n, VectorND* v) {
    v->resize(n);
    for (size_t i = 0; i < n; ++i) (*v)[i] = src[i];
}

void ref_csr_mvm(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                 const double* v, double* out) {
    MatrixCsrD m;
    toCsr(val, rp, ci, rows, &m);
    VectorND vv, r(rows);
    toVectorN(v, rows, &vv);
    FdmCompressedBlas3::mvm(m, vv, &r);
    for (size_t i = 0; i < rows; ++i) out[i] = r[i];
}

void ref_csr_residual(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                      const double* x, const double* b, double* out) {
    MatrixCsrD m;
    toCsr(val, rp, ci, rows, &m);
    VectorND xx, bb, r(rows);
    toVectorN(x, rows, &xx);
    toVectorN(b, rows, &bb);
    FdmCompressedBlas3::residual(m, xx, bb, &r);
    for (size_t i = 0; i < rows; ++i) out[i] = r[i];
}

int ref_csr_cg_solve(const double* val, const size_t* rp, const size_t* ci, size_t rows,
                     const double* b, double* x, unsigned max_iter, double tol, unsigned* iters,
                     double* residual) {
    FdmCompressedLinearSystem3 sys;
    toCsr(val, rp, ci, rows, &sys.A);
    toVectorN(b, rows, &sys.b);
    sys.x.resize(rows, 0.0);
    FdmCgSolver3 solver(max_iter, tol);
    const bool ok = solver.solveCompressed(&sys);
    for (size_t i = 0; i < rows; ++i) x[i] = sys.x[i];
    *iters = solver.lastNumberOfIterations();
    *residual = solver.lastResidual();
    return ok ? 1 : 0;
}

static void fillCells(CellCenteredScalarGrid3* g, const double* d);

// ------------------------------------------------ backward Euler diffusion
// GridBackwardEulerDiffusionSolver3 on a cell-centred grid whose data is f,
// with the SDFs given per data point (a CustomScalarField3 that looks the
// point up), and a linear solver that records the system it is handed.

namespace {
class CaptureSolver final : public FdmLinearSystemSolver3 {
 public:
    Row* A = nullptr;
    double* b = nullptr;
    bool solve(FdmLinearSystem3* s) override {
        const Size3 n = s->A.size();
        for (size_t k = 0; k < n.z; ++k)
            for (size_t j = 0; j < n.y; ++j)
                for (size_t i = 0; i < n.x; ++i) {
                    const size_t q = i + n.x * (j + n.y * k);
                    A[q] = {s->A(i, j, k).center, s->A(i, j, k).right, s->A(i, j, k).up,
                            s->A(i, j, k).front};
                    b[q] = s->b(i, j, k);
                }
        return true;
    }
};
}  // namespace

void ref_diffusion_system(const double* f, const double* bsdf, const double* fsdf, size_t sx,
                          size_t sy, size_t sz, const double* h, double dt, double mu,
                          int dirichlet, double* A, double* b) {
    CellCenteredScalarGrid3 src(Size3(sx, sy, sz), Vector3D(h[0], h[1], h[2]), Vector3D());
    fillCells(&src, f);
    CellCenteredScalarGrid3 dst(src);
    const Vector3D o = src.dataOrigin();
    auto lookup = [&](const double* a) {
        return [=](const Vector3D& p) {
            const size_t i = (size_t)std::llround((p.x - o.x) / h[0]);
            const size_t j = (size_t)std::llround((p.y - o.y) / h[1]);
            const size_t k = (size_t)std::llround((p.z - o.z) / h[2]);
            return a[i + sx * (j + sy * k)];
        };
    };
    CustomScalarField3 bf(lookup(bsdf)), ff(lookup(fsdf));
    GridBackwardEulerDiffusionSolver3 solver(
        dirichlet ? GridBackwardEulerDiffusionSolver3::Dirichlet
                  : GridBackwardEulerDiffusionSolver3::Neumann);
    auto cap = std::make_shared<CaptureSolver>();
    cap->A = reinterpret_cast<Row*>(A);
    cap->b = b;
    solver.setLinearSystemSolver(cap);
    solver.solve(src, mu, dt, &dst, bf, ff);
}

// Jet's own diffusion solve (its default ICCG linear solver) of the same
// grid: dest = the solution.
void ref_diffusion_solve(const double* f, const double* bsdf, const double* fsdf, size_t sx,
                         size_t sy, size_t sz, const double* h, double dt, double mu,
                         int dirichlet, double* dest) {
    CellCenteredScalarGrid3 src(Size3(sx, sy, sz), Vector3D(h[0], h[1], h[2]), Vector3D());
    fillCells(&src, f);
    CellCenteredScalarGrid3 dst(src);
    const Vector3D o = src.dataOrigin();
    auto lookup = [&](const double* a) {
        return [=](const Vector3D& p) {
            const size_t i = (size_t)std::llround((p.x - o.x) / h[0]);
            const size_t j = (size_t)std::llround((p.y - o.y) / h[1]);
            const size_t k = (size_t)std::llround((p.z - o.z) / h[2]);
            return a[i + sx * (j + sy * k)];
        };
    };
    CustomScalarField3 bf(lookup(bsdf)), ff(lookup(fsdf));
    GridBackwardEulerDiffusionSolver3 solver(
        dirichlet ? GridBackwardEulerDiffusionSolver3::Dirichlet
                  : GridBackwardEulerDiffusionSolver3::Neumann);
    solver.solve(src, mu, dt, &dst, bf, ff);
    dst.forEachDataPointIndex([&](size_t i, size_t j, size_t k) {
        dest[i + sx * (j + sy * k)] = dst(i, j, k);
    });
}

// ------------------------------------------------ boundary conditions
// Jet's grids, sampler, gradient, extrapolation, friction helper and rigid
// collider.  GridFractionalBoundaryConditionSolver3 itself cannot be linked
// (its onColliderUpdated constructs SurfaceToImplicit3, which needs
// TriangleMesh3 and the absent tinyobj sources), so ref_constrain_velocity
// drives these compiled parts in the order its constrainVelocity
// (src/jet/grid_fractional_boundary_condition_solver3.cpp:26-232) and the
// blocked variant (src/jet/grid_blocked_boundary_condition_solver3.cpp:23-69)
// call them, with the collider SDF given at the cell centres.

struct RefCollider {
    double lin[3], ang[3], trans[3], friction;
};

static std::shared_ptr<RigidBodyCollider3> makeCollider(const RefCollider* c) {
    auto surf = std::make_shared<Sphere3>(Vector3D(), 1.0);
    surf->transform = Transform3(Vector3D(c->trans[0], c->trans[1], c->trans[2]), QuaternionD());
    auto col = std::make_shared<RigidBodyCollider3>(surf, Vector3D(c->lin[0], c->lin[1], c->lin[2]),
                                                    Vector3D(c->ang[0], c->ang[1], c->ang[2]));
    col->setFrictionCoefficient(c->friction);
    return col;
}

static void fillCells(CellCenteredScalarGrid3* g, const double* d) {
    const Size3 s = g->dataSize();
    g->forEachDataPointIndex([&](size_t i, size_t j, size_t k) {
        (*g)(i, j, k) = d[i + s.x * (j + s.y * k)];
    });
}

static void fillFaces(FaceCenteredGrid3* f, const double* u, const double* v, const double* w) {
    const Size3 su = f->uSize(), sv = f->vSize(), sw = f->wSize();
    f->forEachUIndex([&](size_t i, size_t j, size_t k) { f->u(i, j, k) = u[i + su.x * (j + su.y * k)]; });
    f->forEachVIndex([&](size_t i, size_t j, size_t k) { f->v(i, j, k) = v[i + sv.x * (j + sv.y * k)]; });
    f->forEachWIndex([&](size_t i, size_t j, size_t k) { f->w(i, j, k) = w[i + sw.x * (j + sw.y * k)]; });
}

static void readFaces(const FaceCenteredGrid3& f, double* u, double* v, double* w) {
    const Size3 su = f.uSize(), sv = f.vSize(), sw = f.wSize();
    f.forEachUIndex([&](size_t i, size_t j, size_t k) { u[i + su.x * (j + su.y * k)] = f.u(i, j, k); });
    f.forEachVIndex([&](size_t i, size_t j, size_t k) { v[i + sv.x * (j + sv.y * k)] = f.v(i, j, k); });
    f.forEachWIndex([&](size_t i, size_t j, size_t k) { w[i + sw.x * (j + sw.y * k)] = f.w(i, j, k); });
}

double ref_sample_cell(const double* data, size_t nx, size_t ny, size_t nz, const double* h,
                       const double* o, const double* p) {
    CellCenteredScalarGrid3 g(Size3(nx, ny, nz), Vector3D(h[0], h[1], h[2]), Vector3D(o[0], o[1], o[2]));
    fillCells(&g, data);
    return g.sample(Vector3D(p[0], p[1], p[2]));
}

void ref_gradient_cell(const double* data, size_t nx, size_t ny, size_t nz, const double* h,
                       const double* o, const double* p, double* out) {
    CellCenteredScalarGrid3 g(Size3(nx, ny, nz), Vector3D(h[0], h[1], h[2]), Vector3D(o[0], o[1], o[2]));
    fillCells(&g, data);
    const Vector3D r = g.gradient(Vector3D(p[0], p[1], p[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void ref_extrapolate_to_region(const double* input, const char* valid, size_t sx, size_t sy,
                               size_t sz, unsigned iterations, double* output) {
    Array3<double> in(sx, sy, sz), out(sx, sy, sz);
    Array3<char> va(sx, sy, sz);
    std::memcpy(in.data(), input, sizeof(double) * sx * sy * sz);
    std::memcpy(va.data(), valid, sx * sy * sz);
    extrapolateToRegion(in.constAccessor(), va.constAccessor(), iterations, out.accessor());
    std::memcpy(output, out.data(), sizeof(double) * sx * sy * sz);
}

void ref_project_and_apply_friction(const double* vel, const double* n, double friction,
                                    double* out) {
    const Vector3D r = projectAndApplyFriction(Vector3D(vel[0], vel[1], vel[2]),
                                               Vector3D(n[0], n[1], n[2]), friction);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void ref_collider_velocity(const RefCollider* c, const double* p, double* out) {
    const Vector3D r = makeCollider(c)->velocityAt(Vector3D(p[0], p[1], p[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void ref_constrain_velocity(int blocked, size_t nx, size_t ny, size_t nz, const double* hh,
                            const double* oo, double* uo, double* vo, double* wo,
                            const double* sdfd, const RefCollider* c, unsigned depth,
                            int closed) {
    const Size3 size(nx, ny, nz);
    const Vector3D h(hh[0], hh[1], hh[2]), origin(oo[0], oo[1], oo[2]);
    FaceCenteredGrid3 velocity(size, h, origin);
    fillFaces(&velocity, uo, vo, wo);
    CellCenteredScalarGrid3 sdf(size, h, origin);
    fillCells(&sdf, sdfd);
    RefCollider zero = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, 0};
    auto collider = makeCollider(c ? c : &zero);

    auto u = velocity.uAccessor();
    auto v = velocity.vAccessor();
    auto w = velocity.wAccessor();
    auto uPos = velocity.uPosition();
    auto vPos = velocity.vPosition();
    auto wPos = velocity.wPosition();
    Array3<double> uTemp(u.size()), vTemp(v.size()), wTemp(w.size());
    Array3<char> uMarker(u.size(), 1), vMarker(v.size(), 1), wMarker(w.size(), 1);

    auto init = [&](ArrayAccessor3<double> a, Array3<char>& m, const VectorGrid3::DataPositionFunc& pos,
                    int axis) {
        Vector3D off;
        off[axis] = 0.5 * h[axis];
        a.forEachIndex([&](size_t i, size_t j, size_t k) {
            const Vector3D pt = pos(i, j, k);
            const double phi0 = sdf.sample(pt - off);
            const double phi1 = sdf.sample(pt + off);
            double frac = fractionInsideSdf(phi0, phi1);
            frac = 1.0 - clamp(frac, 0.0, 1.0);
            if (frac > 0.0) {
                m(i, j, k) = 1;
            } else {
                a(i, j, k) = collider->velocityAt(pt)[axis];
                m(i, j, k) = 0;
            }
        });
    };
    init(u, uMarker, uPos, 0);
    init(v, vMarker, vPos, 1);
    init(w, wMarker, wPos, 2);
    extrapolateToRegion(velocity.uConstAccessor(), uMarker.constAccessor(), depth, u);
    extrapolateToRegion(velocity.vConstAccessor(), vMarker.constAccessor(), depth, v);
    extrapolateToRegion(velocity.wConstAccessor(), wMarker.constAccessor(), depth, w);

    auto project = [&](ConstArrayAccessor3<double> a, Array3<double>& t,
                       const VectorGrid3::DataPositionFunc& pos, int axis) {
        a.forEachIndex([&](size_t i, size_t j, size_t k) {
            const Vector3D pt = pos(i, j, k);
            if (isInsideSdf(sdf.sample(pt))) {
                const Vector3D colliderVel = collider->velocityAt(pt);
                const Vector3D vel = velocity.sample(pt);
                const Vector3D g = sdf.gradient(pt);
                if (g.lengthSquared() > 0.0) {
                    const Vector3D n = g.normalized();
                    const Vector3D velr = vel - colliderVel;
                    const Vector3D velt =
                        projectAndApplyFriction(velr, n, collider->frictionCoefficient());
                    t(i, j, k) = (velt + colliderVel)[axis];
                } else {
                    t(i, j, k) = colliderVel[axis];
                }
            } else {
                t(i, j, k) = a(i, j, k);
            }
        });
    };
    project(velocity.uConstAccessor(), uTemp, uPos, 0);
    project(velocity.vConstAccessor(), vTemp, vPos, 1);
    project(velocity.wConstAccessor(), wTemp, wPos, 2);
    u.forEachIndex([&](size_t i, size_t j, size_t k) { u(i, j, k) = uTemp(i, j, k); });
    v.forEachIndex([&](size_t i, size_t j, size_t k) { v(i, j, k) = vTemp(i, j, k); });
    w.forEachIndex([&](size_t i, size_t j, size_t k) { w(i, j, k) = wTemp(i, j, k); });

    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j) {
            if (closed & kDirectionLeft) u(0, j, k) = 0;
            if (closed & kDirectionRight) u(nx, j, k) = 0;
        }
    for (size_t k = 0; k < nz; ++k)
        for (size_t i = 0; i < nx; ++i) {
            if (closed & kDirectionDown) v(i, 0, k) = 0;
            if (closed & kDirectionUp) v(i, ny, k) = 0;
        }
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            if (closed & kDirectionBack) w(i, j, 0) = 0;
            if (closed & kDirectionFront) w(i, j, nz) = 0;
        }

    if (blocked) {
        Array3<char> marker(size);
        marker.forEachIndex([&](size_t i, size_t j, size_t k) {
            marker(i, j, k) = isInsideSdf(sdf(i, j, k)) ? 0 : 1;
        });
        marker.forEachIndex([&](size_t i, size_t j, size_t k) {
            if (marker(i, j, k) != 0) return;
            if (i > 0 && marker(i - 1, j, k) == 1) u(i, j, k) = collider->velocityAt(uPos(i, j, k)).x;
            if (i < nx - 1 && marker(i + 1, j, k) == 1)
                u(i + 1, j, k) = collider->velocityAt(uPos(i + 1, j, k)).x;
            if (j > 0 && marker(i, j - 1, k) == 1) v(i, j, k) = collider->velocityAt(vPos(i, j, k)).y;
            if (j < ny - 1 && marker(i, j + 1, k) == 1)
                v(i, j + 1, k) = collider->velocityAt(vPos(i, j + 1, k)).y;
            if (k > 0 && marker(i, j, k - 1) == 1) w(i, j, k) = collider->velocityAt(wPos(i, j, k)).z;
            if (k < nz - 1 && marker(i, j, k + 1) == 1)
                w(i, j, k + 1) = collider->velocityAt(wPos(i, j, k + 1)).z;
        });
    }
    readFaces(velocity, uo, vo, wo);
}

}  // extern "C"

// GridFractionalSinglePhasePressureSolver2::solve on an nx x ny grid of
// spacing h and origin o: the variational system Jet's buildWeights +
// buildSystem hand its multigrid solver (A of every level as {center, right,
// up, 0} rows, levels back to back, and b of the finest level), captured by a
// recording FdmMgSolver2 that answers with the given pressure p; then Jet's
// applyPressureGradient with that p (uo, vo).  The boundary SDF is looked up
// at the cell centres of the grid plus one ghost cell per side (bsdf,
// (nx + 2) x (ny + 2), which is every point buildWeights samples), the fluid
// SDF at the cell centres, the collider velocity x at the u faces (bu) and y
// at the v faces (bv; NULL: at rest).
namespace {
class CaptureMg2 final : public FdmMgSolver2 {
 public:
    explicit CaptureMg2(size_t levels) : FdmMgSolver2(levels) {}
    Row* A = nullptr;
    double* b = nullptr;
    const double* p = nullptr;
    bool solve(FdmMgLinearSystem2* s) override {
        Row* out = A;
        for (size_t l = 0; l < s->A.levels.size(); ++l) {
            const FdmMatrix2& M = s->A.levels[l];
            const Size2 n = M.size();
            for (size_t j = 0; j < n.y; ++j)
                for (size_t i = 0; i < n.x; ++i)
                    *out++ = {M(i, j).center, M(i, j).right, M(i, j).up, 0.0};
        }
        const Size2 n = s->b.levels[0].size();
        for (size_t j = 0; j < n.y; ++j)
            for (size_t i = 0; i < n.x; ++i) {
                b[i + n.x * j] = s->b.levels[0](i, j);
                s->x.levels[0](i, j) = p[i + n.x * j];
            }
        return true;
    }
};
}  // namespace

extern "C" {

void ref_fr_solve2(size_t nx, size_t ny, const double* h, const double* o, const double* u,
                   const double* v, const double* bsdf, const double* fsdf, const double* bu,
                   const double* bv, size_t max_levels, const double* p, double* A, double* b,
                   double* uo, double* vo) {
    FaceCenteredGrid2 vel(Size2(nx, ny), Vector2D(h[0], h[1]), Vector2D(o[0], o[1]));
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i <= nx; ++i) vel.u(i, j) = u[i + (nx + 1) * j];
    for (size_t j = 0; j <= ny; ++j)
        for (size_t i = 0; i < nx; ++i) vel.v(i, j) = v[i + nx * j];
    // cell index of a cell-centre position (ghost cells at -1 and n)
    auto cell = [=](double x, double org, double hh) {
        return (long)std::llround((x - org) / hh - 0.5);
    };
    CustomScalarField2 boundary([=](const Vector2D& q) {
        const long i = cell(q.x, o[0], h[0]), j = cell(q.y, o[1], h[1]);
        return bsdf[(size_t)(i + 1) + (nx + 2) * (size_t)(j + 1)];
    });
    CustomScalarField2 fluid([=](const Vector2D& q) {
        const long i = cell(q.x, o[0], h[0]), j = cell(q.y, o[1], h[1]);
        return fsdf[(size_t)i + nx * (size_t)j];
    });
    // a u face sits at (i h, (j + 1/2) h), a v face at ((i + 1/2) h, j h)
    CustomVectorField2 collider([=](const Vector2D& q) {
        const double fx = (q.x - o[0]) / h[0], fy = (q.y - o[1]) / h[1];
        const long ix = std::llround(fx), iy = std::llround(fy);
        if (std::fabs(fx - (double)ix) < 0.25) {  // a u face
            const long j = std::llround(fy - 0.5);
            return Vector2D(bu ? bu[(size_t)ix + (nx + 1) * (size_t)j] : 0.0, 0.0);
        }
        const long i = std::llround(fx - 0.5);
        return Vector2D(0.0, bv ? bv[(size_t)i + nx * (size_t)iy] : 0.0);
    });
    auto cap = std::make_shared<CaptureMg2>(max_levels);
    cap->A = reinterpret_cast<Row*>(A);
    cap->b = b;
    cap->p = p;
    GridFractionalSinglePhasePressureSolver2 solver;
    solver.setLinearSystemSolver(cap);
    FaceCenteredGrid2 out(vel);
    solver.solve(vel, 1.0, &out, boundary, collider, fluid, false);
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i <= nx; ++i) uo[i + (nx + 1) * j] = out.u(i, j);
    for (size_t j = 0; j <= ny; ++j)
        for (size_t i = 0; i < nx; ++i) vo[i + nx * j] = out.v(i, j);
}

}  // extern "C"
