// fdm_gpu_solvers3.cpp -- Jet solver classes over the C ABI of jetgpu.h.
//
// Compiled as part of a Jet build (include/jet on the include path, linked
// against libjetgpu.so).  Jet's FdmMatrixRow3 {center, right, up, front} and
// x-fastest Array3 are the library's layout, so systems are handed over as
// plain pointers without conversion.
#include "fdm_gpu_solvers3.h"

#include <jet/fdm_linear_system3.h>
#include <jet/fdm_mg_linear_system3.h>
#include <jet/logging.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace jet {

namespace {

static_assert(sizeof(FdmMatrixRow3) == 4 * sizeof(double),
              "FdmMatrixRow3 must be four packed doubles");

// The library's status -> Jet's error convention.
void check(int status, const jg_solver* s, const char* what) {
    if (status == JG_OK) return;
    std::string msg = std::string(what) + ": " + (s ? jg_last_error(s) : "");
    if (status == JG_ERR_INVALID_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

jg_solver* lazySolver(jg_solver** slot, int device) {
    if (!*slot) {
        const int st = jg_solver_create(device, slot);
        if (st != JG_OK) {
            *slot = nullptr;
            throw std::runtime_error("jetgpu: no usable GPU device " +
                                     std::to_string(device) + " (status " +
                                     std::to_string(st) + ")");
        }
    }
    return *slot;
}

bool sameSize(const Size3& a, const Size3& b) {
    return a.x == b.x && a.y == b.y && a.z == b.z;
}

// Level sizes and row pointers of a multigrid system, checked as Jet's
// solvers would (finest A, x and b of one size).  False: an empty grid.
struct Hierarchy {
    std::vector<size_t> nx, ny, nz;
    std::vector<const double*> rows;
};

bool hierarchyOf(FdmMgLinearSystem3* system, Hierarchy* h) {
    JET_THROW_INVALID_ARG_IF(system == nullptr);
    const size_t L = system->A.levels.size();
    JET_THROW_INVALID_ARG_IF(L == 0 || system->x.levels.empty() ||
                             system->b.levels.empty());
    const Size3 n0 = system->A.levels.front().size();
    JET_THROW_INVALID_ARG_IF(
        !sameSize(n0, system->x.levels.front().size()) ||
        !sameSize(n0, system->b.levels.front().size()));
    if (n0.x * n0.y * n0.z == 0) return false;
    h->nx.resize(L);
    h->ny.resize(L);
    h->nz.resize(L);
    h->rows.resize(L);
    for (size_t l = 0; l < L; ++l) {
        const Size3 n = system->A.levels[l].size();
        h->nx[l] = n.x;
        h->ny[l] = n.y;
        h->nz[l] = n.z;
        h->rows[l] = reinterpret_cast<const double*>(system->A.levels[l].data());
    }
    return true;
}

// Jet's MgParameters and smoother factor as the library takes them.
jg_mg_params paramsOf(const FdmMgSolver3& solver, double tolerance) {
    jg_mg_params p;
    jg_mg_params_default(&p);
    const MgParameters<FdmBlas3>& mg = solver.params();
    p.max_levels = mg.maxNumberOfLevels;
    p.restrict_iter = mg.numberOfRestrictionIter;
    p.correct_iter = mg.numberOfCorrectionIter;
    p.coarsest_iter = mg.numberOfCoarsestIter;
    p.final_iter = mg.numberOfFinalIter;
    p.max_tol = tolerance;
    p.sor = solver.sorFactor();
    p.red_black = solver.useRedBlackOrdering() ? 1 : 0;
    return p;
}

}  // namespace

// ------------------------------------------------------------ CG

FdmGpuCgSolver3::FdmGpuCgSolver3(unsigned int maxNumberOfIterations,
                                 double tolerance, int device)
    : _maxNumberOfIterations(maxNumberOfIterations),
      _tolerance(tolerance),
      _device(device) {
    _info.last_residual = kMaxD;
}

FdmGpuCgSolver3::~FdmGpuCgSolver3() { jg_solver_destroy(_solver); }

bool FdmGpuCgSolver3::solve(FdmLinearSystem3* system) {
    JET_THROW_INVALID_ARG_IF(system == nullptr);
    const Size3 n = system->A.size();
    JET_THROW_INVALID_ARG_IF(!sameSize(n, system->b.size()) ||
                             !sameSize(n, system->x.size()));
    if (n.x * n.y * n.z == 0) {
        _info = jg_solve_info{};
        return true;
    }
    jg_solver* s = lazySolver(&_solver, _device);
    check(jg_cg_solve_host(s, n.x, n.y, n.z,
                           reinterpret_cast<const double*>(system->A.data()),
                           system->b.data(), system->x.data(),
                           _maxNumberOfIterations, _tolerance, &_info),
          s, "FdmGpuCgSolver3::solve");
    return _info.converged != 0;
}

bool FdmGpuCgSolver3::solveCompressed(FdmCompressedLinearSystem3* system) {
    JET_THROW_INVALID_ARG_IF(system == nullptr);
    const MatrixCsrD& A = system->A;
    const size_t rows = A.rows();
    JET_THROW_INVALID_ARG_IF(A.cols() > rows || system->b.size() != rows);
    system->x.resize(rows, 0.0);
    if (rows == 0) {
        _info = jg_solve_info{};
        return true;
    }
    jg_solver* s = lazySolver(&_solver, _device);
    const size_t nnz = A.numberOfNonZeros();
    check(jg_csr_cg_solve_host(s, rows, nnz, nnz ? &*A.nonZeroBegin() : nullptr,
                               &*A.rowPointersBegin(), nnz ? &*A.columnIndicesBegin() : nullptr,
                               system->b.data(), system->x.data(), _maxNumberOfIterations,
                               _tolerance, &_info),
          s, "FdmGpuCgSolver3::solveCompressed");
    return _info.converged != 0;
}

unsigned int FdmGpuCgSolver3::maxNumberOfIterations() const {
    return _maxNumberOfIterations;
}
unsigned int FdmGpuCgSolver3::lastNumberOfIterations() const {
    return _info.iterations;
}
double FdmGpuCgSolver3::tolerance() const { return _tolerance; }
double FdmGpuCgSolver3::lastResidual() const { return _info.last_residual; }
const jg_solve_info& FdmGpuCgSolver3::lastSolveInfo() const { return _info; }

// --------------------------------------------------------- MGPCG

FdmGpuMgpcgSolver3::FdmGpuMgpcgSolver3(
    unsigned int numberOfCgIter, size_t maxNumberOfLevels,
    unsigned int numberOfRestrictionIter, unsigned int numberOfCorrectionIter,
    unsigned int numberOfCoarsestIter, unsigned int numberOfFinalIter,
    double maxTolerance, double sorFactor, bool useRedBlackOrdering, int device)
    : FdmMgSolver3(maxNumberOfLevels, numberOfRestrictionIter,
                   numberOfCorrectionIter, numberOfCoarsestIter,
                   numberOfFinalIter, maxTolerance, sorFactor,
                   useRedBlackOrdering),
      _maxNumberOfIterations(numberOfCgIter),
      _tolerance(maxTolerance),
      _device(device) {
    jg_gpu_options_default(&_options);
    _info.last_residual = kMaxD;
}

FdmGpuMgpcgSolver3::~FdmGpuMgpcgSolver3() { jg_solver_destroy(_solver); }

bool FdmGpuMgpcgSolver3::solve(FdmMgLinearSystem3* system) {
    Hierarchy h;
    if (!hierarchyOf(system, &h)) {
        _info = jg_solve_info{};
        return true;
    }
    const jg_mg_params p = paramsOf(*this, _tolerance);
    jg_solver* s = lazySolver(&_solver, _device);
    check(jg_mgpcg_solve_host(s, h.rows.size(), h.nx.data(), h.ny.data(), h.nz.data(),
                              h.rows.data(), system->b.levels.front().data(),
                              system->x.levels.front().data(),
                              _maxNumberOfIterations, &p, &_options, &_info),
          s, "FdmGpuMgpcgSolver3::solve");
    // Jet's log line (src/jet/fdm_mgpcg_solver3.cpp:75-76)
    JET_INFO << "Residual after solving MGPCG: " << _info.last_residual
             << " Number of MGPCG iterations: " << _info.iterations;
    return _info.converged != 0;
}

unsigned int FdmGpuMgpcgSolver3::maxNumberOfIterations() const {
    return _maxNumberOfIterations;
}
unsigned int FdmGpuMgpcgSolver3::lastNumberOfIterations() const {
    return _info.iterations;
}
double FdmGpuMgpcgSolver3::tolerance() const { return _tolerance; }
double FdmGpuMgpcgSolver3::lastResidual() const { return _info.last_residual; }
void FdmGpuMgpcgSolver3::setGpuOptions(const jg_gpu_options& options) {
    _options = options;
}
const jg_gpu_options& FdmGpuMgpcgSolver3::gpuOptions() const {
    return _options;
}
const jg_solve_info& FdmGpuMgpcgSolver3::lastSolveInfo() const {
    return _info;
}

// ---------------------------------------------------- flat MGPCG

FdmGpuMgpcgFlatSolver3::FdmGpuMgpcgFlatSolver3(
    unsigned int numberOfCgIter, size_t maxNumberOfLevels,
    unsigned int numberOfRestrictionIter, unsigned int numberOfCorrectionIter,
    unsigned int numberOfCoarsestIter, unsigned int numberOfFinalIter,
    double maxTolerance, double sorFactor, bool useRedBlackOrdering, int device)
    : _sorFactor(sorFactor),
      _useRedBlackOrdering(useRedBlackOrdering),
      _maxNumberOfIterations(numberOfCgIter),
      _device(device) {
    _mgParams.maxNumberOfLevels = maxNumberOfLevels;
    _mgParams.numberOfRestrictionIter = numberOfRestrictionIter;
    _mgParams.numberOfCorrectionIter = numberOfCorrectionIter;
    _mgParams.numberOfCoarsestIter = numberOfCoarsestIter;
    _mgParams.numberOfFinalIter = numberOfFinalIter;
    _mgParams.maxTolerance = maxTolerance;
    jg_gpu_options_default(&_options);
    _info.last_residual = kMaxD;
}

FdmGpuMgpcgFlatSolver3::~FdmGpuMgpcgFlatSolver3() { jg_solver_destroy(_solver); }

bool FdmGpuMgpcgFlatSolver3::solve(FdmLinearSystem3* system) {
    JET_THROW_INVALID_ARG_IF(system == nullptr);
    const Size3 n = system->A.size();
    JET_THROW_INVALID_ARG_IF(!sameSize(n, system->b.size()) ||
                             !sameSize(n, system->x.size()));
    if (n.x * n.y * n.z == 0) {
        _info = jg_solve_info{};
        return true;
    }
    jg_mg_params p;
    jg_mg_params_default(&p);
    p.max_levels = _mgParams.maxNumberOfLevels;
    p.restrict_iter = _mgParams.numberOfRestrictionIter;
    p.correct_iter = _mgParams.numberOfCorrectionIter;
    p.coarsest_iter = _mgParams.numberOfCoarsestIter;
    p.final_iter = _mgParams.numberOfFinalIter;
    p.max_tol = _mgParams.maxTolerance;
    p.sor = _sorFactor;
    p.red_black = _useRedBlackOrdering ? 1 : 0;
    jg_solver* s = lazySolver(&_solver, _device);
    check(jg_mgpcg_flat_solve_host(s, n.x, n.y, n.z,
                                   reinterpret_cast<const double*>(system->A.data()),
                                   system->b.data(), system->x.data(),
                                   _maxNumberOfIterations, &p, &_options, &_info),
          s, "FdmGpuMgpcgFlatSolver3::solve");
    JET_INFO << "Residual after solving MGPCG: " << _info.last_residual
             << " Number of MGPCG iterations: " << _info.iterations;
    return _info.converged != 0;
}

const MgParameters<FdmBlas3>& FdmGpuMgpcgFlatSolver3::params() const {
    return _mgParams;
}
unsigned int FdmGpuMgpcgFlatSolver3::maxNumberOfIterations() const {
    return _maxNumberOfIterations;
}
unsigned int FdmGpuMgpcgFlatSolver3::lastNumberOfIterations() const {
    return _info.iterations;
}
double FdmGpuMgpcgFlatSolver3::tolerance() const { return _mgParams.maxTolerance; }
double FdmGpuMgpcgFlatSolver3::lastResidual() const { return _info.last_residual; }
void FdmGpuMgpcgFlatSolver3::setGpuOptions(const jg_gpu_options& options) {
    _options = options;
}
const jg_gpu_options& FdmGpuMgpcgFlatSolver3::gpuOptions() const { return _options; }
const jg_solve_info& FdmGpuMgpcgFlatSolver3::lastSolveInfo() const { return _info; }

// ------------------------------------------------------------ MG

FdmGpuMgSolver3::FdmGpuMgSolver3(size_t maxNumberOfLevels,
                                 unsigned int numberOfRestrictionIter,
                                 unsigned int numberOfCorrectionIter,
                                 unsigned int numberOfCoarsestIter,
                                 unsigned int numberOfFinalIter,
                                 double maxTolerance, double sorFactor,
                                 bool useRedBlackOrdering, int device)
    : FdmMgSolver3(maxNumberOfLevels, numberOfRestrictionIter,
                   numberOfCorrectionIter, numberOfCoarsestIter,
                   numberOfFinalIter, maxTolerance, sorFactor,
                   useRedBlackOrdering),
      _device(device) {
    jg_gpu_options_default(&_options);
    _info.last_residual = kMaxD;
}

FdmGpuMgSolver3::~FdmGpuMgSolver3() { jg_solver_destroy(_solver); }

bool FdmGpuMgSolver3::solve(FdmMgLinearSystem3* system) {
    Hierarchy h;
    if (!hierarchyOf(system, &h)) {
        _info = jg_solve_info{};
        return true;
    }
    const jg_mg_params p = paramsOf(*this, params().maxTolerance);
    jg_solver* s = lazySolver(&_solver, _device);
    check(jg_mg_solve_host(s, h.rows.size(), h.nx.data(), h.ny.data(), h.nz.data(),
                           h.rows.data(), system->b.levels.front().data(),
                           system->x.levels.front().data(), &p, &_options, &_info),
          s, "FdmGpuMgSolver3::solve");
    return _info.converged != 0;
}

double FdmGpuMgSolver3::lastResidual() const { return _info.last_residual; }
void FdmGpuMgSolver3::setGpuOptions(const jg_gpu_options& options) {
    _options = options;
}
const jg_gpu_options& FdmGpuMgSolver3::gpuOptions() const { return _options; }
const jg_solve_info& FdmGpuMgSolver3::lastSolveInfo() const { return _info; }

}  // namespace jet
