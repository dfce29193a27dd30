// fdm_gpu_solvers2.cpp -- FdmGpuCgSolver2 and FdmGpuMgpcgSolver2 over the C ABI
// of jetgpu.h.
#include "fdm_gpu_solvers2.h"

#include <jet/fdm_linear_system2.h>
#include <jet/fdm_mg_linear_system2.h>
#include <jet/logging.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace jet {

namespace {

void check2(int status, const jg_solver* s, const char* what) {
    if (status == JG_OK) return;
    std::string msg = std::string(what) + ": " + (s ? jg_last_error(s) : "");
    if (status == JG_ERR_INVALID_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

jg_solver* lazySolver2(jg_solver** slot, int device) {
    if (!*slot) {
        const int st = jg_solver_create(device, slot);
        if (st != JG_OK) {
            *slot = nullptr;
            throw std::runtime_error("jetgpu: no usable GPU device " + std::to_string(device) +
                                     " (status " + std::to_string(st) + ")");
        }
    }
    return *slot;
}

}  // namespace

FdmGpuCgSolver2::FdmGpuCgSolver2(unsigned int maxNumberOfIterations, double tolerance,
                                 int device)
    : _maxNumberOfIterations(maxNumberOfIterations), _tolerance(tolerance), _device(device) {
    _info.last_residual = kMaxD;
}

FdmGpuCgSolver2::~FdmGpuCgSolver2() { jg_solver_destroy(_solver); }

bool FdmGpuCgSolver2::solve(FdmLinearSystem2* system) {
    JET_THROW_INVALID_ARG_IF(system == nullptr);
    const Size2 n = system->A.size();
    JET_THROW_INVALID_ARG_IF(n.x != system->b.size().x || n.y != system->b.size().y ||
                             n.x != system->x.size().x || n.y != system->x.size().y);
    if (n.x * n.y == 0) {
        _info = jg_solve_info{};
        return true;
    }
    // {center, right, up} -> {center, right, up, front = 0}, x fastest
    std::vector<double> rows(4 * n.x * n.y);
    for (size_t j = 0; j < n.y; ++j)
        for (size_t i = 0; i < n.x; ++i) {
            const FdmMatrixRow2& r = system->A(i, j);
            double* d = &rows[4 * (i + n.x * j)];
            d[0] = r.center;
            d[1] = r.right;
            d[2] = r.up;
            d[3] = 0.0;
        }
    jg_solver* s = lazySolver2(&_solver, _device);
    check2(jg_cg_solve_host(s, n.x, n.y, 1, rows.data(), system->b.data(), system->x.data(),
                            _maxNumberOfIterations, _tolerance, &_info),
           s, "FdmGpuCgSolver2::solve");
    return _info.converged != 0;
}

bool FdmGpuCgSolver2::solveCompressed(FdmCompressedLinearSystem2* system) {
    JET_THROW_INVALID_ARG_IF(system == nullptr);
    const MatrixCsrD& A = system->A;
    const size_t rows = A.rows();
    JET_THROW_INVALID_ARG_IF(A.cols() > rows || system->b.size() != rows);
    system->x.resize(rows, 0.0);
    if (rows == 0) {
        _info = jg_solve_info{};
        return true;
    }
    jg_solver* s = lazySolver2(&_solver, _device);
    const size_t nnz = A.numberOfNonZeros();
    check2(jg_csr_cg_solve_host(s, rows, nnz, nnz ? &*A.nonZeroBegin() : nullptr,
                                &*A.rowPointersBegin(), nnz ? &*A.columnIndicesBegin() : nullptr,
                                system->b.data(), system->x.data(), _maxNumberOfIterations,
                                _tolerance, &_info),
           s, "FdmGpuCgSolver2::solveCompressed");
    return _info.converged != 0;
}

unsigned int FdmGpuCgSolver2::maxNumberOfIterations() const { return _maxNumberOfIterations; }
unsigned int FdmGpuCgSolver2::lastNumberOfIterations() const { return _info.iterations; }
double FdmGpuCgSolver2::tolerance() const { return _tolerance; }
double FdmGpuCgSolver2::lastResidual() const { return _info.last_residual; }
const jg_solve_info& FdmGpuCgSolver2::lastSolveInfo() const { return _info; }

// --------------------------------------------------------- MGPCG

FdmGpuMgpcgSolver2::FdmGpuMgpcgSolver2(unsigned int numberOfCgIter, size_t maxNumberOfLevels,
                                       unsigned int numberOfRestrictionIter,
                                       unsigned int numberOfCorrectionIter,
                                       unsigned int numberOfCoarsestIter,
                                       unsigned int numberOfFinalIter, double maxTolerance,
                                       double sorFactor, bool useRedBlackOrdering, int device)
    : FdmMgSolver2(maxNumberOfLevels, numberOfRestrictionIter, numberOfCorrectionIter,
                   numberOfCoarsestIter, numberOfFinalIter, maxTolerance, sorFactor,
                   useRedBlackOrdering),
      _maxNumberOfIterations(numberOfCgIter),
      _tolerance(maxTolerance),
      _device(device) {
    jg_gpu_options_default(&_options);
    _info.last_residual = kMaxD;
}

FdmGpuMgpcgSolver2::~FdmGpuMgpcgSolver2() { jg_solver_destroy(_solver); }

bool FdmGpuMgpcgSolver2::solve(FdmMgLinearSystem2* system) {
    JET_THROW_INVALID_ARG_IF(system == nullptr);
    const size_t L = system->A.levels.size();
    JET_THROW_INVALID_ARG_IF(L == 0 || system->x.levels.empty() || system->b.levels.empty());
    const Size2 n0 = system->A.levels.front().size();
    JET_THROW_INVALID_ARG_IF(n0 != system->x.levels.front().size() ||
                             n0 != system->b.levels.front().size());
    if (n0.x * n0.y == 0) {
        _info = jg_solve_info{};
        return true;
    }
    // every level as one plane of {center, right, up, front = 0} rows
    std::vector<size_t> nx(L), ny(L), nz(L, 1);
    std::vector<std::vector<double>> rows(L);
    std::vector<const double*> ptrs(L);
    for (size_t l = 0; l < L; ++l) {
        const FdmMatrix2& A = system->A.levels[l];
        const Size2 n = A.size();
        nx[l] = n.x;
        ny[l] = n.y;
        rows[l].resize(4 * n.x * n.y);
        for (size_t j = 0; j < n.y; ++j)
            for (size_t i = 0; i < n.x; ++i) {
                const FdmMatrixRow2& r = A(i, j);
                double* d = &rows[l][4 * (i + n.x * j)];
                d[0] = r.center;
                d[1] = r.right;
                d[2] = r.up;
                d[3] = 0.0;
            }
        ptrs[l] = rows[l].data();
    }
    // Jet's MgParameters and smoother factor, as FdmMgpcgSolver2 uses them
    jg_mg_params p;
    jg_mg_params_default(&p);
    const MgParameters<FdmBlas2>& mg = params();
    p.max_levels = mg.maxNumberOfLevels;
    p.restrict_iter = mg.numberOfRestrictionIter;
    p.correct_iter = mg.numberOfCorrectionIter;
    p.coarsest_iter = mg.numberOfCoarsestIter;
    p.final_iter = mg.numberOfFinalIter;
    p.max_tol = _tolerance;
    p.sor = sorFactor();
    p.red_black = useRedBlackOrdering() ? 1 : 0;

    jg_solver* s = lazySolver2(&_solver, _device);
    check2(jg_mgpcg_solve_host(s, L, nx.data(), ny.data(), nz.data(), ptrs.data(),
                               system->b.levels.front().data(), system->x.levels.front().data(),
                               _maxNumberOfIterations, &p, &_options, &_info),
           s, "FdmGpuMgpcgSolver2::solve");
    // Jet's log line (src/jet/fdm_mgpcg_solver2.cpp:75-76)
    JET_INFO << "Residual after solving MGPCG: " << _info.last_residual
             << " Number of MGPCG iterations: " << _info.iterations;
    return _info.converged != 0;
}

unsigned int FdmGpuMgpcgSolver2::maxNumberOfIterations() const { return _maxNumberOfIterations; }
unsigned int FdmGpuMgpcgSolver2::lastNumberOfIterations() const { return _info.iterations; }
double FdmGpuMgpcgSolver2::tolerance() const { return _tolerance; }
double FdmGpuMgpcgSolver2::lastResidual() const { return _info.last_residual; }
void FdmGpuMgpcgSolver2::setGpuOptions(const jg_gpu_options& options) { _options = options; }
const jg_gpu_options& FdmGpuMgpcgSolver2::gpuOptions() const { return _options; }
const jg_solve_info& FdmGpuMgpcgSolver2::lastSolveInfo() const { return _info; }

}  // namespace jet
