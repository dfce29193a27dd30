// fdm_gpu_solvers2.h -- the 2-D drop-ins for Jet's FdmCgSolver2
// (include/jet/fdm_cg_solver2.h:16-60, src/jet/fdm_cg_solver2.cpp) and
// FdmMgpcgSolver2 (include/jet/fdm_mgpcg_solver2.h:20-85,
// src/jet/fdm_mgpcg_solver2.cpp) over the C ABI of jetgpu.h.
//
// A 2-D system nx x ny is the library's 3-D system nx x ny x 1: each
// FdmMatrixRow2 {center, right, up} becomes {center, right, up, 0}, which
// changes no sum (the missing +z / -z terms are exact zeros), and the
// compressed form is dimension-free.  Semantics as FdmGpuCgSolver3: x is
// zeroed and overwritten, the bool result is Jet's, a size mismatch throws
// std::invalid_argument, a device failure std::runtime_error.  The library's
// stencil reductions bound one plane to about 2^24 cells, so larger 2-D grids are
// rejected (std::runtime_error); their compressed form has no such bound.
#pragma once

#include <jet/fdm_linear_system_solver2.h>
#include <jet/fdm_mg_solver2.h>

#include "jetgpu.h"

namespace jet {

class FdmGpuCgSolver2 final : public FdmLinearSystemSolver2 {
 public:
    FdmGpuCgSolver2(unsigned int maxNumberOfIterations, double tolerance, int device = 0);
    ~FdmGpuCgSolver2() override;
    FdmGpuCgSolver2(const FdmGpuCgSolver2&) = delete;
    FdmGpuCgSolver2& operator=(const FdmGpuCgSolver2&) = delete;

    bool solve(FdmLinearSystem2* system) override;
    bool solveCompressed(FdmCompressedLinearSystem2* system) override;

    unsigned int maxNumberOfIterations() const;
    unsigned int lastNumberOfIterations() const;
    double tolerance() const;
    double lastResidual() const;
    const jg_solve_info& lastSolveInfo() const;

 private:
    unsigned int _maxNumberOfIterations;
    double _tolerance;
    int _device;
    jg_solver* _solver = nullptr;
    jg_solve_info _info{};
};

// MGPCG on a 2-D multigrid system: every level nx x ny becomes a one-plane
// level of the library's hierarchy (rows {center, right, up, 0}), which the
// library solves with FdmMgUtils2's transfers.  The preconditioner is the
// library's symmetric V-cycle (forward pre-sweeps, backward post-sweeps,
// zero initial guess), so iteration counts can differ from Jet's by a few;
// the solution meets the same stop rule.  Options and errors as
// FdmGpuMgpcgSolver3.
class FdmGpuMgpcgSolver2 final : public FdmMgSolver2 {
 public:
    FdmGpuMgpcgSolver2(unsigned int numberOfCgIter, size_t maxNumberOfLevels,
                       unsigned int numberOfRestrictionIter = 5,
                       unsigned int numberOfCorrectionIter = 5,
                       unsigned int numberOfCoarsestIter = 20,
                       unsigned int numberOfFinalIter = 20, double maxTolerance = 1e-9,
                       double sorFactor = 1.5, bool useRedBlackOrdering = false, int device = 0);
    ~FdmGpuMgpcgSolver2() override;
    FdmGpuMgpcgSolver2(const FdmGpuMgpcgSolver2&) = delete;
    FdmGpuMgpcgSolver2& operator=(const FdmGpuMgpcgSolver2&) = delete;

    bool solve(FdmMgLinearSystem2* system) override;

    unsigned int maxNumberOfIterations() const;
    unsigned int lastNumberOfIterations() const;
    double tolerance() const;
    double lastResidual() const;

    void setGpuOptions(const jg_gpu_options& options);
    const jg_gpu_options& gpuOptions() const;
    const jg_solve_info& lastSolveInfo() const;

 private:
    unsigned int _maxNumberOfIterations;
    double _tolerance;
    int _device;
    jg_gpu_options _options;
    jg_solver* _solver = nullptr;
    jg_solve_info _info{};
};

typedef std::shared_ptr<FdmGpuCgSolver2> FdmGpuCgSolver2Ptr;
typedef std::shared_ptr<FdmGpuMgpcgSolver2> FdmGpuMgpcgSolver2Ptr;

}  // namespace jet
