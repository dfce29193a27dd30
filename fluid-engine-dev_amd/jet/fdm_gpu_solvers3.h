// fdm_gpu_solvers3.h -- Jet linear-system solvers backed by the MI355X
// pressure-projection library (include/jetgpu.h).
//
// These are the classes a Jet build adds to run its grid pressure solve on
// the GPU.  They derive from Jet's own solver interfaces, so every caller
// that takes an FdmLinearSystemSolver3Ptr -- GridSinglePhasePressureSolver3,
// GridFractionalSinglePhasePressureSolver3, GridBackwardEulerDiffusionSolver3
// -- uses them unchanged through setLinearSystemSolver():
//
//   FdmGpuCgSolver3     : FdmLinearSystemSolver3
//       replaces FdmCgSolver3 (include/jet/fdm_cg_solver3.h:19-37,
//       src/jet/fdm_cg_solver3.cpp:14-47): same constructor arguments,
//       accessors and bool result;
//   FdmGpuMgSolver3     : FdmMgSolver3
//       replaces FdmMgSolver3 (include/jet/fdm_mg_solver3.h:23-51,
//       src/jet/fdm_mg_solver3.cpp:14-74): one V-cycle per solve;
//   FdmGpuMgpcgFlatSolver3 : FdmLinearSystemSolver3
//       MGPCG for callers that hand over one level (its coarse matrices
//       built on the device from A);
//   FdmGpuMgpcgSolver3  : FdmMgSolver3
//       replaces FdmMgpcgSolver3 (include/jet/fdm_mgpcg_solver3.h:35-56,
//       src/jet/fdm_mgpcg_solver3.cpp:40-80): same constructor arguments and
//       accessors; params() sizes the caller's hierarchy exactly as Jet's
//       pressure solvers expect (src/jet/grid_single_phase_pressure_solver3.
//       cpp:254-256), and the pressure solvers' dynamic_pointer_cast to
//       FdmMgSolver3 (:228) selects the multigrid system.
//
// Semantics kept from Jet: solve() zeroes and overwrites x, returns
// lastResidual <= tolerance || iterations < max; a size mismatch between A,
// x and b throws std::invalid_argument (Jet's JET_THROW_INVALID_ARG_IF,
// include/jet/macros.h:37-39); FdmGpuCgSolver3::solveCompressed() runs the
// same CG on the compressed CSR system (as FdmCgSolver3's does), while the
// multigrid solver's solveCompressed() returns false (the base class no-op,
// include/jet/fdm_linear_system_solver3.h:28, as for Jet's MGPCG) and its
// flat solve is FdmMgSolver3's final no-op.  A failure of
// the device itself (no GPU, out of device memory) throws std::runtime_error
// with the library's message; Jet's CPU solvers have no such failure.
//
// The multigrid preconditioner is the library's symmetric V-cycle (zero
// initial guess, red-black smoothing with the solver's sorFactor), not Jet's
// warm-started lexicographic one; FP64 by default, or FP32 iterative
// refinement to the same tolerance with setGpuOptions().
//
// Device state (the resident system and work vectors) is created on the first
// solve and cached by size, like the scratch vectors Jet's solvers keep as
// members; an instance is used by one thread at a time, as Jet's are.
#ifndef FLUID_ENGINE_DEV_AMD_JET_FDM_GPU_SOLVERS3_H_
#define FLUID_ENGINE_DEV_AMD_JET_FDM_GPU_SOLVERS3_H_

#include <jet/fdm_linear_system_solver3.h>
#include <jet/fdm_mg_solver3.h>

#include <memory>

#include "jetgpu.h"

namespace jet {

//! GPU conjugate gradient (Jet's cg<FdmBlas3> control flow on the device).
class FdmGpuCgSolver3 final : public FdmLinearSystemSolver3 {
 public:
    FdmGpuCgSolver3(unsigned int maxNumberOfIterations, double tolerance,
                    int device = 0);
    ~FdmGpuCgSolver3() override;

    FdmGpuCgSolver3(const FdmGpuCgSolver3&) = delete;
    FdmGpuCgSolver3& operator=(const FdmGpuCgSolver3&) = delete;

    bool solve(FdmLinearSystem3* system) override;
    //! FdmCgSolver3::solveCompressed (src/jet/fdm_cg_solver3.cpp:49-74): the
    //! same CG on the compressed (fluid-rows-only) CSR system.
    bool solveCompressed(FdmCompressedLinearSystem3* system) override;

    unsigned int maxNumberOfIterations() const;
    unsigned int lastNumberOfIterations() const;
    double tolerance() const;
    double lastResidual() const;

    //! Device-side details of the last solve (time, FP64 relative residual).
    const jg_solve_info& lastSolveInfo() const;

 private:
    unsigned int _maxNumberOfIterations;
    double _tolerance;
    int _device;
    jg_solver* _solver = nullptr;
    jg_solve_info _info{};
};

//! GPU multigrid-preconditioned CG.
class FdmGpuMgpcgSolver3 final : public FdmMgSolver3 {
 public:
    FdmGpuMgpcgSolver3(unsigned int numberOfCgIter, size_t maxNumberOfLevels,
                       unsigned int numberOfRestrictionIter = 5,
                       unsigned int numberOfCorrectionIter = 5,
                       unsigned int numberOfCoarsestIter = 20,
                       unsigned int numberOfFinalIter = 20,
                       double maxTolerance = 1e-9, double sorFactor = 1.5,
                       bool useRedBlackOrdering = false, int device = 0);
    ~FdmGpuMgpcgSolver3() override;

    FdmGpuMgpcgSolver3(const FdmGpuMgpcgSolver3&) = delete;
    FdmGpuMgpcgSolver3& operator=(const FdmGpuMgpcgSolver3&) = delete;

    bool solve(FdmMgLinearSystem3* system) override;

    unsigned int maxNumberOfIterations() const;
    unsigned int lastNumberOfIterations() const;
    double tolerance() const;
    double lastResidual() const;

    //! GPU-only options: precision (JG_FP64 / JG_MIXED), stop rule, smoother
    //! relaxation.  Defaults: FP64 and Jet's stop rule (sqrt(r.z) <= tol).
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

//! GPU multigrid solver: one V-cycle per solve, from the caller's x.
//!
//! Replaces FdmMgSolver3 itself (include/jet/fdm_mg_solver3.h:23-51,
//! src/jet/fdm_mg_solver3.cpp:14-74): same constructor arguments and params();
//! solve() runs one V-cycle on the caller's hierarchy starting from the x it
//! holds (Jet does not zero it) and returns Jet's result, the residual norm
//! after the cycle below maxTolerance.  The cycle is the library's symmetric
//! red-black V-cycle, as in FdmGpuMgpcgSolver3.
class FdmGpuMgSolver3 final : public FdmMgSolver3 {
 public:
    FdmGpuMgSolver3(size_t maxNumberOfLevels,
                    unsigned int numberOfRestrictionIter = 5,
                    unsigned int numberOfCorrectionIter = 5,
                    unsigned int numberOfCoarsestIter = 20,
                    unsigned int numberOfFinalIter = 20,
                    double maxTolerance = 1e-9, double sorFactor = 1.5,
                    bool useRedBlackOrdering = false, int device = 0);
    ~FdmGpuMgSolver3() override;

    FdmGpuMgSolver3(const FdmGpuMgSolver3&) = delete;
    FdmGpuMgSolver3& operator=(const FdmGpuMgSolver3&) = delete;

    bool solve(FdmMgLinearSystem3* system) override;

    //! ||b - A x||_2 on the finest level after the last solve.
    double lastResidual() const;

    void setGpuOptions(const jg_gpu_options& options);
    const jg_gpu_options& gpuOptions() const;

    const jg_solve_info& lastSolveInfo() const;

 private:
    int _device;
    jg_gpu_options _options;
    jg_solver* _solver = nullptr;
    jg_solve_info _info{};
};

//! GPU MGPCG for single-level systems (the flat entry).
//!
//! Jet's FdmMgSolver3::solve(FdmLinearSystem3*) is final and does nothing
//! (include/jet/fdm_mg_solver3.h:42), so a caller that builds one level --
//! GridBackwardEulerDiffusionSolver3 (src/jet/
//! grid_backward_euler_diffusion_solver3.cpp:42,71,85,99) -- cannot use
//! multigrid.  This class derives FdmLinearSystemSolver3 instead and takes
//! FdmMgpcgSolver3's constructor arguments
//! (include/jet/fdm_mgpcg_solver3.h:40-45): solve() sizes a hierarchy by
//! Jet's resizeArrayWithFinest rule (halving in x and y only for a one-plane
//! system), builds every coarse matrix on the device from A (2x2x2
//! aggregation scaled to the operator at twice the spacing, jetgpu.h's
//! jg_solver_coarsen) and runs the library's MGPCG.  solve() zeroes and
//! overwrites x and returns Jet's MGPCG result; solveCompressed() returns
//! false (the base class no-op).
class FdmGpuMgpcgFlatSolver3 final : public FdmLinearSystemSolver3 {
 public:
    FdmGpuMgpcgFlatSolver3(unsigned int numberOfCgIter, size_t maxNumberOfLevels,
                           unsigned int numberOfRestrictionIter = 5,
                           unsigned int numberOfCorrectionIter = 5,
                           unsigned int numberOfCoarsestIter = 20,
                           unsigned int numberOfFinalIter = 20,
                           double maxTolerance = 1e-9, double sorFactor = 1.5,
                           bool useRedBlackOrdering = false, int device = 0);
    ~FdmGpuMgpcgFlatSolver3() override;

    FdmGpuMgpcgFlatSolver3(const FdmGpuMgpcgFlatSolver3&) = delete;
    FdmGpuMgpcgFlatSolver3& operator=(const FdmGpuMgpcgFlatSolver3&) = delete;

    bool solve(FdmLinearSystem3* system) override;

    const MgParameters<FdmBlas3>& params() const;
    unsigned int maxNumberOfIterations() const;
    unsigned int lastNumberOfIterations() const;
    double tolerance() const;
    double lastResidual() const;

    void setGpuOptions(const jg_gpu_options& options);
    const jg_gpu_options& gpuOptions() const;

    const jg_solve_info& lastSolveInfo() const;

 private:
    MgParameters<FdmBlas3> _mgParams;
    double _sorFactor;
    bool _useRedBlackOrdering;
    unsigned int _maxNumberOfIterations;
    int _device;
    jg_gpu_options _options;
    jg_solver* _solver = nullptr;
    jg_solve_info _info{};
};

typedef std::shared_ptr<FdmGpuCgSolver3> FdmGpuCgSolver3Ptr;
typedef std::shared_ptr<FdmGpuMgSolver3> FdmGpuMgSolver3Ptr;
typedef std::shared_ptr<FdmGpuMgpcgSolver3> FdmGpuMgpcgSolver3Ptr;
typedef std::shared_ptr<FdmGpuMgpcgFlatSolver3> FdmGpuMgpcgFlatSolver3Ptr;

}  // namespace jet

#endif  // FLUID_ENGINE_DEV_AMD_JET_FDM_GPU_SOLVERS3_H_
