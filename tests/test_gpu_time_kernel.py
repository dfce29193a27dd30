"""jg_solver_time_kernel: every level-0 kernel of the solve, in both
precisions, times to a positive duration with positive algorithmic bytes,
and leaves the resident system (matrices and right-hand side) as it found
it, so a solve after a measurement returns the same solution bit for bit.
The two-sweep kernel is FP32 only and refuses FP64."""
import numpy as np
import pytest

from jgtest import jetgpu, torch_cuda
from systems import random_stencil

pytestmark = pytest.mark.gpu

KERNELS = [jetgpu.JG_KERNEL_RBGS, jetgpu.JG_KERNEL_DIR_SPMV, jetgpu.JG_KERNEL_RESIDUAL,
           jetgpu.JG_KERNEL_UPDATE_XR, jetgpu.JG_KERNEL_DOT, jetgpu.JG_KERNEL_COPY,
           jetgpu.JG_KERNEL_SWEEP, jetgpu.JG_KERNEL_SWEEP2, jetgpu.JG_KERNEL_PRE0,
           jetgpu.JG_KERNEL_POST0]


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


@pytest.mark.parametrize("shape,levels", [((96, 64, 48), 3), ((128, 96, 64), 3)])
def test_time_kernel_keeps_the_system(solver, shape, levels):
    print(f"case time_kernel shape={shape} levels={levels}", flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(71)
    sizes = [(nx >> l, ny >> l, nz >> l) for l in range(levels)]
    solver.set_levels(sizes)
    b = rng.uniform(-1, 1, (nz, ny, nx))
    for l, (a, c, e) in enumerate(sizes):
        solver.upload(l, random_stencil(a, c, e, rng), b if l == 0 else None)
    p = jetgpu.mg_params(max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2)
    before = {}
    for prec in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
        o = jetgpu.gpu_options(precision=prec, stop_rule=jetgpu.JG_STOP_RELATIVE, rel_tol=1e-10)
        info = solver.mgpcg(200, p, o)
        assert info.converged
        before[prec] = (info.iterations, solver.download_x())
    for kid in KERNELS:
        for prec in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
            if kid == jetgpu.JG_KERNEL_SWEEP2 and prec == jetgpu.JG_FP64:
                with pytest.raises(jetgpu.JgError):
                    solver.time_kernel(kid, prec, 2)
                continue
            ms, nbytes = solver.time_kernel(kid, prec, 3)
            assert ms > 0 and nbytes > 0, (kid, prec)
    for prec in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
        o = jetgpu.gpu_options(precision=prec, stop_rule=jetgpu.JG_STOP_RELATIVE, rel_tol=1e-10)
        info = solver.mgpcg(200, p, o)
        it, x = before[prec]
        assert info.iterations == it, prec
        assert np.array_equal(solver.download_x(), x), prec
