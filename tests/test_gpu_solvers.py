"""GPU CG and MGPCG (FP64 and mixed FP32/FP64) against the CPU oracle.

Pass criteria (the solver parity bar): relative residual <= 1e-8 and
||p - p_ref|| / ||p_ref|| <= 1e-6, with p_ref the oracle's FP64 solution of
the same system (modulo the mean on all-Neumann systems).  GPU CG follows
Jet's FdmCgSolver3 control flow, so its iteration count must match the
oracle's to within rounding (+-2).
"""
import numpy as np
import pytest

from jgtest import jetgpu, rel_err, rel_err_mean_free, torch_cuda
from systems import mg_poisson_hierarchy, poisson_helper, random_stencil

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def upload_mg(solver, sys_):
    solver.set_levels([(x.shape[2], x.shape[1], x.shape[0]) for x in sys_.x])
    for l, A in enumerate(sys_.A):
        solver.upload(l, A, sys_.b[0] if l == 0 else None)


def test_cg_2x2_known_answer(solver):
    # Jet Cg.Solve (src/tests/unit_tests/cg_tests.cpp:16-55).
    A = np.zeros((1, 1, 2, 4))
    A[0, 0, 0] = (4.0, 1.0, 0.0, 0.0)
    A[0, 0, 1] = (3.0, 0.0, 0.0, 0.0)
    solver.set_levels([(2, 1, 1)])
    solver.upload(0, A, np.array([[[1.0, 2.0]]]))
    info = solver.cg(0, 0.0)
    assert info.iterations == 0 and info.last_residual == pytest.approx(np.sqrt(5.0), rel=1e-15)
    assert np.all(solver.download_x() == 0.0)
    info = solver.cg(10, 0.0)
    x = solver.download_x().ravel()
    assert x[0] == pytest.approx(1 / 11, rel=1e-14) and x[1] == pytest.approx(7 / 11, rel=1e-14)
    assert info.iterations <= 2 and info.last_residual <= 1e-15


# (the last two take the tiled z-march: partial 64 x 8 x 16 tiles in x, y and z)
@pytest.mark.parametrize("shape,tol", [((3, 3, 3), 1e-9), ((16, 12, 10), 1e-10), ((40, 40, 40), 1e-9),
                                       ((100, 20, 37), 1e-10), ((64, 13, 50), 1e-10)])
def test_cg_matches_oracle(solver, oracle, shape, tol):
    print(f"case cg shape={shape}", flush=True)
    rng = np.random.default_rng(21)
    nx, ny, nz = shape
    A = random_stencil(nx, ny, nz, rng)
    b = rng.uniform(-1, 1, (nz, ny, nx))
    xo, io, ro, oko = oracle.cg_solve(A, b, 1000, tol)
    solver.set_levels([shape])
    solver.upload(0, A, b)
    info = solver.cg(1000, tol)
    x = solver.download_x()
    assert abs(int(info.iterations) - io) <= 2 and bool(info.converged) == oko
    assert info.last_residual <= tol
    assert rel_err(x, xo) <= 1e-6


def test_cg_closed_form_helper(solver):
    n = 24
    A, b = poisson_helper(n, n, n)
    solver.set_levels([(n, n, n)])
    solver.upload(0, A, b)
    info = solver.cg(1000, 1e-11)
    x = solver.download_x()
    d = x + np.arange(n)[None, :, None]
    assert info.converged and np.ptp(d) < 1e-8


def test_cg_max_iterations_false(solver):
    A, b = poisson_helper(12, 12, 12)
    solver.set_levels([(12, 12, 12)])
    solver.upload(0, A, b)
    info = solver.cg(3, 1e-14)
    assert info.iterations == 3 and not info.converged


def test_mgpcg_jet_unit_test(solver):
    # FdmMgpcgSolver3.Solve (src/tests/unit_tests/fdm_mgpcg_solver3_tests.cpp:13-60):
    # MGPCG(50, 4, 5, 5, 10, 10, 1e-4, 1.5) on 32^3 must return true.
    s = mg_poisson_hierarchy(4, 4)
    upload_mg(solver, s)
    p = jetgpu.mg_params(max_levels=4, restrict_iter=5, correct_iter=5, coarsest_iter=10,
                         final_iter=10, max_tol=1e-4, sor=1.5)
    info = solver.mgpcg(50, p)
    assert info.converged and info.last_residual <= 1e-4


def scene_system(oracle, kind, n, levels, seed):
    s = jetgpu.generate_scene_host(jetgpu.scene_desc(kind, n, seed=seed))
    markers, sys_ = oracle.sp_build_mg(s["u"], s["v"], s["w"], (1.0,) * 3, (0.0,) * 3,
                                       s["bsdf"], s["fsdf"], levels)
    return markers, sys_


@pytest.mark.parametrize("kind,n,levels", [(2, 64, 4), (3, 64, 4), (2, 128, 5)])
@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_mgpcg_scene_matches_oracle(solver, oracle, kind, n, levels, precision):
    print(f"case mgpcg kind={kind} n={n} levels={levels} precision={precision}", flush=True)
    from oracle.jet_oracle import MgParams

    markers, s = scene_system(oracle, kind, n, levels, seed=5)
    upload_mg(solver, s)
    # FP64 reference solution of the same system (clean-mode oracle solve,
    # pinned by its true residual).
    ok, it, rel = oracle.mgpcg_clean_solve(s, MgParams.make(max_levels=levels, restrict_iter=2,
                                                             correct_iter=2, coarsest_iter=20,
                                                             final_iter=2), 100, 1e-12)
    assert ok and rel <= 1e-12
    p_ref = s.x[0].copy()
    p = jetgpu.mg_params(max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2, max_tol=1e-12)
    o = jetgpu.gpu_options(precision=precision, stop_rule=jetgpu.JG_STOP_RELATIVE, rel_tol=1e-10)
    info = solver.mgpcg(200, p, o)
    x = solver.download_x()
    assert info.rel_residual <= 1e-8
    assert rel_err(x, p_ref) <= 1e-6
    # Non-fluid cells are identity rows with b = 0: p = 0 there up to the
    # solve tolerance (exactly 0 only when the smoother's SOR factor is 1,
    # since the V-cycle's correction reaches them and a sweep scales them by
    # 1 - omega).
    assert np.max(np.abs(x[markers[0] != 0])) <= 1e-6 * np.max(np.abs(p_ref))
    assert info.iterations <= 60


@pytest.mark.parametrize("precision", [jetgpu.JG_MIXED, jetgpu.JG_FP64])
def test_mgpcg_two_sweep_kernels_bitwise(solver, oracle, precision):
    # The solve with sweeps paired into two-sweep launches (z-marching and
    # brick kernels) gives the solve with single-sweep launches: the same
    # iterations, and x to rounding.  The V-cycles are bit-identical
    # (test_vcycle_two_sweep_kernel_bitwise), but the r.z / r.r sums differ
    # in order (a single last sweep emits them as block partials, a paired
    # one leaves them to a dot kernel); in mixed precision that one-ulp FP64
    # difference usually rounds away in the FP32 inner iterates, but not
    # always, so neither precision is compared bit for bit.
    markers, s = scene_system(oracle, 3, 128, 5, seed=9)
    upload_mg(solver, s)
    p = jetgpu.mg_params(max_levels=5, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2)
    out = {}
    for pairs in (-1, 1):
        o = jetgpu.gpu_options(precision=precision, stop_rule=jetgpu.JG_STOP_RELATIVE,
                               rel_tol=1e-9, sweep_pairs=pairs)
        info = solver.mgpcg(200, p, o)
        assert info.rel_residual <= 1e-8
        out[pairs] = (info.iterations, solver.download_x())
    assert out[-1][0] == out[1][0]
    assert rel_err(out[-1][1], out[1][1]) <= 1e-12


def test_mgpcg_closed_sphere_singular(solver, oracle):
    # All-Neumann (config-1 geometry): compare modulo the mean.
    markers, s = scene_system(oracle, 1, 32, 3, seed=2)
    upload_mg(solver, s)
    xc, it, res, ok = oracle.cg_solve(s.A[0], s.b[0], 3000, 1e-10)
    assert ok
    p = jetgpu.mg_params(max_levels=3, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2, max_tol=1e-10)
    for prec in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
        info = solver.mgpcg(200, p, jetgpu.gpu_options(precision=prec, stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                       rel_tol=1e-10))
        x = solver.download_x()
        fluid = markers[0] == 0
        assert info.rel_residual <= 1e-8
        assert rel_err_mean_free(x, xc, fluid) <= 1e-6


def test_mgpcg_zero_rhs(solver):
    s = mg_poisson_hierarchy(4, 3)
    s.b[0][:] = 0.0
    upload_mg(solver, s)
    info = solver.mgpcg(10, jetgpu.mg_params(max_levels=3))
    assert info.iterations == 0 and info.converged
    assert np.all(solver.download_x() == 0.0)


def test_mgpcg_nonfinite_rhs_is_an_error(solver):
    s = mg_poisson_hierarchy(4, 2)
    s.b[0][0, 0, 0] = np.nan
    upload_mg(solver, s)
    with pytest.raises(jetgpu.JgError):
        solver.mgpcg(10, jetgpu.mg_params(max_levels=2))


def test_mgpcg_mixed_fp32_overflow_is_an_error(solver):
    # b inside FP64's range but beyond FP32's: the FP32 inner solve meets inf, and the mixed solve
    # must report a non-finite value instead of returning x = 0 as converged.  The all-Neumann
    # system needs a b that sums to zero: Jet's +-1 rows, scaled
    s = mg_poisson_hierarchy(4, 2)
    s.b[0][:] *= 1e39
    upload_mg(solver, s)
    p = jetgpu.mg_params(max_levels=2)
    with pytest.raises(jetgpu.JgError):
        solver.mgpcg(20, p, jetgpu.gpu_options(precision=jetgpu.JG_MIXED))
    info = solver.mgpcg(100, p, jetgpu.gpu_options(precision=jetgpu.JG_FP64,
                                                    stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                    rel_tol=1e-9))
    assert info.converged and info.rel_residual <= 1e-9


def test_mgpcg_mixed_refinement_limit_is_not_convergence(solver):
    # one refinement (inner tolerance 1e-4 relative) cannot reach an absolute 1e-10: Jet's bool
    # must be false, as an FP64 solve stopped at max_iter would return
    s = mg_poisson_hierarchy(4, 3)
    upload_mg(solver, s)
    p = jetgpu.mg_params(max_levels=3, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2, max_tol=1e-10)
    info = solver.mgpcg(200, p, jetgpu.gpu_options(precision=jetgpu.JG_MIXED, max_refine=1))
    assert info.refinements == 1 and info.last_residual > 1e-10 and not info.converged
    info = solver.mgpcg(200, p, jetgpu.gpu_options(precision=jetgpu.JG_MIXED))
    assert info.last_residual <= 1e-10 and info.converged
