"""Weighted-Jacobi smoothing in the V-cycle (JG_SMOOTH_JACOBI) and the
one-V-cycle solve (jg_solver_mg, FdmMgSolver3::solve) through the C ABI,
against the oracle.

The Jacobi V-cycle is the oracle's clean Jacobi cycle (zero start,
restrict_iter sweeps x += omega ((b - s) / c - x) in FdmJacobiSolver3::relax's
operand order, src/jet/fdm_jacobi_solver3.cpp:95-112, residual and Jet's
restriction, recursion, Jet's correction, post sweeps): the GPU evaluates the
same kernels one by one, so FP64 agrees to rounding (<= 1e-12 relative, in
practice bit for bit) and FP32 to FP32 rounding.  MGPCG with it must reach the
oracle's 1e-12 solution p_ref to 1e-6, as with the red-black smoother.

jg_solver_mg runs one V-cycle from the caller's x: x1 = x0 + M (b - A x0)
with M the zero-start cycle.  It is checked against the oracle's residual and
clean V-cycle on a random hierarchy, and with Jet's FdmMgSolver3.Solve test
(src/tests/unit_tests/fdm_mg_solver3_tests.cpp:13-69: 6 levels from a 4^3
coarsest level, one cycle must reduce ||b - A x||).
"""
import numpy as np
import pytest

from jgtest import dev, empty, host, jetgpu, rel_err, torch_cuda
from systems import mg_poisson_hierarchy, random_stencil

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def random_hierarchy(shape, levels, seed):
    from oracle.jet_oracle import MgSystem

    rng = np.random.default_rng(seed)
    s = MgSystem()
    nx, ny, nz = shape
    twod = nz == 1 and levels > 1
    for l in range(levels):
        A = random_stencil(nx >> l, ny >> l, 1 if twod else nz >> l, rng)
        s.A.append(A)
        s.b.append(np.zeros(A.shape[:3]))
        s.x.append(np.zeros(A.shape[:3]))
    return s


def upload(solver, s, b=None):
    solver.set_levels([(x.shape[2], x.shape[1], x.shape[0]) for x in s.x])
    for l, A in enumerate(s.A):
        solver.upload(l, A, b if l == 0 else None)


def both_params(**kw):
    from oracle.jet_oracle import MgParams

    return jetgpu.mg_params(**kw), MgParams.make(**kw)


@pytest.mark.parametrize("shape,levels", [((32, 32, 32), 4), ((48, 32, 16), 3), ((16, 16, 16), 1),
                                          ((64, 48, 1), 4)])
@pytest.mark.parametrize("sched", [(2, 2, 20, 2), (1, 3, 5, 2)])
@pytest.mark.parametrize("omega", [6.0 / 7.0, 1.0, 0.5])
def test_jacobi_vcycle_matches_oracle(solver, oracle, shape, levels, sched, omega):
    print(f"case jacobi vcycle shape={shape} levels={levels} sched={sched} omega={omega}", flush=True)
    s = random_hierarchy(shape, levels, seed=81)
    upload(solver, s)
    pre, post, coarsest, final = sched
    gp, op = both_params(restrict_iter=pre, correct_iter=post, coarsest_iter=coarsest,
                         final_iter=final)
    r = np.random.default_rng(82).uniform(-1, 1, s.x[0].shape)
    want = oracle.mg_clean_vcycle_jacobi(s, op, omega, r)
    for precision, tol in ((jetgpu.JG_FP64, 1e-12), (jetgpu.JG_MIXED, 2e-5)):
        z = empty(r.shape)
        solver.mg_vcycle(gp, jetgpu.gpu_options(precision=precision, smoother=jetgpu.JG_SMOOTH_JACOBI,
                                                omega=omega), dev(r), z)
        assert rel_err(host(z), want) <= tol, precision


def test_jacobi_vcycle_zero_centre_rows(solver, oracle):
    # a fluid cell with only solid neighbours is an all-zero row: the Jacobi
    # smoother sets it to 0 (as the red-black one does) instead of dividing
    s = random_hierarchy((16, 16, 16), 2, seed=83)
    s.A[0][3, 4, 5, :] = 0.0
    s.A[0][3, 4, 4, 1] = 0.0
    s.A[0][3, 3, 5, 2] = 0.0
    s.A[0][2, 4, 5, 3] = 0.0
    upload(solver, s)
    gp, op = both_params(restrict_iter=2, correct_iter=2, coarsest_iter=10, final_iter=2)
    r = np.random.default_rng(84).uniform(-1, 1, s.x[0].shape)
    want = oracle.mg_clean_vcycle_jacobi(s, op, 6.0 / 7.0, r)
    z = empty(r.shape)
    solver.mg_vcycle(gp, jetgpu.gpu_options(smoother=jetgpu.JG_SMOOTH_JACOBI), dev(r), z)
    got = host(z)
    assert np.all(np.isfinite(got)) and got[3, 4, 5] == 0.0
    assert rel_err(got, want) <= 1e-12


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_jacobi_mgpcg_sphere_pool_matches_oracle(solver, oracle, precision):
    from oracle.jet_oracle import MgParams

    n, levels = 64, 4
    print(f"case jacobi mgpcg sphere pool n={n} precision={precision}", flush=True)
    sc = jetgpu.generate_scene_host(jetgpu.scene_desc(jetgpu.JG_SCENE_SPHERE_POOL, n, seed=5))
    markers, s = oracle.sp_build_mg(sc["u"], sc["v"], sc["w"], (1.0,) * 3, (0.0,) * 3, sc["bsdf"],
                                    sc["fsdf"], levels)
    upload(solver, s, s.b[0])
    ok, _, rel = oracle.mgpcg_clean_solve(s, MgParams.make(max_levels=levels, restrict_iter=2,
                                                           correct_iter=2, coarsest_iter=20,
                                                           final_iter=2), 100, 1e-12)
    assert ok and rel <= 1e-12
    p_ref = s.x[0].copy()
    p = jetgpu.mg_params(max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2)
    info = solver.mgpcg(200, p, jetgpu.gpu_options(precision=precision,
                                                   smoother=jetgpu.JG_SMOOTH_JACOBI,
                                                   stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                   rel_tol=1e-10))
    assert info.converged and info.rel_residual <= 1e-8
    assert rel_err(solver.download_x(), p_ref) <= 1e-6
    # the oracle's Jacobi-preconditioned solve needs about as many iterations
    _, it_ref, _ = oracle.mgpcg_clean_solve_jacobi(s, MgParams.make(
        max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20, final_iter=2),
        6.0 / 7.0, 200, 1e-10)
    print(f"jacobi mgpcg iterations: gpu {info.iterations}, oracle {it_ref}")
    assert abs(int(info.iterations) - int(it_ref)) <= 3


def test_jacobi_weight_outside_range_is_rejected(solver):
    s = mg_poisson_hierarchy(4, 2)
    upload(solver, s, s.b[0])
    with pytest.raises(jetgpu.JgError, match="weight"):
        solver.mgpcg(10, jetgpu.mg_params(max_levels=2),
                     jetgpu.gpu_options(smoother=jetgpu.JG_SMOOTH_JACOBI, omega=1.2))


# ------------------------------------------------- one V-cycle (FdmMgSolver3)

@pytest.mark.parametrize("precision,tol", [(jetgpu.JG_FP64, 1e-12), (jetgpu.JG_MIXED, 2e-5)])
def test_mg_solve_is_one_cycle_from_x(solver, oracle, precision, tol):
    s = random_hierarchy((32, 32, 32), 4, seed=85)
    rng = np.random.default_rng(86)
    b = rng.uniform(-1, 1, s.x[0].shape)
    x0 = rng.uniform(-1, 1, s.x[0].shape)
    upload(solver, s, b)
    gp, op = both_params(max_levels=4, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2, max_tol=1e-9)
    x = x0.copy()
    info = solver.mg_host(s, b, x, gp, jetgpu.gpu_options(precision=precision))
    r0 = oracle.residual(s.A[0], x0, b)
    want = x0 + oracle.mg_clean_vcycle(s, op, r0)
    assert rel_err(x, want) <= tol
    r1 = np.linalg.norm(oracle.residual(s.A[0], x, b))
    assert info.iterations == 1 and abs(info.last_residual - r1) <= 1e-9 * r1
    assert bool(info.converged) == (info.last_residual < 1e-9)


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_mg_solve_jet_known_answer(solver, oracle, precision):
    # FdmMgSolver3.Solve: 6 levels from 4^3 (128^3 finest), FdmMgSolver3(levels, 5, 5, 20, 20, 1e-9)
    s = mg_poisson_hierarchy(4, 6)
    b = s.b[0]
    x = np.zeros_like(b)
    gp = jetgpu.mg_params(max_levels=6, restrict_iter=5, correct_iter=5, coarsest_iter=20,
                          final_iter=20, max_tol=1e-9)
    norm0 = np.linalg.norm(oracle.residual(s.A[0], x, b))
    info = solver.mg_host(s, b, x, gp, jetgpu.gpu_options(precision=precision))
    norm1 = np.linalg.norm(oracle.residual(s.A[0], x, b))
    print(f"mg known answer ({precision}): ||r|| {norm0:.4e} -> {norm1:.4e}")
    assert norm1 < norm0
    assert abs(info.last_residual - norm1) <= 1e-6 * norm1
    # the next cycle starts from this x (Jet does not zero it) and goes on
    info = solver.mg_host(s, b, x, gp, jetgpu.gpu_options(precision=precision))
    norm2 = np.linalg.norm(oracle.residual(s.A[0], x, b))
    assert norm2 < norm1
