"""GPU 2-D multigrid (FdmMgLinearSystem2 / FdmMgpcgSolver2 through the same
solver: a hierarchy of one-plane levels halving in x and y) against the
oracle: FdmMgUtils2's transfers bit for bit, the V-cycle against the
oracle's clean-mode V-cycle on the same 2-D hierarchy, and MGPCG solutions
against the oracle's FP64 solve."""
import numpy as np
import pytest

from jgtest import dev, empty, host, jetgpu, rel_err, torch_cuda
from systems import mg_poisson_hierarchy2, random_stencil

pytestmark = pytest.mark.gpu

HIER2 = [((64, 64), 4), ((96, 32), 3), ((256, 128), 5), ((30, 18), 2), ((1024, 512), 6)]
SCHEDULES = [(2, 2, 20, 2), (1, 1, 4, 1), (0, 2, 6, 3), (2, 0, 6, 0)]


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def hierarchy2(shape, levels, seed):
    from oracle.jet_oracle import MgSystem

    rng = np.random.default_rng(seed)
    s = MgSystem()
    nx, ny = shape
    for l in range(levels):
        A = random_stencil(nx >> l, ny >> l, 1, rng)
        s.A.append(A)
        s.b.append(np.zeros(A.shape[:3]))
        s.x.append(np.zeros(A.shape[:3]))
    return s


def upload(solver, s, b=None):
    solver.set_levels([(x.shape[2], x.shape[1], 1) for x in s.x])
    for l, A in enumerate(s.A):
        solver.upload(l, A, b if l == 0 else None)


def params(sched, sor=1.3):
    from oracle.jet_oracle import MgParams

    pre, post, coarsest, final = sched
    kw = dict(restrict_iter=pre, correct_iter=post, coarsest_iter=coarsest, final_iter=final, sor=sor)
    return jetgpu.mg_params(**kw), MgParams.make(**kw)


@pytest.mark.parametrize("cshape", [(1, 1), (2, 3), (16, 8), (3, 2), (65, 33)])
def test_restrict2_correct2_bitwise(solver, oracle, cshape):
    print(f"case transfers2 cshape={cshape}", flush=True)
    cnx, cny = cshape
    rng = np.random.default_rng(51)
    f = rng.uniform(-1, 1, (2 * cny, 2 * cnx))
    c = rng.uniform(-1, 1, (cny, cnx))
    out = empty((cny, cnx))
    solver.restrict2(cshape, dev(f), out)
    assert np.array_equal(host(out), oracle.restrict2(f))
    fd = dev(f)
    solver.correct2((2 * cnx, 2 * cny), dev(c), fd)
    assert np.array_equal(host(fd), oracle.correct2(c, f))


@pytest.mark.parametrize("shape,levels", HIER2)
@pytest.mark.parametrize("sched", SCHEDULES)
def test_vcycle2_matches_oracle(solver, oracle, shape, levels, sched):
    print(f"case vcycle2 shape={shape} levels={levels} sched={sched}", flush=True)
    s = hierarchy2(shape, levels, seed=52)
    upload(solver, s)
    gp, op = params(sched)
    r = np.random.default_rng(53).uniform(-1, 1, s.x[0].shape)
    want = oracle.mg_clean_vcycle(s, op, r)
    for prec, tol in ((jetgpu.JG_FP64, 1e-12), (jetgpu.JG_MIXED, 2e-5)):
        z = empty(r.shape)
        solver.mg_vcycle(gp, jetgpu.gpu_options(precision=prec), dev(r), z)
        assert rel_err(host(z), want) <= tol, prec


def apply2(solver, gp, precision, r, **kw):
    z = empty(r.shape)
    solver.mg_vcycle(gp, jetgpu.gpu_options(precision=precision, **kw), dev(r), z)
    return host(z)


@pytest.mark.parametrize("shape,levels", [((256, 128), 5), ((360, 200), 3), ((1024, 512), 6),
                                          ((500, 300), 3), ((130, 66), 2), ((64, 64), 4),
                                          ((30, 18), 2)])
@pytest.mark.parametrize("sched", [(2, 2, 20, 2), (3, 3, 6, 3), (2, 4, 6, 4), (1, 2, 4, 0),
                                   (0, 2, 6, 3), (2, 0, 6, 0), (2, 1, 5, 1), (4, 4, 7, 4)])
def test_vcycle2_two_sweep_kernel_bitwise(solver, oracle, shape, levels, sched):
    # Two sweeps per launch (k_rb_sweep2_plane) must give exactly the
    # single-sweep result (k_rb_sweep_plane, k_residual_plane, k_correct2):
    # pairs that start from zero, from x, or from x plus the prolonged
    # correction; the last pre-smoothing pair writing the residual (118 x 86
    # FP32 / 118 x 54 FP64 interiors, 120 x 88 / 120 x 56 without it);
    # partial tiles in x and y; odd counts leaving a single sweep first (pre)
    # or last (post, coarsest).
    print(f"case two-sweep 2-D shape={shape} levels={levels} sched={sched}", flush=True)
    s = hierarchy2(shape, levels, seed=58)
    upload(solver, s)
    gp, op = params(sched)
    r = np.random.default_rng(59).uniform(-1, 1, s.x[0].shape)
    want = oracle.mg_clean_vcycle(s, op, r)
    for prec, tol in ((jetgpu.JG_MIXED, 2e-5), (jetgpu.JG_FP64, 1e-12)):
        single = apply2(solver, gp, prec, r, sweep_pairs=-1)
        paired = apply2(solver, gp, prec, r, sweep_pairs=1)
        assert np.array_equal(single, paired), prec
        assert np.array_equal(apply2(solver, gp, prec, r), paired), prec
        assert rel_err(paired, want) <= tol, prec


def test_vcycle2_zero_centre_rows(solver, oracle):
    # all-zero rows (a fluid cell walled in by solids) are set to zero by both
    # one-plane smoothers, on every level
    s = hierarchy2((256, 192), 4, seed=60)
    rng = np.random.default_rng(61)
    for A in s.A:
        mask = rng.random(A.shape[:3]) < 0.02
        A[mask] = 0.0
        A[:, :, 1:, 1][mask[:, :, 1:]] = 0.0
        A[:, :, :-1, 1][mask[:, :, 1:]] = 0.0
        A[:, 1:, :, 2][mask[:, 1:, :]] = 0.0
        A[:, :-1, :, 2][mask[:, 1:, :]] = 0.0
    upload(solver, s)
    gp, op = params((2, 2, 20, 2))
    r = rng.uniform(-1, 1, s.x[0].shape)
    zero = s.A[0][..., 0] == 0.0
    r[zero] = 0.0
    want = oracle.mg_clean_vcycle(s, op, r)
    assert np.isfinite(want).all() and np.all(want[zero] == 0.0)
    for prec, tol in ((jetgpu.JG_FP64, 1e-12), (jetgpu.JG_MIXED, 2e-5)):
        single = apply2(solver, gp, prec, r, sweep_pairs=-1)
        paired = apply2(solver, gp, prec, r)
        assert np.array_equal(single, paired)
        assert np.isfinite(paired).all() and np.all(paired[zero] == 0.0)
        assert rel_err(paired, want) <= tol


@pytest.mark.parametrize("precision,tol", [(jetgpu.JG_FP64, 1e-12), (jetgpu.JG_MIXED, 1e-5)])
def test_vcycle2_symmetric_positive_linear(solver, precision, tol):
    print(f"case vcycle2 spd precision={precision}", flush=True)
    s = hierarchy2((128, 64), 4, seed=54)
    upload(solver, s)
    gp, _ = params((2, 2, 20, 2))
    rng = np.random.default_rng(55)
    r1, r2 = rng.uniform(-1, 1, (2,) + s.x[0].shape)

    def apply(r):
        z = empty(r.shape)
        solver.mg_vcycle(gp, jetgpu.gpu_options(precision=precision), dev(r), z)
        return host(z)

    z1, z2 = apply(r1), apply(r2)
    scale = np.linalg.norm(r1) * np.linalg.norm(z2)
    assert abs(np.vdot(r1, z2) - np.vdot(r2, z1)) <= tol * scale
    assert np.vdot(r1, z1) > 0 and np.vdot(r2, z2) > 0
    assert rel_err(apply(2.0 * r1 + r2), 2.0 * z1 + z2) <= 10 * tol
    assert np.array_equal(apply(r1), z1)


@pytest.mark.parametrize("shape,levels", [((128, 128), 6), ((1024, 1024), 8), ((320, 96), 4),
                                          ((2048, 2048), 8)])
@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_mgpcg2_matches_oracle_solution(solver, oracle, shape, levels, precision):
    # Random SPD one-plane hierarchy, solved to 1e-10 relative residual: the
    # solution agrees with the oracle's FP64 clean-mode solve to 1e-10 and
    # the iteration counts are close (the same preconditioner).
    print(f"case mgpcg2 shape={shape} levels={levels} precision={precision}", flush=True)
    s = hierarchy2(shape, levels, seed=56)
    b = np.random.default_rng(57).uniform(-1, 1, s.x[0].shape)
    s.b[0] = b
    upload(solver, s, b)
    gp, op = params((2, 2, 20, 2), sor=1.0)
    info = solver.mgpcg(100, gp, jetgpu.gpu_options(precision=precision,
                                                    stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                    rel_tol=1e-10))
    assert info.rel_residual <= 1e-10
    ok, it, rel = oracle.mgpcg_clean_solve(s, op, 100, 1e-12)
    assert ok
    x = solver.download_x()
    assert rel_err(x, s.x[0]) <= 1e-8
    if precision == jetgpu.JG_FP64:
        assert info.iterations <= it


def test_mgpcg2_jet_kat(solver):
    # FdmMgpcgSolver2.Solve (src/tests/unit_tests/fdm_mgpcg_solver2_tests.cpp:
    # 13-49) on the GPU: 6 levels from 4 x 4, MGPCG(200, 6, 5, 5, 10, 10,
    # 1e-4) converges (Jet's absolute residual rule), in FP64 and mixed.
    s = mg_poisson_hierarchy2(4, 6)
    upload(solver, s, s.b[0])
    gp = jetgpu.mg_params(max_levels=6, restrict_iter=5, correct_iter=5, coarsest_iter=10,
                          final_iter=10, max_tol=1e-4)
    for prec in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
        info = solver.mgpcg(200, gp, jetgpu.gpu_options(precision=prec))
        print(f"mgpcg2 KAT precision={prec}: {info.iterations} iterations, "
              f"residual {info.last_residual:.3e}", flush=True)
        assert info.converged and info.last_residual <= 1e-4 and info.iterations < 200


def test_plane_limit(solver):
    # one plane is reduced in 64 x 4 tiles with at most 65536 partials:
    # 4096^2 is accepted, 8192 x 4096 is rejected before anything is allocated
    solver.set_levels([(4096, 4096, 1), (2048, 2048, 1)])
    with pytest.raises(jetgpu.JgError):
        solver.set_levels([(8192, 4096, 1)])


@pytest.mark.parametrize("precision", [jetgpu.JG_FP64, jetgpu.JG_MIXED])
def test_mg_solve2_jet_kat(solver, oracle, precision):
    # FdmMgSolver2.Solve (src/tests/unit_tests/fdm_mg_solver2_tests.cpp:13-62):
    # one V-cycle from x = 0 with FdmMgSolver2(6, 5, 5, 20, 20) on Jet's test
    # system reduces the residual norm; here the GPU V-cycle of the same
    # hierarchy (z = M b is x after one cycle from zero).
    s = mg_poisson_hierarchy2(4, 6)
    upload(solver, s)
    b = s.b[0]
    gp = jetgpu.mg_params(max_levels=6, restrict_iter=5, correct_iter=5, coarsest_iter=20,
                          final_iter=20, max_tol=1e-9)
    z = empty(b.shape)
    solver.mg_vcycle(gp, jetgpu.gpu_options(precision=precision), dev(b), z)
    x = host(z)
    norm0 = np.linalg.norm(b)
    norm1 = np.linalg.norm(oracle.residual(s.A[0], x, b))
    print(f"mg_solve2 KAT precision={precision}: |r| {norm0:.3e} -> {norm1:.3e}", flush=True)
    assert norm1 < norm0
