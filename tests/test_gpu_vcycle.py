"""GPU V-cycle (the MGPCG preconditioner, z = M r) against the oracle's
clean-mode V-cycle on the same hierarchy, plus the properties PCG relies on.

The GPU evaluates the same sequence of operators as the oracle (Jet's
relaxRedBlack, restriction and prolongation-correction; symmetric sweep
order), but through different kernels: fused sweeps with the prolongation
evaluated separably, and the lower levels in one workgroup (k_bottom) with
w / c precomputed.  FP64 results therefore agree to rounding (<= 1e-12
relative); the FP32 (mixed-precision) V-cycle to FP32 rounding (<= 2e-5).
The hierarchies cover every code path: levels handled by the multi-kernel
path and by the one-workgroup bottom solve (FP32 bottoms start at 4096 cells,
FP64 at 2048), a V-cycle that is entirely one k_bottom launch, a coarsest
level too large for it, non-cubic grids with n.y > n.z (the prolongation's
k clamp), no pre-smoothing (restriction of b, zero-initial fused correction)
and no post-smoothing (separate correction kernel).  Two consecutive sweeps
may run as one launch (k_rb_sweep2 in FP32, the brick kernel on small levels
in both precisions); those paths are checked to be bit-identical to the
single sweeps.
"""
import numpy as np
import pytest

from jgtest import dev, empty, host, jetgpu, rel_err, torch_cuda
from systems import random_stencil

pytestmark = pytest.mark.gpu

HIER = [
    ((32, 32, 32), 4),
    ((48, 32, 16), 3),
    ((32, 64, 16), 3),
    ((16, 16, 16), 3),
    ((8, 8, 8), 2),
    ((64, 64, 64), 5),
    ((32, 32, 32), 1),
    ((24, 16, 8), 2),
]
# (restrict_iter, correct_iter, coarsest_iter, final_iter)
SCHEDULES = [(2, 2, 20, 2), (1, 1, 4, 1), (0, 2, 6, 3), (2, 0, 6, 0)]


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
    for l in range(levels):
        A = random_stencil(nx >> l, ny >> l, nz >> l, rng)
        s.A.append(A)
        s.b.append(np.zeros(A.shape[:3]))
        s.x.append(np.zeros(A.shape[:3]))
    return s


def upload(solver, s):
    solver.set_levels([(x.shape[2], x.shape[1], x.shape[0]) for x in s.x])
    for l, A in enumerate(s.A):
        solver.upload(l, A, None)


def params(sched, sor):
    from oracle.jet_oracle import MgParams

    pre, post, coarsest, final = sched
    kw = dict(restrict_iter=pre, correct_iter=post, coarsest_iter=coarsest, final_iter=final, sor=sor)
    return jetgpu.mg_params(**kw), MgParams.make(**kw)


def apply(solver, gp, precision, r, **kw):
    z = empty(r.shape)
    solver.mg_vcycle(gp, jetgpu.gpu_options(precision=precision, **kw), dev(r), z)
    return host(z)


@pytest.mark.parametrize("shape,levels", HIER)
@pytest.mark.parametrize("sched", SCHEDULES)
def test_vcycle_matches_oracle(solver, oracle, shape, levels, sched):
    print(f"case vcycle shape={shape} levels={levels} sched={sched}", flush=True)
    s = random_hierarchy(shape, levels, seed=31)
    upload(solver, s)
    gp, op = params(sched, 1.3)
    r = np.random.default_rng(32).uniform(-1, 1, s.x[0].shape)
    want = oracle.mg_clean_vcycle(s, op, r)
    assert rel_err(apply(solver, gp, jetgpu.JG_FP64, r), want) <= 1e-12
    assert rel_err(apply(solver, gp, jetgpu.JG_MIXED, r), want) <= 2e-5


@pytest.mark.parametrize("shape,levels", [((32, 32, 32), 4), ((48, 32, 16), 3), ((16, 16, 16), 3)])
@pytest.mark.parametrize("precision,tol", [(jetgpu.JG_FP64, 1e-12), (jetgpu.JG_MIXED, 1e-5)])
def test_vcycle_symmetric_positive_linear(solver, shape, levels, precision, tol):
    # PCG needs M symmetric positive definite: r1.M r2 = r2.M r1, r.M r > 0,
    # and M linear (z = M r is a fixed linear map, no warm start).
    print(f"case vcycle spd shape={shape} precision={precision}", flush=True)
    s = random_hierarchy(shape, levels, seed=33)
    upload(solver, s)
    gp, _ = params((2, 2, 20, 2), 1.3)
    rng = np.random.default_rng(34)
    r1 = rng.uniform(-1, 1, s.x[0].shape)
    r2 = rng.uniform(-1, 1, s.x[0].shape)
    z1 = apply(solver, gp, precision, r1)
    z2 = apply(solver, gp, precision, r2)
    scale = np.linalg.norm(r1) * np.linalg.norm(z2)
    assert abs(np.vdot(r1, z2) - np.vdot(r2, z1)) <= tol * scale
    assert np.vdot(r1, z1) > 0 and np.vdot(r2, z2) > 0
    z12 = apply(solver, gp, precision, 2.0 * r1 + r2)
    assert rel_err(z12, 2.0 * z1 + z2) <= 10 * tol
    # the same input gives the same output (no state carried between calls)
    assert np.array_equal(apply(solver, gp, precision, r1), z1)


@pytest.mark.parametrize("shape,levels", [((64, 64, 64), 5), ((48, 32, 16), 3), ((32, 64, 16), 3),
                                          ((40, 24, 20), 2), ((120, 56, 36), 2), ((96, 64, 48), 3),
                                          ((34, 18, 10), 2), ((130, 70, 44), 2), ((108, 44, 60), 3)])
@pytest.mark.parametrize("sched", [(2, 2, 20, 2), (3, 3, 6, 3), (2, 4, 6, 4), (1, 2, 4, 0)])
def test_vcycle_two_sweep_kernel_bitwise(solver, oracle, shape, levels, sched):
    # Two sweeps per launch wherever two sweeps follow each other must give
    # exactly the single-sweep result: in FP32 the z-marching k_rb_sweep2
    # (levels above 64^3 cells) and the 8^3-brick kernel (below), in FP64 the
    # brick kernel.  Where the pre-smoothing ends with a z-marching pair, that
    # launch also writes the residual for the restriction (54 x 22 interiors),
    # which must equal the separate residual-restriction's bit for bit.
    # Shapes cover partial tiles in x (56- and 54-wide interiors), y (24 and
    # 22 rows) and bricks (sizes not multiples of 8), z extents that are not
    # multiples of the chunk, and pairs that start with a zero-start, a
    # correction or a plain sweep.
    print(f"case two-sweep shape={shape} levels={levels} sched={sched}", flush=True)
    s = random_hierarchy(shape, levels, seed=41)
    upload(solver, s)
    gp, op = params(sched, 1.3)
    r = np.random.default_rng(42).uniform(-1, 1, s.x[0].shape)
    want = oracle.mg_clean_vcycle(s, op, r)
    for prec, tol in ((jetgpu.JG_MIXED, 2e-5), (jetgpu.JG_FP64, 1e-12)):
        single = apply(solver, gp, prec, r, sweep_pairs=-1)
        paired = apply(solver, gp, prec, r, sweep_pairs=1)
        assert np.array_equal(single, paired), prec
        assert rel_err(paired, want) <= tol


@pytest.mark.parametrize("shape,levels", [((32, 32, 32), 4), ((16, 16, 16), 3)])
def test_vcycle_zero_centre_rows(solver, oracle, shape, levels):
    # Fluid cells whose neighbours are all solid have all-zero rows in Jet's
    # single-phase build; the smoothers set them to zero instead of dividing
    # by the zero centre, on every level.
    print(f"case vcycle zero rows shape={shape}", flush=True)
    s = random_hierarchy(shape, levels, seed=36)
    rng = np.random.default_rng(37)
    for A in s.A:
        nz, ny, nx = A.shape[:3]
        mask = rng.random((nz, ny, nx)) < 0.02
        A[mask] = 0.0
        # the neighbours' couplings toward a zero row vanish too (symmetry)
        A[:, :, 1:, 1][mask[:, :, 1:]] = 0.0
        A[:, :, :-1, 1][mask[:, :, 1:]] = 0.0
        A[:, 1:, :, 2][mask[:, 1:, :]] = 0.0
        A[:, :-1, :, 2][mask[:, 1:, :]] = 0.0
        A[1:, :, :, 3][mask[1:, :, :]] = 0.0
        A[:-1, :, :, 3][mask[1:, :, :]] = 0.0
    upload(solver, s)
    gp, op = params((2, 2, 20, 2), 1.3)
    r = rng.uniform(-1, 1, s.x[0].shape)
    zero = s.A[0][..., 0] == 0.0
    r[zero] = 0.0
    want = oracle.mg_clean_vcycle(s, op, r)
    assert np.isfinite(want).all() and np.all(want[zero] == 0.0)
    for prec, tol in ((jetgpu.JG_FP64, 1e-12), (jetgpu.JG_MIXED, 2e-5)):
        got = apply(solver, gp, prec, r)
        assert np.isfinite(got).all() and np.all(got[zero] == 0.0)
        assert rel_err(got, want) <= tol


def test_vcycle_invalid_arguments(solver):
    s = random_hierarchy((8, 8, 8), 2, seed=35)
    upload(solver, s)
    gp, _ = params((2, 2, 4, 2), 1.3)
    r = dev(np.ones((8, 8, 8)))
    with pytest.raises(jetgpu.JgError):
        solver.mg_vcycle(gp, None, r, r)  # r must not alias z
    bad = jetgpu.mg_params(sor=2.0)
    with pytest.raises(jetgpu.JgError):
        solver.mg_vcycle(bad, None, r, empty((8, 8, 8)))
