"""GPU parity of the FdmBlas3 / relaxation / transfer kernels (through the C
ABI) against the CPU oracle on the same seeded inputs.

The elementwise kernels evaluate Jet's expressions in Jet's operand order
with floating-point contraction disabled, so their FP64 results must be
bit-identical to the oracle's (which is bit-identical to Jet's own build).
Dot products are reduced in a different (tree) order: they must agree to
1e-13 relative to sum |a_i b_i| and be deterministic run to run.
"""
import numpy as np
import pytest

from jgtest import dev, empty, host, jetgpu, torch_cuda
from systems import poisson_helper, random_stencil

pytestmark = pytest.mark.gpu

# (the last five: partial 64 x 8 x 16 tiles of the tiled z-march in x, y and z,
# one plane and one line)
SHAPES = [(1, 1, 1), (2, 3, 1), (7, 1, 3), (5, 9, 4), (64, 64, 8), (33, 17, 70), (130, 3, 5),
          (96, 20, 37), (200, 9, 33), (48, 13, 17), (100, 40, 1), (150, 1, 1)]


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_mvm_residual(solver, oracle, shape):
    rng = np.random.default_rng(11)
    nx, ny, nz = shape
    print(f"case mvm/residual shape={shape}", flush=True)
    A = random_stencil(nx, ny, nz, rng, spd=False)
    x = rng.uniform(-1, 1, (nz, ny, nx))
    b = rng.uniform(-1, 1, (nz, ny, nx))
    dA, dx, db, dy = dev(A), dev(x), dev(b), empty((nz, ny, nx))
    solver.mvm(shape, dA, dx, dy)
    assert np.array_equal(host(dy), oracle.mvm(A, x))
    solver.residual(shape, dA, dx, db, dy)
    assert np.array_equal(host(dy), oracle.residual(A, x, b))


def test_mvm_integer_exact(solver, oracle):
    A, _ = poisson_helper(40, 24, 16)
    x = np.random.default_rng(12).integers(-999, 999, (16, 24, 40)).astype(np.float64)
    y = empty((16, 24, 40))
    solver.mvm((40, 24, 16), dev(A), dev(x), y)
    assert np.array_equal(host(y), oracle.mvm(A, x))


@pytest.mark.parametrize("n", [1, 255, 4097, 1 << 20])
def test_dot_axpy(solver, oracle, n):
    rng = np.random.default_rng(13)
    a = rng.uniform(-1, 1, (1, 1, n))
    b = rng.uniform(-1, 1, (1, 1, n))
    got = solver.dot(n, dev(a), dev(b))
    want = oracle.dot(a, b)
    assert abs(got - want) <= 1e-13 * np.sum(np.abs(a * b))
    out = empty((1, 1, n))
    solver.axpy(n, -0.75, dev(a), dev(b), out)
    assert np.array_equal(host(out), oracle.axpy(-0.75, a, b))


def test_dot_is_deterministic(solver):
    rng = np.random.default_rng(14)
    a = dev(rng.uniform(-1, 1, 3_000_001))
    vals = {solver.dot(3_000_001, a, a) for _ in range(5)}
    assert len(vals) == 1


@pytest.mark.parametrize("shape", [(9, 8, 7), (64, 32, 16), (3, 1, 1)])
@pytest.mark.parametrize("sor", [1.0, 1.5])
def test_relax_red_black(solver, oracle, shape, sor):
    # FdmGaussSeidelSolver3::relaxRedBlack: each colour reads only the other
    # colour, so the GPU sweep (same operand order, no contraction) is bitwise Jet's.
    rng = np.random.default_rng(15)
    nx, ny, nz = shape
    A = random_stencil(nx, ny, nz, rng)
    b = rng.uniform(-1, 1, (nz, ny, nx))
    x = rng.uniform(-1, 1, (nz, ny, nx))
    dx = dev(x)
    solver.relax_red_black(shape, dev(A), dev(b), sor, dx)
    assert np.array_equal(host(dx), oracle.gs_relax(A, b, sor, x, red_black=True))


def test_relax_jacobi(solver, oracle):
    rng = np.random.default_rng(16)
    A = random_stencil(20, 12, 9, rng)
    b = rng.uniform(-1, 1, (9, 12, 20))
    x = rng.uniform(-1, 1, (9, 12, 20))
    out = empty((9, 12, 20))
    solver.relax_jacobi((20, 12, 9), dev(A), dev(b), dev(x), out)
    assert np.array_equal(host(out), oracle.jacobi_relax(A, b, x))


@pytest.mark.parametrize("fshape", [(2, 2, 2), (8, 6, 4), (64, 64, 64), (34, 10, 22), (6, 4, 8)])
def test_restrict_correct(solver, oracle, fshape):
    rng = np.random.default_rng(17)
    nx, ny, nz = fshape
    f = rng.uniform(-1, 1, (nz, ny, nx))
    c = rng.uniform(-1, 1, (nz // 2, ny // 2, nx // 2))
    out = empty((nz // 2, ny // 2, nx // 2))
    solver.restrict((nx // 2, ny // 2, nz // 2), dev(f), out)
    assert np.array_equal(host(out), oracle.restrict(f))
    df = dev(f)
    solver.correct(fshape, dev(c), df)
    # Jet clamps the prolongation's k index with n.y: (6, 4, 8) exercises the
    # in-bounds case (n.y < n.z) that must match Jet; (8, 6, 4) the case where
    # Jet would read past the coarse array, clamped at n.z on both sides.
    assert np.array_equal(host(df), oracle.correct(c, f))


def test_invalid_arguments_are_errors(solver):
    with pytest.raises(jetgpu.JgError):
        solver.correct((3, 4, 4), 0, 0)
    with pytest.raises(jetgpu.JgError):
        solver.set_levels([(8, 8, 8), (3, 4, 4)])
