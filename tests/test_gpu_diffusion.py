"""GPU backward Euler diffusion (jg_diffusion_build + jg_solver_cg) against
the oracle: the system bit for bit, the solution within the CG parity bar,
and Jet's GridBackwardEulerDiffusionSolver3.Solve answer."""
import numpy as np
import pytest

from jgtest import dev, jetgpu, torch_cuda
from test_oracle_diffusion import KAT_EDGE, KAT_MID, kat_input, random_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def device_system(solver, shape):
    nz, ny, nx = shape
    A = np.empty((nz, ny, nx, 4))
    b = np.empty((nz, ny, nx))
    jetgpu._memcpy_dtoh(A.reshape(-1), solver.matrix_ptr(0))
    jetgpu._memcpy_dtoh(b.reshape(-1), solver.rhs_ptr())
    return A, b


@pytest.mark.parametrize("n,dirichlet,seed", [((9, 7, 6), False, 1), ((9, 7, 6), True, 2),
                                              ((65, 33, 17), True, 3), ((1, 1, 1), True, 4),
                                              ((64, 64, 64), False, 5)])
def test_diffusion_system_bitwise(solver, oracle, n, dirichlet, seed):
    print(f"case diffusion build n={n} dirichlet={dirichlet}", flush=True)
    f, bs, fs = random_case(n, seed)
    h = (0.5, 0.25, 0.4)
    solver.set_levels([n])
    solver.diffusion_build(dev(f), dev(bs), dev(fs), 0.3, 0.7, h, dirichlet)
    A, b = device_system(solver, f.shape)
    Ao, bo = oracle.diffusion_build(f, bs, fs, 0.3, 0.7, h, dirichlet)
    assert np.array_equal(A, Ao) and np.array_equal(b, bo)


def test_diffusion_kat_on_gpu(solver):
    f, bs, fs = kat_input()
    solver.set_levels([(3, 3, 3)])
    solver.diffusion_build(dev(f), dev(bs), dev(fs), 1.0, 1.0 / 12.0, 1.0)
    solver.cg(200, 1e-12)
    x = solver.download_x()
    np.testing.assert_allclose(x[1], KAT_MID, atol=1e-6)
    np.testing.assert_allclose(x[0], KAT_EDGE, atol=1e-6)


@pytest.mark.parametrize("dirichlet", [False, True])
def test_diffusion_face_component_solve(solver, oracle, dirichlet):
    # the u component of a face-centred grid of 40 x 24 x 16 cells: a data
    # grid of 41 x 24 x 16 points, solved with the GPU CG against the oracle
    print(f"case diffusion u-faces dirichlet={dirichlet}", flush=True)
    f, bs, fs = random_case((41, 24, 16), 11)
    h = (1.0 / 40,) * 3
    solver.set_levels([(41, 24, 16)])
    solver.diffusion_build(dev(f), dev(bs), dev(fs), 0.01, 0.05, h, dirichlet)
    info = solver.cg(2000, 1e-12)
    x = solver.download_x()
    A, b = oracle.diffusion_build(f, bs, fs, 0.01, 0.05, h, dirichlet)
    xo, it, res, ok = oracle.cg_solve(A, b, 2000, 1e-12)
    assert abs(info.iterations - it) <= 2
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-9 * max(1.0, np.abs(xo).max()))
