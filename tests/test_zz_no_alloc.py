"""The no-allocation red-black relax (jg_relax_red_black_noalloc) against the
allocating entry point it mirrors (jg_relax_red_black), through the C ABI.

Both run the same sweep kernel on the same inputs, so their results must be
identical bit for bit; the oracle's relaxRedBlack
(src/jet/fdm_gauss_seidel_solver3.cpp:150-221) is checked too.  Two calls in a
row, the second on a smaller grid, show that the shared scratch array holds
nothing over from the previous call; the size limit is checked at and past
2^21 cells.
"""
import numpy as np
import pytest

from jgtest import dev, host, jetgpu, torch_cuda
from systems import random_stencil

pytestmark = pytest.mark.gpu

MAX_CELLS = 1 << 21


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def _case(shape, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    A = random_stencil(nx, ny, nz, rng)
    b = rng.uniform(-1, 1, (nz, ny, nx))
    x = rng.uniform(-1, 1, (nz, ny, nx))
    return A, b, x


def _both(solver, shape, A, b, x, sor):
    dA, db = dev(A), dev(b)
    ref = dev(x)
    solver.relax_red_black(shape, dA, db, sor, ref)
    got = dev(x)
    solver.relax_red_black_noalloc(shape, dA, db, sor, got)
    return host(got), host(ref)


@pytest.mark.parametrize("sor", [1.0, 1.5])
def test_noalloc_matches_allocating_entry_twice(solver, oracle, sor):
    # a larger grid first, then a smaller one: the second call must not see
    # the first call's copy of x left in the scratch array
    for shape, seed in (((64, 48, 40), 71), ((9, 8, 7), 72)):
        print(f"case relax_red_black_noalloc shape={shape} sor={sor}", flush=True)
        A, b, x = _case(shape, seed)
        got, ref = _both(solver, shape, A, b, x, sor)
        assert np.array_equal(got, ref), shape
        assert np.array_equal(got, oracle.gs_relax(A, b, sor, x, red_black=True)), shape


def test_noalloc_at_the_size_limit(solver):
    shape = (128, 128, 128)  # exactly 2^21 cells
    print(f"case relax_red_black_noalloc shape={shape}", flush=True)
    A, b, x = _case(shape, 73)
    got, ref = _both(solver, shape, A, b, x, 1.3)
    assert np.array_equal(got, ref)
    # and a small grid right after the largest one
    A, b, x = _case((5, 3, 2), 74)
    got, ref = _both(solver, (5, 3, 2), A, b, x, 1.3)
    assert np.array_equal(got, ref)


def test_noalloc_rejects_grids_past_the_scratch(solver):
    import torch

    n = MAX_CELLS + 128 * 128  # 129 x 128 x 128: rejected before any launch
    A = torch.zeros((n * 4,), dtype=torch.float64, device="cuda")
    b = torch.zeros((n,), dtype=torch.float64, device="cuda")
    x = torch.ones((n,), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(jetgpu.JgError, match="scratch"):
        solver.relax_red_black_noalloc((128, 128, 129), A, b, 1.0, x)
    assert bool(torch.all(x == 1.0))
