"""GPU compressed (CSR) path -- FdmCompressedBlas3 mvm / residual and
FdmCgSolver3::solveCompressed through the C ABI -- against the oracle.

mvm and residual are bitwise (same row sums in the same order); CG follows
the solver parity bar: iterations within 2 of the oracle's (the dot-product
summation order differs) and the solution within 1e-8."""
import numpy as np
import pytest

from jgtest import jetgpu, torch_cuda
from systems import poisson_csr_kat

pytestmark = pytest.mark.gpu

BIG = np.finfo(np.float64).max


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def pressure_csr(oracle, n, seed, fluid=0.7, h=0.5, solid=True):
    """Random markers (fluid / air / boundary) and faces, compressed build.
    solid=False leaves out boundary cells: then every fluid row has a
    positive centre and the system is SPD (a fluid cell walled in by
    boundary cells is an all-zero row, on which CG divides by zero, in Jet
    as here)."""
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    other = np.where(rng.random((nz, ny, nx)) < 0.5, 1, 2) if solid else 1
    m = np.where(rng.random((nz, ny, nx)) < fluid, 0, other).astype(np.int8)
    u = rng.uniform(-1, 1, (nz, ny, nx + 1))
    v = rng.uniform(-1, 1, (nz, ny + 1, nx))
    w = rng.uniform(-1, 1, (nz + 1, ny, nx))
    return m, (u, v, w), oracle.sp_build_compressed(m, u, v, w, (h, h, h))


@pytest.mark.parametrize("n,seed", [((14, 11, 9), 1), ((33, 17, 5), 2), ((64, 64, 64), 3), ((1, 1, 1), 4)])
def test_csr_mvm_residual_bitwise(solver, oracle, n, seed):
    print(f"case csr blas n={n}", flush=True)
    m, _, (val, rp, ci, b) = pressure_csr(oracle, n, seed)
    if b.size == 0:
        pytest.skip("no fluid rows")
    x = np.random.default_rng(seed + 10).uniform(-1, 1, b.size)
    assert np.array_equal(solver.csr_mvm(val, rp, ci, x), oracle.csr_mvm(val, rp, ci, x))
    assert np.array_equal(solver.csr_residual(val, rp, ci, x, b), oracle.csr_residual(val, rp, ci, x, b))


@pytest.mark.parametrize("n,seed,tol", [((14, 11, 9), 1, 1e-10), ((40, 32, 24), 2, 1e-9),
                                        ((64, 64, 64), 3, 1e-8)])
def test_csr_cg_matches_oracle(solver, oracle, n, seed, tol):
    print(f"case csr cg n={n}", flush=True)
    _, _, (val, rp, ci, b) = pressure_csr(oracle, n, seed, solid=False)
    x, info = solver.csr_cg(val, rp, ci, b, 2000, tol)
    xo, oko, ito, reso = oracle.csr_cg_solve(val, rp, ci, b, 2000, tol)
    assert bool(info.converged) == oko and abs(info.iterations - ito) <= 2
    assert np.abs(x - xo).max() <= 1e-8 * max(1.0, np.abs(xo).max())
    # the reported relative residual is the true one
    r = oracle.csr_residual(val, rp, ci, x, b)
    assert abs(info.rel_residual - np.linalg.norm(r) / np.linalg.norm(b)) <= 1e-10


def test_csr_cg_kats(solver, oracle):
    # Jet's FdmCgSolver3.SolveCompressed: residual below the tolerance
    val, rp, ci, b = poisson_csr_kat((3, 3, 3))
    x, info = solver.csr_cg(val, rp, ci, b, 100, 1e-9)
    assert info.converged and info.last_residual < 1e-9
    # Jet's SolveFreeSurfaceCompressed: p = 2 - j on the fluid rows
    n = (3, 3, 3)
    u, v, w = np.zeros((3, 3, 4)), np.zeros((3, 4, 3)), np.zeros((4, 3, 3))
    v[:, 1:3, :] = 1.0
    z, y, xx = np.meshgrid(np.arange(3) + 0.5, np.arange(3) + 0.5, np.arange(3) + 0.5, indexing="ij")
    m = oracle.sp_markers(np.full(n, BIG), np.ascontiguousarray(y - 2.0))
    val, rp, ci, b = oracle.sp_build_compressed(m, u, v, w, (1.0, 1.0, 1.0))
    p, info = solver.csr_cg(val, rp, ci, b, 100, 1e-12)
    np.testing.assert_allclose(p, 2.0 - np.argwhere(m == 0)[:, 1], atol=1e-6)


def test_csr_cg_equals_flat_cg_on_fluid_cells(solver, oracle):
    # The compressed system is the flat one restricted to fluid cells, so
    # both GPU CG paths reach the same pressure there.
    print("case csr vs flat", flush=True)
    n = (24, 20, 16)
    m, (u, v, w), (val, rp, ci, b) = pressure_csr(oracle, n, 5, solid=False)
    A, bf = oracle.sp_build_system(m, u, v, w, (0.5, 0.5, 0.5))
    xc, ic = solver.csr_cg(val, rp, ci, b, 2000, 1e-11)
    flat = jetgpu.Solver(0)
    try:
        xf, inf = flat.cg_host(A, bf, 2000, 1e-11)
    finally:
        flat.close()
    np.testing.assert_allclose(xc, xf[m == 0], rtol=0, atol=1e-7 * max(1.0, np.abs(xc).max()))


def test_csr_rejects_malformed_matrices(solver):
    val = np.array([1.0, 2.0])
    with pytest.raises(jetgpu.JgError):
        solver.csr_cg(val, np.array([0, 1, 3], np.uint64), np.array([0, 1], np.uint64),
                      np.ones(2), 10, 1e-6)  # row pointers do not end at nnz
    with pytest.raises(jetgpu.JgError):
        solver.csr_cg(val, np.array([0, 1, 2], np.uint64), np.array([0, 5], np.uint64),
                      np.ones(2), 10, 1e-6)  # column outside the matrix
    with pytest.raises(jetgpu.JgError):
        solver.csr_cg(val, np.array([0, 2, 1], np.uint64), np.array([0, 1], np.uint64),
                      np.ones(2), 10, 1e-6)  # decreasing row pointers


# ------------------------------------------------- device compressed build
from jgtest import dev  # noqa: E402


@pytest.mark.parametrize("kind,n", [(1, 32), (2, 48), (3, 64), (4, 40)])
def test_sp_build_compressed_matches_oracle(solver, oracle, kind, n):
    print(f"case csr build kind={kind} n={n}", flush=True)
    d = jetgpu.scene_desc(kind, n, seed=6)
    s = jetgpu.generate_scene_host(d)
    m = oracle.sp_markers(s["bsdf"], s["fsdf"])
    val, rp, ci, b = oracle.sp_build_compressed(m, s["u"], s["v"], s["w"], (1.0,) * 3)
    solver.set_levels([(n, n, n)])
    rows, nnz = solver.sp_build_compressed(dev(s["u"]), dev(s["v"]), dev(s["w"]), dev(s["bsdf"]),
                                           dev(s["fsdf"]), 1.0)
    assert (rows, nnz) == (b.size, val.size)
    gval, grp, gci, gb, _ = solver.csr_download(rows, nnz)
    assert np.array_equal(gval, val) and np.array_equal(gb, b)
    assert np.array_equal(grp.astype(np.uint64), rp) and np.array_equal(gci.astype(np.uint64), ci)


def test_compressed_projection_matches_flat(solver, oracle):
    # build -> CG -> decompress -> gradient on the device, against the flat
    # CG of the same scene: the same pressure on fluid cells (1e-7), zero
    # elsewhere, and the same projected velocity.
    print("case csr projection", flush=True)
    n = 48
    d = jetgpu.scene_desc(2, n, seed=7)
    s = jetgpu.generate_scene_host(d)
    u, v, w, bs, fs = (dev(s[k]) for k in ("u", "v", "w", "bsdf", "fsdf"))
    solver.set_levels([(n, n, n)])
    rows, nnz = solver.sp_build_compressed(u, v, w, bs, fs, 1.0)
    info = solver.csr_solve(4000, 1e-10)
    assert info.converged
    solver.csr_decompress()
    pc = solver.download_x()
    m = oracle.sp_markers(s["bsdf"], s["fsdf"])
    assert np.all(pc[m != 0] == 0.0)
    flat = jetgpu.Solver(0)
    try:
        flat.set_levels([(n, n, n)])
        flat.sp_build(u, v, w, bs, fs, 1.0)
        fi = flat.cg(4000, 1e-10)
        pf = flat.download_x()
    finally:
        flat.close()
    scale = max(1.0, np.abs(pf[m == 0]).max())
    np.testing.assert_allclose(pc[m == 0], pf[m == 0], rtol=0, atol=1e-7 * scale)
    assert abs(info.iterations - fi.iterations) <= 2


def test_compressed_free_surface_kat_on_device(solver, oracle):
    # Jet's SolveFreeSurfaceCompressed through the resident path
    n = 3
    u, v, w = np.zeros((3, 3, 4)), np.zeros((3, 4, 3)), np.zeros((4, 3, 3))
    v[:, 1:3, :] = 1.0
    z, y, x = np.meshgrid(np.arange(3) + 0.5, np.arange(3) + 0.5, np.arange(3) + 0.5, indexing="ij")
    bs = np.full((3, 3, 3), BIG)
    fs = np.ascontiguousarray(y - 2.0)
    solver.set_levels([(n, n, n)])
    rows, nnz = solver.sp_build_compressed(dev(u), dev(v), dev(w), dev(bs), dev(fs), 1.0)
    assert rows == 18
    solver.csr_solve(100, 1e-12)
    solver.csr_decompress()
    p = solver.download_x()
    for j in range(2):
        np.testing.assert_allclose(p[:, j, :], 2.0 - j, atol=1e-6)
    assert np.all(p[:, 2, :] == 0.0)
