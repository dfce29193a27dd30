"""Small linear systems used by the tests (numpy, Jet layout).

Matrices are (nz, ny, nx, 4) arrays of FdmMatrixRow3 {center, right, up,
front}; vectors are (nz, ny, nx).
"""
import numpy as np


def poisson_helper(nx, ny, nz, scale=1.0, b_scale=1.0):
    """Jet's FdmLinearSystemSolverTestHelper3::buildTestLinearSystem
    (src/tests/unit_tests/fdm_linear_system_solver_test_helper3.h:13-49),
    optionally scaled per level as in fdm_mgpcg_solver3_tests.cpp:20-55
    (A by invdx^2, b by invdx).  All-Neumann; b = +1 at j=0, -1 at j=ny-1,
    so the solution is x = c - j."""
    A = np.zeros((nz, ny, nx, 4))
    b = np.zeros((nz, ny, nx))
    i = np.arange(nx)[None, None, :]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[:, None, None]
    c = np.zeros((nz, ny, nx))
    for t in ((i > 0), (i < nx - 1), (j > 0), (j < ny - 1), (k > 0), (k < nz - 1)):
        c += t.astype(float)
    A[..., 0] = c * scale
    A[..., 1] = -(i < nx - 1).astype(float) * scale * np.ones((nz, ny, 1))
    A[..., 2] = -(j < ny - 1).astype(float) * scale * np.ones((nz, 1, nx))
    A[..., 3] = -(k < nz - 1).astype(float) * scale * np.ones((1, ny, nx))
    b += (j == 0).astype(float) * b_scale
    b -= (j == ny - 1).astype(float) * b_scale
    return np.ascontiguousarray(A), np.ascontiguousarray(b)


def random_stencil(nx, ny, nz, rng, spd=True):
    """Random 7-point symmetric stencil.  With spd=True the diagonal strictly
    dominates (center = sum|offdiag| + U(0.1, 1)), so A is SPD."""
    A = np.zeros((nz, ny, nx, 4))
    A[..., 1] = -rng.uniform(0.0, 1.0, (nz, ny, nx))
    A[..., 2] = -rng.uniform(0.0, 1.0, (nz, ny, nx))
    A[..., 3] = -rng.uniform(0.0, 1.0, (nz, ny, nx))
    A[:, :, nx - 1, 1] = 0.0
    A[:, ny - 1, :, 2] = 0.0
    A[nz - 1, :, :, 3] = 0.0
    if spd:
        s = -(A[..., 1] + A[..., 2] + A[..., 3])
        s[:, :, 1:] += -A[:, :, :-1, 1]
        s[:, 1:, :] += -A[:, :-1, :, 2]
        s[1:, :, :] += -A[:-1, :, :, 3]
        A[..., 0] = s + rng.uniform(0.1, 1.0, (nz, ny, nx))
    else:
        A[..., 0] = rng.uniform(0.0, 1.0, (nz, ny, nx))
    return np.ascontiguousarray(A)


def dense(A):
    """Dense n x n matrix of a stencil (brute force for tiny grids)."""
    nz, ny, nx = A.shape[:3]
    n = nx * ny * nz
    M = np.zeros((n, n))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                M[c, c] = A[k, j, i, 0]
                if i + 1 < nx:
                    M[c, c + 1] = M[c + 1, c] = A[k, j, i, 1]
                if j + 1 < ny:
                    M[c, c + nx] = M[c + nx, c] = A[k, j, i, 2]
                if k + 1 < nz:
                    M[c, c + nx * ny] = M[c + nx * ny, c] = A[k, j, i, 3]
    return M


def mg_poisson_hierarchy(coarsest, levels):
    """FdmMgpcgSolver3.Solve / FdmMgSolver3.Solve test system
    (src/tests/unit_tests/fdm_mgpcg_solver3_tests.cpp:13-60): resizeWithCoarsest,
    A scaled by invdx^2 and b by invdx with invdx = 0.5^l."""
    from oracle.jet_oracle import MgSystem

    sys = MgSystem()
    for l in range(levels):
        n = coarsest << (levels - 1 - l)
        invdx = 0.5 ** l
        A, b = poisson_helper(n, n, n, scale=invdx * invdx, b_scale=invdx)
        sys.A.append(A)
        sys.b.append(b)
        sys.x.append(np.zeros_like(b))
    return sys


def collider_scene(n, seed, spacing=None, origin=(0.1, -0.2, 0.05), radius=0.22,
                   centre=(0.55, 0.25, 0.45), moving=True, friction=0.3):
    """A grid of n = (nx, ny, nz) cells with a sphere collider: cell-centre
    SDF, random face velocities, and the collider's rigid motion as
    (linear, angular, translation, friction)."""
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    h = np.array(spacing if spacing is not None else (1.0 / nx,) * 3, np.float64)
    o = np.array(origin, np.float64)
    z, y, x = np.meshgrid(o[2] + (np.arange(nz) + 0.5) * h[2], o[1] + (np.arange(ny) + 0.5) * h[1],
                          o[0] + (np.arange(nx) + 0.5) * h[0], indexing="ij")
    c = np.array(centre, np.float64)
    sdf = np.ascontiguousarray(np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - radius)
    u = rng.uniform(-1, 1, (nz, ny, nx + 1))
    v = rng.uniform(-1, 1, (nz, ny + 1, nx))
    w = rng.uniform(-1, 1, (nz + 1, ny, nx))
    motion = (((0.3, -0.1, 0.2), (0.5, 1.0, -0.7)) if moving else ((0, 0, 0), (0, 0, 0)))
    return h, o, sdf, (u, v, w), (motion[0], motion[1], tuple(c), friction)


def poisson_csr_kat(n=(3, 3, 3)):
    """The compressed test system of Jet's FdmCgSolver3.SolveCompressed
    (src/tests/unit_tests/fdm_linear_system_solver_test_helper3.h:51-118):
    every cell a row, -1 to each existing neighbour, the neighbour count on
    the diagonal, b = +1 / -1 for missing -y / +y and -z / +z neighbours;
    entries sorted by column as MatrixCsr::addRow stores them."""
    nx, ny, nz = n
    val, ci, rp, b = [], [], [0], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                row = {c: 0.0}
                bb = 0.0
                for (di, dj, dk, miss) in ((-1, 0, 0, 0.0), (1, 0, 0, 0.0), (0, -1, 0, 1.0),
                                           (0, 1, 0, -1.0), (0, 0, -1, 1.0), (0, 0, 1, -1.0)):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        row[c] += 1.0
                        row[ii + nx * (jj + ny * kk)] = -1.0
                    else:
                        bb += miss
                for col in sorted(row):
                    ci.append(col)
                    val.append(row[col])
                rp.append(len(ci))
                b.append(bb)
    return (np.array(val), np.array(rp, np.uint64), np.array(ci, np.uint64), np.array(b))


def csr_from_stencil(A, mask):
    """CSR rows of the cells in mask (x-fastest numbering) from Jet-layout
    stencil rows (centre, right, up, front), couplings to cells outside the
    mask dropped: a reference for the compressed build."""
    nz, ny, nx = mask.shape
    idx = -np.ones(mask.shape, np.int64)
    idx[mask] = np.arange(int(mask.sum()))
    val, ci, rp = [], [], [0]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if not mask[k, j, i]:
                    continue
                row = {idx[k, j, i]: A[k, j, i, 0]}
                for (di, dj, dk, own, comp) in ((1, 0, 0, True, 1), (-1, 0, 0, False, 1),
                                                (0, 1, 0, True, 2), (0, -1, 0, False, 2),
                                                (0, 0, 1, True, 3), (0, 0, -1, False, 3)):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if not (0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz) or not mask[kk, jj, ii]:
                        continue
                    a = A[k, j, i, comp] if own else A[kk, jj, ii, comp]
                    if a != 0.0:
                        row[idx[kk, jj, ii]] = a
                for col in sorted(row):
                    ci.append(col)
                    val.append(row[col])
                rp.append(len(ci))
    return np.array(val), np.array(rp, np.uint64), np.array(ci, np.uint64)


def mg_poisson_hierarchy2(coarsest, levels):
    """FdmMgpcgSolver2.Solve test system (src/tests/unit_tests/
    fdm_mgpcg_solver2_tests.cpp:13-49): resizeWithCoarsest({c, c}, levels), A
    scaled by invdx^2 and b by invdx with invdx = 0.5^l, as one-plane levels
    ((1, ny, nx, 4) rows with front = 0)."""
    from oracle.jet_oracle import MgSystem

    sys = MgSystem()
    for l in range(levels):
        n = coarsest << (levels - 1 - l)
        invdx = 0.5 ** l
        A, b = poisson_helper(n, n, 1, scale=invdx * invdx, b_scale=invdx)
        sys.A.append(A)
        sys.b.append(b)
        sys.x.append(np.zeros_like(b))
    return sys
