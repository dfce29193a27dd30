"""Pins the oracle's coarse operator of the flat-entry MGPCG (or_mg_coarsen).
Jet has no counterpart (FdmMgSolver3::solve(FdmLinearSystem3*) is final and
empty, include/jet/fdm_mg_solver3.h:42), so the pins are mathematical:

- Closed form: GridBackwardEulerDiffusionSolver3's all-fluid Neumann system
  with c = dt mu / h^2 coarsens to the same build at twice the spacing (c / 4),
  bit for bit (c a power of two), in 3-D and on one plane.
- Dense identity: A_c = (1/8) P'MP + (1/16) P'LP (one plane: 1/4, 1/8) with P
  the piecewise-constant injection, M = diag(row sums), L = A - M, evaluated
  with dense matrices on random stencils and on diffusion systems with
  colliders, air and Dirichlet faces.
- Zero row sums (an all-Neumann pressure system) stay zero, and SPD stays SPD.
- The hierarchy it builds makes MGPCG reach a dense direct solve.
- A planted mistake (the coupling part halved twice) fails the closed form.
"""
import numpy as np
import pytest

from systems import dense, poisson_helper, random_stencil

BIG = np.finfo(np.float64).max


def injection(shape):
    """Dense P (fine x coarse): each fine cell to the aggregate holding it."""
    nz, ny, nx = shape
    plane = nz == 1
    cz, cy, cx = (1 if plane else nz // 2), ny // 2, nx // 2
    P = np.zeros((nx * ny * nz, cx * cy * cz))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                K = k if plane else k // 2
                P[i + nx * (j + ny * k), i // 2 + cx * (j // 2 + cy * K)] = 1.0
    return P


def galerkin(A):
    D = dense(A)
    M = np.diag(D.sum(axis=1))
    P = injection(A.shape[:3])
    mass, coup = (0.25, 0.125) if A.shape[0] == 1 else (0.125, 0.0625)
    return mass * (P.T @ M @ P) + coup * (P.T @ (D - M) @ P)


def uniform_diffusion(oracle, shape, c):
    nz, ny, nx = shape
    f = np.zeros(shape)
    A, _ = oracle.diffusion_build(f, np.full(shape, BIG), np.full(shape, -BIG), 1.0, c,
                                  (1.0, 1.0, 1.0))
    return A


@pytest.mark.parametrize("shape", [(8, 6, 10), (4, 4, 4), (2, 2, 2), (1, 8, 12), (1, 2, 2),
                                   (6, 2, 4)])
def test_uniform_diffusion_coarsens_to_the_rediscretised_system(oracle, shape):
    fine = uniform_diffusion(oracle, shape, 2.0)
    nz, ny, nx = shape
    cshape = (1 if nz == 1 else nz // 2, ny // 2, nx // 2)
    want = uniform_diffusion(oracle, cshape, 0.5)
    assert np.array_equal(oracle.mg_coarsen(fine), want)


def test_planted_wrong_scale_is_caught(oracle):
    # a planted mistake in a copy of the formula: the coupling part with
    # 1/32 instead of 1/16 (the piecewise-constant Galerkin product halved)
    fine = uniform_diffusion(oracle, (4, 4, 4), 2.0)
    good = oracle.mg_coarsen(fine)
    D = dense(fine)
    M = np.diag(D.sum(axis=1))
    P = injection((4, 4, 4))
    planted = 0.125 * (P.T @ M @ P) + (1.0 / 32.0) * (P.T @ (D - M) @ P)
    want = dense(uniform_diffusion(oracle, (2, 2, 2), 0.5))
    assert np.allclose(dense(good), want, rtol=0, atol=0)
    assert not np.allclose(planted, want, rtol=1e-3)


@pytest.mark.parametrize("shape,seed", [((4, 6, 8), 1), ((2, 4, 2), 2), ((6, 4, 4), 3),
                                        ((1, 6, 8), 4), ((1, 4, 10), 5)])
def test_dense_identity_random_stencils(oracle, shape, seed):
    nz, ny, nx = shape
    A = random_stencil(nx, ny, nz, np.random.default_rng(seed))
    got = dense(oracle.mg_coarsen(A))
    np.testing.assert_allclose(got, galerkin(A), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("shape,dirichlet,seed", [((6, 8, 4), False, 6), ((6, 8, 4), True, 7),
                                                  ((1, 8, 6), True, 8), ((4, 4, 4), True, 9)])
def test_dense_identity_diffusion_with_colliders(oracle, shape, dirichlet, seed):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, shape)
    bs = np.where(rng.random(shape) < 0.2, -1.0, 1.0)
    fs = np.where(rng.random(shape) < 0.7, -1.0, 1.0)
    A, _ = oracle.diffusion_build(f, bs, fs, 0.3, 0.7, (0.5, 0.25, 0.5), dirichlet)
    got = dense(oracle.mg_coarsen(A))
    np.testing.assert_allclose(got, galerkin(A), rtol=1e-13, atol=1e-13)
    assert np.all(np.linalg.eigvalsh(got) > 0)


def test_zero_row_sums_stay_zero(oracle):
    A, _ = poisson_helper(8, 8, 8)
    C = oracle.mg_coarsen(A)
    Dc = dense(C)
    assert np.max(np.abs(Dc.sum(axis=1))) <= 1e-15 * np.max(np.abs(Dc))
    # and the coarse pressure stencil is the fine one / 4 (the h^-2 scaling)
    np.testing.assert_array_equal(C, poisson_helper(4, 4, 4, scale=0.25)[0])


def test_flat_hierarchy_mgpcg_reaches_direct_solve(oracle):
    from oracle.jet_oracle import MgParams

    rng = np.random.default_rng(11)
    shape = (8, 16, 16)
    f = rng.uniform(-1, 1, shape)
    bs = np.where(rng.random(shape) < 0.1, -1.0, 1.0)
    fs = np.full(shape, -1.0)
    A, b = oracle.diffusion_build(f, bs, fs, 1.0 / 60, 0.05, (1.0 / 16,) * 3)
    s = oracle.mg_flat_hierarchy(A, b, 4)
    assert [x.shape for x in s.x] == [(8, 16, 16), (4, 8, 8), (2, 4, 4), (1, 2, 2)]
    ok, it, rel = oracle.mgpcg_clean_solve(s, MgParams.make(max_levels=4, restrict_iter=2,
                                                            correct_iter=2, coarsest_iter=20,
                                                            final_iter=2), 100, 1e-12)
    want = np.linalg.solve(dense(A), b.ravel()).reshape(shape)
    assert ok and rel <= 1e-12 and it <= 12
    np.testing.assert_allclose(s.x[0], want, rtol=1e-9, atol=1e-11)
    # plain CG on the same system takes several times as many iterations
    _, cg_it, _, _ = oracle.cg_solve(A, b, 500, 1e-12 * np.linalg.norm(b))
    print(f"flat MGPCG {it} iterations, CG {cg_it}")
    assert cg_it > 2 * it
