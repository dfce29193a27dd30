"""Pins the oracle's restatement of Jet's velocity boundary conditions
(GridFractionalBoundaryConditionSolver3 / GridBlockedBoundaryConditionSolver3
::constrainVelocity and the parts they are made of).

The solvers themselves cannot be linked here (their collider update builds a
SurfaceToImplicit3, which needs TriangleMesh3 and the absent tinyobj
sources), so the restatement is pinned by Jet's own known answers
(src/tests/unit_tests/grid_*_boundary_condition_solver3_tests.cpp,
array_utils_tests.cpp ExtrapolateToRegion2), by closed forms, and bitwise
against Jet's compiled grids, samplers, gradient, extrapolation, friction
helper and rigid collider in oracle/_ref, driven in the order
constrainVelocity calls them.
"""
import numpy as np
import pytest

from oracle.jet_oracle import DIR_ALL, DIR_FRONT, DIR_LEFT, DIR_UP, RigidCollider
from systems import collider_scene

BIG = np.finfo(np.float64).max


def kat_domain(oracle, blocked, closed):
    # Jet's ClosedDomain / OpenDomain tests: 10^3 cells, spacing 1, origin -5,
    # no collider (SDF = DBL_MAX), velocity (1, 1, 1), default depth 5.
    n = 10
    sdf = np.full((n, n, n), BIG)
    u, v, w = np.ones((n, n, n + 1)), np.ones((n, n + 1, n)), np.ones((n + 1, n, n))
    return oracle.constrain_velocity(u, v, w, sdf, (1.0, 1.0, 1.0), (-5.0, -5.0, -5.0), None, 5,
                                     closed, blocked)


@pytest.mark.parametrize("blocked", [False, True])
def test_closed_domain_kat(oracle, blocked):
    u, v, w = kat_domain(oracle, blocked, DIR_ALL)
    for a, axis in ((u, 2), (v, 1), (w, 0)):
        want = np.ones_like(a)
        idx = [slice(None)] * 3
        for end in (0, -1):
            idx[axis] = end
            want[tuple(idx)] = 0.0
        assert np.array_equal(a, want)


@pytest.mark.parametrize("blocked", [False, True])
def test_open_domain_kat(oracle, blocked):
    # closed: left, up, front -> u(0), v(ny), w(nz) are zero, nothing else
    u, v, w = kat_domain(oracle, blocked, DIR_LEFT | DIR_UP | DIR_FRONT)
    wu, wv, ww = np.ones_like(u), np.ones_like(v), np.ones_like(w)
    wu[:, :, 0] = 0
    wv[:, -1, :] = 0
    ww[-1, :, :] = 0
    assert np.array_equal(u, wu) and np.array_equal(v, wv) and np.array_equal(w, ww)


def test_extrapolation_kat_in_3d(oracle):
    # Jet's ExtrapolateToRegion2 data (10 x 12, valid block i in [2, 6),
    # j in [3, 10), value i + 10 j, 6 passes), repeated along z: every slice
    # must give the 2-D answer, whose entries are i, j clamped to the block,
    # except where the block is more than 6 steps away (left at 0).
    nx, ny, nz = 10, 12, 4
    data = np.zeros((nz, ny, nx))
    valid = np.zeros((nz, ny, nx), np.int8)
    for j in range(3, 10):
        for i in range(2, 6):
            data[:, j, i] = i + j * 10.0
            valid[:, j, i] = 1
    out = oracle.extrapolate_to_region(data, valid, 6)
    for j in range(ny):
        for i in range(nx):
            ci, cj = min(max(i, 2), 5), min(max(j, 3), 9)
            reach = abs(i - ci) + abs(j - cj) <= 6
            assert np.all(out[:, j, i] == (ci + 10 * cj if reach else 0.0)), (i, j)
    assert out[0, 0, 9] == 0.0  # the one cell of Jet's answer left at zero


def test_sampler_and_gradient_closed_forms(oracle):
    nx, ny, nz = 7, 6, 5
    # dyadic spacing and origin: cell centres are exact in binary
    h, o = np.array([0.5, 0.25, 0.125]), np.array([-1.0, 0.5, 2.0])
    rng = np.random.default_rng(3)
    data = rng.uniform(-1, 1, (nz, ny, nx))
    # at a cell centre the sampler returns the stored value (indexing pinned)
    for (i, j, k) in ((0, 0, 0), (3, 2, 1), (6, 5, 4)):
        p = o + h * (np.array([i, j, k]) + 0.5)
        assert oracle.sample_cell(data, h, o, p) == data[k, j, i]
    # a linear field: interior gradient is its slope
    a = np.array([0.7, -1.3, 2.1])
    z, y, x = np.meshgrid(o[2] + (np.arange(nz) + 0.5) * h[2], o[1] + (np.arange(ny) + 0.5) * h[1],
                          o[0] + (np.arange(nx) + 0.5) * h[0], indexing="ij")
    lin = np.ascontiguousarray(a[0] * x + a[1] * y + a[2] * z)
    p = o + h * np.array([3.3, 2.6, 2.2])
    assert np.allclose(oracle.gradient_cell(lin, h, o, p), a, rtol=1e-12)
    assert np.isclose(oracle.sample_cell(lin, h, o, p), a @ p, rtol=1e-12)


def test_friction_and_rigid_motion_closed_forms(oracle):
    n = np.array([0.0, 0.6, 0.8])
    vel = np.array([1.0, -2.0, 0.5])
    t = oracle.project_and_apply_friction(vel, n, 0.0)
    assert abs(t @ n) < 1e-15 and np.allclose(t, vel - (vel @ n) * n, rtol=0, atol=1e-15)
    # inflow -(v.n) = 0.8 against a tangential speed |t|: friction mu scales
    # by max(1 - mu (-(v.n)) / |t|, 0); a large mu stops the slip
    mu = 0.25
    assert np.isclose(-(vel @ n), 0.8, rtol=1e-15)
    t2 = oracle.project_and_apply_friction(vel, n, mu)
    assert np.allclose(t2, t * (1 - mu * 0.8 / np.linalg.norm(t)), rtol=1e-14)
    assert np.array_equal(oracle.project_and_apply_friction(vel, n, 1e6), np.zeros(3))
    # outflow: no friction
    assert np.array_equal(oracle.project_and_apply_friction(-vel, n, 0.9), -t)
    c = RigidCollider.make((0.3, -0.1, 0.2), (0.5, 1.0, -0.7), (1.0, 2.0, 3.0), 0.4)
    assert np.array_equal(oracle.collider_velocity(c, (1.0, 2.0, 3.0)), np.array([0.3, -0.1, 0.2]))
    r = np.array([0.5, -1.0, 2.0])
    want = np.array([0.3, -0.1, 0.2]) + np.cross([0.5, 1.0, -0.7], r)
    assert np.allclose(oracle.collider_velocity(c, np.array([1.0, 2.0, 3.0]) + r), want, rtol=1e-15)


def test_no_flux_inside_collider(oracle):
    # Inside a resting, frictionless collider the relative velocity keeps no
    # component along the SDF gradient: (u, v, w) . n = 0 at covered faces
    # whose three components were all projected with the same normal (a
    # plane collider, whose SDF gradient is constant).
    nx = ny = nz = 12
    h, o = np.full(3, 1.0 / nx), np.zeros(3)
    nrm = np.array([0.36, 0.48, 0.8])
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) * h[2], (np.arange(ny) + 0.5) * h[1],
                          (np.arange(nx) + 0.5) * h[0], indexing="ij")
    sdf = np.ascontiguousarray(nrm[0] * x + nrm[1] * y + nrm[2] * z - 0.9)
    rng = np.random.default_rng(5)
    vel = [np.full((nz, ny, nx + 1), 1.0), np.full((nz, ny + 1, nx), -0.5),
           np.full((nz + 1, ny, nx), 0.25)]
    vel = [a + 1e-3 * rng.uniform(-1, 1, a.shape) for a in vel]
    u, v, w = oracle.constrain_velocity(*vel, sdf, h, o, None, 2, 0, False)
    # a face deep inside (SDF < 0 with its whole cell neighbourhood inside):
    # the sampled velocity there is projected, and a uniform field stays
    # uniform there, so the normal part of (u, v, w) vanishes
    uu, vv, ww = u[2, 2, 2], v[2, 2, 2], w[2, 2, 2]
    before = np.array([vel[0][2, 2, 2], vel[1][2, 2, 2], vel[2][2, 2, 2]])
    assert abs(np.array([uu, vv, ww]) @ nrm) < 1e-2 * abs(before @ nrm)


@pytest.mark.parametrize("n,blocked,depth,closed,seed", [
    ((16, 12, 10), False, 3, DIR_ALL, 1), ((16, 12, 10), True, 3, DIR_ALL, 2),
    ((20, 16, 14), False, 5, DIR_LEFT | DIR_UP | DIR_FRONT, 3), ((13, 17, 11), True, 0, 0, 4),
    ((24, 24, 24), True, 7, DIR_ALL, 5), ((9, 5, 3), False, 2, DIR_ALL, 6)])
def test_constrain_velocity_matches_jet_parts(oracle, ref, n, blocked, depth, closed, seed):
    h, o, sdf, vel, (lin, ang, tr, mu) = collider_scene(n, seed)
    c = RigidCollider.make(lin, ang, tr, mu)
    got = oracle.constrain_velocity(*vel, sdf, h, o, c, depth, closed, blocked)
    want = ref.constrain_velocity(*vel, sdf, h, o, c, depth, closed, blocked)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    assert sum(int((a != b).sum()) for a, b in zip(got, vel)) > 0


def test_parts_match_jet(oracle, ref):
    rng = np.random.default_rng(7)
    nx, ny, nz = 9, 7, 6
    h, o = np.array([0.5, 0.25, 0.4]), np.array([-1.0, 0.3, 2.0])
    data = rng.uniform(-1, 1, (nz, ny, nx))
    for _ in range(200):
        p = o + (rng.uniform(-0.2, 1.2, 3) * np.array([nx, ny, nz])) * h
        assert oracle.sample_cell(data, h, o, p) == ref.sample_cell(data, h, o, p)
        assert np.array_equal(oracle.gradient_cell(data, h, o, p), ref.gradient_cell(data, h, o, p))
        vv, nn = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        nn /= np.linalg.norm(nn)
        mu = rng.uniform(0, 2)
        assert np.array_equal(oracle.project_and_apply_friction(vv, nn, mu),
                              ref.project_and_apply_friction(vv, nn, mu))
    valid = (rng.random((nz, ny, nx)) < 0.2).astype(np.int8)
    for it in (0, 1, 3, 9):
        assert np.array_equal(oracle.extrapolate_to_region(data, valid, it),
                              ref.extrapolate_to_region(data, valid, it))
    c = RigidCollider.make((0.3, -0.1, 0.2), (0.5, 1.0, -0.7), (1.0, 2.0, 3.0), 0.4)
    p = rng.uniform(-3, 3, 3)
    assert np.array_equal(oracle.collider_velocity(c, p), ref.collider_velocity(c, p))


def test_planted_mistake_is_caught(oracle, ref):
    # The comparison is sharp enough to see a plausible slip: dropping the
    # friction changes the result of the case above.
    h, o, sdf, vel, (lin, ang, tr, mu) = collider_scene((16, 12, 10), 2)
    good = ref.constrain_velocity(*vel, sdf, h, o, RigidCollider.make(lin, ang, tr, mu), 3,
                                  DIR_ALL, True)
    slip = oracle.constrain_velocity(*vel, sdf, h, o, RigidCollider.make(lin, ang, tr, 0.0), 3,
                                     DIR_ALL, True)
    assert not all(np.array_equal(a, b) for a, b in zip(good, slip))
