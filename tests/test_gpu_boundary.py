"""GPU velocity boundary conditions (jg_constrain_velocity) against the
oracle's restatement of GridFractionalBoundaryConditionSolver3 /
GridBlockedBoundaryConditionSolver3::constrainVelocity.

Every step is FP64 in Jet's operand order (samplers, SDF gradient, friction,
extrapolation averages), so the result must be bit-identical."""
import numpy as np
import pytest

from jgtest import dev, host, jetgpu, torch_cuda
from systems import collider_scene

pytestmark = pytest.mark.gpu

BIG = np.finfo(np.float64).max


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


def run_gpu(solver, vel, sdf, h, o, motion, depth, closed, blocked):
    nz, ny, nx = sdf.shape
    solver.set_levels([(nx, ny, nz)])
    u, v, w = (dev(a) for a in vel)
    c = jetgpu.RigidCollider.make(*motion) if motion is not None else None
    solver.constrain_velocity(u, v, w, dev(sdf), tuple(h), tuple(o), c, depth, closed, blocked)
    return host(u), host(v), host(w)


def run_oracle(oracle, vel, sdf, h, o, motion, depth, closed, blocked):
    from oracle.jet_oracle import RigidCollider

    c = RigidCollider.make(*motion) if motion is not None else None
    return oracle.constrain_velocity(*vel, sdf, h, o, c, depth, closed, blocked)


@pytest.mark.parametrize("n,blocked,depth,closed,moving,seed", [
    ((16, 12, 10), False, 3, 63, True, 1), ((16, 12, 10), True, 3, 63, True, 2),
    ((20, 16, 14), False, 5, 1 | 8 | 32, True, 3), ((13, 17, 11), True, 0, 0, True, 4),
    ((33, 29, 31), True, 7, 63, False, 5), ((9, 5, 3), False, 2, 63, True, 6),
    ((64, 48, 40), True, 4, 2 | 4 | 16, True, 7), ((1, 1, 1), True, 1, 63, True, 8)])
def test_constrain_velocity_bitwise(solver, oracle, n, blocked, depth, closed, moving, seed):
    print(f"case constrain n={n} blocked={blocked} depth={depth} closed={closed}", flush=True)
    h, o, sdf, vel, motion = collider_scene(n, seed, moving=moving)
    got = run_gpu(solver, vel, sdf, h, o, motion, depth, closed, blocked)
    want = run_oracle(oracle, vel, sdf, h, o, motion, depth, closed, blocked)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("blocked", [False, True])
@pytest.mark.parametrize("closed", [63, 1 | 8 | 32])
def test_domain_kats_on_gpu(solver, blocked, closed):
    # Jet's ClosedDomain / OpenDomain: no collider, velocity 1, only the
    # closed walls' normal faces become 0.
    print(f"case constrain kat blocked={blocked} closed={closed}", flush=True)
    n = 10
    sdf = np.full((n, n, n), BIG)
    vel = (np.ones((n, n, n + 1)), np.ones((n, n + 1, n)), np.ones((n + 1, n, n)))
    u, v, w = run_gpu(solver, vel, sdf, (1.0, 1.0, 1.0), (-5.0, -5.0, -5.0), None, 5, closed,
                      blocked)
    want = [np.ones_like(a) for a in vel]
    walls = [(0, 2, 0, 1), (0, 2, -1, 2), (1, 1, 0, 4), (1, 1, -1, 8), (2, 0, 0, 16), (2, 0, -1, 32)]
    for comp, axis, end, bit in walls:
        if closed & bit:
            idx = [slice(None)] * 3
            idx[axis] = end
            want[comp][tuple(idx)] = 0.0
    for a, b in zip((u, v, w), want):
        assert np.array_equal(a, b)


def test_constrain_velocity_large_and_repeated(solver, oracle):
    # 128^3 cells, and a second call on the same solver (scratch reused,
    # validity masks rewritten) must give the same answer as a fresh oracle run.
    print("case constrain 128^3 twice", flush=True)
    h, o, sdf, vel, motion = collider_scene((128, 128, 128), 11)
    want = run_oracle(oracle, vel, sdf, h, o, motion, 3, 63, True)
    for _ in range(2):
        got = run_gpu(solver, vel, sdf, h, o, motion, 3, 63, True)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)


def test_constrain_velocity_rejects_bad_arguments(solver):
    import ctypes as C

    solver.set_levels([(8, 8, 8)])
    hh = (C.c_double * 3)(0.0, 1.0, 1.0)
    oo = (C.c_double * 3)(0.0, 0.0, 0.0)
    lib = jetgpu.lib()
    assert lib.jg_constrain_velocity(solver.h, 0, hh, oo, None, None, None, None, None, 1, 63) != 0
