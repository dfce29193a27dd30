"""GPU pressure-system assembly, projection and end-to-end pressure solves
against the oracle restatement of GridSinglePhasePressureSolver3.

Markers and matrix rows are integer/exactly-representable work and must be
bit-identical; the divergence right-hand side is computed with the same
IEEE operations in the same order and must be bit-identical too.  Pressure
and projected velocities follow the solver parity bar (1e-6 relative).
"""
import numpy as np
import pytest

from jgtest import dev, empty, host, jetgpu, rel_err, rel_err_mean_free, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = jetgpu.Solver(0)
    yield s
    s.close()


@pytest.mark.parametrize("kind,n", [(1, 32), (2, 48), (3, 64), (4, 40), (5, 32)])
def test_device_scene_matches_host(solver, kind, n):
    d = jetgpu.scene_desc(kind, n, seed=9)
    h = jetgpu.generate_scene_host(d, vertex=True)
    g = {k: empty(v.shape) for k, v in h.items()}
    solver.generate_scene(d, g["u"], g["v"], g["w"], g["bsdf"], g["fsdf"], g["vertex"])
    for k in h:
        assert np.array_equal(host(g[k]), h[k]), k


@pytest.mark.parametrize("kind,n,levels", [(1, 32, 3), (2, 64, 4), (3, 64, 5), (4, 48, 5)])
def test_sp_build_matches_oracle(solver, oracle, kind, n, levels):
    print(f"case sp_build kind={kind} n={n}", flush=True)
    d = jetgpu.scene_desc(kind, n, seed=4)
    s = jetgpu.generate_scene_host(d)
    markers, sys_ = oracle.sp_build_mg(s["u"], s["v"], s["w"], (1.0,) * 3, (0.0,) * 3,
                                       s["bsdf"], s["fsdf"], levels)
    solver.set_levels([(m.shape[2], m.shape[1], m.shape[0]) for m in markers])
    solver.sp_build(dev(s["u"]), dev(s["v"]), dev(s["w"]), dev(s["bsdf"]), dev(s["fsdf"]), 1.0)
    for l, m in enumerate(markers):
        gm = np.empty(m.shape, np.int8)
        A = np.empty(sys_.A[l].shape)
        copy_to_host(jetgpu.lib().jg_solver_markers(solver.h, l), gm)
        copy_to_host(solver.matrix_ptr(l), A)
        assert np.array_equal(gm, m), l
        assert np.array_equal(A, sys_.A[l]), l
    b = np.empty(sys_.b[0].shape)
    copy_to_host(solver.rhs_ptr(), b)
    assert np.array_equal(b, sys_.b[0])


def copy_to_host(ptr, out):
    """hipMemcpy of a raw device pointer owned by the solver into numpy."""
    import ctypes as C

    import torch

    torch.cuda.synchronize()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device -> host


def _kat_fields():
    u = np.zeros((3, 3, 4))
    v = np.zeros((3, 4, 3))
    w = np.zeros((4, 3, 3))
    v[:, 1:3, :] = 1.0
    return u, v, w


def _cc(fn, n=3):
    z, y, x = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="ij")
    return np.ascontiguousarray(fn(x, y, z), dtype=np.float64)


BIG = np.finfo(np.float64).max


@pytest.mark.parametrize("case", ["single", "boundary", "free", "free_boundary"])
def test_jet_pressure_known_answers(solver, case):
    # GridSinglePhasePressureSolver3.* (src/tests/unit_tests/grid_single_phase_pressure_solver3_tests.cpp)
    # with the GPU CG as the linear solver.
    u, v, w = _kat_fields()
    bs = _cc(lambda x, y, z: -x + 2.0) if "boundary" in case else _cc(lambda x, y, z: BIG + 0 * x)
    fs = _cc(lambda x, y, z: y - 2.0) if "free" in case else _cc(lambda x, y, z: -BIG + 0 * x)
    solver.set_levels([(3, 3, 3)])
    du, dv, dw = dev(u), dev(v), dev(w)
    solver.sp_build(du, dv, dw, dev(bs), dev(fs), 1.0)
    info = solver.cg(100, 1e-10)
    solver.sp_apply_gradient(du, dv, dw, du, dv, dw, 1.0)
    p = solver.download_x()
    uo, vo, wo = host(du), host(dv), host(dw)
    assert np.abs(uo).max() < 1e-6 and np.abs(wo).max() < 1e-6
    expect_v = np.zeros_like(vo)
    if "boundary" in case:
        expect_v[:, 1:3, 2] = 1.0
    np.testing.assert_allclose(vo, expect_v, atol=1e-6)
    cols = slice(0, 2) if "boundary" in case else slice(0, 3)
    if "free" in case:
        for j in range(2):
            np.testing.assert_allclose(p[:, j, cols], 2.0 - j, atol=1e-6)
    else:
        np.testing.assert_allclose(p[:, 1:, cols] - p[:, :-1, cols], -1.0, atol=1e-6)


def test_projection_gradient_matches_oracle(solver, oracle):
    # spacing 1/48: 1/h is not a power of two, so a fused multiply-add in the
    # gradient would show up as a last-bit difference
    n = 48
    h = 1.0 / n
    d = jetgpu.scene_desc(jetgpu.JG_SCENE_SPHERE_POOL, n, seed=6)
    s = jetgpu.generate_scene_host(d)
    markers, sys_ = oracle.sp_build_mg(s["u"], s["v"], s["w"], (h,) * 3, (0.0,) * 3,
                                       s["bsdf"], s["fsdf"], 1)
    solver.set_levels([(n, n, n)])
    du, dv, dw = dev(s["u"]), dev(s["v"]), dev(s["w"])
    solver.sp_build(du, dv, dw, dev(s["bsdf"]), dev(s["fsdf"]), h)
    A = np.empty(sys_.A[0].shape)
    b = np.empty(sys_.b[0].shape)
    copy_to_host(solver.matrix_ptr(0), A)
    copy_to_host(solver.rhs_ptr(), b)
    assert np.array_equal(A, sys_.A[0]) and np.array_equal(b, sys_.b[0])
    info = solver.cg(2000, 1e-9)
    p = solver.download_x()
    # Jet writes only the projected faces of the output grid; the rest keep
    # the output's contents, here a copy of the input (as the oracle does).
    uo, vo, wo = dev(s["u"]), dev(s["v"]), dev(s["w"])
    solver.sp_apply_gradient(du, dv, dw, uo, vo, wo, h)
    ou, ov, ow = oracle.sp_apply_gradient(markers[0], p, s["u"], s["v"], s["w"], (h,) * 3)
    for g, o_ in ((uo, ou), (vo, ov), (wo, ow)):
        assert np.array_equal(host(g), o_)


def test_closed_sphere_64_cg(solver, oracle):
    # The closed-sphere workload at full size: 64^3, sphere r = 16, all
    # Neumann, FdmCgSolver3(2000, 1e-8).  Compared modulo the mean.
    n = 64
    d = jetgpu.scene_desc(jetgpu.JG_SCENE_CLOSED_SPHERE, n, seed=1)
    s = jetgpu.generate_scene_host(d)
    markers, sys_ = oracle.sp_build_mg(s["u"], s["v"], s["w"], (1.0,) * 3, (0.0,) * 3,
                                       s["bsdf"], s["fsdf"], 1)
    xo, io, ro, oko = oracle.cg_solve(sys_.A[0], sys_.b[0], 2000, 1e-8)
    solver.set_levels([(n, n, n)])
    solver.sp_build(dev(s["u"]), dev(s["v"]), dev(s["w"]), dev(s["bsdf"]), dev(s["fsdf"]), 1.0)
    info = solver.cg(2000, 1e-8)
    x = solver.download_x()
    fluid = markers[0] == 0
    assert oko and info.converged and abs(int(info.iterations) - io) <= 3
    assert rel_err_mean_free(x, xo, fluid) <= 1e-6


def test_dam_break_128_mgpcg(solver, oracle):
    # The dam-break workload at full size: 128^3, 5 levels, 2/2 sweeps.
    from oracle.jet_oracle import MgParams

    n, levels = 128, 5
    d = jetgpu.scene_desc(jetgpu.JG_SCENE_DAM_BREAK, n, seed=2)
    s = jetgpu.generate_scene_host(d)
    markers, sys_ = oracle.sp_build_mg(s["u"], s["v"], s["w"], (1.0,) * 3, (0.0,) * 3,
                                       s["bsdf"], s["fsdf"], levels)
    ok, it, rel = oracle.mgpcg_clean_solve(sys_, MgParams.make(max_levels=levels, restrict_iter=2,
                                                                correct_iter=2, coarsest_iter=20,
                                                                final_iter=2), 100, 1e-12)
    assert ok
    solver.set_levels([(m.shape[2], m.shape[1], m.shape[0]) for m in markers])
    solver.sp_build(dev(s["u"]), dev(s["v"]), dev(s["w"]), dev(s["bsdf"]), dev(s["fsdf"]), 1.0)
    p = jetgpu.mg_params(max_levels=levels, restrict_iter=2, correct_iter=2, coarsest_iter=20,
                         final_iter=2)
    for prec in (jetgpu.JG_FP64, jetgpu.JG_MIXED):
        info = solver.mgpcg(200, p, jetgpu.gpu_options(precision=prec, stop_rule=jetgpu.JG_STOP_RELATIVE,
                                                       rel_tol=1e-10))
        x = solver.download_x()
        assert info.rel_residual <= 1e-8
        assert rel_err(x, sys_.x[0]) <= 1e-6, prec
