"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/jetgpu.h declares, and the host scene generator meets its spec."""
import os
import re

import numpy as np
import pytest

from jgtest import ROOT, jetgpu


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "jetgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(jg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = jetgpu.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
        assert name in jetgpu.SIGNATURES, f"{name} has no ctypes signature"
    assert set(jetgpu.SIGNATURES) == set(names)


def test_defaults_match_jet():
    # MgParameters / FdmMgSolver3 defaults (include/jet/mg.h:58-86,
    # include/jet/fdm_mg_solver3.h:24-30).
    p = jetgpu.mg_params()
    assert (p.max_levels, p.restrict_iter, p.correct_iter, p.coarsest_iter, p.final_iter) == (1, 5, 5, 20, 20)
    assert p.max_tol == 1e-9 and p.sor == 1.5 and p.red_black == 0


def test_solver_create_without_gpu_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(jetgpu.JgError):
        jetgpu.Solver(0)


def test_batch_create_without_gpu_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(jetgpu.JgError):
        jetgpu.Batch([(32, 32, 32), (48, 48, 48)], [3, 4])


def test_batch_rejects_bad_arguments():
    import ctypes as C

    lib = jetgpu.lib()
    h = C.c_void_p()
    one = (C.c_size_t * 1)(8)
    assert lib.jg_batch_create(0, 0, one, one, one, one, C.byref(h)) == 1  # no systems
    assert lib.jg_batch_create(0, 1, one, one, one, one, None) == 1        # no output
    assert lib.jg_batch_mgpcg(None, 10, None, None, None, None) == 1
    assert lib.jg_batch_system(None, 0) is None and lib.jg_batch_size(None) == 0


def test_null_solver_rejected_by_new_entry_points():
    # no device work happens before the argument checks
    lib = jetgpu.lib()
    assert lib.jg_csr_cg_solve_host(None, 1, 1, None, None, None, None, None, 10, 1e-6, None) == 1
    assert lib.jg_csr_mvm_host(None, 1, 1, None, None, None, None, None) == 1
    assert lib.jg_csr_residual_host(None, 1, 1, None, None, None, None, None, None) == 1
    assert lib.jg_constrain_velocity(None, 0, None, None, None, None, None, None, None, 1, 63) == 1


@pytest.mark.parametrize("kind,n", [(1, 16), (2, 16), (3, 32), (4, 24), (5, 32)])
def test_scene_generator_deterministic_and_bc(kind, n):
    d = jetgpu.scene_desc(kind, n, seed=7)
    a = jetgpu.generate_scene_host(d, vertex=True)
    b = jetgpu.generate_scene_host(d, vertex=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    u, v, w, bs = a["u"], a["v"], a["w"], a["bsdf"]
    # Domain walls are closed.
    assert np.all(u[:, :, 0] == 0) and np.all(u[:, :, -1] == 0)
    assert np.all(v[:, 0, :] == 0) and np.all(v[:, -1, :] == 0)
    assert np.all(w[0] == 0) and np.all(w[-1] == 0)
    # Faces touching a solid cell are zero.
    solid = bs < 0
    assert np.all(u[:, :, 1:-1][solid[:, :, 1:] | solid[:, :, :-1]] == 0)
    assert np.all(v[:, 1:-1, :][solid[:, 1:, :] | solid[:, :-1, :]] == 0)
    assert np.all(w[1:-1][solid[1:] | solid[:-1]] == 0)
    # Interior open faces are U(-1,1) draws (+ gravity for the dam break).
    lo, hi = (-1 - 9.8 / 60, 1 - 9.8 / 60) if kind == 2 else (-1, 1)
    assert v.min() >= lo and v.max() <= hi and np.abs(u).max() <= 1.0
    # Different seeds differ.
    c = jetgpu.generate_scene_host(jetgpu.scene_desc(kind, n, seed=8))
    assert not np.array_equal(a["u"], c["u"])


def test_scene_sphere_pool_fluid_fraction():
    n = 64
    s = jetgpu.generate_scene_host(jetgpu.scene_desc(jetgpu.JG_SCENE_SPHERE_POOL, n, seed=3))
    fluid = (s["bsdf"] >= 0) & (s["fsdf"] < 0)
    frac = fluid.mean()
    assert 0.59 <= frac <= 0.61 + 1.0 / n
    assert (s["bsdf"] < 0).sum() > 0


def test_scene_closed_sphere_is_consistent():
    # All-Neumann: the discrete divergence sums to zero over the fluid
    # (solid and wall faces are zero), so b is in the range of A.
    n = 32
    s = jetgpu.generate_scene_host(jetgpu.scene_desc(jetgpu.JG_SCENE_CLOSED_SPHERE, n, seed=1))
    u, v, w = s["u"], s["v"], s["w"]
    div = (u[:, :, 1:] - u[:, :, :-1]) + (v[:, 1:, :] - v[:, :-1, :]) + (w[1:] - w[:-1])
    fluid = s["bsdf"] >= 0
    assert abs(div[fluid].sum()) < 1e-9
    assert fluid.mean() > 0.9 and (~fluid).sum() > 0
