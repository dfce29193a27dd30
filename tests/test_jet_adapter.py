"""The C++ solver classes (fluid-engine-dev_amd/jet/fdm_gpu_solvers3.h) behind
Jet's own interfaces, next to Jet's compiled CPU solvers.

oracle/_ref/jet_adapter_check (tests/cpp/jet_adapter_check.cpp, built by
`make -C oracle adapter` where the reference tree is present) links Jet's
FdmCgSolver3 / FdmMgpcgSolver3 objects and the GPU classes and checks:
Jet's error behaviour (std::invalid_argument on size mismatch, false from
solveCompressed and from the multigrid solver's flat solve, params() and the
accessors), selection through dynamic_pointer_cast<FdmMgSolver3>, GPU CG
against Jet's CG (iterations within 2, x within 1e-8), Jet's MGPCG unit test
through the GPU class in FP64 and mixed precision, the closed form
x = (ny - 1)/2 - j of Jet's test systems, GPU MGPCG against Jet's CG,
Jet's FdmMgSolver3.Solve test through the one-V-cycle class FdmGpuMgSolver3
(warm start included), and Jet's MGPCG log line.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "oracle", "_ref", "jet_adapter_check")


def run_check(*args):
    if not os.path.exists(CHECK):
        pytest.skip("oracle/_ref/jet_adapter_check not built (needs the reference tree)")
    p = subprocess.run([CHECK, *args], capture_output=True, text=True, timeout=240)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED" in p.stdout
    return p.stdout


def test_adapter_argument_checks():
    out = run_check("--no-gpu")
    assert out.count("ok  ") == 19


@pytest.mark.gpu
def test_adapter_solves_on_gpu():
    from jgtest import torch_cuda

    if torch_cuda() is None:
        pytest.skip("no GPU")
    out = run_check()
    assert "FAIL" not in out
