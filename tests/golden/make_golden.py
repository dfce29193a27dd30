"""Writes tests/golden/jet_linear.npz: input/output vectors of Jet's own
compiled linear-solver code (oracle/_ref/libjetref.so, built by
`make -C oracle ref` from the reference sources), for the GPU and oracle
parity tests on machines where that build is absent.

Run from the repository root in a tree that has oracle/_ref:
    python tests/golden/make_golden.py

Every array is FP64 in Jet's layout (x fastest; matrices as (nz, ny, nx, 4)
FdmMatrixRow3 {center, right, up, front}).  Contents:
  blas_*      FdmBlas3::mvm and ::residual on a random 9 x 8 x 7 system
              (src/jet/fdm_linear_system3.cpp:79-116);
  relax_*     one FdmGaussSeidelSolver3::relaxRedBlack sweep, SOR 1.3
              (src/jet/fdm_gauss_seidel_solver3.cpp:150-221), and one
              FdmJacobiSolver3::relax step (src/jet/fdm_jacobi_solver3.cpp:95-112),
              on a 12 x 10 x 6 SPD system;
  mg_*        FdmMgUtils3::restrict 14 x 10 x 12 -> 7 x 5 x 6 and ::correct back
              (src/jet/fdm_mg_linear_system3.cpp:43-179);
  cg_*        FdmCgSolver3(300, 1e-10)::solve on a 12 x 11 x 10 SPD system
              (src/jet/fdm_cg_solver3.cpp:20-47): x, iterations, lastResidual.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle.jet_oracle import Ref  # noqa: E402
from systems import random_stencil  # noqa: E402


def main():
    ref = Ref()
    rng = np.random.default_rng(20261202)
    out = {}

    A = random_stencil(9, 8, 7, rng, spd=False)
    x = rng.uniform(-1, 1, (7, 8, 9))
    b = rng.uniform(-1, 1, (7, 8, 9))
    out.update(blas_A=A, blas_x=x, blas_b=b, blas_mvm=ref.mvm(A, x), blas_residual=ref.residual(A, x, b))

    A = random_stencil(12, 10, 6, rng)
    b = rng.uniform(-1, 1, (6, 10, 12))
    x0 = rng.uniform(-1, 1, (6, 10, 12))
    out.update(relax_A=A, relax_b=b, relax_x0=x0,
               relax_rb=ref.gs_relax(A, b, 1.3, x0, red_black=True),
               relax_jacobi=ref.jacobi_relax(A, b, x0))

    # n.y <= n.z: Jet's prolongation clamps k with n.y, in bounds only then
    fine = rng.uniform(-1, 1, (12, 10, 14))
    coarse = rng.uniform(-1, 1, (6, 5, 7))
    out.update(mg_fine=fine, mg_restrict=ref.restrict(fine), mg_coarse=coarse,
               mg_correct=ref.correct(coarse, fine))

    A = random_stencil(12, 11, 10, rng)
    b = rng.uniform(-1, 1, (10, 11, 12))
    xs, it, res, ok = ref.cg_solve(A, b, 300, 1e-10)
    assert ok
    out.update(cg_A=A, cg_b=b, cg_x=xs, cg_iterations=np.array(it), cg_last_residual=np.array(res))

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jet_linear.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {os.path.getsize(path)} bytes; cg iterations {it}, lastResidual {res:.3e}")


if __name__ == "__main__":
    main()
