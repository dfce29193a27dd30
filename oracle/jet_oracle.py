"""ctypes front end of the CPU oracle (test infrastructure only).

Arrays follow Jet's Array3 layout (x fastest): a cell vector is a C-ordered
numpy array of shape (nz, ny, nx); a matrix is (nz, ny, nx, 4) holding
FdmMatrixRow3 {center, right, up, front}; face arrays are u:(nz, ny, nx+1),
v:(nz, ny+1, nx), w:(nz+1, ny, nx).

``Oracle`` wraps ``liboracle.so`` (the C restatement).  ``Ref`` wraps
``_ref/libjetref.so`` (Jet's own linear solvers, compiled from the reference
sources) when it has been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libjetref.so")

FLUID, AIR, BOUNDARY = 0, 1, 2

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_cp = C.c_char_p
_sz = C.c_size_t


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def _c(a):
    assert a.dtype == np.int8 and a.flags.c_contiguous
    return C.cast(a.ctypes.data, _cp)


def _shape3(v):
    nz, ny, nx = v.shape[:3]
    return nx, ny, nz


class MgParams(C.Structure):
    """or_mg_params; defaults are MgParameters' (include/jet/mg.h:58-86) with
    FdmMgSolver3's sorFactor 1.5 (include/jet/fdm_mg_solver3.h:24-30)."""

    _fields_ = [
        ("max_levels", C.c_size_t),
        ("restrict_iter", C.c_uint),
        ("correct_iter", C.c_uint),
        ("coarsest_iter", C.c_uint),
        ("final_iter", C.c_uint),
        ("max_tol", C.c_double),
        ("sor", C.c_double),
        ("red_black", C.c_int),
    ]

    @classmethod
    def make(cls, max_levels=1, restrict_iter=5, correct_iter=5,
             coarsest_iter=20, final_iter=20, max_tol=1e-9, sor=1.5,
             red_black=False):
        return cls(max_levels, restrict_iter, correct_iter, coarsest_iter,
                   final_iter, max_tol, sor, int(bool(red_black)))


MAX_LEVELS = 16


class MgSystemC(C.Structure):
    _fields_ = [
        ("nlev", C.c_size_t),
        ("nx", C.c_size_t * MAX_LEVELS),
        ("ny", C.c_size_t * MAX_LEVELS),
        ("nz", C.c_size_t * MAX_LEVELS),
        ("A", _dp * MAX_LEVELS),
        ("x", _dp * MAX_LEVELS),
        ("b", _dp * MAX_LEVELS),
    ]


@dataclass
class MgSystem:
    """FdmMgLinearSystem3: per-level A, x, b (level 0 finest)."""

    A: list = field(default_factory=list)
    x: list = field(default_factory=list)
    b: list = field(default_factory=list)

    def to_c(self):
        s = MgSystemC()
        s.nlev = len(self.A)
        for l, (A, x, b) in enumerate(zip(self.A, self.x, self.b)):
            nx, ny, nz = _shape3(x)
            s.nx[l], s.ny[l], s.nz[l] = nx, ny, nz
            s.A[l], s.x[l], s.b[l] = _d(A), _d(x), _d(b)
        return s


def _setup(lib):
    lib.or_dot.restype = C.c_double
    lib.or_l2norm.restype = C.c_double
    lib.or_linfnorm.restype = C.c_double
    lib.or_fraction_inside.restype = C.c_double
    lib.or_fraction_inside.argtypes = [C.c_double] * 4
    lib.or_fraction_inside_sdf.restype = C.c_double
    lib.or_fraction_inside_sdf.argtypes = [C.c_double] * 2
    lib.or_mg_solve.argtypes = [C.POINTER(MgSystemC), C.POINTER(MgParams), _dp]
    lib.or_mgpcg_solve.argtypes = [C.POINTER(MgSystemC), C.POINTER(MgParams),
                                   C.c_uint, C.POINTER(C.c_uint), _dp]
    lib.or_mgpcg_clean_solve.argtypes = [C.POINTER(MgSystemC), C.POINTER(MgParams), C.c_uint,
                                         C.c_double, C.POINTER(C.c_uint), _dp]
    lib.or_mg_clean_vcycle.argtypes = [C.POINTER(MgSystemC), C.POINTER(MgParams), _dp, _dp]
    lib.or_mg_clean_vcycle_jacobi.argtypes = [C.POINTER(MgSystemC), C.POINTER(MgParams), C.c_double,
                                              _dp, _dp]
    lib.or_levels_with_finest.restype = C.c_size_t
    lib.or_levels_with_finest2.restype = C.c_size_t
    lib.or_divergence.restype = C.c_double
    return lib


class RigidCollider(C.Structure):
    """RigidBodyCollider3 state: velocities, surface translation, friction."""
    _fields_ = [("linear_velocity", C.c_double * 3), ("angular_velocity", C.c_double * 3),
                ("translation", C.c_double * 3), ("friction", C.c_double)]

    @classmethod
    def make(cls, linear=(0, 0, 0), angular=(0, 0, 0), translation=(0, 0, 0), friction=0.0):
        return cls((C.c_double * 3)(*linear), (C.c_double * 3)(*angular),
                   (C.c_double * 3)(*translation), friction)


# Jet's kDirection* bits (include/jet/constants.h:312-335)
DIR_LEFT, DIR_RIGHT, DIR_DOWN, DIR_UP, DIR_BACK, DIR_FRONT = 1, 2, 4, 8, 16, 32
DIR_ALL = 63


class _BoundaryCalls:
    """Velocity boundary conditions; the same calls on the restatement (or_*)
    and on Jet's compiled parts (ref_*)."""

    _p = "or_"

    def _fn(self, name):
        return getattr(self.lib, self._p + name)

    def sample_cell(self, data, h, origin, p):
        nx, ny, nz = _shape3(data)
        f = self._fn("sample_cell")
        f.restype = C.c_double
        return f(_d(data), _sz(nx), _sz(ny), _sz(nz), _d(np.asarray(h, np.float64)),
                 _d(np.asarray(origin, np.float64)), _d(np.asarray(p, np.float64)))

    def gradient_cell(self, data, h, origin, p):
        nx, ny, nz = _shape3(data)
        out = np.empty(3)
        self._fn("gradient_cell")(_d(data), _sz(nx), _sz(ny), _sz(nz),
                                  _d(np.asarray(h, np.float64)), _d(np.asarray(origin, np.float64)),
                                  _d(np.asarray(p, np.float64)), _d(out))
        return out

    def extrapolate_to_region(self, data, valid, iterations):
        nx, ny, nz = _shape3(data)
        out = np.empty_like(data)
        self._fn("extrapolate_to_region")(_d(data), _c(valid), _sz(nx), _sz(ny), _sz(nz),
                                          C.c_uint(iterations), _d(out))
        return out

    def project_and_apply_friction(self, vel, normal, friction):
        out = np.empty(3)
        self._fn("project_and_apply_friction")(_d(np.asarray(vel, np.float64)),
                                               _d(np.asarray(normal, np.float64)),
                                               C.c_double(friction), _d(out))
        return out

    def collider_velocity(self, collider, p):
        out = np.empty(3)
        self._fn("collider_velocity")(C.byref(collider), _d(np.asarray(p, np.float64)), _d(out))
        return out

    def constrain_velocity(self, u, v, w, sdf, h, origin, collider=None, depth=5,
                           closed=DIR_ALL, blocked=False):
        """Returns the constrained (u, v, w); the inputs are not modified."""
        nx, ny, nz = _shape3(sdf)
        uo, vo, wo = (np.ascontiguousarray(a, np.float64).copy() for a in (u, v, w))
        self._fn("constrain_velocity")(C.c_int(1 if blocked else 0), _sz(nx), _sz(ny), _sz(nz),
                                       _d(np.asarray(h, np.float64)),
                                       _d(np.asarray(origin, np.float64)), _d(uo), _d(vo), _d(wo),
                                       _d(sdf), C.byref(collider) if collider is not None else None,
                                       C.c_uint(depth), C.c_int(closed))
        return uo, vo, wo


class _CsrCalls:
    """Compressed (CSR) systems; the same calls on the restatement (or_*) and
    on Jet's compiled FdmCompressedBlas3 / FdmCgSolver3 (ref_*)."""

    _p = "or_"

    @staticmethod
    def _csr(val, rp, ci):
        val = np.ascontiguousarray(val, np.float64)
        rp = np.ascontiguousarray(rp, np.uint64)
        ci = np.ascontiguousarray(ci, np.uint64)
        return val, rp, ci, rp.ctypes.data_as(C.POINTER(C.c_size_t)), ci.ctypes.data_as(
            C.POINTER(C.c_size_t))

    def csr_mvm(self, val, rp, ci, v):
        val, rp, ci, prp, pci = self._csr(val, rp, ci)
        v = np.ascontiguousarray(v, np.float64)
        out = np.empty_like(v)
        getattr(self.lib, self._p + "csr_mvm")(_d(val), prp, pci, _sz(v.size), _d(v), _d(out))
        return out

    def csr_residual(self, val, rp, ci, x, b):
        val, rp, ci, prp, pci = self._csr(val, rp, ci)
        x = np.ascontiguousarray(x, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        out = np.empty_like(x)
        getattr(self.lib, self._p + "csr_residual")(_d(val), prp, pci, _sz(x.size), _d(x), _d(b),
                                                    _d(out))
        return out

    def csr_cg_solve(self, val, rp, ci, b, max_iter, tol):
        val, rp, ci, prp, pci = self._csr(val, rp, ci)
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        it, res = C.c_uint(), C.c_double()
        ok = getattr(self.lib, self._p + "csr_cg_solve")(_d(val), prp, pci, _sz(b.size), _d(b),
                                                         _d(x), C.c_uint(max_iter), C.c_double(tol),
                                                         C.byref(it), C.byref(res))
        return x, bool(ok), it.value, res.value


def diffusion_c(dt, mu, h):
    """c = dt * mu / h^2 per axis, in the order
    GridBackwardEulerDiffusionSolver3::solve evaluates it
    (src/jet/grid_backward_euler_diffusion_solver3.cpp:33)."""
    s = dt * mu
    return np.array([s / (hh * hh) for hh in h], np.float64)


class Oracle(_BoundaryCalls, _CsrCalls):
    """The C restatement (liboracle.so)."""

    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = _setup(C.CDLL(path))

    # ----------------------------------------------------------- FdmBlas3
    def mvm(self, A, v):
        out = np.empty_like(v)
        self.lib.or_mvm(_d(A), _d(v), _d(out), *map(_sz, _shape3(v)))
        return out

    def residual(self, A, x, b):
        out = np.empty_like(x)
        self.lib.or_residual(_d(A), _d(x), _d(b), _d(out), *map(_sz, _shape3(x)))
        return out

    def dot(self, a, b):
        return self.lib.or_dot(_d(a), _d(b), *map(_sz, _shape3(a)))

    def l2norm(self, a):
        return self.lib.or_l2norm(_d(a), *map(_sz, _shape3(a)))

    def linfnorm(self, a):
        return self.lib.or_linfnorm(_d(a), *map(_sz, _shape3(a)))

    def axpy(self, a, x, y):
        out = np.empty_like(x)
        self.lib.or_axpy(C.c_double(a), _d(x), _d(y), _d(out), *map(_sz, _shape3(x)))
        return out

    # ---------------------------------------------------------------- CG
    def cg_solve(self, A, b, max_iter, tol):
        x = np.zeros_like(b)
        it, res = C.c_uint(), C.c_double()
        ok = self.lib.or_cg_solve(_d(A), _d(b), _d(x), *map(_sz, _shape3(b)),
                                  C.c_uint(max_iter), C.c_double(tol),
                                  C.byref(it), C.byref(res))
        return x, it.value, res.value, bool(ok)

    # ---------------------------------------------------------- smoothers
    def gs_relax(self, A, b, sor, x, red_black=False):
        x = x.copy()
        fn = self.lib.or_gs_relax_rb if red_black else self.lib.or_gs_relax
        fn(_d(A), _d(b), C.c_double(sor), _d(x), *map(_sz, _shape3(x)))
        return x

    def jacobi_relax(self, A, b, x):
        out = np.empty_like(x)
        self.lib.or_jacobi_relax(_d(A), _d(b), _d(x), _d(out), *map(_sz, _shape3(x)))
        return out

    # --------------------------------------------------------- multigrid
    def levels_with_finest(self, shape_xyz, max_levels):
        lx, ly, lz = (C.c_size_t * MAX_LEVELS)(), (C.c_size_t * MAX_LEVELS)(), (C.c_size_t * MAX_LEVELS)()
        n = self.lib.or_levels_with_finest(*map(_sz, shape_xyz), _sz(max_levels), lx, ly, lz)
        return [(lx[l], ly[l], lz[l]) for l in range(n)]

    def restrict(self, fine):
        nz, ny, nx = fine.shape
        out = np.empty((nz // 2, ny // 2, nx // 2))
        self.lib.or_mg_restrict(_d(fine), _d(out), _sz(nx // 2), _sz(ny // 2), _sz(nz // 2))
        return out

    def correct(self, coarse, fine):
        fine = fine.copy()
        self.lib.or_mg_correct(_d(coarse), _d(fine), *map(_sz, _shape3(fine)))
        return fine

    def levels_with_finest2(self, shape_xy, max_levels):
        """FdmMgUtils2::resizeArrayWithFinest: [(nx, ny), ...] finest first."""
        lx, ly = (C.c_size_t * MAX_LEVELS)(), (C.c_size_t * MAX_LEVELS)()
        n = self.lib.or_levels_with_finest2(*map(_sz, shape_xy), _sz(max_levels), lx, ly)
        return [(lx[l], ly[l]) for l in range(n)]

    def restrict2(self, fine):
        """FdmMgUtils2::restrict of an (ny, nx) array."""
        ny, nx = fine.shape
        fine = np.ascontiguousarray(fine, dtype=np.float64)
        out = np.empty((ny // 2, nx // 2))
        self.lib.or_mg_restrict2(_d(fine), _d(out), _sz(nx // 2), _sz(ny // 2))
        return out

    def correct2(self, coarse, fine):
        """FdmMgUtils2::correct: fine + prolongation of coarse, (ny, nx) arrays."""
        fine = np.array(fine, dtype=np.float64)
        coarse = np.ascontiguousarray(coarse, dtype=np.float64)
        ny, nx = fine.shape
        self.lib.or_mg_correct2(_d(coarse), _d(fine), _sz(nx), _sz(ny))
        return fine

    def mg_solve(self, sys: MgSystem, params: MgParams):
        """FdmMgSolver3::solve -- updates sys.x in place; returns (ok, norm)."""
        s = sys.to_c()
        norm = C.c_double()
        ok = self.lib.or_mg_solve(C.byref(s), C.byref(params), C.byref(norm))
        return bool(ok), norm.value

    def mgpcg_solve(self, sys: MgSystem, params: MgParams, max_iter):
        """FdmMgpcgSolver3::solve -- sys.x[0] is the solution."""
        s = sys.to_c()
        it, res = C.c_uint(), C.c_double()
        ok = self.lib.or_mgpcg_solve(C.byref(s), C.byref(params), C.c_uint(max_iter),
                                     C.byref(it), C.byref(res))
        return bool(ok), it.value, res.value

    def mgpcg_clean_solve(self, sys: MgSystem, params: MgParams, max_iter, rel_tol):
        """Clean-mode FP64 reference solve; returns (ok, iters, true_rel_residual)."""
        s = sys.to_c()
        it, res = C.c_uint(), C.c_double()
        ok = self.lib.or_mgpcg_clean_solve(C.byref(s), C.byref(params), C.c_uint(max_iter),
                                           C.c_double(rel_tol), C.byref(it), C.byref(res))
        return bool(ok), it.value, res.value

    def mgpcg_clean_solve_jacobi(self, sys: MgSystem, params: MgParams, omega, max_iter, rel_tol):
        """The clean solve with the weighted-Jacobi V-cycle; (ok, iters, true_rel_residual)."""
        s = sys.to_c()
        it, res = C.c_uint(), C.c_double()
        ok = self.lib.or_mgpcg_clean_solve_jacobi(C.byref(s), C.byref(params), C.c_double(omega),
                                                  C.c_uint(max_iter), C.c_double(rel_tol),
                                                  C.byref(it), C.byref(res))
        return bool(ok), it.value, res.value

    def mg_clean_vcycle_jacobi(self, sys: MgSystem, params: MgParams, omega, r):
        """z = M r with the weighted-Jacobi clean V-cycle."""
        s = sys.to_c()
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        self.lib.or_mg_clean_vcycle_jacobi(C.byref(s), C.byref(params), C.c_double(omega), _d(r),
                                           _d(z))
        return z

    def mg_clean_vcycle(self, sys: MgSystem, params: MgParams, r):
        """z = M r: one application of the clean-mode symmetric V-cycle."""
        s = sys.to_c()
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        self.lib.or_mg_clean_vcycle(C.byref(s), C.byref(params), _d(r), _d(z))
        return z

    # ------------------------------------------------ single-phase build
    def mg_coarsen(self, A):
        """or_mg_coarsen: the aggregated coarse rows of one level (2x2x2, or
        2x2 for a one-plane level)."""
        nz, ny, nx = A.shape[:3]
        A = np.ascontiguousarray(A, dtype=np.float64)
        out = np.empty((1 if nz == 1 else nz // 2, ny // 2, nx // 2, 4))
        self.lib.or_mg_coarsen(_d(A), _sz(nx), _sz(ny), _sz(nz), _d(out))
        return out

    def mg_flat_hierarchy(self, A0, b0, max_levels):
        """The flat-entry MGPCG's hierarchy of one system: Jet's level sizing
        (FdmMgUtils3 / FdmMgUtils2 for one plane), the coarse rows aggregated
        level by level from A0; b on level 0, x zero."""
        nz, ny, nx = A0.shape[:3]
        if nz == 1:
            sizes = [(a, b, 1) for a, b in self.levels_with_finest2((nx, ny), max_levels)]
        else:
            sizes = self.levels_with_finest((nx, ny, nz), max_levels)
        s = MgSystem()
        A = np.ascontiguousarray(A0, dtype=np.float64)
        for l, (lx, ly, lz) in enumerate(sizes):
            if l > 0:
                A = self.mg_coarsen(A)
            assert A.shape[:3] == (lz, ly, lx)
            s.A.append(A)
            s.b.append(np.ascontiguousarray(b0, dtype=np.float64).copy() if l == 0
                       else np.zeros((lz, ly, lx)))
            s.x.append(np.zeros((lz, ly, lx)))
        return s

    def sp_markers(self, bsdf, fsdf):
        m = np.empty(bsdf.shape, np.int8)
        self.lib.or_sp_markers(_d(bsdf), _d(fsdf), _c(m), _sz(m.size))
        return m

    def sp_coarsen_markers(self, fine):
        nz, ny, nx = fine.shape
        out = np.empty((nz // 2, ny // 2, nx // 2), np.int8)
        self.lib.or_sp_coarsen_markers(_c(fine), _c(out), _sz(nx // 2), _sz(ny // 2), _sz(nz // 2))
        return out

    def sp_build_system(self, markers, u, v, w, h):
        nx, ny, nz = _shape3(markers)
        A = np.empty((nz, ny, nx, 4))
        b = np.empty((nz, ny, nx))
        hh = (C.c_double * 3)(*h)
        self.lib.or_sp_build_system(_c(markers), _d(u), _d(v), _d(w), _sz(nx), _sz(ny), _sz(nz),
                                    hh, _d(A), _d(b))
        return A, b

    def diffusion_build(self, f, bsdf, fsdf, dt, mu, h, dirichlet=False):
        """Backward Euler diffusion system (A rows, b) of one data grid."""
        nx, ny, nz = _shape3(f)
        A = np.empty(f.shape + (4,))
        b = np.empty_like(f)
        self.lib.or_diffusion_build(_d(f), _d(bsdf), _d(fsdf), _sz(nx), _sz(ny), _sz(nz),
                                    _d(diffusion_c(dt, mu, h)), C.c_int(1 if dirichlet else 0),
                                    _d(A), _d(b))
        return A, b

    def sp_build_compressed(self, markers, u, v, w, h):
        """Compressed buildSingleSystem: (values, row_ptr, col_idx, b)."""
        nx, ny, nz = _shape3(markers)
        hh = np.asarray(h, np.float64)
        nnz = C.c_size_t()
        f = self.lib.or_sp_build_compressed
        f.restype = C.c_size_t
        args = (_c(markers), _d(u), _d(v), _d(w), _sz(nx), _sz(ny), _sz(nz), _d(hh))
        rows = f(*args, None, None, None, None, C.byref(nnz))
        val = np.empty(nnz.value)
        rp = np.empty(rows + 1, np.uint64)
        ci = np.empty(nnz.value, np.uint64)
        b = np.empty(rows)
        sz_p = C.POINTER(C.c_size_t)
        f(*args, _d(val), rp.ctypes.data_as(sz_p), ci.ctypes.data_as(sz_p), _d(b), C.byref(nnz))
        return val, rp, ci, b

    def face_fill_coarse(self, u, v, w, h, origin):
        nz, ny, nx = w.shape[0] - 1, w.shape[1], w.shape[2]
        cx, cy, cz = nx // 2, ny // 2, nz // 2
        ch = tuple(2.0 * x for x in h)
        cu = np.empty((cz, cy, cx + 1))
        cv = np.empty((cz, cy + 1, cx))
        cw = np.empty((cz + 1, cy, cx))
        self.lib.or_face_fill_coarse(_d(u), _d(v), _d(w), _sz(nx), _sz(ny), _sz(nz),
                                     (C.c_double * 3)(*h), (C.c_double * 3)(*origin),
                                     _d(cu), _d(cv), _d(cw), _sz(cx), _sz(cy), _sz(cz),
                                     (C.c_double * 3)(*ch))
        return cu, cv, cw, ch

    def sp_apply_gradient(self, markers, p, u, v, w, h):
        uo, vo, wo = u.copy(), v.copy(), w.copy()
        nx, ny, nz = _shape3(markers)
        self.lib.or_sp_apply_gradient(_c(markers), _d(p), _d(u), _d(v), _d(w), _d(uo), _d(vo),
                                      _d(wo), _sz(nx), _sz(ny), _sz(nz), (C.c_double * 3)(*h))
        return uo, vo, wo

    # ------------------------------------------ 2-D single phase ((ny, nx) arrays)
    def sp_coarsen_markers2(self, fine):
        ny, nx = fine.shape
        out = np.empty((ny // 2, nx // 2), np.int8)
        self.lib.or_sp_coarsen_markers2(_c(np.ascontiguousarray(fine)), _c(out), _sz(nx // 2),
                                        _sz(ny // 2))
        return out

    def sp_build_system2(self, markers, u, v, h):
        """-> (A as (1, ny, nx, 4) one-plane rows, b as (1, ny, nx)); u None: A only."""
        ny, nx = markers.shape
        A = np.empty((1, ny, nx, 4))
        b = np.zeros((1, ny, nx))
        self.lib.or_sp_build_system2(_c(np.ascontiguousarray(markers)),
                                     _d(u) if u is not None else None,
                                     _d(v) if v is not None else None, _sz(nx), _sz(ny),
                                     (C.c_double * 2)(*h), _d(A),
                                     _d(b) if u is not None else None)
        return A, b

    def sp_build_compressed2(self, markers, u, v, h):
        """Compressed 2-D buildSingleSystem: (values, row_ptr, col_idx, b)."""
        ny, nx = markers.shape
        hh = np.asarray(h, np.float64)
        nnz = C.c_size_t()
        f = self.lib.or_sp_build_compressed2
        f.restype = C.c_size_t
        args = (_c(np.ascontiguousarray(markers)), _d(u), _d(v), _sz(nx), _sz(ny), _d(hh))
        rows = f(*args, None, None, None, None, C.byref(nnz))
        val = np.empty(nnz.value)
        rp = np.empty(rows + 1, np.uint64)
        ci = np.empty(nnz.value, np.uint64)
        b = np.empty(rows)
        sz_p = C.POINTER(C.c_size_t)
        f(*args, _d(val), rp.ctypes.data_as(sz_p), ci.ctypes.data_as(sz_p), _d(b), C.byref(nnz))
        return val, rp, ci, b

    def sp_apply_gradient2(self, markers, p, u, v, h):
        uo, vo = u.copy(), v.copy()
        ny, nx = markers.shape
        self.lib.or_sp_apply_gradient2(_c(np.ascontiguousarray(markers)), _d(p), _d(u), _d(v),
                                       _d(uo), _d(vo), _sz(nx), _sz(ny), (C.c_double * 2)(*h))
        return uo, vo

    def sp_build_mg2(self, u, v, h, bsdf, fsdf, max_levels):
        """GridSinglePhasePressureSolver2::buildMarkers + buildSystem (src/jet/
        grid_single_phase_pressure_solver2.cpp:218-354) on (ny, nx) cell-centred
        SDF samples: (markers per level, MgSystem of one-plane levels).  Coarse
        levels carry A only (their b is what the V-cycle restricts)."""
        ny, nx = bsdf.shape
        sizes = self.levels_with_finest2((nx, ny), max_levels)
        markers = [self.sp_markers(np.ascontiguousarray(bsdf), np.ascontiguousarray(fsdf))]
        for _ in sizes[1:]:
            markers.append(self.sp_coarsen_markers2(markers[-1]))
        sys = MgSystem()
        hh = tuple(h)
        for l, m in enumerate(markers):
            A, b = self.sp_build_system2(m, u if l == 0 else None, v if l == 0 else None, hh)
            sys.A.append(A)
            sys.b.append(b)
            sys.x.append(np.zeros_like(b))
            hh = (2.0 * hh[0], 2.0 * hh[1])
        return markers, sys

    def sp_build_mg(self, u, v, w, h, origin, bsdf, fsdf, max_levels):
        """GridSinglePhasePressureSolver3::buildMarkers + buildSystem
        (src/jet/grid_single_phase_pressure_solver3.cpp:248-399): returns
        (markers per level, MgSystem)."""
        nz, ny, nx = bsdf.shape
        sizes = self.levels_with_finest((nx, ny, nz), max_levels)
        markers = [self.sp_markers(bsdf, fsdf)]
        for _ in sizes[1:]:
            markers.append(self.sp_coarsen_markers(markers[-1]))
        sys = MgSystem()
        cu, cv, cw, ch = u, v, w, tuple(h)
        for l, m in enumerate(markers):
            if l > 0:
                cu, cv, cw, ch = self.face_fill_coarse(cu, cv, cw, ch, origin)
            A, b = self.sp_build_system(m, cu, cv, cw, ch)
            sys.A.append(A)
            sys.b.append(b)
            sys.x.append(np.zeros_like(b))
        return markers, sys

    # -------------------------------------------------- fractional build
    def fraction_inside_sdf(self, a, b):
        return self.lib.or_fraction_inside_sdf(a, b)

    def fraction_inside(self, bl, br, tl, tr):
        return self.lib.or_fraction_inside(bl, br, tl, tr)

    def fr_face_weights(self, vertex_sdf):
        nz, ny, nx = (s - 1 for s in vertex_sdf.shape)
        uw = np.empty((nz, ny, nx + 1), np.float32)
        vw = np.empty((nz, ny + 1, nx), np.float32)
        ww = np.empty((nz + 1, ny, nx), np.float32)
        self.lib.or_fr_face_weights(_d(vertex_sdf), _sz(nx), _sz(ny), _sz(nz), _f(uw), _f(vw), _f(ww))
        return uw, vw, ww

    def fr_restrict_float(self, fine, coarse_shape):
        out = np.empty(coarse_shape, np.float32)
        fz, fy, fx = fine.shape
        cz, cy, cx = coarse_shape
        self.lib.or_fr_restrict_float(_f(fine), _sz(fx), _sz(fy), _sz(fz), _f(out), _sz(cx), _sz(cy), _sz(cz))
        return out

    def fr_build_system(self, fsdf, uw, vw, ww, u, v, w, h, bu=None, bv=None, bw=None):
        nx, ny, nz = _shape3(fsdf)
        A = np.empty((nz, ny, nx, 4))
        b = np.empty((nz, ny, nx))
        nul = C.cast(None, _dp)
        self.lib.or_fr_build_system(_f(fsdf), _f(uw), _f(vw), _f(ww), _d(u), _d(v), _d(w),
                                    _d(bu) if bu is not None else nul,
                                    _d(bv) if bv is not None else nul,
                                    _d(bw) if bw is not None else nul,
                                    _sz(nx), _sz(ny), _sz(nz), (C.c_double * 3)(*h), _d(A), _d(b))
        return A, b

    def fr_apply_gradient(self, fsdf, uw, vw, ww, p, u, v, w, h):
        uo, vo, wo = u.copy(), v.copy(), w.copy()
        nx, ny, nz = _shape3(fsdf)
        self.lib.or_fr_apply_gradient(_f(fsdf), _f(uw), _f(vw), _f(ww), _d(p), _d(u), _d(v), _d(w),
                                      _d(uo), _d(vo), _d(wo), _sz(nx), _sz(ny), _sz(nz),
                                      (C.c_double * 3)(*h))
        return uo, vo, wo

    def fr_build_mg(self, u, v, w, h, origin, vertex_sdf, fluid_sdf_cc, max_levels,
                    bu=None, bv=None, bw=None):
        """GridFractionalSinglePhasePressureSolver3::buildWeights + buildSystem
        (src/jet/grid_fractional_single_phase_pressure_solver3.cpp:520-713).
        fluid_sdf_cc is cast to float as Jet stores it (:548-551)."""
        nz, ny, nx = fluid_sdf_cc.shape
        sizes = self.levels_with_finest((nx, ny, nz), max_levels)
        fs = [np.ascontiguousarray(fluid_sdf_cc, dtype=np.float32)]
        uw, vw, ww = self.fr_face_weights(vertex_sdf)
        UW, VW, WW = [uw], [vw], [ww]
        for (cx, cy, cz) in sizes[1:]:
            fs.append(self.fr_restrict_float(fs[-1], (cz, cy, cx)))
            UW.append(self.fr_restrict_float(UW[-1], (cz, cy, cx + 1)))
            VW.append(self.fr_restrict_float(VW[-1], (cz, cy + 1, cx)))
            WW.append(self.fr_restrict_float(WW[-1], (cz + 1, cy, cx)))
        sys = MgSystem()
        cu, cv, cw, ch = u, v, w, tuple(h)
        for l in range(len(sizes)):
            if l > 0:
                # Coarse right-hand sides are overwritten by the V-cycle's
                # restriction before use, so a moving boundary is only
                # applied on the finest level here.
                cu, cv, cw, ch = self.face_fill_coarse(cu, cv, cw, ch, origin)
            A, b = self.fr_build_system(fs[l], UW[l], VW[l], WW[l], cu, cv, cw, ch,
                                        bu if l == 0 else None, bv if l == 0 else None,
                                        bw if l == 0 else None)
            sys.A.append(A)
            sys.b.append(b)
            sys.x.append(np.zeros_like(b))
        return (fs, UW, VW, WW), sys

    # ---------------------------------------------- fractional build, 2-D
    def fr_face_weights2(self, bsdf_ghost):
        """Face weights of GridFractionalSinglePhasePressureSolver2::buildWeights
        from the boundary SDF at the cell centres of the grid plus one ghost
        cell per side, an (ny + 2, nx + 2) array."""
        ny, nx = (s - 2 for s in bsdf_ghost.shape)
        uw = np.empty((ny, nx + 1), np.float32)
        vw = np.empty((ny + 1, nx), np.float32)
        self.lib.or_fr_face_weights2(_d(bsdf_ghost), _sz(nx), _sz(ny), _f(uw), _f(vw))
        return uw, vw

    def fr_restrict_float2(self, fine, coarse_shape):
        out = np.empty(coarse_shape, np.float32)
        fy, fx = fine.shape
        cy, cx = coarse_shape
        self.lib.or_fr_restrict_float2(_f(fine), _sz(fx), _sz(fy), _f(out), _sz(cx), _sz(cy))
        return out

    def fr_build_system2(self, fsdf, uw, vw, u, v, h, bu=None, bv=None):
        """(A as (1, ny, nx, 4) rows with front = 0, b as (1, ny, nx))."""
        ny, nx = fsdf.shape
        A = np.empty((1, ny, nx, 4))
        b = np.empty((1, ny, nx))
        nul = C.cast(None, _dp)
        self.lib.or_fr_build_system2(_f(fsdf), _f(uw), _f(vw), _d(u), _d(v),
                                     _d(bu) if bu is not None else nul,
                                     _d(bv) if bv is not None else nul,
                                     _sz(nx), _sz(ny), (C.c_double * 2)(*h), _d(A), _d(b))
        return A, b

    def fr_apply_gradient2(self, fsdf, uw, vw, p, u, v, h):
        uo, vo = u.copy(), v.copy()
        ny, nx = fsdf.shape
        self.lib.or_fr_apply_gradient2(_f(fsdf), _f(uw), _f(vw), _d(p), _d(u), _d(v), _d(uo),
                                       _d(vo), _sz(nx), _sz(ny), (C.c_double * 2)(*h))
        return uo, vo

    def fr_build_mg2(self, u, v, h, bsdf_ghost, fluid_sdf_cc, max_levels, bu=None, bv=None):
        """GridFractionalSinglePhasePressureSolver2::buildWeights + buildSystem
        (src/jet/grid_fractional_single_phase_pressure_solver2.cpp:414-563) on
        (ny, nx) cell-centred samples: ((fluid SDF, u weights, v weights) per
        level, MgSystem of one-plane levels).  Coarse levels carry A only:
        their right-hand sides are what the V-cycle restricts."""
        ny, nx = fluid_sdf_cc.shape
        sizes = self.levels_with_finest2((nx, ny), max_levels)
        fs = [np.ascontiguousarray(fluid_sdf_cc, dtype=np.float32)]
        uw, vw = self.fr_face_weights2(np.ascontiguousarray(bsdf_ghost, dtype=np.float64))
        UW, VW = [uw], [vw]
        for (cx, cy) in sizes[1:]:
            fs.append(self.fr_restrict_float2(fs[-1], (cy, cx)))
            UW.append(self.fr_restrict_float2(UW[-1], (cy, cx + 1)))
            VW.append(self.fr_restrict_float2(VW[-1], (cy + 1, cx)))
        sys = MgSystem()
        for l, (cx, cy) in enumerate(sizes):
            hl = (h[0] * 2 ** l, h[1] * 2 ** l)
            if l == 0:
                A, b = self.fr_build_system2(fs[0], UW[0], VW[0], np.ascontiguousarray(u),
                                             np.ascontiguousarray(v), hl, bu, bv)
            else:
                A, _ = self.fr_build_system2(fs[l], UW[l], VW[l], np.zeros((cy, cx + 1)),
                                             np.zeros((cy + 1, cx)), hl)
                b = np.zeros((1, cy, cx))
            sys.A.append(A)
            sys.b.append(b)
            sys.x.append(np.zeros_like(b))
        return (fs, UW, VW), sys


class Ref(_BoundaryCalls, _CsrCalls):
    """Jet's own compiled linear solvers (oracle/_ref/libjetref.so), and the
    compiled grids, samplers, extrapolation, friction helper and rigid
    collider that Jet's velocity boundary conditions are made of."""

    _p = "ref_"

    def __init__(self, path=REF_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        lib.ref_dot.restype = C.c_double
        lib.ref_linfnorm.restype = C.c_double
        lib.ref_fraction_inside.restype = C.c_double
        lib.ref_fraction_inside.argtypes = [C.c_double] * 4
        lib.ref_fraction_inside_sdf.restype = C.c_double
        lib.ref_fraction_inside_sdf.argtypes = [C.c_double] * 2
        lib.ref_levels_with_finest.restype = C.c_size_t
        lib.ref_levels_with_finest2.restype = C.c_size_t
        lib.ref_quiet()
        self.lib = lib

    def mvm(self, A, v):
        out = np.empty_like(v)
        self.lib.ref_mvm(_d(A), _d(v), _d(out), *map(_sz, _shape3(v)))
        return out

    def residual(self, A, x, b):
        out = np.empty_like(x)
        self.lib.ref_residual(_d(A), _d(x), _d(b), _d(out), *map(_sz, _shape3(x)))
        return out

    def dot(self, a, b):
        return self.lib.ref_dot(_d(a), _d(b), *map(_sz, _shape3(a)))

    def linfnorm(self, a):
        return self.lib.ref_linfnorm(_d(a), *map(_sz, _shape3(a)))

    def axpy(self, a, x, y):
        out = np.empty_like(x)
        self.lib.ref_axpy(C.c_double(a), _d(x), _d(y), _d(out), *map(_sz, _shape3(x)))
        return out

    def cg_solve(self, A, b, max_iter, tol):
        x = np.zeros_like(b)
        it, res = C.c_uint(), C.c_double()
        ok = self.lib.ref_cg_solve(_d(A), _d(b), _d(x), *map(_sz, _shape3(b)), C.c_uint(max_iter),
                                   C.c_double(tol), C.byref(it), C.byref(res))
        return x, it.value, res.value, bool(ok)

    def gs_relax(self, A, b, sor, x, red_black=False):
        x = x.copy()
        self.lib.ref_gs_relax(_d(A), _d(b), C.c_double(sor), _d(x), *map(_sz, _shape3(x)),
                              C.c_int(int(red_black)))
        return x

    def jacobi_relax(self, A, b, x):
        out = np.empty_like(x)
        self.lib.ref_jacobi_relax(_d(A), _d(b), _d(x), _d(out), *map(_sz, _shape3(x)))
        return out

    def restrict(self, fine):
        nz, ny, nx = fine.shape
        out = np.empty((nz // 2, ny // 2, nx // 2))
        self.lib.ref_mg_restrict(_d(fine), _d(out), _sz(nx // 2), _sz(ny // 2), _sz(nz // 2))
        return out

    def correct(self, coarse, fine):
        fine = fine.copy()
        self.lib.ref_mg_correct(_d(coarse), _d(fine), *map(_sz, _shape3(fine)))
        return fine

    def levels_with_finest(self, shape_xyz, max_levels):
        lx, ly, lz = (C.c_size_t * MAX_LEVELS)(), (C.c_size_t * MAX_LEVELS)(), (C.c_size_t * MAX_LEVELS)()
        n = self.lib.ref_levels_with_finest(*map(_sz, shape_xyz), _sz(max_levels), lx, ly, lz)
        return [(lx[l], ly[l], lz[l]) for l in range(n)]

    def _mg_args(self, sys: MgSystem):
        n = len(sys.A)
        SZ = C.c_size_t * n
        nx = SZ(*[_shape3(x)[0] for x in sys.x])
        ny = SZ(*[_shape3(x)[1] for x in sys.x])
        nz = SZ(*[_shape3(x)[2] for x in sys.x])
        A = (_dp * n)(*[_d(a) for a in sys.A])
        X = (_dp * n)(*[_d(a) for a in sys.x])
        B = (_dp * n)(*[_d(a) for a in sys.b])
        return _sz(n), nx, ny, nz, A, X, B

    def mg_solve(self, sys: MgSystem, p: MgParams):
        ok = self.lib.ref_mg_solve(*self._mg_args(sys), _sz(p.max_levels), C.c_uint(p.restrict_iter),
                                   C.c_uint(p.correct_iter), C.c_uint(p.coarsest_iter),
                                   C.c_uint(p.final_iter), C.c_double(p.max_tol),
                                   C.c_double(p.sor), C.c_int(p.red_black))
        return bool(ok)

    def mgpcg_solve(self, sys: MgSystem, p: MgParams, max_iter):
        it, res = C.c_uint(), C.c_double()
        ok = self.lib.ref_mgpcg_solve(*self._mg_args(sys), C.c_uint(max_iter), _sz(p.max_levels),
                                      C.c_uint(p.restrict_iter), C.c_uint(p.correct_iter),
                                      C.c_uint(p.coarsest_iter), C.c_uint(p.final_iter),
                                      C.c_double(p.max_tol), C.c_double(p.sor),
                                      C.c_int(p.red_black), C.byref(it), C.byref(res))
        return bool(ok), it.value, res.value

    def levels_with_finest2(self, shape_xy, max_levels):
        lx, ly = (C.c_size_t * MAX_LEVELS)(), (C.c_size_t * MAX_LEVELS)()
        n = self.lib.ref_levels_with_finest2(*map(_sz, shape_xy), _sz(max_levels), lx, ly)
        return [(lx[l], ly[l]) for l in range(n)]

    def restrict2(self, fine):
        ny, nx = fine.shape
        fine = np.ascontiguousarray(fine, dtype=np.float64)
        out = np.empty((ny // 2, nx // 2))
        self.lib.ref_mg_restrict2(_d(fine), _d(out), _sz(nx // 2), _sz(ny // 2))
        return out

    def correct2(self, coarse, fine):
        fine = np.array(fine, dtype=np.float64)
        coarse = np.ascontiguousarray(coarse, dtype=np.float64)
        ny, nx = fine.shape
        self.lib.ref_mg_correct2(_d(coarse), _d(fine), _sz(nx), _sz(ny))
        return fine

    def mgpcg_solve2(self, sys: MgSystem, p: MgParams, max_iter):
        """FdmMgpcgSolver2::solve on a hierarchy of one-plane levels."""
        n, nx, ny, _, A, X, B = self._mg_args(sys)
        it, res = C.c_uint(), C.c_double()
        ok = self.lib.ref_mgpcg_solve2(n, nx, ny, A, X, B, C.c_uint(max_iter), _sz(p.max_levels),
                                       C.c_uint(p.restrict_iter), C.c_uint(p.correct_iter),
                                       C.c_uint(p.coarsest_iter), C.c_uint(p.final_iter),
                                       C.c_double(p.max_tol), C.c_double(p.sor),
                                       C.c_int(p.red_black), C.byref(it), C.byref(res))
        return bool(ok), it.value, res.value

    def fraction_inside_sdf(self, a, b):
        return self.lib.ref_fraction_inside_sdf(a, b)

    def fr_solve2(self, u, v, h, origin, bsdf_ghost, fluid_sdf_cc, max_levels, p, bu=None, bv=None):
        """Jet's compiled GridFractionalSinglePhasePressureSolver2::solve with a
        recording multigrid solver: (A of every level as (1, ny_l, nx_l, 4) rows,
        b of the finest level, and the projected (uo, vo) of the given pressure p)."""
        ny, nx = fluid_sdf_cc.shape
        sizes = Oracle().levels_with_finest2((nx, ny), max_levels)
        A = np.empty(sum(cx * cy for cx, cy in sizes) * 4)
        b = np.empty((1, ny, nx))
        uo = np.empty((ny, nx + 1))
        vo = np.empty((ny + 1, nx))
        nul = C.cast(None, _dp)
        self.lib.ref_fr_solve2(_sz(nx), _sz(ny), (C.c_double * 2)(*h), (C.c_double * 2)(*origin),
                               _d(np.ascontiguousarray(u, np.float64)),
                               _d(np.ascontiguousarray(v, np.float64)),
                               _d(np.ascontiguousarray(bsdf_ghost, np.float64)),
                               _d(np.ascontiguousarray(fluid_sdf_cc, np.float64)),
                               _d(bu) if bu is not None else nul, _d(bv) if bv is not None else nul,
                               _sz(max_levels), _d(np.ascontiguousarray(p, np.float64)), _d(A),
                               _d(b), _d(uo), _d(vo))
        levels, off = [], 0
        for cx, cy in sizes:
            levels.append(A[off:off + cx * cy * 4].reshape(1, cy, cx, 4))
            off += cx * cy * 4
        return levels, b, uo, vo

    def diffusion_system(self, f, bsdf, fsdf, dt, mu, h, dirichlet=False):
        """The system Jet's GridBackwardEulerDiffusionSolver3 builds (A, b)."""
        nx, ny, nz = _shape3(f)
        A = np.empty(f.shape + (4,))
        b = np.empty_like(f)
        self.lib.ref_diffusion_system(_d(f), _d(bsdf), _d(fsdf), _sz(nx), _sz(ny), _sz(nz),
                                      _d(np.asarray(h, np.float64)), C.c_double(dt),
                                      C.c_double(mu), C.c_int(1 if dirichlet else 0), _d(A), _d(b))
        return A, b

    def diffusion_solve(self, f, bsdf, fsdf, dt, mu, h, dirichlet=False):
        """Jet's own diffusion solve (default ICCG) -> the new field."""
        nx, ny, nz = _shape3(f)
        out = np.empty_like(f)
        self.lib.ref_diffusion_solve(_d(f), _d(bsdf), _d(fsdf), _sz(nx), _sz(ny), _sz(nz),
                                     _d(np.asarray(h, np.float64)), C.c_double(dt), C.c_double(mu),
                                     C.c_int(1 if dirichlet else 0), _d(out))
        return out

    def fraction_inside(self, bl, br, tl, tr):
        return self.lib.ref_fraction_inside(bl, br, tl, tr)


def have_ref():
    return os.path.exists(REF_PATH)
