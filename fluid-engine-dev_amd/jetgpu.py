"""ctypes binding of libjetgpu (include/jetgpu.h) for tests and benchmarks.

This module is plumbing: it declares the C ABI and offers small helpers that
pass torch device tensors (or numpy host arrays) to it.  All computation is
in lib/libjetgpu.so; there is no fallback -- importing this module raises if
the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libjetgpu.so")

JG_OK = 0
ERRORS = {1: "invalid argument", 2: "out of memory", 3: "HIP error", 4: "non-finite value",
          5: "no device", 6: "unsupported"}
JG_FP64, JG_MIXED = 0, 1
JG_STOP_JET, JG_STOP_RELATIVE = 0, 1
JG_SMOOTH_RBGS, JG_SMOOTH_JACOBI = 0, 1
(JG_KERNEL_RBGS, JG_KERNEL_DIR_SPMV, JG_KERNEL_RESIDUAL, JG_KERNEL_UPDATE_XR, JG_KERNEL_DOT,
 JG_KERNEL_COPY, JG_KERNEL_SWEEP, JG_KERNEL_SWEEP2, JG_KERNEL_PRE0, JG_KERNEL_POST0) = range(10)
(JG_SCENE_CLOSED_SPHERE, JG_SCENE_DAM_BREAK, JG_SCENE_SPHERE_POOL, JG_SCENE_BATCH_MEMBER,
 JG_SCENE_TORUS_POOL) = 1, 2, 3, 4, 5


class MgParams(C.Structure):
    _fields_ = [("max_levels", C.c_size_t), ("restrict_iter", C.c_uint),
                ("correct_iter", C.c_uint), ("coarsest_iter", C.c_uint),
                ("final_iter", C.c_uint), ("max_tol", C.c_double), ("sor", C.c_double),
                ("red_black", C.c_int)]


class GpuOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("stop_rule", C.c_int), ("rel_tol", C.c_double),
                ("smoother", C.c_int), ("omega", C.c_double), ("max_refine", C.c_uint),
                ("inner_rel_tol", C.c_double), ("check_every", C.c_uint),
                ("sweep_pairs", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_uint), ("refinements", C.c_uint),
                ("last_residual", C.c_double), ("rel_residual", C.c_double),
                ("converged", C.c_int), ("solve_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SceneDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_size_t), ("h", C.c_double),
                ("seed", C.c_ulonglong), ("fluid_fraction", C.c_double)]


class RigidCollider(C.Structure):
    """jg_rigid_collider: RigidBodyCollider3 velocities, surface translation, friction."""
    _fields_ = [("linear_velocity", C.c_double * 3), ("angular_velocity", C.c_double * 3),
                ("translation", C.c_double * 3), ("friction", C.c_double)]

    @classmethod
    def make(cls, linear=(0, 0, 0), angular=(0, 0, 0), translation=(0, 0, 0), friction=0.0):
        return cls((C.c_double * 3)(*linear), (C.c_double * 3)(*angular),
                   (C.c_double * 3)(*translation), friction)


# closed domain walls (Jet's kDirection* bits)
JG_DIR_LEFT, JG_DIR_RIGHT, JG_DIR_DOWN, JG_DIR_UP, JG_DIR_BACK, JG_DIR_FRONT = 1, 2, 4, 8, 16, 32
JG_DIR_ALL = 63

_vp = C.c_void_p
_sz = C.c_size_t
_d = C.c_double
_u = C.c_uint

# name -> (restype, argtypes): every symbol declared in include/jetgpu.h.
SIGNATURES = {
    "jg_mg_params_default": (None, [C.POINTER(MgParams)]),
    "jg_gpu_options_default": (None, [C.POINTER(GpuOptions)]),
    "jg_solver_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "jg_solver_destroy": (None, [_vp]),
    "jg_last_error": (C.c_char_p, [_vp]),
    "jg_solver_stream": (_vp, [_vp]),
    "jg_solver_set_levels": (C.c_int, [_vp, _sz, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)]),
    "jg_solver_matrix": (_vp, [_vp, _sz]),
    "jg_solver_rhs": (_vp, [_vp]),
    "jg_solver_solution": (_vp, [_vp]),
    "jg_solver_upload": (C.c_int, [_vp, _sz, _vp, _vp]),
    "jg_solver_download_x": (C.c_int, [_vp, _vp]),
    "jg_solver_matrices_changed": (C.c_int, [_vp]),
    "jg_solver_cg": (C.c_int, [_vp, _u, _d, C.POINTER(SolveInfo)]),
    "jg_solver_mgpcg": (C.c_int, [_vp, _u, C.POINTER(MgParams), C.POINTER(GpuOptions),
                                  C.POINTER(SolveInfo)]),
    "jg_mg_vcycle": (C.c_int, [_vp, C.POINTER(MgParams), C.POINTER(GpuOptions), _vp, _vp]),
    "jg_solver_mg": (C.c_int, [_vp, C.POINTER(MgParams), C.POINTER(GpuOptions),
                               C.POINTER(SolveInfo)]),
    "jg_mg_solve_host": (C.c_int, [_vp, _sz, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz),
                                   C.POINTER(_vp), _vp, _vp, C.POINTER(MgParams),
                                   C.POINTER(GpuOptions), C.POINTER(SolveInfo)]),
    "jg_cg_solve_host": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp, _vp, _u, _d, C.POINTER(SolveInfo)]),
    "jg_mgpcg_flat_solve_host": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp, _vp, _u,
                                           C.POINTER(MgParams), C.POINTER(GpuOptions),
                                           C.POINTER(SolveInfo)]),
    "jg_solver_coarsen": (C.c_int, [_vp]),
    "jg_csr_cg_solve_host": (C.c_int, [_vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _u, _d,
                                       C.POINTER(SolveInfo)]),
    "jg_csr_mvm_host": (C.c_int, [_vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp]),
    "jg_sp_build_compressed": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d, C.POINTER(_sz),
                                         C.POINTER(_sz)]),
    "jg_solver_csr_cg": (C.c_int, [_vp, _u, _d, C.POINTER(SolveInfo)]),
    "jg_sp_build_compressed2": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _d, _d, C.POINTER(_sz),
                                          C.POINTER(_sz)]),
    "jg_diffusion_build": (C.c_int, [_vp, _vp, _vp, _vp, _d, _d, _d, _d, _d, C.c_int]),
    "jg_csr_decompress": (C.c_int, [_vp]),
    "jg_csr_array": (_vp, [_vp, C.c_int]),
    "jg_csr_index": (_vp, [_vp, C.c_int]),
    "jg_csr_residual_host": (C.c_int, [_vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jg_mgpcg_solve_host": (C.c_int, [_vp, _sz, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz),
                                      C.POINTER(_vp), _vp, _vp, _u, C.POINTER(MgParams),
                                      C.POINTER(GpuOptions), C.POINTER(SolveInfo)]),
    "jg_blas_mvm": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp, _vp]),
    "jg_blas_residual": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp, _vp, _vp]),
    "jg_blas_dot": (C.c_int, [_vp, _sz, _vp, _vp, C.POINTER(_d)]),
    "jg_blas_axpy": (C.c_int, [_vp, _sz, _d, _vp, _vp, _vp]),
    "jg_relax_red_black": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp, _d, _vp]),
    "jg_relax_red_black_noalloc": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp, _d, _vp]),
    "jg_relax_jacobi": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp, _vp, _vp]),
    "jg_mg_restrict": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp]),
    "jg_mg_correct": (C.c_int, [_vp, _sz, _sz, _sz, _vp, _vp]),
    "jg_mg_restrict2": (C.c_int, [_vp, _sz, _sz, _vp, _vp]),
    "jg_sp_build2": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_double]),
    "jg_sp_apply_gradient2": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_double]),
    "jg_mg_correct2": (C.c_int, [_vp, _sz, _sz, _vp, _vp]),
    "jg_solver_time_kernel": (C.c_int, [_vp, C.c_int, C.c_int, _u, C.POINTER(_d), C.POINTER(_d)]),
    "jg_sp_build": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d]),
    "jg_sp_apply_gradient": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d]),
    "jg_solver_markers": (_vp, [_vp, _sz]),
    "jg_fr_build": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d]),
    "jg_fr_apply_gradient": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d]),
    "jg_fr_array": (_vp, [_vp, _sz, C.c_int]),
    "jg_fr_build2": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d]),
    "jg_fr_apply_gradient2": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _d, _d]),
    "jg_constrain_velocity": (C.c_int, [_vp, C.c_int, C.POINTER(_d), C.POINTER(_d), _vp, _vp, _vp,
                                        _vp, C.POINTER(RigidCollider), _u, C.c_int]),
    "jg_batch_create": (C.c_int, [C.c_int, _sz, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz),
                                  C.POINTER(_sz), C.POINTER(_vp)]),
    "jg_batch_destroy": (None, [_vp]),
    "jg_batch_last_error": (C.c_char_p, [_vp]),
    "jg_batch_size": (_sz, [_vp]),
    "jg_batch_system": (_vp, [_vp, _sz]),
    "jg_batch_mgpcg": (C.c_int, [_vp, _u, C.POINTER(MgParams), C.POINTER(GpuOptions),
                                 C.POINTER(SolveInfo), C.POINTER(_d)]),
    "jg_scene_generate_host": (C.c_int, [C.POINTER(SceneDesc), _vp, _vp, _vp, _vp, _vp, _vp]),
    "jg_scene_generate_device": (C.c_int, [_vp, C.POINTER(SceneDesc), _vp, _vp, _vp, _vp, _vp, _vp]),
}


def load(path: str = LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() or `make -C fluid-engine-dev_amd`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def mg_params(**kw):
    p = MgParams()
    lib().jg_mg_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gpu_options(**kw):
    o = GpuOptions()
    lib().jg_gpu_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def scene_desc(kind, n, seed=0, h=1.0, fluid_fraction=0.6):
    return SceneDesc(kind, n, h, seed, fluid_fraction)


def generate_scene_host(desc: SceneDesc, vertex=False):
    """Host generator -> dict of numpy arrays (Jet layout)."""
    n = desc.n
    out = {
        "u": np.empty((n, n, n + 1)), "v": np.empty((n, n + 1, n)), "w": np.empty((n + 1, n, n)),
        "bsdf": np.empty((n, n, n)), "fsdf": np.empty((n, n, n)),
    }
    vs = np.empty((n + 1, n + 1, n + 1)) if vertex else None
    ptr = lambda a: a.ctypes.data_as(_vp) if a is not None else None
    st = lib().jg_scene_generate_host(C.byref(desc), ptr(out["u"]), ptr(out["v"]), ptr(out["w"]),
                                      ptr(out["bsdf"]), ptr(out["fsdf"]), ptr(vs))
    if st != JG_OK:
        raise RuntimeError(f"jg_scene_generate_host: {ERRORS.get(st, st)}")
    if vertex:
        out["vertex"] = vs
    return out


class JgError(RuntimeError):
    pass


class Batch:
    """Owns one jg_batch: independent systems solved by one replayed graph."""

    def __init__(self, sizes, levels, device=0):
        n = len(sizes)
        arr = lambda v: (_sz * n)(*v)
        h = _vp()
        st = lib().jg_batch_create(device, n, arr([s[0] for s in sizes]), arr([s[1] for s in sizes]),
                                   arr([s[2] for s in sizes]), arr(levels), C.byref(h))
        if st != JG_OK:
            raise JgError(f"jg_batch_create: {ERRORS.get(st, st)}")
        self.h = h
        self.systems = [Solver.wrap(lib().jg_batch_system(h, i)) for i in range(n)]
        for sys_, (nx, ny, nz), lv in zip(self.systems, sizes, levels):
            sys_.sizes = [(nx >> l, ny >> l, nz >> l) for l in range(lv)]

    def close(self):
        if self.h:
            lib().jg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mgpcg(self, max_iter, params: MgParams, options: GpuOptions | None = None):
        """-> (list of SolveInfo, total device ms)"""
        n = len(self.systems)
        infos = (SolveInfo * n)()
        ms = _d()
        st = lib().jg_batch_mgpcg(self.h, max_iter, C.byref(params),
                                  C.byref(options) if options is not None else None, infos,
                                  C.byref(ms))
        if st != JG_OK:
            msg = lib().jg_batch_last_error(self.h)
            raise JgError(f"batch mgpcg: {ERRORS.get(st, st)}: {msg.decode() if msg else ''}")
        return list(infos), ms.value


class Solver:
    """Owns one jg_solver handle (one HIP stream on one device)."""

    @classmethod
    def wrap(cls, handle):
        """A view of a solver owned elsewhere (a batch member): never destroyed here."""
        obj = cls.__new__(cls)
        obj.h = _vp(handle)
        obj._owned = False
        return obj

    def __init__(self, device=0):
        h = _vp()
        st = lib().jg_solver_create(device, C.byref(h))
        if st != JG_OK:
            raise JgError(f"jg_solver_create: {ERRORS.get(st, st)}")
        self.h = h
        self._owned = True

    def close(self):
        if self.h and getattr(self, "_owned", True):
            lib().jg_solver_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st, what):
        if st != JG_OK:
            msg = lib().jg_last_error(self.h)
            raise JgError(f"{what}: {ERRORS.get(st, st)}: {msg.decode() if msg else ''}")

    # -------------------------------------------------- resident system
    def set_levels(self, sizes):
        n = len(sizes)
        SZ = _sz * n
        self.check(lib().jg_solver_set_levels(self.h, n, SZ(*[s[0] for s in sizes]),
                                              SZ(*[s[1] for s in sizes]), SZ(*[s[2] for s in sizes])),
                   "set_levels")
        self.sizes = list(sizes)

    def matrix_ptr(self, level):
        return lib().jg_solver_matrix(self.h, level)

    def rhs_ptr(self):
        return lib().jg_solver_rhs(self.h)

    def solution_ptr(self):
        return lib().jg_solver_solution(self.h)

    def upload(self, level, A, b=None):
        A = np.ascontiguousarray(A, dtype=np.float64)
        bp = np.ascontiguousarray(b, dtype=np.float64).ctypes.data_as(_vp) if b is not None else None
        self.check(lib().jg_solver_upload(self.h, level, A.ctypes.data_as(_vp), bp), "upload")

    def download_x(self):
        nx, ny, nz = self.sizes[0]
        x = np.empty((nz, ny, nx))
        self.check(lib().jg_solver_download_x(self.h, x.ctypes.data_as(_vp)), "download_x")
        return x

    def cg(self, max_iter, tol):
        info = SolveInfo()
        self.check(lib().jg_solver_cg(self.h, max_iter, tol, C.byref(info)), "cg")
        return info

    def cg_host(self, A, b, max_iter, tol):
        """jg_cg_solve_host on host arrays: A (nz, ny, nx, 4) Jet rows, b (nz, ny, nx)."""
        import numpy as np

        A = np.ascontiguousarray(A, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        nz, ny, nx = b.shape
        x = np.empty_like(b)
        info = SolveInfo()
        self.check(lib().jg_cg_solve_host(self.h, nx, ny, nz, A.ctypes.data, b.ctypes.data,
                                          x.ctypes.data, max_iter, tol, C.byref(info)), "cg_host")
        return x, info

    def mgpcg_flat_host(self, A, b, max_iter, params: MgParams, options: GpuOptions | None = None):
        """jg_mgpcg_flat_solve_host: MGPCG on a single-level system (host
        arrays A (nz, ny, nx, 4), b (nz, ny, nx)); the hierarchy and its
        coarse matrices are made on the device.  Returns (x, info)."""
        import numpy as np

        A = np.ascontiguousarray(A, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        nz, ny, nx = b.shape
        x = np.empty_like(b)
        info = SolveInfo()
        self.check(lib().jg_mgpcg_flat_solve_host(self.h, nx, ny, nz, A.ctypes.data, b.ctypes.data,
                                                  x.ctypes.data, max_iter, C.byref(params),
                                                  C.byref(options) if options is not None else None,
                                                  C.byref(info)), "mgpcg_flat_host")
        return x, info

    def coarsen(self):
        """jg_solver_coarsen: the coarse matrices of every level from level 0's."""
        self.check(lib().jg_solver_coarsen(self.h), "coarsen")

    def csr_cg(self, values, row_ptr, col_idx, b, max_iter, tol):
        """FdmCgSolver3::solveCompressed on host CSR arrays (numpy: float64
        values, uint64 row pointers / column indices); returns (x, info)."""
        import numpy as np

        values = np.ascontiguousarray(values, np.float64)
        row_ptr = np.ascontiguousarray(row_ptr, np.uint64)
        col_idx = np.ascontiguousarray(col_idx, np.uint64)
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        info = SolveInfo()
        self.check(lib().jg_csr_cg_solve_host(self.h, b.size, values.size, values.ctypes.data,
                                              row_ptr.ctypes.data, col_idx.ctypes.data,
                                              b.ctypes.data, x.ctypes.data, max_iter, tol,
                                              C.byref(info)), "csr_cg")
        return x, info

    def sp_build_compressed(self, u, v, w, bsdf, fsdf, h):
        """Resident compressed system of the finest level; returns (rows, nnz)."""
        rows, nnz = _sz(), _sz()
        self.check(lib().jg_sp_build_compressed(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(bsdf),
                                                _ptr(fsdf), h, h, h, C.byref(rows), C.byref(nnz)),
                   "sp_build_compressed")
        return rows.value, nnz.value

    def sp_build_compressed2(self, u, v, bsdf, fsdf, h):
        """Resident compressed 2-D system of a one-plane finest level; (rows, nnz)."""
        rows, nnz = _sz(), _sz()
        self.check(lib().jg_sp_build_compressed2(self.h, _ptr(u), _ptr(v), _ptr(bsdf), _ptr(fsdf),
                                                 h[0], h[1], C.byref(rows), C.byref(nnz)),
                   "sp_build_compressed2")
        return rows.value, nnz.value

    def diffusion_build(self, f, bsdf, fsdf, dt, mu, h, dirichlet=False):
        hx, hy, hz = (h, h, h) if isinstance(h, (int, float)) else h
        self.check(lib().jg_diffusion_build(self.h, _ptr(f), _ptr(bsdf), _ptr(fsdf), dt, mu, hx, hy,
                                            hz, 1 if dirichlet else 0), "diffusion_build")

    def csr_solve(self, max_iter, tol):
        info = SolveInfo()
        self.check(lib().jg_solver_csr_cg(self.h, max_iter, tol, C.byref(info)), "csr_solve")
        return info

    def csr_decompress(self):
        self.check(lib().jg_csr_decompress(self.h), "csr_decompress")

    def csr_download(self, rows, nnz):
        """Host copies (values, row_ptr, col_idx, b, x) of the resident compressed system."""
        import numpy as np

        out = (np.empty(nnz), np.empty(rows + 1, np.int32), np.empty(nnz, np.int32),
               np.empty(rows), np.empty(rows))
        srcs = (lib().jg_csr_array(self.h, 0), lib().jg_csr_index(self.h, 0),
                lib().jg_csr_index(self.h, 1), lib().jg_csr_array(self.h, 1),
                lib().jg_csr_array(self.h, 2))
        import torch

        torch.cuda.synchronize()
        for a, p in zip(out, srcs):
            if a.size:
                _memcpy_dtoh(a, p)
        return out

    def csr_mvm(self, values, row_ptr, col_idx, v):
        import numpy as np

        values, row_ptr, col_idx, v = _csr_host(values, row_ptr, col_idx, v)
        out = np.empty_like(v)
        self.check(lib().jg_csr_mvm_host(self.h, v.size, values.size, values.ctypes.data,
                                         row_ptr.ctypes.data, col_idx.ctypes.data, v.ctypes.data,
                                         out.ctypes.data), "csr_mvm")
        return out

    def csr_residual(self, values, row_ptr, col_idx, x, b):
        import numpy as np

        values, row_ptr, col_idx, x = _csr_host(values, row_ptr, col_idx, x)
        b = np.ascontiguousarray(b, np.float64)
        out = np.empty_like(x)
        self.check(lib().jg_csr_residual_host(self.h, x.size, values.size, values.ctypes.data,
                                              row_ptr.ctypes.data, col_idx.ctypes.data,
                                              x.ctypes.data, b.ctypes.data, out.ctypes.data),
                   "csr_residual")
        return out

    def mgpcg(self, max_iter, params: MgParams, options: GpuOptions | None = None):
        info = SolveInfo()
        self.check(lib().jg_solver_mgpcg(self.h, max_iter, C.byref(params),
                                         C.byref(options) if options is not None else None,
                                         C.byref(info)), "mgpcg")
        return info

    def mg(self, params: MgParams, options: GpuOptions | None = None):
        """One V-cycle from the resident x (FdmMgSolver3::solve)."""
        info = SolveInfo()
        self.check(lib().jg_solver_mg(self.h, C.byref(params),
                                      C.byref(options) if options is not None else None,
                                      C.byref(info)), "mg")
        return info

    def mg_host(self, system, b, x, params: MgParams, options: GpuOptions | None = None):
        """jg_mg_solve_host: one V-cycle on host arrays (system.A levels as
        (nz, ny, nx, 4) rows); x (finest shape) is the start and is overwritten."""
        import numpy as np

        L = len(system.A)
        rows = [np.ascontiguousarray(A, np.float64) for A in system.A]
        SZ = _sz * L
        nx = SZ(*[A.shape[2] for A in rows])
        ny = SZ(*[A.shape[1] for A in rows])
        nz = SZ(*[A.shape[0] for A in rows])
        ptrs = (_vp * L)(*[A.ctypes.data for A in rows])
        b = np.ascontiguousarray(b, np.float64)
        assert x.flags.c_contiguous and x.dtype == np.float64 and x.shape == b.shape
        info = SolveInfo()
        self.check(lib().jg_mg_solve_host(self.h, L, nx, ny, nz, ptrs, b.ctypes.data, x.ctypes.data,
                                          C.byref(params),
                                          C.byref(options) if options is not None else None,
                                          C.byref(info)), "mg_host")
        self.sizes = [(A.shape[2], A.shape[1], A.shape[0]) for A in rows]
        return info

    def mg_vcycle(self, params: MgParams, options: GpuOptions | None, r, z):
        """z = M r (device vectors of the finest level's size)."""
        self.check(lib().jg_mg_vcycle(self.h, C.byref(params),
                                      C.byref(options) if options is not None else None,
                                      _ptr(r), _ptr(z)), "mg_vcycle")

    def time_kernel(self, kernel, precision, reps=20):
        """(average launch ms, algorithmic bytes per launch) of a level-0 kernel."""
        ms, nbytes = _d(), _d()
        self.check(lib().jg_solver_time_kernel(self.h, kernel, precision, reps, C.byref(ms),
                                               C.byref(nbytes)), "time_kernel")
        return ms.value, nbytes.value

    # ------------------------------------------------ pressure assembly
    def sp_build(self, u, v, w, bsdf, fsdf, h):
        self.check(lib().jg_sp_build(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(bsdf), _ptr(fsdf),
                                     h, h, h), "sp_build")

    def sp_apply_gradient(self, u, v, w, uo, vo, wo, h):
        self.check(lib().jg_sp_apply_gradient(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(uo),
                                              _ptr(vo), _ptr(wo), h, h, h), "sp_apply_gradient")

    def fr_build(self, u, v, w, vertex_sdf, fsdf, h, bu=None, bv=None, bw=None):
        self.check(lib().jg_fr_build(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(vertex_sdf), _ptr(fsdf),
                                     _ptr(bu), _ptr(bv), _ptr(bw), h, h, h), "fr_build")

    def fr_apply_gradient(self, u, v, w, uo, vo, wo, h):
        self.check(lib().jg_fr_apply_gradient(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(uo),
                                              _ptr(vo), _ptr(wo), h, h, h), "fr_apply_gradient")

    def constrain_velocity(self, u, v, w, sdf, h, origin, collider=None, depth=5,
                           closed=JG_DIR_ALL, blocked=False):
        """Velocity boundary conditions in place on device face arrays u, v, w
        (the finest level's size), collider SDF at the cell centres."""
        hh = (_d * 3)(*h)
        oo = (_d * 3)(*origin)
        self.check(lib().jg_constrain_velocity(self.h, 1 if blocked else 0, hh, oo, _ptr(u), _ptr(v),
                                               _ptr(w), _ptr(sdf),
                                               C.byref(collider) if collider is not None else None,
                                               depth, closed), "constrain_velocity")

    def fr_build2(self, u, v, bsdf_ghost, fsdf, h, bu=None, bv=None):
        self.check(lib().jg_fr_build2(self.h, _ptr(u), _ptr(v), _ptr(bsdf_ghost), _ptr(fsdf),
                                      _ptr(bu), _ptr(bv), h[0], h[1]), "fr_build2")

    def fr_apply_gradient2(self, u, v, uo, vo, h):
        self.check(lib().jg_fr_apply_gradient2(self.h, _ptr(u), _ptr(v), _ptr(uo), _ptr(vo), h[0],
                                               h[1]), "fr_apply_gradient2")

    def fr_array(self, level, which):
        """Device pointer of a level's face weights (which 0-2) or fluid SDF (3)."""
        return lib().jg_fr_array(self.h, level, which)

    def generate_scene(self, desc, u=None, v=None, w=None, bsdf=None, fsdf=None, vertex=None):
        self.check(lib().jg_scene_generate_device(self.h, C.byref(desc), _ptr(u), _ptr(v), _ptr(w),
                                                  _ptr(bsdf), _ptr(fsdf), _ptr(vertex)), "scene")

    # ------------------------------------------------- device BLAS entries
    def mvm(self, shape, A, x, y):
        nx, ny, nz = shape
        self.check(lib().jg_blas_mvm(self.h, nx, ny, nz, _ptr(A), _ptr(x), _ptr(y)), "mvm")

    def residual(self, shape, A, x, b, r):
        nx, ny, nz = shape
        self.check(lib().jg_blas_residual(self.h, nx, ny, nz, _ptr(A), _ptr(x), _ptr(b), _ptr(r)),
                   "residual")

    def dot(self, n, a, b):
        out = _d()
        self.check(lib().jg_blas_dot(self.h, n, _ptr(a), _ptr(b), C.byref(out)), "dot")
        return out.value

    def axpy(self, n, a, x, y, out):
        self.check(lib().jg_blas_axpy(self.h, n, a, _ptr(x), _ptr(y), _ptr(out)), "axpy")

    def relax_red_black(self, shape, A, b, sor, x):
        nx, ny, nz = shape
        self.check(lib().jg_relax_red_black(self.h, nx, ny, nz, _ptr(A), _ptr(b), sor, _ptr(x)),
                   "relax_red_black")

    def relax_red_black_noalloc(self, shape, A, b, sor, x):
        nx, ny, nz = shape
        self.check(lib().jg_relax_red_black_noalloc(self.h, nx, ny, nz, _ptr(A), _ptr(b), sor,
                                                    _ptr(x)), "relax_red_black_noalloc")

    def relax_jacobi(self, shape, A, b, x, xt):
        nx, ny, nz = shape
        self.check(lib().jg_relax_jacobi(self.h, nx, ny, nz, _ptr(A), _ptr(b), _ptr(x), _ptr(xt)),
                   "relax_jacobi")

    def restrict(self, cshape, fine, coarse):
        nx, ny, nz = cshape
        self.check(lib().jg_mg_restrict(self.h, nx, ny, nz, _ptr(fine), _ptr(coarse)), "restrict")

    def correct(self, fshape, coarse, fine):
        nx, ny, nz = fshape
        self.check(lib().jg_mg_correct(self.h, nx, ny, nz, _ptr(coarse), _ptr(fine)), "correct")

    def sp_build2(self, u, v, bsdf, fsdf, h):
        """GridSinglePhasePressureSolver2's markers and system on the resident
        2-D hierarchy (device arrays; h = (hx, hy))."""
        self.check(lib().jg_sp_build2(self.h, _ptr(u), _ptr(v), _ptr(bsdf), _ptr(fsdf), h[0], h[1]),
                   "sp_build2")

    def sp_apply_gradient2(self, u, v, uo, vo, h):
        self.check(lib().jg_sp_apply_gradient2(self.h, _ptr(u), _ptr(v), _ptr(uo), _ptr(vo), h[0],
                                               h[1]), "sp_apply_gradient2")

    def restrict2(self, cshape, fine, coarse):
        """FdmMgUtils2::restrict; cshape = (cnx, cny)."""
        cnx, cny = cshape
        self.check(lib().jg_mg_restrict2(self.h, cnx, cny, _ptr(fine), _ptr(coarse)), "restrict2")

    def correct2(self, fshape, coarse, fine):
        """FdmMgUtils2::correct; fshape = (fnx, fny)."""
        fnx, fny = fshape
        self.check(lib().jg_mg_correct2(self.h, fnx, fny, _ptr(coarse), _ptr(fine)), "correct2")


class _DeviceView:
    """A device address seen as a 1-D array (__cuda_array_interface__)."""

    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False),
                                         "version": 2}


def _memcpy_dtoh(a, ptr):
    """Device address ptr -> host numpy array a (same length and dtype)."""
    import torch

    t = torch.as_tensor(_DeviceView(ptr, a.size, a.dtype.str), device="cuda")
    a[...] = t.cpu().numpy()


def _csr_host(values, row_ptr, col_idx, v):
    import numpy as np

    return (np.ascontiguousarray(values, np.float64), np.ascontiguousarray(row_ptr, np.uint64),
            np.ascontiguousarray(col_idx, np.uint64), np.ascontiguousarray(v, np.float64))


def _ptr(t):
    """Device pointer of a torch tensor, an int address, or None."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()
