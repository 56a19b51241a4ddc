"""MI355X-native hot path of DoubleAdapParabolicFEM: assembly, equilibrated-flux
estimator and MINRES for the space-time heat equation.

This package is a thin ctypes mirror of the C ABI in ``include/apf.h``
(library ``libapf.so``, HIP kernels for gfx950).  The public objects follow the
reference's call chain (algorithm/adaptive_driver.hpp:317-336 ->
compute_adaptive_test_space_iteration, src/finite_element/time_slabs/
time_slab_adaptive_loop.hpp:727):

* :class:`Space`   -- X/Y spaces, DOF numbering, K pattern, vertex patches
                      (host integer work, no GPU needed);
* :class:`Context` -- one HIP device + stream;
* :class:`System`  -- device-resident K, b, solution and estimator buffers:
  ``assemble`` (assemble_main_system_blocks), ``minres``
  (solve_minres_with_parabolic_graph_norm), ``estimate`` (slab
  reconstruction + solve_all_patches + flux indicators).

There is no CPU fallback: if ``libapf.so`` is missing the import of the
library raises, and device calls raise when no HIP device is present.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libapf.so")

APF_OK, APF_BAD_ARG, APF_NOT_CONVERGED, APF_RESIDUAL_TOO_LARGE, APF_SINGULAR, APF_HIP_ERROR, APF_NO_DEVICE = range(7)
APF_FP64, APF_FP32_REFINE = 0, 1
APF_ASSEMBLE_KRON, APF_ASSEMBLE_SCATTER = 0, 1
APF_PRECOND_DIAGONAL, APF_PRECOND_MULTIGRID = 0, 1
APF_SLAB_FROM_SOLUTION, APF_SLAB_SOLVE = 0, 1
APF_SLAB_STOP_NONE, APF_SLAB_STOP_TOL, APF_SLAB_STOP_FLOOR, APF_SLAB_STOP_MAX_IT = -1, 0, 1, 2
APF_CSR_FULL, APF_CSR_TRIANGLE = 0, 1
APF_NOALLOC_MAX_N = 262144  # include/apf.h

_lib = None


class APFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[apf status {code}] {msg}")
        self.code = code


class NotConverged(APFError):
    """Mirror of la::concepts::PreconditionedSolveFailure
    (src/linear_algebra/preconditioners/solve_failure.hpp:7-25)."""

    def __init__(self, code, msg, stats=None):
        super().__init__(code, msg)
        self.stats = stats


class _MeshDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("p", ctypes.c_int), ("n_vertices", ctypes.c_int),
                ("vertices", ctypes.POINTER(ctypes.c_double)), ("n_cells", ctypes.c_int),
                ("cells", ctypes.POINTER(ctypes.c_int)), ("n_times", ctypes.c_int),
                ("times", ctypes.POINTER(ctypes.c_double))]


class _SpaceInfo(ctypes.Structure):
    _fields_ = [("nY", ctypes.c_int64), ("nX", ctypes.c_int64), ("n", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("dim", ctypes.c_int), ("p", ctypes.c_int), ("Ns", ctypes.c_int), ("Nt", ctypes.c_int),
                ("nf", ctypes.c_int), ("n_nodes", ctypes.c_int), ("n_space_cells", ctypes.c_int),
                ("n_intervals", ctypes.c_int), ("n_vertices", ctypes.c_int), ("nnz_spatial", ctypes.c_int),
                ("n_patches", ctypes.c_int), ("max_patch_cells", ctypes.c_int), ("max_patch_flux", ctypes.c_int),
                ("max_patch_scalar", ctypes.c_int), ("n_bottom_points", ctypes.c_int)]


class _Coeff(ctypes.Structure):
    _fields_ = [("M", ctypes.POINTER(ctypes.c_double)), ("ell", ctypes.c_double),
                ("u0_bottom", ctypes.POINTER(ctypes.c_double))]


class _MinresStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("true_relative_residual", ctypes.c_double), ("estimate", ctypes.c_double),
                ("seconds", ctypes.c_double)]


class _EstimateOut(ctypes.Structure):
    _fields_ = [("lambda_y2", ctypes.c_double), ("trace2", ctypes.c_double), ("eta2", ctypes.c_double),
                ("flux2", ctypes.c_double), ("recon2", ctypes.c_double), ("y_est2", ctypes.c_double),
                ("div_residual_l2", ctypes.c_double), ("max_patch_residual", ctypes.c_double),
                ("seconds_patches", ctypes.c_double), ("seconds_diagnostics", ctypes.c_double),
                ("extra_refinements", ctypes.c_int), ("slab_iterations", ctypes.c_int),
                ("slab_relative_residual", ctypes.c_double),
                ("seconds_slab", ctypes.c_double), ("slab_stop", ctypes.c_int), ("fullpiv_patches", ctypes.c_int),
                ("seconds_total", ctypes.c_double)]


class _GEstimateOut(ctypes.Structure):
    _fields_ = [("lambda_difference_squared", ctypes.c_double), ("lambda_difference", ctypes.c_double),
                ("raw", ctypes.c_double), ("coefficient_difference_squared", ctypes.c_double),
                ("solver_relative_residual", ctypes.c_double), ("solver_iterations", ctypes.c_int),
                ("g_true_dofs", ctypes.c_int64), ("seconds", ctypes.c_double), ("solver_stop", ctypes.c_int)]


class _CellSpaceDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("p", ctypes.c_int), ("n_vertices", ctypes.c_int64),
                ("vertices", ctypes.c_void_p), ("n_cells", ctypes.c_int64), ("cells", ctypes.c_void_p),
                ("times", ctypes.c_void_p), ("restr_ptr", ctypes.c_void_p), ("restr_dof", ctypes.c_void_p),
                ("restr_weight", ctypes.c_void_p), ("n_dofs", ctypes.c_int64)]


class _GeneralCoeff(ctypes.Structure):
    _fields_ = [("grid_n", ctypes.c_int), ("M_grid", ctypes.c_void_p), ("ell", ctypes.c_double),
                ("u0_bottom", ctypes.c_void_p)]


class _DoerflerDiag(ctypes.Structure):
    _fields_ = [("total", ctypes.c_double), ("target", ctypes.c_double), ("cutoff_value", ctypes.c_double),
                ("cutoff_margin", ctypes.c_double), ("max_indicator", ctypes.c_double),
                ("cutoff_index", ctypes.c_int64), ("near_tie_count", ctypes.c_int64)]


EXPORTED = {
    "apf_space_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(_MeshDesc), ctypes.c_int]),
    "apf_space_destroy": (None, [ctypes.c_void_p]),
    "apf_space_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_SpaceInfo)]),
    "apf_space_bottom_points": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_space_cell_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_space_free_coords": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_space_pattern": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "apf_space_patch_sizes": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "apf_space_add_coarse_mesh": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_MeshDesc)]),
    "apf_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "apf_ctx_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "apf_ctx_destroy": (None, [ctypes.c_void_p]),
    "apf_system_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "apf_system_destroy": (None, [ctypes.c_void_p]),
    "apf_assemble_main": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Coeff), ctypes.c_int]),
    "apf_system_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "apf_system_download_pattern": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "apf_system_set_solution": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_minres_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(_MinresStats)]),
    "apf_minres_solve_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.POINTER(_MinresStats)]),
    "apf_system_mg_setup_seconds": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "apf_system_use_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "apf_bench_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "apf_bench_store": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "apf_bench_spmv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "apf_bench_kron_apply": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "apf_bench_mg_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]),
    "apf_system_matvec": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "apf_estimate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_EstimateOut),
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "apf_estimate_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(_EstimateOut), ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "apf_g_estimate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                      ctypes.POINTER(_GEstimateOut)]),
    "apf_system_timings": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_estimate_phase_timings": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_combine_estimator": (ctypes.c_int, [ctypes.c_int64] + [ctypes.c_void_p] * 4),
    "apf_doerfler_mark": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(_DoerflerDiag)]),
    "apf_y_estimator_threshold": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                 ctypes.POINTER(ctypes.c_double)]),
    "apf_csr_matrix_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "apf_csr_matrix_destroy": (None, [ctypes.c_void_p]),
    "apf_csr_matrix_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64)]),
    "apf_csr_minres_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(_MinresStats)]),
    "apf_csr_spmv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "apf_csr_block_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_double)]),
    "apf_csr_block_pcg_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.POINTER(_MinresStats)]),
    "apf_csr_minres_solve_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    ctypes.POINTER(_MinresStats)]),
    "apf_general_pattern": (ctypes.c_int, [ctypes.POINTER(_CellSpaceDesc), ctypes.POINTER(_CellSpaceDesc),
                                           ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "apf_general_system_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_CellSpaceDesc),
                                                 ctypes.POINTER(_CellSpaceDesc), ctypes.c_void_p,
                                                 ctypes.POINTER(ctypes.c_void_p)]),
    "apf_general_system_destroy": (None, [ctypes.c_void_p]),
    "apf_general_system_info": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int64)] * 4
                                + [ctypes.POINTER(ctypes.c_int)]),
    "apf_general_bottom_points": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_general_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_GeneralCoeff),
                                            ctypes.POINTER(ctypes.c_double)]),
    "apf_general_download": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "apf_general_matrix": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "apf_general_estimator_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.POINTER(ctypes.c_void_p)]),
    "apf_general_estimator_destroy": (None, [ctypes.c_void_p]),
    "apf_general_estimator_create_sliced": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                           ctypes.c_void_p, ctypes.c_void_p,
                                                           ctypes.POINTER(ctypes.c_void_p)]),
    "apf_general_estimator_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(_Coeff)]),
    "apf_general_estimate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(_EstimateOut), ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "apf_device_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]),
    "apf_device_free": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_device_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "apf_device_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "apf_mesh1d_create": (ctypes.c_int, [ctypes.c_double] * 4 + [ctypes.POINTER(ctypes.c_void_p)]),
    "apf_mesh1d_destroy": (None, [ctypes.c_void_p]),
    "apf_mesh1d_create_slab": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                                              ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)]),
    "apf_mesh1d_info": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int64)] * 3),
    "apf_mesh1d_vertices": (ctypes.c_int, [ctypes.c_void_p] * 3),
    "apf_mesh1d_cells": (ctypes.c_int, [ctypes.c_void_p] * 6),
    "apf_fespace1d_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]),
    "apf_fespace1d_destroy": (None, [ctypes.c_void_p]),
    "apf_fespace1d_refine": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "apf_fespace1d_info": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int64)] * 4),
    "apf_fespace1d_active_cells": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_fespace1d_cell_dofs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "apf_fespace1d_constraints": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "apf_fespace1d_restriction": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "apf_fespace1d_ancestors": (ctypes.c_int, [ctypes.c_void_p] * 3),
    "apf_fespace1d_neighbours": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                ctypes.c_void_p]),
    "apf_general_x_indicators": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(_GeneralCoeff),
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "apf_general_graph_norm": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_GeneralCoeff),
                                              ctypes.POINTER(ctypes.c_void_p)]),
    "apf_csr_matrix_download": (ctypes.c_int, [ctypes.c_void_p] * 4),
}


def lib():
    """Load libapf.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; build it with `python -m doubleadapparabolicfem_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTED.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(code, ctx=None, what=""):
    if code != APF_OK:
        msg = lib().apf_last_error(ctx).decode(errors="replace")
        raise APFError(code, f"{what}: {msg}" if what else msg)


@dataclass
class Mesh:
    """Tensor space-time mesh: spatial vertices [nv, dim], cells [ns, dim+1], times [m+1]."""
    dim: int
    vertices: np.ndarray
    cells: np.ndarray
    times: np.ndarray


def _mesh_desc(mesh: Mesh, p: int):
    v = np.ascontiguousarray(mesh.vertices, dtype=np.float64).reshape(-1)
    c = np.ascontiguousarray(mesh.cells, dtype=np.int32).reshape(-1)
    t = np.ascontiguousarray(mesh.times if mesh.times is not None else [0.0, 1.0], dtype=np.float64)
    md = _MeshDesc(mesh.dim, p, len(v) // mesh.dim, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                   len(c) // (mesh.dim + 1), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                   len(t), t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return md, (v, c, t)


class Space:
    """X/Y spaces of degree p on a tensor mesh, with the CSR pattern of K.

    ``coarse``: optional nested coarser spatial meshes (fine to coarse) for the
    multigrid MINRES preconditioner."""

    def __init__(self, mesh: Mesh, p: int, pattern: bool = True, coarse=()):
        L = lib()
        self.mesh = mesh
        md, self._keep = _mesh_desc(mesh, p)
        h = ctypes.c_void_p()
        _check(L.apf_space_create(ctypes.byref(h), ctypes.byref(md), 1 if pattern else 0), None, "apf_space_create")
        self.handle = h
        self.n_coarse = 0
        for cm in coarse:
            cmd, keep = _mesh_desc(cm, p)
            _check(L.apf_space_add_coarse_mesh(h, ctypes.byref(cmd)), None, "apf_space_add_coarse_mesh")
            self.n_coarse += 1
        info = _SpaceInfo()
        _check(L.apf_space_get_info(h, ctypes.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in _SpaceInfo._fields_}
        self.p = p
        self.dim = mesh.dim

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.apf_space_destroy(self.handle)
            self.handle = None

    def __getattr__(self, k):
        info = self.__dict__.get("info")
        if info is not None and k in info:
            return info[k]
        raise AttributeError(k)

    def bottom_points(self):
        nb = self.info["n_bottom_points"]
        out = np.zeros((self.info["n_space_cells"], nb, self.dim))
        _check(lib().apf_space_bottom_points(self.handle, _ptr(out)))
        return out

    def cell_free(self):
        out = np.zeros((self.info["n_space_cells"], self.info["Ns"]), dtype=np.int32)
        _check(lib().apf_space_cell_free(self.handle, _ptr(out)))
        return out

    def free_coords(self):
        out = np.zeros((self.info["nf"], self.dim))
        _check(lib().apf_space_free_coords(self.handle, _ptr(out)))
        return out

    def pattern(self):
        rowptr = np.zeros(self.info["n"] + 1, dtype=np.int64)
        col = np.zeros(self.info["nnz"], dtype=np.int32)
        _check(lib().apf_space_pattern(self.handle, _ptr(rowptr), _ptr(col)))
        return rowptr, col

    def patch_sizes(self):
        n = self.info["n_patches"]
        nf, ns_, nc, bd = (np.zeros(n, dtype=np.int32) for _ in range(4))
        _check(lib().apf_space_patch_sizes(self.handle, _ptr(nf), _ptr(ns_), _ptr(nc), _ptr(bd)))
        return nf, ns_, nc, bd


class Context:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().apf_ctx_create(ctypes.byref(h), device), None, "apf_ctx_create")
        self.handle = h

    def bench_stream(self, nbytes=1 << 30, reps=10):
        """Copy-kernel bandwidth in GB/s (read + write)."""
        v = ctypes.c_double()
        _check(lib().apf_bench_stream(self.handle, int(nbytes), reps, ctypes.byref(v)), self.handle)
        return v.value

    def bench_store(self, nbytes=1 << 30, reps=10):
        """Store-only kernel bandwidth in GB/s (bytes written)."""
        v = ctypes.c_double()
        _check(lib().apf_bench_store(self.handle, int(nbytes), reps, ctypes.byref(v)), self.handle)
        return v.value

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.apf_ctx_destroy(self.handle)
            self.handle = None


@dataclass
class Coefficients:
    """M per spatial cell ([ns, dim, dim] or None = identity), constant ell, u0 callable on [n, dim]."""
    M: np.ndarray | None
    ell: float
    u0: object


class System:
    def __init__(self, ctx: Context, space: Space):
        self.ctx, self.space = ctx, space
        h = ctypes.c_void_p()
        _check(lib().apf_system_create(ctx.handle, space.handle, ctypes.byref(h)), ctx.handle, "apf_system_create")
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.apf_system_destroy(self.handle)
            self.handle = None

    def assemble(self, coeff: Coefficients, mode: int = APF_ASSEMBLE_KRON):
        sp = self.space
        bp = sp.bottom_points()
        u0b = np.ascontiguousarray(coeff.u0(bp.reshape(-1, sp.dim)).reshape(bp.shape[:2]), dtype=np.float64)
        M = None if coeff.M is None else np.ascontiguousarray(coeff.M, dtype=np.float64).reshape(-1)
        c = _Coeff(M.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if M is not None else None, float(coeff.ell),
                   u0b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        self._keep = (M, u0b)
        _check(lib().apf_assemble_main(self.handle, ctypes.byref(c), mode), self.ctx.handle, "apf_assemble_main")

    def download(self, values=True, rhs=True):
        v = np.zeros(self.space.nnz) if values else None
        b = np.zeros(self.space.n) if rhs else None
        _check(lib().apf_system_download(self.handle, _ptr(v), _ptr(b)), self.ctx.handle)
        return v, b

    def download_pattern(self):
        """K's CSR pattern as held on the device (generated there): (rowptr, col)."""
        rowptr = np.zeros(self.space.n + 1, dtype=np.int64)
        col = np.zeros(self.space.nnz, dtype=np.int32)
        _check(lib().apf_system_download_pattern(self.handle, _ptr(rowptr), _ptr(col)), self.ctx.handle)
        return rowptr, col

    def set_solution(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().apf_system_set_solution(self.handle, _ptr(x)), self.ctx.handle)

    def minres(self, tol=1e-10, max_it=20000, check_interval=100, precond=APF_FP64, x0=None, raise_on_fail=True,
               kind=None, download=True):
        """Preconditioned MINRES; kind None -> multigrid when coarse meshes were given.  With
        download=False the solution stays on the device (returned x is None)."""
        st = _MinresStats()
        x = np.zeros(self.space.n) if download else None
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        if kind is None:
            kind = APF_PRECOND_MULTIGRID if self.space.n_coarse else APF_PRECOND_DIAGONAL
        code = lib().apf_minres_solve_ex(self.handle, tol, max_it, check_interval, kind, precond, _ptr(x0), _ptr(x),
                                         ctypes.byref(st))
        stats = {f: getattr(st, f) for f, _ in _MinresStats._fields_}
        if code == APF_NOT_CONVERGED and raise_on_fail:
            raise NotConverged(code, lib().apf_last_error(self.ctx.handle).decode(), stats)
        if code not in (APF_OK, APF_NOT_CONVERGED):
            _check(code, self.ctx.handle, "apf_minres_solve")
        return x, stats

    def estimate(self, prec=APF_FP64, refine_steps=2, per_cell=True, slab=APF_SLAB_FROM_SOLUTION):
        """Equilibrated-flux estimator on the current solution; slab=APF_SLAB_SOLVE
        solves the slab reconstruction instead of taking lambda."""
        out = _EstimateOut()
        n = self.space.n_space_cells * self.space.n_intervals
        fl = np.zeros(n) if per_cell else None
        rs = np.zeros(n) if per_cell else None
        rc = np.zeros(n) if per_cell else None
        _check(lib().apf_estimate_ex(self.handle, prec, refine_steps, slab, ctypes.byref(out), _ptr(fl), _ptr(rs),
                                     _ptr(rc)), self.ctx.handle, "apf_estimate")
        res = {f: getattr(out, f) for f, _ in _EstimateOut._fields_}
        res.update(flux_by_cell=fl, resid_by_cell=rs, recon_by_cell=rc)
        return res

    def g_estimate(self, gsys, tol=1e-12, max_it=100):
        """G-estimator of this system's solution (algorithm/adaptive_driver.hpp:2680-2905).
        gsys: a degree p+1 system on the same mesh, assembled with the same
        coefficients, with coarse meshes for its multigrid; its solution vector is
        overwritten."""
        out = _GEstimateOut()
        _check(lib().apf_g_estimate(self.handle, gsys.handle, tol, max_it, ctypes.byref(out)), self.ctx.handle,
               "apf_g_estimate")
        return {f: getattr(out, f) for f, _ in _GEstimateOut._fields_}

    def use_graph(self, on=True):
        _check(lib().apf_system_use_graph(self.handle, 1 if on else 0), self.ctx.handle)

    def bench_spmv(self, reps=20):
        v = ctypes.c_double()
        _check(lib().apf_bench_spmv(self.handle, reps, ctypes.byref(v)), self.ctx.handle)
        return v.value

    def bench_kron_apply(self, reps=20):
        v = ctypes.c_double()
        _check(lib().apf_bench_kron_apply(self.handle, reps, ctypes.byref(v)), self.ctx.handle)
        return v.value

    def bench_mg_step(self, reps=20, prec=APF_FP64):
        """(ms per launch, algorithmic bytes per launch) of the fine-level Chebyshev step"""
        v, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().apf_bench_mg_step(self.handle, reps, prec, ctypes.byref(v), ctypes.byref(b)), self.ctx.handle)
        return v.value, b.value

    def matvec(self, x, matrix_free=False):
        """y = K x: the assembled CSR K, or the same entries from K's Kronecker factors"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _check(lib().apf_system_matvec(self.handle, _ptr(x), _ptr(y), 1 if matrix_free else 0), self.ctx.handle,
               "apf_system_matvec")
        return y

    def mg_setup_seconds(self):
        v = ctypes.c_double()
        _check(lib().apf_system_mg_setup_seconds(self.handle, ctypes.byref(v)))
        return v.value

    def timings(self):
        t = np.zeros(4)
        _check(lib().apf_system_timings(self.handle, _ptr(t)))
        return dict(assembly_ms=t[0], patch_ms=t[1], diagnostics_ms=t[2], spmv_ms=t[3])

    def phase_timings(self):
        """Device milliseconds of the last estimate's patch phases."""
        t = np.zeros(3)
        _check(lib().apf_estimate_phase_timings(self.handle, _ptr(t)))
        return dict(factor_ms=t[0], rhs_ms=t[1], solve_ms=t[2])


class CsrMatrix:
    """A symmetric sparse matrix on the device, from the compressed arrays of a caller's
    Eigen::SparseMatrix<double, ColMajor, int> (outer [n+1], inner, values: int32, int32,
    float64) or from a scipy.sparse matrix.  storage=APF_CSR_TRIANGLE reads only the
    entries with inner >= outer and completes them symmetrically."""

    def __init__(self, ctx: Context, A=None, outer=None, inner=None, values=None, storage=APF_CSR_FULL):
        if A is not None:
            A = A.tocsr()
            A.sort_indices()
            outer, inner, values = A.indptr, A.indices, A.data
        outer = np.ascontiguousarray(outer, dtype=np.int32)
        inner = np.ascontiguousarray(inner, dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().apf_csr_matrix_create(ctx.handle, len(outer) - 1, _ptr(outer), _ptr(inner), _ptr(values),
                                           storage, ctypes.byref(h)), ctx.handle, "apf_csr_matrix_create")
        self.handle = h
        n, nnz = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().apf_csr_matrix_info(h, ctypes.byref(n), ctypes.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.apf_csr_matrix_destroy(self.handle)
            self.handle = None

    def download(self):
        """(rowptr int64 [n+1], col int32, values) of the device matrix"""
        rowptr = np.zeros(self.n + 1, np.int64)
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz)
        _check(lib().apf_csr_matrix_download(self.handle, _ptr(rowptr), _ptr(col), _ptr(val)), self.ctx.handle,
               "apf_csr_matrix_download")
        return rowptr, col, val

    def minres(self, b, n_lambda, hhat=None, tol=1e-10, max_it=20000, check_interval=100, degree=8,
               precond=APF_FP32_REFINE, x0=None, raise_on_fail=True):
        """MINRES on this matrix (apf_csr_minres_solve); returns (x, stats)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        st = _MinresStats()
        code = lib().apf_csr_minres_solve(self.ctx.handle, self.handle, n_lambda,
                                          hhat.handle if hhat is not None else None, _ptr(b), _ptr(x), _ptr(x0),
                                          tol, max_it, check_interval, degree, precond, ctypes.byref(st))
        stats = {f: getattr(st, f) for f, _ in _MinresStats._fields_}
        if code == APF_NOT_CONVERGED and raise_on_fail:
            raise NotConverged(code, lib().apf_last_error(self.ctx.handle).decode(), stats)
        if code not in (APF_OK, APF_NOT_CONVERGED):
            _check(code, self.ctx.handle, "apf_csr_minres_solve")
        return x, stats


    def spmv(self, x):
        """A x (apf_csr_spmv)"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.n,):
            raise ValueError("spmv: x must have n entries")
        y = np.zeros(self.n)
        _check(lib().apf_csr_spmv(self.ctx.handle, self.handle, _ptr(x), _ptr(y)), self.ctx.handle, "apf_csr_spmv")
        return y

    def block_energy(self, d, n_block):
        """d^T A[:n_block, :n_block] d (apf_csr_block_energy)"""
        d = np.ascontiguousarray(d, dtype=np.float64)
        if d.shape != (n_block,):
            raise ValueError("block_energy: d must have n_block entries")
        out = ctypes.c_double()
        _check(lib().apf_csr_block_energy(self.ctx.handle, self.handle, n_block, _ptr(d), ctypes.byref(out)),
               self.ctx.handle, "apf_csr_block_energy")
        return out.value

    def block_pcg(self, b, n_block, tol=1e-12, max_it=20000, check_interval=20, degree=8, x0=None,
                  raise_on_fail=True):
        """Conjugate gradients on the SPD leading block (apf_csr_block_pcg_solve); returns (x, stats)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (n_block,):
            raise ValueError("block_pcg: b must have n_block entries")
        x = np.zeros(n_block)
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        st = _MinresStats()
        code = lib().apf_csr_block_pcg_solve(self.ctx.handle, self.handle, n_block, _ptr(b), _ptr(x), _ptr(x0), tol,
                                             max_it, check_interval, degree, ctypes.byref(st))
        stats = {f: getattr(st, f) for f, _ in _MinresStats._fields_}
        if code == APF_NOT_CONVERGED and raise_on_fail:
            raise NotConverged(code, lib().apf_last_error(self.ctx.handle).decode(), stats)
        if code not in (APF_OK, APF_NOT_CONVERGED):
            _check(code, self.ctx.handle, "apf_csr_block_pcg_solve")
        return x, stats


class DeviceArray:
    """A float64 array in device memory (apf_device_alloc), for the device vectors of
    apf_csr_minres_solve_noalloc."""

    def __init__(self, ctx: Context, n: int):
        self.ctx, self.n = ctx, int(n)
        h = ctypes.c_void_p()
        _check(lib().apf_device_alloc(ctx.handle, 8 * self.n, ctypes.byref(h)), ctx.handle, "apf_device_alloc")
        self.ptr = h

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        _check(lib().apf_device_upload(self.ctx.handle, self.ptr, _ptr(a), 8 * self.n), self.ctx.handle)

    def download(self):
        out = np.empty(self.n)
        _check(lib().apf_device_download(self.ctx.handle, _ptr(out), self.ptr, 8 * self.n), self.ctx.handle)
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.apf_device_free(self.ctx.handle, self.ptr)
            self.ptr = None


def csr_minres_noalloc(K: "CsrMatrix", d_b: DeviceArray, d_x: DeviceArray, n_lambda, tol=1e-10, max_it=20000,
                       check_interval=100, degree=8, precond=APF_FP32_REFINE, use_x0=False, raise_on_fail=True):
    """apf_csr_minres_solve_noalloc on device vectors; returns the stats (x stays in d_x)."""
    st = _MinresStats()
    code = lib().apf_csr_minres_solve_noalloc(K.ctx.handle, K.handle, n_lambda, d_b.ptr, d_x.ptr,
                                              1 if use_x0 else 0, tol, max_it, check_interval, degree, precond,
                                              ctypes.byref(st))
    stats = {f: getattr(st, f) for f, _ in _MinresStats._fields_}
    if code == APF_NOT_CONVERGED and raise_on_fail:
        raise NotConverged(code, lib().apf_last_error(K.ctx.handle).decode(), stats)
    if code not in (APF_OK, APF_NOT_CONVERGED):
        _check(code, K.ctx.handle, "apf_csr_minres_solve_noalloc")
    return stats


def _cell_space_desc(cs):
    keep = [np.ascontiguousarray(cs.vertices, dtype=np.float64), np.ascontiguousarray(cs.cells, dtype=np.int32),
            np.ascontiguousarray(cs.times, dtype=np.float64), np.ascontiguousarray(cs.restr_ptr, dtype=np.int64),
            np.ascontiguousarray(cs.restr_dof, dtype=np.int32), np.ascontiguousarray(cs.restr_w, dtype=np.float64)]
    d = _CellSpaceDesc(cs.dim, cs.p, len(keep[0]), _ptr(keep[0]), len(keep[1]), _ptr(keep[1]), _ptr(keep[2]),
                       _ptr(keep[3]), _ptr(keep[4]), _ptr(keep[5]), cs.n_dofs)
    return d, keep


def general_pattern(mesh):
    """Host-side CSR pattern of K for a non-tensor mesh (layered.LayeredMesh or any object
    with X, Y cell spaces and y_ancestor); no GPU needed."""
    dx, kx = _cell_space_desc(mesh.X)
    dy, ky = _cell_space_desc(mesh.Y)
    anc = np.ascontiguousarray(mesh.y_ancestor, dtype=np.int32)
    nnz = ctypes.c_int64()
    _check(lib().apf_general_pattern(ctypes.byref(dx), ctypes.byref(dy), _ptr(anc), ctypes.byref(nnz), None, None),
           None, "apf_general_pattern")
    n = mesh.X.n_dofs + mesh.Y.n_dofs
    rowptr = np.zeros(n + 1, dtype=np.int64)
    col = np.zeros(nnz.value, dtype=np.int32)
    _check(lib().apf_general_pattern(ctypes.byref(dx), ctypes.byref(dy), _ptr(anc), ctypes.byref(nnz), _ptr(rowptr),
                                     _ptr(col)), None, "apf_general_pattern")
    return rowptr, col


class GeneralSystem:
    """Saddle-point system of a space-time mesh that is not a tensor product: cells with
    restriction rows (hanging nodes) and Y cells inside X ancestors (apf_general_*)."""

    def __init__(self, ctx: Context, mesh):
        self.ctx, self.mesh = ctx, mesh
        dx, kx = _cell_space_desc(mesh.X)
        dy, ky = _cell_space_desc(mesh.Y)
        anc = np.ascontiguousarray(mesh.y_ancestor, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().apf_general_system_create(ctx.handle, ctypes.byref(dx), ctypes.byref(dy), _ptr(anc),
                                               ctypes.byref(h)), ctx.handle, "apf_general_system_create")
        self.handle = h
        n, nl, nnz, nb = (ctypes.c_int64() for _ in range(4))
        nq = ctypes.c_int()
        _check(lib().apf_general_system_info(h, ctypes.byref(n), ctypes.byref(nl), ctypes.byref(nnz), ctypes.byref(nb),
                                             ctypes.byref(nq)))
        self.n, self.nY, self.nnz, self.n_bottom, self.n_bottom_points = n.value, nl.value, nnz.value, nb.value, nq.value

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.apf_general_system_destroy(self.handle)
            self.handle = None

    def bottom_points(self):
        out = np.zeros((self.n_bottom, self.n_bottom_points, self.mesh.dim))
        _check(lib().apf_general_bottom_points(self.handle, _ptr(out)))
        return out

    def assemble(self, u0, M_grid=None, grid_n=0, ell=0.0):
        """M_grid: [grid_n^dim, dim, dim] tensors of a piecewise-constant grid (None: identity).
        Returns the device time of the assembly in ms."""
        pts = self.bottom_points()
        u0b = np.ascontiguousarray(u0(pts.reshape(-1, self.mesh.dim)).reshape(self.n_bottom, -1), dtype=np.float64)
        Mg = None if M_grid is None else np.ascontiguousarray(M_grid, dtype=np.float64).ravel()
        c = _GeneralCoeff(grid_n if Mg is not None else 0, _ptr(Mg), ell, _ptr(u0b))
        ms = ctypes.c_double()
        _check(lib().apf_general_assemble(self.handle, ctypes.byref(c), ctypes.byref(ms)), self.ctx.handle,
               "apf_general_assemble")
        return ms.value

    def download(self):
        rowptr = np.zeros(self.n + 1, dtype=np.int64)
        col = np.zeros(self.nnz, dtype=np.int32)
        vals = np.zeros(self.nnz)
        b = np.zeros(self.n)
        _check(lib().apf_general_download(self.handle, _ptr(rowptr), _ptr(col), _ptr(vals), _ptr(b)), self.ctx.handle)
        return rowptr, col, vals, b

    def rhs(self):
        b = np.zeros(self.n)
        _check(lib().apf_general_download(self.handle, None, None, None, _ptr(b)), self.ctx.handle)
        return b

    def x_indicators(self, x, u0, M_grid=None, grid_n=0):
        """(lambda_Y^2 per Y cell, initial-trace mismatch^2 per X cell) of x = [lambda; u]
        (apf_general_x_indicators)."""
        pts = self.bottom_points()
        u0b = np.ascontiguousarray(u0(pts.reshape(-1, self.mesh.dim)).reshape(self.n_bottom, -1), dtype=np.float64)
        Mg = None if M_grid is None else np.ascontiguousarray(M_grid, dtype=np.float64).ravel()
        c = _GeneralCoeff(grid_n if Mg is not None else 0, _ptr(Mg), 0.0, _ptr(u0b))
        xx = np.ascontiguousarray(x, dtype=np.float64)
        if xx.shape != (self.n,):
            raise ValueError(f"x must have length {self.n}")
        lam2 = np.zeros(len(self.mesh.y_ancestor))
        tr2 = np.zeros(self.mesh.X.n_cells)
        _check(lib().apf_general_x_indicators(self.handle, _ptr(xx), ctypes.byref(c), _ptr(lam2), _ptr(tr2)),
               self.ctx.handle, "apf_general_x_indicators")
        return lam2, tr2

    def graph_norm(self, M_grid=None, grid_n=0):
        """Hhat = -C + B_dt diag(A_y)^-1 B_dt^T + A_x as a CsrMatrix (apf_general_graph_norm),
        after assemble()"""
        Mg = None if M_grid is None else np.ascontiguousarray(M_grid, dtype=np.float64).ravel()
        c = _GeneralCoeff(grid_n if Mg is not None else 0, _ptr(Mg), 0.0, None)
        h = ctypes.c_void_p()
        _check(lib().apf_general_graph_norm(self.handle, ctypes.byref(c), ctypes.byref(h)), self.ctx.handle,
               "apf_general_graph_norm")
        M = CsrMatrix.__new__(CsrMatrix)
        M.ctx, M.handle = self.ctx, h
        n, nnz = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().apf_csr_matrix_info(h, ctypes.byref(n), ctypes.byref(nnz)))
        M.n, M.nnz = n.value, nnz.value
        return M

    def matrix(self):
        """a CsrMatrix view of K (valid while this system lives)"""
        h = ctypes.c_void_p()
        _check(lib().apf_general_matrix(self.handle, ctypes.byref(h)), self.ctx.handle)
        M = CsrMatrix.__new__(CsrMatrix)
        M.ctx, M.handle, M.n, M.nnz = self.ctx, h, self.n, self.nnz
        M._owner = self
        return M


class GeneralEstimator:
    """Estimator of a GeneralSystem's solution, slab group by slab group
    (apf_general_estimator_*).  groups: [(tensor Mesh of the group, coarse meshes for the
    slab solve, first Y cell)]."""

    def __init__(self, gs: GeneralSystem, groups):
        """groups: [(tensor Mesh, coarse meshes, first Y cell)], or, for Y cells sliced into
        several slabs (apf_general_estimator_create_sliced), [(tensor Mesh, coarse meshes,
        source Y cell of each group cell, slab-major)]"""
        self.gs = gs
        self.spaces = [Space(m, gs.mesh.p, coarse=coarse) for m, coarse, _ in groups]
        arr = (ctypes.c_void_p * len(groups))(*[sp.handle.value for sp in self.spaces])
        h = ctypes.c_void_p()
        if np.isscalar(groups[0][2]):
            first = np.ascontiguousarray([f for _, _, f in groups], dtype=np.int64)
            _check(lib().apf_general_estimator_create(gs.handle, len(groups), arr, _ptr(first), ctypes.byref(h)),
                   gs.ctx.handle, "apf_general_estimator_create")
        else:
            src = [np.asarray(s, dtype=np.int32) for _, _, s in groups]
            ptr = np.ascontiguousarray(np.concatenate([[0], np.cumsum([len(s) for s in src])]), dtype=np.int64)
            cat = np.ascontiguousarray(np.concatenate(src), dtype=np.int32)
            _check(lib().apf_general_estimator_create_sliced(gs.handle, len(groups), arr, _ptr(ptr), _ptr(cat),
                                                             ctypes.byref(h)),
                   gs.ctx.handle, "apf_general_estimator_create_sliced")
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.apf_general_estimator_destroy(self.handle)
            self.handle = None

    def assemble(self, u0, M_grid=None, grid_n=0, ell=0.0):
        """each group's tensor system: M per cell from the grid (cells lie inside grid cells),
        u0 at the bottom points of the group that starts at t = 0"""
        dim = self.gs.mesh.dim
        self._keep = []
        for g, sp in enumerate(self.spaces):
            m = sp.mesh
            if M_grid is None:
                M = None
            else:
                cen = m.vertices[m.cells].mean(axis=1)
                ix = np.minimum((cen[:, 0] * grid_n).astype(np.int64), grid_n - 1)
                idx = ix if dim == 1 else np.minimum((cen[:, 1] * grid_n).astype(np.int64), grid_n - 1) * grid_n + ix
                M = np.ascontiguousarray(np.asarray(M_grid).reshape(-1, dim * dim)[idx], dtype=np.float64)
            pts = sp.bottom_points()
            u0b = (u0(pts.reshape(-1, dim)).reshape(pts.shape[0], pts.shape[1]) if m.times[0] == 0.0
                   else np.zeros(pts.shape[:2]))
            u0b = np.ascontiguousarray(u0b, dtype=np.float64)
            self._keep += [M, u0b]
            dp = ctypes.POINTER(ctypes.c_double)
            c = _Coeff(M.ctypes.data_as(dp) if M is not None else None, float(ell), u0b.ctypes.data_as(dp))
            _check(lib().apf_general_estimator_assemble(self.handle, g, ctypes.byref(c)), self.gs.ctx.handle,
                   "apf_general_estimator_assemble")

    def estimate(self, x, prec=APF_FP32_REFINE, refine_steps=2, per_cell=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = _EstimateOut()
        n = self.gs.mesh.Y.n_cells
        fl = np.zeros(n) if per_cell else None
        rs = np.zeros(n) if per_cell else None
        rc = np.zeros(n) if per_cell else None
        _check(lib().apf_general_estimate(self.handle, _ptr(x), prec, refine_steps, ctypes.byref(out), _ptr(fl),
                                          _ptr(rs), _ptr(rc)), self.gs.ctx.handle, "apf_general_estimate")
        res = {f: getattr(out, f) for f, _ in _EstimateOut._fields_}
        res.update(flux_by_cell=fl, resid_by_cell=rs, recon_by_cell=rc)
        return res


def combine_estimator(flux2, recon2, resid2=None):
    """estimator^2 = flux^2 + reconstruction^2 per cell (host C++;
    time_slab_estimator_combination.hpp:88-250).  Raises APFError on a negative
    or non-finite squared quantity."""
    f = np.ascontiguousarray(flux2, dtype=np.float64)
    r = np.ascontiguousarray(recon2, dtype=np.float64)
    d = None if resid2 is None else np.ascontiguousarray(resid2, dtype=np.float64)
    if r.shape != f.shape or (d is not None and d.shape != f.shape):
        raise ValueError("combine_estimator: shape mismatch")
    out = np.empty_like(f)
    _check(lib().apf_combine_estimator(f.size, _ptr(f), _ptr(r), _ptr(d), _ptr(out)))
    return out


def doerfler_mark(values, theta, ids=None, near_tie_tolerance=0.0):
    """Deterministic Doerfler marking (host C++; time_slab_cellwise_errors.hpp:156-280).
    Returns (marked ids in marking order, diagnostics dict)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    i = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
    if i is not None and i.shape != v.shape:
        raise ValueError("doerfler_mark: ids and values differ in length")
    marked = np.empty(max(v.size, 1), dtype=np.int64)
    nm = ctypes.c_int64()
    dg = _DoerflerDiag()
    _check(lib().apf_doerfler_mark(v.size, _ptr(i), _ptr(v), float(theta), float(near_tie_tolerance), _ptr(marked),
                                   ctypes.byref(nm), ctypes.byref(dg)))
    diag = {name: getattr(dg, name) for name, _ in _DoerflerDiag._fields_}
    return marked[:nm.value].copy(), diag


def y_estimator_threshold(rho, lambda_y2, trace2):
    """rho^2 (lambda_Y^2 + (1 + 1/(rho+1/2)^2) trace^2) (indicator_utilities.hpp:475-492)."""
    out = ctypes.c_double()
    _check(lib().apf_y_estimator_threshold(float(rho), float(lambda_y2), float(trace2), ctypes.byref(out)))
    return out.value
