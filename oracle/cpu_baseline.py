"""CPU baseline of the hot path (test infrastructure only: bench.py's cpu_baseline leg).

Drives oracle/cpu_baseline.cpp, the FP64 C++/OpenMP restatement of
* the element assembly q-loop and scatter of dense_local_main_blocks_2d.hpp:176-361;
* the coupled (slab, vertex) patch solves of dense_local_error_blocks_2d.hpp:1060-1143
  (partial-pivot LU, one solve, the 1e-10 residual acceptance) at the real orders
  N = (p+1) (n_flux + n_scalar) + (p+1) [interior] of the workload's patches;
* MINRES with the Chebyshev-Jacobi block preconditioner of apf_csr_minres_solve,
on the host's cores with a stated thread count.  The reference itself cannot be built
here (it needs Eigen 3.4), so this restatement is the baseline ("port").
"""
from __future__ import annotations

import ctypes
import os
import platform
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_baseline.cpp")
LIB = os.path.join(HERE, "libapf_cpu_baseline.so")  # built in-tree; travels with the tree
_lib = None


def build():
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-o", LIB + ".tmp", SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline build failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.cpu_pattern.restype = ctypes.c_int64
        L.cpu_pattern.argtypes = [ctypes.c_int64, ctypes.c_int, P, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, P,
                                  P, P]
        L.cpu_assemble.restype = ctypes.c_double
        L.cpu_assemble.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P, ctypes.c_int, P, P,
                                   ctypes.c_int64, ctypes.c_int, P, P, P, P, P, ctypes.c_int64, P, ctypes.c_int]
        L.cpu_patch_solves.restype = ctypes.c_double
        L.cpu_patch_solves.argtypes = [ctypes.c_int, ctypes.c_int64, P, ctypes.POINTER(ctypes.c_double)]
        L.cpu_minres.restype = ctypes.c_double
        L.cpu_minres.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, P, P, P, P, P, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def host_info():
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        model = platform.processor()
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": nproc, "cpu_count": os.cpu_count(), "lscpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


class TensorCase:
    """Mesh arrays, maps, tables and pattern of a tensor workload (oracle numbering,
    bit-identical to the product's)."""

    def __init__(self, stm, p):
        from . import mesh as OM
        from .basis import spacetime_lagrange
        from .quadrature import cell_rule

        self.stm, self.p = stm, p
        d = stm.space.dim
        self.dim = d
        od = OM.spacetime_dofs(stm, p)
        self.nX, self.nY = od.nX, od.nY
        ns, m = stm.space.n_cells, stm.n_intervals
        nc = ns * m
        V = stm.space.vertices[stm.space.cells]                   # [ns, d+1, d]
        geo = np.zeros((nc, 8))
        k = np.repeat(np.arange(m), ns)
        s = np.tile(np.arange(ns), m)
        geo[:, :(d + 1) * d] = V[s].reshape(nc, -1)
        geo[:, 6] = stm.times[k]
        geo[:, 7] = stm.times[k + 1]
        self.geo = geo
        self.M = np.broadcast_to(np.eye(d).ravel(), (nc, d * d)).copy()
        self.ymap = np.ascontiguousarray(od.y_map, dtype=np.int32)
        self.xmap = np.ascontiguousarray(od.x_map, dtype=np.int32)
        self.nl = self.ymap.shape[1]
        pts, w = cell_rule(d, p)
        vals, grads = spacetime_lagrange(d, p, pts)
        self.w = np.ascontiguousarray(w)
        self.phi = np.ascontiguousarray(vals)
        self.dphi = np.ascontiguousarray(grads)
        self.bottom = np.arange(ns, dtype=np.int32)
        L = lib()
        n = self.nX + self.nY
        self.rowptr = np.zeros(n + 1, dtype=np.int64)
        nnz = L.cpu_pattern(nc, self.nl, _p(self.ymap), _p(self.xmap), self.nY, self.nX, ns, _p(self.bottom),
                            _p(self.rowptr), None)
        self.col = np.zeros(nnz, dtype=np.int32)
        L.cpu_pattern(nc, self.nl, _p(self.ymap), _p(self.xmap), self.nY, self.nX, ns, _p(self.bottom),
                      _p(self.rowptr), _p(self.col))
        self.val = np.zeros(nnz)

    def assemble(self, threads, reps=1):
        return lib().cpu_assemble(self.dim, threads, len(self.geo), _p(self.geo), _p(self.M), self.nl, _p(self.ymap),
                                  _p(self.xmap), self.nY, len(self.w), _p(self.w), _p(self.phi), _p(self.dphi),
                                  _p(self.rowptr), _p(self.col), len(self.val), _p(self.val), reps)

    def patch_orders(self):
        """coupled order N of every spatial vertex patch (shared by all slabs)"""
        from .estimator import build_patches_spatial

        pp = self.p + 1
        out = []
        for pt in build_patches_spatial(self.stm.space, self.p):
            out.append(pp * (pt.n_flux + pt.n_scalar) + (0 if pt.boundary else pp))
        return np.array(out, dtype=np.int32)


def patch_solves(orders, threads):
    st = ctypes.c_double()
    t = lib().cpu_patch_solves(threads, len(orders), _p(orders), ctypes.byref(st))
    return t, st.value


def minres(rowptr, col, val, b, nY, threads, iters, degree=8, check_interval=None):
    """at most `iters` iterations (true residual every check_interval, default: at the end)"""
    x = np.zeros(len(b))
    it = ctypes.c_int()
    rel = ctypes.c_double()
    t = lib().cpu_minres(threads, len(b), nY, _p(rowptr), _p(col), _p(val), _p(b), _p(x), 1e-10, iters,
                         check_interval or iters, degree, ctypes.byref(it), ctypes.byref(rel))
    return t, it.value, rel.value
