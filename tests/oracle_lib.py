"""ctypes loader of the oracle (oracle/_build/liboracle.so): the CPU restatement
of FasterCap's solve path used as the checker.  Only tests, smoke() and the
bench's cpu_baseline leg import this module."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from fastercap_amd._ffi import JobDesc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# FASTERCAP_ORACLE_LIB names another build (the sanitizer build of tests/sanitize.sh)
LIB = os.environ.get("FASTERCAP_ORACLE_LIB") or os.path.join(ORACLE_DIR, "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        jd = C.POINTER(JobDesc)
        L.orc_potential_opt_3d.argtypes = [dp, fp]
        L.orc_potential_num_3d.argtypes = [dp, fp, C.c_int]
        L.orc_enfield_num_3d.argtypes = [dp, fp, dp, C.c_int]
        L.orc_potential_opt_2d.argtypes = [dp, fp]
        L.orc_enfield_2d.argtypes = [dp, fp, dp]
        for f in ("orc_potential_opt_3d", "orc_potential_num_3d", "orc_enfield_num_3d", "orc_potential_opt_2d",
                  "orc_enfield_2d", "orc_pot_estimate", "orc_self"):
            getattr(L, f).restype = C.c_double
        L.orc_auto_3d.argtypes = [fp]
        L.orc_mutual_half_3d.argtypes = [fp, fp, C.c_int]
        L.orc_mutual_dhalf_3d.argtypes = [fp, fp, dp, C.c_double, C.c_int]
        L.orc_mutual_full_3d.argtypes = [fp, fp, C.c_int]
        L.orc_mutual_dfull_3d.argtypes = [fp, fp, dp, C.c_int]
        for f in ("orc_auto_3d", "orc_mutual_half_3d", "orc_mutual_dhalf_3d", "orc_mutual_full_3d",
                  "orc_mutual_dfull_3d"):
            getattr(L, f).restype = C.c_double
        L.orc_pot_estimate.argtypes = [jd, C.c_int, C.c_int]
        L.orc_self.argtypes = [jd, C.c_int]
        L.orc_fill.argtypes = [jd, dp, dp, C.c_int]
        L.orc_mvp.argtypes = [jd, dp, dp, dp, dp, C.c_int]
        L.orc_solve.argtypes = [jd, dp, dp, C.c_double, C.c_int, dp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_int), C.c_int]
        L.orc_solve_rows.argtypes = [jd, dp, dp, C.c_double, C.c_int, C.c_int, dp, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
        L.orc_compressed_dense.argtypes = [jd, dp, dp, dp]
        L.orc_leaf_dense.argtypes = [jd, dp, C.c_int]
        L.orc_link_coverage.argtypes = [jd]
        L.orc_link_coverage.restype = C.c_longlong
        ip = C.POINTER(C.c_int)
        llp = C.POINTER(C.c_longlong)
        L.orc_block_partition.argtypes = [jd, C.c_int, ip, ip]
        L.orc_block_precond.argtypes = [jd, C.c_int, ip, ip, llp, dp]
        L.orc_block_apply.argtypes = [C.c_int, ip, llp, dp, dp, dp]
        L.orc_block_apply.restype = None
        L.orc_solve_block.argtypes = [jd, dp, dp, C.c_double, C.c_int, dp, ip, ip, ip, C.c_int]
        L.orc_node_self.argtypes = [jd, C.c_int]
        L.orc_node_self.restype = C.c_double
        L.orc_super_partition.argtypes = [jd, C.c_int, ip, ip, ip]
        L.orc_super_precond.argtypes = [jd, C.c_int, ip, ip, ip, dp]
        L.orc_super_apply.argtypes = [jd, C.c_int, ip, ip, dp, dp, dp]
        L.orc_super_apply.restype = None
        L.orc_solve_super.argtypes = [jd, dp, dp, C.c_double, C.c_int, dp, ip, ip, ip, C.c_int]
        L.orc_self_cplx.argtypes = [jd, C.c_int, dp, dp]
        L.orc_self_cplx.restype = None
        L.orc_mvp_cplx.argtypes = [jd, dp, dp, dp, dp, dp, C.c_int]
        L.orc_solve_cplx.argtypes = [jd, dp, dp, dp, C.c_double, C.c_int, dp, ip, ip, ip, C.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def potential_opt_3d(r, v9):
    r = np.ascontiguousarray(r, dtype=np.float64)
    v = np.ascontiguousarray(v9, dtype=np.float32)
    return lib().orc_potential_opt_3d(_d(r), _f(v))


def potential_num_3d(r, v9, rule):
    r = np.ascontiguousarray(r, dtype=np.float64)
    v = np.ascontiguousarray(v9, dtype=np.float32)
    return lib().orc_potential_num_3d(_d(r), _f(v), rule)


def enfield_num_3d(r, v9, n, rule):
    r = np.ascontiguousarray(r, dtype=np.float64)
    v = np.ascontiguousarray(v9, dtype=np.float32)
    n = np.ascontiguousarray(n, dtype=np.float64)
    return lib().orc_enfield_num_3d(_d(r), _f(v), _d(n), rule)


def potential_opt_2d(r, v6):
    r = np.ascontiguousarray(r, dtype=np.float64)
    v = np.ascontiguousarray(v6, dtype=np.float32)
    return lib().orc_potential_opt_2d(_d(r), _f(v))


def enfield_2d(r, v6, n):
    r = np.ascontiguousarray(r, dtype=np.float64)
    v = np.ascontiguousarray(v6, dtype=np.float32)
    n = np.ascontiguousarray(n, dtype=np.float64)
    return lib().orc_enfield_2d(_d(r), _f(v), _d(n))


def _v(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def auto_3d(v9):
    v = _v(v9)
    return lib().orc_auto_3d(_f(v))


def mutual_half_3d(v1, v2, rule):
    a, b = _v(v1), _v(v2)
    return lib().orc_mutual_half_3d(_f(a), _f(b), rule)


def mutual_dhalf_3d(v1, v2, n, h, rule):
    a, b = _v(v1), _v(v2)
    n = np.ascontiguousarray(n, dtype=np.float64)
    return lib().orc_mutual_dhalf_3d(_f(a), _f(b), _d(n), h, rule)


def mutual_full_3d(v1, v2, order):
    a, b = _v(v1), _v(v2)
    return lib().orc_mutual_full_3d(_f(a), _f(b), order)


def mutual_dfull_3d(v1, v2, n, rule):
    a, b = _v(v1), _v(v2)
    n = np.ascontiguousarray(n, dtype=np.float64)
    return lib().orc_mutual_dfull_3d(_f(a), _f(b), _d(n), rule)


def fill(mesh, nthreads=0):
    coef = np.empty(mesh.L, dtype=np.float64)
    selfc = np.empty(mesh.N, dtype=np.float64)
    lib().orc_fill(C.byref(mesh.desc), _d(coef), _d(selfc), nthreads)
    return coef, selfc


def mvp(mesh, coef, selfc, q, nthreads=0):
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.empty_like(q)
    lib().orc_mvp(C.byref(mesh.desc), _d(coef), _d(selfc), _d(q), _d(v), nthreads)
    return v


def solve(mesh, coef, selfc, tol=1e-8, precond=1, nthreads=0):
    ncond = int((mesh.group_is_diel == 0).sum())
    cap = np.zeros(ncond * ncond, dtype=np.float64)
    iters = np.zeros(ncond, dtype=np.int32)
    rows, cols = C.c_int(), C.c_int()
    lib().orc_solve(C.byref(mesh.desc), _d(coef), _d(selfc), tol, precond, _d(cap), C.byref(rows), C.byref(cols),
                    iters.ctypes.data_as(C.POINTER(C.c_int)), nthreads)
    r, c = rows.value, cols.value
    return cap[: r * c].reshape(r, c), iters[:r]


def solve_rows(mesh, coef, selfc, max_rows, tol=1e-8, precond=1, nthreads=0):
    ncond = int((mesh.group_is_diel == 0).sum())
    cap = np.zeros(ncond * ncond, dtype=np.float64)
    iters = np.zeros(ncond, dtype=np.int32)
    rows, cols = C.c_int(), C.c_int()
    lib().orc_solve_rows(C.byref(mesh.desc), _d(coef), _d(selfc), tol, precond, max_rows, _d(cap), C.byref(rows),
                         C.byref(cols), iters.ctypes.data_as(C.POINTER(C.c_int)), nthreads)
    return cap.reshape(ncond, ncond)[:max_rows], iters[:max_rows]


def compressed_dense(mesh, coef, selfc):
    P = np.empty((mesh.N, mesh.N), dtype=np.float64)
    lib().orc_compressed_dense(C.byref(mesh.desc), _d(coef), _d(selfc), _d(P))
    return P


def leaf_dense(mesh, nthreads=0):
    P = np.empty((mesh.N, mesh.N), dtype=np.float64)
    lib().orc_leaf_dense(C.byref(mesh.desc), _d(P), nthreads)
    return P


def link_coverage_errors(mesh):
    return int(lib().orc_link_coverage(C.byref(mesh.desc)))


def pot_estimate(mesh, i, j):
    return lib().orc_pot_estimate(C.byref(mesh.desc), i, j)


class BlockPrecond:
    """Block preconditioner of one mesh: blocks[b] = leaf range [starts[b],
    starts[b+1]), inverse P[b] (n x n, row-major)."""

    def __init__(self, mesh, bsize):
        n = mesh.N
        self.starts = np.zeros(n + 1, dtype=np.int32)
        self.poff = np.zeros(n + 1, dtype=np.int64)
        self.P = np.zeros(n * max(2, min(128, bsize)), dtype=np.float64)
        nb = C.c_int()
        lib().orc_block_precond(C.byref(mesh.desc), bsize, self.starts.ctypes.data_as(C.POINTER(C.c_int)),
                                C.byref(nb), self.poff.ctypes.data_as(C.POINTER(C.c_longlong)), _d(self.P))
        self.nblocks = nb.value
        self.starts = self.starts[: self.nblocks + 1].copy()
        self.poff = self.poff[: self.nblocks + 1].copy()

    def block(self, b):
        n = int(self.starts[b + 1] - self.starts[b])
        return self.P[self.poff[b]: self.poff[b] + n * n].reshape(n, n)

    def apply(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty_like(q)
        lib().orc_block_apply(self.nblocks, self.starts.ctypes.data_as(C.POINTER(C.c_int)),
                              self.poff.ctypes.data_as(C.POINTER(C.c_longlong)), _d(self.P), _d(q), _d(out))
        return out


def block_partition(mesh, bsize):
    starts = np.zeros(mesh.N + 1, dtype=np.int32)
    nb = C.c_int()
    lib().orc_block_partition(C.byref(mesh.desc), bsize, starts.ctypes.data_as(C.POINTER(C.c_int)), C.byref(nb))
    return starts[: nb.value + 1].copy()


def solve_block(mesh, coef, selfc, bsize, tol=1e-8, nthreads=0):
    ncond = int((mesh.group_is_diel == 0).sum())
    cap = np.zeros(ncond * ncond, dtype=np.float64)
    iters = np.zeros(ncond, dtype=np.int32)
    rows, cols = C.c_int(), C.c_int()
    lib().orc_solve_block(C.byref(mesh.desc), _d(coef), _d(selfc), tol, bsize, _d(cap), C.byref(rows),
                          C.byref(cols), iters.ctypes.data_as(C.POINTER(C.c_int)), nthreads)
    r, c = rows.value, cols.value
    return cap[: r * c].reshape(r, c), iters[:r]


class SuperPrecond:
    """Two-level preconditioner of one mesh: m elements (tree nodes) with leaf
    ranges [starts[s], starts[s+1]) and the inverse Sinv of their matrix;
    available is False when the mesh has at least dim tree roots."""

    def __init__(self, mesh, dim):
        ip = C.POINTER(C.c_int)
        self.mesh = mesh
        self.nodes = np.zeros(max(dim, 1), dtype=np.int32)
        self.starts = np.zeros(max(dim, 1) + 1, dtype=np.int32)
        Sinv = np.zeros(dim * dim, dtype=np.float64)
        m = C.c_int()
        rc = lib().orc_super_precond(C.byref(mesh.desc), dim, self.nodes.ctypes.data_as(ip),
                                     self.starts.ctypes.data_as(ip), C.byref(m), _d(Sinv))
        self.available = rc == 0
        self.m = m.value
        self.nodes = self.nodes[: self.m].copy()
        self.starts = self.starts[: self.m + 1].copy()
        self.Sinv = Sinv[: self.m * self.m].reshape(self.m, self.m).copy()

    def apply(self, q):
        ip = C.POINTER(C.c_int)
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty_like(q)
        S = np.ascontiguousarray(self.Sinv)
        lib().orc_super_apply(C.byref(self.mesh.desc), self.m, self.nodes.ctypes.data_as(ip),
                              self.starts.ctypes.data_as(ip), _d(S), _d(q), _d(out))
        return out


def node_self(mesh, node):
    return lib().orc_node_self(C.byref(mesh.desc), node)


def solve_super(mesh, coef, selfc, dim, tol=1e-8, nthreads=0):
    ncond = int((mesh.group_is_diel == 0).sum())
    cap = np.zeros(ncond * ncond, dtype=np.float64)
    iters = np.zeros(ncond, dtype=np.int32)
    rows, cols = C.c_int(), C.c_int()
    lib().orc_solve_super(C.byref(mesh.desc), _d(coef), _d(selfc), tol, dim, _d(cap), C.byref(rows),
                          C.byref(cols), iters.ctypes.data_as(C.POINTER(C.c_int)), nthreads)
    r, c = rows.value, cols.value
    return cap[: r * c].reshape(r, c), iters[:r]



# ------------------------------------------------------------------ complex permittivities
def self_cplx(mesh):
    """(re [N], im [N]) self terms of a complex job."""
    re = np.empty(mesh.N)
    im = np.empty(mesh.N)
    a, b = C.c_double(), C.c_double()
    for i in range(mesh.N):
        lib().orc_self_cplx(C.byref(mesh.desc), i, C.byref(a), C.byref(b))
        re[i], im[i] = a.value, b.value
    return re, im


def fill_cplx(mesh, nthreads=0):
    """(coef [L], self_re [N], self_im [N]) of a complex job."""
    coef, _ = fill(mesh, nthreads)
    sre, sim = self_cplx(mesh)
    return coef, sre, sim


def mvp_cplx(mesh, coef, sre, sim, q, nthreads=0):
    """[Re | Im] product of a complex job, q of length 2N."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.empty_like(q)
    lib().orc_mvp_cplx(C.byref(mesh.desc), _d(coef), _d(sre), _d(sim), _d(q), _d(v), nthreads)
    return v


def solve_cplx(mesh, coef, sre, sim, tol=1e-8, precond=1, nthreads=0):
    """Complex C matrix (rows x n_cond) and iteration counts of a complex job."""
    ncond = int((mesh.group_is_diel == 0).sum())
    cap = np.zeros(2 * ncond * ncond, dtype=np.float64)
    iters = np.zeros(ncond, dtype=np.int32)
    rows, cols = C.c_int(), C.c_int()
    lib().orc_solve_cplx(C.byref(mesh.desc), _d(coef), _d(sre), _d(sim), tol, precond, _d(cap), C.byref(rows),
                         C.byref(cols), iters.ctypes.data_as(C.POINTER(C.c_int)), nthreads)
    r, c = rows.value, cols.value
    return cap[: r * c].reshape(r, c) + 1j * cap[r * c: 2 * r * c].reshape(r, c), iters[:r]
