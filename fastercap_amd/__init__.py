"""MI355X (gfx950) FasterCap BEM solve path.

Host side (mesh builder, include/fastercap_host.h) + device C-ABI
(include/fastercap_amd.h), wrapped with ctypes.  The hot path (link fill,
hierarchical product, GMRES) only runs on the GPU: `Device` raises if the HIP
library is missing or no GPU is visible; there is no CPU fallback.

Python mirror of the reference interfaces:
  Mesh.build(path, ...)       CAutoRefine ReadFastCapFile/BuildSuperHierarchy/
                              AutoRefinePanels/links (Solver/Autorefine.cpp)
  Device.upload/fill/mvp      AutoRefineLinks fill, CMultHier::MultiplyMatByVec_fast
  Device.solve                CSolveCap::Solve (Solver/SolveCapacitance.cpp:1138-1425)
  Device.solve_batch, Engine  several independent solves advanced together
  run(argv)                   FasterCapConsole::main + CSolveCap::Run (manual mode)
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _ffi
from ._ffi import (
    FCG_MVP_FP32,
    FCG_MVP_FP32_ACC64,
    FCG_MVP_FP64,
    FCG_PREC_FP64,
    FCG_PREC_MIXED,
    FCG_PRECOND_BLOCK,
    FCG_PRECOND_JACOBI,
    FCG_PRECOND_NONE,
    FCG_PRECOND_SUPER,
    FCG_SCHEME_COLLOCATION,
    FCG_SCHEME_GALERKIN,
    HostOptions,
    HostStats,
    JobDesc,
    JobMem,
    LinkParams,
    NodeGeom,
    SolveOpts,
    SolveStats,
)

__all__ = [
    "Mesh",
    "Device",
    "Engine",
    "SolveResult",
    "BatchSolve",
    "FasterCapError",
    "run",
    "FCG_MVP_FP64",
    "FCG_MVP_FP32",
    "FCG_MVP_FP32_ACC64",
    "FCG_PREC_FP64",
    "FCG_PREC_MIXED",
    "FCG_PRECOND_BLOCK",
    "FCG_PRECOND_JACOBI",
    "FCG_PRECOND_NONE",
    "FCG_PRECOND_SUPER",
    "FCG_SCHEME_COLLOCATION",
    "FCG_SCHEME_GALERKIN",
    "TWO_PI_TIMES_E0",
]

TWO_PI_TIMES_E0 = 5.563250279878e-11  # Solver/Potential.h:42


class FasterCapError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"FC error {code}: {msg}")
        self.code = code


def _np(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


class Mesh:
    """Refined geometry + link structure produced by the host builder."""

    def __init__(self, handle: int):
        self._h = handle
        self.lib = _ffi.host_lib()
        self.desc = JobDesc()
        rc = self.lib.fch_get_desc(C.c_void_p(handle), C.byref(self.desc))
        if rc != 0:
            raise FasterCapError(rc, "fch_get_desc failed")
        st = HostStats()
        self.lib.fch_get_stats(C.c_void_p(handle), C.byref(st))
        self.stats = st
        d = self.desc
        self.dim = d.dim
        self.N = d.n_leaves
        self.M = d.n_nodes
        self.L = d.n_links
        self.G = d.n_groups
        self.row_ptr = _np(d.row_ptr, self.M + 1, np.int64)
        self.col = _np(d.col, self.L, np.int32)
        self.left = _np(d.left, self.M, np.int32)
        self.right = _np(d.right, self.M, np.int32)
        self.parent = _np(d.parent, self.M, np.int32)
        self.leaf_lo = _np(d.leaf_lo, self.M, np.int32)
        self.leaf_hi = _np(d.leaf_hi, self.M, np.int32)
        self.leaf_node = _np(d.leaf_node, self.N, np.int32)
        self.group_is_diel = _np(d.group_is_diel, self.G, np.int32)
        self.n_cond = int((self.group_is_diel == 0).sum())  # conductor groups: the C matrix order
        self.group_leaf_start = _np(d.group_leaf_start, self.G + 1, np.int32)
        self.group_outperm = _np(d.group_outperm, self.G, np.float64)
        self.group_inperm = _np(d.group_inperm, self.G, np.float64)
        self.leaf_outperm = _np(d.leaf_outperm, self.N, np.float64)
        # complex permittivities (the job solves the real form of the complex system)
        self.complex = bool(d.group_outperm_im)
        if self.complex:
            self.group_outperm_im = _np(d.group_outperm_im, self.G, np.float64)
            self.group_inperm_im = _np(d.group_inperm_im, self.G, np.float64)
            self.leaf_outperm_im = _np(d.leaf_outperm_im, self.N, np.float64)
        raw = np.ctypeslib.as_array(C.cast(d.nodes, C.POINTER(C.c_uint8)), shape=(self.M * C.sizeof(NodeGeom),))
        self.nodes_raw = raw
        self.group_names = [self.lib.fch_group_name(C.c_void_p(handle), g).decode() for g in range(self.G)]

    @classmethod
    def build(cls, path: str, mesh_eps: float = 0.1, eps_ratio: float = 1.0, curv_coeff: float = 3.0,
              device: "Device | None" = None, refine_on_device: bool = False) -> "Mesh":
        """Refine the geometry of `path` and discover its links: on the host, or
        with `device` the links on that GPU (fcg_links_discover) and, with
        refine_on_device, the panel refinement too (fcg_refine).  The mesh is the
        same either way."""
        lib = _ffi.host_lib()
        o = HostOptions()
        lib.fch_default_options(C.byref(o))
        o.mesh_eps, o.eps_ratio, o.curv_coeff = mesh_eps, eps_ratio, curv_coeff
        o.links_on_device = 1 if device is not None else 0
        o.refine_on_device = 1 if (device is not None and refine_on_device) else 0
        h = C.c_void_p()
        rc = lib.fch_build(str(path).encode(), C.byref(o), C.byref(h))
        if rc != 0:
            raise FasterCapError(rc, lib.fch_last_error().decode())
        if device is not None:
            try:
                if o.refine_on_device:
                    device.refine(h.value)
                device.discover_links(h.value)
            except BaseException:
                lib.fch_free(h)
                raise
        return cls(h.value)

    def node_geom(self) -> np.ndarray:
        dt = np.dtype(
            [
                ("v", np.float32, 9),
                ("c", np.float32, 3),
                ("flags", np.uint32),
                ("cond", np.int32),
                ("maxside", np.float64),
                ("area", np.float64),
                ("dn", np.float64, 3),
            ]
        )
        assert dt.itemsize == 96
        return self.nodes_raw.view(dt)

    @property
    def scheme(self) -> int:
        """Discretization scheme of the job this mesh uploads as (fcg_job_desc.scheme):
        FCG_SCHEME_COLLOCATION (also 0) or FCG_SCHEME_GALERKIN (3D; -g)."""
        return int(self.desc.scheme)

    @scheme.setter
    def scheme(self, value: int):
        self.desc.scheme = int(value)

    @property
    def conductor_names(self):
        return [n for n, d in zip(self.group_names, self.group_is_diel) if not d]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self.lib.fch_free(C.c_void_p(h))
            self._h = None


@dataclass(slots=True)
class SolveResult:
    # (slots and no default factories: a batch of a thousand results is built in
    # about a millisecond)
    cap: np.ndarray  # rows x n_cond, unscaled (2D: multiply by 2*pi*e0 for F/m)
    iterations: list | None = None
    residual: float = 0.0
    t_solve_s: float = 0.0
    t_mvp_s: float = 0.0
    n_mvp: int = 0
    charges: np.ndarray | None = None
    converged: bool = True  # every rhs met the tolerance (mixed: the FP64 residual check)
    cycles: int = 0  # inner GMRES cycles (mixed: refinement cycles)
    # with want_residuals: per rhs, |r0| then |g_i| / |b| after each Arnoldi step (-r)
    residuals: list | None = None
    t_precond_s: float = 0.0
    mem: dict | None = None  # device bytes of the solve (gmres, vectors, product)
    t_fill_s: float = 0.0  # set by callers that time the fill (command line -i)
    job_mem: dict | None = None  # fcg_job_memory of the job (command line -i)


class BatchSolve:
    """Results of a batch of solves as arrays: caps holds every job's C matrix
    (rows x cols, row-major; a complex job's imaginary parts follow its real
    ones) job after job, rows and cols their shapes, cplx the complex flags and
    stats the fcg_solve_stats records (a structured numpy array: iterations,
    final_residual, converged, cycles, t_solve_s, n_mvp, ...)."""

    __slots__ = ("caps", "rows", "cols", "cplx", "stats")

    def __init__(self, caps, rows, cols, cplx, stats):
        self.caps, self.rows, self.cols, self.cplx, self.stats = caps, rows, cols, cplx, stats

    def __len__(self):
        return len(self.rows)

    def results(self) -> list:
        """One SolveResult per job (the stats through plain lists: per-record
        ctypes or numpy access costs milliseconds for a thousand jobs)."""
        a = self.stats
        rl, cl = self.rows.tolist(), self.cols.tolist()
        rmax = max(max(rl, default=1), 1)
        its = a["iterations"][:, : min(rmax, 64)].tolist()
        resid, conv, cyc = a["final_residual"].tolist(), a["converged"].tolist(), a["cycles"].tolist()
        tsol, nmvp = a["t_solve_s"].tolist(), a["n_mvp"].tolist()
        caps = self.caps
        out, off = [], 0
        for cx, r, c, it, re, cv, cy, ts, nm in zip(self.cplx, rl, cl, its, resid, conv, cyc, tsol, nmvp):
            n = r * c
            capv = caps[off : off + n].reshape(r, c)
            if cx:
                capv = capv + 1j * caps[off + n : off + 2 * n].reshape(r, c)
                off += 2 * n
            else:
                off += n
            # positional: cap, iterations, residual, t_solve_s, t_mvp_s, n_mvp, charges, converged, cycles
            out.append(SolveResult(capv, it[: min(r, 64)], re, ts, 0.0, nm, None, bool(cv), cy))
        return out


_STATS_DTYPE = None


def _stats_dtype():
    """numpy dtype of fcg_solve_stats (built once)."""
    global _STATS_DTYPE
    if _STATS_DTYPE is None:
        _STATS_DTYPE = np.ctypeslib.as_array(SolveStats()).dtype
    return _STATS_DTYPE


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Device:
    """One fcg context on a HIP device."""

    def __init__(self, device: int = 0):
        self.lib = _ffi.device_lib()
        h = C.c_void_p()
        rc = self.lib.fcg_ctx_create(device, C.byref(h))
        if rc != 0:
            raise FasterCapError(rc, f"cannot create a context on HIP device {device}")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.fcg_last_error(self._h)
            raise FasterCapError(rc, f"{what}: {msg.decode() if msg else ''}")

    def refine(self, mesh_handle: int) -> int:
        """Refine the coarse tree of a mesh built with refine_on_device on this
        device and finish its build; returns the node count."""
        hl = _ffi.host_lib()
        mh = C.c_void_p(mesh_handle)
        nodes = C.POINTER(_ffi.RefineNode)()
        tops = C.POINTER(C.c_int32)()
        n = C.c_int()
        p = _ffi.RefineParams()
        if hl.fch_get_coarse_tree(mh, C.byref(nodes), C.byref(n), C.byref(tops), C.byref(p)) != 0:
            raise FasterCapError(65, "mesh is not waiting for a refinement")
        t0 = time.perf_counter()
        n_out, lev = C.c_int(), C.c_int32()
        eps = C.c_double()
        self._check(self.lib.fcg_refine(self._h, nodes, n.value, tops, C.byref(p), C.byref(n_out), C.byref(eps),
                                        C.byref(lev)), "fcg_refine")
        out = (_ffi.RefineNode * n_out.value)()
        self._check(self.lib.fcg_refine_fetch(self._h, out), "fcg_refine_fetch")
        rc = hl.fch_set_refined_tree(mh, out, n_out.value, eps.value, lev.value, time.perf_counter() - t0)
        if rc != 0:
            raise FasterCapError(rc, "fch_set_refined_tree: " + hl.fch_last_error().decode())
        return n_out.value

    def discover_links(self, mesh_handle: int) -> int:
        """Discover the links of a mesh built with links_on_device on this device
        and install them in the mesh; returns the link count."""
        hl = _ffi.host_lib()
        mh = C.c_void_p(mesh_handle)
        d = JobDesc()
        p = LinkParams()
        if hl.fch_get_desc(mh, C.byref(d)) != 0 or hl.fch_get_link_params(mh, C.byref(p)) != 0:
            raise FasterCapError(65, "bad mesh handle")
        t0 = time.perf_counter()
        n = C.c_int64()
        lev = C.c_int32()
        self._check(self.lib.fcg_links_discover(self._h, C.byref(d), C.byref(p), C.byref(n), C.byref(lev)),
                    "fcg_links_discover")
        row = np.empty(d.n_nodes + 1, dtype=np.int64)
        col = np.empty(max(1, n.value), dtype=np.int32)
        self._check(self.lib.fcg_links_fetch(self._h, row.ctypes.data_as(C.POINTER(C.c_int64)),
                                             col.ctypes.data_as(C.POINTER(C.c_int32))), "fcg_links_fetch")
        rc = hl.fch_set_links(mh, row.ctypes.data_as(C.POINTER(C.c_int64)), col.ctypes.data_as(C.POINTER(C.c_int32)),
                              n.value, lev.value, time.perf_counter() - t0)
        if rc != 0:
            raise FasterCapError(rc, "fch_set_links rejected the discovered links")
        return n.value

    def upload(self, mesh: Mesh) -> int:
        j = C.c_int()
        self._check(self.lib.fcg_job_upload(self._h, C.byref(mesh.desc), C.byref(j)), "fcg_job_upload")
        return j.value

    def release(self, job: int):
        self._check(self.lib.fcg_job_release(self._h, job), "fcg_job_release")

    def fill(self, job: int, precision: int = FCG_PREC_FP64):
        self._check(self.lib.fcg_fill(self._h, job, precision), "fcg_fill")

    def fill_batch(self, jobs, precision: int = FCG_PREC_FP64):
        jobs = list(jobs)
        ja = (C.c_int * len(jobs))(*jobs)
        self._check(self.lib.fcg_fill_batch(self._h, ja, len(jobs), precision), "fcg_fill_batch")

    def coeffs(self, job: int, mesh: Mesh):
        coef = np.empty(mesh.L, dtype=np.float64)
        self_c = np.empty(mesh.N, dtype=np.float64)
        self._check(self.lib.fcg_download_coeffs(self._h, job, _dptr(coef), _dptr(self_c)), "fcg_download_coeffs")
        return coef, self_c

    def mvp(self, job: int, q: np.ndarray) -> np.ndarray:
        q = np.ascontiguousarray(q, dtype=np.float64)
        k = 1 if q.ndim == 1 else q.shape[0]
        v = np.empty_like(q)
        self._check(self.lib.fcg_mvp(self._h, job, _dptr(q), _dptr(v), k), "fcg_mvp")
        return v

    def mvp_prec(self, job: int, q: np.ndarray, mvp_prec: int) -> np.ndarray:
        """Product in one of the three precisions (FCG_MVP_*); the mixed ones need
        a mixed fill of the job."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        k = 1 if q.ndim == 1 else q.shape[0]
        v = np.empty_like(q)
        self._check(self.lib.fcg_mvp_prec(self._h, job, mvp_prec, _dptr(q), _dptr(v), k), "fcg_mvp_prec")
        return v

    def precond(self, job: int, q: np.ndarray) -> np.ndarray:
        q = np.ascontiguousarray(q, dtype=np.float64)
        k = 1 if q.ndim == 1 else q.shape[0]
        v = np.empty_like(q)
        self._check(self.lib.fcg_precond_apply(self._h, job, _dptr(q), _dptr(v), k), "fcg_precond_apply")
        return v

    def block_precond(self, job: int, q: np.ndarray, block_size: int = 32) -> np.ndarray:
        """B q with the job's block preconditioner (blocks of at most block_size leaves)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        k = 1 if q.ndim == 1 else q.shape[0]
        v = np.empty_like(q)
        self._check(self.lib.fcg_block_precond_apply(self._h, job, block_size, _dptr(q), _dptr(v), k),
                    "fcg_block_precond_apply")
        return v

    def super_precond(self, job: int, q: np.ndarray, dim: int = 128) -> np.ndarray:
        """S q with the job's two-level preconditioner of dimension dim; raises
        FcgError when it is unavailable (at least dim conductor and dielectric
        roots)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        k = 1 if q.ndim == 1 else q.shape[0]
        v = np.empty_like(q)
        self._check(self.lib.fcg_super_precond_apply(self._h, job, dim, _dptr(q), _dptr(v), k),
                    "fcg_super_precond_apply")
        return v

    def solve(
        self,
        job: int,
        mesh: Mesh,
        tol: float = 0.01,
        max_iter: int = 1000,
        precond: int = FCG_PRECOND_JACOBI,
        precision: int = FCG_PREC_FP64,
        want_charges: bool = False,
        max_refine: int = 8,
        precond_size: int = 0,
        want_residuals: bool = False,
    ) -> SolveResult:
        ncond = mesh.n_cond
        rows_max = ncond
        nf = 2 if mesh.complex else 1  # complex jobs: real then imaginary part
        cap = np.zeros(nf * rows_max * ncond, dtype=np.float64)
        charges = np.zeros(nf * rows_max * mesh.N, dtype=np.float64) if want_charges else None
        o = SolveOpts(tol=tol, max_iter=max_iter, precond=precond, precision=precision, max_refine=max_refine,
                      precond_size=precond_size)
        hist = None
        if want_residuals:
            hcap = max_iter * (max_refine + 1) + 1
            hist = np.zeros(rows_max * hcap, dtype=np.float64)
            o.residuals = _dptr(hist)
            o.residuals_cap = hcap
        st = SolveStats()
        rows, cols = C.c_int(), C.c_int()
        self._check(
            self.lib.fcg_solve(
                self._h,
                job,
                C.byref(o),
                _dptr(cap),
                C.byref(rows),
                C.byref(cols),
                _dptr(charges) if charges is not None else None,
                C.byref(st),
            ),
            "fcg_solve",
        )
        r, c = rows.value, cols.value
        if mesh.complex:
            capv = cap[: r * c].reshape(r, c) + 1j * cap[r * c: 2 * r * c].reshape(r, c)
            if charges is not None:
                ch = charges[: 2 * r * mesh.N].reshape(r, 2, mesh.N)
                charges = ch[:, 0, :] + 1j * ch[:, 1, :]
        else:
            capv = cap[: r * c].reshape(r, c)
            if charges is not None:
                charges = charges[: r * mesh.N].reshape(r, mesh.N)
        return SolveResult(
            cap=capv,
            iterations=[st.iterations[i] for i in range(min(r, 64))],
            residual=st.final_residual,
            converged=bool(st.converged),
            cycles=int(st.cycles),
            t_solve_s=st.t_solve_s,
            t_mvp_s=st.t_mvp_s,
            n_mvp=st.n_mvp,
            charges=charges,
            residuals=None if hist is None else [
                hist[q * o.residuals_cap: q * o.residuals_cap + min(st.iterations[q] + 1, o.residuals_cap)].copy()
                for q in range(min(r, 64))],
            t_precond_s=st.t_precond_s,
            mem={"gmres": st.mem_gmres, "vectors": st.mem_vectors, "product": st.mem_product},
        )

    def memory(self, job: int) -> dict:
        """Device bytes a resident job holds by category (fcg_job_memory)."""
        m = JobMem()
        self._check(self.lib.fcg_job_memory(self._h, job, C.byref(m)), "fcg_job_memory")
        return {"hierarchy": m.hierarchy, "links": m.links, "conductors": m.conductors, "precond": m.precond}

    def solve_batch_arrays(
        self,
        jobs,
        meshes,
        tol: float = 0.01,
        max_iter: int = 1000,
        precond: int = FCG_PRECOND_JACOBI,
        precision: int = FCG_PREC_FP64,
        max_refine: int = 8,
        precond_size: int = 0,
    ) -> "BatchSolve":
        """Solve several filled jobs together (one launch per stage for all of
        them); the results as arrays (BatchSolve), with no per-job objects."""
        jobs = list(jobs)
        ncs = [m.n_cond for m in meshes]
        cplx = [bool(m.complex) for m in meshes]
        caps = np.zeros(sum(n * n * (2 if c else 1) for n, c in zip(ncs, cplx)), dtype=np.float64)
        ja = (C.c_int * len(jobs))(*jobs)
        rows = (C.c_int * len(jobs))()
        cols = (C.c_int * len(jobs))()
        sts = (SolveStats * len(jobs))()
        o = SolveOpts(tol=tol, max_iter=max_iter, precond=precond, precision=precision, max_refine=max_refine,
                      precond_size=precond_size)
        self._check(
            self.lib.fcg_solve_batch(self._h, ja, len(jobs), C.byref(o), _dptr(caps), rows, cols, sts),
            "fcg_solve_batch",
        )
        return BatchSolve(caps, np.ctypeslib.as_array(rows).copy(), np.ctypeslib.as_array(cols).copy(), cplx,
                          np.frombuffer(sts, dtype=_stats_dtype()))

    def solve_batch(self, jobs, meshes, **kw):
        """solve_batch_arrays with one SolveResult per job, in order."""
        return self.solve_batch_arrays(jobs, meshes, **kw).results()

    def footprint(self, job: int, precision: int = FCG_PREC_FP64, iters: int = 64):
        """(resident bytes of the job, bytes one solve of it allocates with a
        Krylov basis of `iters` vectors)."""
        a, b = C.c_int64(), C.c_int64()
        self._check(
            self.lib.fcg_job_footprint(self._h, job, precision, iters, C.byref(a), C.byref(b)), "fcg_job_footprint"
        )
        return a.value, b.value

    def is_complex(self, job: int) -> bool:
        v = C.c_int()
        self._check(self.lib.fcg_job_is_complex(self._h, job, C.byref(v)), "fcg_job_is_complex")
        return bool(v.value)

    def conductors(self, job: int) -> int:
        n = C.c_int()
        self._check(self.lib.fcg_job_conductors(self._h, job, C.byref(n)), "fcg_job_conductors")
        return n.value

    def engine(self, tol: float = 0.01, max_iter: int = 1000, precond: int = FCG_PRECOND_JACOBI,
               precision: int = FCG_PREC_FP64, max_refine: int = 8, precond_size: int = 0) -> "Engine":
        return Engine(self, SolveOpts(tol=tol, max_iter=max_iter, precond=precond, precision=precision,
                                      max_refine=max_refine, precond_size=precond_size))

    def mvp_timing(self, job: int, k: int, reps: int = 20) -> dict:
        t = _ffi.MvpTiming()
        self._check(self.lib.fcg_mvp_timing_run(self._h, job, k, reps, C.byref(t)), "fcg_mvp_timing_run")
        return {f: getattr(t, f) for f, _ in _ffi.MvpTiming._fields_}

    def mvp_timing_batch(self, jobs, mvp_prec: int = FCG_MVP_FP64, reps: int = 10) -> dict:
        jobs = list(jobs)
        ja = (C.c_int * len(jobs))(*jobs)
        t = _ffi.MvpTiming()
        self._check(self.lib.fcg_mvp_timing_batch(self._h, ja, len(jobs), mvp_prec, reps, C.byref(t)),
                    "fcg_mvp_timing_batch")
        return {f: getattr(t, f) for f, _ in _ffi.MvpTiming._fields_}

    def profile(self, enable: int = -1, reset: bool = False) -> dict:
        """Gather/product device time and algorithmic bytes of engine solves, per
        product precision ("fp64", "fp32", "fp32_acc64"); enable 1/0 switches the
        HIP-event accounting on/off, -1 leaves it."""
        p = _ffi.Profile()
        self._check(self.lib.fcg_ctx_profile(self._h, enable, 1 if reset else 0, C.byref(p)), "fcg_ctx_profile")
        out = {}
        for i, name in enumerate(("fp64", "fp32", "fp32_acc64")):
            out[name] = {
                "gather_ms": p.gather_ms[i],
                "gather_bytes": p.gather_bytes[i],
                "gather_launches": p.gather_launches[i],
                "gather_kernels": p.gather_kernels[i],
                "mvp_ms": p.mvp_ms[i],
                "mvp_launches": p.mvp_launches[i],
            }
        out["engine"] = {"steps": p.steps, "host_plan_ms": p.host_plan_ms, "host_wait_ms": p.host_wait_ms,
                         "host_post_ms": p.host_post_ms}
        return out

    def autotune_batch(self, jobs, mvp_prec: int = FCG_MVP_FP64, reps: int = 3):
        """Pick the gather shape of each rhs class of products over several jobs
        (engine steps); returns (shape of the largest class, gather ms)."""
        jobs = list(jobs)
        ja = (C.c_int * len(jobs))(*jobs)
        v, ms = C.c_int(), C.c_double()
        self._check(self.lib.fcg_autotune_batch(self._h, ja, len(jobs), mvp_prec, reps, C.byref(v), C.byref(ms)),
                    "fcg_autotune_batch")
        return v.value, ms.value

    def set_batch_variant(self, mvp_prec: int, variant: int):
        """Pin one gather shape for every rhs class (variant < 0: class defaults)."""
        self._check(self.lib.fcg_set_batch_variant(self._h, mvp_prec, variant), "fcg_set_batch_variant")

    def set_gather_variant(self, job: int, variant: int) -> int:
        n = C.c_int()
        self._check(self.lib.fcg_set_gather_variant(self._h, job, variant, C.byref(n)), "fcg_set_gather_variant")
        return n.value

    def autotune(self, job: int, k: int, reps: int = 3):
        v, ms = C.c_int(), C.c_double()
        self._check(self.lib.fcg_autotune(self._h, job, k, reps, C.byref(v), C.byref(ms)), "fcg_autotune")
        return v.value, ms.value

    def close(self):
        if getattr(self, "_h", None):
            self.lib.fcg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """Continuous-batching solver: submit filled jobs at any time, step, poll
    finished results.  Every step advances every active solve by one GMRES
    stage with one launch per kernel for all of them."""

    def __init__(self, dev: Device, opts: SolveOpts):
        self.dev = dev
        self.lib = dev.lib
        self.opts = opts
        h = C.c_void_p()
        dev._check(self.lib.fcg_engine_create(dev._h, C.byref(opts), C.byref(h)), "fcg_engine_create")
        self._h = h

    def submit(self, job: int, tag: int):
        self.dev._check(self.lib.fcg_engine_submit(self._h, job, tag), "fcg_engine_submit")

    def step(self):
        """Advance every active solve one stage; returns (active, ready)."""
        a, r = C.c_int(), C.c_int()
        self.dev._check(self.lib.fcg_engine_step(self._h, C.byref(a), C.byref(r)), "fcg_engine_step")
        return a.value, r.value

    def poll(self, max_cols: int = 256):
        """Next finished solve as (tag, job, SolveResult), or None."""
        cap = np.zeros(max_cols * max_cols, dtype=np.float64)
        tag, job, rows, cols = C.c_int64(), C.c_int(), C.c_int(), C.c_int()
        st = SolveStats()
        rc = self.lib.fcg_engine_poll(self._h, C.byref(tag), C.byref(job), _dptr(cap), cap.size, C.byref(rows),
                                      C.byref(cols), C.byref(st))
        if rc != 0:
            return None
        r, c = rows.value, cols.value
        capv = cap[: r * c].reshape(r, c).copy()
        if self.dev.is_complex(job.value):
            capv = capv + 1j * cap[r * c : 2 * r * c].reshape(r, c)
        res = SolveResult(
            cap=capv,
            iterations=[st.iterations[q] for q in range(min(r, 64))],
            residual=st.final_residual,
            converged=bool(st.converged),
            cycles=int(st.cycles),
            t_solve_s=st.t_solve_s,
            n_mvp=st.n_mvp,
        )
        return tag.value, job.value, res

    def bytes(self) -> int:
        b = C.c_int64()
        self.dev._check(self.lib.fcg_engine_bytes(self._h, C.byref(b)), "fcg_engine_bytes")
        return b.value

    def close(self):
        if getattr(self, "_h", None):
            self.lib.fcg_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(argv):
    """Manual- or -a-mode solve with the reference's switches; prints the matrix as
    FasterCap does and returns (mesh, SolveResult).  FasterCapConsole::ParseCmdLine
    (FasterCapConsole.cpp:136-357) + CSolveCap::Run (Solver/SolveCapacitance.cpp:86-256)."""
    from . import cli

    return cli.run(cli.parse_args(argv))
