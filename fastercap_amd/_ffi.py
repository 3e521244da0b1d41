"""ctypes declarations of the C-ABI structs of include/fastercap_amd.h and
include/fastercap_host.h, and loaders for the two in-tree libraries."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
# another build of the host library (the sanitizer build of tests/sanitize.sh)
HOST_LIB_ENV = "FASTERCAP_HOST_LIB"


class NodeGeom(C.Structure):
    _fields_ = [
        ("v", C.c_float * 9),
        ("c", C.c_float * 3),
        ("flags", C.c_uint32),
        ("cond", C.c_int32),
        ("maxside", C.c_double),
        ("area", C.c_double),
        ("dn", C.c_double * 3),
    ]


assert C.sizeof(NodeGeom) == 96


class JobDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("n_leaves", C.c_int32),
        ("n_nodes", C.c_int32),
        ("n_groups", C.c_int32),
        ("n_links", C.c_int64),
        ("nodes", C.POINTER(NodeGeom)),
        ("left", C.POINTER(C.c_int32)),
        ("right", C.POINTER(C.c_int32)),
        ("parent", C.POINTER(C.c_int32)),
        ("leaf_lo", C.POINTER(C.c_int32)),
        ("leaf_hi", C.POINTER(C.c_int32)),
        ("leaf_node", C.POINTER(C.c_int32)),
        ("row_ptr", C.POINTER(C.c_int64)),
        ("col", C.POINTER(C.c_int32)),
        ("group_is_diel", C.POINTER(C.c_int32)),
        ("group_leaf_start", C.POINTER(C.c_int32)),
        ("group_outperm", C.POINTER(C.c_double)),
        ("group_inperm", C.POINTER(C.c_double)),
        ("leaf_outperm", C.POINTER(C.c_double)),
        ("group_outperm_im", C.POINTER(C.c_double)),  # NULL for real permittivities
        ("group_inperm_im", C.POINTER(C.c_double)),
        ("leaf_outperm_im", C.POINTER(C.c_double)),
        ("scheme", C.c_int32),
    ]


class LinkParams(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("reserved", C.c_int32),
        ("link_eps", C.c_double),
        ("max_side", C.c_double),
        ("max_area", C.c_double),
        ("max_length", C.c_double),
    ]


class RefineNode(C.Structure):
    _fields_ = [
        ("v", C.c_float * 9),
        ("c", C.c_float * 3),
        ("n", C.c_float * 3),
        ("pad0", C.c_float),
        ("dim", C.c_double),
        ("maxside", C.c_double),
        ("left", C.c_int32),
        ("right", C.c_int32),
        ("group", C.c_int32),
        ("type", C.c_uint8),
        ("max_side_idx", C.c_uint8),
        ("diel_index", C.c_uint8),
        ("pad1", C.c_uint8),
    ]


class RefineParams(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("n_groups", C.c_int32),
        ("mesh_eps", C.c_double),
        ("curv", C.c_double),
        ("max_side", C.c_double),
        ("max_area", C.c_double),
        ("max_length", C.c_double),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("max_iter", C.c_int32),
        ("precond", C.c_int32),
        ("precision", C.c_int32),
        ("max_refine", C.c_int32),
        ("precond_size", C.c_int32),
        ("residuals", C.POINTER(C.c_double)),
        ("residuals_cap", C.c_int32),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32 * 64),
        ("n_rhs", C.c_int32),
        ("max_iterations", C.c_int32),
        ("final_residual", C.c_double),
        ("t_fill_s", C.c_double),
        ("t_solve_s", C.c_double),
        ("t_mvp_s", C.c_double),
        ("n_mvp", C.c_int64),
        ("converged", C.c_int32),
        ("cycles", C.c_int32),
        ("t_precond_s", C.c_double),
        ("mem_gmres", C.c_int64),
        ("mem_vectors", C.c_int64),
        ("mem_product", C.c_int64),
    ]


class JobMem(C.Structure):
    _fields_ = [
        ("hierarchy", C.c_int64),
        ("links", C.c_int64),
        ("conductors", C.c_int64),
        ("precond", C.c_int64),
    ]


class MvpTiming(C.Structure):
    _fields_ = [
        ("ms_scan", C.c_double),
        ("ms_charge", C.c_double),
        ("ms_gather", C.c_double),
        ("ms_down", C.c_double),
        ("ms_total", C.c_double),
        ("reps", C.c_int32),
        ("k", C.c_int32),
        ("variant", C.c_int32),
    ]


class Profile(C.Structure):
    _fields_ = [
        ("gather_ms", C.c_double * 3),
        ("gather_bytes", C.c_double * 3),
        ("gather_launches", C.c_int64 * 3),
        ("mvp_ms", C.c_double * 3),
        ("mvp_launches", C.c_int64 * 3),
        ("steps", C.c_int64),
        ("host_plan_ms", C.c_double),
        ("host_wait_ms", C.c_double),
        ("host_post_ms", C.c_double),
        ("gather_kernels", C.c_int64 * 3),
    ]


class HostOptions(C.Structure):
    _fields_ = [
        ("mesh_eps", C.c_double),
        ("eps_ratio", C.c_double),
        ("curv_coeff", C.c_double),
        ("verbose", C.c_int32),
        ("links_on_device", C.c_int32),
        ("refine_on_device", C.c_int32),
    ]


class HostStats(C.Structure):
    _fields_ = [
        ("input_raw_cond_panels", C.c_int64),
        ("input_raw_diel_panels", C.c_int64),
        ("input_panels", C.c_int64),
        ("panels", C.c_int64),
        ("nodes", C.c_int64),
        ("links", C.c_int64),
        ("n_conductors", C.c_int32),
        ("n_dielectrics", C.c_int32),
        ("dim", C.c_int32),
        ("max_level", C.c_int32),
        ("max_mesh_eps", C.c_double),
        ("t_read_s", C.c_double),
        ("t_superh_s", C.c_double),
        ("t_discretize_s", C.c_double),
        ("t_links_s", C.c_double),
    ]


# FC_* codes (Solver/SolverGlobal.h:39-49) plus this library's own
FCG_OK = 0
FCG_PREC_FP64 = 0
FCG_PREC_MIXED = 1
FCG_PRECOND_NONE = 0
FCG_PRECOND_JACOBI = 1
FCG_PRECOND_BLOCK = 2
FCG_PRECOND_SUPER = 4
FCG_PRECOND_HIER = 8
FCG_SCHEME_COLLOCATION = 1
FCG_SCHEME_GALERKIN = 2
FCG_MVP_FP64 = 0
FCG_MVP_FP32 = 1
FCG_MVP_FP32_ACC64 = 2
FCG_NODE_LEAF = 1
FCG_NODE_DIEL = 2
FCG_NODE_SUPER = 128

# every symbol the device header declares
DEVICE_SYMBOLS = (
    "fcg_ctx_create",
    "fcg_ctx_destroy",
    "fcg_job_upload",
    "fcg_job_release",
    "fcg_fill",
    "fcg_download_coeffs",
    "fcg_mvp",
    "fcg_mvp_device",
    "fcg_mvp_device_noalloc",
    "fcg_precond_apply",
    "fcg_block_precond_apply",
    "fcg_super_precond_apply",
    "fcg_solve",
    "fcg_job_is_complex",
    "fcg_job_conductors",
    "fcg_last_error",
    "fcg_mvp_timing_run",
    "fcg_autotune",
    "fcg_set_gather_variant",
    "fcg_mvp_prec",
    "fcg_solve_batch",
    "fcg_engine_create",
    "fcg_engine_submit",
    "fcg_engine_step",
    "fcg_engine_poll",
    "fcg_engine_bytes",
    "fcg_engine_destroy",
    "fcg_job_footprint",
    "fcg_fill_batch",
    "fcg_mvp_timing_batch",
    "fcg_ctx_profile",
    "fcg_autotune_batch",
    "fcg_set_batch_variant",
    "fcg_links_discover",
    "fcg_links_fetch",
    "fcg_job_upload_links",
    "fcg_refine",
    "fcg_refine_fetch",
    "fcg_job_memory",
)
HOST_SYMBOLS = (
    "fch_default_options",
    "fch_build",
    "fch_get_desc",
    "fch_get_stats",
    "fch_group_name",
    "fch_last_error",
    "fch_free",
    "fch_get_link_params",
    "fch_set_links",
    "fch_get_coarse_tree",
    "fch_set_refined_tree",
    "fch_write_fastcap",
)


def _lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


_host = None
_dev = None


def host_lib() -> C.CDLL:
    global _host
    if _host is None:
        path = os.environ.get(HOST_LIB_ENV) or _lib_path("libfastercap_host.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = C.CDLL(path)
        lib.fch_default_options.argtypes = [C.POINTER(HostOptions)]
        lib.fch_default_options.restype = None
        lib.fch_build.argtypes = [C.c_char_p, C.POINTER(HostOptions), C.POINTER(C.c_void_p)]
        lib.fch_build.restype = C.c_int
        lib.fch_get_desc.argtypes = [C.c_void_p, C.POINTER(JobDesc)]
        lib.fch_get_desc.restype = C.c_int
        lib.fch_get_stats.argtypes = [C.c_void_p, C.POINTER(HostStats)]
        lib.fch_get_stats.restype = C.c_int
        lib.fch_group_name.argtypes = [C.c_void_p, C.c_int]
        lib.fch_group_name.restype = C.c_char_p
        lib.fch_last_error.argtypes = []
        lib.fch_last_error.restype = C.c_char_p
        lib.fch_free.argtypes = [C.c_void_p]
        lib.fch_free.restype = None
        lib.fch_get_link_params.argtypes = [C.c_void_p, C.POINTER(LinkParams)]
        lib.fch_get_link_params.restype = C.c_int
        lib.fch_set_links.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int64,
                                      C.c_int32, C.c_double]
        lib.fch_set_links.restype = C.c_int
        lib.fch_get_coarse_tree.argtypes = [C.c_void_p, C.POINTER(C.POINTER(RefineNode)), C.POINTER(C.c_int),
                                            C.POINTER(C.POINTER(C.c_int32)), C.POINTER(RefineParams)]
        lib.fch_get_coarse_tree.restype = C.c_int
        lib.fch_set_refined_tree.argtypes = [C.c_void_p, C.POINTER(RefineNode), C.c_int, C.c_double, C.c_int32,
                                             C.c_double]
        lib.fch_set_refined_tree.restype = C.c_int
        lib.fch_write_fastcap.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_double)]
        lib.fch_write_fastcap.restype = C.c_int
        _host = lib
    return _host


def device_lib() -> C.CDLL:
    """Load the gfx950 library; raises if it is missing (no CPU fallback exists)."""
    global _dev
    if _dev is None:
        path = _lib_path("libfcg.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; the HIP path is required (run __graft_entry__.build())")
        lib = C.CDLL(path)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        lib.fcg_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        lib.fcg_ctx_destroy.argtypes = [vp]
        lib.fcg_job_upload.argtypes = [vp, C.POINTER(JobDesc), C.POINTER(C.c_int)]
        lib.fcg_job_release.argtypes = [vp, C.c_int]
        lib.fcg_fill.argtypes = [vp, C.c_int, C.c_int]
        lib.fcg_download_coeffs.argtypes = [vp, C.c_int, dp, dp]
        lib.fcg_mvp.argtypes = [vp, C.c_int, dp, dp, C.c_int]
        lib.fcg_mvp_device.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        lib.fcg_mvp_device_noalloc.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        lib.fcg_precond_apply.argtypes = [vp, C.c_int, dp, dp, C.c_int]
        lib.fcg_block_precond_apply.argtypes = [vp, C.c_int, C.c_int, dp, dp, C.c_int]
        lib.fcg_super_precond_apply.argtypes = [vp, C.c_int, C.c_int, dp, dp, C.c_int]
        lib.fcg_solve.argtypes = [
            vp,
            C.c_int,
            C.POINTER(SolveOpts),
            dp,
            C.POINTER(C.c_int),
            C.POINTER(C.c_int),
            dp,
            C.POINTER(SolveStats),
        ]
        lib.fcg_job_conductors.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
        lib.fcg_job_is_complex.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
        lib.fcg_mvp_timing_run.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(MvpTiming)]
        lib.fcg_set_gather_variant.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int)]
        lib.fcg_autotune.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        ip = C.POINTER(C.c_int)
        i64p = C.POINTER(C.c_int64)
        lib.fcg_autotune_batch.argtypes = [vp, ip, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(C.c_double)]
        lib.fcg_set_batch_variant.argtypes = [vp, C.c_int, C.c_int]
        lib.fcg_ctx_profile.argtypes = [vp, C.c_int, C.c_int, C.POINTER(Profile)]
        lib.fcg_mvp_timing_batch.argtypes = [vp, ip, C.c_int, C.c_int, C.c_int, C.POINTER(MvpTiming)]
        lib.fcg_fill_batch.argtypes = [vp, ip, C.c_int, C.c_int]
        lib.fcg_mvp_prec.argtypes = [vp, C.c_int, C.c_int, dp, dp, C.c_int]
        lib.fcg_solve_batch.argtypes = [vp, ip, C.c_int, C.POINTER(SolveOpts), dp, ip, ip, C.POINTER(SolveStats)]
        lib.fcg_engine_create.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(vp)]
        lib.fcg_engine_submit.argtypes = [vp, C.c_int, C.c_int64]
        lib.fcg_engine_step.argtypes = [vp, ip, ip]
        lib.fcg_engine_poll.argtypes = [vp, i64p, ip, dp, C.c_int, ip, ip, C.POINTER(SolveStats)]
        lib.fcg_engine_bytes.argtypes = [vp, i64p]
        lib.fcg_engine_destroy.argtypes = [vp]
        lib.fcg_job_footprint.argtypes = [vp, C.c_int, C.c_int, C.c_int, i64p, i64p]
        lib.fcg_links_discover.argtypes = [vp, C.POINTER(JobDesc), C.POINTER(LinkParams), i64p,
                                           C.POINTER(C.c_int32)]
        lib.fcg_links_fetch.argtypes = [vp, i64p, C.POINTER(C.c_int32)]
        lib.fcg_job_upload_links.argtypes = [vp, C.POINTER(JobDesc), C.POINTER(C.c_int)]
        lib.fcg_refine.argtypes = [vp, C.POINTER(RefineNode), C.c_int, C.POINTER(C.c_int32), C.POINTER(RefineParams),
                                   C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        lib.fcg_refine_fetch.argtypes = [vp, C.POINTER(RefineNode)]
        lib.fcg_job_memory.argtypes = [vp, C.c_int, C.POINTER(JobMem)]
        lib.fcg_last_error.argtypes = [vp]
        lib.fcg_last_error.restype = C.c_char_p
        for name in DEVICE_SYMBOLS:
            if name != "fcg_last_error":
                getattr(lib, name).restype = C.c_int
        _dev = lib
    return _dev
