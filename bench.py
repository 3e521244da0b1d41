"""Benchmark of the MI355X FasterCap solve path (one GPU).

Workloads (--workload):
  batch  (default) batched uneven solves: 1024 seeded 2D three-line cross-sections
         (~300-380 segments) plus 256 seeded 3D coplanar plate pairs (pitch
         U[1.5,3], 2k-33k panels).  A step is one extraction of the whole batch:
         the batched link fill and the batched GMRES of every conductor of every
         job, advanced together.  `value` is the mixed-precision path (FP32
         far-field links and products, FP64 near field / Arnoldi / refinement);
         the plain FP64 path (one job at a time, as the reference solves) and the
         batched FP64 path are timed on the same batch beside it.
  bus    8-conductor crossing bus (bars 1x1x9, unit 1e-6 m, -m600: 74k panels,
         48M links); a step is fill + GMRES of the 8 conductors.
  boxes  32 random boxes over an eps 3.9 interface (mixed vs FP64).
  plates two 1 m plates, 0.1 m gap.

Meshes and link structures are built on the host and uploaded before the timed
region.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads: one per core, packed (read when the oracle's
# OpenMP runtime starts, which is after this point)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

PEAK_HBM_GBS = 8000.0
ACCURACY_BAR = 1e-4  # largest per-entry relative C error against the FP64 path (tol 1e-8)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(xs, p):
    s = sorted(xs)
    return s[min(len(s) - 1, int(round(p / 100.0 * (len(s) - 1))))]


def c_errors(caps, refs):
    """(max per-entry relative error, max error relative to the row diagonal)."""
    import numpy as np

    pe, pd = 0.0, 0.0
    for c, r in zip(caps, refs):
        d = np.abs(c - r)
        pe = max(pe, float(np.max(d / np.abs(r))))
        pd = max(pd, float(np.max(d / np.abs(np.diag(r))[:, None])))
    return pe, pd


def cpu_info():
    """CPU model, sockets and cores of this host (lscpu), the OpenMP placement."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    sockets = int(info.get("Socket(s)", "0") or 0)
    cps = int(info.get("Core(s) per socket", "0") or 0)
    return {"model": info.get("Model name", "unknown"), "sockets": sockets, "cores_per_socket": cps,
            "threads_per_core": int(info.get("Thread(s) per core", "0") or 0),
            "host_cores": sockets * cps, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def copy_ceiling(mib=1024, reps=5):
    """Achievable HBM rate of a plain device-to-device copy of `mib` MiB (read +
    write bytes / time), through the HIP runtime the library links."""
    import ctypes as C

    from fastercap_amd import _ffi

    _ffi.device_lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if h.hipMalloc(C.byref(a), n) != 0 or h.hipMalloc(C.byref(b), n) != 0:
        return None
    try:
        h.hipMemcpy(b, a, n, 3)  # warm
        h.hipDeviceSynchronize()
        t = time.perf_counter()
        for _ in range(reps):
            if h.hipMemcpy(b, a, n, 3) != 0:  # hipMemcpyDeviceToDevice
                return None
        h.hipDeviceSynchronize()
        dt = (time.perf_counter() - t) / reps
    finally:
        h.hipFree(a)
        h.hipFree(b)
    return {"gbps": 2 * n / dt / 1e9, "bytes": n, "reps": reps,
            "kind": "hipMemcpy device-to-device, read + write bytes"}


def ncond(mesh):
    return mesh.n_cond


def cpu_sample(meshes, idx, tol, super_dim=0):
    """Oracle fill + full GMRES of the sampled jobs on all host cores (Jacobi, or
    the two-level preconditioner of dimension super_dim when it is positive)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    ncores = len(os.sched_getaffinity(0))
    a = time.perf_counter()
    for i in idx:
        co, so = orc.fill(meshes[i], nthreads=ncores)
        if super_dim > 0:
            orc.solve_super(meshes[i], co, so, super_dim, tol=tol, nthreads=ncores)
        else:
            orc.solve(meshes[i], co, so, tol=tol, nthreads=ncores)
    return time.perf_counter() - a, ncores


def bench_batch(args, fc, dev, tmp):
    import numpy as np

    from fastercap_amd import workloads

    spec = workloads.batch_uneven(tmp, n2d=args.n2d, n3d=args.n3d)
    t0 = time.perf_counter()
    meshes = workloads.build_meshes(spec, threads=16)
    log(f"built {len(meshes)} meshes in {time.perf_counter() - t0:.1f}s")
    Ns = [m.N for m in meshes]
    Ls = [m.L for m in meshes]
    jobs = [dev.upload(m) for m in meshes]
    nj = len(jobs)
    log(f"uploaded: panels {sum(Ns)} (min {min(Ns)} max {max(Ns)}), links {sum(Ls)}")
    MIX, F64 = fc.FCG_PREC_MIXED, fc.FCG_PREC_FP64
    tol = args.tol
    # preconditioner of every solve of this workload; a fill drops the two-level
    # inverse, so each step's solve builds it again (one launch for the batch)
    two = args.precond == "two-level"
    PC = dict(precond=fc.FCG_PRECOND_SUPER if two else fc.FCG_PRECOND_JACOBI,
              precond_size=args.precond_size if two else 0)

    # products over several jobs use the context's batch gather shape; the default
    # shape is what fcg_autotune_batch picks for this batch, so the timed run makes
    # no extra tuning launches (the rocprof averages of the same command then cover
    # the solve launches only)
    tuned = {}
    if args.gather_variant >= 0:  # one pinned FP32 gather shape for every rhs class (comparisons)
        dev.set_batch_variant(fc.FCG_MVP_FP32, args.gather_variant)
    if args.only_mixed:  # profiling aid: account for every product of the process
        dev.profile(enable=1, reset=True)
    for _ in range(max(1, args.warmup)):
        dev.fill_batch(jobs, MIX)
        dev.solve_batch(jobs, meshes, tol=tol, precision=MIX, **PC)
    log("warmup done")
    lat = []
    # Python's collector is paused over the timed steps (as the standard library's
    # timeit does): a full collection of the interpreter's objects otherwise lands
    # in some step; the solver itself is native code and unaffected
    gc.collect()
    gc.disable()
    dev.profile(enable=1, reset=not args.only_mixed)
    t_start = time.perf_counter()
    res = None
    for s in range(args.steps):
        a = time.perf_counter()
        dev.fill_batch(jobs, MIX)
        b = time.perf_counter()
        # every job's C matrix and statistics as arrays (BatchSolve)
        res = dev.solve_batch_arrays(jobs, meshes, tol=tol, precision=MIX, **PC)
        c = time.perf_counter()
        lat = res.stats["t_solve_s"].tolist()
        log(f"step {s}: fill {1e3 * (b - a):.1f} ms, batched mixed solve {1e3 * (c - b):.1f} ms")
    t_mixed = time.perf_counter() - t_start
    gc.enable()
    prof = dev.profile(enable=0, reset=True)
    value = args.steps * nj / t_mixed
    res = res.results()  # one SolveResult per job, for the checks below
    mixed_caps = [r.cap for r in res]
    iters_mixed = [max(r.iterations) for r in res]
    unconverged = sum(1 for r in res if not r.converged)
    cycles_mixed = [r.cycles for r in res]
    if args.only_mixed:  # profiling aid: the mixed solves alone, warm-up included
        log(f"all products of the process: {prof}")
        return None

    # plain FP64 path: one job at a time (fill + solve), as the reference runs
    dev.fill_batch(jobs, F64)
    for j, m in zip(jobs, meshes):  # warm
        dev.solve(j, m, tol=tol, **PC)
    a = time.perf_counter()
    plain = []
    for j, m in zip(jobs, meshes):
        dev.fill(j, F64)
        plain.append(dev.solve(j, m, tol=tol, **PC))
    t_plain = time.perf_counter() - a
    log(f"plain FP64, one job at a time: {t_plain:.2f}s ({nj / t_plain:.1f} solves/s)")
    tuned["fp64"] = dev.autotune_batch(jobs, fc.FCG_MVP_FP64)
    dev.solve_batch(jobs, meshes, tol=tol, precision=F64, **PC)  # warm
    a = time.perf_counter()
    dev.fill_batch(jobs, F64)
    bres = dev.solve_batch_arrays(jobs, meshes, tol=tol, precision=F64, **PC)
    t_b64 = time.perf_counter() - a
    bres = bres.results()
    log(f"batched FP64: {t_b64:.2f}s ({nj / t_b64:.1f} solves/s)")
    ref = dev.solve_batch(jobs, meshes, tol=1e-8, precision=F64, **PC)
    e_conv = c_errors(mixed_caps, [r.cap for r in ref])
    e_plain = c_errors(mixed_caps, [r.cap for r in plain])
    e_plain_conv = c_errors([r.cap for r in plain], [r.cap for r in ref])
    unconverged += sum(1 for r in plain + bres + ref if not r.converged)
    log(f"mixed vs FP64 tol 1e-8: per-entry {e_conv[0]:.3e} vs-diag {e_conv[1]:.3e}; "
        f"plain FP64 at tol {tol} vs tol 1e-8: {e_plain_conv[0]:.3e}")
    # the mixed path's accuracy floor: mixed at tol 1e-8 against FP64 at tol 1e-8
    dev.fill_batch(jobs, MIX)
    floor = dev.solve_batch(jobs, meshes, tol=1e-8, precision=MIX, **PC)
    unconverged += sum(1 for r in floor if not r.converged)
    e_floor = c_errors([r.cap for r in floor], [r.cap for r in ref])
    log(f"mixed floor (tol 1e-8 vs FP64 tol 1e-8): per-entry {e_floor[0]:.3e} vs-diag {e_floor[1]:.3e}; "
        f"unconverged solves {unconverged}")
    bar_met = e_conv[0] <= ACCURACY_BAR and unconverged == 0

    # dominant kernel: the FP32 gather of the inner products of the timed solves
    # (HIP events around the gather of every product, which is one launch per rhs
    # class of the product's jobs; bytes per product from the active jobs)
    g = prof["fp32"]
    achieved = g["gather_bytes"] / (g["gather_ms"] * 1e-3) / 1e9
    # HBM bytes per product: the ratio of HBM bytes (rocprofv3 FETCH_SIZE and
    # WRITE_SIZE passes on this workload, scaled by a plain copy's readings) to
    # algorithmic bytes, stored under profiles/, times this run's algorithmic bytes
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "gather_traffic_batch.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            ratio = json.load(f).get("traffic_over_alg")
        if ratio:
            traffic = ratio * g["gather_bytes"] / max(1, g["gather_launches"])
    log(f"timed-region products: {prof}")

    pc_name = f"two-level, dim {PC['precond_size']}" if two else "Jacobi"
    cpu = None
    if not args.no_cpu:
        stride = max(1, nj // args.cpu_jobs)
        idx = list(range(0, nj, stride))
        t_cpu, ncores = cpu_sample(meshes, idx, tol, PC["precond_size"])
        cpu = {
            "value": len(idx) / t_cpu,
            "unit": "solves/s",
            "cores": ncores,
            "kind": "port",
            "sample": f"oracle fill + FP64 GMRES ({pc_name}) of every conductor of every {stride}th job "
                      f"({len(idx)} of {nj} jobs, {t_cpu:.1f}s)",
            "host": cpu_info(),
        }
        log(f"cpu baseline: {cpu}")

    ceiling = copy_ceiling()
    metric = ("solves/s (batched uneven jobs, mixed precision, <=1e-4 rel C error vs FP64 path)" if bar_met else
              "solves/s (batched uneven jobs, mixed precision; accuracy bar NOT met, see accuracy_bar_met)")
    return {
        "metric": metric,
        "value": value,
        "unit": "solves/s",
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "dtype": "f32+f64",
        "config": {
            "workload": f"{args.n2d} 2D three-line cross-sections (-m0.02) + {args.n3d} 3D coplanar plate pairs "
                        f"(pitch U[1.5,3], -m log-uniform [0.0025,0.01])",
            "jobs": nj,
            "panels_total": int(sum(Ns)),
            "panels_min": int(min(Ns)),
            "panels_max": int(max(Ns)),
            "links_total": int(sum(Ls)),
            "gmres_tol": tol,
            "preconditioner": pc_name + (" (FasterCap -ps%d), inverse built again after each step's fill" % PC["precond_size"]
                                         if two else " (FasterCap default)"),
            "precision": "mixed: FP32 far-field links and inner products, FP64 near field, Arnoldi, refinement",
        },
        "latency_ms": {"p50": 1e3 * pct(lat, 50), "p99": 1e3 * pct(lat, 99)},
        "batch_gather": "one launch per rhs class (2D lines k=3 -> stride 4, plates k=2 -> stride 2), "
                        "default per-class shapes for the mixed solves, tuned per class for batched FP64",
        "fp64_plain_solves_per_s": nj / t_plain,
        "fp64_batched_solves_per_s": nj / t_b64,
        "speedup_vs_fp64_plain": value / (nj / t_plain),
        "gmres_iterations_max_per_job": {"p50": pct(iters_mixed, 50), "max": max(iters_mixed)},
        "accuracy_bar_met": bar_met,
        "accuracy_bar": f"max per-entry |C_mixed - C_FP64| / |C_FP64| <= {ACCURACY_BAR:g} (FP64 at tol 1e-8), "
                        "every solve converged",
        "unconverged_solves": unconverged,
        "refine_cycles_mixed": {"p50": pct(cycles_mixed, 50), "max": max(cycles_mixed)},
        "c_rel_err_mixed_vs_fp64_tol1e-8": {"max_per_entry": e_conv[0], "max_vs_diagonal": e_conv[1]},
        "mixed_floor_per_entry": e_floor[0],
        "mixed_floor": "mixed at tol 1e-8 against FP64 at tol 1e-8 (max per entry)",
        "c_rel_err_fp64_plain_vs_fp64_tol1e-8": {"max_per_entry": e_plain_conv[0], "max_vs_diagonal": e_plain_conv[1]},
        "c_rel_err_mixed_vs_fp64_plain_same_tol": {"max_per_entry": e_plain[0], "max_vs_diagonal": e_plain[1]},
        "mvp": {
            "inner_fp32": {"launches": g["mvp_launches"], "avg_ms": g["mvp_ms"] / max(1, g["mvp_launches"])},
            "residual_fp32_coef_fp64_acc": {
                "launches": prof["fp32_acc64"]["mvp_launches"],
                "avg_ms": prof["fp32_acc64"]["mvp_ms"] / max(1, prof["fp32_acc64"]["mvp_launches"])},
        },
        "roofline": {
            "kernel": "k_gather_staged<float,float,float,*> (inner FP32 products of the timed batched solves)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "traffic_source": "profiles/gather_traffic_batch.md" if traffic else None,
            "products": g["gather_launches"],
            "kernel_launches": g["gather_kernels"],
            "avg_gather_ms_per_product": g["gather_ms"] / max(1, g["gather_launches"]),
            "alg_bytes_per_product": g["gather_bytes"] / max(1, g["gather_launches"]),
            "achievable": ceiling["gbps"] if ceiling else None,
            "achievable_source": ceiling,
        },
        "cpu_baseline": cpu,
    }


def bench_serve(args, fc, dev, tmp):
    """Serving stream: Poisson arrivals over a pool of resident jobs."""
    import numpy as np

    from fastercap_amd import serving, workloads

    t0 = time.perf_counter()
    pool = workloads.serve_pool(tmp, n_jobs=args.serve_jobs, n_max=args.serve_max_panels, threads=16)
    meshes = [m for m, _, _ in pool]
    kinds = [k for _, k, _ in pool]
    log(f"built {len(meshes)} serving jobs in {time.perf_counter() - t0:.1f}s: panels "
        f"{min(m.N for m in meshes)}..{max(m.N for m in meshes)}, links {sum(m.L for m in meshes)}")
    # progress lines at every stage: the largest pools take minutes to set up
    jobs = []
    for i, m in enumerate(meshes):
        jobs.append(dev.upload(m))
        if m.N > 100_000 or i == len(meshes) - 1:
            log(f"uploaded {i + 1}/{len(meshes)} jobs (this one {m.N} panels, {m.L} links)")
    MIX, F64 = fc.FCG_PREC_MIXED, fc.FCG_PREC_FP64
    dev.fill_batch(jobs, F64)
    log("FP64 fill done")
    dev.autotune_batch(jobs, fc.FCG_MVP_FP64)
    log("FP64 gather shapes tuned")
    refs = {j: r.cap for j, r in zip(jobs, dev.solve_batch(jobs, meshes, tol=1e-8, precision=F64))}
    log("reference FP64 solves (tol 1e-8) done")
    dev.fill_batch(jobs, MIX)
    dev.autotune_batch(jobs, fc.FCG_MVP_FP32)
    dev.autotune_batch(jobs, fc.FCG_MVP_FP32_ACC64)
    log("mixed fill and gather shapes done")
    # closed-batch capacity: every job once, all together
    dev.solve_batch(jobs, meshes, tol=args.tol, precision=MIX)
    a = time.perf_counter()
    dev.solve_batch(jobs, meshes, tol=args.tol, precision=MIX)
    cap_rate = len(jobs) / (time.perf_counter() - a)
    log(f"closed-batch capacity {cap_rate:.1f} solves/s")

    def serve(rate, window, seed):
        srv = serving.Server(dev, tol=args.tol, precision=MIX)
        for j in jobs:
            srv.add_job(j)
        arr = serving.poisson_arrivals(rate, max(8, int(rate * window)), jobs, seed)
        rep = srv.run(arr)
        srv.close()
        errs = [c_errors([r.result.cap], [refs[r.job]]) for r in rep.requests]
        unconv = sum(1 for r in rep.requests if not r.result.converged)
        if not errs:
            return rep, (float("nan"), float("nan")), unconv
        return rep, (max(e[0] for e in errs), max(e[1] for e in errs)), unconv

    serve(cap_rate, 2.0, 1)  # warm
    log("serving warm-up done")
    dev.profile(enable=1, reset=True)
    rate = args.serve_load * cap_rate
    rep, err, unc = serve(rate, args.serve_window, 7)
    prof = dev.profile(enable=0, reset=True)
    log(f"load {args.serve_load:.2f}x capacity ({rate:.1f}/s): {rep.completed} solves in {rep.elapsed_s:.1f}s = "
        f"{rep.solves_per_s:.1f}/s, latency {rep.latency_ms}, max active {rep.max_active}, C err {err}")
    half, err_h, unc_h = serve(0.5 * cap_rate, args.serve_window / 2, 11)
    bar_met = max(err[0], err_h[0]) <= ACCURACY_BAR and unc + unc_h == 0
    log(f"half load: {half.solves_per_s:.1f}/s, latency {half.latency_ms}")
    g = prof["fp32"]
    achieved = g["gather_bytes"] / (g["gather_ms"] * 1e-3) / 1e9 if g["gather_ms"] > 0 else 0.0
    return {
        "metric": ("served solves/s (Poisson arrivals over resident jobs, mixed precision, <=1e-4 rel C error)"
                   if bar_met else "served solves/s (Poisson arrivals, mixed precision; accuracy bar NOT met)"),
        "accuracy_bar_met": bar_met,
        "unconverged_solves": unc + unc_h,
        "value": rep.solves_per_s,
        "unit": "solves/s",
        "ms_per_step": 1e3 * rep.elapsed_s / max(1, rep.steps),
        "dtype": "f32+f64",
        "config": {
            "workload": f"{len(jobs)} resident jobs from the plate / crossing-bus / random-box generators, half over "
                        f"the eps 3.9 interface, panels log-uniform in [1000, {args.serve_max_panels}]",
            "jobs": len(jobs),
            "kinds": {k: kinds.count(k) for k in set(kinds)},
            "panels_total": int(sum(m.N for m in meshes)),
            "links_total": int(sum(m.L for m in meshes)),
            "arrival_rate_per_s": rate,
            "load_vs_closed_batch_capacity": args.serve_load,
            "requests": rep.completed,
            "gmres_tol": args.tol,
        },
        "latency_ms": rep.latency_ms,
        "queue_ms": rep.queue_ms,
        "half_load": {"rate_per_s": 0.5 * cap_rate, "solves_per_s": half.solves_per_s, "latency_ms": half.latency_ms},
        "closed_batch_capacity_solves_per_s": cap_rate,
        "max_active": rep.max_active,
        "c_rel_err_vs_fp64_tol1e-8": {"max_per_entry": max(err[0], err_h[0]), "max_vs_diagonal": max(err[1], err_h[1])},
        "roofline": {"kernel": "k_gather_staged<float,float,float,*> (inner FP32 products while serving)", "bound": "hbm",
                     "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": None, "products": g["gather_launches"], "kernel_launches": g["gather_kernels"],
                     "avg_gather_ms_per_product": g["gather_ms"] / max(1, g["gather_launches"])},
        "cpu_baseline": None,
    }


def single_workload(name, d):
    from fastercap_amd import workloads

    if name == "bus":
        return workloads.crossing_bus(d), 600.0, "crossing bus 8 bars 1x1x9 um, -m600"
    if name == "plates":
        return workloads.two_plates(d), 0.01, "two 1x1 m plates 0.1 m gap, -m0.01"
    if name == "boxes":
        return (workloads.random_boxes(d, n=32, seed=0), 3.4e-6,
                "32 random boxes (sides log-uniform [0.2,2] in a 20^3 cube) over an eps 3.9 interface, -m3.4e-6")
    raise SystemExit(f"unknown workload {name}")


def bench_single(args, fc, dev, tmp):
    import numpy as np

    path, m_default, desc = single_workload(args.workload, tmp)
    m = args.mesh_eps if args.mesh_eps is not None else m_default
    t0 = time.perf_counter()
    mesh = fc.Mesh.build(path, mesh_eps=m)
    t_build = time.perf_counter() - t0
    k = ncond(mesh)
    log(f"mesh: N={mesh.N} M={mesh.M} L={mesh.L} conductors={k} build={t_build:.2f}s")
    prec = fc.FCG_PREC_MIXED if args.precision == "mixed" else fc.FCG_PREC_FP64
    job = dev.upload(mesh)
    dev.fill(job, prec)
    for _ in range(args.warmup):
        dev.fill(job, prec)
        dev.solve(job, mesh, tol=args.tol, precision=prec)
    log("warmup done")
    lat, fills, solves = [], [], []
    dev.profile(enable=1, reset=True)
    t_start = time.perf_counter()
    for s in range(args.steps):
        a = time.perf_counter()
        dev.fill(job, prec)
        b = time.perf_counter()
        res = dev.solve(job, mesh, tol=args.tol, precision=prec)
        c = time.perf_counter()
        lat.append(c - a)
        fills.append(b - a)
        solves.append(c - b)
        log(f"step {s}: fill {1e3 * (b - a):.1f} ms solve {1e3 * (c - b):.1f} ms iters {res.iterations}")
    t_total = time.perf_counter() - t_start
    prof = dev.profile(enable=0, reset=True)
    value = args.steps / t_total
    cap = res.cap
    dev.fill(job, fc.FCG_PREC_FP64)
    ref = dev.solve(job, mesh, tol=1e-8).cap
    e_pe, e_pd = c_errors([cap], [ref])
    log(f"C error vs tol 1e-8 FP64 solve: per-entry {e_pe:.3e}, relative to diagonal {e_pd:.3e}")
    g = prof["fp32" if prec == fc.FCG_PREC_MIXED else "fp64"]
    achieved = g["gather_bytes"] / (g["gather_ms"] * 1e-3) / 1e9
    log(f"timed-region products: {prof}")
    tm = dev.mvp_timing(job, k, reps=args.reps) if prec == fc.FCG_PREC_FP64 else None
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc

        ncores = len(os.sched_getaffinity(0))
        a = time.perf_counter()
        ocoef, oself = orc.fill(mesh, nthreads=ncores)
        t_fill_cpu = time.perf_counter() - a
        a = time.perf_counter()
        ocap, _ = orc.solve_rows(mesh, ocoef, oself, args.cpu_rows, tol=args.tol, nthreads=ncores)
        t_rows = time.perf_counter() - a
        t_cpu = t_fill_cpu + t_rows * (k / args.cpu_rows)
        cpu = {
            "value": 1.0 / t_cpu,
            "unit": "solves/s",
            "cores": ncores,
            "kind": "port",
            "sample": f"oracle fill of all {mesh.L} links + GMRES of {args.cpu_rows} of {k} right-hand sides "
                      f"(fill {t_fill_cpu:.2f}s, rhs {t_rows:.2f}s, scaled to {k} rhs)",
        }
        log(f"cpu baseline: {cpu}")
    return {
        "metric": f"solves/s (fill + GMRES of all conductors, {args.precision} path)",
        "value": value,
        "unit": "solves/s",
        "ms_per_step": 1e3 * t_total / args.steps,
        "dtype": "f64" if prec == fc.FCG_PREC_FP64 else "f32+f64",
        "config": {"workload": desc, "panels": mesh.N, "nodes": mesh.M, "links": mesh.L, "conductors": k,
                   "gmres_tol": args.tol, "precision": args.precision},
        "latency_ms": {"p50": 1e3 * statistics.median(lat), "p99": 1e3 * pct(lat, 99)},
        "fill_ms": 1e3 * statistics.median(fills),
        "solve_ms": 1e3 * statistics.median(solves),
        "gmres_iterations": res.iterations,
        "mvp": {"k": k, "launches": g["mvp_launches"], "avg_ms": g["mvp_ms"] / max(1, g["mvp_launches"]),
                "stage_ms": None if tm is None else {s: tm["ms_" + s] for s in ("scan", "charge", "gather", "down")},
                "gather_variant": None if tm is None else tm["variant"]},
        "c_rel_err_vs_fp64_tol1e-8": {"max_per_entry": e_pe, "max_vs_diagonal": e_pd},
        # jobs of at most 65536 nodes gather with their tiles' sources staged in LDS
        "roofline": {"kernel": ("k_gather_staged" if mesh.M <= 65536 else "k_gather") + " (products of the timed solves)",
                     "bound": "hbm", "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "products": g["gather_launches"], "kernel_launches": g["gather_kernels"],
                     "avg_gather_ms_per_product": g["gather_ms"] / max(1, g["gather_launches"]),
                     "alg_bytes_per_product": g["gather_bytes"] / max(1, g["gather_launches"])},
        "cpu_baseline": cpu,
        "host_build_s": t_build,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch", choices=["batch", "serve", "bus", "plates", "boxes"])
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--serve-max-panels", type=int, default=1_000_000)
    ap.add_argument("--serve-load", type=float, default=1.5, help="arrival rate / closed-batch capacity")
    ap.add_argument("--serve-window", type=float, default=20.0, help="seconds of arrivals")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"], help="single-job workloads")
    ap.add_argument("--precond", default="two-level", choices=["two-level", "jacobi"],
                    help="batch: preconditioner of every solve")
    ap.add_argument("--precond-size", type=int, default=128, help="batch: two-level dimension")
    ap.add_argument("--mesh-eps", type=float, default=None)
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--n2d", type=int, default=1024)
    ap.add_argument("--n3d", type=int, default=256)
    ap.add_argument("--reps", type=int, default=10, help="timed repetitions of the product for the roofline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-mixed", action="store_true", help="batch: run only the timed mixed solves (profiling)")
    ap.add_argument("--gather-variant", type=int, default=-1,
                    help="batch: pin the FP32 gather shape of every rhs class (fcg_set_batch_variant)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only copy this many MiB device-to-device (reference kernel for memory counters) and exit")
    ap.add_argument("--cpu-rows", type=int, default=1, help="single-job CPU sample: right-hand sides")
    ap.add_argument("--cpu-jobs", type=int, default=1280, help="batch CPU sample: jobs (evenly strided)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; how the path would shard is described in "
                                   "the design notes, not run", "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    if args.tol is None:
        # the batch and serving workloads are quoted at <=1e-4 per-entry C error
        # against FP64, which GMRES tol 1e-6 meets on the uneven batch
        args.tol = 1e-6 if args.workload in ("batch", "serve", "boxes") else 1e-4

    if args.calibrate_copy_mib > 0:
        import torch

        n = args.calibrate_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": 4 * n, "copies": 3}), flush=True)
        return

    import fastercap_amd as fc

    tmp = tempfile.mkdtemp(prefix="fcbench")
    dev = fc.Device(0)
    if args.workload == "batch":
        out = bench_batch(args, fc, dev, tmp)
        if out is None:
            dev.close()
            return
    elif args.workload == "serve":
        out = bench_serve(args, fc, dev, tmp)
    else:
        out = bench_single(args, fc, dev, tmp)
    line = {
        "metric": out.pop("metric"),
        "value": out.pop("value"),
        "unit": out.pop("unit"),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": out.pop("dtype"),
        "data": "synthetic",
        "config": out.pop("config"),
    }
    line.update(out)
    print(json.dumps(line), flush=True)
    dev.close()
    if line.get("unconverged_solves"):
        log(f"{line['unconverged_solves']} solves ended without meeting their tolerance")
        sys.exit(1)


if __name__ == "__main__":
    main()
