"""The benchmark workloads at test size: the batched uneven jobs (2D three-line
cross-sections and coplanar plate pairs, as in the batch benchmark) against the
oracle, and the serving stream (Poisson arrivals over jobs from the plate,
crossing-bus and random-box generators, half over the eps 3.9 interface) with
admission by HBM footprint."""
import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import serving, workloads

pytestmark = pytest.mark.gpu


def _per_entry(cap, ref):
    return float(np.max(np.abs(cap - ref) / np.abs(ref)))


@pytest.fixture(scope="module")
def dev():
    d = fc.Device(0)
    yield d
    d.close()


def test_uneven_batch_against_oracle(dev, geo_dir):
    """16 seeded 2D three-line jobs and 4 coplanar plate pairs (pitch U[1.5, 3]) at
    the coarse end of the benchmark's mesh range, solved as one batch: FP64 equals
    the oracle per entry to 1e-6 and mixed precision (tol 1e-8) equals FP64 per
    entry to 1e-4."""
    spec = workloads.batch_uneven(geo_dir + "/ub", n2d=16, n3d=4, m3d_range=(0.008, 0.01))
    meshes = workloads.build_meshes(spec, threads=8)
    print(f"uneven batch: panels {[m.N for m in meshes]}", flush=True)
    assert any(m.dim == 2 for m in meshes) and any(m.dim == 3 for m in meshes)
    jobs = [dev.upload(m) for m in meshes]
    dev.fill_batch(jobs, fc.FCG_PREC_FP64)
    f64 = dev.solve_batch(jobs, meshes, tol=1e-10, precision=fc.FCG_PREC_FP64)
    worst_o = 0.0
    for (path, _, kind), m, r in zip(spec, meshes, f64):
        co, so = orc.fill(m)
        cap, _ = orc.solve(m, co, so, tol=1e-10)
        e = _per_entry(r.cap, cap)
        worst_o = max(worst_o, e)
        assert r.converged
        assert e < 1e-6, (kind, m.N, e)
    dev.fill_batch(jobs, fc.FCG_PREC_MIXED)
    mix = dev.solve_batch(jobs, meshes, tol=1e-8, precision=fc.FCG_PREC_MIXED)
    dev.fill_batch(jobs, fc.FCG_PREC_FP64)
    ref = dev.solve_batch(jobs, meshes, tol=1e-8, precision=fc.FCG_PREC_FP64)
    worst_m = 0.0
    for m, a, b in zip(meshes, mix, ref):
        assert a.converged
        worst_m = max(worst_m, _per_entry(a.cap, b.cap))
    print(f"uneven batch: FP64 vs oracle per entry {worst_o:.2e}; mixed vs FP64 (tol 1e-8) {worst_m:.2e}", flush=True)
    assert worst_m < 1e-4
    for j in jobs:
        dev.release(j)


def test_serving_stream_admission_and_results(dev, geo_dir):
    """Poisson arrivals over 8 resident jobs (plates / crossing bus / boxes, half
    over the interface, 1k-6k panels) with an HBM budget that holds about two
    solves at once: admitted bytes stay within the budget, requests queue and
    overlap, and every result equals its job's alone-solve per entry to 1e-4."""
    pool = workloads.serve_pool(geo_dir + "/sv", n_jobs=8, n_min=1000, n_max=6000, threads=8)
    meshes = [m for m, _, _ in pool]
    kinds = [k for _, k, _ in pool]
    print(f"serving pool: {list(zip(kinds, [m.N for m in meshes]))}", flush=True)
    assert kinds.count("boxes_interface") == 4
    jobs = [dev.upload(m) for m in meshes]
    dev.fill_batch(jobs, fc.FCG_PREC_MIXED)
    tol = 1e-6
    alone = {j: dev.solve(j, m, tol=tol, precision=fc.FCG_PREC_MIXED) for j, m in zip(jobs, meshes)}
    srv = serving.Server(dev, tol=tol, precision=fc.FCG_PREC_MIXED)
    for j in jobs:
        srv.add_job(j)
    big = max(srv.solve_bytes.values())
    resident = sum(srv.resident.values())
    srv.hbm_budget = resident + int(2.2 * big)
    # arrivals faster than the solves: requests wait for HBM
    arr = serving.poisson_arrivals(400.0, 24, jobs, seed=3)
    rep = srv.run(arr)
    srv.close()
    print(f"serving: {rep.completed} done in {rep.elapsed_s:.2f}s, max active {rep.max_active}, "
          f"queue {rep.queue_ms}, latency {rep.latency_ms}, peak admitted {rep.peak_solve_bytes} "
          f"peak device {rep.peak_engine_bytes} budget {srv.hbm_budget - resident}", flush=True)
    assert rep.completed == 24
    assert rep.resident_bytes + rep.peak_solve_bytes <= srv.hbm_budget
    assert rep.peak_engine_bytes <= rep.peak_solve_bytes  # the estimates cover what the solves hold
    assert rep.max_active >= 2  # solves overlapped on the device
    assert rep.queue_ms["p99"] > 0.0  # and requests waited for room
    for r in rep.requests:
        assert r.result.converged
        e = _per_entry(r.result.cap, alone[r.job].cap)
        assert e < 1e-4, (r.rid, r.job, e)
    for j in jobs:
        dev.release(j)
