"""GPU tests of the mixed-precision path, the batched solve and the continuous
batching engine, plus 2D and dielectric geometries, all through the C-ABI and
checked against the oracle (or a closed form)."""
import math
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads
from test_golden_cpu import two_cubes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = fc.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def meshes(geo_dir):
    """Three different jobs: stacked plates, the two-cubes example, a box cluster
    over a dielectric interface."""
    os.makedirs(os.path.join(geo_dir, "c"), exist_ok=True)
    return {
        "plates": fc.Mesh.build(workloads.two_plates(geo_dir + "/p"), mesh_eps=0.02),
        "cubes": fc.Mesh.build(two_cubes(geo_dir + "/c"), mesh_eps=0.05),
        "boxes": fc.Mesh.build(workloads.random_boxes(geo_dir + "/b", n=8, seed=1), mesh_eps=0.1),
    }


def _c_err(cap, ref):
    """Largest per-entry relative difference |C - C_ref| / |C_ref| (the metric's
    error measure); the difference relative to each row's diagonal is printed."""
    d = np.abs(cap - ref)
    print(f"  C error: per entry {np.max(d / np.abs(ref)):.3e}, "
          f"vs row diagonal {np.max(d / np.abs(np.diag(ref))[:, None]):.3e}", flush=True)
    return float(np.max(d / np.abs(ref)))


def test_mixed_fill_close_to_fp64(dev, meshes):
    m = meshes["plates"]
    j64 = dev.upload(m)
    dev.fill(j64, fc.FCG_PREC_FP64)
    c64, s64 = dev.coeffs(j64, m)
    j32 = dev.upload(m)
    dev.fill(j32, fc.FCG_PREC_MIXED)
    c32, s32 = dev.coeffs(j32, m)
    rel = np.abs(c32 - c64) / np.abs(c64)
    print(f"mixed fill: max rel {rel.max():.3e} median {np.median(rel):.3e}", flush=True)
    # FP32 storage (2^-24) plus FP32 far-field arithmetic
    assert rel.max() < 2e-6
    np.testing.assert_array_equal(s32, s64)  # self terms stay FP64
    dev.release(j64)
    dev.release(j32)


@pytest.mark.parametrize("k", [1, 4, 9])
def test_mvp_precisions_against_oracle(dev, meshes, k):
    m = meshes["cubes"]
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_MIXED)
    ocoef, oself = orc.fill(m)
    q = np.random.default_rng(30 + k).standard_normal((k, m.N))
    vo = np.stack([orc.mvp(m, ocoef, oself, q[r]) for r in range(k)])
    scale = np.abs(vo).max()
    v32 = dev.mvp_prec(job, q, fc.FCG_MVP_FP32)
    v3264 = dev.mvp_prec(job, q, fc.FCG_MVP_FP32_ACC64)
    e32 = np.abs(v32 - vo).max() / scale
    e3264 = np.abs(v3264 - vo).max() / scale
    print(f"mvp k={k}: fp32 {e32:.3e} fp32-coef/fp64-acc {e3264:.3e}", flush=True)
    assert e32 < 2e-5
    assert e3264 < 2e-6
    with pytest.raises(fc.FasterCapError):
        dev.mvp_prec(job, q, fc.FCG_MVP_FP64)  # no FP64 coefficients on this job
    dev.release(job)


@pytest.mark.parametrize("name", ["plates", "cubes", "boxes"])
def test_mixed_solve_within_1e4_of_fp64(dev, meshes, name):
    m = meshes[name]
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_FP64)
    ref = dev.solve(job, m, tol=1e-8)
    dev.fill(job, fc.FCG_PREC_MIXED)
    mix = dev.solve(job, m, tol=1e-6, precision=fc.FCG_PREC_MIXED)
    err = _c_err(mix.cap, ref.cap)
    print(f"{name}: mixed C err per entry {err:.3e} residual {mix.residual:.3e} products {mix.n_mvp} "
          f"cycles {mix.cycles}", flush=True)
    assert err < 1e-4
    assert mix.converged and mix.residual < 1e-6
    dev.release(job)


@pytest.mark.parametrize("name", ["cubes", "boxes"])
def test_mixed_accuracy_floor(dev, meshes, name):
    """Mixed precision at tol 1e-8 against FP64 at tol 1e-8: the refinement's
    residual uses FP32 far-field and FP64 near-field coefficients, so what is left
    is the FP32 rounding of the far field."""
    m = meshes[name]
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_FP64)
    ref = dev.solve(job, m, tol=1e-8)
    dev.fill(job, fc.FCG_PREC_MIXED)
    mix = dev.solve(job, m, tol=1e-8, precision=fc.FCG_PREC_MIXED)
    err = _c_err(mix.cap, ref.cap)
    print(f"{name}: mixed floor per entry {err:.3e}, cycles {mix.cycles}, residual {mix.residual:.3e}", flush=True)
    assert mix.converged
    assert err < 1e-6
    dev.release(job)


def _near_mask(m):
    """Links of the analytic PotEstimateOpt branch in descriptor order (3D: centroid
    distance below one max side; 2D: below four), computed as the fill classes are."""
    g = m.node_geom()
    rows = np.repeat(np.arange(m.M), np.diff(m.row_ptr))
    c1, c2 = g["c"][rows].astype(np.float64), g["c"][m.col].astype(np.float64)
    d = c1 - c2
    if m.dim == 2:
        dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
    else:
        dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    ratio = dist / np.maximum(g["maxside"][rows], g["maxside"][m.col])
    return ratio < (4.0 if m.dim == 2 else 1.0)


@pytest.mark.parametrize("name", ["cubes", "boxes", "lines2d"])
def test_residual_product_uses_fp64_near_field(dev, meshes, geo_dir, name):
    """The refinement residual's product (FCG_MVP_FP32_ACC64) equals the oracle's
    product with the FP64 coefficient on every analytic-branch link and the
    device's FP32 coefficient on every other link (FP64 charges and sums)."""
    if name == "lines2d":
        m = fc.Mesh.build(workloads.coplanar_lines_2d(geo_dir + "/nl", seed=2), mesh_eps=0.02)
    else:
        m = meshes[name]
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_MIXED)
    c32, s32 = dev.coeffs(job, m)
    ocoef, oself = orc.fill(m)
    near = _near_mask(m)
    assert near.any() and not near.all()
    cmix = np.where(near, ocoef, c32)
    q = np.random.default_rng(7).standard_normal((3, m.N))
    v = dev.mvp_prec(job, q, fc.FCG_MVP_FP32_ACC64)
    for r in range(3):
        want = orc.mvp(m, cmix, oself, q[r])
        only32 = orc.mvp(m, c32, oself, q[r])
        e = np.max(np.abs(v[r] - want)) / np.max(np.abs(want))
        e32 = np.max(np.abs(v[r] - only32)) / np.max(np.abs(want))
        print(f"{name} rhs {r}: vs FP64-near/FP32-far {e:.2e}, vs all-FP32 coefficients {e32:.2e}", flush=True)
        assert e < 1e-12
        assert e32 > 10 * max(e, 1e-16)  # the near-field correction is there
    dev.release(job)


def test_unconverged_solves_are_flagged(dev, meshes):
    m = meshes["cubes"]
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_MIXED)
    short = dev.solve(job, m, tol=1e-12, precision=fc.FCG_PREC_MIXED, max_refine=1)
    assert not short.converged and short.residual > 1e-12 and short.cycles == 1
    ok = dev.solve(job, m, tol=1e-7, precision=fc.FCG_PREC_MIXED)
    assert ok.converged and ok.residual < 1e-7
    dev.fill(job, fc.FCG_PREC_FP64)
    capped = dev.solve(job, m, tol=1e-10, max_iter=3)
    assert not capped.converged
    assert dev.solve(job, m, tol=1e-6).converged
    dev.release(job)


def test_precond_bit_set_as_the_reference_reads_it(dev, meshes):
    """fcg_solve_opts.precond takes CAutoRefGlobalVars::m_ucPrecondType unchanged:
    JACOBI|BLOCK runs the block preconditioner, JACOBI|SUPER the two-level one,
    HIER is refused."""
    m = meshes["cubes"]
    job = dev.upload(m)
    dev.fill(job)
    blk = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_BLOCK, precond_size=16)
    blk_j = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_BLOCK | fc.FCG_PRECOND_JACOBI, precond_size=16)
    assert blk.iterations == blk_j.iterations
    np.testing.assert_array_equal(blk.cap, blk_j.cap)
    sup = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_SUPER, precond_size=64)
    sup_j = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_SUPER | fc.FCG_PRECOND_JACOBI, precond_size=64)
    assert sup.iterations == sup_j.iterations
    np.testing.assert_array_equal(sup.cap, sup_j.cap)
    with pytest.raises(fc.FasterCapError):
        dev.solve(job, m, tol=1e-8, precond=8)  # AUTOREFINE_PRECOND_HIER
    dev.release(job)


@pytest.mark.parametrize("precision", [fc.FCG_PREC_FP64, fc.FCG_PREC_MIXED])
def test_batch_matches_single_solves(dev, meshes, precision):
    names = ["plates", "cubes", "boxes"]
    ms = [meshes[n] for n in names]
    jobs = [dev.upload(m) for m in ms]
    for j in jobs:
        dev.fill(j, precision)
    alone = [dev.solve(j, m, tol=1e-7, precision=precision) for j, m in zip(jobs, ms)]
    batch = dev.solve_batch(jobs, ms, tol=1e-7, precision=precision)
    # FP64: the same products up to summation order; mixed: FP32 products whose
    # summation order depends on the gather shape (a job's own alone, its rhs
    # class's in a batch), both refined to tol, within the mixed path's 1e-4
    bound = 1e-6 if precision == fc.FCG_PREC_FP64 else 1e-4
    for n, a, b in zip(names, alone, batch):
        err = _c_err(b.cap, a.cap)
        print(f"batch {n}: iters alone {a.iterations} batch {b.iterations} diff {err:.3e}", flush=True)
        assert err < bound
        assert max(abs(x - y) for x, y in zip(a.iterations, b.iterations)) <= 1
    for j in jobs:
        dev.release(j)


def test_engine_staggered_submissions(dev, meshes):
    ms = [meshes["cubes"], meshes["boxes"], meshes["plates"], meshes["cubes"]]
    jobs = [dev.upload(m) for m in ms[:3]]
    jobs.append(jobs[0])  # the same job twice at once: each solve has its own scratch
    for j in jobs[:3]:
        dev.fill(j)
    alone = [dev.solve(j, m, tol=1e-7) for j, m in zip(jobs, ms)]
    eng = dev.engine(tol=1e-7)
    got = {}
    pending = list(range(4))
    step = 0
    while len(got) < 4:
        # one new request every third step
        if pending and step % 3 == 0:
            t = pending.pop(0)
            eng.submit(jobs[t], 100 + t)
        active, ready = eng.step()
        while True:
            r = eng.poll()
            if r is None:
                break
            tag, job, res = r
            got[tag - 100] = (job, res)
        step += 1
        assert step < 10000
        if not pending:
            assert active > 0 or ready > 0 or len(got) == 4
    assert eng.bytes() == 0
    for t in range(4):
        job, res = got[t]
        assert job == jobs[t]
        err = _c_err(res.cap, alone[t].cap)
        print(f"engine tag {t}: diff {err:.3e} iters {res.iterations}", flush=True)
        assert err < 1e-6
    eng.close()
    for j in jobs[:3]:
        dev.release(j)


def test_dielectric_boxes_match_oracle(dev, meshes):
    m = meshes["boxes"]
    job = dev.upload(m)
    dev.fill(job)
    res = dev.solve(job, m, tol=1e-8)
    ocoef, oself = orc.fill(m)
    cap, iters = orc.solve(m, ocoef, oself, tol=1e-8)
    print(f"boxes: N={m.N} gpu iters {res.iterations} oracle {list(iters)}", flush=True)
    assert np.max(np.abs(res.cap - cap) / np.abs(np.diag(cap))[:, None]) < 1e-6
    dev.release(job)


@pytest.mark.parametrize("seed", [0, 1])
def test_2d_lines_match_oracle(dev, geo_dir, seed):
    m = fc.Mesh.build(workloads.coplanar_lines_2d(geo_dir + f"/l{seed}", seed=seed), mesh_eps=0.05)
    job = dev.upload(m)
    dev.fill(job)
    ocoef, oself = orc.fill(m)
    coef, selfc = dev.coeffs(job, m)
    assert np.max(np.abs(coef - ocoef) / np.abs(ocoef)) < 1e-12
    q = np.random.default_rng(seed).standard_normal((3, m.N))
    v = dev.mvp(job, q)
    for r in range(3):
        vo = orc.mvp(m, ocoef, oself, q[r])
        assert np.max(np.abs(v[r] - vo)) / np.max(np.abs(vo)) < 1e-12
    res = dev.solve(job, m, tol=1e-8)
    cap, iters = orc.solve(m, ocoef, oself, tol=1e-8)
    print(f"2D lines seed {seed}: rows {res.cap.shape} gpu iters {res.iterations} oracle {list(iters)}", flush=True)
    assert res.cap.shape == cap.shape == (3, 4)
    assert np.max(np.abs(res.cap - cap) / np.abs(np.diag(cap))[:, None]) < 1e-6
    dev.release(job)


def test_2d_coax_closed_form(dev, geo_dir):
    a, b = 1.0, 2.0
    m = fc.Mesh.build(workloads.coax_2d(geo_dir + "/coax", a=a, b=b, n=96), mesh_eps=0.01)
    job = dev.upload(m)
    dev.fill(job)
    res = dev.solve(job, m, tol=1e-8)
    exact = 1.0 / math.log(b / a)  # in units of 2 pi e0 per metre
    print(f"coax: C11 {res.cap[0, 0]:.6f} exact {exact:.6f}", flush=True)
    assert abs(res.cap[0, 0] - exact) / exact < 1e-3
    assert abs(res.cap[0, 1] + res.cap[0, 0]) / exact < 1e-9  # zero total charge
    dev.release(job)


def test_footprint_reports_bytes(dev, meshes):
    m = meshes["boxes"]
    job = dev.upload(m)
    res64, s64 = dev.footprint(job, fc.FCG_PREC_FP64, iters=32)
    res32, s32 = dev.footprint(job, fc.FCG_PREC_MIXED, iters=32)
    # FP64: 8 B per link; mixed: 4 B per link plus the FP64 near-field side list
    # (8 B coefficient difference, 4 B column, 4 B link index per analytic-branch
    # link, 8 B per row pointer)
    near = int(_near_mask(m).sum())
    assert res64 > 8 * m.L
    assert res32 >= 4 * m.L + 16 * near + 8 * (m.M + 1)
    assert res32 - 4 * m.L - 16 * near - 8 * (m.M + 1) == res64 - 8 * m.L
    assert s64 > 8 * m.N * 32
    dev.release(job)


@pytest.mark.parametrize("which", ["coated_coax_2d", "sphere_in_shell"])
def test_curved_dielectrics_match_oracle(dev, geo_dir, which):
    d = os.path.join(geo_dir, "gpu_" + which)
    if which == "coated_coax_2d":
        m = fc.Mesh.build(workloads.coated_coax_2d(d), mesh_eps=0.05)
    else:
        m = fc.Mesh.build(workloads.sphere_in_shell(d, b=3.0, subdiv=2), mesh_eps=0.05)
    job = dev.upload(m)
    dev.fill(job)
    ocoef, oself = orc.fill(m)
    coef, selfc = dev.coeffs(job, m)
    # field coefficients of nearly tangential pairs cancel to ~1e-7 of the largest,
    # so the bound mixes a relative and an absolute term
    assert np.all(np.abs(coef - ocoef) <= 1e-12 * np.abs(ocoef) + 1e-14 * np.abs(ocoef).max())
    res = dev.solve(job, m, tol=1e-10, max_iter=2000)
    cap, iters = orc.solve(m, ocoef, oself, tol=1e-10)
    print(f"{which}: N={m.N} gpu {res.cap.ravel()} oracle {cap.ravel()} iters {res.iterations} {list(iters)}",
          flush=True)
    assert np.max(np.abs(res.cap - cap) / np.abs(np.diag(cap))[:, None]) < 1e-6
    dev.release(job)
