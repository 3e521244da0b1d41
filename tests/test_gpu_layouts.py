"""GPU tests of the storage and launch layouts: 16-bit and 32-bit source indices
(jobs of at most / more than 65536 nodes), one gather launch per rhs class and
index width inside a batch, and a job whose rhs count needs the wide lane
shapes.  Products are checked against the oracle, batched solves against the
same jobs solved alone."""
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = fc.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def jobs_meshes(geo_dir):
    os.makedirs(geo_dir, exist_ok=True)
    big = fc.Mesh.build(workloads.crossing_bus(os.path.join(geo_dir, "wide_bus")), mesh_eps=600.0)
    small = fc.Mesh.build(workloads.two_plates(os.path.join(geo_dir, "small_p"), gap=0.2), mesh_eps=0.02)
    many = fc.Mesh.build(workloads.random_boxes(os.path.join(geo_dir, "many_b"), n=20, seed=3,
                                                with_interface=False), mesh_eps=0.2)
    return {"big": big, "small": small, "many": many}


def _c_err(cap, ref):
    return float(np.max(np.abs(cap - ref) / np.abs(np.diag(ref))[:, None]))


def test_index_widths_cover_both_sides(jobs_meshes):
    # the layout switch is at 65536 nodes: one job on each side
    assert jobs_meshes["big"].M > 65536
    assert jobs_meshes["small"].M <= 65536
    assert len(jobs_meshes["many"].conductor_names) >= 17  # rhs stride 32: the widest class


@pytest.mark.parametrize("name", ["big", "small", "many"])
def test_products_match_oracle(dev, jobs_meshes, name):
    m = jobs_meshes[name]
    print(f"layout product: {name} N={m.N} M={m.M} L={m.L}", flush=True)
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_FP64)
    coef, selfc = dev.coeffs(job, m)
    rng = np.random.default_rng(11)
    k = 3
    q = rng.standard_normal((k, m.N))
    v = dev.mvp(job, q)
    for r in range(k):
        vo = orc.mvp(m, coef, selfc, q[r])
        np.testing.assert_allclose(v[r], vo, rtol=1e-12, atol=1e-12 * np.abs(vo).max())
    # mixed: FP32 coefficients and charges against the FP64 product
    dev.fill(job, fc.FCG_PREC_MIXED)
    v32 = dev.mvp_prec(job, q, fc.FCG_MVP_FP32)
    rel = np.abs(v32 - v).max() / np.abs(v).max()
    print(f"  fp32 product rel {rel:.3e}", flush=True)
    assert rel < 2e-5
    dev.release(job)


@pytest.mark.parametrize("precision", [fc.FCG_PREC_FP64, fc.FCG_PREC_MIXED])
def test_batch_over_classes_and_widths_matches_alone(dev, jobs_meshes, precision):
    # rhs strides 8 (bus), 2 (plates) and 32 (20 boxes); 32-bit and 16-bit indices
    # in one batch
    names = ["big", "small", "many", "small"]
    ms = [jobs_meshes[n] for n in names]
    jobs = [dev.upload(m) for m in ms]
    for j in jobs:
        dev.fill(j, precision)
    alone = [dev.solve(j, m, tol=1e-7, precision=precision) for j, m in zip(jobs, ms)]
    batch = dev.solve_batch(jobs, ms, tol=1e-7, precision=precision)
    for n, a, b in zip(names, alone, batch):
        err = _c_err(b.cap, a.cap)
        print(f"batch layout {n}: iters alone {max(a.iterations)} batch {max(b.iterations)} diff {err:.3e}",
              flush=True)
        assert err < 1e-6
    for j in jobs:
        dev.release(j)
