"""GPU parity: the gfx950 fill, hierarchical product, Jacobi and GMRES solve,
called through the C-ABI, against the oracle on the same meshes."""
import json
import math
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads
from test_golden_cpu import two_cubes

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def dev():
    d = fc.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def plates(geo_dir):
    mesh = fc.Mesh.build(workloads.two_plates(geo_dir), mesh_eps=0.02)
    return mesh


def _filled(dev, mesh):
    job = dev.upload(mesh)
    dev.fill(job)
    return job


def test_fill_matches_oracle(dev, plates):
    print(f"fill parity: plates N={plates.N} L={plates.L}", flush=True)
    job = _filled(dev, plates)
    coef, selfc = dev.coeffs(job, plates)
    ocoef, oself = orc.fill(plates)
    # FP64 with device libm: per-coefficient relative difference <= 1e-12
    assert np.max(np.abs(coef - ocoef) / np.abs(ocoef)) < 1e-12
    assert np.max(np.abs(selfc - oself) / np.abs(oself)) < 1e-12
    dev.release(job)


@pytest.mark.parametrize("k", [1, 3, 9, 12])
def test_mvp_matches_oracle(dev, plates, k):
    print(f"mvp parity: k={k}", flush=True)
    job = _filled(dev, plates)
    ocoef, oself = orc.fill(plates)
    rng = np.random.default_rng(k)
    q = rng.standard_normal((k, plates.N))
    v = dev.mvp(job, q)
    for r in range(k):
        vo = orc.mvp(plates, ocoef, oself, q[r])
        assert np.max(np.abs(v[r] - vo)) / np.max(np.abs(vo)) < 1e-12
    dev.release(job)


@pytest.mark.parametrize("k", [1, 2, 5, 8, 17])
def test_every_gather_variant_matches_oracle(dev, plates, k):
    job = _filled(dev, plates)
    ocoef, oself = orc.fill(plates)
    q = np.random.default_rng(100 + k).standard_normal((k, plates.N))
    vo = np.stack([orc.mvp(plates, ocoef, oself, q[r]) for r in range(k)])
    nvar = dev.set_gather_variant(job, -1)
    for var in range(nvar):
        print(f"gather variant {var} k={k}", flush=True)
        dev.set_gather_variant(job, var)
        v = dev.mvp(job, q)
        assert np.max(np.abs(v - vo)) / np.max(np.abs(vo)) < 1e-12, var
    best, ms = dev.autotune(job, k)
    assert 0 <= best < nvar and ms > 0
    dev.release(job)


def test_mvp_unit_columns_match_compressed_matrix(dev, geo_dir):
    mesh = fc.Mesh.build(two_cubes(geo_dir), mesh_eps=0.02)
    print(f"mvp columns: cubes N={mesh.N}", flush=True)
    job = _filled(dev, mesh)
    ocoef, oself = orc.fill(mesh)
    P = orc.compressed_dense(mesh, ocoef, oself)
    E = np.eye(mesh.N)[:8]
    V = dev.mvp(job, E)
    for i in range(8):
        np.testing.assert_allclose(V[i], P[:, i], rtol=1e-12, atol=1e-12 * np.abs(P[:, i]).max())
    dev.release(job)


def test_precond_is_jacobi(dev, plates):
    job = _filled(dev, plates)
    _, oself = orc.fill(plates)
    q = np.random.default_rng(5).standard_normal((2, plates.N))
    pq = dev.precond(job, q)
    np.testing.assert_allclose(pq, q / oself[None, :], rtol=1e-15)
    dev.release(job)


def test_solve_matches_oracle_fp64(dev, plates):
    job = _filled(dev, plates)
    res = dev.solve(job, plates, tol=1e-8)
    ocoef, oself = orc.fill(plates)
    cap, iters = orc.solve(plates, ocoef, oself, tol=1e-8)
    print(f"solve: gpu iters {res.iterations} oracle {list(iters)}", flush=True)
    assert np.max(np.abs(res.cap - cap) / np.abs(cap)) < 1e-6
    dev.release(job)


def test_two_cubes_auto_chain_on_gpu(dev, geo_dir):
    gold = json.load(open(os.path.join(HERE, "golden", "two_cubes_doc.json")))
    path = two_cubes(geo_dir)
    m = 1e32
    for it in gold["iterations"]:
        mesh = fc.Mesh.build(path, mesh_eps=m)
        job = _filled(dev, mesh)
        res = dev.solve(job, mesh, tol=gold["gmres_tol"])
        assert res.iterations == [it["gmres_iters"]] * 2
        np.testing.assert_allclose(res.cap, np.array(it["cap"]), rtol=1e-5)
        dev.release(job)
        m = mesh.stats.max_mesh_eps / math.sqrt(2.0)
