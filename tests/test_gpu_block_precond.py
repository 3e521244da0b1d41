"""GPU tests of the block preconditioner (-pb): the device-built block inverses
applied to random vectors against the oracle's, and block-preconditioned solves
(FP64, mixed, batched) against the oracle and against Jacobi, through the C-ABI."""
import os
import subprocess

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads
from test_gpu_cli import _last_matrix

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fastercap_amd", "bin",
                   "fastercap")


@pytest.fixture(scope="module")
def dev():
    d = fc.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def meshes(geo_dir):
    os.makedirs(geo_dir, exist_ok=True)
    return {
        "plates": fc.Mesh.build(workloads.two_plates(os.path.join(geo_dir, "bp_plates")), mesh_eps=0.02),
        "boxes": fc.Mesh.build(workloads.random_boxes(os.path.join(geo_dir, "bp_boxes"), n=6, seed=4), mesh_eps=0.15),
        "coax2d": fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "bp_coax")), mesh_eps=0.05),
    }


def _c_err(cap, ref):
    return float(np.max(np.abs(cap - ref) / np.abs(np.diag(ref))[:, None]))


@pytest.mark.parametrize("bsize", [2, 32, 128])  # 128: blocks above 64 leaves are factored in global memory
@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_block_apply_matches_oracle(dev, meshes, name, bsize):
    m = meshes[name]
    print(f"block apply: {name} N={m.N} b={bsize}", flush=True)
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_FP64)
    bp = orc.BlockPrecond(m, bsize)
    q = np.random.default_rng(5).standard_normal((3, m.N))
    got = dev.block_precond(job, q, bsize)
    for r in range(3):
        ref = bp.apply(q[r])
        np.testing.assert_allclose(got[r], ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
    dev.release(job)


@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_block_solve_matches_oracle_and_jacobi(dev, meshes, name):
    m = meshes[name]
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_FP64)
    coef, selfc = dev.coeffs(job, m)
    blk = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_BLOCK, precond_size=32)
    jac = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_JACOBI)
    cap_o, it_o = orc.solve_block(m, coef, selfc, 32, tol=1e-8)
    e_o, e_j = _c_err(blk.cap, cap_o), _c_err(blk.cap, jac.cap)
    print(f"{name}: block iters {list(blk.iterations)} oracle {list(it_o)} Jacobi {list(jac.iterations)}; "
          f"C vs oracle {e_o:.2e} vs Jacobi {e_j:.2e}", flush=True)
    assert e_o < 1e-6
    assert e_j < 1e-6
    assert max(abs(int(a) - int(b)) for a, b in zip(blk.iterations, it_o)) <= 1
    dev.release(job)


def test_block_mixed_and_batched(dev, meshes):
    names = ["plates", "boxes", "coax2d"]
    ms = [meshes[n] for n in names]
    jobs = [dev.upload(m) for m in ms]
    for j in jobs:
        dev.fill(j, fc.FCG_PREC_FP64)
    ref = [dev.solve(j, m, tol=1e-8, precond=fc.FCG_PRECOND_BLOCK, precond_size=16) for j, m in zip(jobs, ms)]
    batch = dev.solve_batch(jobs, ms, tol=1e-8, precond=fc.FCG_PRECOND_BLOCK, precond_size=16)
    for n, a, b in zip(names, ref, batch):
        assert _c_err(b.cap, a.cap) < 1e-6, n
    for j in jobs:
        dev.fill(j, fc.FCG_PREC_MIXED)
    mixed = dev.solve_batch(jobs, ms, tol=1e-5, precision=fc.FCG_PREC_MIXED, precond=fc.FCG_PRECOND_BLOCK,
                            precond_size=16)
    for n, a, b in zip(names, ref, mixed):
        err = _c_err(b.cap, a.cap)
        print(f"mixed block {n}: C err {err:.2e} iters {list(b.iterations)}", flush=True)
        assert err < 1e-4, n
    for j in jobs:
        dev.release(j)


def test_binary_block_switch(geo_dir):
    path = workloads.two_plates(os.path.join(geo_dir, "bp_cli"))
    runs = {}
    for sw in ("-pj", "-pb32"):
        out = subprocess.run([BIN, path, "-m0.05", "-t1e-8", sw], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        runs[sw] = _last_matrix(out.stdout)[1]
    np.testing.assert_allclose(runs["-pb32"], runs["-pj"], rtol=1e-6)
