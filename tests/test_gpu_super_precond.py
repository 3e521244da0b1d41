"""GPU tests of the two-level preconditioner (-ps): the device-built element
inverse applied to random vectors against the oracle's, two-level
preconditioned solves (FP64, mixed, batched) against the oracle and against
Jacobi, the unavailable case (at least as many tree roots as the dimension,
solved unpreconditioned as the reference does) and the command line, through
the C-ABI."""
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
        "plates": fc.Mesh.build(workloads.two_plates(os.path.join(geo_dir, "sp_plates")), mesh_eps=0.02),
        "boxes": fc.Mesh.build(workloads.random_boxes(os.path.join(geo_dir, "sp_boxes"), n=6, seed=4), mesh_eps=0.15),
        "coax2d": fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "sp_coax")), mesh_eps=0.05),
        "many": fc.Mesh.build(workloads.random_boxes(os.path.join(geo_dir, "sp_many"), n=20, seed=3,
                                                     with_interface=False), mesh_eps=0.3),
    }


def _c_err(cap, ref):
    return float(np.max(np.abs(cap - ref) / np.abs(np.diag(ref))[:, None]))


@pytest.mark.parametrize("dim", [16, 128, 1024])
@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_super_apply_matches_oracle(dev, meshes, name, dim):
    m = meshes[name]
    sp = orc.SuperPrecond(m, dim)
    print(f"two-level apply: {name} N={m.N} dim={dim} elements={sp.m}", flush=True)
    assert sp.available
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_FP64)
    q = np.random.default_rng(6).standard_normal((3, m.N))
    got = dev.super_precond(job, q, dim)
    for r in range(3):
        ref = sp.apply(q[r])
        np.testing.assert_allclose(got[r], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    dev.release(job)


@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_super_solve_matches_oracle_and_jacobi(dev, meshes, name):
    m = meshes[name]
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_FP64)
    coef, selfc = dev.coeffs(job, m)
    sup = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_SUPER, precond_size=128)
    jac = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_JACOBI)
    cap_o, it_o = orc.solve_super(m, coef, selfc, 128, tol=1e-8)
    e_o, e_j = _c_err(sup.cap, cap_o), _c_err(sup.cap, jac.cap)
    print(f"{name}: two-level iters {list(sup.iterations)} oracle {list(it_o)} Jacobi {list(jac.iterations)}; "
          f"C vs oracle {e_o:.2e} vs Jacobi {e_j:.2e}", flush=True)
    assert e_o < 1e-6
    assert e_j < 1e-6
    assert max(abs(int(a) - int(b)) for a, b in zip(sup.iterations, it_o)) <= 1
    dev.release(job)


def test_super_mixed_and_batched(dev, meshes):
    names = ["plates", "boxes", "coax2d"]
    ms = [meshes[n] for n in names]
    jobs = [dev.upload(m) for m in ms]
    for j in jobs:
        dev.fill(j, fc.FCG_PREC_FP64)
    ref = [dev.solve(j, m, tol=1e-8, precond=fc.FCG_PRECOND_SUPER, precond_size=64) for j, m in zip(jobs, ms)]
    batch = dev.solve_batch(jobs, ms, tol=1e-8, precond=fc.FCG_PRECOND_SUPER, precond_size=64)
    for n, a, b in zip(names, ref, batch):
        assert _c_err(b.cap, a.cap) < 1e-6, n
    for j in jobs:
        dev.fill(j, fc.FCG_PREC_MIXED)
    mixed = dev.solve_batch(jobs, ms, tol=1e-5, precision=fc.FCG_PREC_MIXED, precond=fc.FCG_PRECOND_SUPER,
                            precond_size=64)
    for n, a, b in zip(names, ref, mixed):
        err = _c_err(b.cap, a.cap)
        print(f"mixed two-level {n}: C err {err:.2e} iters {list(b.iterations)}", flush=True)
        assert err < 1e-4, n
    for j in jobs:
        dev.release(j)


def test_super_batched_inverses_match_oracle(dev, meshes):
    """The engine builds the inverses of all starting solves in one launch, its
    jobs in the batch's row-stride order (2D, 3D with 2 and 6 conductors here);
    each job's inverse, applied right after, is the oracle's.  A refill drops
    the inverse and the next solve builds it again to the same C."""
    names = ["coax2d", "plates", "boxes", "many"]
    ms = [meshes[n] for n in names]
    jobs = [dev.upload(m) for m in ms]
    for j in jobs:
        dev.fill(j, fc.FCG_PREC_FP64)
    first = dev.solve_batch(jobs, ms, tol=1e-8, precond=fc.FCG_PRECOND_SUPER, precond_size=64)
    rng = np.random.default_rng(8)
    for n, m, j in zip(names, ms, jobs):
        sp = orc.SuperPrecond(m, 64)
        print(f"batched two-level inverse: {n} N={m.N} elements={sp.m} available={sp.available}", flush=True)
        if not sp.available:
            continue
        q = rng.standard_normal((2, m.N))
        got = dev.super_precond(j, q, 64)
        for r in range(2):
            ref = sp.apply(q[r])
            np.testing.assert_allclose(got[r], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max(), err_msg=n)
    for j in jobs:
        dev.fill(j, fc.FCG_PREC_FP64)
    again = dev.solve_batch(jobs, ms, tol=1e-8, precond=fc.FCG_PRECOND_SUPER, precond_size=64)
    for n, a, b in zip(names, first, again):
        assert _c_err(b.cap, a.cap) < 1e-12, n
        assert list(a.iterations) == list(b.iterations), n
    for j in jobs:
        dev.release(j)


def test_super_unavailable_solves_unpreconditioned(dev, meshes):
    m = meshes["many"]
    roots = int((m.parent < 0).sum())
    assert roots >= 16
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_FP64)
    with pytest.raises(fc.FasterCapError):
        dev.super_precond(job, np.ones(m.N), 16)
    sup = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_SUPER, precond_size=16)
    non = dev.solve(job, m, tol=1e-8, precond=fc.FCG_PRECOND_NONE)
    np.testing.assert_allclose(sup.cap, non.cap, rtol=1e-12, atol=1e-12 * np.abs(non.cap).max())
    assert list(sup.iterations) == list(non.iterations)
    dev.release(job)


def test_binary_two_level_and_auto_switches(geo_dir):
    path = workloads.two_plates(os.path.join(geo_dir, "sp_cli"))
    runs = {}
    for sw in ("-pj", "-ps128", "-ap"):
        out = subprocess.run([BIN, path, "-m0.05", "-t1e-8", sw], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        runs[sw] = _last_matrix(out.stdout)[1]
    np.testing.assert_allclose(runs["-ps128"], runs["-pj"], rtol=1e-6)
    np.testing.assert_allclose(runs["-ap"], runs["-pj"], rtol=1e-6)
