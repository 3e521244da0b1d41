"""Complex permittivities on the GPU (lossy dielectrics, the reference's
AUTOREFINE_CPLX_PERM path), through the C-ABI against the oracle's complex
restatement and the closed forms that pin it (tests/test_complex_cpu.py)."""
import math
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads

pytestmark = pytest.mark.gpu

E0 = 8.854187817e-12


@pytest.fixture(scope="module")
def dev():
    d = fc.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def cmeshes(geo_dir):
    return {
        "shell": fc.Mesh.build(workloads.sphere_in_shell(os.path.join(geo_dir, "gc_shell"), b=3.0,
                                                         eps_shell=3.0 - 1.5j, subdiv=2), mesh_eps=0.05),
        "coax": fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "gc_coax"), eps_in=2.0 - 1.0j,
                                                       eps_out=1.5 - 0.2j), mesh_eps=0.05),
        "lossycoax": fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "gc_lcoax"), eps_in=2.0 - 1.0j),
                                   mesh_eps=0.1),
    }


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("name", ["shell", "coax"])
@pytest.mark.parametrize("k", [1, 3])
def test_complex_product_matches_oracle(dev, cmeshes, name, k):
    """[R -D; D R] on [Re | Im] vectors, k at a time, 2D balance rows included."""
    m = cmeshes[name]
    job = dev.upload(m)
    assert dev.is_complex(job)
    dev.fill(job)
    coef, sre, sim = orc.fill_cplx(m)
    q = np.random.default_rng(k).standard_normal((k, 2 * m.N))
    v = dev.mvp(job, q)
    for r in range(k):
        ref = orc.mvp_cplx(m, coef, sre, sim, q[r])
        e = np.max(np.abs(v[r] - ref)) / np.max(np.abs(ref))
        print(f"{name} k={k} rhs {r}: complex product rel err {e:.2e}", flush=True)
        assert e < 1e-12
    # Jacobi on both halves by the real self term
    p = dev.precond(job, q)
    np.testing.assert_allclose(p, q / np.concatenate([sre, sre])[None, :], rtol=1e-13)
    dev.release(job)


@pytest.mark.parametrize("name", ["shell", "coax"])
def test_complex_solve_matches_oracle(dev, cmeshes, name):
    m = cmeshes[name]
    job = dev.upload(m)
    dev.fill(job)
    res = dev.solve(job, m, tol=1e-10, want_charges=True)
    coef, sre, sim = orc.fill_cplx(m)
    cap, iters = orc.solve_cplx(m, coef, sre, sim, tol=1e-10)
    print(f"{name}: gpu C {res.cap.ravel()} oracle {cap.ravel()} iters {res.iterations} {list(iters)}", flush=True)
    assert np.iscomplexobj(res.cap) and res.cap.shape == cap.shape
    assert _rel(res.cap, cap) < 1e-6
    assert res.converged
    # the charges sum to the matrix with the complex outer permittivities
    w = m.leaf_outperm + 1j * m.leaf_outperm_im
    gs = m.group_leaf_start
    conds = [(gs[g], gs[g + 1]) for g in range(m.G) if not m.group_is_diel[g]]
    for r in range(res.cap.shape[0]):
        for c, (lo, hi) in enumerate(conds):
            s = (w[lo:hi] * res.charges[r, lo:hi]).sum()
            assert abs(s - res.cap[r, c]) <= 1e-12 * abs(res.cap[r, r])
    dev.release(job)


def test_lossy_coax_closed_form(dev, cmeshes):
    """The reference's 2D coated coax with a lossy insulator (eps_in = 2 - 1j):
    C = C1 C2 / (C1 + C2) within 2.5% from the hierarchical solve."""
    m = cmeshes["lossycoax"]
    job = dev.upload(m)
    dev.fill(job)
    res = dev.solve(job, m, tol=1e-10)
    eps_in = 2.0 - 1.0j
    c1 = 2 * math.pi * E0 * eps_in / math.log(1.5)
    c2 = 2 * math.pi * E0 / math.log(2.0 / 1.5)
    exact = c1 * c2 / (c1 + c2)
    got = res.cap[0, 0] * fc.TWO_PI_TIMES_E0
    print(f"lossy coax: {got} exact {exact}", flush=True)
    assert abs(got - exact) < 0.025 * abs(exact)
    dev.release(job)


def test_complex_mixed_and_batched(dev, cmeshes, geo_dir):
    """Mixed precision on complex jobs within 1e-4 per entry of FP64 (tol 1e-8);
    complex and real jobs in one batch each equal their alone solves."""
    plates = fc.Mesh.build(workloads.two_plates(geo_dir + "/gc_p"), mesh_eps=0.05)
    ms = [cmeshes["shell"], plates, cmeshes["coax"]]
    jobs = [dev.upload(m) for m in ms]
    for j in jobs:
        dev.fill(j)
    alone = [dev.solve(j, m, tol=1e-8) for j, m in zip(jobs, ms)]
    batch = dev.solve_batch(jobs, ms, tol=1e-8)
    for m, a, b in zip(ms, alone, batch):
        assert np.iscomplexobj(b.cap) == m.complex
        assert _rel(b.cap, a.cap) < 1e-6
    dev.fill_batch(jobs, fc.FCG_PREC_MIXED)
    mixed = dev.solve_batch(jobs, ms, tol=1e-8, precision=fc.FCG_PREC_MIXED)
    for m, a, b in zip(ms, alone, mixed):
        e = _rel(b.cap, a.cap)
        print(f"mixed vs FP64 ({'complex' if m.complex else 'real'} N={m.N}): per entry {e:.2e}", flush=True)
        assert b.converged and e < 1e-4
    for j in jobs:
        dev.release(j)


def test_complex_preconditioner_rules(dev, cmeshes):
    """Jacobi and none run; the block and two-level preconditioners (which the
    reference applies to the first half of the real form only) are refused."""
    m = cmeshes["coax"]
    job = dev.upload(m)
    dev.fill(job)
    jac = dev.solve(job, m, tol=1e-9)
    non = dev.solve(job, m, tol=1e-9, precond=fc.FCG_PRECOND_NONE)
    assert _rel(non.cap, jac.cap) < 1e-6
    for pc in (fc.FCG_PRECOND_BLOCK, fc.FCG_PRECOND_SUPER):
        with pytest.raises(fc.FasterCapError):
            dev.solve(job, m, tol=1e-9, precond=pc)
    dev.release(job)
