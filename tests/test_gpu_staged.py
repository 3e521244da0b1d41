"""GPU tests of the staged FP32 gather (charge windows of each 128-row tile in
LDS, sources outside the window read from global memory), through the C-ABI.

Every staged lane shape is compared with the same shape reading all charges
from global memory (same link order and sums, so the two agree to FP32
rounding of identical operations) and with the oracle's FP64 product within
the FP32 product tolerance.  The plate pair has 16382 nodes, more than any
window (4096 / kp rows), so both the LDS and the global branch are taken."""
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
def plates8k(geo_dir):
    return fc.Mesh.build(workloads.two_plates(geo_dir + "/p8k", pitch=2.0), mesh_eps=0.005)


@pytest.fixture(scope="module")
def oracle_fill(plates8k):
    return orc.fill(plates8k)


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_staged_shapes_match_global_shapes_and_oracle(dev, plates8k, oracle_fill, k):
    m = plates8k
    assert m.M > 4096 and m.M <= 65536  # windows smaller than the job, 16-bit indices
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_MIXED)
    ocoef, oself = oracle_fill
    q = np.random.default_rng(700 + k).standard_normal((k, m.N))
    vo = np.stack([orc.mvp(m, ocoef, oself, q[r]) for r in range(k)])
    scale = np.abs(vo).max()
    nvar = dev.set_gather_variant(job, -1)
    assert nvar % 2 == 0
    nshape = nvar // 2
    worst_pair = worst_oracle = 0.0
    for var in range(nshape):
        print(f"staged gather shape {var} k={k}", flush=True)
        dev.set_gather_variant(job, var)
        vg = dev.mvp_prec(job, q, fc.FCG_MVP_FP32)
        dev.set_gather_variant(job, nshape + var)
        vs = dev.mvp_prec(job, q, fc.FCG_MVP_FP32)
        pair = float(np.abs(vs - vg).max() / scale)
        orac = float(np.abs(vs - vo).max() / scale)
        worst_pair, worst_oracle = max(worst_pair, pair), max(worst_oracle, orac)
        assert pair < 1e-6, (var, pair)
        assert orac < 2e-5, (var, orac)
    print(f"k={k}: staged vs global {worst_pair:.2e}, staged vs oracle {worst_oracle:.2e}", flush=True)
    dev.release(job)


def test_staged_default_in_mixed_solves(dev, plates8k):
    """A mixed solve with a staged shape pinned equals the same solve with its
    global-memory shape (same iteration count, C to 1e-9 relative)."""
    m = plates8k
    job = dev.upload(m)
    dev.fill(job, fc.FCG_PREC_MIXED)
    nshape = dev.set_gather_variant(job, -1) // 2
    # the default shape of two right-hand sides: one link per lane, pairs of rhs, 8 rows per wavefront
    dev.set_gather_variant(job, 20)
    a = dev.solve(job, m, tol=1e-6, precision=fc.FCG_PREC_MIXED)
    dev.set_gather_variant(job, nshape + 20)
    b = dev.solve(job, m, tol=1e-6, precision=fc.FCG_PREC_MIXED)
    assert a.converged and b.converged
    assert abs(max(a.iterations) - max(b.iterations)) <= 1
    assert np.max(np.abs(a.cap - b.cap) / np.abs(a.cap)) < 1e-9
    dev.release(job)


@pytest.mark.parametrize("prec", ["fp64", "acc64"])
@pytest.mark.parametrize("k", [1, 2, 5])
def test_staged_shapes_in_fp64_products(dev, plates8k, oracle_fill, prec, k):
    """The staged gather with FP64 charges: FP64 products (FP64 coefficients)
    and the refinement residual's product (FP32 coefficients, FP64 charges and
    sums), each shape against its global-memory form and the oracle."""
    m = plates8k
    job = dev.upload(m)
    mode, fill, tol = {"fp64": (fc.FCG_MVP_FP64, fc.FCG_PREC_FP64, 1e-12),
                       "acc64": (fc.FCG_MVP_FP32_ACC64, fc.FCG_PREC_MIXED, 2e-6)}[prec]
    dev.fill(job, fill)
    ocoef, oself = oracle_fill
    q = np.random.default_rng(900 + k).standard_normal((k, m.N))
    vo = np.stack([orc.mvp(m, ocoef, oself, q[r]) for r in range(k)])
    scale = np.abs(vo).max()
    nshape = dev.set_gather_variant(job, -1) // 2
    for var in range(nshape):
        dev.set_gather_variant(job, var)
        vg = dev.mvp_prec(job, q, mode)
        dev.set_gather_variant(job, nshape + var)
        vs = dev.mvp_prec(job, q, mode)
        pair = float(np.abs(vs - vg).max() / scale)
        orac = float(np.abs(vs - vo).max() / scale)
        assert pair < 1e-13, (var, pair)
        assert orac < tol, (var, orac)
    dev.release(job)
