"""fcg_mvp_device_noalloc: the product on device buffers with its scratch in the
library's static device array gives fcg_mvp_device's result, call after call."""
import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def _products(dev, job, q, k):
    """(fcg_mvp_device, fcg_mvp_device_noalloc) of the [k][N] host array q."""
    dq = DeviceArray.from_host(q)
    v1, v2 = DeviceArray(q.shape), DeviceArray(q.shape)
    v2.fill_nan()
    lib = dev.lib
    dev._check(lib.fcg_mvp_device(dev._h, job, dq.ptr, v1.ptr, k, None), "fcg_mvp_device")
    dev._check(lib.fcg_mvp_device_noalloc(dev._h, job, dq.ptr, v2.ptr, k, None), "fcg_mvp_device_noalloc")
    return v1.to_host(), v2.to_host()


def test_noalloc_product_equals_mvp_device(geo_dir):
    print("no-alloc: plates -m0.05 k=3, then 2D lines -m0.02 k=1", flush=True)
    plates = fc.Mesh.build(workloads.two_plates(geo_dir + "/na_p"), mesh_eps=0.05)
    lines = fc.Mesh.build(workloads.coplanar_lines_2d(geo_dir + "/na_l", seed=5), mesh_eps=0.02)
    dev = fc.Device(0)
    try:
        jp, jl = dev.upload(plates), dev.upload(lines)
        dev.fill(jp)
        dev.fill(jl)
        rng = np.random.default_rng(11)
        # first call: the larger input
        q = rng.standard_normal((3, plates.N))
        v1, v2 = _products(dev, jp, q, 3)
        assert np.isfinite(v2).all()
        np.testing.assert_array_equal(v2, v1)  # same kernels, same order: identical
        co, so = orc.fill(plates)
        ref = np.stack([orc.mvp(plates, co, so, q[r]) for r in range(3)])
        assert np.max(np.abs(v2 - ref)) / np.max(np.abs(ref)) < 1e-12
        # second call: a smaller input on another job (the array keeps the first
        # call's contents; nothing may depend on them)
        q = rng.standard_normal((1, lines.N))
        v1, v2 = _products(dev, jl, q, 1)
        assert np.isfinite(v2).all()
        np.testing.assert_array_equal(v2, v1)
        co, so = orc.fill(lines)
        ref = orc.mvp(lines, co, so, q[0])
        assert np.max(np.abs(v2[0] - ref)) / np.max(np.abs(ref)) < 1e-12
        # a product larger than the static array is refused before any launch
        big = 4096
        dq = DeviceArray((big, plates.N))
        rc = dev.lib.fcg_mvp_device_noalloc(dev._h, jp, dq.ptr, dq.ptr, big, None)
        assert rc == 65  # FCG_INVALID_ARGUMENT
        assert b"no-alloc limit" in dev.lib.fcg_last_error(dev._h)
    finally:
        dev.close()
