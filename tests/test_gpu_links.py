"""GPU tests of link discovery on the device (fcg_links_discover): the link set
of every mesh must equal the host builder's (the rows compared after sorting each
by source node), on 3D and 2D meshes, with and without dielectric interfaces,
and on a mesh whose traversal needs several work-stack slices; a solve on the
device-discovered mesh must reproduce the host-built one."""
import os

import numpy as np
import pytest

import fastercap_amd as fc
from fastercap_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = fc.Device(0)
    yield d
    d.close()


def _cases(geo_dir):
    os.makedirs(geo_dir, exist_ok=True)
    return {
        "plates": (workloads.two_plates(os.path.join(geo_dir, "lk_plates")), 0.02),
        "boxes_interface": (workloads.random_boxes(os.path.join(geo_dir, "lk_boxes"), n=6, seed=4), 0.15),
        "coax2d": (workloads.coated_coax_2d(os.path.join(geo_dir, "lk_coax")), 0.02),
        "sphere_shell": (workloads.sphere_in_shell(os.path.join(geo_dir, "lk_sphere")), 0.05),
        "bus": (workloads.crossing_bus(os.path.join(geo_dir, "lk_bus")), 600.0),  # 48M links: many slices
        # rows of up to 11177 links: the long-row sort (rows beyond the LDS sort)
        "boxes_long_rows": (workloads.random_boxes(os.path.join(geo_dir, "lk_boxes16"), n=16, seed=0), 1e-5),
    }


def _sorted_rows(m):
    rows = np.repeat(np.arange(m.M, dtype=np.int64), np.diff(m.row_ptr))
    order = np.lexsort((m.col, rows))
    return m.col[order]


@pytest.mark.parametrize("name", ["plates", "boxes_interface", "coax2d", "sphere_shell", "bus", "boxes_long_rows"])
def test_device_links_equal_host_links(dev, geo_dir, name):
    path, eps = _cases(geo_dir)[name]
    host = fc.Mesh.build(path, mesh_eps=eps)
    devm = fc.Mesh.build(path, mesh_eps=eps, device=dev)
    print(f"links {name}: N={host.N} M={host.M} L={host.L} host {host.stats.t_links_s:.3f}s "
          f"device {devm.stats.t_links_s:.3f}s", flush=True)
    assert (devm.N, devm.M, devm.L) == (host.N, host.M, host.L)
    np.testing.assert_array_equal(devm.row_ptr, host.row_ptr)
    np.testing.assert_array_equal(devm.col, _sorted_rows(host))
    assert devm.stats.links == host.stats.links
    assert devm.stats.max_level == host.stats.max_level
    if name == "boxes_long_rows":
        assert np.diff(host.row_ptr).max() > 4096


def test_solve_on_device_links_matches_host_mesh(dev, geo_dir):
    path, eps = _cases(geo_dir)["boxes_interface"]
    res = []
    for kw in ({}, {"device": dev}):
        m = fc.Mesh.build(path, mesh_eps=eps, **kw)
        job = dev.upload(m)
        dev.fill(job, fc.FCG_PREC_FP64)
        res.append(dev.solve(job, m, tol=1e-8).cap)
        dev.release(job)
    np.testing.assert_allclose(res[1], res[0], rtol=1e-12, atol=1e-12 * np.abs(res[0]).max())


def test_two_cubes_chain_on_device_links_matches_reference_output(dev, geo_dir):
    """The reference help's -a0.01 two-cubes run with the links discovered on the
    device: panel and link counts, recursion depth (the traversal's own levels
    count) and the printed capacitance digits of every iteration."""
    import json
    import math

    from test_golden_cpu import GOLDEN, two_cubes

    gold = json.load(open(os.path.join(GOLDEN, "two_cubes_doc.json")))
    path = two_cubes(geo_dir)
    m = 1e32
    for it in gold["iterations"]:
        mesh = fc.Mesh.build(path, mesh_eps=m, device=dev)
        assert mesh.N == it["panels"]
        assert mesh.L == it["links_computed"]
        assert mesh.stats.max_level == it["max_level"]
        job = dev.upload(mesh)
        dev.fill(job, fc.FCG_PREC_FP64)
        res = dev.solve(job, mesh, tol=gold["gmres_tol"])
        dev.release(job)
        assert list(res.iterations) == [it["gmres_iters"]] * 2
        for a, b in zip(res.cap.ravel(), np.array(it["cap"]).ravel()):
            assert float("%.6g" % a) == pytest.approx(b, rel=1e-12), (a, b)
        m = mesh.stats.max_mesh_eps / math.sqrt(2.0)


def _upload_on_device_links(dev, path, eps):
    """Tree from the host builder, links discovered on the device and kept
    there, job created from them (fcg_job_upload_links)."""
    import ctypes as C

    from fastercap_amd import _ffi

    lib = _ffi.host_lib()
    o = _ffi.HostOptions()
    lib.fch_default_options(C.byref(o))
    o.mesh_eps, o.links_on_device = eps, 1
    h = C.c_void_p()
    assert lib.fch_build(path.encode(), C.byref(o), C.byref(h)) == 0
    tree = fc.Mesh(h.value)
    p = _ffi.LinkParams()
    assert lib.fch_get_link_params(h, C.byref(p)) == 0
    n, lev = C.c_int64(), C.c_int32()
    dev._check(dev.lib.fcg_links_discover(dev._h, C.byref(tree.desc), C.byref(p), C.byref(n), C.byref(lev)),
               "fcg_links_discover")
    job = C.c_int()
    dev._check(dev.lib.fcg_job_upload_links(dev._h, C.byref(tree.desc), C.byref(job)), "fcg_job_upload_links")
    return tree, job.value, n.value


@pytest.mark.parametrize("name", ["plates", "coax2d", "sphere_shell", "bus", "boxes_long_rows"])
def test_device_layout_equals_host_layout(dev, geo_dir, name):
    import ctypes as C

    path, eps = _cases(geo_dir)[name]
    host = fc.Mesh.build(path, mesh_eps=eps)
    ja = dev.upload(host)
    tree, jb, L = _upload_on_device_links(dev, path, eps)
    assert L == host.L
    rng = np.random.default_rng(17)
    q = rng.standard_normal((3, host.N))
    for prec in (fc.FCG_PREC_FP64, fc.FCG_PREC_MIXED):
        dev.fill(ja, prec)
        dev.fill(jb, prec)
        mp = fc.FCG_MVP_FP64 if prec == fc.FCG_PREC_FP64 else fc.FCG_MVP_FP32
        va, vb = dev.mvp_prec(ja, q, mp), dev.mvp_prec(jb, q, mp)
        # the same stored layout and fill order: the same bits
        np.testing.assert_array_equal(va, vb)
    dev.fill(ja, fc.FCG_PREC_FP64)
    dev.fill(jb, fc.FCG_PREC_FP64)
    ca, sa = dev.coeffs(ja, host)
    cb = np.empty(L)
    sb = np.empty(host.N)
    dp = C.POINTER(C.c_double)
    dev._check(dev.lib.fcg_download_coeffs(dev._h, jb, cb.ctypes.data_as(dp), sb.ctypes.data_as(dp)),
               "fcg_download_coeffs")
    # device-layout jobs return each row by source node id
    rows = np.repeat(np.arange(host.M, dtype=np.int64), np.diff(host.row_ptr))
    order = np.lexsort((host.col, rows))
    np.testing.assert_array_equal(cb, ca[order])
    np.testing.assert_array_equal(sb, sa)
    ra = dev.solve(ja, host, tol=1e-8)
    rb = dev.solve(jb, host, tol=1e-8)
    np.testing.assert_array_equal(ra.cap, rb.cap)
    dev.release(ja)
    dev.release(jb)
