"""GPU tests of adaptive panel refinement on the device (fcg_refine): the refined
mesh must equal the host builder's exactly (every node record bit for bit, the
tree, the leaf order, the links, the largest accepted relative estimate and the
recursion depth) on 3D and 2D meshes, with planar and curved dielectric
interfaces, and the reference help's two-cubes chain must reproduce its
published output."""
import json
import math
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
        "plates": (workloads.two_plates(os.path.join(geo_dir, "rf_plates")), 0.02),
        "boxes_interface": (workloads.random_boxes(os.path.join(geo_dir, "rf_boxes"), n=6, seed=4), 0.15),
        "coax2d": (workloads.coated_coax_2d(os.path.join(geo_dir, "rf_coax")), 0.02),
        "sphere_shell": (workloads.sphere_in_shell(os.path.join(geo_dir, "rf_sphere")), 0.05),
        "lines2d": (workloads.coplanar_lines_2d(os.path.join(geo_dir, "rf_lines")), 0.02),
        "bus": (workloads.crossing_bus(os.path.join(geo_dir, "rf_bus")), 600.0),
        "boxes16": (workloads.random_boxes(os.path.join(geo_dir, "rf_boxes16"), n=16, seed=0), 1e-5),
    }


def _sorted_cols(m):
    rows = np.repeat(np.arange(m.M, dtype=np.int64), np.diff(m.row_ptr))
    return m.col[np.lexsort((m.col, rows))]


@pytest.mark.parametrize("name", ["plates", "boxes_interface", "coax2d", "sphere_shell", "lines2d", "bus",
                                  "boxes16"])
def test_device_refinement_equals_host(dev, geo_dir, name):
    path, eps = _cases(geo_dir)[name]
    host = fc.Mesh.build(path, mesh_eps=eps)
    devm = fc.Mesh.build(path, mesh_eps=eps, device=dev, refine_on_device=True)
    print(f"refine {name}: N={host.N} M={host.M} L={host.L} discretize host {host.stats.t_discretize_s:.3f}s "
          f"device {devm.stats.t_discretize_s:.3f}s", flush=True)
    assert (devm.N, devm.M, devm.L) == (host.N, host.M, host.L)
    assert devm.nodes_raw.tobytes() == host.nodes_raw.tobytes()
    for f in ("left", "right", "parent", "leaf_lo", "leaf_hi", "leaf_node", "row_ptr", "group_leaf_start",
              "leaf_outperm"):
        np.testing.assert_array_equal(getattr(devm, f), getattr(host, f), err_msg=f)
    np.testing.assert_array_equal(devm.col, _sorted_cols(host))
    assert devm.stats.max_mesh_eps == host.stats.max_mesh_eps
    assert devm.stats.max_level == host.stats.max_level
    assert devm.stats.panels == host.stats.panels and devm.stats.links == host.stats.links


def test_two_cubes_chain_refined_on_device_matches_reference_output(dev, geo_dir):
    from test_golden_cpu import GOLDEN, two_cubes

    gold = json.load(open(os.path.join(GOLDEN, "two_cubes_doc.json")))
    path = two_cubes(geo_dir)
    m = 1e32
    for it in gold["iterations"]:
        mesh = fc.Mesh.build(path, mesh_eps=m, device=dev, refine_on_device=True)
        assert mesh.N == it["panels"]
        assert mesh.L == it["links_computed"]
        assert mesh.stats.max_level == it["max_level"]
        assert float("%.6g" % mesh.stats.max_mesh_eps) == pytest.approx(it["max_mesh_eps"], rel=1e-9)
        job = dev.upload(mesh)
        dev.fill(job, fc.FCG_PREC_FP64)
        res = dev.solve(job, mesh, tol=gold["gmres_tol"])
        dev.release(job)
        assert list(res.iterations) == [it["gmres_iters"]] * 2
        for a, b in zip(res.cap.ravel(), np.array(it["cap"]).ravel()):
            assert float("%.6g" % a) == pytest.approx(b, rel=1e-12), (a, b)
        m = mesh.stats.max_mesh_eps / math.sqrt(2.0)


@pytest.mark.parametrize("name", ["plates", "coax2d", "boxes_interface"])
def test_binary_device_build_prints_host_build_output(geo_dir, name):
    # the command line's default (refinement and links on the device) against
    # -glinks-host (the host builder's passes): the same statistics and matrix
    import subprocess

    binp = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fastercap_amd", "bin",
                        "fastercap")
    path, eps = _cases(geo_dir)[name]
    outs = []
    for extra in ([], ["-glinks-host"]):
        r = subprocess.run([binp, path, f"-m{eps}", "-t1e-6"] + extra, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        keep = [ln for ln in r.stdout.splitlines() if not ln.startswith(("Total", "Input file", "Running"))]
        outs.append("\n".join(keep))
    assert outs[0] == outs[1]
