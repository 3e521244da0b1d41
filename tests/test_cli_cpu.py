"""Command-line parsing and matrix printout of the drop-in CLI (no GPU)."""
import os

import numpy as np
import pytest

from fastercap_amd import cli


def test_defaults_match_reference():
    a = cli.parse_args(["in.lst"])
    # CAutoRefGlobalVars::Reset (Solver/SolverGlobal.cpp:53-90)
    assert (a.mesh_eps, a.eps_ratio, a.tol, a.auto, a.auto_err, a.precond) == (0.1, 1.0, 0.01, False, 0.01, 1)
    assert a.curv_coeff == 3.0


def test_switches():
    a = cli.parse_args(["-a0.005", "-m0.02", "-mc2", "-d0.5", "-t1e-4", "-pn", "-v", "-b", "x.txt", "-gmixed"])
    assert a.auto and a.auto_err == 0.005 and a.mesh_eps == 0.02 and a.curv_coeff == 2.0
    assert a.eps_ratio == 0.5 and a.tol == 1e-4 and a.precond == 0 and a.verbose and a.precision == 1
    assert cli.parse_args(["x", "-mc0.5"]).curv_coeff == 1.0  # clamped as the reference does


@pytest.mark.parametrize("bad", [["x", "-ph"], ["x", "-psx"], ["x", "-ps-3"], ["x", "-pbx"], ["x", "-mabc"], [], ["a", "b"]])
def test_errors(bad):
    with pytest.raises(cli.CommandLineError):
        cli.parse_args(bad)
    assert cli.main(bad) == cli.COMMAND_LINE_ERROR


def test_block_preconditioner_switch():
    a = cli.parse_args(["x", "-pb16"])
    assert a.precond == 2 and a.block_size == 16
    # the reference clamps the dimension to 2..128 with a warning
    assert cli.parse_args(["x", "-pb500"]).block_size == 128
    assert cli.parse_args(["x", "-pb1"]).block_size == 2
    # a later -pj or -pn replaces it
    assert cli.parse_args(["x", "-pb16", "-pj"]).precond == 1


def test_two_level_preconditioner_switch(capsys):
    a = cli.parse_args(["x", "-ps512"])
    assert a.precond == 4 and a.super_dim == 512  # AUTOREFINE_PRECOND_SUPER
    # the reference clamps the dimension to 16..4096 with a warning
    assert cli.parse_args(["x", "-ps5000"]).super_dim == 4096
    assert cli.parse_args(["x", "-ps3"]).super_dim == 16
    # with Jacobi, in either order, only the two-level preconditioner is kept
    capsys.readouterr()
    assert cli.parse_args(["x", "-pj", "-ps64"]).precond == 4
    assert cli.parse_args(["x", "-ps64", "-pj"]).precond == 4
    assert capsys.readouterr().err.count("keeping only the two-levels preconditioner") == 2
    assert cli.parse_args(["x", "-ps64", "-pn"]).precond == 0
    assert cli.parse_args(["x", "-ps64", "-pb8"]).precond == 2


@pytest.mark.parametrize("links,cond,want", [
    (0, 5, (1, 128)), (250_999, 2, (1, 128)), (251_000, 2, (4, 128)),
    (2_500_999, 2, (4, 128)), (2_501_000, 2, (4, 512)), (5_000_999, 2, (4, 512)),
    (5_001_000, 2, (4, 1024)), (10_000_000_000, 1, (4, 1024)),
])
def test_auto_preconditioner_rule(links, cond, want):
    # CSolveCap::AutoSetPrecondType (Solver/SolveCapacitance.cpp:580-608): links are
    # counted in whole thousands before the product with the conductor count
    assert cli.auto_preconditioner(links, cond) == want


def test_matrix_printout():
    cap = np.array([[1.5e-10, -2.5e-11, -1e-11], [-2.5e-11, 1.4e-10, -1.1e-11]])
    txt = cli.format_cap(cap[:, :2], ["g1_A", "g2_B"], 3)
    assert txt.splitlines()[:3] == ["Capacitance matrix is:", "Dimension 2 x 2", "g1_A  1.5e-10 -2.5e-11 "]
    # 2D: scaled by 2 pi e0, balance conductor column dropped
    txt = cli.format_cap(np.array([[1.0, -1.0]]), ["inner", "outer"], 2)
    assert txt.splitlines()[1] == "Dimension 1 x 1"
    assert txt.splitlines()[2] == "inner  5.56325e-11 "


def test_host_build_without_links_and_set_links(tmp_path):
    # links_on_device stops the host builder before its link passes; fch_set_links
    # installs rows found elsewhere and rejects malformed ones
    import ctypes as C

    import fastercap_amd as fc
    from fastercap_amd import _ffi, workloads

    path = workloads.two_plates(str(tmp_path))
    ref = fc.Mesh.build(path, mesh_eps=0.1)
    lib = _ffi.host_lib()
    o = _ffi.HostOptions()
    lib.fch_default_options(C.byref(o))
    o.mesh_eps, o.links_on_device = 0.1, 1
    h = C.c_void_p()
    assert lib.fch_build(path.encode(), C.byref(o), C.byref(h)) == 0
    bare = fc.Mesh(h.value)
    assert (bare.N, bare.M, bare.L) == (ref.N, ref.M, 0)
    p = _ffi.LinkParams()
    assert lib.fch_get_link_params(h, C.byref(p)) == 0
    assert p.dim == 3 and p.link_eps == 0.1 and p.max_side > 0 and p.max_area > 0
    i64 = C.POINTER(C.c_int64)
    i32 = C.POINTER(C.c_int32)
    row, col = ref.row_ptr.copy(), ref.col.copy()
    bad = row.copy()
    bad[1] = -1
    assert lib.fch_set_links(h, bad.ctypes.data_as(i64), col.ctypes.data_as(i32), ref.L, 0, 0.0) != 0
    badc = col.copy()
    badc[0] = ref.M
    assert lib.fch_set_links(h, row.ctypes.data_as(i64), badc.ctypes.data_as(i32), ref.L, 0, 0.0) != 0
    assert lib.fch_set_links(h, row.ctypes.data_as(i64), col.ctypes.data_as(i32), ref.L, 99, 0.5) == 0
    full = fc.Mesh(h.value)
    bare._h = None  # the second wrapper owns the handle
    assert full.L == ref.L and full.stats.links == ref.L and full.stats.t_links_s == 0.5
    assert full.stats.max_level == 99
    np.testing.assert_array_equal(full.col, ref.col)


def test_refined_geometry_file_rebuilds_the_refined_panels(tmp_path):
    # fch_write_fastcap (-o) writes FastCap input: read back without refinement it
    # gives the same panels (count, total area per conductor) as the mesh it came from
    import ctypes as C

    import fastercap_amd as fc
    from fastercap_amd import _ffi, workloads

    path = workloads.two_plates(str(tmp_path))
    m = fc.Mesh.build(path, mesh_eps=0.05)
    lib = _ffi.host_lib()
    assert lib.fch_write_fastcap(C.c_void_p(m._h), path.encode(), b"ref", None) == 0
    lst = path[: path.rfind(".")] + "_ref.lst"
    assert os.path.exists(lst)
    back = fc.Mesh.build(lst, mesh_eps=1e32)
    assert back.N == m.N
    area = lambda mm, g: sum(mm.node_geom()["area"][mm.leaf_node[mm.group_leaf_start[g]:mm.group_leaf_start[g + 1]]])
    for g in range(m.G):
        assert abs(area(back, g) / area(m, g) - 1) < 1e-6


def test_galerkin_switches():
    """-g selects the Galerkin scheme; -r and -i the residual and time/memory dumps."""
    a = cli.parse_args(["x.txt", "-g", "-r", "-i"])
    assert a.galerkin and a.dump_resid and a.dump_time
    assert not cli.parse_args(["x.txt"]).galerkin
    assert "Solution scheme (-g): Galerkin," in cli.format_params(a, 0.01)
    assert "Solution scheme (-g): Collocation," in cli.format_params(cli.parse_args(["x.txt"]), 0.01)
