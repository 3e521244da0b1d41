"""The Galerkin scheme (-g) on the GPU, through the C-ABI against the oracle's
Galerkin restatement (pinned in tests/test_galerkin_cpu.py): the fill of every
branch on conductor and dielectric rows, the self terms, the preconditioners'
blocks, FP64 and mixed solves, the 2D fallback to collocation and the command
lines."""
import io
import math
import os
import subprocess

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import cli, workloads
from test_golden_cpu import two_cubes

pytestmark = pytest.mark.gpu

E0 = 8.854187817e-12
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "fastercap_amd", "bin", "fastercap")


@pytest.fixture(scope="module")
def dev():
    d = fc.Device(0)
    yield d
    d.close()


def _far_dielectric(d):
    """Two stacked conductor plates and a dielectric square 15 m above them: links
    from the square's rows reach the far-field branch."""
    os.makedirs(d, exist_ok=True)
    files = {
        "plateA.txt": "0 plate A\nQ A 0 0 0  1 0 0  1 1 0  0 1 0\n",
        "plateB.txt": "0 plate B\nQ B 0 0 0.1  1 0 0.1  1 1 0.1  0 1 0.1\n",
        "diel.txt": "0 far dielectric square\nQ D 0 0 15  1 0 15  1 1 15  0 1 15\n",
        "geo.lst": "* two plates and a far dielectric square\nC plateA.txt 1.0 0 0 0\nC plateB.txt 1.0 0 0 0\n"
                   "D diel.txt 1.0 3.0 0 0 0 0.5 0.5 -5\n",
    }
    for name, text in files.items():
        with open(os.path.join(d, name), "w") as f:
            f.write(text)
    return os.path.join(d, "geo.lst")


@pytest.fixture(scope="module")
def gmeshes(geo_dir):
    d = os.path.join(geo_dir, "gal")
    os.makedirs(d, exist_ok=True)
    # small interaction coefficients (-d) spread the links over every distance branch
    ms = {
        "cubes": fc.Mesh.build(two_cubes(d), mesh_eps=0.05),
        "shell": fc.Mesh.build(workloads.sphere_in_shell(os.path.join(geo_dir, "gal_shell"), b=2.0, eps_shell=3.0,
                                                         subdiv=2), mesh_eps=0.1, eps_ratio=0.02),
        "far": fc.Mesh.build(_far_dielectric(os.path.join(geo_dir, "gal_far")), mesh_eps=0.1, eps_ratio=0.01),
        "boxes": fc.Mesh.build(workloads.random_boxes(os.path.join(geo_dir, "gal_boxes"), n=4, seed=5), mesh_eps=0.1),
    }
    for m in ms.values():
        m.scheme = fc.FCG_SCHEME_GALERKIN
    return ms


def _branch_counts(m):
    """links per (dielectric row, PotEstimateOpt branch) of a 3D mesh"""
    raw = np.frombuffer(m.nodes_raw.tobytes(), dtype=np.uint8).reshape(m.M, 96)
    c = raw[:, 36:48].copy().view(np.float32).reshape(m.M, 3).astype(np.float64)
    flags = raw[:, 48:52].copy().view(np.uint32).ravel()
    ms = raw[:, 56:64].copy().view(np.float64).ravel()
    rp = np.ctypeslib.as_array(m.desc.row_ptr, shape=(m.M + 1,))
    col = np.ctypeslib.as_array(m.desc.col, shape=(m.L,))
    rows = np.repeat(np.arange(m.M), np.diff(rp))
    ratio = np.linalg.norm(c[rows] - c[col], axis=1) / np.maximum(ms[rows], ms[col])
    br = np.digitize(ratio, [1.0, 2.0, 4.0, 10.0])
    diel = (flags[rows] & 2) != 0
    return {(int(d), int(b)): int(((diel == d) & (br == b)).sum()) for d in (0, 1) for b in range(5)}


@pytest.mark.parametrize("name", ["cubes", "shell", "far", "boxes"])
def test_galerkin_fill_matches_oracle(dev, gmeshes, name):
    m = gmeshes[name]
    job = dev.upload(m)
    try:
        dev.fill(job)
        coef, selfc = dev.coeffs(job, m)
        ocoef, oself = orc.fill(m)
        print(f"galerkin fill {name}: N={m.N} L={m.L} branches {_branch_counts(m)}", flush=True)
        np.testing.assert_allclose(coef, ocoef, rtol=1e-12, atol=1e-12 * np.max(np.abs(ocoef)))
        np.testing.assert_allclose(selfc, oself, rtol=1e-12)
        # the scheme is applied: the conductor self terms are Auto, not the centroid potential
        m.scheme = fc.FCG_SCHEME_COLLOCATION
        try:
            ccoef, cself = orc.fill(m)
        finally:
            m.scheme = fc.FCG_SCHEME_GALERKIN
        assert np.max(np.abs(selfc / cself - 1.0)) > 1e-3
    finally:
        dev.release(job)


def test_galerkin_fill_covers_every_branch(gmeshes):
    """The meshes together take every Galerkin branch on conductor and dielectric rows."""
    tot = {}
    for m in gmeshes.values():
        for k, v in _branch_counts(m).items():
            tot[k] = tot.get(k, 0) + v
    assert all(tot[(d, b)] > 0 for d in (0, 1) for b in range(5)), tot


@pytest.mark.parametrize("name", ["cubes", "shell", "far", "boxes"])
def test_galerkin_solve_matches_oracle(dev, gmeshes, name):
    m = gmeshes[name]
    job = dev.upload(m)
    try:
        dev.fill(job)
        res = dev.solve(job, m, tol=1e-8)
        ocoef, oself = orc.fill(m)
        cap, _ = orc.solve(m, ocoef, oself, tol=1e-8)
        err = np.max(np.abs(res.cap - cap) / np.abs(cap))
        print(f"galerkin solve {name}: max per-entry rel err vs oracle {err:.3e}", flush=True)
        assert err < 1e-6
    finally:
        dev.release(job)


def test_galerkin_preconditioners_and_mixed(dev, gmeshes):
    """Block and two-level preconditioners build their blocks with the Galerkin
    integrals: the same C as Jacobi; the mixed path within 1e-4 per entry of FP64."""
    m = gmeshes["boxes"]
    job = dev.upload(m)
    try:
        dev.fill(job)
        ref = dev.solve(job, m, tol=1e-10).cap
        for pc, size in ((fc.FCG_PRECOND_BLOCK, 32), (fc.FCG_PRECOND_SUPER, 64)):
            got = dev.solve(job, m, tol=1e-10, precond=pc, precond_size=size).cap
            assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-6, pc
        dev.fill(job, fc.FCG_PREC_MIXED)
        mixed = dev.solve(job, m, tol=1e-8, precision=fc.FCG_PREC_MIXED)
        assert mixed.converged
        assert np.max(np.abs(mixed.cap - ref) / np.abs(ref)) < 1e-4
    finally:
        dev.release(job)


def test_galerkin_sphere_capacitance(dev, geo_dir):
    d = os.path.join(geo_dir, "gal_ball")
    workloads.sphere_in_shell(d, subdiv=3)
    path = os.path.join(d, "ball_only.lst")
    with open(path, "w") as f:
        f.write("* conductor sphere\nC ball.txt 1.0 0 0 0\n")
    m = fc.Mesh.build(path, mesh_eps=0.02)
    m.scheme = fc.FCG_SCHEME_GALERKIN
    job = dev.upload(m)
    try:
        dev.fill(job)
        cap = dev.solve(job, m, tol=1e-10).cap[0, 0]
        assert abs(cap / (4 * math.pi * E0) - 1.0) < 0.01
    finally:
        dev.release(job)


def test_scheme_is_collocation_in_2d(dev, geo_dir):
    m = fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "gal_coax")), mesh_eps=0.1)
    j0 = dev.upload(m)
    m.scheme = fc.FCG_SCHEME_GALERKIN
    j1 = dev.upload(m)
    try:
        dev.fill(j0)
        dev.fill(j1)
        c0, s0 = dev.coeffs(j0, m)
        c1, s1 = dev.coeffs(j1, m)
        assert np.array_equal(c0, c1) and np.array_equal(s0, s1)
    finally:
        dev.release(j0)
        dev.release(j1)


def _last_matrix(text):
    lines = text.splitlines()
    start = max(i for i, ln in enumerate(lines) if ln.startswith("Capacitance matrix is:"))
    m, n = [int(x) for x in lines[start + 1].split()[1::2]]
    rows = [lines[start + 2 + i].split() for i in range(m)]
    return np.array([[float(v) for v in r[1:1 + n]] for r in rows])


@pytest.mark.parametrize("python_cli", [False, True])
def test_command_line_galerkin(dev, geo_dir, python_cli):
    """-g prints the Galerkin scheme and solves with it (the matrix equals the
    library's Galerkin solve); on a 2D input it warns and solves by collocation."""
    d = os.path.join(geo_dir, f"gal_cli{int(python_cli)}")
    os.makedirs(d, exist_ok=True)
    path = two_cubes(d)
    args = [path, "-m0.05", "-t1e-8", "-g"]
    if python_cli:
        buf = io.StringIO()
        cli.run(cli.parse_args(args), out=buf)
        text = buf.getvalue()
    else:
        out = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        text = out.stdout
    assert "Solution scheme (-g): Galerkin," in text
    m = fc.Mesh.build(path, mesh_eps=0.05)
    m.scheme = fc.FCG_SCHEME_GALERKIN
    job = dev.upload(m)
    try:
        dev.fill(job)
        ref = dev.solve(job, m, tol=1e-8).cap
    finally:
        dev.release(job)
    np.testing.assert_allclose(_last_matrix(text), ref, rtol=1e-5)
    coax = workloads.coated_coax_2d(os.path.join(d, "coax2d"))
    if python_cli:
        buf = io.StringIO()
        cli.run(cli.parse_args([coax, "-m0.1", "-g"]), out=buf)
        text2 = buf.getvalue()
    else:
        out = subprocess.run([BIN, coax, "-m0.1", "-g"], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        assert "Galerkin scheme yet not supported in 2D" in out.stderr
        text2 = out.stdout
    assert "Solution scheme (-g): Collocation," in text2
