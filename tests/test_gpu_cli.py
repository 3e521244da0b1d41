"""The command-line drop-ins on the GPU: the reference help's documented
`FasterCap cubes.txt -a0.01` run (hlp/Output/Output_ch01.htm) through the
`fastercap` binary and the Python mirror, and a 2D run through the binary."""
import io
import json
import os
import subprocess

import numpy as np
import pytest

from fastercap_amd import cli, workloads
from test_golden_cpu import two_cubes

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "fastercap_amd", "bin", "fastercap")


def _last_matrix(text):
    lines = text.splitlines()
    start = max(i for i, l in enumerate(lines) if l.startswith("Capacitance matrix is:"))
    m, n = [int(x) for x in lines[start + 1].split()[1::2]]
    rows = [lines[start + 2 + i].split() for i in range(m)]
    return [r[0] for r in rows], np.array([[float(v) for v in r[1:1 + n]] for r in rows])


def test_binary_auto_two_cubes(geo_dir):
    d = os.path.join(geo_dir, "cli_cubes")
    os.makedirs(d, exist_ok=True)
    path = two_cubes(d)
    gold = json.load(open(os.path.join(HERE, "golden", "two_cubes_doc.json")))
    print(f"running {BIN} {path} -a0.01", flush=True)
    out = subprocess.run([BIN, path, "-a0.01"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    names, cap = _last_matrix(out.stdout)
    assert names == ["g1_mycube", "g2_mycube"]  # the help's printout (hlp/Output/Output_ch01.htm:65-66)
    assert out.stdout.count("Iteration number #") == len(gold["iterations"])
    # every iteration's "Solve statistics" block as the reference help prints it
    lines = out.stdout.splitlines()
    blocks = [[m for m in lines[i + 1:i + 8]] for i, ln in enumerate(lines) if ln == "Solve statistics:"]
    assert [b for b in blocks] == [it["stats_lines"] for it in gold["iterations"]]
    np.testing.assert_allclose(cap, np.array(gold["iterations"][-1]["cap"]), rtol=1e-5)


def test_python_cli_matches_binary(geo_dir):
    d = os.path.join(geo_dir, "cli_cubes2")
    os.makedirs(d, exist_ok=True)
    path = two_cubes(d)
    buf = io.StringIO()
    mesh, res = cli.run(cli.parse_args([path, "-m0.05", "-t1e-6"]), out=buf)
    out = subprocess.run([BIN, path, "-m0.05", "-t1e-6"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    _, cap_bin = _last_matrix(out.stdout)
    _, cap_py = _last_matrix(buf.getvalue())
    np.testing.assert_allclose(cap_py, cap_bin, rtol=1e-5)


def test_binary_2d_coated_coax(geo_dir):
    path = workloads.coated_coax_2d(os.path.join(geo_dir, "cli_coax"))
    out = subprocess.run([BIN, path, "-m0.1", "-t1e-8"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    names, cap = _last_matrix(out.stdout)
    assert cap.shape == (1, 1)  # balance conductor's column dropped, as the reference prints 2D results
    assert abs(cap[0, 0] / 113.4e-12 - 1.0) < 0.025


def _complex_matrix(text):
    lines = text.splitlines()
    start = max(i for i, l in enumerate(lines) if l.startswith("Capacitance matrix is:"))
    m, n = [int(x) for x in lines[start + 1].split()[1::2]]
    rows = [lines[start + 2 + i].split() for i in range(m)]
    return np.array([[complex(v) for v in r[1:1 + n]] for r in rows])


@pytest.mark.parametrize("python_cli", [False, True])
def test_lossy_coax_complex_printout(geo_dir, python_cli):
    """A lossy insulator ('2.0-j1.0' in the input file): the matrix is printed
    as re+imj (OutputCapMtx's complex format) in F/m, the CSV likewise, and the
    value is the closed form's within 2.5%."""
    import math

    d = os.path.join(geo_dir, f"cli_lossy_{int(python_cli)}")
    path = workloads.coated_coax_2d(d, eps_in=2.0 - 1.0j)
    args = [path, "-m0.1", "-t1e-8", "-e"]
    if python_cli:
        buf = io.StringIO()
        cli.run(cli.parse_args(args), out=buf)
        text = buf.getvalue()
    else:
        out = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        text = out.stdout
    cap = _complex_matrix(text)
    e0 = 8.854187817e-12
    c1 = 2 * math.pi * e0 * (2.0 - 1.0j) / math.log(1.5)
    c2 = 2 * math.pi * e0 / math.log(2.0 / 1.5)
    exact = c1 * c2 / (c1 + c2)
    assert cap.shape == (1, 1)
    assert abs(cap[0, 0] - exact) < 0.025 * abs(exact)
    base = path[: path.rfind(".")]
    csv = [complex(x.strip()) for x in open(base + ".csv").read().split(",")]
    assert abs(csv[0] - cap[0, 0]) <= 1e-5 * abs(cap[0, 0])


def test_binary_rejects_unsupported_switch(geo_dir):
    out = subprocess.run([BIN, "x.txt", "-ph"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 64


def _tri_area(v):
    a, b, c = np.array(v[0:3]), np.array(v[3:6]), np.array(v[6:9])
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a))


@pytest.mark.parametrize("python_cli", [False, True])
def test_file_outputs_csv_geometry_charges(geo_dir, python_cli):
    # -e writes the printed matrix to <input>.csv; -o the refined geometry; -c each
    # conductor solve's charge densities, whose area integral over conductor j's
    # panels is the matrix entry (row of the solve, column j; outer permittivity 1)
    d = os.path.join(geo_dir, f"cli_out_{int(python_cli)}")
    os.makedirs(d, exist_ok=True)
    path = two_cubes(d)
    args = [path, "-m0.05", "-t1e-8", "-e", "-o", "-c"]
    if python_cli:
        buf = io.StringIO()
        cli.run(cli.parse_args(args), out=buf)
        text = buf.getvalue()
    else:
        r = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        text = r.stdout
    names, cap = _last_matrix(text)
    base = path[: path.rfind(".")]
    csv = np.array([[float(x) for x in ln.split(",")] for ln in open(base + ".csv").read().splitlines()])
    np.testing.assert_allclose(csv, cap, rtol=1e-5)
    assert os.path.exists(base + "_ref.lst")
    for i, name in enumerate(names):
        lst = base + "_" + name + ".lst"
        assert os.path.exists(lst), lst
        for j, other in enumerate(names):
            total = 0.0
            for ln in open(base + "_" + other + "_" + name + "_0.txt"):
                if not ln.startswith("T "):
                    continue
                f = ln.split()
                total += float(f[-1]) * _tri_area([float(x) for x in f[2:11]])
            assert abs(total - cap[i, j]) <= 1e-5 * abs(cap[i, i]), (name, other, total, cap[i, j])


def _run_cli(args, python_cli):
    if python_cli:
        buf = io.StringIO()
        cli.run(cli.parse_args(args), out=buf)
        return buf.getvalue()
    out = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout


@pytest.mark.parametrize("python_cli", [False, True])
def test_residual_and_time_dump(geo_dir, python_cli):
    """-r (SolveCapacitance.cpp:2374-2376, 2511-2513): each GMRES step number is
    preceded by the residual before it (|r0| first, then |g_i| / |b|, which
    cannot grow), and the final residual ends the line; -i (:704-720): the phase
    times and the device memory by category follow the statistics.  Neither
    changes the matrix or the step count."""
    d = os.path.join(geo_dir, f"cli_dump_{int(python_cli)}")
    os.makedirs(d, exist_ok=True)
    path = two_cubes(d)
    args = [path, "-m0.05", "-t1e-6"]
    print(f"cli -r -i python={python_cli} {path}", flush=True)
    plain = _run_cli(args, python_cli)
    dump = _run_cli(args + ["-r", "-i"], python_cli)
    np.testing.assert_allclose(_last_matrix(dump)[1], _last_matrix(plain)[1], rtol=1e-12)
    pl = [ln.split()[2:] for ln in plain.splitlines() if ln.startswith("GMRES Iteration:")]
    dl = [ln.split()[2:] for ln in dump.splitlines() if ln.startswith("GMRES Iteration:")]
    assert len(pl) == len(dl) == 2
    for p, t in zip(pl, dl):
        n = len(p)
        assert n > 3 and len(t) == 2 * n + 1
        assert [int(x) for x in t[1:2 * n:2]] == list(range(n))
        res = [float(x) for x in t[0:2 * n:2]]
        assert res[0] > 0.0
        rel = np.array(res[1:])  # printed to 3 significant digits
        assert np.all(rel[1:] <= rel[:-1] * (1 + 5e-3))
        assert 0.0 <= float(t[-1]) < 5e-4  # "%.3f" of a residual below the tolerance
    lines = dump.splitlines()
    k = max(i for i, ln in enumerate(lines) if ln.startswith("Max Mesh relative refinement value"))
    block = lines[k + 1:k + 14]
    keys = ["Time for reading input file", "Time for building super hierarchy", "Time for discretization",
            "Time for building potential matrix", "Time for precond calculation", "Time for gmres solving",
            "Memory allocated for panel hierarchy", "Memory allocated for links structure",
            "Memory allocated for conductor list", "Memory allocated for Gmres",
            "Memory allocated for preconditioner", "Memory allocated for charge vectors",
            "Memory allocated for hierarchical multiplication"]
    assert [ln.split(":")[0] for ln in block] == keys
    vals = {ln.split(":")[0]: float(ln.split(": ")[1].split()[0].rstrip("s")) for ln in block}
    assert all(vals[k] >= 0.0 for k in keys[:6])
    assert vals["Time for gmres solving"] > 0.0
    for k in keys[6:]:
        if k != "Memory allocated for preconditioner":  # Jacobi: the self terms only
            assert vals[k] > 0, k
    assert "Time for reading input file" not in plain


def test_residual_history_matches_oracle(geo_dir):
    """The first value of the residual history is |b - A x0| with x0 = P b (Jacobi),
    as gmresPrecondSFastAll computes it before the loop (SolveCapacitance.cpp:
    2302-2323), checked against the oracle's product; the rest are the GMRES
    estimates |g_i| / |b|: non-increasing, one per step, the last below tol."""
    import fastercap_amd as fc
    import oracle_lib as orc

    path = workloads.two_plates(os.path.join(geo_dir, "resid_hist"))
    mesh = fc.Mesh.build(path, mesh_eps=0.05)
    dev = fc.Device(0)
    try:
        job = dev.upload(mesh)
        dev.fill(job)
        print(f"residual history: plates N={mesh.N}", flush=True)
        res = dev.solve(job, mesh, tol=1e-8, precond=fc.FCG_PRECOND_JACOBI, want_residuals=True)
        ocoef, oself = orc.fill(mesh)
        conds = [g for g in range(mesh.G) if not mesh.group_is_diel[g]]
        assert len(res.residuals) == len(conds)
        for q, g in enumerate(conds):
            b = np.zeros(mesh.N)
            b[mesh.group_leaf_start[g]:mesh.group_leaf_start[g + 1]] = 1.0
            x0 = b / oself
            r0 = b - orc.mvp(mesh, ocoef, oself, x0)
            h = res.residuals[q]
            assert len(h) == res.iterations[q] + 1
            assert abs(h[0] / np.linalg.norm(r0) - 1.0) < 1e-10
            assert np.all(np.diff(h[1:]) <= 1e-14 * h[1])
            assert h[-1] < 1e-8 and h[-2] >= 1e-8
    finally:
        dev.close()
