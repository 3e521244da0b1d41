"""Seeded FastCap-format geometry generators for the benchmark / test workloads.

Each generator writes plain FastCap input files (the format of
Solver/Autorefine.cpp:4746-5988) into a directory and returns the path of the
top-level file, so the host builder reads exactly what the reference would.
"""
from __future__ import annotations

import math
import os

import numpy as np

BASE_SEED = 20261203

# Face vertex order of a unit box, as in the reference's own cube example
# (hlp/InputFile/3D/InputFile_3D_ch02.htm).
_UNIT_BOX_FACES = (
    ((1, 1, 0), (1, 0, 0), (1, 0, 1), (1, 1, 1)),
    ((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)),
    ((1, 0, 0), (0, 0, 0), (0, 0, 1), (1, 0, 1)),
    ((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)),
    ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)),
    ((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)),
)


def _fmt(x: float) -> str:
    return repr(float(x))


def _perm(e) -> str:
    """A permittivity in FastCap syntax: real, or complex as 're-jim' / 're+jim'
    (the reference's parser, Solver/Autorefine.cpp:4840-4860; a lossy medium is
    eps' - j eps'')."""
    if isinstance(e, complex):
        return f"{e.real!r}{'-' if e.imag < 0 else '+'}j{abs(e.imag)!r}"
    return repr(float(e))


def box_quads(name: str, lo, hi, scale: float = 1.0):
    """Six 'Q' statements of an axis-aligned box [lo, hi]."""
    lines = []
    for face in _UNIT_BOX_FACES:
        pts = []
        for corner in face:
            p = [(lo[a] if corner[a] == 0 else hi[a]) * scale for a in range(3)]
            pts.append(" ".join(_fmt(c) for c in p))
        lines.append(f"Q {name}  " + "  ".join(pts))
    return lines


def two_plates(outdir: str, gap: float = 0.1, pitch: float | None = None) -> str:
    """Two 1 m x 1 m zero-thickness plates, one quad each.  Without `pitch` they are
    stacked at vertical distance `gap`; with `pitch` they are coplanar at that
    lateral centre pitch."""
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, "plates.txt")
    if pitch is None:
        a = "Q A 0 0 0  1 0 0  1 1 0  0 1 0"
        b = f"Q B 0 0 {gap!r}  1 0 {gap!r}  1 1 {gap!r}  0 1 {gap!r}"
    else:
        a = "Q A 0 0 0  1 0 0  1 1 0  0 1 0"
        b = f"Q B {pitch!r} 0 0  {pitch + 1.0!r} 0 0  {pitch + 1.0!r} 1 0  {pitch!r} 1 0"
    with open(path, "w") as f:
        f.write("* two parallel plates (FastCap format)\n" + a + "\n" + b + "\n")
    return path


def crossing_bus(outdir: str, unit: float = 1e-6) -> str:
    """8 bars of section 1x1 and length 9: 4 along x on z in [0,1] at y = 0,2,4,6,
    4 along y on z in [2,3] at x = 0,2,4,6; each face one quad."""
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, "bus.txt")
    lines = ["* 8-conductor crossing bus, unit 1e-6 m"]
    for i, y in enumerate((0, 2, 4, 6)):
        lines += box_quads(f"x{i}", (-1, y, 0), (8, y + 1, 1), unit)
    for i, x in enumerate((0, 2, 4, 6)):
        lines += box_quads(f"y{i}", (x, -1, 2), (x + 1, 8, 3), unit)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def random_boxes(outdir: str, n: int = 32, seed: int = 0, with_interface: bool = True, span: float = 20.0,
                 min_gap: float = 0.05) -> str:
    """`n` axis-aligned boxes, sides log-uniform in [0.2, 2], centres uniform in a
    span^3 cube above z = 0, overlaps (gap < min_gap) rejected; optionally a planar
    dielectric interface at z = 0 (eps 3.9 below, 1 above) meshed as one quad."""
    rng = np.random.default_rng(BASE_SEED + seed)
    boxes = []
    tries = 0
    while len(boxes) < n:
        tries += 1
        if tries > 100000:
            raise RuntimeError("cannot place boxes")
        side = np.exp(rng.uniform(math.log(0.2), math.log(2.0), size=3))
        c = rng.uniform(0.0, span, size=3)
        lo, hi = c - side / 2, c + side / 2
        if lo[2] < min_gap:
            continue
        ok = True
        for (blo, bhi) in boxes:
            gap = np.maximum(np.maximum(blo - hi, lo - bhi), 0.0)
            if np.all(gap < min_gap):
                ok = False
                break
        if ok:
            boxes.append((lo, hi))
    os.makedirs(outdir, exist_ok=True)
    main = os.path.join(outdir, "boxes.lst")
    lines = ["* random box cluster (FastCap format)"]
    if with_interface:
        lines.append("D plane 1.0 3.9 0 0 0 10 10 30")
    for i, (lo, hi) in enumerate(boxes):
        lines += box_quads(f"b{i}", lo, hi)
    if with_interface:
        lines.append("End")
        lines.append("File plane")
        lines.append(f"Q iface {-0.5 * span!r} {-0.5 * span!r} 0  {1.5 * span!r} {-0.5 * span!r} 0  "
                     f"{1.5 * span!r} {1.5 * span!r} 0  {-0.5 * span!r} {1.5 * span!r} 0")
        lines.append("End")
    with open(main, "w") as f:
        f.write("\n".join(lines) + "\n")
    return main


def coplanar_lines_2d(outdir: str, seed: int = 0, height: float = 1.0, ground_t: float = 1.0) -> str:
    """2D cross-section: 3 rectangular lines (width, spacing, thickness ~ U[0.5,3])
    at `height` over a ground line of width 30 and thickness `ground_t`, each
    rectangle as 4 'S' segments."""
    rng = np.random.default_rng(BASE_SEED + seed)
    w = rng.uniform(0.5, 3.0, size=3)
    s = rng.uniform(0.5, 3.0, size=2)
    t = rng.uniform(0.5, 3.0, size=3)
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, "lines2d.txt")
    lines = ["0 2D coplanar lines over ground (FastCap 2D format)"]

    def rect(name, x0, y0, x1, y1):
        x0, y0, x1, y1 = (float(v) for v in (x0, y0, x1, y1))
        return [
            f"S {name} {x0!r} {y0!r} {x1!r} {y0!r}",
            f"S {name} {x1!r} {y0!r} {x1!r} {y1!r}",
            f"S {name} {x1!r} {y1!r} {x0!r} {y1!r}",
            f"S {name} {x0!r} {y1!r} {x0!r} {y0!r}",
        ]

    total = w.sum() + s.sum()
    x = -total / 2
    for i in range(3):
        lines += rect(f"l{i}", x, ground_t + height, x + w[i], ground_t + height + t[i])
        x += w[i] + (s[i] if i < 2 else 0.0)
    lines += rect("gnd", -15.0, 0.0, 15.0, ground_t)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def coax_2d(outdir: str, a: float = 1.0, b: float = 2.0, n: int = 96) -> str:
    """2D cross-section of a coaxial line: inner circle of radius `a` and outer
    circle of radius `b`, each an `n`-gon of 'S' segments (closed-form
    capacitance 2 pi e0 / ln(b/a) per metre, up to the polygon approximation)."""
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, "coax2d.txt")
    lines = ["0 2D coaxial line (FastCap 2D format)"]
    for name, r in (("inner", a), ("outer", b)):
        for i in range(n):
            t0, t1 = 2 * math.pi * i / n, 2 * math.pi * (i + 1) / n
            lines.append(f"S {name} {r * math.cos(t0)!r} {r * math.sin(t0)!r} {r * math.cos(t1)!r} {r * math.sin(t1)!r}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def batch_uneven(outdir: str, n2d: int = 1024, n3d: int = 256, seed: int = 0, m2d: float = 0.02,
                 m3d_range=(0.0025, 0.01)):
    """Batch of uneven extraction jobs: `n2d` seeded 2D three-line cross-sections
    (coplanar_lines_2d, mesh_eps m2d: about 330-380 segments each) and `n3d`
    seeded coplanar plate pairs at a lateral pitch uniform in [1.5, 3] with mesh_eps
    log-uniform in m3d_range (about 2k-33k panels).  Returns a list of
    (path, mesh_eps, kind) tuples, 2D jobs first."""
    rng = np.random.default_rng(BASE_SEED + 7919 * seed + 1)
    jobs = []
    for i in range(n2d):
        d = os.path.join(outdir, f"l2d_{i}")
        jobs.append((coplanar_lines_2d(d, seed=seed * 100003 + i), m2d, "2d"))
    lo, hi = math.log(m3d_range[0]), math.log(m3d_range[1])
    for i in range(n3d):
        pitch = float(rng.uniform(1.5, 3.0))
        m = float(math.exp(rng.uniform(lo, hi)))
        d = os.path.join(outdir, f"p3d_{i}")
        jobs.append((two_plates(d, pitch=pitch), m, "3d"))
    return jobs


def build_meshes(jobs, threads: int = 16):
    """Build the meshes of (path, mesh_eps, kind) jobs on a thread pool (the host
    builder releases the interpreter lock and keeps no shared state)."""
    from concurrent.futures import ThreadPoolExecutor

    from . import Mesh

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda j: Mesh.build(j[0], mesh_eps=j[1]), jobs))


def _circle_segments(name: str, r: float, n: int):
    lines = []
    for i in range(n):
        t0, t1 = 2 * math.pi * i / n, 2 * math.pi * (i + 1) / n
        lines.append(f"S {name} {r * math.cos(t0)!r} {r * math.sin(t0)!r} {r * math.cos(t1)!r} {r * math.sin(t1)!r}")
    return lines


def coated_coax_2d(outdir: str, a: float = 0.1, b: float = 0.15, c: float = 0.2, eps_in=2.0,
                   eps_out=1.0, n: int = 40) -> str:
    """2D coated coaxial cable (the example of the reference's 2D input-file help):
    inner conductor radius a, dielectric interface at radius b (eps_in inside,
    eps_out outside), outer conductor radius c; each circle an n-gon.  Closed form
    C = C1 C2 / (C1 + C2), C1 = 2 pi e0 eps_in / ln(b/a), C2 = 2 pi e0 eps_out / ln(c/b)."""
    os.makedirs(outdir, exist_ok=True)
    files = {}
    for tag, r in (("inner", a), ("iface", b), ("outer", c)):
        p = os.path.join(outdir, f"{tag}.txt")
        with open(p, "w") as f:
            f.write(f"0 circle radius {r!r}\n" + "\n".join(_circle_segments(tag, r, n)) + "\n")
        files[tag] = os.path.basename(p)
    main = os.path.join(outdir, "coated_coax.lst")
    with open(main, "w") as f:
        f.write("* 2D coated coaxial cable\n")
        f.write(f"C {files['inner']} {_perm(eps_in)}  0.0 0.0\n")
        # the reference point (centre) lies in the region of the first permittivity
        f.write(f"D {files['iface']} {_perm(eps_in)} {_perm(eps_out)}  0.0 0.0  0.0 0.0\n")
        f.write(f"C {files['outer']} {_perm(eps_out)}  0.0 0.0\n")
    return main


def _icosphere(subdiv: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [tuple(c / math.sqrt(1 + t * t) for c in v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                p = [(verts[i][k] + verts[j][k]) / 2 for k in range(3)]
                n = math.sqrt(sum(c * c for c in p))
                verts.append(tuple(c / n for c in p))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return verts, faces


def sphere_in_shell(outdir: str, a: float = 1.0, b: float = 2.0, eps_shell=3.0, subdiv: int = 2) -> str:
    """Conductor sphere of radius a inside a concentric dielectric sphere of radius b
    (eps_shell between them, vacuum outside), both icospheres of 20*4^subdiv 'T'
    panels.  Closed form C = 4 pi e0 / ((1/a - 1/b)/eps_shell + 1/b)."""
    os.makedirs(outdir, exist_ok=True)
    verts, faces = _icosphere(subdiv)

    def write(name, r):
        p = os.path.join(outdir, f"{name}.txt")
        with open(p, "w") as f:
            f.write(f"0 sphere radius {r!r}\n")
            for tri in faces:
                pts = "  ".join(" ".join(repr(r * c) for c in verts[i]) for i in tri)
                f.write(f"T {name} {pts}\n")
        return os.path.basename(p)

    fc = write("ball", a)
    fd = write("shell", b)
    main = os.path.join(outdir, "sphere_shell.lst")
    with open(main, "w") as f:
        f.write("* conductor sphere in a dielectric shell\n")
        f.write(f"C {fc} {_perm(eps_shell)}  0 0 0\n")
        f.write(f"D {fd} {_perm(eps_shell)} 1.0  0 0 0  0 0 0\n")
    return main


def _sized_mesh(make_path, target, m0, exponent=1.6, tries=3):
    """Build a mesh whose panel count is near `target` by adjusting mesh_eps from a
    first guess m0 with the power law N ~ m^-exponent (a few builds)."""
    from . import Mesh

    m = m0
    mesh = Mesh.build(make_path(), mesh_eps=m)
    for _ in range(tries):
        ratio = mesh.N / float(target)
        if 0.7 < ratio < 1.4:
            break
        m = m * ratio ** (1.0 / exponent)
        mesh = Mesh.build(make_path(), mesh_eps=m)
    return mesh, m


def serve_pool(outdir: str, n_jobs: int = 32, seed: int = 0, n_min: int = 1000, n_max: int = 200_000,
               threads: int = 16):
    """Jobs for the serving stream, drawn from the plate, crossing-bus and
    random-box generators: half in vacuum, half over the eps 3.9 interface, with
    target panel counts log-uniform in [n_min, n_max].  Returns a list of
    (mesh, kind, mesh_eps)."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(BASE_SEED + 104729 * seed + 5)
    specs = []
    for i in range(n_jobs):
        target = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        if i % 2 == 1:
            kind = "boxes_interface"
        else:
            kind = ("plates", "bus", "boxes_vacuum")[int(rng.integers(0, 3))]
        specs.append((i, kind, target, int(rng.integers(0, 1 << 30)), float(rng.uniform(0.05, 0.2)),
                      int(rng.choice([8, 16, 32]))))

    def build(spec):
        i, kind, target, sd, gap, nbox = spec
        d = os.path.join(outdir, f"serve_{i}")
        if kind == "plates":
            mk, m0 = (lambda: two_plates(d, gap=gap)), 0.02
        elif kind == "bus":
            mk, m0 = (lambda: crossing_bus(d)), 2000.0
        elif kind == "boxes_vacuum":
            mk, m0 = (lambda: random_boxes(d, n=nbox, seed=sd, with_interface=False)), 0.05
        else:
            mk, m0 = (lambda: random_boxes(d, n=nbox, seed=sd)), 1e-4
        mesh, m = _sized_mesh(mk, target, m0)
        return mesh, kind, m

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(build, specs))
