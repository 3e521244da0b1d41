"""Python mirror of the `fastercap` command line (fastercap_amd/csrc/host/cli.cpp).

Options and defaults follow FasterCapConsole::ParseCmdLine (FasterCapConsole.cpp:136-357)
and CAutoRefGlobalVars::Reset (Solver/SolverGlobal.cpp:53-90); the -a loop follows
CSolveCap::SolveMain (Solver/SolveCapacitance.cpp:294-477) and the matrix output
CSolveCap::OutputCapMtx (:724-809).

    python -m fastercap_amd.cli cubes.txt -a0.01
"""
from __future__ import annotations

import math
import sys
import time
from dataclasses import dataclass

import numpy as np

COMMAND_LINE_ERROR = 64
MAX_AUTO_ITERATIONS = 128  # SOLVE_MAX_AUTO_ITERATIONS
AUTO_INCREMENT = 1.1       # SOLVE_AUTO_INCREMENT_FACTOR


class CommandLineError(ValueError):
    pass


@dataclass
class Args:
    file: str = ""
    mesh_eps: float = 0.1
    eps_ratio: float = 1.0
    curv_coeff: float = 3.0
    tol: float = 0.01
    auto: bool = False
    auto_err: float = 0.01
    auto_precond: bool = False
    precond: int = 1  # FCG_PRECOND_JACOBI
    block_size: int = 32  # -pb<n> (SolverGlobal.cpp:60)
    super_dim: int = 128  # -ps<n> (SolverGlobal.cpp:59)
    jacobi_set: bool = False  # -pj given (kept only without -ps, SolveCapacitance.cpp:122-125)
    precision: int = 0  # FCG_PREC_FP64
    verbose: bool = False
    device: int = 0
    host_links: bool = False  # -glinks-host: refinement and link discovery on the host builder
    out_csv: bool = False  # -e
    out_geo: bool = False  # -o
    out_charge: bool = False  # -c
    dump_resid: bool = False  # -r
    dump_time: bool = False  # -i
    galerkin: bool = False  # -g (3D)
    stop: bool = False  # -b? / -bv


def _num(s: str, opt: str) -> float:
    try:
        return float(s)
    except ValueError:
        raise CommandLineError(f"bad value in parameter '{opt}'") from None


def parse_args(argv) -> Args:
    a = Args()
    for opt in argv:
        if not opt.startswith("-"):
            if a.file:
                raise CommandLineError(f"more than one input file ('{a.file}', '{opt}')")
            a.file = opt
        elif opt in ("-b?", "-bv"):
            a.stop = True
        elif opt == "-b":
            pass
        elif opt == "-v":
            a.verbose = True
        elif opt == "-ap":
            a.auto_precond = True
        elif opt == "-pj":
            if a.precond != 4:  # FCG_PRECOND_SUPER
                a.precond = 1
            a.jacobi_set = True
        elif opt == "-pn":
            a.precond = 0
            a.jacobi_set = False
        elif opt.startswith("-pb"):
            # the reference adds the block preconditioner to the current choice and then
            # runs a flexible GMRES it reports as unsupported; here -pb selects the block
            # preconditioner alone (as -pn -pb<n> there)
            v = _num(opt[3:], opt)
            if v != int(v) or v < 0:
                raise CommandLineError(f"bad preconditioner dimension '{opt[3:]}'")
            a.precond = 2  # FCG_PRECOND_BLOCK
            a.block_size = int(v)
            if a.block_size > 128:
                print("Warning: maximum block preconditioner dimension is 128, using 128", file=sys.stderr)
                a.block_size = 128
            elif a.block_size < 2:
                print("Warning: minimum block preconditioner dimension is 2, using 2", file=sys.stderr)
                a.block_size = 2
        elif opt.startswith("-ps"):
            # the two-level preconditioner replaces Jacobi (SolveCapacitance.cpp:122-125)
            # and, here, a block preconditioner given before it
            v = _num(opt[3:], opt)
            if v != int(v) or v < 0:
                raise CommandLineError(f"bad preconditioner dimension '{opt[3:]}'")
            a.precond = 4  # FCG_PRECOND_SUPER
            a.super_dim = int(min(v, 1e9))
            if a.super_dim > 4096:
                print("Warning: maximum two-levels preconditioner dimension is 4096, using 4096", file=sys.stderr)
                a.super_dim = 4096
            elif a.super_dim < 16:
                print("Warning: minimum two-levels preconditioner dimension is 16, using 16", file=sys.stderr)
                a.super_dim = 16
        elif opt == "-glinks-host":
            a.host_links = True
        elif opt == "-e":
            a.out_csv = True
        elif opt == "-o":
            a.out_geo = True
        elif opt == "-c":
            a.out_charge = True
        elif opt == "-r":
            a.dump_resid = True
        elif opt == "-i":
            a.dump_time = True
        elif opt == "-gmixed":
            a.precision = 1
        elif opt == "-g":
            a.galerkin = True
        elif opt.startswith("-gdev"):
            a.device = int(_num(opt[5:], opt))
        elif opt.startswith("-mc"):
            a.curv_coeff = _num(opt[3:], opt)
        elif opt.startswith("-m"):
            a.mesh_eps = _num(opt[2:], opt)
        elif opt.startswith("-d"):
            a.eps_ratio = _num(opt[2:], opt)
        elif opt.startswith("-t"):
            a.tol = _num(opt[2:], opt)
        elif opt.startswith("-a"):
            a.auto = True
            if len(opt) > 2:
                a.auto_err = _num(opt[2:], opt)
        else:
            raise CommandLineError(f"unsupported parameter '{opt}' (this build covers the solve path only)")
    if a.precond == 4 and a.jacobi_set:
        print("Warning: both two-levels preconditioner and Jacobi preconditioner specified, keeping only the "
              "two-levels preconditioner", file=sys.stderr)
    if not a.file and not a.stop:
        raise CommandLineError("no input file")
    if a.curv_coeff < 1.0:
        a.curv_coeff = 1.0
    return a


def _entry(x, scale):
    """One matrix entry as the reference prints it: %g, complex as %g%+gj."""
    if np.iscomplexobj(x):
        return "%g%+gj" % (x.real * scale, x.imag * scale)
    return "%g" % (x * scale)


def format_cap(cap: np.ndarray, names, dim: int, verbose: bool = False) -> str:
    """The reference's matrix printout (Solver/SolveCapacitance.cpp:724-809): 2D
    values in F/m with the balance conductor's column dropped unless verbose,
    complex entries as re+imj."""
    from . import TWO_PI_TIMES_E0

    M, N = cap.shape
    if dim == 2 and not verbose and N > 1:
        N -= 1
    scale = TWO_PI_TIMES_E0 if dim == 2 else 1.0
    lines = ["Capacitance matrix is:", f"Dimension {M} x {N}"]
    for i in range(M):
        lines.append(f"{names[i]}  " + "".join(_entry(cap[i, j], scale) + " " for j in range(N)))
    return "\n".join(lines) + "\n"


def auto_preconditioner(mutual_links: int, conductors: int):
    """(precond, size) of CSolveCap::AutoSetPrecondType (Solver/SolveCapacitance.cpp:
    580-608): from thousands of mutual links times conductors, Jacobi up to 500,
    then the two-level preconditioner of dimension 128 / 512 / 1024 above 500 /
    5000 / 10000."""
    lx = (mutual_links // 1000) * conductors
    if lx <= 500:
        return 1, 128
    return 4, (128 if lx <= 5000 else 512 if lx <= 10000 else 1024)


def preconditioner(a: Args, mesh):
    if a.auto_precond:
        return auto_preconditioner(int(mesh.stats.links), len(mesh.conductor_names))
    return a.precond, (a.block_size if a.precond == 2 else a.super_dim)


def format_stats(mesh, a: Args | None = None, res=None) -> str:
    """CSolveCap::OutputSolveStats (Solver/SolveCapacitance.cpp:684-721); with -i the
    phase times and the device memory of the job and its solve (:704-720)."""
    st = mesh.stats
    n, links = int(st.panels), int(st.links) + int(st.panels)
    ratio = np.float32(100.0) - np.float32(links) / np.float32(n * n) * 100.0
    extra = ""
    if a is not None and a.dump_time and res is not None:
        jm = res.job_mem or {}
        mem = res.mem or {}
        extra = (f"Time for reading input file: {st.t_read_s:f}s\n"
                 f"Time for building super hierarchy: {st.t_superh_s:f}s\n"
                 f"Time for discretization: {st.t_discretize_s:f}s\n"
                 f"Time for building potential matrix: {st.t_links_s + res.t_fill_s:f}s\n"
                 f"Time for precond calculation: {res.t_precond_s:f}s\n"
                 f"Time for gmres solving: {res.t_solve_s:f}s\n"
                 f"Memory allocated for panel hierarchy: {jm.get('hierarchy', 0)} bytes\n"
                 f"Memory allocated for links structure: {jm.get('links', 0)} bytes\n"
                 f"Memory allocated for conductor list: {jm.get('conductors', 0)} bytes\n"
                 f"Memory allocated for Gmres: {mem.get('gmres', 0)} bytes\n"
                 f"Memory allocated for preconditioner: {jm.get('precond', 0)} bytes\n"
                 f"Memory allocated for charge vectors: {mem.get('vectors', 0)} bytes\n"
                 f"Memory allocated for hierarchical multiplication: {mem.get('product', 0)} bytes\n")
    return ("\nSolve statistics:\n"
            f"Number of input panels: {st.input_raw_cond_panels + st.input_raw_diel_panels} of which "
            f"{st.input_raw_cond_panels} conductors and {st.input_raw_diel_panels} dielectric\n"
            f"Number of input panels to solver engine: {st.input_panels}\n"
            f"Number of panels after refinement: {n}\n"
            f"Number of potential estimates: {int(st.links) // 2}\n"
            f"Number of links: {links} (uncompressed {n * n}, compression ratio is {ratio:.1f}%)\n"
            f"Max recursion level: {st.max_level}\n"
            f"Max Mesh relative refinement value: {st.max_mesh_eps:g}\n" + extra)


def format_params(a: Args, tol: float) -> str:
    """CSolveCap::OutputSolveParams (Solver/SolveCapacitance.cpp:610-650)."""
    lines = []
    if a.out_geo:
        lines.append("Output refined geometry in FastCap compatible format (-o)")
    if a.out_charge:
        lines.append("Output charge densities information (-c)")
    if a.out_csv:
        lines.append("Output capacitance matrix to file (-e)")
    scheme = "Galerkin" if a.galerkin else "Collocation"
    lines += [f"Solution scheme (-g): {scheme}, GMRES tolerance (-t): {tol:g}",
              "Out-of-core free memory to link memory condition (-f): 5",
              f"Potential interaction coefficient to mesh refinement ratio (-d): {a.eps_ratio:g}",
              f"Mesh curvature (-mc): {a.curv_coeff:g}"]
    return "\n".join(lines) + "\n"


def _precond_text(precond: int, size: int) -> str:
    return {0: "None", 1: "Jacobi ", 2: f"Block, block preconditioner dimension (-pb): {size} ",
            4: f"Two-levels, two-levels preconditioner dimension (-ps): {size} "}[precond]


def _report_solve(out, a: Args, mesh, res):
    """The link, preconditioner and GMRES lines of one solve (SolveComputeLinks,
    OutputSolvePrecondType, gmresPrecondSFastAll's iteration line)."""
    precond, size = preconditioner(a, mesh)
    out.write("***************************************\nComputing the links.. \n"
              f"Number of panels after refinement: {mesh.N}\nNumber of links to be computed: {mesh.L}\n"
              "Done computing links\n***************************************\n"
              f"Precond Type(s) (-p): {_precond_text(precond, size)}\n")
    for q, it in enumerate(res.iterations):
        h = res.residuals[q] if res.residuals is not None else None
        if h is None:
            out.write("GMRES Iteration: " + "".join(f"{i} " for i in range(it)) + "\n")
        else:  # -r: the residual before each step, the final one at the end
            out.write("GMRES Iteration: " + "".join(f"{h[i]:.3g} {i} " for i in range(it)) + f"{h[it]:.3f} \n")


def solve_file(dev, a: Args, mesh_eps: float, tol: float, out=None):
    from . import Mesh

    mesh = Mesh.build(a.file, mesh_eps=mesh_eps, eps_ratio=a.eps_ratio, curv_coeff=a.curv_coeff,
                      device=None if a.host_links else dev, refine_on_device=not a.host_links)
    if a.galerkin and mesh.dim == 3:
        from . import FCG_SCHEME_GALERKIN

        mesh.scheme = FCG_SCHEME_GALERKIN
    job = dev.upload(mesh)
    try:
        t0 = time.perf_counter()
        dev.fill(job, a.precision)
        t_fill = time.perf_counter() - t0
        precond, size = preconditioner(a, mesh)
        res = dev.solve(job, mesh, tol=tol, precond=precond, precision=a.precision, precond_size=size,
                        want_charges=a.out_charge, want_residuals=a.dump_resid)
        res.t_fill_s = t_fill
        res.job_mem = dev.memory(job)
        if out is not None:
            _report_solve(out, a, mesh, res)
    finally:
        dev.release(job)
    return mesh, res


def write_csv(path: str, cap: np.ndarray, dim: int, out=sys.stdout):
    """CSolveCap::OutputCapMtxToFile (Solver/SolveCapacitance.cpp:835-911)."""
    from . import TWO_PI_TIMES_E0

    base = path[: path.rfind(".")] if "." in path else path
    name = base + ".csv"
    M, N = cap.shape
    if dim == 2 and N > 1:
        N -= 1
    scale = TWO_PI_TIMES_E0 if dim == 2 else 1.0
    out.write(f'Generating capacitance matrix file "{name}"\n')
    with open(name, "w") as f:
        for i in range(M):
            f.write(", ".join(_entry(cap[i, j], scale) for j in range(N)) + "\n")


def write_outputs(a: Args, mesh, res, out=sys.stdout):
    """The file outputs after the last solve (Solver/SolveCapacitance.cpp:533-566)."""
    import ctypes as C

    from . import _ffi

    if a.out_csv:
        write_csv(a.file, res.cap, mesh.dim, out)
    lib = _ffi.host_lib()
    h = C.c_void_p(mesh._h)
    if a.out_charge and res.charges is not None:
        out.write("\n")
        for i, name in enumerate(mesh.conductor_names[: res.cap.shape[0]]):
            out.write(f"Outputting charge densities for conductor {name}..\n")
            # complex solves: the real part of the charge densities
            q = np.ascontiguousarray(np.real(res.charges[i]), dtype=np.float64)
            lib.fch_write_fastcap(h, a.file.encode(), name.encode(), q.ctypes.data_as(C.POINTER(C.c_double)))
            out.write("Done\n")
    if a.out_geo:
        out.write("\nOutputting refined geometry in FastCap compatible format..\n")
        lib.fch_write_fastcap(h, a.file.encode(), b"ref", None)
        out.write("Done\n")


def run(a: Args, out=sys.stdout):
    """Manual or -a run; returns (mesh, SolveResult) of the final iteration."""
    from . import Device

    if a.galerkin:
        # the Galerkin scheme is 3D only: 2D inputs are solved by collocation
        # (Solver/SolveCapacitance.cpp:159-161)
        from . import Mesh

        if Mesh.build(a.file, mesh_eps=1e32).dim == 2:
            print("Warning: Galerkin scheme yet not supported in 2D, moving to Collocation", file=sys.stderr)
            a.galerkin = False
    dev = Device(a.device)
    try:
        if not a.auto:
            out.write("\n" + format_params(a, a.tol))
            out.write("***************************************\nRefining the geometry.. \n"
                      f"Refinement completed\nMesh refinement (-m): {a.mesh_eps:g}\n")
            mesh, res = solve_file(dev, a, a.mesh_eps, a.tol, out)
            out.write(format_cap(res.cap, mesh.conductor_names, mesh.dim, a.verbose))
            out.write(format_stats(mesh, a, res))
            write_outputs(a, mesh, res, out)
            return mesh, res
        tol = a.auto_err / 2.0
        out.write(f"Auto calculation with max error: {a.auto_err:g}\n"
                  "Remark: Auto option overrides all other Manual settings\n\n" + format_params(a, tol))
        out.write("\nIteration number #0 ***************************\n\n"
                  "***************************************\nRefining the geometry.. \n"
                  f"Refinement completed\nMesh refinement (-m): {1e32:g}\n")
        mesh, res = solve_file(dev, a, 1e32, tol, out)
        out.write(format_cap(res.cap, mesh.conductor_names, mesh.dim, a.verbose))
        out.write(format_stats(mesh, a, res))
        old_p, old_l = mesh.N, mesh.L
        for i in range(1, MAX_AUTO_ITERATIONS):
            out.write(f"\nIteration number #{i} ***************************\n\n"
                      "***************************************\nIncreasing the geometric refinement.. \n")
            m = mesh.stats.max_mesh_eps
            for _ in range(MAX_AUTO_ITERATIONS):
                m = m / math.sqrt(2.0)
                nmesh, nres = solve_file(dev, a, m, tol)
                if nmesh.L > old_l * AUTO_INCREMENT and nmesh.N > old_p * AUTO_INCREMENT:
                    break
                out.write(f"Delta in refined panel and link count w.r.t. previous iteration less than "
                          f"{int(round((AUTO_INCREMENT - 1.0) * 100.0))}%\n(Panels # {nmesh.N}, Links # {nmesh.L})\n"
                          "Automatically increasing the refinement parameters\n")
                m = nmesh.stats.max_mesh_eps
            out.write(f"Refinement completed\nMesh refinement (-m): {m:g}\n")
            _report_solve(out, a, nmesh, nres)
            diff = np.linalg.norm(nres.cap - res.cap) / np.linalg.norm(nres.cap)
            out.write(format_cap(nres.cap, nmesh.conductor_names, nmesh.dim, a.verbose))
            out.write(f"Weighted Frobenius norm of the difference between capacitance (auto option): {diff:g}\n")
            out.write(format_stats(nmesh, a, nres))
            mesh, res = nmesh, nres
            old_p, old_l = mesh.N, mesh.L
            if diff <= a.auto_err:
                break
        write_outputs(a, mesh, res, out)
        return mesh, res
    finally:
        dev.close()


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    try:
        a = parse_args(argv)
    except CommandLineError as e:
        print(f"Error: {e}", file=sys.stderr)
        return COMMAND_LINE_ERROR
    if a.stop:
        print("fastercap (gfx950 solve path), FasterCap 6.0.7 compatible input and output")
        return 0
    t0 = time.perf_counter()
    run(a)
    print(f"Total time: {time.perf_counter() - t0:g} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
