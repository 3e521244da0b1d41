// fastercap: command-line drop-in of FasterCap's console solve on the gfx950 path.
//
// Options and defaults follow FasterCapConsole::ParseCmdLine
// (FasterCapConsole.cpp:136-357) and CAutoRefGlobalVars::Reset
// (Solver/SolverGlobal.cpp:53-90); the run follows CSolveCap::Run / SolveMain
// (Solver/SolveCapacitance.cpp:86-578), including the -a loop, and the matrix is
// printed as CSolveCap::OutputCapMtx does (Solver/SolveCapacitance.cpp:724-809).
// The mesh and links come from the host builder (include/fastercap_host.h), the
// fill and the solve from the device library (include/fastercap_amd.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fastercap_amd.h"
#include "fastercap_host.h"

namespace {

constexpr double kTwoPiE0 = 5.563250279878e-11;  // Solver/Potential.h:42
constexpr int kMaxAutoIterations = 128;          // SOLVE_MAX_AUTO_ITERATIONS
constexpr double kAutoIncrement = 1.1;           // SOLVE_AUTO_INCREMENT_FACTOR

struct Options {
    std::string file;
    double mesh_eps = 0.1, eps_ratio = 1.0, curv = 3.0, tol = 0.01, auto_err = 0.01;
    bool autom = false, verbose = false, auto_precond = false;
    int precond = FCG_PRECOND_JACOBI;
    int block_size = 32;   // -pb<n> (SolverGlobal.cpp:60)
    int super_dim = 128;   // -ps<n> (SolverGlobal.cpp:59)
    bool jacobi_set = false;  // -pj given (kept only without -ps, SolveCapacitance.cpp:122-125)
    int precision = FCG_PREC_FP64;
    int device = 0;
    bool host_links = false;  // -glinks-host: refinement and link discovery on the host builder
    bool out_csv = false;     // -e: matrix to <input>.csv
    bool out_geo = false;     // -o: refined geometry in FastCap format
    bool out_charge = false;  // -c: charge densities per conductor in FastCap format
    bool dump_resid = false;  // -r: GMRES residual at each iteration
    bool dump_time = false;   // -i: phase times and memory
    bool galerkin = false;    // -g: Galerkin scheme (3D)
};

void usage() {
    std::printf(
        "Usage: fastercap <input file> [options]\n"
        "  -a<err>  Automatically refine, stop when the relative C-matrix change is below err (default 0.01)\n"
        "  -ap      Automatic preconditioner choice (Jacobi or two-level, from links x conductors)\n"
        "  -m<eps>  Mesh relative refinement value (default 0.1)\n"
        "  -mc<c>   Mesh curvature coefficient (default 3)\n"
        "  -t<tol>  GMRES iteration tolerance (default 0.01)\n"
        "  -d<r>    Direct potential interaction coefficient to mesh refinement ratio (default 1)\n"
        "  -pj      Jacobi preconditioner (default)\n"
        "  -pb<n>   Block preconditioner, blocks of at most n panels (2..128)\n"
        "  -ps<n>   Two-level preconditioner of dimension n (16..4096)\n"
        "  -pn      No preconditioner\n"
        "  -e       Also write the capacitance matrix to <input>.csv\n"
        "  -o       Write the refined geometry in FastCap format (<input>_ref.lst)\n"
        "  -c       Write each conductor's charge densities in FastCap format (<input>_<conductor>.lst)\n"
        "  -r       Print the GMRES residual at each iteration\n"
        "  -i       Print phase times and device memory\n"
        "  -v       Verbose output\n"
        "  -b       Console mode (always)\n"
        "  -b?      This text\n"
        "  -bv      Print only the version\n"
        "  -g       Galerkin scheme (3D; 2D inputs are solved by collocation)\n"
        "  -gmixed  Mixed precision (FP32 far field and products, FP64 refinement)\n"
        "  -gdev<n> HIP device (default 0)\n"
        "  -glinks-host  Refine the panels and discover the links on the host (default: on the device,\n"
        "                the same mesh and links)\n");
}

// returns 0 or FC_COMMAND_LINE_ERROR
int parse(int argc, char** argv, Options& o, bool& stop) {
    stop = false;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (a[0] != '-') {
            if (!o.file.empty()) {
                std::fprintf(stderr, "Error: more than one input file ('%s', '%s')\n", o.file.c_str(), a);
                return FCG_COMMAND_LINE_ERROR;
            }
            o.file = a;
            continue;
        }
        auto num = [&](const char* s, double* v) {
            char* end = nullptr;
            *v = std::strtod(s, &end);
            return end != s && *end == '\0';
        };
        bool ok = true;
        if (!std::strcmp(a, "-b?")) {
            usage();
            stop = true;
            return 0;
        } else if (!std::strcmp(a, "-bv")) {
            std::printf("fastercap (gfx950 solve path), FasterCap 6.0.7 compatible input and output\n");
            stop = true;
            return 0;
        } else if (!std::strcmp(a, "-b")) {
        } else if (!std::strcmp(a, "-v")) {
            o.verbose = true;
        } else if (!std::strcmp(a, "-ap")) {
            o.auto_precond = true;
        } else if (!std::strcmp(a, "-pj")) {
            if (o.precond != FCG_PRECOND_SUPER) o.precond = FCG_PRECOND_JACOBI;
            o.jacobi_set = true;
        } else if (!std::strcmp(a, "-pn")) {
            o.precond = FCG_PRECOND_NONE;
            o.jacobi_set = false;
        } else if (!std::strncmp(a, "-pb", 3)) {
            // the reference adds the block preconditioner to the current choice and
            // then runs a flexible GMRES that it reports as unsupported; here -pb
            // selects the block preconditioner alone (as -pn -pb<n> there)
            double v;
            ok = num(a + 3, &v) && v >= 0 && v == (int)v;
            if (ok) {
                o.precond = FCG_PRECOND_BLOCK;
                o.block_size = (int)v;
                if (o.block_size > 128) {
                    std::fprintf(stderr, "Warning: maximum block preconditioner dimension is 128, using 128\n");
                    o.block_size = 128;
                } else if (o.block_size < 2) {
                    std::fprintf(stderr, "Warning: minimum block preconditioner dimension is 2, using 2\n");
                    o.block_size = 2;
                }
            }
        } else if (!std::strncmp(a, "-ps", 3)) {
            // the two-level preconditioner replaces Jacobi (SolveCapacitance.cpp:
            // 122-125) and, here, a block preconditioner given before it
            double v;
            ok = num(a + 3, &v) && v >= 0 && v == (int)v;
            if (ok) {
                o.precond = FCG_PRECOND_SUPER;
                o.super_dim = (int)std::min(v, 1e9);
                if (o.super_dim > 4096) {
                    std::fprintf(stderr, "Warning: maximum two-levels preconditioner dimension is 4096, using 4096\n");
                    o.super_dim = 4096;
                } else if (o.super_dim < 16) {
                    std::fprintf(stderr, "Warning: minimum two-levels preconditioner dimension is 16, using 16\n");
                    o.super_dim = 16;
                }
            }
        } else if (!std::strcmp(a, "-glinks-host")) {
            o.host_links = true;
        } else if (!std::strcmp(a, "-e")) {
            o.out_csv = true;
        } else if (!std::strcmp(a, "-o")) {
            o.out_geo = true;
        } else if (!std::strcmp(a, "-c")) {
            o.out_charge = true;
        } else if (!std::strcmp(a, "-r")) {
            o.dump_resid = true;
        } else if (!std::strcmp(a, "-i")) {
            o.dump_time = true;
        } else if (!std::strcmp(a, "-gmixed")) {
            o.precision = FCG_PREC_MIXED;
        } else if (!std::strcmp(a, "-g")) {
            o.galerkin = true;
        } else if (!std::strncmp(a, "-gdev", 5)) {
            double v;
            ok = num(a + 5, &v);
            o.device = (int)v;
        } else if (!std::strncmp(a, "-mc", 3)) {
            ok = num(a + 3, &o.curv);
        } else if (a[1] == 'm') {
            ok = num(a + 2, &o.mesh_eps);
        } else if (a[1] == 'd') {
            ok = num(a + 2, &o.eps_ratio);
        } else if (a[1] == 't') {
            ok = num(a + 2, &o.tol);
        } else if (a[1] == 'a') {
            o.autom = true;
            ok = a[2] == '\0' || num(a + 2, &o.auto_err);
        } else {
            std::fprintf(stderr, "Error: unsupported parameter '%s' (this build covers the solve path only)\n", a);
            return FCG_COMMAND_LINE_ERROR;
        }
        if (!ok) {
            std::fprintf(stderr, "Error: bad value in parameter '%s'\n", a);
            return FCG_COMMAND_LINE_ERROR;
        }
    }
    if (o.precond == FCG_PRECOND_SUPER && o.jacobi_set)
        std::fprintf(stderr, "Warning: both two-levels preconditioner and Jacobi preconditioner specified, keeping only "
                             "the two-levels preconditioner\n");
    if (o.file.empty() && !stop) {
        std::fprintf(stderr, "Error: no input file\n");
        usage();
        return FCG_COMMAND_LINE_ERROR;
    }
    return 0;
}

struct Result {
    std::shared_ptr<fch_mesh> mesh;  // kept for the -o / -c writers
    double mesh_eps = 0.0;
    std::vector<double> charges;     // [rows][N] with -c ([rows][2N], Re | Im, when complex)
    std::vector<double> cap;         // rows x cols, then the imaginary part when complex
    int rows = 0, cols = 0, dim = 3;
    bool cplx = false;               // complex permittivities (m_ucHasCmplxPerm)
    std::vector<std::string> names;
    fch_stats hs{};
    fcg_solve_stats ss{};
    fcg_job_mem mem{};
    double t_fill_s = 0.0;
};

// refinement and links at the given mesh eps (RefineGeoAndCountLinks +
// SolveComputeLinks); the mesh is kept in r
int build_mesh(fcg_ctx* ctx, const Options& o, double mesh_eps, Result& r) {
    fch_options ho;
    fch_default_options(&ho);
    ho.mesh_eps = mesh_eps;
    ho.eps_ratio = o.eps_ratio;
    ho.curv_coeff = o.curv;
    ho.verbose = o.verbose ? 1 : 0;
    ho.links_on_device = o.host_links ? 0 : 1;
    ho.refine_on_device = o.host_links ? 0 : 1;
    fch_mesh* mesh = nullptr;
    int rc = fch_build(o.file.c_str(), &ho, &mesh);
    if (rc != 0) {
        std::fprintf(stderr, "Error: %s\n", fch_last_error());
        return rc;
    }
    if (!o.host_links) {
        // the panel refinement on the device (fcg_refine), handed back to the builder
        const auto t0 = std::chrono::steady_clock::now();
        const fcg_refine_node* coarse = nullptr;
        const int32_t* tops = nullptr;
        int n = 0, n_out = 0;
        fcg_refine_params rp;
        double eps = 0.0;
        int32_t lev = 0;
        rc = fch_get_coarse_tree(mesh, &coarse, &n, &tops, &rp);
        if (rc == FCG_OK) rc = fcg_refine(ctx, coarse, n, tops, &rp, &n_out, &eps, &lev);
        std::vector<fcg_refine_node> refined;
        if (rc == FCG_OK) {
            refined.resize((size_t)n_out);
            rc = fcg_refine_fetch(ctx, refined.data());
        }
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc == FCG_OK) rc = fch_set_refined_tree(mesh, refined.data(), n_out, eps, lev, t);
        if (rc != FCG_OK) {
            std::fprintf(stderr, "Error: refinement: %s %s\n", fcg_last_error(ctx), fch_last_error());
            fch_free(mesh);
            return rc;
        }
    }
    fcg_job_desc d;
    fch_get_desc(mesh, &d);
    if (!o.host_links) {
        // the link structure from the device (fcg_links_discover); it stays there
        // for fcg_job_upload_links, the mesh records only its statistics
        const auto t0 = std::chrono::steady_clock::now();
        fcg_link_params lp;
        fch_get_link_params(mesh, &lp);
        int64_t L = 0;
        int32_t maxlev = 0;
        rc = fcg_links_discover(ctx, &d, &lp, &L, &maxlev);
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc == FCG_OK) rc = fch_set_links(mesh, nullptr, nullptr, L, maxlev, t);
        if (rc != FCG_OK) {
            std::fprintf(stderr, "Error: link discovery: %s\n", fcg_last_error(ctx));
            fch_free(mesh);
            return rc;
        }
    }
    fch_get_stats(mesh, &r.hs);
    r.dim = d.dim;
    r.names.clear();
    for (int g = 0; g < d.n_groups; ++g)
        if (!d.group_is_diel[g]) r.names.push_back(fch_group_name(mesh, g));
    r.mesh = std::shared_ptr<fch_mesh>(mesh, fch_free);
    r.mesh_eps = mesh_eps;
    return FCG_OK;
}

const char* precond_line(int precond, int size) {
    static char buf[160];
    if (precond == FCG_PRECOND_NONE)
        std::snprintf(buf, sizeof buf, "None");
    else if (precond == FCG_PRECOND_JACOBI)
        std::snprintf(buf, sizeof buf, "Jacobi ");
    else if (precond == FCG_PRECOND_BLOCK)
        std::snprintf(buf, sizeof buf, "Block, block preconditioner dimension (-pb): %d ", size);
    else
        std::snprintf(buf, sizeof buf, "Two-levels, two-levels preconditioner dimension (-ps): %d ", size);
    return buf;
}

// upload, fill and solve of the mesh in r (SolveForCapacitance), with the
// reference's link and preconditioner banners and per-solve GMRES lines
int solve_mesh(fcg_ctx* ctx, const Options& o, double tol, Result& r) {
    fch_mesh* mesh = r.mesh.get();
    fcg_job_desc d;
    fch_get_desc(mesh, &d);
    int rc = FCG_OK;
    // SolveComputeLinks (Solver/SolveCapacitance.cpp:1025-1041)
    std::printf("***************************************\nComputing the links.. \n");
    std::printf("Number of panels after refinement: %lld\n", (long long)r.hs.panels);
    std::printf("Number of links to be computed: %lld\n", (long long)r.hs.links);
    std::printf("Done computing links\n***************************************\n");
    int precond = o.precond, size = o.precond == FCG_PRECOND_BLOCK ? o.block_size : o.super_dim;
    if (o.auto_precond) {
        // CSolveCap::AutoSetPrecondType (Solver/SolveCapacitance.cpp:580-608):
        // thousands of mutual links times conductors
        const long long lx = (long long)(r.hs.links / 1000) * (long long)r.names.size();
        precond = lx <= 500 ? FCG_PRECOND_JACOBI : FCG_PRECOND_SUPER;
        size = lx <= 5000 ? 128 : lx <= 10000 ? 512 : 1024;
        if (o.verbose) {
            if (precond == FCG_PRECOND_JACOBI)
                std::printf("Automatic preconditioner: Jacobi\n");
            else
                std::printf("Automatic preconditioner: two-levels, dimension %d\n", size);
        }
    }
    std::printf("Precond Type(s) (-p): %s\n", precond_line(precond, size));
    int job = -1;
    if (o.galerkin && d.dim == 3) d.scheme = FCG_SCHEME_GALERKIN;
    rc = o.host_links ? fcg_job_upload(ctx, &d, &job) : fcg_job_upload_links(ctx, &d, &job);
    const auto tf = std::chrono::steady_clock::now();
    if (rc == FCG_OK) rc = fcg_fill(ctx, job, o.precision);
    r.t_fill_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tf).count();
    // -r: the residual history of every right-hand side (|r0|, then |g_i| / |b|)
    constexpr int kHistCap = 1000 * 9 + 1;  // max_iter x (max_refine + 1) + 1
    std::vector<double> hist;
    if (rc == FCG_OK) {
        fcg_solve_opts so{tol, 1000, precond, o.precision, 8, size, nullptr, 0};
        if (o.dump_resid) {
            hist.assign((size_t)d.n_groups * kHistCap, 0.0);
            so.residuals = hist.data();
            so.residuals_cap = kHistCap;
        }
        const int nc = (int)r.names.size();
        r.cplx = d.group_outperm_im != nullptr;
        const size_t nf = r.cplx ? 2 : 1;
        r.cap.assign(nf * nc * nc, 0.0);
        if (o.out_charge) r.charges.assign(nf * (size_t)nc * (size_t)d.n_leaves, 0.0);
        rc = fcg_solve(ctx, job, &so, r.cap.data(), &r.rows, &r.cols, o.out_charge ? r.charges.data() : nullptr,
                       &r.ss);
    }
    if (rc == FCG_OK) rc = fcg_job_memory(ctx, job, &r.mem);
    // one GMRES line per right-hand side (gmresPrecondSFastAll,
    // Solver/SolveCapacitance.cpp:2359-2514): the step numbers, with -r each
    // preceded by the residual before it and the final residual at the end
    if (rc == FCG_OK)
        for (int q = 0; q < r.rows && q < 64; ++q) {
            const int n = r.ss.iterations[q];
            const double* h = hist.empty() ? nullptr : hist.data() + (size_t)q * kHistCap;
            std::printf("GMRES Iteration: ");
            for (int i = 0; i < n; ++i) {
                if (h && i < kHistCap) std::printf("%.3g ", h[i]);
                std::printf("%d ", i);
            }
            if (h) std::printf("%.3f ", h[std::min(n, kHistCap - 1)]);
            std::printf("\n");
        }
    if (rc != FCG_OK) std::fprintf(stderr, "Error: %s\n", fcg_last_error(ctx));
    if (job >= 0) fcg_job_release(ctx, job);
    return rc;
}

// CSolveCap::OutputCapMtxToFile (Solver/SolveCapacitance.cpp:835-911): <input
// without extension>.csv, ", " between columns, 2D in F/m without the balance column
void write_csv(const Options& o, const Result& r) {
    std::string base = o.file;
    const size_t dot = base.rfind('.');
    if (dot != std::string::npos) base.resize(dot);
    const std::string name = base + ".csv";
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "Warning: cannot open \"%s\" for writing, skipping\n", name.c_str());
        return;
    }
    std::printf("Generating capacitance matrix file \"%s\"\n", name.c_str());
    int N = r.cols;
    if (r.dim == 2) {
        if (N > 1)
            --N;
        else
            std::fprintf(stderr, "Error: only one input conductor defined for a 2D problem. Result is meaningless.");
    }
    const double scale = r.dim == 2 ? kTwoPiE0 : 1.0;
    const size_t im = (size_t)r.rows * r.cols;  // offset of the imaginary part
    for (int i = 0; i < r.rows; ++i) {
        for (int j = 0; j < N; ++j) {
            const size_t e = (size_t)i * r.cols + j;
            if (r.cplx)
                std::fprintf(f, "%g%+gj", r.cap[e] * scale, r.cap[im + e] * scale);
            else
                std::fprintf(f, "%g", r.cap[e] * scale);
            if (j < N - 1) std::fprintf(f, ", ");
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

// the file outputs of SolveMain after the last solve (Solver/SolveCapacitance.cpp:533-566)
void write_outputs(const Options& o, const Result& r) {
    if (o.out_csv) write_csv(o, r);
    if (o.out_charge && r.mesh) {
        std::printf("\n");
        for (int i = 0; i < r.rows && i < (int)r.names.size(); ++i) {
            std::printf("Outputting charge densities for conductor %s..\n", r.names[(size_t)i].c_str());
            // complex: the real part of the charge densities
            const size_t stride = r.charges.size() / std::max<size_t>(r.names.size(), 1);
            fch_write_fastcap(r.mesh.get(), o.file.c_str(), r.names[(size_t)i].c_str(),
                              r.charges.data() + (size_t)i * stride);
            std::printf("Done\n");
        }
    }
    if (o.out_geo && r.mesh) {
        std::printf("\nOutputting refined geometry in FastCap compatible format..\n");
        fch_write_fastcap(r.mesh.get(), o.file.c_str(), "ref", nullptr);
        std::printf("Done\n");
    }
}

void print_cap(const Options& o, const Result& r) {
    std::printf("Capacitance matrix is:\n");
    int M = r.rows, N = r.cols;
    if (r.dim == 2 && !o.verbose) {
        if (N > 1)
            --N;  // the balance conductor's column
        else
            std::fprintf(stderr, "Error: only one input conductor defined for a 2D problem. Result is meaningless.");
    }
    std::printf("Dimension %d x %d\n", M, N);
    const double scale = r.dim == 2 ? kTwoPiE0 : 1.0;
    const size_t imo = (size_t)r.rows * r.cols;  // offset of the imaginary part
    auto im = [&](int i, int j) { return r.cplx ? r.cap[imo + (size_t)i * r.cols + j] : 0.0; };
    for (int i = 0; i < M; ++i) {
        std::printf("%s  ", i < (int)r.names.size() ? r.names[(size_t)i].c_str() : "?");
        for (int j = 0; j < N; ++j) {
            if (r.cplx)
                std::printf("%g%+gj ", r.cap[(size_t)i * r.cols + j] * scale, im(i, j) * scale);
            else
                std::printf("%g ", r.cap[(size_t)i * r.cols + j] * scale);
        }
        std::printf("\n");
    }
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            if (i != j && r.cap[(size_t)i * r.cols + j] >= 0.0)
                std::fprintf(stderr, "Warning: capacitance matrix has a non-negative off-diagonal element at row %d col %d\n",
                             i + 1, j + 1);
    // diagonal dominance on the moduli (real matrices: the absolute values)
    for (int i = 0; i < M; ++i) {
        double sum = 0.0;
        for (int j = 0; j < N; ++j)
            if (i != j) sum += std::hypot(r.cap[(size_t)i * r.cols + j], im(i, j));
        if (std::hypot(r.cap[(size_t)i * r.cols + i], im(i, i)) < sum)
            std::fprintf(stderr, "Warning: capacitance matrix is not diagonally dominant due to row %d\n", i + 1);
    }
    std::printf("\n");
}

// CSolveCap::OutputSolveStats (Solver/SolveCapacitance.cpp:684-721)
void print_stats(const Options& o, const Result& r) {
    const long long N = r.hs.panels, L = r.hs.links + r.hs.panels;
    std::printf("\nSolve statistics:\n");
    std::printf("Number of input panels: %lld of which %lld conductors and %lld dielectric\n",
                (long long)(r.hs.input_raw_cond_panels + r.hs.input_raw_diel_panels),
                (long long)r.hs.input_raw_cond_panels, (long long)r.hs.input_raw_diel_panels);
    std::printf("Number of input panels to solver engine: %lld\n", (long long)r.hs.input_panels);
    std::printf("Number of panels after refinement: %lld\n", N);
    std::printf("Number of potential estimates: %lld\n", (long long)(r.hs.links / 2));
    std::printf("Number of links: %lld (uncompressed %lld, compression ratio is %.1f%%)\n", L, N * N,
                100.0f - ((float)L) / ((float)(N * N)) * 100.0);
    std::printf("Max recursion level: %d\n", r.hs.max_level);
    std::printf("Max Mesh relative refinement value: %g\n", r.hs.max_mesh_eps);
    if (!o.dump_time) return;
    // -i (Solver/SolveCapacitance.cpp:704-720): the phases of this path and the
    // device memory of the job and its solve
    std::printf("Time for reading input file: %fs\n", r.hs.t_read_s);
    std::printf("Time for building super hierarchy: %fs\n", r.hs.t_superh_s);
    std::printf("Time for discretization: %fs\n", r.hs.t_discretize_s);
    std::printf("Time for building potential matrix: %fs\n", r.hs.t_links_s + r.t_fill_s);
    std::printf("Time for precond calculation: %fs\n", r.ss.t_precond_s);
    std::printf("Time for gmres solving: %fs\n", r.ss.t_solve_s);
    std::printf("Memory allocated for panel hierarchy: %lld bytes\n", (long long)r.mem.hierarchy);
    std::printf("Memory allocated for links structure: %lld bytes\n", (long long)r.mem.links);
    std::printf("Memory allocated for conductor list: %lld bytes\n", (long long)r.mem.conductors);
    std::printf("Memory allocated for Gmres: %lld bytes\n", (long long)r.ss.mem_gmres);
    std::printf("Memory allocated for preconditioner: %lld bytes\n", (long long)r.mem.precond);
    std::printf("Memory allocated for charge vectors: %lld bytes\n", (long long)r.ss.mem_vectors);
    std::printf("Memory allocated for hierarchical multiplication: %lld bytes\n", (long long)r.ss.mem_product);
}

// CSolveCap::PrintTime (Solver/SolveCapacitance.cpp)
void print_time(double t) {
    long sec = (long)t, min = sec / 60, hr = min / 60;
    const long days = hr / 24;
    sec -= min * 60;
    min -= hr * 60;
    hr -= days * 24;
    std::printf("%fs (%ld days, %ld hours, %ld mins, %ld s)\n", t, days, hr, min, sec);
}

// CSolveCap::OutputSolveParams (Solver/SolveCapacitance.cpp:610-650)
void print_params(const Options& o, double tol, bool galerkin) {
    if (o.out_geo) std::printf("Output refined geometry in FastCap compatible format (-o)\n");
    if (o.out_charge) std::printf("Output charge densities information (-c)\n");
    if (o.out_csv) std::printf("Output capacitance matrix to file (-e)\n");
    std::printf("Solution scheme (-g): %s, GMRES tolerance (-t): %g\n", galerkin ? "Galerkin" : "Collocation", tol);
    std::printf("Out-of-core free memory to link memory condition (-f): %g\n", 5.0);
    std::printf("Potential interaction coefficient to mesh refinement ratio (-d): %g\n", o.eps_ratio);
    std::printf("Mesh curvature (-mc): %g\n", o.curv);
}

double frob(const std::vector<double>& a) {
    double s = 0.0;
    for (double v : a) s += v * v;
    return std::sqrt(s);
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    bool stop = false;
    int rc = parse(argc, argv, o, stop);
    if (rc != 0 || stop) return rc;
    if (o.curv < 1.0) {
        std::fprintf(stderr, "Warning: curvature coefficient %f is below the minimum value of 1.0, setting to 1.0\n",
                     o.curv);
        o.curv = 1.0;
    }
    std::printf("Running fastercap (gfx950 solve path)\nInput file: %s\n", o.file.c_str());
    fcg_ctx* ctx = nullptr;
    if (fcg_ctx_create(o.device, &ctx) != FCG_OK) {
        std::fprintf(stderr, "Error: cannot open HIP device %d\n", o.device);
        return FCG_DEVICE_ERROR;
    }
    // the Galerkin scheme is 3D only: the reference warns and solves 2D inputs by
    // collocation (Solver/SolveCapacitance.cpp:159-161)
    bool galerkin = o.galerkin;
    if (galerkin) {
        fch_options ho;
        fch_default_options(&ho);
        fch_mesh* probe = nullptr;
        ho.mesh_eps = 1e32;
        ho.links_on_device = 0;
        ho.refine_on_device = 0;
        if (fch_build(o.file.c_str(), &ho, &probe) == 0) {
            fch_stats st;
            fch_get_stats(probe, &st);
            if (st.dim == 2) {
                std::fprintf(stderr, "Warning: Galerkin scheme yet not supported in 2D, moving to Collocation\n");
                galerkin = false;
                o.galerkin = false;
            }
            fch_free(probe);
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    using clk = std::chrono::steady_clock;
    auto since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
    Result r;
    if (!o.autom) {
        // manual run (Solver/SolveCapacitance.cpp:479-531)
        std::printf("\n");
        print_params(o, o.tol, galerkin);
        std::printf("***************************************\nRefining the geometry.. \n");
        rc = build_mesh(ctx, o, o.mesh_eps, r);
        if (rc == FCG_OK) {
            std::printf("Refinement completed\nMesh refinement (-m): %g\n", o.mesh_eps);
            rc = solve_mesh(ctx, o, o.tol, r);
        }
        if (rc == FCG_OK) {
            print_cap(o, r);
            print_stats(o, r);
        }
    } else {
        // -a run (Solver/SolveCapacitance.cpp:294-477)
        std::printf("Auto calculation with max error: %g\n", o.auto_err);
        std::printf("Remark: Auto option overrides all other Manual settings\n");
        const double tol = o.auto_err / 2.0;
        std::printf("\n");
        print_params(o, tol, galerkin);
        std::printf("\nIteration number #0 ***************************\n\n");
        auto ti = clk::now();
        std::printf("***************************************\nRefining the geometry.. \n");
        rc = build_mesh(ctx, o, 1e32, r);
        if (rc == FCG_OK) {
            std::printf("Refinement completed\nMesh refinement (-m): %g\n", 1e32);
            rc = solve_mesh(ctx, o, tol, r);
        }
        if (rc == FCG_OK) {
            print_cap(o, r);
            print_stats(o, r);
            std::printf("Iteration time: ");
            print_time(since(ti));
        }
        // CAutoRefine::GetLinksNum: mutual links only
        long long old_panels = r.hs.panels, old_links = r.hs.links;
        for (int i = 1; rc == FCG_OK && i < kMaxAutoIterations; ++i) {
            std::printf("\nIteration number #%d ***************************\n\n", i);
            ti = clk::now();
            std::printf("***************************************\nIncreasing the geometric refinement.. \n");
            Result nr;
            double m = r.hs.max_mesh_eps;
            for (int j = 0; j < kMaxAutoIterations; ++j) {
                m = (j == 0 ? r.hs.max_mesh_eps : nr.hs.max_mesh_eps) / std::sqrt(2.0);
                rc = build_mesh(ctx, o, m, nr);
                if (rc != FCG_OK) break;
                const long long np = nr.hs.panels, nl = nr.hs.links;
                if (nl > old_links * kAutoIncrement && np > old_panels * kAutoIncrement) break;
                std::printf("Delta in refined panel and link count w.r.t. previous iteration less than %d%%\n",
                            (int)((kAutoIncrement - 1.0) * 100.0));
                std::printf("(Panels # %lld, Links # %lld)\n", np, nl);
                if (o.verbose) std::printf("(Mesh relative refinement value (-m): %g)\n", m);
                std::printf("Automatically increasing the refinement parameters\n");
            }
            if (rc != FCG_OK) break;
            std::printf("Refinement completed\nMesh refinement (-m): %g\n", m);
            rc = solve_mesh(ctx, o, tol, nr);
            if (rc != FCG_OK) break;
            std::vector<double> diff(nr.cap.size());
            for (size_t q = 0; q < diff.size(); ++q) diff[q] = nr.cap[q] - r.cap[q];
            const double rel = frob(diff) / frob(nr.cap);
            print_cap(o, nr);
            std::printf("Weighted Frobenius norm of the difference between capacitance (auto option): %g\n", rel);
            print_stats(o, nr);
            std::printf("Iteration time: ");
            print_time(since(ti));
            old_panels = nr.hs.panels;
            old_links = nr.hs.links;
            r = nr;
            if (rel <= o.auto_err) break;
        }
    }
    if (rc == FCG_OK) write_outputs(o, r);
    std::printf("\nTotal time: ");
    print_time(since(t0));
    fcg_ctx_destroy(ctx);
    return rc;
}
