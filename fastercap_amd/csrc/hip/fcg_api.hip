// C-ABI of the gfx950 FasterCap solve path (declarations and reference citations
// in include/fastercap_amd.h).  Owns the HBM-resident job state and drives the
// kernels of kernels.hip on one stream per context (default priority).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kernels.h"

using fcg::DD;
using fcg::JobView;
using fcg::MvpScratch;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        bytes = n;
        if (n == 0) return hipSuccess;
        return hipMalloc(&p, n);
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

struct PinnedBuf {
    void* p = nullptr;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t alloc(size_t n) { return hipHostMalloc(&p, n, hipHostMallocDefault); }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

struct Job {
    int dim = 3, N = 0, M = 0, G = 0;
    int64_t L = 0;
    DevBuf nodes, left, right, parent, leaf_lo, leaf_hi, leaf_node, row_ptr, col, gout, gin, coef, self, diag;
    DevBuf leaf_w, col_range;
    // host copies of the group table
    std::vector<int32_t> group_is_diel, group_leaf_start;
    std::vector<double> leaf_outperm;
    int first_cond_leaf = 0;
    bool filled = false;
    int gather_variant = 0;
    int tuned_k = 0;  // rhs count the gather variant was tuned for (0 = not tuned)
    // product scratch
    DevBuf prefix, block_sums, charge, pot;
    MvpScratch sc{};
    JobView view() const {
        JobView v{};
        v.dim = dim;
        v.N = N;
        v.M = M;
        v.G = G;
        v.L = L;
        v.nodes = nodes.as<fcg_node_geom>();
        v.left = left.as<int32_t>();
        v.right = right.as<int32_t>();
        v.parent = parent.as<int32_t>();
        v.leaf_lo = leaf_lo.as<int32_t>();
        v.leaf_hi = leaf_hi.as<int32_t>();
        v.leaf_node = leaf_node.as<int32_t>();
        v.row_ptr = row_ptr.as<int64_t>();
        v.col = col.as<int32_t>();
        v.group_outperm = gout.as<double>();
        v.group_inperm = gin.as<double>();
        v.coef = coef.as<double>();
        v.self = self.as<double>();
        v.diag = diag.as<double>();
        v.first_cond_leaf = first_cond_leaf;
        v.gather_variant = gather_variant;
        return v;
    }
};

}  // namespace

struct fcg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Job>> jobs;
    std::string err;
};

namespace {

int fail(fcg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHECK(ctx, call)                                                                              \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            return fail(ctx, e_ == hipErrorOutOfMemory ? FCG_OUT_OF_MEMORY : FCG_DEVICE_ERROR,           \
                        std::string(#call) + ": " + hipGetErrorString(e_));                              \
        }                                                                                                \
    } while (0)

Job* get_job(fcg_ctx* ctx, int job) {
    if (!ctx || job < 0 || (size_t)job >= ctx->jobs.size()) return nullptr;
    return ctx->jobs[(size_t)job].get();
}

template <typename T>
hipError_t upload(DevBuf& b, const T* src, size_t n, hipStream_t s) {
    hipError_t e = b.alloc(n * sizeof(T));
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
}

int ensure_scratch(fcg_ctx* ctx, Job& j, int k) {
    const int nblk = (j.N + fcg::kScanTile - 1) / fcg::kScanTile;
    if (j.sc.kmax >= k && j.sc.nblk >= nblk) return FCG_OK;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHECK(ctx, j.prefix.alloc(sizeof(DD) * (size_t)k * (j.N + 1)));
    HIPCHECK(ctx, j.block_sums.alloc(sizeof(DD) * (size_t)k * (nblk > 1 ? nblk : 1) + 16 * (size_t)k));
    HIPCHECK(ctx, j.charge.alloc(sizeof(double) * (size_t)k * j.M));
    HIPCHECK(ctx, j.pot.alloc(sizeof(double) * (size_t)k * j.M));
    j.sc.prefix = j.prefix.as<DD>();
    j.sc.block_sums = j.block_sums.as<DD>();
    j.sc.charge = j.charge.as<double>();
    j.sc.pot = j.pot.as<double>();
    j.sc.nblk = nblk;
    j.sc.kmax = k;
    return FCG_OK;
}

// Validate the structural invariants the kernels rely on (bounds of every index).
int validate_desc(fcg_ctx* ctx, const fcg_job_desc* d) {
    if (!d || !d->nodes || !d->left || !d->right || !d->parent || !d->leaf_lo || !d->leaf_hi || !d->leaf_node ||
        !d->row_ptr || (d->n_links > 0 && !d->col) || !d->group_is_diel || !d->group_leaf_start ||
        !d->group_outperm || !d->group_inperm || !d->leaf_outperm)
        return fail(ctx, FCG_INVALID_ARGUMENT, "null array in job descriptor");
    if (d->dim != 2 && d->dim != 3) return fail(ctx, FCG_INVALID_ARGUMENT, "dim must be 2 or 3");
    const int N = d->n_leaves, M = d->n_nodes;
    if (N < 1 || M < N || d->n_groups < 1 || d->n_links < 0)
        return fail(ctx, FCG_INVALID_ARGUMENT, "bad sizes in job descriptor");
    if (d->row_ptr[0] != 0 || d->row_ptr[M] != d->n_links)
        return fail(ctx, FCG_INVALID_ARGUMENT, "row_ptr does not span the links");
    for (int i = 0; i < M; ++i) {
        if (d->row_ptr[i + 1] < d->row_ptr[i]) return fail(ctx, FCG_INVALID_ARGUMENT, "row_ptr not monotone");
        const bool leaf = d->left[i] < 0;
        if (leaf != (d->right[i] < 0)) return fail(ctx, FCG_INVALID_ARGUMENT, "node with a single child");
        if (!leaf && (d->left[i] >= i || d->right[i] >= i))
            return fail(ctx, FCG_INVALID_ARGUMENT, "children must precede parents (post-order)");
        if (d->parent[i] >= M || (d->parent[i] >= 0 && d->parent[i] <= i))
            return fail(ctx, FCG_INVALID_ARGUMENT, "parent index out of order");
        if (d->leaf_lo[i] < 0 || d->leaf_hi[i] > N || d->leaf_lo[i] >= d->leaf_hi[i])
            return fail(ctx, FCG_INVALID_ARGUMENT, "bad leaf range");
        if (d->nodes[i].cond < 0 || d->nodes[i].cond >= d->n_groups)
            return fail(ctx, FCG_INVALID_ARGUMENT, "bad node group");
        if (d->nodes[i].maxside <= 0.0) return fail(ctx, FCG_INVALID_ARGUMENT, "non-positive node size");
    }
    for (int k = 0; k < N; ++k) {
        const int nd = d->leaf_node[k];
        if (nd < 0 || nd >= M || d->left[nd] >= 0 || d->leaf_lo[nd] != k)
            return fail(ctx, FCG_INVALID_ARGUMENT, "bad leaf_node map");
    }
    for (int64_t l = 0; l < d->n_links; ++l)
        if (d->col[l] < 0 || d->col[l] >= M) return fail(ctx, FCG_INVALID_ARGUMENT, "link source out of range");
    if (d->group_leaf_start[0] != 0 || d->group_leaf_start[d->n_groups] != N)
        return fail(ctx, FCG_INVALID_ARGUMENT, "group leaf ranges do not span the leaves");
    return FCG_OK;
}

}  // namespace

extern "C" {

int fcg_ctx_create(int device, fcg_ctx** out) {
    if (!out) return FCG_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return FCG_DEVICE_ERROR;
    if (hipSetDevice(device) != hipSuccess) return FCG_DEVICE_ERROR;
    auto c = std::make_unique<fcg_ctx>();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return FCG_DEVICE_ERROR;
    *out = c.release();
    return FCG_OK;
}

int fcg_ctx_destroy(fcg_ctx* ctx) {
    if (!ctx) return FCG_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->jobs.clear();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FCG_OK;
}

const char* fcg_last_error(fcg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fcg_job_upload(fcg_ctx* ctx, const fcg_job_desc* d, int* job) {
    if (!ctx || !job) return FCG_INVALID_ARGUMENT;
    int rc = validate_desc(ctx, d);
    if (rc != FCG_OK) return rc;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    auto j = std::make_unique<Job>();
    j->dim = d->dim;
    j->N = d->n_leaves;
    j->M = d->n_nodes;
    j->G = d->n_groups;
    j->L = d->n_links;
    hipStream_t s = ctx->stream;
    HIPCHECK(ctx, upload(j->nodes, d->nodes, (size_t)j->M, s));
    HIPCHECK(ctx, upload(j->left, d->left, (size_t)j->M, s));
    HIPCHECK(ctx, upload(j->right, d->right, (size_t)j->M, s));
    HIPCHECK(ctx, upload(j->parent, d->parent, (size_t)j->M, s));
    HIPCHECK(ctx, upload(j->leaf_lo, d->leaf_lo, (size_t)j->M, s));
    HIPCHECK(ctx, upload(j->leaf_hi, d->leaf_hi, (size_t)j->M, s));
    HIPCHECK(ctx, upload(j->leaf_node, d->leaf_node, (size_t)j->N, s));
    HIPCHECK(ctx, upload(j->row_ptr, d->row_ptr, (size_t)j->M + 1, s));
    HIPCHECK(ctx, upload(j->col, d->col, (size_t)j->L, s));
    HIPCHECK(ctx, upload(j->gout, d->group_outperm, (size_t)j->G, s));
    HIPCHECK(ctx, upload(j->gin, d->group_inperm, (size_t)j->G, s));
    HIPCHECK(ctx, upload(j->leaf_w, d->leaf_outperm, (size_t)j->N, s));
    HIPCHECK(ctx, j->coef.alloc(sizeof(double) * (size_t)(j->L > 0 ? j->L : 1)));
    HIPCHECK(ctx, j->self.alloc(sizeof(double) * (size_t)j->N));
    HIPCHECK(ctx, j->diag.alloc(sizeof(double) * (size_t)j->M));
    HIPCHECK(ctx, hipMemsetAsync(j->diag.p, 0, sizeof(double) * (size_t)j->M, s));
    j->group_is_diel.assign(d->group_is_diel, d->group_is_diel + j->G);
    j->group_leaf_start.assign(d->group_leaf_start, d->group_leaf_start + j->G + 1);
    j->leaf_outperm.assign(d->leaf_outperm, d->leaf_outperm + j->N);
    j->first_cond_leaf = j->N;
    std::vector<int32_t> ranges;
    for (int g = 0; g < j->G; ++g) {
        if (j->group_is_diel[(size_t)g]) continue;
        if (j->first_cond_leaf == j->N) j->first_cond_leaf = j->group_leaf_start[(size_t)g];
        ranges.push_back(j->group_leaf_start[(size_t)g]);
        ranges.push_back(j->group_leaf_start[(size_t)g + 1]);
    }
    if (ranges.empty()) return fail(ctx, FCG_INVALID_ARGUMENT, "job has no conductor");
    HIPCHECK(ctx, upload(j->col_range, ranges.data(), ranges.size(), s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    ctx->jobs.push_back(std::move(j));
    *job = (int)ctx->jobs.size() - 1;
    return FCG_OK;
}

int fcg_job_release(fcg_ctx* ctx, int job) {
    Job* j = get_job(ctx, job);
    if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->jobs[(size_t)job].reset();
    return FCG_OK;
}

int fcg_job_conductors(fcg_ctx* ctx, int job, int* n_cond) {
    Job* j = get_job(ctx, job);
    if (!j || !n_cond) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    int n = 0;
    for (int32_t d : j->group_is_diel) n += d ? 0 : 1;
    *n_cond = n;
    return FCG_OK;
}

int fcg_fill(fcg_ctx* ctx, int job, int precision) {
    Job* j = get_job(ctx, job);
    if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    if (precision != FCG_PREC_FP64 && precision != FCG_PREC_MIXED)
        return fail(ctx, FCG_INVALID_ARGUMENT, "unknown precision mode");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    JobView v = j->view();
    DevBuf link_row;
    HIPCHECK(ctx, link_row.alloc(sizeof(int32_t) * (size_t)(j->L > 0 ? j->L : 1)));
    HIPCHECK(ctx, fcg::launch_link_rows(v, link_row.as<int32_t>(), s));
    HIPCHECK(ctx, fcg::launch_fill(v, link_row.as<int32_t>(), s));
    HIPCHECK(ctx, fcg::launch_self(v, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    j->filled = true;
    return FCG_OK;
}

int fcg_download_coeffs(fcg_ctx* ctx, int job, double* coef, double* self_coef) {
    Job* j = get_job(ctx, job);
    if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    if (!j->filled) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    if (coef && j->L > 0)
        HIPCHECK(ctx, hipMemcpyAsync(coef, j->coef.p, sizeof(double) * (size_t)j->L, hipMemcpyDeviceToHost,
                                     ctx->stream));
    if (self_coef)
        HIPCHECK(ctx, hipMemcpyAsync(self_coef, j->self.p, sizeof(double) * (size_t)j->N, hipMemcpyDeviceToHost,
                                     ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

int fcg_mvp_device(fcg_ctx* ctx, int job, const double* d_q, double* d_v, int k, void* stream) {
    Job* j = get_job(ctx, job);
    if (!j || !d_q || !d_v || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int rc = ensure_scratch(ctx, *j, k);
    if (rc != FCG_OK) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    HIPCHECK(ctx, fcg::launch_mvp(j->view(), j->sc, d_q, d_v, k, 0, s));
    return FCG_OK;
}

int fcg_mvp(fcg_ctx* ctx, int job, const double* q, double* v, int k) {
    Job* j = get_job(ctx, job);
    if (!j || !q || !v || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DevBuf dq, dv;
    const size_t n = (size_t)j->N * k;
    HIPCHECK(ctx, upload(dq, q, n, ctx->stream));
    HIPCHECK(ctx, dv.alloc(sizeof(double) * n));
    int rc = fcg_mvp_device(ctx, job, dq.as<double>(), dv.as<double>(), k, nullptr);
    if (rc != FCG_OK) return rc;
    HIPCHECK(ctx, hipMemcpyAsync(v, dv.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

int fcg_mvp_timing_run(fcg_ctx* ctx, int job, int k, int reps, fcg_mvp_timing* out) {
    Job* j = get_job(ctx, job);
    if (!j || !out || k < 1 || reps < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int rc = ensure_scratch(ctx, *j, k);
    if (rc != FCG_OK) return rc;
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)j->N * k;
    DevBuf dq, dv;
    HIPCHECK(ctx, dq.alloc(sizeof(double) * n));
    HIPCHECK(ctx, dv.alloc(sizeof(double) * n));
    HIPCHECK(ctx, fcg::launch_random(dq.as<double>(), n, 12345, s));
    struct Ev {
        hipEvent_t e[5] = {};
        ~Ev() {
            for (auto x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    for (auto& x : ev.e) HIPCHECK(ctx, hipEventCreate(&x));
    // one untimed product to warm caches and code objects
    HIPCHECK(ctx, fcg::launch_mvp(j->view(), j->sc, dq.as<double>(), dv.as<double>(), k, 0, s));
    double acc[4] = {0, 0, 0, 0}, tot = 0;
    for (int r = 0; r < reps; ++r) {
        HIPCHECK(ctx, fcg::launch_mvp(j->view(), j->sc, dq.as<double>(), dv.as<double>(), k, 0, s, ev.e));
        HIPCHECK(ctx, hipEventSynchronize(ev.e[4]));
        for (int st = 0; st < 4; ++st) {
            float ms = 0;
            HIPCHECK(ctx, hipEventElapsedTime(&ms, ev.e[st], ev.e[st + 1]));
            acc[st] += ms;
        }
        float ms = 0;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ev.e[0], ev.e[4]));
        tot += ms;
    }
    out->ms_scan = acc[0] / reps;
    out->ms_charge = acc[1] / reps;
    out->ms_gather = acc[2] / reps;
    out->ms_down = acc[3] / reps;
    out->ms_total = tot / reps;
    out->reps = reps;
    out->k = k;
    out->variant = j->gather_variant;
    return FCG_OK;
}

int fcg_set_gather_variant(fcg_ctx* ctx, int job, int variant, int* n_variants) {
    Job* j = get_job(ctx, job);
    if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    if (n_variants) *n_variants = fcg::gather_variant_count();
    if (variant < 0) return FCG_OK;  // query only
    if (variant >= fcg::gather_variant_count()) return fail(ctx, FCG_INVALID_ARGUMENT, "bad gather variant");
    j->gather_variant = variant;
    j->tuned_k = -1;  // pinned by the caller: fcg_solve will not re-tune
    return FCG_OK;
}

int fcg_autotune(fcg_ctx* ctx, int job, int k, int reps, int* variant, double* best_ms) {
    Job* j = get_job(ctx, job);
    if (!j || k < 1 || reps < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int rc = ensure_scratch(ctx, *j, k);
    if (rc != FCG_OK) return rc;
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)j->N * k;
    DevBuf dq, dv;
    HIPCHECK(ctx, dq.alloc(sizeof(double) * n));
    HIPCHECK(ctx, dv.alloc(sizeof(double) * n));
    HIPCHECK(ctx, fcg::launch_random(dq.as<double>(), n, 777, s));
    struct Ev {
        hipEvent_t e[5] = {};
        ~Ev() {
            for (auto x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    for (auto& x : ev.e) HIPCHECK(ctx, hipEventCreate(&x));
    int best = 0;
    double best_t = 1e300;
    for (int var = 0; var < fcg::gather_variant_count(); ++var) {
        j->gather_variant = var;
        HIPCHECK(ctx, fcg::launch_mvp(j->view(), j->sc, dq.as<double>(), dv.as<double>(), k, 0, s));
        double t = 0.0;
        for (int r = 0; r < reps; ++r) {
            HIPCHECK(ctx, fcg::launch_mvp(j->view(), j->sc, dq.as<double>(), dv.as<double>(), k, 0, s, ev.e));
            HIPCHECK(ctx, hipEventSynchronize(ev.e[4]));
            float ms = 0;
            HIPCHECK(ctx, hipEventElapsedTime(&ms, ev.e[2], ev.e[3]));
            t += ms;
        }
        t /= reps;
        if (t < best_t) {
            best_t = t;
            best = var;
        }
    }
    j->gather_variant = best;
    j->tuned_k = k;
    if (variant) *variant = best;
    if (best_ms) *best_ms = best_t;
    return FCG_OK;
}

int fcg_precond_apply(fcg_ctx* ctx, int job, const double* q, double* pq, int k) {
    Job* j = get_job(ctx, job);
    if (!j || !q || !pq || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DevBuf dq, dp;
    const size_t n = (size_t)j->N * k;
    HIPCHECK(ctx, upload(dq, q, n, ctx->stream));
    HIPCHECK(ctx, dp.alloc(sizeof(double) * n));
    HIPCHECK(ctx, fcg::launch_jacobi(j->view(), dq.as<double>(), dp.as<double>(), k, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(pq, dp.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

// Right-preconditioned GMRES on all conductor right-hand sides at once
// (gmresPrecondSFastAll, Solver/SolveCapacitance.cpp:2288-2574): x0 = P b,
// r0 = b - A x0, Arnoldi with classical Gram-Schmidt and one re-orthogonalisation
// pass, Givens rotations per right-hand side on the host, the reference stop rule
// |g_{i+1}| / |b| < tol per right-hand side, x = x0 + P (Q z).
int fcg_solve(fcg_ctx* ctx, int job, const fcg_solve_opts* opts, double* cap, int* rows_out, int* cols_out,
              double* charges, fcg_solve_stats* stats) {
    using clk = std::chrono::steady_clock;
    Job* j = get_job(ctx, job);
    if (!j || !opts || !cap) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    if (opts->precond != FCG_PRECOND_NONE && opts->precond != FCG_PRECOND_JACOBI)
        return fail(ctx, FCG_INVALID_ARGUMENT, "unsupported preconditioner");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int N = j->N;
    const bool jac = opts->precond == FCG_PRECOND_JACOBI;
    auto t_start = clk::now();

    // right-hand sides: one per conductor; 2D skips the conductor that ends the leaf list
    std::vector<std::pair<int, int>> rhs;  // leaf range raised to 1 V
    int ncol = 0;
    for (int g = 0; g < j->G; ++g) {
        if (j->group_is_diel[(size_t)g]) continue;
        ++ncol;
        const int lo = j->group_leaf_start[(size_t)g], hi = j->group_leaf_start[(size_t)g + 1];
        if (j->dim == 2 && hi == N) continue;
        rhs.emplace_back(lo, hi);
    }
    const int k = (int)rhs.size();
    if (k == 0) return fail(ctx, FCG_INVALID_ARGUMENT, "no conductor right-hand side");
    int cap_iter = opts->max_iter > 0 ? opts->max_iter : 1000;
    if (cap_iter > N) cap_iter = N;
    int rc = ensure_scratch(ctx, *j, k);
    if (rc != FCG_OK) return rc;
    if (j->tuned_k != k && j->tuned_k >= 0 && (rc = fcg_autotune(ctx, job, k, 3, nullptr, nullptr)) != FCG_OK)
        return rc;
    JobView v = j->view();
    const size_t NK = (size_t)N * k;

    DevBuf b, x0, r, w, tmp, part, hbuf, nrm, dz, dcap;
    HIPCHECK(ctx, b.alloc(sizeof(double) * NK));
    HIPCHECK(ctx, x0.alloc(sizeof(double) * NK));
    HIPCHECK(ctx, r.alloc(sizeof(double) * NK));
    HIPCHECK(ctx, w.alloc(sizeof(double) * NK));
    HIPCHECK(ctx, tmp.alloc(sizeof(double) * NK));
    const int nchunks = fcg::dots_blocks(N);
    {
        std::vector<double> hb(NK, 0.0);
        for (int q = 0; q < k; ++q)
            for (int n = rhs[(size_t)q].first; n < rhs[(size_t)q].second; ++n) hb[(size_t)q * N + n] = 1.0;
        HIPCHECK(ctx, hipMemcpyAsync(b.p, hb.data(), sizeof(double) * NK, hipMemcpyHostToDevice, s));
    }
    // Krylov basis: [k][N] vectors allocated in chunks; device table of their addresses
    constexpr int kChunk = 64;
    std::vector<std::unique_ptr<DevBuf>> chunks;
    std::vector<double*> hptrs;
    hptrs.reserve((size_t)cap_iter + 2 + kChunk);
    DevBuf qptrs;
    HIPCHECK(ctx, qptrs.alloc(sizeof(double*) * (size_t)(cap_iter + 2 + kChunk)));
    auto ensure_basis = [&](size_t count) -> int {
        while (hptrs.size() < count) {
            chunks.emplace_back(new DevBuf());
            HIPCHECK(ctx, chunks.back()->alloc(sizeof(double) * NK * kChunk));
            const size_t base = hptrs.size();
            for (int c = 0; c < kChunk; ++c) hptrs.push_back(chunks.back()->as<double>() + (size_t)c * NK);
            // new table entries are not read by any enqueued kernel yet
            HIPCHECK(ctx, hipMemcpy(qptrs.as<double*>() + base, hptrs.data() + base, sizeof(double*) * kChunk,
                                    hipMemcpyHostToDevice));
        }
        return FCG_OK;
    };
    const int maxvals = (cap_iter + 2) * k;
    // h1 | h2 | |w|^2 of one Arnoldi step, contiguous so that one copy brings them back
    HIPCHECK(ctx, part.alloc(sizeof(double) * (size_t)maxvals * nchunks));
    HIPCHECK(ctx, hbuf.alloc(sizeof(double) * ((size_t)maxvals * 2 + k)));
    HIPCHECK(ctx, nrm.alloc(sizeof(double) * (size_t)k * 2));
    PinnedBuf hstage;
    HIPCHECK(ctx, hstage.alloc(sizeof(double) * ((size_t)maxvals * 2 + 2 * k)));
    int64_t n_mvp = 0;
    double t_mvp = 0.0;
    struct Events {
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } evs;
    HIPCHECK(ctx, hipEventCreate(&evs.a));
    HIPCHECK(ctx, hipEventCreate(&evs.b));
    hipEvent_t ev0 = evs.a, ev1 = evs.b;

    // squared norms of the k vectors of `vec` into out (synchronises the stream)
    DevBuf selfptr;
    HIPCHECK(ctx, selfptr.alloc(sizeof(double*)));
    double** norm_src = reinterpret_cast<double**>(hstage.as<double>() + (size_t)maxvals * 2 + k);
    auto norms = [&](const double* vec, std::vector<double>& out) -> int {
        HIPCHECK(ctx, hipStreamSynchronize(s));
        *norm_src = const_cast<double*>(vec);
        HIPCHECK(ctx, hipMemcpyAsync(selfptr.p, norm_src, sizeof(double*), hipMemcpyHostToDevice, s));
        HIPCHECK(ctx, fcg::launch_block_dots(selfptr.as<double* const>(), 1, vec, N, k, part.as<double>(), nchunks, s));
        HIPCHECK(ctx, fcg::launch_reduce_partials(part.as<double>(), nchunks, k, nrm.as<double>(), s));
        HIPCHECK(ctx, hipMemcpyAsync(hstage.p, nrm.p, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        HIPCHECK(ctx, hipStreamSynchronize(s));
        out.assign(hstage.as<double>(), hstage.as<double>() + k);
        for (auto& x : out) x = std::sqrt(x);
        return FCG_OK;
    };
    // divide the k vectors of `in` by the values `d` (host) into `out`
    auto scale_by = [&](const double* in, const std::vector<double>& d, double* out) -> int {
        HIPCHECK(ctx, hipStreamSynchronize(s));
        double* st = hstage.as<double>() + (size_t)maxvals * 2;
        for (int q = 0; q < k; ++q) st[q] = d[(size_t)q];
        HIPCHECK(ctx, hipMemcpyAsync(nrm.p, st, sizeof(double) * k, hipMemcpyHostToDevice, s));
        HIPCHECK(ctx, fcg::launch_scale(in, nrm.as<double>(), out, N, k, s));
        return FCG_OK;
    };

    std::vector<double> normb, normr;
    if (jac) {
        HIPCHECK(ctx, fcg::launch_jacobi(v, b.as<double>(), x0.as<double>(), k, s));
        HIPCHECK(ctx, fcg::launch_mvp(v, j->sc, x0.as<double>(), w.as<double>(), k, 0, s));
        ++n_mvp;
        HIPCHECK(ctx, fcg::launch_sub(b.as<double>(), w.as<double>(), r.as<double>(), N, k, s));
        if ((rc = norms(r.as<double>(), normr)) != FCG_OK) return rc;
        if ((rc = norms(b.as<double>(), normb)) != FCG_OK) return rc;
    } else {
        HIPCHECK(ctx, hipMemsetAsync(x0.p, 0, sizeof(double) * NK, s));
        HIPCHECK(ctx, hipMemcpyAsync(r.p, b.p, sizeof(double) * NK, hipMemcpyDeviceToDevice, s));
        if ((rc = norms(r.as<double>(), normr)) != FCG_OK) return rc;
        normb = normr;
    }
    if ((rc = ensure_basis(1)) != FCG_OK) return rc;
    if ((rc = scale_by(r.as<double>(), normr, hptrs[0])) != FCG_OK) return rc;

    // per-rhs Givens state
    std::vector<std::vector<double>> H((size_t)k);  // column i at H[r][i*ld .. i*ld+i]
    std::vector<std::vector<double>> cs((size_t)k), sn((size_t)k), g((size_t)k);
    std::vector<int> done((size_t)k, 0), iters((size_t)k, -1);
    std::vector<double> rel((size_t)k, 0.0);
    const int ld = cap_iter + 1;
    for (int q = 0; q < k; ++q) {
        g[(size_t)q].assign((size_t)cap_iter + 2, 0.0);
        g[(size_t)q][0] = normr[(size_t)q];
        rel[(size_t)q] = normr[(size_t)q] / normb[(size_t)q];
        if (rel[(size_t)q] < opts->tol) done[(size_t)q] = 1;  // the initial guess is good enough
    }
    auto all_done = [&]() {
        for (int d : done)
            if (!d) return false;
        return true;
    };
    std::vector<double> hn((size_t)k);
    int i = 0;
    for (i = 0; i < cap_iter && !all_done(); ++i) {
        const int nv = i + 1;
        if ((rc = ensure_basis((size_t)nv + 1)) != FCG_OK) return rc;
        double* h1 = hbuf.as<double>();
        double* h2 = h1 + (size_t)nv * k;
        double* w2 = h2 + (size_t)nv * k;
        // w = A P q_i
        HIPCHECK(ctx, hipEventRecord(ev0, s));
        HIPCHECK(ctx, fcg::launch_mvp(v, j->sc, hptrs[(size_t)i], w.as<double>(), k, jac ? 1 : 0, s));
        HIPCHECK(ctx, hipEventRecord(ev1, s));
        ++n_mvp;
        // classical Gram-Schmidt, two passes; the second pass also returns |w|^2
        HIPCHECK(ctx, fcg::launch_block_dots(qptrs.as<double* const>(), nv, w.as<double>(), N, k, part.as<double>(),
                                             nchunks, s));
        HIPCHECK(ctx, fcg::launch_reduce_partials(part.as<double>(), nchunks, nv * k, h1, s));
        HIPCHECK(ctx, fcg::launch_basis_update(qptrs.as<double* const>(), nv, h1, w.as<double>(), N, k, s));
        HIPCHECK(ctx, fcg::launch_block_dots(qptrs.as<double* const>(), nv, w.as<double>(), N, k, part.as<double>(),
                                             nchunks, s));
        HIPCHECK(ctx, fcg::launch_reduce_partials(part.as<double>(), nchunks, nv * k, h2, s));
        HIPCHECK(ctx, fcg::launch_basis_update_norm(qptrs.as<double* const>(), nv, h2, w.as<double>(), N, k,
                                                    part.as<double>(), nchunks, s));
        HIPCHECK(ctx, fcg::launch_reduce_partials(part.as<double>(), nchunks, k, w2, s));
        HIPCHECK(ctx, hipMemcpyAsync(hstage.p, h1, sizeof(double) * ((size_t)2 * nv * k + k), hipMemcpyDeviceToHost,
                                     s));
        HIPCHECK(ctx, hipStreamSynchronize(s));
        {
            float ms = 0.0f;
            HIPCHECK(ctx, hipEventElapsedTime(&ms, ev0, ev1));
            t_mvp += 1e-3 * ms;
        }
        const double* hh1 = hstage.as<double>();
        const double* hh2 = hh1 + (size_t)nv * k;
        const double* hw2 = hh2 + (size_t)nv * k;
        for (int q = 0; q < k; ++q) hn[(size_t)q] = std::sqrt(hw2[q]);
        for (int q = 0; q < k; ++q) {
            if (done[(size_t)q]) continue;
            std::vector<double>& Hq = H[(size_t)q];
            if (Hq.size() < (size_t)(i + 1) * ld) Hq.resize((size_t)(i + 1) * ld, 0.0);
            double* colv = Hq.data() + (size_t)i * ld;
            for (int jj = 0; jj <= i; ++jj) colv[jj] = hh1[(size_t)jj * k + q] + hh2[(size_t)jj * k + q];
            auto& c = cs[(size_t)q];
            auto& sv = sn[(size_t)q];
            for (int jj = 0; jj < i; ++jj) {
                const double t1 = colv[jj], t2 = colv[jj + 1];
                colv[jj] = c[(size_t)jj] * t1 - sv[(size_t)jj] * t2;
                colv[jj + 1] = c[(size_t)jj] * t2 + sv[(size_t)jj] * t1;
            }
            const double t1 = colv[i], t2 = hn[(size_t)q];
            const double len = std::sqrt(t1 * t1 + t2 * t2);
            c.push_back(t1 / len);
            sv.push_back(-t2 / len);
            colv[i] = c[(size_t)i] * t1 - sv[(size_t)i] * t2;
            const double gi = g[(size_t)q][(size_t)i];
            g[(size_t)q][(size_t)i] = c[(size_t)i] * gi;
            g[(size_t)q][(size_t)i + 1] = sv[(size_t)i] * gi;
            rel[(size_t)q] = std::fabs(g[(size_t)q][(size_t)i + 1]) / normb[(size_t)q];
            iters[(size_t)q] = i;
            if (rel[(size_t)q] < opts->tol) done[(size_t)q] = 1;
        }
        if (all_done()) break;
        if ((rc = scale_by(w.as<double>(), hn, hptrs[(size_t)i + 1])) != FCG_OK) return rc;
    }
    // back substitution per rhs, z[jj][r] (zero beyond each rhs's own count)
    const int nvec_total = (int)hptrs.size();
    std::vector<double> z((size_t)nvec_total * k, 0.0);
    int max_it = 0;
    for (int q = 0; q < k; ++q) {
        const int it = iters[(size_t)q];
        if (it < 0) continue;
        if (it + 1 > max_it) max_it = it + 1;
        std::vector<double> zz(g[(size_t)q].begin(), g[(size_t)q].begin() + it + 1);
        const std::vector<double>& Hq = H[(size_t)q];
        for (int ii = it; ii >= 0; --ii) {
            zz[(size_t)ii] = zz[(size_t)ii] / Hq[(size_t)ii * ld + ii];
            for (int jj = ii - 1; jj >= 0; --jj) zz[(size_t)jj] -= Hq[(size_t)ii * ld + jj] * zz[(size_t)ii];
        }
        for (int jj = 0; jj <= it; ++jj) z[(size_t)jj * k + q] = zz[(size_t)jj];
    }
    HIPCHECK(ctx, dz.alloc(sizeof(double) * z.size()));
    HIPCHECK(ctx, hipMemcpyAsync(dz.p, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice, s));
    HIPCHECK(ctx, hipMemsetAsync(tmp.p, 0, sizeof(double) * NK, s));
    if (max_it > 0)
        HIPCHECK(ctx, fcg::launch_basis_combine(qptrs.as<double* const>(), max_it, dz.as<double>(), tmp.as<double>(),
                                                N, k, s));
    // x = x0 + P y
    double* xout = r.as<double>();
    if (jac) {
        HIPCHECK(ctx, fcg::launch_jacobi(v, tmp.as<double>(), xout, k, s));
    } else {
        HIPCHECK(ctx, hipMemcpyAsync(xout, tmp.p, sizeof(double) * NK, hipMemcpyDeviceToDevice, s));
    }
    HIPCHECK(ctx, fcg::launch_add(x0.as<double>(), xout, xout, N, k, s));
    HIPCHECK(ctx, dcap.alloc(sizeof(double) * (size_t)k * ncol));
    HIPCHECK(ctx, fcg::launch_charge_sums(xout, j->leaf_w.as<double>(), j->col_range.as<int32_t>(), N, k, ncol,
                                          dcap.as<double>(), s));
    HIPCHECK(ctx, hipMemcpyAsync(cap, dcap.p, sizeof(double) * (size_t)k * ncol, hipMemcpyDeviceToHost, s));
    if (charges)
        HIPCHECK(ctx, hipMemcpyAsync(charges, xout, sizeof(double) * NK, hipMemcpyDeviceToHost, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    if (rows_out) *rows_out = k;
    if (cols_out) *cols_out = ncol;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->n_rhs = k;
        double worst = 0.0;
        int mx = 0;
        for (int q = 0; q < k; ++q) {
            const int it = iters[(size_t)q] + 1;
            if (q < 64) stats->iterations[q] = it;
            if (it > mx) mx = it;
            if (rel[(size_t)q] > worst) worst = rel[(size_t)q];
        }
        stats->max_iterations = mx;
        stats->final_residual = worst;
        stats->n_mvp = n_mvp;
        stats->t_mvp_s = t_mvp;
        stats->t_solve_s = std::chrono::duration<double>(clk::now() - t_start).count();
    }
    return FCG_OK;
}

}  // extern "C"
