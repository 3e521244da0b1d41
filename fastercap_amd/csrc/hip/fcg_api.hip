// C-ABI of the gfx950 FasterCap solve path (declarations and reference citations
// in include/fastercap_amd.h).  Owns the HBM-resident jobs and a batched GMRES
// engine that advances any set of solves together, one launch per stage per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kernels.h"

using fcg::BatchView;
using fcg::BJob;
using fcg::Chunk;
using fcg::KJob;
using fcg::KrylovView;
using fcg::Seg;

namespace {

using clk = std::chrono::steady_clock;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        release();
        bytes = n;
        if (n == 0) return hipSuccess;
        return hipMalloc(&p, n);
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
};

struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    void* dev = nullptr;  // the device's address of a coherent buffer
    hipError_t alloc(size_t n, bool coherent = false) {
        if (p) (void)hipHostFree(p);
        p = dev = nullptr;
        bytes = n;
        hipError_t e = hipHostMalloc(&p, n, coherent ? hipHostMallocCoherent : hipHostMallocDefault);
        if (e == hipSuccess && coherent) e = hipHostGetDevicePointer(&dev, p, 0);
        return e;
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

// Caching device allocator for solve state: freed blocks are kept per rounded size
// and reused by later solves (serving keeps allocating and releasing solves).
struct Pool {
    std::map<size_t, std::vector<void*>> free_;
    int64_t held = 0, in_use = 0;
    static size_t round(size_t n) {
        if (n <= 4096) return 4096;
        size_t p2 = 1;
        while (p2 < n) p2 <<= 1;
        const size_t step = std::max<size_t>(p2 / 8, 4096);
        return (n + step - 1) / step * step;
    }
    hipError_t get(size_t n, void** out, size_t* rounded) {
        const size_t r = round(n);
        *rounded = r;
        auto it = free_.find(r);
        if (it != free_.end() && !it->second.empty()) {
            *out = it->second.back();
            it->second.pop_back();
            in_use += (int64_t)r;
            return hipSuccess;
        }
        hipError_t e = hipMalloc(out, r);
        if (e == hipSuccess) {
            held += (int64_t)r;
            in_use += (int64_t)r;
        }
        return e;
    }
    void put(void* p, size_t r) {
        if (!p) return;
        free_[r].push_back(p);
        in_use -= (int64_t)r;
    }
    ~Pool() {
        for (auto& kv : free_)
            for (void* p : kv.second) (void)hipFree(p);
    }
};

// Pool-backed device block; released to the pool (not to HIP) when dropped.
struct PBuf {
    void* p = nullptr;
    size_t r = 0;
    Pool* pool = nullptr;
    PBuf() = default;
    PBuf(const PBuf&) = delete;
    PBuf& operator=(const PBuf&) = delete;
    ~PBuf() { release(); }
    void release() {
        if (p && pool) pool->put(p, r);
        p = nullptr;
        r = 0;
    }
    hipError_t get(Pool* pl, size_t n) {
        release();
        pool = pl;
        return pl->get(n, &p, &r);
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

// Per-step tables: slabs of device memory with paired pinned staging,
// bump-allocated and reset once the step's work has completed.
struct Tables {
    struct Slab {
        DevBuf dev;
        PinnedBuf host;
        size_t used = 0;
    };
    std::vector<std::unique_ptr<Slab>> slabs;
    size_t cur = 0;
    hipError_t take(size_t bytes, char** h, char** d) {
        const size_t need = (bytes + 255) & ~(size_t)255;
        while (cur < slabs.size() && slabs[cur]->used + need > slabs[cur]->dev.bytes) ++cur;
        if (cur == slabs.size()) {
            auto sl = std::make_unique<Slab>();
            const size_t sz = std::max<size_t>(need, (size_t)8 << 20);
            hipError_t e = sl->dev.alloc(sz);
            if (e != hipSuccess) return e;
            e = sl->host.alloc(sz, true);
            if (e != hipSuccess) return e;
            slabs.push_back(std::move(sl));
        }
        Slab& sl = *slabs[cur];
        *h = sl.host.as<char>() + sl.used;
        *d = sl.dev.as<char>() + sl.used;
        sl.used += need;
        return hipSuccess;
    }
    // copy host data into a fresh device region: enqueued on s at once, or, in
    // deferred mode, by the next flush (one copy per slab for a whole step)
    bool defer = false;
    hipError_t push(const void* data, size_t bytes, hipStream_t s, void** dptr) {
        char *h, *d;
        hipError_t e = take(bytes, &h, &d);
        if (e != hipSuccess) return e;
        *dptr = d;
        if (!bytes) return hipSuccess;
        std::memcpy(h, data, bytes);
        if (defer) return hipSuccess;
        return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s);
    }
    // deferred mode: one copy kernel per slab on s, the stream of the step's
    // kernels (slab sizes are multiples of 256 bytes)
    hipError_t flush(hipStream_t s) {
        for (auto& sl : slabs) {
            if (!sl->used) continue;
            hipError_t e = fcg::launch_copy_in(sl->host.dev, sl->dev.p, sl->used, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    // pinned host area for a device-to-host copy (valid until reset)
    hipError_t host_area(size_t bytes, void** hptr) {
        char *h, *d;
        hipError_t e = take(bytes, &h, &d);
        *hptr = h;
        return e;
    }
    void reset() {
        for (auto& s : slabs) s->used = 0;
        cur = 0;
    }
};

// Scratch of one hierarchical product on k right-hand sides: node charges and
// node potentials.  Every concurrently running solve owns one.
struct Scratch {
    int k = 0, kp = 0;
    PBuf charge64, charge32, pot64, pot32, lt;
    size_t bytes() const { return charge64.r + charge32.r + pot64.r + pot32.r + lt.r; }
};

struct Job {
    int dim = 3, N = 0, M = 0, G = 0;
    int scheme = FCG_SCHEME_COLLOCATION;  // discretization (fcg_job_desc::scheme)
    int64_t L = 0;
    DevBuf nodes, left, right, parent, leaf_lo, leaf_hi, leaf_node, row_ptr, col, link_row, gout, gin, leaf_w,
        col_range, self, diag, coef64, coef32, perm, iperm, rorder, fill_order, fill_start, rhs_range, col16, grow,
        src_geom, fill_rs;
    // tree passes of the product (kernels.h BJob): inner nodes per chunk and
    // level, top nodes per level, top flags, group roots
    DevBuf tb_rec, tb_ppm, tb_lev, tb_lvoff, tt_node, tt_rec, tt_rows, tt_lev, ttop, troots, lpath;
    int tt_nlev = 0, ntop = 0, nroots = 0, lpath_len = 0, lpath_top = 0;
    // host copy of the row pointers and, for every stored link, its offset within
    // its row in the descriptor's order (the device keeps each row sorted by source)
    std::vector<int64_t> h_row_ptr;
    std::vector<int64_t> fill_cnt;  // links per fill class
    // host copies of the tree (block preconditioner partition)
    std::vector<int32_t> h_left, h_right, h_parent, h_leaf_lo, h_leaf_hi;
    // block preconditioner of the last block size asked for (0: none yet)
    struct BlockPre {
        int size = 0, nb = 0;
        DevBuf bstart, boff, P;
    } bpre;
    // two-level preconditioner of the last dimension asked for (0: none yet)
    struct SuperPre {
        int dim = 0, m = 0;  // dim: Sinv valid for this dimension (0: to be built)
        int part_dim = 0;    // dimension of the partition, snode .. w
        bool avail = false;
        DevBuf snode, sstart, sidx, Sinv, w;
    } spre;
    // complex permittivities: imaginary parts of the group permittivities and of
    // the leaf charge weights, the imaginary self terms (filled with the real ones)
    bool cplx = false;
    DevBuf gout_im, gin_im, leaf_w_im, self_im;
    // length of a solve's vectors: [Re | Im] for complex jobs
    int vlen() const { return cplx ? 2 * N : N; }
    // product vectors per right-hand side: a complex vector is two real ones
    int pk(int k) const { return cplx ? 2 * k : k; }
    // staged gather: source list of every gather tile and each link's index in
    // its tile's list (built before the first product or solve; jobs with 16-bit
    // source indices only); max_tile_src is the longest list, -1 until built
    DevBuf tsrc_off, tsrc, lsrc;
    int max_tile_src = -1;
    // FP64 near-field side list (built on the first mixed fill; kernels.h BJob)
    DevBuf near_ptr, near_col, near_link, near_dc;
    int64_t Ln = 0;
    bool near_ready = false;
    std::vector<uint32_t> link_pos;  // descriptor-order offset of each stored link (empty: rows by node id)
    std::vector<int32_t> h_iperm;     // charge row -> node
    bool filled64 = false, filled32 = false;
    std::vector<int32_t> group_is_diel, group_leaf_start;
    std::vector<std::pair<int, int>> rhs;  // leaf range raised to 1 V per right-hand side
    int ncol = 0;
    int first_cond_leaf = 0;
    Scratch scr;  // product scratch of the stand-alone entry points
    // gather lane shape per product precision (index fcg::MvpPrec) and whether it
    // was chosen by measurement (1), pinned by the caller (-1) or is the default (0)
    int gather_variant[3] = {0, 0, 0};
    int tuned[3] = {0, 0, 0};

    int64_t resident_bytes() const {
        return (int64_t)(nodes.bytes + left.bytes + right.bytes + parent.bytes + leaf_lo.bytes + leaf_hi.bytes +
                         leaf_node.bytes + row_ptr.bytes + col.bytes + link_row.bytes + gout.bytes + gin.bytes +
                         leaf_w.bytes + col_range.bytes + self.bytes + diag.bytes + coef64.bytes + coef32.bytes +
                         perm.bytes + iperm.bytes + rorder.bytes + fill_order.bytes + fill_start.bytes + rhs_range.bytes +
                         col16.bytes + grow.bytes + src_geom.bytes + fill_rs.bytes + tb_rec.bytes +
                         tb_ppm.bytes + tb_lev.bytes + lpath.bytes +
                         tb_lvoff.bytes + tt_node.bytes + tt_rec.bytes + tt_rows.bytes + tt_lev.bytes + ttop.bytes + troots.bytes +
                         bpre.bstart.bytes + bpre.boff.bytes + bpre.P.bytes + spre.snode.bytes +
                         spre.sstart.bytes + spre.sidx.bytes + spre.Sinv.bytes + spre.w.bytes +
                         near_ptr.bytes + near_col.bytes + near_link.bytes + near_dc.bytes + gout_im.bytes +
                         gin_im.bytes + leaf_w_im.bytes + self_im.bytes + tsrc_off.bytes + tsrc.bytes +
                         lsrc.bytes + scr.bytes());
    }

    BJob rec(const Scratch& sc, const double* q, double* v, const float* q32 = nullptr) const {
        BJob b{};
        b.N = N;
        b.M = M;
        b.k = sc.k;
        b.kp = sc.kp;
        b.dim = dim;
        b.scheme = scheme;
        b.first_cond_leaf = first_cond_leaf;
        b.L = L;
        b.nodes = nodes.as<fcg_node_geom>();
        b.row_ptr = row_ptr.as<int64_t>();
        b.col = col.as<int32_t>();
        b.col16 = col16.as<uint16_t>();
        b.perm = perm.as<int32_t>();
        b.iperm = iperm.as<int32_t>();
        b.rorder = rorder.as<int32_t>();
        b.tb_rec = tb_rec.as<int4>();
        b.tb_ppm = tb_ppm.as<int32_t>();
        b.lpath = lpath.as<int32_t>();
        b.lpath_len = lpath_len;
        b.lpath_top = lpath_top;
        b.tb_lev = tb_lev.as<int32_t>();
        b.tb_lvoff = tb_lvoff.as<int32_t>();
        b.tt_node = tt_node.as<int32_t>();
        b.tt_rec = tt_rec.as<int4>();
        b.tt_rows = tt_rows.as<int4>();
        b.tt_lev = tt_lev.as<int32_t>();
        b.ttop = ttop.as<uint8_t>();
        b.troots = troots.as<int32_t>();
        b.tt_nlev = tt_nlev;
        b.ntop = ntop;
        b.nroots = nroots;
        b.grow = grow.as<fcg::GRow>();
        b.tsrc_off = max_tile_src >= 0 ? tsrc_off.as<int32_t>() : nullptr;
        b.tsrc = tsrc.as<uint16_t>();
        b.lsrc = lsrc.as<uint16_t>();
        b.src_geom = src_geom.as<fcg::SrcGeom>();
        b.fill_order = fill_order.as<uint32_t>();
        b.fill_rs = fill_rs.as<int2>();
        b.fill_start = fill_start.as<int64_t>();
        b.link_row = link_row.as<int32_t>();
        b.coef64 = coef64.as<double>();
        b.coef32 = coef32.as<float>();
        b.near_ptr = near_ready ? near_ptr.as<int64_t>() : nullptr;
        b.near_col = near_col.as<int32_t>();
        b.near_link = near_link.as<uint32_t>();
        b.near_dc = near_dc.as<double>();
        b.diag = diag.as<double>();
        b.self = self.as<double>();
        b.self_im = cplx ? self_im.as<double>() : nullptr;
        b.gout_im = gout_im.as<double>();
        b.gin_im = gin_im.as<double>();
        b.left = left.as<int32_t>();
        b.right = right.as<int32_t>();
        b.leaf_lo = leaf_lo.as<int32_t>();
        b.leaf_hi = leaf_hi.as<int32_t>();
        b.parent = parent.as<int32_t>();
        b.leaf_node = leaf_node.as<int32_t>();
        b.gout = gout.as<double>();
        b.gin = gin.as<double>();
        b.charge64 = sc.charge64.as<double>();
        b.charge32 = sc.charge32.as<float>();
        b.pot64 = sc.pot64.as<double>();
        b.pot32 = sc.pot32.as<float>();
        b.lt = sc.lt.as<double>();
        b.q = q;
        b.q32 = q32;
        b.v = v;
        return b;
    }
};

}  // namespace

// Scratch of fcg_mvp_device_noalloc: one static device array, present on every
// GPU the library runs on until the process ends.  Host code never uses its name
// as an address: each device's copy has its own address, read once per device
// with hipGetSymbolAddress (scratch_address) and passed to the kernels.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on `device`, which must be the current device
hipError_t scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// The job table of a no-alloc product, written into the scratch array by a
// kernel from its by-value argument (no staging buffer, no copy call).
struct NoAllocTable {
    BJob job;
    int64_t off[10];  // link, tree-chunk, node, leaf and rhs offsets of a batch of one
};
static_assert(sizeof(NoAllocTable) % 8 == 0, "table copied in 8-byte words");
constexpr size_t kNoAllocTableBytes = (sizeof(NoAllocTable) + 255) & ~(size_t)255;

__global__ void k_noalloc_table(NoAllocTable t, uint64_t* __restrict__ dst) {
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&t);
    for (unsigned i = threadIdx.x; i < sizeof(NoAllocTable) / 8; i += blockDim.x) dst[i] = src[i];
}
}  // namespace

struct fcg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t down_stream = nullptr;  // engine result downloads (default priority)
    void* noalloc = nullptr;            // this device's address of `scratch`
    bool profiling = false;
    fcg_profile prof{};
    // gather lane shape of products over several jobs, per product precision and
    // rhs class (-1: the class's default shape)
    int batch_variant[3][fcg::kGatherClasses];
    Pool pool;  // declared before everything that returns blocks to it
    std::vector<std::unique_ptr<Job>> jobs;
    std::string err;
    Tables tabs;  // tables of stand-alone calls (reset per call)
    // engine step tables kept between engines (pinned staging is slow to allocate)
    std::vector<std::unique_ptr<Tables>> spare_tables;
    // result of the last fcg_links_discover, until fcg_links_fetch
    struct LinkOut {
        std::vector<int64_t> row_ptr;
        DevBuf col;
        int64_t L = 0;
        bool ready = false;
        void reset() {
            row_ptr = std::vector<int64_t>();
            col.release();
            L = 0;
            ready = false;
        }
    } links;
    // result of the last fcg_refine, until fcg_refine_fetch
    struct RefineOut {
        DevBuf pool;
        int n = 0;
        bool ready = false;
    } refined;

    fcg_ctx() {
        for (auto& row : batch_variant)
            for (int& v : row) v = -1;
    }
};

namespace {

int fail(fcg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHECK(ctx, call)                                                                    \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            return fail(ctx, e_ == hipErrorOutOfMemory ? FCG_OUT_OF_MEMORY : FCG_DEVICE_ERROR, \
                        std::string(#call) + ": " + hipGetErrorString(e_));                    \
        }                                                                                      \
    } while (0)

#define RCCHECK(call)                  \
    do {                               \
        int rc_ = (call);              \
        if (rc_ != FCG_OK) return rc_; \
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

// product scratch for k right-hand sides of an (N leaves, M nodes) job
int ensure_scratch(fcg_ctx* ctx, Scratch& sc, int N, int M, int k, bool need32) {
    (void)N;
    const int kp = fcg::padded_k(k);
    if (sc.k == k && (!need32 || sc.charge32.p)) return FCG_OK;
    // buffers already in use may be read by queued kernels: drain before replacing
    if (sc.k != 0) HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    Pool* pl = &ctx->pool;
    if (sc.k != k) {
        HIPCHECK(ctx, sc.charge64.get(pl, sizeof(double) * (size_t)kp * M));
        HIPCHECK(ctx, sc.pot64.get(pl, sizeof(double) * (size_t)kp * M));
        HIPCHECK(ctx, sc.lt.get(pl, sizeof(double) * 2 * (size_t)k));
        sc.charge32.release();
        sc.pot32.release();
    }
    if (need32 && !sc.charge32.p) {
        HIPCHECK(ctx, sc.charge32.get(pl, sizeof(float) * (size_t)kp * M));
        HIPCHECK(ctx, sc.pot32.get(pl, sizeof(float) * (size_t)kp * M));
    }
    sc.k = k;
    sc.kp = kp;
    return FCG_OK;
}

// Upload a batch table for the given (job, q, v) triples into `tabs`.
struct BatchEntry {
    Job* j;
    const Scratch* sc;
    const double* q;
    double* v;
    const float* q32 = nullptr;  // FP32 input instead of q
};
// Jobs are listed by row stride (stable), so every rhs class owns one contiguous
// range of batch rows; `plan` (optional) receives those ranges with no lane shape
// chosen yet (see choose_variants).
int make_batch(fcg_ctx* ctx, Tables& tabs, const std::vector<BatchEntry>& ents, BatchView* out,
               fcg::GatherPlan* plan = nullptr, std::vector<int>* order_out = nullptr) {
    std::memset(out, 0, sizeof *out);
    if (plan) plan->n = 0;
    const int n = (int)ents.size();
    if (n == 0) return FCG_OK;
    std::vector<int> order((size_t)n);
    for (int i = 0; i < n; ++i) order[(size_t)i] = i;
    // by row stride, then 16-bit before 32-bit source indices
    auto wide = [&](int i) { return ents[(size_t)i].j->col16.p ? 0 : 1; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const int ka = ents[(size_t)a].sc->kp, kb = ents[(size_t)b].sc->kp;
        return ka != kb ? ka < kb : wide(a) < wide(b);
    });
    if (order_out) *order_out = order;
    std::vector<BJob> recs((size_t)n);
    std::vector<int64_t> off((size_t)6 * (n + 1), 0);
    int64_t* lo = off.data();
    int64_t* to = lo + (n + 1);
    int64_t* no = to + (n + 1);
    int64_t* fo = no + (n + 1);
    int64_t* ko = fo + (n + 1);
    int64_t* go = ko + (n + 1);
    int kmax = 0, any2d = 0, anycplx = 0, kpmin = 1 << 30;
    for (int i = 0; i < n; ++i) {
        const BatchEntry& e = ents[(size_t)order[(size_t)i]];
        const Job& J = *e.j;
        const Scratch& sc = *e.sc;
        recs[(size_t)i] = J.rec(sc, e.q, e.v, e.q32);
        lo[i + 1] = lo[i] + J.L;
        to[i + 1] = to[i] + fcg::tree_chunks(J.M);
        no[i + 1] = no[i] + J.M;
        fo[i + 1] = fo[i] + J.N;
        ko[i + 1] = ko[i] + sc.k;
        go[i + 1] = go[i] + fcg::gather_tiles(J.M);
        kmax = std::max(kmax, sc.k);
        kpmin = std::min(kpmin, std::max(sc.kp, 1));
        any2d |= (J.dim == 2);
        anycplx |= J.cplx ? 1 : 0;
        if (plan) {
            const int cls = fcg::gather_class(std::max(sc.kp, 1));
            const int i16 = J.col16.p ? 1 : 0;
            if (plan->n == 0 || plan->seg[plan->n - 1].cls != cls || plan->seg[plan->n - 1].idx16 != i16) {
                fcg::GatherSeg& g = plan->seg[plan->n++];
                g.row_lo = no[i];
                g.tile_lo = go[i];
                g.staged = 1;
                g.stage_floats = 1;
                g.kpmin = std::max(sc.kp, 1);
                g.kmax = 0;
                g.cls = cls;
                g.variant = -1;
                g.idx16 = i16;
            }
            fcg::GatherSeg& g = plan->seg[plan->n - 1];
            g.row_hi = no[i + 1];
            g.tile_hi = go[i + 1];
            g.kmax = std::max(g.kmax, sc.k);
            // staged when every job has its lists (launch_mvp checks that the
            // longest one fits in LDS in the product's charge type)
            const int64_t fl = (int64_t)J.max_tile_src * std::max(sc.kp, 1);
            if (J.max_tile_src < 0 || fl > (int64_t)fcg::kStageMaxBytes) g.staged = 0;
            else g.stage_floats = (int)std::max<int64_t>(g.stage_floats, fl);
        }
    }
    void *dr, *doff;
    HIPCHECK(ctx, tabs.push(recs.data(), sizeof(BJob) * n, ctx->stream, &dr));
    HIPCHECK(ctx, tabs.push(off.data(), sizeof(int64_t) * off.size(), ctx->stream, &doff));
    int64_t* d = reinterpret_cast<int64_t*>(doff);
    out->jobs = reinterpret_cast<const BJob*>(dr);
    out->njobs = n;
    out->link_off = d;
    out->tchunk_off = d + (n + 1);
    out->node_off = d + 2 * (n + 1);
    out->leaf_off = d + 3 * (n + 1);
    out->rhs_off = d + 4 * (n + 1);
    out->gtile_off = d + 5 * (n + 1);
    out->links = lo[n];
    out->tchunks = to[n];
    out->nodes = no[n];
    out->leaves = fo[n];
    out->rhs = ko[n];
    out->gtiles = go[n];
    out->kmax = kmax;
    out->kpmin = kpmin;
    out->any2d = any2d;
    out->anycplx = anycplx;
    return FCG_OK;
}

// Lane shape of every class range of a product plan: a batch of one job uses the
// job's own shape, a batch of several the context's shape for each class.
void choose_variants(const fcg_ctx* ctx, int prec, const std::vector<BatchEntry>& ents, fcg::GatherPlan& plan) {
    for (int i = 0; i < plan.n; ++i) {
        fcg::GatherSeg& g = plan.seg[i];
        if (ents.size() == 1) {
            g.variant = ents[0].j->gather_variant[prec];
        } else {
            const int v = ctx->batch_variant[prec][g.cls];
            g.variant = v >= 0 ? v : fcg::default_gather_variant(prec, g.cls);
        }
    }
}

// Algorithmic bytes of the gather of one product over a batch (see fcg_profile).
double gather_alg_bytes(const std::vector<BatchEntry>& ents, int prec) {
    const double bc = prec == fcg::kMvp64 ? 8.0 : 4.0;
    const double bn = prec == fcg::kMvp32 ? 4.0 : 8.0;
    double b = 0.0;
    for (const auto& e : ents) {
        const double bi = e.j->col16.p ? 2.0 : 4.0;  // source index width
        b += (double)e.j->L * (bc + bi) + 16.0 * e.j->M + 8.0 + 2.0 * e.j->M * (double)e.sc->k * bn;
        // the residual product's FP64 near-field correction: its side list
        if (prec == fcg::kMvp32Acc && e.j->near_ready) b += 12.0 * (double)e.j->Ln + 8.0 * (e.j->M + 1);
    }
    return b;
}

// Events around one product of the engine when the context is profiling.
struct ProfEvents {
    hipEvent_t e[5] = {};
    bool on = false;
    int prec = 0, kernels = 0;
    double bytes = 0.0;
    ~ProfEvents() {
        for (auto x : e)
            if (x) (void)hipEventDestroy(x);
    }
    hipError_t arm(bool enable, int p, double b, int nkernels) {
        on = enable;
        prec = p;
        bytes = b;
        kernels = nkernels;
        if (!on) return hipSuccess;
        for (auto& x : e)
            if (!x) {
                hipError_t r = hipEventCreate(&x);
                if (r != hipSuccess) return r;
            }
        return hipSuccess;
    }
    hipEvent_t* events() { return on ? e : nullptr; }
    // after the stream has been synchronised
    void collect(fcg_profile& pr) {
        if (!on) return;
        float g = 0.f, t = 0.f;
        if (hipEventElapsedTime(&g, e[2], e[3]) != hipSuccess) return;
        if (hipEventElapsedTime(&t, e[0], e[4]) != hipSuccess) return;
        pr.gather_ms[prec] += g;
        pr.gather_bytes[prec] += bytes;
        pr.gather_launches[prec] += 1;
        pr.gather_kernels[prec] += kernels;
        pr.mvp_ms[prec] += t;
        pr.mvp_launches[prec] += 1;
        on = false;
    }
};

// Validate the structural invariants the kernels rely on (bounds of every index).
int validate_desc(fcg_ctx* ctx, const fcg_job_desc* d, bool check_cols = true) {
    if (!d || !d->nodes || !d->left || !d->right || !d->parent || !d->leaf_lo || !d->leaf_hi || !d->leaf_node ||
        !d->row_ptr || (check_cols && d->n_links > 0 && !d->col) || !d->group_is_diel || !d->group_leaf_start ||
        !d->group_outperm || !d->group_inperm || !d->leaf_outperm)
        return fail(ctx, FCG_INVALID_ARGUMENT, "null array in job descriptor");
    if (d->dim != 2 && d->dim != 3) return fail(ctx, FCG_INVALID_ARGUMENT, "dim must be 2 or 3");
    if (d->scheme != 0 && d->scheme != FCG_SCHEME_COLLOCATION && d->scheme != FCG_SCHEME_GALERKIN)
        return fail(ctx, FCG_INVALID_ARGUMENT, "unknown discretization scheme");
    const int N = d->n_leaves, M = d->n_nodes;
    if (N < 1 || M < N || d->n_groups < 1 || d->n_links < 0)
        return fail(ctx, FCG_INVALID_ARGUMENT, "bad sizes in job descriptor");
    if (d->n_links > (int64_t)UINT32_MAX)
        return fail(ctx, FCG_INVALID_ARGUMENT, "more than 2^32 links in one job (link indices are 32-bit)");
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
        if (!leaf && (d->leaf_lo[i] != d->leaf_lo[d->left[i]] || d->leaf_hi[i] != d->leaf_hi[d->right[i]] ||
                      d->leaf_hi[d->left[i]] != d->leaf_lo[d->right[i]]))
            return fail(ctx, FCG_INVALID_ARGUMENT, "leaf ranges do not nest");
        if (d->nodes[i].cond < 0 || d->nodes[i].cond >= d->n_groups)
            return fail(ctx, FCG_INVALID_ARGUMENT, "bad node group");
        if (!(d->nodes[i].maxside > 0.0)) return fail(ctx, FCG_INVALID_ARGUMENT, "non-positive node size");
    }
    for (int k = 0; k < N; ++k) {
        const int nd = d->leaf_node[k];
        if (nd < 0 || nd >= M || d->left[nd] >= 0 || d->leaf_lo[nd] != k)
            return fail(ctx, FCG_INVALID_ARGUMENT, "bad leaf_node map");
    }
    if (check_cols)
        for (int64_t l = 0; l < d->n_links; ++l)
            if (d->col[l] < 0 || d->col[l] >= M) return fail(ctx, FCG_INVALID_ARGUMENT, "link source out of range");
    if (d->group_leaf_start[0] != 0 || d->group_leaf_start[d->n_groups] != N)
        return fail(ctx, FCG_INVALID_ARGUMENT, "group leaf ranges do not span the leaves");
    for (int g = 0; g < d->n_groups; ++g)
        if (d->group_leaf_start[g + 1] <= d->group_leaf_start[g])
            return fail(ctx, FCG_INVALID_ARGUMENT, "empty group");
    return FCG_OK;
}

constexpr int64_t kGatherRowChunk = 4096;
// spare entries after a job's link index and coefficient arrays (room for
// vector loads of the link streams; nothing reads them at present)
constexpr int64_t kLinkSpare = 4;

// Charge-row layout of a job: nodes grouped by tree depth (root first), post-order
// within a depth, so the sources one row links to (mostly nodes of similar size
// near the row) sit in neighbouring rows; every row's links are then sorted by
// source row.  Returns perm (node -> row), iperm, the gather's row order, the
// sorted source rows and each stored link's offset within its row in the
// descriptor's order.
void build_tree_layout(const fcg_job_desc* d, std::vector<int32_t>& perm, std::vector<int32_t>& iperm,
                       std::vector<int32_t>& rorder) {
    const int M = d->n_nodes;
    std::vector<int32_t> depth((size_t)M, 0);
    int maxd = 0;
    for (int i = M - 1; i >= 0; --i) {  // parents follow their children
        const int p = d->parent[i];
        depth[(size_t)i] = p >= 0 ? depth[(size_t)p] + 1 : 0;
        maxd = std::max(maxd, depth[(size_t)i]);
    }
    std::vector<int64_t> start((size_t)maxd + 2, 0);
    for (int i = 0; i < M; ++i) start[(size_t)depth[(size_t)i] + 1]++;
    for (int t = 0; t <= maxd; ++t) start[(size_t)t + 1] += start[(size_t)t];
    perm.assign((size_t)M, 0);
    iperm.assign((size_t)M, 0);
    for (int i = 0; i < M; ++i) {
        const int64_t slot = start[(size_t)depth[(size_t)i]]++;
        perm[(size_t)i] = (int32_t)slot;
        iperm[(size_t)slot] = i;
    }
    {
        // gather row order: post-order chunks of kGatherRowChunk nodes (each about
        // a subtree, so the rows one XCD walks stay spatially close and their
        // sources stay in its L2), charge-row order inside a chunk (neighbouring
        // rows of a wavefront are same-depth neighbours with overlapping sources)
        const int64_t C = kGatherRowChunk;
        const int64_t nc = (M + C - 1) / C;
        std::vector<int64_t> cs((size_t)nc + 1, 0);
        for (int i = 0; i < M; ++i) cs[(size_t)(i / C) + 1]++;
        for (int64_t c = 0; c < nc; ++c) cs[(size_t)c + 1] += cs[(size_t)c];
        rorder.assign((size_t)M, 0);
        for (int sl = 0; sl < M; ++sl) {
            const int node = iperm[(size_t)sl];
            rorder[(size_t)cs[(size_t)(node / C)]++] = node;
        }
    }
}

// Tables of the product's tree passes (kernels.h kTreeChunk): post-order nodes in
// chunks of kTreeChunk; a node whose subtree (post-order range [start, n]) lies in
// its chunk is inner, the others are top nodes (their subtrees straddle a chunk
// boundary; a top node's parent is a top node).  Inner nodes are listed chunk by
// chunk grouped by depth, top nodes grouped by depth, both ascending.
int upload_tree_passes(fcg_ctx* ctx, Job& j, const fcg_job_desc* d, const std::vector<int32_t>& perm) {
    hipStream_t s = ctx->stream;
    const int M = d->n_nodes;
    constexpr int C = fcg::kTreeChunk;
    std::vector<int32_t> depth((size_t)M, 0), start((size_t)M, 0), roots;
    int maxd = 0;
    for (int i = M - 1; i >= 0; --i) {  // parents follow their children
        const int p = d->parent[i];
        depth[(size_t)i] = p >= 0 ? depth[(size_t)p] + 1 : 0;
        maxd = std::max(maxd, depth[(size_t)i]);
        if (p < 0) roots.push_back(i);
    }
    for (int i = 0; i < M; ++i) {
        int st = i;
        if (d->left[i] >= 0) st = std::min(st, start[(size_t)d->left[i]]);
        if (d->right[i] >= 0) st = std::min(st, start[(size_t)d->right[i]]);
        start[(size_t)i] = st;
    }
    std::vector<uint8_t> top((size_t)M, 0);
    for (int i = 0; i < M; ++i) top[(size_t)i] = start[(size_t)i] < (i / C) * C ? 1 : 0;
    const int nc = (int)fcg::tree_chunks(M);
    std::vector<int32_t> tb_node, tb_lev, tb_lvoff((size_t)nc + 1, 0), tt_node, tt_lev;
    std::vector<std::vector<int32_t>> byd((size_t)maxd + 1);
    for (int t = 0; t < nc; ++t) {
        for (auto& v : byd) v.clear();
        for (int i = t * C; i < std::min(M, (t + 1) * C); ++i)
            if (!top[(size_t)i]) byd[(size_t)depth[(size_t)i]].push_back(i);
        for (const auto& v : byd) {
            if (v.empty()) continue;
            tb_lev.push_back((int32_t)tb_node.size());
            tb_node.insert(tb_node.end(), v.begin(), v.end());
        }
        tb_lev.push_back((int32_t)tb_node.size());
        tb_lvoff[(size_t)t + 1] = (int32_t)tb_lev.size();
    }
    for (auto& v : byd) v.clear();
    for (int i = 0; i < M; ++i)
        if (top[(size_t)i]) byd[(size_t)depth[(size_t)i]].push_back(i);
    for (const auto& v : byd) {
        if (v.empty()) continue;
        tt_lev.push_back((int32_t)tt_node.size());
        tt_node.insert(tt_node.end(), v.begin(), v.end());
    }
    j.tt_nlev = (int)tt_lev.size();
    tt_lev.push_back((int32_t)tt_node.size());
    j.ntop = (int)tt_node.size();
    j.nroots = (int)roots.size();
    // per top node: its slot's parent and top children slots (-1: not a top node)
    std::vector<int32_t> slot((size_t)M, -1);
    for (size_t i = 0; i < tt_node.size(); ++i) slot[(size_t)tt_node[i]] = (int32_t)i;
    std::vector<int4> rec(std::max<size_t>(tt_node.size(), 1)), rows(rec.size(), make_int4(0, -1, -1, -1));
    for (size_t i = 0; i < tt_node.size(); ++i) {
        const int n = tt_node[i];
        const int p = d->parent[n], lc = d->left[n], rc = d->right[n];
        rec[i] = make_int4(n, p >= 0 ? slot[(size_t)p] : -1, lc >= 0 ? slot[(size_t)lc] : -1,
                           rc >= 0 ? slot[(size_t)rc] : -1);
        rows[i] = make_int4(perm[(size_t)n], lc >= 0 ? perm[(size_t)lc] : -1, rc >= 0 ? perm[(size_t)rc] : -1,
                            p >= 0 ? perm[(size_t)p] : -1);
    }
    // records of the inner nodes in list order (kernels.h BJob tb_rec / tb_ppm)
    std::vector<int4> tbr(std::max<size_t>(tb_node.size(), 1));
    std::vector<int32_t> ppm(std::max<size_t>(tb_node.size(), 1), -1);
    for (size_t i = 0; i < tb_node.size(); ++i) {
        const int n = tb_node[i], c0 = (n / C) * C;
        const int lc = d->left[n], rc = d->right[n], p = d->parent[n];
        const bool lds_parent = p >= c0 && p < c0 + C && !top[(size_t)p];
        auto pack = [](int lo, int hi) { return (int32_t)((uint32_t)(lo & 0xffff) | ((uint32_t)hi << 16)); };
        tbr[i] = make_int4(pack(n - c0, lc >= 0 ? lc - c0 : -1), pack(rc >= 0 ? rc - c0 : -1, lds_parent ? p - c0 : -1),
                           perm[(size_t)n], lc < 0 ? d->leaf_lo[n] : -1);
        if (p >= 0 && !lds_parent) ppm[i] = perm[(size_t)p];
    }
    // the path of the last leaf's node to its root as charge rows, root first
    // (2D fix-up values); its top nodes come first (a top node's parent is one)
    std::vector<int32_t> lp;
    for (int n = d->leaf_node[d->n_leaves - 1]; n >= 0; n = d->parent[n]) lp.push_back(n);
    std::reverse(lp.begin(), lp.end());
    j.lpath_len = (int)lp.size();
    j.lpath_top = 0;
    while (j.lpath_top < j.lpath_len && top[(size_t)lp[(size_t)j.lpath_top]]) ++j.lpath_top;
    for (int32_t& n : lp) n = perm[(size_t)n];
    if (tt_node.empty()) tt_node.push_back(0);
    HIPCHECK(ctx, upload(j.tb_rec, tbr.data(), tbr.size(), s));
    HIPCHECK(ctx, upload(j.tb_ppm, ppm.data(), ppm.size(), s));
    HIPCHECK(ctx, upload(j.lpath, lp.data(), lp.size(), s));
    HIPCHECK(ctx, upload(j.tb_lev, tb_lev.data(), tb_lev.size(), s));
    HIPCHECK(ctx, upload(j.tb_lvoff, tb_lvoff.data(), tb_lvoff.size(), s));
    HIPCHECK(ctx, upload(j.tt_node, tt_node.data(), tt_node.size(), s));
    HIPCHECK(ctx, upload(j.tt_rec, rec.data(), rec.size(), s));
    HIPCHECK(ctx, upload(j.tt_rows, rows.data(), rows.size(), s));
    HIPCHECK(ctx, upload(j.tt_lev, tt_lev.data(), tt_lev.size(), s));
    HIPCHECK(ctx, upload(j.ttop, top.data(), top.size(), s));
    HIPCHECK(ctx, upload(j.troots, roots.data(), roots.size(), s));
    HIPCHECK(ctx, hipStreamSynchronize(s));  // the host vectors are released here
    return FCG_OK;
}

void build_layout(const fcg_job_desc* d, std::vector<int32_t>& perm, std::vector<int32_t>& iperm,
                  std::vector<int32_t>& rorder, std::vector<int32_t>& col_sorted, std::vector<uint32_t>& link_pos) {
    const int M = d->n_nodes;
    build_tree_layout(d, perm, iperm, rorder);
    const int64_t L = d->n_links;
    col_sorted.assign((size_t)L, 0);
    link_pos.assign((size_t)L, 0);
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t) {
        th.emplace_back([&, t] {
            // rows split so that every thread gets about L/nth links
            const int64_t lo_l = L * t / nth, hi_l = L * (t + 1) / nth;
            int r = (int)(std::upper_bound(d->row_ptr, d->row_ptr + M + 1, lo_l) - d->row_ptr) - 1;
            if (lo_l == 0) r = 0;
            std::vector<std::pair<int32_t, uint32_t>> tmp;
            for (; r < M && d->row_ptr[r] < hi_l; ++r) {
                const int64_t b = d->row_ptr[r], e = d->row_ptr[r + 1];
                if (b < lo_l) continue;  // owned by the previous thread
                tmp.clear();
                for (int64_t l = b; l < e; ++l) tmp.emplace_back(perm[(size_t)d->col[l]], (uint32_t)(l - b));
                std::sort(tmp.begin(), tmp.end());
                for (int64_t l = b; l < e; ++l) {
                    col_sorted[(size_t)l] = tmp[(size_t)(l - b)].first;
                    link_pos[(size_t)l] = tmp[(size_t)(l - b)].second;
                }
            }
        });
    }
    for (auto& x : th) x.join();
}

// Fill class of a link (kernels.h): the branch of PotEstimateOpt
// (Solver/Autorefine.cpp:3474-3733, 2D :4156-4325) from the centroid distance over
// the larger max side, computed in FP64 in the reference's order of operations,
// then whether the row is a dielectric panel, then the dimension.  The fill
// kernels evaluate exactly this branch.
using fcg::kFillClasses;
int fill_class(const fcg_node_geom& a, const fcg_node_geom& b, int dim) {
    const double dx = (double)a.c[0] - (double)b.c[0], dy = (double)a.c[1] - (double)b.c[1];
    const double rmax = std::max(a.maxside, b.maxside);
    const int diel = (a.flags & FCG_NODE_DIEL) ? 1 : 0;
    if (dim == 2) return 10 + 2 * diel + (std::sqrt(dx * dx + dy * dy) / rmax < 4.0 ? 0 : 1);
    const double dz = (double)a.c[2] - (double)b.c[2];
    const double ratio = std::sqrt(dx * dx + dy * dy + dz * dz) / rmax;
    const int br = ratio < 1.0 ? 0 : ratio < 2.0 ? 1 : ratio < 4.0 ? 2 : ratio < 10.0 ? 3 : 4;
    return 5 * diel + br;
}

// Fill order of a job's stored links: grouped by fill_class (stable, so links of
// one class keep their row order); cnt_out receives the links per class.
void build_fill_order(const fcg_job_desc* d, const std::vector<int32_t>& iperm,
                      const std::vector<int32_t>& col_sorted, std::vector<uint32_t>& order,
                      std::vector<int64_t>& cnt_out) {
    const int M = d->n_nodes;
    const int64_t L = d->n_links;
    order.assign((size_t)L, 0);
    cnt_out.assign(kFillClasses, 0);
    if (L == 0) return;
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<uint8_t> cls((size_t)L);
    std::vector<int64_t> cnt((size_t)nth * kFillClasses, 0);
    auto row_range = [&](int t, int& r0, int& r1) {  // rows split by link count
        r0 = (int)(std::lower_bound(d->row_ptr, d->row_ptr + M + 1, L * t / nth) - d->row_ptr);
        r1 = (int)(std::lower_bound(d->row_ptr, d->row_ptr + M + 1, L * (t + 1) / nth) - d->row_ptr);
        if (t == nth - 1) r1 = M;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t) {
        th.emplace_back([&, t] {
            int r0, r1;
            row_range(t, r0, r1);
            for (int r = r0; r < r1; ++r)
                for (int64_t l = d->row_ptr[r]; l < d->row_ptr[r + 1]; ++l) {
                    const int c = fill_class(d->nodes[r], d->nodes[iperm[(size_t)col_sorted[(size_t)l]]], d->dim);
                    cls[(size_t)l] = (uint8_t)c;
                    cnt[(size_t)t * kFillClasses + c]++;
                }
        });
    }
    for (auto& x : th) x.join();
    th.clear();
    // class-major, thread-minor offsets keep the partition stable
    std::vector<int64_t> off((size_t)nth * kFillClasses, 0);
    int64_t run = 0;
    for (int c = 0; c < kFillClasses; ++c)
        for (int t = 0; t < nth; ++t) {
            off[(size_t)t * kFillClasses + c] = run;
            run += cnt[(size_t)t * kFillClasses + c];
            cnt_out[(size_t)c] += cnt[(size_t)t * kFillClasses + c];
        }
    for (int t = 0; t < nth; ++t) {
        th.emplace_back([&, t] {
            int r0, r1;
            row_range(t, r0, r1);
            int64_t* o = &off[(size_t)t * kFillClasses];
            for (int64_t l = d->row_ptr[r0]; l < d->row_ptr[r1]; ++l) order[(size_t)o[cls[(size_t)l]]++] = (uint32_t)l;
        });
    }
    for (auto& x : th) x.join();
}

// Rows and sources of a job's links in fill order, built on its first fill (one
// gather over the links), so that the fill kernels read them coalesced.
int ensure_fill_rs(fcg_ctx* ctx, Job& j) {
    if (j.fill_rs.p || j.L == 0) return FCG_OK;
    HIPCHECK(ctx, j.fill_rs.alloc(sizeof(int2) * (size_t)j.L));
    HIPCHECK(ctx, fcg::launch_fill_rs(j.fill_order.as<uint32_t>(), j.link_row.as<int32_t>(), j.col.as<int32_t>(), j.L,
                                      j.fill_rs.as<int2>(), ctx->stream));
    return FCG_OK;
}

// FP64 near-field side list of a job, built on its first mixed fill: the links of
// the analytic PotEstimateOpt branch (3D fill classes 0 and 5, centroid distance
// below one max side; 2D classes 10 and 12, below four), whose log and atan terms
// cancel, row by row in stored link order.  Their fill_order ranges are flagged,
// counted per row, scanned on the host and scattered.
int ensure_near(fcg_ctx* ctx, Job& j) {
    if (j.near_ready) return FCG_OK;
    hipStream_t s = ctx->stream;
    const int M = j.M;
    int64_t fstart[kFillClasses + 1];
    fstart[0] = 0;
    for (int c = 0; c < kFillClasses; ++c) fstart[c + 1] = fstart[c] + j.fill_cnt[(size_t)c];
    DevBuf flag, cnt;
    HIPCHECK(ctx, flag.alloc((size_t)std::max<int64_t>(j.L, 1)));
    HIPCHECK(ctx, hipMemsetAsync(flag.p, 0, (size_t)std::max<int64_t>(j.L, 1), s));
    int64_t want = 0;
    for (int c = 0; c < kFillClasses; ++c) {
        if (!fcg::near_class(c) || j.fill_cnt[(size_t)c] == 0) continue;
        HIPCHECK(ctx, fcg::launch_near_flag(j.fill_order.as<uint32_t>(), fstart[c], j.fill_cnt[(size_t)c],
                                            flag.as<uint8_t>(), s));
        want += j.fill_cnt[(size_t)c];
    }
    HIPCHECK(ctx, cnt.alloc(sizeof(int32_t) * (size_t)M));
    HIPCHECK(ctx, fcg::launch_near_count(j.row_ptr.as<int64_t>(), flag.as<uint8_t>(), M, cnt.as<int32_t>(), s));
    std::vector<int32_t> h((size_t)M);
    HIPCHECK(ctx, hipMemcpyAsync(h.data(), cnt.p, sizeof(int32_t) * (size_t)M, hipMemcpyDeviceToHost, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    std::vector<int64_t> ptr((size_t)M + 1, 0);
    for (int i = 0; i < M; ++i) ptr[(size_t)i + 1] = ptr[(size_t)i] + h[(size_t)i];
    const int64_t Ln = ptr[(size_t)M];
    if (Ln != want) return fail(ctx, FCG_DEVICE_ERROR, "near-field side list: counts do not add up");
    HIPCHECK(ctx, upload(j.near_ptr, ptr.data(), ptr.size(), s));
    HIPCHECK(ctx, j.near_col.alloc(sizeof(int32_t) * (size_t)Ln));
    HIPCHECK(ctx, j.near_link.alloc(sizeof(uint32_t) * (size_t)Ln));
    HIPCHECK(ctx, j.near_dc.alloc(sizeof(double) * (size_t)Ln));
    if (Ln > 0)
        HIPCHECK(ctx, fcg::launch_near_scatter(j.row_ptr.as<int64_t>(), flag.as<uint8_t>(), j.col.as<int32_t>(), M,
                                               j.near_ptr.as<int64_t>(), j.near_col.as<int32_t>(),
                                               j.near_link.as<uint32_t>(), s));
    HIPCHECK(ctx, hipStreamSynchronize(s));  // flag, cnt and the host vectors are released here
    j.Ln = Ln;
    j.near_ready = true;
    return FCG_OK;
}

// Source lists of a job's gather tiles for the staged gather (once per job,
// two launches and a host scan of the tile counts): only for jobs with 16-bit
// source indices, once the gather rows are known.
int ensure_tile_sources(fcg_ctx* ctx, Job& j) {
    if (j.max_tile_src >= 0 || !j.col16.p || !j.grow.p || j.M < 1 || j.M > 65536) return FCG_OK;
    hipStream_t s = ctx->stream;
    const int64_t nt = fcg::gather_tiles(j.M);
    DevBuf cnt;
    HIPCHECK(ctx, cnt.alloc(sizeof(int32_t) * (size_t)nt));
    HIPCHECK(ctx, fcg::launch_tile_sources_count(j.grow.as<fcg::GRow>(), j.col16.as<uint16_t>(), j.M,
                                                 cnt.as<int32_t>(), s));
    std::vector<int32_t> h((size_t)nt), off((size_t)nt + 1, 0);
    HIPCHECK(ctx, hipMemcpyAsync(h.data(), cnt.p, sizeof(int32_t) * (size_t)nt, hipMemcpyDeviceToHost, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    int mx = 0;
    for (int64_t t = 0; t < nt; ++t) {
        off[(size_t)t + 1] = off[(size_t)t] + h[(size_t)t];
        mx = std::max(mx, h[(size_t)t]);
    }
    HIPCHECK(ctx, upload(j.tsrc_off, off.data(), off.size(), s));
    HIPCHECK(ctx, j.tsrc.alloc(sizeof(uint16_t) * (size_t)std::max<int32_t>(off[(size_t)nt], 1)));
    HIPCHECK(ctx, j.lsrc.alloc(sizeof(uint16_t) * (size_t)(j.L + kLinkSpare)));
    HIPCHECK(ctx, hipMemsetAsync(j.lsrc.p, 0, sizeof(uint16_t) * (size_t)(j.L + kLinkSpare), s));
    HIPCHECK(ctx, fcg::launch_tile_sources_fill(j.grow.as<fcg::GRow>(), j.col16.as<uint16_t>(), j.M,
                                                j.tsrc_off.as<int32_t>(), j.tsrc.as<uint16_t>(),
                                                j.lsrc.as<uint16_t>(), s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    j.max_tile_src = mx;
    return FCG_OK;
}

// Block preconditioner partition (RecurseComputePrecond,
// Solver/SolveCapacitance.cpp:1094-1104, 1876-1993): every conductor's tree (roots
// in leaf order) is walked depth first, left first; the first subtree met with at
// most bsize leaves is one block of consecutive leaves.  A leaf met before any
// block is open (its parent has more than bsize leaves) is a block of its own;
// the reference opens a block there that it never computes.
std::vector<int32_t> block_partition(const Job& j, int bsize) {
    std::vector<int32_t> roots;
    for (int i = 0; i < j.M; ++i)
        if (j.h_parent[(size_t)i] < 0) roots.push_back(i);
    std::sort(roots.begin(), roots.end(),
              [&](int a, int b) { return j.h_leaf_lo[(size_t)a] < j.h_leaf_lo[(size_t)b]; });
    std::vector<int32_t> starts, stack;
    for (int r : roots) {
        stack.push_back(r);
        while (!stack.empty()) {
            const int a = stack.back();
            stack.pop_back();
            if (j.h_leaf_hi[(size_t)a] - j.h_leaf_lo[(size_t)a] <= bsize || j.h_left[(size_t)a] < 0) {
                starts.push_back(j.h_leaf_lo[(size_t)a]);
                continue;
            }
            stack.push_back(j.h_right[(size_t)a]);
            stack.push_back(j.h_left[(size_t)a]);
        }
    }
    starts.push_back(j.N);
    return starts;
}

int clamp_block_size(int b) { return b <= 0 ? 32 : std::min(std::max(b, 2), fcg::kMaxBlockPrecond); }

// the job's block preconditioner for block size bsize (built on first use; needs
// the self terms of a fill)
int ensure_block_precond(fcg_ctx* ctx, Job& j, int bsize) {
    bsize = clamp_block_size(bsize);
    if (j.bpre.size == bsize) return FCG_OK;
    if (!j.filled64 && !j.filled32) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    hipStream_t s = ctx->stream;
    const std::vector<int32_t> starts = block_partition(j, bsize);
    const int nb = (int)starts.size() - 1;
    std::vector<int64_t> off((size_t)nb + 1, 0);
    int maxn = 0;
    for (int b = 0; b < nb; ++b) {
        const int64_t n = starts[(size_t)b + 1] - starts[(size_t)b];
        off[(size_t)b + 1] = off[(size_t)b] + n * n;
        maxn = std::max(maxn, (int)n);
    }
    HIPCHECK(ctx, hipStreamSynchronize(s));  // the buffers below may be in use; the table arena is reset
    j.bpre.size = 0;
    HIPCHECK(ctx, upload(j.bpre.bstart, starts.data(), starts.size(), s));
    HIPCHECK(ctx, upload(j.bpre.boff, off.data(), off.size(), s));
    HIPCHECK(ctx, j.bpre.P.alloc(sizeof(double) * (size_t)std::max<int64_t>(off.back(), 1)));
    DevBuf work;
    if (maxn > 64) HIPCHECK(ctx, work.alloc(sizeof(double) * (size_t)off.back()));
    ctx->tabs.reset();
    BatchView bv;
    RCCHECK(make_batch(ctx, ctx->tabs, {BatchEntry{&j, &j.scr, nullptr, nullptr}}, &bv));
    HIPCHECK(ctx, fcg::launch_block_setup(bv, nb, j.bpre.bstart.as<int32_t>(), j.bpre.boff.as<int64_t>(),
                                          j.bpre.P.as<double>(), work.as<double>(), s));
    HIPCHECK(ctx, hipStreamSynchronize(s));  // work and the host vectors are released below
    j.bpre.size = bsize;
    j.bpre.nb = nb;
    return FCG_OK;
}

// Apply table of the block preconditioner over several solves.
struct PEnt {
    const Job* j;
    const void* in;
    bool in32;
    double* out;
    int k;
};
int make_pview(fcg_ctx* ctx, Tables& tabs, const std::vector<PEnt>& ents, fcg::PView* pv) {
    std::memset(pv, 0, sizeof *pv);
    if (ents.empty()) return FCG_OK;
    std::vector<fcg::PJob> pj;
    std::vector<int64_t> off(ents.size() + 1, 0);
    for (size_t i = 0; i < ents.size(); ++i) {
        const PEnt& e = ents[i];
        fcg::PJob r{};
        r.in = e.in;
        r.out = e.out;
        r.brange = e.j->rhs_range.as<int32_t>();
        r.bstart = e.j->bpre.bstart.as<int32_t>();
        r.boff = e.j->bpre.boff.as<int64_t>();
        r.P = e.j->bpre.P.as<double>();
        r.N = e.j->N;
        r.k = e.k;
        r.in32 = e.in32 ? 1 : 0;
        pj.push_back(r);
        off[i + 1] = off[i] + e.j->bpre.nb;
    }
    void *dj, *doff;
    HIPCHECK(ctx, tabs.push(pj.data(), sizeof(fcg::PJob) * pj.size(), ctx->stream, &dj));
    HIPCHECK(ctx, tabs.push(off.data(), sizeof(int64_t) * off.size(), ctx->stream, &doff));
    pv->jobs = reinterpret_cast<const fcg::PJob*>(dj);
    pv->njobs = (int)ents.size();
    pv->blk_off = reinterpret_cast<const int64_t*>(doff);
    pv->nblk = off.back();
    return FCG_OK;
}

// Two-level preconditioner elements (MakePreconditioner, Solver/SolveCapacitance.cpp:
// 1067-1104, and the super branch of RecurseComputePrecond, :1876-1993): with R
// tree roots and dimension dim, every root's tree is walked to level
// L = log2(dim / R) (integer division; root level 0); a non-leaf at level L is
// one element standing for its leaves, a leaf at level <= L is one.  Elements
// follow leaf order.  Unavailable (false) when R >= dim.
bool super_partition(const Job& j, int dim, std::vector<int32_t>& nodes, std::vector<int32_t>& starts) {
    std::vector<int32_t> roots;
    for (int i = 0; i < j.M; ++i)
        if (j.h_parent[(size_t)i] < 0) roots.push_back(i);
    std::sort(roots.begin(), roots.end(),
              [&](int a, int b) { return j.h_leaf_lo[(size_t)a] < j.h_leaf_lo[(size_t)b]; });
    nodes.clear();
    starts.clear();
    const unsigned R = (unsigned)roots.size();
    if (R >= (unsigned)dim) return false;
    const int L = (int)(unsigned)(std::log((double)((unsigned)dim / R)) / std::log(2.0));
    std::vector<std::pair<int, int>> stack;
    for (int r : roots) {
        stack.push_back({r, 0});
        while (!stack.empty()) {
            const auto [a, lev] = stack.back();
            stack.pop_back();
            if (j.h_left[(size_t)a] < 0 || lev == L) {
                nodes.push_back(a);
                starts.push_back(j.h_leaf_lo[(size_t)a]);
                continue;
            }
            stack.push_back({j.h_right[(size_t)a], lev + 1});
            stack.push_back({j.h_left[(size_t)a], lev + 1});
        }
    }
    starts.push_back(j.N);
    return true;
}

int clamp_super_dim(int d) { return d <= 0 ? 128 : std::min(std::max(d, 16), 4096); }

// Structure of the job's two-level preconditioner of dimension dim: the element
// partition, its device arrays and the leaf weights.  These depend on the tree
// alone and are kept until another dimension is asked for.
int super_structure(fcg_ctx* ctx, Job& j, int dim) {
    dim = clamp_super_dim(dim);
    if (j.spre.part_dim == dim) return FCG_OK;
    hipStream_t s = ctx->stream;
    std::vector<int32_t> nodes, starts;
    const bool avail = super_partition(j, dim, nodes, starts);
    if (avail) {
        // the elements are ascending, non-empty leaf ranges covering [0, N), as
        // the collection (sstart) and the distribution (sidx) read them
        bool ok = starts.size() == nodes.size() + 1 && starts.front() == 0 && starts.back() == j.N;
        for (size_t e = 0; ok && e + 1 < starts.size(); ++e) ok = starts[e] < starts[e + 1];
        if (!ok) return fail(ctx, FCG_GENERIC_ERROR, "two-level preconditioner: bad element partition");
    }
    HIPCHECK(ctx, hipStreamSynchronize(s));  // the buffers below may be in use; the table arena is reset
    j.spre.dim = 0;
    j.spre.part_dim = 0;
    j.spre.avail = avail;
    j.spre.m = (int)nodes.size();
    if (avail) {
        const int m = j.spre.m;
        std::vector<int32_t> sidx((size_t)j.N);
        for (int e = 0; e < m; ++e)
            for (int i = starts[(size_t)e]; i < starts[(size_t)e + 1]; ++i) sidx[(size_t)i] = e;
        HIPCHECK(ctx, upload(j.spre.snode, nodes.data(), nodes.size(), s));
        HIPCHECK(ctx, upload(j.spre.sstart, starts.data(), starts.size(), s));
        HIPCHECK(ctx, upload(j.spre.sidx, sidx.data(), sidx.size(), s));
        HIPCHECK(ctx, j.spre.Sinv.alloc(sizeof(double) * (size_t)m * m));
        HIPCHECK(ctx, j.spre.w.alloc(sizeof(double) * (size_t)j.N));
        ctx->tabs.reset();
        BatchView bv;
        RCCHECK(make_batch(ctx, ctx->tabs, {BatchEntry{&j, &j.scr, nullptr, nullptr}}, &bv));
        HIPCHECK(ctx, fcg::launch_super_weights(bv, j.N, j.spre.snode.as<int32_t>(), j.spre.sidx.as<int32_t>(),
                                                j.spre.w.as<double>(), s));
        HIPCHECK(ctx, hipStreamSynchronize(s));  // the host vectors are released below
    }
    j.spre.part_dim = dim;
    return FCG_OK;
}

// The job's two-level preconditioner of dimension dim, its inverse included
// (built on first use after each fill; needs a fill).  At most kSuperLds
// elements: one workgroup in LDS; more: Gauss-Jordan over the whole device.
int ensure_super_precond(fcg_ctx* ctx, Job& j, int dim) {
    dim = clamp_super_dim(dim);
    if (!j.filled64 && !j.filled32) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    RCCHECK(super_structure(ctx, j, dim));
    if (j.spre.dim == dim) return FCG_OK;
    if (j.spre.avail) {
        hipStream_t s = ctx->stream;
        const int m = j.spre.m;
        HIPCHECK(ctx, hipStreamSynchronize(s));  // the table arena is reset
        ctx->tabs.reset();
        BatchView bv;
        RCCHECK(make_batch(ctx, ctx->tabs, {BatchEntry{&j, &j.scr, nullptr, nullptr}}, &bv));
        if (m <= fcg::kSuperLds) {
            const fcg::SSetup st{j.spre.snode.as<int32_t>(), j.spre.Sinv.as<double>(), m};
            void* d;
            HIPCHECK(ctx, ctx->tabs.push(&st, sizeof st, s, &d));
            HIPCHECK(ctx, fcg::launch_super_invert_lds(bv, reinterpret_cast<const fcg::SSetup*>(d), 1, s));
        } else {
            DevBuf A;
            HIPCHECK(ctx, A.alloc(sizeof(double) * (size_t)m * m));
            HIPCHECK(ctx, fcg::launch_super_invert_global(bv, m, j.spre.snode.as<int32_t>(), A.as<double>(),
                                                          j.spre.Sinv.as<double>(), s));
            HIPCHECK(ctx, hipStreamSynchronize(s));  // A is released below
        }
    }
    j.spre.dim = dim;
    return FCG_OK;
}

// Queue the two-level preconditioner inverses of jobs (all with a structure of
// dimension dim and at most kSuperLds elements) as one launch, on the tables
// of tabs; q receives the launch.
int queue_super_inverses(fcg_ctx* ctx, Tables& tabs, const std::vector<Job*>& jobs,
                         std::vector<std::function<hipError_t()>>& q) {
    if (jobs.empty()) return FCG_OK;
    std::vector<BatchEntry> ents;
    for (Job* j : jobs) ents.push_back(BatchEntry{j, &j->scr, nullptr, nullptr});
    BatchView bv;
    std::vector<int> order;
    RCCHECK(make_batch(ctx, tabs, ents, &bv, nullptr, &order));
    std::vector<fcg::SSetup> st;
    for (int i : order) {
        const Job* j = jobs[(size_t)i];
        st.push_back(fcg::SSetup{j->spre.snode.as<int32_t>(), j->spre.Sinv.as<double>(), j->spre.m});
    }
    void* d;
    HIPCHECK(ctx, tabs.push(st.data(), sizeof(fcg::SSetup) * st.size(), ctx->stream, &d));
    const auto* tab = reinterpret_cast<const fcg::SSetup*>(d);
    const int n = (int)st.size();
    hipStream_t s = ctx->stream;
    q.push_back([=] { return fcg::launch_super_invert_lds(bv, tab, n, s); });
    return FCG_OK;
}

// Apply table of the two-level preconditioner over several solves.
struct SEnt {
    const Job* j;
    const void* in;
    bool in32;
    double* out;
    int k;
    double* scratch;  // [2][k][m]
};
int make_sview(fcg_ctx* ctx, Tables& tabs, const std::vector<SEnt>& ents, fcg::SView* sv) {
    std::memset(sv, 0, sizeof *sv);
    if (ents.empty()) return FCG_OK;
    std::vector<fcg::SJob> sj;
    const size_t n1 = ents.size() + 1;
    std::vector<int64_t> off(4 * n1, 0);
    int64_t* eo = off.data();
    int64_t* lo = eo + n1;
    int64_t* no = lo + n1;  // collection lanes, whole wavefronts per solve
    int64_t* mo = no + n1;  // elements, whole workgroups of 64 per solve
    for (size_t i = 0; i < ents.size(); ++i) {
        const SEnt& e = ents[i];
        const Job::SuperPre& sp = e.j->spre;
        if ((int64_t)e.k * e.j->N >= ((int64_t)1 << 31))
            return fail(ctx, FCG_INVALID_ARGUMENT, "two-level preconditioner: k x N exceeds 2^31");
        fcg::SJob r{};
        r.in = e.in;
        r.out = e.out;
        r.brange = e.j->rhs_range.as<int32_t>();
        r.sstart = sp.sstart.as<int32_t>();
        r.sidx = sp.sidx.as<int32_t>();
        r.w = sp.w.as<double>();
        r.Sinv = sp.Sinv.as<double>();
        r.pot = e.scratch;
        r.chg = e.scratch + (size_t)e.k * sp.m;
        r.N = e.j->N;
        r.k = e.k;
        r.m = sp.m;
        r.in32 = e.in32 ? 1 : 0;
        // lanes per element: the power of two at or above the mean leaves per
        // element, at most a wavefront
        const int per = (e.j->N + sp.m - 1) / std::max(sp.m, 1);
        r.lpe = 1;
        while (r.lpe < per && r.lpe < 64) r.lpe <<= 1;
        // k_super_collect indexes a solve's lanes in 32 bits
        if (((int64_t)e.k * sp.m * r.lpe + 63) / 64 * 64 >= ((int64_t)1 << 31))
            return fail(ctx, FCG_INVALID_ARGUMENT, "two-level preconditioner: collection lanes exceed 2^31");
        sj.push_back(r);
        eo[i + 1] = eo[i] + (int64_t)e.k * sp.m;
        lo[i + 1] = lo[i] + (int64_t)e.k * e.j->N;
        no[i + 1] = no[i] + ((int64_t)e.k * sp.m * r.lpe + 63) / 64 * 64;
        mo[i + 1] = mo[i] + ((int64_t)sp.m + 63) / 64 * 64;
    }
    void *dj, *doff;
    HIPCHECK(ctx, tabs.push(sj.data(), sizeof(fcg::SJob) * sj.size(), ctx->stream, &dj));
    HIPCHECK(ctx, tabs.push(off.data(), sizeof(int64_t) * off.size(), ctx->stream, &doff));
    sv->jobs = reinterpret_cast<const fcg::SJob*>(dj);
    sv->njobs = (int)ents.size();
    const int64_t* d = reinterpret_cast<const int64_t*>(doff);
    sv->el_off = d;
    sv->lf_off = d + n1;
    sv->ln_off = d + 2 * n1;
    sv->em_off = d + 3 * n1;
    sv->nel = eo[ents.size()];
    sv->nlf = lo[ents.size()];
    sv->nln = no[ents.size()];
    sv->nem = mo[ents.size()];
    return FCG_OK;
}

// One stand-alone product on device buffers of job j (batch of one) on k
// right-hand sides, each a [vlen()] vector (complex jobs: two product vectors).
int run_mvp(fcg_ctx* ctx, Job& j, int prec, const double* dq, double* dv, int k, int jacobi, hipStream_t s,
            hipEvent_t* ev = nullptr) {
    if (prec == fcg::kMvp64 && !j.filled64) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled in FP64");
    if (prec != fcg::kMvp64 && !j.filled32) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled in mixed mode");
    RCCHECK(ensure_scratch(ctx, j.scr, j.N, j.M, j.pk(k), prec == fcg::kMvp32));
    RCCHECK(ensure_tile_sources(ctx, j));
    HIPCHECK(ctx, hipStreamSynchronize(s));  // the table arena is reused below
    ctx->tabs.reset();
    BatchView bv;
    fcg::GatherPlan gp;
    const std::vector<BatchEntry> ents{BatchEntry{&j, &j.scr, dq, dv}};
    RCCHECK(make_batch(ctx, ctx->tabs, ents, &bv, &gp));
    choose_variants(ctx, prec, ents, gp);
    HIPCHECK(ctx, fcg::launch_mvp(bv, prec, jacobi, gp, s, ev));
    return FCG_OK;
}

}  // namespace

// ======================================================================== engine
namespace {

enum Phase { kInit = 0, kResid = 1, kArn = 2, kUpdate = 3, kFinish = 4, kDone = 5 };

struct Solve {
    int job_id = -1;
    Job* J = nullptr;
    int64_t tag = 0;
    int k = 0, N = 0, ncol = 0;
    bool mixed = false, jac = true, want_x = false;
    bool block = false;   // block preconditioner: P applied by its own kernel, u = P q scratch
    bool twolev = false;  // two-level preconditioner, likewise; spw = its [2][k][m] scratch
    PBuf u, spw;
    PBuf x, r, w;
    Scratch scr;
    // Krylov basis: blocks of vectors and their addresses (sent with each step's tables)
    std::vector<std::unique_ptr<PBuf>> qblocks;
    std::vector<void*> hq;
    bool b32 = false;  // FP32 basis vectors (inner solves of the mixed path)
    int phase = kInit;
    int i = 0;  // Arnoldi step index within the cycle (the next product uses Q[i])
    int cycles = 0;
    bool converged = false;     // every rhs met the tolerance (see fcg_solve_stats)
    bool need_scale = false;    // Q[0] = r / d pending (cycle start)
    std::vector<double> d;      // scale divisors per rhs
    std::vector<double> normb, beta, thr;
    // Arnoldi state (Hessenberg columns, rotations, g) on the device: one DCGS2
    // record per rhs (kernels.h dcgs_rec) of hcap columns; gres: the last
    // residual estimate |g| of each rhs
    PBuf hd;
    int hcap = 0;
    std::vector<double> gres;
    std::vector<int> done, last_i, total_steps;
    int nz = 0;  // columns of the pending solution update
    int cap_iter = 0;
    int64_t n_mvp = 0;
    double rel = 0.0;
    // residual history per rhs (fcg_solve_opts::residuals): |r0|, then |g_i| / |b|
    bool want_hist = false;
    std::vector<std::vector<double>> hist;
    clk::time_point t0;
    int hoff = 0, ooff = 0;  // the step's reduction and host-output offsets in the step buffer
    int64_t bytes() const {
        int64_t s = (int64_t)(x.r + r.r + w.r + u.r + spw.r + hd.r + scr.bytes());
        for (auto& q : qblocks) s += (int64_t)q->r;
        return s;
    }
};

struct Ready {
    int64_t tag;
    int job;
    std::vector<double> cap, x;
    std::vector<std::vector<double>> hist;
    int rows, cols;
    fcg_solve_stats stats;
};

}  // namespace

namespace {
// One part of an engine's solves.  The engine alternates between its cohorts, so
// the host's bookkeeping and planning of one cohort overlap the device work of
// the other; each cohort's own steps stay strictly ordered (plan, launch, wait,
// bookkeeping), so results do not depend on the split.
struct Cohort {
    std::vector<std::unique_ptr<Solve>> active;
    // step tables: slabs of device memory and pinned staging, lent by the
    // context for the engine's life so that engines do not allocate them
    Tables* tabs = nullptr;
    PBuf partial_a, partial_b, hbuf;
    PBuf coefb, zbuf;  // DCGS2 update coefficients and solution-update coefficients of a step
    ProfEvents pe_res, pe_arn;
    hipEvent_t ev = nullptr;      // end of the step in flight
    hipEvent_t kev = nullptr;     // end of the step's kernels
    bool inflight = false;
    int64_t load = 0;  // links x rhs of the active solves (balances submissions)
    // the step in flight, as its bookkeeping needs it
    std::vector<Solve*> res, arn;
    std::vector<std::pair<Solve*, int64_t>> fin;
    std::vector<double*> fin_x;
    double* fin_host = nullptr;
    double* hh = nullptr;
    PBuf fin_dev;
    clk::time_point t_plan, t_wait;
    Cohort() = default;
    Cohort(const Cohort&) = delete;
    Cohort& operator=(const Cohort&) = delete;
    ~Cohort() {
        if (ev) (void)hipEventDestroy(ev);
        if (kev) (void)hipEventDestroy(kev);
    }
};
constexpr int kCohorts = 2;
}  // namespace

struct fcg_engine {
    fcg_ctx* ctx = nullptr;
    fcg_solve_opts opts{};
    Cohort co[kCohorts];
    std::deque<Ready> ready;
    bool want_x = false;
    bool want_hist = false;  // record residual histories (fcg_solve with opts->residuals)
    size_t n_active() const {
        size_t n = 0;
        for (const Cohort& c : co) n += c.active.size();
        return n;
    }
    // a step still in flight (an error path) must end before the cohorts' buffers
    // and staging are freed
    ~fcg_engine() {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(ctx->down_stream);
        for (Cohort& c : co)
            if (c.tabs) {
                c.tabs->reset();
                c.tabs->defer = false;
                ctx->spare_tables.emplace_back(c.tabs);
                c.tabs = nullptr;
            }
    }
};

namespace {

// step buffers of a cohort grow geometrically: a step's size creeps up with the
// Arnoldi index.  No wait: the cohort's previous step has completed before its
// next one is planned, so its buffers are idle (the other cohort's work in flight
// uses buffers of its own).
int ensure_pbuf(fcg_ctx* ctx, PBuf& b, size_t bytes) {
    if (b.p && b.r >= bytes) return FCG_OK;
    HIPCHECK(ctx, b.get(&ctx->pool, std::max(bytes, 2 * b.r)));
    return FCG_OK;
}

// make sure basis vector index `idx` exists (device memory from the pool only)
// Device Arnoldi records of a solve for `cols` Hessenberg columns, grown by
// doubling.  The record layout does not depend on the capacity (dcgs_rec), so a
// larger one starts with a copy of the smaller; the copy is ordered on the
// compute stream, which is synchronised before the old block returns to the
// pool (growth is rare: 48, 96, 192, ... columns).
int ensure_hstate(fcg_ctx* ctx, Solve& S, int cols, hipStream_t s) {
    if (cols <= S.hcap) return FCG_OK;
    const int cap = std::max(cols, std::max(2 * S.hcap, 48));
    const int64_t st_new = fcg::dcgs_rec(cap), st_old = fcg::dcgs_rec(S.hcap);
    PBuf nb;
    HIPCHECK(ctx, nb.get(&ctx->pool, sizeof(double) * (size_t)(st_new * S.k)));
    if (S.hcap > 0) {
        HIPCHECK(ctx, hipMemcpy2DAsync(nb.p, sizeof(double) * (size_t)st_new, S.hd.p, sizeof(double) * (size_t)st_old,
                                       sizeof(double) * (size_t)st_old, (size_t)S.k, hipMemcpyDeviceToDevice, s));
        HIPCHECK(ctx, hipStreamSynchronize(s));
    }
    std::swap(S.hd.p, nb.p);
    std::swap(S.hd.r, nb.r);
    std::swap(S.hd.pool, nb.pool);
    S.hcap = cap;
    return FCG_OK;
}

int ensure_basis(fcg_ctx* ctx, Solve& S, int idx) {
    const size_t vec = (size_t)S.k * S.N;
    const size_t elem = S.b32 ? sizeof(float) : sizeof(double);
    while ((int)S.hq.size() <= idx) {
        // blocks of about 64 MiB, 4 to 64 vectors
        const size_t per = std::max<size_t>(4, std::min<size_t>(64, ((size_t)64 << 20) / (vec * 8 + 1)));
        S.qblocks.emplace_back(new PBuf());
        HIPCHECK(ctx, S.qblocks.back()->get(&ctx->pool, elem * vec * per));
        for (size_t c = 0; c < per; ++c) S.hq.push_back(S.qblocks.back()->as<char>() + c * vec * elem);
    }
    return FCG_OK;
}

KJob kjob_of(Solve& S, void* const* qtab, double* w, int nv, int zoff) {
    KJob J{};
    J.Q = qtab;
    J.w = w;
    J.x = S.x.as<double>();
    J.r = S.r.as<double>();
    J.brange = S.J->rhs_range.as<int32_t>();
    J.self = S.J->self.as<double>();
    J.N = S.N;
    J.nleaf = S.J->N;
    J.k = S.k;
    J.nv = nv;
    J.hoff = S.hoff;
    J.zoff = zoff;
    J.nz = S.nz;
    J.jacobi = S.jac ? 1 : 0;
    J.y = (S.block || S.twolev) ? S.u.as<double>() : nullptr;
    return J;
}

// Krylov tables for a list of solves
struct KEnt {
    Solve* S;
    double* w;
    int nv;
    int zoff;
    int coff = 0;      // DCGS2 update: coefficient offset
    bool next = false;  // DCGS2 update: write the next provisional vector
};
int make_krylov(fcg_ctx* ctx, Tables& tabs, const std::vector<KEnt>& ents, KrylovView* kv) {
    std::memset(kv, 0, sizeof *kv);
    std::vector<KJob> kj;
    std::vector<Seg> segs;
    std::vector<Chunk> chunks;
    int maxnv = 0;
    // basis address tables of every entry, back to back
    std::vector<void*> qaddr;
    std::vector<size_t> qoff;
    for (size_t e = 0; e < ents.size(); ++e) {
        Solve& S = *ents[e].S;
        // slots 0..nv, the next one for a DCGS2 update, 0..nz-1 for a solution update
        const size_t need = std::min(S.hq.size(), (size_t)std::max(ents[e].nv + (ents[e].next ? 2 : 1), S.nz));
        qoff.push_back(qaddr.size());
        qaddr.insert(qaddr.end(), S.hq.begin(), S.hq.begin() + need);
    }
    void* qd = nullptr;
    HIPCHECK(ctx, tabs.push(qaddr.data(), sizeof(void*) * qaddr.size(), ctx->stream, &qd));
    void* const* qbase = reinterpret_cast<void* const*>(qd);
    // per rhs: Arnoldi flags (met the tolerance), last columns and cycle-start
    // norms, for the DCGS2 kernels
    std::vector<int32_t> iv;
    std::vector<double> bv;
    std::vector<size_t> ioff;
    for (size_t e = 0; e < ents.size(); ++e) {
        const Solve& S = *ents[e].S;
        ioff.push_back(iv.size());
        for (int q = 0; q < S.k; ++q) iv.push_back(q < (int)S.done.size() ? S.done[(size_t)q] : 0);
        for (int q = 0; q < S.k; ++q) iv.push_back(q < (int)S.last_i.size() ? S.last_i[(size_t)q] : -1);
        for (int q = 0; q < S.k; ++q) bv.push_back(q < (int)S.beta.size() ? S.beta[(size_t)q] : 0.0);
    }
    void *id = nullptr, *bd = nullptr;
    HIPCHECK(ctx, tabs.push(iv.data(), sizeof(int32_t) * std::max<size_t>(iv.size(), 1), ctx->stream, &id));
    HIPCHECK(ctx, tabs.push(bv.data(), sizeof(double) * std::max<size_t>(bv.size(), 1), ctx->stream, &bd));
    for (size_t e = 0; e < ents.size(); ++e) {
        Solve& S = *ents[e].S;
        kj.push_back(kjob_of(S, qbase + qoff[e], ents[e].w, ents[e].nv, ents[e].zoff));
        KJob& J = kj.back();
        J.coff = ents[e].coff;
        J.next = ents[e].next ? 1 : 0;
        J.hd = S.hd.as<double>();
        J.hstride = fcg::dcgs_rec(S.hcap);
        J.ooff = S.ooff;
        J.flags = reinterpret_cast<const int32_t*>(id) + ioff[e];
        J.last = J.flags + S.k;
        J.beta = reinterpret_cast<const double*>(bd) + ioff[e] / 2;
        maxnv = std::max(maxnv, ents[e].nv);
        kv->maxnz = std::max(kv->maxnz, S.nz);
        for (int r = 0; r < S.k; ++r) {
            Seg sg;
            sg.job = (int)e;
            sg.r = r;
            sg.c_lo = (int)chunks.size();
            for (int n0 = 0; n0 < S.N; n0 += fcg::kChunk)
                chunks.push_back(Chunk{(int)segs.size(), n0, std::min(S.N, n0 + fcg::kChunk)});
            sg.c_hi = (int)chunks.size();
            segs.push_back(sg);
        }
    }
    void *a, *b2, *c;
    HIPCHECK(ctx, tabs.push(kj.data(), sizeof(KJob) * kj.size(), ctx->stream, &a));
    HIPCHECK(ctx, tabs.push(segs.data(), sizeof(Seg) * segs.size(), ctx->stream, &b2));
    HIPCHECK(ctx, tabs.push(chunks.data(), sizeof(Chunk) * chunks.size(), ctx->stream, &c));
    kv->kjobs = reinterpret_cast<const KJob*>(a);
    kv->segs = reinterpret_cast<const Seg*>(b2);
    kv->chunks = reinterpret_cast<const Chunk*>(c);
    kv->nsegs = (int)segs.size();
    kv->nchunks = (int)chunks.size();
    kv->maxnv = maxnv;
    kv->basis32 = ents.empty() ? 0 : (ents[0].S->b32 ? 1 : 0);  // one engine, one precision
    return FCG_OK;
}

// Relative accuracy an inner FP32 cycle is asked for at most: a few times the
// FP32 rounding of the stored coefficients, charges and basis vectors.
constexpr double kInnerFloor = 2e-7;

void start_cycle(fcg_engine* E, Solve& S, const double* beta) {
    const double tol = E->opts.tol;
    S.beta.assign(beta, beta + S.k);
    S.thr.assign((size_t)S.k, 0.0);
    S.done.assign((size_t)S.k, 0);
    S.last_i.assign((size_t)S.k, -1);
    S.gres.assign((size_t)S.k, 0.0);
    for (int q = 0; q < S.k; ++q) {
        if (!S.mixed) {
            // reference stop rule |g_{i+1}| / |b| < tol (SolveCapacitance.cpp:2493-2500)
            S.thr[(size_t)q] = tol * S.normb[(size_t)q];
        } else {
            // inner FP32 solve: the reference stop rule at half the tolerance, so
            // that the FP64 residual check after the cycle (which differs from the
            // inner estimate by the FP32 product error, about 1e-7 relative with
            // FP32 coefficients and charges summed in FP64) passes at once; never
            // below a small multiple of that error
            const double rel_in = std::max(0.5 * tol * S.normb[(size_t)q] / beta[q], kInnerFloor);
            S.thr[(size_t)q] = rel_in * beta[q];
        }
        if (beta[q] / S.normb[(size_t)q] < tol) S.done[(size_t)q] = 1;
    }
    if (S.want_hist && S.cycles == 0) {
        S.hist.assign((size_t)S.k, {});
        for (int q = 0; q < S.k; ++q) S.hist[(size_t)q].push_back(beta[q]);
    }
    S.i = 0;
    S.cycles++;
    S.need_scale = true;
    S.d.assign(beta, beta + S.k);
    S.phase = kArn;
    bool any = false;
    for (int d : S.done) any |= !d;
    if (!any) S.phase = kFinish;  // the current solution already meets the tolerance
}

// back substitution of every rhs; z[jj][r], zero beyond each rhs's own count
// The cycle ends: its solution update uses the least squares over each rhs's
// columns (k_dcgs_lsq on the device), nz = the longest.
void prepare_update(Solve& S) {
    int nz = 0;
    for (int q = 0; q < S.k; ++q) nz = std::max(nz, S.last_i[(size_t)q] + 1);
    S.nz = nz;
    S.phase = kUpdate;
}

// Givens update of one Arnoldi step of every rhs (h1 + h2 = column, |w| = subdiagonal)

int engine_submit(fcg_engine* E, int job, int64_t tag) {
    fcg_ctx* ctx = E->ctx;
    Job* J = get_job(ctx, job);
    if (!J) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    const bool mixed = E->opts.precision == FCG_PREC_MIXED;
    if (mixed ? !J->filled32 : !J->filled64) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled in this precision");
    if (J->rhs.empty()) return fail(ctx, FCG_INVALID_ARGUMENT, "job has no conductor right-hand side");
    auto S = std::make_unique<Solve>();
    S->job_id = job;
    S->J = J;
    S->tag = tag;
    S->k = (int)J->rhs.size();
    S->N = J->vlen();  // complex: [Re | Im] vectors of the real 2N system
    S->ncol = J->ncol;
    S->mixed = mixed;
    S->b32 = mixed;
    if (J->cplx && E->opts.precond != FCG_PRECOND_NONE && E->opts.precond != FCG_PRECOND_JACOBI)
        return fail(ctx, FCG_INVALID_ARGUMENT,
                    "complex permittivities: Jacobi or no preconditioner (the reference's block and two-level "
                    "preconditioners handle the real form's first half only)");
    S->jac = E->opts.precond == FCG_PRECOND_JACOBI;
    S->block = E->opts.precond == FCG_PRECOND_BLOCK;
    if (E->opts.precond == FCG_PRECOND_SUPER) {
        RCCHECK(super_structure(ctx, *J, E->opts.precond_size));
        // unavailable (at least dim roots): the reference drops it and solves
        // unpreconditioned
        S->twolev = J->spre.avail;
        // inverses of at most kSuperLds elements are built by the cohort's first
        // step, for all its starting solves in one launch; larger ones here
        if (S->twolev && J->spre.m > fcg::kSuperLds) RCCHECK(ensure_super_precond(ctx, *J, E->opts.precond_size));
    }
    S->want_x = E->want_x;
    S->want_hist = E->want_hist;
    S->cap_iter = std::min(E->opts.max_iter > 0 ? E->opts.max_iter : 1000, J->N);
    RCCHECK(ensure_scratch(ctx, S->scr, J->N, J->M, J->pk(S->k), mixed));
    RCCHECK(ensure_tile_sources(ctx, *J));
    const size_t vec = (size_t)S->k * S->N;
    HIPCHECK(ctx, S->x.get(&ctx->pool, sizeof(double) * vec));
    HIPCHECK(ctx, S->r.get(&ctx->pool, sizeof(double) * vec));
    HIPCHECK(ctx, S->w.get(&ctx->pool, sizeof(double) * vec));
    if (S->block) {
        RCCHECK(ensure_block_precond(ctx, *J, E->opts.precond_size));
        HIPCHECK(ctx, S->u.get(&ctx->pool, sizeof(double) * vec));
    }
    if (S->twolev) {
        HIPCHECK(ctx, S->u.get(&ctx->pool, sizeof(double) * vec));
        HIPCHECK(ctx, S->spw.get(&ctx->pool, sizeof(double) * 2 * (size_t)S->k * J->spre.m));
    }
    // unit potential on each conductor's panels (SolveCapacitance.cpp:1209-1211); the
    // kernels form b from the job's leaf ranges, so only its norm is kept here
    S->normb.assign((size_t)S->k, 0.0);
    for (int q = 0; q < S->k; ++q) {
        const auto rg = J->rhs[(size_t)q];
        S->normb[(size_t)q] = std::sqrt((double)(rg.second - rg.first));
    }
    RCCHECK(ensure_basis(ctx, *S, 1));
    S->total_steps.assign((size_t)S->k, 0);
    S->t0 = clk::now();
    S->phase = kInit;
    Cohort* C = &E->co[0];
    for (Cohort& c : E->co)
        if (c.load < C->load) C = &c;
    C->load += J->L * (int64_t)S->k;
    C->active.push_back(std::move(S));
    return FCG_OK;
}

// Plan the next step of cohort C and queue it; its bookkeeping is engine_finish.
int engine_launch(fcg_engine* E, Cohort& C) {
    fcg_ctx* ctx = E->ctx;
    hipStream_t s = ctx->stream;
    C.t_plan = clk::now();
    C.tabs->reset();  // the cohort's previous step has completed (engine_finish)
    C.tabs->defer = true;
    const bool mixed = E->opts.precision == FCG_PREC_MIXED;
    // Every table of the step is written to pinned staging first and reaches the
    // device in one copy; the launches are queued here and issued after that copy.
    // The results go back to the host after the step's last kernel, on a stream of
    // their own (dq), so the other cohort's next kernels do not wait for them.
    std::vector<std::function<hipError_t()>> q, dq;
    // two-level inverses still to build for the step's starting solves (one
    // launch, queued ahead of their first use).  A job listed here is marked as
    // having its inverse; until the step's launches have all been issued that
    // is not sure, so every error exit clears the marks and the next use builds
    // the inverse again.
    std::vector<Job*> sinv;
    struct MarkGuard {
        std::vector<Job*>& jobs;
        bool armed;
        ~MarkGuard() {
            if (armed)
                for (Job* j : jobs) j->spre.dim = 0;
        }
    } sinv_guard{sinv, true};
    // The download stream pays off only while another cohort has work in flight;
    // alone (a single solve), the copies stay on the compute stream.
    bool overlap = false;
    for (const Cohort& c : E->co) overlap |= (&c != &C && c.inflight);
    const hipStream_t ds = overlap ? ctx->down_stream : s;

    // ---- A: cycle starts' first basis vectors, solution updates and starts
    {
        std::vector<KEnt> sc;
        std::vector<double> dvals;
        for (auto& up : C.active) {
            Solve& S = *up;
            if (S.phase != kArn || !S.need_scale) continue;
            RCCHECK(ensure_basis(ctx, S, S.i + 1));
            sc.push_back(KEnt{&S, S.r.as<double>(), S.i, 0});
            for (int q = 0; q < S.k; ++q) dvals.push_back(S.d[(size_t)q]);
        }
        if (!sc.empty()) {
            KrylovView kv;
            RCCHECK(make_krylov(ctx, *C.tabs, sc, &kv));
            void* dd;
            HIPCHECK(ctx, C.tabs->push(dvals.data(), sizeof(double) * dvals.size(), s, &dd));
            q.push_back([=] { return fcg::launch_scale_new(kv, reinterpret_cast<double*>(dd), s); });
            for (auto& e : sc) e.S->need_scale = false;
        }
        std::vector<KEnt> up;
        std::vector<PEnt> upp;  // block preconditioner: x += P (Q z)
        std::vector<SEnt> ups;  // two-level preconditioner, likewise
        int64_t nzall = 0;
        for (auto& u : C.active) {
            Solve& S = *u;
            if (S.phase != kUpdate) continue;
            up.push_back(KEnt{&S, S.w.as<double>(), 0, (int)nzall});
            nzall += (int64_t)S.nz * S.k;
            if (S.block) upp.push_back(PEnt{S.J, S.u.as<double>(), false, S.x.as<double>(), S.k});
            if (S.twolev) ups.push_back(SEnt{S.J, S.u.as<double>(), false, S.x.as<double>(), S.k, S.spw.as<double>()});
        }
        if (!up.empty()) {
            KrylovView kv;
            RCCHECK(make_krylov(ctx, *C.tabs, up, &kv));
            RCCHECK(ensure_pbuf(ctx, C.zbuf, sizeof(double) * (size_t)std::max<int64_t>(nzall, 1)));
            double* dz = C.zbuf.as<double>();
            // the least squares of every ended cycle, then x += P (Q z)
            q.push_back([=] { return fcg::launch_dcgs_lsq(kv, dz, s); });
            q.push_back([=] { return fcg::launch_xupdate(kv, dz, s); });
            if (!upp.empty()) {
                fcg::PView pv;
                RCCHECK(make_pview(ctx, *C.tabs, upp, &pv));
                q.push_back([=] { return fcg::launch_block_apply(pv, fcg::kPreAdd, s); });
            }
            if (!ups.empty()) {
                fcg::SView sv;
                RCCHECK(make_sview(ctx, *C.tabs, ups, &sv));
                q.push_back([=] { return fcg::launch_super_apply(sv, fcg::kPreAdd, s); });
            }
            for (auto& e : up) {
                e.S->nz = 0;
                e.S->phase = e.S->mixed ? kResid : kFinish;
            }
        }
        std::vector<KEnt> init;
        std::vector<PEnt> initp;  // block preconditioner: x0 = P b
        std::vector<SEnt> inits;  // two-level preconditioner, likewise
        const int sdim = clamp_super_dim(E->opts.precond_size);
        for (auto& u : C.active) {
            if (u->phase != kInit || !u->twolev || u->J->spre.dim == sdim) continue;
            if (u->J->spre.m > fcg::kSuperLds) {
                RCCHECK(ensure_super_precond(ctx, *u->J, sdim));
                continue;
            }
            u->J->spre.dim = sdim;  // marks the job as queued (a job may start several solves)
            sinv.push_back(u->J);
        }
        RCCHECK(queue_super_inverses(ctx, *C.tabs, sinv, q));
        for (auto& u : C.active) {
            if (u->phase != kInit) continue;
            if (u->block)
                initp.push_back(PEnt{u->J, nullptr, false, u->x.as<double>(), u->k});
            else if (u->twolev)
                inits.push_back(SEnt{u->J, nullptr, false, u->x.as<double>(), u->k, u->spw.as<double>()});
            else
                init.push_back(KEnt{u.get(), u->w.as<double>(), 0, 0});
            u->phase = kResid;
        }
        if (!init.empty()) {
            KrylovView kv;
            RCCHECK(make_krylov(ctx, *C.tabs, init, &kv));
            q.push_back([=] { return fcg::launch_init_x(kv, s); });
        }
        if (!initp.empty()) {
            fcg::PView pv;
            RCCHECK(make_pview(ctx, *C.tabs, initp, &pv));
            q.push_back([=] { return fcg::launch_block_apply(pv, fcg::kPreFromB, s); });
        }
        if (!inits.empty()) {
            fcg::SView sv;
            RCCHECK(make_sview(ctx, *C.tabs, inits, &sv));
            q.push_back([=] { return fcg::launch_super_apply(sv, fcg::kPreFromB, s); });
        }
    }
    // finishing solves: conductor charge sums (SolveCapacitance.cpp:1305-1366), one
    // launch and one copy for all of them
    std::vector<std::pair<Solve*, int64_t>> fin;
    std::vector<double*> fin_x;
    std::vector<fcg::ChargeSumJob> csj;
    int64_t fin_vals = 0, fin_blk = 0;
    for (auto& u : C.active) {
        Solve& S = *u;
        if (S.phase != kFinish) continue;
        const int64_t nsum = (int64_t)S.k * S.ncol;  // blocks: one per (rhs, conductor)
        const int64_t nval = S.J->cplx ? 2 * nsum : nsum;  // complex: real sums, then imaginary
        csj.push_back(fcg::ChargeSumJob{S.x.as<double>(), S.J->leaf_w.as<double>(),
                                        S.J->cplx ? S.J->leaf_w_im.as<double>() : nullptr,
                                        S.J->col_range.as<int32_t>(), S.J->N, S.k, S.ncol, fin_blk, fin_vals});
        fin.emplace_back(&S, fin_vals);
        fin_vals += nval;
        fin_blk += nsum;
        double* hx = nullptr;
        if (S.want_x) {
            void* hxv;
            HIPCHECK(ctx, C.tabs->host_area(sizeof(double) * (size_t)S.k * S.N, &hxv));
            hx = reinterpret_cast<double*>(hxv);
            const void* xs = S.x.p;
            const size_t nx = sizeof(double) * (size_t)S.k * S.N;
            dq.push_back([=] { return hipMemcpyAsync(hx, xs, nx, hipMemcpyDeviceToHost, ds); });
        }
        fin_x.push_back(hx);
    }
    double* fin_host = nullptr;
    PBuf& fin_dev = C.fin_dev;
    if (!csj.empty()) {
        void *dt, *h;
        HIPCHECK(ctx, C.tabs->push(csj.data(), sizeof(fcg::ChargeSumJob) * csj.size(), s, &dt));
        HIPCHECK(ctx, C.tabs->host_area(sizeof(double) * (size_t)fin_vals, &h));
        HIPCHECK(ctx, fin_dev.get(&ctx->pool, sizeof(double) * (size_t)fin_vals));
        fin_host = reinterpret_cast<double*>(h);
        const auto* jt = reinterpret_cast<const fcg::ChargeSumJob*>(dt);
        const int nj = (int)csj.size();
        const int64_t nb = fin_blk;
        double* out = fin_dev.as<double>();
        const size_t nbytes = sizeof(double) * (size_t)fin_vals;
        q.push_back([=] { return fcg::launch_charge_sums(jt, nj, nb, out, s); });
        dq.push_back([=] { return hipMemcpyAsync(h, out, nbytes, hipMemcpyDeviceToHost, ds); });
    }

    // ---- B: refinement residuals r = b - A x
    std::vector<Solve*> res, arn;
    for (auto& u : C.active) {
        if (u->phase == kResid) res.push_back(u.get());
        if (u->phase == kArn) arn.push_back(u.get());
    }
    // step buffer: what the host reads back first (residual norms, then each
    // Arnoldi rhs's status and residual estimate), then the Arnoldi projections
    int hoff = 0;
    for (Solve* S : res) {
        S->hoff = hoff;
        hoff += S->k;
    }
    for (Solve* S : arn) {
        S->ooff = hoff;
        hoff += 2 * S->k;
    }
    const int hdown = hoff;
    for (Solve* S : arn) {
        S->hoff = hoff;
        hoff += (2 * S->i + 3) * S->k;  // DCGS2 projection of step i
    }
    RCCHECK(ensure_pbuf(ctx, C.hbuf, sizeof(double) * (size_t)std::max(hoff, 1)));
    if (!res.empty()) {
        std::vector<BatchEntry> be;
        std::vector<KEnt> ke;
        for (Solve* S : res) {
            be.push_back(BatchEntry{S->J, &S->scr, S->x.as<double>(), S->w.as<double>()});
            ke.push_back(KEnt{S, S->w.as<double>(), 0, 0});
            S->n_mvp++;
        }
        BatchView bv;
        fcg::GatherPlan gp;
        RCCHECK(make_batch(ctx, *C.tabs, be, &bv, &gp));
        const int prec = mixed ? fcg::kMvp32Acc : fcg::kMvp64;
        choose_variants(ctx, prec, be, gp);
        HIPCHECK(ctx, C.pe_res.arm(ctx->profiling, prec, ctx->profiling ? gather_alg_bytes(be, prec) : 0.0, gp.n));
        hipEvent_t* ev = C.pe_res.events();
        q.push_back([=] { return fcg::launch_mvp(bv, prec, 0, gp, s, ev); });
        KrylovView kv;
        RCCHECK(make_krylov(ctx, *C.tabs, ke, &kv));
        RCCHECK(ensure_pbuf(ctx, C.partial_a, sizeof(double) * (size_t)kv.nchunks));
        double* pa = C.partial_a.as<double>();
        double* hb = C.hbuf.as<double>();
        q.push_back([=] { return fcg::launch_resid(kv, pa, s); });
        q.push_back([=] { return fcg::launch_reduce_norm(kv, pa, hb, s); });
    }

    // ---- C: one Arnoldi step of every active cycle: the product on the
    // provisional vector, its projection, the Hessenberg column and the update
    // that completes the basis vector (DCGS2, kernels.hip k_dcgs_*)
    if (!arn.empty()) {
        std::vector<BatchEntry> be;
        std::vector<KEnt> ke;
        std::vector<PEnt> pe;  // block preconditioner: the product runs on u = P q_i
        std::vector<SEnt> se;  // two-level preconditioner, likewise
        int64_t ncoef = 0;
        for (Solve* S : arn) {
            RCCHECK(ensure_basis(ctx, *S, S->i + 1));
            RCCHECK(ensure_hstate(ctx, *S, S->i + 1, s));
            void* qi = S->hq[(size_t)S->i];
            if (S->block || S->twolev) {
                if (S->block)
                    pe.push_back(PEnt{S->J, qi, S->b32, S->u.as<double>(), S->k});
                else
                    se.push_back(SEnt{S->J, qi, S->b32, S->u.as<double>(), S->k, S->spw.as<double>()});
                be.push_back(BatchEntry{S->J, &S->scr, S->u.as<double>(), S->w.as<double>(), nullptr});
            } else {
                be.push_back(BatchEntry{S->J, &S->scr, S->b32 ? nullptr : static_cast<const double*>(qi),
                                        S->w.as<double>(), S->b32 ? static_cast<const float*>(qi) : nullptr});
            }
            KEnt e{S, S->w.as<double>(), S->i, 0};  // final vectors 0..i-1, provisional i
            e.coff = (int)ncoef;
            e.next = true;
            ncoef += (int64_t)(2 * S->i + 3) * S->k;
            ke.push_back(e);
            S->n_mvp++;
        }
        if (!pe.empty()) {
            fcg::PView pv;
            RCCHECK(make_pview(ctx, *C.tabs, pe, &pv));
            q.push_back([=] { return fcg::launch_block_apply(pv, fcg::kPreSet, s); });
        }
        if (!se.empty()) {
            fcg::SView sv;
            RCCHECK(make_sview(ctx, *C.tabs, se, &sv));
            q.push_back([=] { return fcg::launch_super_apply(sv, fcg::kPreSet, s); });
        }
        BatchView bv;
        fcg::GatherPlan gp;
        RCCHECK(make_batch(ctx, *C.tabs, be, &bv, &gp));
        const int prec = mixed ? fcg::kMvp32 : fcg::kMvp64;
        choose_variants(ctx, prec, be, gp);
        HIPCHECK(ctx, C.pe_arn.arm(ctx->profiling, prec, ctx->profiling ? gather_alg_bytes(be, prec) : 0.0, gp.n));
        hipEvent_t* ev = C.pe_arn.events();
        const int jac = arn[0]->jac ? 1 : 0;
        q.push_back([=] { return fcg::launch_mvp(bv, prec, jac, gp, s, ev); });
        KrylovView kv;
        RCCHECK(make_krylov(ctx, *C.tabs, ke, &kv));
        RCCHECK(ensure_pbuf(ctx, C.partial_b,
                            sizeof(double) * (size_t)kv.nchunks * 4 * (2 * (size_t)kv.maxnv + 3)));
        double* pb = C.partial_b.as<double>();
        double* hb = C.hbuf.as<double>();
        RCCHECK(ensure_pbuf(ctx, C.coefb, sizeof(double) * (size_t)std::max<int64_t>(ncoef, 1)));
        double* cb = C.coefb.as<double>();
        q.push_back([=] { return fcg::launch_dcgs_dots(kv, pb, s); });
        q.push_back([=] { return fcg::launch_dcgs_coef(kv, pb, hb, cb, s); });
        q.push_back([=] { return fcg::launch_dcgs_update(kv, cb, s); });
    }

    // ---- D: one table copy, the queued launches, one copy back, one synchronisation
    double* hh = nullptr;
    if (hdown > 0) {
        void* h;
        HIPCHECK(ctx, C.tabs->host_area(sizeof(double) * (size_t)hdown, &h));
        hh = reinterpret_cast<double*>(h);
        const void* src = C.hbuf.p;
        const size_t nb = sizeof(double) * (size_t)hdown;
        dq.push_back([=] { return hipMemcpyAsync(hh, src, nb, hipMemcpyDeviceToHost, ds); });
    }
    // The tables go up by a copy kernel on the compute stream, right after the
    // other cohort's kernels in flight (a copy on the copy engine would make the
    // compute stream wait on another queue's signal, about 50 us per step).
    HIPCHECK(ctx, C.tabs->flush(s));
    for (auto& f : q) HIPCHECK(ctx, f());
    sinv_guard.armed = false;  // every inverse launch of the step is queued
    if (ds != s) {
        if (!C.kev) HIPCHECK(ctx, hipEventCreateWithFlags(&C.kev, hipEventDisableTiming));
        HIPCHECK(ctx, hipEventRecord(C.kev, s));
        HIPCHECK(ctx, hipStreamWaitEvent(ds, C.kev, 0));
    }
    for (auto& f : dq) HIPCHECK(ctx, f());
    if (!C.ev) HIPCHECK(ctx, hipEventCreateWithFlags(&C.ev, hipEventDisableTiming));
    HIPCHECK(ctx, hipEventRecord(C.ev, ds));  // the kernels and the downloads of the step
    C.inflight = true;
    C.res = std::move(res);
    C.arn = std::move(arn);
    C.fin = std::move(fin);
    C.fin_x = std::move(fin_x);
    C.fin_host = fin_host;
    C.hh = hh;
    C.t_wait = clk::now();
    return FCG_OK;
}

// Wait for cohort C's step in flight and do its bookkeeping.
int engine_finish(fcg_engine* E, Cohort& C) {
    fcg_ctx* ctx = E->ctx;
    const auto t_wait = clk::now();
    HIPCHECK(ctx, hipEventSynchronize(C.ev));
    C.inflight = false;
    const auto t_post = clk::now();
    C.pe_res.collect(ctx->prof);
    C.pe_arn.collect(ctx->prof);
    const auto& res = C.res;
    const auto& arn = C.arn;
    const auto& fin = C.fin;
    const auto& fin_x = C.fin_x;
    const double* fin_host = C.fin_host;
    const double* hh = C.hh;

    // ---- host bookkeeping
    for (size_t f = 0; f < fin.size(); ++f) {
        Solve& S = *fin[f].first;
        Ready rd;
        rd.tag = S.tag;
        rd.job = S.job_id;
        rd.rows = S.k;
        rd.cols = S.ncol;
        rd.cap.assign(fin_host + fin[f].second,
                      fin_host + fin[f].second + (size_t)S.k * S.ncol * (S.J->cplx ? 2 : 1));
        if (fin_x[f]) rd.x.assign(fin_x[f], fin_x[f] + (size_t)S.k * S.N);
        std::memset(&rd.stats, 0, sizeof rd.stats);
        rd.stats.n_rhs = S.k;
        int mx = 0;
        for (int q = 0; q < S.k; ++q) {
            if (q < 64) rd.stats.iterations[q] = S.total_steps[(size_t)q];
            mx = std::max(mx, S.total_steps[(size_t)q]);
        }
        rd.stats.max_iterations = mx;
        rd.stats.final_residual = S.rel;
        rd.stats.n_mvp = S.n_mvp;
        rd.stats.converged = S.converged ? 1 : 0;
        rd.stats.cycles = S.cycles;
        rd.stats.t_solve_s = std::chrono::duration<double>(clk::now() - S.t0).count();
        rd.stats.mem_gmres = (int64_t)S.hd.r;
        for (const auto& qb : S.qblocks) rd.stats.mem_gmres += (int64_t)qb->r;
        rd.stats.mem_vectors = (int64_t)(S.x.r + S.r.r + S.w.r + S.u.r + S.spw.r);
        rd.stats.mem_product = (int64_t)S.scr.bytes();
        rd.hist = std::move(S.hist);
        E->ready.push_back(std::move(rd));
        S.phase = kDone;
    }
    for (Solve* S : res) {
        std::vector<double> beta((size_t)S->k);
        double worst = 0.0;
        bool all = true;
        for (int q = 0; q < S->k; ++q) {
            beta[(size_t)q] = std::sqrt(hh[S->hoff + q]);
            const double rel = beta[(size_t)q] / S->normb[(size_t)q];
            worst = std::max(worst, rel);
            if (!(rel < E->opts.tol)) all = false;
        }
        if (S->mixed) {
            S->rel = worst;
            S->converged = all;
        }
        // a mixed solve that reaches max_refine cycles finishes unconverged
        // (stats.converged = 0, final_residual above tol)
        if (S->mixed && (all || S->cycles >= std::max(1, E->opts.max_refine)))
            S->phase = kFinish;
        else
            start_cycle(E, *S, beta.data());
    }
    for (Solve* S : arn) {
        // per rhs: status (0 iterated, 1 met the tolerance before, 2 the
        // provisional vector lay in the basis, 3 exact breakdown) and |g_(i+1)|
        const double* out = hh + S->ooff;
        for (int q = 0; q < S->k; ++q) {
            const double st = out[2 * q], gr = out[2 * q + 1];
            if (st == 1.0) continue;
            if (st == 2.0) {
                S->done[(size_t)q] = 1;
                continue;
            }
            S->total_steps[(size_t)q]++;
            S->last_i[(size_t)q] = S->i;
            S->gres[(size_t)q] = gr;
            if (S->want_hist && (size_t)q < S->hist.size()) S->hist[(size_t)q].push_back(gr / S->normb[(size_t)q]);
            if (gr < S->thr[(size_t)q] || st == 3.0) S->done[(size_t)q] = 1;
        }
        bool all = true;
        double worst = 0.0;
        for (int q = 0; q < S->k; ++q) {
            if (!S->done[(size_t)q]) all = false;
            const int it = S->last_i[(size_t)q];
            if (it >= 0) worst = std::max(worst, S->gres[(size_t)q] / S->normb[(size_t)q]);
        }
        if (!S->mixed) {
            S->rel = worst;
            S->converged = all;
        }
        // basis vector i is final (the step's update also wrote the next
        // provisional vector, slot i + 1)
        if (all || S->i + 1 >= S->cap_iter)
            prepare_update(*S);
        else
            S->i += 1;
    }
    // drop finished solves (their device memory returns to the pool)
    for (auto& u : C.active)
        if (u->phase == kDone) C.load -= u->J->L * (int64_t)u->k;
    C.active.erase(std::remove_if(C.active.begin(), C.active.end(),
                                  [](const std::unique_ptr<Solve>& u) { return u->phase == kDone; }),
                   C.active.end());
    C.res.clear();
    C.arn.clear();
    C.fin.clear();
    C.fin_x.clear();
    if (ctx->profiling) {
        const auto t_end = clk::now();
        ctx->prof.host_plan_ms += std::chrono::duration<double, std::milli>(C.t_wait - C.t_plan).count();
        ctx->prof.host_wait_ms += std::chrono::duration<double, std::milli>(t_post - t_wait).count();
        ctx->prof.host_post_ms += std::chrono::duration<double, std::milli>(t_end - t_post).count();
    }
    return FCG_OK;
}

// One engine step: every cohort's step in flight is finished and its next one
// queued, cohort after cohort, so the device runs one cohort while the host
// serves the other.
int engine_step(fcg_engine* E) {
    for (Cohort& C : E->co) {
        if (C.inflight) RCCHECK(engine_finish(E, C));
        if (!C.active.empty()) RCCHECK(engine_launch(E, C));
    }
    if (E->ctx->profiling) E->ctx->prof.steps += 1;
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
    if (scratch_address(device, &c->noalloc) != hipSuccess || !c->noalloc) return FCG_DEVICE_ERROR;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return FCG_DEVICE_ERROR;
    if (hipStreamCreateWithFlags(&c->down_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        return FCG_DEVICE_ERROR;
    }
    *out = c.release();
    return FCG_OK;
}

int fcg_ctx_destroy(fcg_ctx* ctx) {
    if (!ctx) return FCG_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->down_stream);
    hipStream_t s = ctx->stream, ds = ctx->down_stream;
    delete ctx;
    (void)hipStreamDestroy(s);
    (void)hipStreamDestroy(ds);
    return FCG_OK;
}

const char* fcg_last_error(fcg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

}  // extern "C"

namespace {
// tree arrays of an upload (both layouts)
int upload_tree(fcg_ctx* ctx, const fcg_job_desc* d, Job& j) {
    j.dim = d->dim;
    // Galerkin in 3D; 2D jobs are solved by collocation (SolveCapacitance.cpp:159-161)
    j.scheme = d->dim == 3 && d->scheme == FCG_SCHEME_GALERKIN ? FCG_SCHEME_GALERKIN : FCG_SCHEME_COLLOCATION;
    j.N = d->n_leaves;
    j.M = d->n_nodes;
    j.G = d->n_groups;
    j.L = d->n_links;
    hipStream_t s = ctx->stream;
    HIPCHECK(ctx, upload(j.nodes, d->nodes, (size_t)j.M, s));
    HIPCHECK(ctx, upload(j.left, d->left, (size_t)j.M, s));
    HIPCHECK(ctx, upload(j.right, d->right, (size_t)j.M, s));
    HIPCHECK(ctx, upload(j.parent, d->parent, (size_t)j.M, s));
    HIPCHECK(ctx, upload(j.leaf_lo, d->leaf_lo, (size_t)j.M, s));
    HIPCHECK(ctx, upload(j.leaf_hi, d->leaf_hi, (size_t)j.M, s));
    j.h_left.assign(d->left, d->left + j.M);
    j.h_right.assign(d->right, d->right + j.M);
    j.h_parent.assign(d->parent, d->parent + j.M);
    j.h_leaf_lo.assign(d->leaf_lo, d->leaf_lo + j.M);
    j.h_leaf_hi.assign(d->leaf_hi, d->leaf_hi + j.M);
    HIPCHECK(ctx, upload(j.leaf_node, d->leaf_node, (size_t)j.N, s));
    return FCG_OK;
}

// gather row records (fcg::GRow) of a job: for every row in gather order, its
// link range, charge row and node (link offsets fit in 32 bits: validate_desc)
int upload_grow(fcg_ctx* ctx, Job& j, const int64_t* row_ptr, const std::vector<int32_t>& perm,
                const std::vector<int32_t>& rorder) {
    std::vector<fcg::GRow> g((size_t)j.M);
    for (int i = 0; i < j.M; ++i) {
        const int node = rorder[(size_t)i];
        g[(size_t)i] = fcg::GRow{(uint32_t)row_ptr[node], (uint32_t)row_ptr[node + 1], perm[(size_t)node], node};
    }
    HIPCHECK(ctx, upload(j.grow, g.data(), g.size(), ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));  // g is released here
    return FCG_OK;
}

// source geometry of a job by charge row (fcg::SrcGeom: the fill's link sources)
int upload_src_geom(fcg_ctx* ctx, Job& j, const fcg_node_geom* nodes, const std::vector<int32_t>& perm) {
    std::vector<fcg::SrcGeom> g((size_t)j.M);
    for (int i = 0; i < j.M; ++i) {
        fcg::SrcGeom& e = g[(size_t)perm[(size_t)i]];
        std::memcpy(e.v, nodes[i].v, sizeof e.v);
        std::memcpy(e.c, nodes[i].c, sizeof e.c);
    }
    HIPCHECK(ctx, upload(j.src_geom, g.data(), g.size(), ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));  // g is released here
    return FCG_OK;
}

// groups, right-hand sides, self and diagonal storage and the link rows (both
// layouts); returns synchronised
int upload_finish(fcg_ctx* ctx, const fcg_job_desc* d, Job& j) {
    hipStream_t s = ctx->stream;
    HIPCHECK(ctx, upload(j.gout, d->group_outperm, (size_t)j.G, s));
    HIPCHECK(ctx, upload(j.gin, d->group_inperm, (size_t)j.G, s));
    HIPCHECK(ctx, upload(j.leaf_w, d->leaf_outperm, (size_t)j.N, s));
    if (d->group_outperm_im || d->group_inperm_im || d->leaf_outperm_im) {
        if (!d->group_outperm_im || !d->group_inperm_im || !d->leaf_outperm_im)
            return fail(ctx, FCG_INVALID_ARGUMENT, "complex permittivities need all three imaginary arrays");
        j.cplx = true;
        HIPCHECK(ctx, upload(j.gout_im, d->group_outperm_im, (size_t)j.G, s));
        HIPCHECK(ctx, upload(j.gin_im, d->group_inperm_im, (size_t)j.G, s));
        HIPCHECK(ctx, upload(j.leaf_w_im, d->leaf_outperm_im, (size_t)j.N, s));
        HIPCHECK(ctx, j.self_im.alloc(sizeof(double) * (size_t)j.N));
    }
    HIPCHECK(ctx, j.link_row.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(j.L, 1)));
    HIPCHECK(ctx, j.self.alloc(sizeof(double) * (size_t)j.N));
    HIPCHECK(ctx, j.diag.alloc(sizeof(double) * (size_t)j.M));
    HIPCHECK(ctx, hipMemsetAsync(j.diag.p, 0, sizeof(double) * (size_t)j.M, s));
    j.group_is_diel.assign(d->group_is_diel, d->group_is_diel + j.G);
    j.group_leaf_start.assign(d->group_leaf_start, d->group_leaf_start + j.G + 1);
    j.first_cond_leaf = j.N;
    std::vector<int32_t> ranges;
    for (int g = 0; g < j.G; ++g) {
        if (j.group_is_diel[(size_t)g]) continue;
        const int lo = j.group_leaf_start[(size_t)g], hi = j.group_leaf_start[(size_t)g + 1];
        if (j.first_cond_leaf == j.N) j.first_cond_leaf = lo;
        ranges.push_back(lo);
        ranges.push_back(hi);
        ++j.ncol;
        // 2D: the conductor that ends the leaf list is the balance row (SolveCapacitance.cpp:1233-1240)
        if (!(j.dim == 2 && hi == j.N)) j.rhs.emplace_back(lo, hi);
    }
    if (ranges.empty()) return fail(ctx, FCG_INVALID_ARGUMENT, "job has no conductor");
    {
        const int cls = fcg::gather_class(fcg::padded_k(std::max(1, (int)j.rhs.size())));
        for (int p = 0; p < 3; ++p) j.gather_variant[p] = fcg::default_gather_variant(p, cls);
    }
    HIPCHECK(ctx, upload(j.col_range, ranges.data(), ranges.size(), s));
    {
        std::vector<int32_t> rr;
        for (const auto& rg : j.rhs) {
            rr.push_back(rg.first);
            rr.push_back(rg.second);
        }
        if (rr.empty()) rr.assign(2, 0);
        HIPCHECK(ctx, upload(j.rhs_range, rr.data(), rr.size(), s));
        HIPCHECK(ctx, hipStreamSynchronize(s));  // rr and ranges are released at the end of this block
    }
    HIPCHECK(ctx, hipStreamSynchronize(s));
    // destination node of every link (used by the fill)
    ctx->tabs.reset();
    BatchView bv;
    RCCHECK(make_batch(ctx, ctx->tabs, {BatchEntry{&j, &j.scr, nullptr, nullptr}}, &bv));
    HIPCHECK(ctx, fcg::launch_link_rows(bv, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    return FCG_OK;
}

// sorts every row of col (device) ascending: rows up to the workgroup sort's
// LDS capacity in LDS, the longer ones in a scratch copy (one launch each)
int sort_rows(fcg_ctx* ctx, const DevBuf& drow, const std::vector<int64_t>& row_ptr, int32_t* col, int M) {
    hipStream_t s = ctx->stream;
    HIPCHECK(ctx, fcg::launch_links_sort_rows(drow.as<int64_t>(), col, M, s));
    const int cap = fcg::links_sort_lds_max();
    std::vector<int32_t> rows;
    std::vector<int64_t> off(1, 0);
    for (int i = 0; i < M; ++i) {
        const int64_t n = row_ptr[(size_t)i + 1] - row_ptr[(size_t)i];
        if (n <= cap) continue;
        if (n > (int64_t)INT32_MAX / 2) return fail(ctx, FCG_INVALID_ARGUMENT, "link row too long");
        int64_t P = 2;
        while (P < n) P <<= 1;
        rows.push_back(i);
        off.push_back(off.back() + P);
    }
    if (rows.empty()) return FCG_OK;
    DevBuf drows, doff, scratch;
    HIPCHECK(ctx, upload(drows, rows.data(), rows.size(), s));
    HIPCHECK(ctx, upload(doff, off.data(), off.size(), s));
    HIPCHECK(ctx, scratch.alloc(sizeof(int32_t) * (size_t)off.back()));
    HIPCHECK(ctx, fcg::launch_links_sort_long(drow.as<int64_t>(), drows.as<int32_t>(), (int)rows.size(),
                                              doff.as<int64_t>(), col, scratch.as<int32_t>(), s));
    HIPCHECK(ctx, hipStreamSynchronize(s));  // the scratch and the lists are released here
    return FCG_OK;
}
}  // namespace

extern "C" {

int fcg_job_upload(fcg_ctx* ctx, const fcg_job_desc* d, int* job) {
    if (!ctx || !job) return FCG_INVALID_ARGUMENT;
    RCCHECK(validate_desc(ctx, d));
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    auto j = std::make_unique<Job>();
    hipStream_t s = ctx->stream;
    RCCHECK(upload_tree(ctx, d, *j));
    HIPCHECK(ctx, upload(j->row_ptr, d->row_ptr, (size_t)j->M + 1, s));
    {
        std::vector<int32_t> perm, iperm, rorder, cols;
        build_layout(d, perm, iperm, rorder, cols, j->link_pos);
        HIPCHECK(ctx, upload(j->perm, perm.data(), perm.size(), s));
        HIPCHECK(ctx, upload(j->iperm, iperm.data(), iperm.size(), s));
        HIPCHECK(ctx, upload(j->rorder, rorder.data(), rorder.size(), s));
        RCCHECK(upload_grow(ctx, *j, d->row_ptr, perm, rorder));
        RCCHECK(upload_src_geom(ctx, *j, d->nodes, perm));
        RCCHECK(upload_tree_passes(ctx, *j, d, perm));
        std::vector<uint32_t> forder;
        build_fill_order(d, iperm, cols, forder, j->fill_cnt);
        if (!forder.empty()) HIPCHECK(ctx, upload(j->fill_order, forder.data(), forder.size(), s));
        std::vector<int64_t> fstart((size_t)kFillClasses + 1, 0);
        for (int c = 0; c < kFillClasses; ++c) fstart[(size_t)c + 1] = fstart[(size_t)c] + j->fill_cnt[(size_t)c];
        HIPCHECK(ctx, upload(j->fill_start, fstart.data(), fstart.size(), s));
        cols.resize(cols.size() + kLinkSpare, 0);
        HIPCHECK(ctx, upload(j->col, cols.data(), cols.size(), s));
        if (j->M <= 65536) {
            std::vector<uint16_t> c16(cols.size());
            for (size_t l = 0; l < cols.size(); ++l) c16[l] = (uint16_t)cols[l];
            HIPCHECK(ctx, upload(j->col16, c16.data(), c16.size(), s));
            HIPCHECK(ctx, hipStreamSynchronize(s));  // c16 is released here
        }
        j->h_row_ptr.assign(d->row_ptr, d->row_ptr + j->M + 1);
        j->h_iperm = iperm;
        HIPCHECK(ctx, hipStreamSynchronize(s));  // the host vectors above are released here
    }
    RCCHECK(upload_finish(ctx, d, *j));
    ctx->jobs.push_back(std::move(j));
    *job = (int)ctx->jobs.size() - 1;
    return FCG_OK;
}

int fcg_job_upload_links(fcg_ctx* ctx, const fcg_job_desc* d, int* job) {
    if (!ctx || !job || !d) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!ctx->links.ready || ctx->links.row_ptr.size() != (size_t)d->n_nodes + 1)
        return fail(ctx, FCG_INVALID_ARGUMENT, "no discovered links for a tree of this size");
    fcg_job_desc dd = *d;
    dd.row_ptr = ctx->links.row_ptr.data();
    dd.n_links = ctx->links.L;
    dd.col = nullptr;
    RCCHECK(validate_desc(ctx, &dd, false));
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    auto j = std::make_unique<Job>();
    hipStream_t s = ctx->stream;
    RCCHECK(upload_tree(ctx, &dd, *j));
    const int M = j->M;
    const int64_t L = j->L;
    HIPCHECK(ctx, upload(j->row_ptr, dd.row_ptr, (size_t)M + 1, s));
    std::vector<int32_t> perm, iperm, rorder;
    build_tree_layout(&dd, perm, iperm, rorder);
    HIPCHECK(ctx, upload(j->perm, perm.data(), perm.size(), s));
    HIPCHECK(ctx, upload(j->iperm, iperm.data(), iperm.size(), s));
    HIPCHECK(ctx, upload(j->rorder, rorder.data(), rorder.size(), s));
    RCCHECK(upload_grow(ctx, *j, dd.row_ptr, perm, rorder));
    RCCHECK(upload_src_geom(ctx, *j, dd.nodes, perm));
    RCCHECK(upload_tree_passes(ctx, *j, &dd, perm));
    // stored columns: the source's charge row, every row ascending
    HIPCHECK(ctx, j->col.alloc(sizeof(int32_t) * (size_t)(L + kLinkSpare)));
    HIPCHECK(ctx, hipMemsetAsync(j->col.as<int32_t>() + L, 0, sizeof(int32_t) * kLinkSpare, s));
    HIPCHECK(ctx, fcg::launch_perm_cols(ctx->links.col.as<int32_t>(), L, j->perm.as<int32_t>(), j->col.as<int32_t>(), s));
    RCCHECK(sort_rows(ctx, j->row_ptr, ctx->links.row_ptr, j->col.as<int32_t>(), M));
    if (M <= 65536) {
        HIPCHECK(ctx, j->col16.alloc(sizeof(uint16_t) * (size_t)(L + kLinkSpare)));
        HIPCHECK(ctx, fcg::launch_col16(j->col.as<int32_t>(), L + kLinkSpare, j->col16.as<uint16_t>(), s));
    }
    HIPCHECK(ctx, hipStreamSynchronize(s));
    j->h_row_ptr = std::move(ctx->links.row_ptr);
    j->h_iperm = std::move(iperm);
    ctx->links.reset();  // the discovered rows are consumed
    RCCHECK(upload_finish(ctx, &dd, *j));
    // fill order: links grouped by fill class, link order kept inside a class
    j->fill_cnt.assign(kFillClasses, 0);
    std::vector<int64_t> fstart((size_t)kFillClasses + 1, 0);
    if (L > 0) {
        DevBuf cls, hist, off;
        HIPCHECK(ctx, cls.alloc((size_t)L));
        HIPCHECK(ctx, fcg::launch_link_fill_class(j->link_row.as<int32_t>(), j->col.as<int32_t>(),
                                                  j->iperm.as<int32_t>(), j->nodes.as<fcg_node_geom>(), j->dim, L,
                                                  cls.as<uint8_t>(), s));
        const int64_t nb = (L + fcg::class_chunk() - 1) / fcg::class_chunk();
        HIPCHECK(ctx, hist.alloc(sizeof(int32_t) * (size_t)(nb * kFillClasses)));
        HIPCHECK(ctx, fcg::launch_class_hist(cls.as<uint8_t>(), L, hist.as<int32_t>(), s));
        std::vector<int32_t> h((size_t)(nb * kFillClasses));
        HIPCHECK(ctx, hipMemcpyAsync(h.data(), hist.p, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost, s));
        HIPCHECK(ctx, hipStreamSynchronize(s));
        std::vector<int64_t> o((size_t)(nb * kFillClasses));
        int64_t run = 0;
        for (int c = 0; c < kFillClasses; ++c) {
            fstart[(size_t)c] = run;
            for (int64_t b = 0; b < nb; ++b) {
                o[(size_t)(b * kFillClasses + c)] = run;
                run += h[(size_t)(b * kFillClasses + c)];
            }
            j->fill_cnt[(size_t)c] = run - fstart[(size_t)c];
        }
        if (run != L) return fail(ctx, FCG_DEVICE_ERROR, "fill order: class counts do not add up");
        HIPCHECK(ctx, upload(off, o.data(), o.size(), s));
        HIPCHECK(ctx, j->fill_order.alloc(sizeof(uint32_t) * (size_t)L));
        HIPCHECK(ctx, fcg::launch_class_scatter(cls.as<uint8_t>(), L, off.as<int64_t>(), j->fill_order.as<uint32_t>(), s));
        HIPCHECK(ctx, hipStreamSynchronize(s));  // cls, hist, off and the host vectors are released here
    }
    fstart[(size_t)kFillClasses] = L;
    HIPCHECK(ctx, upload(j->fill_start, fstart.data(), fstart.size(), s));
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
    *n_cond = j->ncol;
    return FCG_OK;
}

int fcg_job_is_complex(fcg_ctx* ctx, int job, int* is_complex) {
    Job* j = get_job(ctx, job);
    if (!j || !is_complex) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    *is_complex = j->cplx ? 1 : 0;
    return FCG_OK;
}

int fcg_job_footprint(fcg_ctx* ctx, int job, int precision, int iters, int64_t* resident, int64_t* solve) {
    Job* j = get_job(ctx, job);
    if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    const int64_t k = (int64_t)j->rhs.size();
    int64_t res = j->resident_bytes();
    if (precision == FCG_PREC_MIXED && !j->coef32.p) res += 4 * j->L;
    if (precision == FCG_PREC_MIXED && !j->near_ready) {
        int64_t ln = 0;
        for (int c = 0; c < kFillClasses; ++c)
            if (fcg::near_class(c)) ln += j->fill_cnt[(size_t)c];
        res += 16 * ln + 8 * ((int64_t)j->M + 1);
    }
    if (precision != FCG_PREC_MIXED && !j->coef64.p) res += 8 * j->L;
    if (resident) *resident = res;
    // vectors b, x, r, w and the basis, plus the solve's own product scratch
    const int64_t pk = j->pk((int)k);
    const int64_t scratch = pk * ((int64_t)(j->N + 1) * 16 + (int64_t)j->M * (precision == FCG_PREC_MIXED ? 24 : 16));
    if (solve) *solve = (int64_t)8 * k * j->vlen() * (4 + std::max(iters, 1) + 1) + scratch;
    return FCG_OK;
}

int fcg_fill_batch(fcg_ctx* ctx, const int* jobs, int njobs, int precision) {
    if (!ctx || !jobs || njobs < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (precision != FCG_PREC_FP64 && precision != FCG_PREC_MIXED)
        return fail(ctx, FCG_INVALID_ARGUMENT, "unknown precision mode");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    std::vector<Job*> js;
    for (int i = 0; i < njobs; ++i) {
        Job* j = get_job(ctx, jobs[i]);
        if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
        js.push_back(j);
    }
    {
        // a job listed more than once (a sorted copy: batches hold thousands of jobs)
        std::vector<Job*> sorted(js);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(ctx, FCG_INVALID_ARGUMENT, "job listed twice");
    }
    std::vector<BatchEntry> ents;
    for (Job* j : js) {
        if (precision == FCG_PREC_MIXED && !j->coef32.p)
            HIPCHECK(ctx, j->coef32.alloc(sizeof(float) * (size_t)(j->L + kLinkSpare)));
        if (precision == FCG_PREC_MIXED) RCCHECK(ensure_near(ctx, *j));
        RCCHECK(ensure_fill_rs(ctx, *j));
        if (precision == FCG_PREC_FP64 && !j->coef64.p)
            HIPCHECK(ctx, j->coef64.alloc(sizeof(double) * (size_t)(j->L + kLinkSpare)));
        ents.push_back(BatchEntry{j, &j->scr, nullptr, nullptr});
    }
    HIPCHECK(ctx, hipStreamSynchronize(s));
    ctx->tabs.reset();
    // the tables of every fill go up in one copy per slab, ahead of the launches
    ctx->tabs.defer = true;
    struct Undefer {
        Tables& t;
        ~Undefer() { t.defer = false; }
    } undefer{ctx->tabs};
    std::vector<std::function<hipError_t()>> q;
    // one fill per scheme present (the Galerkin kernels are instantiations of their own)
    for (const int scheme : {FCG_SCHEME_COLLOCATION, FCG_SCHEME_GALERKIN}) {
        std::vector<BatchEntry> part;
        for (const auto& e : ents)
            if (e.j->scheme == scheme) part.push_back(e);
        if (part.empty()) continue;
        BatchView bv;
        std::vector<int> order;
        RCCHECK(make_batch(ctx, ctx->tabs, part, &bv, nullptr, &order));
        // per class, the prefix of the jobs' link counts in the batch's job order
        std::array<const int64_t*, kFillClasses> offp;
        std::array<int64_t, kFillClasses> total;
        for (int c = 0; c < kFillClasses; ++c) {
            std::vector<int64_t> off(order.size() + 1, 0);
            for (size_t i = 0; i < order.size(); ++i)
                off[i + 1] = off[i] + part[(size_t)order[i]].j->fill_cnt[(size_t)c];
            total[c] = off.back();
            void* d;
            HIPCHECK(ctx, ctx->tabs.push(off.data(), sizeof(int64_t) * off.size(), s, &d));
            offp[c] = reinterpret_cast<const int64_t*>(d);
        }
        // mixed: the analytic classes from the near-field side lists
        const int64_t* near_off = nullptr;
        int64_t near_total = 0;
        if (precision == FCG_PREC_MIXED) {
            std::vector<int64_t> off(order.size() + 1, 0);
            for (size_t i = 0; i < order.size(); ++i) off[i + 1] = off[i] + part[(size_t)order[i]].j->Ln;
            near_total = off.back();
            void* d;
            HIPCHECK(ctx, ctx->tabs.push(off.data(), sizeof(int64_t) * off.size(), s, &d));
            near_off = reinterpret_cast<const int64_t*>(d);
        }
        const bool gal = scheme == FCG_SCHEME_GALERKIN;
        q.push_back([=] { return fcg::launch_fill(bv, offp.data(), total.data(), precision, near_off, near_total, gal, s); });
    }
    // no wait: later work on the context's stream runs after the fill, and every
    // reuse of the context's tables waits for the stream first
    HIPCHECK(ctx, ctx->tabs.flush(s));
    for (auto& f : q) HIPCHECK(ctx, f());
    for (auto& e : ents) {
        if (precision == FCG_PREC_MIXED)
            e.j->filled32 = true;
        else
            e.j->filled64 = true;
        // a fill starts a new solve: its two-level inverse is built again on first
        // use, as the reference builds the preconditioner after each fill
        e.j->spre.dim = 0;
    }
    return FCG_OK;
}

int fcg_fill(fcg_ctx* ctx, int job, int precision) { return fcg_fill_batch(ctx, &job, 1, precision); }

int fcg_download_coeffs(fcg_ctx* ctx, int job, double* coef, double* self_coef) {
    Job* j = get_job(ctx, job);
    if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    if (!j->filled64 && !j->filled32) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (coef && j->L > 0) {
        // stored order is sorted by source within each row; return descriptor order
        std::vector<double> tmp((size_t)j->L);
        if (j->filled64) {
            HIPCHECK(ctx, hipMemcpy(tmp.data(), j->coef64.p, sizeof(double) * (size_t)j->L, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t32((size_t)j->L);
            HIPCHECK(ctx, hipMemcpy(t32.data(), j->coef32.p, sizeof(float) * (size_t)j->L, hipMemcpyDeviceToHost));
            for (int64_t l = 0; l < j->L; ++l) tmp[(size_t)l] = t32[(size_t)l];
        }
        if (j->link_pos.empty()) {
            // rows discovered on the device: descriptor order is by source node id
            std::vector<int32_t> cols((size_t)j->L);
            HIPCHECK(ctx, hipMemcpy(cols.data(), j->col.p, sizeof(int32_t) * (size_t)j->L, hipMemcpyDeviceToHost));
            std::vector<std::pair<int32_t, int64_t>> row;
            for (int r = 0; r < j->M; ++r) {
                const int64_t b = j->h_row_ptr[(size_t)r], e = j->h_row_ptr[(size_t)r + 1];
                row.clear();
                for (int64_t l = b; l < e; ++l) row.emplace_back(j->h_iperm[(size_t)cols[(size_t)l]], l);
                std::sort(row.begin(), row.end());
                for (int64_t q = b; q < e; ++q) coef[q] = tmp[(size_t)row[(size_t)(q - b)].second];
            }
        } else {
            for (int r = 0; r < j->M; ++r)
                for (int64_t l = j->h_row_ptr[(size_t)r]; l < j->h_row_ptr[(size_t)r + 1]; ++l)
                    coef[j->h_row_ptr[(size_t)r] + j->link_pos[(size_t)l]] = tmp[(size_t)l];
        }
    }
    if (self_coef) HIPCHECK(ctx, hipMemcpy(self_coef, j->self.p, sizeof(double) * (size_t)j->N, hipMemcpyDeviceToHost));
    return FCG_OK;
}

int fcg_mvp_device(fcg_ctx* ctx, int job, const double* d_q, double* d_v, int k, void* stream) {
    Job* j = get_job(ctx, job);
    if (!j || !d_q || !d_v || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    return run_mvp(ctx, *j, fcg::kMvp64, d_q, d_v, k, 0, stream ? (hipStream_t)stream : ctx->stream);
}

int fcg_mvp_device_noalloc(fcg_ctx* ctx, int job, const double* d_q, double* d_v, int k, void* stream) {
    Job* j = get_job(ctx, job);
    if (!j || !d_q || !d_v || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled64) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled in FP64");
    if (!ctx->noalloc) return fail(ctx, FCG_DEVICE_ERROR, "no scratch address on this device");
    k = j->pk(k);  // complex jobs: two product vectors per right-hand side
    // scratch layout: job table, node charges and potentials [M][kp], 2D fix-up
    // values [2k]
    const int N = j->N, M = j->M, kp = fcg::padded_k(k);
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_charge = kNoAllocTableBytes;
    const size_t o_pot = o_charge + al(sizeof(double) * (size_t)kp * M);
    const size_t o_lt = o_pot + al(sizeof(double) * (size_t)kp * M);
    const size_t need = o_lt + al(sizeof(double) * 2 * (size_t)k);
    if (need > kNoAllocScratchBytes)
        return fail(ctx, FCG_INVALID_ARGUMENT,
                    "product needs " + std::to_string(need) + " bytes of scratch, the no-alloc limit is " +
                        std::to_string(kNoAllocScratchBytes));
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    char* base = static_cast<char*>(ctx->noalloc);
    Scratch sc;  // sizes only: the buffers are the scratch array's
    sc.k = k;
    sc.kp = kp;
    NoAllocTable t{};
    t.job = j->rec(sc, d_q, d_v);
    t.job.charge64 = reinterpret_cast<double*>(base + o_charge);
    t.job.pot64 = reinterpret_cast<double*>(base + o_pot);
    t.job.lt = reinterpret_cast<double*>(base + o_lt);
    t.job.charge32 = nullptr;
    t.job.pot32 = nullptr;
    const int64_t sizes[5] = {j->L, fcg::tree_chunks(M), M, N, k};
    for (int i = 0; i < 5; ++i) {
        t.off[2 * i] = 0;
        t.off[2 * i + 1] = sizes[i];
    }
    k_noalloc_table<<<1, 64, 0, s>>>(t, reinterpret_cast<uint64_t*>(base));
    HIPCHECK(ctx, hipGetLastError());
    const auto* tab = reinterpret_cast<const NoAllocTable*>(base);
    BatchView bv{};
    bv.jobs = &tab->job;
    bv.njobs = 1;
    bv.link_off = tab->off;
    bv.tchunk_off = tab->off + 2;
    bv.node_off = tab->off + 4;
    bv.leaf_off = tab->off + 6;
    bv.rhs_off = tab->off + 8;
    bv.links = j->L;
    bv.tchunks = fcg::tree_chunks(M);
    bv.nodes = M;
    bv.leaves = N;
    bv.rhs = k;
    bv.kmax = k;
    bv.kpmin = kp;
    bv.any2d = j->dim == 2;
    bv.anycplx = j->cplx ? 1 : 0;
    // the lane plan of a batch of this one job (make_batch + choose_variants)
    fcg::GatherPlan gp;
    gp.n = 1;
    fcg::GatherSeg& g = gp.seg[0];
    g.row_lo = 0;
    g.row_hi = M;
    g.tile_lo = 0;
    g.tile_hi = fcg::gather_tiles(M);
    g.staged = 0;  // FP64 product
    g.stage_floats = 0;
    g.kpmin = kp;
    g.kmax = k;
    g.cls = fcg::gather_class(kp);
    g.idx16 = j->col16.p ? 1 : 0;
    g.variant = j->gather_variant[fcg::kMvp64];
    HIPCHECK(ctx, fcg::launch_mvp(bv, fcg::kMvp64, 0, gp, s, nullptr));
    return FCG_OK;
}

int fcg_mvp_prec(fcg_ctx* ctx, int job, int mvp_prec, const double* q, double* v, int k) {
    Job* j = get_job(ctx, job);
    if (!j || !q || !v || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (mvp_prec < FCG_MVP_FP64 || mvp_prec > FCG_MVP_FP32_ACC64)
        return fail(ctx, FCG_INVALID_ARGUMENT, "bad product precision");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DevBuf dq, dv;
    const size_t n = (size_t)j->vlen() * k;
    HIPCHECK(ctx, upload(dq, q, n, ctx->stream));
    HIPCHECK(ctx, dv.alloc(sizeof(double) * n));
    RCCHECK(run_mvp(ctx, *j, mvp_prec, dq.as<double>(), dv.as<double>(), k, 0, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(v, dv.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

int fcg_mvp(fcg_ctx* ctx, int job, const double* q, double* v, int k) {
    return fcg_mvp_prec(ctx, job, FCG_MVP_FP64, q, v, k);
}

namespace {
struct Events5 {
    hipEvent_t e[5] = {};
    ~Events5() {
        for (auto x : e)
            if (x) (void)hipEventDestroy(x);
    }
};
}  // namespace

int fcg_mvp_timing_run(fcg_ctx* ctx, int job, int k, int reps, fcg_mvp_timing* out) {
    Job* j = get_job(ctx, job);
    if (!j || !out || k < 1 || reps < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled64 && !j->filled32) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    const int prec = j->filled64 ? fcg::kMvp64 : fcg::kMvp32;
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)j->vlen() * k;
    DevBuf dq, dv;
    HIPCHECK(ctx, dq.alloc(sizeof(double) * n));
    HIPCHECK(ctx, dv.alloc(sizeof(double) * n));
    HIPCHECK(ctx, fcg::launch_random(dq.as<double>(), n, 12345, s));
    Events5 ev;
    for (auto& x : ev.e) HIPCHECK(ctx, hipEventCreate(&x));
    RCCHECK(run_mvp(ctx, *j, prec, dq.as<double>(), dv.as<double>(), k, 0, s));
    double acc[4] = {0, 0, 0, 0}, tot = 0;
    for (int r = 0; r < reps; ++r) {
        RCCHECK(run_mvp(ctx, *j, prec, dq.as<double>(), dv.as<double>(), k, 0, s, ev.e));
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
    out->variant = j->gather_variant[prec];
    return FCG_OK;
}

namespace {
// One timed product of time_batch: seg >= 0 gives class range `seg` the shape
// `var` (the other ranges keep the context's), seg == kAllRanges gives every range
// the shape `var`, seg == kAsPlanned keeps the context's shapes.
struct Trial {
    int seg, var;
};
constexpr int kAsPlanned = -1, kAllRanges = -2;

// Time batched products over `jobs` (each on as many random right-hand sides as it
// has conductor solves) for every trial; `plan_out` receives the class ranges.
int time_batch(fcg_ctx* ctx, const int* jobs, int njobs, int mvp_prec, int reps, const std::vector<Trial>& trials,
               std::vector<fcg_mvp_timing>& outs, fcg::GatherPlan* plan_out = nullptr) {
    if (!ctx || !jobs || njobs < 1 || reps < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (mvp_prec < FCG_MVP_FP64 || mvp_prec > FCG_MVP_FP32_ACC64)
        return fail(ctx, FCG_INVALID_ARGUMENT, "bad product precision");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    std::vector<std::unique_ptr<DevBuf>> qs, vs;
    std::vector<Job*> js;
    int ktot = 0;
    for (int i = 0; i < njobs; ++i) {
        Job* j = get_job(ctx, jobs[i]);
        if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
        if (std::find(js.begin(), js.end(), j) != js.end()) return fail(ctx, FCG_INVALID_ARGUMENT, "job listed twice");
        if (mvp_prec == FCG_MVP_FP64 ? !j->filled64 : !j->filled32)
            return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled in this precision");
        const int k = std::max(1, (int)j->rhs.size());
        RCCHECK(ensure_scratch(ctx, j->scr, j->N, j->M, j->pk(k), mvp_prec == FCG_MVP_FP32));
        RCCHECK(ensure_tile_sources(ctx, *j));
        const size_t n = (size_t)j->vlen() * k;
        qs.emplace_back(new DevBuf());
        vs.emplace_back(new DevBuf());
        HIPCHECK(ctx, qs.back()->alloc(sizeof(double) * n));
        HIPCHECK(ctx, vs.back()->alloc(sizeof(double) * n));
        HIPCHECK(ctx, fcg::launch_random(qs.back()->as<double>(), n, 4242 + (uint64_t)i, s));
        js.push_back(j);
        ktot += k;
    }
    HIPCHECK(ctx, hipStreamSynchronize(s));
    ctx->tabs.reset();
    std::vector<BatchEntry> ents;
    for (size_t i = 0; i < js.size(); ++i)
        ents.push_back(BatchEntry{js[i], &js[i]->scr, qs[i]->as<double>(), vs[i]->as<double>()});
    BatchView bv;
    fcg::GatherPlan base;
    RCCHECK(make_batch(ctx, ctx->tabs, ents, &bv, &base));
    choose_variants(ctx, mvp_prec, ents, base);
    if (plan_out) *plan_out = base;
    Events5 ev;
    for (auto& x : ev.e) HIPCHECK(ctx, hipEventCreate(&x));
    outs.clear();
    for (const Trial& t : trials) {
        fcg::GatherPlan gp = base;
        if (t.seg == kAllRanges) {
            for (int i = 0; i < gp.n; ++i) gp.seg[i].variant = t.var;
        } else if (t.seg >= 0) {
            if (t.seg >= gp.n) return fail(ctx, FCG_INVALID_ARGUMENT, "bad class range");
            gp.seg[t.seg].variant = t.var;
        }
        HIPCHECK(ctx, fcg::launch_mvp(bv, mvp_prec, 0, gp, s));
        double acc[4] = {0, 0, 0, 0}, tot = 0;
        for (int r = 0; r < reps; ++r) {
            HIPCHECK(ctx, fcg::launch_mvp(bv, mvp_prec, 0, gp, s, ev.e));
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
        fcg_mvp_timing o{};
        o.ms_scan = acc[0] / reps;
        o.ms_charge = acc[1] / reps;
        o.ms_gather = acc[2] / reps;
        o.ms_down = acc[3] / reps;
        o.ms_total = tot / reps;
        o.reps = reps;
        o.k = ktot;
        o.variant = gp.n > 0 ? gp.seg[t.seg >= 0 ? t.seg : 0].variant : 0;
        outs.push_back(o);
    }
    HIPCHECK(ctx, hipStreamSynchronize(s));
    return FCG_OK;
}
}  // namespace

int fcg_mvp_timing_batch(fcg_ctx* ctx, const int* jobs, int njobs, int mvp_prec, int reps, fcg_mvp_timing* out) {
    if (!out || !jobs || njobs < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    std::vector<fcg_mvp_timing> outs;
    RCCHECK(time_batch(ctx, jobs, njobs, mvp_prec, reps, {Trial{kAsPlanned, 0}}, outs));
    *out = outs[0];
    return FCG_OK;
}

int fcg_autotune_batch(fcg_ctx* ctx, const int* jobs, int njobs, int mvp_prec, int reps, int* variant,
                       double* best_ms) {
    if (!jobs || njobs < 2) return fail(ctx, FCG_INVALID_ARGUMENT, "a batch needs at least two jobs");
    // the class ranges are separate launches: tune each with the others as they are
    fcg::GatherPlan gp;
    std::vector<fcg_mvp_timing> outs;
    RCCHECK(time_batch(ctx, jobs, njobs, mvp_prec, 1, {Trial{kAsPlanned, 0}}, outs, &gp));
    std::vector<Trial> trials;
    const int nv = fcg::gather_variant_count();
    for (int i = 0; i < gp.n; ++i)
        for (int v = 0; v < nv; ++v) trials.push_back(Trial{i, v});
    RCCHECK(time_batch(ctx, jobs, njobs, mvp_prec, reps, trials, outs));
    int64_t most = -1;
    for (int i = 0; i < gp.n; ++i) {
        int best = 0;
        for (int v = 1; v < nv; ++v)
            if (outs[(size_t)(i * nv + v)].ms_gather < outs[(size_t)(i * nv + best)].ms_gather) best = v;
        ctx->batch_variant[mvp_prec][gp.seg[i].cls] = best;
        const int64_t rows = gp.seg[i].row_hi - gp.seg[i].row_lo;
        if (rows > most) {
            most = rows;
            if (variant) *variant = best;
        }
    }
    RCCHECK(time_batch(ctx, jobs, njobs, mvp_prec, reps, {Trial{kAsPlanned, 0}}, outs));
    if (best_ms) *best_ms = outs[0].ms_gather;
    return FCG_OK;
}

int fcg_set_batch_variant(fcg_ctx* ctx, int mvp_prec, int variant) {
    if (!ctx) return FCG_INVALID_ARGUMENT;
    if (mvp_prec < FCG_MVP_FP64 || mvp_prec > FCG_MVP_FP32_ACC64)
        return fail(ctx, FCG_INVALID_ARGUMENT, "bad product precision");
    if (variant >= fcg::gather_variant_count()) return fail(ctx, FCG_INVALID_ARGUMENT, "bad gather variant");
    for (int& v : ctx->batch_variant[mvp_prec]) v = variant < 0 ? -1 : variant;
    return FCG_OK;
}

int fcg_ctx_profile(fcg_ctx* ctx, int enable, int reset, fcg_profile* out) {
    if (!ctx) return FCG_INVALID_ARGUMENT;
    if (out) *out = ctx->prof;
    if (reset) ctx->prof = fcg_profile{};
    if (enable >= 0) ctx->profiling = enable != 0;
    return FCG_OK;
}

int fcg_set_gather_variant(fcg_ctx* ctx, int job, int variant, int* n_variants) {
    Job* j = get_job(ctx, job);
    if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
    if (n_variants) *n_variants = fcg::gather_variant_count();
    if (variant < 0) return FCG_OK;  // query only
    if (variant >= fcg::gather_variant_count()) return fail(ctx, FCG_INVALID_ARGUMENT, "bad gather variant");
    for (int p = 0; p < 3; ++p) {
        j->gather_variant[p] = variant;
        j->tuned[p] = -1;  // pinned by the caller
    }
    return FCG_OK;
}

namespace {
// measure every gather shape of job j in product precision prec and keep the fastest
int autotune_prec(fcg_ctx* ctx, Job* j, int prec, int k, int reps, int* variant, double* best_ms) {
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)j->vlen() * k;
    DevBuf dq, dv;
    HIPCHECK(ctx, dq.alloc(sizeof(double) * n));
    HIPCHECK(ctx, dv.alloc(sizeof(double) * n));
    HIPCHECK(ctx, fcg::launch_random(dq.as<double>(), n, 777, s));
    Events5 ev;
    for (auto& x : ev.e) HIPCHECK(ctx, hipEventCreate(&x));
    int best = 0;
    double best_t = 1e300;
    const int nv = fcg::gather_variant_count();
    for (int var = 0; var < nv; ++var) {
        j->gather_variant[prec] = var;
        RCCHECK(run_mvp(ctx, *j, prec, dq.as<double>(), dv.as<double>(), k, 0, s));
        double t = 0.0;
        for (int r = 0; r < reps; ++r) {
            RCCHECK(run_mvp(ctx, *j, prec, dq.as<double>(), dv.as<double>(), k, 0, s, ev.e));
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
    j->gather_variant[prec] = best;
    j->tuned[prec] = 1;
    if (variant) *variant = best;
    if (best_ms) *best_ms = best_t;
    return FCG_OK;
}
}  // namespace

int fcg_autotune(fcg_ctx* ctx, int job, int k, int reps, int* variant, double* best_ms) {
    Job* j = get_job(ctx, job);
    if (!j || k < 1 || reps < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled64 && !j->filled32) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    // every filled precision; the values reported are those of the first one tuned
    bool first = true;
    for (int prec = 0; prec < 3; ++prec) {
        if (prec == fcg::kMvp64 ? !j->filled64 : !j->filled32) continue;
        int v = 0;
        double ms = 0.0;
        RCCHECK(autotune_prec(ctx, j, prec, k, reps, &v, &ms));
        if (first) {
            if (variant) *variant = v;
            if (best_ms) *best_ms = ms;
            first = false;
        }
    }
    return FCG_OK;
}

int fcg_precond_apply(fcg_ctx* ctx, int job, const double* q, double* pq, int k) {
    Job* j = get_job(ctx, job);
    if (!j || !q || !pq || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!j->filled64 && !j->filled32) return fail(ctx, FCG_INVALID_ARGUMENT, "job not filled");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DevBuf dq, dp;
    const size_t n = (size_t)j->vlen() * k;
    HIPCHECK(ctx, upload(dq, q, n, ctx->stream));
    HIPCHECK(ctx, dp.alloc(sizeof(double) * n));
    // complex: both halves divided by the real self term (SolveCapacitance.cpp:3508-3524)
    HIPCHECK(ctx, fcg::launch_jacobi(dq.as<double>(), j->self.as<double>(), j->N, j->pk(k), dp.as<double>(),
                                     ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(pq, dp.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

int fcg_block_precond_apply(fcg_ctx* ctx, int job, int block_size, const double* q, double* pq, int k) {
    Job* j = get_job(ctx, job);
    if (!j || !q || !pq || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (j->cplx) return fail(ctx, FCG_INVALID_ARGUMENT, "block preconditioner: real permittivities only");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    RCCHECK(ensure_block_precond(ctx, *j, block_size));
    DevBuf dq, dp;
    const size_t n = (size_t)j->N * k;
    HIPCHECK(ctx, upload(dq, q, n, ctx->stream));
    HIPCHECK(ctx, dp.alloc(sizeof(double) * n));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));  // the table arena is reset below
    ctx->tabs.reset();
    fcg::PView pv;
    RCCHECK(make_pview(ctx, ctx->tabs, {PEnt{j, dq.p, false, dp.as<double>(), k}}, &pv));
    HIPCHECK(ctx, fcg::launch_block_apply(pv, fcg::kPreSet, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(pq, dp.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

int fcg_super_precond_apply(fcg_ctx* ctx, int job, int dim, const double* q, double* pq, int k) {
    Job* j = get_job(ctx, job);
    if (!j || !q || !pq || k < 1) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (j->cplx) return fail(ctx, FCG_INVALID_ARGUMENT, "two-level preconditioner: real permittivities only");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    RCCHECK(ensure_super_precond(ctx, *j, dim));
    if (!j->spre.avail)
        return fail(ctx, FCG_INVALID_ARGUMENT, "two-level preconditioner unavailable: at least as many roots as its dimension");
    DevBuf dq, dp, sw;
    const size_t n = (size_t)j->N * k;
    HIPCHECK(ctx, upload(dq, q, n, ctx->stream));
    HIPCHECK(ctx, dp.alloc(sizeof(double) * n));
    HIPCHECK(ctx, sw.alloc(sizeof(double) * 2 * (size_t)k * j->spre.m));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));  // the table arena is reset below
    ctx->tabs.reset();
    fcg::SView sv;
    RCCHECK(make_sview(ctx, ctx->tabs, {SEnt{j, dq.p, false, dp.as<double>(), k, sw.as<double>()}}, &sv));
    HIPCHECK(ctx, fcg::launch_super_apply(sv, fcg::kPreSet, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(pq, dp.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

int fcg_engine_create(fcg_ctx* ctx, const fcg_solve_opts* opts, fcg_engine** out) {
    if (!ctx || !opts || !out) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    // the reference's m_ucPrecondType bit set (SolveCapacitance.cpp:122-125, 1092-1108)
    // reduced to the one preconditioner this engine runs
    const int pbits = opts->precond;
    if (pbits & ~(FCG_PRECOND_JACOBI | FCG_PRECOND_BLOCK | FCG_PRECOND_SUPER))
        return fail(ctx, FCG_INVALID_ARGUMENT,
                    (pbits & FCG_PRECOND_HIER) ? "hierarchical preconditioner (-ph) is not built"
                                               : "unsupported preconditioner bits");
    int precond = FCG_PRECOND_NONE;
    if (pbits & FCG_PRECOND_SUPER)
        precond = FCG_PRECOND_SUPER;
    else if (pbits & FCG_PRECOND_BLOCK)
        precond = FCG_PRECOND_BLOCK;
    else if (pbits & FCG_PRECOND_JACOBI)
        precond = FCG_PRECOND_JACOBI;
    if (opts->precision != FCG_PREC_FP64 && opts->precision != FCG_PREC_MIXED)
        return fail(ctx, FCG_INVALID_ARGUMENT, "unknown precision mode");
    if (!(opts->tol > 0.0)) return fail(ctx, FCG_INVALID_ARGUMENT, "tolerance must be positive");
    auto e = std::make_unique<fcg_engine>();
    e->ctx = ctx;
    for (Cohort& c : e->co) {
        if (ctx->spare_tables.empty()) {
            c.tabs = new Tables();
        } else {
            c.tabs = ctx->spare_tables.back().release();
            ctx->spare_tables.pop_back();
        }
    }
    e->opts = *opts;
    e->opts.precond = precond;
    *out = e.release();
    return FCG_OK;
}

int fcg_engine_submit(fcg_engine* eng, int job, int64_t tag) {
    if (!eng) return FCG_INVALID_ARGUMENT;
    HIPCHECK(eng->ctx, hipSetDevice(eng->ctx->device));
    return engine_submit(eng, job, tag);
}

int fcg_engine_step(fcg_engine* eng, int* active, int* ready) {
    if (!eng) return FCG_INVALID_ARGUMENT;
    HIPCHECK(eng->ctx, hipSetDevice(eng->ctx->device));
    bool any = false;
    for (const Cohort& c : eng->co) any |= !c.active.empty() || c.inflight;
    if (any) RCCHECK(engine_step(eng));
    if (active) *active = (int)eng->n_active();
    if (ready) *ready = (int)eng->ready.size();
    return FCG_OK;
}

int fcg_engine_poll(fcg_engine* eng, int64_t* tag, int* job, double* cap, int cap_len, int* rows, int* cols,
                    fcg_solve_stats* stats) {
    if (!eng) return FCG_INVALID_ARGUMENT;
    if (eng->ready.empty()) return FCG_INVALID_ARGUMENT;
    Ready& r = eng->ready.front();
    if (cap && cap_len < (int)r.cap.size()) return fail(eng->ctx, FCG_INVALID_ARGUMENT, "result buffer too small");
    if (tag) *tag = r.tag;
    if (job) *job = r.job;
    if (cap) std::memcpy(cap, r.cap.data(), sizeof(double) * r.cap.size());
    if (rows) *rows = r.rows;
    if (cols) *cols = r.cols;
    if (stats) *stats = r.stats;
    eng->ready.pop_front();
    return FCG_OK;
}

int fcg_engine_bytes(fcg_engine* eng, int64_t* bytes) {
    if (!eng || !bytes) return FCG_INVALID_ARGUMENT;
    int64_t b = 0;
    for (const Cohort& c : eng->co)
        for (auto& s : c.active) b += s->bytes();
    *bytes = b;
    return FCG_OK;
}

int fcg_engine_destroy(fcg_engine* eng) {
    if (!eng) return FCG_INVALID_ARGUMENT;
    delete eng;  // waits for the engine's work in flight
    return FCG_OK;
}

namespace {
// run an engine over `jobs` to completion; results in job order
int run_batch(fcg_ctx* ctx, const int* jobs, int njobs, const fcg_solve_opts* opts, bool want_x, double* caps,
              int* rows, int* cols, fcg_solve_stats* stats, double* charges, bool want_hist = false) {
    fcg_engine* e = nullptr;
    RCCHECK(fcg_engine_create(ctx, opts, &e));
    std::unique_ptr<fcg_engine> guard(e);
    e->want_x = want_x;
    e->want_hist = want_hist && opts->residuals && opts->residuals_cap > 0;
    std::vector<size_t> off((size_t)njobs + 1, 0);
    for (int i = 0; i < njobs; ++i) {
        Job* j = get_job(ctx, jobs[i]);
        if (!j) return fail(ctx, FCG_INVALID_ARGUMENT, "bad job id");
        off[(size_t)i + 1] = off[(size_t)i] + j->rhs.size() * (size_t)j->ncol * (j->cplx ? 2 : 1);
    }
    for (int i = 0; i < njobs; ++i) RCCHECK(fcg_engine_submit(e, jobs[i], i));
    int got = 0;
    while (got < njobs) {
        int active = 0, ready = 0;
        RCCHECK(fcg_engine_step(e, &active, &ready));
        while (!e->ready.empty()) {
            Ready& r = e->ready.front();
            const int i = (int)r.tag;
            std::memcpy(caps + off[(size_t)i], r.cap.data(), sizeof(double) * r.cap.size());
            if (rows) rows[i] = r.rows;
            if (cols) cols[i] = r.cols;
            if (stats) stats[i] = r.stats;
            if (charges && !r.x.empty()) std::memcpy(charges, r.x.data(), sizeof(double) * r.x.size());
            if (e->want_hist)  // a single job (fcg_solve)
                for (size_t q = 0; q < r.hist.size(); ++q) {
                    const size_t n = std::min(r.hist[q].size(), (size_t)opts->residuals_cap);
                    std::memcpy(opts->residuals + q * (size_t)opts->residuals_cap, r.hist[q].data(), sizeof(double) * n);
                }
            e->ready.pop_front();
            ++got;
        }
        if (active == 0 && got < njobs) return fail(ctx, FCG_GENERIC_ERROR, "batch ended with missing results");
    }
    return FCG_OK;
}
}  // namespace

int fcg_solve_batch(fcg_ctx* ctx, const int* jobs, int njobs, const fcg_solve_opts* opts, double* caps, int* rows,
                    int* cols, fcg_solve_stats* stats) {
    if (!ctx || !jobs || njobs < 1 || !opts || !caps) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    return run_batch(ctx, jobs, njobs, opts, false, caps, rows, cols, stats, nullptr);
}

int fcg_solve(fcg_ctx* ctx, int job, const fcg_solve_opts* opts, double* cap, int* rows_out, int* cols_out,
              double* charges, fcg_solve_stats* stats) {
    Job* j = get_job(ctx, job);
    if (!j || !opts || !cap) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    // the preconditioner is built ahead of the iterations (the engine finds it
    // built), so that its time is the reference's "precond calculation" phase
    const auto tp0 = clk::now();
    if (!j->cplx && (opts->precond & FCG_PRECOND_SUPER)) {
        RCCHECK(ensure_super_precond(ctx, *j, opts->precond_size));
    } else if (!j->cplx && (opts->precond & FCG_PRECOND_BLOCK)) {
        RCCHECK(ensure_block_precond(ctx, *j, opts->precond_size));
    }
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    const double tp = std::chrono::duration<double>(clk::now() - tp0).count();
    // the gather shapes are the rhs class defaults set at upload; measuring every
    // shape costs about a hundred products, more than a one-off solve, so it is
    // left to fcg_autotune / fcg_autotune_batch
    RCCHECK(run_batch(ctx, &job, 1, opts, charges != nullptr, cap, rows_out, cols_out, stats, charges, true));
    if (stats) stats->t_precond_s = tp;
    return FCG_OK;
}

int fcg_job_memory(fcg_ctx* ctx, int job, fcg_job_mem* out) {
    Job* j = get_job(ctx, job);
    if (!j || !out) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    out->hierarchy = (int64_t)(j->nodes.bytes + j->left.bytes + j->right.bytes + j->parent.bytes + j->leaf_lo.bytes +
                               j->leaf_hi.bytes + j->leaf_node.bytes + j->gout.bytes + j->gin.bytes + j->gout_im.bytes +
                               j->gin_im.bytes + j->perm.bytes + j->iperm.bytes + j->rorder.bytes + j->tb_rec.bytes +
                               j->tb_ppm.bytes + j->tb_lev.bytes + j->tb_lvoff.bytes + j->tt_node.bytes +
                               j->tt_rec.bytes + j->tt_rows.bytes + j->tt_lev.bytes + j->ttop.bytes + j->troots.bytes + j->lpath.bytes +
                               j->self.bytes + j->self_im.bytes + j->diag.bytes + j->src_geom.bytes);
    out->links = (int64_t)(j->row_ptr.bytes + j->col.bytes + j->col16.bytes + j->link_row.bytes + j->coef64.bytes +
                           j->coef32.bytes + j->fill_order.bytes + j->fill_start.bytes + j->fill_rs.bytes +
                           j->grow.bytes + j->tsrc_off.bytes + j->tsrc.bytes + j->lsrc.bytes + j->near_ptr.bytes +
                           j->near_col.bytes + j->near_link.bytes + j->near_dc.bytes);
    out->conductors = (int64_t)(j->leaf_w.bytes + j->leaf_w_im.bytes + j->col_range.bytes + j->rhs_range.bytes);
    out->precond = (int64_t)(j->bpre.bstart.bytes + j->bpre.boff.bytes + j->bpre.P.bytes + j->spre.snode.bytes +
                             j->spre.sstart.bytes + j->spre.sidx.bytes + j->spre.Sinv.bytes + j->spre.w.bytes);
    return FCG_OK;
}

// ---------------------------------------------------------------- link discovery
// CAutoRefine::ComputeLinks and the record pass of AutoRefineLinks
// (Solver/Autorefine.cpp:1132-1171, 481-756) on the device: a work stack of node
// pairs processed in slices of kLinkSlice items (links.hip), then the
// compressed rows, each sorted by source node.
namespace {
constexpr int64_t kLinkSlice = 8ll << 20;

// b holds at least need bytes, keeping its first keep bytes
int ensure_bytes(fcg_ctx* ctx, DevBuf& b, size_t need, size_t keep, hipStream_t s) {
    if (b.bytes >= need) return FCG_OK;
    DevBuf nb;
    HIPCHECK(ctx, nb.alloc(std::max(need, b.bytes + b.bytes / 2)));
    if (keep) HIPCHECK(ctx, hipMemcpyAsync(nb.p, b.p, keep, hipMemcpyDeviceToDevice, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    b.swap(nb);
    return FCG_OK;
}
}  // namespace

int fcg_links_discover(fcg_ctx* ctx, const fcg_job_desc* d, const fcg_link_params* lp, int64_t* n_links,
                       int32_t* max_level) {
    if (!ctx || !d || !lp || !n_links) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    const int M = d->n_nodes;
    if (M < 1 || !d->nodes || !d->left || !d->right || !d->parent || (d->dim != 2 && d->dim != 3) ||
        lp->dim != d->dim || !(lp->link_eps > 0) || !(lp->max_side > 0 || d->dim == 2) ||
        !(lp->max_area > 0 || d->dim == 2) || !(lp->max_length > 0 || d->dim == 3))
        return fail(ctx, FCG_INVALID_ARGUMENT, "bad tree or link parameters");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    ctx->links.reset();
    // post order (children before their parent) bounds the traversal
    for (int i = 0; i < M; ++i) {
        const int32_t l = d->left[i], r = d->right[i];
        if ((l < 0) != (r < 0) || l >= i || r >= i)
            return fail(ctx, FCG_INVALID_ARGUMENT, "bad tree: children must precede their parent");
    }
    DevBuf nodes, left, right;
    HIPCHECK(ctx, upload(nodes, d->nodes, (size_t)M, s));
    HIPCHECK(ctx, upload(left, d->left, (size_t)M, s));
    HIPCHECK(ctx, upload(right, d->right, (size_t)M, s));
    // ComputeLinks: every pair of roots, then every root's self item
    std::vector<int32_t> roots;
    for (int i = 0; i < M; ++i)
        if (d->parent[i] < 0) roots.push_back(i);
    std::vector<uint64_t> init;
    for (size_t i = 0; i < roots.size(); ++i)
        for (size_t j = i + 1; j < roots.size(); ++j)
            init.push_back(((uint64_t)(uint32_t)roots[i] << 32) | (uint32_t)roots[j]);
    for (int32_t r : roots) init.push_back((1ull << 63) | ((uint64_t)(uint32_t)r << 32));
    // the items' recursion levels (both kinds start at 1, as the top calls)
    const std::vector<uint8_t> init_lev(init.size(), 1);
    DevBuf stack, slev, out, olev, links, ctr;
    const size_t cap0 = std::max<size_t>(init.size(), (size_t)kLinkSlice * 4);
    HIPCHECK(ctx, stack.alloc(sizeof(uint64_t) * cap0));
    HIPCHECK(ctx, slev.alloc(cap0));
    HIPCHECK(ctx, hipMemcpyAsync(stack.p, init.data(), sizeof(uint64_t) * init.size(), hipMemcpyHostToDevice, s));
    HIPCHECK(ctx, hipMemcpyAsync(slev.p, init_lev.data(), init_lev.size(), hipMemcpyHostToDevice, s));
    HIPCHECK(ctx, out.alloc(sizeof(uint64_t) * 3 * (size_t)kLinkSlice));
    HIPCHECK(ctx, olev.alloc(3 * (size_t)kLinkSlice));
    HIPCHECK(ctx, links.alloc(sizeof(int2) * std::max<size_t>((size_t)kLinkSlice, (size_t)M * 64)));
    HIPCHECK(ctx, ctr.alloc(3 * sizeof(unsigned long long)));
    unsigned long long* d_out = ctr.as<unsigned long long>();
    unsigned long long* d_link = d_out + 1;
    int* d_lev = reinterpret_cast<int*>(d_out + 2);
    HIPCHECK(ctx, hipMemsetAsync(ctr.p, 0, ctr.bytes, s));
    const fcg::LinkParams P{d->dim, lp->link_eps, lp->max_side, lp->max_area, lp->max_length};
    int64_t top = (int64_t)init.size();
    unsigned long long nlinks = 0;
    while (top > 0) {
        const int64_t n = std::min(top, kLinkSlice);
        RCCHECK(ensure_bytes(ctx, links, sizeof(int2) * (size_t)(nlinks + (unsigned long long)n),
                             sizeof(int2) * (size_t)nlinks, s));
        HIPCHECK(ctx, hipMemsetAsync(d_out, 0, sizeof(unsigned long long), s));
        HIPCHECK(ctx, fcg::launch_links_step(stack.as<uint64_t>() + (top - n), slev.as<uint8_t>() + (top - n), n,
                                             out.as<uint64_t>(), olev.as<uint8_t>(), d_out, links.as<int2>(), d_link,
                                             d_lev, nodes.as<fcg_node_geom>(), left.as<int32_t>(),
                                             right.as<int32_t>(), P, s));
        unsigned long long h[2];
        HIPCHECK(ctx, hipMemcpyAsync(h, ctr.p, sizeof h, hipMemcpyDeviceToHost, s));
        HIPCHECK(ctx, hipStreamSynchronize(s));
        if (h[0] > 3ull * (unsigned long long)n || h[1] < nlinks || h[1] - nlinks > (unsigned long long)n)
            return fail(ctx, FCG_DEVICE_ERROR, "link discovery: inconsistent step counts");
        nlinks = h[1];
        top -= n;
        RCCHECK(ensure_bytes(ctx, stack, sizeof(uint64_t) * (size_t)(top + (int64_t)h[0]),
                             sizeof(uint64_t) * (size_t)top, s));
        RCCHECK(ensure_bytes(ctx, slev, (size_t)(top + (int64_t)h[0]), (size_t)top, s));
        if (h[0]) {
            HIPCHECK(ctx, hipMemcpyAsync(stack.as<uint64_t>() + top, out.p, sizeof(uint64_t) * h[0],
                                         hipMemcpyDeviceToDevice, s));
            HIPCHECK(ctx, hipMemcpyAsync(slev.as<uint8_t>() + top, olev.p, h[0], hipMemcpyDeviceToDevice, s));
        }
        top += (int64_t)h[0];
    }
    int maxlev = 0;
    HIPCHECK(ctx, hipMemcpyAsync(&maxlev, d_lev, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    stack.release();
    slev.release();
    out.release();
    olev.release();
    // compressed rows: degrees, offsets (host scan), scatter, per-row sort
    const int64_t Pn = (int64_t)nlinks, L = 2 * Pn;
    DevBuf deg, cursor, col;
    HIPCHECK(ctx, deg.alloc(sizeof(int32_t) * (size_t)M));
    HIPCHECK(ctx, hipMemsetAsync(deg.p, 0, deg.bytes, s));
    HIPCHECK(ctx, fcg::launch_links_degree(links.as<int2>(), Pn, deg.as<int32_t>(), s));
    std::vector<int32_t> hdeg((size_t)M);
    HIPCHECK(ctx, hipMemcpyAsync(hdeg.data(), deg.p, sizeof(int32_t) * (size_t)M, hipMemcpyDeviceToHost, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    std::vector<int64_t> row_ptr((size_t)M + 1, 0);
    for (int i = 0; i < M; ++i) row_ptr[(size_t)i + 1] = row_ptr[(size_t)i] + hdeg[(size_t)i];
    if (row_ptr[(size_t)M] != L) return fail(ctx, FCG_DEVICE_ERROR, "link discovery: degree sum mismatch");
    DevBuf drow;
    HIPCHECK(ctx, upload(drow, row_ptr.data(), row_ptr.size(), s));
    HIPCHECK(ctx, cursor.alloc(sizeof(unsigned long long) * (size_t)M));
    HIPCHECK(ctx, hipMemcpyAsync(cursor.p, drow.p, sizeof(int64_t) * (size_t)M, hipMemcpyDeviceToDevice, s));
    HIPCHECK(ctx, col.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(L, 1)));
    HIPCHECK(ctx, fcg::launch_links_scatter(links.as<int2>(), Pn, cursor.as<unsigned long long>(),
                                            col.as<int32_t>(), s));
    RCCHECK(sort_rows(ctx, drow, row_ptr, col.as<int32_t>(), M));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    ctx->links.row_ptr = std::move(row_ptr);
    ctx->links.col.swap(col);
    ctx->links.L = L;
    ctx->links.ready = true;
    *n_links = L;
    if (max_level) *max_level = maxlev;
    return FCG_OK;
}

int fcg_links_fetch(fcg_ctx* ctx, int64_t* row_ptr, int32_t* col) {
    if (!ctx || !row_ptr || (!col && ctx->links.L > 0)) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!ctx->links.ready) return fail(ctx, FCG_INVALID_ARGUMENT, "no discovered links to fetch");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    std::memcpy(row_ptr, ctx->links.row_ptr.data(), sizeof(int64_t) * ctx->links.row_ptr.size());
    if (ctx->links.L > 0) {
        HIPCHECK(ctx, hipMemcpyAsync(col, ctx->links.col.p, sizeof(int32_t) * (size_t)ctx->links.L,
                                     hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    ctx->links.reset();
    return FCG_OK;
}

// ---------------------------------------------------------------- adaptive refinement
// AutoRefinePanels' discretization (Solver/Autorefine.cpp:293-479, 2638-2921) on
// the device: the pairs of group tops, then the groups' self passes round by
// round (depth by depth), each pass's pair work drained from a work stack in
// slices of evaluate / split / expand.
int fcg_refine(fcg_ctx* ctx, const fcg_refine_node* nodes, int n_nodes, const int32_t* tops,
               const fcg_refine_params* p, int* n_out, double* max_mesh_eps, int32_t* max_level) {
    if (!ctx || !nodes || n_nodes < 1 || !tops || !p || !n_out) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if ((p->dim != 2 && p->dim != 3) || p->n_groups < 1 || !(p->mesh_eps > 0) || !(p->curv >= 1.0))
        return fail(ctx, FCG_INVALID_ARGUMENT, "bad refinement parameters");
    // pre-order (children after their parent) bounds every walk
    for (int i = 0; i < n_nodes; ++i) {
        const int l = nodes[i].left, r = nodes[i].right;
        if ((l < 0) != (r < 0) || (l >= 0 && (l <= i || r <= i || l >= n_nodes || r >= n_nodes)))
            return fail(ctx, FCG_INVALID_ARGUMENT, "bad tree: children must follow their parent");
    }
    for (int g = 0; g < p->n_groups; ++g)
        if (tops[g] < 0 || tops[g] >= n_nodes) return fail(ctx, FCG_INVALID_ARGUMENT, "bad group top");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    ctx->refined.pool.release();
    ctx->refined.ready = false;
    const fcg::RefineParams P{p->dim, p->mesh_eps, p->curv, p->max_side, p->max_area, p->max_length};
    const int64_t S = kLinkSlice;
    // node pool and split flags (grown together; new flags start at 0)
    DevBuf pool, want;
    int64_t cap = std::max<int64_t>(2 * (int64_t)n_nodes, 1 << 20);
    HIPCHECK(ctx, pool.alloc(sizeof(fcg_refine_node) * (size_t)cap));
    HIPCHECK(ctx, want.alloc(sizeof(int32_t) * (size_t)cap));
    HIPCHECK(ctx, hipMemcpyAsync(pool.p, nodes, sizeof(fcg_refine_node) * (size_t)n_nodes, hipMemcpyHostToDevice, s));
    HIPCHECK(ctx, hipMemsetAsync(want.p, 0, want.bytes, s));
    // counters: pool size, pending, stack output, next-round size, max eps bits, max level
    DevBuf ctr;
    HIPCHECK(ctx, ctr.alloc(6 * sizeof(unsigned long long)));
    HIPCHECK(ctx, hipMemsetAsync(ctr.p, 0, ctr.bytes, s));
    unsigned long long* c_pool = ctr.as<unsigned long long>();
    unsigned long long* c_pend = c_pool + 1;
    unsigned long long* c_out = c_pool + 2;
    unsigned long long* c_next = c_pool + 3;
    unsigned long long* c_eps = c_pool + 4;
    int* c_lev = reinterpret_cast<int*>(c_pool + 5);
    const unsigned long long n0 = (unsigned long long)n_nodes;
    HIPCHECK(ctx, hipMemcpyAsync(c_pool, &n0, sizeof n0, hipMemcpyHostToDevice, s));
    int64_t pool_n = n_nodes;
    auto grow_pool = [&](int64_t need) -> int {
        if (need <= cap) return FCG_OK;
        const int64_t nc = std::max(need, cap + cap / 2);
        DevBuf np, nw;
        HIPCHECK(ctx, np.alloc(sizeof(fcg_refine_node) * (size_t)nc));
        HIPCHECK(ctx, nw.alloc(sizeof(int32_t) * (size_t)nc));
        HIPCHECK(ctx, hipMemcpyAsync(np.p, pool.p, sizeof(fcg_refine_node) * (size_t)pool_n, hipMemcpyDeviceToDevice, s));
        HIPCHECK(ctx, hipMemsetAsync(nw.p, 0, nw.bytes, s));
        HIPCHECK(ctx, hipMemcpyAsync(nw.p, want.p, sizeof(int32_t) * (size_t)pool_n, hipMemcpyDeviceToDevice, s));
        HIPCHECK(ctx, hipStreamSynchronize(s));
        pool.swap(np);
        want.swap(nw);
        cap = nc;
        return FCG_OK;
    };
    DevBuf stack, slev, pend, plev, out, olev;
    HIPCHECK(ctx, stack.alloc(sizeof(uint64_t) * (size_t)(4 * S)));
    HIPCHECK(ctx, slev.alloc((size_t)(4 * S)));
    HIPCHECK(ctx, pend.alloc(sizeof(uint64_t) * (size_t)S));
    HIPCHECK(ctx, plev.alloc((size_t)S));
    HIPCHECK(ctx, out.alloc(sizeof(uint64_t) * (size_t)(2 * S)));
    HIPCHECK(ctx, olev.alloc((size_t)(2 * S)));
    int64_t top = 0;
    // push n items from out / olev onto the stack
    auto push_out = [&](int64_t n) -> int {
        if (n <= 0) return FCG_OK;
        RCCHECK(ensure_bytes(ctx, stack, sizeof(uint64_t) * (size_t)(top + n), sizeof(uint64_t) * (size_t)top, s));
        RCCHECK(ensure_bytes(ctx, slev, (size_t)(top + n), (size_t)top, s));
        HIPCHECK(ctx, hipMemcpyAsync(stack.as<uint64_t>() + top, out.p, sizeof(uint64_t) * (size_t)n,
                                     hipMemcpyDeviceToDevice, s));
        HIPCHECK(ctx, hipMemcpyAsync(slev.as<uint8_t>() + top, olev.p, (size_t)n, hipMemcpyDeviceToDevice, s));
        top += n;
        return FCG_OK;
    };
    // run the stack's pair work to completion
    auto drain = [&]() -> int {
        while (top > 0) {
            const int64_t n = std::min(top, S);
            top -= n;
            HIPCHECK(ctx, hipMemsetAsync(c_pend, 0, sizeof(unsigned long long), s));
            HIPCHECK(ctx, fcg::launch_refine_eval(stack.as<uint64_t>() + top, slev.as<uint8_t>() + top, n,
                                                  pool.as<fcg_refine_node>(), P, want.as<int32_t>(),
                                                  pend.as<uint64_t>(), plev.as<uint8_t>(), c_pend, c_eps, c_lev, s));
            unsigned long long h[2];
            HIPCHECK(ctx, hipMemcpyAsync(h, c_pool, sizeof h, hipMemcpyDeviceToHost, s));
            HIPCHECK(ctx, hipStreamSynchronize(s));
            const int64_t np = (int64_t)h[1];
            pool_n = (int64_t)h[0];
            if (np > n || pool_n < n_nodes) return fail(ctx, FCG_DEVICE_ERROR, "refinement: inconsistent step counts");
            if (np == 0) continue;
            if (pool_n + 2 * np > (int64_t)INT32_MAX) return fail(ctx, FCG_OUT_OF_MEMORY, "refinement: too many nodes");
            RCCHECK(grow_pool(pool_n + 2 * np));
            HIPCHECK(ctx, fcg::launch_refine_split(pend.as<uint64_t>(), np, pool.as<fcg_refine_node>(),
                                                   want.as<int32_t>(), c_pool, P.dim, s));
            HIPCHECK(ctx, hipMemsetAsync(c_out, 0, sizeof(unsigned long long), s));
            HIPCHECK(ctx, fcg::launch_refine_expand(pend.as<uint64_t>(), plev.as<uint8_t>(), np,
                                                    pool.as<fcg_refine_node>(), out.as<uint64_t>(),
                                                    olev.as<uint8_t>(), c_out, s));
            RCCHECK(push_out(2 * np));  // every pending item gives two pairs
        }
        return FCG_OK;
    };
    // the pairs of group tops (DiscretizeMutual from AutoRefinePanels), level 1
    {
        std::vector<uint64_t> init;
        for (int i = 0; i < p->n_groups; ++i)
            for (int j = i + 1; j < p->n_groups; ++j)
                init.push_back(((uint64_t)(uint32_t)tops[i] << 32) | (uint32_t)tops[j]);
        const std::vector<uint8_t> lev(init.size(), 1);
        RCCHECK(ensure_bytes(ctx, stack, sizeof(uint64_t) * init.size(), 0, s));
        RCCHECK(ensure_bytes(ctx, slev, init.size(), 0, s));
        if (!init.empty()) {
            HIPCHECK(ctx, hipMemcpyAsync(stack.p, init.data(), sizeof(uint64_t) * init.size(), hipMemcpyHostToDevice, s));
            HIPCHECK(ctx, hipMemcpyAsync(slev.p, lev.data(), lev.size(), hipMemcpyHostToDevice, s));
        }
        top = (int64_t)init.size();
        RCCHECK(drain());
    }
    // DiscretizeSelf, round by round: the tops (split when a leaf) at level 1,
    // then every round's children after the round's pair work has finished
    int maxlev = 0;
    {
        DevBuf front, next;
        std::vector<int32_t> t0(tops, tops + p->n_groups);
        std::sort(t0.begin(), t0.end());
        t0.erase(std::unique(t0.begin(), t0.end()), t0.end());
        HIPCHECK(ctx, upload(front, t0.data(), t0.size(), s));
        int64_t nf = (int64_t)t0.size();
        unsigned long long pc = 0;
        HIPCHECK(ctx, hipMemcpyAsync(&pc, c_pool, sizeof pc, hipMemcpyDeviceToHost, s));
        HIPCHECK(ctx, hipStreamSynchronize(s));
        pool_n = (int64_t)pc;
        RCCHECK(grow_pool(pool_n + 2 * nf));
        HIPCHECK(ctx, fcg::launch_refine_split_tops(front.as<int32_t>(), (int)nf, pool.as<fcg_refine_node>(), c_pool,
                                                    P.dim, s));
        for (int lev = 1; nf > 0; ++lev) {
            maxlev = std::max(maxlev, std::min(lev, 255));
            RCCHECK(ensure_bytes(ctx, next, sizeof(int32_t) * (size_t)(2 * nf), 0, s));
            RCCHECK(ensure_bytes(ctx, out, sizeof(uint64_t) * (size_t)nf, 0, s));
            RCCHECK(ensure_bytes(ctx, olev, (size_t)nf, 0, s));
            HIPCHECK(ctx, hipMemsetAsync(c_out, 0, 2 * sizeof(unsigned long long), s));  // c_out, c_next
            HIPCHECK(ctx, fcg::launch_refine_self(front.as<int32_t>(), nf, lev, pool.as<fcg_refine_node>(),
                                                  out.as<uint64_t>(), olev.as<uint8_t>(), c_out, next.as<int32_t>(),
                                                  c_next, s));
            unsigned long long h[2];
            HIPCHECK(ctx, hipMemcpyAsync(h, c_out, sizeof h, hipMemcpyDeviceToHost, s));
            HIPCHECK(ctx, hipStreamSynchronize(s));
            if ((int64_t)h[0] > nf || h[1] != 2 * h[0])
                return fail(ctx, FCG_DEVICE_ERROR, "refinement: inconsistent self round");
            RCCHECK(push_out((int64_t)h[0]));
            RCCHECK(drain());
            front.swap(next);
            nf = (int64_t)h[1];
        }
    }
    unsigned long long h[6];
    HIPCHECK(ctx, hipMemcpyAsync(h, ctr.p, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHECK(ctx, hipStreamSynchronize(s));
    double eps = 0.0;
    std::memcpy(&eps, &h[4], sizeof eps);
    int dlev = 0;
    std::memcpy(&dlev, &h[5], sizeof dlev);
    ctx->refined.pool.swap(pool);
    ctx->refined.n = (int)h[0];
    ctx->refined.ready = true;
    *n_out = (int)h[0];
    if (max_mesh_eps) *max_mesh_eps = eps;
    if (max_level) *max_level = std::max(maxlev, dlev);
    return FCG_OK;
}

int fcg_refine_fetch(fcg_ctx* ctx, fcg_refine_node* out) {
    if (!ctx || !out) return fail(ctx, FCG_INVALID_ARGUMENT, "bad argument");
    if (!ctx->refined.ready) return fail(ctx, FCG_INVALID_ARGUMENT, "no refined tree to fetch");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    HIPCHECK(ctx, hipMemcpyAsync(out, ctx->refined.pool.p, sizeof(fcg_refine_node) * (size_t)ctx->refined.n,
                                 hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->refined.pool.release();
    ctx->refined.ready = false;
    ctx->refined.n = 0;
    return FCG_OK;
}

}  // extern "C"
