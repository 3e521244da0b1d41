// gfx950 kernels of the FasterCap solve path, batched over jobs.
//
// Every kernel works on a batch of jobs described by a device table of BJob
// records and per-kernel work offsets (cumulative work units per job, searched
// with a short binary search).  A batch of one job is the single-solve path.
// Host-side launch wrappers are declared here and implemented in kernels.hip;
// the C-ABI (fcg_api.hip) owns the memory and the tables.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/fastercap_amd.h"

namespace fcg {

// Double-double value used by the prefix scan of the up-pass.
// Tree passes of the product (up: node charges, down: leaf potentials) work on
// chunks of kTreeChunk consecutive post-order nodes.  A node whose subtree lies in
// its chunk is inner: its chunk's block computes it level by level in LDS.  The
// others (the nodes whose subtrees straddle a chunk boundary: the top of the tree,
// closed under parents) are done by one block per job.
constexpr int kTreeChunk = 1024;
constexpr int kTreeRhs = 4;      // right-hand sides per LDS pass of the tree kernels
constexpr int kTreeTopLds = 1024;  // top nodes a job's top pass holds in LDS (more: global memory)
inline int64_t tree_chunks(int M) { return ((int64_t)M + kTreeChunk - 1) / kTreeChunk; }
constexpr int kChunk = 1024;     // Krylov vector chunk per block
constexpr int kNumXcd = 8;       // MI355X accelerator complex dies

// Precision of one hierarchical product.
enum MvpPrec : int {
    kMvp64 = 0,     // FP64 coefficients, FP64 charges and accumulation
    kMvp32 = 1,     // FP32 coefficients, charges and gather sums (inner solves)
    kMvp32Acc = 2,  // FP32 coefficients, FP64 charges and accumulation (refinement residual)
};

// One destination row of the gather, in gather row order: its links
// [beg, end) in the stored link arrays, its charge / potential row and its node.
struct GRow {
    uint32_t beg, end;
    int32_t prow, node;
};

// Vertices and centroid of a node as a link source (48 bytes, three 16-byte
// loads), stored by charge row: the fill reads a link's source through its
// stored column directly.
struct alignas(16) SrcGeom {
    float v[9];
    float c[3];
};

// Device record of one job in a batch.
struct BJob {
    int N, M, k, dim;
    int scheme;           // FCG_SCHEME_GALERKIN for a 3D Galerkin job, else collocation
    int kp;               // node-major row stride: k rounded up to a power of two
    int first_cond_leaf;  // 2D zero-total-charge fix-up start
    int64_t L;
    const fcg_node_geom* nodes;
    const int64_t* row_ptr;
    const int32_t* col;       // [L] source slot (permuted node index), ascending within a row
    const uint16_t* col16;    // the same in 16 bits when M <= 65536, else null
    const int32_t* perm;      // [M] node -> slot of its charge / potential row
    const int32_t* iperm;     // [M] slot -> node
    const int32_t* rorder;    // [M] gather row order -> node
    // tree passes: the inner nodes of every chunk grouped by depth (ascending),
    // one record per node, tb_rec[tb_lev[tb_lvoff[t] + l] ...] the nodes of level
    // l of chunk t (the chunk's last tb_lev entry ends its list); a record holds
    // x = slot | left child slot << 16, y = right child slot | parent slot << 16
    // (slots are the node's offset in its chunk, -1 for none or a top parent),
    // z = charge row (perm), w = leaf index (-1 for internal nodes); tb_ppm: the
    // charge row of a top parent (-1 otherwise); the top nodes by depth, level l
    // at tt_node[tt_lev[l] ..]; ttop[n] = 1 for a top node; the group roots;
    // lpath: the charge rows on the path of the last leaf's node to its root,
    // root first, its first lpath_top entries top nodes (2D fix-up)
    const int4* tb_rec;
    const int32_t* tb_ppm;
    const int32_t* lpath;
    int lpath_len, lpath_top;
    const int32_t* tb_lev;
    const int32_t* tb_lvoff;   // [chunks + 1]
    const int32_t* tt_node;
    const int4* tt_rec;        // per top node: (node, slot of its parent, of its left child, of its right
                               // child), -1 where that node is not a top node (slots index tt_node)
    const int4* tt_rows;       // per top node: the charge rows (perm) of itself, its left child, its
                               // right child and its parent, -1 for none
    const int32_t* tt_lev;     // [tt_nlev + 1]
    const uint8_t* ttop;       // [M]
    const int32_t* troots;     // [nroots]
    int tt_nlev, ntop, nroots;
    const GRow* grow;         // [M] the rows in gather row order
    // staged gather (null until built): tsrc[tsrc_off[t] ..] lists the source
    // charge rows of tile t ascending; lsrc[l] is link l's index in its tile's list
    const int32_t* tsrc_off;  // [ceil(M / kGatherTileRows) + 1]
    const uint16_t* tsrc;
    const uint16_t* lsrc;     // [L + 4]
    const SrcGeom* src_geom;  // [M] source geometry by charge row
    const uint32_t* fill_order;  // [L] links grouped by fill class
    const int2* fill_rs;         // [L] (destination node, source charge row) of each fill_order entry
    const int64_t* fill_start;   // [kFillClasses + 1] first fill_order entry of each class
    const int32_t* link_row;  // [L] destination node of each link (fill only)
    double* coef64;           // [L] or null
    float* coef32;            // [L] or null
    // FP64 near-field side list (mixed precision): the links of the analytic
    // PotEstimateOpt branch, row by row in stored link order; near_dc is the
    // difference between the FP64 coefficient and its FP32 copy in coef32
    const int64_t* near_ptr;    // [M+1] or null
    const int32_t* near_col;    // [Ln] source charge row
    const uint32_t* near_link;  // [Ln] stored link index
    double* near_dc;            // [Ln]
    double* diag;             // [M] self coefficient of leaves, 0 elsewhere
    double* self;             // [N]
    // complex permittivities: the product's 2k vectors are k pairs [Re | Im]
    // of [k][2N] vectors, coupled by the imaginary self terms self_im [N]
    // (null for real jobs); gout_im / gin_im [groups] are the imaginary parts
    double* self_im;
    const double* gout_im;
    const double* gin_im;
    const int32_t* left;
    const int32_t* right;
    const int32_t* leaf_lo;
    const int32_t* leaf_hi;
    const int32_t* parent;
    const int32_t* leaf_node;
    const double* gout;
    const double* gin;
    // product scratch sized for k right-hand sides
    double* charge64;
    float* charge32;  // [M][kp], row perm[node]
    double* pot64;
    float* pot32;  // [M][kp], row perm[node]
    double* lt;    // [2k] 2D fix-up values (k_down_top)
    // current product operands ([k][N] each): the input is q, or q32 when it is
    // an FP32 basis vector (then q is null)
    const double* q;
    const float* q32;
    double* v;
};

// Work offsets of a batch: cumulative work units per job (length njobs + 1).
struct BatchView {
    const BJob* jobs;
    int njobs;
    const int64_t* link_off;  // links
    const int64_t* tchunk_off;  // tree chunks (kTreeChunk post-order nodes)
    const int64_t* node_off;  // nodes
    const int64_t* leaf_off;  // leaves
    const int64_t* rhs_off;   // right-hand sides
    const int64_t* gtile_off;  // staged-gather tiles (kGatherTileRows rows of one job)
    int64_t links, tchunks, nodes, leaves, rhs, gtiles;
    int kmax;   // largest k in the batch
    int kpmin;  // smallest row stride in the batch
    int any2d;
    int anycplx;  // some job has complex permittivities
};

// Krylov bookkeeping of one job for the batched GMRES kernels.
struct KJob {
    void* const* Q;         // device array of basis vector addresses ([k][N] each, FP32 when basis32)
    double* w;              // work vector [k][N]
    double* x;              // current solution [k][N]
    double* r;              // residual [k][N]
    const int32_t* brange;  // [2k] right-hand side r is 1 on leaves [brange[2r], brange[2r+1]), 0 elsewhere
    const double* self;  // self coefficients [nleaf] (Jacobi)
    int N, k, nv;        // N: vector length (2 nleaf for complex jobs); nv = basis vectors to project
    int nleaf;           // leaves (the self term of element n is self[n mod nleaf])
    int hoff;            // offset of this job's block in the step buffer
    int zoff, nz;        // solution-update coefficients z[zoff + jj*k + r], jj < nz
    int jacobi;
    double* y;           // block preconditioner: the update writes Q z here instead of adding to x
    int coff, next;      // DCGS2 update: coefficient offset, and whether the next vector is written
    // DCGS2 Arnoldi state on the device, per rhs at hd + r * hstride (dcgs_rec):
    // the Hessenberg columns, their rotations and the rotated right-hand side
    double* hd;
    int64_t hstride;
    int ooff;                // step outputs for the host: [k][2] = status, |g_(nv+1)|
    const int32_t* flags;    // [k] 1: the rhs met the tolerance before this step
    const int32_t* last;     // [k] last Arnoldi column of each rhs (least squares), -1: none
    const double* beta;      // [k] |r| of the cycle's start (first step)
};

// Per (job, rhs) DCGS2 record, laid out independently of its capacity so that a
// larger one starts with a copy of the smaller: [0] the last column's provisional
// subdiagonal, [1] g at that column before its rotation, [2] g_0, then column c
// at dcgs_rec(c): the unrotated column (rows 0..c+1), the rotated one (rows
// 0..c), its rotation (cos, sin) and g_(c+1).
__host__ __device__ inline int64_t dcgs_rec(int c) { return 3 + (int64_t)c * (c + 5); }

// Segment (job, rhs) and chunk tables for the Krylov kernels.
struct Seg {
    int job, r, c_lo, c_hi;  // chunk range [c_lo, c_hi)
};
struct Chunk {
    int seg, n0, n1;
};
struct KrylovView {
    const KJob* kjobs;
    const Seg* segs;
    const Chunk* chunks;
    int nsegs, nchunks, maxnv;
    int maxnz;  // the most solution-update columns of a job (k_dcgs_lsq)
    int basis32;  // the basis vectors are FP32 (inner solves of the mixed path)
};

// Fill classes: which integral branch of PotEstimateOpt a link takes (brackets of
// centroid distance over the larger max side), on a dielectric row or not, 3D or
// 2D.  3D: 5 diel + branch (0 analytic, 1 rule 5, 2 rule 3, 3 rule 2, 4 point);
// 2D: 10 + 2 diel + (0 analytic, 1 point).  Each class is filled by a kernel of
// its own.
constexpr int kFillClasses = 14;

// fill of every link and self term of the batch; mixed stores FP32 coefficients.
// off[c] (device, njobs + 1 entries) is the prefix over the batch's jobs of their
// class-c link counts, total[c] its last entry.
hipError_t launch_link_rows(const BatchView& b, hipStream_t s);
// In mixed precision the analytic classes (near_class) are filled from the jobs'
// near-field side lists instead: near_off (device, njobs + 1) is the prefix of
// their near link counts, near_total its last entry.
// galerkin: every job of the batch takes the Galerkin scheme's 3D integrals
hipError_t launch_fill(const BatchView& b, const int64_t* const* off, const int64_t* total, int precision,
                       const int64_t* near_off, int64_t near_total, bool galerkin, hipStream_t s);
inline bool near_class(int c) { return c == 0 || c == 5 || c == 10 || c == 12; }
// Near-field side list of one job: flag[l] = 1 for the fill_order entries
// [p0, p0 + n); per-row counts of flagged links; the rows' flagged links
// (charge-row column, link index) scattered at near_ptr.
hipError_t launch_near_flag(const uint32_t* fill_order, int64_t p0, int64_t n, uint8_t* flag, hipStream_t s);
// rs[g] = (link_row[l], col[l]) for l = fill_order[g]: the fill reads its links'
// rows and sources in fill order, coalesced
hipError_t launch_fill_rs(const uint32_t* fill_order, const int32_t* link_row, const int32_t* col, int64_t L,
                          int2* rs, hipStream_t s);
hipError_t launch_near_count(const int64_t* row_ptr, const uint8_t* flag, int M, int32_t* cnt, hipStream_t s);
hipError_t launch_near_scatter(const int64_t* row_ptr, const uint8_t* flag, const int32_t* col, int M,
                               const int64_t* near_ptr, int32_t* near_col, uint32_t* near_link, hipStream_t s);

// Rhs classes of the gather: jobs whose node-major row stride is kp = 2^c share
// class c (strides beyond the last class share it).  A batch lists its jobs
// ordered by stride and then by source index width, so each (class, width) owns
// one contiguous range of batch rows and the gather runs one launch per range,
// each with its own lane shape and index width.
constexpr int kGatherClasses = 12;
inline int gather_class(int kp) {
    int c = 0;
    while ((1 << c) < kp && c < kGatherClasses - 1) ++c;
    return c;
}
struct GatherSeg {
    int64_t row_lo, row_hi;    // batch rows [row_lo, row_hi)
    int64_t tile_lo, tile_hi;  // its staged-gather tiles [tile_lo, tile_hi)
    int kpmin, kmax;           // smallest row stride and largest rhs count in the range
    int cls, variant;          // rhs class and gather lane shape
    int idx16;                 // every job of the range has 16-bit source indices
    int staged;                // every job of the range has its tile source lists (staged gather)
    int stage_floats;          // the range's largest tile source list times kp (LDS elements)
};

// Staged gather (products of jobs with 16-bit source indices).  A job's
// gather rows are cut into tiles of kGatherTileRows consecutive rows (same-depth
// neighbours).  The block of a tile first copies the charges of the tile's
// distinct sources into LDS and then reads every link's charge from there, by
// the link's 16-bit index in that list; the loop's memory loads are then the
// coefficient and index streams alone.  The lists are built once per job
// (launch_tile_sources_count / _fill); a tile's list times the row stride in
// the charge type must fit kStageMaxBytes of LDS, or the range gathers from
// global memory.
constexpr int kGatherTileRows = 128;
constexpr size_t kStageMaxBytes = 65536;
inline int64_t gather_tiles(int M) { return ((int64_t)M + kGatherTileRows - 1) / kGatherTileRows; }
// cnt[t] = distinct sources of tile t (jobs of at most 65536 nodes)
hipError_t launch_tile_sources_count(const GRow* grow, const uint16_t* col16, int M, int32_t* cnt, hipStream_t s);
// the lists at tsrc_off and every link's index in its tile's list
hipError_t launch_tile_sources_fill(const GRow* grow, const uint16_t* col16, int M, const int32_t* tsrc_off,
                                    uint16_t* tsrc, uint16_t* lsrc, hipStream_t s);
struct GatherPlan {
    int n = 0;
    GatherSeg seg[2 * kGatherClasses];
};

// v = P q on every job of the batch (operands in BJob::q / BJob::v).  jacobi
// divides the input by the self terms first.  The plan's segments cover every
// batch row once and pick each range's gather lane layout; `ev` (5 events or
// null) times the stages.
hipError_t launch_mvp(const BatchView& b, int prec, int jacobi, const GatherPlan& plan, hipStream_t s,
                      hipEvent_t* ev = nullptr);
int gather_variant_count();
// whether a variant stages the tiles' source charges in LDS (FP32 products)
bool gather_variant_staged(int variant);
// gather lane shape used for rhs class `cls` in product precision `prec` unless
// a job or context pins another one (measured best shapes of short-row batches)
int default_gather_variant(int prec, int cls);
// power-of-two row stride of the node-major arrays for k right-hand sides
inline int padded_k(int k) {
    int p = 1;
    while (p < k) p <<= 1;
    return p;
}

// Krylov helpers (batched over segments).
// hbuf[hoff + 2 nv k + r] = the sum of the segment's chunk partials (|r|^2 of k_resid)
hipError_t launch_reduce_norm(const KrylovView& kv, const double* partial, double* hbuf, hipStream_t s);
// Q_nv[seg] = w[seg] / d[seg]   (d per segment, 0 -> zero vector)
hipError_t launch_scale_new(const KrylovView& kv, const double* d, hipStream_t s);
// Arnoldi with delayed reorthogonalisation (DCGS2).  Step buffer layout per job
// at hoff: [2 nv + 3][k] = Q^T v [nv], Q^T w [nv], v.v, v.w, w.w (v: slot nv,
// w: the product); partial holds nchunks * 4 * (2 maxnv + 3) doubles.
hipError_t launch_dcgs_dots(const KrylovView& kv, double* partial, hipStream_t s);
// the final q in slot nv and, for jobs with `next`, the next provisional vector
// in slot nv + 1; coefficients per (job, rhs) at coef[coff + r (2 nv + 3)]
hipError_t launch_dcgs_update(const KrylovView& kv, const double* coef, hipStream_t s);
// per (job, rhs): the projection summed from the partials of launch_dcgs_dots
// into hbuf (at hoff), then the step's Hessenberg columns, rotations and update
// coefficients, and the host's outputs at hbuf[ooff + 2 r] (status: 0 iterated,
// 1 met the tolerance before, 2 the provisional vector lies in the basis; then
// |g_(nv+1)|)
hipError_t launch_dcgs_coef(const KrylovView& kv, const double* partial, double* hbuf, double* coef, hipStream_t s);
// per (job, rhs): the least-squares solution over columns 0..last of the rotated
// system, z[zoff + jj k + r] for jj < nz (0 beyond last)
hipError_t launch_dcgs_lsq(const KrylovView& kv, double* z, hipStream_t s);
// Block preconditioner (ComputePrecondVectFast block branch,
// Solver/SolveCapacitance.cpp:3527-3544): blocks of consecutive leaves
// [bstart[b], bstart[b+1]) of at most 128 leaves, each block's inverse stored
// column-major at P + boff[b].
constexpr int kMaxBlockPrecond = 128;
// Build and invert every block of one job (the batch view holds that job only):
// the block's collocation matrix (self terms on the diagonal, link coefficients
// elsewhere), LU with partial pivoting, inverse.  work: scratch of boff[nb]
// doubles, used by blocks of more than 64 leaves.
hipError_t launch_block_setup(const BatchView& one_job, int nb, const int32_t* bstart, const int64_t* boff, double* P,
                              double* work, hipStream_t s);
// Apply of the block preconditioner to several solves in one launch.
struct PJob {
    const void* in;          // [k][N] FP64, or FP32 when in32
    double* out;             // [k][N]
    const int32_t* brange;   // conductor leaf ranges (mode kPreFromB: in = b)
    const int32_t* bstart;
    const int64_t* boff;
    const double* P;
    int N, k, in32;
};
struct PView {
    const PJob* jobs;
    int njobs;
    const int64_t* blk_off;  // cumulative block counts (njobs + 1)
    int64_t nblk;
};
enum PreMode { kPreSet = 0, kPreAdd = 1, kPreFromB = 2 };  // out = P in, out += P in, out = P b
hipError_t launch_block_apply(const PView& pv, int mode, hipStream_t s);

// Two-level ("super") preconditioner: m elements (tree nodes snode[m]) over
// consecutive leaf ranges, the inverse Sinv of their m x m collocation matrix
// (column-major), and per leaf its element sidx[i] and weight w[i] (leaf
// dimension over element dimension).
//   launch_super_weights: w of one job (structural, with the partition).
//   launch_super_invert_lds: Sinv of many jobs with m <= kSuperLds, one
//     workgroup per job, the matrix built and inverted in place in LDS.
//   launch_super_invert_global: Sinv of one job of any m; A (m x m scratch),
//     Gauss-Jordan over the whole device, 2m launches.
constexpr int kSuperLds = 128;
struct SSetup {
    const int32_t* snode;  // [m]
    double* Sinv;          // [m][m] column-major
    int m;
};
hipError_t launch_super_weights(const BatchView& one_job, int N, const int32_t* snode, const int32_t* sidx, double* w,
                                hipStream_t s);
hipError_t launch_super_invert_lds(const BatchView& bv, const SSetup* tab, int njobs, hipStream_t s);
hipError_t launch_super_invert_global(const BatchView& one_job, int m, const int32_t* snode, double* A, double* Sinv,
                                      hipStream_t s);
struct SJob {
    const void* in;  // [k][N] FP64, or FP32 when in32
    double* out;     // [k][N]
    const int32_t* brange;
    const int32_t* sstart;  // [m + 1]
    const int32_t* sidx;    // [N]
    const double* w;        // [N]
    const double* Sinv;     // [m][m] column-major
    double* pot;            // [k][m] scratch of this solve
    double* chg;            // [k][m]
    int N, k, m, in32;
    int lpe;                // lanes per element in the collection (power of two, at most 64)
};
struct SView {
    const SJob* jobs;
    int njobs;
    const int64_t* el_off;  // cumulative k * m per solve
    const int64_t* lf_off;  // cumulative k * N per solve
    const int64_t* ln_off;  // cumulative collection lanes per solve: k * m * lpe rounded up to 64
    const int64_t* em_off;  // cumulative m per solve, each rounded up to 64
    int64_t nel, nlf, nln, nem;
};
hipError_t launch_super_apply(const SView& sv, int mode, hipStream_t s);

// x = P b (Jacobi) or b
hipError_t launch_init_x(const KrylovView& kv, hipStream_t s);
// r = b - w with |r|^2 chunk partials (the product A x is in w)
hipError_t launch_resid(const KrylovView& kv, double* partial, hipStream_t s);
// x += P (sum_jj z[jj] Q_jj)
hipError_t launch_xupdate(const KrylovView& kv, const double* z, hipStream_t s);

// elementwise / small helpers
hipError_t launch_random(double* out, size_t n, uint64_t seed, hipStream_t s);
hipError_t launch_add(const double* x, const double* y, double* out, size_t n, hipStream_t s);
hipError_t launch_sub(const double* x, const double* y, double* out, size_t n, hipStream_t s);
hipError_t launch_jacobi(const double* q, const double* self, int N, int k, double* pq, hipStream_t s);
// y[r][n] += sum_jj z[jj][r] Q_jj[r][n] for jj < nvec
// Conductor charge sums of several finished solves in one launch: for each job,
// out[off + r ncol + c] = sum over the leaves of conductor column c of w[n] x[r][n].
struct ChargeSumJob {
    const double* x;         // [k][N] solution ([k][2N], Re | Im, with w_im)
    const double* w;         // [N] leaf weights
    const double* w_im;      // [N] their imaginary parts (complex jobs; else null): the
                             // imaginary sums follow the real ones at off + k ncol
    const int32_t* col_range;  // [2 ncol] leaf range of every conductor column
    int N, k, ncol;
    int64_t blk;  // first block of this job (ascending, first entry 0)
    int64_t off;  // first output value of this job
};
hipError_t launch_copy_in(const void* src, void* dst, size_t bytes, hipStream_t s);
hipError_t launch_charge_sums(const ChargeSumJob* jobs, int njobs, int64_t nblocks, double* out, hipStream_t s);
// squared norms of the k segments of a [k][N] vector into out[k]
hipError_t launch_sqnorms(const double* x, int N, int k, double* out, hipStream_t s);

// Link discovery (links.hip).  A work item is a node pair (a << 32 | b) or, with
// bit 63 set, the self item of node a.
struct LinkParams {
    int dim;
    double eps;  // link eps (eps ratio x mesh eps)
    double max_side, max_area, max_length;
};
hipError_t launch_links_step(const uint64_t* in, const uint8_t* in_lev, int64_t n, uint64_t* out, uint8_t* out_lev,
                             unsigned long long* out_cnt, int2* links, unsigned long long* link_cnt, int* max_lev,
                             const fcg_node_geom* nodes, const int32_t* left, const int32_t* right,
                             const LinkParams& P, hipStream_t s);
hipError_t launch_links_degree(const int2* links, int64_t P, int32_t* deg, hipStream_t s);
hipError_t launch_links_scatter(const int2* links, int64_t P, unsigned long long* cursor, int32_t* col,
                                hipStream_t s);
// sorts every row of at most links_sort_lds_max() entries; the longer ones are
// listed in rows[nrows] for launch_links_sort_long, with scratch offsets (each
// row padded to a power of two)
hipError_t launch_links_sort_rows(const int64_t* row_ptr, int32_t* col, int M, hipStream_t s);
int links_sort_lds_max();
hipError_t launch_links_sort_long(const int64_t* row_ptr, const int32_t* rows, int nrows, const int64_t* scr_off,
                                  int32_t* col, int32_t* scratch, hipStream_t s);

// Adaptive refinement (links.hip).  A work item is a node pair (a << 32 | b),
// bit 63 set inside one group (curvature coefficient applies).
struct RefineParams {
    int dim;
    double mesh_eps, curv;
    double max_side, max_area, max_length;
};
hipError_t launch_refine_eval(const uint64_t* in, const uint8_t* in_lev, int64_t n, const fcg_refine_node* nodes,
                              const RefineParams& P, int32_t* want, uint64_t* pend, uint8_t* pend_lev,
                              unsigned long long* pend_cnt, unsigned long long* max_eps, int* max_lev, hipStream_t s);
hipError_t launch_refine_split(const uint64_t* pend, int64_t n, fcg_refine_node* nodes, int32_t* want,
                               unsigned long long* pool_cnt, int dim, hipStream_t s);
hipError_t launch_refine_expand(const uint64_t* pend, const uint8_t* pend_lev, int64_t n,
                                const fcg_refine_node* nodes, uint64_t* out, uint8_t* out_lev,
                                unsigned long long* out_cnt, hipStream_t s);
hipError_t launch_refine_split_tops(const int32_t* list, int n, fcg_refine_node* nodes, unsigned long long* pool_cnt,
                                    int dim, hipStream_t s);
hipError_t launch_refine_self(const int32_t* front, int64_t n, int lev, const fcg_refine_node* nodes, uint64_t* out,
                              uint8_t* out_lev, unsigned long long* out_cnt, int32_t* next,
                              unsigned long long* next_cnt, hipStream_t s);
// device layout of discovered links: stored column = charge row of the source;
// fill class per link; stable partition by class (kClassChunk links per
// histogram block, offsets class-major then block); 16-bit columns
hipError_t launch_perm_cols(const int32_t* col_node, int64_t L, const int32_t* perm, int32_t* out, hipStream_t s);
hipError_t launch_link_fill_class(const int32_t* link_row, const int32_t* col, const int32_t* iperm,
                                  const fcg_node_geom* nodes, int dim, int64_t L, uint8_t* cls, hipStream_t s);
int class_chunk();
hipError_t launch_class_hist(const uint8_t* cls, int64_t L, int32_t* hist, hipStream_t s);
hipError_t launch_class_scatter(const uint8_t* cls, int64_t L, const int64_t* off, uint32_t* order, hipStream_t s);
hipError_t launch_col16(const int32_t* col, int64_t n, uint16_t* c16, hipStream_t s);

}  // namespace fcg
