// gfx950 kernels of the FasterCap BEM solve path, batched over jobs.
//
//  fill        one link per lane: PotEstimateOpt(row node, source node)
//              (Solver/Autorefine.cpp:3474-3733, 2D :4156-4325); self terms
//              (SelfPotential, Solver/Autorefine.cpp:2358-2506).  The mixed mode
//              evaluates the analytic near-field integrals in FP64 and the
//              quadrature and point branches in FP32, storing FP32.
//  product     v = P q (Solver/MultiplyHierarchical.cpp:72-330) in three stages:
//              1. up-pass: node charges as subtree sums, level by level over
//                 chunks of post-order nodes (the chunk-straddling top of the
//                 tree by one block per job), node-major [M][kp]
//              2. CSR gather: lanes = (link slot, rhs), each 128-row tile's
//                 source charges staged in LDS, coalesced coefficient and index
//                 streams, XCD-aware tile ranges
//              3. down-pass: path sums of the node potentials, top nodes first,
//                 then level by level over the chunks
//  krylov      batched projections and updates over (job, rhs) segments for
//              GMRES with delayed reorthogonalisation (Solver/SolveCapacitance.cpp:2288-2574)
#include "kernels.h"

#include <algorithm>

#include "../common/potential.h"

namespace fcg {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int find_job(const int64_t* __restrict__ off, int njobs, int64_t g) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// find_job for a g that is the same on every lane of the calling wavefront (all
// 64 lanes active): a 64-way search, each step testing 64 evenly spaced offsets at
// once, so a batch of up to 4096 jobs takes two dependent loads instead of the
// dozen of a binary search.  off ascends with off[0] = 0 <= g; the result is the
// last index i < njobs with off[i] <= g.
__device__ __forceinline__ int find_job_wave(const int64_t* __restrict__ off, int njobs, int64_t g) {
    const int lane = threadIdx.x & 63;
    int lo = 0, hi = njobs;  // the answer lies in [lo, hi)
    while (hi - lo > 1) {
        const int step = (hi - lo + 63) >> 6;
        const int idx = lo + lane * step;
        const bool ok = idx < hi && off[idx] <= g;  // true on a prefix of the lanes (lane 0 at least)
        const unsigned long long m = __ballot(ok);
        const int last = 63 - __clzll((long long)m);
        lo += last * step;
        hi = min(hi, lo + step);
    }
    return lo;
}

// find_job for a block of consecutive work items starting at g0 (every thread
// of the block must call it, before any thread leaves): one wavefront search per
// block, staged in LDS, then a few steps forward per thread; a block that spans
// many jobs falls back to the per-thread search
__device__ __forceinline__ int find_job_block(const int64_t* __restrict__ off, int njobs, int64_t g0, int64_t g) {
    __shared__ int jb0;
    if (threadIdx.x < 64) {
        const int j = find_job_wave(off, njobs, g0);
        if (threadIdx.x == 0) jb0 = j;
    }
    __syncthreads();
    int jb = jb0;
    for (int step = 0; step < 4; ++step) {
        if (jb + 1 >= njobs || off[jb + 1] > g) return jb;
        ++jb;
    }
    return find_job(off, njobs, g);
}

static inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    __syncthreads();
    return s;
}

// ------------------------------------------------------------------ fill
// Link coefficient of PotEstimateOpt (Solver/Autorefine.cpp:3474-3733, 2D
// :4156-4325) for one fill class: branch BR (0 analytic, 1 rule 5, 2 rule 3,
// 3 rule 2, 4 point), evaluated in T.  The branch itself is chosen at upload on
// the FP64 distance ratio, exactly as the reference computes it.
enum FillBranch { kBrAnalytic = 0, kBrRule5 = 1, kBrRule3 = 2, kBrRule2 = 3, kBrPoint = 4 };

// p2: the source, a fcg_node_geom or its SrcGeom copy (vertices, centroid)
template <typename T, int DIM, bool DIEL, int BR, bool GAL = false, typename G2>
__device__ __forceinline__ T link_coef(const fcg_node_geom& p1, const G2& p2) {
    if constexpr (GAL && DIM == 3) {
        // Galerkin branches of PotEstimateOpt (Solver/Autorefine.cpp:3570-3674)
        const Vec3<T> dist{(T)p1.c[0] - (T)p2.c[0], (T)p1.c[1] - (T)p2.c[1], (T)p1.c[2] - (T)p2.c[2]};
        const T k4pe0 = (T)FCG_FOUR_PI_TIMES_E0;
        if constexpr (!DIEL) {
            if constexpr (BR == kBrAnalytic) return mutual_half_3d<T, 17>(p1.v, p2.v);
            if constexpr (BR == kBrRule5) return mutual_half_3d<T, 10>(p1.v, p2.v);
            if constexpr (BR == kBrRule3) return mutual_full_3d<T, 4>(p1.v, p2.v);
            if constexpr (BR == kBrRule2) return mutual_full_3d<T, 2>(p1.v, p2.v);
            if constexpr (BR == kBrPoint) return (T)1 / (k4pe0 * norm3(dist));
        } else {
            const Vec3<T> n1{(T)p1.dn[0], (T)p1.dn[1], (T)p1.dn[2]};
            if constexpr (BR == kBrAnalytic) return mutual_dhalf_3d<T, 17>(p1.v, p2.v, n1, norm3(dist) / (T)20);
            if constexpr (BR == kBrRule5) return mutual_dfull_3d<T, 10>(p1.v, p2.v, n1);
            if constexpr (BR == kBrRule3) return mutual_dfull_3d<T, 5>(p1.v, p2.v, n1);
            if constexpr (BR == kBrRule2) return mutual_dfull_3d<T, 0>(p1.v, p2.v, n1);
            if constexpr (BR == kBrPoint) {
                const T rd = norm3(dist);
                return dot3(dist, n1) / (k4pe0 * rd * rd * rd);
            }
        }
    } else if constexpr (DIM == 2) {
        const T dx = (T)p1.c[0] - (T)p2.c[0], dy = (T)p1.c[1] - (T)p2.c[1];
        if constexpr (!DIEL) {
            if constexpr (BR == kBrAnalytic)
                return potential_opt_2d<T>((T)p1.c[0], (T)p1.c[1], p2.v);
            else
                return -log(sqrt(dx * dx + dy * dy));
        } else {
            if constexpr (BR == kBrAnalytic) {
                return enfield_2d<T>((T)p1.c[0], (T)p1.c[1], p2.v, (T)p1.dn[0], (T)p1.dn[1]);
            } else {
                const T rd = sqrt(dx * dx + dy * dy);
                return (dx * (T)p1.dn[0] + dy * (T)p1.dn[1]) / (rd * rd);
            }
        }
    } else {
        const Vec3<T> c1{(T)p1.c[0], (T)p1.c[1], (T)p1.c[2]};
        const Vec3<T> dist{c1.x - (T)p2.c[0], c1.y - (T)p2.c[1], c1.z - (T)p2.c[2]};
        const T k4pe0 = (T)FCG_FOUR_PI_TIMES_E0;
        if constexpr (!DIEL) {
            if constexpr (BR == kBrAnalytic) return potential_opt_3d<T>(c1, p2.v) / k4pe0;
            if constexpr (BR == kBrRule5) return potential_quad_3d<T, 5>(c1, p2.v) / k4pe0;
            if constexpr (BR == kBrRule3) return potential_quad_3d<T, 3>(c1, p2.v) / k4pe0;
            if constexpr (BR == kBrRule2) return potential_quad_3d<T, 2>(c1, p2.v) / k4pe0;
            if constexpr (BR == kBrPoint) return (T)1 / (k4pe0 * norm3(dist));
        } else {
            const Vec3<T> n1{(T)p1.dn[0], (T)p1.dn[1], (T)p1.dn[2]};
            if constexpr (BR == kBrAnalytic) {
                // divided difference of the potential along the normal (h = rd / 20)
                const T h = norm3(dist) / (T)20;
                const Vec3<T> ep{c1.x + n1.x * h, c1.y + n1.y * h, c1.z + n1.z * h};
                const Vec3<T> em{c1.x - n1.x * h, c1.y - n1.y * h, c1.z - n1.z * h};
                const T pp = potential_opt_3d<T>(ep, p2.v);
                const T pn = potential_opt_3d<T>(em, p2.v);
                return (pn - pp) / ((T)2 * h * k4pe0);
            }
            if constexpr (BR == kBrRule5) return enfield_quad_3d<T, 5>(c1, p2.v, n1) / k4pe0;
            if constexpr (BR == kBrRule3) return enfield_quad_3d<T, 3>(c1, p2.v, n1) / k4pe0;
            if constexpr (BR == kBrRule2) return enfield_quad_3d<T, 2>(c1, p2.v, n1) / k4pe0;
            if constexpr (BR == kBrPoint) {
                const T rd = norm3(dist);
                return dot3(dist, n1) / (k4pe0 * rd * rd * rd);
            }
        }
    }
}

// One fill class over the batch.  Mixed: the analytic near-field integrals, whose
// log and atan terms cancel, stay in FP64; the quadrature and point branches run
// in FP32.  Every mixed coefficient is stored in FP32, and that rounding bounds
// all branches alike.  The quadrature and point classes take kFillPer links per
// thread (a block's links are consecutive class entries, thread t taking entries
// t, t + 256, ...): the index chain of every link (fill order -> link -> source
// charge row -> source node) is loaded before any integral is evaluated, so the
// chains' memory latencies overlap.  The analytic classes (FP64, a few percent of
// the links) keep one link per thread.
constexpr int kFillPer = 4;
template <int BR>
constexpr int fill_per() {
    return BR == kBrAnalytic ? 1 : kFillPer;
}

template <int DIM, bool DIEL, int BR, bool MIXED, bool GAL>
__device__ __forceinline__ void fill_one(const BJob& J, int64_t l, const fcg_node_geom& p1, const SrcGeom& p2) {
    if constexpr (MIXED) {
        // Galerkin: every branch integrated in FP64 (double quadratures of
        // nearby panels), stored in FP32
        if constexpr (BR == kBrAnalytic || GAL)
            J.coef32[l] = (float)link_coef<double, DIM, DIEL, BR, GAL>(p1, p2);
        else
            J.coef32[l] = link_coef<float, DIM, DIEL, BR>(p1, p2);
    } else {
        J.coef64[l] = link_coef<double, DIM, DIEL, BR, GAL>(p1, p2);
    }
}

template <int DIM, bool DIEL, int BR, bool MIXED, bool GAL>
__global__ void __launch_bounds__(256) k_fill(BatchView bv, const int64_t* __restrict__ off, int64_t total, int cls) {
    constexpr int P = fill_per<BR>();
    const int64_t g0 = (int64_t)blockIdx.x * blockDim.x * P;
    const int64_t gf = g0 + threadIdx.x;
    int jb = find_job_block(off, bv.njobs, g0, gf < total ? gf : total - 1);
    int64_t l[P];
    int row[P], src[P];
    // a block whose entries all lie in one job (the common case) reads the job's
    // record once, with scalar loads, and no entry searches for its job
    const int ju = __builtin_amdgcn_readfirstlane(jb);
    const int64_t glast = min(g0 + (int64_t)blockDim.x * P, total) - 1;
    if (ju + 1 >= bv.njobs || off[ju + 1] > glast) {
        const BJob& J = bv.jobs[ju];
        const int64_t fb = J.fill_start[cls] - off[ju];
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int64_t g = gf + (int64_t)i * blockDim.x;
            l[i] = -1;
            if (g < total) {
                l[i] = J.fill_order[fb + g];
                const int2 rs = J.fill_rs[fb + g];  // the source's charge row indexes src_geom
                row[i] = rs.x;
                src[i] = rs.y;
            }
        }
        if constexpr (MIXED && !GAL && BR != kBrAnalytic) {
            // FP32 quadrature and point classes: every entry's geometry in
            // registers before the first store (a store may alias the records as
            // far as the compiler knows); the FP64 bodies keep one entry at a time
            fcg_node_geom a1[P];
            SrcGeom a2[P];
#pragma unroll
            for (int i = 0; i < P; ++i)
                if (l[i] >= 0) {
                    a1[i] = J.nodes[row[i]];
                    a2[i] = J.src_geom[src[i]];
                }
#pragma unroll
            for (int i = 0; i < P; ++i)
                if (l[i] >= 0) fill_one<DIM, DIEL, BR, MIXED, GAL>(J, l[i], a1[i], a2[i]);
        } else {
#pragma unroll
            for (int i = 0; i < P; ++i)
                if (l[i] >= 0) fill_one<DIM, DIEL, BR, MIXED, GAL>(J, l[i], J.nodes[row[i]], J.src_geom[src[i]]);
        }
        return;
    }
    int jl[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        const int64_t g = gf + (int64_t)i * blockDim.x;
        l[i] = -1;
        jl[i] = jb;
        if (g < total) {
            while (jb + 1 < bv.njobs && off[jb + 1] <= g) ++jb;  // entries ascend with i
            const BJob& J = bv.jobs[jb];
            jl[i] = jb;
            const int64_t f = J.fill_start[cls] + (g - off[jb]);
            l[i] = J.fill_order[f];
            const int2 rs = J.fill_rs[f];  // the source's charge row indexes src_geom
            row[i] = rs.x;
            src[i] = rs.y;
        }
    }
#pragma unroll
    for (int i = 0; i < P; ++i)
        if (l[i] >= 0) {
            const BJob& J = bv.jobs[jl[i]];
            fill_one<DIM, DIEL, BR, MIXED, GAL>(J, l[i], J.nodes[row[i]], J.src_geom[src[i]]);
        }
}

// Mixed precision, analytic branch: the coefficient integrated in FP64 once,
// stored as its FP32 copy (read by every FP32 product) and, in the job's near-field
// side list, the difference to FP64 (added by the refinement residual's product).
// One near link per lane; the row node's dielectric flag picks the kernel of
// PotEstimateOpt's conductor or dielectric branch as the fill class does.
template <bool GAL>
__global__ void __launch_bounds__(256) k_fill_near(BatchView bv, const int64_t* __restrict__ off, int64_t total) {
    const int64_t g0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t g = g0 + threadIdx.x;
    const int jb = find_job_block(off, bv.njobs, g0, g < total ? g : total - 1);
    if (g >= total) return;
    const BJob& J = bv.jobs[jb];
    const int64_t sl = g - off[jb];
    const uint32_t l = J.near_link[sl];
    const fcg_node_geom& p1 = J.nodes[J.link_row[l]];
    const SrcGeom& p2 = J.src_geom[J.near_col[sl]];
    const bool diel = (p1.flags & FCG_NODE_DIEL) != 0;
    double c;
    if (J.dim == 2)
        c = diel ? link_coef<double, 2, true, kBrAnalytic>(p1, p2) : link_coef<double, 2, false, kBrAnalytic>(p1, p2);
    else
        c = diel ? link_coef<double, 3, true, kBrAnalytic, GAL>(p1, p2)
                 : link_coef<double, 3, false, kBrAnalytic, GAL>(p1, p2);
    const float c32 = (float)c;
    J.coef32[l] = c32;
    J.near_dc[sl] = c - (double)c32;
}

__global__ void k_near_flag(const uint32_t* __restrict__ fill_order, int64_t p0, int64_t n, uint8_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[fill_order[p0 + i]] = 1;
}

__global__ void k_near_count(const int64_t* __restrict__ row_ptr, const uint8_t* __restrict__ flag, int M,
                             int32_t* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int32_t c = 0;
    for (int64_t l = row_ptr[i]; l < row_ptr[i + 1]; ++l) c += flag[l];
    cnt[i] = c;
}

__global__ void k_near_scatter(const int64_t* __restrict__ row_ptr, const uint8_t* __restrict__ flag,
                               const int32_t* __restrict__ col, int M, const int64_t* __restrict__ near_ptr,
                               int32_t* __restrict__ near_col, uint32_t* __restrict__ near_link) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int64_t o = near_ptr[i];
    for (int64_t l = row_ptr[i]; l < row_ptr[i + 1]; ++l)
        if (flag[l]) {
            near_col[o] = col[l];
            near_link[o] = (uint32_t)l;
            ++o;
        }
}

__global__ void k_link_rows(BatchView bv) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= bv.nodes) return;
    const int jb = find_job(bv.node_off, bv.njobs, g);
    const BJob& J = bv.jobs[jb];
    const int i = (int)(g - bv.node_off[jb]);
    int32_t* lr = const_cast<int32_t*>(J.link_row);
    for (int64_t l = J.row_ptr[i]; l < J.row_ptr[i + 1]; ++l) lr[l] = i;
}

// SelfPotential of any tree node (leaf, refined parent or super node)
// (Solver/Autorefine.cpp:2358-2506)
__device__ double node_self(const BJob& J, const fcg_node_geom& nd) {
    if (nd.flags & FCG_NODE_DIEL) {
        const double eb = J.gout[nd.cond], ea = J.gin[nd.cond];
        if (J.dim == 3) return (ea + eb) / (FCG_TWO_TIMES_E0 * (eb - ea) * nd.area);
        return FCG_PI * (ea + eb) / ((eb - ea) * nd.area);
    }
    if (J.dim == 3) {
        if (J.scheme == FCG_SCHEME_GALERKIN) return auto_3d<double>(nd.v);  // Autorefine.cpp:2399-2401
        const Vec3<double> c{(double)nd.c[0], (double)nd.c[1], (double)nd.c[2]};
        return potential_opt_3d<double>(c, nd.v) / FCG_FOUR_PI_TIMES_E0;
    }
    return potential_opt_2d<double>((double)nd.c[0], (double)nd.c[1], nd.v);
}

// Self term of a dielectric panel of a complex job (SelfPotential's complex
// branch, Solver/Autorefine.cpp:2377-2393, 2D :2454-2471): the complex number
// (ea + eb) / (2 e0 (eb - ea) A) in 3D, pi (ea + eb) / ((eb - ea) L) in 2D, divided
// as numerator times the conjugate denominator over its squared modulus.
__device__ void diel_self_cplx(const BJob& J, const fcg_node_geom& nd, double* re, double* im) {
    const double sr = J.gin[nd.cond] + J.gout[nd.cond], si = J.gin_im[nd.cond] + J.gout_im[nd.cond];
    const double dr = J.gout[nd.cond] - J.gin[nd.cond], di = J.gout_im[nd.cond] - J.gin_im[nd.cond];
    const double scale = J.dim == 3 ? 1.0 / (FCG_TWO_TIMES_E0 * nd.area) : FCG_PI / nd.area;
    const double mod2 = dr * dr + di * di;
    *re = scale * (sr * dr + si * di) / mod2;
    *im = scale * (si * dr - sr * di) / mod2;
}

__global__ void k_self(BatchView bv) {
    const int64_t g0 = (int64_t)blockIdx.x * blockDim.x, g = g0 + threadIdx.x;
    const int jb = find_job_block(bv.leaf_off, bv.njobs, g0, g < bv.leaves ? g : bv.leaves - 1);
    if (g >= bv.leaves) return;
    const BJob& J = bv.jobs[jb];
    const int k = (int)(g - bv.leaf_off[jb]);
    const fcg_node_geom& nd = J.nodes[J.leaf_node[k]];
    double s, si = 0.0;
    if (J.self_im && (nd.flags & FCG_NODE_DIEL))
        diel_self_cplx(J, nd, &s, &si);
    else
        s = node_self(J, nd);
    J.self[k] = s;
    J.diag[J.leaf_node[k]] = s;  // non-leaf entries stay zero
    if (J.self_im) J.self_im[k] = si;
}

// ------------------------------------------------------------------ tree passes
// Up-pass (ComputePanelCharges_fast, Solver/MultiplyHierarchical.cpp:72-127): the
// charge of a node is the sum of its leaves' inputs, so a node's charge is its
// children's sum.  Down-pass (ComputeLeafPotentials_fast, :231-283): a leaf's
// result is the sum of the potentials of the nodes on its path to the root.  Both
// run over chunks of kTreeChunk consecutive post-order nodes: the inner nodes of a
// chunk (subtree inside the chunk) level by level in LDS, one block per chunk,
// and the top nodes (subtrees across chunk boundaries) by one block per job.
// Sums are accumulated in FP64; charges and potentials are stored as TQ, one
// kp-row per node at row perm[node] (padding columns zero).
constexpr int kTreeThreads = 512;      // per chunk (inner kernels): three workgroups fit a CU's LDS
constexpr int kTreeTopThreads = 1024;  // per job (top kernels)

__device__ __forceinline__ double scan_input(const BJob& J, int r, int n, int jacobi) {
    const size_t i = (size_t)r * J.N + n;
    const double x = J.q32 ? (double)J.q32[i] : J.q[i];
    return jacobi ? x / J.self[n] : x;
}

// Inner node charges of one chunk.  The chunk's records go to LDS and the
// leaves' inputs are read in one parallel pass; the levels (deepest first) then
// add children in LDS alone, and one parallel pass writes every node's charge row.
__device__ __forceinline__ int tree_lo(int v) { return (int)(int16_t)(v & 0xffff); }
__device__ __forceinline__ int tree_hi(int v) { return v >> 16; }

template <typename TQ>
__global__ void __launch_bounds__(kTreeThreads) k_up_inner(BatchView bv, int jacobi) {
    __shared__ double tacc[kTreeChunk * kTreeRhs];
    __shared__ int4 rec[kTreeChunk];
    const int64_t g = blockIdx.x;
    const int jb = find_job_wave(bv.tchunk_off, bv.njobs, g);
    const BJob& J = bv.jobs[jb];
    const int t = (int)(g - bv.tchunk_off[jb]);
    const int k = J.k, kp = J.kp;
    TQ* out = (sizeof(TQ) == 8) ? reinterpret_cast<TQ*>(J.charge64) : reinterpret_cast<TQ*>(J.charge32);
    const int l0 = J.tb_lvoff[t], l1 = J.tb_lvoff[t + 1];  // l1 - l0 - 1 levels
    const int e0 = J.tb_lev[l0], e1 = J.tb_lev[l1 - 1];    // the chunk's list
    for (int i = (int)threadIdx.x; i < e1 - e0; i += kTreeThreads) rec[i] = J.tb_rec[e0 + i];
    __syncthreads();
    for (int rg = 0; rg < kp; rg += kTreeRhs) {
        const int nr = min(kTreeRhs, kp - rg);
        for (int i = (int)threadIdx.x; i < e1 - e0; i += kTreeThreads) {  // the leaves' inputs
            const int4 rc = rec[i];
            if (rc.w < 0) continue;
            const int sl = tree_lo(rc.x);
            for (int r = 0; r < nr; ++r)
                tacc[sl * kTreeRhs + r] = rg + r < k ? scan_input(J, rg + r, rc.w, jacobi) : 0.0;
        }
        __syncthreads();
        for (int lv = l1 - 2; lv >= l0; --lv) {
            for (int i = J.tb_lev[lv] - e0 + (int)threadIdx.x; i < J.tb_lev[lv + 1] - e0; i += kTreeThreads) {
                const int4 rc = rec[i];
                if (rc.w >= 0) continue;
                const int sl = tree_lo(rc.x), lc = tree_hi(rc.x), rch = tree_lo(rc.y);
                for (int r = 0; r < nr; ++r)
                    tacc[sl * kTreeRhs + r] = (lc >= 0 ? tacc[lc * kTreeRhs + r] : 0.0) +
                                              (rch >= 0 ? tacc[rch * kTreeRhs + r] : 0.0);
            }
            __syncthreads();
        }
        for (int i = (int)threadIdx.x; i < e1 - e0; i += kTreeThreads) {
            const int4 rc = rec[i];
            const int sl = tree_lo(rc.x);
            TQ* o = out + (size_t)rc.z * kp + rg;
            for (int r = 0; r < nr; ++r) o[r] = (TQ)tacc[sl * kTreeRhs + r];
        }
        __syncthreads();  // tacc is reused by the next group
    }
}

// Top node charges of one job (one block).  Each top node starts from its inner
// children's charges (global: k_up_inner), then the levels (deepest first) add
// the top children in LDS; every top node's row is written at the end.  Jobs
// with more than kTreeTopLds top nodes take the same steps in global memory.
template <typename TQ>
__global__ void __launch_bounds__(kTreeTopThreads) k_up_top(BatchView bv) {
    __shared__ double tacc[kTreeTopLds * kTreeRhs];
    const BJob& J = bv.jobs[blockIdx.x];
    const int k = J.k, kp = J.kp, nt = J.ntop;
    if (nt == 0) return;
    TQ* ch = (sizeof(TQ) == 8) ? reinterpret_cast<TQ*>(J.charge64) : reinterpret_cast<TQ*>(J.charge32);
    if (nt > kTreeTopLds) {
        for (int lv = J.tt_nlev - 1; lv >= 0; --lv) {
            for (int i = J.tt_lev[lv] + (int)threadIdx.x; i < J.tt_lev[lv + 1]; i += kTreeTopThreads) {
                const int4 rw = J.tt_rows[i];  // rows of the node and its children
                TQ* o = ch + (size_t)rw.x * kp;
                for (int r = 0; r < kp; ++r) {
                    double c = 0.0;
                    if (r < k) {
                        if (rw.y >= 0) c += (double)ch[(size_t)rw.y * kp + r];
                        if (rw.z >= 0) c += (double)ch[(size_t)rw.z * kp + r];
                    }
                    o[r] = (TQ)c;
                }
            }
            __syncthreads();
        }
        return;
    }
    for (int rg = 0; rg < kp; rg += kTreeRhs) {
        const int nr = min(kTreeRhs, kp - rg);
        for (int i = (int)threadIdx.x; i < nt; i += kTreeTopThreads) {  // inner children
            const int4 rc = J.tt_rec[i];
            const int4 rw = J.tt_rows[i];
            for (int r = 0; r < nr; ++r) {
                double c = 0.0;
                if (rg + r < k) {
                    if (rw.y >= 0 && rc.z < 0) c += (double)ch[(size_t)rw.y * kp + rg + r];
                    if (rw.z >= 0 && rc.w < 0) c += (double)ch[(size_t)rw.z * kp + rg + r];
                }
                tacc[i * kTreeRhs + r] = c;
            }
        }
        __syncthreads();
        for (int lv = J.tt_nlev - 1; lv >= 0; --lv) {
            for (int i = J.tt_lev[lv] + (int)threadIdx.x; i < J.tt_lev[lv + 1]; i += kTreeTopThreads) {
                const int4 rc = J.tt_rec[i];
                for (int r = 0; r < nr; ++r) {
                    double c = tacc[i * kTreeRhs + r];
                    if (rc.z >= 0) c += tacc[rc.z * kTreeRhs + r];
                    if (rc.w >= 0) c += tacc[rc.w * kTreeRhs + r];
                    tacc[i * kTreeRhs + r] = c;
                }
            }
            __syncthreads();
        }
        for (int i = (int)threadIdx.x; i < nt; i += kTreeTopThreads) {
            TQ* o = ch + (size_t)J.tt_rows[i].x * kp + rg;
            for (int r = 0; r < nr; ++r) o[r] = (TQ)tacc[i * kTreeRhs + r];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ gather
// One wavefront per destination node.  Lane = (slot, rhs group): KR lanes share
// one link slot and each loads VR consecutive right-hand sides of the source row
// with one vector load, so a link's charges are one contiguous kp-row of the
// node-major charge array; the 64/KR slots walk the row's links in stride, U
// links per slot in flight, and a shuffle tree over the slots reduces each
// right-hand side.  Charge rows are laid out depth-major (nodes of one tree depth
// together, post-order within a depth) and each row's links are sorted by source
// row, so neighbouring links of one instruction often share a cache line.  Blocks
// are remapped so that each of the 8 XCDs owns one contiguous range of
// destination rows: post-order rows are spatially coherent, so the charges one
// XCD gathers stay in its own L2.
constexpr int kGatherWaves = 4;  // rows per 256-thread block

template <typename T, int VR>
struct alignas(sizeof(T) * VR) VecT {
    T v[VR];
};

template <typename TC, typename TQ, typename TA, typename TI, int KR, int VR, int U, int RPW>
__global__ void __launch_bounds__(256) k_gather(BatchView bv, int blocks_per_xcd, int64_t row_lo, int64_t row_hi) {
    constexpr int G = 64 / RPW;  // lanes per destination row
    constexpr int SL = G / KR;   // link slots per row
    constexpr int W = KR * VR;
    constexpr int Q = U >= 8 ? 2 : 1;  // aligned quads of links in flight per slot
    using V = VecT<TQ, VR>;
    using CI = VecT<TI, 4>;
    using CC = VecT<TC, 4>;
    const int64_t gb = (int64_t)(blockIdx.x % kNumXcd) * blocks_per_xcd + (blockIdx.x / kNumXcd);
    const int lane = threadIdx.x & 63;
    // global row index over the batch; a lane group never straddles two rows
    const int64_t grow = row_lo + (gb * kGatherWaves + (threadIdx.x >> 6)) * RPW + lane / G;
    const int64_t grow0 = row_lo + gb * kGatherWaves * RPW;
    const int jb = find_job_block(bv.node_off, bv.njobs, grow0 < row_hi ? grow0 : row_hi - 1,
                                  grow < row_hi ? grow : row_hi - 1);
    if (grow >= row_hi) return;
    const BJob& J = bv.jobs[jb];
    // one record gives the row's link range, charge / potential row and node
    const GRow rr = J.grow[grow - bv.node_off[jb]];
    const int gl = lane % G;
    const int slot = gl / KR;
    const int rl = (gl % KR) * VR;
    const int k = J.k, kp = J.kp;
    const TC* coef = (sizeof(TC) == 8) ? reinterpret_cast<const TC*>(J.coef64) : reinterpret_cast<const TC*>(J.coef32);
    const TQ* charge =
        (sizeof(TQ) == 8) ? reinterpret_cast<const TQ*>(J.charge64) : reinterpret_cast<const TQ*>(J.charge32);
    TQ* pot = (sizeof(TQ) == 8) ? reinterpret_cast<TQ*>(J.pot64) : reinterpret_cast<TQ*>(J.pot32);
    // source indices: 16 bits for jobs of at most 65536 nodes (a quarter of the
    // FP32 link stream), 32 bits otherwise
    const TI* cix = (sizeof(TI) == 2) ? reinterpret_cast<const TI*>(J.col16) : reinterpret_cast<const TI*>(J.col);
    const int64_t beg = rr.beg, end = rr.end;
    const TA sf = (TA)J.diag[rr.node];
    const size_t orow = (size_t)rr.prow * kp;
    // The link streams are read as aligned quads: one vector load brings 4
    // indices, one the 4 coefficients (the arrays start 256-byte aligned and
    // carry 4 spare entries, so the quad of the last link is in bounds); links of
    // a quad outside [beg, end) get weight 0 and charge row 0.  A slot takes Q
    // quads per step, the row's slots consecutive quads.
    const int64_t q0 = beg >> 2, q1 = (end + 3) >> 2;
    for (int r0 = 0; r0 < k; r0 += W) {
        const bool active = r0 + rl < k;  // lanes past the last rhs load nothing
        // the row's own charges (self term), read ahead of the links
        TQ selfc[VR];
#pragma unroll
        for (int v = 0; v < VR; ++v)
            selfc[v] = (slot == 0 && r0 + rl + v < k) ? charge[orow + r0 + rl + v] : (TQ)0;
        TA acc[VR];
#pragma unroll
        for (int v = 0; v < VR; ++v) acc[v] = (TA)0;
        if (active) {
            // Software pipeline: the next step's quads are loaded while the
            // charges of the current ones arrive.
            int c[4 * Q];
            TA w[4 * Q];
            auto load_quads = [&](int64_t qs, int* cc, TA* ww) {
#pragma unroll
                for (int qq = 0; qq < Q; ++qq) {
                    const int64_t q = qs + (int64_t)qq * SL;
                    if (q < q1) {
                        const CI iv = *reinterpret_cast<const CI*>(cix + 4 * q);
                        const CC cv = *reinterpret_cast<const CC*>(coef + 4 * q);
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int64_t ll = 4 * q + t;
                            const bool in = ll >= beg && ll < end;
                            cc[4 * qq + t] = in ? (int)iv.v[t] : 0;
                            ww[4 * qq + t] = in ? (TA)cv.v[t] : (TA)0;
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            cc[4 * qq + t] = 0;
                            ww[4 * qq + t] = (TA)0;
                        }
                    }
                }
            };
            load_quads(q0 + slot, c, w);
            for (int64_t qs = q0 + slot; qs < q1; qs += (int64_t)Q * SL) {
                V ch[4 * Q];
#pragma unroll
                for (int u = 0; u < 4 * Q; ++u) ch[u] = *reinterpret_cast<const V*>(charge + (size_t)c[u] * kp + r0 + rl);
                int cn[4 * Q];
                TA wn[4 * Q];
                load_quads(qs + (int64_t)Q * SL, cn, wn);
#pragma unroll
                for (int u = 0; u < 4 * Q; ++u)
#pragma unroll
                    for (int v = 0; v < VR; ++v) acc[v] += w[u] * (TA)ch[u].v[v];
#pragma unroll
                for (int u = 0; u < 4 * Q; ++u) {
                    c[u] = cn[u];
                    w[u] = wn[u];
                }
            }
        }
        // reduce over the slots of this row's lane group
#pragma unroll
        for (int v = 0; v < VR; ++v)
#pragma unroll
            for (int off = KR; off < G; off <<= 1) acc[v] += __shfl_xor(acc[v], off, 64);
        if (slot == 0) {
#pragma unroll
            for (int v = 0; v < VR; ++v)
                if (r0 + rl + v < k) pot[orow + r0 + rl + v] = (TQ)(sf * (TA)selfc[v] + acc[v]);
        }
    }
}

// ------------------------------------------------------------------ staged gather
// One block per tile of kGatherTileRows gather rows of one job; TC, TQ, TA are
// the coefficient, charge and accumulation types of k_gather.  The block copies
// the charges of the tile's distinct sources (its
// list tsrc, ascending charge rows) into LDS, then walks the tile's rows
// kGatherWaves * RPW at a time with the lane shape of k_gather, reading each
// link's charge from LDS by the link's index in the list (lsrc).  The rows of a
// tile are same-depth neighbours (gather row order): 128 rows of a 32768-panel
// plate pair have about 1000 distinct sources for 23000 links.  With no charge
// loads from global memory in the loop, the link quads can be fetched two steps
// ahead (the memory counter waits on them alone).
template <typename TC, typename TQ, typename TA, int KR, int VR, int U, int RPW>
__global__ void __launch_bounds__(256) k_gather_staged(BatchView bv, int blocks_per_xcd, int64_t tile_lo, int64_t tile_hi) {
    constexpr int G = 64 / RPW;  // lanes per destination row
    constexpr int SL = G / KR;   // link slots per row
    constexpr int W = KR * VR;
    constexpr int Q = U >= 8 ? 2 : 1;
    constexpr int kPass = kGatherWaves * RPW;  // rows of one pass of the block
    using V = VecT<TQ, VR>;
    using CI = VecT<uint16_t, 4>;
    using CC = VecT<TC, 4>;
    extern __shared__ __align__(16) unsigned char stage_raw[];  // the largest list of the launch times kp
    TQ* stage = reinterpret_cast<TQ*>(stage_raw);
    const int64_t tg = tile_lo + (int64_t)(blockIdx.x % kNumXcd) * blocks_per_xcd + (blockIdx.x / kNumXcd);
    if (tg >= tile_hi) return;  // the whole block: tg is uniform
    const int jb = find_job_wave(bv.gtile_off, bv.njobs, tg);
    const BJob& J = bv.jobs[jb];
    const int t = (int)(tg - bv.gtile_off[jb]);
    const int k = J.k, kp = J.kp, M = J.M;
    const TQ* charge =
        (sizeof(TQ) == 8) ? reinterpret_cast<const TQ*>(J.charge64) : reinterpret_cast<const TQ*>(J.charge32);
    {
        const int s0 = J.tsrc_off[t];
        const int nf = (J.tsrc_off[t + 1] - s0) * kp;
        const uint16_t* src = J.tsrc + s0;
        const int lk = __ffs(kp) - 1;  // kp is a power of two
        for (int f = threadIdx.x; f < nf; f += 256) stage[f] = charge[(size_t)src[f >> lk] * kp + (f & (kp - 1))];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int gl = lane % G;
    const int slot = gl / KR;
    const int rl = (gl % KR) * VR;
    const TC* coef = (sizeof(TC) == 8) ? reinterpret_cast<const TC*>(J.coef64) : reinterpret_cast<const TC*>(J.coef32);
    const uint16_t* cix = J.lsrc;
    TQ* pot = (sizeof(TQ) == 8) ? reinterpret_cast<TQ*>(J.pot64) : reinterpret_cast<TQ*>(J.pot32);
    const int r_beg = t * kGatherTileRows;
    const int r_end = min(M, r_beg + kGatherTileRows);
    for (int row = r_beg + (int)(threadIdx.x >> 6) * RPW + lane / G; row < r_end; row += kPass) {
        const GRow rr = J.grow[row];
        // link offsets fit in 32 bits (GRow), so the quad arithmetic is 32-bit
        const uint32_t beg = rr.beg, end = rr.end;
        const TA sf = (TA)J.diag[rr.node];
        const size_t orow = (size_t)rr.prow * kp;
        const int q0 = (int)(beg >> 2), q1 = (int)((end + 3) >> 2);
        for (int r0 = 0; r0 < k; r0 += W) {
            const bool active = r0 + rl < k;
            TQ selfc[VR];
#pragma unroll
            for (int v = 0; v < VR; ++v) selfc[v] = (slot == 0 && r0 + rl + v < k) ? charge[orow + r0 + rl + v] : (TQ)0;
            TA acc[VR];
#pragma unroll
            for (int v = 0; v < VR; ++v) acc[v] = (TA)0;
            if (active) {
                // Two-stage ring of raw link quads, unrolled by two so that no
                // register copy waits on a load in flight: while step j is
                // consumed (charges from LDS), the quads of step j + 1 are in
                // flight and those of step j + 2 are issued.  Quads are unpacked
                // (masked to the row's links) only where they are used.
                struct Raw {
                    CI iv[Q];
                    CC cv[Q];
                };
                auto fetch = [&](int qs, Raw& r) {
#pragma unroll
                    for (int qq = 0; qq < Q; ++qq) {
                        const int q = qs + qq * SL;
                        if (q < q1) {
                            r.iv[qq] = *reinterpret_cast<const CI*>(cix + 4 * (size_t)q);
                            r.cv[qq] = *reinterpret_cast<const CC*>(coef + 4 * (size_t)q);
                        } else {
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                r.iv[qq].v[u] = 0;
                                r.cv[qq].v[u] = (TC)0;
                            }
                        }
                    }
                };
                auto consume = [&](int qs, const Raw& r) {
                    int c[4 * Q];
                    TA w[4 * Q];
#pragma unroll
                    for (int qq = 0; qq < Q; ++qq)
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const uint32_t ll = 4u * (uint32_t)(qs + qq * SL) + (uint32_t)u;
                            const bool in = ll >= beg && ll < end;
                            c[4 * qq + u] = in ? (int)r.iv[qq].v[u] : 0;  // a row with links has list entry 0
                            w[4 * qq + u] = in ? (TA)r.cv[qq].v[u] : (TA)0;
                        }
                    V ch[4 * Q];
#pragma unroll
                    for (int u = 0; u < 4 * Q; ++u) ch[u] = *reinterpret_cast<const V*>(stage + c[u] * kp + r0 + rl);
#pragma unroll
                    for (int u = 0; u < 4 * Q; ++u)
#pragma unroll
                        for (int v = 0; v < VR; ++v) acc[v] += w[u] * (TA)ch[u].v[v];
                };
                constexpr int kStep = Q * SL;
                Raw ra, rb;
                int qa = q0 + slot, qb = qa + kStep;
                fetch(qa, ra);
                fetch(qb, rb);
                while (qa < q1) {
                    consume(qa, ra);
                    qa = qb + kStep;
                    fetch(qa, ra);
                    if (qb >= q1) break;
                    consume(qb, rb);
                    qb = qa + kStep;
                    fetch(qb, rb);
                }
            }
#pragma unroll
            for (int v = 0; v < VR; ++v)
#pragma unroll
                for (int off = KR; off < G; off <<= 1) acc[v] += __shfl_xor(acc[v], off, 64);
            if (slot == 0) {
#pragma unroll
                for (int v = 0; v < VR; ++v)
                    if (r0 + rl + v < k) pot[orow + r0 + rl + v] = (TQ)(sf * (TA)selfc[v] + acc[v]);
            }
        }
    }
}

// Source lists of a job's tiles (one block per tile).  The tile's sources are
// marked in a bitmap over the job's charge rows (M <= 65536: 8 KB of LDS); the
// count is the bitmap's population, the list its set bits in order, and a
// link's index the population below its source's bit.
constexpr int kTileBitWords = 65536 / 32;

__device__ void tile_source_bitmap(const GRow* __restrict__ grow, const uint16_t* __restrict__ col16, int M, int t,
                                   uint32_t* bits) {
    const int nw = (M + 31) >> 5;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) bits[i] = 0u;
    __syncthreads();
    const int r1 = min(M, (t + 1) * kGatherTileRows);
    for (int r = t * kGatherTileRows + (int)(threadIdx.x >> 6); r < r1; r += (int)(blockDim.x >> 6)) {
        const GRow g = grow[r];
        for (uint32_t l = g.beg + (threadIdx.x & 63); l < g.end; l += 64) {
            const int c = col16[l];
            atomicOr(&bits[c >> 5], 1u << (c & 31));
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_tile_src_count(const GRow* __restrict__ grow,
                                                        const uint16_t* __restrict__ col16, int M,
                                                        int32_t* __restrict__ cnt) {
    __shared__ uint32_t bits[kTileBitWords];
    __shared__ int wsum[kGatherWaves];
    tile_source_bitmap(grow, col16, M, blockIdx.x, bits);
    const int nw = (M + 31) >> 5;
    int n = 0;
    for (int i = threadIdx.x; i < nw; i += 256) n += __popc(bits[i]);
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < kGatherWaves; ++w) tot += wsum[w];
        cnt[blockIdx.x] = tot;
    }
}

__global__ void __launch_bounds__(256) k_tile_src_fill(const GRow* __restrict__ grow,
                                                       const uint16_t* __restrict__ col16, int M,
                                                       const int32_t* __restrict__ tsrc_off,
                                                       uint16_t* __restrict__ tsrc, uint16_t* __restrict__ lsrc) {
    constexpr int kPer = kTileBitWords / 256;  // bitmap words per thread
    __shared__ uint32_t bits[kTileBitWords];
    __shared__ int pre[kTileBitWords];  // set bits below each word
    __shared__ int wsum[kGatherWaves];
    const int t = blockIdx.x;
    tile_source_bitmap(grow, col16, M, t, bits);
    // exclusive prefix of the words' populations: each thread owns kPer
    // consecutive words; the thread totals are scanned by shuffles inside the
    // wavefronts and the wavefront totals in LDS
    const int nw = (M + 31) >> 5;
    const int w0 = threadIdx.x * kPer;
    int mine = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) mine += (w0 + i < nw) ? __popc(bits[w0 + i]) : 0;
    const int lane = threadIdx.x & 63;
    int incl = mine;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads();
    int base = incl - mine;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) base += wsum[w];
    const int s0 = tsrc_off[t];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int wi = w0 + i;
        if (wi >= nw) break;
        pre[wi] = base;
        uint32_t b = bits[wi];
        int r = base;
        while (b) {
            const int bit = __ffs(b) - 1;
            tsrc[s0 + r++] = (uint16_t)(wi * 32 + bit);
            b &= b - 1u;
        }
        base = r;
    }
    __syncthreads();
    const int r1 = min(M, (t + 1) * kGatherTileRows);
    for (int row = t * kGatherTileRows + (int)(threadIdx.x >> 6); row < r1; row += kGatherWaves) {
        const GRow g = grow[row];
        for (uint32_t l = g.beg + lane; l < g.end; l += 64) {
            const int c = col16[l];
            lsrc[l] = (uint16_t)(pre[c >> 5] + __popc(bits[c >> 5] & ((1u << (c & 31)) - 1u)));
        }
    }
}

hipError_t launch_tile_sources_count(const GRow* grow, const uint16_t* col16, int M, int32_t* cnt, hipStream_t s) {
    if (M < 1 || M > 65536) return hipErrorInvalidValue;
    k_tile_src_count<<<(unsigned)gather_tiles(M), 256, 0, s>>>(grow, col16, M, cnt);
    return hipGetLastError();
}

hipError_t launch_tile_sources_fill(const GRow* grow, const uint16_t* col16, int M, const int32_t* tsrc_off,
                                    uint16_t* tsrc, uint16_t* lsrc, hipStream_t s) {
    if (M < 1 || M > 65536) return hipErrorInvalidValue;
    k_tile_src_fill<<<(unsigned)gather_tiles(M), 256, 0, s>>>(grow, col16, M, tsrc_off, tsrc, lsrc);
    return hipGetLastError();
}

struct GatherCfg {
    int kr, vr, u, rpw;
};
// requested lane shapes: lanes per link, rhs per lane, links in flight per slot,
// rows per wavefront; gather_shape fits them to the batch's row strides
static const GatherCfg kGatherCfgs[] = {
    {2, 4, 4, 1}, {1, 8, 4, 1}, {4, 2, 4, 1}, {1, 8, 2, 1}, {2, 4, 2, 1}, {8, 1, 4, 1}, {4, 2, 2, 1},
    {1, 4, 4, 1}, {16, 1, 2, 1}, {2, 2, 4, 1}, {4, 8, 2, 1}, {8, 4, 2, 1}, {16, 2, 2, 1}, {2, 2, 4, 2},
    {1, 4, 4, 2}, {2, 4, 4, 2}, {4, 2, 4, 2}, {1, 2, 4, 4}, {2, 2, 4, 4}, {1, 2, 2, 4}, {1, 2, 4, 8},
    {2, 2, 4, 8}, {1, 1, 4, 8}, {1, 4, 4, 8}, {1, 2, 2, 8}, {1, 2, 4, 16}, {1, 1, 4, 16}, {1, 2, 2, 16},
    {1, 2, 8, 8}, {1, 4, 8, 8}, {1, 2, 8, 4},
};
constexpr int kNumGatherCfgs = sizeof(kGatherCfgs) / sizeof(kGatherCfgs[0]);
// variants [0, kNumGatherCfgs) read charges from global memory; the next
// kNumGatherCfgs are the same lane shapes with the tile's window staged in LDS
// (k_gather_staged), used where the range allows it and the shape otherwise
int gather_variant_count() { return 2 * kNumGatherCfgs; }
bool gather_variant_staged(int variant) { return variant >= kNumGatherCfgs; }

// VR must divide every job's row stride (vector loads stay inside one row), and
// lanes beyond the largest rhs count are dropped
static GatherCfg gather_shape(int variant, int kpmin, int kmax) {
    if (variant >= kNumGatherCfgs) variant -= kNumGatherCfgs;
    GatherCfg c = kGatherCfgs[(variant >= 0 && variant < kNumGatherCfgs) ? variant : 0];
    const int G = 64 / c.rpw;
    while (c.vr > kpmin) {
        c.vr /= 2;
        c.kr = c.kr * 2 > G / 4 ? c.kr : c.kr * 2;
        if (c.kr > 16) c.kr = 16;
    }
    while (c.kr > 1 && (c.kr / 2) * c.vr >= kmax) c.kr /= 2;
    return c;
}

static int find_cfg(int kr, int vr, int u, int rpw) {
    for (int i = 0; i < kNumGatherCfgs; ++i) {
        const GatherCfg& g = kGatherCfgs[i];
        if (g.kr == kr && g.vr == vr && g.u == u && g.rpw == rpw) return i;
    }
    return 0;
}

int default_gather_variant(int prec, int cls) {
    // fastest shapes measured per rhs class: few rhs want several rows per wavefront
    // (rows average 70-160 links there), many rhs want one row per wavefront with
    // the lanes spread over the rhs.  Every default stages the tiles' source
    // charges in LDS (on the batch: FP32 gather 1.05 -> 0.80 ms, residual 1.59 ->
    // 1.08 ms, FP64 1.47 -> 1.20 ms); a range whose lists do not fit in LDS
    // gathers with the same shape from global memory.
    const bool f32 = prec == kMvp32;
    const int st = kNumGatherCfgs;
    switch (cls) {
        case 0:
        case 1: return st + find_cfg(1, 2, 4, 8);
        case 2: return st + (f32 ? find_cfg(1, 4, 4, 8) : find_cfg(2, 2, 4, 8));
        case 3: return st + find_cfg(2, 4, 4, 2);
        case 4: return st + find_cfg(8, 4, 2, 1);
        default: return st + (f32 ? find_cfg(4, 8, 2, 1) : find_cfg(8, 4, 2, 1));
    }
}

template <typename TC, typename TQ, typename TA>
static hipError_t gather_dispatch_staged(const BatchView& bv, const GatherSeg& sg, hipStream_t s) {
    const GatherCfg c = gather_shape(sg.variant, sg.kpmin, sg.kmax);
    const int64_t blocks = sg.tile_hi - sg.tile_lo;
    const int64_t per_xcd = (blocks + kNumXcd - 1) / kNumXcd;
    const unsigned grid = (unsigned)(per_xcd * kNumXcd);
    if (grid == 0) return hipSuccess;
    const size_t lds = sizeof(TQ) * (size_t)sg.stage_floats;
    if (sg.stage_floats < 1 || lds > kStageMaxBytes) return hipErrorInvalidConfiguration;
#define FCG_S(KR, VR, U, R)                                                                                    \
    if (c.kr == KR && c.vr == VR && c.u == U && c.rpw == R) {                                                  \
        k_gather_staged<TC, TQ, TA, KR, VR, U, R><<<grid, 256, lds, s>>>(bv, (int)per_xcd, sg.tile_lo, sg.tile_hi); \
        return hipGetLastError();                                                                              \
    }
    // every shape gather_shape can return
    FCG_S(1, 1, 2, 1) FCG_S(1, 1, 4, 1) FCG_S(1, 2, 2, 1) FCG_S(1, 2, 4, 1) FCG_S(1, 4, 2, 1)
    FCG_S(1, 4, 4, 1) FCG_S(1, 8, 2, 1) FCG_S(1, 8, 4, 1) FCG_S(2, 1, 2, 1) FCG_S(2, 1, 4, 1)
    FCG_S(2, 2, 2, 1) FCG_S(2, 2, 4, 1) FCG_S(2, 4, 2, 1) FCG_S(2, 4, 4, 1) FCG_S(2, 8, 2, 1)
    FCG_S(4, 1, 2, 1) FCG_S(4, 1, 4, 1) FCG_S(4, 2, 2, 1) FCG_S(4, 2, 4, 1) FCG_S(4, 4, 2, 1)
    FCG_S(4, 8, 2, 1) FCG_S(8, 1, 2, 1) FCG_S(8, 1, 4, 1) FCG_S(8, 2, 2, 1) FCG_S(8, 4, 2, 1)
    FCG_S(16, 1, 2, 1) FCG_S(16, 2, 2, 1) FCG_S(1, 1, 4, 2) FCG_S(1, 2, 4, 2) FCG_S(1, 4, 4, 2)
    FCG_S(2, 1, 4, 2) FCG_S(2, 2, 4, 2) FCG_S(2, 4, 4, 2) FCG_S(4, 1, 4, 2) FCG_S(4, 2, 4, 2)
    FCG_S(8, 1, 4, 2) FCG_S(1, 1, 2, 4) FCG_S(1, 1, 4, 4) FCG_S(1, 2, 2, 4) FCG_S(1, 2, 4, 4)
    FCG_S(2, 1, 2, 4) FCG_S(2, 1, 4, 4) FCG_S(2, 2, 4, 4) FCG_S(4, 1, 4, 4) FCG_S(1, 1, 2, 8)
    FCG_S(1, 1, 4, 8) FCG_S(1, 2, 2, 8) FCG_S(1, 2, 4, 8) FCG_S(1, 4, 4, 8) FCG_S(2, 1, 2, 8)
    FCG_S(2, 1, 4, 8) FCG_S(2, 2, 4, 8) FCG_S(1, 1, 2, 16) FCG_S(1, 1, 4, 16) FCG_S(1, 2, 2, 16)
    FCG_S(1, 2, 4, 16) FCG_S(1, 2, 8, 8) FCG_S(1, 4, 8, 8) FCG_S(1, 2, 8, 4) FCG_S(1, 1, 8, 8)
    FCG_S(1, 1, 8, 4) FCG_S(2, 1, 8, 4) FCG_S(2, 1, 8, 8) FCG_S(1, 4, 8, 4) FCG_S(2, 2, 8, 8) FCG_S(2, 2, 8, 4)
#undef FCG_S
    return hipErrorInvalidConfiguration;
}

template <typename TC, typename TQ, typename TA>
static hipError_t gather_dispatch(const BatchView& bv, const GatherSeg& sg, hipStream_t s) {
    const GatherCfg c = gather_shape(sg.variant, sg.kpmin, sg.kmax);
    const int64_t rows = sg.row_hi - sg.row_lo;
    const int64_t rows_per_block = (int64_t)kGatherWaves * c.rpw;
    const int64_t blocks = (rows + rows_per_block - 1) / rows_per_block;
    const int64_t per_xcd = (blocks + kNumXcd - 1) / kNumXcd;
    const unsigned grid = (unsigned)(per_xcd * kNumXcd);
    if (grid == 0) return hipSuccess;
#define FCG_G(KR, VR, U, R)                                                                    \
    if (c.kr == KR && c.vr == VR && c.u == U && c.rpw == R) {                                  \
        if (sg.idx16)                                                                          \
            k_gather<TC, TQ, TA, uint16_t, KR, VR, U, R><<<grid, 256, 0, s>>>(bv, (int)per_xcd, sg.row_lo, sg.row_hi); \
        else                                                                                   \
            k_gather<TC, TQ, TA, int32_t, KR, VR, U, R><<<grid, 256, 0, s>>>(bv, (int)per_xcd, sg.row_lo, sg.row_hi); \
        return hipGetLastError();                                                              \
    }
    // every shape gather_shape can return for the table above
    FCG_G(1, 1, 2, 1) FCG_G(1, 1, 4, 1) FCG_G(1, 2, 2, 1) FCG_G(1, 2, 4, 1) FCG_G(1, 4, 2, 1)
    FCG_G(1, 4, 4, 1) FCG_G(1, 8, 2, 1) FCG_G(1, 8, 4, 1) FCG_G(2, 1, 2, 1) FCG_G(2, 1, 4, 1)
    FCG_G(2, 2, 2, 1) FCG_G(2, 2, 4, 1) FCG_G(2, 4, 2, 1) FCG_G(2, 4, 4, 1) FCG_G(2, 8, 2, 1)
    FCG_G(4, 1, 2, 1) FCG_G(4, 1, 4, 1) FCG_G(4, 2, 2, 1) FCG_G(4, 2, 4, 1) FCG_G(4, 4, 2, 1)
    FCG_G(4, 8, 2, 1) FCG_G(8, 1, 2, 1) FCG_G(8, 1, 4, 1) FCG_G(8, 2, 2, 1) FCG_G(8, 4, 2, 1)
    FCG_G(16, 1, 2, 1) FCG_G(16, 2, 2, 1) FCG_G(1, 1, 4, 2) FCG_G(1, 2, 4, 2) FCG_G(1, 4, 4, 2)
    FCG_G(2, 1, 4, 2) FCG_G(2, 2, 4, 2) FCG_G(2, 4, 4, 2) FCG_G(4, 1, 4, 2) FCG_G(4, 2, 4, 2)
    FCG_G(8, 1, 4, 2) FCG_G(1, 1, 2, 4) FCG_G(1, 1, 4, 4) FCG_G(1, 2, 2, 4) FCG_G(1, 2, 4, 4)
    FCG_G(2, 1, 2, 4) FCG_G(2, 1, 4, 4) FCG_G(2, 2, 4, 4) FCG_G(4, 1, 4, 4) FCG_G(1, 1, 2, 8)
    FCG_G(1, 1, 4, 8) FCG_G(1, 2, 2, 8) FCG_G(1, 2, 4, 8) FCG_G(1, 4, 4, 8) FCG_G(2, 1, 2, 8)
    FCG_G(2, 1, 4, 8) FCG_G(2, 2, 4, 8) FCG_G(1, 1, 2, 16) FCG_G(1, 1, 4, 16) FCG_G(1, 2, 2, 16)
    FCG_G(1, 2, 4, 16) FCG_G(1, 2, 8, 8) FCG_G(1, 4, 8, 8) FCG_G(1, 2, 8, 4) FCG_G(1, 1, 8, 8)
    FCG_G(1, 1, 8, 4) FCG_G(2, 1, 8, 4) FCG_G(2, 1, 8, 8) FCG_G(1, 4, 8, 4) FCG_G(2, 2, 8, 8) FCG_G(2, 2, 8, 4)
#undef FCG_G
    return hipErrorInvalidConfiguration;
}

// ------------------------------------------------------------------ down-pass
// Top nodes of one job (one block): each top node's potential row becomes the sum
// of the potentials on its path to the root.  The rows go to LDS, the levels
// (shallowest first) add the parent's sum in LDS, and the rows are written back;
// jobs with more than kTreeTopLds top nodes take the same steps in global memory.
template <typename TQ>
__global__ void __launch_bounds__(kTreeTopThreads) k_down_top(BatchView bv) {
    __shared__ double tacc[kTreeTopLds * kTreeRhs];
    const BJob& J = bv.jobs[blockIdx.x];
    const int k = J.k, kp = J.kp, nt = J.ntop;
    TQ* pot = (sizeof(TQ) == 8) ? reinterpret_cast<TQ*>(J.pot64) : reinterpret_cast<TQ*>(J.pot32);
    if (J.dim == 2) {
        // 2D fix-up values, read by k_down_inner (MultiplyMatByVec_fast,
        // Solver/MultiplyHierarchical.cpp:305-330): lt[2r] = the last leaf's result,
        // the potentials on its path summed root first in the order the two passes
        // sum them (the top part rounded to TQ as its row is stored, at every level
        // where the top rows stay in global memory); lt[2r + 1] = the total input
        // charge, the group roots' charge rows of this product
        const TQ* ch = (sizeof(TQ) == 8) ? reinterpret_cast<const TQ*>(J.charge64)
                                         : reinterpret_cast<const TQ*>(J.charge32);
        for (int r = (int)threadIdx.x; r < k; r += kTreeTopThreads) {
            double v = 0.0;
            int a = 0;
            for (; a < J.lpath_top; ++a) {
                v += (double)pot[(size_t)J.lpath[a] * kp + r];
                if (nt > kTreeTopLds) v = (double)(TQ)v;
            }
            v = (double)(TQ)v;
            for (; a < J.lpath_len; ++a) v += (double)pot[(size_t)J.lpath[a] * kp + r];
            double t = 0.0;
            for (int i = 0; i < J.nroots; ++i) t += (double)ch[(size_t)J.perm[J.troots[i]] * kp + r];
            J.lt[2 * r] = v;
            J.lt[2 * r + 1] = t;
        }
        __syncthreads();  // the top pass below rewrites rows read here
    }
    if (nt == 0) return;
    if (nt > kTreeTopLds) {
        for (int lv = 0; lv < J.tt_nlev; ++lv) {
            for (int i = J.tt_lev[lv] + (int)threadIdx.x; i < J.tt_lev[lv + 1]; i += kTreeTopThreads) {
                const int4 rw = J.tt_rows[i];  // rows of the node and its parent
                if (rw.w < 0) continue;
                TQ* o = pot + (size_t)rw.x * kp;
                const TQ* a = pot + (size_t)rw.w * kp;
                for (int r = 0; r < k; ++r) o[r] = (TQ)((double)o[r] + (double)a[r]);
            }
            __syncthreads();
        }
        return;
    }
    for (int rg = 0; rg < k; rg += kTreeRhs) {
        const int nr = min(kTreeRhs, k - rg);
        for (int i = (int)threadIdx.x; i < nt; i += kTreeTopThreads) {
            const TQ* o = pot + (size_t)J.tt_rows[i].x * kp + rg;
            for (int r = 0; r < nr; ++r) tacc[i * kTreeRhs + r] = (double)o[r];
        }
        __syncthreads();
        for (int lv = 0; lv < J.tt_nlev; ++lv) {
            for (int i = J.tt_lev[lv] + (int)threadIdx.x; i < J.tt_lev[lv + 1]; i += kTreeTopThreads) {
                const int ps = J.tt_rec[i].y;
                if (ps < 0) continue;
                for (int r = 0; r < nr; ++r) tacc[i * kTreeRhs + r] += tacc[ps * kTreeRhs + r];
            }
            __syncthreads();
        }
        for (int i = (int)threadIdx.x; i < nt; i += kTreeTopThreads) {
            TQ* o = pot + (size_t)J.tt_rows[i].x * kp + rg;
            for (int r = 0; r < nr; ++r) o[r] = (TQ)tacc[i * kTreeRhs + r];
        }
        __syncthreads();
    }
}

// Inner nodes of one chunk.  One parallel pass puts every inner node's potential
// in LDS, plus its parent's path sum when the parent is a top node (k_down_top);
// the levels (shallowest first) add the parent's sum in LDS alone; one parallel
// pass writes the leaves' sums to v, finishing them there:
//  * complex jobs (MultiplyMatByVec_fast's complex branch,
//    Solver/MultiplyHierarchical.cpp:332-455): each [Re | Im] pair of results is
//    coupled through the imaginary self terms, v_re -= D u_im, v_im += D u_re, u
//    the (Jacobi-scaled) input (D is zero on conductor panels);
//  * 2D jobs (:305-330): every conductor row from the first conductor leaf on,
//    but the last, subtracts the last leaf's result, and the last row becomes the
//    total input charge (both from k_down_top, lt).
template <typename TQ>
__global__ void __launch_bounds__(kTreeThreads) k_down_inner(BatchView bv, int jacobi) {
    __shared__ double tacc[kTreeChunk * kTreeRhs];
    __shared__ int4 rec[kTreeChunk];
    __shared__ double fixv[2 * kTreeRhs];  // 2D: lt of this group of right-hand sides
    const int64_t g = blockIdx.x;
    const int jb = find_job_wave(bv.tchunk_off, bv.njobs, g);
    const BJob& J = bv.jobs[jb];
    const int t = (int)(g - bv.tchunk_off[jb]);
    const int k = J.k, kp = J.kp, N = J.N;
    const bool d2 = J.dim == 2;
    const TQ* pot = (sizeof(TQ) == 8) ? reinterpret_cast<const TQ*>(J.pot64) : reinterpret_cast<const TQ*>(J.pot32);
    const int l0 = J.tb_lvoff[t], l1 = J.tb_lvoff[t + 1];
    const int e0 = J.tb_lev[l0], e1 = J.tb_lev[l1 - 1];
    for (int i = (int)threadIdx.x; i < e1 - e0; i += kTreeThreads) rec[i] = J.tb_rec[e0 + i];
    __syncthreads();
    for (int rg = 0; rg < k; rg += kTreeRhs) {
        const int nr = min(kTreeRhs, k - rg);
        if (d2 && (int)threadIdx.x < 2 * nr) fixv[threadIdx.x] = J.lt[2 * rg + threadIdx.x];
        for (int i = (int)threadIdx.x; i < e1 - e0; i += kTreeThreads) {
            const int4 rc = rec[i];
            const int sl = tree_lo(rc.x);
            const TQ* own = pot + (size_t)rc.z * kp + rg;
            const int pm = J.tb_ppm[e0 + i];
            const TQ* pp = pm >= 0 ? pot + (size_t)pm * kp + rg : nullptr;
            for (int r = 0; r < nr; ++r) tacc[sl * kTreeRhs + r] = (double)own[r] + (pp ? (double)pp[r] : 0.0);
        }
        __syncthreads();
        for (int lv = l0; lv < l1 - 1; ++lv) {
            for (int i = J.tb_lev[lv] - e0 + (int)threadIdx.x; i < J.tb_lev[lv + 1] - e0; i += kTreeThreads) {
                const int4 rc = rec[i];
                const int sl = tree_lo(rc.x), ps = tree_hi(rc.y);
                if (ps < 0) continue;
                for (int r = 0; r < nr; ++r) tacc[sl * kTreeRhs + r] += tacc[ps * kTreeRhs + r];
            }
            __syncthreads();
        }
        for (int i = (int)threadIdx.x; i < e1 - e0; i += kTreeThreads) {
            const int4 rc = rec[i];
            const int n = rc.w;
            if (n < 0) continue;
            const int sl = tree_lo(rc.x);
            const bool fx = d2 && (n >= J.first_cond_leaf || n == N - 1);  // a 2D fix-up row
            const double c = J.self_im ? J.self_im[n] : 0.0;
            if (c == 0.0) {
                for (int r = 0; r < nr; ++r) {
                    double x = tacc[sl * kTreeRhs + r];
                    if (fx) x = n < N - 1 ? x - fixv[2 * r] : fixv[2 * r + 1];
                    J.v[(size_t)(rg + r) * N + n] = x;
                }
            } else {  // complex job: the pairs (rg + r, rg + r + 1), rg and kTreeRhs even
                const double sc = jacobi ? 1.0 / J.self[n] : 1.0;
                for (int r = 0; r + 1 < nr; r += 2) {
                    const size_t ore = (size_t)(rg + r) * N + n, oim = ore + N;
                    const double ure = (J.q32 ? (double)J.q32[ore] : J.q[ore]) * sc;
                    const double uim = (J.q32 ? (double)J.q32[oim] : J.q[oim]) * sc;
                    double xr = tacc[sl * kTreeRhs + r] - c * uim;
                    double xi = tacc[sl * kTreeRhs + r + 1] + c * ure;
                    if (fx) {
                        xr = n < N - 1 ? xr - fixv[2 * r] : fixv[2 * r + 1];
                        xi = n < N - 1 ? xi - fixv[2 * r + 2] : fixv[2 * r + 3];
                    }
                    J.v[ore] = xr;
                    J.v[oim] = xi;
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ launchers
hipError_t launch_link_rows(const BatchView& b, hipStream_t s) {
    if (b.nodes == 0) return hipSuccess;
    k_link_rows<<<blocks_for((size_t)b.nodes, 256), 256, 0, s>>>(b);
    return hipGetLastError();
}

template <bool MIXED, bool GAL>
static void fill_class(const BatchView& b, const int64_t* off, int64_t n, int c, hipStream_t s) {
#define FCG_F(C, DIM, DIEL, BR) \
    case C:                                                                                              \
        k_fill<DIM, DIEL, BR, MIXED, GAL && DIM == 3><<<blocks_for((size_t)n, 256 * fill_per<BR>()), 256, 0, s>>>(b, off, n, c); \
        break;
    switch (c) {
        FCG_F(0, 3, false, 0) FCG_F(1, 3, false, 1) FCG_F(2, 3, false, 2) FCG_F(3, 3, false, 3) FCG_F(4, 3, false, 4)
        FCG_F(5, 3, true, 0) FCG_F(6, 3, true, 1) FCG_F(7, 3, true, 2) FCG_F(8, 3, true, 3) FCG_F(9, 3, true, 4)
        FCG_F(10, 2, false, 0) FCG_F(11, 2, false, 4) FCG_F(12, 2, true, 0) FCG_F(13, 2, true, 4)
        default: break;
    }
#undef FCG_F
}

hipError_t launch_fill(const BatchView& b, const int64_t* const* off, const int64_t* total, int precision,
                       const int64_t* near_off, int64_t near_total, bool galerkin, hipStream_t s) {
    const bool mixed = precision == FCG_PREC_MIXED;
    for (int c = 0; c < kFillClasses; ++c) {
        if (total[c] <= 0) continue;
        if (mixed && near_off && near_class(c)) continue;  // filled from the side lists below
        if (mixed)
            galerkin ? fill_class<true, true>(b, off[c], total[c], c, s) : fill_class<true, false>(b, off[c], total[c], c, s);
        else
            galerkin ? fill_class<false, true>(b, off[c], total[c], c, s)
                     : fill_class<false, false>(b, off[c], total[c], c, s);
    }
    if (mixed && near_off && near_total > 0) {
        if (galerkin)
            k_fill_near<true><<<blocks_for((size_t)near_total, 256), 256, 0, s>>>(b, near_off, near_total);
        else
            k_fill_near<false><<<blocks_for((size_t)near_total, 256), 256, 0, s>>>(b, near_off, near_total);
    }
    if (b.leaves > 0) k_self<<<blocks_for((size_t)b.leaves, 256), 256, 0, s>>>(b);
    return hipGetLastError();
}

__global__ void k_fill_rs(const uint32_t* __restrict__ fill_order, const int32_t* __restrict__ link_row,
                          const int32_t* __restrict__ col, int64_t L, int2* __restrict__ rs) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= L) return;
    const uint32_t l = fill_order[g];
    rs[g] = make_int2(link_row[l], col[l]);
}
hipError_t launch_fill_rs(const uint32_t* fill_order, const int32_t* link_row, const int32_t* col, int64_t L,
                          int2* rs, hipStream_t s) {
    if (L <= 0) return hipSuccess;
    k_fill_rs<<<blocks_for((size_t)L, 256), 256, 0, s>>>(fill_order, link_row, col, L, rs);
    return hipGetLastError();
}
hipError_t launch_near_flag(const uint32_t* fill_order, int64_t p0, int64_t n, uint8_t* flag, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_near_flag<<<blocks_for((size_t)n, 256), 256, 0, s>>>(fill_order, p0, n, flag);
    return hipGetLastError();
}

hipError_t launch_near_count(const int64_t* row_ptr, const uint8_t* flag, int M, int32_t* cnt, hipStream_t s) {
    k_near_count<<<blocks_for((size_t)M, 256), 256, 0, s>>>(row_ptr, flag, M, cnt);
    return hipGetLastError();
}

hipError_t launch_near_scatter(const int64_t* row_ptr, const uint8_t* flag, const int32_t* col, int M,
                               const int64_t* near_ptr, int32_t* near_col, uint32_t* near_link, hipStream_t s) {
    k_near_scatter<<<blocks_for((size_t)M, 256), 256, 0, s>>>(row_ptr, flag, col, M, near_ptr, near_col, near_link);
    return hipGetLastError();
}

// FP64 near-field correction of the refinement residual's product (FP32
// coefficients, FP64 charges and potentials): each node adds to its potentials
// the near links' FP64-minus-FP32 coefficient differences times the source
// charges, so the near field enters in FP64 and the far field in FP32.  One node
// per lane, its near links in stored order (deterministic sums).
template <int KC>
__global__ void k_near_correct(BatchView bv) {
    const int64_t g0 = (int64_t)blockIdx.x * blockDim.x, g = g0 + threadIdx.x;
    const int jb = find_job_block(bv.node_off, bv.njobs, g0, g < bv.nodes ? g : bv.nodes - 1);
    if (g >= bv.nodes) return;
    const BJob& J = bv.jobs[jb];
    if (!J.near_ptr) return;
    const int i = (int)(g - bv.node_off[jb]);
    const int64_t b = J.near_ptr[i], e = J.near_ptr[i + 1];
    if (b == e) return;
    double* pot = J.pot64 + (size_t)J.perm[i] * J.kp;
    for (int r0 = 0; r0 < J.k; r0 += KC) {
        double acc[KC];
#pragma unroll
        for (int r = 0; r < KC; ++r) acc[r] = 0.0;
        for (int64_t sl = b; sl < e; ++sl) {
            const double w = J.near_dc[sl];
            const double* ch = J.charge64 + (size_t)J.near_col[sl] * J.kp + r0;
#pragma unroll
            for (int r = 0; r < KC; ++r)
                if (r0 + r < J.k) acc[r] += w * ch[r];
        }
#pragma unroll
        for (int r = 0; r < KC; ++r)
            if (r0 + r < J.k) pot[r0 + r] += acc[r];
    }
}


// staged when the variant asks for it, every job has its tile source lists and
// the longest list fits in LDS in the product's charge type
static bool seg_staged(const GatherSeg& sg, int prec) {
    return gather_variant_staged(sg.variant) && sg.staged && sg.idx16 &&
           (size_t)sg.stage_floats * (prec == kMvp32 ? 4 : 8) <= kStageMaxBytes;
}

// The up-pass (chunks, then the top nodes of every job), the gather (one launch
// per rhs class range; the refinement residual adds its near-field correction),
// the down-pass (the top nodes of every job with the 2D fix-up values, then the
// chunks, which also couple complex jobs and fix up 2D jobs).
hipError_t launch_mvp(const BatchView& b, int prec, int jacobi, const GatherPlan& plan, hipStream_t s, hipEvent_t* ev) {
    if (b.njobs == 0) return hipSuccess;
    if (ev) (void)hipEventRecord(ev[0], s);
    if (prec == kMvp32)
        k_up_inner<float><<<(unsigned)b.tchunks, kTreeThreads, 0, s>>>(b, jacobi);
    else
        k_up_inner<double><<<(unsigned)b.tchunks, kTreeThreads, 0, s>>>(b, jacobi);
    if (ev) (void)hipEventRecord(ev[1], s);
    if (prec == kMvp32)
        k_up_top<float><<<(unsigned)b.njobs, kTreeTopThreads, 0, s>>>(b);
    else
        k_up_top<double><<<(unsigned)b.njobs, kTreeTopThreads, 0, s>>>(b);
    if (ev) (void)hipEventRecord(ev[2], s);
    for (int i = 0; i < plan.n; ++i) {
        hipError_t ge;
        const GatherSeg& sg = plan.seg[i];
        const bool st = seg_staged(sg, prec);
        if (prec == kMvp64)
            ge = st ? gather_dispatch_staged<double, double, double>(b, sg, s)
                    : gather_dispatch<double, double, double>(b, sg, s);
        else if (prec == kMvp32)
            ge = st ? gather_dispatch_staged<float, float, float>(b, sg, s)
                    : gather_dispatch<float, float, float>(b, sg, s);
        else
            ge = st ? gather_dispatch_staged<float, double, double>(b, sg, s)
                    : gather_dispatch<float, double, double>(b, sg, s);
        if (ge != hipSuccess) return ge;
    }
    if (prec == kMvp32Acc) k_near_correct<8><<<blocks_for((size_t)b.nodes, 256), 256, 0, s>>>(b);
    if (ev) (void)hipEventRecord(ev[3], s);
    if (prec == kMvp32) {
        k_down_top<float><<<(unsigned)b.njobs, kTreeTopThreads, 0, s>>>(b);
        k_down_inner<float><<<(unsigned)b.tchunks, kTreeThreads, 0, s>>>(b, jacobi);
    } else {
        k_down_top<double><<<(unsigned)b.njobs, kTreeTopThreads, 0, s>>>(b);
        k_down_inner<double><<<(unsigned)b.tchunks, kTreeThreads, 0, s>>>(b, jacobi);
    }
    if (ev) (void)hipEventRecord(ev[4], s);
    return hipGetLastError();
}

// ------------------------------------------------------------------ krylov (batched)
constexpr int kKThreads = 256;

constexpr int kKPer = kChunk / kKThreads;  // chunk elements per thread

// ---- Arnoldi with delayed reorthogonalisation (DCGS2)
// Classical Gram-Schmidt applied twice, with the second pass of basis vector j
// folded into the projection of step j + 1 (Bielich, Langou, Thomas,
// Swirydowicz, Yamazaki, Boman, "Low-synch Gram-Schmidt with delayed
// reorthogonalization for Krylov solvers", Parallel Computing 112, 2022): the
// product runs on the provisional vector v_j (basis slot nv = j), one block
// projection gives every dot product the host needs, and one update pass writes
// the final q_j and the next provisional vector.  The basis is read twice per
// step instead of four times.
constexpr int kDcgsWaves = kKThreads / 64;

// grid (nchunks): the chunk's dot products of v (slot nv) and of the product w
// with the nv final basis vectors, and v.v, v.w, w.w.  Each wavefront writes its
// own sums: partial[(c * kDcgsWaves + wave) * P + e], P = 2 maxnv + 3, with
//   e < nv: Q_e . v,   nv <= e < 2 nv: Q_(e-nv) . w,   2 nv: v.v,   2 nv + 1: v.w,   2 nv + 2: w.w
template <typename TB>
__global__ void __launch_bounds__(kKThreads) k_dcgs_dots(KrylovView kv, double* __restrict__ partial) {
    const int P = 2 * kv.maxnv + 3;
    const int c = blockIdx.x;
    const Chunk ch = kv.chunks[c];
    const Seg sg = kv.segs[ch.seg];
    const KJob kj = kv.kjobs[sg.job];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nv = kj.nv;
    const size_t ro = (size_t)sg.r * kj.N;
    const TB* v = reinterpret_cast<TB* const*>(kj.Q)[nv] + ro;
    const double* w = kj.w + ro;
    double* out = partial + ((size_t)c * kDcgsWaves + wv) * P;
    double vr[kKPer], wr[kKPer];
    double svv = 0.0, svw = 0.0, sww = 0.0;
#pragma unroll
    for (int t = 0; t < kKPer; ++t) {
        const int n = ch.n0 + threadIdx.x + t * kKThreads;
        vr[t] = n < ch.n1 ? (double)v[n] : 0.0;
        wr[t] = n < ch.n1 ? w[n] : 0.0;
        svv += vr[t] * vr[t];
        svw += vr[t] * wr[t];
        sww += wr[t] * wr[t];
    }
    // four basis vectors at a time, each read once for both dot products
    constexpr int kG = 4;
    for (int j0 = 0; j0 < nv; j0 += kG) {
        double a[kG], b[kG];
#pragma unroll
        for (int g = 0; g < kG; ++g) {
            a[g] = b[g] = 0.0;
            if (j0 + g < nv) {
                const TB* q = reinterpret_cast<TB* const*>(kj.Q)[j0 + g] + ro;
#pragma unroll
                for (int t = 0; t < kKPer; ++t) {
                    const int n = ch.n0 + threadIdx.x + t * kKThreads;
                    if (n < ch.n1) {
                        const double qq = (double)q[n];
                        a[g] += qq * vr[t];
                        b[g] += qq * wr[t];
                    }
                }
            }
        }
        // the 8 sums over the wavefront by recursive halving: at lane distance
        // 32, 16 and 8 each lane keeps half of its values and sends the other
        // half, so lane bits 5..3 pick one value, then 4, 2 and 1 finish it
        // (10 shuffles instead of 48)
        double x[2 * kG];
#pragma unroll
        for (int g = 0; g < kG; ++g) {
            x[g] = a[g];
            x[kG + g] = b[g];
        }
#pragma unroll
        for (int h = kG, off = 32; h >= 1; h >>= 1, off >>= 1) {
            const bool up = (lane & off) != 0;
#pragma unroll
            for (int i = 0; i < h; ++i) {
                const double keep = up ? x[h + i] : x[i];
                const double send = up ? x[i] : x[h + i];
                x[i] = keep + __shfl_xor(send, off, 64);
            }
        }
        double y = x[0];
        for (int off = 4; off > 0; off >>= 1) y += __shfl_xor(y, off, 64);
        const int idx = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);  // value of this lane
        if ((lane & 7) == 0) {
            const int g = idx & (kG - 1);
            if (j0 + g < nv) out[(idx < kG ? 0 : nv) + j0 + g] = y;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        svv += __shfl_xor(svv, off, 64);
        svw += __shfl_xor(svw, off, 64);
        sww += __shfl_xor(sww, off, 64);
    }
    if (lane == 0) {
        out[2 * nv] = svv;
        out[2 * nv + 1] = svw;
        out[2 * nv + 2] = sww;
    }
}

// grid (nchunks): the update that completes a DCGS2 step, with the host's
// coefficients per (job, rhs) at coef[coff + r (2 nv + 3)]: s[nv], a[nv], then
// inv_nu, cw, cv.  The final basis vector
//     q = (v - sum_{i<nv} s_i Q_i) inv_nu                   over v in slot nv,
// and, when `next` is set, the next provisional vector
//     v' = cw w - sum_{i<nv} a_i Q_i - cv v                 in slot nv + 1.
// Each thread keeps its elements of both in registers across the basis vectors.
template <typename TB>
__global__ void __launch_bounds__(kKThreads) k_dcgs_update(KrylovView kv, const double* __restrict__ coef) {
    const int c = blockIdx.x;
    const Chunk ch = kv.chunks[c];
    const Seg sg = kv.segs[ch.seg];
    const KJob kj = kv.kjobs[sg.job];
    const int nv = kj.nv;
    const size_t ro = (size_t)sg.r * kj.N;
    const double* cf = coef + kj.coff + (size_t)sg.r * (2 * nv + 3);
    const double inv_nu = cf[2 * nv], cw = cf[2 * nv + 1], cv = cf[2 * nv + 2];
    const bool next = kj.next != 0;
    TB* vq = reinterpret_cast<TB* const*>(kj.Q)[nv] + ro;
    const double* w = kj.w + ro;
    double aq[kKPer], av[kKPer];
#pragma unroll
    for (int t = 0; t < kKPer; ++t) {
        const int n = ch.n0 + threadIdx.x + t * kKThreads;
        const double vv = n < ch.n1 ? (double)vq[n] : 0.0;
        const double ww = (next && n < ch.n1) ? w[n] : 0.0;
        aq[t] = vv;
        av[t] = cw * ww - cv * vv;
    }
#pragma unroll 4
    for (int i = 0; i < nv; ++i) {
        const double si = cf[i], ai = cf[nv + i];
        const TB* q = reinterpret_cast<TB* const*>(kj.Q)[i] + ro;
#pragma unroll
        for (int t = 0; t < kKPer; ++t) {
            const int n = ch.n0 + threadIdx.x + t * kKThreads;
            if (n < ch.n1) {
                const double qq = (double)q[n];
                aq[t] -= si * qq;
                av[t] -= ai * qq;
            }
        }
    }
    TB* vn = next ? reinterpret_cast<TB* const*>(kj.Q)[nv + 1] + ro : nullptr;
#pragma unroll
    for (int t = 0; t < kKPer; ++t) {
        const int n = ch.n0 + threadIdx.x + t * kKThreads;
        if (n < ch.n1) {
            vq[n] = (TB)(aq[t] * inv_nu);
            if (next) vn[n] = (TB)av[t];
        }
    }
}

// Rotation of column c of a DCGS2 record: the rotations 0..c-1 applied to the
// unrotated column (rows 0..c+1), the rotated rows 0..c stored beside it, and the
// rotation c that zeroes row c+1 (GMRES's Givens rotation, as the reference's
// gmresPrecondSFastAll: cos = a / |(a, b)|, sin = -b / |(a, b)|).
__device__ void dcgs_rotate(double* __restrict__ D, int c) {
    const double* H = D + dcgs_rec(c);
    double* R = D + dcgs_rec(c) + c + 2;
    double a = H[0];
    for (int r = 0; r < c; ++r) {
        const double* rr = D + dcgs_rec(r) + 2 * r + 3;
        const double cs = rr[0], sn = rr[1];
        const double b = H[r + 1];
        R[r] = cs * a - sn * b;
        a = cs * b + sn * a;
    }
    const double b = H[c + 1];
    const double len = sqrt(a * a + b * b);
    const double cs = len > 0.0 ? a / len : 1.0, sn = len > 0.0 ? -b / len : 0.0;
    R[c] = cs * a - sn * b;
    double* rc = D + dcgs_rec(c) + 2 * c + 3;
    rc[0] = cs;
    rc[1] = sn;
}
__device__ __forceinline__ double& dcgs_g(double* D, int c) { return c == 0 ? D[2] : D[dcgs_rec(c - 1) + 2 * c + 3]; }

// LDS of k_dcgs_coef when it fits in 48 KB (else 0: the projection goes through
// the step buffer and the rotations read the record in global memory): the
// projection of one rhs (2 maxnv + 3 values), then for the rotations the (cs, sn)
// pairs (2 maxnv + 2) and the two columns they rotate (2 maxnv + 3)
__host__ __device__ inline size_t coef_lds_bytes(int maxnv) {
    const size_t b = sizeof(double) * (6 * (size_t)maxnv + 8);
    return b <= 49152 ? b : 0;
}

// dcgs_rotate with column c's entries in H (LDS) and the rotations 0..c-1 in rot
// (LDS, cs and sn interleaved); the rotated rows and the new rotation go to the
// record in D as dcgs_rotate writes them, the new rotation to rot too
__device__ void dcgs_rotate_lds(double* __restrict__ D, int c, const double* H, double* rot) {
    double* R = D + dcgs_rec(c) + c + 2;
    double a = H[0];
    for (int r = 0; r < c; ++r) {
        const double cs = rot[2 * r], sn = rot[2 * r + 1];
        const double b = H[r + 1];
        R[r] = cs * a - sn * b;
        a = cs * b + sn * a;
    }
    const double b = H[c + 1];
    const double len = sqrt(a * a + b * b);
    const double cs = len > 0.0 ? a / len : 1.0, sn = len > 0.0 ? -b / len : 0.0;
    R[c] = cs * a - sn * b;
    double* rc = D + dcgs_rec(c) + 2 * c + 3;
    rc[0] = cs;
    rc[1] = sn;
    rot[2 * c] = cs;
    rot[2 * c + 1] = sn;
}

// grid (nsegs), one wavefront per (job, rhs): the host part of a DCGS2 step, on
// the device.  With s = Q^T v_j, z = Q^T w and the norms from the projection:
// nu^2 = v.v - s.s gives the final q_j = (v_j - Q s) / nu.  Since v_j = nu q_j +
// Q s, column j-1 (A P q_(j-1) = Q h + btil v_j) becomes final: h += btil s and the
// subdiagonal btil nu.  Then A P q_j = (w - Q_j t) / nu with t = H s over the final
// columns, so column j is h = ([z; (v.w - s.z) / nu] - t) / nu, with the
// provisional subdiagonal btil_j = |A P q_j - Q_j h| from the norms, and the next
// provisional vector is v_(j+1) = (w / nu - Q_j (t / nu + h)) / btil_j; its
// coefficients over Q_(j-1), v_j and w go to coef for k_dcgs_update.  Lane 0 then
// rotates the final column j-1 and, provisionally, column j, and reports
// |g_(j+1)|.  A rhs that met the tolerance before gets coefficients that keep its
// vectors.  The wavefront first sums its segment's projection from the chunk and
// wavefront partials (in chunk order) into hbuf.
__global__ void __launch_bounds__(64) k_dcgs_coef(KrylovView kv, const double* __restrict__ partial,
                                                  double* __restrict__ hbuf, double* __restrict__ coef) {
    __shared__ double tj_sh;
    // the segment's projection, in LDS when the launch gave room for 2 maxnv + 3
    // values (else in hbuf at hoff)
    extern __shared__ double hvl[];
    const bool in_lds = coef_lds_bytes(kv.maxnv) > 0;
    const Seg sg = kv.segs[blockIdx.x];
    const KJob kj = kv.kjobs[sg.job];
    const int j = kj.nv, k = kj.k, r = sg.r, lane = threadIdx.x;
    {
        // the partials in (chunk, wavefront) order, eight loads in flight
        const int P = 2 * kv.maxnv + 3;
        const size_t base = (size_t)sg.c_lo * kDcgsWaves;
        const int n = (sg.c_hi - sg.c_lo) * kDcgsWaves;
        for (int e = lane; e < 2 * j + 3; e += 64) {
            const double* pp = partial + e;
            double s = 0.0;
            int t = 0;
            for (; t + 8 <= n; t += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = pp[(base + t + u) * P];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += v[u];
            }
            for (; t < n; ++t) s += pp[(base + t) * P];
            if (in_lds)
                hvl[e] = s;
            else
                hbuf[(size_t)kj.hoff + (size_t)e * k + r] = s;
        }
        __syncthreads();
    }
    const double* hv = hbuf + kj.hoff;
    auto at = [&](int e) { return in_lds ? hvl[e] : hv[(size_t)e * k + r]; };
    double* cf = coef + kj.coff + (size_t)r * (2 * j + 3);
    double* out = hbuf + kj.ooff + 2 * r;
    double* D = kj.hd + (size_t)r * kj.hstride;
    auto keep = [&](double status) {
        for (int e = lane; e < 2 * j + 3; e += 64) cf[e] = (e == 2 * j) ? 1.0 : 0.0;
        if (lane == 0) {
            out[0] = status;
            out[1] = 0.0;
        }
    };
    if (kj.flags[r]) {
        keep(1.0);
        return;
    }
    double ss = 0.0, sz = 0.0;
    for (int i = lane; i < j; i += 64) {
        ss += at(i) * at(i);
        sz += at(i) * at(j + i);
    }
    for (int off = 32; off > 0; off >>= 1) {
        ss += __shfl_xor(ss, off, 64);
        sz += __shfl_xor(sz, off, 64);
    }
    const double vv = at(2 * j), vw = at(2 * j + 1), ww = at(2 * j + 2);
    const double nu2 = vv - ss;
    if (!(vv > 0.0) || !(nu2 > 1e-28 * vv)) {
        keep(2.0);
        return;
    }
    const double nu = sqrt(nu2);
    if (j == 0) {
        if (lane == 0) D[2] = kj.beta[r] * nu;  // r = beta v_0 = beta nu q_0
    } else {
        double* Hp = D + dcgs_rec(j - 1);
        const double bt0 = D[0];
        for (int rr = lane; rr < j; rr += 64) Hp[rr] += bt0 * at(rr);
        if (lane == 0) Hp[j] = bt0 * nu;
    }
    __syncthreads();
    double* Hj = D + dcgs_rec(j);
    double tq = 0.0, tt = 0.0, hh = 0.0;
    for (int rr = lane; rr <= j; rr += 64) {
        double tr = 0.0;
#pragma unroll 4
        for (int c = rr > 0 ? rr - 1 : 0; c < j; ++c) tr += D[dcgs_rec(c) + rr] * at(c);
        const double qw = rr < j ? at(j + rr) : (vw - sz) / nu;
        const double hr = (qw - tr) / nu;
        Hj[rr] = hr;
        if (rr < j) cf[j + rr] = tr;  // t_rr, turned into the coefficient below
        else tj_sh = tr;
        tq += tr * qw;
        tt += tr * tr;
        hh += hr * hr;
    }
    for (int off = 32; off > 0; off >>= 1) {
        tq += __shfl_xor(tq, off, 64);
        tt += __shfl_xor(tt, off, 64);
        hh += __shfl_xor(hh, off, 64);
    }
    __syncthreads();
    const double b2 = (ww - 2.0 * tq + tt) / nu2 - hh;
    const double bt = b2 > 0.0 ? sqrt(b2) : 0.0;
    const double cj = tj_sh / nu + Hj[j];
    for (int i = lane; i < j; i += 64) {
        const double si = at(i);
        cf[i] = si;
        cf[j + i] = bt > 0.0 ? (cf[j + i] / nu + Hj[i] - cj / nu * si) / bt : 0.0;
    }
    if (lane == 0) {
        Hj[j + 1] = bt;
        D[0] = bt;
        cf[2 * j] = 1.0 / nu;
        cf[2 * j + 1] = bt > 0.0 ? 1.0 / (nu * bt) : 0.0;
        cf[2 * j + 2] = bt > 0.0 ? cj / (nu * bt) : 0.0;
    }
    __syncthreads();
    // the serial rotations of lane 0 read LDS: the stored rotations 0..j-2 and
    // the columns j-1 and j, loaded by the whole wavefront first
    double* rot_l = hvl + 2 * kv.maxnv + 3;
    double* col_a = rot_l + 2 * kv.maxnv + 2;  // column j-1, rows 0..j
    double* col_b = col_a + kv.maxnv + 1;  // column j, rows 0..j+1
    if (in_lds) {
        for (int q = lane; q < j - 1; q += 64) {
            const double* rr = D + dcgs_rec(q) + 2 * q + 3;
            rot_l[2 * q] = rr[0];
            rot_l[2 * q + 1] = rr[1];
        }
        if (j >= 1)
            for (int q = lane; q <= j; q += 64) col_a[q] = D[dcgs_rec(j - 1) + q];
        for (int q = lane; q <= j + 1; q += 64) col_b[q] = D[dcgs_rec(j) + q];
        __syncthreads();
    }
    if (lane == 0) {
        if (j >= 1) {
            if (in_lds)
                dcgs_rotate_lds(D, j - 1, col_a, rot_l);
            else
                dcgs_rotate(D, j - 1);
            const double* rc = D + dcgs_rec(j - 1) + 2 * (j - 1) + 3;
            const double gp = D[1];
            dcgs_g(D, j - 1) = rc[0] * gp;
            dcgs_g(D, j) = rc[1] * gp;
        }
        if (in_lds)
            dcgs_rotate_lds(D, j, col_b, rot_l);
        else
            dcgs_rotate(D, j);
        const double* rc = D + dcgs_rec(j) + 2 * j + 3;
        const double gp = dcgs_g(D, j);
        D[1] = gp;
        dcgs_g(D, j) = rc[0] * gp;
        dcgs_g(D, j + 1) = rc[1] * gp;
        out[0] = bt > 0.0 ? 0.0 : 3.0;  // 3: exact breakdown, the residual is final
        out[1] = fabs(rc[1] * gp);
    }
}

// grid (nsegs), one wavefront per (job, rhs): back substitution of the rotated
// Hessenberg system over columns 0..last (R y = g), y into z (0 beyond last)
// LDS of k_dcgs_lsq when it fits in 48 KB (else 0: the substitution works on the
// record and z in global memory): y (maxnz) and the rotated columns 0..maxnz-1
// packed, column ii at ii (ii + 1) / 2
__host__ __device__ inline size_t lsq_lds_bytes(int maxnz) {
    const size_t m = (size_t)(maxnz > 0 ? maxnz : 1);
    const size_t b = sizeof(double) * (m + m * (m + 1) / 2);
    return b <= 49152 ? b : 0;
}

__global__ void __launch_bounds__(64) k_dcgs_lsq(KrylovView kv, double* __restrict__ z) {
    extern __shared__ double lsq_l[];
    const Seg sg = kv.segs[blockIdx.x];
    const KJob kj = kv.kjobs[sg.job];
    const int k = kj.k, r = sg.r, lane = threadIdx.x;
    const int it = kj.last[r];
    double* D = kj.hd + (size_t)r * kj.hstride;
    double* y = z + kj.zoff + r;  // y[jj * k]
    if (lsq_lds_bytes(kv.maxnz) > 0 && it < kv.maxnz) {
        // y and the triangle in LDS (the whole wavefront loads them), then the
        // same substitution in LDS
        double* yl = lsq_l;
        double* Rl = lsq_l + (kv.maxnz > 0 ? kv.maxnz : 1);
        for (int jj = lane; jj <= it; jj += 64) yl[jj] = dcgs_g(D, jj);
        for (int ii = 0; ii <= it; ++ii) {
            const double* R = D + dcgs_rec(ii) + ii + 2;
            for (int jj = lane; jj <= ii; jj += 64) Rl[ii * (ii + 1) / 2 + jj] = R[jj];
        }
        __syncthreads();
        for (int ii = it; ii >= 0; --ii) {
            const double* R = Rl + ii * (ii + 1) / 2;
            const double yi = yl[ii] / R[ii];
            __syncthreads();
            if (lane == 0) yl[ii] = yi;
            for (int jj = lane; jj < ii; jj += 64) yl[jj] -= R[jj] * yi;
            __syncthreads();
        }
        for (int jj = lane; jj < kj.nz; jj += 64) y[(size_t)jj * k] = jj <= it ? yl[jj] : 0.0;
        return;
    }
    for (int jj = lane; jj < kj.nz; jj += 64) y[(size_t)jj * k] = jj <= it ? dcgs_g(D, jj) : 0.0;
    __syncthreads();
    for (int ii = it; ii >= 0; --ii) {
        const double* R = D + dcgs_rec(ii) + ii + 2;  // rotated column ii, rows 0..ii
        const double yi = y[(size_t)ii * k] / R[ii];
        __syncthreads();
        if (lane == 0) y[(size_t)ii * k] = yi;
        for (int jj = lane; jj < ii; jj += 64) y[(size_t)jj * k] -= R[jj] * yi;
        __syncthreads();
    }
}

// grid (nsegs): hbuf[hoff + 2*nv*k + r] = sum of the segment's |w|^2 partials
__global__ void k_reduce_norm(KrylovView kv, const double* __restrict__ partial, double* __restrict__ hbuf) {
    __shared__ double sh[4];
    const Seg sg = kv.segs[blockIdx.x];
    const KJob kj = kv.kjobs[sg.job];
    double s = 0.0;
    for (int c = sg.c_lo + threadIdx.x; c < sg.c_hi; c += blockDim.x) s += partial[c];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) hbuf[(size_t)kj.hoff + (size_t)2 * kj.nv * kj.k + sg.r] = s;
}

// grid (nchunks): Q_nv[r] = w[r] / d[seg]
template <typename TB>
__global__ void k_scale_new(KrylovView kv, const double* __restrict__ d) {
    const int c = blockIdx.x;
    const Chunk ch = kv.chunks[c];
    const Seg sg = kv.segs[ch.seg];
    const KJob kj = kv.kjobs[sg.job];
    const double den = d[ch.seg];
    const double* w = kj.w + (size_t)sg.r * kj.N;
    TB* q = reinterpret_cast<TB* const*>(kj.Q)[kj.nv] + (size_t)sg.r * kj.N;
    for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += blockDim.x) q[n] = (TB)(den != 0.0 ? w[n] / den : 0.0);
}

__global__ void k_init_x(KrylovView kv) {
    const Chunk ch = kv.chunks[blockIdx.x];
    const Seg sg = kv.segs[ch.seg];
    const KJob kj = kv.kjobs[sg.job];
    const size_t o = (size_t)sg.r * kj.N;
    const int lo = kj.brange[2 * sg.r], hi = kj.brange[2 * sg.r + 1];
    for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += blockDim.x) {
        const double b = (n >= lo && n < hi) ? 1.0 : 0.0;
        kj.x[o + n] = kj.jacobi ? b / kj.self[n < kj.nleaf ? n : n - kj.nleaf] : b;
    }
}

__global__ void __launch_bounds__(kKThreads) k_resid(KrylovView kv, double* __restrict__ partial) {
    __shared__ double sh[kKThreads / 64];
    const Chunk ch = kv.chunks[blockIdx.x];
    const Seg sg = kv.segs[ch.seg];
    const KJob kj = kv.kjobs[sg.job];
    const size_t o = (size_t)sg.r * kj.N;
    const int lo = kj.brange[2 * sg.r], hi = kj.brange[2 * sg.r + 1];
    double ss = 0.0;
    for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += kKThreads) {
        const double rr = ((n >= lo && n < hi) ? 1.0 : 0.0) - kj.w[o + n];
        kj.r[o + n] = rr;
        ss += rr * rr;
    }
    ss = block_sum(ss, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = ss;
}

template <typename TB>
__global__ void k_xupdate(KrylovView kv, const double* __restrict__ z) {
    const Chunk ch = kv.chunks[blockIdx.x];
    const Seg sg = kv.segs[ch.seg];
    const KJob kj = kv.kjobs[sg.job];
    const size_t o = (size_t)sg.r * kj.N;
    const double* zz = z + kj.zoff + sg.r;
    for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += blockDim.x) {
        double acc = 0.0;
        for (int jj = 0; jj < kj.nz; ++jj) {
            const double c = zz[(size_t)jj * kj.k];
            if (c != 0.0) acc += c * (double)reinterpret_cast<TB* const*>(kj.Q)[jj][o + n];
        }
        if (kj.y)
            kj.y[o + n] = acc;
        else
            kj.x[o + n] += kj.jacobi ? acc / kj.self[n < kj.nleaf ? n : n - kj.nleaf] : acc;
    }
}

hipError_t launch_init_x(const KrylovView& kv, hipStream_t s) {
    if (kv.nchunks == 0) return hipSuccess;
    k_init_x<<<kv.nchunks, 256, 0, s>>>(kv);
    return hipGetLastError();
}
hipError_t launch_resid(const KrylovView& kv, double* partial, hipStream_t s) {
    if (kv.nchunks == 0) return hipSuccess;
    k_resid<<<kv.nchunks, kKThreads, 0, s>>>(kv, partial);
    return hipGetLastError();
}
hipError_t launch_xupdate(const KrylovView& kv, const double* z, hipStream_t s) {
    if (kv.nchunks == 0) return hipSuccess;
    if (kv.basis32)
        k_xupdate<float><<<kv.nchunks, 256, 0, s>>>(kv, z);
    else
        k_xupdate<double><<<kv.nchunks, 256, 0, s>>>(kv, z);
    return hipGetLastError();
}

hipError_t launch_dcgs_dots(const KrylovView& kv, double* partial, hipStream_t s) {
    if (kv.nchunks == 0) return hipSuccess;
    if (kv.basis32)
        k_dcgs_dots<float><<<kv.nchunks, kKThreads, 0, s>>>(kv, partial);
    else
        k_dcgs_dots<double><<<kv.nchunks, kKThreads, 0, s>>>(kv, partial);
    return hipGetLastError();
}
hipError_t launch_dcgs_update(const KrylovView& kv, const double* coef, hipStream_t s) {
    if (kv.nchunks == 0) return hipSuccess;
    if (kv.basis32)
        k_dcgs_update<float><<<kv.nchunks, kKThreads, 0, s>>>(kv, coef);
    else
        k_dcgs_update<double><<<kv.nchunks, kKThreads, 0, s>>>(kv, coef);
    return hipGetLastError();
}
hipError_t launch_dcgs_coef(const KrylovView& kv, const double* partial, double* hbuf, double* coef, hipStream_t s) {
    if (kv.nsegs == 0) return hipSuccess;
    k_dcgs_coef<<<kv.nsegs, 64, coef_lds_bytes(kv.maxnv), s>>>(kv, partial, hbuf, coef);
    return hipGetLastError();
}
hipError_t launch_dcgs_lsq(const KrylovView& kv, double* z, hipStream_t s) {
    if (kv.nsegs == 0) return hipSuccess;
    k_dcgs_lsq<<<kv.nsegs, 64, lsq_lds_bytes(kv.maxnz), s>>>(kv, z);
    return hipGetLastError();
}
hipError_t launch_reduce_norm(const KrylovView& kv, const double* partial, double* hbuf, hipStream_t s) {
    if (kv.nsegs == 0) return hipSuccess;
    k_reduce_norm<<<kv.nsegs, 256, 0, s>>>(kv, partial, hbuf);
    return hipGetLastError();
}
hipError_t launch_scale_new(const KrylovView& kv, const double* d, hipStream_t s) {
    if (kv.nchunks == 0) return hipSuccess;
    if (kv.basis32)
        k_scale_new<float><<<kv.nchunks, 256, 0, s>>>(kv, d);
    else
        k_scale_new<double><<<kv.nchunks, 256, 0, s>>>(kv, d);
    return hipGetLastError();
}

// ------------------------------------------------------------------ small helpers
__global__ void k_random(double* __restrict__ out, size_t n, uint64_t seed) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    // splitmix64
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (t + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    out[t] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
hipError_t launch_random(double* out, size_t n, uint64_t seed, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_random<<<blocks_for(n, 256), 256, 0, s>>>(out, n, seed);
    return hipGetLastError();
}

__global__ void k_add(const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = x[t] + y[t];
}
__global__ void k_sub(const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = x[t] - y[t];
}
hipError_t launch_add(const double* x, const double* y, double* out, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_add<<<blocks_for(n, 256), 256, 0, s>>>(x, y, out, n);
    return hipGetLastError();
}
hipError_t launch_sub(const double* x, const double* y, double* out, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_sub<<<blocks_for(n, 256), 256, 0, s>>>(x, y, out, n);
    return hipGetLastError();
}

__global__ void k_jacobi(const double* __restrict__ q, const double* __restrict__ self, int N, int k,
                         double* __restrict__ pq) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * k) return;
    pq[t] = q[t] / self[t % N];
}
hipError_t launch_jacobi(const double* q, const double* self, int N, int k, double* pq, hipStream_t s) {
    k_jacobi<<<blocks_for((size_t)N * k, 256), 256, 0, s>>>(q, self, N, k, pq);
    return hipGetLastError();
}


// grid (ncol, k)
// one block per (job, rhs, conductor column)
__global__ void k_charge_sums(const ChargeSumJob* __restrict__ jobs, int njobs, double* __restrict__ out) {
    __shared__ double sh[4];
    const int64_t b = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].blk <= b) lo = mid; else hi = mid - 1;
    }
    const ChargeSumJob& J = jobs[lo];
    const int64_t loc = b - J.blk;
    const int r = (int)(loc / J.ncol), c = (int)(loc % J.ncol);
    if (!J.w_im) {
        double s = 0.0;
        for (int n = J.col_range[2 * c] + threadIdx.x; n < J.col_range[2 * c + 1]; n += blockDim.x)
            s += J.w[n] * J.x[(size_t)r * J.N + n];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) out[J.off + loc] = s;
        return;
    }
    // complex: the complex outer permittivity times the complex charge
    // (SolveCapacitance.cpp:1342-1354), x = [Re(N) | Im(N)] per rhs
    const double* xr = J.x + (size_t)r * 2 * J.N;
    double sr = 0.0, si = 0.0;
    for (int n = J.col_range[2 * c] + threadIdx.x; n < J.col_range[2 * c + 1]; n += blockDim.x) {
        sr += J.w[n] * xr[n] - J.w_im[n] * xr[J.N + n];
        si += J.w_im[n] * xr[n] + J.w[n] * xr[J.N + n];
    }
    sr = block_sum(sr, sh);
    si = block_sum(si, sh);
    if (threadIdx.x == 0) {
        out[J.off + loc] = sr;
        out[J.off + (int64_t)J.k * J.ncol + loc] = si;
    }
}
hipError_t launch_charge_sums(const ChargeSumJob* jobs, int njobs, int64_t nblocks, double* out, hipStream_t s) {
    if (njobs == 0 || nblocks == 0) return hipSuccess;
    k_charge_sums<<<(unsigned)nblocks, 256, 0, s>>>(jobs, njobs, out);
    return hipGetLastError();
}

// Step tables from pinned host staging to device memory, as a kernel on the
// stream that uses them: a copy on another engine would need a cross-queue
// dependency, which costs tens of microseconds at every step boundary.  src is
// host-coherent (fine-grained) memory written by the host before the launch.
__global__ void __launch_bounds__(256) k_copy_in(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

hipError_t launch_copy_in(const void* src, void* dst, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    if ((bytes & 15) || ((uintptr_t)src & 15) || ((uintptr_t)dst & 15)) return hipErrorInvalidValue;
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
    k_copy_in<<<grid, 256, 0, s>>>(reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), n);
    return hipGetLastError();
}

// grid (k): one block per segment
__global__ void k_sqnorms(const double* __restrict__ x, int N, double* __restrict__ out) {
    __shared__ double sh[4];
    const double* xr = x + (size_t)blockIdx.x * N;
    double s = 0.0;
    for (int n = threadIdx.x; n < N; n += blockDim.x) s += xr[n] * xr[n];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}
hipError_t launch_sqnorms(const double* x, int N, int k, double* out, hipStream_t s) {
    k_sqnorms<<<k, 256, 0, s>>>(x, N, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ block preconditioner
// Link coefficient with the branch chosen at run time (setup only): the ratio is
// computed in FP64 without contractions, as the reference and the upload-time
// classification compute it.
template <bool GAL>
__device__ __noinline__ double pair_coef_s(const fcg_node_geom& p1, const fcg_node_geom& p2, int dim) {
    double ratio;
    {
#pragma clang fp contract(off)
        const double dx = (double)p1.c[0] - (double)p2.c[0], dy = (double)p1.c[1] - (double)p2.c[1];
        const double dz = dim == 3 ? (double)p1.c[2] - (double)p2.c[2] : 0.0;
        const double rmax = p1.maxside > p2.maxside ? p1.maxside : p2.maxside;
        ratio = (dim == 3 ? sqrt(dx * dx + dy * dy + dz * dz) : sqrt(dx * dx + dy * dy)) / rmax;
    }
    const bool diel = (p1.flags & FCG_NODE_DIEL) != 0;
    if (dim == 2) {
        if (ratio < 4.0)
            return diel ? link_coef<double, 2, true, kBrAnalytic>(p1, p2) : link_coef<double, 2, false, kBrAnalytic>(p1, p2);
        return diel ? link_coef<double, 2, true, kBrPoint>(p1, p2) : link_coef<double, 2, false, kBrPoint>(p1, p2);
    }
    if (diel) {
        if (ratio < 1.0) return link_coef<double, 3, true, kBrAnalytic, GAL>(p1, p2);
        if (ratio < 2.0) return link_coef<double, 3, true, kBrRule5, GAL>(p1, p2);
        if (ratio < 4.0) return link_coef<double, 3, true, kBrRule3, GAL>(p1, p2);
        if (ratio < 10.0) return link_coef<double, 3, true, kBrRule2, GAL>(p1, p2);
        return link_coef<double, 3, true, kBrPoint, GAL>(p1, p2);
    }
    if (ratio < 1.0) return link_coef<double, 3, false, kBrAnalytic, GAL>(p1, p2);
    if (ratio < 2.0) return link_coef<double, 3, false, kBrRule5, GAL>(p1, p2);
    if (ratio < 4.0) return link_coef<double, 3, false, kBrRule3, GAL>(p1, p2);
    if (ratio < 10.0) return link_coef<double, 3, false, kBrRule2, GAL>(p1, p2);
    return link_coef<double, 3, false, kBrPoint, GAL>(p1, p2);
}

// the scheme of the job decides the branches' integrals (3D)
__device__ __forceinline__ double pair_coef(const BJob& J, const fcg_node_geom& p1, const fcg_node_geom& p2) {
    return J.dim == 3 && J.scheme == FCG_SCHEME_GALERKIN ? pair_coef_s<true>(p1, p2, 3) : pair_coef_s<false>(p1, p2, J.dim);
}

constexpr int kSetupThreads = 256;
constexpr int kSetupLds = 64;  // blocks of at most this many leaves are factored in LDS

// LU with partial pivoting (the first row of largest |entry| in the column, as
// InvertMatrix, Solver/SolveCapacitance.cpp:2116-2258) of the n x n row-major
// matrix A, then its inverse, column k from the unit vector e_k, written
// column-major to out.  Run by a whole workgroup of kSetupThreads threads.
__device__ void lu_invert(double* A, int n, int* piv, double* out, double* rv, int* ri) {
    const int t = threadIdx.x;
    for (int j = 0; j < n; ++j) {
        double best = -1.0;
        int bi = n;
        for (int i = j + t; i < n; i += kSetupThreads) {
            const double v = fabs(A[(size_t)i * n + j]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        rv[t] = best;
        ri[t] = bi;
        __syncthreads();
        for (int off = kSetupThreads / 2; off > 0; off >>= 1) {
            if (t < off) {
                const double v2 = rv[t + off];
                const int i2 = ri[t + off];
                if (v2 > rv[t] || (v2 == rv[t] && i2 < ri[t])) {
                    rv[t] = v2;
                    ri[t] = i2;
                }
            }
            __syncthreads();
        }
        const int p = ri[0];
        if (t == 0) piv[j] = p;
        if (p != j)
            for (int c = t; c < n; c += kSetupThreads) {
                const double tmp = A[(size_t)j * n + c];
                A[(size_t)j * n + c] = A[(size_t)p * n + c];
                A[(size_t)p * n + c] = tmp;
            }
        __syncthreads();
        if (j < n - 1) {
            const double rcp = 1.0 / A[(size_t)j * n + j];
            for (int i = j + 1 + t; i < n; i += kSetupThreads) A[(size_t)i * n + j] *= rcp;
            __syncthreads();
            const int m = n - 1 - j;
            for (int64_t e = t; e < (int64_t)m * m; e += kSetupThreads) {
                const int i = j + 1 + (int)(e / m), c = j + 1 + (int)(e % m);
                A[(size_t)i * n + c] -= A[(size_t)i * n + j] * A[(size_t)j * n + c];
            }
        }
        __syncthreads();
    }
    for (int k = t; k < n; k += kSetupThreads) {
        double* y = out + (size_t)k * n;
        for (int i = 0; i < n; ++i) y[i] = (i == k) ? 1.0 : 0.0;
        for (int i = 0; i < n; ++i) {
            const int ip = piv[i];
            double sum = y[ip];
            y[ip] = y[i];
            for (int c = 0; c < i; ++c) sum -= A[(size_t)i * n + c] * y[c];
            y[i] = sum;
        }
        for (int i = n - 1; i >= 0; --i) {
            double sum = y[i];
            for (int c = i + 1; c < n; ++c) sum -= A[(size_t)i * n + c] * y[c];
            y[i] = sum / A[(size_t)i * n + i];
        }
    }
}

// one workgroup per block (ComputeBlockPrecond, Solver/SolveCapacitance.cpp:2068-2113)
__global__ void __launch_bounds__(kSetupThreads) k_block_setup(BatchView bv, const int32_t* __restrict__ bstart,
                                                               const int64_t* __restrict__ boff,
                                                               double* __restrict__ P, double* __restrict__ work) {
    __shared__ double sh[kSetupLds * kSetupLds];
    __shared__ double rv[kSetupThreads];
    __shared__ int ri[kSetupThreads];
    __shared__ int piv[kMaxBlockPrecond];
    const BJob& J = bv.jobs[0];
    const int b = blockIdx.x;
    const int b0 = bstart[b], n = bstart[b + 1] - b0;
    double* A = n <= kSetupLds ? sh : work + boff[b];  // row-major n x n
    for (int e = threadIdx.x; e < n * n; e += kSetupThreads) {
        const int i = e / n, j = e - i * n;
        A[e] = (i == j) ? J.self[b0 + i] : pair_coef(J, J.nodes[J.leaf_node[b0 + i]], J.nodes[J.leaf_node[b0 + j]]);
    }
    __syncthreads();
    lu_invert(A, n, piv, P + boff[b], rv, ri);
}

// Two-level preconditioner matrix (ComputeSuperPrecond,
// Solver/SolveCapacitance.cpp:2007-2066): element i's node self term on the
// diagonal, link coefficients between element nodes elsewhere; column-major
// m x m (A[j * m + i] = row i, column j).  X is set to the identity.
__global__ void k_super_build(BatchView bv, const int32_t* __restrict__ snode, int m, double* __restrict__ A,
                              double* __restrict__ X) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)m * m) return;
    const BJob& J = bv.jobs[0];
    const int j = (int)(e / m), i = (int)(e % m);
    const fcg_node_geom& ni = J.nodes[snode[i]];
    A[e] = (i == j) ? node_self(J, ni) : pair_coef(J, ni, J.nodes[snode[j]]);
    X[e] = (i == j) ? 1.0 : 0.0;
}

// per leaf: its element's index weight (leaf dimension over element dimension)
__global__ void k_super_weights(BatchView bv, const int32_t* __restrict__ snode, const int32_t* __restrict__ sidx,
                                double* __restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const BJob& J = bv.jobs[0];
    if (i >= J.N) return;
    w[i] = J.nodes[J.leaf_node[i]].area / J.nodes[snode[sidx[i]]].area;
}

// Gauss-Jordan inversion of the column-major n x n matrix A into X (which holds
// the identity on entry), spread over the whole device: per column j, one
// workgroup picks the pivot (the first row of largest |entry| at or below the
// diagonal, the choice of InvertMatrix's partial pivoting,
// Solver/SolveCapacitance.cpp:2116-2258), swaps it into row j and scales that
// row; then every other row is cleared in column j by a grid of workgroups.
// Column j of A is left in place (it holds the multipliers) and never read
// again.  X ends as the inverse, column-major.
__global__ void __launch_bounds__(kSetupThreads) k_gj_pivot(double* __restrict__ A, double* __restrict__ X, int n,
                                                            int j) {
    __shared__ double rv[kSetupThreads];
    __shared__ int ri[kSetupThreads];
    const int t = threadIdx.x;
    double best = -1.0;
    int bi = n;
    for (int i = j + t; i < n; i += kSetupThreads) {
        const double v = fabs(A[(size_t)j * n + i]);
        if (v > best) {
            best = v;
            bi = i;
        }
    }
    rv[t] = best;
    ri[t] = bi;
    __syncthreads();
    for (int off = kSetupThreads / 2; off > 0; off >>= 1) {
        if (t < off) {
            const double v2 = rv[t + off];
            const int i2 = ri[t + off];
            if (v2 > rv[t] || (v2 == rv[t] && i2 < ri[t])) {
                rv[t] = v2;
                ri[t] = i2;
            }
        }
        __syncthreads();
    }
    const int p = ri[0];
    const double pv = A[(size_t)j * n + p];
    __syncthreads();  // every thread holds the pivot before column j changes
    const double rcp = 1.0 / pv;
    // swap rows j and p and scale the new row j (columns < j of A are unit
    // vectors by now and are left alone); with p == j the second store wins
    for (int c = j + t; c < n; c += kSetupThreads) {
        double* M = A + (size_t)c * n;
        const double a = M[j], b = M[p];
        M[p] = a;
        M[j] = b * rcp;
    }
    for (int c = t; c < n; c += kSetupThreads) {
        double* M = X + (size_t)c * n;
        const double a = M[j], b = M[p];
        M[p] = a;
        M[j] = b * rcp;
    }
}

constexpr int kGjCols = 8;  // columns per thread in the elimination
__global__ void __launch_bounds__(256) k_gj_elim(double* __restrict__ A, double* __restrict__ X, int n, int j) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || i == j) return;
    const double f = A[(size_t)j * n + i];
    // columns j+1..n-1 of A, then 0..n-1 of X: index c' in [0, 2n - j - 1)
    const int c0 = blockIdx.y * kGjCols;
    const int ncol = 2 * n - j - 1;
#pragma unroll
    for (int q = 0; q < kGjCols; ++q) {
        const int c = c0 + q;
        if (c >= ncol) break;
        double* M = c < n - j - 1 ? A + (size_t)(j + 1 + c) * n : X + (size_t)(c - (n - j - 1)) * n;
        M[i] -= f * M[j];
    }
}

// Two-level preconditioner inverse of job blockIdx.x (m <= kSuperLds), one
// workgroup per job: the element matrix (as k_super_build) is built in LDS and
// inverted in place by Gauss-Jordan elimination with the pivot choice of
// InvertMatrix (Solver/SolveCapacitance.cpp:2116-2258: first row of largest
// |entry| at or below the diagonal).  Per column j: the pivot row p is swapped
// into row j and scaled (its column-j entry taken as 1 first), every other row
// i loses f_i times it (its column-j entry taken as 0 first), which leaves
// column j of the inverse of the row-permuted matrix in place; at the end the
// columns are swapped back in reverse order of the row swaps.
constexpr int kSuperThreads = 1024;
__global__ void __launch_bounds__(kSuperThreads) k_super_invert_lds(BatchView bv, const SSetup* __restrict__ tab) {
    __shared__ double A[kSuperLds * kSuperLds];  // column-major: A[c * m + i] = row i, column c
    __shared__ double rowj[kSuperLds], colj[kSuperLds];
    __shared__ int perm[kSuperLds];
    static_assert(kSuperThreads % kSuperLds == 0, "whole columns per thread group");
    const BJob& J = bv.jobs[blockIdx.x];
    const SSetup st = tab[blockIdx.x];
    const int m = st.m, t = threadIdx.x, mm = m * m;
    for (int e = t; e < mm; e += kSuperThreads) {
        const int c = e / m, i = e - c * m;
        const fcg_node_geom& ni = J.nodes[st.snode[i]];
        A[e] = (i == c) ? node_self(J, ni) : pair_coef(J, ni, J.nodes[st.snode[c]]);
    }
    __syncthreads();
    for (int j = 0; j < m; ++j) {
        if (t < 64) {  // one wavefront picks the pivot row
            double best = -1.0;
            int bi = m;
            for (int i = j + t; i < m; i += 64) {
                const double v = fabs(A[j * m + i]);
                if (v > best) {
                    best = v;
                    bi = i;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double v2 = __shfl_down(best, off);
                const int i2 = __shfl_down(bi, off);
                if (v2 > best || (v2 == best && i2 < bi)) {
                    best = v2;
                    bi = i2;
                }
            }
            if (t == 0) perm[j] = bi;
        }
        __syncthreads();
        const int p = perm[j];
        const double rcp = 1.0 / A[j * m + p];
        double a = 0.0, b = 0.0, f = 0.0;
        if (t < m) {  // thread t: column t of rows j and p, row t of column j (after the swap)
            a = A[t * m + j];
            b = A[t * m + p];
            f = A[j * m + (t == p ? j : t)];
        }
        __syncthreads();
        if (t < m) {
            const double r = (t == j ? 1.0 : b) * rcp;
            rowj[t] = r;
            colj[t] = f;
            if (p != j) A[t * m + p] = a;
            A[t * m + j] = r;
        }
        __syncthreads();
        // thread t: row t % kSuperLds of the columns congruent to t / kSuperLds
        // modulo kSuperThreads / kSuperLds (no division by m in the loop)
        const int ei = t % kSuperLds;
        if (ei < m && ei != j) {
            const double f = colj[ei];
            for (int c = t / kSuperLds; c < m; c += kSuperThreads / kSuperLds) {
                double* a = &A[c * m + ei];
                *a = (c == j ? 0.0 : *a) - f * rowj[c];
            }
        }
        __syncthreads();
    }
    for (int j = m - 1; j >= 0; --j) {
        const int p = perm[j];  // uniform over the workgroup
        if (p == j) continue;
        if (t < m) {
            const double x = A[j * m + t];
            A[j * m + t] = A[p * m + t];
            A[p * m + t] = x;
        }
        __syncthreads();
    }
    for (int e = t; e < mm; e += kSuperThreads) st.Sinv[e] = A[e];
}

hipError_t launch_super_weights(const BatchView& one_job, int N, const int32_t* snode, const int32_t* sidx, double* w,
                                hipStream_t s) {
    if (N <= 0) return hipSuccess;
    k_super_weights<<<blocks_for((size_t)N, 256), 256, 0, s>>>(one_job, snode, sidx, w);
    return hipGetLastError();
}

hipError_t launch_super_invert_lds(const BatchView& bv, const SSetup* tab, int njobs, hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    k_super_invert_lds<<<(unsigned)njobs, kSuperThreads, 0, s>>>(bv, tab);
    return hipGetLastError();
}

hipError_t launch_super_invert_global(const BatchView& one_job, int m, const int32_t* snode, double* A, double* Sinv,
                                      hipStream_t s) {
    if (m <= 0) return hipSuccess;
    k_super_build<<<blocks_for((size_t)m * m, 256), 256, 0, s>>>(one_job, snode, m, A, Sinv);
    for (int j = 0; j < m; ++j) {
        k_gj_pivot<<<1, kSetupThreads, 0, s>>>(A, Sinv, m, j);
        const dim3 grid((unsigned)((m + 255) / 256), (unsigned)((2 * m - j - 1 + kGjCols - 1) / kGjCols));
        if (grid.y > 0) k_gj_elim<<<grid, 256, 0, s>>>(A, Sinv, m, j);
    }
    return hipGetLastError();
}

hipError_t launch_block_setup(const BatchView& one_job, int nb, const int32_t* bstart, const int64_t* boff, double* P,
                              double* work, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    k_block_setup<<<(unsigned)nb, kSetupThreads, 0, s>>>(one_job, bstart, boff, P, work);
    return hipGetLastError();
}

// one wavefront per (solve, block): lanes over the block's rows, up to kPreRhs
// right-hand sides per pass so each pass reads the block's inverse once
constexpr int kPreRhs = 8;
template <int MODE>
__global__ void __launch_bounds__(256) k_block_apply(PView pv) {
    const int64_t gb = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gb >= pv.nblk) return;  // whole wavefronts
    const int jb = find_job_wave(pv.blk_off, pv.njobs, gb);
    const PJob& J = pv.jobs[jb];
    const int b = (int)(gb - pv.blk_off[jb]);
    const int b0 = J.bstart[b], n = J.bstart[b + 1] - b0;
    const double* Pb = J.P + J.boff[b];
    const int lane = threadIdx.x & 63;
    const double* in64 = static_cast<const double*>(J.in);
    const float* in32 = static_cast<const float*>(J.in);
    for (int r0 = 0; r0 < J.k; r0 += kPreRhs) {
        double acc[kPreRhs][2];
#pragma unroll
        for (int r = 0; r < kPreRhs; ++r) acc[r][0] = acc[r][1] = 0.0;
        for (int j = 0; j < n; ++j) {
            const double p0 = lane < n ? Pb[(size_t)j * n + lane] : 0.0;
            const double p1 = lane + 64 < n ? Pb[(size_t)j * n + lane + 64] : 0.0;
#pragma unroll
            for (int r = 0; r < kPreRhs; ++r) {
                if (r0 + r >= J.k) break;
                const size_t o = (size_t)(r0 + r) * J.N + b0 + j;
                double x;
                if (MODE == kPreFromB) {
                    const int leaf = b0 + j;
                    x = (leaf >= J.brange[2 * (r0 + r)] && leaf < J.brange[2 * (r0 + r) + 1]) ? 1.0 : 0.0;
                } else {
                    x = J.in32 ? (double)in32[o] : in64[o];
                }
                acc[r][0] += p0 * x;
                acc[r][1] += p1 * x;
            }
        }
#pragma unroll
        for (int r = 0; r < kPreRhs; ++r) {
            if (r0 + r >= J.k) break;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = lane + 64 * h;
                if (i < n) {
                    double* o = J.out + (size_t)(r0 + r) * J.N + b0 + i;
                    if (MODE == kPreAdd)
                        *o += acc[r][h];
                    else
                        *o = acc[r][h];
                }
            }
        }
    }
}

hipError_t launch_block_apply(const PView& pv, int mode, hipStream_t s) {
    if (pv.nblk == 0) return hipSuccess;
    const unsigned grid = (unsigned)((pv.nblk + 3) / 4);
    if (mode == kPreSet)
        k_block_apply<kPreSet><<<grid, 256, 0, s>>>(pv);
    else if (mode == kPreAdd)
        k_block_apply<kPreAdd><<<grid, 256, 0, s>>>(pv);
    else
        k_block_apply<kPreFromB><<<grid, 256, 0, s>>>(pv);
    return hipGetLastError();
}

// Two-level preconditioner apply (ComputePrecondVectFast super branch,
// Solver/SolveCapacitance.cpp:3545-3592), batched over solves.
// collect: a group of lpe lanes per (solve, rhs, element): pot = sum of w_i in_i
// over the element's leaves.  lpe follows the solve's leaves per element; each
// solve's lanes start on a wavefront boundary, so the groups are aligned and a
// group's shuffles stay inside it (padding lanes form whole groups and write
// nothing).
template <int MODE>
__global__ void __launch_bounds__(256) k_super_collect(SView sv) {
    const int64_t g0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t g = g0 + threadIdx.x;
    const int jb = find_job_block(sv.ln_off, sv.njobs, g0, g < sv.nln ? g : sv.nln - 1);
    if (g >= sv.nln) return;
    const SJob& J = sv.jobs[jb];
    // a solve's lanes fit in 32 bits: the host rejects solves whose padded lane
    // count (k m lpe rounded up to a wavefront) reaches 2^31
    const int local = (int)(g - sv.ln_off[jb]);
    const int lpe = J.lpe;
    const int grp = local >> __builtin_ctz(lpe);  // lpe is a power of two
    const int sub = local & (lpe - 1);
    const bool live = grp < J.k * J.m;
    const int r = live ? grp / J.m : 0, e = live ? grp - r * J.m : 0;
    const double* in64 = static_cast<const double*>(J.in);
    const float* in32 = static_cast<const float*>(J.in);
    double acc = 0.0;
    if (live) {
        // four leaves per lane in flight; the sum runs over i ascending as one
        // leaf per step would
        constexpr int kU = 4;
        const int i1 = J.sstart[e + 1];
        for (int ib = J.sstart[e] + sub; ib < i1; ib += kU * lpe) {
            double x[kU], wv[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int i = ib + u * lpe;
                const bool in = i < i1;
                if (MODE == kPreFromB)
                    x[u] = (in && i >= J.brange[2 * r] && i < J.brange[2 * r + 1]) ? 1.0 : 0.0;
                else
                    x[u] = !in ? 0.0 : J.in32 ? (double)in32[(size_t)r * J.N + i] : in64[(size_t)r * J.N + i];
                wv[u] = in ? J.w[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (ib + u * lpe < i1) acc += wv[u] * x[u];
        }
    }
    for (int off = lpe >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (live && sub == 0) J.pot[(size_t)r * J.m + e] = acc;
}

// multiply: chg = sum over t != s of Sinv[s][t] pot[t] (the diagonal is applied
// per leaf in the distribution).  A workgroup of 4 wavefronts takes 64 elements s
// of one solve (each solve's elements padded to whole workgroups); wavefront q
// sums over the quarter q of t, the right-hand sides in passes of kSuperRhs over
// Sinv's rows, and the partial sums meet in LDS.
constexpr int kSuperRhs = 4;
__global__ void __launch_bounds__(256) k_super_mult(SView sv) {
    __shared__ double part[3][kSuperRhs][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t b0 = (int64_t)blockIdx.x * 64;
    if (b0 >= sv.nem) return;  // uniform over the workgroup
    const int jb = find_job_wave(sv.em_off, sv.njobs, b0);
    const SJob& J = sv.jobs[jb];
    const int m = J.m, e = (int)(b0 - sv.em_off[jb]) + lane;
    const bool live = e < m;
    const int tq = (m + 3) / 4, t0 = q * tq, t1 = min(m, t0 + tq);
    const double* __restrict__ S = J.Sinv + (live ? e : 0);  // S[t * m] = Sinv[e][t]
    for (int r0 = 0; r0 < J.k; r0 += kSuperRhs) {
        const int nr = min(kSuperRhs, J.k - r0);
        const double* pot = J.pot + (size_t)r0 * m;
        double acc[kSuperRhs] = {};
        if (live) {
            // kU steps of t in flight; the sums run over t ascending as one step
            // at a time would
            constexpr int kU = 8;
            for (int tb = t0; tb < t1; tb += kU) {
                double a[kU], p[kU][kSuperRhs];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int t = tb + u;
                    a[u] = (t < t1 && t != e) ? S[(size_t)t * m] : 0.0;
#pragma unroll
                    for (int r = 0; r < kSuperRhs; ++r) p[u][r] = (t < t1 && r < nr) ? pot[(size_t)r * m + t] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    if (tb + u < t1) {
#pragma unroll
                        for (int r = 0; r < kSuperRhs; ++r)
                            if (r < nr) acc[r] += a[u] * p[u][r];
                    }
            }
        }
        if (q > 0) {
#pragma unroll
            for (int r = 0; r < kSuperRhs; ++r) part[q - 1][r][lane] = acc[r];
        }
        __syncthreads();
        if (q == 0 && live) {
#pragma unroll
            for (int r = 0; r < kSuperRhs; ++r)
                if (r < nr)
                    J.chg[(size_t)(r0 + r) * m + e] = acc[r] + part[0][r][lane] + part[1][r][lane] + part[2][r][lane];
        }
        __syncthreads();
    }
}

// distribute: kDistPer (solve, rhs, leaf) items per thread, at g0 + threadIdx.x +
// u 256, every item's loads issued before the first result is written
constexpr int kDistPer = 4;
template <int MODE>
__global__ void __launch_bounds__(256) k_super_dist(SView sv) {
    const int64_t g0 = (int64_t)blockIdx.x * 256 * kDistPer;
    int jb = find_job_block(sv.lf_off, sv.njobs, g0, min(g0 + (int64_t)threadIdx.x, sv.nlf - 1));
    int jj[kDistPer], ii[kDistPer], rr[kDistPer];
#pragma unroll
    for (int u = 0; u < kDistPer; ++u) {
        const int64_t g = g0 + threadIdx.x + (int64_t)u * 256;
        jj[u] = -1;
        if (g < sv.nlf) {
            while (jb + 1 < sv.njobs && sv.lf_off[jb + 1] <= g) ++jb;  // items ascend with u
            const SJob& J = sv.jobs[jb];
            const int local = (int)(g - sv.lf_off[jb]);  // k * N < 2^31 (checked when the table is built)
            jj[u] = jb;
            rr[u] = local / J.N;
            ii[u] = local - rr[u] * J.N;
        }
    }
    int ee[kDistPer];
    double x[kDistPer], wv[kDistPer];
#pragma unroll
    for (int u = 0; u < kDistPer; ++u) {
        if (jj[u] < 0) continue;
        const SJob& J = sv.jobs[jj[u]];
        const int r = rr[u], i = ii[u];
        const size_t o = (size_t)r * J.N + i;
        ee[u] = J.sidx[i];
        wv[u] = J.w[i];
        if (MODE == kPreFromB)
            x[u] = (i >= J.brange[2 * r] && i < J.brange[2 * r + 1]) ? 1.0 : 0.0;
        else
            x[u] = J.in32 ? (double)static_cast<const float*>(J.in)[o] : static_cast<const double*>(J.in)[o];
    }
#pragma unroll
    for (int u = 0; u < kDistPer; ++u) {
        if (jj[u] < 0) continue;
        const SJob& J = sv.jobs[jj[u]];
        const int r = rr[u], i = ii[u], e = ee[u];
        const size_t o = (size_t)r * J.N + i;
        const double v = wv[u] * (J.chg[(size_t)r * J.m + e] + J.Sinv[(size_t)e * J.m + e] * x[u]);
        if (MODE == kPreAdd)
            J.out[o] += v;
        else
            J.out[o] = v;
    }
}

hipError_t launch_super_apply(const SView& sv, int mode, hipStream_t s) {
    if (sv.nel == 0) return hipSuccess;
    const unsigned gc = blocks_for((size_t)sv.nln, 256);
    if (mode == kPreFromB)
        k_super_collect<kPreFromB><<<gc, 256, 0, s>>>(sv);
    else
        k_super_collect<kPreSet><<<gc, 256, 0, s>>>(sv);
    k_super_mult<<<(unsigned)(sv.nem / 64), 256, 0, s>>>(sv);
    const unsigned gd = blocks_for((size_t)sv.nlf, 256 * kDistPer);
    if (mode == kPreFromB)
        k_super_dist<kPreFromB><<<gd, 256, 0, s>>>(sv);
    else if (mode == kPreAdd)
        k_super_dist<kPreAdd><<<gd, 256, 0, s>>>(sv);
    else
        k_super_dist<kPreSet><<<gd, 256, 0, s>>>(sv);
    return hipGetLastError();
}

}  // namespace fcg
