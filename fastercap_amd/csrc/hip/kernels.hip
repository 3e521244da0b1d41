// gfx950 kernels of the FasterCap BEM solve path.
//
//  fill        one link per lane: PotEstimateOpt(row node, source node)
//              (Solver/Autorefine.cpp:3474-3733, 2D :4156-4325), self terms
//              (SelfPotential, Solver/Autorefine.cpp:2358-2506)
//  product     v = P q (Solver/MultiplyHierarchical.cpp:72-330) in four stages:
//              1. double-double prefix scan of the (Jacobi-scaled) input over the
//                 DFS leaf order; every node's charge is the difference of two
//                 prefixes because each subtree owns a contiguous leaf range
//              2. node charges (leaves copy their own input exactly)
//              3. CSR gather: a group of lanes per destination node, coalesced
//                 coef/col streaming, source charges gathered from L2, lane
//                 shuffle reduction
//              4. down-pass: each leaf sums the potentials of its ancestor chain
//  krylov      batched dot products / updates over [k][N] blocks for CGS2 GMRES
#include "kernels.h"

#include "../common/potential.h"

namespace fcg {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return DD{s, e};
}
__device__ __forceinline__ DD dd_quick(double a, double b) {
    const double s = a + b;
    return DD{s, b - (s - a)};
}
// accurate double-double addition (two two-sums, two renormalisations)
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = dd_two_sum(a.hi, b.hi);
    DD t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_quick(s.hi, s.lo);
    s.lo += t.lo;
    return dd_quick(s.hi, s.lo);
}
__device__ __forceinline__ DD dd_neg(DD a) { return DD{-a.hi, -a.lo}; }
__device__ __forceinline__ double dd_diff(DD a, DD b) {
    const DD d = dd_add(a, dd_neg(b));
    return d.hi + d.lo;
}

// ------------------------------------------------------------------ fill
template <typename T>
__device__ T pot_estimate(const fcg_node_geom& p1, const fcg_node_geom& p2, int dim) {
    const bool diel1 = (p1.flags & FCG_NODE_DIEL) != 0;
    const T rmax = (T)(p1.maxside > p2.maxside ? p1.maxside : p2.maxside);
    if (dim == 2) {
        const T dx = (T)p1.c[0] - (T)p2.c[0], dy = (T)p1.c[1] - (T)p2.c[1];
        const T rd = sqrt(dx * dx + dy * dy);
        if (!diel1) {
            if (rd / rmax < (T)4.0) return potential_opt_2d<T>((T)p1.c[0], (T)p1.c[1], p2.v);
            return -log(rd);
        }
        if (rd / rmax < (T)4.0) return enfield_2d<T>((T)p1.c[0], (T)p1.c[1], p2.v, (T)p1.dn[0], (T)p1.dn[1]);
        return (dx * (T)p1.dn[0] + dy * (T)p1.dn[1]) / (rd * rd);
    }
    const Vec3<T> c1{(T)p1.c[0], (T)p1.c[1], (T)p1.c[2]};
    const Vec3<T> dist{c1.x - (T)p2.c[0], c1.y - (T)p2.c[1], c1.z - (T)p2.c[2]};
    const T rd = norm3(dist);
    const T ratio = rd / rmax;
    const T k4pe0 = (T)FCG_FOUR_PI_TIMES_E0;
    if (!diel1) {
        if (ratio < (T)1.0) return potential_opt_3d<T>(c1, p2.v) / k4pe0;
        if (ratio < (T)2.0) return potential_num_3d<T>(c1, p2.v, 5) / k4pe0;
        if (ratio < (T)4.0) return potential_num_3d<T>(c1, p2.v, 3) / k4pe0;
        if (ratio < (T)10.0) return potential_num_3d<T>(c1, p2.v, 2) / k4pe0;
        return (T)1 / (k4pe0 * rd);
    }
    const Vec3<T> n1{(T)p1.dn[0], (T)p1.dn[1], (T)p1.dn[2]};
    if (ratio < (T)1.0) {
        const T h = rd / (T)20;
        const Vec3<T> ep{c1.x + n1.x * h, c1.y + n1.y * h, c1.z + n1.z * h};
        const Vec3<T> em{c1.x - n1.x * h, c1.y - n1.y * h, c1.z - n1.z * h};
        const T pp = potential_opt_3d<T>(ep, p2.v);
        const T pn = potential_opt_3d<T>(em, p2.v);
        return (pn - pp) / ((T)2 * h * k4pe0);
    }
    if (ratio < (T)2.0) return enfield_num_3d<T>(c1, p2.v, n1, 5) / k4pe0;
    if (ratio < (T)4.0) return enfield_num_3d<T>(c1, p2.v, n1, 3) / k4pe0;
    if (ratio < (T)10.0) return enfield_num_3d<T>(c1, p2.v, n1, 2) / k4pe0;
    return dot3(dist, n1) / (k4pe0 * rd * rd * rd);
}

__global__ void k_link_rows(const int64_t* __restrict__ row_ptr, int M, int32_t* __restrict__ link_row) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int64_t b = row_ptr[i], e = row_ptr[i + 1];
    for (int64_t l = b; l < e; ++l) link_row[l] = i;
}

__global__ void __launch_bounds__(256) k_fill(const fcg_node_geom* __restrict__ nodes,
                                              const int32_t* __restrict__ link_row,
                                              const int32_t* __restrict__ col, int64_t L, int dim,
                                              double* __restrict__ coef) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    const int i = link_row[l];
    const int j = col[l];
    coef[l] = pot_estimate<double>(nodes[i], nodes[j], dim);
}

__global__ void k_self(const fcg_node_geom* __restrict__ nodes, const int32_t* __restrict__ leaf_node,
                       const double* __restrict__ gout, const double* __restrict__ gin, int N, int dim,
                       double* __restrict__ self, double* __restrict__ diag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const fcg_node_geom& g = nodes[leaf_node[k]];
    double s;
    if (g.flags & FCG_NODE_DIEL) {
        const double eb = gout[g.cond], ea = gin[g.cond];
        if (dim == 3)
            s = (ea + eb) / (FCG_TWO_TIMES_E0 * (eb - ea) * g.area);
        else
            s = FCG_PI * (ea + eb) / ((eb - ea) * g.area);
    } else if (dim == 3) {
        const Vec3<double> c{(double)g.c[0], (double)g.c[1], (double)g.c[2]};
        s = potential_opt_3d<double>(c, g.v) / FCG_FOUR_PI_TIMES_E0;
    } else {
        s = potential_opt_2d<double>((double)g.c[0], (double)g.c[1], g.v);
    }
    self[k] = s;
    diag[leaf_node[k]] = s;  // non-leaf entries stay zero
}

// ------------------------------------------------------------------ up-pass scan
// Block-level exclusive scan of one DD per thread (Hillis-Steele over LDS).
template <int NT>
__device__ DD block_exclusive_scan(DD v, DD* sh, DD* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        DD add = DD{0.0, 0.0};
        if (t >= off) add = sh[t - off];
        __syncthreads();
        if (t >= off) sh[t] = dd_add(sh[t], add);
        __syncthreads();
    }
    const DD excl = (t > 0) ? sh[t - 1] : DD{0.0, 0.0};
    *total = sh[NT - 1];
    __syncthreads();
    return excl;
}

constexpr int kScanThreads = 256;
constexpr int kScanPer = kScanTile / kScanThreads;  // 8

__device__ __forceinline__ double scan_input(const double* q, const double* self, int N, int r, int n, int jacobi) {
    const double x = q[(size_t)r * N + n];
    return jacobi ? x / self[n] : x;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const double* __restrict__ q,
                                                              const double* __restrict__ self, int N, int jacobi,
                                                              DD* __restrict__ block_sums, int nblk) {
    __shared__ DD sh[kScanThreads];
    const int r = blockIdx.y, b = blockIdx.x;
    const int base = b * kScanTile + threadIdx.x * kScanPer;
    DD acc{0.0, 0.0};
    for (int i = 0; i < kScanPer; ++i) {
        const int n = base + i;
        if (n < N) acc = dd_add(acc, DD{scan_input(q, self, N, r, n, jacobi), 0.0});
    }
    DD total;
    block_exclusive_scan<kScanThreads>(acc, sh, &total);
    if (threadIdx.x == 0) block_sums[(size_t)r * nblk + b] = total;
}

// exclusive scan of the block sums of every right-hand side (one block per rhs)
__global__ void __launch_bounds__(1024) k_scan_top(DD* __restrict__ block_sums, int nblk) {
    __shared__ DD sh[1024];
    const int r = blockIdx.x;
    DD* bs = block_sums + (size_t)r * nblk;
    const int per = (nblk + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    DD acc{0.0, 0.0};
    for (int i = 0; i < per; ++i)
        if (b0 + i < nblk) acc = dd_add(acc, bs[b0 + i]);
    DD total;
    DD off = block_exclusive_scan<1024>(acc, sh, &total);
    for (int i = 0; i < per; ++i) {
        if (b0 + i < nblk) {
            const DD v = bs[b0 + i];
            bs[b0 + i] = off;
            off = dd_add(off, v);
        }
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_final(const double* __restrict__ q,
                                                             const double* __restrict__ self, int N, int jacobi,
                                                             const DD* __restrict__ block_sums, int nblk,
                                                             DD* __restrict__ prefix) {
    __shared__ DD sh[kScanThreads];
    const int r = blockIdx.y, b = blockIdx.x;
    const int base = b * kScanTile + threadIdx.x * kScanPer;
    double x[kScanPer];
    DD acc{0.0, 0.0};
    for (int i = 0; i < kScanPer; ++i) {
        const int n = base + i;
        x[i] = (n < N) ? scan_input(q, self, N, r, n, jacobi) : 0.0;
        acc = dd_add(acc, DD{x[i], 0.0});
    }
    DD total;
    DD run = block_exclusive_scan<kScanThreads>(acc, sh, &total);
    run = dd_add(run, block_sums[(size_t)r * nblk + b]);
    DD* P = prefix + (size_t)r * (N + 1);
    for (int i = 0; i < kScanPer; ++i) {
        const int n = base + i;
        if (n < N) P[n] = run;
        run = dd_add(run, DD{x[i], 0.0});
        if (n == N - 1) P[N] = run;
    }
}

// node charges, node-major [M][k]
__global__ void k_node_charge(const int32_t* __restrict__ left, const int32_t* __restrict__ leaf_lo,
                              const int32_t* __restrict__ leaf_hi, const DD* __restrict__ prefix,
                              const double* __restrict__ q, const double* __restrict__ self, int N, int M, int k,
                              int jacobi, double* __restrict__ charge) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int lo = leaf_lo[i], hi = leaf_hi[i];
    const bool leaf = left[i] < 0;
    for (int r = 0; r < k; ++r) {
        double c;
        if (leaf) {
            c = scan_input(q, self, N, r, lo, jacobi);
        } else {
            const DD* P = prefix + (size_t)r * (N + 1);
            c = dd_diff(P[hi], P[lo]);
        }
        charge[(size_t)i * k + r] = c;
    }
}

// ------------------------------------------------------------------ gather
// One wavefront per destination node.  Lane = (slot, rhs): KR lanes share one link
// slot and each reads its own right-hand side, so a link's KR source charges are
// one contiguous [k]-row of the node-major charge array (one or two cache lines
// per slot instead of one per lane); the 64/KR slots walk the row's links in
// stride, and a shuffle tree over the slots reduces each right-hand side.
// Blocks are remapped so that each of the 8 XCDs owns one contiguous range of
// destination rows: post-order rows are spatially coherent, so the charges one
// XCD gathers stay resident in its own L2.
constexpr int kNumXcd = 8;
constexpr int kGatherWaves = 4;  // rows per 256-thread block

// KR lanes per link slot, VR right-hand sides per lane (KR*VR >= the rhs chunk),
// SL = 64/KR slots, U links per slot in flight per iteration.
template <int KR, int VR, int U>
__global__ void __launch_bounds__(256) k_gather(const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                                const double* __restrict__ coef, const double* __restrict__ diag,
                                                const double* __restrict__ charge, int M, int k, int r0,
                                                int blocks_per_xcd, double* __restrict__ pot) {
    constexpr int SL = 64 / KR;
    const int b = (int)(blockIdx.x % kNumXcd) * blocks_per_xcd + (int)(blockIdx.x / kNumXcd);
    const int row = b * kGatherWaves + (int)(threadIdx.x >> 6);
    if (row >= M) return;
    const int lane = threadIdx.x & 63;
    const int slot = lane / KR;
    const int rb = (lane % KR) * VR;  // first rhs (relative to r0) of this lane
    const int kc = min(k - r0, KR * VR);
    // every load that does not depend on the link list is issued up front
    const int64_t beg = row_ptr[row], end = row_ptr[row + 1];
    const double sf = diag[row];
    double own[VR];
#pragma unroll
    for (int v = 0; v < VR; ++v) own[v] = (slot == 0 && rb + v < kc) ? charge[(size_t)row * k + r0 + rb + v] : 0.0;
    const double* src = charge + r0 + rb;
    double acc[VR];
#pragma unroll
    for (int v = 0; v < VR; ++v) acc[v] = 0.0;
    for (int64_t l = beg + slot; l < end; l += U * SL) {
        int c[U];
        double w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t ll = l + (int64_t)u * SL;
            const bool ok = ll < end;
            c[u] = ok ? col[ll] : 0;
            w[u] = ok ? coef[ll] : 0.0;
        }
        double ch[U][VR];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < VR; ++v) ch[u][v] = (rb + v < kc) ? src[(size_t)c[u] * k + v] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < VR; ++v) acc[v] += w[u] * ch[u][v];
    }
#pragma unroll
    for (int v = 0; v < VR; ++v)
#pragma unroll
        for (int off = KR; off < 64; off <<= 1) acc[v] += __shfl_xor(acc[v], off, 64);
    if (slot == 0) {
#pragma unroll
        for (int v = 0; v < VR; ++v)
            if (rb + v < kc) pot[(size_t)row * k + r0 + rb + v] = sf * own[v] + acc[v];
    }
}

// Gather variants: (KR, VR, U) triples per rhs-chunk width; gather_variant picks one.
struct GatherCfg {
    int kr, vr, u;
};
static const GatherCfg kGatherCfgs[] = {
    {8, 1, 2}, {4, 2, 2}, {4, 2, 4}, {8, 1, 4}, {2, 4, 2}, {16, 1, 2}, {1, 8, 2}, {2, 4, 4},
};
constexpr int kNumGatherCfgs = sizeof(kGatherCfgs) / sizeof(kGatherCfgs[0]);

int gather_variant_count() { return kNumGatherCfgs; }

template <int KR, int VR, int U>
static void launch_gather_t(const JobView& j, const MvpScratch& sc, int k, int r0, unsigned grid, int per_xcd,
                            hipStream_t s) {
    k_gather<KR, VR, U><<<grid, 256, 0, s>>>(j.row_ptr, j.col, j.coef, j.diag, sc.charge, j.M, k, r0, per_xcd,
                                             sc.pot);
}

// width = KR*VR rhs per pass
static int gather_width(int variant, int k) {
    const GatherCfg& c = kGatherCfgs[variant];
    int w = c.kr * c.vr;
    // narrow rhs counts: shrink the lane group so padding lanes stay few
    while (w / 2 >= k && w > c.vr) w /= 2;
    return w;
}

// ------------------------------------------------------------------ down-pass
template <int KC>
__global__ void k_down(const int32_t* __restrict__ leaf_node, const int32_t* __restrict__ parent,
                       const double* __restrict__ pot, int N, int k, int r0, double* __restrict__ v) {
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= N) return;
    const int kc = (k - r0) < KC ? (k - r0) : KC;
    double acc[KC];
#pragma unroll
    for (int r = 0; r < KC; ++r) acc[r] = 0.0;
    int a = leaf_node[leaf];
    while (a >= 0) {
        const double* p = pot + (size_t)a * k + r0;
#pragma unroll
        for (int r = 0; r < KC; ++r)
            if (r < kc) acc[r] += p[r];
        a = parent[a];
    }
#pragma unroll
    for (int r = 0; r < KC; ++r)
        if (r < kc) v[(size_t)(r0 + r) * N + leaf] = acc[r];
}

__global__ void k_2d_last(const double* __restrict__ v, const DD* __restrict__ prefix, int N, int k,
                          double* __restrict__ last_total) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= k) return;
    last_total[2 * r] = v[(size_t)r * N + N - 1];
    const DD t = prefix[(size_t)r * (N + 1) + N];
    last_total[2 * r + 1] = t.hi + t.lo;
}

__global__ void k_2d_fix(double* __restrict__ v, const double* __restrict__ last_total, int N, int first) {
    const int r = blockIdx.y;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double* vr = v + (size_t)r * N;
    if (n >= first && n < N - 1) vr[n] -= last_total[2 * r];
    if (n == N - 1) vr[n] = 1.0 * last_total[2 * r + 1];
}

// ------------------------------------------------------------------ krylov helpers
__global__ void k_jacobi(const double* __restrict__ q, const double* __restrict__ self, int N, int k,
                         double* __restrict__ pq) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * k) return;
    pq[t] = q[t] / self[t % N];
}

constexpr int kDotThreads = 256;
constexpr int kDotChunk = 4096;

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

// grid: (chunks, nvec*k). partial[(jj*k + r) * nchunks + chunk]
__global__ void __launch_bounds__(kDotThreads) k_block_dots(double* const* __restrict__ basis,
                                                            const double* __restrict__ w, int N, int k,
                                                            double* __restrict__ partial, int nchunks) {
    __shared__ double sh[kDotThreads / 64];
    const int pair = blockIdx.y;
    const int jj = pair / k, r = pair % k;
    const double* q = basis[jj] + (size_t)r * N;
    const double* wr = w + (size_t)r * N;
    const int n0 = blockIdx.x * kDotChunk;
    const int n1 = min(N, n0 + kDotChunk);
    double s = 0.0;
    for (int n = n0 + threadIdx.x; n < n1; n += kDotThreads) s += q[n] * wr[n];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(size_t)pair * nchunks + blockIdx.x] = s;
}

__global__ void k_reduce_partials(const double* __restrict__ partial, int nchunks, int nvals,
                                  double* __restrict__ out) {
    __shared__ double sh[4];
    const int v = blockIdx.x;
    if (v >= nvals) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < nchunks; i += blockDim.x) s += partial[(size_t)v * nchunks + i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[v] = s;
}

// w[r][n] -= sum_jj h[jj][r] Q_jj[r][n] on one chunk of one rhs per block, and the
// chunk's partial sum of the updated w^2 (grid (nchunks, k)).
__global__ void __launch_bounds__(kDotThreads) k_basis_update_norm(double* const* __restrict__ basis, int nvec,
                                                                   const double* __restrict__ h,
                                                                   double* __restrict__ w, int N, int k,
                                                                   double* __restrict__ partial, int nchunks) {
    __shared__ double sh[kDotThreads / 64];
    const int r = blockIdx.y;
    const int n0 = blockIdx.x * kDotChunk;
    const int n1 = min(N, n0 + kDotChunk);
    double ss = 0.0;
    for (int n = n0 + threadIdx.x; n < n1; n += kDotThreads) {
        const size_t t = (size_t)r * N + n;
        double acc = w[t];
        for (int jj = 0; jj < nvec; ++jj) acc -= h[(size_t)jj * k + r] * basis[jj][t];
        w[t] = acc;
        ss += acc * acc;
    }
    ss = block_sum(ss, sh);
    if (threadIdx.x == 0) partial[(size_t)r * nchunks + blockIdx.x] = ss;
}

// h layout [nvec][k]
__global__ void k_basis_update(double* const* __restrict__ basis, int nvec, const double* __restrict__ h,
                               double* __restrict__ w, int N, int k) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * k) return;
    const int r = (int)(t / N);
    double acc = w[t];
    for (int jj = 0; jj < nvec; ++jj) acc -= h[(size_t)jj * k + r] * basis[jj][t];
    w[t] = acc;
}

__global__ void k_basis_combine(double* const* __restrict__ basis, int nvec, const double* __restrict__ z,
                                double* __restrict__ out, int N, int k) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * k) return;
    const int r = (int)(t / N);
    double acc = 0.0;
    for (int jj = 0; jj < nvec; ++jj) {
        const double zz = z[(size_t)jj * k + r];
        if (zz != 0.0) acc += basis[jj][t] * zz;
    }
    out[t] += acc;
}

__global__ void k_scale(const double* __restrict__ a, const double* __restrict__ d, double* __restrict__ out, int N,
                        int k) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * k) return;
    const double den = d[t / N];
    out[t] = den != 0.0 ? a[t] / den : 0.0;
}

__global__ void k_add(const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ out,
                      size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = x[t] + y[t];
}

__global__ void k_sub(const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ out,
                      size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = x[t] - y[t];
}

// grid (ncol, k): one block per (rhs, conductor column); column c owns leaves
// [col_range[2c], col_range[2c+1])
__global__ void k_charge_sums(const double* __restrict__ x, const double* __restrict__ wgt,
                              const int32_t* __restrict__ col_range, int N, int k, int ncol,
                              double* __restrict__ out) {
    __shared__ double sh[4];
    const int c = blockIdx.x, r = blockIdx.y;
    double s = 0.0;
    for (int n = col_range[2 * c] + threadIdx.x; n < col_range[2 * c + 1]; n += blockDim.x)
        s += wgt[n] * x[(size_t)r * N + n];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[(size_t)r * ncol + c] = s;
}

// ------------------------------------------------------------------ launchers
static inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_link_rows(const JobView& j, int32_t* link_row, hipStream_t s) {
    if (j.M == 0) return hipSuccess;
    k_link_rows<<<blocks_for(j.M, 256), 256, 0, s>>>(j.row_ptr, j.M, link_row);
    return hipGetLastError();
}

hipError_t launch_fill(const JobView& j, const int32_t* link_row, hipStream_t s) {
    if (j.L == 0) return hipSuccess;
    k_fill<<<blocks_for((size_t)j.L, 256), 256, 0, s>>>(j.nodes, link_row, j.col, j.L, j.dim, j.coef);
    return hipGetLastError();
}

hipError_t launch_self(const JobView& j, hipStream_t s) {
    k_self<<<blocks_for(j.N, 256), 256, 0, s>>>(j.nodes, j.leaf_node, j.group_outperm, j.group_inperm, j.N, j.dim,
                                              j.self, j.diag);
    return hipGetLastError();
}

static hipError_t gather_dispatch(const JobView& j, const MvpScratch& sc, int k, hipStream_t s) {
    const int rows_blocks = (j.M + kGatherWaves - 1) / kGatherWaves;
    const int per_xcd = (rows_blocks + kNumXcd - 1) / kNumXcd;
    const unsigned grid = (unsigned)(per_xcd * kNumXcd);
    const int variant = (j.gather_variant >= 0 && j.gather_variant < kNumGatherCfgs) ? j.gather_variant : 0;
    const GatherCfg& c = kGatherCfgs[variant];
    const int width = gather_width(variant, k);
    const int kr = width / c.vr;
    for (int r0 = 0; r0 < k; r0 += width) {
        // dispatch on (kr, vr, u)
#define FCG_G(KR, VR, U)                                          \
    if (kr == KR && c.vr == VR && c.u == U) {                      \
        launch_gather_t<KR, VR, U>(j, sc, k, r0, grid, per_xcd, s); \
        continue;                                                   \
    }
        FCG_G(1, 1, 2) FCG_G(2, 1, 2) FCG_G(4, 1, 2) FCG_G(8, 1, 2) FCG_G(16, 1, 2) FCG_G(32, 1, 2) FCG_G(64, 1, 2)
        FCG_G(1, 1, 4) FCG_G(2, 1, 4) FCG_G(4, 1, 4) FCG_G(8, 1, 4) FCG_G(16, 1, 4) FCG_G(32, 1, 4) FCG_G(64, 1, 4)
        FCG_G(1, 2, 2) FCG_G(2, 2, 2) FCG_G(4, 2, 2) FCG_G(8, 2, 2) FCG_G(16, 2, 2) FCG_G(32, 2, 2) FCG_G(64, 2, 2)
        FCG_G(1, 2, 4) FCG_G(2, 2, 4) FCG_G(4, 2, 4) FCG_G(8, 2, 4) FCG_G(16, 2, 4) FCG_G(32, 2, 4) FCG_G(64, 2, 4)
        FCG_G(1, 4, 2) FCG_G(2, 4, 2) FCG_G(4, 4, 2) FCG_G(8, 4, 2) FCG_G(16, 4, 2) FCG_G(32, 4, 2) FCG_G(64, 4, 2)
        FCG_G(1, 4, 4) FCG_G(2, 4, 4) FCG_G(4, 4, 4) FCG_G(8, 4, 4) FCG_G(16, 4, 4) FCG_G(32, 4, 4) FCG_G(64, 4, 4)
        FCG_G(1, 8, 2) FCG_G(2, 8, 2) FCG_G(4, 8, 2) FCG_G(8, 8, 2) FCG_G(16, 8, 2) FCG_G(32, 8, 2) FCG_G(64, 8, 2)
#undef FCG_G
        return hipErrorInvalidConfiguration;  // no instantiation for this (kr, vr, u)
    }
    return hipSuccess;
}

hipError_t launch_mvp(const JobView& j, MvpScratch& sc, const double* q, double* v, int k, int jacobi,
                      hipStream_t s, hipEvent_t* ev) {
    const int N = j.N, M = j.M;
    const int nblk = (N + kScanTile - 1) / kScanTile;
    if (k > sc.kmax || nblk > sc.nblk) return hipErrorInvalidValue;
    if (ev) (void)hipEventRecord(ev[0], s);
    k_scan_reduce<<<dim3(nblk, k), kScanThreads, 0, s>>>(q, j.self, N, jacobi, sc.block_sums, nblk);
    k_scan_top<<<k, 1024, 0, s>>>(sc.block_sums, nblk);
    k_scan_final<<<dim3(nblk, k), kScanThreads, 0, s>>>(q, j.self, N, jacobi, sc.block_sums, nblk, sc.prefix);
    if (ev) (void)hipEventRecord(ev[1], s);
    k_node_charge<<<blocks_for(M, 256), 256, 0, s>>>(j.left, j.leaf_lo, j.leaf_hi, sc.prefix, q, j.self, N, M, k,
                                                     jacobi, sc.charge);
    if (ev) (void)hipEventRecord(ev[2], s);
    {
        const hipError_t ge = gather_dispatch(j, sc, k, s);
        if (ge != hipSuccess) return ge;
    }
    if (ev) (void)hipEventRecord(ev[3], s);
    for (int r0 = 0; r0 < k; r0 += 8) {
        if (k - r0 <= 4)
            k_down<4><<<blocks_for(N, 256), 256, 0, s>>>(j.leaf_node, j.parent, sc.pot, N, k, r0, v);
        else
            k_down<8><<<blocks_for(N, 256), 256, 0, s>>>(j.leaf_node, j.parent, sc.pot, N, k, r0, v);
    }
    if (j.dim == 2) {
        // reuse the block-sum area for the per-rhs (last, total) pairs
        double* lt = reinterpret_cast<double*>(sc.block_sums);
        k_2d_last<<<blocks_for(k, 64), 64, 0, s>>>(v, sc.prefix, N, k, lt);
        k_2d_fix<<<dim3(blocks_for(N, 256), k), 256, 0, s>>>(v, lt, N, j.first_cond_leaf);
    }
    if (ev) (void)hipEventRecord(ev[4], s);
    return hipGetLastError();
}

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

hipError_t launch_jacobi(const JobView& j, const double* q, double* pq, int k, hipStream_t s) {
    k_jacobi<<<blocks_for((size_t)j.N * k, 256), 256, 0, s>>>(q, j.self, j.N, k, pq);
    return hipGetLastError();
}

int dots_blocks(int N) { return (N + kDotChunk - 1) / kDotChunk; }

hipError_t launch_block_dots(double* const* basis, int nvec, const double* w, int N, int k, double* partial,
                             int nchunks, hipStream_t s) {
    k_block_dots<<<dim3(nchunks, nvec * k), kDotThreads, 0, s>>>(basis, w, N, k, partial, nchunks);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* partial, int nchunks, int nvals, double* out, hipStream_t s) {
    k_reduce_partials<<<nvals, 256, 0, s>>>(partial, nchunks, nvals, out);
    return hipGetLastError();
}

hipError_t launch_basis_update(double* const* basis, int nvec, const double* h, double* w, int N, int k,
                               hipStream_t s) {
    k_basis_update<<<blocks_for((size_t)N * k, 256), 256, 0, s>>>(basis, nvec, h, w, N, k);
    return hipGetLastError();
}

hipError_t launch_basis_update_norm(double* const* basis, int nvec, const double* h, double* w, int N, int k,
                                   double* partial, int nchunks, hipStream_t s) {
    k_basis_update_norm<<<dim3(nchunks, k), kDotThreads, 0, s>>>(basis, nvec, h, w, N, k, partial, nchunks);
    return hipGetLastError();
}

hipError_t launch_basis_combine(double* const* basis, int nvec, const double* z, double* out, int N, int k,
                                hipStream_t s) {
    k_basis_combine<<<blocks_for((size_t)N * k, 256), 256, 0, s>>>(basis, nvec, z, out, N, k);
    return hipGetLastError();
}

hipError_t launch_scale(const double* a, const double* scale, double* out, int N, int k, hipStream_t s) {
    k_scale<<<blocks_for((size_t)N * k, 256), 256, 0, s>>>(a, scale, out, N, k);
    return hipGetLastError();
}

hipError_t launch_add(const double* x, const double* y, double* out, int N, int k, hipStream_t s) {
    k_add<<<blocks_for((size_t)N * k, 256), 256, 0, s>>>(x, y, out, (size_t)N * k);
    return hipGetLastError();
}

hipError_t launch_sub(const double* x, const double* y, double* out, int N, int k, hipStream_t s) {
    k_sub<<<blocks_for((size_t)N * k, 256), 256, 0, s>>>(x, y, out, (size_t)N * k);
    return hipGetLastError();
}

hipError_t launch_charge_sums(const double* x, const double* leaf_weight, const int32_t* col_range, int N, int k,
                              int ncol, double* out, hipStream_t s) {
    k_charge_sums<<<dim3(ncol, k), 256, 0, s>>>(x, leaf_weight, col_range, N, k, ncol, out);
    return hipGetLastError();
}

}  // namespace fcg
