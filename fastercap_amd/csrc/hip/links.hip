// Link discovery on the device: the dual-tree recursion of CAutoRefine::ComputeLinks
// (Solver/Autorefine.cpp:1132-1171) with RefineMutual (:3096-3235 3D, :3239-3447
// 2D), RefineSelf (:2115-2229, structure part) and RefineCriteria (:2928-3003 3D,
// :3007-3092 2D), followed by the compressed-row build of the link record pass of
// AutoRefineLinks (:481-756).
//
// The recursion becomes a work stack of node pairs processed in slices: a thread
// takes one pair, evaluates the refinement criterion and either records a link or
// pushes the two child pairs (a self item pushes the pair of its children and the
// two children's self items).  Every decision must equal the host builder's, so
// this file is compiled without floating-point contraction (no fused
// multiply-add), evaluates every expression in the host's order and uses the
// same quadrature routines.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "../common/potential.h"
#include "kernels.h"

namespace fcg {
namespace {

struct D3 {
    double x, y, z;
};
__device__ inline D3 centroid_d(const fcg_node_geom& g) { return D3{(double)g.c[0], (double)g.c[1], (double)g.c[2]}; }
__device__ inline double dot_d(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// The fields of a node that RefineCriteria reads.
struct CritNode {
    D3 c;            // centroid (from float storage)
    const float* v;  // vertices
    double r;        // 3D: longest side; 2D: length
    double dim;      // area (3D) or length (2D)
    bool diel;
    D3 dn;           // dielectric normal (GetDielNormal)
};
__device__ inline CritNode crit_node(const fcg_node_geom& g) {
    return CritNode{centroid_d(g), g.v, g.maxside, g.area, (g.flags & FCG_NODE_DIEL) != 0,
                    D3{g.dn[0], g.dn[1], g.dn[2]}};
}

// RefineCriteria (3D and 2D) with curvature coefficient ccoeff: e1 / e2 receive
// the two relative potential estimates; true when the pair must be refined.
__device__ bool refine_test(const CritNode& p1, const CritNode& p2, int dim, double eps, double ccoeff,
                            double max_side, double max_area, double max_length, double* e1o, double* e2o) {
    const D3 c1 = p1.c, c2 = p2.c;
    const D3 dist{c1.x - c2.x, c1.y - c2.y, c1.z - c2.z};
    const double rdist = sqrt(dot_d(dist, dist));
    const double r1 = p1.r, r2 = p2.r;
    const double rmax = (r1 > r2) ? r1 : r2;
    const bool d1 = p1.diel, d2 = p2.diel;
    const D3 n1 = p1.dn, n2 = p2.dn;
    const D3 mdist{-dist.x, -dist.y, -dist.z};
    double pot1, pot2, e1, e2;
    if (dim == 3) {
        if (rdist / rmax < 2.0) {
            const Vec3<double> r1v{c1.x, c1.y, c1.z}, r2v{c2.x, c2.y, c2.z};
            pot1 = d1 ? enfield_num_3d<double>(r1v, p2.v, Vec3<double>{n1.x, n1.y, n1.z}, 2)
                      : potential_num_3d<double>(r1v, p2.v, 2);
            pot2 = d2 ? enfield_num_3d<double>(r2v, p1.v, Vec3<double>{n2.x, n2.y, n2.z}, 2)
                      : potential_num_3d<double>(r2v, p1.v, 2);
        } else {
            pot1 = d1 ? dot_d(dist, n1) / (rdist * rdist * rdist) : 1.0 / rdist;
            pot2 = d2 ? dot_d(mdist, n2) / (rdist * rdist * rdist) : 1.0 / rdist;
        }
        e1 = d1 ? fabs(pot1 * p2.dim * p2.r / (max_side * max_area * ccoeff))
                : fabs(pot1 * p2.dim / (max_area * ccoeff));
        e2 = d2 ? fabs(pot2 * p1.dim * p1.r / (max_side * max_area * ccoeff))
                : fabs(pot2 * p1.dim / (max_area * ccoeff));
    } else {
        if (rdist / rmax < 2.0) {
            pot1 = d1 ? enfield_2d<double>(c1.x, c1.y, p2.v, n1.x, n1.y) : potential_opt_2d<double>(c1.x, c1.y, p2.v);
            pot2 = d2 ? enfield_2d<double>(c2.x, c2.y, p1.v, n2.x, n2.y) : potential_opt_2d<double>(c2.x, c2.y, p1.v);
        } else {
            pot1 = d1 ? dot_d(dist, n1) / (rdist * rdist) : -log(rdist);
            pot2 = d2 ? dot_d(mdist, n2) / (rdist * rdist) : -log(rdist);
        }
        e1 = d1 ? fabs(pot1 * p2.dim * p2.dim / (max_length * max_length * ccoeff))
                : fabs((pot1 + 0.0) * p2.dim / (max_length * ccoeff));
        e2 = d2 ? fabs(pot2 * p1.dim * p1.dim / (max_length * max_length * ccoeff))
                : fabs((pot2 + 0.0) * p1.dim / (max_length * ccoeff));
    }
    *e1o = e1;
    *e2o = e2;
    return e1 > eps || e2 > eps;
}

// the link pass: curvature coefficient 1
__device__ bool link_refine(const fcg_node_geom& p1, const fcg_node_geom& p2, const LinkParams& P) {
    double e1, e2;
    return refine_test(crit_node(p1), crit_node(p2), P.dim, P.eps, 1.0, P.max_side, P.max_area, P.max_length, &e1,
                       &e2);
}

// Reserve cnt slots per lane from a global counter with one atomic per
// wavefront; returns the lane's first slot.  All 64 lanes must call it.
__device__ unsigned long long wave_reserve(unsigned long long* ctr, unsigned cnt) {
    const int lane = threadIdx.x & 63;
    unsigned incl = cnt;
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    const unsigned total = __shfl(incl, 63, 64);
    unsigned long long base = 0;
    if (lane == 63 && total) base = atomicAdd(ctr, (unsigned long long)total);
    base = __shfl(base, 63, 64);
    return base + (incl - cnt);
}

constexpr uint64_t kSelfBit = 1ull << 63;
__device__ inline uint64_t pair_item(int32_t a, int32_t b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; }

// One slice of the work stack.  out receives at most 3 items per input, links
// at most one pair per input (the caller sizes both).  Every item carries its
// recursion level (in_lev / out_lev: RefineMutual and RefineSelf count one per
// call); max_lev receives the deepest.
__global__ void __launch_bounds__(256) k_links_step(const uint64_t* __restrict__ in,
                                                    const uint8_t* __restrict__ in_lev, int64_t n,
                                                    uint64_t* __restrict__ out, uint8_t* __restrict__ out_lev,
                                                    unsigned long long* out_cnt, int2* __restrict__ links,
                                                    unsigned long long* link_cnt, int* max_lev,
                                                    const fcg_node_geom* __restrict__ nodes,
                                                    const int32_t* __restrict__ left,
                                                    const int32_t* __restrict__ right, LinkParams P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t e0 = 0, e1 = 0, e2 = 0;
    unsigned ne = 0;
    int2 lk{0, 0};
    unsigned nl = 0;
    int lev = 0;
    if (i < n) {
        lev = in_lev[i];
        const uint64_t it = in[i];
        const int32_t a = (int32_t)((it >> 32) & 0x7fffffffu), b = (int32_t)(uint32_t)it;
        if (it & kSelfBit) {
            const int32_t l = left[a];
            if (l >= 0) {
                const int32_t r = right[a];
                e0 = pair_item(l, r);
                e1 = kSelfBit | pair_item(l, 0);
                e2 = kSelfBit | pair_item(r, 0);
                ne = 3;
            }
        } else {
            const int32_t l1 = left[a], l2 = left[b];
            bool force = false;
            if (l1 >= 0 || l2 >= 0) force = link_refine(nodes[a], nodes[b], P);
            if (force) {
                int which;
                if (l1 >= 0 && l2 < 0)
                    which = 1;
                else if (l1 < 0 && l2 >= 0)
                    which = 2;
                else if (P.dim == 3)
                    which = (nodes[a].area > nodes[b].area) ? 1 : 2;
                else
                    which = (nodes[a].maxside > nodes[b].maxside) ? 1 : 2;
                if (which == 1) {
                    e0 = pair_item(l1, b);
                    e1 = pair_item(right[a], b);
                } else {
                    e0 = pair_item(a, l2);
                    e1 = pair_item(a, right[b]);
                }
                ne = 2;
            } else {
                lk = int2{a, b};
                nl = 1;
            }
        }
    }
    const unsigned long long eo = wave_reserve(out_cnt, ne);
    const uint8_t cl = (uint8_t)(lev < 255 ? lev + 1 : 255);
    if (ne > 0) {
        out[eo] = e0;
        out_lev[eo] = cl;
    }
    if (ne > 1) {
        out[eo + 1] = e1;
        out_lev[eo + 1] = cl;
    }
    if (ne > 2) {
        out[eo + 2] = e2;
        out_lev[eo + 2] = cl;
    }
    const unsigned long long lo = wave_reserve(link_cnt, nl);
    if (nl) links[lo] = lk;
    for (int d = 32; d > 0; d >>= 1) lev = max(lev, __shfl_xor(lev, d, 64));
    if ((threadIdx.x & 63) == 0 && lev > 0) atomicMax(max_lev, lev);
}

// entries per node: a link (a, b) is an entry of row a (source b) and of row b
__global__ void k_links_degree(const int2* __restrict__ links, int64_t P, int32_t* __restrict__ deg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int2 l = links[i];
    atomicAdd(&deg[l.x], 1);
    atomicAdd(&deg[l.y], 1);
}

__global__ void k_links_scatter(const int2* __restrict__ links, int64_t P, unsigned long long* __restrict__ cursor,
                                int32_t* __restrict__ col) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int2 l = links[i];
    col[atomicAdd(&cursor[l.x], 1ull)] = l.y;
    col[atomicAdd(&cursor[l.y], 1ull)] = l.x;
}

// Sort every row of at most kSortLds entries ascending: one workgroup per row,
// bitonic network in LDS over the row padded to a power of two.
constexpr int kSortLds = 4096;
__global__ void __launch_bounds__(256) k_links_sort_rows(const int64_t* __restrict__ row_ptr, int32_t* __restrict__ col,
                                                         int M) {
    __shared__ int32_t s[kSortLds];
    const int row = blockIdx.x;
    if (row >= M) return;
    const int64_t b = row_ptr[row];
    const int n = (int)(row_ptr[row + 1] - b);
    if (n < 2 || n > kSortLds) return;
    int P = 2;
    while (P < n) P <<= 1;
    for (int i = threadIdx.x; i < P; i += 256) s[i] = i < n ? col[b + i] : INT32_MAX;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += 256) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const int32_t x = s[i], y = s[ixj];
                    if (((i & k) == 0) ? (x > y) : (x < y)) {
                        s[i] = y;
                        s[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += 256) col[b + i] = s[i];
}

// Rows longer than kSortLds: one workgroup per row, the same bitonic network on
// a scratch copy of the row padded to a power of two (scr_off: each listed
// row's scratch offset).
__global__ void __launch_bounds__(1024) k_links_sort_long(const int64_t* __restrict__ row_ptr,
                                                          const int32_t* __restrict__ rows,
                                                          const int64_t* __restrict__ scr_off,
                                                          int32_t* __restrict__ col, int32_t* __restrict__ scratch) {
    const int r = rows[blockIdx.x];
    const int64_t b = row_ptr[r];
    const int n = (int)(row_ptr[r + 1] - b);
    int P = 2;
    while (P < n) P <<= 1;
    int32_t* S = scratch + scr_off[blockIdx.x];
    for (int i = threadIdx.x; i < P; i += 1024) S[i] = i < n ? col[b + i] : INT32_MAX;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += 1024) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const int32_t x = S[i], y = S[ixj];
                    if (((i & k) == 0) ? (x > y) : (x < y)) {
                        S[i] = y;
                        S[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += 1024) col[b + i] = S[i];
}

// ---------------------------------------------------------------- refinement
// CAutoRefine::DiscretizeMutual / DiscretizeSelf (Solver/Autorefine.cpp:2638-2921)
// with CAutoPanel::Subdivide / CAutoSegment::Subdivide (AutoPanel.cpp:382-523,
// AutoSegment.cpp:152-230) on a growing node pool.

__device__ inline D3 f3d(const float* f) { return D3{(double)f[0], (double)f[1], (double)f[2]}; }
__device__ inline void d3f(float* f, D3 a) {
    f[0] = (float)a.x;
    f[1] = (float)a.y;
    f[2] = (float)a.z;
}
__device__ inline double mod_d(D3 a) { return sqrt(dot_d(a, a)); }
__device__ inline D3 sub_d(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 add_d(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 div_d(D3 a, double s) { return D3{a.x / s, a.y / s, a.z / s}; }
constexpr double kPanelEpsD = 1E-12;  // AUTOPANEL_EPS, AutoElement.h:58

// GetDielNormal (AutoPanel.h:109-134, AutoSegment.h:107-133)
__device__ D3 diel_normal(const fcg_refine_node& e) {
    D3 u;
    if (e.type & FCG_NODE_SUPER) {
        const double m = mod_d(f3d(e.n));
        u = (m < kPanelEpsD) ? D3{0, 0, 0} : div_d(f3d(e.n), m);
    } else {
        u = f3d(e.n);
    }
    if ((e.type & FCG_NODE_OUTPERM_NORMAL_DIR) != FCG_NODE_OUTPERM_NORMAL_DIR) u = D3{-u.x, -u.y, -u.z};
    return u;
}

// GetGeoNormal (AutoPanel.h:95-107, AutoSegment.h:92-105): a super node's normal
// from its vertices (Newell's method accumulated in float storage), else the
// stored one
__device__ D3 geo_normal(const fcg_refine_node& e, int dim) {
    if (!(e.type & FCG_NODE_SUPER)) return f3d(e.n);
    float u[3] = {0.f, 0.f, 0.f};
    if (dim == 3) {
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1 > 2) ? 0 : i + 1;
            const D3 a = f3d(e.v + 3 * i), b = f3d(e.v + 3 * j);
            u[0] = (float)((double)u[0] + (a.y - b.y) * (a.z + b.z));
            u[1] = (float)((double)u[1] + (a.z - b.z) * (a.x + b.x));
            u[2] = (float)((double)u[2] + (a.x - b.x) * (a.y + b.y));
        }
    } else {
        u[0] = (float)((double)e.v[1] - (double)e.v[4]);
        u[1] = (float)((double)e.v[3] - (double)e.v[0]);
        u[2] = 0.f;
    }
    const double nm = mod_d(f3d(u));
    if (nm > kPanelEpsD) d3f(u, div_d(f3d(u), nm));
    return f3d(u);
}

__device__ inline CritNode crit_refine(const fcg_refine_node& e, int dim) {
    return CritNode{f3d(e.c), e.v, dim == 3 ? e.maxside : e.dim, e.dim, (e.type & FCG_NODE_DIEL) != 0,
                    diel_normal(e)};
}

// CAutoPanel::MaxSide (AutoPanel.cpp:91-131)
__device__ void max_side3(fcg_refine_node& e, double l0, double l1, double l2) {
    if (l0 >= l1 && l0 >= l2) {
        e.max_side_idx = 0;
        e.maxside = l0;
    } else if (l1 >= l0 && l1 >= l2) {
        e.max_side_idx = 1;
        e.maxside = l1;
    } else {
        e.max_side_idx = 2;
        e.maxside = l2;
    }
}

// split leaf p into nodes[slot] and nodes[slot + 1] (longest-side bisection in 3D,
// halving in 2D), then link them under p
__device__ void split_node(fcg_refine_node* nodes, int p, int slot, int dim) {
    fcg_refine_node P = nodes[p];
    fcg_refine_node L = P, R = P;
    L.left = L.right = R.left = R.right = -1;
    if (dim == 3) {
        const D3 v0 = f3d(P.v), v1 = f3d(P.v + 3), v2 = f3d(P.v + 6);
        double ls[3], rs[3];
        const double half = P.maxside / 2.0;
        // child vertex slots: (a, mid, b) and (mid, c, b) around the split side
        if (P.max_side_idx == 0) {
            const D3 mid = div_d(add_d(v1, v0), 2.0);
            const double nsl = mod_d(sub_d(v2, mid));
            for (int k = 0; k < 3; ++k) L.v[k] = P.v[k];
            d3f(L.v + 3, mid);
            for (int k = 0; k < 3; ++k) L.v[6 + k] = P.v[6 + k];
            ls[0] = half;
            ls[1] = nsl;
            ls[2] = mod_d(sub_d(v0, v2));
            d3f(R.v, mid);
            for (int k = 0; k < 3; ++k) R.v[3 + k] = P.v[3 + k];
            for (int k = 0; k < 3; ++k) R.v[6 + k] = P.v[6 + k];
            rs[0] = half;
            rs[1] = mod_d(sub_d(v2, v1));
            rs[2] = nsl;
        } else if (P.max_side_idx == 1) {
            const D3 mid = div_d(add_d(v2, v1), 2.0);
            const double nsl = mod_d(sub_d(v0, mid));
            for (int k = 0; k < 3; ++k) L.v[k] = P.v[3 + k];
            d3f(L.v + 3, mid);
            for (int k = 0; k < 3; ++k) L.v[6 + k] = P.v[k];
            ls[0] = half;
            ls[1] = nsl;
            ls[2] = mod_d(sub_d(v1, v0));
            d3f(R.v, mid);
            for (int k = 0; k < 3; ++k) R.v[3 + k] = P.v[6 + k];
            for (int k = 0; k < 3; ++k) R.v[6 + k] = P.v[k];
            rs[0] = half;
            rs[1] = mod_d(sub_d(v0, v2));
            rs[2] = nsl;
        } else {
            const D3 mid = div_d(add_d(v0, v2), 2.0);
            const double nsl = mod_d(sub_d(v1, mid));
            for (int k = 0; k < 3; ++k) L.v[k] = P.v[6 + k];
            d3f(L.v + 3, mid);
            for (int k = 0; k < 3; ++k) L.v[6 + k] = P.v[3 + k];
            ls[0] = half;
            ls[1] = nsl;
            ls[2] = mod_d(sub_d(v2, v1));
            d3f(R.v, mid);
            for (int k = 0; k < 3; ++k) R.v[3 + k] = P.v[k];
            for (int k = 0; k < 3; ++k) R.v[6 + k] = P.v[3 + k];
            rs[0] = half;
            rs[1] = mod_d(sub_d(v1, v0));
            rs[2] = nsl;
        }
        L.dim = R.dim = P.dim / 2.0;
        d3f(L.c, div_d(add_d(add_d(f3d(L.v), f3d(L.v + 3)), f3d(L.v + 6)), 3.0));
        d3f(R.c, div_d(add_d(add_d(f3d(R.v), f3d(R.v + 3)), f3d(R.v + 6)), 3.0));
        max_side3(L, ls[0], ls[1], ls[2]);
        max_side3(R, rs[0], rs[1], rs[2]);
    } else {
        const D3 mid = div_d(add_d(f3d(P.v + 3), f3d(P.v)), 2.0);
        d3f(L.v + 3, mid);
        d3f(R.v, mid);
        L.dim = R.dim = P.dim / 2.0;
        L.maxside = L.dim;
        R.maxside = R.dim;
        d3f(L.c, div_d(add_d(f3d(L.v), f3d(L.v + 3)), 2.0));
        d3f(R.c, div_d(add_d(f3d(R.v), f3d(R.v + 3)), 2.0));
    }
    nodes[slot] = L;
    nodes[slot + 1] = R;
    nodes[p].left = slot;
    nodes[p].right = slot + 1;
    nodes[p].type = (uint8_t)(P.type & ~FCG_NODE_LEAF);
}

// refinement items: a (bits 32-62), b (bits 0-30), bit 63 the self-pass
// curvature flag, bit 31 (pending items only) "split b" instead of "split a"
constexpr uint64_t kCurvBit = 1ull << 63, kSplitB = 1ull << 31;

__device__ inline void max_u64_wave(unsigned long long* dst, unsigned long long v) {
    for (int d = 32; d > 0; d >>= 1) {
        const unsigned long long o = __shfl_xor(v, d, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(dst, v);
}

// evaluate one slice: pairs that need refinement become pending items (their
// larger node marked for splitting when it is a leaf); the others update the
// largest relative estimate of non-super nodes (bits of a non-negative double)
__global__ void __launch_bounds__(256) k_refine_eval(const uint64_t* __restrict__ in, const uint8_t* __restrict__ in_lev,
                                                     int64_t n, const fcg_refine_node* __restrict__ nodes,
                                                     RefineParams P, int32_t* __restrict__ want,
                                                     uint64_t* __restrict__ pend, uint8_t* __restrict__ pend_lev,
                                                     unsigned long long* pend_cnt, unsigned long long* max_eps,
                                                     int* max_lev) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned np = 0;
    uint64_t pi = 0;
    int lev = 0;
    unsigned long long eb = 0;
    if (i < n) {
        lev = in_lev[i];
        const uint64_t it = in[i];
        const int a = (int)((it >> 32) & 0x7fffffffu), b = (int)(it & 0x7fffffffu);
        const fcg_refine_node& p1 = nodes[a];
        const fcg_refine_node& p2 = nodes[b];
        double cc = 1.0;
        if (it & kCurvBit) {
            const D3 g1 = geo_normal(p1, P.dim), g2 = geo_normal(p2, P.dim);
            cc = (P.curv - 1.0) * (dot_d(g1, g2) + 1.0) + 1.0;
        }
        double e1, e2;
        const bool force = refine_test(crit_refine(p1, P.dim), crit_refine(p2, P.dim), P.dim, P.mesh_eps, cc,
                                       P.max_side, P.max_area, P.max_length, &e1, &e2);
        if (force) {
            const bool split_a = p1.dim > p2.dim;
            const int big = split_a ? a : b;
            if (nodes[big].left < 0) want[big] = 1;
            pi = (it & ~kSplitB) | (split_a ? 0ull : kSplitB);
            np = 1;
        } else {
            double m = 0.0;
            if (!(p1.type & FCG_NODE_SUPER) && e1 > m) m = e1;
            if (!(p2.type & FCG_NODE_SUPER) && e2 > m) m = e2;
            eb = (unsigned long long)__double_as_longlong(m);
        }
    }
    const unsigned long long o = wave_reserve(pend_cnt, np);
    if (np) {
        pend[o] = pi;
        pend_lev[o] = (uint8_t)lev;
    }
    max_u64_wave(max_eps, eb);
    for (int d = 32; d > 0; d >>= 1) lev = max(lev, __shfl_xor(lev, d, 64));
    if ((threadIdx.x & 63) == 0 && lev > 0) atomicMax(max_lev, lev);
}

// split every marked leaf once (the first pending item that claims it)
__global__ void k_refine_split(const uint64_t* __restrict__ pend, int64_t n, fcg_refine_node* __restrict__ nodes,
                               int32_t* __restrict__ want, unsigned long long* pool_cnt, int dim) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t it = pend[i];
    const int big = (it & kSplitB) ? (int)(it & 0x7fffffffu) : (int)((it >> 32) & 0x7fffffffu);
    if (atomicExch(&want[big], 0) != 1) return;
    const int slot = (int)atomicAdd(pool_cnt, 2ull);
    split_node(nodes, big, slot, dim);
}

// the two child pairs of every pending item (its larger node now split)
__global__ void __launch_bounds__(256) k_refine_expand(const uint64_t* __restrict__ pend,
                                                       const uint8_t* __restrict__ pend_lev, int64_t n,
                                                       const fcg_refine_node* __restrict__ nodes,
                                                       uint64_t* __restrict__ out, uint8_t* __restrict__ out_lev,
                                                       unsigned long long* out_cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned ne = 0;
    uint64_t e0 = 0, e1 = 0;
    uint8_t lv = 0;
    if (i < n) {
        const uint64_t it = pend[i];
        const uint64_t curv = it & kCurvBit;
        const int a = (int)((it >> 32) & 0x7fffffffu), b = (int)(it & 0x7fffffffu);
        lv = (uint8_t)(pend_lev[i] < 255 ? pend_lev[i] + 1 : 255);
        if (it & kSplitB) {
            e0 = curv | pair_item(a, nodes[b].left);
            e1 = curv | pair_item(a, nodes[b].right);
        } else {
            e0 = curv | pair_item(nodes[a].left, b);
            e1 = curv | pair_item(nodes[a].right, b);
        }
        ne = 2;
    }
    const unsigned long long o = wave_reserve(out_cnt, ne);
    if (ne) {
        out[o] = e0;
        out[o + 1] = e1;
        out_lev[o] = lv;
        out_lev[o + 1] = lv;
    }
}

// one self round: split the listed tops that are leaves (first round only),
// then each non-leaf node gives the pair of its children (curvature flag set)
// to the work stack and its children to the next round
__global__ void k_refine_split_tops(const int32_t* __restrict__ list, int n, fcg_refine_node* __restrict__ nodes,
                                    unsigned long long* pool_cnt, int dim) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = list[i];
    if (nodes[p].left >= 0) return;
    const int slot = (int)atomicAdd(pool_cnt, 2ull);
    split_node(nodes, p, slot, dim);
}

__global__ void __launch_bounds__(256) k_refine_self(const int32_t* __restrict__ front, int64_t n, int lev,
                                                     const fcg_refine_node* __restrict__ nodes,
                                                     uint64_t* __restrict__ out, uint8_t* __restrict__ out_lev,
                                                     unsigned long long* out_cnt, int32_t* __restrict__ next,
                                                     unsigned long long* next_cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned ne = 0;
    int l = -1, r = -1;
    if (i < n) {
        const int p = front[i];
        l = nodes[p].left;
        r = nodes[p].right;
        ne = l >= 0 ? 1 : 0;
    }
    const unsigned long long o = wave_reserve(out_cnt, ne);
    if (ne) {
        out[o] = kCurvBit | pair_item(l, r);
        out_lev[o] = (uint8_t)(lev < 255 ? lev + 1 : 255);
    }
    const unsigned long long q = wave_reserve(next_cnt, 2 * ne);
    if (ne) {
        next[q] = l;
        next[q + 1] = r;
    }
}

// ---------------------------------------------------------------- device layout
// of a job whose links were discovered here (the upload's host-side layout
// build, done on the device)

__global__ void k_perm_cols(const int32_t* __restrict__ col_node, int64_t L, const int32_t* __restrict__ perm,
                            int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L) out[i] = perm[col_node[i]];
}

// the upload's fill class of a link (destination a, source b): the branch of
// PotEstimateOpt from the centroid distance over the larger max side in FP64,
// then the dielectric row, then the dimension (same expression order as the
// host's, no contraction)
__device__ int fill_class_dev(const fcg_node_geom& a, const fcg_node_geom& b, int dim) {
    const double dx = (double)a.c[0] - (double)b.c[0], dy = (double)a.c[1] - (double)b.c[1];
    const double rmax = a.maxside > b.maxside ? a.maxside : b.maxside;
    const int diel = (a.flags & FCG_NODE_DIEL) ? 1 : 0;
    if (dim == 2) return 10 + 2 * diel + (sqrt(dx * dx + dy * dy) / rmax < 4.0 ? 0 : 1);
    const double dz = (double)a.c[2] - (double)b.c[2];
    const double ratio = sqrt(dx * dx + dy * dy + dz * dz) / rmax;
    const int br = ratio < 1.0 ? 0 : ratio < 2.0 ? 1 : ratio < 4.0 ? 2 : ratio < 10.0 ? 3 : 4;
    return 5 * diel + br;
}

__global__ void k_link_fill_class(const int32_t* __restrict__ link_row, const int32_t* __restrict__ col,
                                  const int32_t* __restrict__ iperm, const fcg_node_geom* __restrict__ nodes,
                                  int dim, int64_t L, uint8_t* __restrict__ cls) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L) cls[i] = (uint8_t)fill_class_dev(nodes[link_row[i]], nodes[iperm[col[i]]], dim);
}

// Stable partition of the links by class: per chunk of kClassChunk links a
// class histogram, then (with the host's exclusive offsets, class-major and
// chunk-minor) a scatter that keeps link order inside every class.
constexpr int kClassChunk = 4096;
__global__ void __launch_bounds__(256) k_class_hist(const uint8_t* __restrict__ cls, int64_t L,
                                                    int32_t* __restrict__ hist) {
    __shared__ int32_t h[kFillClasses];
    if (threadIdx.x < kFillClasses) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kClassChunk;
    for (int k = threadIdx.x; k < kClassChunk; k += 256) {
        const int64_t l = b0 + k;
        if (l < L) atomicAdd(&h[cls[l]], 1);
    }
    __syncthreads();
    if (threadIdx.x < kFillClasses) hist[(int64_t)blockIdx.x * kFillClasses + threadIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(256) k_class_scatter(const uint8_t* __restrict__ cls, int64_t L,
                                                       const int64_t* __restrict__ off,
                                                       uint32_t* __restrict__ order) {
    __shared__ int64_t run[kFillClasses];
    __shared__ int32_t wc[4][kFillClasses];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t < kFillClasses) run[t] = off[(int64_t)blockIdx.x * kFillClasses + t];
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kClassChunk;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int tile = 0; tile < kClassChunk; tile += 256) {
        const int64_t l = b0 + tile + t;
        const int c = l < L ? (int)cls[l] : -1;
        int rank = 0;
        for (int cc = 0; cc < kFillClasses; ++cc) {
            const uint64_t m = __ballot(c == cc);
            if (c == cc) rank = __popcll(m & below);
            if (lane == 0) wc[w][cc] = __popcll(m);
        }
        __syncthreads();
        if (c >= 0) {
            int64_t base = run[c];
            for (int v = 0; v < w; ++v) base += wc[v][c];
            order[base + rank] = (uint32_t)l;
        }
        __syncthreads();
        if (t < kFillClasses) run[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
        __syncthreads();
    }
}

__global__ void k_col16(const int32_t* __restrict__ col, int64_t n, uint16_t* __restrict__ c16) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c16[i] = (uint16_t)col[i];
}

}  // namespace

int links_sort_lds_max() { return kSortLds; }

hipError_t launch_refine_eval(const uint64_t* in, const uint8_t* in_lev, int64_t n, const fcg_refine_node* nodes,
                              const RefineParams& P, int32_t* want, uint64_t* pend, uint8_t* pend_lev,
                              unsigned long long* pend_cnt, unsigned long long* max_eps, int* max_lev, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_refine_eval<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, in_lev, n, nodes, P, want, pend, pend_lev, pend_cnt,
                                                               max_eps, max_lev);
    return hipGetLastError();
}

hipError_t launch_refine_split(const uint64_t* pend, int64_t n, fcg_refine_node* nodes, int32_t* want,
                               unsigned long long* pool_cnt, int dim, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_refine_split<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pend, n, nodes, want, pool_cnt, dim);
    return hipGetLastError();
}

hipError_t launch_refine_expand(const uint64_t* pend, const uint8_t* pend_lev, int64_t n,
                                const fcg_refine_node* nodes, uint64_t* out, uint8_t* out_lev,
                                unsigned long long* out_cnt, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_refine_expand<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pend, pend_lev, n, nodes, out, out_lev, out_cnt);
    return hipGetLastError();
}

hipError_t launch_refine_split_tops(const int32_t* list, int n, fcg_refine_node* nodes, unsigned long long* pool_cnt,
                                    int dim, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_refine_split_tops<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(list, n, nodes, pool_cnt, dim);
    return hipGetLastError();
}

hipError_t launch_refine_self(const int32_t* front, int64_t n, int lev, const fcg_refine_node* nodes, uint64_t* out,
                              uint8_t* out_lev, unsigned long long* out_cnt, int32_t* next,
                              unsigned long long* next_cnt, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_refine_self<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(front, n, lev, nodes, out, out_lev, out_cnt, next,
                                                               next_cnt);
    return hipGetLastError();
}
int class_chunk() { return kClassChunk; }

hipError_t launch_links_sort_long(const int64_t* row_ptr, const int32_t* rows, int nrows, const int64_t* scr_off,
                                  int32_t* col, int32_t* scratch, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    k_links_sort_long<<<(unsigned)nrows, 1024, 0, s>>>(row_ptr, rows, scr_off, col, scratch);
    return hipGetLastError();
}

hipError_t launch_perm_cols(const int32_t* col_node, int64_t L, const int32_t* perm, int32_t* out, hipStream_t s) {
    if (L <= 0) return hipSuccess;
    k_perm_cols<<<(unsigned)((L + 255) / 256), 256, 0, s>>>(col_node, L, perm, out);
    return hipGetLastError();
}

hipError_t launch_link_fill_class(const int32_t* link_row, const int32_t* col, const int32_t* iperm,
                                  const fcg_node_geom* nodes, int dim, int64_t L, uint8_t* cls, hipStream_t s) {
    if (L <= 0) return hipSuccess;
    k_link_fill_class<<<(unsigned)((L + 255) / 256), 256, 0, s>>>(link_row, col, iperm, nodes, dim, L, cls);
    return hipGetLastError();
}

hipError_t launch_class_hist(const uint8_t* cls, int64_t L, int32_t* hist, hipStream_t s) {
    if (L <= 0) return hipSuccess;
    k_class_hist<<<(unsigned)((L + kClassChunk - 1) / kClassChunk), 256, 0, s>>>(cls, L, hist);
    return hipGetLastError();
}

hipError_t launch_class_scatter(const uint8_t* cls, int64_t L, const int64_t* off, uint32_t* order, hipStream_t s) {
    if (L <= 0) return hipSuccess;
    k_class_scatter<<<(unsigned)((L + kClassChunk - 1) / kClassChunk), 256, 0, s>>>(cls, L, off, order);
    return hipGetLastError();
}

hipError_t launch_col16(const int32_t* col, int64_t n, uint16_t* c16, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_col16<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(col, n, c16);
    return hipGetLastError();
}

hipError_t launch_links_step(const uint64_t* in, const uint8_t* in_lev, int64_t n, uint64_t* out, uint8_t* out_lev,
                             unsigned long long* out_cnt, int2* links, unsigned long long* link_cnt, int* max_lev,
                             const fcg_node_geom* nodes, const int32_t* left, const int32_t* right,
                             const LinkParams& P, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_links_step<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, in_lev, n, out, out_lev, out_cnt, links, link_cnt,
                                                              max_lev, nodes, left, right, P);
    return hipGetLastError();
}

hipError_t launch_links_degree(const int2* links, int64_t P, int32_t* deg, hipStream_t s) {
    if (P <= 0) return hipSuccess;
    k_links_degree<<<(unsigned)((P + 255) / 256), 256, 0, s>>>(links, P, deg);
    return hipGetLastError();
}

hipError_t launch_links_scatter(const int2* links, int64_t P, unsigned long long* cursor, int32_t* col,
                                hipStream_t s) {
    if (P <= 0) return hipSuccess;
    k_links_scatter<<<(unsigned)((P + 255) / 256), 256, 0, s>>>(links, P, cursor, col);
    return hipGetLastError();
}

hipError_t launch_links_sort_rows(const int64_t* row_ptr, int32_t* col, int M, hipStream_t s) {
    if (M <= 0) return hipSuccess;
    k_links_sort_rows<<<(unsigned)M, 256, 0, s>>>(row_ptr, col, M);
    return hipGetLastError();
}

}  // namespace fcg
