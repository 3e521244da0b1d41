// gfx950 kernels of the FasterCap solve path.  Host-side launch wrappers are
// declared here and implemented in kernels.hip; the C-ABI (fcg_api.hip) owns the
// memory and calls them on the context stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/fastercap_amd.h"

namespace fcg {

// Double-double value used by the prefix scan of the up-pass.
struct DD {
    double hi, lo;
};

// Device view of one job (all pointers are device memory).
struct JobView {
    int dim, N, M, G;
    int64_t L;
    const fcg_node_geom* nodes;
    const int32_t* left;
    const int32_t* right;
    const int32_t* parent;
    const int32_t* leaf_lo;
    const int32_t* leaf_hi;
    const int32_t* leaf_node;
    const int32_t* node_depth;
    const int64_t* row_ptr;
    const int32_t* col;
    const double* group_outperm;
    const double* group_inperm;
    double* coef;    // [L]
    double* self;    // [N]
    double* diag;    // [M] self coefficient of leaf nodes, 0 elsewhere
    int first_cond_leaf;  // 2D zero-total-charge fix-up start
    int gather_variant;   // index into the gather configuration table
};

// Scratch of one hierarchical product on k right-hand sides.
struct MvpScratch {
    DD* prefix;       // [k][N+1]
    DD* block_sums;   // [k][nblk]
    double* charge;   // [M][k]
    double* pot;      // [M][k]
    int nblk;
    int kmax;
};

constexpr int kScanTile = 2048;  // elements per scan block (256 threads x 8)

// Fill
hipError_t launch_link_rows(const JobView& j, int32_t* link_row, hipStream_t s);
hipError_t launch_fill(const JobView& j, const int32_t* link_row, hipStream_t s);
hipError_t launch_self(const JobView& j, hipStream_t s);

// Product v = P q on k vectors ([k][N] layout). If jacobi != 0 the input is
// first divided by the self coefficients (right preconditioning fused in).
// `ev`, when given, holds 5 events recorded before the scan and after the scan,
// the charges, the gather and the down-pass (stage timing).
hipError_t launch_mvp(const JobView& j, MvpScratch& sc, const double* q, double* v, int k, int jacobi,
                      hipStream_t s, hipEvent_t* ev = nullptr);
// Deterministic pseudo-random fill of n doubles in [-1, 1).
hipError_t launch_random(double* out, size_t n, uint64_t seed, hipStream_t s);

// Elementwise / Krylov helpers on [k][N] blocks.
hipError_t launch_jacobi(const JobView& j, const double* q, double* pq, int k, hipStream_t s);
// partial[b][jj][r] = sum over block b of Q_jj[r][n] * w[r][n]; Q_jj = basis[jj] ([k][N])
hipError_t launch_block_dots(double* const* basis, int nvec, const double* w, int N, int k, double* partial,
                             int nblocks, hipStream_t s);
hipError_t launch_reduce_partials(const double* partial, int nblocks, int nvals, double* out, hipStream_t s);
// w[r][n] -= sum_jj h[jj][r] * Q_jj[r][n]
hipError_t launch_basis_update(double* const* basis, int nvec, const double* h, double* w, int N, int k,
                               hipStream_t s);
// Same update, plus partial[r][chunk] = sum over the chunk of the updated w^2.
hipError_t launch_basis_update_norm(double* const* basis, int nvec, const double* h, double* w, int N, int k,
                                    double* partial, int nchunks, hipStream_t s);
// out[r][n] = a[r][n] / d[r] (0 where d[r] == 0)
hipError_t launch_scale(const double* a, const double* scale, double* out, int N, int k, hipStream_t s);
// out[r][n] = x[r][n] + y[r][n]
hipError_t launch_add(const double* x, const double* y, double* out, int N, int k, hipStream_t s);
// out[r][n] = x[r][n] - y[r][n]
hipError_t launch_sub(const double* x, const double* y, double* out, int N, int k, hipStream_t s);
// out[r][n] += sum_jj z[jj][r] * Q_jj[r][n]
hipError_t launch_basis_combine(double* const* basis, int nvec, const double* z, double* out, int N, int k,
                                hipStream_t s);
// out[r][c] = sum over leaves n of conductor column c of leaf_weight[n] * x[r][n];
// col_range[2c], col_range[2c+1] bound the leaves of column c
hipError_t launch_charge_sums(const double* x, const double* leaf_weight, const int32_t* col_range, int N, int k,
                              int ncol, double* out, hipStream_t s);

int dots_blocks(int N);
int gather_variant_count();

}  // namespace fcg
