/*
 * oracle.h - CPU restatement of FasterCap 6.0.7's BEM solve path.
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker for the gfx950 path and the
 * reported CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never calls it.
 *
 * It consumes the flat job arrays of include/fastercap_amd.h (built by the host
 * mesh builder) and restates, independently of the device code:
 *   orc_potential_opt_3d   Solver/Potential.cpp:1561-1711
 *   orc_potential_num_3d   Solver/Potential.cpp:2077-2153 (rules :112-206)
 *   orc_enfield_num_3d     Solver/Potential.cpp:2172-2254
 *   orc_potential_opt_2d   Solver/Potential.cpp:1866-1949
 *   orc_enfield_2d         Solver/Potential.cpp:1967-2060
 *   orc_auto_3d            Solver/Potential.cpp:994-1017 (Galerkin self term)
 *   orc_mutual_half_3d     Solver/Potential.cpp:2485-2625 (rules 10, 17 :345-389, 717-800)
 *   orc_mutual_dhalf_3d    Solver/Potential.cpp:2802-2964
 *   orc_mutual_full_3d     Solver/Potential.cpp:3138-3347, 3790-4003
 *   orc_mutual_dfull_3d    Solver/Potential.cpp:4197-4377, 3349-3577
 *   orc_pot_estimate       CAutoRefine::PotEstimateOpt one-sided 3D/2D, collocation and Galerkin
 *                          (Solver/Autorefine.cpp:3474-3733, 4156-4325)
 *   orc_self               CAutoRefine::SelfPotential (Solver/Autorefine.cpp:2358-2506)
 *   orc_fill               fill loop of AutoRefineLinks (Solver/Autorefine.cpp:760-840), OpenMP over nodes
 *   orc_mvp                CMultHier::MultiplyMatByVec_fast (Solver/MultiplyHierarchical.cpp:72-330)
 *   orc_solve              CSolveCap::Solve + gmresPrecondSFastAll + Jacobi
 *                          (Solver/SolveCapacitance.cpp:1138-1425, 2288-2574, 3489-3526)
 *   orc_compressed_dense   the compressed matrix column by column (DebugDumpPotMtx pattern,
 *                          Solver/SolveCapacitance.cpp:3734-3795), built directly from the links
 *   orc_leaf_dense         the same matrix with every leaf pair evaluated directly (no
 *                          hierarchical approximation), for closed-form checks
 *   orc_link_coverage      link-partition invariant: every ordered leaf pair i != j is covered
 *                          by exactly one link
 *   orc_block_partition    leaf blocks of the block preconditioner (RecurseComputePrecond,
 *                          Solver/SolveCapacitance.cpp:1094-1104, 1876-1993)
 *   orc_block_precond      each block's collocation matrix inverted (ComputeBlockPrecond,
 *                          InvertMatrix, :2068-2258)
 *   orc_block_apply        Pq = block-diagonal product (ComputePrecondVectFast block branch,
 *                          :3527-3544)
 *   orc_solve_block        orc_solve with the block preconditioner
 *   orc_super_partition    two-level preconditioner elements (MakePreconditioner and the super
 *                          branch of RecurseComputePrecond, Solver/SolveCapacitance.cpp:1067-1104,
 *                          1876-1993)
 *   orc_super_precond      their collocation matrix inverted (ComputeSuperPrecond, :2007-2066)
 *   orc_super_apply        collect / multiply / distribute (ComputePrecondVectFast, :3545-3592)
 *   orc_solve_super        orc_solve with the two-level preconditioner
 *   orc_node_self          SelfPotential of any tree node (leaf, refined parent or super node)
 *   orc_self_cplx          SelfPotential of a complex job's leaf (Solver/Autorefine.cpp:2358-2393, 2435-2471)
 *   orc_mvp_cplx           MultiplyMatByVec_fast complex branch [R -D; D R] (MultiplyHierarchical.cpp:332-455)
 *   orc_solve_cplx         CSolveCap::Solve of a complex job: GMRES on the real 2N system, Jacobi on both
 *                          halves, complex charge sums (SolveCapacitance.cpp:1138-1366, 3498-3526)
 */
#ifndef FASTERCAP_ORACLE_H
#define FASTERCAP_ORACLE_H

#include "../include/fastercap_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

double orc_potential_opt_3d(const double* r, const float* v);
double orc_potential_num_3d(const double* r, const float* v, int rule);
double orc_enfield_num_3d(const double* r, const float* v, const double* n, int rule);
double orc_potential_opt_2d(const double* r, const float* v);
double orc_enfield_2d(const double* r, const float* v, const double* n);

/* Galerkin scheme (3D): Auto, MutualHalfNumerical, MutualDHalfNumerical, the product-rule
 * double quadratures (order 2: Mutual_2thOrd_FullNum, 4: Mutual_4thOrd_FullNum) and
 * MutualDFullNumerical (rule 0: MutualD_2thOrd_FullNum), all divided by 4 pi e0 */
double orc_auto_3d(const float* v);
double orc_mutual_half_3d(const float* v1, const float* v2, int rule);
double orc_mutual_dhalf_3d(const float* v1, const float* v2, const double* n, double h, int rule);
double orc_mutual_full_3d(const float* v1, const float* v2, int order);
double orc_mutual_dfull_3d(const float* v1, const float* v2, const double* n, int rule);

double orc_pot_estimate(const fcg_job_desc* d, int node_i, int node_j);
double orc_self(const fcg_job_desc* d, int leaf);

/* coef [L], self [N]; nthreads <= 0 uses the OpenMP default. Returns 0. */
int orc_fill(const fcg_job_desc* d, double* coef, double* self_coef, int nthreads);

/* v = P q for one vector (N), 2D fix-ups included. Returns 0. */
int orc_mvp(const fcg_job_desc* d, const double* coef, const double* self_coef, const double* q, double* v,
            int nthreads);

/* C matrix (rows x n_cond, row-major), 3D: rows = n_cond, 2D: n_cond - 1, unscaled.
 * iters[rows] receives GMRES iteration counts. precond: 0 none, 1 Jacobi. */
int orc_solve(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int precond,
              double* cap, int* rows, int* cols, int* iters, int nthreads);

/* Same, stopping after the first `max_rows` right-hand sides (bounded CPU baseline samples). */
int orc_solve_rows(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int precond,
                   int max_rows, double* cap, int* rows, int* cols, int* iters, int nthreads);

/* Dense compressed matrix P[N][N] (row-major) assembled from the links. */
int orc_compressed_dense(const fcg_job_desc* d, const double* coef, const double* self_coef, double* P);

/* Leaf-level dense matrix P[N][N]: every leaf pair evaluated directly with the
 * same estimate as a leaf-to-leaf link (no hierarchical approximation), self
 * terms on the diagonal. */
int orc_leaf_dense(const fcg_job_desc* d, double* P, int nthreads);

/* Returns the number of ordered leaf pairs whose link cover count differs from the
 * expected value (1 for i != j, 0 for i == j); 0 means the links partition the matrix. */
long long orc_link_coverage(const fcg_job_desc* d);

/* Block preconditioner of block size bsize (2..128).  starts[0..nblocks] are the
 * blocks' first leaves (capacity n_leaves + 1); P holds each block's inverse,
 * row-major, block after block, at poff[b] (capacity n_leaves * bsize doubles,
 * nblocks + 1 offsets).  A leaf reached before any block is open forms a block of
 * its own (the reference leaves that block unset). */
int orc_block_partition(const fcg_job_desc* d, int bsize, int* starts, int* nblocks);
int orc_block_precond(const fcg_job_desc* d, int bsize, int* starts, int* nblocks, long long* poff, double* P);
void orc_block_apply(int nblocks, const int* starts, const long long* poff, const double* P, const double* q,
                     double* out);
int orc_solve_block(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int bsize,
                    double* cap, int* rows, int* cols, int* iters, int nthreads);

/* Two-level preconditioner of dimension dim: m elements (nodes[m], leaf ranges
 * starts[0..m], capacity dim and dim + 1), Sinv m x m row-major (capacity dim^2).
 * Returns 1 when unavailable (at least dim tree roots). */
double orc_node_self(const fcg_job_desc* d, int node);
int orc_super_partition(const fcg_job_desc* d, int dim, int* nodes, int* starts, int* m);
int orc_super_precond(const fcg_job_desc* d, int dim, int* nodes, int* starts, int* m, double* Sinv);
void orc_super_apply(const fcg_job_desc* d, int m, const int* nodes, const int* starts, const double* Sinv,
                     const double* q, double* out);
int orc_solve_super(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int dim,
                    double* cap, int* rows, int* cols, int* iters, int nthreads);

/* Complex permittivities (jobs whose desc carries the *_im arrays). */
void orc_self_cplx(const fcg_job_desc* d, int leaf, double* re, double* im);
/* q, v [Re(N) | Im(N)] */
int orc_mvp_cplx(const fcg_job_desc* d, const double* coef, const double* self_re, const double* self_im,
                 const double* q, double* v, int nthreads);
/* cap: real matrix (rows x n_cond) then the imaginary one */
int orc_solve_cplx(const fcg_job_desc* d, const double* coef, const double* self_re, const double* self_im,
                   double tol, int precond, double* cap, int* rows, int* cols, int* iters, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
