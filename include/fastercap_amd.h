/*
 * fastercap_amd.h - C-ABI of the MI355X (gfx950) BEM solve path of FasterCap.
 *
 * The library owns device memory for "jobs": one refined panel hierarchy with its
 * destination-sorted link list.  A host solver (the mesh builder of this package,
 * or FasterCap's own CSolveCap) uploads the hierarchy once, then calls the fill,
 * the hierarchical matrix-vector product and the preconditioned GMRES through the
 * entry points below.  All sizes and pointers are plain C; no exceptions cross the
 * boundary; every function returns an FC_* code (0 = success) as FasterCap does
 * (Solver/SolverGlobal.h:39-49).
 *
 * Reference interfaces each entry point replaces are cited per function.
 */
#ifndef FASTERCAP_AMD_H
#define FASTERCAP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Discretization schemes: the reference's AUTOREFINE_COLLOCATION / AUTOREFINE_GALERKIN
 * values (Solver/SolverGlobal.h:72-73). */
#define FCG_SCHEME_COLLOCATION 1
#define FCG_SCHEME_GALERKIN 2

/* Return codes: identical values to FasterCap's FC_* (Solver/SolverGlobal.h:39-49). */
#define FCG_OK 0
#define FCG_GENERIC_ERROR 1
#define FCG_COMMAND_LINE_ERROR 64
#define FCG_CANNOT_OPEN_FILE 66
#define FCG_OUT_OF_MEMORY 71
#define FCG_FILE_ERROR 74
#define FCG_EXCEPTION_ERROR 98
#define FCG_USER_BREAK 125
/* Codes of this library that have no FasterCap counterpart. */
#define FCG_INVALID_ARGUMENT 65
#define FCG_DEVICE_ERROR 96

/* Node flag bits: same values as CAutoElement::m_ucType (Solver/AutoElement.h:44-48). */
#define FCG_NODE_LEAF 1u
#define FCG_NODE_DIEL 2u
#define FCG_NODE_OUTPERM_NORMAL_DIR 4u
#define FCG_NODE_OUTPERM_ELEMENT_LEVEL 8u
#define FCG_NODE_SUPER 128u

/* Precision modes of the link coefficients and of the product. */
#define FCG_PREC_FP64 0  /* every coefficient and accumulation in FP64 */
#define FCG_PREC_MIXED 1 /* FP32 far-field coefficients/products, FP64 near field + refinement */

/* Preconditioner bit flags: values of AUTOREFINE_PRECOND_* (Solver/SolverGlobal.h:59-68),
 * so a caller passes CAutoRefGlobalVars::m_ucPrecondType unchanged.  fcg_solve_opts.precond
 * is read as the reference reads that bit set (Solver/SolveCapacitance.cpp:122-125,
 * 1092-1108): SUPER with anything else is SUPER alone (the reference drops JACOBI
 * with a warning), BLOCK with JACOBI is BLOCK alone (the reference then runs a
 * flexible GMRES it reports as unsupported), JACOBI is JACOBI, 0 is none.  HIER
 * (the hierarchical preconditioner, -ph) is not built: FCG_INVALID_ARGUMENT. */
#define FCG_PRECOND_NONE 0
#define FCG_PRECOND_JACOBI 1
#define FCG_PRECOND_BLOCK 2 /* block preconditioner (-pb<n>), fcg_solve_opts.precond_size */
#define FCG_PRECOND_SUPER 4 /* two-level preconditioner (-ps<n>), fcg_solve_opts.precond_size */
#define FCG_PRECOND_HIER 8  /* reserved: the reference's value for -ph, rejected here */

/*
 * Geometry of one tree node (leaf panel, refined parent or super node), 96 bytes.
 * 3D: v = triangle vertices (x,y,z) x 3; 2D: v[0..2] and v[3..5] are the segment
 * end points (z = 0).  Values are the reference's stored float geometry
 * (CAutoPanel::m_clsVertex / m_clsCentroid, Solver/AutoPanel.h:196-199).
 */
typedef struct fcg_node_geom {
    float v[9];       /* vertices */
    float c[3];       /* centroid */
    uint32_t flags;   /* FCG_NODE_* */
    int32_t cond;     /* index of the owning conductor / dielectric group */
    double maxside;   /* GetMaxSideLen() (3D) or GetLength() (2D) */
    double area;      /* m_dDimension: area (3D) or length (2D) */
    double dn[3];     /* GetDielNormal(): unit normal towards the outer permittivity */
} fcg_node_geom;

/*
 * One job = one refined geometry.  Nodes are numbered in the post-order of the
 * reference's RecurseIndex (Solver/Autorefine.cpp:1173-1220); leaves are numbered
 * in left-first DFS order across the groups, so every node owns the contiguous
 * leaf range [leaf_lo, leaf_hi).  Row r of the link CSR holds the links of node r
 * in the order the reference's RefineMutual recursion recorded them; col holds
 * the source node ids.
 */
typedef struct fcg_job_desc {
    int32_t dim;               /* 3 or 2 */
    int32_t n_leaves;          /* N */
    int32_t n_nodes;           /* M */
    int32_t n_groups;          /* conductors + dielectric interfaces, deque order */
    int64_t n_links;           /* L (mutual links, self terms excluded) */
    const fcg_node_geom* nodes; /* [M] */
    const int32_t* left;       /* [M] left child node id, -1 for leaves */
    const int32_t* right;      /* [M] right child node id, -1 for leaves */
    const int32_t* parent;     /* [M] parent node id, -1 for group roots */
    const int32_t* leaf_lo;    /* [M] */
    const int32_t* leaf_hi;    /* [M] */
    const int32_t* leaf_node;  /* [N] node id of leaf k */
    const int64_t* row_ptr;    /* [M+1] */
    const int32_t* col;        /* [L] */
    const int32_t* group_is_diel;     /* [n_groups] */
    const int32_t* group_leaf_start;  /* [n_groups+1] leaf ranges of the groups */
    const double* group_outperm;      /* [n_groups] outer permittivity (real part) */
    const double* group_inperm;       /* [n_groups] inner permittivity (dielectrics) */
    const double* leaf_outperm;       /* [N] permittivity multiplying the leaf charge in C */
    /* Imaginary parts of the three permittivity arrays above, or all NULL when
     * every permittivity is real.  A job with them is complex (the reference's
     * AUTOREFINE_CPLX_PERM, Solver/Autorefine.cpp:4840-4951): its solves run on
     * the real form [R -D; D R] of the complex system (R the hierarchical matrix,
     * D the imaginary self terms of the dielectric panels), every right-hand side
     * and charge vector is [Re(N) | Im(N)], and every capacitance matrix comes
     * back as its real part followed by its imaginary part
     * (Solver/MultiplyHierarchical.cpp:332-455, SolveCapacitance.cpp:1305-1366). */
    const double* group_outperm_im;
    const double* group_inperm_im;
    const double* leaf_outperm_im;
    /* Discretization scheme (-g; m_cScheme, Solver/SolverGlobal.h:72-73): FCG_SCHEME_GALERKIN
     * takes the Galerkin branches of PotEstimateOpt (Solver/Autorefine.cpp:3570-3674) and
     * SelfPotential (:2399-2401) for the fill, the preconditioners' blocks and the self terms;
     * 0 or FCG_SCHEME_COLLOCATION collocation.  3D only: the reference solves 2D jobs by
     * collocation (Solver/SolveCapacitance.cpp:159-161), and so does this library. */
    int32_t scheme;
} fcg_job_desc;

typedef struct fcg_solve_opts {
    double tol;        /* GMRES relative tolerance (-t), reference stop rule */
    int32_t max_iter;  /* cap on Arnoldi steps (reference: 1000) */
    int32_t precond;   /* FCG_PRECOND_* */
    int32_t precision; /* FCG_PREC_FP64 or FCG_PREC_MIXED */
    int32_t max_refine;/* outer FP64 refinement sweeps in FCG_PREC_MIXED */
    int32_t precond_size; /* FCG_PRECOND_BLOCK: most leaves per block (-pb<n>; 0 = 32, clamped to 2..128);
                             FCG_PRECOND_SUPER: dimension (-ps<n>; 0 = 128, clamped to 16..4096) */
    /* Optional residual history of fcg_solve, the values the reference prints with -r
     * (Solver/SolveCapacitance.cpp:2374-2376, 2511-2513): residuals[q * residuals_cap + i]
     * for right-hand side q holds, at i = 0, the norm of the initial residual |r0| and, at
     * i >= 1, the relative residual estimate |g_i| / |b| after Arnoldi step i (the last of
     * them the final one); at most residuals_cap values per right-hand side, the count is
     * stats->iterations[q] + 1.  Mixed precision: the inner cycles' estimates in order.
     * NULL: not recorded (fcg_solve_batch and the engine ignore it). */
    double* residuals;
    int32_t residuals_cap;
} fcg_solve_opts;

typedef struct fcg_solve_stats {
    int32_t iterations[64]; /* GMRES iterations per right-hand side (first 64) */
    int32_t n_rhs;
    int32_t max_iterations;
    double final_residual;  /* largest relative residual estimate at exit */
    double t_fill_s, t_solve_s, t_mvp_s;
    int64_t n_mvp;          /* number of hierarchical products executed */
    int32_t converged;      /* 1: every right-hand side met the tolerance (FP64: the GMRES
                               estimate; mixed: the FP64 residual of the last refinement
                               check); 0: the iteration cap or max_refine ended the solve */
    int32_t cycles;         /* inner GMRES cycles (mixed: refinement cycles) */
    /* phase time and device memory of the solve, the figures the reference prints
     * with -i (Solver/SolveCapacitance.cpp:704-720) */
    double t_precond_s;     /* preconditioner set-up (fcg_solve builds it before the
                               iterations; 0 in batched solves, which build it in a step) */
    int64_t mem_gmres;      /* Krylov basis and Arnoldi records */
    int64_t mem_vectors;    /* solution, residual and work vectors */
    int64_t mem_product;    /* product scratch: node charges and potentials */
} fcg_solve_stats;

typedef struct fcg_ctx fcg_ctx;

/* Create a context on HIP device `device`. No reference counterpart (the reference
 * state lives in CSolveCap / CMultHier objects, Solver/SolveCapacitance.h:53-191). */
int fcg_ctx_create(int device, fcg_ctx** out);
int fcg_ctx_destroy(fcg_ctx* ctx);

/* Device memory a resident job holds, by the categories the reference prints with -i
 * (Solver/SolveCapacitance.cpp:713-719): the panel hierarchy (tree, geometry, tree-pass
 * tables), the link structure (rows, sources, coefficients, fill and gather tables), the
 * conductor list (leaf ranges, charge weights) and the preconditioner arrays. */
typedef struct fcg_job_mem {
    int64_t hierarchy, links, conductors, precond;
} fcg_job_mem;
int fcg_job_memory(fcg_ctx* ctx, int job, fcg_job_mem* out);

/* Upload a job's hierarchy and link structure into HBM.
 * Replaces the in-memory link chunk allocation of CAutoRefine::AutoRefineLinks
 * (Solver/Autorefine.cpp:481-706) and the tree the MVP walks. */
int fcg_job_upload(fcg_ctx* ctx, const fcg_job_desc* desc, int* job);
int fcg_job_release(fcg_ctx* ctx, int job);

/* Compute the self coefficients and every link coefficient on the device.
 * Replaces the fill loop of CAutoRefine::AutoRefineLinks (Solver/Autorefine.cpp:760-840)
 * calling PotEstimateOpt (Solver/Autorefine.cpp:3450-3733, 2D :4156-4325), and the self
 * terms of RefineSelf/SelfPotential (Solver/Autorefine.cpp:2115-2229, 2358-2506). */
int fcg_fill(fcg_ctx* ctx, int job, int precision);
/* fcg_fill of several jobs in one launch (same arithmetic per link).
 * A fill starts a new solve: the job's two-level preconditioner inverse is
 * built again on its next use (its element partition is kept). */
int fcg_fill_batch(fcg_ctx* ctx, const int* jobs, int njobs, int precision);

/* Copy the filled coefficients to host memory (coef [L] and self [N], either may be NULL). */
int fcg_download_coeffs(fcg_ctx* ctx, int job, double* coef, double* self_coef);

/* v = P q for k right-hand sides, host buffers laid out vector after vector ([k][N];
 * complex jobs [k][2N], each vector [Re | Im], P the real form [R -D; D R]).
 * Replaces CMultHier::MultiplyMatByVec_fast (Solver/MultiplyHierarchical.cpp:285-458). */
int fcg_mvp(fcg_ctx* ctx, int job, const double* q, double* v, int k);

/* Same product on device-resident buffers, enqueued on `stream` (a hipStream_t, NULL =
 * the context stream).  No host synchronisation. */
int fcg_mvp_device(fcg_ctx* ctx, int job, const double* d_q, double* d_v, int k, void* stream);

/* fcg_mvp_device that allocates nothing.  No reference counterpart: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same product, kernels and result as
 * fcg_mvp_device (FP64 job, k right-hand sides, device buffers, enqueued on
 * `stream`, NULL = the context stream, no host synchronisation); its scratch
 * (the job table, leaf prefix sums 16 k (N+1) B, scan tile sums, node charges
 * and potentials 2 x 8 kp M B with kp = k rounded up to a power of two) lives in
 * one static device array of 32 MiB per GPU, held until the process ends,
 * instead of the library's pool.  No hipMalloc, hipFree or allocating call.
 * FCG_INVALID_ARGUMENT when the product needs more than that array (about
 * 16 k N + 32 kp N bytes for M ~ 2N); FCG_INVALID_ARGUMENT as well when the job
 * is not filled in FP64.  Every call on one GPU shares its array: calls that
 * overlap on one GPU (different streams without ordering) race and must be
 * serialised by the caller. */
int fcg_mvp_device_noalloc(fcg_ctx* ctx, int job, const double* d_q, double* d_v, int k, void* stream);

/* Product precisions of fcg_mvp_prec. */
#define FCG_MVP_FP64 0       /* FP64 coefficients, charges and accumulation */
#define FCG_MVP_FP32 1       /* FP32 coefficients, charges and accumulation (inner solves, mixed path) */
#define FCG_MVP_FP32_ACC64 2 /* FP32 coefficients, FP64 charges and accumulation (refinement residual) */

/* v = P q in a given precision (the job must be filled in FCG_PREC_FP64 for
 * FCG_MVP_FP64 and in FCG_PREC_MIXED for the two FP32-coefficient modes). */
int fcg_mvp_prec(fcg_ctx* ctx, int job, int mvp_prec, const double* q, double* v, int k);

/* Pq = P^-1 q (Jacobi) for k vectors, host buffers (complex jobs: [k][2N], both
 * halves divided by the real self term, as the reference does).
 * Replaces CSolveCap::ComputePrecondVectFast (Solver/SolveCapacitance.cpp:3458-3526). */
int fcg_precond_apply(fcg_ctx* ctx, int job, const double* q, double* pq, int k);

/* Pq = B q for k vectors (host buffers) with the block preconditioner B of block
 * size block_size (0 = 32, clamped to 2..128): blocks of consecutive leaves, each
 * the inverse of its dense collocation matrix.  Built on first use (the job must
 * be filled) and kept for later calls and solves.
 * Replaces CSolveCap::RecurseComputePrecond, ComputeBlockPrecond, InvertMatrix
 * (Solver/SolveCapacitance.cpp:1094-1104, 1876-2258) and the block branch of
 * ComputePrecondVectFast (:3527-3544).  A leaf met before any block is open (its
 * parent has more than block_size leaves) is a block of its own; the reference
 * opens a block there that it never computes. */
int fcg_block_precond_apply(fcg_ctx* ctx, int job, int block_size, const double* q, double* pq, int k);

/* Pq = S q for k vectors (host buffers) with the two-level preconditioner of
 * dimension dim (0 = 128, clamped to 16..4096): the tree nodes at depth
 * log2(dim / roots) of every conductor and dielectric tree (or shallower
 * leaves), the inverse of their collocation matrix, applied by collecting each
 * node's leaves (area weighted), multiplying by the off-diagonal part and
 * distributing back with the diagonal term.  Built on first use (the job must be
 * filled).  FCG_INVALID_ARGUMENT when unavailable (at least dim roots; solves
 * then run unpreconditioned, as the reference does).
 * Replaces MakePreconditioner, the super branch of RecurseComputePrecond,
 * ComputeSuperPrecond and InvertMatrix (Solver/SolveCapacitance.cpp:1067-1104,
 * 1876-2066, 2116-2258) and the super branch of ComputePrecondVectFast
 * (:3545-3592). */
int fcg_super_precond_apply(fcg_ctx* ctx, int job, int dim, const double* q, double* pq, int k);

/* ---------------------------------------------------------------- link discovery */

/* Link criterion parameters of a refined mesh (fch_get_link_params). */
typedef struct fcg_link_params {
    int32_t dim;        /* 3 or 2 */
    int32_t reserved;
    double link_eps;    /* eps ratio x mesh eps (-d, -m; 0 -> 0.01) */
    double max_side;    /* 3D: longest panel side of the input */
    double max_area;    /* 3D: largest panel area of the input */
    double max_length;  /* 2D: longest segment of the input */
} fcg_link_params;

/* Discover the link structure of a refined tree on the device: every pair of
 * group roots and every root's own tree are walked as CAutoRefine::ComputeLinks
 * does, a pair becoming a link when RefineCriteria (curvature coefficient 1) no
 * longer asks for refinement or both are leaves, else the larger node is split.
 * The criterion is evaluated as the host builder evaluates it (FP64, no fused
 * operations), so the link set equals the host builder's.  Uses d->dim, n_nodes,
 * nodes, left, right (children before their parent) and parent (-1 at roots);
 * the CSR fields are ignored.  *n_links receives L and *max_level (may be NULL)
 * the deepest recursion level of the traversal (the top calls are level 1); the
 * rows stay on the device until fcg_links_fetch.
 * Replaces ComputeLinks, RefineMutual, RefineSelf (structure part),
 * RefineCriteria and the record pass of AutoRefineLinks
 * (Solver/Autorefine.cpp:1132-1171, 3096-3447, 2115-2229, 2928-3092, 481-756). */
int fcg_links_discover(fcg_ctx* ctx, const fcg_job_desc* d, const fcg_link_params* lp, int64_t* n_links,
                       int32_t* max_level);

/* Copy the last discovery's rows to the host: row_ptr[n_nodes + 1], col[L],
 * every row sorted by source node id (the host builder records each row in
 * traversal order; the device layout sorts rows anyway).  Frees them on the
 * device. */
int fcg_links_fetch(fcg_ctx* ctx, int64_t* row_ptr, int32_t* col);

/* Create a job from a refined tree (d, CSR fields ignored) and the rows of the
 * last fcg_links_discover on it, which it consumes: the rows never leave the
 * device.  The stored layout (rows sorted by source charge row, fill classes
 * and fill order, 16-bit columns) is built on the device and equals the one
 * fcg_job_upload builds from the same links.  fcg_download_coeffs returns each
 * row ordered by source node id.  Replaces, with fcg_job_upload, the
 * AutoRefineLinks storage set-up (Solver/Autorefine.cpp:481-706). */
int fcg_job_upload_links(fcg_ctx* ctx, const fcg_job_desc* d, int* job);

/* ---------------------------------------------------------------- adaptive refinement */

/* One node of a tree under refinement (the host builder's element record):
 * children are indices into the same array, -1 for leaves. */
typedef struct fcg_refine_node {
    float v[9];           /* vertices (2D: two) */
    float c[3];           /* centroid */
    float n[3];           /* stored normal */
    float pad0;
    double dim;           /* area (3D) or length (2D) */
    double maxside;       /* 3D: longest side */
    int32_t left, right;  /* children, -1 for leaves */
    int32_t group;        /* owning group (deque order) */
    uint8_t type;         /* FCG_NODE_* flags */
    uint8_t max_side_idx; /* 3D: index of the longest side */
    uint8_t diel_index;   /* outer-medium index of a conductor panel */
    uint8_t pad1;
} fcg_refine_node;

typedef struct fcg_refine_params {
    int32_t dim;         /* 3 or 2 */
    int32_t n_groups;
    double mesh_eps;     /* -m */
    double curv;         /* -mc, clamped to >= 1 */
    double max_side;     /* 3D: longest side of the group tops */
    double max_area;     /* 3D: largest group top area */
    double max_length;   /* 2D: longest group top */
} fcg_refine_params;

/* Refine a super hierarchy on the device as CAutoRefine::AutoRefinePanels does
 * with DiscretizeMutual / DiscretizeSelf: every pair of group tops, then every
 * group's own tree, a pair's larger node split (longest-side bisection, 2D
 * halving) while RefineCriteria with the mesh eps (and, inside one group, the
 * curvature coefficient) asks for it.  A node's self pass runs after every pass
 * that can split it, as in the reference's recursion, so the refined tree is the
 * host builder's.  nodes[n_nodes] with tops[n_groups] (the group roots); *n_out
 * receives the refined node count (the input nodes keep their indices, new ones
 * follow), *max_mesh_eps the largest relative estimate of an accepted pair of
 * non-super nodes and *max_level the deepest recursion level.  The refined nodes
 * stay on the device until fcg_refine_fetch.
 * Replaces AutoRefinePanels' discretization, DiscretizeMutual, DiscretizeSelf,
 * RefineCriteria and CAutoPanel/CAutoSegment::Subdivide
 * (Solver/Autorefine.cpp:293-479, 2638-2921, 2928-3092; AutoPanel.cpp:382-523;
 * AutoSegment.cpp:152-230). */
int fcg_refine(fcg_ctx* ctx, const fcg_refine_node* nodes, int n_nodes, const int32_t* tops,
               const fcg_refine_params* p, int* n_out, double* max_mesh_eps, int32_t* max_level);

/* Copy the refined nodes (n_out of the last fcg_refine) to the host and free
 * them on the device. */
int fcg_refine_fetch(fcg_ctx* ctx, fcg_refine_node* out);

/* Solve every conductor right-hand side of the job with right-preconditioned GMRES
 * and assemble the Maxwell capacitance matrix cap[rows][n_cond] (row-major; rows =
 * n_cond in 3D, n_cond-1 in 2D; 2D values unscaled, as the reference keeps them
 * before OutputCapMtx).  Complex jobs (fcg_job_desc *_im arrays): cap holds the
 * real matrix followed by the imaginary one (the reference's cRe, cIm), charges
 * [rows][2N] ([Re | Im]); Jacobi or no preconditioner only.  Replaces
 * CSolveCap::Solve + gmresPrecondSFastAll (Solver/SolveCapacitance.cpp:1138-1425,
 * 2288-2574). */
int fcg_solve(fcg_ctx* ctx, int job, const fcg_solve_opts* opts, double* cap, int* rows, int* cols,
              double* charges /* [rows][N] or NULL */, fcg_solve_stats* stats);

/* Average per-stage device time of the hierarchical product, measured with HIP
 * events on the context stream over `reps` products of k pseudo-random device
 * vectors.  No reference counterpart: plays the role of the reference's phase
 * timers (Solver/SolveCapacitance.cpp:704-720) for the product alone. */
typedef struct fcg_mvp_timing {
    double ms_scan;    /* up-pass: every node charge (one kernel; top nodes included) */
    double ms_charge;  /* 0: the node charges are part of the up-pass */
    double ms_gather;  /* link gather (the dominant kernel; with each job's top stage of
                          the down-pass when the gather is staged) */
    double ms_down;    /* down-pass (+ complex coupling, 2D fix-up) */
    double ms_total;   /* whole product */
    int32_t reps, k;
    int32_t variant;   /* gather configuration in use */
} fcg_mvp_timing;
int fcg_mvp_timing_run(fcg_ctx* ctx, int job, int k, int reps, fcg_mvp_timing* out);

/* Device time and algorithmic bytes of the gathers made by solves on this
 * context, accumulated per product precision (index FCG_MVP_*) while profiling is
 * on (HIP events around the gather of every product of the solve engine; a
 * product's gather is one kernel launch per rhs class of its jobs, counted in
 * gather_kernels).  Algorithmic bytes of one product's gather = sum over its jobs
 * of L (b_coef + b_idx) + 16 M + 8 + 2 M k b_node (b_coef, b_node = 8 for FP64
 * and 4 for FP32 coefficients / node values; b_idx = 2 for jobs of at most 65536
 * nodes, whose source indices are 16-bit, else 4).  No reference counterpart (a
 * measurement hook). */
typedef struct fcg_profile {
    double gather_ms[3];
    double gather_bytes[3];
    int64_t gather_launches[3];
    double mvp_ms[3];        /* whole products (scan, charges, gather, down-pass) */
    int64_t mvp_launches[3];
    int64_t steps;           /* engine steps */
    double host_plan_ms;     /* host time building tables and queueing launches */
    double host_wait_ms;     /* host time blocked on the step's synchronisation */
    double host_post_ms;     /* host time in Givens updates and result bookkeeping */
    int64_t gather_kernels[3]; /* gather kernel launches (gather_launches counts products) */
} fcg_profile;
/* enable: 1 on, 0 off, -1 unchanged; reset clears the totals after copying them to out (may be NULL). */
int fcg_ctx_profile(fcg_ctx* ctx, int enable, int reset, fcg_profile* out);
/* The same for one batched product over several filled jobs, each on as many
 * random right-hand sides as it has conductor solves, in precision mvp_prec
 * (FCG_MVP_*); out->k is the total right-hand side count. */
int fcg_mvp_timing_batch(fcg_ctx* ctx, const int* jobs, int njobs, int mvp_prec, int reps, fcg_mvp_timing* out);

/* Time every gather configuration (lanes per link, right-hand sides per lane,
 * links in flight) on this job with k right-hand sides and keep the fastest for
 * the job's later products.  Without it a job gathers with its rhs class's
 * default configuration; a sweep costs about a hundred products, so it pays
 * only for jobs solved many times.  No reference counterpart (the reference has
 * a single CPU loop). */
int fcg_autotune(fcg_ctx* ctx, int job, int k, int reps, int* variant, double* best_ms);

/* Pin the gather configuration of a job (variant < 0 only queries the count). */
int fcg_set_gather_variant(fcg_ctx* ctx, int job, int variant, int* n_variants);

/* Products over several jobs gather each rhs class (jobs whose padded conductor
 * count is the same power of two) in a launch of its own.  Time every gather
 * configuration on each class present in a product over the filled jobs (at
 * least two) in precision mvp_prec (FCG_MVP_*) and keep each class's fastest for
 * the context's products over several jobs (the solve engine's steps).  *variant
 * receives the choice of the class with the most rows, *best_ms the gather time
 * with every class's choice.  No reference counterpart. */
int fcg_autotune_batch(fcg_ctx* ctx, const int* jobs, int njobs, int mvp_prec, int reps, int* variant,
                       double* best_ms);
/* Pin the gather configuration of every rhs class of the context's products over
 * several jobs; variant < 0 restores each class's default configuration. */
int fcg_set_batch_variant(fcg_ctx* ctx, int mvp_prec, int variant);

/* Solve several jobs together: every GMRES step of every job is advanced by one
 * batched launch per stage.  caps receives the per-job matrices back to back
 * (rows[i] x cols[i] each, in job order); stats may be NULL or hold njobs entries.
 * Replaces the per-conductor sequential loop of CSolveCap::Solve
 * (Solver/SolveCapacitance.cpp:1198-1419) extended across jobs. */
int fcg_solve_batch(fcg_ctx* ctx, const int* jobs, int njobs, const fcg_solve_opts* opts, double* caps, int* rows,
                    int* cols, fcg_solve_stats* stats);

/*
 * Serving engine: continuous batching of solves.  Jobs submitted while others are
 * in flight join at the next step; each step advances every active solve by one
 * Arnoldi step (or one refinement residual) with one launch per stage for all of
 * them.  Finished results are popped with fcg_engine_poll.  No reference
 * counterpart (the reference solves one geometry per process run).
 */
typedef struct fcg_engine fcg_engine;
int fcg_engine_create(fcg_ctx* ctx, const fcg_solve_opts* opts, fcg_engine** out);
int fcg_engine_submit(fcg_engine* eng, int job, int64_t tag);
/* Advance all active solves by one step; returns the number still active and the
 * number of finished results waiting to be polled. */
int fcg_engine_step(fcg_engine* eng, int* active, int* ready);
/* Pop one finished result; returns FCG_OK, or FCG_INVALID_ARGUMENT when none is ready. */
int fcg_engine_poll(fcg_engine* eng, int64_t* tag, int* job, double* cap, int cap_len, int* rows, int* cols,
                    fcg_solve_stats* stats);
/* Device bytes held by the engine's active solves (admission control). */
int fcg_engine_bytes(fcg_engine* eng, int64_t* bytes);
int fcg_engine_destroy(fcg_engine* eng);

/* Device bytes a job occupies (structure + coefficients of `precision`) and an
 * estimate of the solve working set for `iters` Arnoldi steps. */
int fcg_job_footprint(fcg_ctx* ctx, int job, int precision, int iters, int64_t* resident, int64_t* solve);

/* Number of conductors (non-dielectric groups) of a job. */
int fcg_job_conductors(fcg_ctx* ctx, int job, int* n_cond);

/* 1 when the job has complex permittivities (its vectors are [Re | Im] and its
 * capacitance matrices come back as real then imaginary part), else 0.  No
 * reference counterpart (the reference's m_ucHasCmplxPerm, Solver/SolverGlobal.h:169). */
int fcg_job_is_complex(fcg_ctx* ctx, int job, int* is_complex);

/* Last error message of the context (static storage, valid until the next call). */
const char* fcg_last_error(fcg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* FASTERCAP_AMD_H */
