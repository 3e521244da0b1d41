/*
 * fastercap_host.h - host side of the MI355X FasterCap solve path.
 *
 * Reads a FastCap-format geometry file, builds the super hierarchy, adaptively
 * discretizes the panels and discovers the hierarchical links, exactly as the
 * reference's CAutoRefine does before its link-fill loop, then flattens the result
 * into the plain arrays of fcg_job_desc (include/fastercap_amd.h).  This is the
 * caller of the device C-ABI; it performs no coefficient fill and no solve.
 *
 * Reference interfaces mirrored:
 *   fch_build   ReadFastCapFile + BuildSuperHierarchy + AutoRefinePanels + the link
 *               record pass of AutoRefineLinks (Solver/Autorefine.cpp:4568-4674,
 *               6410-6544, 293-479, 481-756; RecurseIndex :1173-1220)
 *   fch_options CAutoRefGlobalVars mesh fields and defaults (Solver/SolverGlobal.cpp:238-279)
 *
 * With fch_options.links_on_device the builder stops after the discretization
 * and the node indexing; the caller discovers the links on the device
 * (fcg_links_discover with fch_get_link_params, then fcg_links_fetch) and hands
 * them back with fch_set_links.
 *
 * With fch_options.refine_on_device it stops earlier, after the super
 * hierarchy: fch_get_coarse_tree gives the tree and parameters for fcg_refine,
 * and fch_set_refined_tree takes the refined tree back and finishes the build
 * (node indexing, no link passes: the links then come from the device as
 * above).  fch_get_desc fails on a mesh waiting for its refined tree.
 */
#ifndef FASTERCAP_HOST_H
#define FASTERCAP_HOST_H

#include "fastercap_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fch_options {
    double mesh_eps;   /* -m  m_dMeshEps (default 0.1) */
    double eps_ratio;  /* -d  m_dEpsRatio (default 1.0); link eps = eps_ratio * mesh_eps */
    double curv_coeff; /* -mc m_dMeshCurvCoeff (default 3.0, clamped to >= 1) */
    int32_t verbose;   /* -v */
    int32_t links_on_device; /* 1: no link passes here; see fch_set_links */
    int32_t refine_on_device; /* 1: stop after the super hierarchy; see fch_get_coarse_tree */
} fch_options;

typedef struct fch_stats {
    int64_t input_raw_cond_panels, input_raw_diel_panels, input_panels;
    int64_t panels, nodes, links;     /* N, M, L (links exclude self terms) */
    int32_t n_conductors, n_dielectrics, dim;
    int32_t max_level;                /* deepest recursion level reached */
    double max_mesh_eps;              /* Max Mesh relative refinement value */
    double t_read_s, t_superh_s, t_discretize_s, t_links_s;
} fch_stats;

typedef struct fch_mesh fch_mesh;

void fch_default_options(fch_options* opts);

/* Build a refined mesh with its link structure. Returns an FC_* code. */
int fch_build(const char* path, const fch_options* opts, fch_mesh** out);

/* Parameters of the link criterion of this mesh (for fcg_links_discover). */
int fch_get_link_params(const fch_mesh* mesh, fcg_link_params* params);

/* Install a link structure discovered elsewhere: row_ptr[nodes + 1] (row_ptr[0]
 * = 0, non-decreasing, row_ptr[nodes] = n_links) and col[n_links] (node ids).
 * max_level (the traversal's deepest recursion level) and t_links_s go into the
 * statistics.  With row_ptr and col both NULL only the statistics are recorded
 * (the rows stay on the device, fcg_job_upload_links). */
int fch_set_links(fch_mesh* mesh, const int64_t* row_ptr, const int32_t* col, int64_t n_links, int32_t max_level,
                  double t_links_s);

/* The coarse tree of a mesh built with refine_on_device (pre-order: children
 * after their parent), its group tops and the refinement parameters; valid while
 * the mesh waits for fch_set_refined_tree. */
int fch_get_coarse_tree(const fch_mesh* mesh, const fcg_refine_node** nodes, int* n_nodes, const int32_t** tops,
                        fcg_refine_params* params);

/* Finish a refine_on_device build with the refined tree (fcg_refine_fetch: the
 * coarse nodes first, in their order, then the new ones), the largest accepted
 * relative estimate and the recursion depth of the refinement. */
int fch_set_refined_tree(fch_mesh* mesh, const fcg_refine_node* nodes, int n_nodes, double max_mesh_eps,
                         int32_t max_level, double t_discretize_s);

/* Point desc at the mesh's arrays (valid while the mesh lives). */
int fch_get_desc(const fch_mesh* mesh, fcg_job_desc* desc);

int fch_get_stats(const fch_mesh* mesh, fch_stats* stats);

/* Write the refined mesh in FastCap format, as the reference's -o (suffix
 * "ref", charges NULL) and -c (suffix = a conductor's name, charges = that
 * conductor's solve, [panels], raw; written as densities, 2D scaled by 2 pi e0):
 * <filein without extension>_<suffix>.lst naming one panel file per conductor
 * surface (outer medium) and per dielectric.  Replaces
 * CAutoRefine::OutputFastCapFile, OutputPanelFile and OutputPanelTree
 * (Solver/Autorefine.cpp:1230-1470). */
int fch_write_fastcap(const fch_mesh* mesh, const char* filein, const char* suffix, const double* charges);

/* Names of the groups in deque order (dielectrics first, as the reference orders them). */
const char* fch_group_name(const fch_mesh* mesh, int group);

/* Last error text of fch_build (thread-local static storage). */
const char* fch_last_error(void);

void fch_free(fch_mesh* mesh);

#ifdef __cplusplus
}
#endif
#endif /* FASTERCAP_HOST_H */
