/*
 * mhdg host-side mesh setup: complex construction, geometry, boundary classification, background field
 * and seeded synthetic states.  These run once per mesh on the host (the reference does the same work in
 * its preprocessing and solver constructors); the per-step path is in mhdg.h.
 */
#ifndef MHDG_MESH_H
#define MHDG_MESH_H

#include <stdint.h>

#include "mhdg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhdg_mesh mhdg_mesh;

/* Synthetic shell (no reference counterpart; the reference generates hexahedral boxes and cubed-sphere
 * shells, src/uns/generate/*.cpp): lat-lon net n_lat x n_lon, n_r geometric radial layers on
 * [r_in, r_out] with thickness ratio `ratio`, interior nodes jittered by `jitter` of the local spacing,
 * every hexahedron split into 6 tetrahedra (3 at the poles).  Tags: 1 = inner, 2 = outer. */
int mhdg_mesh_shell(int n_lat, int n_lon, int n_r, double r_in, double r_out, double ratio, double jitter,
                    uint64_t seed, mhdg_mesh **out);

/* Generic polyhedral mesh: cell c has face loops [cell_ptr[c], cell_ptr[c+1]), loop l has the nodes
 * loop_nodes[loop_ptr[l] .. loop_ptr[l+1]) counter-clockwise seen from outside (3 or 4 nodes).
 * Replaces uns::ComplexBuilder::add_cell/tag_face/finish (src/uns/complex/ComplexBuilder.cpp:45-134)
 * followed by uns::compute_geometry (src/uns/geometry/Geometry.cpp:17-169). */
int mhdg_mesh_from_cells(int32_t n_nodes, const double *xyz, int64_t n_cells, const int32_t *cell_ptr,
                         const int32_t *loop_ptr, const int32_t *loop_nodes, const int32_t *material,
                         int64_t n_tags, const int32_t *tag_ptr, const int32_t *tag_nodes, const int32_t *tag_id,
                         mhdg_mesh **out);
void mhdg_mesh_free(mhdg_mesh *m);

/* Partition-ordered numbering for one GPU.  mhdg_mesh_hilbert_order: the cells sorted by the Hilbert key of their
 * centres (the reference's SFC partitioner, src/uns/partition/Partitioner.cpp:37-108: 20-bit coordinates in the
 * bounding box of the centres, Skilling's transpose, ties by id): order[k] = the cell placed at k.
 * mhdg_mesh_renumber_cells: the same mesh (nodes, orientation, tags) with its cells in that order and its faces and
 * edges numbered by first appearance, the reference builder's rule (ComplexBuilder.cpp:45-134); every cell keeps
 * its face loops, so a face keeps its loop and an edge its node pair: new c2f row k = old c2f row order[k] up to
 * the face ids, with equal signs.  Tetrahedral shells (mhdg_mesh_shell) only. */
int mhdg_mesh_hilbert_order(const mhdg_mesh *m, int32_t *order);
int mhdg_mesh_renumber_cells(const mhdg_mesh *m, const int32_t *order, mhdg_mesh **out);
const char *mhdg_mesh_error(void);

/* counts[0..7] = nodes, edges, faces, cells, nnz(f2e), nnz(c2f), nnz(e2f), nnz(edge->cell) */
int mhdg_mesh_counts(const mhdg_mesh *m, int64_t counts[8]);
/* Pointer to an internal array by name (valid until mhdg_mesh_free): "xyz", "tets", "e2n.ptr", "e2n.idx",
 * "e2n.sgn", likewise f2e/c2f/e2f/f2c, "face_tag", "material", "face_kind", "edge_mode", and every
 * geometry array of mhdg_mesh_desc ("face_area", ..., "star1", "star2", "star3", "edge_dual_area",
 * "cell_min_tet").  *count receives the element count; *elem_size the element size in bytes. */
int mhdg_mesh_array(const mhdg_mesh *m, const char *name, const void **ptr, int64_t *count, int32_t *elem_size);

/* Boundary classification (mhd::BoundaryInfo::build, Common/boundary/BoundaryInfo.cpp:9-49).
 * tag_type[t] for tag id t (0..n_tag_types-1): 0 wall, 1 inflow, 2 outflow, 3 farfield;
 * untagged boundary faces use `default_type`. */
int mhdg_mesh_boundary(mhdg_mesh *m, const int32_t *tag_type, int32_t n_tag_types, int32_t default_type,
                       const double u_inf[3]);
/* Fills a descriptor that points into the mesh's arrays (requires mhdg_mesh_boundary). */
int mhdg_mesh_desc_fill(const mhdg_mesh *m, mhdg_mesh_desc *desc);

/* Static background (mercury::set_background, MercuryUns/initial/Initial.cpp:66-84): PhiAdd = B_imf.A_f
 * plus d1 of the dipole potential A = m x (r-c)/(|r-c|^2+soft^2)^{3/2} along the edges; BaddCell its
 * reconstructed cell vectors. */
int mhdg_mesh_background(const mhdg_mesh *m, const double B_imf[3], int dipole, const double dip_m[3],
                         const double dip_c[3], double soft, double *PhiAdd, double *BaddCell);
/* Background with a crustal field of point dipoles (mars::MarsSolver::build_background_,
 * MarsUns/driver/MarsSolver.cpp:169-200; MARS::DipoleModel, MarsUns/physics/MarsDipoles.cpp:63-107):
 * PhiAdd = B_imf.A_f + d1(E_A), E_A the edge integrals of A = sum_i m_i x (x-c_i)/(|x-c_i|^2+core_i^2)^{3/2}
 * (compensated sum over the dipoles in order, two-point Gauss rule on each edge, dipoles farther than `cutoff`
 * skipped, cutoff 0 = none); BaddCell its reconstructed cell vectors.  pos, mom 3n; core n (nondimensional). */
int mhdg_mesh_background_dipoles(const mhdg_mesh *m, const double B_imf[3], int32_t n, const double *pos,
                                 const double *mom, const double *core, double cutoff, double *PhiAdd,
                                 double *BaddCell);
/* The edge integrals alone for edges from xa[3e] to xb[3e] (MarsDipoles.cpp:92-107). */
int mhdg_dipole_edge_integrals(int32_t n, const double *pos, const double *mom, const double *core, double cutoff,
                               int64_t n_edges, const double *xa, const double *xb, double *out);
/* uns::reconstruct_cell_vector on the host (Reconstruct.cpp:8-46). */
int mhdg_mesh_cell_vector(const mhdg_mesh *m, const double *flux, double *b);
/* Seeded synthetic state: free stream fs[species] in every cell with relative Gaussian noise `noise` on
 * rho and p and noise*|u| on each velocity component, and an induced flux Phi = d1(A) from a seeded edge
 * potential scaled so that rms(Phi/|A_f|) = noise * rms|BaddCell|.  mt19937_64(seed). */
int mhdg_mesh_seeded_state(const mhdg_mesh *m, const mhdg_config *cfg, double noise, uint64_t seed,
                           const double *BaddCell, double *U_H, double *U_Na, double *Phi);

/* ------------------------------------------------------------------------ UNSBIN mesh and field files */
/* The reference's container (include/uns/io/BinContainer.h:1-6, docs/FileFormats.md:3-9).  Complex files as
 * uns::io::write_local_mesh / read_local_mesh (src/uns/io/MeshIO.cpp:38-115) for one rank holding the whole
 * mesh: the stored incidences, numbering and orientation are used as they are, geometry is recomputed, and
 * the stored topology hash is verified.  A part file of a several-rank partition is refused. */
int mhdg_mesh_read_unsbin(const char *path, mhdg_mesh **out);
/* The n part files part_00000.unsb .. of an n-rank partition (uns::io::write_local_mesh of each
 * uns::Distributor::build(r), src/uns/parallel/Distributor.cpp:115-221, written by the reference's preprocessing
 * pipeline, Preprocess/Pipeline.cpp:133-156) merged into the one global complex they were cut from, in its global
 * numbering: every entity's row comes from the file of its owner (the first n_owned entities of each dimension), its
 * column ids mapped through that file's global ids, signs as stored; each file's topology hash and every ghost copy are
 * checked against the merge. Geometry is recomputed. No counterpart in the reference, whose ranks each read one part:
 * it lets one GPU run the reference's partitioned preprocessing output. */
int mhdg_mesh_read_unsbin_parts(const char *const *paths, int32_t n_parts, mhdg_mesh **out);
int mhdg_mesh_write_unsbin(const mhdg_mesh *m, const char *path);
/* uns::topology_hash (src/uns/complex/CellComplex.cpp:56-70). */
int mhdg_mesh_topology_hash(const mhdg_mesh *m, uint64_t *hash);
/* Face-tag names, newline-joined, index = tag id (0 reserved): the `tag_names` record. */
int mhdg_mesh_tag_names(const mhdg_mesh *m, char *buf, int buflen);
int mhdg_mesh_set_tag_names(mhdg_mesh *m, const char *newline_joined);

/* Field files as uns::io::write_fields / read_fields (MeshIO.cpp:117-151): entity-numbered arrays
 * (dim 0 node, 1 edge, 2 face, 3 cell; ncomp values per entity) and the topology hash of their complex. */
typedef struct mhdg_field {
    const char *name;
    int32_t dim, ncomp;
    int64_t count;          /* number of doubles */
    const double *values;
} mhdg_field;
int mhdg_fields_write(const char *path, uint64_t hash, const mhdg_field *fields, int32_t n_fields);
typedef struct mhdg_fields mhdg_fields;
int mhdg_fields_read(const char *path, mhdg_fields **out, uint64_t *hash, int32_t *n_fields);
/* Field i of a file read by mhdg_fields_read; the pointers stay valid until mhdg_fields_free. */
int mhdg_fields_get(const mhdg_fields *f, int32_t i, mhdg_field *field);
void mhdg_fields_free(mhdg_fields *f);

/* VTK XML output as uns::io::write_vtu / write_pvtu / write_pvd (src/uns/io/VtkWriter.h:20-26,
 * VtkWriter.cpp:42-116): every cell a polyhedron, raw appended binary payloads; byte-identical with the
 * reference's writer for the same mesh and arrays.  values: ncomp doubles per cell (cell arrays) or per
 * node (point arrays). */
typedef struct mhdg_vtk_array {
    const char *name;
    int32_t ncomp;
    const double *values;
} mhdg_vtk_array;
int mhdg_mesh_write_vtu(const mhdg_mesh *m, const char *path, int32_t n_cell_arrays, const mhdg_vtk_array *cell_arrays,
                        int32_t n_point_arrays, const mhdg_vtk_array *point_arrays);
int mhdg_write_pvtu(const char *path, int32_t n_pieces, const char *const *pieces, int32_t n_cell_arrays,
                    const mhdg_vtk_array *cell_arrays, int32_t n_point_arrays, const mhdg_vtk_array *point_arrays);
int mhdg_write_pvd(const char *path, int32_t n, const double *times, const char *const *files);

/* Agglomeration hierarchy of the multigrid preconditioner (precision 2 of mhdg_ind_solve / mhdg_step), built on the
 * host as the device context builds it: levels of node aggregates by heavy-edge matching (the reference's
 * agglomeration, MercuryUns/multigrid/Multigrid.cpp:219-282, on the node graph), coarse unknowns = pairs of adjacent
 * aggregates, prolongation +-1 per finer edge (0 inside an aggregate and on Dirichlet edges), Galerkin coarse
 * operators M_l (diagonal) and K_l (CSR) of I + dt K, K the resistive curl-curl of ImplicitInduction::apply_F.
 * cfg supplies mg_levels, mg_agglom_passes and the resistivity; the descriptor must carry e2n.  For tests and
 * inspection; no reference counterpart (the reference has no induction multigrid). */
typedef struct mhdg_mg mhdg_mg;
int mhdg_mg_build(const mhdg_mesh_desc *mesh, const mhdg_config *cfg, const int8_t *dirichlet, mhdg_mg **out);
void mhdg_mg_free(mhdg_mg *mg);
int mhdg_mg_levels(const mhdg_mg *mg);
/* n unknowns, n_nodes aggregates (level 0: the mesh's edges and nodes), k_nnz entries of K_l (0 on level 0) */
int mhdg_mg_level_size(const mhdg_mg *mg, int level, int32_t *n, int32_t *n_nodes, int64_t *k_nnz);
/* Level >= 1: map (n of the finer level: signed coarse index, v < 0 = -1 on ~v, INT32_MIN none), mass (n), K_l CSR
 * (k_ptr n+1, k_col and k_val k_nnz).  Level 0: mass only.  Any output may be NULL. */
int mhdg_mg_level_arrays(const mhdg_mg *mg, int level, int32_t *map, double *mass, int32_t *k_ptr, int32_t *k_col, double *k_val);
/* End nodes of every unknown of a level (tail, head: n each; level 0 the mesh's edge orientation) and, level >= 1, the
 * aggregate of every node of the finer level (agg: n_nodes of the finer level).  Any output may be NULL. */
int mhdg_mg_level_nodes(const mhdg_mg *mg, int level, int32_t *tail, int32_t *head, int32_t *agg);

#ifdef __cplusplus
}
#endif
#endif
