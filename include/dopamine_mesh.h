/* Mesh setup C ABI of libdopamine_hip.so.
 *
 * Builds the cell-centred unstructured finite-volume mesh that every hot-path kernel reads, from a
 * Gmsh-style element list (nodes, volume elements, boundary elements tagged by physical group).
 * The construction follows the reference's mesh pipeline:
 *   - face matching and per-element geometry: src/mesh/mod_gmsh_reader.f90:239-616, 737-977
 *   - cell->face adjacency, face geometry weights: src/mesh/mod_mesh.f90:215-259, 437-509
 *   - periodic patch pairing: src/bc/mod_bc_setup.f90:134-330
 *   - least-squares gradient stencil and weights: src/mesh/mod_mesh.f90:513-758
 * so that a mesh built here is the mesh the reference builds from the same .msh contents.
 *
 * Conventions: every index is 0-based (the reference is 1-based); a face neighbour of -1 marks a
 * boundary face; boundary face b is global face n_int_faces + b; boundary faces are contiguous per
 * patch. Arrays of 3-vectors are stored component-fastest (x0 y0 z0 x1 y1 z1 ...).
 */
#ifndef DOPAMINE_MESH_H
#define DOPAMINE_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary-condition ids, identical to the reference (src/mod_parameters.f90:43-54). */
enum {
  DM_BC_INLET = 1,
  DM_BC_OUTLET = 2,
  DM_BC_WALL_NOSLIP = 3,
  DM_BC_WALL_MODEL = 4,
  DM_BC_SLIP = 5,
  DM_BC_PERIODIC = 6,
  DM_BC_SYMMETRY = 7,
  DM_BC_FIXEDTEMP = 8,
  DM_BC_ZEROGRADIENT = 9,
  DM_BC_WALL_VELOCITY = 10,
  DM_BC_ABL_INLET = 11
};

/* Gmsh element type codes accepted by the builder (src/mesh/mod_gmsh_reader.f90:14-23). */
enum { DM_GMSH_TRI3 = 2, DM_GMSH_QUAD4 = 3, DM_GMSH_TET4 = 4, DM_GMSH_HEX8 = 5, DM_GMSH_PRISM6 = 6, DM_GMSH_PYR5 = 7 };

typedef struct dm_mesh_input {
  int64_t n_nodes;
  const double* coords;       /* 3*n_nodes */
  int64_t n_vol;              /* volume elements, in file order */
  const int32_t* vol_type;    /* Gmsh type code per volume element */
  const int64_t* vol_offset;  /* n_vol+1 offsets into vol_nodes */
  const int32_t* vol_nodes;   /* 0-based node indices, Gmsh node order */
  int64_t n_bnd;              /* boundary (2-D) elements, in file order */
  const int32_t* bnd_type;
  const int64_t* bnd_offset;  /* n_bnd+1 */
  const int32_t* bnd_nodes;
  const int32_t* bnd_group;   /* 0-based 2-D physical group of each boundary element */
  int32_t n_groups;           /* number of 2-D physical groups */
  const int32_t* group_bc;    /* bc id per physical group (patch type after name matching) */
} dm_mesh_input;

/* Read-only view of a built mesh; every pointer stays valid until dm_mesh_free. */
typedef struct dm_mesh_view {
  int32_t n_cells, n_faces, n_int_faces, n_bnd_faces, n_patches, n_vertices;
  int64_t n_ls;               /* total least-squares stencil entries */
  int32_t max_cell_faces;
  int32_t ls_degenerate;      /* cells whose LS normal matrix needed regularisation */
  /* cells */
  const double* cell_centroid;      /* 3*nc */
  const double* cell_volume;        /* nc */
  const int32_t* cell_face_offset;  /* nc+1 */
  const int32_t* cell_faces;        /* cell_face_offset[nc] entries, ascending face order per cell */
  /* faces */
  const int32_t* face_owner;        /* nf */
  const int32_t* face_neighbour;    /* nf, -1 on boundary faces */
  const int32_t* face_patch;        /* nf, -1 on interior faces */
  const double* face_area;
  const double* face_centre;        /* 3*nf */
  const double* face_normal;        /* 3*nf unit, owner -> neighbour / outward */
  const double* face_sf;            /* 3*nf area vector */
  const double* face_dcf;
  const double* face_gc;            /* owner interpolation weight; 1 on boundary faces */
  const double* face_skew;
  const double* face_eps;           /* 3*nf skewness offset; 0 on boundary faces */
  /* patches */
  const int32_t* patch_bc;          /* np */
  const int32_t* patch_start;       /* first boundary-face index of each patch */
  const int32_t* patch_nfaces;
  /* periodic pairing, per boundary face */
  const int32_t* periodic_pair;     /* partner boundary face or -1 */
  const double* periodic_pair_dcf;  /* partner owner->face distance (own d_cf when unpaired) */
  const double* periodic_gc;        /* owner weight d_pair/(d_own+d_pair+SMALL); 0.5 when unpaired */
  /* least-squares stencil */
  const int32_t* ls_offset;         /* nc+1 */
  const int32_t* ls_cell;           /* neighbour cell, or -1 for a boundary pseudo-neighbour */
  const int32_t* ls_face;           /* global face of a boundary pseudo-neighbour, else -1 */
  const double* ls_weight;          /* 3*n_ls */
  /* vertices (for output and tests) */
  const double* vertex_coords;      /* 3*n_vertices */
  const int64_t* cell_vertex_offset;
  const int32_t* cell_vertices;
  /* face point lists, ordered so that the right-hand normal points along face_normal */
  const int64_t* face_node_offset;  /* nf+1 */
  const int32_t* face_nodes;
} dm_mesh_view;

/* Reverse Cuthill-McKee order of a built mesh's cells over its face graph (interior faces and
 * periodic pairs), order[new] = old, for renumbering the cells of shuffled meshes before the
 * final build. No counterpart in the reference, whose multi-rank partitioner orders cells along a
 * Morton curve instead (build_morton_partition, src/mesh/mod_partitioner.f90:39). Returns 0, or
 * -1 with dm_last_error(). */
int dm_mesh_rcm_order(const void* mesh, int32_t* order);

/* Build a mesh. Returns an opaque handle, or NULL with a message available from dm_last_error(). */
void* dm_mesh_build(const dm_mesh_input* in);
/* Fill *view with the arrays of a built mesh; returns 0. */
int dm_mesh_get_view(const void* mesh, dm_mesh_view* view);
void dm_mesh_free(void* mesh);
/* Last error message of this library (thread-local); empty when none. */
const char* dm_last_error(void);

/* Face-based (polyhedral) input, the OpenFOAM polyMesh layout: internal faces first, each with
 * owner < neighbour, then the boundary faces, every one inside exactly one patch range.
 * Geometry follows the reference's polyhedral path (src/mesh/mod_mesh.f90:1336-1461: fan
 * triangulated face area, normal and area-weighted centre; cell volume and centroid from signed
 * tets about a per-cell reference vertex), then the same adjacency, weights, periodic pairing and
 * least-squares stencil as dm_mesh_build. Faces keep their input order. */
typedef struct dm_poly_input {
  int64_t n_points;
  const double* points;        /* 3*n_points */
  int64_t n_faces;
  const int64_t* face_offset;  /* n_faces+1 offsets into face_nodes */
  const int32_t* face_nodes;   /* 0-based point labels, right-hand normal owner -> neighbour / outward */
  int64_t n_int_faces;
  const int32_t* owner;        /* n_faces */
  const int32_t* neighbour;    /* n_int_faces */
  int32_t n_patches;
  const int64_t* patch_start;  /* first (global) face of each patch */
  const int64_t* patch_nfaces;
  const int32_t* patch_bc;     /* bc id per patch (patch type after name matching) */
} dm_poly_input;
/* NULL with dm_last_error() on a negative cell index, owner >= neighbour, a point label out of
 * range, a boundary face in no patch or in two, or a patch range outside the boundary faces. */
void* dm_mesh_build_poly(const dm_poly_input* in);

/* OpenFOAM polyMesh reader (foam_read, src/mesh/mod_foam_reader.f90:16-97): points, faces,
 * owner, neighbour, boundary of a constant/polyMesh directory, ASCII or binary (label 32/64,
 * scalar 32/64, LSB/MSB, compact or plain face lists). The view is a dm_poly_input whose patch_bc
 * holds the bc id of each patch's `type` entry (bc_type_to_id, :1144-1167); the case
 * configuration may override it by patch name. NULL with dm_foam_last_error() on an unreadable
 * or malformed file and on the reference's topology checks (negative cell index, owner >=
 * neighbour, unmapped boundary face, patch range outside the boundary faces, odd number of
 * cyclic patches). */
void* dm_foam_read(const char* polymesh_dir);
int dm_foam_get_view(const void* file, dm_poly_input* view);
const char* dm_foam_patch_name(const void* file, int patch);
const char* dm_foam_patch_type(const void* file, int patch);
/* bc id of an OpenFOAM patch type (bc_type_to_id; unknown types give zeroGradient). */
int32_t dm_foam_bc_of_type(const char* type);
void dm_foam_free(void* file);
const char* dm_foam_last_error(void);

/* Gmsh v4 ASCII reader (gmsh_read, src/mesh/mod_gmsh_reader.f90:45-236): nodes in file order,
 * 3-D elements as volume elements in file order, 2-D elements as boundary elements tagged by
 * their entity's first physical group, numbered among the 2-D physical groups in $PhysicalNames
 * order. The view's arrays are the builder's dm_mesh_input fields (group_bc is the caller's:
 * patch types come from the case configuration, by group name). */
typedef struct dm_gmsh_view {
  int64_t n_nodes;
  const double* coords;
  int64_t n_vol;
  const int32_t* vol_type;
  const int64_t* vol_offset;
  const int32_t* vol_nodes;
  int64_t n_bnd;
  const int32_t* bnd_type;
  const int64_t* bnd_offset;
  const int32_t* bnd_nodes;
  const int32_t* bnd_group;
  int32_t n_groups;
} dm_gmsh_view;
/* NULL on error (dm_gmsh_last_error()): unreadable file, not v4 ASCII, unsupported element type,
 * a boundary element outside every 2-D physical group. */
void* dm_gmsh_read(const char* path);
int dm_gmsh_get_view(const void* file, dm_gmsh_view* view);
/* Name of 2-D physical group g, or NULL. */
const char* dm_gmsh_group_name(const void* file, int group);
void dm_gmsh_free(void* file);
const char* dm_gmsh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
