/*
 * xcal_mesh.h -- host-side mesh ingest and generation (C ABI).
 *
 * These builders produce meshes in the exact array layout of XCALibre's
 * Mesh2/Mesh3 types (src/Mesh/Mesh_0_types.jl:135-185): 1-based Int64 indices,
 * Float64 geometry, boundary-first face numbering, internal faces with
 * ownerCells[1] < ownerCells[2].  The arrays are what the Julia side would
 * hand to xcal_mesh_create() (xcal.h), so the device library never depends on
 * which builder produced a mesh.
 *
 * Builders (each restates the reference's own algorithm, same loop orders):
 *   xcalm_foam_from_arrays -- FOAM3D_mesh on in-memory polyMesh data
 *                             (src/FoamMesh/FoamMesh_1_read.jl:24-41,
 *                              FoamMesh_2_connect.jl, FoamMesh_3_generate.jl,
 *                              FoamMesh_4_geometry.jl)
 *   xcalm_read_foam        -- FOAM3D_mesh(path; scale) on an ASCII polyMesh dir
 *                             (src/FoamMesh/FoamMesh_5_build.jl:21-27)
 *   xcalm_box3d            -- seeded jittered + randomly renumbered hex box,
 *                             emitted as polyMesh data and sent through the
 *                             FOAM path above (no reference counterpart: this
 *                             is the generator of the 1M-cell test mesh)
 *   xcalm_block2d          -- BlockMesher2D generate! + UNV2.update_mesh_format
 *                             (src/Mesh/BlockMesher2D/2_builder.jl,
 *                              3_connectivity.jl, 4_geometry.jl,
 *                              src/UNV2/UNV2_4_update_format.jl)
 *   xcalm_read_unv2        -- UNV2D_mesh(path; scale)
 *                             (src/UNV2/UNV2_1_loader.jl, UNV2_2_geometry.jl,
 *                              UNV2_3_builder.jl)
 *
 * Errors: every function returns 0 on success and a negative code on error;
 * xcalm_last_error() gives a thread-local message.
 */
#ifndef XCAL_MESH_H
#define XCAL_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xcalm_mesh xcalm_mesh;

const char* xcalm_last_error(void);

/* polyMesh arrays, 0-based as in the OpenFOAM files.
 * face_nodes_ptr has n_faces+1 entries (CSR over face_nodes).
 * owner has n_faces entries, neighbour n_ifaces entries.
 * Patches are given in file order with 0-based startFace. */
int xcalm_foam_from_arrays(int64_t n_points, const double* points,
                           int64_t n_faces, const int64_t* face_nodes_ptr,
                           const int64_t* face_nodes,
                           const int64_t* owner, int64_t n_ifaces,
                           const int64_t* neighbour,
                           int32_t n_patches, const char* const* patch_names,
                           const int64_t* patch_start, const int64_t* patch_nfaces,
                           xcalm_mesh** out);

int xcalm_read_foam(const char* polymesh_dir, double scale, xcalm_mesh** out);

/* Hex box [0,Lx]x[0,Ly]x[0,Lz] with nx*ny*nz cells.  Interior node coordinates
 * are moved by u*jitter*h along each axis where the node is not on a boundary
 * plane of that axis (u uniform in [-1,1) from xoshiro256** seeded through
 * SplitMix64 with seed_jitter).  Cells are renumbered by a Fisher-Yates
 * permutation drawn from seed_perm (seed_perm == 0 keeps natural order).
 * Patches: inlet (x=0), outlet (x=Lx), walls (the four side planes). */
int xcalm_box3d(int64_t nx, int64_t ny, int64_t nz,
                double Lx, double Ly, double Lz,
                double jitter, uint64_t seed_jitter, uint64_t seed_perm,
                xcalm_mesh** out);

/* BlockMesher2D.  points: n_points*3 coordinates (z ignored = 0 plane).
 * edges: n_edges*3 int64 (p1, p2, ncells), 1-based point ids, straight
 * linear lines as line!(points, p1, p2, ncells).  blocks: n_blocks*4 edge ids
 * as quad(edges, [e1,e2,e3,e4]) (1-based).  Patches: names plus a CSR list of
 * edge ids (patch_edge_ptr has n_patches+1 entries, 0-based offsets). */
int xcalm_block2d(int64_t n_points, const double* points,
                  int64_t n_edges, const int64_t* edges,
                  int64_t n_blocks, const int64_t* blocks,
                  int32_t n_patches, const char* const* patch_names,
                  const int64_t* patch_edge_ptr, const int64_t* patch_edges,
                  xcalm_mesh** out);

int xcalm_read_unv2(const char* path, double scale, xcalm_mesh** out);

/* Accessors.  Names: "cell_centre"(3n) "cell_volume"(n) "cell_faces_range"(2n)
 * "cell_nodes_range"(2n) "cell_nodes" "cell_faces" "cell_neighbours"
 * "cell_nsign" "face_nodes_range"(2nf) "face_nodes" "face_owner"(2nf)
 * "face_centre"(3nf) "face_normal"(3nf) "face_e"(3nf) "face_area"(nf)
 * "face_delta"(nf) "face_weight"(nf) "boundary_cellsID"(nbf)
 * "patch_range"(2np) "node_coords"(3nn).
 * *is_float is set to 1 for double arrays, 0 for int64 arrays. */
int xcalm_array(const xcalm_mesh* m, const char* name, const void** ptr,
                int64_t* len, int32_t* is_float);
int xcalm_sizes(const xcalm_mesh* m, int32_t* dim, int64_t* n_cells,
                int64_t* n_faces, int64_t* n_bfaces, int64_t* n_nodes,
                int32_t* n_patches);
const char* xcalm_patch_name(const xcalm_mesh* m, int32_t i);
void xcalm_free(xcalm_mesh* m);

#ifdef __cplusplus
}
#endif
#endif
