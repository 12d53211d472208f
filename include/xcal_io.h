/*
 * xcal_io.h -- result writers for cell fields (C ABI, host code).
 *
 * The reference hands results back as VTK files: a legacy ASCII file for 2D
 * meshes (src/IOFormats/VTK/VTK_writer.jl:6-124) and an XML unstructured
 * grid of polyhedra for 3D meshes (src/IOFormats/VTK/VTK_writer_3D.jl:16-224).
 * These functions write the same files, line for line, from the mesh arrays
 * of xcal_mesh.h (1-based Int64 indices, interleaved xyz Float64 geometry)
 * and host copies of the fields (download them with xcal_field_download or
 * xcal_simple_get_field first; the reference copies device fields to the
 * host the same way, VTK_writer_3D.jl:3-7).
 *
 * Numbers are written as Julia's print(::Float64) writes them: the shortest
 * decimal digits that read back to the same double, in plain form
 * ("0.0001", "123456.0") for 1e-4 <= |x| < 1e6, else as "1.25e-5" or
 * "1.0e6"; "0.0", "-0.0", "NaN", "Inf", "-Inf" for the special values.
 *
 * Errors: 0 on success, negative on error; xcalw_last_error() gives a
 * thread-local message.
 */
#ifndef XCAL_IO_H
#define XCAL_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xcalw_writer xcalw_writer;

const char* xcalw_last_error(void);

/* The text Julia's print(::Float64) gives for x, NUL-terminated, into buf
 * (at least 32 bytes).  Returns the length. */
int32_t xcalw_format_double(double x, char* buf);

/* initialise_writer(VTK(), mesh::Mesh2) (VTK_writer.jl:4).  The reference
 * keeps nothing here and rebuilds the mesh part of the file on every write;
 * this writer builds it once (same bytes). */
int xcalw_vtk2d_init(int64_t n_nodes, const double* node_coords, int64_t n_cells,
                     const int64_t* cell_nodes_range, const int64_t* cell_nodes,
                     xcalw_writer** out);

/* initialise_writer(VTK(), mesh::Mesh3) (VTK_writer_3D.jl:16-146): the
 * header with points, polyhedron connectivity, offsets, the faces of every
 * cell (boundary faces included, each ordered so that its first three nodes
 * give an outward normal) and faceoffsets, plus the footer. */
int xcalw_vtu3d_init(int64_t n_nodes, const double* node_coords, int64_t n_cells,
                     const int64_t* cell_nodes_range, const int64_t* cell_nodes,
                     const double* cell_centre, int64_t n_faces,
                     const int64_t* face_nodes_range, const int64_t* face_nodes,
                     const int64_t* face_owner, const double* face_centre,
                     xcalw_writer** out);

/* One cell field: ncomp 1 (scalar), 3 (vector: x, y, z) or 9 (tensor: xx,
 * xy, xz, yx, yy, yz, zx, zy, zz); values holds ncomp blocks of n_cells
 * doubles, component after component. */
typedef struct {
  const char* label;
  int32_t ncomp;
  const double* values;
} xcalw_field;

/* write_results(iteration, time, mesh, meshData, BCs, args...; suffix)
 * (VTK_writer.jl:6-124, VTK_writer_3D.jl:150-224): writes
 * "iteration_<iteration><suffix>.vtk|.vtu" when time equals iteration, else
 * "time_<iteration><suffix>...", into directory (NULL: the working
 * directory, as the reference does). */
int xcalw_write_results(const xcalw_writer* w, int64_t iteration, double time,
                        const char* suffix, const char* directory, int32_t n_fields,
                        const xcalw_field* fields);

/* The file name write_results would use (without directory), into buf of
 * cap bytes.  Returns the length, or negative when it does not fit. */
int32_t xcalw_result_name(const xcalw_writer* w, int64_t iteration, double time,
                          const char* suffix, char* buf, int32_t cap);

void xcalw_free(xcalw_writer* w);

#ifdef __cplusplus
}
#endif

#endif /* XCAL_IO_H */
