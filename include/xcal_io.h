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
 * host the same way, VTK_writer_3D.jl:3-7).  The OpenFOAM case writer
 * (src/IOFormats/OpenFOAM/OpenFOAM_writer.jl) is below the VTK ones.
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

/* ---- OpenFOAM case output (src/IOFormats/OpenFOAM/OpenFOAM_writer.jl) ---- */

/* initialise_writer(OpenFOAM(), mesh::Mesh3) (OpenFOAM_writer.jl:1-185):
 * creates <directory>/XCALibre.foam (empty, for ParaView) and writes
 * <directory>/constant/polyMesh/{points, faces, owner, neighbour, boundary}
 * with internal faces first, then boundary faces in patch order; points
 * print with C's "%g" as the reference's @sprintf does.  directory NULL: the
 * working directory.  patch_names[i] and patch_range (1-based first/last
 * boundary face of each patch) describe the patches.  A 2D mesh is an
 * error, as in the reference (OpenFOAM_writer.jl:187-189). */
int xcalw_foam_init(const char* directory, int32_t dim, int64_t n_nodes,
                    const double* node_coords, int64_t n_cells, int64_t n_faces,
                    int64_t n_bfaces, const int64_t* face_nodes_range,
                    const int64_t* face_nodes, const int64_t* face_owner, int32_t n_patches,
                    const char* const* patch_names, const int64_t* patch_range,
                    xcalw_writer** out);

/* One boundary condition of a field: patch (0-based), kind (an xcal_bc_kind
 * of xcal.h), value (1 or 3 components as the field).  The file entry
 * follows _foam_boundary_entry (OpenFOAM_writer.jl:277-359): Dirichlet and
 * a vector Wall give fixedValue with the value, Empty gives empty, every
 * other kind zeroGradient. */
typedef struct {
  int32_t patch;
  int32_t kind;
  double value[3];
} xcalw_foam_bc;

/* One field: ncomp 1 (volScalarField) or 3 (volVectorField), values as
 * xcalw_field, and its boundary conditions in the order they print. */
typedef struct {
  const char* label;
  int32_t ncomp;
  const double* values;
  int32_t n_bcs;
  const xcalw_foam_bc* bcs;
} xcalw_foam_field;

/* write_results(iteration, time, mesh, ::FOAMWriter, BCs, args...)
 * (OpenFOAM_writer.jl:191-273): one file per field in the time directory
 * "<iteration>" when time equals iteration, else "%.8f" of time. */
int xcalw_foam_write_results(const xcalw_writer* w, int64_t iteration, double time,
                             int32_t n_fields, const xcalw_foam_field* fields);

/* The time directory name write_results uses, into buf of cap bytes.
 * Returns the length, or negative when it does not fit. */
int32_t xcalw_foam_time_dir(int64_t iteration, double time, char* buf, int32_t cap);

void xcalw_free(xcalw_writer* w);

#ifdef __cplusplus
}
#endif

#endif /* XCAL_IO_H */
