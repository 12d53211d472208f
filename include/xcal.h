/*
 * xcal.h -- C ABI of libxcalhip.so, the gfx950 (MI355X) implementation of
 * XCALibre's per-iteration finite-volume SIMPLE hot path.
 *
 * The entry points are what a Julia package extension (a sibling of
 * ext/XCALibre_AMDExt.jl) binds with `ccall` behind XCALibre's own API:
 * `(...) -> ScalarEquation(phi, BCs)`, `-> VectorEquation(psi, BCs)` and
 * `run!(model, config; pref)`.  Each declaration cites the reference
 * function it replaces.  Conventions:
 *   - Mesh arrays arrive in the reference layout (1-based Int64 indices,
 *     Float64 geometry, boundary-first faces; src/Mesh/Mesh_0_types.jl).
 *     Host pointers are borrowed for the duration of the call only.
 *   - The library owns every device allocation behind opaque handles.
 *     A *_destroy call frees only its own handle and never reads the parent
 *     handles, so handles may be released in any order (garbage collectors
 *     finalise in arbitrary order); after a parent is destroyed its children
 *     may only be destroyed.
 *   - Every call returns 0 on success, < 0 on error (message from
 *     xcal_last_error(), thread-local), > 0 for warnings (XCAL_ITMAX).
 *   - Calls on one context are ordered on its HIP stream; calls that return
 *     host scalars synchronise.  Contexts must not be shared between threads
 *     without external locking.
 *   - Cell numbering, CSR pattern (columns ascending within a row) and face
 *     numbering are the reference's; CSR arrays cross the boundary 0-based.
 */
#ifndef XCAL_H
#define XCAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XCAL_OK 0
#define XCAL_ITMAX 1          /* warning: Krylov itmax reached (Solve_1_api.jl:302) */
#define XCAL_ERR_ARG -1
#define XCAL_ERR_HIP -2
#define XCAL_ERR_UNSUPPORTED -3
#define XCAL_ERR_NOMEM -4
#define XCAL_ERR_STATE -5

typedef struct xcal_ctx xcal_ctx;
typedef struct xcal_mesh xcal_mesh;
typedef struct xcal_field xcal_field;
typedef struct xcal_eqn xcal_eqn;
typedef struct xcal_simple xcal_simple;

/* ---------------------------------------------------------------- mesh in */
/* The fields of Mesh2/Mesh3 that the hot path reads
 * (src/Mesh/Mesh_0_types.jl:135-185).  dim = 2 for Mesh2, 3 for Mesh3. */
typedef struct {
  int32_t dim;
  int32_t n_patches;
  int64_t n_cells, n_faces, n_bfaces, n_cell_faces;
  const double* cell_centre;        /* 3*n_cells                         */
  const double* cell_volume;        /* n_cells                           */
  const int64_t* cell_faces_range;  /* 2*n_cells: first,last (1-based)   */
  const int64_t* cell_faces;        /* n_cell_faces, 1-based face ids    */
  const int64_t* cell_neighbours;   /* n_cell_faces, 1-based cell ids    */
  const int64_t* cell_nsign;        /* n_cell_faces, +1 owner / -1 neigh */
  const int64_t* face_owner;        /* 2*n_faces: ownerCells (1-based)   */
  const double* face_centre;        /* 3*n_faces                         */
  const double* face_normal;        /* 3*n_faces                         */
  const double* face_e;             /* 3*n_faces                         */
  const double* face_area;          /* n_faces                           */
  const double* face_delta;         /* n_faces                           */
  const double* face_weight;        /* n_faces                           */
  const int64_t* boundary_cellsID;  /* n_bfaces (1-based)                */
  const int64_t* patch_range;       /* 2*n_patches: IDs_range (1-based)  */
} xcal_mesh_host;

/* ------------------------------------------------------- operator spec */
/* Operator types (ModelFramework_0_types.jl:18-78, schemes in
 * Discretise_1_schemes.jl). */
enum {
  XCAL_TIME_STEADY = 0,       /* Time{SteadyState}          :12-22 */
  XCAL_TIME_EULER = 1,        /* Time{Euler}                :25-42 */
  XCAL_LAPLACIAN_LINEAR = 2,  /* Laplacian{Linear}          :68-110 */
  XCAL_DIV_LINEAR = 3,        /* Divergence{Linear}         :114-133 */
  XCAL_DIV_UPWIND = 4,        /* Divergence{Upwind}         :137-151 */
  XCAL_DIV_LUST = 5,          /* Divergence{LUST}           :155-179 */
  XCAL_DIV_BOUNDED_UPWIND = 6,/* Divergence{BoundedUpwind}  :183-199 */
  XCAL_SI = 7,                /* Si (implicit source)       :204-218 */
  /* Time{CrankNicolson} (:44-62): assembled as Time{Euler}.  An equation
   * whose first term is Crank-Nicolson takes the explicit step
   * x = 2 x - x_old after each of its Krylov solves (solve_system!,
   * Solve_1_api.jl:291-296). */
  XCAL_TIME_CRANK_NICOLSON = 8
};

/* One operator term: Operator(flux, phi, sign, type).  flux is a face scalar
 * field for Laplacian/Divergence and a cell scalar field for Time/Si; a NULL
 * flux means ConstantScalar(flux_const). */
typedef struct {
  int32_t type;
  int32_t sign; /* +1 or -1 */
  const xcal_field* flux;
  double flux_const;
} xcal_term;

/* Src(field, sign) (ModelFramework_0_types.jl:80-88); the field's kind
 * (scalar or vector) must match the equation. */
typedef struct {
  const xcal_field* field;
  int32_t sign;
} xcal_source;

/* Boundary conditions (files under Discretise/boundary_conditions/).  One entry per
 * mesh patch, in patch order, as `assign(region, (U=[...], ...))` produces. */
enum {
  XCAL_BC_DIRICHLET = 0,
  XCAL_BC_NEUMANN = 1,
  XCAL_BC_ZEROGRADIENT = 2,
  XCAL_BC_EXTRAPOLATED = 3,
  XCAL_BC_WALL = 4,
  XCAL_BC_SYMMETRY = 5,
  /* wall functions of the k-omega model (Discretise_0_types.jl:38-160) with
   * the reference's default coefficients kappa 0.41, beta1 0.075, cmu 0.09,
   * B 5.2, E 9.8.  Assembly: KWallFunction as Zerogradient, OmegaWallFunction
   * as Extrapolated (Discretise_3_boundary_conditions.jl); face values are
   * zero-gradient (zerogradient_interpolation.jl:1). */
  XCAL_BC_KWALL = 6,
  XCAL_BC_OMEGAWALL = 7,
  XCAL_BC_NUTWALL = 8,
  /* Empty (empty.jl:46-106): no coefficient from any operator, face values
   * zero (empty_interpolation.jl:1-16).  Outlet (outlet.jl) acts as
   * XCAL_BC_ZEROGRADIENT in pressure-based solvers and is passed as such. */
  XCAL_BC_EMPTY = 9
};
typedef struct {
  int32_t kind;
  int32_t patch;    /* 0-based mesh patch index */
  double value[3];  /* scalar BCs use value[0] */
} xcal_bc;

/* Linear solver settings: SolverSetup (Solve_1_api.jl:8-83). */
enum { XCAL_SOLVER_CG = 0, XCAL_SOLVER_BICGSTAB = 1 };
enum { XCAL_PRECON_NONE = 0, XCAL_PRECON_JACOBI = 1, XCAL_PRECON_DILU = 2 };
enum { XCAL_PREC_FP64 = 0,       /* plain FP64 Krylov                         */
       XCAL_PREC_FP32_IR = 1,    /* FP32 inner Krylov in FP64 refinement loop */
       /* FP64 CG (SpMV, vectors, dot products, residual recurrence) with
        * the DILU sweeps computed in FP32 from FP32 factors, on single
        * meshes in the cone sweep form.  BiCGSTAB (its left-preconditioned
        * recurrence runs through the sweeps), other sweep forms and other
        * preconditioners keep FP64 under this mode; batch meshes reject it
        * (XCAL_ERR_UNSUPPORTED). */
       XCAL_PREC_FP32_PRECON = 2 };
typedef struct {
  int32_t solver;
  int32_t precon;
  int32_t precision;
  int32_t itmax;
  double convergence;
  double relax;
  double atol;
  double rtol;
  double ir_inner_rtol; /* FP32_IR: relative tolerance of each inner solve (0: 1e-6) */
  /* SolverSetup.smoother = JacobiSmoother(; loops, omega)
   * (Smoothers_jacobi.jl:36-51): `smoother_loops` weighted-Jacobi sweeps of
   * the field before the Krylov solve (apply_smoother!, :54-110, called by
   * solve_system!, Solve_1_api.jl:285).  0 loops: no smoother (nothing). */
  int32_t smoother_loops;
  double smoother_omega;
} xcal_solver_setup;

typedef struct {
  int32_t niter;     /* Krylov iterations (summed over IR sweeps)        */
  int32_t nouter;    /* IR sweeps (1 for FP64)                           */
  double residual;   /* residual() of Solve_1_api.jl:441-464            */
} xcal_solve_stats;

/* ------------------------------------------------------------ context */
const char* xcal_last_error(void);
int32_t xcal_abi_version(void);
int xcal_ctx_create(int32_t device, xcal_ctx** out);
int xcal_ctx_destroy(xcal_ctx* ctx);
int xcal_ctx_synchronize(xcal_ctx* ctx);

/* -------------------------------------------------------------- mesh */
/* Replaces adapt(backend, mesh) plus the host CSR build of
 * sparse_matrix_connectivity / _build_A (ModelFramework_0_types.jl:106,
 * 223-242; ext/XCALibre_AMDExt.jl:25-28).  Builds the device SoA mesh, the
 * CSR pattern, face->nonzero slot maps (replacing every spindex search,
 * ModelFramework_0_types.jl:246-258) and the DILU level schedule. */
int xcal_mesh_create(xcal_ctx* ctx, const xcal_mesh_host* host, xcal_mesh** out);
int xcal_mesh_destroy(xcal_mesh* mesh);
/* n_cells, nnz, number of forward/backward DILU levels, device bytes */
int xcal_mesh_info(const xcal_mesh* mesh, int64_t* n_cells, int64_t* nnz,
                   int32_t* fwd_levels, int32_t* bwd_levels, int64_t* device_bytes);
/* Entries per row of the pipelined one-workgroup DILU sweeps (forward,
 * backward), or -1 where the sweeps step level by level.  Batch meshes: the
 * per-case sweeps, -1 if any case exceeds their limits (more than 256 rows
 * in a level, 2048 levels or 8 entries per row).  Single meshes: the widest
 * pipelined run of small levels.  No reference counterpart (reports a device
 * code path). */
int xcal_mesh_sweep_form(const xcal_mesh* mesh, int32_t* fwd, int32_t* bwd);
/* 1 where the pipelined sweeps of a batch mesh (forward, backward) read a
 * dependency more than 4096 rows back, beyond their LDS ring, from global
 * memory; 0 otherwise.  No reference counterpart (reports a device code
 * path). */
int xcal_mesh_sweep_far(const xcal_mesh* mesh, int32_t* fwd, int32_t* bwd);
/* Rows of a level per thread (1, 2 or 4; -1 before the first DILU solve) of
 * the pipelined batch sweeps, forward and backward, as the built-in tuner
 * chose them by timing every choice on the first DILU solve on a mesh of this
 * shape (XCAL_PIPE_ROWS=1|2|4 fixes them), and its timings: us[6] = mean µs
 * per launch, forward at 1, 2, 4 rows per thread, then backward (0 where not
 * timed).  Results do not depend on the choice.  No reference counterpart. */
int xcal_mesh_pipe_tuning(const xcal_mesh* mesh, int32_t* fwd_rows, int32_t* bwd_rows,
                          double* us);
/* Cone form of the DILU sweeps (single meshes): rows per block, forward
 * blocks, and the cone positions of all blocks for the forward and backward
 * sweeps; all 0 where the sweeps launch level by level.  Each sweep is then
 * one launch per band of consecutive levels (one band where the cones fit,
 * as on randomly numbered meshes), in which the workgroup of a block of
 * cells sweeps the block and every row of the band it depends on, in the
 * reference's order, reading earlier bands' rows from global memory.  No
 * reference counterpart (reports a device code path;
 * forward_substitution!/backward_substitution!,
 * preconditioners_1_DILU.jl:84-114, are what the sweeps compute). */
int xcal_mesh_cone_form(const xcal_mesh* mesh, int32_t* block_rows, int32_t* blocks,
                        int64_t* fwd_rows, int64_t* bwd_rows);
/* Bands (launches) per sweep of the cone form: forward, backward (0: level
 * launches).  No reference counterpart. */
int xcal_mesh_cone_bands(const xcal_mesh* mesh, int32_t* fwd_bands, int32_t* bwd_bands);
/* CSR pattern, 0-based (rowptr n+1, colval nnz) */
int xcal_mesh_csr_pattern(const xcal_mesh* mesh, int32_t* rowptr, int32_t* colval);

/* Mesh ingest: the matrix connectivity of a mesh in the reference numbering,
 * built by device kernels (on_device = 1) or by the host loops (0); the two
 * are equal element for element.  Replaces sparse_matrix_connectivity +
 * sparsecsr (ModelFramework_0_types.jl:139-140, 223-242) and the spindex
 * searches of assembly (:246-258).  xcal_mesh_create builds its mesh with the
 * device form (XCAL_HOST_INGEST=1 selects the host loops); this handle
 * exposes the arrays for checks and timing.  The device form runs on the
 * context's stream and synchronises it. */
typedef struct xcal_connectivity xcal_connectivity;
int xcal_connectivity_create(xcal_ctx* ctx, const xcal_mesh_host* mesh, int32_t on_device,
                             xcal_connectivity** out);
/* nnz, cell-face pairs, forward and backward DILU level counts, wall time
 * of the build (host clock, uploads and downloads included) and, for the
 * device form, the kernel time between events on the stream (ms) */
int xcal_connectivity_info(const xcal_connectivity* c, int64_t* nnz, int64_t* n_cell_faces,
                           int32_t* nlev_f, int32_t* nlev_b, double* wall_ms, double* device_ms);
/* which: 0 rowptr (n+1), 1 colval (nnz), 2 diagonal slot (n), 3 slot of each
 * cell_faces entry (n_cell_faces, faces_range order), 4 slot of each entry's
 * transpose (nnz, -1 if absent), 5 owner-row slot of each face (n_faces, -1
 * on boundary faces), 6 forward DILU level (n), 7 backward DILU level (n) */
int xcal_connectivity_get(const xcal_connectivity* c, int32_t which, int32_t* host);
int xcal_connectivity_destroy(xcal_connectivity* c);

/* Batched form (no single reference counterpart: it runs many independent
 * cases of the reference's mesh/equation/run! path at once).  The cases'
 * meshes become one device mesh whose fields, equations and SIMPLE runs
 * cover every case; each case keeps its own Krylov iterations and stops on
 * its own, and a case's results do not depend on which other cases share
 * the batch.  All cases must have the same patches (the BC table is shared).
 * Fields of a batch mesh cross the ABI as the cases' reference-numbered
 * arrays concatenated (cells: case 0's cells, then case 1's, ...; faces
 * likewise, each case boundary-first). */
int xcal_mesh_create_batch(xcal_ctx* ctx, const xcal_mesh_host* const* hosts, int32_t ncase,
                           xcal_mesh** out);
/* ncase and the ncase+1 cell offsets of the cases (either may be NULL) */
int xcal_mesh_cases(const xcal_mesh* mesh, int32_t* ncase, int64_t* cell_offsets);

/* ------------------------------------------------------------- fields */
/* ScalarField/VectorField/FaceScalarField/FaceVectorField/TensorField
 * (src/Fields/Fields_0_types.jl:56-203), SoA per component. */
enum {
  XCAL_CELL_SCALAR = 0,
  XCAL_CELL_VECTOR = 1,
  XCAL_FACE_SCALAR = 2,
  XCAL_FACE_VECTOR = 3,
  XCAL_CELL_TENSOR = 4
};
int xcal_field_create(xcal_mesh* mesh, int32_t kind, xcal_field** out);
int xcal_field_destroy(xcal_field* f);
int xcal_field_upload(xcal_field* f, int32_t component, const double* host);
int xcal_field_download(const xcal_field* f, int32_t component, double* host);
/* initialise!(field, value) (Fields_0_types.jl:376-402) */
int xcal_field_fill(xcal_field* f, int32_t component, double value);

/* ---------------------------------------------------------- equations */
/* ScalarEquation(phi, BCs) / VectorEquation(psi, BCs)
 * (ModelFramework_0_types.jl:110-215). */
int xcal_eqn_create(xcal_mesh* mesh, int32_t vector, xcal_eqn** out);
int xcal_eqn_destroy(xcal_eqn* e);
/* discretise! (Discretise_2_generated_distretisation.jl:3-167): cell-gather
 * assembly into nzval (A for scalar, A0 for vector) and b (bx,by,bz).
 * prev: the field (scalar or vector) used by Time terms. */
int xcal_discretise(xcal_eqn* e, const xcal_term* terms, int32_t nterms,
                    const xcal_source* sources, int32_t nsources, const xcal_field* prev);
/* discretise! with config.runtime.dt (Discretise_1_schemes.jl:25-42,
 * Time{Euler}: a_P += V/dt, b += prev V/dt); dt > 0. */
int xcal_discretise_dt(xcal_eqn* e, const xcal_term* terms, int32_t nterms,
                       const xcal_source* sources, int32_t nsources, const xcal_field* prev,
                       double dt);
/* update_equation! (Discretise_2_generated_distretisation.jl:305-330): A <- A0 */
int xcal_update_equation(xcal_eqn* e);
/* apply_boundary_conditions! (Discretise_5_apply_bcs.jl:5-123) for the
 * given component (0 for scalar equations, 0..2 for vector equations).
 * phi is the equation's unknown (read by Extrapolated/Wall/Symmetry). */
int xcal_apply_bcs(xcal_eqn* e, const xcal_term* terms, int32_t nterms,
                   const xcal_bc* bcs, int32_t nbcs, const xcal_field* phi,
                   int32_t component);
/* implicit_relaxation_diagdom! (Solve_1_api.jl:370-411) */
int xcal_relax_implicit_diagdom(xcal_eqn* e, const xcal_field* phi, int32_t component,
                                double alpha);
/* setReference! (Solve_1_api.jl:414-439); cell is 1-based */
int xcal_set_reference(xcal_eqn* e, double pref, int64_t cell);
/* copy A (which=0) or A0 (which=1) in CSR order and b of `component` out */
int xcal_eqn_download(const xcal_eqn* e, int32_t which, int32_t component, double* nzval,
                      double* b);
/* update_preconditioner! (preconditioners_2_functions.jl:65-124) */
int xcal_precond_update(xcal_eqn* e, int32_t precon);
/* z = P \ r with the equation's current preconditioner: ldiv!(x, P, b) for
 * DILU (preconditioners_1_DILU.jl:116-120, ext/XCALibre_AMDExt.jl:149-156),
 * mul!(z, opDiagonal(1/|a_ii|), r) for Jacobi.  r and z: cell scalar fields
 * or components of cell vector fields. */
int xcal_precond_apply(xcal_eqn* e, const xcal_field* r, int32_t rcomp, xcal_field* z,
                       int32_t zcomp);
/* xcal_precond_apply without allocating: no counterpart in the reference.  It
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The same kernels and results as
 * xcal_precond_apply, but the DILU sweeps stage through a 16 MiB device array
 * of the library (held by each GPU it runs on until the process ends)
 * instead of the equation's workspace, so no hipMalloc or hipFree is called.
 * Meshes of at most 1048576 rows; larger ones return XCAL_ERR_ARG.  Calls
 * that overlap on one GPU share the array and race: serialise them.  The
 * equation's preconditioner must have been set up (xcal_precond_update). */
int xcal_precond_apply_noalloc(xcal_eqn* e, const xcal_field* r, int32_t rcomp, xcal_field* z,
                               int32_t zcomp);
/* y = A x with the equation's A (mul! on SparseXCSR, Multithread/spmvm.jl:85-91) */
int xcal_spmv(xcal_eqn* e, const xcal_field* x, int32_t xcomp, xcal_field* y, int32_t ycomp);
/* solve_system! (Solve_1_api.jl:268-307): Krylov solve warm-started from
 * phi, copy into phi, residual.  Returns XCAL_ITMAX when itmax was hit. */
/* solve_system! for every case of a batch mesh: stats[ncase] */
int xcal_solve_system_batch(xcal_eqn* e, const xcal_solver_setup* s, xcal_field* phi,
                            int32_t component, xcal_solve_stats* stats);
int xcal_solve_system(xcal_eqn* e, const xcal_solver_setup* s, xcal_field* phi,
                      int32_t component, xcal_solve_stats* stats);

/* ---------------------------------------------------------- calculate */
/* interpolate! (Calculate_2_interpolation.jl:40-76, 347-397) */
int xcal_interpolate(xcal_field* face, const xcal_field* cell);
/* correct_boundaries! (boundary_conditions/1_boundary_interpolation.jl:3-37) */
int xcal_correct_boundaries(xcal_field* face, const xcal_field* cell, const xcal_bc* bcs,
                            int32_t nbcs);
/* green_gauss! (Calculate_1_green_gauss.jl:5-173): scalar->vector or
 * vector->tensor, selected by the field kinds */
int xcal_green_gauss(xcal_field* grad, const xcal_field* face);
/* grad!{Gauss} = interpolate! + correct_boundaries! + green_gauss!
 * (Calculate_0_gradient.jl:105-117) */
int xcal_grad(xcal_field* grad, xcal_field* face, const xcal_field* cell, const xcal_bc* bcs,
              int32_t nbcs);
/* grad! with Grad{Midpoint} (Calculate_0_gradient.jl:121-250):
 * interpolate_midpoint! (face = (P + N)/2), correct_boundaries!,
 * green_gauss!, then twice correct_interpolation! (internal faces:
 * face = (P + N)/2 + (grad_P + grad_N)/2 . (r_f - (r_P + r_N)/2)) and
 * green_gauss!.  Scalar cell field -> vector gradient, or vector cell field
 * -> tensor gradient; `face` holds the final face values. */
int xcal_grad_midpoint(xcal_field* grad, xcal_field* face, const xcal_field* cell,
                       const xcal_bc* bcs, int32_t nbcs);
/* div! for a FaceScalarField (Calculate_0_divergence.jl:99-171) */
int xcal_div(xcal_field* cell, const xcal_field* face);
/* flux! (Solvers_0_functions.jl:41-64): mdot_f = psi_f . S_f */
int xcal_flux(xcal_field* mdot, const xcal_field* face_vector);
/* explicit_relaxation! (Solve_1_api.jl:317-333) */
int xcal_relax_explicit(xcal_field* phi, const xcal_field* phi0, double alpha);

/* ------------------------------------------------------ SIMPLE stages */
/* inverse_diagonal! (Solvers_0_functions.jl:96-122) */
int xcal_inverse_diagonal(xcal_field* rD, const xcal_eqn* e);
/* remove_pressure_source! (Solvers_0_functions.jl:156-181) */
int xcal_remove_pressure_source(xcal_eqn* e, const xcal_field* gradp, int32_t source_sign);
/* H! (Solvers_0_functions.jl:184-249) */
int xcal_H(xcal_field* Hv, const xcal_field* U, const xcal_eqn* e);
/* correct_mass_flux! + correct_boundary_mass_flux!
 * (Solvers_1_SIMPLE.jl:340-388, 474-557).  p_flux/p_sign: the pressure
 * Laplacian term's flux (rDf) and sign. */
int xcal_correct_mass_flux(xcal_field* mdot, const xcal_eqn* peqn, const xcal_field* p,
                           const xcal_field* p_flux, int32_t p_sign, const xcal_bc* bcs,
                           int32_t nbcs);
/* correct_velocity! (Solvers_0_functions.jl:126-152) */
int xcal_correct_velocity(xcal_field* U, const xcal_field* Hv, const xcal_field* gradp,
                          const xcal_field* rD);
/* nonorthogonal_face_correction (Solvers_1_SIMPLE.jl:267-321): adds
 * (flux gradf).T_hat of every internal face to b of its owner and subtracts
 * it from b of its neighbour (gradp: the pressure gradient, flux: rDf) */
int xcal_nonorthogonal_correction(xcal_eqn* p_eqn, const xcal_field* gradp, const xcal_field* p,
                                  const xcal_field* flux);

/* ------------------------------------------------------ SIMPLE driver */
/* The steady incompressible SIMPLE loop of Solvers_1_SIMPLE.jl:103-263 with
 * Gauss or Midpoint gradients and setup.ncorrectors non-orthogonal
 * correctors, entirely on the device (run!/simple!,
 * Solvers_3_solver_dispatch.jl:120-135). */
typedef struct {
  double nu;
  int32_t div_scheme;   /* XCAL_DIV_* for the momentum convection term */
  int32_t has_pref;     /* run!(...; pref) given                       */
  double pref;
  xcal_solver_setup U;
  xcal_solver_setup p;
  const xcal_bc* bc_U;  /* n_patches entries */
  const xcal_bc* bc_p;  /* n_patches entries */
  int32_t n_bcs;
  int32_t grad_p;       /* XCAL_GRAD_* of grad!(∇p) (schemes.p.gradient)   */
  int32_t grad_U;       /* XCAL_GRAD_* of grad!(∇U) (schemes.U.gradient)   */
  int32_t time_U;       /* XCAL_TIME_STEADY, _EULER or _CRANK_NICOLSON:
                           Time{schemes.U.time}(U) of the momentum equation
                           (Solvers_1_SIMPLE.jl:68-74) */
  double dt;            /* runtime.dt of Time{Euler/CrankNicolson} (> 0 then) */
  int32_t ncorrectors;  /* run!(...; ncorrectors): non-orthogonal correctors of
                           the pressure (Solvers_1_SIMPLE.jl:194-206,
                           Solvers_2_PISO.jl:148-165) */
} xcal_simple_setup;
enum { XCAL_GRAD_GAUSS = 0, XCAL_GRAD_MIDPOINT = 1 };

enum { XCAL_SIMPLE_U = 0, XCAL_SIMPLE_P = 1, XCAL_SIMPLE_MDOTF = 2, XCAL_SIMPLE_K = 3,
       XCAL_SIMPLE_OMEGA = 4, XCAL_SIMPLE_NUT = 5 };

/* RANS{KOmega} (RANS_kOmega.jl:41-247): k and omega transport with
 * production from the strain rate, wall functions, bound! and nut = k/omega;
 * the momentum Laplacian then uses nueff = nu + nutf (update_nueff!,
 * Solvers_0_functions.jl:5-36).  Turbulence residuals join the convergence
 * test of the SIMPLE loop (Solvers_1_SIMPLE.jl:226-233). */
enum { XCAL_TURB_LAMINAR = 0, XCAL_TURB_KOMEGA = 1 };
typedef struct {
  int32_t model;
  double beta_star, alpha1, beta1, sigma_k, sigma_omega; /* 0.09, 0.52, 0.072, 0.5, 0.5 */
  int32_t div_scheme;   /* XCAL_DIV_* of the k and omega convection terms */
  xcal_solver_setup k, omega;
  const xcal_bc* bc_k;     /* n_patches entries each */
  const xcal_bc* bc_omega;
  const xcal_bc* bc_nut;
} xcal_turbulence_setup;

int xcal_simple_create(xcal_mesh* mesh, const xcal_simple_setup* setup, xcal_simple** out);
int xcal_simple_destroy(xcal_simple* s);
/* turbulence model of a SIMPLE run (laminar unless set) */
int xcal_simple_set_turbulence(xcal_simple* s, const xcal_turbulence_setup* t);
/* U: 3*n_cells (x block, y block, z block); p, k, omega, nut: n_cells */
int xcal_simple_set_field(xcal_simple* s, int32_t which, const double* host);
int xcal_simple_get_field(const xcal_simple* s, int32_t which, double* host);
/* Run up to `iterations` SIMPLE iterations (stops early on convergence as
 * the reference does).  residuals: 4*iterations doubles laid out
 * [Ux..., Uy..., Uz..., p...] (the NamedTuple of Solvers_1_SIMPLE.jl:262);
 * done: iterations performed.  krylov_iters (optional, 4*iterations ints)
 * gets the Krylov iteration counts of the Ux, Uy, Uz and p solves. */
int xcal_simple_run(xcal_simple* s, int32_t iterations, double* residuals, int32_t* done,
                    int32_t* krylov_iters);
/* Whether the last xcal_simple_run stopped on the convergence test of
 * Solvers_1_SIMPLE.jl:226-233 (every residual below its SolverSetup's
 * convergence, and the turbulence model's state.converged): 1, else 0
 * (iteration limit or the end of the requested iterations). */
int xcal_simple_converged(const xcal_simple* s, int32_t* converged);
/* Transient incompressible PISO (Solvers_2_PISO.jl:40-211; run! of a
 * Transient model, Solvers_3_solver_dispatch.jl:186-190, calls it with
 * inner_loops = 2): `steps` time steps of setup.dt, each one momentum solve
 * then `inner_loops` pressure correctors, the last one unrelaxed.  Laminar
 * models; every step runs (no convergence stop).  residuals, done and
 * krylov_iters as for xcal_simple_run, per time step (p: the last
 * corrector's solve). */
int xcal_piso_run(xcal_simple* s, int32_t steps, int32_t inner_loops, double* residuals,
                  int32_t* done, int32_t* krylov_iters);
/* Batch meshes: residuals and krylov_iters laid out [case][field][iteration]
 * (ncase*4*iterations); the run stops when every case has converged.
 * xcal_simple_run on a batch reports the largest value over the cases. */
int xcal_simple_run_batch(xcal_simple* s, int32_t iterations, double* residuals, int32_t* done,
                          int32_t* krylov_iters);
/* Batch meshes: kinematic viscosity of each case (setup.nu is the default) */
int xcal_simple_set_case_nu(xcal_simple* s, const double* nu);
/* Device time (ms, HIP events) spent in the last xcal_simple_run by stage:
 * [0] momentum assembly+BCs, [1] momentum solves, [2] pressure assembly,
 * [3] pressure solve, [4] gradients/interpolation/flux/corrections, [5] total */
int xcal_simple_stage_times(const xcal_simple* s, double* ms6);

/* Serving (no single reference counterpart: it runs many independent
 * run!(model, config) cases).  Requests are built at submission (the matrix
 * connectivity on the device, as xcal_connectivity_create; the rest on the
 * host) and wait in a FIFO queue.  Each of `lanes` scheduler threads (one by
 * default, each on its own stream) keeps its cases in flight as one batch
 * mesh and steps them together, every launch of a SIMPLE iteration covering
 * all of them (continuous batching); the lanes' launches overlap on the
 * device.  The queue head joins when its device footprint fits the HBM
 * budget left by the cases in flight and fewer than `slots` cases are in
 * flight over all lanes.  It joins the lane whose cases are closest to it
 * in size (largest case within a factor 4, else an empty lane), among lanes
 * that are empty or whose settings (SIMPLE, turbulence, boundary
 * conditions) equal its own.  A case leaves after the iteration in which it converges
 * (the run!/simple! convergence test), reaches max_iterations, or shows a
 * non-finite residual (failed); its fields are then kept on the host until
 * released.  A case's result does not depend on which other cases are in
 * flight. */
typedef struct xcal_server xcal_server;
typedef struct {
  int32_t device;
  int32_t slots;          /* most cases in flight together, 1..256 */
  int64_t hbm_budget;     /* bytes of device memory the cases may occupy */
  int32_t max_iterations; /* SIMPLE iterations per case at most */
  /* batches stepped concurrently, 0..16 (0: one).  Each lane has its own
   * stream, as has the submission path: past the process's hardware queues
   * (GPU_MAX_HW_QUEUES, 4 by default) streams share queues, and 4 lanes
   * measured slower than 2. */
  int32_t lanes;
} xcal_server_config;
enum { XCAL_REQ_QUEUED = 0, XCAL_REQ_RUNNING = 1, XCAL_REQ_DONE = 2, XCAL_REQ_FAILED = 3 };
typedef struct {
  int32_t state;
  int32_t iterations;   /* SIMPLE iterations performed */
  int64_t cells;
  int64_t footprint;    /* admitted device bytes (estimate) */
  double t_submit, t_start, t_end; /* seconds since server creation */
  int32_t converged;    /* finished on the convergence test (not max_iterations) */
  int32_t lane;         /* the lane that admitted it (-1 while queued) */
  int64_t batch_rebuilds; /* batches the server has formed so far */
} xcal_request_status;
int xcal_server_create(const xcal_server_config* cfg, xcal_server** out);
/* waits for the batch in flight; queued cases are dropped */
int xcal_server_destroy(xcal_server* s);
/* turb: NULL or XCAL_TURB_LAMINAR for a laminar case.  Initial fields may be
 * NULL (zero start): U0 3*n_cells (x/y/z blocks), p0, k0, omega0, nut0 n_cells
 * (the last three only with a turbulence model). */
int xcal_server_submit(xcal_server* s, const xcal_mesh_host* mesh, const xcal_simple_setup* setup,
                       const xcal_turbulence_setup* turb, const double* U0, const double* p0,
                       const double* k0, const double* omega0, const double* nut0, int64_t* id);
int xcal_server_poll(xcal_server* s, int64_t id, xcal_request_status* status);
/* id > 0: wait for that request; id <= 0: until nothing is queued or running.
 * timeout_s <= 0 waits without limit; returns 2 on timeout. */
int xcal_server_wait(xcal_server* s, int64_t id, double timeout_s);
/* which: XCAL_SIMPLE_U, _P, or (turbulent cases) _K, _OMEGA, _NUT */
int xcal_server_fetch(xcal_server* s, int64_t id, int32_t which, double* host);
int xcal_server_release(xcal_server* s, int64_t id);

/* Kernel timing with HIP events on the context's stream (no reference
 * counterpart; measurement support).  While a kernel's timing is enabled,
 * every 8th launch of it, starting with the first after the enable, is
 * bracketed by an event pair.  Kernels (each enabled independently):
 *   0 = CG SpMV (p = z + beta p; q = A p; p.q),
 *   1 = forward DILU sweep (fused x, r update), 2 = backward DILU sweep.
 * xcal_kernel_timing_read returns the kernel's timed launches and their
 * summed milliseconds since its last enable. */
int xcal_kernel_timing_enable(xcal_ctx* ctx, int32_t kernel, int32_t enable);
int xcal_kernel_timing_read(xcal_ctx* ctx, int32_t kernel, int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif
