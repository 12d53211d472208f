/*
 * oracle.h -- CPU restatement of XCALibre's FP64 SIMPLE hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * gfx950 implementation (libxcalhip.so) and the timed CPU baseline; it is
 * never part of the product path.  Each function restates the reference
 * function named beside it, in the same loop order and floating-point
 * evaluation order (built with -ffp-contract=off).  Cell and face loops that
 * the reference runs as independent work-items may run in OpenMP threads;
 * every accumulation that the reference performs with atomics or in a
 * sequential loop is performed sequentially here in ascending index order.
 *
 * Index conventions: all ids crossing this ABI are 0-based; the mesh comes
 * in as xcal_mesh_host (1-based, reference layout) and is converted once.
 */
#ifndef XCAL_ORACLE_H
#define XCAL_ORACLE_H

#include <stdint.h>

#include "xcal.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh orc_mesh;

/* operator term over host arrays; flux == NULL means constant flux_const */
typedef struct {
  int32_t type; /* XCAL_TIME_STEADY ... XCAL_SI */
  int32_t sign;
  const double* flux;
  double flux_const;
} orc_term;

/* source term; y,z are used by vector equations only */
typedef struct {
  const double* x;
  const double* y;
  const double* z;
  int32_t sign;
} orc_source;

const char* orc_last_error(void);
int orc_set_threads(int32_t n);

int orc_mesh_create(const xcal_mesh_host* host, orc_mesh** out);
void orc_mesh_destroy(orc_mesh* m);
int64_t orc_nnz(const orc_mesh* m);
/* CSR pattern (0-based): rowptr n+1, colval nnz */
void orc_csr(const orc_mesh* m, int64_t* rowptr, int64_t* colval);

/* discretise! scalar (Discretise_2_generated_distretisation.jl:89-167) */
int orc_discretise_scalar(const orc_mesh* m, const orc_term* t, int32_t nt,
                          const orc_source* s, int32_t ns, const double* prev, double dt,
                          double* nzval, double* b);
/* discretise! vector (Discretise_2_generated_distretisation.jl:3-87) -> A0 */
int orc_discretise_vector(const orc_mesh* m, const orc_term* t, int32_t nt,
                          const orc_source* s, int32_t ns, const double* prevx,
                          const double* prevy, const double* prevz, double dt, double* nzval0,
                          double* bx, double* by, double* bz);
/* apply_boundary_conditions! (Discretise_5_apply_bcs.jl:5-123).
 * component: -1 scalar equation, 0..2 vector component.  phi: the unknown
 * (x,y,z arrays for vectors; phix only for scalars). */
int orc_apply_bcs(const orc_mesh* m, const orc_term* t, int32_t nt, const xcal_bc* bcs,
                  int32_t nbc, int32_t component, const double* phix, const double* phiy,
                  const double* phiz, double* nzval, double* b);
/* implicit_relaxation_diagdom! (Solve_1_api.jl:370-411) */
void orc_relax_implicit_diagdom(const orc_mesh* m, double* nzval, double* b,
                                const double* field, double alpha);
/* setReference! (Solve_1_api.jl:414-439) */
void orc_set_reference(const orc_mesh* m, double* nzval, double* b, double pref, int64_t cell);
/* SpMV y = A x (Multithread/spmvm.jl:35-59) */
void orc_spmv(const orc_mesh* m, const double* nzval, const double* x, double* y);
/* residual (Solve_1_api.jl:441-464) */
double orc_residual(const orc_mesh* m, const double* nzval, const double* b, const double* x);
/* DILU (preconditioners_1_DILU.jl:43-120): D from A, then z = P \ r */
void orc_dilu_diagonal(const orc_mesh* m, const double* nzval, double* D);
void orc_dilu_apply(const orc_mesh* m, const double* nzval, const double* D, const double* r,
                    double* z);
/* Krylov CG / BiCGSTAB as called from solve_system! (Solve_1_api.jl:268-307).
 * x: warm start in, solution out.  Returns 1 when itmax was reached. */
int orc_krylov(const orc_mesh* m, int32_t solver, int32_t precon, const double* nzval,
               const double* b, double* x, int32_t itmax, double atol, double rtol,
               int32_t* niter);

/* apply_smoother!(JacobiSmoother) (Smoothers_jacobi.jl:54-110): `loops`
 * weighted-Jacobi sweeps of x, each reading the previous sweep's values */
void orc_smooth_jacobi(const orc_mesh* m, const double* nzval, const double* b, double* x,
                       int32_t loops, double omega);
/* solve_system! (Solve_1_api.jl:268-307) around the Krylov call: the setup's
 * smoother on x, the Krylov solve warm-started from it, and with cn != 0
 * (the equation's first term is Time{CrankNicolson}) x = 2 x - x_old. */
int orc_solve_system(const orc_mesh* m, const xcal_solver_setup* setup, int32_t cn,
                     const double* nzval, const double* b, double* x, int32_t* niter);

/* interpolate! scalar (Calculate_2_interpolation.jl:40-76) */
void orc_interpolate(const orc_mesh* m, const double* phi, double* phif);
/* correct_boundaries! (1_boundary_interpolation.jl); vector: component 0..2
 * arrays given via phix/phiy/phiz and outputs fx/fy/fz; scalar: fy==NULL */
int orc_correct_boundaries(const orc_mesh* m, const xcal_bc* bcs, int32_t nbc,
                           const double* phix, const double* phiy, const double* phiz,
                           double* fx, double* fy, double* fz);
/* green_gauss! scalar -> vector (Calculate_1_green_gauss.jl:5-78) */
void orc_green_gauss(const orc_mesh* m, const double* phif, double* gx, double* gy,
                     double* gz);
/* green_gauss! vector -> tensor (Calculate_1_green_gauss.jl:82-173);
 * t: 9 arrays xx,xy,xz,yx,yy,yz,zx,zy,zz */
void orc_green_gauss_tensor(const orc_mesh* m, const double* fx, const double* fy,
                            const double* fz, double* const* t);
/* grad! with Grad{Midpoint} (Calculate_0_gradient.jl:121-250): phi holds 1
 * (scalar) or 3 (vector) cell arrays, phif as many face arrays (the final
 * face values), grad 3 or 9 arrays (tensor row p = component p). */
int orc_grad_midpoint(const orc_mesh* m, const xcal_bc* bcs, int32_t nbc, int32_t ncomp,
                      const double* const* phi, double* const* phif, double* const* grad);
/* div! (Calculate_0_divergence.jl:99-171) */
void orc_div(const orc_mesh* m, const double* psif, double* div);
/* flux! (Solvers_0_functions.jl:41-64) */
void orc_flux(const orc_mesh* m, const double* fx, const double* fy, const double* fz,
              double* mdot);
/* H! (Solvers_0_functions.jl:184-249) */
void orc_H(const orc_mesh* m, const double* nzval, const double* bx, const double* by,
           const double* bz, const double* Ux, const double* Uy, const double* Uz,
           double* Hx, double* Hy, double* Hz);
/* correct_mass_flux! + correct_boundary_mass_flux! (Solvers_1_SIMPLE.jl:340-557) */
int orc_correct_mass_flux(const orc_mesh* m, const double* p_nzval, const double* p,
                          const double* pflux, int32_t psign, const xcal_bc* bcs,
                          int32_t nbc, double* mdot);

/* nonorthogonal_face_correction (Solvers_1_SIMPLE.jl:267-336): b of each
 * internal face's owner += (flux gradf).T_hat, of its neighbour -= (added
 * per cell in face-list order; the reference's atomics fix no order) */
int orc_nonorth_correction(const orc_mesh* m, const double* gx, const double* gy,
                           const double* gz, const double* p, const double* flux, double* b);

/* ---- pieces of turbulence!(KOmegaModel) (RANS_kOmega.jl:175-235) ---- */
/* magnitude2! of S = (gradU + gradU')/2 times 2 (RANS_tensor_algebra.jl:68):
 * Pw = alpha1 * 2 S:S, Pk = nut * 2 S:S.  gU: 9 arrays, gU[3*j+q] = dU_j/dx_q */
void orc_kw_production(const orc_mesh* m, const double* const* gU, const double* nut,
                       double alpha1, double* Pk, double* Pw);
/* correct_production! (RANS_functions.jl:62-124), wall velocity 0 */
int orc_kw_wall_production(const orc_mesh* m, const xcal_bc* bc_k, int32_t nbc, double nu,
                           const double* Ux, const double* Uy, const double* Uz,
                           const double* k, double* Pk);
/* Dw = beta1 omega, Dk = beta* omega (cells); mueff of omega and k (faces) */
void orc_kw_coeffs(const orc_mesh* m, const double* omega, const double* nutf, double nu,
                   double beta1, double beta_star, double sigma_omega, double sigma_k, double* Dw,
                   double* Dk, double* mw, double* mk);
/* constrain_equation!/constrain! (RANS_functions.jl:262-356) */
int orc_kw_constrain(const orc_mesh* m, const xcal_bc* bc_omega, int32_t nbc, double nu,
                     const double* k, double* nzval, double* b);
/* bound! (RANS_functions.jl:2-38) in place */
void orc_bound(const orc_mesh* m, double* v);
/* nut = k/omega, nutf by interpolate!/correct_boundaries!, then
 * correct_eddy_viscosity! on NutWallFunction faces (RANS_functions.jl:126-200) */
int orc_kw_nut(const orc_mesh* m, const xcal_bc* bc_nut, int32_t nbc, double nu, const double* k,
               const double* omega, double* nut, double* nutf);
/* wall-function constants for viscosity nu: cmu^1/4, y+_lam (10 fixed-point
 * iterations of the log law, RANS_functions.jl:40-43), kappa, E */
void orc_kw_wall_constants(double nu, double* out4);

/* Whole steady SIMPLE run (Solvers_1_SIMPLE.jl:45-263, laminar).
 * U: 3 arrays (x,y,z) in/out, p in/out.  residuals: 4*iterations
 * [Ux.., Uy.., Uz.., p..]; krylov_iters: 4*iterations (may be NULL);
 * mdotf: n_faces output (may be NULL). */
int orc_simple_run(const orc_mesh* m, const xcal_simple_setup* setup, double* Ux, double* Uy,
                   double* Uz, double* p, int32_t iterations, double* residuals,
                   int32_t* done, int32_t* krylov_iters, double* mdotf);
/* Transient PISO run (Solvers_2_PISO.jl:40-211, laminar): `steps` time
 * steps of setup->dt with `inner_loops` pressure correctors each, the last
 * one unrelaxed; every step is taken.  Arguments as for orc_simple_run. */
int orc_piso_run(const orc_mesh* m, const xcal_simple_setup* setup, double* Ux, double* Uy,
                 double* Uz, double* p, int32_t steps, int32_t inner_loops, double* residuals,
                 int32_t* done, int32_t* krylov_iters, double* mdotf);
/* the same with the k-omega model (RANS_kOmega.jl:175-235) when t is given;
 * k, omega, nut (n_cells each) are in/out */
int orc_simple_run_turb(const orc_mesh* m, const xcal_simple_setup* setup,
                        const xcal_turbulence_setup* t, double* Ux, double* Uy, double* Uz,
                        double* p, double* k, double* omega, double* nut, int32_t iterations,
                        double* residuals, int32_t* done, int32_t* krylov_iters, double* mdotf);

#ifdef __cplusplus
}
#endif
#endif
