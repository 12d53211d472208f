// Incompressible pressure-velocity drivers on the device: the steady loop of
// SIMPLE (Solvers_1_SIMPLE.jl:103-263) behind simple!/run!
// (Solvers_1_SIMPLE.jl:28-42, Solvers_3_solver_dispatch.jl:120-135), and the
// transient loop of PISO (Solvers_2_PISO.jl:40-211) behind run! of a
// Transient model (Solvers_3_solver_dispatch.jl:186-190).  Laminar or
// k-omega (SIMPLE) viscosity, Gauss or Midpoint gradients, and `ncorrectors`
// non-orthogonal correctors of the pressure.  All fields stay resident in
// HBM; the host only sequences launches and reads back residual scalars.
#include <algorithm>
#include <cstring>
#include <vector>

#include "device_common.h"

using namespace xcal;

struct xcal_simple {
  xcal_mesh* mesh = nullptr;
  xcal_simple_setup setup{};
  std::vector<xcal_bc> bcU, bcp;
  BCsK bU{}, bp{};
  // fields (Solvers_1_SIMPLE.jl:57-64, 127-137)
  xcal_field *U = nullptr, *p = nullptr, *Uf = nullptr, *pf = nullptr, *gradp = nullptr,
             *mdotf = nullptr, *rDf = nullptr, *nueff = nullptr, *divHv = nullptr, *Hv = nullptr,
             *rD = nullptr, *prev = nullptr, *gradU = nullptr;
  xcal_eqn *Ueqn = nullptr, *peqn = nullptr;
  hipEvent_t ev[8]{};
  double stage_ms[6]{};
  bool initialised = false;
  // serving policy: end the run at the first non-finite residual
  bool stop_nonfinite = false, diverged = false;
  // the last run stopped on the convergence test of Solvers_1_SIMPLE.jl:226-233
  // (every residual, the turbulence model's included, below its setting)
  bool converged = false;
  // k-omega model (RANS_kOmega.jl), off unless set
  bool turb = false;
  xcal_turbulence_setup tset{};
  std::vector<xcal_bc> bcK, bcW, bcN;
  BCsK bK{}, bW{}, bN{};
  xcal_field *k = nullptr, *omega = nullptr, *nut = nullptr, *nutf = nullptr, *Pk = nullptr,
             *Pw = nullptr, *Dk = nullptr, *Dw = nullptr, *mk = nullptr, *mw = nullptr,
             *tmp = nullptr;
  double res_k = 0.0, res_w = 0.0;
  // per case of the last iteration: k and omega residuals, the convergence
  // test of Solvers_1_SIMPLE.jl:226-233 (turbulence included), and any
  // non-finite residual
  std::vector<double> case_res_k, case_res_w;
  std::vector<char> case_conv, case_bad;
  // batch meshes: viscosity per case, reference cell (device id) per case
  std::vector<double> case_nu;
  xcal::DBuf<double> case_nu_d;
  xcal::DBuf<int32_t> ref_cells;
  ~xcal_simple() {
    for (xcal_field* f : {U, p, Uf, pf, gradp, mdotf, rDf, nueff, divHv, Hv, rD, prev, gradU, k,
                          omega, nut, nutf, Pk, Pw, Dk, Dw, mk, mw, tmp})
      if (f) xcal_field_destroy(f);
    if (Ueqn) xcal_eqn_destroy(Ueqn);
    if (peqn) xcal_eqn_destroy(peqn);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

// update_nueff! (laminar): nueff = nu, per case on a batch mesh
__global__ void k_fill_by_case(double* __restrict__ f, const int32_t* __restrict__ face_case,
                               const double* __restrict__ nu, int64_t nf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nf) f[i] = nu[face_case[i]];
}

// setReference! in the reference cell of every case
__global__ void k_set_reference_cases(const int32_t* __restrict__ diag, double* nzval, double* b,
                                      double pref, const int32_t* __restrict__ cells, int ncase) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncase) return;
  const int cell = cells[k];
  const int c = diag[cell];
  b[cell] = nzval[c] * pref;
  nzval[c] += nzval[c];
}

void update_nueff(xcal_simple* S) {
  xcal_mesh* m = S->mesh;
  if (S->turb)
    turb_nueff(m, S->setup.nu, S->nutf->comp(0), S->nueff->comp(0));  // nueff = nu + nutf
  else if (m->batch)
    hipLaunchKernelGGL(k_fill_by_case, dim3(grid_for(m->nf)), dim3(kBlock), 0, m->ctx->stream,
                       S->nueff->comp(0), m->face_case.p, S->case_nu_d.p, m->nf);
  else
    fill(m->ctx->stream, S->nueff->comp(0), m->nf, S->setup.nu);
  XCAL_HIP(hipGetLastError());
}

xcal_field* mk(xcal_mesh* m, int kind) {
  xcal_field* f = nullptr;
  int rc = xcal_field_create(m, kind, &f);
  if (rc) throw Error(rc, xcal_last_error());
  return f;
}

void check_setup(const xcal_mesh* m, const xcal_simple_setup* s) {
  XCAL_CHECK(s->n_bcs == m->n_patches, "one boundary condition per patch is required for U and p");
  XCAL_CHECK(s->div_scheme >= XCAL_DIV_LINEAR && s->div_scheme <= XCAL_DIV_BOUNDED_UPWIND,
             "momentum divergence scheme must be a Divergence type");
  for (const xcal_solver_setup* ss : {&s->U, &s->p}) {
    XCAL_CHECK(ss->solver == XCAL_SOLVER_CG || ss->solver == XCAL_SOLVER_BICGSTAB, "unknown solver");
    XCAL_CHECK(ss->precon >= XCAL_PRECON_NONE && ss->precon <= XCAL_PRECON_DILU,
               "unknown preconditioner");
    XCAL_CHECK(ss->relax > 0.0, "relaxation factor must be positive");
    XCAL_CHECK(ss->itmax >= 0, "itmax must be >= 0");
  }
  for (int i = 0; i < s->n_bcs; ++i)
    XCAL_CHECK(s->bc_U[i].kind != XCAL_BC_NEUMANN, "Neumann is not defined for vector fields");
  for (int g : {s->grad_p, s->grad_U})
    XCAL_CHECK(g == XCAL_GRAD_GAUSS || g == XCAL_GRAD_MIDPOINT, "gradient scheme: Gauss or Midpoint");
  XCAL_CHECK(s->time_U == XCAL_TIME_STEADY || s->time_U == XCAL_TIME_EULER ||
                 s->time_U == XCAL_TIME_CRANK_NICOLSON,
             "momentum time scheme: SteadyState, Euler or CrankNicolson");
  XCAL_CHECK(s->time_U == XCAL_TIME_STEADY || s->dt > 0.0,
             "Time{Euler} and Time{CrankNicolson} need runtime.dt > 0");
  for (const xcal_solver_setup* ss : {&s->U, &s->p})
    XCAL_CHECK(ss->smoother_loops >= 0, "smoother loops must be >= 0");
  XCAL_CHECK(s->ncorrectors >= 0, "ncorrectors must be >= 0");
}

// grad!(∇p, pf, p, BCs.p) (Calculate_0_gradient.jl:105-109)
void grad_p(xcal_simple* S) {
  xcal_mesh* m = S->mesh;
  const double* ph[1] = {S->p->comp(0)};
  double* fc[1] = {S->pf->comp(0)};
  if (S->setup.grad_p == XCAL_GRAD_MIDPOINT) {  // Grad{Midpoint} (:235-244)
    grad_midpoint(m, S->bp, ph, fc, 1, S->gradp->comp(0));
    return;
  }
  interpolate(m, S->p->comp(0), S->pf->comp(0));
  correct_boundaries(m, S->bp, ph, fc, 1);
  green_gauss(m, S->pf->comp(0), S->gradp->comp(0), S->gradp->comp(1), S->gradp->comp(2));
}

// interpolate!(Uf, V) + correct_boundaries!(Uf, V, BCs.U)
void interp_vec(xcal_simple* S, const xcal_field* V) {
  xcal_mesh* m = S->mesh;
  for (int c = 0; c < 3; ++c) interpolate(m, V->comp(c), S->Uf->comp(c));
  const double* ph[3] = {V->comp(0), V->comp(1), V->comp(2)};
  double* fc[3] = {S->Uf->comp(0), S->Uf->comp(1), S->Uf->comp(2)};
  correct_boundaries(m, S->bU, ph, fc, 3);
}

}  // namespace

namespace xcal {
void simple_stop_nonfinite(xcal_simple* S, bool on) { S->stop_nonfinite = on; }
bool simple_diverged(const xcal_simple* S) { return S->diverged; }
}  // namespace xcal

extern "C" {

int xcal_simple_create(xcal_mesh* m, const xcal_simple_setup* s, xcal_simple** out) {
  return guarded([&] {
    XCAL_CHECK(m && s && out, "NULL argument");
    check_setup(m, s);
    auto S = std::make_unique<xcal_simple>();
    S->mesh = m;
    S->setup = *s;
    S->bcU.assign(s->bc_U, s->bc_U + s->n_bcs);
    S->bcp.assign(s->bc_p, s->bc_p + s->n_bcs);
    S->setup.bc_U = S->bcU.data();
    S->setup.bc_p = S->bcp.data();
    S->bU = make_bcs(m, S->bcU.data(), s->n_bcs);
    S->bp = make_bcs(m, S->bcp.data(), s->n_bcs);
    S->U = mk(m, XCAL_CELL_VECTOR);
    S->p = mk(m, XCAL_CELL_SCALAR);
    S->Uf = mk(m, XCAL_FACE_VECTOR);
    S->pf = mk(m, XCAL_FACE_SCALAR);
    S->gradp = mk(m, XCAL_CELL_VECTOR);
    S->mdotf = mk(m, XCAL_FACE_SCALAR);
    S->rDf = mk(m, XCAL_FACE_SCALAR);
    S->nueff = mk(m, XCAL_FACE_SCALAR);
    S->divHv = mk(m, XCAL_CELL_SCALAR);
    S->Hv = mk(m, XCAL_CELL_VECTOR);
    S->rD = mk(m, XCAL_CELL_SCALAR);
    S->prev = mk(m, XCAL_CELL_SCALAR);
    S->gradU = mk(m, XCAL_CELL_TENSOR);
    int rc = xcal_eqn_create(m, 1, &S->Ueqn);
    if (rc) throw Error(rc, xcal_last_error());
    rc = xcal_eqn_create(m, 0, &S->peqn);
    if (rc) throw Error(rc, xcal_last_error());
    fill(m->ctx->stream, S->rDf->comp(0), m->nf, 1.0);  // initialise!(rDf, 1.0)
    if (m->batch) {
      S->case_nu.assign(m->ncase, s->nu);
      S->case_nu_d.alloc(m->ncase);
      XCAL_HIP(hipMemcpy(S->case_nu_d.p, S->case_nu.data(), sizeof(double) * m->ncase,
                         hipMemcpyHostToDevice));
      std::vector<int32_t> refs(m->ncase);
      for (int k = 0; k < m->ncase; ++k) refs[k] = m->cperm[m->case_cell[k]];  // cell 1 of case k
      S->ref_cells.alloc(m->ncase);
      XCAL_HIP(hipMemcpy(S->ref_cells.p, refs.data(), sizeof(int32_t) * m->ncase,
                         hipMemcpyHostToDevice));
    }
    for (auto& e : S->ev) XCAL_HIP(hipEventCreate(&e));
    XCAL_HIP(hipStreamSynchronize(m->ctx->stream));
    *out = S.release();
    return 0;
  });
}

int xcal_simple_destroy(xcal_simple* s) {
  return guarded([&] {
    delete s;
    return 0;
  });
}

int xcal_simple_set_field(xcal_simple* S, int32_t which, const double* host) {
  return guarded([&] {
    XCAL_CHECK(S && host, "NULL argument");
    xcal_field* f = which == XCAL_SIMPLE_U       ? S->U
                    : which == XCAL_SIMPLE_P     ? S->p
                    : which == XCAL_SIMPLE_K     ? S->k
                    : which == XCAL_SIMPLE_OMEGA ? S->omega
                    : which == XCAL_SIMPLE_NUT   ? S->nut
                                                 : nullptr;
    XCAL_CHECK(f, "only U, p and (with a turbulence model) k, omega, nut can be set");
    for (int c = 0; c < f->ncomp; ++c) {
      int rc = xcal_field_upload(f, c, host + (int64_t)c * f->len);
      if (rc) throw Error(rc, xcal_last_error());
    }
    S->initialised = false;
    return 0;
  });
}

int xcal_simple_get_field(const xcal_simple* S, int32_t which, double* host) {
  return guarded([&] {
    XCAL_CHECK(S && host, "NULL argument");
    const xcal_field* f = which == XCAL_SIMPLE_U       ? S->U
                          : which == XCAL_SIMPLE_P     ? S->p
                          : which == XCAL_SIMPLE_MDOTF ? S->mdotf
                          : which == XCAL_SIMPLE_K     ? S->k
                          : which == XCAL_SIMPLE_OMEGA ? S->omega
                          : which == XCAL_SIMPLE_NUT   ? S->nut
                                                       : nullptr;
    XCAL_CHECK(f, "unknown field selector (or no turbulence model)");
    for (int c = 0; c < f->ncomp; ++c) {
      int rc = xcal_field_download(f, c, host + (int64_t)c * f->len);
      if (rc) throw Error(rc, xcal_last_error());
    }
    return 0;
  });
}

}  // extern "C"

namespace {

// turbulence!(KOmegaModel) (RANS_kOmega.jl:175-235), after the velocity
// correction; gradU and Uf are current.  Returns whether k and omega met
// their convergence criteria.
// k and omega of case k met their criteria at the last turbulence step
bool case_turb_conv(const xcal_simple* S, int k) {
  const xcal_turbulence_setup& T = S->tset;
  return S->case_res_k[k] < T.k.convergence && S->case_res_w[k] < T.omega.convergence;
}

bool komega_step(xcal_simple* S) {
  xcal_mesh* m = S->mesh;
  const xcal_turbulence_setup& T = S->tset;
  const double nu = S->setup.nu;
  xcal_eqn* e = S->peqn;  // k and omega share the pressure equation's storage
  turb_production(m, S->gradU->comp(0), S->nut->comp(0), T.alpha1, S->Pk->comp(0), S->Pw->comp(0));
  turb_wall_production(m, S->bK, nu, S->U->comp(0), S->U->comp(1), S->U->comp(2), S->k->comp(0),
                       S->Pk->comp(0));
  turb_coeffs(m, S->omega->comp(0), S->nutf->comp(0), nu, T.beta1, T.beta_star, T.sigma_omega,
              T.sigma_k, S->Dw->comp(0), S->Dk->comp(0), S->mw->comp(0), S->mk->comp(0));
  auto transport = [&](xcal_field* phi, xcal_field* D, xcal_field* mu, xcal_field* P,
                       const BCsK& bc, const xcal_solver_setup& ss, bool constrain) {
    xcal_term t[4] = {{XCAL_TIME_STEADY, 1, nullptr, 1.0},
                      {T.div_scheme, 1, S->mdotf, 0.0},
                      {XCAL_LAPLACIAN_LINEAR, -1, mu, 0.0},
                      {XCAL_SI, 1, D, 0.0}};
    TermsK TT = make_terms(t, 4);
    SourcesK SS{};
    SS.n = 1;
    SS.sign[0] = 1;
    SS.x[0] = P->comp(0);
    discretise(e, TT, SS, phi, 1.0);
    apply_bcs(e, TT, bc, phi, 0);
    relax_diagdom(e, phi->comp(0), e->bc(0), ss.relax);
    if (constrain) turb_omega_constrain(e, bc, nu, S->k->comp(0));
    precond_update(e, ss.precon);
    std::vector<xcal_solve_stats> st(m->ncase);
    if (m->batch) solve_system_batch(e, &ss, phi->comp(0), e->bc(0), st.data());
    else solve_system(e, &ss, phi->comp(0), e->bc(0), st.data());
    turb_bound(m, phi->comp(0), S->tmp->comp(0));
    std::vector<double> r(m->ncase);
    for (int k = 0; k < m->ncase; ++k) r[k] = st[k].residual;
    return r;
  };
  S->case_res_w = transport(S->omega, S->Dw, S->mw, S->Pw, S->bW, T.omega, true);
  S->case_res_k = transport(S->k, S->Dk, S->mk, S->Pk, S->bK, T.k, false);
  S->res_w = *std::max_element(S->case_res_w.begin(), S->case_res_w.end());
  S->res_k = *std::max_element(S->case_res_k.begin(), S->case_res_k.end());
  turb_nut(m, S->k->comp(0), S->omega->comp(0), S->nut->comp(0));
  interpolate(m, S->nut->comp(0), S->nutf->comp(0));
  const double* ph[1] = {S->nut->comp(0)};
  double* fc[1] = {S->nutf->comp(0)};
  correct_boundaries(m, S->bN, ph, fc, 1);
  turb_nut_wall(m, S->bN, nu, S->k->comp(0), S->nutf->comp(0));
  bool all = true;
  for (int k = 0; k < m->ncase; ++k) all = all && case_turb_conv(S, k);
  return all;
}

// initial calculations (Solvers_1_SIMPLE.jl:150-156), once after the fields
// were set
void initialise(xcal_simple* S) {
  if (S->initialised) return;
  xcal_mesh* m = S->mesh;
  interp_vec(S, S->U);
  flux(m, S->Uf->comp(0), S->Uf->comp(1), S->Uf->comp(2), S->mdotf->comp(0));
  grad_p(S);
  update_nueff(S);  // update_nueff! (laminar)
  S->initialised = true;
}

// The SIMPLE loop (inner == 0), or PISO's (inner >= 1 pressure correctors
// per time step, Solvers_2_PISO.jl:107-211).  R[(k*4 + f)*iterations + it]:
// residual of field f (Ux, Uy, Uz, p) of case k at iteration it; KIT
// likewise for Krylov iterations.  SIMPLE stops when every case has
// converged; PISO runs every step.
int run_impl(xcal_simple* S, int32_t iterations, std::vector<double>& R, std::vector<int32_t>& KIT,
             int inner) {
  {
    xcal_mesh* m = S->mesh;
    const int K = m->ncase;
    S->converged = false;
    R.assign((size_t)K * 4 * iterations, 0.0);
    KIT.assign((size_t)K * 4 * iterations, 0);
    std::vector<xcal_solve_stats> cst(K);
    auto solve = [&](xcal_eqn* e, const xcal_solver_setup& ss, double* x, const double* b) {
      if (m->batch) return solve_system_batch(e, &ss, x, b, cst.data());
      return solve_system(e, &ss, x, b, cst.data());
    };
    hipStream_t st = m->ctx->stream;
    const xcal_simple_setup& cfg = S->setup;
    const int64_t n = m->n;
    XCAL_HIP(hipSetDevice(m->ctx->device));
    // operator definitions (Solvers_1_SIMPLE.jl:68-78)
    xcal_term Ut[3] = {{cfg.time_U, 1, nullptr, 1.0},
                       {cfg.div_scheme, 1, S->mdotf, 0.0},
                       {XCAL_LAPLACIAN_LINEAR, -1, S->nueff, 0.0}};
    xcal_term pt[1] = {{XCAL_LAPLACIAN_LINEAR, -1, S->rDf, 0.0}};
    TermsK UT = make_terms(Ut, 3), PT = make_terms(pt, 1);
    SourcesK US{};
    US.n = 1;
    US.sign[0] = -1;
    US.x[0] = S->gradp->comp(0);
    US.y[0] = S->gradp->comp(1);
    US.z[0] = S->gradp->comp(2);
    SourcesK PS{};
    PS.n = 1;
    PS.sign[0] = -1;
    PS.x[0] = S->divHv->comp(0);
    xcal_eqn* Ue = S->Ueqn;
    xcal_eqn* pe = S->peqn;
    const bool is3d = m->dim == 3;
    const double dt = cfg.time_U != XCAL_TIME_STEADY ? cfg.dt : 1.0;
    const int ncorr = inner > 0 ? inner : 1;
    std::memset(S->stage_ms, 0, sizeof(S->stage_ms));

    initialise(S);
    float acc[6] = {0, 0, 0, 0, 0, 0};
    auto mark = [&](int k) { XCAL_HIP(hipEventRecord(S->ev[k], st)); };
    int it = 0;
    XCAL_HIP(hipEventRecord(S->ev[6], st));
    for (it = 0; it < iterations; ++it) {
      // solve_equation! vector (Solve_1_api.jl:233-266)
      mark(0);
      discretise(Ue, UT, US, S->U, dt);
      mark(1);
      float tA = 0.0f, tS = 0.0f;
      for (int c = 0; c < (is3d ? 3 : 2); ++c) {
        mark(2);
        copy_nz(m, Ue->nzval.p, Ue->nzval0.p);  // update_equation!
        apply_bcs(Ue, UT, S->bU, S->U, c);
        relax_diagdom(Ue, S->U->comp(c), Ue->bc(c), cfg.U.relax);
        if (c == 0) precond_update(Ue, cfg.U.precon);
        mark(3);
        solve(Ue, cfg.U, S->U->comp(c), Ue->bc(c));
        mark(4);
        for (int k = 0; k < K; ++k) {
          R[((size_t)k * 4 + c) * iterations + it] = cst[k].residual;
          KIT[((size_t)k * 4 + c) * iterations + it] = cst[k].niter;
        }
        float a, b;
        XCAL_HIP(hipEventSynchronize(S->ev[4]));
        XCAL_HIP(hipEventElapsedTime(&a, S->ev[2], S->ev[3]));
        XCAL_HIP(hipEventElapsedTime(&b, S->ev[3], S->ev[4]));
        tA += a;
        tS += b;
      }
      {
        float a;
        XCAL_HIP(hipEventElapsedTime(&a, S->ev[0], S->ev[1]));
        acc[0] += a + tA;
        acc[1] += tS;
      }
      // pressure correction (Solvers_1_SIMPLE.jl:170-185; PISO :117-121 then
      // `inner` correctors :123-175)
      mark(0);
      inverse_diagonal(Ue, S->rD->comp(0));
      interpolate(m, S->rD->comp(0), S->rDf->comp(0));
      remove_pressure_source(Ue, S->gradp->comp(0), S->gradp->comp(1), S->gradp->comp(2), -1);
      for (int ic = 0; ic < ncorr; ++ic) {
        if (ic > 0) mark(0);
        Hop(Ue, S->U->comp(0), S->U->comp(1), S->U->comp(2), S->Hv->comp(0), S->Hv->comp(1),
            S->Hv->comp(2));
        interp_vec(S, S->Hv);
        flux(m, S->Uf->comp(0), S->Uf->comp(1), S->Uf->comp(2), S->mdotf->comp(0));
        divergence(m, S->mdotf->comp(0), S->divHv->comp(0));
        XCAL_HIP(hipMemcpyAsync(S->prev->comp(0), S->p->comp(0), sizeof(double) * n,
                                hipMemcpyDeviceToDevice, st));
        mark(1);
        // solve_equation! scalar (Solve_1_api.jl:214-231)
        discretise(pe, PT, PS, S->p, 1.0);
        apply_bcs(pe, PT, S->bp, S->p, 0);
        if (cfg.has_pref) {  // reference cell 1 (of every case)
          if (m->batch)
            hipLaunchKernelGGL(k_set_reference_cases, dim3(grid_for(K)), dim3(kBlock), 0, st,
                               m->diag.p, pe->nzval.p, pe->bc(0), cfg.pref, S->ref_cells.p, K);
          else
            set_reference(pe, cfg.pref, m->cperm[0]);
        }
        precond_update(pe, cfg.p.precon);
        mark(2);
        solve(pe, cfg.p, S->p->comp(0), pe->bc(0));
        mark(3);
        for (int k = 0; k < K; ++k) {
          R[((size_t)k * 4 + 3) * iterations + it] = cst[k].residual;
          KIT[((size_t)k * 4 + 3) * iterations + it] = cst[k].niter;
        }
        // explicit_relaxation!: PISO's last corrector is unrelaxed
        const double prelax = inner > 0 && ic == ncorr - 1 ? 1.0 : cfg.p.relax;
        relax_explicit(m, n, S->p->comp(0), S->prev->comp(0), prelax);
        grad_p(S);
        // non-orthogonal correctors (Solvers_1_SIMPLE.jl:194-206; PISO
        // :148-165 also sets the reference, updates the preconditioner and
        // leaves its last corrector unrelaxed)
        for (int jc = 0; jc < cfg.ncorrectors; ++jc) {
          discretise(pe, PT, PS, S->p, 1.0);
          apply_bcs(pe, PT, S->bp, S->p, 0);
          if (inner > 0 && cfg.has_pref) {
            if (m->batch)
              hipLaunchKernelGGL(k_set_reference_cases, dim3(grid_for(K)), dim3(kBlock), 0, st,
                                 m->diag.p, pe->nzval.p, pe->bc(0), cfg.pref, S->ref_cells.p, K);
            else
              set_reference(pe, cfg.pref, m->cperm[0]);
          }
          nonorthogonal_correction(pe, S->gradp->comp(0), S->gradp->comp(1), S->gradp->comp(2),
                                   S->p->comp(0), S->rDf->comp(0));
          if (inner > 0) precond_update(pe, cfg.p.precon);
          solve(pe, cfg.p, S->p->comp(0), pe->bc(0));
          for (int k = 0; k < K; ++k) {
            R[((size_t)k * 4 + 3) * iterations + it] = cst[k].residual;
            KIT[((size_t)k * 4 + 3) * iterations + it] = cst[k].niter;
          }
          const double nrelax =
              inner > 0 && jc == cfg.ncorrectors - 1 ? 1.0 : cfg.p.relax;
          relax_explicit(m, n, S->p->comp(0), S->prev->comp(0), nrelax);
          grad_p(S);
        }
        correct_mass_flux(pe, S->p->comp(0), S->rDf->comp(0), -1, S->bp, S->mdotf->comp(0));
        correct_velocity(m, S->U->comp(0), S->U->comp(1), S->U->comp(2), S->Hv->comp(0),
                         S->Hv->comp(1), S->Hv->comp(2), S->gradp->comp(0), S->gradp->comp(1),
                         S->gradp->comp(2), S->rD->comp(0));
        if (ic < ncorr - 1) {  // stage times of the correctors before the last
          mark(4);
          float a, b, c, d;
          XCAL_HIP(hipEventSynchronize(S->ev[4]));
          XCAL_HIP(hipEventElapsedTime(&a, S->ev[0], S->ev[1]));
          XCAL_HIP(hipEventElapsedTime(&b, S->ev[1], S->ev[2]));
          XCAL_HIP(hipEventElapsedTime(&c, S->ev[2], S->ev[3]));
          XCAL_HIP(hipEventElapsedTime(&d, S->ev[3], S->ev[4]));
          acc[4] += a + d;
          acc[2] += b;
          acc[3] += c;
        }
      }
      // turbulence!: grad!(gradU) (RANS_laminar.jl:73-79, RANS_kOmega.jl:192)
      if (cfg.grad_U == XCAL_GRAD_MIDPOINT) {  // Grad{Midpoint} of U
        const double* ph[3] = {S->U->comp(0), S->U->comp(1), S->U->comp(2)};
        double* fc[3] = {S->Uf->comp(0), S->Uf->comp(1), S->Uf->comp(2)};
        grad_midpoint(m, S->bU, ph, fc, 3, S->gradU->comp(0));
      } else {
        interp_vec(S, S->U);
        green_gauss_tensor(m, S->Uf->comp(0), S->Uf->comp(1), S->Uf->comp(2), S->gradU->comp(0));
      }
      const bool turb_conv = S->turb ? komega_step(S) : true;
      update_nueff(S);  // update_nueff!
      mark(4);
      {
        float a, b, c, d;
        XCAL_HIP(hipEventSynchronize(S->ev[4]));
        XCAL_HIP(hipEventElapsedTime(&a, S->ev[0], S->ev[1]));
        XCAL_HIP(hipEventElapsedTime(&b, S->ev[1], S->ev[2]));
        XCAL_HIP(hipEventElapsedTime(&c, S->ev[2], S->ev[3]));
        XCAL_HIP(hipEventElapsedTime(&d, S->ev[3], S->ev[4]));
        acc[4] += a + d;
        acc[2] += b;
        acc[3] += c;
      }
      bool conv = turb_conv;
      S->case_conv.assign(K, 0);
      S->case_bad.assign(K, 0);
      for (int k = 0; k < K; ++k) {
        const double* r = &R[(size_t)k * 4 * iterations];
        bool bad = false;
        for (int f = 0; f < 4; ++f) bad = bad || !std::isfinite(r[(size_t)f * iterations + it]);
        if (S->turb)
          bad = bad || !std::isfinite(S->case_res_k[k]) || !std::isfinite(S->case_res_w[k]);
        const bool zc = is3d ? r[2 * iterations + it] <= cfg.U.convergence : true;
        const bool ck = r[it] <= cfg.U.convergence && r[iterations + it] <= cfg.U.convergence &&
                        zc && r[3 * iterations + it] <= cfg.p.convergence &&
                        (!S->turb || case_turb_conv(S, k));
        S->case_conv[k] = ck;
        S->case_bad[k] = bad;
        conv = conv && ck;
      }
      if (S->stop_nonfinite) {
        for (int k = 0; k < K; ++k) S->diverged = S->diverged || S->case_bad[k];
        if (S->diverged) {
          ++it;
          break;
        }
      }
      if (inner > 0) conv = false;  // a transient run takes every time step
      if (conv) {
        S->converged = true;
        ++it;
        break;
      }
    }
    XCAL_HIP(hipEventRecord(S->ev[7], st));
    XCAL_HIP(hipEventSynchronize(S->ev[7]));
    float tot = 0.0f;
    XCAL_HIP(hipEventElapsedTime(&tot, S->ev[6], S->ev[7]));
    for (int k = 0; k < 5; ++k) S->stage_ms[k] = acc[k];
    S->stage_ms[5] = tot;
    return it;
  }
}

// One case: R holds [Ux..., Uy..., Uz..., p...]; a batch reports the largest
// residual and Krylov count over its cases.
void report(const xcal_simple* S, int32_t iterations, const std::vector<double>& R,
            const std::vector<int32_t>& KIT, double* residuals, int32_t* kit) {
  const int K = S->mesh->ncase;
  for (int f = 0; f < 4; ++f)
    for (int i = 0; i < iterations; ++i) {
      double r = 0.0;
      int32_t q = 0;
      for (int k = 0; k < K; ++k) {
        r = std::max(r, R[((size_t)k * 4 + f) * iterations + i]);
        q = std::max(q, KIT[((size_t)k * 4 + f) * iterations + i]);
      }
      residuals[(size_t)f * iterations + i] = r;
      if (kit) kit[(size_t)f * iterations + i] = q;
    }
}

}  // namespace

namespace xcal {

// One SIMPLE iteration of every case of S (the serving engine steps its
// batch this way).  res: 6 per case (Ux, Uy, Uz, p, k, omega residuals).
void simple_step(xcal_simple* S, std::vector<double>& res) {
  std::vector<double> R;
  std::vector<int32_t> KIT;
  run_impl(S, 1, R, KIT, 0);
  const int K = S->mesh->ncase;
  res.assign((size_t)6 * K, 0.0);
  for (int k = 0; k < K; ++k) {
    for (int f = 0; f < 4; ++f) res[6 * k + f] = R[(size_t)k * 4 + f];
    if (S->turb) {
      res[6 * k + 4] = S->case_res_k[k];
      res[6 * k + 5] = S->case_res_w[k];
    }
  }
}
bool simple_case_converged(const xcal_simple* S, int k) { return S->case_conv.at(k) != 0; }
bool simple_case_bad(const xcal_simple* S, int k) { return S->case_bad.at(k) != 0; }
void simple_initialise(xcal_simple* S) {
  XCAL_HIP(hipSetDevice(S->mesh->ctx->device));
  initialise(S);
}

// Case ks of src becomes case kd of dst: the fields a SIMPLE iteration
// carries to the next (U, p, grad p, mdotf, nueff; k, omega, nut, nutf with
// k-omega) are copied on the device.  Both hold the same case mesh, in its
// own device numbering, on one context.
void simple_copy_case(xcal_simple* dst, int kd, const xcal_simple* src, int ks) {
  const xcal_mesh *md = dst->mesh, *ms = src->mesh;
  XCAL_CHECK(md->ctx == ms->ctx, "cases copied between contexts");
  const int64_t nc = ms->case_cell[ks + 1] - ms->case_cell[ks];
  const int64_t nb = ms->case_bface[ks + 1] - ms->case_bface[ks];
  const int64_t ni = ms->case_iface[ks + 1] - ms->case_iface[ks];
  XCAL_CHECK(md->case_cell[kd + 1] - md->case_cell[kd] == nc &&
                 md->case_bface[kd + 1] - md->case_bface[kd] == nb &&
                 md->case_iface[kd + 1] - md->case_iface[kd] == ni,
             "copied case has another mesh");
  XCAL_CHECK(dst->turb == src->turb, "copied case has another turbulence model");
  hipStream_t st = md->ctx->stream;
  auto cells = [&](xcal_field* d, const xcal_field* s) {
    for (int c = 0; c < s->ncomp; ++c)
      XCAL_HIP(hipMemcpyAsync(d->comp(c) + md->case_cell[kd], s->comp(c) + ms->case_cell[ks],
                              sizeof(double) * nc, hipMemcpyDeviceToDevice, st));
  };
  auto faces = [&](xcal_field* d, const xcal_field* s) {
    XCAL_HIP(hipMemcpyAsync(d->comp(0) + md->case_bface[kd], s->comp(0) + ms->case_bface[ks],
                            sizeof(double) * nb, hipMemcpyDeviceToDevice, st));
    XCAL_HIP(hipMemcpyAsync(d->comp(0) + md->case_iface[kd], s->comp(0) + ms->case_iface[ks],
                            sizeof(double) * ni, hipMemcpyDeviceToDevice, st));
  };
  cells(dst->U, src->U);
  cells(dst->p, src->p);
  cells(dst->gradp, src->gradp);
  faces(dst->mdotf, src->mdotf);
  faces(dst->nueff, src->nueff);
  if (src->turb) {
    cells(dst->k, src->k);
    cells(dst->omega, src->omega);
    cells(dst->nut, src->nut);
    faces(dst->nutf, src->nutf);
  }
}

// Field `which` (XCAL_SIMPLE_*) of case k in its reference numbering.
void simple_case_field(const xcal_simple* S, int which, int k, double* host) {
  const xcal_mesh* m = S->mesh;
  const xcal_field* f = which == XCAL_SIMPLE_U       ? S->U
                        : which == XCAL_SIMPLE_P     ? S->p
                        : which == XCAL_SIMPLE_K     ? S->k
                        : which == XCAL_SIMPLE_OMEGA ? S->omega
                        : which == XCAL_SIMPLE_NUT   ? S->nut
                                                     : nullptr;
  XCAL_CHECK(f, "unknown cell field selector (or no turbulence model)");
  const int64_t c0 = m->case_cell[k], nk = m->case_cell[k + 1] - c0;
  const std::vector<int32_t> perm(m->cperm.begin() + c0, m->cperm.begin() + c0 + nk);
  for (int c = 0; c < f->ncomp; ++c)
    download_permuted(m, perm, f->comp(c), host + (int64_t)c * nk, nk, f->len);
}

}  // namespace xcal

extern "C" {

int xcal_simple_run(xcal_simple* S, int32_t iterations, double* residuals, int32_t* done,
                    int32_t* kit) {
  return guarded([&] {
    XCAL_CHECK(S && residuals && done && iterations >= 0, "bad arguments");
    std::vector<double> R;
    std::vector<int32_t> KIT;
    const int it = run_impl(S, iterations, R, KIT, 0);
    report(S, iterations, R, KIT, residuals, kit);
    *done = it;
    return 0;
  });
}

int xcal_simple_converged(const xcal_simple* S, int32_t* converged) {
  return guarded([&] {
    XCAL_CHECK(S && converged, "NULL argument");
    *converged = S->converged ? 1 : 0;
    return 0;
  });
}

int xcal_piso_run(xcal_simple* S, int32_t steps, int32_t inner_loops, double* residuals,
                  int32_t* done, int32_t* kit) {
  return guarded([&] {
    XCAL_CHECK(S && residuals && done && steps >= 0, "bad arguments");
    XCAL_CHECK(inner_loops >= 1, "PISO needs at least one pressure corrector");
    XCAL_CHECK(!S->turb, "PISO runs laminar models");
    std::vector<double> R;
    std::vector<int32_t> KIT;
    const int it = run_impl(S, steps, R, KIT, inner_loops);
    report(S, steps, R, KIT, residuals, kit);
    *done = it;
    return 0;
  });
}

int xcal_simple_run_batch(xcal_simple* S, int32_t iterations, double* residuals, int32_t* done,
                          int32_t* kit) {
  return guarded([&] {
    XCAL_CHECK(S && residuals && done && iterations >= 0, "bad arguments");
    std::vector<double> R;
    std::vector<int32_t> KIT;
    *done = run_impl(S, iterations, R, KIT, 0);
    std::copy(R.begin(), R.end(), residuals);
    if (kit) std::copy(KIT.begin(), KIT.end(), kit);
    return 0;
  });
}

int xcal_simple_set_turbulence(xcal_simple* S, const xcal_turbulence_setup* t) {
  return guarded([&] {
    XCAL_CHECK(S && t, "NULL argument");
    xcal_mesh* m = S->mesh;
    if (t->model == XCAL_TURB_LAMINAR) {
      S->turb = false;
      S->initialised = false;
      return 0;
    }
    XCAL_CHECK(t->model == XCAL_TURB_KOMEGA, "unknown turbulence model");
    if (m->batch)  // the k-omega kernels take one viscosity
      for (double nu : S->case_nu)
        XCAL_CHECK(nu == S->setup.nu, "k-omega on a batch mesh needs one viscosity for all cases");
    XCAL_CHECK(t->bc_k && t->bc_omega && t->bc_nut, "k, omega and nut need boundary conditions");
    XCAL_CHECK(t->div_scheme >= XCAL_DIV_LINEAR && t->div_scheme <= XCAL_DIV_BOUNDED_UPWIND,
               "k/omega divergence scheme must be a Divergence type");
    for (const xcal_solver_setup* ss : {&t->k, &t->omega}) {
      XCAL_CHECK(ss->solver == XCAL_SOLVER_CG || ss->solver == XCAL_SOLVER_BICGSTAB, "unknown solver");
      XCAL_CHECK(ss->precon >= XCAL_PRECON_NONE && ss->precon <= XCAL_PRECON_DILU,
                 "unknown preconditioner");
      XCAL_CHECK(ss->relax > 0.0, "relaxation factor must be positive");
    }
    const int np = m->n_patches;
    S->tset = *t;
    S->bcK.assign(t->bc_k, t->bc_k + np);
    S->bcW.assign(t->bc_omega, t->bc_omega + np);
    S->bcN.assign(t->bc_nut, t->bc_nut + np);
    S->tset.bc_k = S->bcK.data();
    S->tset.bc_omega = S->bcW.data();
    S->tset.bc_nut = S->bcN.data();
    S->bK = make_bcs(m, S->bcK.data(), np);
    S->bW = make_bcs(m, S->bcW.data(), np);
    S->bN = make_bcs(m, S->bcN.data(), np);
    if (!S->k) {
      S->k = mk(m, XCAL_CELL_SCALAR);
      S->omega = mk(m, XCAL_CELL_SCALAR);
      S->nut = mk(m, XCAL_CELL_SCALAR);
      S->nutf = mk(m, XCAL_FACE_SCALAR);
      S->Pk = mk(m, XCAL_CELL_SCALAR);
      S->Pw = mk(m, XCAL_CELL_SCALAR);
      S->Dk = mk(m, XCAL_CELL_SCALAR);
      S->Dw = mk(m, XCAL_CELL_SCALAR);
      S->mk = mk(m, XCAL_FACE_SCALAR);
      S->mw = mk(m, XCAL_FACE_SCALAR);
      S->tmp = mk(m, XCAL_CELL_SCALAR);
    }
    S->turb = true;
    S->initialised = false;
    return 0;
  });
}

int xcal_simple_set_case_nu(xcal_simple* S, const double* nu) {
  return guarded([&] {
    XCAL_CHECK(S && nu, "NULL argument");
    xcal_mesh* m = S->mesh;
    XCAL_CHECK(m->batch, "per-case viscosity applies to batch meshes");
    for (int k = 0; k < m->ncase; ++k) XCAL_CHECK(nu[k] > 0.0, "viscosity must be positive");
    S->case_nu.assign(nu, nu + m->ncase);
    XCAL_HIP(hipMemcpy(S->case_nu_d.p, nu, sizeof(double) * m->ncase, hipMemcpyHostToDevice));
    S->initialised = false;
    return 0;
  });
}


int xcal_simple_stage_times(const xcal_simple* S, double* ms6) {
  return guarded([&] {
    XCAL_CHECK(S && ms6, "NULL argument");
    for (int k = 0; k < 6; ++k) ms6[k] = S->stage_ms[k];
    return 0;
  });
}

}  // extern "C"
