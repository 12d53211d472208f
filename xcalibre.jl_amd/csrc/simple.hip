// Steady incompressible SIMPLE driver on the device: the loop of
// SIMPLE (Solvers_1_SIMPLE.jl:103-263) behind simple!/run!
// (Solvers_1_SIMPLE.jl:28-42, Solvers_3_solver_dispatch.jl:120-135) with a
// laminar viscosity model (RANS_laminar.jl), Gauss gradients and no
// non-orthogonal correctors.  All fields stay resident in HBM; the host only
// sequences launches and reads back residual scalars.
#include <cstring>

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
  ~xcal_simple() {
    for (xcal_field* f : {U, p, Uf, pf, gradp, mdotf, rDf, nueff, divHv, Hv, rD, prev, gradU})
      if (f) xcal_field_destroy(f);
    if (Ueqn) xcal_eqn_destroy(Ueqn);
    if (peqn) xcal_eqn_destroy(peqn);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

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
}

// grad!(∇p, pf, p, BCs.p) (Calculate_0_gradient.jl:105-109)
void grad_p(xcal_simple* S) {
  xcal_mesh* m = S->mesh;
  interpolate(m, S->p->comp(0), S->pf->comp(0));
  const double* ph[1] = {S->p->comp(0)};
  double* fc[1] = {S->pf->comp(0)};
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
    xcal_field* f = which == XCAL_SIMPLE_U ? S->U : (which == XCAL_SIMPLE_P ? S->p : nullptr);
    XCAL_CHECK(f, "only U and p can be set");
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
    const xcal_field* f = which == XCAL_SIMPLE_U ? S->U
                          : which == XCAL_SIMPLE_P ? S->p
                          : which == XCAL_SIMPLE_MDOTF ? S->mdotf
                                                       : nullptr;
    XCAL_CHECK(f, "unknown field selector");
    for (int c = 0; c < f->ncomp; ++c) {
      int rc = xcal_field_download(f, c, host + (int64_t)c * f->len);
      if (rc) throw Error(rc, xcal_last_error());
    }
    return 0;
  });
}

int xcal_simple_run(xcal_simple* S, int32_t iterations, double* residuals, int32_t* done,
                    int32_t* kit) {
  return guarded([&] {
    XCAL_CHECK(S && residuals && done && iterations >= 0, "bad arguments");
    xcal_mesh* m = S->mesh;
    hipStream_t st = m->ctx->stream;
    const xcal_simple_setup& cfg = S->setup;
    const int64_t n = m->n;
    XCAL_HIP(hipSetDevice(m->ctx->device));
    // operator definitions (Solvers_1_SIMPLE.jl:68-78)
    xcal_term Ut[3] = {{XCAL_TIME_STEADY, 1, nullptr, 1.0},
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
    std::memset(S->stage_ms, 0, sizeof(S->stage_ms));

    // initial calculations (Solvers_1_SIMPLE.jl:150-156)
    if (!S->initialised) {
      interp_vec(S, S->U);
      flux(m, S->Uf->comp(0), S->Uf->comp(1), S->Uf->comp(2), S->mdotf->comp(0));
      grad_p(S);
      fill(st, S->nueff->comp(0), m->nf, cfg.nu);  // update_nueff! (laminar)
      S->initialised = true;
    }
    float acc[6] = {0, 0, 0, 0, 0, 0};
    auto mark = [&](int k) { XCAL_HIP(hipEventRecord(S->ev[k], st)); };
    int it = 0;
    XCAL_HIP(hipEventRecord(S->ev[6], st));
    for (it = 0; it < iterations; ++it) {
      // solve_equation! vector (Solve_1_api.jl:233-266)
      mark(0);
      discretise(Ue, UT, US, S->U, 1.0);
      mark(1);
      double res[3] = {0.0, 0.0, 0.0};
      int itc[3] = {0, 0, 0};
      float tA = 0.0f, tS = 0.0f;
      for (int c = 0; c < (is3d ? 3 : 2); ++c) {
        mark(2);
        copy_nz(m, Ue->nzval.p, Ue->nzval0.p);  // update_equation!
        apply_bcs(Ue, UT, S->bU, S->U, c);
        relax_diagdom(Ue, S->U->comp(c), Ue->bc(c), cfg.U.relax);
        if (c == 0) precond_update(Ue, cfg.U.precon);
        mark(3);
        xcal_solve_stats ss{};
        solve_system(Ue, &cfg.U, S->U->comp(c), Ue->bc(c), &ss);
        mark(4);
        res[c] = ss.residual;
        itc[c] = ss.niter;
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
      // pressure correction (Solvers_1_SIMPLE.jl:170-185)
      mark(0);
      inverse_diagonal(Ue, S->rD->comp(0));
      interpolate(m, S->rD->comp(0), S->rDf->comp(0));
      remove_pressure_source(Ue, S->gradp->comp(0), S->gradp->comp(1), S->gradp->comp(2), -1);
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
      if (cfg.has_pref) set_reference(pe, cfg.pref, m->cperm[0]);  // reference cell 1
      precond_update(pe, cfg.p.precon);
      mark(2);
      xcal_solve_stats ps{};
      solve_system(pe, &cfg.p, S->p->comp(0), pe->bc(0), &ps);
      mark(3);
      relax_explicit(m, n, S->p->comp(0), S->prev->comp(0), cfg.p.relax);
      grad_p(S);
      correct_mass_flux(pe, S->p->comp(0), S->rDf->comp(0), -1, S->bp, S->mdotf->comp(0));
      correct_velocity(m, S->U->comp(0), S->U->comp(1), S->U->comp(2), S->Hv->comp(0),
                       S->Hv->comp(1), S->Hv->comp(2), S->gradp->comp(0), S->gradp->comp(1),
                       S->gradp->comp(2), S->rD->comp(0));
      // turbulence! laminar (RANS_laminar.jl:73-79): tensor gradient of U
      interp_vec(S, S->U);
      green_gauss_tensor(m, S->Uf->comp(0), S->Uf->comp(1), S->Uf->comp(2), S->gradU->comp(0));
      fill(st, S->nueff->comp(0), m->nf, cfg.nu);  // update_nueff!
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
      residuals[it] = res[0];
      residuals[iterations + it] = res[1];
      residuals[2 * iterations + it] = res[2];
      residuals[3 * iterations + it] = ps.residual;
      if (kit) {
        kit[it] = itc[0];
        kit[iterations + it] = itc[1];
        kit[2 * iterations + it] = itc[2];
        kit[3 * iterations + it] = ps.niter;
      }
      const bool zc = is3d ? res[2] <= cfg.U.convergence : true;
      if (res[0] <= cfg.U.convergence && res[1] <= cfg.U.convergence && zc &&
          ps.residual <= cfg.p.convergence) {
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
    *done = it;
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
