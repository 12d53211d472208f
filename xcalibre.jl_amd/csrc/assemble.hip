// Cell-gather assembly of Time/Divergence/Laplacian/Source operators into CSR
// nzval and b, boundary conditions, relaxation and reference pressure.
//
// Replaces _discretise_scalar_model!/_discretise_vector_model!
// (Discretise_2_generated_distretisation.jl:39-167) and
// apply_boundary_conditions_kernel! (Discretise_5_apply_bcs.jl:84-123).
// Each cell's work-item visits its internal faces in the reference's
// faces_range order and writes through precomputed CSR slots (no spindex
// searches).  Boundary faces are gathered per cell in ascending face id, so
// the reference's atomic += on the diagonal and b become deterministic
// per-cell sums with no atomics.
#include "device_common.h"

namespace xcal {

template <bool VEC>
__global__ void __launch_bounds__(kBlock)
    k_discretise(MeshK m, TermsK T, SourcesK S, const double* __restrict__ px,
                 const double* __restrict__ py, const double* __restrict__ pz, double dt,
                 double* __restrict__ nzval, double* __restrict__ b0, double* __restrict__ b1,
                 double* __restrict__ b2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  double ac_sum = 0.0;
  // each face's values for every term are loaded for a group of faces at
  // once, then the faces are summed in list order.  Only Laplacian and
  // Divergence terms read a face flux; a term without a face flux field
  // (a constant, or the cell field of a Time or Si term, which has no face
  // contribution) takes the weight in its place and never uses it.
  struct Rec {
    double ns, w, lapEf, delta, J[kMaxTerms];
    int32_t slot;
  };
  cell_faces(
      m.cf_ptr[i], m.cf_ptr[i + 1],
      [&](int32_t k) {
        const int f = m.cf_face[k];
        Rec r;
        r.ns = (double)m.cf_sign[k];
        r.slot = m.cf_slot[k];
        // the face's geometry streams in cf order; only the fluxes gather
        r.w = m.cf_w[k];
        r.lapEf = m.cf_lapEf[k];
        r.delta = m.cf_delta[k];
#pragma unroll
        for (int t = 0; t < kMaxTerms; ++t) {
          const bool face_flux = t < T.n && T.t[t].flux && face_term(T.t[t].type);
          r.J[t] = face_flux ? T.t[t].flux[f] : r.w;
        }
        return r;
      },
      [&](const Rec& r, bool ok) {
        double AC = 0.0, AN = 0.0;  // _scheme! (:173-197)
#pragma unroll
        for (int t = 0; t < kMaxTerms; ++t) {
          if (t < T.n) {
            const double J = T.t[t].flux ? r.J[t] : T.t[t].flux_const;
            double ac, an;
            scheme_vals(T.t[t], J, r.ns, r.w, r.lapEf, r.delta, ac, an);
            AC += ac;
            AN += an;
          }
        }
        ac_sum = ok ? ac_sum + AC : ac_sum;
        if (ok) nzval[r.slot] = AN;
      });
  double AC = 0.0, BX = 0.0, BY = 0.0, BZ = 0.0;  // _scheme_source! (:200-252)
  for (int t = 0; t < T.n; ++t) {
    double ac, bx;
    scheme_source(m, T.t[t], i, px, dt, ac, bx);
    if (VEC) {
      double by, bz;
      scheme_source(m, T.t[t], i, py, dt, ac, by);
      scheme_source(m, T.t[t], i, pz, dt, ac, bz);
      BY += by;
      BZ += bz;
    }
    AC += ac;
    BX += bx;
  }
  nzval[m.diag[i]] = ac_sum + AC;
  const double V = m.vol[i];
  double Bx = 0.0, By = 0.0, Bz = 0.0;  // _sources! (:255-294)
  for (int s = 0; s < S.n; ++s) {
    const double sg = (double)S.sign[s];
    Bx += (sg * S.x[s][i]) * V;
    if (VEC) {
      By += (sg * S.y[s][i]) * V;
      Bz += (sg * S.z[s][i]) * V;
    }
  }
  b0[i] = Bx + BX;
  if (VEC) {
    b1[i] = By + BY;
    b2[i] = Bz + BZ;
  }
}

__global__ void __launch_bounds__(kBlock)
    k_apply_bcs(MeshK m, TermsK T, BCsK B, int comp, const double* __restrict__ px,
                const double* __restrict__ py, const double* __restrict__ pz,
                double* __restrict__ nzval, double* __restrict__ b) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m.nbcell) return;
  const int c = m.bc_cells[s];
  const int dg = m.diag[c];
  double diag = nzval[dg], bc = b[c];
  for (int32_t k = m.bc_ptr[s]; k < m.bc_ptr[s + 1]; ++k) {
    const int f = m.bc_face[k];
    const int p = m.bface_patch[f];
    if (p < 0) continue;
    const BCK& bk = B.p[p];
    if (bk.kind < 0) continue;
    double AP = 0.0, BP = 0.0;  // apply! (Discretise_5_apply_bcs.jl:132-157)
    for (int t = 0; t < T.n; ++t) {
      double ap, bp;
      bc_coeffs(m, bk, T.t[t], f, c, comp, px, py, pz, ap, bp);
      AP += ap;
      BP += bp;
    }
    diag += AP;
    bc += BP;
  }
  nzval[dg] = diag;
  b[c] = bc;
}

// _implicit_relaxation_diagdom! (Solve_1_api.jl:388-411)
__global__ void __launch_bounds__(kBlock)
    k_relax_diagdom(int64_t n, const int32_t* __restrict__ rb, const int32_t* __restrict__ re,
                    const int32_t* __restrict__ diag, double* __restrict__ nzval,
                    double* __restrict__ b, const double* __restrict__ field, double alpha) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sumv = 0.0;
  for (int32_t k = rb[i]; k < re[i]; k += kSlice) sumv += fabs(nzval[k]);
  const int c = diag[i];
  const double D0 = nzval[c];
  sumv -= fabs(D0);
  const double Dmax = jmax(fabs(D0), sumv) / alpha;
  nzval[c] = Dmax;
  b[i] += (Dmax - D0) * field[i];
}

// _setReference! (Solve_1_api.jl:431-439)
__global__ void k_set_reference(const int32_t* diag, double* nzval, double* b, double pref,
                                int64_t cell) {
  const int c = diag[cell];
  b[cell] = nzval[c] * pref;
  nzval[c] += nzval[c];
}

void discretise(xcal_eqn* e, const TermsK& t, const SourcesK& s, const xcal_field* prev, double dt) {
  xcal_mesh* m = e->mesh;
  e->cn = t.cn_first != 0;
  MeshK mk = meshk(m);
  double* nz = e->vector ? e->nzval0.p : e->nzval.p;
  const double* px = prev ? prev->comp(0) : nullptr;
  if (e->vector) {
    const double* py = prev ? prev->comp(1) : nullptr;
    const double* pz = prev ? prev->comp(2) : nullptr;
    hipLaunchKernelGGL(k_discretise<true>, dim3(grid_for(m->n)), dim3(kBlock), 0, m->ctx->stream,
                       mk, t, s, px, py, pz, dt, nz, e->bc(0), e->bc(1), e->bc(2));
  } else {
    hipLaunchKernelGGL(k_discretise<false>, dim3(grid_for(m->n)), dim3(kBlock), 0,
                       m->ctx->stream, mk, t, s, px, px, px, dt, nz, e->bc(0), nullptr, nullptr);
  }
  XCAL_HIP(hipGetLastError());
}

void apply_bcs(xcal_eqn* e, const TermsK& t, const BCsK& b, const xcal_field* phi, int comp) {
  xcal_mesh* m = e->mesh;
  if (m->nbcell == 0) return;
  const double* px = phi->comp(0);
  const double* py = phi->ncomp == 3 ? phi->comp(1) : px;
  const double* pz = phi->ncomp == 3 ? phi->comp(2) : px;
  double* bb = e->vector ? e->bc(comp) : e->bc(0);
  hipLaunchKernelGGL(k_apply_bcs, dim3(grid_for(m->nbcell)), dim3(kBlock), 0, m->ctx->stream,
                     meshk(m), t, b, e->vector ? comp : -1, px, py, pz, e->nzval.p, bb);
  XCAL_HIP(hipGetLastError());
}

void relax_diagdom(xcal_eqn* e, const double* field, double* b, double alpha) {
  xcal_mesh* m = e->mesh;
  hipLaunchKernelGGL(k_relax_diagdom, dim3(grid_for(m->n)), dim3(kBlock), 0, m->ctx->stream, m->n,
                     m->rbeg.p, m->rend.p, m->diag.p, e->nzval.p, b, field, alpha);
  XCAL_HIP(hipGetLastError());
}

void set_reference(xcal_eqn* e, double pref, int64_t cell) {
  xcal_mesh* m = e->mesh;
  hipLaunchKernelGGL(k_set_reference, dim3(1), dim3(1), 0, m->ctx->stream, m->diag.p, e->nzval.p,
                     e->bc(0), pref, cell);
  XCAL_HIP(hipGetLastError());
}

void copy_nz(xcal_mesh* m, double* dst, const double* src) {
  XCAL_HIP(hipMemcpyAsync(dst, src, sizeof(double) * m->nzlen, hipMemcpyDeviceToDevice,
                          m->ctx->stream));
}

static SourcesK make_sources(const xcal_eqn* e, const xcal_source* s, int ns) {
  XCAL_CHECK(ns >= 0 && ns <= kMaxSources, "number of sources must be 0..2");
  SourcesK k{};
  k.n = ns;
  for (int i = 0; i < ns; ++i) {
    XCAL_CHECK(s[i].field != nullptr, "source field is NULL");
    XCAL_CHECK(s[i].field->mesh == e->mesh, "source field belongs to another mesh");
    XCAL_CHECK(s[i].field->kind == (e->vector ? XCAL_CELL_VECTOR : XCAL_CELL_SCALAR),
               "source field kind does not match the equation");
    k.sign[i] = s[i].sign;
    k.x[i] = s[i].field->comp(0);
    k.y[i] = e->vector ? s[i].field->comp(1) : nullptr;
    k.z[i] = e->vector ? s[i].field->comp(2) : nullptr;
  }
  return k;
}

static void check_terms(const xcal_eqn* e, const xcal_term* t, int nt) {
  for (int i = 0; i < nt; ++i) {
    if (!t[i].flux) continue;
    XCAL_CHECK(t[i].flux->mesh == e->mesh, "term flux belongs to another mesh");
    bool cellwise = t[i].type == XCAL_TIME_STEADY || t[i].type == XCAL_TIME_EULER ||
                    t[i].type == XCAL_TIME_CRANK_NICOLSON || t[i].type == XCAL_SI;
    XCAL_CHECK(t[i].flux->kind == (cellwise ? XCAL_CELL_SCALAR : XCAL_FACE_SCALAR),
               "term flux must be a face scalar field (Laplacian/Divergence) or a cell scalar "
               "field (Time/Si)");
  }
}

}  // namespace xcal

using namespace xcal;

extern "C" {

int xcal_eqn_create(xcal_mesh* m, int32_t vector, xcal_eqn** out) {
  return guarded([&] {
    XCAL_CHECK(m && out, "NULL argument");
    XCAL_HIP(hipSetDevice(m->ctx->device));
    auto e = std::make_unique<xcal_eqn>();
    e->mesh = m;
    e->vector = vector != 0;
    e->nzval.alloc(m->nzlen);
    if (e->vector) e->nzval0.alloc(m->nzlen);
    e->b.alloc((e->vector ? 3 : 1) * m->n);
    XCAL_HIP(hipMemsetAsync(e->nzval.p, 0, e->nzval.bytes(), m->ctx->stream));
    if (e->vector) XCAL_HIP(hipMemsetAsync(e->nzval0.p, 0, e->nzval0.bytes(), m->ctx->stream));
    XCAL_HIP(hipMemsetAsync(e->b.p, 0, e->b.bytes(), m->ctx->stream));
    XCAL_HIP(hipStreamSynchronize(m->ctx->stream));
    *out = e.release();
    return 0;
  });
}

int xcal_eqn_destroy(xcal_eqn* e) {
  return guarded([&] {
    if (e) {
      (void)hipDeviceSynchronize();
      delete e;
    }
    return 0;
  });
}

int xcal_discretise(xcal_eqn* e, const xcal_term* terms, int32_t nterms, const xcal_source* src,
                    int32_t nsrc, const xcal_field* prev) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    check_terms(e, terms, nterms);
    TermsK t = make_terms(terms, nterms);
    SourcesK s = make_sources(e, src, nsrc);
    bool needs_prev = false;
    for (int i = 0; i < nterms; ++i)
      needs_prev |= terms[i].type == XCAL_TIME_EULER || terms[i].type == XCAL_TIME_CRANK_NICOLSON;
    if (prev) check_field(prev, e->mesh, e->vector ? XCAL_CELL_VECTOR : XCAL_CELL_SCALAR, "prev");
    XCAL_CHECK(!needs_prev || prev, "Time{Euler} needs the previous field");
    XCAL_CHECK(!needs_prev, "Time{Euler} needs runtime.dt: use xcal_discretise_dt");
    discretise(e, t, s, prev, 1.0);
    return 0;
  });
}

int xcal_discretise_dt(xcal_eqn* e, const xcal_term* terms, int32_t nterms, const xcal_source* src,
                       int32_t nsrc, const xcal_field* prev, double dt) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    XCAL_CHECK(dt > 0.0, "dt must be positive");
    check_terms(e, terms, nterms);
    TermsK t = make_terms(terms, nterms);
    SourcesK s = make_sources(e, src, nsrc);
    bool needs_prev = false;
    for (int i = 0; i < nterms; ++i)
      needs_prev |= terms[i].type == XCAL_TIME_EULER || terms[i].type == XCAL_TIME_CRANK_NICOLSON;
    if (prev) check_field(prev, e->mesh, e->vector ? XCAL_CELL_VECTOR : XCAL_CELL_SCALAR, "prev");
    XCAL_CHECK(!needs_prev || prev, "Time{Euler} needs the previous field");
    discretise(e, t, s, prev, dt);
    return 0;
  });
}

int xcal_update_equation(xcal_eqn* e) {
  return guarded([&] {
    XCAL_CHECK(e && e->vector, "update_equation! applies to vector equations");
    copy_nz(e->mesh, e->nzval.p, e->nzval0.p);
    return 0;
  });
}

int xcal_apply_bcs(xcal_eqn* e, const xcal_term* terms, int32_t nterms, const xcal_bc* bcs,
                   int32_t nbcs, const xcal_field* phi, int32_t comp) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    check_terms(e, terms, nterms);
    TermsK t = make_terms(terms, nterms);
    BCsK b = make_bcs(e->mesh, bcs, nbcs);
    check_field(phi, e->mesh, e->vector ? XCAL_CELL_VECTOR : XCAL_CELL_SCALAR, "phi");
    if (e->vector) {
      XCAL_CHECK(comp >= 0 && comp <= 2, "vector component must be 0..2");
      for (int i = 0; i < nbcs; ++i)
        XCAL_CHECK(bcs[i].kind != XCAL_BC_NEUMANN, "Neumann is not defined for vector fields");
    } else {
      XCAL_CHECK(comp == 0, "scalar equations take component 0");
    }
    apply_bcs(e, t, b, phi, comp);
    return 0;
  });
}

int xcal_relax_implicit_diagdom(xcal_eqn* e, const xcal_field* phi, int32_t comp, double alpha) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    check_field(phi, e->mesh, e->vector ? XCAL_CELL_VECTOR : XCAL_CELL_SCALAR, "phi");
    XCAL_CHECK(comp >= 0 && comp < (e->vector ? 3 : 1), "component out of range");
    relax_diagdom(e, phi->comp(comp), e->bc(e->vector ? comp : 0), alpha);
    return 0;
  });
}

int xcal_set_reference(xcal_eqn* e, double pref, int64_t cell) {
  return guarded([&] {
    XCAL_CHECK(e && !e->vector, "setReference! applies to scalar equations");
    XCAL_CHECK(cell >= 1 && cell <= e->mesh->n, "reference cell out of range");
    set_reference(e, pref, e->mesh->cperm[cell - 1]);
    return 0;
  });
}

int xcal_eqn_download(const xcal_eqn* e, int32_t which, int32_t comp, double* nzval, double* b) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    XCAL_CHECK(which == 0 || (which == 1 && e->vector), "which: 0 = A, 1 = A0 (vector only)");
    XCAL_CHECK(comp >= 0 && comp < (e->vector ? 3 : 1), "component out of range");
    const xcal_mesh* m = e->mesh;
    if (nzval) download_permuted(m, m->nzperm, which ? e->nzval0.p : e->nzval.p, nzval, m->nnz,
                                  m->nzlen);
    if (b) download_permuted(m, m->cperm, e->bc(comp), b, m->n);
    return 0;
  });
}

}  // extern "C"
