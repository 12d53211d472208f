// CPU restatement of XCALibre's FP64 SIMPLE hot path -- test infrastructure.
//
// This file is the parity oracle for libxcalhip.so and the timed CPU
// baseline.  It is not part of the product.  Every function restates the
// reference function cited beside it with the same loop order and the same
// floating-point evaluation order (compile with -ffp-contract=off).
//
// Parity status per part (see the tests that pin each):
//   mesh/CSR pattern ............ pinned (reference mesh-count fixtures,
//                                  closed form pattern = diag U neighbours)
//   discretise!/BCs ............. pinned (unit_test_laplace.jl golden A/b,
//                                  closed forms on uniform meshes)
//   interpolate!/green_gauss!/div! pinned (exact for linear fields, closed
//                                  cells)
//   DILU ........................ pinned (test_DILU.jl identities)
//   Krylov CG/BiCGSTAB .......... textbook + recalled Krylov.jl conventions;
//                                  pinned against dense LU solves only:
//                                  iterate sequences are parity unpinned
//   SIMPLE end-to-end ........... pinned by physical assertions
//                                  (2d_incompressible_laminar_BFS.jl:93 style)
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
int g_threads = 0;

// Julia max/min for Float64 (NaN-propagating, +0.0 > -0.0)
inline double jmax(double a, double b) {
  if (a != a || b != b) return std::numeric_limits<double>::quiet_NaN();
  if (a == b) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}
inline double jeps(double x) {
  double a = std::fabs(x);
  return std::nextafter(a, std::numeric_limits<double>::infinity()) - a;
}

struct V3 {
  double x, y, z;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

}  // namespace

struct orc_mesh {
  int dim;
  int64_t n, nf, nbf;
  std::vector<double> vol;
  std::vector<V3> cc;
  std::vector<int64_t> cf_first, cf_last;  // inclusive, 0-based into cell_* arrays
  std::vector<int64_t> cell_faces, cell_nbr, cell_nsign;
  std::vector<int64_t> own1, own2;
  std::vector<V3> fc, fn, fe;
  std::vector<double> area, delta, weight;
  std::vector<int64_t> bcell;
  std::vector<int64_t> pfirst, plast;  // inclusive 0-based face ids
  std::vector<int64_t> rowptr, colval;
  std::vector<int64_t> diag;  // spindex(i,i) precomputed only where the reference
                              // stores it (DILU's Di, diagonal_indices!)
};

namespace {

#define OMP_FOR _Pragma("omp parallel for schedule(static) if(n > 4096)")

// spindex (ModelFramework_0_types.jl:246-258); -1 when absent
inline int64_t spindex(const orc_mesh& m, int64_t i, int64_t j) {
  for (int64_t k = m.rowptr[i]; k < m.rowptr[i + 1]; ++k)
    if (m.colval[k] == j) return k;
  return -1;
}

inline double term_flux(const orc_term& t, int64_t i) {
  return t.flux ? t.flux[i] : t.flux_const;
}

// scheme! (Discretise_1_schemes.jl), per internal face visit
inline void scheme(const orc_mesh& m, const orc_term& t, int64_t fID, double ns, double& ac,
                   double& an) {
  const double sign = (double)t.sign;
  switch (t.type) {
    case XCAL_TIME_STEADY:
    case XCAL_TIME_EULER:
    case XCAL_TIME_CRANK_NICOLSON:
    case XCAL_SI:
      ac = 0.0;
      an = 0.0;
      return;
    case XCAL_LAPLACIAN_LINEAR: {  // :68-104
      const double area = m.area[fID];
      V3 nrm = m.fn[fID];
      V3 Sf{(ns * area) * nrm.x, (ns * area) * nrm.y, (ns * area) * nrm.z};
      V3 e{ns * m.fe[fID].x, ns * m.fe[fID].y, ns * m.fe[fID].z};
      double c = dot(Sf, Sf) / dot(Sf, e);
      V3 Ef{c * e.x, c * e.y, c * e.z};
      double Ef_mag = norm(Ef);
      double ap = (sign * (term_flux(t, fID) * Ef_mag)) / m.delta[fID];
      ac = -ap;
      an = ap;
      return;
    }
    case XCAL_DIV_LINEAR: {  // :114-129
      double w = m.weight[fID];
      w = 0.5 + ns * (w - 0.5);
      double ap = sign * (term_flux(t, fID) * ns);
      ac = ap * w;
      an = ap * (1.0 - w);
      return;
    }
    case XCAL_DIV_UPWIND: {  // :137-147
      double ap = sign * (term_flux(t, fID) * ns);
      ac = jmax(ap, 0.0);
      an = -jmax(-ap, 0.0);
      return;
    }
    case XCAL_DIV_LUST: {  // :155-175
      double w = m.weight[fID];
      if (std::signbit(ns)) w = 1.0 - w;
      double ap = sign * (term_flux(t, fID) * ns);
      double acL = ap * w, anL = ap * (1.0 - w);
      double acU = jmax(ap, 0.0), anU = -jmax(-ap, 0.0);
      ac = 0.75 * acL + 0.25 * acU;
      an = 0.75 * anL + 0.25 * anU;
      return;
    }
    case XCAL_DIV_BOUNDED_UPWIND: {  // :183-195
      double ap = sign * (term_flux(t, fID) * ns);
      ac = jmax(-ap, 0.0);
      an = -jmax(-ap, 0.0);
      return;
    }
  }
  throw std::runtime_error("unknown operator type");
}

// scheme_source! (Discretise_1_schemes.jl), per cell
inline void scheme_source(const orc_mesh& m, const orc_term& t, int64_t cID, const double* prev,
                          double dt, double& ac, double& b) {
  switch (t.type) {
    case XCAL_TIME_EULER:             // :30-42
    case XCAL_TIME_CRANK_NICOLSON: {  // :52-62, the same coefficients
      double vol_rdt = (term_flux(t, cID) * m.vol[cID]) / dt;
      ac = vol_rdt;
      b = prev[cID] * vol_rdt;
      return;
    }
    case XCAL_SI: {  // :208-218
      double flux = ((double)t.sign * term_flux(t, cID)) * m.vol[cID];
      ac = flux;
      b = 0.0;
      return;
    }
    default:
      ac = 0.0;
      b = 0.0;
      return;
  }
}

// BC functors (Discretise/boundary_conditions/*.jl); component -1 = scalar
void bc_coeffs(const orc_mesh& m, const xcal_bc& bc, const orc_term& t, int64_t fID,
               int64_t cID, int comp, const double* phix, const double* phiy,
               const double* phiz, double& ap_out, double& bp_out) {
  const double sign = (double)t.sign;
  const bool vec = comp >= 0;
  const double value = vec ? bc.value[comp] : bc.value[0];
  const double area = m.area[fID], delta = m.delta[fID];
  auto phi_c = [&]() -> double {
    return vec ? (comp == 0 ? phix : (comp == 1 ? phiy : phiz))[cID] : phix[cID];
  };
  ap_out = 0.0;
  bp_out = 0.0;
  if (t.type == XCAL_TIME_STEADY || t.type == XCAL_TIME_EULER ||
      t.type == XCAL_TIME_CRANK_NICOLSON || t.type == XCAL_SI)
    return;
  const bool lap = t.type == XCAL_LAPLACIAN_LINEAR;
  const bool divLUL = t.type == XCAL_DIV_LINEAR || t.type == XCAL_DIV_UPWIND ||
                      t.type == XCAL_DIV_LUST;
  const bool divBU = t.type == XCAL_DIV_BOUNDED_UPWIND;
  const double J = term_flux(t, fID);
  switch (bc.kind) {
    case XCAL_BC_DIRICHLET:  // dirichlet.jl:19-99
      if (lap) {
        double flux = (J * area) / delta;
        double ap = sign * (-flux);
        ap_out = ap;
        bp_out = ap * value;
      } else if (divLUL) {
        double flux = -J;
        double ap = sign * flux;
        bp_out = ap * value;
      } else if (divBU) {
        double ap = sign * J;
        ap_out = jmax(-ap, 0.0);
        double an = -jmax(-ap, 0.0);
        bp_out = -an * value;
      }
      return;
    case XCAL_BC_NEUMANN:  // neumann.jl:23-63 (scalar only)
      if (vec) throw std::runtime_error("Neumann is defined for scalar fields only");
      if (lap) {
        bp_out = (J * area) * value;
      } else if (divLUL || divBU) {
        double ap = sign * J;
        ap_out = ap;
        bp_out = ((-value) * ap) * delta;
      }
      return;
    case XCAL_BC_EMPTY:  // empty.jl:46-106: (0, 0) for every operator
      return;
    case XCAL_BC_ZEROGRADIENT:  // zerogradient.jl:25-76
    case XCAL_BC_KWALL:         // KWallFunction, Discretise_3_boundary_conditions.jl:11-95
    case XCAL_BC_NUTWALL:
      if (divLUL) ap_out = sign * J;
      else if (divBU) {
        double ap = sign * J;
        ap_out = ap - J;
      }
      return;
    case XCAL_BC_OMEGAWALL:       // OmegaWallFunction, same file :27-123
    case XCAL_BC_EXTRAPOLATED:  // extrapolated.jl:21-72
      if (lap) {
        double flux = ((-J) * area) / delta;
        double ap = sign * flux;
        ap_out = ap;
        bp_out = ap * phi_c();
      } else if (divLUL) {
        ap_out = sign * J;
      } else if (divBU) {
        double ap = sign * J;
        ap_out = ap - J;
      }
      return;
    case XCAL_BC_WALL:  // wall.jl:26-120
      if (vec) {
        if (lap) {
          V3 n = m.fn[fID];
          double flux = (J * area) / delta;
          double ap = sign * (-flux);
          V3 vb{bc.value[0], bc.value[1], bc.value[2]};
          V3 vc{phix[cID], phiy[cID], phiz[cID]};
          double vcn = dot(vc, n), vbn = dot(vb, n);
          V3 vc_n{vcn * n.x, vcn * n.y, vcn * n.z};
          V3 vb_n{vbn * n.x, vbn * n.y, vbn * n.z};
          V3 vb_p{vb.x - vb_n.x, vb.y - vb_n.y, vb.z - vb_n.z};
          ap_out = ap;
          bp_out = ap * (vb_p[comp] + vc_n[comp]);
        }
        // Divergence (all schemes) on vectors: 0, 0
      } else {
        if (divLUL) {
          ap_out = sign * J;
        } else if (divBU) {
          double ap = sign * J;
          ap_out = jmax(-ap, 0.0);
          double an = -jmax(-ap, 0.0);
          bp_out = -an * value;
        }
      }
      return;
    case XCAL_BC_SYMMETRY:  // symmetry.jl:23-151
      if (vec) {
        V3 n = m.fn[fID];
        V3 vc{phix[cID], phiy[cID], phiz[cID]};
        double vcn = dot(vc, n);
        V3 vn{vcn * n.x, vcn * n.y, vcn * n.z};
        if (lap) {
          double flux = (J * area) / delta;
          double ap = sign * (-flux);
          V3 vp{vc.x - vn.x, vc.y - vn.y, vc.z - vn.z};
          ap_out = ap;
          bp_out = ap * vp[comp];
        } else if (divLUL || divBU) {
          double ap = sign * J;
          ap_out = ap;
          bp_out = ap * vn[comp];
        }
      } else {
        if (divLUL || divBU) ap_out = sign * J;
      }
      return;
  }
  throw std::runtime_error("unknown boundary condition kind");
}

const xcal_bc* find_bc(const orc_mesh& m, const xcal_bc* bcs, int nbc, int64_t fID) {
  for (int k = 0; k < nbc; ++k) {
    int p = bcs[k].patch;
    if (p < 0 || p >= (int)m.pfirst.size()) throw std::runtime_error("bad patch index in BC");
    if (m.pfirst[p] <= fID && fID <= m.plast[p]) return &bcs[k];
  }
  return nullptr;
}

template <class Fn>
int guard(Fn fn) {
  try {
    return fn();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// ---------------------------------------------------------------- Krylov
struct Op {
  const orc_mesh& m;
  const double* nzval;
  int precon;
  std::vector<double> S;  // Jacobi 1/|a_ii|
  std::vector<double> D;  // DILU diagonal
  void apply_A(const double* x, double* y) const { orc_spmv(&m, nzval, x, y); }
  // mulorldiv!(z, M, r, ldiv)
  void apply_M(const double* r, double* z) const {
    const int64_t n = m.n;
    if (precon == XCAL_PRECON_JACOBI) {
      for (int64_t i = 0; i < n; ++i) z[i] = S[i] * r[i];
    } else if (precon == XCAL_PRECON_DILU) {
      orc_dilu_apply(&m, nzval, D.data(), r, z);
    } else {
      std::memcpy(z, r, sizeof(double) * n);
    }
  }
};

double kdot(int64_t n, const double* a, const double* b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
double knorm(int64_t n, const double* a) { return std::sqrt(kdot(n, a, a)); }

// Krylov.jl cg! with warm start (restated; Saad Alg. 9.1 with M applied as
// mulorldiv!, stopping at rNorm <= atol + rtol*rNorm0, rNorm = sqrt(r'z))
int krylov_cg(const Op& op, const double* b, double* x0, int itmax, double atol, double rtol,
              int& niter) {
  const int64_t n = op.m.n;
  std::vector<double> r(n), z(n), p(n), Ap(n), x(n, 0.0);
  op.apply_A(x0, r.data());
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
  op.apply_M(r.data(), z.data());
  p = z;
  double gamma = kdot(n, r.data(), z.data());
  double rNorm = std::sqrt(gamma);
  niter = 0;
  if (gamma == 0.0) return 0;  // x = x0
  if (itmax == 0) itmax = (int)(2 * n);
  double pNorm2 = gamma;
  const double eps_ = std::numeric_limits<double>::epsilon();
  const double epsilon = atol + rtol * rNorm;
  bool solved = rNorm <= epsilon, tired = niter >= itmax, zero_curv = false;
  while (!(solved || tired || zero_curv)) {
    op.apply_A(p.data(), Ap.data());
    double pAp = kdot(n, p.data(), Ap.data());
    if (pAp <= eps_ * pNorm2 && std::fabs(pAp) <= eps_ * pNorm2) {
      zero_curv = true;
      continue;
    }
    double alpha = gamma / pAp;
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
    for (int64_t i = 0; i < n; ++i) r[i] = r[i] + (-alpha) * Ap[i];
    op.apply_M(r.data(), z.data());
    double gamma_next = kdot(n, r.data(), z.data());
    rNorm = std::sqrt(gamma_next);
    bool mach = rNorm + 1.0 <= 1.0;
    solved = rNorm <= epsilon || mach;
    if (!solved) {
      double beta = gamma_next / gamma;
      pNorm2 = gamma_next + beta * beta * pNorm2;
      gamma = gamma_next;
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    ++niter;
    tired = niter >= itmax;
  }
  for (int64_t i = 0; i < n; ++i) x0[i] = x[i] + x0[i];
  return niter >= itmax ? 1 : 0;
}

// Krylov.jl bicgstab! with warm start, left preconditioner M, shadow c = b
int krylov_bicgstab(const Op& op, const double* b, double* x0, int itmax, double atol,
                    double rtol, int& niter) {
  const int64_t n = op.m.n;
  std::vector<double> r0(n), r(n), p(n), v(n, 0.0), s(n, 0.0), q(n), t(n), d(n), x(n, 0.0);
  op.apply_A(x0, r0.data());
  for (int64_t i = 0; i < n; ++i) r0[i] = b[i] - r0[i];
  op.apply_M(r0.data(), r.data());
  p = r;
  double rNorm = knorm(n, r.data());
  niter = 0;
  if (rNorm == 0.0) return 0;
  if (itmax == 0) itmax = (int)(2 * n);
  const double epsilon = atol + rtol * rNorm;
  const double* c = b;
  double next_rho = kdot(n, c, r.data());
  if (next_rho == 0.0) return 0;  // breakdown <c, r0> = 0: x = x0
  bool solved = rNorm <= epsilon, tired = niter >= itmax, breakdown = false;
  while (!(solved || tired || breakdown)) {
    ++niter;
    double rho = next_rho;
    op.apply_A(p.data(), q.data());        // y = p (N = I)
    op.apply_M(q.data(), v.data());
    double alpha = rho / kdot(n, c, v.data());
    for (int64_t i = 0; i < n; ++i) s[i] = r[i] + (-alpha) * v[i];
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
    op.apply_A(s.data(), d.data());        // z = s
    op.apply_M(d.data(), t.data());
    double omega = kdot(n, t.data(), s.data()) / kdot(n, t.data(), t.data());
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + omega * s[i];
    for (int64_t i = 0; i < n; ++i) r[i] = s[i] + (-omega) * t[i];
    next_rho = kdot(n, c, r.data());
    double beta = (next_rho / rho) * (alpha / omega);
    for (int64_t i = 0; i < n; ++i) p[i] = p[i] + (-omega) * v[i];
    for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    rNorm = knorm(n, r.data());
    bool mach = rNorm + 1.0 <= 1.0;
    solved = rNorm <= epsilon || mach;
    tired = niter >= itmax;
    breakdown = (alpha == 0.0 || alpha != alpha);
  }
  for (int64_t i = 0; i < n; ++i) x0[i] = x[i] + x0[i];
  return niter >= itmax ? 1 : 0;
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_set_threads(int32_t n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_mesh_create(const xcal_mesh_host* h, orc_mesh** out) {
  return guard([&] {
    auto* m = new orc_mesh();
    m->dim = h->dim;
    m->n = h->n_cells;
    m->nf = h->n_faces;
    m->nbf = h->n_bfaces;
    const int64_t n = m->n, nf = m->nf;
    m->vol.assign(h->cell_volume, h->cell_volume + n);
    m->cc.resize(n);
    for (int64_t i = 0; i < n; ++i)
      m->cc[i] = V3{h->cell_centre[3 * i], h->cell_centre[3 * i + 1], h->cell_centre[3 * i + 2]};
    m->cf_first.resize(n);
    m->cf_last.resize(n);
    for (int64_t i = 0; i < n; ++i) {
      m->cf_first[i] = h->cell_faces_range[2 * i] - 1;
      m->cf_last[i] = h->cell_faces_range[2 * i + 1] - 1;
    }
    m->cell_faces.resize(h->n_cell_faces);
    m->cell_nbr.resize(h->n_cell_faces);
    m->cell_nsign.resize(h->n_cell_faces);
    for (int64_t k = 0; k < h->n_cell_faces; ++k) {
      m->cell_faces[k] = h->cell_faces[k] - 1;
      m->cell_nbr[k] = h->cell_neighbours[k] - 1;
      m->cell_nsign[k] = h->cell_nsign[k];
    }
    m->own1.resize(nf);
    m->own2.resize(nf);
    m->fc.resize(nf);
    m->fn.resize(nf);
    m->fe.resize(nf);
    for (int64_t f = 0; f < nf; ++f) {
      m->own1[f] = h->face_owner[2 * f] - 1;
      m->own2[f] = h->face_owner[2 * f + 1] - 1;
      m->fc[f] = V3{h->face_centre[3 * f], h->face_centre[3 * f + 1], h->face_centre[3 * f + 2]};
      m->fn[f] = V3{h->face_normal[3 * f], h->face_normal[3 * f + 1], h->face_normal[3 * f + 2]};
      m->fe[f] = V3{h->face_e[3 * f], h->face_e[3 * f + 1], h->face_e[3 * f + 2]};
    }
    m->area.assign(h->face_area, h->face_area + nf);
    m->delta.assign(h->face_delta, h->face_delta + nf);
    m->weight.assign(h->face_weight, h->face_weight + nf);
    m->bcell.resize(m->nbf);
    for (int64_t f = 0; f < m->nbf; ++f) m->bcell[f] = h->boundary_cellsID[f] - 1;
    for (int p = 0; p < h->n_patches; ++p) {
      m->pfirst.push_back(h->patch_range[2 * p] - 1);
      m->plast.push_back(h->patch_range[2 * p + 1] - 1);
    }
    // sparse_matrix_connectivity (ModelFramework_0_types.jl:223-242) then
    // sparsecsr(i, j, v, n, n): columns ascending, duplicates combined
    m->rowptr.assign(n + 1, 0);
    std::vector<std::vector<int64_t>> rows(n);
    for (int64_t c = 0; c < n; ++c) {
      rows[c].push_back(c);
      for (int64_t k = m->cf_first[c]; k <= m->cf_last[c]; ++k) rows[c].push_back(m->cell_nbr[k]);
      std::sort(rows[c].begin(), rows[c].end());
      rows[c].erase(std::unique(rows[c].begin(), rows[c].end()), rows[c].end());
      m->rowptr[c + 1] = m->rowptr[c] + (int64_t)rows[c].size();
    }
    m->colval.reserve(m->rowptr[n]);
    for (auto& r : rows) m->colval.insert(m->colval.end(), r.begin(), r.end());
    // diagonal_indices! (preconditioners_1_DILU.jl:34-41)
    m->diag.resize(n);
    for (int64_t i = 0; i < n; ++i) m->diag[i] = spindex(*m, i, i);
    *out = m;
    return 0;
  });
}

void orc_mesh_destroy(orc_mesh* m) { delete m; }
int64_t orc_nnz(const orc_mesh* m) { return m->rowptr[m->n]; }
void orc_csr(const orc_mesh* m, int64_t* rowptr, int64_t* colval) {
  std::copy(m->rowptr.begin(), m->rowptr.end(), rowptr);
  std::copy(m->colval.begin(), m->colval.end(), colval);
}

int orc_discretise_scalar(const orc_mesh* mp, const orc_term* t, int32_t nt,
                          const orc_source* s, int32_t ns_, const double* prev, double dt,
                          double* nzval, double* b) {
  return guard([&] {
    const orc_mesh& m = *mp;
    const int64_t n = m.n;
    const int64_t nnz = m.rowptr[n];
    for (int64_t i = 0; i < nnz; ++i) nzval[i] = 0.0;
    // _discretise_scalar_model! (Discretise_2_generated_distretisation.jl:125-167)
    OMP_FOR
    for (int64_t i = 0; i < n; ++i) {
      int64_t cIndex = spindex(m, i, i);
      double ac_sum = 0.0;
      for (int64_t fi = m.cf_first[i]; fi <= m.cf_last[i]; ++fi) {
        int64_t fID = m.cell_faces[fi];
        double nsg = (double)m.cell_nsign[fi];
        int64_t nID = m.cell_nbr[fi];
        int64_t nIndex = spindex(m, i, nID);
        double AC = 0.0, AN = 0.0;  // _scheme! :173-197
        for (int k = 0; k < nt; ++k) {
          double ac, an;
          scheme(m, t[k], fID, nsg, ac, an);
          AC += ac;
          AN += an;
        }
        ac_sum += AC;
        nzval[nIndex] = AN;
      }
      double AC = 0.0, B = 0.0;  // _scheme_source! :200-230
      for (int k = 0; k < nt; ++k) {
        double ac, bb;
        scheme_source(m, t[k], i, prev, dt, ac, bb);
        AC += ac;
        B += bb;
      }
      nzval[cIndex] = ac_sum + AC;
      double B2 = 0.0;  // _sources! :255-270
      for (int k = 0; k < ns_; ++k) B2 += ((double)s[k].sign * s[k].x[i]) * m.vol[i];
      b[i] = B2 + B;
    }
    return 0;
  });
}

int orc_discretise_vector(const orc_mesh* mp, const orc_term* t, int32_t nt,
                          const orc_source* s, int32_t ns_, const double* px,
                          const double* py, const double* pz, double dt, double* nzval0,
                          double* bx, double* by, double* bz) {
  return guard([&] {
    const orc_mesh& m = *mp;
    const int64_t n = m.n;
    const int64_t nnz = m.rowptr[n];
    for (int64_t i = 0; i < nnz; ++i) nzval0[i] = 0.0;
    // _discretise_vector_model! (Discretise_2_generated_distretisation.jl:39-87)
    OMP_FOR
    for (int64_t i = 0; i < n; ++i) {
      int64_t cIndex = spindex(m, i, i);
      double ac_sum = 0.0;
      for (int64_t fi = m.cf_first[i]; fi <= m.cf_last[i]; ++fi) {
        int64_t fID = m.cell_faces[fi];
        double nsg = (double)m.cell_nsign[fi];
        int64_t nID = m.cell_nbr[fi];
        int64_t nIndex = spindex(m, i, nID);
        double AC = 0.0, AN = 0.0;
        for (int k = 0; k < nt; ++k) {
          double ac, an;
          scheme(m, t[k], fID, nsg, ac, an);
          AC += ac;
          AN += an;
        }
        ac_sum += AC;
        nzval0[nIndex] = AN;
      }
      double AC = 0.0, BX = 0.0, BY = 0.0, BZ = 0.0;
      for (int k = 0; k < nt; ++k) {
        double ac, b1, b2, b3;
        scheme_source(m, t[k], i, px, dt, ac, b1);
        scheme_source(m, t[k], i, py, dt, ac, b2);
        scheme_source(m, t[k], i, pz, dt, ac, b3);
        AC += ac;
        BX += b1;
        BY += b2;
        BZ += b3;
      }
      nzval0[cIndex] = ac_sum + AC;
      double Bx = 0.0, By = 0.0, Bz = 0.0;  // _sources! :271-293
      for (int k = 0; k < ns_; ++k) {
        double sg = (double)s[k].sign;
        Bx += (sg * s[k].x[i]) * m.vol[i];
        By += (sg * s[k].y[i]) * m.vol[i];
        Bz += (sg * s[k].z[i]) * m.vol[i];
      }
      bx[i] = BX + Bx;
      by[i] = BY + By;
      bz[i] = BZ + Bz;
    }
    return 0;
  });
}

int orc_apply_bcs(const orc_mesh* mp, const orc_term* t, int32_t nt, const xcal_bc* bcs,
                  int32_t nbc, int32_t comp, const double* phix, const double* phiy,
                  const double* phiz, double* nzval, double* b) {
  return guard([&] {
    const orc_mesh& m = *mp;
    // one work-item per boundary face, unrolled patch `if` chain; the
    // atomic += of :114-115 performed in ascending face order
    for (int64_t fID = 0; fID < m.nbf; ++fID) {
      const xcal_bc* bc = find_bc(m, bcs, nbc, fID);
      if (!bc) continue;
      int64_t cellID = m.bcell[fID];
      int64_t z = spindex(m, cellID, cellID);
      double AP = 0.0, BP = 0.0;  // apply! :132-157
      for (int k = 0; k < nt; ++k) {
        double ap, bp;
        bc_coeffs(m, *bc, t[k], fID, cellID, comp, phix, phiy, phiz, ap, bp);
        AP += ap;
        BP += bp;
      }
      nzval[z] += AP;
      b[cellID] += BP;
    }
    return 0;
  });
}

void orc_relax_implicit_diagdom(const orc_mesh* mp, double* nzval, double* b,
                                const double* field, double alpha) {
  const orc_mesh& m = *mp;
  const int64_t n = m.n;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double sumv = 0.0;
    int64_t c = spindex(m, i, i);
    for (int64_t k = m.rowptr[i]; k < m.rowptr[i + 1]; ++k) sumv += std::fabs(nzval[k]);
    sumv -= std::fabs(nzval[c]);
    double D0 = nzval[c];
    double Dmax = jmax(std::fabs(D0), sumv) / alpha;
    nzval[c] = Dmax;
    b[i] += (Dmax - D0) * field[i];
  }
}

void orc_set_reference(const orc_mesh* m, double* nzval, double* b, double pref, int64_t cell) {
  int64_t c = spindex(*m, cell, cell);
  b[cell] = nzval[c] * pref;
  nzval[c] += nzval[c];
}

void orc_spmv(const orc_mesh* mp, const double* nzval, const double* x, double* y) {
  const orc_mesh& m = *mp;
  const int64_t n = m.n;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double accu = 0.0;
    for (int64_t k = m.rowptr[i]; k < m.rowptr[i + 1]; ++k) accu += nzval[k] * x[m.colval[k]];
    y[i] = accu;
  }
}

double orc_residual(const orc_mesh* mp, const double* nzval, const double* b, const double* x) {
  const int64_t n = mp->n;
  std::vector<double> Fx(n);
  orc_spmv(mp, nzval, x, Fx.data());
  double sumR = 0.0, nb2 = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double d = b[i] - Fx[i];
    sumR += d * d;
    nb2 += b[i] * b[i];
  }
  double normb = std::sqrt(nb2);
  double den = normb > jeps(normb) ? normb : 1.0;
  return std::sqrt(sumR) / den;
}

void orc_dilu_diagonal(const orc_mesh* mp, const double* nzval, double* D) {
  const orc_mesh& m = *mp;
  for (int64_t i = 0; i < m.n; ++i) D[i] = nzval[m.diag[i]];  // extract_diagonal!
  for (int64_t i = 0; i < m.n; ++i)                          // update_dilu_diagonal! :43-75
    for (int64_t k = m.rowptr[i]; k < m.rowptr[i + 1]; ++k) {
      int64_t j = m.colval[k];
      if (j > i) {
        int64_t k2 = spindex(m, j, i);
        if (k2 >= 0) D[j] = D[j] - (nzval[k] * nzval[k2]) / D[i];
      }
    }
}

void orc_dilu_apply(const orc_mesh* mp, const double* nzval, const double* D, const double* r,
                    double* z) {
  const orc_mesh& m = *mp;
  const int64_t n = m.n;
  for (int64_t i = 0; i < n; ++i) {  // forward_substitution! :84-98
    double sum = 0.0;
    for (int64_t k = m.rowptr[i]; k < m.diag[i]; ++k) {
      int64_t j = m.colval[k];
      sum += (nzval[k] * z[j]) * (1.0 / D[j]);
    }
    z[i] = r[i] - sum;
  }
  for (int64_t i = n - 1; i >= 0; --i) {  // backward_substitution! :100-114
    double sum = 0.0;
    for (int64_t k = m.diag[i] + 1; k < m.rowptr[i + 1]; ++k) sum += nzval[k] * z[m.colval[k]];
    z[i] = (z[i] - sum) / D[i];
  }
}

int orc_krylov(const orc_mesh* mp, int32_t solver, int32_t precon, const double* nzval,
               const double* b, double* x, int32_t itmax, double atol, double rtol,
               int32_t* niter) {
  return guard([&] {
    Op op{*mp, nzval, precon, {}, {}};
    const int64_t n = mp->n;
    if (precon == XCAL_PRECON_JACOBI) {  // update_Jacobi! :86-94
      op.S.resize(n);
      for (int64_t i = 0; i < n; ++i) op.S[i] = 1.0 / std::fabs(nzval[mp->diag[i]]);
    } else if (precon == XCAL_PRECON_DILU) {
      op.D.resize(n);
      orc_dilu_diagonal(mp, nzval, op.D.data());
    }
    int it = 0, rc;
    if (solver == XCAL_SOLVER_CG) rc = krylov_cg(op, b, x, itmax, atol, rtol, it);
    else if (solver == XCAL_SOLVER_BICGSTAB) rc = krylov_bicgstab(op, b, x, itmax, atol, rtol, it);
    else throw std::runtime_error("unknown solver");
    *niter = it;
    return rc;
  });
}

void orc_smooth_jacobi(const orc_mesh* mp, const double* nzval, const double* b, double* x,
                       int32_t loops, double omega) {
  const orc_mesh& m = *mp;
  const int64_t n = m.n;
  std::vector<double> xt(n);  // smoother.x_temp
  for (int32_t l = 0; l < loops; ++l) {
    OMP_FOR
    for (int64_t i = 0; i < n; ++i) {  // multiply_row (Smoothers_jacobi.jl:97-110)
      double sum = 0.0;
      for (int64_t k = m.rowptr[i]; k < m.rowptr[i + 1]; ++k) sum += nzval[k] * x[m.colval[k]];
      const double a_ii = nzval[m.diag[i]];
      const double xi = x[i];
      sum -= a_ii * xi;
      const double rD = 1.0 / a_ii;
      xt[i] = (omega * rD) * (b[i] - sum) + (1.0 - omega) * xi;
    }
    std::copy(xt.begin(), xt.end(), x);
  }
}

int orc_solve_system(const orc_mesh* mp, const xcal_solver_setup* ss, int32_t cn,
                     const double* nzval, const double* b, double* x, int32_t* niter) {
  if (ss->smoother_loops < 0) {
    g_err = "smoother loops must be >= 0";
    return -1;
  }
  if (ss->smoother_loops > 0)
    orc_smooth_jacobi(mp, nzval, b, x, ss->smoother_loops, ss->smoother_omega);
  std::vector<double> xold;
  if (cn) xold.assign(x, x + mp->n);
  const int rc = orc_krylov(mp, ss->solver, ss->precon, nzval, b, x, ss->itmax, ss->atol,
                            ss->rtol, niter);
  if (rc < 0) return rc;
  if (cn)
    for (int64_t i = 0; i < mp->n; ++i) x[i] = 2.0 * x[i] - xold[i];
  return rc;
}

void orc_interpolate(const orc_mesh* mp, const double* phi, double* phif) {
  const orc_mesh& m = *mp;
  const int64_t n = m.nf;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double w = m.weight[i];
    double omw = 1.0 - w;
    phif[i] = w * phi[m.own1[i]] + omw * phi[m.own2[i]];
  }
}

int orc_correct_boundaries(const orc_mesh* mp, const xcal_bc* bcs, int32_t nbc,
                           const double* phix, const double* phiy, const double* phiz,
                           double* fx, double* fy, double* fz) {
  return guard([&] {
    const orc_mesh& m = *mp;
    const bool vec = fy != nullptr;
    for (int64_t fID = 0; fID < m.nbf; ++fID) {
      for (int k = 0; k < nbc; ++k) {
        const xcal_bc& bc = bcs[k];
        if (!(m.pfirst[bc.patch] <= fID && fID <= m.plast[bc.patch])) continue;
        int64_t c = m.bcell[fID];
        switch (bc.kind) {
          case XCAL_BC_DIRICHLET:  // dirichlet_interpolation.jl
            fx[fID] = bc.value[0];
            if (vec) {
              fy[fID] = bc.value[1];
              fz[fID] = bc.value[2];
            }
            break;
          case XCAL_BC_NEUMANN:  // neumann_interpolation.jl
            if (vec) throw std::runtime_error(
                "Neumann boundary condition for vector fields is not implemented yet.");
            fx[fID] = phix[c] + m.delta[fID] * bc.value[0];
            break;
          case XCAL_BC_WALL:  // wall_interpolation.jl
            if (vec) {
              fx[fID] = bc.value[0];
              fy[fID] = bc.value[1];
              fz[fID] = bc.value[2];
            } else {
              fx[fID] = phix[c];
            }
            break;
          case XCAL_BC_ZEROGRADIENT:
          case XCAL_BC_KWALL:
          case XCAL_BC_OMEGAWALL:
          case XCAL_BC_NUTWALL:
          case XCAL_BC_EXTRAPOLATED:  // zerogradient_interpolation.jl:1
            fx[fID] = phix[c];
            if (vec) {
              fy[fID] = phiy[c];
              fz[fID] = phiz[c];
            }
            break;
          case XCAL_BC_EMPTY:  // empty_interpolation.jl:1-16
            fx[fID] = 0.0;
            if (vec) {
              fy[fID] = 0.0;
              fz[fID] = 0.0;
            }
            break;
          case XCAL_BC_SYMMETRY:  // symmetry_interpolation.jl
            if (vec) {
              V3 n = m.fn[fID];
              V3 pc{phix[c], phiy[c], phiz[c]};
              double pn = dot(pc, n);
              fx[fID] = pc.x - pn * n.x;
              fy[fID] = pc.y - pn * n.y;
              fz[fID] = pc.z - pn * n.z;
            } else {
              fx[fID] = phix[c];
            }
            break;
          default:
            throw std::runtime_error("unknown boundary condition kind");
        }
      }
    }
    return 0;
  });
}

void orc_green_gauss(const orc_mesh* mp, const double* phif, double* gx, double* gy,
                     double* gz) {
  const orc_mesh& m = *mp;
  const int64_t n = m.n;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {  // _green_gauss! :27-54
    double rx = 0.0, ry = 0.0, rz = 0.0;
    for (int64_t fi = m.cf_first[i]; fi <= m.cf_last[i]; ++fi) {
      int64_t f = m.cell_faces[fi];
      double ns = (double)m.cell_nsign[fi];
      double a = m.area[f];
      V3 nr = m.fn[f];
      double v = phif[f];
      rx += v * ((a * nr.x) * ns);
      ry += v * ((a * nr.y) * ns);
      rz += v * ((a * nr.z) * ns);
    }
    double V = m.vol[i];
    gx[i] = rx / V;
    gy[i] = ry / V;
    gz[i] = rz / V;
  }
  for (int64_t f = 0; f < m.nbf; ++f) {  // boundary_faces_contribution! :58-78
    int64_t c = m.own1[f];
    double a = m.area[f], V = m.vol[c], v = phif[f];
    V3 nr = m.fn[f];
    gx[c] += (v * (a * nr.x)) / V;
    gy[c] += (v * (a * nr.y)) / V;
    gz[c] += (v * (a * nr.z)) / V;
  }
}

void orc_green_gauss_tensor(const orc_mesh* mp, const double* fx, const double* fy,
                            const double* fz, double* const* T) {
  const orc_mesh& m = *mp;
  const int64_t n = m.n;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {  // _green_gauss_vector! :106-151
    double r[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t fi = m.cf_first[i]; fi <= m.cf_last[i]; ++fi) {
      int64_t f = m.cell_faces[fi];
      double ns = (double)m.cell_nsign[fi];
      double a = m.area[f];
      V3 nr = m.fn[f];
      double S[3] = {(a * nr.x) * ns, (a * nr.y) * ns, (a * nr.z) * ns};
      double P[3] = {fx[f], fy[f], fz[f]};
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) r[3 * p + q] += P[p] * S[q];
    }
    double V = m.vol[i];
    for (int k = 0; k < 9; ++k) T[k][i] = r[k] / V;
  }
  for (int64_t f = 0; f < m.nbf; ++f) {  // boundary_faces_contribution_vector! :153-173
    int64_t c = m.own1[f];
    double a = m.area[f], V = m.vol[c];
    V3 nr = m.fn[f];
    double S[3] = {a * nr.x, a * nr.y, a * nr.z};
    double P[3] = {fx[f], fy[f], fz[f]};
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) T[3 * p + q][c] += (P[p] * S[q]) / V;
  }
}

int orc_grad_midpoint(const orc_mesh* mp, const xcal_bc* bcs, int32_t nbc, int32_t ncomp,
                      const double* const* phi, double* const* phif, double* const* grad) {
  return guard([&] {
    const orc_mesh& m = *mp;
    if (ncomp != 1 && ncomp != 3) throw std::runtime_error("grad!: scalar or vector field");
    for (int c = 0; c < ncomp; ++c)  // interpolate_midpoint! (:121-193)
      for (int64_t i = 0; i < m.nf; ++i) {
        const double p1 = phi[c][m.own1[i]];
        phif[c][i] = 0.5 * (p1 + phi[c][m.own2[i]]);
      }
    int rc = ncomp == 1 ? orc_correct_boundaries(mp, bcs, nbc, phi[0], nullptr, nullptr, phif[0],
                                                 nullptr, nullptr)
                        : orc_correct_boundaries(mp, bcs, nbc, phi[0], phi[1], phi[2], phif[0],
                                                 phif[1], phif[2]);
    if (rc < 0) throw std::runtime_error(g_err);
    auto gauss = [&] {
      if (ncomp == 1) orc_green_gauss(mp, phif[0], grad[0], grad[1], grad[2]);
      else orc_green_gauss_tensor(mp, phif[0], phif[1], phif[2], grad);
    };
    gauss();
    for (int it = 0; it < 2; ++it) {  // correct_interpolation! (:195-249) + green_gauss!
      for (int64_t i = m.nbf; i < m.nf; ++i) {
        const int64_t c1 = m.own1[i], c2 = m.own2[i];
        const V3 rf = m.fc[i], rP = m.cc[c1], rN = m.cc[c2];
        const double R[3] = {rf.x - 0.5 * (rP.x + rN.x), rf.y - 0.5 * (rP.y + rN.y),
                             rf.z - 0.5 * (rP.z + rN.z)};
        for (int p = 0; p < ncomp; ++p) {
          const double phip = 0.5 * (phi[p][c1] + phi[p][c2]);
          double g[3];
          for (int q = 0; q < 3; ++q) g[q] = 0.5 * (grad[3 * p + q][c1] + grad[3 * p + q][c2]);
          phif[p][i] = phip + ((g[0] * R[0] + g[1] * R[1]) + g[2] * R[2]);
        }
      }
      gauss();
    }
    return 0;
  });
}

void orc_div(const orc_mesh* mp, const double* psif, double* div) {
  const orc_mesh& m = *mp;
  const int64_t n = m.n;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {  // div_noS_kernel! :129-154
    double red = 0.0;
    for (int64_t fi = m.cf_first[i]; fi <= m.cf_last[i]; ++fi)
      red += psif[m.cell_faces[fi]] * (double)m.cell_nsign[fi];
    div[i] = red / m.vol[i];
  }
  for (int64_t f = 0; f < m.nbf; ++f) {  // :158-171
    int64_t c = m.own1[f];
    div[c] += psif[f] / m.vol[c];
  }
}

void orc_flux(const orc_mesh* mp, const double* fx, const double* fy, const double* fz,
              double* mdot) {
  const orc_mesh& m = *mp;
  const int64_t n = m.nf;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double a = m.area[i];
    V3 nr = m.fn[i];
    V3 Sf{a * nr.x, a * nr.y, a * nr.z};
    mdot[i] = fx[i] * Sf.x + fy[i] * Sf.y + fz[i] * Sf.z;
  }
}

void orc_H(const orc_mesh* mp, const double* nzval, const double* bx, const double* by,
           const double* bz, const double* Ux, const double* Uy, const double* Uz, double* Hx,
           double* Hy, double* Hz) {
  const orc_mesh& m = *mp;
  const int64_t n = m.n;
  OMP_FOR
  for (int64_t i = 0; i < n; ++i) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int64_t k = m.rowptr[i]; k < m.rowptr[i + 1]; ++k) {
      int64_t j = m.colval[k];
      double v = nzval[k];
      sx += v * Ux[j];
      sy += v * Uy[j];
      sz += v * Uz[j];
    }
    double D = nzval[spindex(m, i, i)];
    sx -= D * Ux[i];
    sy -= D * Uy[i];
    sz -= D * Uz[i];
    double rD = 1.0 / D;
    Hx[i] = (bx[i] - sx) * rD;
    Hy[i] = (by[i] - sy) * rD;
    Hz[i] = (bz[i] - sz) * rD;
  }
}

int orc_correct_mass_flux(const orc_mesh* mp, const double* pnz, const double* p,
                          const double* pflux, int32_t psign, const xcal_bc* bcs, int32_t nbc,
                          double* mdot) {
  return guard([&] {
    const orc_mesh& m = *mp;
    const int64_t n = m.nf;
    OMP_FOR
    for (int64_t f = m.nbf; f < n; ++f) {  // _correct_mass_flux! :371-388
      int64_t c1 = m.own1[f], c2 = m.own2[f];
      double aN = pnz[spindex(m, c1, c2)];
      mdot[f] += aN * (p[c2] - p[c1]);
    }
    for (int64_t f = 0; f < m.nbf; ++f) {  // _correct_boundary_mass_flux! :490-557
      const xcal_bc* bc = find_bc(m, bcs, nbc, f);
      if (!bc) continue;
      int64_t c = m.bcell[f];
      if (bc->kind == XCAL_BC_DIRICHLET) {
        double flux = (pflux[f] * m.area[f]) / m.delta[f];
        double ap = (double)psign * (-flux);
        mdot[f] += ap * (p[c] - bc->value[0]);
      } else if (bc->kind == XCAL_BC_NEUMANN) {
        mdot[f] -= (pflux[f] * m.area[f]) * bc->value[0];
      }
    }
    return 0;
  });
}

int orc_nonorth_correction(const orc_mesh* mp, const double* gx, const double* gy,
                           const double* gz, const double* p, const double* flux, double* b) {
  return guard([&] {
    const orc_mesh& m = *mp;
    std::vector<double> corr(m.nf, 0.0);
    for (int64_t f = m.nbf; f < m.nf; ++f) {  // _nonorthogonal_face_correction :285-321
      const int64_t c1 = m.own1[f], c2 = m.own2[f];
      const V3 xf = m.fc[f], xC = m.cc[c1], xN = m.cc[c2], nr = m.fn[f], e = m.fe[f];
      const double area = m.area[f];
      // correction_weight (:323-336)
      const V3 c1_f{xf.x - xC.x, xf.y - xC.y, xf.z - xC.z};
      const V3 c1_c2{xN.x - xC.x, xN.y - xC.y, xN.z - xC.z};
      const double q = dot(c1_f, c1_c2) / dot(c1_c2, c1_c2);
      const V3 f_prime{xC.x - q * (xC.x - xN.x), xC.y - q * (xC.y - xN.y),
                       xC.z - q * (xC.z - xN.z)};
      const double w = norm(V3{xN.x - f_prime.x, xN.y - f_prime.y, xN.z - f_prime.z}) /
                       norm(c1_c2);
      const V3 gi{w * gx[c1] + (1.0 - w) * gx[c2], w * gy[c1] + (1.0 - w) * gy[c2],
                  w * gz[c1] + (1.0 - w) * gz[c2]};
      const double s = (p[c2] - p[c1]) / m.delta[f] - dot(gi, e);
      const V3 gf{gi.x + s * e.x, gi.y + s * e.y, gi.z + s * e.z};
      const V3 dPN = c1_c2;
      const double nn = norm(nr);
      const double fac = (nn * nn) / dot(dPN, nr);
      const V3 T{area * nr.x - (dPN.x * fac) * area, area * nr.y - (dPN.y * fac) * area,
                 area * nr.z - (dPN.z * fac) * area};
      const double J = flux[f];
      corr[f] = dot(V3{J * gf.x, J * gf.y, J * gf.z}, T);  // flux*gradf.T_hat
    }
    // b[owner] += c, b[neighbour] -= c; the reference's atomics have no fixed
    // order, this restatement takes each cell's face-list order
    for (int64_t i = 0; i < m.n; ++i)
      for (int64_t fi = m.cf_first[i]; fi <= m.cf_last[i]; ++fi)
        b[i] += (double)m.cell_nsign[fi] * corr[m.cell_faces[fi]];
    return 0;
  });
}


// ------------------------------------------------------------- k-omega
// turbulence!(KOmegaModel) (RANS_kOmega.jl:175-235) with the wall functions
// of RANS_functions.jl; rho = 1.  Wall-function face loops run in ascending
// boundary-face order (patch order), so the last wall face of a cell wins.
namespace {

int patch_of(const orc_mesh& m, int64_t f) {
  for (size_t p = 0; p < m.pfirst.size(); ++p)
    if (f >= m.pfirst[p] && f <= m.plast[p]) return (int)p;
  return -1;
}

struct Wall {  // KWallFunction/OmegaWallFunction/NutWallFunction defaults
  double kappa = 0.41, beta1 = 0.075, cmu = 0.09, E = 9.8, cmu25, yPlusLam, nu;
  explicit Wall(double nu_) : nu(nu_) {
    cmu25 = std::pow(cmu, 0.25);
    double yL = 11.0;  // y_plus_laminar (RANS_functions.jl:40-43)
    for (int i = 0; i < 10; ++i) yL = std::log(std::max(yL * E, 1.0)) / kappa;
    yPlusLam = yL;
  }
  double yplus(double k, double y) const { return ((cmu25 * y) * std::sqrt(k)) / nu; }
  double nutw(double yp) const {
    return jmax(nu * ((yp * kappa) / std::log(jmax(E * yp, 1.0 + 1e-4)) - 1.0), 0.0);
  }
};

// bound! (RANS_functions.jl:2-38), from the values before the call
void bound(const orc_mesh& m, double* v, double* tmp) {
  const int64_t n = m.n;
  std::copy(v, v + n, tmp);
  const double mzero = std::numeric_limits<double>::epsilon();
  for (int64_t i = 0; i < n; ++i) {
    double sum = 0.0;
    int cnt = 0;
    for (int64_t k = m.cf_first[i]; k <= m.cf_last[i]; ++k) {
      sum += jmax(tmp[m.cell_nbr[k]], mzero);
      ++cnt;
    }
    const double average = sum / (double)cnt;
    const double a = std::signbit(tmp[i]) ? average : std::copysign(0.0, average);
    v[i] = jmax(jmax(tmp[i], a), mzero);
  }
}

}  // namespace

// magnitude2!(Pk, S, 2) with S = (gradU + gradU')/2 (RANS_tensor_algebra.jl:68),
// then Pw = alpha1 * 2S:S and Pk = nut * 2S:S (RANS_kOmega.jl:193-197)
void orc_kw_production(const orc_mesh* mp, const double* const* gU, const double* nut,
                       double alpha1, double* Pk, double* Pw) {
  const int64_t n = mp->n;
  for (int64_t i = 0; i < n; ++i) {
    double sum = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int q = 0; q < 3; ++q) {
        const double sj = 0.5 * (gU[3 * j + q][i] + gU[3 * q + j][i]);
        sum += sj * sj;
      }
    const double P = sum * 2.0;
    Pw[i] = (1.0 * alpha1) * P;
    Pk[i] = (1.0 * nut[i]) * P;
  }
}

// correct_production! (RANS_functions.jl:62-124) on KWallFunction patches,
// wall velocity 0: the cell's Pk becomes the log-law production
int orc_kw_wall_production(const orc_mesh* mp, const xcal_bc* bc_k, int32_t nbc, double nu,
                           const double* Ux, const double* Uy, const double* Uz,
                           const double* k, double* Pk) {
  return guard([&] {
    const orc_mesh& m = *mp;
    if (nbc != (int32_t)m.pfirst.size()) throw std::runtime_error("one BC per patch");
    const Wall W(nu);
    for (int64_t f = 0; f < m.nbf; ++f) {
      const int p = patch_of(m, f);
      if (p < 0 || bc_k[p].kind != XCAL_BC_KWALL) continue;
      const int64_t c = m.bcell[f];
      const double delta = m.delta[f];
      const double uStar = W.cmu25 * std::sqrt(k[c]);
      const double dUdy = uStar / (W.kappa * delta);
      const double yp = W.yplus(k[c], delta);
      const double nutw = W.nutw(yp);
      const V3 nrm = m.fn[f];
      const double dx = Ux[c] - 0.0, dy = Uy[c] - 0.0, dz = Uz[c] - 0.0;
      const double dn = dx * nrm.x + dy * nrm.y + dz * nrm.z;
      const double gx = (dx - dn * nrm.x) / delta, gy = (dy - dn * nrm.y) / delta,
                   gz = (dz - dn * nrm.z) / delta;
      const double mag = std::sqrt(gx * gx + gy * gy + gz * gz);
      Pk[c] = yp > W.yPlusLam ? ((W.nu + nutw) * mag) * dUdy : 0.0;
    }
    return 0;
  });
}

// The remaining coefficients of turbulence!(KOmegaModel) (RANS_kOmega.jl:198-206):
// implicit destruction Dw = beta1 omega and Dk = beta* omega, face
// diffusivities mueffw = nu + sigma_omega nutf and mueffk = nu + sigma_k nutf
void orc_kw_coeffs(const orc_mesh* mp, const double* w, const double* nutf, double nu,
                   double beta1, double beta_star, double sigma_omega, double sigma_k, double* Dw,
                   double* Dk, double* mw, double* mk) {
  const int64_t n = mp->n, nf = mp->nf;
  for (int64_t i = 0; i < n; ++i) {
    Dw[i] = (1.0 * beta1) * w[i];
    Dk[i] = (1.0 * beta_star) * w[i];
  }
  for (int64_t f = 0; f < nf; ++f) {
    mw[f] = 1.0 * (nu + sigma_omega * nutf[f]);
    mk[f] = 1.0 * (nu + sigma_k * nutf[f]);
  }
}

// constrain_equation!/constrain! (RANS_functions.jl:262-356): the row of each
// OmegaWallFunction face's cell becomes an identity row with b = omega_wall,
// the viscous-sublayer value 6 nu / (beta1 y^2) below y+_lam and the log-law
// value sqrt(k) / (cmu^1/4 kappa y) above (last face of a cell wins)
int orc_kw_constrain(const orc_mesh* mp, const xcal_bc* bc_w, int32_t nbc, double nu,
                     const double* k, double* A, double* b) {
  return guard([&] {
    const orc_mesh& m = *mp;
    if (nbc != (int32_t)m.pfirst.size()) throw std::runtime_error("one BC per patch");
    const Wall W(nu);
    for (int64_t f = 0; f < m.nbf; ++f) {
      const int p = patch_of(m, f);
      if (p < 0 || bc_w[p].kind != XCAL_BC_OMEGAWALL) continue;
      const int64_t c = m.bcell[f];
      const double y = m.delta[f];
      const double wvis = (6.0 * W.nu) / (W.beta1 * (y * y));
      const double wlog = std::sqrt(k[c]) / ((W.cmu25 * W.kappa) * y);
      const double wc = W.yplus(k[c], y) > W.yPlusLam ? wlog : wvis;
      for (int64_t q = m.rowptr[c]; q < m.rowptr[c + 1]; ++q) A[q] = 0.0;
      A[spindex(m, c, c)] = 1.0;
      b[c] = wc;
    }
    return 0;
  });
}

void orc_bound(const orc_mesh* mp, double* v) {
  std::vector<double> tmp(mp->n);
  bound(*mp, v, tmp.data());
}

// nut = k / omega, its face values (interpolate!, correct_boundaries!) and
// correct_eddy_viscosity!'s wall value on NutWallFunction faces
// (RANS_functions.jl:126-200): nu (y+ kappa / log(E y+) - 1) above y+_lam, 0 below
int orc_kw_nut(const orc_mesh* mp, const xcal_bc* bc_nut, int32_t nbc, double nu,
               const double* k, const double* w, double* nut, double* nutf) {
  return guard([&] {
    const orc_mesh& m = *mp;
    if (nbc != (int32_t)m.pfirst.size()) throw std::runtime_error("one BC per patch");
    const Wall W(nu);
    for (int64_t i = 0; i < m.n; ++i) nut[i] = k[i] / w[i];
    orc_interpolate(&m, nut, nutf);
    const int rc = orc_correct_boundaries(&m, bc_nut, nbc, nut, nullptr, nullptr, nutf, nullptr,
                                          nullptr);
    if (rc < 0) throw std::runtime_error(g_err);
    for (int64_t f = 0; f < m.nbf; ++f) {
      const int p = patch_of(m, f);
      if (p < 0 || bc_nut[p].kind != XCAL_BC_NUTWALL) continue;
      const int64_t c = m.bcell[f];
      const double yp = W.yplus(k[c], m.delta[f]);
      nutf[f] = yp > W.yPlusLam ? W.nutw(yp) : 0.0;
    }
    return 0;
  });
}

void orc_kw_wall_constants(double nu, double* out4) {
  const Wall W(nu);
  out4[0] = W.cmu25;
  out4[1] = W.yPlusLam;
  out4[2] = W.kappa;
  out4[3] = W.E;
}

}  // extern "C"

static bool komega_step(const orc_mesh& m, const xcal_simple_setup& S,
                        const xcal_turbulence_setup& T, const double* Ux, const double* Uy,
                        const double* Uz, const double* mdotf, double* const* gU, double* k,
                        double* w, double* nut, double* nutf, double* Pk, double* Pw, double* Dk,
                        double* Dw, double* mk, double* mw, double* tmp, double* A, double* b) {
  const int nb = S.n_bcs;
  auto chk = [](int rc) {
    if (rc < 0) throw std::runtime_error(g_err);
  };
  orc_kw_production(&m, gU, nut, T.alpha1, Pk, Pw);
  chk(orc_kw_wall_production(&m, T.bc_k, nb, S.nu, Ux, Uy, Uz, k, Pk));
  orc_kw_coeffs(&m, w, nutf, S.nu, T.beta1, T.beta_star, T.sigma_omega, T.sigma_k, Dw, Dk, mw,
                mk);
  auto transport = [&](double* phi, double* D, double* mu, double* P, const xcal_bc* bc,
                       const xcal_solver_setup& ss, bool constrain) {
    orc_term t[4] = {{XCAL_TIME_STEADY, 1, nullptr, 1.0},
                     {T.div_scheme, 1, mdotf, 0.0},
                     {XCAL_LAPLACIAN_LINEAR, -1, mu, 0.0},
                     {XCAL_SI, 1, D, 0.0}};
    orc_source src[1] = {{P, nullptr, nullptr, 1}};
    chk(orc_discretise_scalar(&m, t, 4, src, 1, phi, 1.0, A, b));
    chk(orc_apply_bcs(&m, t, 4, bc, nb, -1, phi, nullptr, nullptr, A, b));
    orc_relax_implicit_diagdom(&m, A, b, phi, ss.relax);
    if (constrain) chk(orc_kw_constrain(&m, bc, nb, S.nu, k, A, b));
    int32_t it = 0;
    chk(orc_solve_system(&m, &ss, 0, A, b, phi, &it));
    const double r = orc_residual(&m, A, b, phi);
    bound(m, phi, tmp);
    return r;
  };
  const double rw = transport(w, Dw, mw, Pw, T.bc_omega, T.omega, true);
  const double rk = transport(k, Dk, mk, Pk, T.bc_k, T.k, false);
  chk(orc_kw_nut(&m, T.bc_nut, nb, S.nu, k, w, nut, nutf));
  return rk < T.k.convergence && rw < T.omega.convergence;
}



namespace {

// The SIMPLE loop (inner == 0; Solvers_1_SIMPLE.jl:103-263) or PISO's (inner
// >= 1 pressure correctors per time step, the last unrelaxed, every step
// taken; Solvers_2_PISO.jl:107-175).
int run_core(const orc_mesh* mp, const xcal_simple_setup* S, const xcal_turbulence_setup* T,
             double* Ux, double* Uy, double* Uz, double* p, double* kf_, double* wf_,
             double* nut_, int32_t iterations, double* residuals, int32_t* done, int32_t* kit,
             double* mdot_out, int inner) {
  return guard([&] {
    const orc_mesh& m = *mp;
    const int64_t n = m.n, nf = m.nf, nnz = m.rowptr[n];
    const int nb = S->n_bcs;
    // setup_incompressible_solvers (Solvers_1_SIMPLE.jl:45-101)
    std::vector<double> mdotf(nf, 0.0), rDf(nf, 1.0), nueff(nf, 0.0), divHv(n, 0.0);
    std::vector<double> gpx(n), gpy(n), gpz(n), pf(nf);
    std::vector<double> Ufx(nf), Ufy(nf), Ufz(nf);
    std::vector<double> Hx(n), Hy(n), Hz(n), rD(n), prev(n);
    std::vector<double> A0(nnz), A(nnz), bx(n), by(n), bz(n), pA(nnz), pb(n);
    std::vector<std::vector<double>> gradU(9, std::vector<double>(n));
    double* gU[9];
    for (int k = 0; k < 9; ++k) gU[k] = gradU[k].data();
    if (S->time_U != XCAL_TIME_STEADY && S->time_U != XCAL_TIME_EULER &&
        S->time_U != XCAL_TIME_CRANK_NICOLSON)
      throw std::runtime_error("momentum time scheme: SteadyState, Euler or CrankNicolson");
    if (S->time_U != XCAL_TIME_STEADY && !(S->dt > 0.0))
      throw std::runtime_error("Time{Euler} and Time{CrankNicolson} need runtime.dt > 0");
    if (S->ncorrectors < 0) throw std::runtime_error("ncorrectors must be >= 0");
    const double dt = S->time_U != XCAL_TIME_STEADY ? S->dt : 1.0;
    // solve_system!'s Crank-Nicolson step applies to the momentum equation,
    // whose first term is Time{schemes.U.time} (Solve_1_api.jl:291-296)
    const int32_t Ucn = S->time_U == XCAL_TIME_CRANK_NICOLSON;
    const int ncorr = inner > 0 ? inner : 1;
    orc_term Uterms[3] = {{S->time_U, 1, nullptr, 1.0},
                          {S->div_scheme, 1, mdotf.data(), 0.0},
                          {XCAL_LAPLACIAN_LINEAR, -1, nueff.data(), 0.0}};
    orc_source Usrc[1] = {{gpx.data(), gpy.data(), gpz.data(), -1}};
    orc_term pterms[1] = {{XCAL_LAPLACIAN_LINEAR, -1, rDf.data(), 0.0}};
    orc_source psrc[1] = {{divHv.data(), nullptr, nullptr, -1}};
    auto chk = [](int rc) {
      if (rc < 0) throw std::runtime_error(g_err);
    };
    auto grad_p = [&] {  // grad! (Calculate_0_gradient.jl:105-109, Midpoint :235-244)
      if (S->grad_p == XCAL_GRAD_MIDPOINT) {
        const double* ph[1] = {p};
        double* fc[1] = {pf.data()};
        double* g[3] = {gpx.data(), gpy.data(), gpz.data()};
        chk(orc_grad_midpoint(&m, S->bc_p, nb, 1, ph, fc, g));
        return;
      }
      orc_interpolate(&m, p, pf.data());
      chk(orc_correct_boundaries(&m, S->bc_p, nb, p, nullptr, nullptr, pf.data(), nullptr, nullptr));
      orc_green_gauss(&m, pf.data(), gpx.data(), gpy.data(), gpz.data());
    };
    auto interp_vec = [&](const double* x, const double* y, const double* z) {
      orc_interpolate(&m, x, Ufx.data());
      orc_interpolate(&m, y, Ufy.data());
      orc_interpolate(&m, z, Ufz.data());
      chk(orc_correct_boundaries(&m, S->bc_U, nb, x, y, z, Ufx.data(), Ufy.data(), Ufz.data()));
    };
    // k-omega state (RANS_kOmega.jl): nutf starts at zero
    const bool turb = T && T->model == XCAL_TURB_KOMEGA;
    std::vector<double> nutf(nf, 0.0), Pk(n), Pw(n), Dk(n), Dw(n), mk(nf), mw(nf), tmp(n);
    const double nu = S->nu;
    auto update_nueff = [&] {  // update_nueff! (Solvers_0_functions.jl:5-36)
      for (int64_t f = 0; f < nf; ++f) nueff[f] = turb ? nu + nutf[f] : nu;
    };
    // initial calculations (Solvers_1_SIMPLE.jl:150-156)
    interp_vec(Ux, Uy, Uz);
    orc_flux(&m, Ufx.data(), Ufy.data(), Ufz.data(), mdotf.data());
    grad_p();
    update_nueff();
    const bool is3d = m.dim == 3;
    int it = 0;
    double* U[3] = {Ux, Uy, Uz};
    double* bU[3] = {bx.data(), by.data(), bz.data()};
    for (it = 0; it < iterations; ++it) {
      // solve_equation! vector (Solve_1_api.jl:233-266)
      chk(orc_discretise_vector(&m, Uterms, 3, Usrc, 1, Ux, Uy, Uz, dt, A0.data(), bx.data(),
                                by.data(), bz.data()));
      double res[3] = {0.0, 0.0, 0.0};
      int itc[3] = {0, 0, 0};
      for (int c = 0; c < (is3d ? 3 : 2); ++c) {
        A = A0;  // update_equation!
        chk(orc_apply_bcs(&m, Uterms, 3, S->bc_U, nb, c, Ux, Uy, Uz, A.data(), bU[c]));
        orc_relax_implicit_diagdom(&m, A.data(), bU[c], U[c], S->U.relax);
        int32_t ni = 0;
        chk(orc_solve_system(&m, &S->U, Ucn, A.data(), bU[c], U[c], &ni));
        itc[c] = ni;
        res[c] = orc_residual(&m, A.data(), bU[c], U[c]);
      }
      // pressure correction (Solvers_1_SIMPLE.jl:170-185)
      for (int64_t i = 0; i < n; ++i) rD[i] = m.vol[i] / A[m.diag[i]];  // inverse_diagonal!
      orc_interpolate(&m, rD.data(), rDf.data());
      for (int64_t i = 0; i < n; ++i) {  // remove_pressure_source! (sign -1)
        double V = m.vol[i];
        bx[i] -= (-1.0 * gpx[i]) * V;
        by[i] -= (-1.0 * gpy[i]) * V;
        bz[i] -= (-1.0 * gpz[i]) * V;
      }
      double rp = 0.0;
      int32_t pit = 0;
      for (int ic = 0; ic < ncorr; ++ic) {  // PISO's corrector loop (Solvers_2_PISO.jl:123-175)
        orc_H(&m, A.data(), bx.data(), by.data(), bz.data(), Ux, Uy, Uz, Hx.data(), Hy.data(),
              Hz.data());
        interp_vec(Hx.data(), Hy.data(), Hz.data());
        orc_flux(&m, Ufx.data(), Ufy.data(), Ufz.data(), mdotf.data());
        orc_div(&m, mdotf.data(), divHv.data());
        // pressure equation (Solve_1_api.jl:214-231)
        std::copy(p, p + n, prev.begin());
        chk(orc_discretise_scalar(&m, pterms, 1, psrc, 1, p, 1.0, pA.data(), pb.data()));
        chk(orc_apply_bcs(&m, pterms, 1, S->bc_p, nb, -1, p, nullptr, nullptr, pA.data(),
                          pb.data()));
        if (S->has_pref) orc_set_reference(&m, pA.data(), pb.data(), S->pref, 0);
        chk(orc_solve_system(&m, &S->p, 0, pA.data(), pb.data(), p, &pit));
        rp = orc_residual(&m, pA.data(), pb.data(), p);
        // explicit_relaxation!: PISO's last corrector uses 1.0 (:141-145)
        const double prelax = inner > 0 && ic == ncorr - 1 ? 1.0 : S->p.relax;
        for (int64_t i = 0; i < n; ++i) p[i] = prev[i] + prelax * (p[i] - prev[i]);
        grad_p();
        // non-orthogonal correctors (Solvers_1_SIMPLE.jl:194-206; PISO :148-165
        // also sets the reference, updates the preconditioner and leaves its
        // last corrector unrelaxed)
        for (int jc = 0; jc < S->ncorrectors; ++jc) {
          chk(orc_discretise_scalar(&m, pterms, 1, psrc, 1, p, 1.0, pA.data(), pb.data()));
          chk(orc_apply_bcs(&m, pterms, 1, S->bc_p, nb, -1, p, nullptr, nullptr, pA.data(),
                            pb.data()));
          if (inner > 0 && S->has_pref) orc_set_reference(&m, pA.data(), pb.data(), S->pref, 0);
          chk(orc_nonorth_correction(&m, gpx.data(), gpy.data(), gpz.data(), p, rDf.data(),
                                     pb.data()));
          chk(orc_solve_system(&m, &S->p, 0, pA.data(), pb.data(), p, &pit));
          rp = orc_residual(&m, pA.data(), pb.data(), p);
          const double nrelax = inner > 0 && jc == S->ncorrectors - 1 ? 1.0 : S->p.relax;
          for (int64_t i = 0; i < n; ++i) p[i] = prev[i] + nrelax * (p[i] - prev[i]);
          grad_p();
        }
        chk(orc_correct_mass_flux(&m, pA.data(), p, rDf.data(), -1, S->bc_p, nb, mdotf.data()));
        for (int64_t i = 0; i < n; ++i) {  // correct_velocity!
          double r = rD[i];
          Ux[i] = Hx[i] - gpx[i] * r;
          Uy[i] = Hy[i] - gpy[i] * r;
          Uz[i] = Hz[i] - gpz[i] * r;
        }
      }
      // turbulence!: grad!(gradU) (RANS_laminar.jl:73-79, RANS_kOmega.jl:192)
      if (S->grad_U == XCAL_GRAD_MIDPOINT) {
        const double* ph[3] = {Ux, Uy, Uz};
        double* fc[3] = {Ufx.data(), Ufy.data(), Ufz.data()};
        chk(orc_grad_midpoint(&m, S->bc_U, nb, 3, ph, fc, gU));
      } else {
        interp_vec(Ux, Uy, Uz);
        orc_green_gauss_tensor(&m, Ufx.data(), Ufy.data(), Ufz.data(), gU);
      }
      bool turb_conv = true;
      if (turb)
        turb_conv = komega_step(m, *S, *T, Ux, Uy, Uz, mdotf.data(), gU, kf_, wf_, nut_,
                                nutf.data(), Pk.data(), Pw.data(), Dk.data(), Dw.data(), mk.data(),
                                mw.data(), tmp.data(), pA.data(), pb.data());
      update_nueff();
      residuals[it] = res[0];
      residuals[iterations + it] = res[1];
      residuals[2 * iterations + it] = res[2];
      residuals[3 * iterations + it] = rp;
      if (kit) {
        kit[it] = itc[0];
        kit[iterations + it] = itc[1];
        kit[2 * iterations + it] = itc[2];
        kit[3 * iterations + it] = pit;
      }
      bool zc = is3d ? res[2] <= S->U.convergence : true;
      if (inner == 0 && res[0] <= S->U.convergence && res[1] <= S->U.convergence && zc &&
          rp <= S->p.convergence && turb_conv) {
        ++it;
        break;
      }
    }
    *done = it;
    if (mdot_out) std::copy(mdotf.begin(), mdotf.end(), mdot_out);
    return 0;
  });
}

}  // namespace

extern "C" {

int orc_simple_run_turb(const orc_mesh* mp, const xcal_simple_setup* S,
                        const xcal_turbulence_setup* T, double* Ux, double* Uy, double* Uz,
                        double* p, double* kf_, double* wf_, double* nut_, int32_t iterations,
                        double* residuals, int32_t* done, int32_t* kit, double* mdot_out) {
  return run_core(mp, S, T, Ux, Uy, Uz, p, kf_, wf_, nut_, iterations, residuals, done, kit,
                  mdot_out, 0);
}

int orc_simple_run(const orc_mesh* mp, const xcal_simple_setup* S, double* Ux, double* Uy,
                   double* Uz, double* p, int32_t iterations, double* residuals, int32_t* done,
                   int32_t* kit, double* mdot_out) {
  return run_core(mp, S, nullptr, Ux, Uy, Uz, p, nullptr, nullptr, nullptr, iterations,
                  residuals, done, kit, mdot_out, 0);
}

int orc_piso_run(const orc_mesh* mp, const xcal_simple_setup* S, double* Ux, double* Uy,
                 double* Uz, double* p, int32_t steps, int32_t inner_loops, double* residuals,
                 int32_t* done, int32_t* kit, double* mdot_out) {
  if (inner_loops < 1) {
    g_err = "PISO needs at least one pressure corrector";
    return -1;
  }
  return run_core(mp, S, nullptr, Ux, Uy, Uz, p, nullptr, nullptr, nullptr, steps, residuals,
                  done, kit, mdot_out, inner_loops);
}

}  // extern "C"
