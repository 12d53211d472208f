// Face interpolation, boundary face values, Green-Gauss gradients,
// divergence, face fluxes and the SIMPLE pressure-velocity coupling stages.
//
// Cell kernels gather over the cell's internal faces in the reference's
// faces_range order; the reference's boundary-face atomics
// (Calculate_1_green_gauss.jl:74-76,161-171; Calculate_0_divergence.jl:168)
// are replaced by a second pass over cells that own boundary faces, adding
// their contributions in ascending face id.
#include "device_common.h"

namespace xcal {

// interpolate_Scalar! (Calculate_2_interpolation.jl:58-76), all faces
__global__ void __launch_bounds__(kBlock)
    k_interpolate(int64_t nf, const int32_t* __restrict__ own, const int32_t* __restrict__ nei,
                  const double* __restrict__ w, const double* __restrict__ phi,
                  double* __restrict__ phif) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double wf = w[f];
  const double omw = 1.0 - wf;
  phif[f] = wf * phi[own[f]] + omw * phi[nei[f]];
}

// interpolate_midpoint! (Calculate_0_gradient.jl:121-193): every face, the
// two cells' mean (a boundary face's two cells are the same cell)
__global__ void __launch_bounds__(kBlock)
    k_interp_midpoint(int64_t nf, const int32_t* __restrict__ own, const int32_t* __restrict__ nei,
                      const double* __restrict__ phi, double* __restrict__ phif) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double p1 = phi[own[f]];
  phif[f] = 0.5 * (p1 + phi[nei[f]]);
}

// correct_interpolation! (Calculate_0_gradient.jl:195-249) over internal
// faces: face = (P + N)/2 + g . Ri with g = (grad_P + grad_N)/2 and
// Ri = r_f - (r_P + r_N)/2, for each of the field's ncomp components; grad
// holds 3 ncomp components of n values (row p of the tensor: component 3p+q)
__global__ void __launch_bounds__(kBlock)
    k_correct_interp(int64_t nbf, int64_t nf, int64_t n, int ncomp, const int32_t* __restrict__ own,
                     const int32_t* __restrict__ nei, const double* __restrict__ rix,
                     const double* __restrict__ riy, const double* __restrict__ riz,
                     const double* __restrict__ px, const double* __restrict__ py,
                     const double* __restrict__ pz, const double* __restrict__ grad,
                     double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz) {
  const int64_t f = nbf + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int64_t c1 = own[f], c2 = nei[f];
  const double R[3] = {rix[f], riy[f], riz[f]};
  const double* P[3] = {px, py, pz};
  double* F[3] = {fx, fy, fz};
  for (int p = 0; p < ncomp; ++p) {
    const double phip = 0.5 * (P[p][c1] + P[p][c2]);
    double g[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double* gq = grad + (int64_t)(3 * p + q) * n;
      g[q] = 0.5 * (gq[c1] + gq[c2]);
    }
    F[p][f] = phip + ((g[0] * R[0] + g[1] * R[1]) + g[2] * R[2]);
  }
}

// boundary_interpolation! (1_boundary_interpolation.jl and the
// *_interpolation.jl files), one work-item per boundary face
__global__ void __launch_bounds__(kBlock)
    k_correct_boundaries(MeshK m, BCsK B, int ncomp, const double* __restrict__ px,
                         const double* __restrict__ py, const double* __restrict__ pz,
                         double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nbf) return;
  const int p = m.bface_patch[f];
  if (p < 0) return;
  const BCK& b = B.p[p];
  const int c = m.own[f];
  const bool vec = ncomp == 3;
  switch (b.kind) {
    case XCAL_BC_DIRICHLET:
      fx[f] = b.v[0];
      if (vec) { fy[f] = b.v[1]; fz[f] = b.v[2]; }
      break;
    case XCAL_BC_NEUMANN:  // scalar only (vector rejected on the host)
      fx[f] = px[c] + m.delta[f] * b.v[0];
      break;
    case XCAL_BC_WALL:
      if (vec) { fx[f] = b.v[0]; fy[f] = b.v[1]; fz[f] = b.v[2]; }
      else fx[f] = px[c];
      break;
    case XCAL_BC_ZEROGRADIENT:
    case XCAL_BC_EXTRAPOLATED:
    case XCAL_BC_KWALL:
    case XCAL_BC_OMEGAWALL:
    case XCAL_BC_NUTWALL:
      fx[f] = px[c];
      if (vec) { fy[f] = py[c]; fz[f] = pz[c]; }
      break;
    case XCAL_BC_EMPTY:  // empty_interpolation.jl:1-16
      fx[f] = 0.0;
      if (vec) { fy[f] = 0.0; fz[f] = 0.0; }
      break;
    case XCAL_BC_SYMMETRY:
      if (vec) {
        const double nx = m.nx[f], ny = m.ny[f], nz = m.nz[f];
        const double pn = px[c] * nx + py[c] * ny + pz[c] * nz;
        fx[f] = px[c] - pn * nx;
        fy[f] = py[c] - pn * ny;
        fz[f] = pz[c] - pn * nz;
      } else {
        fx[f] = px[c];
      }
      break;
    default:
      break;
  }
}

// _green_gauss! (Calculate_1_green_gauss.jl:27-54)
__global__ void __launch_bounds__(kBlock)
    k_green_gauss(MeshK m, const double* __restrict__ phif, double* __restrict__ gx,
                  double* __restrict__ gy, double* __restrict__ gz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  double rx = 0.0, ry = 0.0, rz = 0.0;
  struct Rec {
    double v, sx, sy, sz;  // sx, sy, sz: S_f nsign, streamed in cf order
  };
  cell_faces(
      m.cf_ptr[i], m.cf_ptr[i + 1],
      [&](int32_t k) {
        const int f = m.cf_face[k];
        return Rec{phif[f], m.cf_sx[k], m.cf_sy[k], m.cf_sz[k]};
      },
      [&](const Rec& r, bool ok) {
        const double tx = r.v * r.sx, ty = r.v * r.sy, tz = r.v * r.sz;
        rx = ok ? rx + tx : rx;
        ry = ok ? ry + ty : ry;
        rz = ok ? rz + tz : rz;
      });
  const double V = m.vol[i];
  gx[i] = rx / V;
  gy[i] = ry / V;
  gz[i] = rz / V;
}

// boundary_faces_contribution! (Calculate_1_green_gauss.jl:58-78)
__global__ void __launch_bounds__(kBlock)
    k_green_gauss_bnd(MeshK m, const double* __restrict__ phif, double* __restrict__ gx,
                      double* __restrict__ gy, double* __restrict__ gz) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m.nbcell) return;
  const int c = m.bc_cells[s];
  const double V = m.vol[c];
  double x = gx[c], y = gy[c], z = gz[c];
  for (int32_t k = m.bc_ptr[s]; k < m.bc_ptr[s + 1]; ++k) {
    const int f = m.bc_face[k];
    const double v = phif[f];
    x += (v * m.sfx[f]) / V;
    y += (v * m.sfy[f]) / V;
    z += (v * m.sfz[f]) / V;
  }
  gx[c] = x;
  gy[c] = y;
  gz[c] = z;
}

// _green_gauss_vector! (Calculate_1_green_gauss.jl:106-151); t0 holds the
// nine components xx,xy,xz,yx,yy,yz,zx,zy,zz with stride n
__global__ void __launch_bounds__(kBlock)
    k_green_gauss_tensor(MeshK m, const double* __restrict__ fx, const double* __restrict__ fy,
                         const double* __restrict__ fz, double* __restrict__ t0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  double r[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  struct Rec {
    double S[3], P[3];
  };
  cell_faces(
      m.cf_ptr[i], m.cf_ptr[i + 1],
      [&](int32_t k) {
        const int f = m.cf_face[k];
        return Rec{{m.cf_sx[k], m.cf_sy[k], m.cf_sz[k]}, {fx[f], fy[f], fz[f]}};
      },
      [&](const Rec& e, bool ok) {
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const double t = e.P[p] * e.S[q];
            r[3 * p + q] = ok ? r[3 * p + q] + t : r[3 * p + q];
          }
      });
  const double V = m.vol[i];
#pragma unroll
  for (int k = 0; k < 9; ++k) t0[(int64_t)k * m.n + i] = r[k] / V;
}

// boundary_faces_contribution_vector! (Calculate_1_green_gauss.jl:153-173)
__global__ void __launch_bounds__(kBlock)
    k_green_gauss_tensor_bnd(MeshK m, const double* __restrict__ fx, const double* __restrict__ fy,
                             const double* __restrict__ fz, double* __restrict__ t0) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m.nbcell) return;
  const int c = m.bc_cells[s];
  const double V = m.vol[c];
  double r[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) r[k] = t0[(int64_t)k * m.n + c];
  for (int32_t k = m.bc_ptr[s]; k < m.bc_ptr[s + 1]; ++k) {
    const int f = m.bc_face[k];
    const double S[3] = {m.sfx[f], m.sfy[f], m.sfz[f]};
    const double P[3] = {fx[f], fy[f], fz[f]};
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) r[3 * p + q] += (P[p] * S[q]) / V;
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) t0[(int64_t)k * m.n + c] = r[k];
}

// div_noS_kernel! + boundary contribution (Calculate_0_divergence.jl:129-171)
__global__ void __launch_bounds__(kBlock)
    k_div(MeshK m, const double* __restrict__ psif, double* __restrict__ div) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  double red = 0.0;
  cell_faces(
      m.cf_ptr[i], m.cf_ptr[i + 1],
      [&](int32_t k) { return psif[m.cf_face[k]] * (double)m.cf_sign[k]; },
      [&](double t, bool ok) { red = ok ? red + t : red; });
  div[i] = red / m.vol[i];
}
__global__ void __launch_bounds__(kBlock)
    k_div_bnd(MeshK m, const double* __restrict__ psif, double* __restrict__ div) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m.nbcell) return;
  const int c = m.bc_cells[s];
  const double V = m.vol[c];
  double d = div[c];
  for (int32_t k = m.bc_ptr[s]; k < m.bc_ptr[s + 1]; ++k) d += psif[m.bc_face[k]] / V;
  div[c] = d;
}

// flux_kernel! (Solvers_0_functions.jl:51-64)
__global__ void __launch_bounds__(kBlock)
    k_flux(int64_t nf, const double* __restrict__ sfx, const double* __restrict__ sfy,
           const double* __restrict__ sfz, const double* __restrict__ fx,
           const double* __restrict__ fy, const double* __restrict__ fz, double* __restrict__ mdot) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  mdot[f] = fx[f] * sfx[f] + fy[f] * sfy[f] + fz[f] * sfz[f];
}

// explicit_relaxation_kernel! (Solve_1_api.jl:327-333)
__global__ void __launch_bounds__(kBlock)
    k_relax_explicit(int64_t n, double* __restrict__ phi, const double* __restrict__ phi0,
                     double alpha) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  phi[i] = phi0[i] + alpha * (phi[i] - phi0[i]);
}

// _inverse_diagonal! (Solvers_0_functions.jl:108-122)
__global__ void __launch_bounds__(kBlock)
    k_inverse_diagonal(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ nz,
                       const double* __restrict__ vol, double* __restrict__ rD) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rD[i] = vol[i] / nz[diag[i]];
}

// _remove_pressure_source! (Solvers_0_functions.jl:170-181)
__global__ void __launch_bounds__(kBlock)
    k_remove_pressure_source(int64_t n, const double* __restrict__ vol, double sign,
                             const double* __restrict__ gx, const double* __restrict__ gy,
                             const double* __restrict__ gz, double* __restrict__ bx,
                             double* __restrict__ by, double* __restrict__ bz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double V = vol[i];
  bx[i] -= (sign * gx[i]) * V;
  by[i] -= (sign * gy[i]) * V;
  bz[i] -= (sign * gz[i]) * V;
}

// _H! (Solvers_0_functions.jl:201-249)
__global__ void __launch_bounds__(kBlock)
    k_H(int64_t n, const int32_t* __restrict__ rb, const int32_t* __restrict__ re,
        const int32_t* __restrict__ colval, const int32_t* __restrict__ diag,
        const double* __restrict__ nz,
        const double* __restrict__ bx, const double* __restrict__ by, const double* __restrict__ bz,
        const double* __restrict__ Ux, const double* __restrict__ Uy, const double* __restrict__ Uz,
        double* __restrict__ Hx, double* __restrict__ Hy, double* __restrict__ Hz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  // kG entries at a time, every load issued unconditionally (slots past the
  // row's last entry re-read it and are left out of the sums)
  constexpr int kG = 8;
  const int32_t k1 = re[i];
  const int32_t klast = k1 - kSlice;
  for (int32_t k0 = rb[i]; k0 < k1; k0 += kG * kSlice) {
    int col[kG];
    double v[kG], ux[kG], uy[kG], uz[kG];
#pragma unroll
    for (int u = 0; u < kG; ++u) {
      const int32_t k = min(k0 + u * kSlice, klast);
      col[u] = colval[k];
      v[u] = nz[k];
    }
#pragma unroll
    for (int u = 0; u < kG; ++u) {
      ux[u] = Ux[col[u]];
      uy[u] = Uy[col[u]];
      uz[u] = Uz[col[u]];
    }
#pragma unroll
    for (int u = 0; u < kG; ++u) {
      const bool in = k0 + u * kSlice < k1;
      const double tx = v[u] * ux[u], ty = v[u] * uy[u], tz = v[u] * uz[u];
      sx = in ? sx + tx : sx;
      sy = in ? sy + ty : sy;
      sz = in ? sz + tz : sz;
    }
  }
  const double D = nz[diag[i]];
  sx -= D * Ux[i];
  sy -= D * Uy[i];
  sz -= D * Uz[i];
  const double rD = 1.0 / D;
  Hx[i] = (bx[i] - sx) * rD;
  Hy[i] = (by[i] - sy) * rD;
  Hz[i] = (bz[i] - sz) * rD;
}

// _correct_mass_flux! (Solvers_1_SIMPLE.jl:371-388), internal faces
__global__ void __launch_bounds__(kBlock)
    k_correct_mass_flux(MeshK m, const double* __restrict__ pnz, const double* __restrict__ p,
                        double* __restrict__ mdot) {
  const int64_t f = m.nbf + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nf) return;
  const double aN = pnz[m.fslot[f]];
  mdot[f] += aN * (p[m.nei[f]] - p[m.own[f]]);
}

// _correct_boundary_mass_flux! (Solvers_1_SIMPLE.jl:490-557)
__global__ void __launch_bounds__(kBlock)
    k_correct_boundary_mass_flux(MeshK m, BCsK B, const double* __restrict__ p,
                                 const double* __restrict__ pflux, double psign,
                                 double* __restrict__ mdot) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nbf) return;
  const int pt = m.bface_patch[f];
  if (pt < 0) return;
  const BCK& b = B.p[pt];
  if (b.kind == XCAL_BC_DIRICHLET) {
    const double flux = (pflux[f] * m.area[f]) / m.delta[f];
    const double ap = psign * (-flux);
    mdot[f] += ap * (p[m.own[f]] - b.v[0]);
  } else if (b.kind == XCAL_BC_NEUMANN) {
    mdot[f] -= (pflux[f] * m.area[f]) * b.v[0];
  }
}

// _correct_velocity! (Solvers_0_functions.jl:136-152)
__global__ void __launch_bounds__(kBlock)
    k_correct_velocity(int64_t n, double* __restrict__ Ux, double* __restrict__ Uy,
                       double* __restrict__ Uz, const double* __restrict__ Hx,
                       const double* __restrict__ Hy, const double* __restrict__ Hz,
                       const double* __restrict__ gx, const double* __restrict__ gy,
                       const double* __restrict__ gz, const double* __restrict__ rD) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rD[i];
  Ux[i] = Hx[i] - gx[i] * r;
  Uy[i] = Hy[i] - gy[i] * r;
  Uz[i] = Hz[i] - gz[i] * r;
}

#define LAUNCH(kern, n, ...)                                                           \
  do {                                                                                 \
    if ((n) > 0)                                                                       \
      hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(kBlock), 0, m->ctx->stream,    \
                         __VA_ARGS__);                                                 \
    XCAL_HIP(hipGetLastError());                                                       \
  } while (0)

void interpolate(const xcal_mesh* m, const double* phi, double* phif) {
  LAUNCH(k_interpolate, m->nf, m->nf, m->own.p, m->nei.p, m->weight.p, phi, phif);
}

void correct_boundaries(const xcal_mesh* m, const BCsK& b, const double* const* phi,
                        double* const* phif, int ncomp) {
  const double* px = phi[0];
  const double* py = ncomp == 3 ? phi[1] : px;
  const double* pz = ncomp == 3 ? phi[2] : px;
  double* fx = phif[0];
  double* fy = ncomp == 3 ? phif[1] : nullptr;
  double* fz = ncomp == 3 ? phif[2] : nullptr;
  LAUNCH(k_correct_boundaries, m->nbf, meshk(m), b, ncomp, px, py, pz, fx, fy, fz);
}

void green_gauss(const xcal_mesh* m, const double* phif, double* gx, double* gy, double* gz) {
  MeshK mk = meshk(m);
  LAUNCH(k_green_gauss, m->n, mk, phif, gx, gy, gz);
  LAUNCH(k_green_gauss_bnd, m->nbcell, mk, phif, gx, gy, gz);
}

void green_gauss_tensor(const xcal_mesh* m, const double* fx, const double* fy, const double* fz,
                        double* t0) {
  MeshK mk = meshk(m);
  LAUNCH(k_green_gauss_tensor, m->n, mk, fx, fy, fz, t0);
  LAUNCH(k_green_gauss_tensor_bnd, m->nbcell, mk, fx, fy, fz, t0);
}

// grad! with Grad{Midpoint} (Calculate_0_gradient.jl:235-244) of a scalar
// (ncomp 1, grad = 3 components) or vector (ncomp 3, grad = 9 components)
void grad_midpoint(const xcal_mesh* m, const BCsK& b, const double* const* phi,
                   double* const* phif, int ncomp, double* grad) {
  for (int c = 0; c < ncomp; ++c)
    LAUNCH(k_interp_midpoint, m->nf, m->nf, m->own.p, m->nei.p, phi[c], phif[c]);
  correct_boundaries(m, b, phi, phif, ncomp);
  const int64_t n = m->n;
  auto gauss = [&] {
    if (ncomp == 1) green_gauss(m, phif[0], grad, grad + n, grad + 2 * n);
    else green_gauss_tensor(m, phif[0], phif[1], phif[2], grad);
  };
  gauss();
  const double* px = phi[0];
  const double* py = ncomp == 3 ? phi[1] : px;
  const double* pz = ncomp == 3 ? phi[2] : px;
  double* fx = phif[0];
  double* fy = ncomp == 3 ? phif[1] : fx;
  double* fz = ncomp == 3 ? phif[2] : fx;
  for (int it = 0; it < 2; ++it) {
    LAUNCH(k_correct_interp, m->nf - m->nbf, m->nbf, m->nf, n, ncomp, m->own.p, m->nei.p,
           m->rix.p, m->riy.p, m->riz.p, px, py, pz, (const double*)grad, fx, fy, fz);
    gauss();
  }
}

void divergence(const xcal_mesh* m, const double* psif, double* div) {
  MeshK mk = meshk(m);
  LAUNCH(k_div, m->n, mk, psif, div);
  LAUNCH(k_div_bnd, m->nbcell, mk, psif, div);
}

void flux(const xcal_mesh* m, const double* fx, const double* fy, const double* fz, double* mdot) {
  LAUNCH(k_flux, m->nf, m->nf, m->sfx.p, m->sfy.p, m->sfz.p, fx, fy, fz, mdot);
}

void relax_explicit(const xcal_mesh* m, int64_t len, double* phi, const double* phi0, double alpha) {
  LAUNCH(k_relax_explicit, len, len, phi, phi0, alpha);
}

void inverse_diagonal(const xcal_eqn* e, double* rD) {
  const xcal_mesh* m = e->mesh;
  LAUNCH(k_inverse_diagonal, m->n, m->n, m->diag.p, e->nzval.p, m->vol.p, rD);
}

void remove_pressure_source(xcal_eqn* e, const double* gx, const double* gy, const double* gz,
                            int sign) {
  const xcal_mesh* m = e->mesh;
  LAUNCH(k_remove_pressure_source, m->n, m->n, m->vol.p, (double)sign, gx, gy, gz, e->bc(0),
         e->bc(1), e->bc(2));
}

void Hop(const xcal_eqn* e, const double* Ux, const double* Uy, const double* Uz, double* Hx,
         double* Hy, double* Hz) {
  const xcal_mesh* m = e->mesh;
  LAUNCH(k_H, m->n, m->n, m->rbeg.p, m->rend.p, m->colval.p, m->diag.p, e->nzval.p, e->bc(0), e->bc(1),
         e->bc(2), Ux, Uy, Uz, Hx, Hy, Hz);
}

void correct_mass_flux(const xcal_eqn* pe, const double* p, const double* pflux, int psign,
                       const BCsK& b, double* mdot) {
  const xcal_mesh* m = pe->mesh;
  MeshK mk = meshk(m);
  LAUNCH(k_correct_mass_flux, m->nf - m->nbf, mk, pe->nzval.p, p, mdot);
  LAUNCH(k_correct_boundary_mass_flux, m->nbf, mk, b, p, pflux, (double)psign, mdot);
}

// nonorthogonal_face_correction (Solvers_1_SIMPLE.jl:267-321), per internal
// face: gradf = gradi + ((p_N - p_P)/delta - gradi.e) e with gradi the
// weighted cell gradients, and the correction (flux gradf).T_hat, in the
// reference's operation order (T_hat and the weight precomputed per face).
__global__ void __launch_bounds__(kBlock)
    k_nonorth_face(MeshK m, const double* __restrict__ ncw, const double* __restrict__ ntx,
                   const double* __restrict__ nty, const double* __restrict__ ntz,
                   const double* __restrict__ gx, const double* __restrict__ gy,
                   const double* __restrict__ gz, const double* __restrict__ p,
                   const double* __restrict__ flux, double* __restrict__ corr) {
  const int64_t f = m.nbf + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nf) return;
  const int c1 = m.own[f], c2 = m.nei[f];
  const double w = ncw[f], w1 = 1.0 - w;
  const double gix = w * gx[c1] + w1 * gx[c2];
  const double giy = w * gy[c1] + w1 * gy[c2];
  const double giz = w * gz[c1] + w1 * gz[c2];
  const double ex = m.ex[f], ey = m.ey[f], ez = m.ez[f];
  const double s = (p[c2] - p[c1]) / m.delta[f] - (gix * ex + giy * ey + giz * ez);
  const double J = flux[f];
  const double fx = J * (gix + s * ex), fy = J * (giy + s * ey), fz = J * (giz + s * ez);
  corr[f] = fx * ntx[f] + fy * nty[f] + fz * ntz[f];
}

// b[owner] += correction, b[neighbour] -= correction.  The reference adds
// with atomics in no fixed order; here each cell adds its faces' corrections
// in its face-list order.
__global__ void __launch_bounds__(kBlock)
    k_nonorth_cell(MeshK m, const double* __restrict__ corr, double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  double acc = b[i];
  cell_faces(
      m.cf_ptr[i], m.cf_ptr[i + 1],
      [&](int32_t k) { return (double)m.cf_sign[k] * corr[m.cf_face[k]]; },
      [&](double t, bool ok) { acc = ok ? acc + t : acc; });
  b[i] = acc;
}

void nonorthogonal_correction(xcal_eqn* pe, const double* gx, const double* gy, const double* gz,
                              const double* p, const double* flux) {
  const xcal_mesh* m = pe->mesh;
  if (pe->ncflux.n < m->nf) pe->ncflux.alloc(m->nf);
  MeshK mk = meshk(m);
  LAUNCH(k_nonorth_face, m->nf - m->nbf, mk, m->ncw.p, m->ntx.p, m->nty.p, m->ntz.p, gx, gy, gz,
         p, flux, pe->ncflux.p);
  LAUNCH(k_nonorth_cell, m->n, mk, pe->ncflux.p, pe->bc(0));
}

void correct_velocity(const xcal_mesh* m, double* Ux, double* Uy, double* Uz, const double* Hx,
                      const double* Hy, const double* Hz, const double* gx, const double* gy,
                      const double* gz, const double* rD) {
  LAUNCH(k_correct_velocity, m->n, m->n, Ux, Uy, Uz, Hx, Hy, Hz, gx, gy, gz, rD);
}

}  // namespace xcal

using namespace xcal;

extern "C" {

int xcal_interpolate(xcal_field* face, const xcal_field* cell) {
  return guarded([&] {
    XCAL_CHECK(face && cell && face->mesh == cell->mesh, "fields must share a mesh");
    XCAL_CHECK((face->kind == XCAL_FACE_SCALAR && cell->kind == XCAL_CELL_SCALAR) ||
                   (face->kind == XCAL_FACE_VECTOR && cell->kind == XCAL_CELL_VECTOR),
               "interpolate!: field kinds do not match");
    for (int c = 0; c < cell->ncomp; ++c) interpolate(cell->mesh, cell->comp(c), face->comp(c));
    return 0;
  });
}

int xcal_correct_boundaries(xcal_field* face, const xcal_field* cell, const xcal_bc* bcs,
                            int32_t nbcs) {
  return guarded([&] {
    XCAL_CHECK(face && cell && face->mesh == cell->mesh, "fields must share a mesh");
    XCAL_CHECK((face->kind == XCAL_FACE_SCALAR && cell->kind == XCAL_CELL_SCALAR) ||
                   (face->kind == XCAL_FACE_VECTOR && cell->kind == XCAL_CELL_VECTOR),
               "correct_boundaries!: field kinds do not match");
    if (cell->ncomp == 3)
      for (int i = 0; i < nbcs; ++i)
        XCAL_CHECK(bcs[i].kind != XCAL_BC_NEUMANN,
                   "Neumann boundary condition for vector fields is not implemented yet.");
    BCsK b = make_bcs(cell->mesh, bcs, nbcs);
    const double* ph[3] = {cell->comp(0), cell->ncomp == 3 ? cell->comp(1) : nullptr,
                           cell->ncomp == 3 ? cell->comp(2) : nullptr};
    double* fc[3] = {face->comp(0), face->ncomp == 3 ? face->comp(1) : nullptr,
                     face->ncomp == 3 ? face->comp(2) : nullptr};
    correct_boundaries(cell->mesh, b, ph, fc, cell->ncomp);
    return 0;
  });
}

int xcal_green_gauss(xcal_field* grad, const xcal_field* face) {
  return guarded([&] {
    XCAL_CHECK(grad && face && grad->mesh == face->mesh, "fields must share a mesh");
    if (face->kind == XCAL_FACE_SCALAR && grad->kind == XCAL_CELL_VECTOR) {
      green_gauss(face->mesh, face->comp(0), grad->comp(0), grad->comp(1), grad->comp(2));
    } else if (face->kind == XCAL_FACE_VECTOR && grad->kind == XCAL_CELL_TENSOR) {
      green_gauss_tensor(face->mesh, face->comp(0), face->comp(1), face->comp(2), grad->comp(0));
    } else {
      throw Error(XCAL_ERR_ARG, "green_gauss!: scalar->vector or vector->tensor only");
    }
    return 0;
  });
}

int xcal_grad(xcal_field* grad, xcal_field* face, const xcal_field* cell, const xcal_bc* bcs,
              int32_t nbcs) {
  int rc = xcal_interpolate(face, cell);
  if (rc) return rc;
  rc = xcal_correct_boundaries(face, cell, bcs, nbcs);
  if (rc) return rc;
  return xcal_green_gauss(grad, face);
}

int xcal_grad_midpoint(xcal_field* grad, xcal_field* face, const xcal_field* cell,
                       const xcal_bc* bcs, int32_t nbcs) {
  return guarded([&] {
    XCAL_CHECK(grad && face && cell && (bcs || nbcs == 0), "NULL argument");
    xcal_mesh* m = cell->mesh;
    XCAL_CHECK(grad->mesh == m && face->mesh == m, "fields must share a mesh");
    const bool scalar = cell->kind == XCAL_CELL_SCALAR;
    XCAL_CHECK((scalar && face->kind == XCAL_FACE_SCALAR && grad->kind == XCAL_CELL_VECTOR) ||
                   (cell->kind == XCAL_CELL_VECTOR && face->kind == XCAL_FACE_VECTOR &&
                    grad->kind == XCAL_CELL_TENSOR),
               "grad!: ScalarField -> VectorField or VectorField -> TensorField");
    XCAL_CHECK(nbcs == m->n_patches, "one boundary condition per patch is required");
    for (int i = 0; i < nbcs && !scalar; ++i)
      XCAL_CHECK(bcs[i].kind != XCAL_BC_NEUMANN, "Neumann is not defined for vector fields");
    const BCsK b = make_bcs(m, bcs, nbcs);
    const int nc = scalar ? 1 : 3;
    const double* ph[3] = {cell->comp(0), nullptr, nullptr};
    double* fc[3] = {face->comp(0), nullptr, nullptr};
    for (int c = 1; c < nc; ++c) {
      ph[c] = cell->comp(c);
      fc[c] = face->comp(c);
    }
    grad_midpoint(m, b, ph, fc, nc, grad->comp(0));
    return 0;
  });
}

int xcal_div(xcal_field* cell, const xcal_field* face) {
  return guarded([&] {
    XCAL_CHECK(cell && face && cell->mesh == face->mesh, "fields must share a mesh");
    XCAL_CHECK(cell->kind == XCAL_CELL_SCALAR && face->kind == XCAL_FACE_SCALAR,
               "div!: FaceScalarField -> ScalarField");
    divergence(cell->mesh, face->comp(0), cell->comp(0));
    return 0;
  });
}

int xcal_flux(xcal_field* mdot, const xcal_field* face_vector) {
  return guarded([&] {
    XCAL_CHECK(mdot && face_vector && mdot->mesh == face_vector->mesh, "fields must share a mesh");
    XCAL_CHECK(mdot->kind == XCAL_FACE_SCALAR && face_vector->kind == XCAL_FACE_VECTOR,
               "flux!: FaceVectorField -> FaceScalarField");
    flux(mdot->mesh, face_vector->comp(0), face_vector->comp(1), face_vector->comp(2), mdot->comp(0));
    return 0;
  });
}

int xcal_relax_explicit(xcal_field* phi, const xcal_field* phi0, double alpha) {
  return guarded([&] {
    XCAL_CHECK(phi && phi0 && phi->mesh == phi0->mesh && phi->kind == phi0->kind,
               "explicit_relaxation!: fields must match");
    relax_explicit(phi->mesh, (int64_t)phi->ncomp * phi->len, phi->comp(0), phi0->comp(0), alpha);
    return 0;
  });
}

int xcal_inverse_diagonal(xcal_field* rD, const xcal_eqn* e) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    check_field(rD, e->mesh, XCAL_CELL_SCALAR, "rD");
    inverse_diagonal(e, rD->comp(0));
    return 0;
  });
}

int xcal_remove_pressure_source(xcal_eqn* e, const xcal_field* gradp, int32_t source_sign) {
  return guarded([&] {
    XCAL_CHECK(e && e->vector, "remove_pressure_source! applies to the momentum equation");
    check_field(gradp, e->mesh, XCAL_CELL_VECTOR, "gradp");
    remove_pressure_source(e, gradp->comp(0), gradp->comp(1), gradp->comp(2), source_sign);
    return 0;
  });
}

int xcal_H(xcal_field* Hv, const xcal_field* U, const xcal_eqn* e) {
  return guarded([&] {
    XCAL_CHECK(e && e->vector, "H! needs the momentum equation");
    check_field(Hv, e->mesh, XCAL_CELL_VECTOR, "Hv");
    check_field(U, e->mesh, XCAL_CELL_VECTOR, "U");
    Hop(e, U->comp(0), U->comp(1), U->comp(2), Hv->comp(0), Hv->comp(1), Hv->comp(2));
    return 0;
  });
}

int xcal_correct_mass_flux(xcal_field* mdot, const xcal_eqn* pe, const xcal_field* p,
                           const xcal_field* pflux, int32_t psign, const xcal_bc* bcs,
                           int32_t nbcs) {
  return guarded([&] {
    XCAL_CHECK(pe && !pe->vector, "correct_mass_flux! needs the pressure equation");
    check_field(mdot, pe->mesh, XCAL_FACE_SCALAR, "mdotf");
    check_field(p, pe->mesh, XCAL_CELL_SCALAR, "p");
    check_field(pflux, pe->mesh, XCAL_FACE_SCALAR, "rDf");
    BCsK b = make_bcs(pe->mesh, bcs, nbcs);
    correct_mass_flux(pe, p->comp(0), pflux->comp(0), psign, b, mdot->comp(0));
    return 0;
  });
}

int xcal_nonorthogonal_correction(xcal_eqn* pe, const xcal_field* gradp, const xcal_field* p,
                                  const xcal_field* flux) {
  return guarded([&] {
    XCAL_CHECK(pe && !pe->vector, "the correction applies to a scalar (pressure) equation");
    check_field(gradp, pe->mesh, XCAL_CELL_VECTOR, "grad");
    check_field(p, pe->mesh, XCAL_CELL_SCALAR, "p");
    check_field(flux, pe->mesh, XCAL_FACE_SCALAR, "flux");
    nonorthogonal_correction(pe, gradp->comp(0), gradp->comp(1), gradp->comp(2), p->comp(0),
                             flux->comp(0));
    return 0;
  });
}

int xcal_correct_velocity(xcal_field* U, const xcal_field* Hv, const xcal_field* gradp,
                          const xcal_field* rD) {
  return guarded([&] {
    XCAL_CHECK(U, "U is NULL");
    const xcal_mesh* m = U->mesh;
    check_field(U, m, XCAL_CELL_VECTOR, "U");
    check_field(Hv, m, XCAL_CELL_VECTOR, "Hv");
    check_field(gradp, m, XCAL_CELL_VECTOR, "gradp");
    check_field(rD, m, XCAL_CELL_SCALAR, "rD");
    correct_velocity(m, U->comp(0), U->comp(1), U->comp(2), Hv->comp(0), Hv->comp(1), Hv->comp(2),
                     gradp->comp(0), gradp->comp(1), gradp->comp(2), rD->comp(0));
    return 0;
  });
}

}  // extern "C"
