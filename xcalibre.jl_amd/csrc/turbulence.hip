// Kernels of the k-omega model (RANS_kOmega.jl:175-235, RANS_functions.jl):
// production from the strain rate, wall-function production and omega
// constraint, model coefficients, bound!, nut = k/omega, the nut wall
// function and nueff.  Face loops over wall-function patches run per
// boundary cell in ascending face order, so where a cell has several wall
// faces the last one wins, as in the reference's sequential order.
#include "device_common.h"

namespace xcal {

// magnitude2!(Pk, S, scale 2) with S = (gradU + gradU')/2 (Fields_0_types.jl:336-339,
// RANS_tensor_algebra.jl:68-98), then Pw = (rho a1) Pk and Pk = (rho nut) Pk
// (RANS_kOmega.jl:196-197), rho = 1.  gradU[3p+q] holds d U_p / d x_q.
__global__ void __launch_bounds__(kBlock)
    k_production(int64_t n, const double* __restrict__ g, const double* __restrict__ nut,
                 double alpha1, double* __restrict__ Pk, double* __restrict__ Pw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double G[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) G[c] = g[(int64_t)c * n + i];
  double sum = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double s = 0.5 * (G[3 * j + k] + G[3 * k + j]);
      sum += s * s;
    }
  const double P = sum * 2.0;
  Pw[i] = (1.0 * alpha1) * P;
  Pk[i] = (1.0 * nut[i]) * P;
}

struct WallK {
  double kappa, beta1, cmu, cmu25, E, yPlusLam, nu;
};

// y_plus, nut_wall (RANS_functions.jl:40-60)
__device__ inline double wf_yplus(const WallK& w, double k, double y) {
  return ((w.cmu25 * y) * sqrt(k)) / w.nu;
}
__device__ inline double wf_nutw(const WallK& w, double yplus) {
  return jmax(w.nu * ((yplus * w.kappa) / log(jmax(w.E * yplus, 1.0 + 1e-4)) - 1.0), 0.0);
}

// set_production! for KWallFunction patches (RANS_functions.jl:89-124), Uw = 0
__global__ void __launch_bounds__(kBlock)
    k_wall_production(MeshK m, BCsK B, WallK w, const double* __restrict__ Ux,
                      const double* __restrict__ Uy, const double* __restrict__ Uz,
                      const double* __restrict__ k, double* __restrict__ Pk) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m.nbcell) return;
  const int c = m.bc_cells[s];
  for (int32_t j = m.bc_ptr[s]; j < m.bc_ptr[s + 1]; ++j) {
    const int f = m.bc_face[j];
    const int p = m.bface_patch[f];
    if (p < 0 || B.p[p].kind != XCAL_BC_KWALL) continue;
    const double delta = m.delta[f];
    const double uStar = w.cmu25 * sqrt(k[c]);
    const double dUdy = uStar / (w.kappa * delta);
    const double yplus = wf_yplus(w, k[c], delta);
    const double nutw = wf_nutw(w, yplus);
    const double nx = m.nx[f], ny = m.ny[f], nz = m.nz[f];
    const double dx = Ux[c] - 0.0, dy = Uy[c] - 0.0, dz = Uz[c] - 0.0;
    const double dn = dx * nx + dy * ny + dz * nz;
    const double gx = (dx - dn * nx) / delta, gy = (dy - dn * ny) / delta,
                 gz = (dz - dn * nz) / delta;
    const double mag = sqrt(gx * gx + gy * gy + gz * gz);
    Pk[c] = yplus > w.yPlusLam ? ((w.nu + nutw) * mag) * dUdy : 0.0;
  }
}

// constrain! for OmegaWallFunction patches (RANS_functions.jl:302-356): the
// row of a wall cell becomes the identity and b = omega_wall
__global__ void __launch_bounds__(kBlock)
    k_omega_constrain(MeshK m, BCsK B, WallK w, const int32_t* __restrict__ rb,
                      const int32_t* __restrict__ re, const int32_t* __restrict__ diag,
                      const double* __restrict__ k, double* __restrict__ nz, double* __restrict__ b) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m.nbcell) return;
  const int c = m.bc_cells[s];
  bool hit = false;
  double wc = 0.0;
  for (int32_t j = m.bc_ptr[s]; j < m.bc_ptr[s + 1]; ++j) {
    const int f = m.bc_face[j];
    const int p = m.bface_patch[f];
    if (p < 0 || B.p[p].kind != XCAL_BC_OMEGAWALL) continue;
    const double y = m.delta[f];
    const double wvis = (6.0 * w.nu) / (w.beta1 * (y * y));
    const double wlog = sqrt(k[c]) / ((w.cmu25 * w.kappa) * y);
    const double yplus = wf_yplus(w, k[c], y);
    wc = yplus > w.yPlusLam ? wlog : wvis;
    hit = true;
  }
  if (!hit) return;
  for (int32_t q = rb[c]; q < re[c]; q += kSlice) nz[q] = 0.0;
  nz[diag[c]] = 1.0;
  b[c] = wc;
}

// Dw = (rho b1) omega, Dk = (rho b*) omega  (RANS_kOmega.jl:199,201)
__global__ void __launch_bounds__(kBlock)
    k_turb_cell_coeffs(int64_t n, const double* __restrict__ omega, double beta1, double beta_star,
                       double* __restrict__ Dw, double* __restrict__ Dk) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Dw[i] = (1.0 * beta1) * omega[i];
  Dk[i] = (1.0 * beta_star) * omega[i];
}

// mueff = rhof (nuf + sigma nutf)  (RANS_kOmega.jl:200,202)
__global__ void __launch_bounds__(kBlock)
    k_turb_face_coeffs(int64_t nf, const double* __restrict__ nutf, double nu, double sigma_w,
                       double sigma_k, double* __restrict__ mw, double* __restrict__ mk) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  mw[f] = 1.0 * (nu + sigma_w * nutf[f]);
  mk[f] = 1.0 * (nu + sigma_k * nutf[f]);
}

// bound! (RANS_functions.jl:2-38), every cell from the values before the call
__global__ void __launch_bounds__(kBlock)
    k_bound(MeshK m, const double* __restrict__ v0, double* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  const double mzero = 2.220446049250313e-16;
  double sum = 0.0;
  int cnt = 0;
  for (int32_t q = m.cf_ptr[i]; q < m.cf_ptr[i + 1]; ++q) {
    sum += jmax(v0[m.cf_nbr[q]], mzero);
    ++cnt;
  }
  const double average = sum / (double)cnt;
  const double vi = v0[i];
  const double a = __builtin_signbit(vi) ? average : copysign(0.0, average);
  v[i] = jmax(jmax(vi, a), mzero);
}

// nut = k / omega
__global__ void __launch_bounds__(kBlock)
    k_nut(int64_t n, const double* __restrict__ k, const double* __restrict__ omega,
          double* __restrict__ nut) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) nut[i] = k[i] / omega[i];
}

// correct_nut_wall! for NutWallFunction patches (RANS_functions.jl:166-200)
__global__ void __launch_bounds__(kBlock)
    k_nut_wall(MeshK m, BCsK B, WallK w, const double* __restrict__ k, double* __restrict__ nutf) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nbf) return;
  const int p = m.bface_patch[f];
  if (p < 0 || B.p[p].kind != XCAL_BC_NUTWALL) return;
  const int c = m.own[f];
  const double yplus = wf_yplus(w, k[c], m.delta[f]);
  nutf[f] = yplus > w.yPlusLam ? wf_nutw(w, yplus) : 0.0;
}

// update_nueff! turbulent (Solvers_0_functions.jl:31-36)
__global__ void __launch_bounds__(kBlock)
    k_nueff_turbulent(int64_t nf, double nu, const double* __restrict__ nutf,
                      double* __restrict__ nueff) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf) nueff[f] = nu + nutf[f];
}

// y_plus_laminar (RANS_functions.jl:40-43)
double y_plus_laminar(double E, double kappa) {
  double yL = 11.0;
  for (int i = 0; i < 10; ++i) yL = std::log(std::max(yL * E, 1.0)) / kappa;
  return yL;
}

static WallK wallk(double nu) {
  WallK w;
  w.kappa = 0.41;
  w.beta1 = 0.075;
  w.cmu = 0.09;
  w.cmu25 = std::pow(0.09, 0.25);
  w.E = 9.8;
  w.yPlusLam = y_plus_laminar(w.E, w.kappa);
  w.nu = nu;
  return w;
}

#define TLAUNCH(kern, n, ...)                                                               \
  do {                                                                                      \
    if ((n) > 0)                                                                            \
      hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(kBlock), 0, m->ctx->stream, __VA_ARGS__); \
    XCAL_HIP(hipGetLastError());                                                            \
  } while (0)

void turb_production(const xcal_mesh* m, const double* gradU, const double* nut, double alpha1,
                     double* Pk, double* Pw) {
  TLAUNCH(k_production, m->n, m->n, gradU, nut, alpha1, Pk, Pw);
}
void turb_wall_production(const xcal_mesh* m, const BCsK& bk, double nu, const double* Ux,
                          const double* Uy, const double* Uz, const double* k, double* Pk) {
  TLAUNCH(k_wall_production, m->nbcell, meshk(m), bk, wallk(nu), Ux, Uy, Uz, k, Pk);
}
void turb_omega_constrain(xcal_eqn* e, const BCsK& bw, double nu, const double* k) {
  const xcal_mesh* m = e->mesh;
  TLAUNCH(k_omega_constrain, m->nbcell, meshk(m), bw, wallk(nu), m->rbeg.p, m->rend.p, m->diag.p,
          k, e->nzval.p, e->bc(0));
}
void turb_coeffs(const xcal_mesh* m, const double* omega, const double* nutf, double nu,
                 double beta1, double beta_star, double sigma_w, double sigma_k, double* Dw,
                 double* Dk, double* mw, double* mk) {
  TLAUNCH(k_turb_cell_coeffs, m->n, m->n, omega, beta1, beta_star, Dw, Dk);
  TLAUNCH(k_turb_face_coeffs, m->nf, m->nf, nutf, nu, sigma_w, sigma_k, mw, mk);
}
void turb_bound(const xcal_mesh* m, double* v, double* tmp) {
  XCAL_HIP(hipMemcpyAsync(tmp, v, sizeof(double) * m->n, hipMemcpyDeviceToDevice, m->ctx->stream));
  TLAUNCH(k_bound, m->n, meshk(m), tmp, v);
}
void turb_nut(const xcal_mesh* m, const double* k, const double* omega, double* nut) {
  TLAUNCH(k_nut, m->n, m->n, k, omega, nut);
}
void turb_nut_wall(const xcal_mesh* m, const BCsK& bn, double nu, const double* k, double* nutf) {
  TLAUNCH(k_nut_wall, m->nbf, meshk(m), bn, wallk(nu), k, nutf);
}
void turb_nueff(const xcal_mesh* m, double nu, const double* nutf, double* nueff) {
  TLAUNCH(k_nueff_turbulent, m->nf, m->nf, nu, nutf, nueff);
}

}  // namespace xcal
