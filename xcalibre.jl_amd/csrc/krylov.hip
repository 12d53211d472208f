// Linear algebra of the SIMPLE hot path on gfx950: CSR SpMV, fused BLAS-1,
// deterministic device reductions, Jacobi and level-scheduled DILU
// preconditioners, and the CG / BiCGSTAB iterations that solve_system!
// (Solve_1_api.jl:268-307) runs through Krylov.jl.
//
// Krylov.jl (compat 0.10, Project.toml:49) is not vendored in the reference;
// the iterations below restate its cg!/bicgstab! as called by XCALibre:
// warm start x0 = current field (solve on the correction, add x0 at the end),
// preconditioner M applied as mulorldiv!(z, M, r, ldiv), BiCGSTAB with left
// preconditioning and shadow residual c = b, stopping when the monitored
// residual norm (sqrt(r'z) for CG, ||M^-1 r|| for BiCGSTAB) drops below
// atol + rtol * (its initial value), or itmax iterations.
#include <cmath>

#include "device_common.h"

namespace xcal {

// ---------------------------------------------------------------- SpMV
// One work-item per row; sequential accumulation in column order, as the
// reference's threaded xmul! (Multithread/spmvm.jl:44-55).
__global__ void __launch_bounds__(kBlock)
    k_spmv(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colval,
           const double* __restrict__ nz, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = xcd_row();
  if (i >= n) return;
  double acc = 0.0;
  for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) acc += nz[k] * x[colval[k]];
  y[i] = acc;
}

// r = b - A x
__global__ void __launch_bounds__(kBlock)
    k_residual_vec(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colval,
                   const double* __restrict__ nz, const double* __restrict__ x,
                   const double* __restrict__ b, double* __restrict__ r) {
  const int64_t i = xcd_row();
  if (i >= n) return;
  double acc = 0.0;
  for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) acc += nz[k] * x[colval[k]];
  r[i] = b[i] - acc;
}

// ----------------------------------------------------------- reductions
// Two dot products in one pass; partial sums per block, then the last block
// to finish (ticket) adds the partials in block order: the result depends
// only on n and the grid size, never on scheduling.
struct DotArgs {
  const double *a0, *b0, *a1, *b1;
};

__device__ inline void store_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double load_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void finish_reduction(double s0, double s1, int nval, double* partial,
                                        unsigned* ticket, double* result) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    store_agent(&partial[2 * blockIdx.x], s0);
    if (nval > 1) store_agent(&partial[2 * blockIdx.x + 1], s1);
    __threadfence();
    unsigned t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    double t0 = 0.0, t1 = 0.0;
    for (unsigned b = 0; b < gridDim.x; ++b) {
      t0 += load_agent(&partial[2 * b]);
      if (nval > 1) t1 += load_agent(&partial[2 * b + 1]);
    }
    result[0] = t0;
    if (nval > 1) result[1] = t1;
    *ticket = 0u;
    __threadfence();
  }
}

__global__ void __launch_bounds__(kBlock)
    k_dot2(int64_t n, DotArgs d, int nval, double* partial, unsigned* ticket, double* result) {
  __shared__ double lds[kBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    s0 += d.a0[i] * d.b0[i];
    if (nval > 1) s1 += d.a1[i] * d.b1[i];
  }
  s0 = block_sum(s0, lds);
  if (nval > 1) s1 = block_sum(s1, lds);
  finish_reduction(s0, s1, nval, partial, ticket, result);
}

// residual() of Solve_1_api.jl:441-464: sum (b - A x)^2 and sum b^2
__global__ void __launch_bounds__(kBlock)
    k_residual_sums(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colval,
                    const double* __restrict__ nz, const double* __restrict__ x,
                    const double* __restrict__ b, double* partial, unsigned* ticket, double* result) {
  __shared__ double lds[kBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) acc += nz[k] * x[colval[k]];
    const double r = b[i] - acc;
    s0 += r * r;
    s1 += b[i] * b[i];
  }
  s0 = block_sum(s0, lds);
  s1 = block_sum(s1, lds);
  finish_reduction(s0, s1, 2, partial, ticket, result);
}

// ------------------------------------------------------------- BLAS-1
__global__ void __launch_bounds__(kBlock)
    k_sub_from(int64_t n, const double* __restrict__ b, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}
// y = y + a*x   (kaxpy!)
__global__ void __launch_bounds__(kBlock)
    k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + a * x[i];
}
// x += a p ; r += (-a) q  (the CG update pair)
__global__ void __launch_bounds__(kBlock)
    k_cg_update(int64_t n, double a, const double* __restrict__ p, const double* __restrict__ q,
                double* __restrict__ x, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = x[i] + a * p[i];
  r[i] = r[i] + (-a) * q[i];
}
// y = x + b*y  (kaxpby!(1, x, b, y))
__global__ void __launch_bounds__(kBlock)
    k_xpby(int64_t n, const double* __restrict__ x, double b, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + b * y[i];
}
// out = u + a*v
__global__ void __launch_bounds__(kBlock)
    k_waxpy(int64_t n, const double* __restrict__ u, double a, const double* __restrict__ v,
            double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = u[i] + a * v[i];
}
// BiCGSTAB tail: x += w s; r = s + (-w) t; p = p + (-w) v; (p = r + beta p later)
__global__ void __launch_bounds__(kBlock)
    k_bicg_update(int64_t n, double w, const double* __restrict__ s, const double* __restrict__ t,
                  const double* __restrict__ v, double* __restrict__ x, double* __restrict__ r,
                  double* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = x[i] + w * s[i];
  r[i] = s[i] + (-w) * t[i];
  p[i] = p[i] + (-w) * v[i];
}
__global__ void __launch_bounds__(kBlock)
    k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// ------------------------------------------------------ preconditioners
// update_Jacobi! (preconditioners_2_functions.jl:86-94)
__global__ void __launch_bounds__(kBlock)
    k_jacobi_update(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ nz,
                    double* __restrict__ S) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) S[i] = 1.0 / fabs(nz[diag[i]]);
}
// opDiagonal(S) * r
__global__ void __launch_bounds__(kBlock)
    k_jacobi_apply(int64_t n, const double* __restrict__ S, const double* __restrict__ r,
                   double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = S[i] * r[i];
}

// extract_diagonal! (preconditioners_1_DILU.jl:17-23)
__global__ void __launch_bounds__(kBlock)
    k_extract_diag(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ nz,
                   double* __restrict__ D) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) D[i] = nz[diag[i]];
}
// update_dilu_diagonal! (preconditioners_1_DILU.jl:43-75) in pull form for
// the rows of one forward level: D_j -= (a_ij a_ji)/D_i over i<j in
// ascending i -- the order in which the reference's outer i loop reaches j.
__global__ void __launch_bounds__(kBlock)
    k_dilu_diag_level(const int32_t* __restrict__ rows, int32_t cnt,
                      const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colval,
                      const int32_t* __restrict__ diag, const int32_t* __restrict__ trans,
                      const double* __restrict__ nz, double* __restrict__ D) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int j = rows[t];
  double Dj = D[j];
  for (int32_t k = rowptr[j]; k < diag[j]; ++k) {
    const int i = colval[k];
    const int kt = trans[k];
    if (kt >= 0) Dj = Dj - (nz[kt] * nz[k]) / D[i];
  }
  D[j] = Dj;
}
__global__ void __launch_bounds__(kBlock)
    k_recip(int64_t n, const double* __restrict__ D, double* __restrict__ rD) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rD[i] = 1.0 / D[i];
}
// forward_substitution! (preconditioners_1_DILU.jl:84-98), one level
__global__ void __launch_bounds__(kBlock)
    k_dilu_fwd(const int32_t* __restrict__ rows, int32_t cnt, const int32_t* __restrict__ rowptr,
               const int32_t* __restrict__ colval, const int32_t* __restrict__ diag,
               const double* __restrict__ nz, const double* __restrict__ rD,
               const double* __restrict__ r, double* __restrict__ z) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  double sum = 0.0;
  for (int32_t k = rowptr[i]; k < diag[i]; ++k) {
    const int j = colval[k];
    sum += (nz[k] * z[j]) * rD[j];
  }
  z[i] = r[i] - sum;
}
// backward_substitution! (preconditioners_1_DILU.jl:100-114), one level
__global__ void __launch_bounds__(kBlock)
    k_dilu_bwd(const int32_t* __restrict__ rows, int32_t cnt, const int32_t* __restrict__ rowptr,
               const int32_t* __restrict__ colval, const int32_t* __restrict__ diag,
               const double* __restrict__ nz, const double* __restrict__ D, double* __restrict__ z) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  double sum = 0.0;
  for (int32_t k = diag[i] + 1; k < rowptr[i + 1]; ++k) sum += nz[k] * z[colval[k]];
  z[i] = (z[i] - sum) / D[i];
}

// ----------------------------------------------------------- host side
struct Ctx {
  const xcal_mesh* m;
  hipStream_t s;
  int64_t n;
};

static void launch_check() { XCAL_HIP(hipGetLastError()); }

void spmv(const xcal_eqn* e, const double* x, double* y) {
  const xcal_mesh* m = e->mesh;
  hipLaunchKernelGGL(k_spmv, dim3(grid_for(m->n)), dim3(kBlock), 0, m->ctx->stream, m->n,
                     m->rowptr.p, m->colval.p, e->nzval.p, x, y);
  launch_check();
}

static int red_grid(int64_t n) { return std::min<int>(grid_for(n), kRedBlocks); }

// returns {a0.b0, a1.b1} (host values; synchronises)
[[maybe_unused]] static void dots(const xcal_eqn* e, const double* a0, const double* b0, const double* a1,
                 const double* b1, double* out) {
  const xcal_mesh* m = e->mesh;
  xcal_ctx* c = m->ctx;
  int nval = a1 ? 2 : 1;
  hipLaunchKernelGGL(k_dot2, dim3(red_grid(m->n)), dim3(kBlock), 0, c->stream, m->n,
                     DotArgs{a0, b0, a1, b1}, nval, c->red_partial.p, c->red_ticket.p,
                     c->red_result.p);
  launch_check();
  XCAL_HIP(hipMemcpyAsync(c->host_pinned, c->red_result.p, sizeof(double) * 2,
                          hipMemcpyDeviceToHost, c->stream));
  XCAL_HIP(hipStreamSynchronize(c->stream));
  out[0] = c->host_pinned[0];
  if (nval > 1) out[1] = c->host_pinned[1];
}

double residual(xcal_eqn* e, const double* x, const double* b) {
  const xcal_mesh* m = e->mesh;
  xcal_ctx* c = m->ctx;
  hipLaunchKernelGGL(k_residual_sums, dim3(red_grid(m->n)), dim3(kBlock), 0, c->stream, m->n,
                     m->rowptr.p, m->colval.p, e->nzval.p, x, b, c->red_partial.p, c->red_ticket.p,
                     c->red_result.p);
  launch_check();
  XCAL_HIP(hipMemcpyAsync(c->host_pinned, c->red_result.p, sizeof(double) * 2,
                          hipMemcpyDeviceToHost, c->stream));
  XCAL_HIP(hipStreamSynchronize(c->stream));
  const double sumR = c->host_pinned[0];
  const double normb = std::sqrt(c->host_pinned[1]);
  const double a = std::fabs(normb);
  const double eps_nb = std::nextafter(a, INFINITY) - a;
  const double den = normb > eps_nb ? normb : 1.0;
  return std::sqrt(sumR) / den;
}

void precond_update(xcal_eqn* e, int precon) {
  xcal_mesh* m = e->mesh;
  hipStream_t s = m->ctx->stream;
  const int64_t n = m->n;
  e->precon = precon;
  if (precon == XCAL_PRECON_JACOBI) {
    if (e->pS.n != n) e->pS.alloc(n);
    hipLaunchKernelGGL(k_jacobi_update, dim3(grid_for(n)), dim3(kBlock), 0, s, n, m->diag.p,
                       e->nzval.p, e->pS.p);
    launch_check();
  } else if (precon == XCAL_PRECON_DILU) {
    if (e->pD.n != n) {
      e->pD.alloc(n);
      e->prD.alloc(n);
    }
    hipLaunchKernelGGL(k_extract_diag, dim3(grid_for(n)), dim3(kBlock), 0, s, n, m->diag.p,
                       e->nzval.p, e->pD.p);
    launch_check();
    for (int l = 1; l < m->nlev_f; ++l) {  // level 0 rows have no lower neighbours
      int32_t a = m->lev_ptr_f[l], cnt = m->lev_ptr_f[l + 1] - a;
      hipLaunchKernelGGL(k_dilu_diag_level, dim3(grid_for(cnt)), dim3(kBlock), 0, s,
                         m->lev_rows_f.p + a, cnt, m->rowptr.p, m->colval.p, m->diag.p,
                         m->trans.p, e->nzval.p, e->pD.p);
      launch_check();
    }
    hipLaunchKernelGGL(k_recip, dim3(grid_for(n)), dim3(kBlock), 0, s, n, e->pD.p, e->prD.p);
    launch_check();
  } else if (precon != XCAL_PRECON_NONE) {
    throw Error(XCAL_ERR_UNSUPPORTED, "unknown preconditioner");
  }
}

// mulorldiv!(z, M, r, ldiv)
void apply_M_levels(const xcal_eqn* e, const double* r, double* z) {
  const xcal_mesh* m = e->mesh;
  hipStream_t s = m->ctx->stream;
  const int64_t n = m->n;
  if (e->precon == XCAL_PRECON_JACOBI) {
    hipLaunchKernelGGL(k_jacobi_apply, dim3(grid_for(n)), dim3(kBlock), 0, s, n, e->pS.p, r, z);
    launch_check();
  } else if (e->precon == XCAL_PRECON_DILU) {
    for (int l = 0; l < m->nlev_f; ++l) {
      int32_t a = m->lev_ptr_f[l], cnt = m->lev_ptr_f[l + 1] - a;
      hipLaunchKernelGGL(k_dilu_fwd, dim3(grid_for(cnt)), dim3(kBlock), 0, s, m->lev_rows_f.p + a,
                         cnt, m->rowptr.p, m->colval.p, m->diag.p, e->nzval.p, e->prD.p, r, z);
    }
    for (int l = 0; l < m->nlev_b; ++l) {
      int32_t a = m->lev_ptr_b[l], cnt = m->lev_ptr_b[l + 1] - a;
      hipLaunchKernelGGL(k_dilu_bwd, dim3(grid_for(cnt)), dim3(kBlock), 0, s, m->lev_rows_b.p + a,
                         cnt, m->rowptr.p, m->colval.p, m->diag.p, e->nzval.p, e->pD.p, z);
    }
    launch_check();
  } else {
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r, z);
    launch_check();
  }
}

}  // namespace xcal

using namespace xcal;

extern "C" {

int xcal_precond_update(xcal_eqn* e, int32_t precon) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    precond_update(e, precon);
    return 0;
  });
}

static const double* cell_comp(const xcal_eqn* e, const xcal_field* f, int c, const char* what) {
  XCAL_CHECK(f && f->mesh == e->mesh, std::string(what) + ": field missing or on another mesh");
  XCAL_CHECK(f->kind == XCAL_CELL_SCALAR || f->kind == XCAL_CELL_VECTOR,
             std::string(what) + ": must be a cell field");
  XCAL_CHECK(c >= 0 && c < f->ncomp, std::string(what) + ": component out of range");
  return f->comp(c);
}

int xcal_spmv(xcal_eqn* e, const xcal_field* x, int32_t xc, xcal_field* y, int32_t yc) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    const double* xp = cell_comp(e, x, xc, "x");
    double* yp = const_cast<double*>(cell_comp(e, y, yc, "y"));
    XCAL_CHECK(xp != yp, "x and y must not alias");
    spmv(e, xp, yp);
    return 0;
  });
}

}  // extern "C"
