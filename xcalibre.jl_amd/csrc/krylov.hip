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
#include "sweep_cone.h"
#include "sweep_pipe.h"

namespace xcal {

// ---------------------------------------------------------------- SpMV
// One work-item per row; sequential accumulation in column order, as the
// reference's threaded xmul! (Multithread/spmvm.jl:44-55).  The sliced-ELL
// storage makes each of a wavefront's loads of entry t one contiguous
// segment.
__global__ void __launch_bounds__(kBlock)
    k_spmv(MatK A, const double* __restrict__ x, double* __restrict__ y) {
  int64_t end;
  const int64_t i = block_row(A, end);
  if (i >= end) return;
  y[i] = row_sum(A, i, [&](int j) { return x[j]; });
}

// ----------------------------------------------------------- reductions
// residual() of Solve_1_api.jl:441-464: per-workgroup partial sums of
// (b - A x)^2 and b^2, then one workgroup adds them in slot order, so the
// result depends only on n, never on scheduling.
__global__ void __launch_bounds__(kBlock)
    k_residual_sums(MatK A, const double* __restrict__ x, const double* __restrict__ b,
                    double* __restrict__ partial) {
  __shared__ double lds[kBlock / 64];
  int64_t end;
  const int64_t i = block_row(A, end);
  double s0 = 0.0, s1 = 0.0;
  if (i < end) {
    const double acc = row_sum(A, i, [&](int j) { return x[j]; });
    const double r = b[i] - acc;
    s0 = r * r;
    s1 = b[i] * b[i];
  }
  s0 = block_sum(s0, lds);
  s1 = block_sum(s1, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s0;
    partial[2 * blockIdx.x + 1] = s1;
  }
}

__global__ void __launch_bounds__(kRunBlock)
    k_sum_partials(const double* __restrict__ partial, int count, double* __restrict__ out) {
  __shared__ double lds[kRunBlock / 64];
  double t0 = 0.0, t1 = 0.0;
  for (int b = threadIdx.x; b < count; b += blockDim.x) {
    t0 += partial[2 * b];
    t1 += partial[2 * b + 1];
  }
  t0 = block_sum(t0, lds);
  t1 = block_sum(t1, lds);
  if (threadIdx.x == 0) {
    out[0] = t0;
    out[1] = t1;
  }
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
// row j: D_j -= (a_ij a_ji)/D_i over i<j in ascending i -- the order in which
// the reference's outer i loop reaches j.  Forward levels are contiguous row
// ranges, so a level is rows [row0, row0 + cnt).
__device__ inline void dilu_diag_row(const MatK& A, const int32_t* __restrict__ trans, double* D,
                                     int j) {
  double Dj = D[j];
  for (int32_t k = A.rb[j]; k < A.diag[j]; k += kSlice) {
    const int i = A.colval[k];
    const int kt = trans[k];
    if (kt >= 0) Dj = Dj - (A.nz[kt] * A.nz[k]) / D[i];
  }
  D[j] = Dj;
}
__global__ void __launch_bounds__(kBlock)
    k_dilu_diag_level(MatK A, const int32_t* __restrict__ trans, double* D, int32_t row0,
                      int32_t cnt) {
  int32_t end;
  const int32_t t = level_pos(cnt, end);
  if (t < end) dilu_diag_row(A, trans, D, row0 + t);
}
__global__ void __launch_bounds__(kRunBlock)
    k_dilu_diag_run(MatK A, const int32_t* __restrict__ trans, double* D,
                    const int32_t* __restrict__ lev_ptr, int lev0, int lev1) {
  for (int l = lev0; l < lev1; ++l) {
    const int32_t b = lev_ptr[l + 1];
    for (int32_t j = lev_ptr[l] + threadIdx.x; j < b; j += blockDim.x) dilu_diag_row(A, trans, D, j);
    __syncthreads();
  }
}
// update_dilu_diagonal! in the cone form (sweep_cone.h), one launch per band
template <int WK, int P>
__global__ void __launch_bounds__(kConeBlock)
    k_dilu_diag_cone(ConeK C, const double* __restrict__ prod, const double* __restrict__ ajj,
                     double* D) {
  __shared__ ConeLDS<double> L;
  diag_cone<WK, P>(C, prod, ajj, D, L);
}
// a_ji a_ij per cone entry (0 where the transpose is absent, which a face
// matrix's symmetric pattern never has), and a_ij itself
__global__ void __launch_bounds__(kBlock)
    k_cone_products(int64_t n, const int32_t* __restrict__ slot, const int32_t* __restrict__ trans,
                    const double* __restrict__ nz, double* __restrict__ out,
                    double* __restrict__ a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int32_t k = slot[e], kt = trans[k];
  const double aij = nz[k];
  out[e] = kt >= 0 ? nz[kt] * aij : 0.0;
  a[e] = aij;
}

// update_dilu_diagonal! over a run of small levels, pipelined (sweep_pipe.h)
template <bool FAR, int WK>
__global__ void __launch_bounds__(kPipeBlock)
    k_dilu_diag_run_pipe(double* D, const double* __restrict__ prod,
                         const int32_t* __restrict__ Lcol, const int32_t* __restrict__ lev_ptr,
                         const int32_t* __restrict__ meta, int lev0, int lev1) {
  __shared__ PipeLDS<double> L;
  diag_pipe<FAR, WK, pipe_depth<WK>()>(D, prod, Lcol, lev_ptr, meta, lev0, lev1 - lev0, true, L);
}

// out[e] = v[col[e]] (0 on padding)
__global__ void __launch_bounds__(kBlock)
    k_gather_cols(int64_t total, const int32_t* __restrict__ col, const double* __restrict__ v,
                  double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total) out[e] = col[e] >= 0 ? v[col[e]] : 0.0;
}
// values of the packed triangular factors: out[e] = nz[slot[e]] (0 on padding)
__global__ void __launch_bounds__(kBlock)
    k_gather_slots(int64_t total, const int32_t* __restrict__ slot, const double* __restrict__ nz,
                   double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total) out[e] = slot[e] >= 0 ? nz[slot[e]] : 0.0;
}
// Db[p] = D[rows_b[p]]
__global__ void __launch_bounds__(kBlock)
    k_gather_rows(int64_t n, const int32_t* __restrict__ rows, const double* __restrict__ D,
                  double* __restrict__ Db) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) Db[p] = D[rows[p]];
}
// out[p] = D of cone position p's row (crow holds row or ~row)
__global__ void __launch_bounds__(kBlock)
    k_gather_cone_rows(int64_t total, const int32_t* __restrict__ crow,
                       const double* __restrict__ D, double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total) return;
  const int32_t c = crow[p];
  out[p] = D[c >= 0 ? c : ~c];
}
__global__ void __launch_bounds__(kBlock)
    k_recip(int64_t n, const double* __restrict__ D, double* __restrict__ rD) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rD[i] = 1.0 / D[i];
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
  hipLaunchKernelGGL(k_spmv, dim3(m->nxblk), dim3(kBlock), 0, m->ctx->stream,
                     matk(m, e->nzval.p), x, y);
  launch_check();
}

double residual(xcal_eqn* e, const double* x, const double* b) {
  const xcal_mesh* m = e->mesh;
  xcal_ctx* c = m->ctx;
  const int grid = m->nxblk;
  if (e->kpart.n < 2 * (int64_t)grid) {
    XCAL_HIP(hipStreamSynchronize(c->stream));  // the old buffer may still be in use
    e->kpart.alloc(2 * (int64_t)grid);
  }
  hipLaunchKernelGGL(k_residual_sums, dim3(grid), dim3(kBlock), 0, c->stream, matk(m, e->nzval.p),
                     x, b, e->kpart.p);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRunBlock), 0, c->stream, e->kpart.p, grid,
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
    const MatK A = matk(m, e->nzval.p);
    if (!m->batch && m->pipe_runs) {  // a_ij a_ji per packed lower entry
      if (e->Lprod.n != m->Lslot.n) e->Lprod.alloc(m->Lslot.n);
      hipLaunchKernelGGL(k_lower_products, dim3(grid_for(m->Lslot.n)), dim3(kBlock), 0, s,
                         m->Lslot.n, m->Lslot.p, m->trans.p, e->nzval.p, e->Lprod.p);
    }
    if (m->batch) precond_update_batch(e);  // one workgroup per case
    else if (m->cone) {  // one launch per forward band; a_jj and a_ji a_ij in cone order
      const xcal_mesh::Cone& F = m->cone_f;
      if (e->cfd.n != F.rows) e->cfd.alloc(F.rows);  // a_jj here, 1/D after the update
      if (e->cfa.n != F.entries) {  // cfr: products here, 1/D_j after
        e->cfa.alloc(F.entries);
        e->cfr.alloc(F.entries);
      }
      hipLaunchKernelGGL(k_gather_cone_rows, dim3(grid_for(F.rows)), dim3(kBlock), 0, s, F.rows,
                         F.crow.p, e->pD.p, e->cfd.p);
      hipLaunchKernelGGL(k_cone_products, dim3(grid_for(F.entries)), dim3(kBlock), 0, s,
                         F.entries, F.slot.p, m->trans.p, e->nzval.p, e->cfr.p, e->cfa.p);
      with_cone_form(F.wk, !m->cone_deep, [&](auto WK, auto P) {
        for (size_t k = 0; k + 1 < F.band.size(); ++k)
          hipLaunchKernelGGL((k_dilu_diag_cone<decltype(WK)::value, decltype(P)::value>),
                             dim3(F.band[k + 1] - F.band[k]), dim3(kConeBlock), 0, s,
                             cone_k(m, F, F.band[k]), (const double*)e->cfr.p,
                             (const double*)e->cfd.p, e->pD.p);
      });
      launch_check();
    } else
    for (const SweepStep& st : m->plan_f) {
      if (st.lev1 == 1) continue;  // level 0 rows have no lower neighbours
      if (st.grid == 1 && st.pipe > 0) {
        auto pipe = [&](auto kern) {
          hipLaunchKernelGGL(kern, dim3(1), dim3(kPipeBlock), 0, s, e->pD.p,
                             (const double*)e->Lprod.p, m->Lcol.p, m->lev_ptr_f_d.p,
                             m->lmeta_f_d.p, st.lev0, st.lev1);
        };
        with_pipe_form(st.pipe, st.far, [&](auto WK, auto FAR) {
          pipe(k_dilu_diag_run_pipe<decltype(FAR)::value, decltype(WK)::value>);
        });
      } else if (st.grid == 1)
        hipLaunchKernelGGL(k_dilu_diag_run, dim3(1), dim3(kRunBlock), 0, s, A, m->trans.p, e->pD.p,
                           m->lev_ptr_f_d.p, st.lev0, st.lev1);
      else
        hipLaunchKernelGGL(k_dilu_diag_level, dim3(st.grid), dim3(kBlock), 0, s, A, m->trans.p,
                           e->pD.p, st.row0, st.row1 - st.row0);
      launch_check();
    }
    hipLaunchKernelGGL(k_recip, dim3(grid_for(n)), dim3(kBlock), 0, s, n, e->pD.p, e->prD.p);
    if (m->cone) {  // the cone sweeps' factors: a_ij, 1/D_j per entry; 1/D, D per position
      const xcal_mesh::Cone &F = m->cone_f, &B = m->cone_b;
      if (e->cfa.n != F.entries) {
        e->cfa.alloc(F.entries);
        e->cfr.alloc(F.entries);
      }
      if (e->cba.n != B.entries) e->cba.alloc(B.entries);
      if (e->cbd.n != B.rows) e->cbd.alloc(B.rows);
      // cfa (a_ij per forward entry) was gathered with the diagonal update's products
      hipLaunchKernelGGL(k_gather_cols, dim3(grid_for(F.entries)), dim3(kBlock), 0, s, F.entries,
                         F.dep.p, e->prD.p, e->cfr.p);
      hipLaunchKernelGGL(k_gather_slots, dim3(grid_for(B.entries)), dim3(kBlock), 0, s, B.entries,
                         B.slot.p, e->nzval.p, e->cba.p);
      hipLaunchKernelGGL(k_gather_cone_rows, dim3(grid_for(B.rows)), dim3(kBlock), 0, s, B.rows,
                         B.crow.p, e->pD.p, e->cbd.p);
      if (e->cfd.n != F.rows) e->cfd.alloc(F.rows);
      hipLaunchKernelGGL(k_gather_cone_rows, dim3(grid_for(F.rows)), dim3(kBlock), 0, s, F.rows,
                         F.crow.p, e->prD.p, e->cfd.p);
      launch_check();
      return;
    }
    if (m->pipe_wf > 0 || m->pipe_runs) {  // 1/D per packed lower entry, for the pipelined sweeps
      if (e->Lrd.n != m->Lcol.n) e->Lrd.alloc(m->Lcol.n);
      if (m->Lcol.n > 0)
        hipLaunchKernelGGL(k_gather_cols, dim3(grid_for(m->Lcol.n)), dim3(kBlock), 0, s, m->Lcol.n,
                           m->Lcol.p, e->prD.p, e->Lrd.p);
    }
    if (e->Lnz.n != m->Lslot.n) e->Lnz.alloc(m->Lslot.n);
    if (e->Unz.n != m->Uslot.n) e->Unz.alloc(m->Uslot.n);
    if (e->Db.n != n) e->Db.alloc(n);
    if (m->Lslot.n > 0)
      hipLaunchKernelGGL(k_gather_slots, dim3(grid_for(m->Lslot.n)), dim3(kBlock), 0, s, m->Lslot.n,
                         m->Lslot.p, e->nzval.p, e->Lnz.p);
    if (m->Uslot.n > 0)
      hipLaunchKernelGGL(k_gather_slots, dim3(grid_for(m->Uslot.n)), dim3(kBlock), 0, s, m->Uslot.n,
                         m->Uslot.p, e->nzval.p, e->Unz.p);
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(n)), dim3(kBlock), 0, s, n, m->lev_rows_b.p,
                       e->pD.p, e->Db.p);
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
    throw Error(XCAL_ERR_STATE, "internal: DILU is applied through the level sweeps");
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
