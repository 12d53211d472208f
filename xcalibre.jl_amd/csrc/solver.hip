// Fused, device-resident Krylov solves for solve_system!
// (Solve_1_api.jl:268-307): CG and BiCGSTAB (the iterations of Krylov.jl's
// cg!/bicgstab!, compat 0.10, restated) with Jacobi or DILU preconditioning.
//
// Design for gfx950:
//  * All Krylov scalars live in device memory (KS).  A reduction writes one
//    partial sum per workgroup with plain stores; a one-workgroup finalize
//    kernel then adds the partials in slot order and updates the scalars
//    (alpha, beta, convergence flags) in place.  The kernel boundary orders
//    the two, so no device-scope fences or atomics are needed (a fence in
//    every workgroup writes back and invalidates the XCD's L2, which costs
//    more than the kernel's own memory traffic).  Every kernel returns at
//    entry once the solve has stopped, so the host enqueues batches of
//    iterations and reads the state once per batch.
//  * BLAS-1 work is fused into the SpMV and the preconditioner sweeps:
//      CG+DILU per iteration: [p = z + beta p ; q = A p ; p.q]
//                             [x += a p ; r -= a q ; forward DILU sweep]
//                             [backward DILU sweep ; r.z]
//  * DILU sweeps are level-synchronous: each level is a set of independent
//    rows; large levels get a launch each, runs of small levels are swept by
//    one workgroup with a barrier between levels.  The summation order in
//    every row is the reference's (ascending column), so the sweeps are
//    bit-identical to forward_substitution!/backward_substitution!
//    (preconditioners_1_DILU.jl:84-114).
//  * Reductions are deterministic: fixed partial-sum slots per workgroup,
//    summed in a fixed order.
#include <cmath>
#include <cstring>
#include <functional>

#include "solver_common.h"
#include "sweep_cone.h"
#include "sweep_pipe.h"

namespace xcal {

// Sums `count` partial pairs in slot order (reduce_partials) and runs the
// scalar update `fin`.
__global__ void __launch_bounds__(kBlock)
    k_finalize(const double* __restrict__ partial, int count, KS* ks, int fin) {
  __shared__ double lds[kBlock / 64];
  KS k;
  if (threadIdx.x == 0) k = *ks;  // loaded while the partials load
  double t0, t1;
  reduce_partials<16>(partial, count, t0, t1, lds);
  if (threadIdx.x == 0 && k.stop == STOP_NONE) {
    finalize(&k, fin, t0, t1);
    *ks = k;
  }
}

// The finalize of a fused chain's last kernel, from its state `in` into `out`.
__global__ void __launch_bounds__(kBlock)
    k_finalize_to(const double* __restrict__ partial, int count, const KS* in, KS* out, int fin) {
  __shared__ double lds[kBlock / 64];
  const KS k = kstate(KStep{in, out, partial, count, fin}, lds);
  (void)k;
}

// --------------------------------------------------------- static kernels
// r = b - A x0 (r = b when x0 is NULL: zero start) ; x = 0 ; pold = 0
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_start(MatKT<T> A, const T* __restrict__ b, const T* __restrict__ x0, T* __restrict__ r,
            T* __restrict__ x, T* __restrict__ pold) {
  int64_t end;
  const int64_t i = block_row(A, end);
  if (i >= end) return;
  if (x0) {
    const T acc = row_sum(A, i, [&](int j) { return x0[j]; });
    r[i] = b[i] - acc;
  } else {
    r[i] = b[i];
  }
  x[i] = 0;
  pold[i] = 0;
}

// z = S r (Jacobi) or z = r; partials r.z  [CG start]   or
// r = S r0 ; partials r.r, c.r               [BiCGSTAB start]
template <bool BICG, class T>
__global__ void __launch_bounds__(kBlock)
    k_precond_diag_start(int64_t n, const T* __restrict__ S, T* __restrict__ r, T* __restrict__ z,
                         const T* __restrict__ c, double* partial) {
  const int64_t i = xcd_row();
  double s0 = 0.0, s1 = 0.0;
  if (i < n) {
    if (BICG) {
      const T ri = S ? S[i] * r[i] : r[i];
      r[i] = ri;
      s0 = dd(ri, ri);
      s1 = dd(c[i], ri);
    } else {
      const T zi = S ? S[i] * r[i] : r[i];
      z[i] = zi;
      s0 = dd(r[i], zi);
    }
  }
  store_partial<kBlock>(s0, s1, partial, blockIdx.x);
}

// CG: p = z + beta p, over the rows of A's row blocks (the rows each
// workgroup of k_cg_spmv reads as its own).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_p(MatKT<T> A, const T* __restrict__ z, const T* __restrict__ pold, T* __restrict__ pnew,
           const KS* ks) {
  const bool live = !stopped(ks);
  const T beta = (T)ks->beta;
  int64_t end;
  const int64_t i = block_row(A, end);
  if (i < end && live) pnew[i] = z[i] + beta * pold[i];
}

// CG with DILU on large meshes: the backward sweep's pending r.z finalize
// (FIN_CG_ITER from its partials, kstate: every workgroup reduces them in the
// same order and workgroup 0 stores the state to K.out), then p = z + beta p
// over the rows of A's row blocks blockIdx.x + u gridDim.x, u < kPFinU.  The
// grid is a multiple of 8, so row block b is handled on the XCD of the SpMV
// workgroup that reads it.  The rows' loads are issued before the
// reduction, which they do not depend on.
constexpr int kPFinU = 4;
inline int pfin_grid(int nblk) { return ((nblk + kPFinU - 1) / kPFinU + 7) / 8 * 8; }
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_p_fin(MatKT<T> A, int nblk, const T* __restrict__ z, const T* __restrict__ pold,
               T* __restrict__ pnew, KStep K) {
  __shared__ double lds[kBlock / 64];
  int64_t i[kPFinU];
  T zv[kPFinU], pv[kPFinU];
#pragma unroll
  for (int u = 0; u < kPFinU; ++u) {
    const int b = (int)blockIdx.x + u * (int)gridDim.x;
    const int bb = b < nblk ? b : 0;
    const int64_t r = (int64_t)A.blk[2 * bb] + threadIdx.x;
    i[u] = b < nblk && r < A.blk[2 * bb + 1] ? r : -1;
  }
#pragma unroll
  for (int u = 0; u < kPFinU; ++u) {
    const int64_t ii = i[u] >= 0 ? i[u] : 0;
    zv[u] = z[ii];
    pv[u] = pold[ii];
  }
  const KS k = kstate(K, lds);
  if (k.stop != STOP_NONE) return;  // uniform over the grid
  const T beta = (T)k.beta;
#pragma unroll
  for (int u = 0; u < kPFinU; ++u)
    if (i[u] >= 0) pnew[i[u]] = zv[u] + beta * pv[u];
}

// CG in one launch for meshes below kSplitRows rows: p = z + beta p on the
// fly for every neighbour ; q = A p ; p.q.  There the launch costs more than
// the second gather per entry.
template <int G, class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_pspmv(MatKT<T> A, const T* __restrict__ z, const T* __restrict__ pold,
               T* __restrict__ pnew, T* __restrict__ q, double* partial, KStep K) {
  __shared__ double lds[kBlock / 64];
  int64_t end;
  const int64_t i = block_row(A, end);
  const bool in = i < end;
  const int64_t ic = in ? i : 0;
  // the row's loads go out before the state is formed (a fused finalize)
  const RowPair<G, T> R = row_pair_load<G>(A, ic, z, pold);
  const T zi = z[ic], poi = pold[ic];
  const KS k = kstate(K, lds);
  const bool live = k.stop == STOP_NONE;
  const T beta = (T)k.beta;
  double s0 = 0.0;
  if (in) {
    const T acc = row_pair_sum(A, R, z, pold, [&](T u, T w) { return u + beta * w; });
    const T pi = zi + beta * poi;
    if (live) {
      pnew[i] = pi;
      q[i] = acc;
    }
    s0 = dd(pi, acc);
  }
  store_partial<kBlock>(s0, 0.0, partial, blockIdx.x);
}

// CG: q = A p ; p.q.  p is formed by k_cg_p beforehand, so each entry
// gathers one value (forming p on the fly for every neighbour doubles the
// gathers, and the gathers bound this kernel).  Rows are addressed through
// their slice (row_sum_slice): two dependent round trips instead of three.
template <int G, class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_spmv(MatKT<T> A, const T* __restrict__ p, T* __restrict__ q, double* partial,
              const KS* ks) {
  const bool live = !stopped(ks);
  int64_t end;
  const int64_t i = block_row(A, end);
  double s0 = 0.0;
  if (i < end) {
    const T pi = p[i];  // issued with the row's first loads
    const auto pj = [&](int j) { return p[j]; };
    // A.rw and A.sinfo are uniform over the launch: a scalar branch
    const T acc = A.sinfo && A.rw <= G ? row_sum_slice<G>(A, i, pj) : row_sum<G>(A, i, pj);
    if (live) q[i] = acc;
    s0 = dd(pi, acc);
  }
  store_partial<kBlock>(s0, 0.0, partial, blockIdx.x);
}

// CG with a diagonal preconditioner: x += a p ; r -= a q ; z = S r ; r.z
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_update_diag(int64_t n, const T* __restrict__ S, const T* __restrict__ p,
                     const T* __restrict__ q, T* __restrict__ x, T* __restrict__ r,
                     T* __restrict__ z, double* partial, KStep K) {
  __shared__ double lds[kBlock / 64];
  const int64_t i = xcd_row();
  const bool in = i < n;
  const int64_t c = in ? i : 0;
  // the row's loads go out before the state is formed (a fused finalize)
  const T xo = x[c], po = p[c], ro = r[c], qo = q[c], so = S ? S[c] : T(1);
  const KS k = kstate(K, lds);
  const bool live = k.stop == STOP_NONE;
  const T a = (T)k.alpha;
  double s0 = 0.0;
  if (in) {
    const T xi = xo + a * po;
    const T ri = ro + (-a) * qo;
    const T zi = S ? so * ri : ri;
    if (live) {
      x[i] = xi;
      r[i] = ri;
      z[i] = zi;
    }
    s0 = dd(ri, zi);
  }
  store_partial<kBlock>(s0, 0.0, partial, blockIdx.x);
}

// BiCGSTAB: pnew = r + beta pold (on the fly) ; q = A pnew ; v = S q (diag
// preconditioner, fused when VDIAG) ; c.v
template <int G, bool VDIAG, class T>
__global__ void __launch_bounds__(kBlock)
    k_bi_spmv_p(MatKT<T> A, const T* __restrict__ r, const T* __restrict__ pold,
                T* __restrict__ pnew, const T* __restrict__ S, T* __restrict__ qv,
                const T* __restrict__ c, double* partial, KStep K) {
  __shared__ double lds[kBlock / 64];
  int64_t end;
  const int64_t i = block_row(A, end);
  const bool in = i < end;
  const int64_t ic = in ? i : 0;
  // the row's loads go out before the state is formed (a fused finalize)
  const RowPair<G, T> R = row_pair_load<G>(A, ic, r, pold);
  const T ri = r[ic], poi = pold[ic];
  const KS k = kstate(K, lds);
  const bool live = k.stop == STOP_NONE;
  const T beta = (T)k.beta;
  double s0 = 0.0;
  if (in) {
    const T acc = row_pair_sum(A, R, r, pold, [&](T u, T w) { return u + beta * w; });
    const T pi = ri + beta * poi;
    const T vi = VDIAG && S ? S[i] * acc : acc;
    if (live) {
      pnew[i] = pi;
      qv[i] = vi;
    }
    if (VDIAG) s0 = dd(c[i], vi);
  }
  if (VDIAG) store_partial<kBlock>(s0, 0.0, partial, blockIdx.x);
}

// BiCGSTAB: s = r - a v (on the fly) ; d = A s ; x += a p ;
// t = S d and the dots t.s, t.t when TDIAG
template <int G, bool TDIAG, class T>
__global__ void __launch_bounds__(kBlock)
    k_bi_spmv_s(MatKT<T> A, const T* __restrict__ r, const T* __restrict__ v,
                const T* __restrict__ p, T* __restrict__ s, T* __restrict__ x,
                const T* __restrict__ S, T* __restrict__ dt, double* partial, KStep K) {
  __shared__ double lds[kBlock / 64];
  int64_t end;
  const int64_t i = block_row(A, end);
  const bool in = i < end;
  const int64_t ic = in ? i : 0;
  // the row's loads go out before the state is formed (a fused finalize)
  const RowPair<G, T> R = row_pair_load<G>(A, ic, r, v);
  const T ri = r[ic], vi = v[ic], xo = x[ic], po = p[ic];
  const KS k = kstate(K, lds);
  const bool live = k.stop == STOP_NONE;
  const T a = (T)k.alpha;
  double s0 = 0.0, s1 = 0.0;
  if (in) {
    const T acc = row_pair_sum(A, R, r, v, [&](T u, T w) { return u + (-a) * w; });
    const T si = ri + (-a) * vi;
    const T xi = xo + a * po;
    const T ti = TDIAG && S ? S[i] * acc : acc;
    if (live) {
      s[i] = si;
      x[i] = xi;
      dt[i] = ti;
    }
    if (TDIAG) {
      s0 = dd(ti, si);
      s1 = dd(ti, ti);
    }
  }
  if (TDIAG) store_partial<kBlock>(s0, s1, partial, blockIdx.x);
}

// BiCGSTAB tail: x += w s ; r = s - w t ; p = p - w v ; c.r and r.r
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_bi_tail(int64_t n, const T* __restrict__ s, const T* __restrict__ t, const T* __restrict__ v,
              T* __restrict__ x, T* __restrict__ r, T* __restrict__ p, const T* __restrict__ c,
              double* partial, KStep K) {
  __shared__ double lds[kBlock / 64];
  const int64_t i = xcd_row();
  const bool in = i < n;
  const int64_t j = in ? i : 0;
  // the row's loads go out before the state is formed (a fused finalize)
  const T xo = x[j], so = s[j], to = t[j], po = p[j], vo = v[j], co = c[j];
  const KS k = kstate(K, lds);
  const bool live = k.stop == STOP_NONE;
  const T w = (T)k.omega;
  double s0 = 0.0, s1 = 0.0;
  if (in) {
    const T xi = xo + w * so;
    const T ri = so + (-w) * to;
    const T pi = po + (-w) * vo;
    if (live) {
      x[i] = xi;
      r[i] = ri;
      p[i] = pi;
    }
    s0 = dd(co, ri);
    s1 = dd(ri, ri);
  }
  store_partial<kBlock>(s0, s1, partial, blockIdx.x);
}

// x0 = x0 + x (warm start, Krylov.jl adds the start vector last)
__global__ void __launch_bounds__(kBlock)
    k_add(int64_t n, const double* __restrict__ x, double* __restrict__ x0) {
  const int64_t i = xcd_row();
  if (i < n) x0[i] = x[i] + x0[i];
}

// One weighted-Jacobi sweep of apply_smoother! (Smoothers_jacobi.jl:72-81,
// multiply_row :97-110): the row sum over every entry in column order, less
// the diagonal's product, then omega/a_ii (b - sum) + (1 - omega) x_i.
// Every row reads the previous sweep's values (xin) and writes xout.
template <int G>
__global__ void __launch_bounds__(kBlock)
    k_jacobi_smooth(MatK A, const double* __restrict__ b, const double* __restrict__ xin,
                    double* __restrict__ xout, double omega) {
  int64_t end;
  const int64_t i = block_row(A, end);
  if (i >= end) return;
  double sum = row_sum<G>(A, i, [&](int j) { return xin[j]; });
  const double aii = A.nz[A.diag[i]];
  const double xi = xin[i];
  sum -= aii * xi;
  const double rD = 1.0 / aii;
  xout[i] = (omega * rD) * (b[i] - sum) + (1.0 - omega) * xi;
}

// Crank-Nicolson's explicit step after the solve (Solve_1_api.jl:291-296):
// x = 2 x - x_old, x_old the field the Krylov solve started from.
__global__ void __launch_bounds__(kBlock)
    k_cn_step(int64_t n, const double* __restrict__ xold, double* __restrict__ x) {
  const int64_t i = xcd_row();
  if (i < n) x[i] = 2.0 * x[i] - xold[i];
}

// --------------------------------------------- mixed-precision refinement
// r = b - A x in FP64, stored as FP32 for the inner solve; partials of
// |r|^2 and |b|^2
__global__ void __launch_bounds__(kBlock)
    k_ir_residual(MatK A, const double* __restrict__ x, const double* __restrict__ b,
                  float* __restrict__ r32, double* partial) {
  int64_t end;
  const int64_t i = block_row(A, end);
  double s0 = 0.0, s1 = 0.0;
  if (i < end) {
    const double acc = row_sum(A, i, [&](int j) { return x[j]; });
    const double r = b[i] - acc;
    r32[i] = (float)r;
    s0 = r * r;
    s1 = b[i] * b[i];
  }
  store_partial<kBlock>(s0, s1, partial, blockIdx.x);
}

__global__ void __launch_bounds__(kRunBlock)
    k_sum2(const double* __restrict__ partial, int count, double* __restrict__ out) {
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

// x += d (FP32 correction into the FP64 solution)
__global__ void __launch_bounds__(kBlock)
    k_ir_update(int64_t n, const float* __restrict__ d, double* __restrict__ x) {
  const int64_t i = xcd_row();
  if (i < n) x[i] = x[i] + (double)d[i];
}

__global__ void __launch_bounds__(kBlock)
    k_to_float(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = xcd_row();
  if (i < n) b[i] = (float)a[i];
}

// b = a * rd rounded to FP32 (the folded forward factor a_ij / D_j)
__global__ void __launch_bounds__(kBlock)
    k_fold_float(int64_t n, const double* __restrict__ a, const double* __restrict__ rd,
                 float* __restrict__ b) {
  const int64_t i = xcd_row();
  if (i < n) b[i] = (float)(a[i] * rd[i]);
}

// b = 1 / a rounded to FP32
__global__ void __launch_bounds__(kBlock)
    k_recip_float(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = xcd_row();
  if (i < n) b[i] = (float)(1.0 / a[i]);
}

// one level: rows [row0, row0 + cnt) (forward levels are contiguous)
template <bool UPD, int WK, class T>
__global__ void __launch_bounds__(kBlock)
    k_fwd_level(FwdK<T> W, int32_t row0, int32_t cnt, int32_t base, int32_t w) {
  int32_t end;
  const int32_t t = level_pos(cnt, end);
  if (t < end) fwd_row<UPD, WK, T, true>(W, row0 + t, (int64_t)base + t, cnt, w, T(0), true);
}

// a run of small levels [lev0, lev1) in one workgroup
template <bool UPD, class T>
__global__ void __launch_bounds__(kRunBlock)
    k_fwd_run(FwdK<T> W, const int32_t* __restrict__ lev_ptr, const int32_t* __restrict__ meta,
              int lev0, int lev1) {
  if (stopped(W.ks)) return;  // uniform over the workgroup
  const T a = UPD ? (T)W.ks->alpha : T(0);
  for (int l = lev0; l < lev1; ++l) {
    const int32_t r0 = lev_ptr[l], cnt = lev_ptr[l + 1] - r0, base = meta[2 * l], w = meta[2 * l + 1];
    for (int32_t q = threadIdx.x; q < cnt; q += blockDim.x)
      fwd_row<UPD, 0>(W, r0 + q, (int64_t)base + q, cnt, w, a, true);
    __syncthreads();
  }
}

template <int DOT, int WK, class T>
__global__ void __launch_bounds__(kBlock)
    k_bwd_level(BwdK<T> W, int32_t pos0, int32_t cnt, int32_t base, int32_t w, int slot) {
  int32_t end;
  const int32_t t = level_pos(cnt, end);
  double s0 = 0.0, s1 = 0.0;
  if (t < end) bwd_row<DOT, WK, T, true>(W, pos0 + t, (int64_t)base + t, cnt, w, true, s0, s1);
  if (DOT != 0) store_partial<kBlock>(s0, s1, W.partial, slot + blockIdx.x);
}

template <int DOT, class T>
__global__ void __launch_bounds__(kRunBlock)
    k_bwd_run(BwdK<T> W, const int32_t* __restrict__ lev_ptr, const int32_t* __restrict__ meta,
              int lev0, int lev1, int slot) {
  if (stopped(W.ks)) return;
  double s0 = 0.0, s1 = 0.0;
  for (int l = lev0; l < lev1; ++l) {
    const int32_t p0 = lev_ptr[l], cnt = lev_ptr[l + 1] - p0, base = meta[2 * l], w = meta[2 * l + 1];
    for (int32_t q = threadIdx.x; q < cnt; q += blockDim.x)
      bwd_row<DOT, 0>(W, p0 + q, (int64_t)base + q, cnt, w, true, s0, s1);
    __syncthreads();
  }
  if (DOT != 0) store_partial<kRunBlock>(s0, s1, W.partial, slot);
}

// Pipelined runs (sweep_pipe.h): one workgroup per run of small levels,
// with the ring preloaded from the launches before it.  UPD runs have had
// x += a p, r -= a q done by k_cg_xr over their rows; their input is u = r.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_xr(int32_t row0, int32_t row1, T* __restrict__ x, const T* __restrict__ p,
            T* __restrict__ r, const T* __restrict__ q, const KS* ks) {
  const int32_t i = row0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (stopped(ks) || i >= row1) return;
  const T a = (T)ks->alpha;
  const T xi = x[i] + a * p[i];
  const T ri = r[i] + (-a) * q[i];
  x[i] = xi;
  r[i] = ri;
}

template <bool UPD, bool FAR, class T, int WK>
__global__ void __launch_bounds__(kPipeBlock)
    k_fwd_run_pipe(FwdK<T> W, const T* __restrict__ Lrd, const int32_t* __restrict__ lev_ptr,
                   const int32_t* __restrict__ meta, int lev0, int lev1) {
  __shared__ PipeLDS<T> L;
  if (stopped(W.ks)) return;  // uniform over the workgroup
  fwd_pipe<true, FAR, T, WK, pipe_depth<WK>()>(W, UPD ? W.r : W.u, Lrd, lev_ptr, meta, lev0,
                                          lev1 - lev0, true, L);
}

template <int DOT, bool FAR, class T, int WK>
__global__ void __launch_bounds__(kPipeBlock)
    k_bwd_run_pipe(BwdK<T> W, const int32_t* __restrict__ Upos, const int32_t* __restrict__ lev_ptr,
                   const int32_t* __restrict__ meta, int lev0, int lev1, int slot) {
  __shared__ PipeLDS<T> L;
  if (stopped(W.ks)) return;
  double s0[1] = {0.0}, s1[1] = {0.0};
  bwd_pipe<DOT, FAR, T, WK, pipe_depth<WK>()>(W, Upos, lev_ptr, meta, lev0, lev1 - lev0, true, L, s0,
                                         s1);
  if (DOT != 0) store_partial<kPipeBlock>(s0[0], s1[0], W.partial, slot);
}

// Cone form (sweep_cone.h): one workgroup per block of the launch's band.
// Both return at entry once the solve has stopped.  The sweep computes in T
// over vectors of type V; Yt is the forward output in T.  With K.fin set the
// forward sweep first does the SpMV's p.q finalize itself (kstate: every
// workgroup from the same partials, workgroup 0 storing the state to K.out)
// and takes alpha from it; otherwise it reads the state W.ks.
static_assert(kConeBlock == kBlock, "kstate reduces with kBlock threads");
template <bool UPD, class T, int WK, int P, class V>
__global__ void __launch_bounds__(kConeBlock)
    k_fwd_cone(ConeK C, const T* __restrict__ a, const T* __restrict__ rd, FwdK<V> W, T* Yt,
               KStep K) {
  __shared__ ConeLDS<T, !cone_folded<T>()> L;
  __shared__ double lds[kConeBlock / 64];
  double al = 0.0;
  if (K.fin != FIN_NONE) {
    const KS k = kstate(K, lds);
    if (k.stop != STOP_NONE) return;  // uniform over the grid
    al = k.alpha;
  } else if (W.ks) {
    if (stopped(W.ks)) return;  // uniform over the workgroup
    al = W.ks->alpha;
  }
  fwd_cone<UPD, T, WK, P, V>(C, a, rd, W, Yt, L, al);
}

template <int DOT, class T, int WK, int P, class V>
__global__ void __launch_bounds__(kConeBlock)
    k_bwd_cone(ConeK C, const T* __restrict__ a, const T* __restrict__ d, BwdK<V> W,
               const T* Yt) {
  __shared__ ConeLDS<T> L;
  if (W.ks && stopped(W.ks)) return;
  bwd_cone<DOT, T, WK, P, V>(C, a, d, W, Yt, L);
}


// ------------------------------------------------------------ host side
// One Krylov solve's view of an equation in value type T: the matrix, the
// preconditioner arrays and 12 n-vectors of workspace.
template <class T>
struct Solver {
  xcal_eqn* e;
  const xcal_mesh* m;
  hipStream_t s;
  int64_t n;
  MatKT<T> A;
  const T *D, *rD, *Sj;  // DILU D and 1/D; Jacobi 1/|a_ii|
  const T *Lnz, *Unz, *Db;  // packed DILU factors, D in backward-level order
  const T* Lrd;             // 1/D per packed lower entry (pipelined runs)
  // cone form: a_ij and 1/D_j per forward entry, 1/D per forward position,
  // a_ij per backward entry, D per backward position
  const T *cfa, *cfr, *cfd, *cba, *cbd;
  // FP32 preconditioner of an FP64 solve (XCAL_PREC_FP32_PRECON, cone form):
  // folded FP32 factors a_ij / D_j, a_ij and 1 / D of the cone sweeps
  bool pc32 = false;
  const float *cfa32 = nullptr, *cba32 = nullptr, *cbd32 = nullptr;
  T* work;
  KS* ks;
  double* partial;
  KS* fks;        // two states for fused finalizes
  double* fpart;  // two partial areas of kFuseBlocks pairs for fused finalizes
  int grid;   // elementwise kernels: grid_for(n) blocks in XCD order
  int gridA;  // gathering kernels: the mesh's row blocks (xcal_mesh::xblk)
  int* cg_batch;
  int* bi_batch;
};

static void ensure_workspace(xcal_eqn* e, bool fp32) {
  const xcal_mesh* m = e->mesh;
  const int64_t n = m->n;
  const int64_t slots =
      std::max<int64_t>(std::max<int64_t>(grid_for(n), m->nxblk), m->sweep_slots) + 8;
  const bool grow = (e->work.n > 0 && e->work.n < 12 * n) ||
                    (e->kpart.n > 0 && e->kpart.n < 2 * slots) ||
                    (fp32 && e->work32.n > 0 && e->work32.n < 12 * n) ||
                    (fp32 && e->nz32.n > 0 && e->nz32.n < m->nzlen);
  if (grow) XCAL_HIP(hipStreamSynchronize(m->ctx->stream));  // replaced buffers may be in use
  if (e->work.n < 12 * n) e->work.alloc(12 * n);
  if (e->kpart.n < 2 * slots) e->kpart.alloc(2 * slots);
  if (e->kstate.n < (int64_t)sizeof(KS)) e->kstate.alloc(sizeof(KS));
  if (e->fpart.n < 4 * kFuseBlocks) e->fpart.alloc(4 * kFuseBlocks);
  if (e->fstate.n < (int64_t)(2 * sizeof(KS))) e->fstate.alloc(2 * sizeof(KS));
  if (!e->kstate_host) XCAL_HIP(hipHostMalloc((void**)&e->kstate_host, sizeof(KS), 0));
  if (fp32) {
    if (e->work32.n < 12 * n) e->work32.alloc(12 * n);
    if (e->nz32.n < m->nzlen) e->nz32.alloc(m->nzlen);
    if (e->pD32.n < n) {
      e->pD32.alloc(n);
      e->prD32.alloc(n);
      e->pS32.alloc(n);
    }
  }
}

static Solver<double> make_solver(xcal_eqn* e) {
  ensure_workspace(e, false);
  Solver<double> S;
  S.e = e;
  S.m = e->mesh;
  S.s = e->mesh->ctx->stream;
  S.n = e->mesh->n;
  S.A = matk(S.m, (const double*)e->nzval.p);
  S.D = e->pD.p;
  S.rD = e->prD.p;
  S.Sj = e->pS.p;
  S.Lnz = e->Lnz.p;
  S.Unz = e->Unz.p;
  S.Db = e->Db.p;
  S.Lrd = e->Lrd.p;
  S.cfa = e->cfa.p;
  S.cfr = e->cfr.p;
  S.cfd = e->cfd.p;
  S.cba = e->cba.p;
  S.cbd = e->cbd.p;
  S.work = e->work.p;
  S.ks = (KS*)e->kstate.p;
  S.partial = e->kpart.p;
  S.fks = (KS*)e->fstate.p;
  S.fpart = e->fpart.p;
  S.grid = grid_for(S.n);
  S.gridA = S.m->nxblk;
  S.cg_batch = &e->cg_batch;
  S.bi_batch = &e->bi_batch;
  return S;
}

// The FP32 cone sweeps' factors, folded (cone_folded) from the FP64 ones:
// a_ij / D_j per forward entry, a_ij per backward entry, 1 / D per backward
// position.
static void cone_factors32(xcal_eqn* e) {
  hipStream_t s = e->mesh->ctx->stream;
  if (e->cfa32.n != e->cfa.n) e->cfa32.alloc(e->cfa.n);
  if (e->cfa.n > 0)
    hipLaunchKernelGGL(k_fold_float, dim3(grid_for(e->cfa.n)), dim3(kBlock), 0, s, e->cfa.n,
                       (const double*)e->cfa.p, (const double*)e->cfr.p, e->cfa32.p);
  if (e->cba32.n != e->cba.n) e->cba32.alloc(e->cba.n);
  if (e->cba.n > 0)
    hipLaunchKernelGGL(k_to_float, dim3(grid_for(e->cba.n)), dim3(kBlock), 0, s, e->cba.n,
                       e->cba.p, e->cba32.p);
  if (e->cbd32.n != e->cbd.n) e->cbd32.alloc(e->cbd.n);
  if (e->cbd.n > 0)
    hipLaunchKernelGGL(k_recip_float, dim3(grid_for(e->cbd.n)), dim3(kBlock), 0, s, e->cbd.n,
                       (const double*)e->cbd.p, e->cbd32.p);
  XCAL_HIP(hipGetLastError());
}

// FP32 copies of the matrix and the preconditioner, made from the FP64 ones
static Solver<float> make_solver32(xcal_eqn* e) {
  ensure_workspace(e, true);
  const xcal_mesh* m = e->mesh;
  hipStream_t s = m->ctx->stream;
  const int64_t n = m->n;
  hipLaunchKernelGGL(k_to_float, dim3(grid_for(m->nzlen)), dim3(kBlock), 0, s, m->nzlen,
                     e->nzval.p, e->nz32.p);
  if (e->precon == XCAL_PRECON_DILU) {
    hipLaunchKernelGGL(k_to_float, dim3(grid_for(n)), dim3(kBlock), 0, s, n, e->pD.p, e->pD32.p);
    hipLaunchKernelGGL(k_to_float, dim3(grid_for(n)), dim3(kBlock), 0, s, n, e->prD.p, e->prD32.p);
  }
  if (e->precon == XCAL_PRECON_DILU && !m->cone) {  // the level sweeps' packed factors
    if (e->Lnz32.n != e->Lnz.n) e->Lnz32.alloc(e->Lnz.n);
    if (e->Unz32.n != e->Unz.n) e->Unz32.alloc(e->Unz.n);
    if (e->Db32.n != n) e->Db32.alloc(n);
    hipLaunchKernelGGL(k_to_float, dim3(grid_for(n)), dim3(kBlock), 0, s, n, e->Db.p, e->Db32.p);
    if (e->Lnz.n > 0)
      hipLaunchKernelGGL(k_to_float, dim3(grid_for(e->Lnz.n)), dim3(kBlock), 0, s, e->Lnz.n,
                         e->Lnz.p, e->Lnz32.p);
    if (e->Unz.n > 0)
      hipLaunchKernelGGL(k_to_float, dim3(grid_for(e->Unz.n)), dim3(kBlock), 0, s, e->Unz.n,
                         e->Unz.p, e->Unz32.p);
    if (e->Lrd.n > 0) {
      if (e->Lrd32.n != e->Lrd.n) e->Lrd32.alloc(e->Lrd.n);
      hipLaunchKernelGGL(k_to_float, dim3(grid_for(e->Lrd.n)), dim3(kBlock), 0, s, e->Lrd.n,
                         e->Lrd.p, e->Lrd32.p);
    }
  } else if (e->precon == XCAL_PRECON_DILU) {  // the cone sweeps' packed factors
    cone_factors32(e);
  } else if (e->precon == XCAL_PRECON_JACOBI) {
    hipLaunchKernelGGL(k_to_float, dim3(grid_for(n)), dim3(kBlock), 0, s, n, e->pS.p, e->pS32.p);
  }
  XCAL_HIP(hipGetLastError());
  Solver<float> S;
  S.e = e;
  S.m = m;
  S.s = s;
  S.n = n;
  S.A = matk(m, (const float*)e->nz32.p);
  S.D = e->pD32.p;
  S.rD = e->prD32.p;
  S.Sj = e->pS32.p;
  S.Lnz = e->Lnz32.p;
  S.Unz = e->Unz32.p;
  S.Db = e->Db32.p;
  S.Lrd = e->Lrd32.p;
  S.cfa = e->cfa32.p;
  S.cfr = e->cfa32.p;  // unused by the folded FP32 forward cone sweep
  S.cfd = e->cfa32.p;  // likewise
  S.cba = e->cba32.p;
  S.cbd = e->cbd32.p;
  S.work = e->work32.p;
  S.ks = (KS*)e->kstate.p;
  S.partial = e->kpart.p;
  S.fks = (KS*)e->fstate.p;
  S.fpart = e->fpart.p;
  S.grid = grid_for(n);
  S.gridA = m->nxblk;
  S.cg_batch = &e->ir_pred[0][0];
  S.bi_batch = &e->ir_pred[1][0];
  return S;
}

template <class T>
static void finalize_launch(const Solver<T>& S, int count, int fin) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, S.s, S.partial, count, S.ks, fin);
}

// Fused finalizes (kFuseBlocks).  Each kernel of the chain reads the state
// `cur` and the pending partials, writes its own partials to the other area
// and the updated state to the other slot; flush() runs the last pending
// finalize into S.ks before the host reads the state.  Off, every kernel
// reads S.ks and writes S.partial, with a one-workgroup finalize between.
template <class T>
struct Fuse {
  const Solver<T>& S;
  bool on;
  KS* cur;
  int kslot = 0, pslot = 0;
  const double* pend_p = nullptr;
  int pend_n = 0, pend_fin = FIN_NONE;
  Fuse(const Solver<T>& s, bool fuse) : S(s), on(fuse), cur(s.ks) {}
  // partial area and state step of the next kernel of the chain
  double* part() {
    if (!on) return S.partial;
    double* p = S.fpart + 2 * kFuseBlocks * pslot;
    pslot ^= 1;
    return p;
  }
  KStep step() {
    if (!on) return kstep_plain(S.ks);
    KS* out = S.fks + kslot;
    kslot ^= 1;
    const KStep K{cur, out, pend_p, pend_n, pend_fin};
    cur = out;
    pend_fin = FIN_NONE;
    return K;
  }
  // the kernel just launched left `count` partials at p for update `fin`
  void done(const double* p, int count, int fin) {
    if (!on) {
      finalize_launch(S, count, fin);
      return;
    }
    pend_p = p;
    pend_n = count;
    pend_fin = fin;
  }
  void flush() {
    if (!on || (pend_fin == FIN_NONE && cur == S.ks)) return;
    hipLaunchKernelGGL(k_finalize_to, dim3(1), dim3(kBlock), 0, S.s, pend_p, pend_n,
                       (const KS*)cur, S.ks, pend_fin);
    cur = S.ks;
    pend_fin = FIN_NONE;
  }
};

// forward sweep over the launch plan; K (cone form, CG): a p.q finalize
// the sweep does first, its state then in K.out
template <bool UPD, class T>
static void sweep_fwd(const Solver<T>& S, const FwdK<T>& W, const KStep* K = nullptr) {
  const xcal_mesh* m = S.m;
  if (m->cone) {  // y to W.Yb in device order
    XCAL_CHECK(W.Yb != (UPD ? W.r : W.u) && (!UPD || (W.rout != W.r && W.rout != W.Yb)),
               "internal: cone sweep operands alias");
    const std::vector<int32_t>& band = m->cone_f.band;
    with_cone_form(m->cone_f.wk, !m->cone_deep, [&](auto WK, auto P) {
      constexpr int wk = decltype(WK)::value, p = decltype(P)::value;
      const KStep Kp = K ? *K : KStep{nullptr, nullptr, nullptr, 0, FIN_NONE};
      for (size_t k = 0; k + 1 < band.size(); ++k) {  // one launch per band of levels
        const ConeK C = cone_k(m, m->cone_f, band[k]);
        const dim3 grid(band[k + 1] - band[k]);
        if (std::is_same<T, double>::value && S.pc32)  // y in float, in the Yb vector's storage
          hipLaunchKernelGGL((k_fwd_cone<UPD, float, wk, p, T>), grid, dim3(kConeBlock), 0, S.s,
                             C, S.cfa32, S.cfa32, W, reinterpret_cast<float*>(W.Yb), Kp);
        else
          hipLaunchKernelGGL((k_fwd_cone<UPD, T, wk, p, T>), grid, dim3(kConeBlock), 0, S.s, C,
                             S.cfa, S.cfd, W, W.Yb, Kp);
      }
    });
    return;
  }
  for (const SweepStep& st : m->plan_f) {
    if (st.grid == 1 && st.pipe > 0) {
      if (UPD)
        hipLaunchKernelGGL(k_cg_xr<T>, dim3(grid_for(st.row1 - st.row0)), dim3(kBlock), 0, S.s,
                           st.row0, st.row1, W.x, W.p, W.r, W.q, W.ks);
      auto pipe = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(1), dim3(kPipeBlock), 0, S.s, W, S.Lrd, m->lev_ptr_f_d.p,
                           m->lmeta_f_d.p, st.lev0, st.lev1);
      };
      with_pipe_form(st.pipe, st.far, [&](auto WK, auto FAR) {
        pipe(k_fwd_run_pipe<UPD, decltype(FAR)::value, T, decltype(WK)::value>);
      });
    } else if (st.grid == 1)
      hipLaunchKernelGGL((k_fwd_run<UPD, T>), dim3(1), dim3(kRunBlock), 0, S.s, W,
                         m->lev_ptr_f_d.p, m->lmeta_f_d.p, st.lev0, st.lev1);
    else {
      const int w = m->lmeta_f[2 * st.lev0 + 1];
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(st.grid), dim3(kBlock), 0, S.s, W, st.row0,
                           st.row1 - st.row0, m->lmeta_f[2 * st.lev0], w);
      };
      switch (entry_bound(w)) {
        case 2: go(k_fwd_level<UPD, 2, T>); break;
        case 4: go(k_fwd_level<UPD, 4, T>); break;
        case 8: go(k_fwd_level<UPD, 8, T>); break;
        default: go(k_fwd_level<UPD, 0, T>);
      }
    }
  }
}

// backward sweep; with DOT != 0 its partials fill slots [0, sweep_slots)
template <int DOT, class T>
static void sweep_bwd(const Solver<T>& S, const BwdK<T>& W) {
  const xcal_mesh* m = S.m;
  if (m->cone) {  // y from W.Yb in device order; partials in slots [0, backward blocks)
    XCAL_CHECK(W.Z != W.Yb && (DOT == 0 || W.dv != W.Z), "internal: cone sweep operands alias");
    const std::vector<int32_t>& band = m->cone_b.band;
    with_cone_form(m->cone_b.wk, !m->cone_deep, [&](auto WK, auto P) {
      constexpr int wk = decltype(WK)::value, p = decltype(P)::value;
      for (size_t k = 0; k + 1 < band.size(); ++k) {  // one launch per band of levels
        const ConeK C = cone_k(m, m->cone_b, band[k]);
        const dim3 grid(band[k + 1] - band[k]);
        if (std::is_same<T, double>::value && S.pc32)
          hipLaunchKernelGGL((k_bwd_cone<DOT, float, wk, p, T>), grid, dim3(kConeBlock), 0, S.s,
                             C, S.cba32, S.cbd32, W, reinterpret_cast<const float*>(W.Yb));
        else
          hipLaunchKernelGGL((k_bwd_cone<DOT, T, wk, p, T>), grid, dim3(kConeBlock), 0, S.s, C,
                             S.cba, S.cbd, W, W.Yb);
      }
    });
    return;
  }
  for (const SweepStep& st : m->plan_b) {
    if (st.grid == 1 && st.pipe > 0) {
      auto pipe = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(1), dim3(kPipeBlock), 0, S.s, W, m->Upos.p,
                           m->lev_ptr_b_d.p, m->lmeta_b_d.p, st.lev0, st.lev1, st.slot);
      };
      with_pipe_form(st.pipe, st.far, [&](auto WK, auto FAR) {
        pipe(k_bwd_run_pipe<DOT, decltype(FAR)::value, T, decltype(WK)::value>);
      });
    } else if (st.grid == 1)
      hipLaunchKernelGGL((k_bwd_run<DOT, T>), dim3(1), dim3(kRunBlock), 0, S.s, W,
                         m->lev_ptr_b_d.p, m->lmeta_b_d.p, st.lev0, st.lev1, st.slot);
    else {
      const int w = m->lmeta_b[2 * st.lev0 + 1];
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(st.grid), dim3(kBlock), 0, S.s, W, st.row0,
                           st.row1 - st.row0, m->lmeta_b[2 * st.lev0], w, st.slot);
      };
      switch (entry_bound(w)) {
        case 2: go(k_bwd_level<DOT, 2, T>); break;
        case 4: go(k_bwd_level<DOT, 4, T>); break;
        case 8: go(k_bwd_level<DOT, 8, T>); break;
        default: go(k_bwd_level<DOT, 0, T>);
      }
    }
  }
}

// Y, Yb: two n-vectors of workspace (forward output in device and in
// backward-level order)
template <class T>
static FwdK<T> fwdk(const Solver<T>& S, const T* u, T* Y, T* Yb, const KS* ks) {
  return FwdK<T>{S.m->Lcol.p, S.Lnz, S.rD, u, Y, Yb, S.m->posb.p, nullptr, nullptr, nullptr,
                 nullptr, ks};
}
template <class T>
static BwdK<T> bwdk(const Solver<T>& S, const T* Yb, T* Z, const T* dv, const KS* ks) {
  return BwdK<T>{S.m->Ucol.p, S.Unz, S.m->lev_rows_b.p, S.Db, Yb, Z, dv, S.partial, ks};
}

// z = M^-1 u through the sweeps (DILU), used by precond_apply; Y and Yb are
// two n-vectors of staging (forward output in device and backward order).
static void dilu_apply(const Solver<double>& S, const double* u, double* z, double* Y,
                       double* Yb) {
  const KS* live = (const KS*)S.m->ctx->nostop.p;  // a state that never stops
  sweep_fwd<false>(S, fwdk(S, u, Y, Yb, live));
  sweep_bwd<0>(S, bwdk(S, (const double*)Yb, z, (const double*)nullptr, live));
  XCAL_HIP(hipGetLastError());
}

// z = M^-1 r with the staging vectors Y and Yb given by the caller: no
// device memory is allocated (xcal_precond_apply_noalloc).
void precond_apply_staged(xcal_eqn* e, const double* r, double* z, double* Y, double* Yb) {
  if (e->precon != XCAL_PRECON_DILU) {
    apply_M_levels(e, r, z);
    return;
  }
  Solver<double> S{};  // the factors only: the sweeps use no Krylov workspace
  S.e = e;
  S.m = e->mesh;
  S.s = e->mesh->ctx->stream;
  S.n = e->mesh->n;
  S.rD = e->prD.p;
  S.Lnz = e->Lnz.p;
  S.Unz = e->Unz.p;
  S.Db = e->Db.p;
  S.Lrd = e->Lrd.p;
  S.cfa = e->cfa.p;
  S.cfr = e->cfr.p;
  S.cfd = e->cfd.p;
  S.cba = e->cba.p;
  S.cbd = e->cbd.p;
  dilu_apply(S, r, z, Y, Yb);
}

template <class T>
static KS read_state(Solver<T>& S) {
  XCAL_HIP(hipMemcpyAsync(S.e->kstate_host, S.ks, sizeof(KS), hipMemcpyDeviceToHost, S.s));
  XCAL_HIP(hipStreamSynchronize(S.s));
  prof_resolve(S.m->ctx);
  KS k;
  std::memcpy(&k, S.e->kstate_host, sizeof(KS));
  return k;
}

template <class T>
static void init_state(Solver<T>& S, int itmax, double atol, double rtol) {
  KS k{};
  k.atol = atol;
  k.rtol = rtol;
  k.itmax = itmax;
  k.stop = STOP_NONE;
  std::memcpy(S.e->kstate_host, &k, sizeof(KS));
  XCAL_HIP(hipMemcpyAsync(S.ks, S.e->kstate_host, sizeof(KS), hipMemcpyHostToDevice, S.s));
}

// Enqueues iterations in batches and reads the device state after each.
// Iterations enqueued after the solve has stopped still load their data
// (only their stores are predicated off), so the first batch is sized from
// the previous solve of the same equation (`*pred` iterations; 0 = none) and
// later batches are short.
template <class T>
static int run_batches(Solver<T>& S, int* pred, int tail, const std::function<void()>& iteration,
                       KS& out, const std::function<void()>& flush = nullptr) {
  int batch = *pred > 0 ? std::max(1, (*pred * 9) / 10) : 8;
  for (;;) {
    for (int b = 0; b < batch; ++b) iteration();
    if (flush) flush();
    XCAL_HIP(hipGetLastError());
    out = read_state(S);
    if (out.stop != STOP_NONE) break;
    batch = tail;
  }
  *pred = out.iter;
  return out.stop;
}

// Krylov.jl cg! with warm start x0 (solution = x0 + correction, written to
// x0).  With x0 == NULL the start is zero and the solution is left in
// *xsol (a workspace vector).  Returns 1 when itmax was reached.
template <class T>
static int cg(Solver<T>& S, int precon, const T* b, T* x0, int itmax, double atol, double rtol,
              int& niter, T** xsol) {
  const int64_t n = S.n;
  T* w = S.work;
  T *r = w, *x = w + n, *pa = w + 2 * n, *pb = w + 3 * n, *q = w + 4 * n, *zd = w + 5 * n,
    *Y = w + 6 * n, *Yb = w + 7 * n;
  init_state(S, itmax, atol, rtol);
  hipLaunchKernelGGL(k_start<T>, dim3(S.gridA), dim3(kBlock), 0, S.s, S.A, b, (const T*)x0, r, x, pa);
  const bool dilu = precon == XCAL_PRECON_DILU;
  const T* Sd = precon == XCAL_PRECON_JACOBI ? S.Sj : nullptr;
  const int sweep_slots = S.m->sweep_slots;
  if (dilu) {
    sweep_fwd<false>(S, fwdk(S, (const T*)r, Y, Yb, (const KS*)S.ks));
    sweep_bwd<1>(S, bwdk(S, (const T*)Yb, zd, (const T*)r, (const KS*)S.ks));
    finalize_launch(S, sweep_slots, FIN_CG_INIT);
  } else {
    hipLaunchKernelGGL((k_precond_diag_start<false, T>), dim3(S.grid), dim3(kBlock), 0, S.s, n, Sd,
                       r, zd, (const T*)nullptr, S.partial);
    finalize_launch(S, S.grid, FIN_CG_INIT);
  }
  int parity = 0;
  // The cone sweeps read the old r of other workgroups' rows while the fused
  // update writes the new one, so r alternates between two vectors there.
  T* rv[2] = {r, w + 8 * n};
  int rpar = 0;
  const bool split = n >= kSplitRows;
  Fuse<T> F(S, S.m->fuse && !dilu && !split && S.gridA <= kFuseBlocks && S.grid <= kFuseBlocks);
  // Cone sweeps on a large mesh: no separate finalize.  The backward
  // sweep's r.z finalize is left pending and done by k_cg_p_fin at the start
  // of the next iteration, the SpMV's p.q finalize by the forward sweep; each
  // moves the state between S.ks and S.fks (`cur` holds it), and flush()
  // brings it back to S.ks before the host reads it.
  const bool pfin = dilu && split && S.m->cone && sweep_slots <= 1024;
  bool pend = false;
  KS* cur = S.ks;
  auto other = [&](KS* k) { return k == S.ks ? S.fks : S.ks; };
  auto iteration = [&]() {
    T* pold = parity ? pb : pa;
    T* pnew = parity ? pa : pb;
    parity ^= 1;
    xcal_ctx* c = S.m->ctx;
    if (split && pend) {
      const KStep Kp{cur, other(cur), S.partial, sweep_slots, FIN_CG_ITER};
      hipLaunchKernelGGL(k_cg_p_fin<T>, dim3(pfin_grid(S.gridA)), dim3(kBlock), 0, S.s, S.A,
                         S.gridA, (const T*)zd, (const T*)pold, pnew, Kp);
      cur = other(cur);
      pend = false;
    } else if (split) {
      hipLaunchKernelGGL(k_cg_p<T>, dim3(S.gridA), dim3(kBlock), 0, S.s, S.A, (const T*)zd,
                         (const T*)pold, pnew, (const KS*)cur);
    }
    double* pp = F.part();
    const KStep K = F.step();
    timed(c, 0, [&] {
      with_row_group(S.A.rw, [&](auto G) {
        constexpr int g = decltype(G)::value;
        if (split)
          hipLaunchKernelGGL((k_cg_spmv<g, T>), dim3(S.gridA), dim3(kBlock), 0, S.s, S.A,
                             (const T*)pnew, q, S.partial, (const KS*)cur);
        else
          hipLaunchKernelGGL((k_cg_pspmv<g, T>), dim3(S.gridA), dim3(kBlock), 0, S.s, S.A,
                             (const T*)zd, (const T*)pold, pnew, q, pp, K);
      });
    });
    if (!pfin) F.done(pp, S.gridA, FIN_CG_PAP);
    if (dilu) {
      FwdK<T> W = fwdk(S, (const T*)nullptr, Y, Yb, (const KS*)cur);
      W.p = pnew;
      W.q = q;
      W.x = x;
      W.r = rv[rpar];
      if (S.m->cone) rpar ^= 1;
      W.rout = rv[rpar];
      const T* rnew = rv[rpar];
      if (pfin) {  // the forward sweep does the p.q finalize
        const KStep Kq{cur, other(cur), S.partial, S.gridA, FIN_CG_PAP};
        cur = other(cur);
        timed(c, 1, [&] { sweep_fwd<true>(S, W, &Kq); });
        timed(c, 2, [&] { sweep_bwd<1>(S, bwdk(S, (const T*)Yb, zd, rnew, (const KS*)cur)); });
        pend = true;
      } else {
        timed(c, 1, [&] { sweep_fwd<true>(S, W); });
        timed(c, 2, [&] { sweep_bwd<1>(S, bwdk(S, (const T*)Yb, zd, rnew, (const KS*)S.ks)); });
        finalize_launch(S, sweep_slots, FIN_CG_ITER);
      }
    } else {
      double* pu = F.part();
      const KStep Ku = F.step();
      hipLaunchKernelGGL(k_cg_update_diag<T>, dim3(S.grid), dim3(kBlock), 0, S.s, n, Sd,
                         (const T*)pnew, (const T*)q, x, r, zd, pu, Ku);
      F.done(pu, S.grid, FIN_CG_ITER);
    }
  };
  KS st = read_state(S);
  if (st.stop == STOP_NONE)
    run_batches(S, S.cg_batch, 4, iteration, st, [&] {
      F.flush();
      if (pend || cur != S.ks)  // the pending r.z finalize, or a copy, into S.ks
        hipLaunchKernelGGL(k_finalize_to, dim3(1), dim3(kBlock), 0, S.s,
                           (const double*)S.partial, sweep_slots, (const KS*)cur, S.ks,
                           (int)(pend ? FIN_CG_ITER : FIN_NONE));
      pend = false;
      cur = S.ks;
    });
  if (x0) hipLaunchKernelGGL(k_add, dim3(S.grid), dim3(kBlock), 0, S.s, n, (const double*)x,
                             (double*)x0);
  if (xsol) *xsol = x;
  XCAL_HIP(hipGetLastError());
  niter = st.iter;
  return (st.stop == STOP_ITMAX || niter >= itmax) ? 1 : 0;
}

// Krylov.jl bicgstab! (left preconditioning, shadow residual c = b); x0 as
// in cg.
template <class T>
static int bicgstab(Solver<T>& S, int precon, const T* b, T* x0, int itmax, double atol,
                    double rtol, int& niter, T** xsol) {
  const int64_t n = S.n;
  T* w = S.work;
  T *r = w, *x = w + n, *pa = w + 2 * n, *pb = w + 3 * n, *v = w + 4 * n, *sv = w + 5 * n,
    *t = w + 6 * n, *d = w + 7 * n, *Y = w + 8 * n, *Yb = w + 9 * n;
  init_state(S, itmax, atol, rtol);
  hipLaunchKernelGGL(k_start<T>, dim3(S.gridA), dim3(kBlock), 0, S.s, S.A, b, (const T*)x0, r, x, pa);
  const bool dilu = precon == XCAL_PRECON_DILU;
  const T* Sd = precon == XCAL_PRECON_JACOBI ? S.Sj : nullptr;
  const int sweep_slots = S.m->sweep_slots;
  const KS* ks = S.ks;
  if (dilu) {
    // r = M r0 (in place: the backward sweep reads only Y) ; ||r||^2, <b, r>
    sweep_fwd<false>(S, fwdk(S, (const T*)r, Y, Yb, ks));
    sweep_bwd<4>(S, bwdk(S, (const T*)Yb, r, b, ks));
    finalize_launch(S, sweep_slots, FIN_BI_INIT);
  } else {
    hipLaunchKernelGGL((k_precond_diag_start<true, T>), dim3(S.grid), dim3(kBlock), 0, S.s, n, Sd, r,
                       (T*)nullptr, b, S.partial);
    finalize_launch(S, S.grid, FIN_BI_INIT);
  }
  int parity = 0;
  Fuse<T> F(S, S.m->fuse && !dilu && S.gridA <= kFuseBlocks && S.grid <= kFuseBlocks);
  auto iteration = [&]() {
    T* pold = parity ? pb : pa;
    T* pnew = parity ? pa : pb;
    parity ^= 1;
    if (dilu) {
      with_row_group(S.A.rw, [&](auto G) {  // d holds q
        constexpr int g = decltype(G)::value;
        hipLaunchKernelGGL((k_bi_spmv_p<g, false, T>), dim3(S.gridA), dim3(kBlock), 0,
                           S.s, S.A,
                           (const T*)r, (const T*)pold, pnew, (const T*)nullptr, d, b, S.partial,
                           kstep_plain(ks));
      });
      sweep_fwd<false>(S, fwdk(S, (const T*)d, Y, Yb, ks));
      sweep_bwd<2>(S, bwdk(S, (const T*)Yb, v, b, ks));
      finalize_launch(S, sweep_slots, FIN_BI_ALPHA);
      with_row_group(S.A.rw, [&](auto G) {
        constexpr int g = decltype(G)::value;
        hipLaunchKernelGGL((k_bi_spmv_s<g, false, T>), dim3(S.gridA), dim3(kBlock), 0,
                           S.s, S.A,
                           (const T*)r, (const T*)v, (const T*)pnew, sv, x, (const T*)nullptr, d,
                           S.partial, kstep_plain(ks));
      });
      sweep_fwd<false>(S, fwdk(S, (const T*)d, Y, Yb, ks));
      sweep_bwd<3>(S, bwdk(S, (const T*)Yb, t, (const T*)sv, ks));
      finalize_launch(S, sweep_slots, FIN_BI_OMEGA);
    } else {
      double* pp = F.part();
      const KStep Kp = F.step();
      with_row_group(S.A.rw, [&](auto G) {
        constexpr int g = decltype(G)::value;
        hipLaunchKernelGGL((k_bi_spmv_p<g, true, T>), dim3(S.gridA), dim3(kBlock), 0,
                           S.s, S.A,
                           (const T*)r, (const T*)pold, pnew, Sd, v, b, pp, Kp);
      });
      F.done(pp, S.gridA, FIN_BI_ALPHA);
      double* ps = F.part();
      const KStep Ks = F.step();
      with_row_group(S.A.rw, [&](auto G) {
        constexpr int g = decltype(G)::value;
        hipLaunchKernelGGL((k_bi_spmv_s<g, true, T>), dim3(S.gridA), dim3(kBlock), 0,
                           S.s, S.A,
                           (const T*)r, (const T*)v, (const T*)pnew, sv, x, Sd, t, ps, Ks);
      });
      F.done(ps, S.gridA, FIN_BI_OMEGA);
    }
    double* pt = F.part();
    const KStep Kt = F.step();
    hipLaunchKernelGGL(k_bi_tail<T>, dim3(S.grid), dim3(kBlock), 0, S.s, n, (const T*)sv,
                       (const T*)t, (const T*)v, x, r, pnew, b, pt, Kt);
    F.done(pt, S.grid, FIN_BI_END);
  };
  KS st = read_state(S);
  if (st.stop == STOP_NONE) run_batches(S, S.bi_batch, 2, iteration, st, [&] { F.flush(); });
  if (x0) hipLaunchKernelGGL(k_add, dim3(S.grid), dim3(kBlock), 0, S.s, n, (const double*)x,
                             (double*)x0);
  if (xsol) *xsol = x;
  XCAL_HIP(hipGetLastError());
  niter = st.iter;
  return (st.stop == STOP_ITMAX || niter >= itmax) ? 1 : 0;
}

// |b - A x|^2 and |b|^2 in FP64; r32 = (float)(b - A x)
static void ir_residual(Solver<double>& S, const double* x, const double* b, float* r32,
                        double& rr, double& bb) {
  xcal_ctx* c = S.m->ctx;
  hipLaunchKernelGGL(k_ir_residual, dim3(S.gridA), dim3(kBlock), 0, S.s, S.A, x, b, r32, S.partial);
  hipLaunchKernelGGL(k_sum2, dim3(1), dim3(kRunBlock), 0, S.s, S.partial, S.gridA, c->red_result.p);
  XCAL_HIP(hipGetLastError());
  XCAL_HIP(hipMemcpyAsync(c->host_pinned, c->red_result.p, 2 * sizeof(double),
                          hipMemcpyDeviceToHost, S.s));
  XCAL_HIP(hipStreamSynchronize(S.s));
  rr = c->host_pinned[0];
  bb = c->host_pinned[1];
}

// Mixed precision: FP64 iterative refinement around FP32 Krylov solves.
//   r = b - A x (FP64 SpMV) ; solve A d = r in FP32 (same Krylov method and
//   preconditioner, to ir_inner_rtol) ; x += d (FP64).
// Stops when |b - A x| <= atol + rtol |b - A x0| (true residual, FP64), or
// on itmax summed inner iterations, or when a sweep no longer reduces the
// residual (FP32 exhausted); the last two return the itmax warning.
static int ir_solve(xcal_eqn* e, const xcal_solver_setup* st, int itmax, double* x,
                    const double* b, int& niter, int& nouter) {
  Solver<double> S = make_solver(e);
  Solver<float> S32 = make_solver32(e);
  float* r32 = e->work32.p + 11 * S.n;  // not used by the FP32 Krylov workspace (10 n)
  // an inner solve need not be tighter than the outer tolerance
  const double inner_rtol = std::max(st->ir_inner_rtol > 0 ? st->ir_inner_rtol : 1e-6, st->rtol);
  double rr, bb;
  ir_residual(S, x, b, r32, rr, bb);
  const double tol = st->atol + st->rtol * std::sqrt(rr);
  double res = std::sqrt(rr), prev = INFINITY;
  niter = 0;
  nouter = 0;
  int rc = 0;
  // A loose tolerance (rtol >= 1e-3) lies well inside FP32's reach: one FP32
  // solve with the FP64 solver's own stopping rule (atol + rtol times its
  // initial preconditioned norm) replaces the refinement loop.
  const bool single = st->rtol >= kIrSingleSweep;
  while (res > tol) {
    if (niter >= itmax || nouter >= 100 || res >= 0.5 * prev || (single && nouter == 1)) {
      rc = niter >= itmax ? 1 : (single ? 0 : 1);
      break;
    }
    int ni = 0;
    float* d = nullptr;
    S32.cg_batch = e->pred_slot(x, 16 + 2 * std::min(nouter, 15));
    S32.bi_batch = e->pred_slot(x, 17 + 2 * std::min(nouter, 15));
    const double ia = single ? st->atol : 0.0;
    // a sweep reduces its residual by inner_rtol, or only by what the outer
    // tolerance still asks (with a factor 2 of margin for the FP32
    // recurrence's drift from the true residual): the last sweep stops
    // there instead of overshooting by decades
    const double rt = single ? inner_rtol : std::max(inner_rtol, std::min(0.5 * tol / res, 0.5));
    if (st->solver == XCAL_SOLVER_CG)
      cg(S32, st->precon, (const float*)r32, (float*)nullptr, itmax - niter, ia, rt, ni, &d);
    else
      bicgstab(S32, st->precon, (const float*)r32, (float*)nullptr, itmax - niter, ia, rt, ni, &d);
    niter += ni;
    hipLaunchKernelGGL(k_ir_update, dim3(S.grid), dim3(kBlock), 0, S.s, S.n, (const float*)d, x);
    prev = res;
    ir_residual(S, x, b, r32, rr, bb);
    res = std::sqrt(rr);
    ++nouter;
  }
  return rc;
}

// The steps of solve_system! (Solve_1_api.jl:268-307) around the Krylov
// solve: apply_smoother! on the field first (:285), and for an equation whose
// first term is Time{CrankNicolson}, x = 2 x - x_old after it (:291-296).
// Workspace: vectors 10 and 11 of the equation's FP64 Krylov workspace,
// which no Krylov form uses.
void solve_prologue(xcal_eqn* e, const xcal_solver_setup* st, double* x, const double* b) {
  XCAL_CHECK(st->smoother_loops >= 0, "smoother loops must be >= 0");
  const xcal_mesh* m = e->mesh;
  const int64_t n = m->n;
  if (st->smoother_loops == 0 && !e->cn) return;
  hipStream_t s = m->ctx->stream;
  if (e->work.n < 12 * n) {
    XCAL_HIP(hipStreamSynchronize(s));  // a replaced buffer may be in use
    e->work.alloc(12 * n);
  }
  if (st->smoother_loops > 0) {
    const MatK A = matk(m, (const double*)e->nzval.p);
    double* src = x;
    double* dst = e->work.p + 10 * n;
    for (int l = 0; l < st->smoother_loops; ++l) {
      with_row_group(A.rw, [&](auto G) {
        hipLaunchKernelGGL((k_jacobi_smooth<decltype(G)::value>), dim3(m->nxblk), dim3(kBlock), 0,
                           s, A, b, (const double*)src, dst, st->smoother_omega);
      });
      std::swap(src, dst);
    }
    if (src != x)
      XCAL_HIP(hipMemcpyAsync(x, src, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  }
  if (e->cn)
    XCAL_HIP(hipMemcpyAsync(e->work.p + 11 * n, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  XCAL_HIP(hipGetLastError());
}

void solve_epilogue(xcal_eqn* e, double* x) {
  if (!e->cn) return;
  const xcal_mesh* m = e->mesh;
  hipLaunchKernelGGL(k_cn_step, dim3(grid_for(m->n)), dim3(kBlock), 0, m->ctx->stream, m->n,
                     (const double*)(e->work.p + 11 * m->n), x);
  XCAL_HIP(hipGetLastError());
}

int solve_system(xcal_eqn* e, const xcal_solver_setup* st, double* x, const double* b,
                 xcal_solve_stats* stats) {
  XCAL_CHECK(st->precision == XCAL_PREC_FP64 || st->precision == XCAL_PREC_FP32_IR ||
                 st->precision == XCAL_PREC_FP32_PRECON,
             "unknown precision mode");
  XCAL_CHECK(st->solver == XCAL_SOLVER_CG || st->solver == XCAL_SOLVER_BICGSTAB,
             "unknown Krylov solver");
  XCAL_CHECK(st->precon == e->precon,
             "preconditioner of the solver setup has not been updated on this equation");
  if (e->mesh->batch) {  // batch mesh: the worst case is reported
    if (st->precision == XCAL_PREC_FP32_PRECON)
      throw Error(XCAL_ERR_UNSUPPORTED,
                  "the FP32 preconditioner form runs on single meshes; batch meshes take FP32_IR");
    std::vector<xcal_solve_stats> cs(e->mesh->ncase);
    const int rc = solve_system_batch(e, st, x, b, cs.data());
    if (stats) {
      *stats = cs[0];
      for (const auto& c : cs) {
        stats->niter = std::max(stats->niter, c.niter);
        stats->nouter = std::max(stats->nouter, c.nouter);
        stats->residual = std::max(stats->residual, c.residual);
      }
    }
    return rc;
  }
  const int itmax = st->itmax == 0 ? (int)(2 * e->mesh->n) : st->itmax;
  int niter = 0, nouter = 1, rc;
  solve_prologue(e, st, x, b);
  if (st->precision == XCAL_PREC_FP32_IR) {
    rc = ir_solve(e, st, itmax, x, b, niter, nouter);
  } else {
    Solver<double> S = make_solver(e);
    // CG only: its residual recurrence r -= alpha A p uses the FP64 matrix,
    // so FP32 sweeps change the search directions and not the residual.
    // Left-preconditioned BiCGSTAB updates M^-1 r through the sweeps, where
    // FP32 rounding would move the residual itself (1.7e-8 in the solution
    // at rtol 1e-12 on a 110 592-cell box), so it keeps FP64 sweeps.
    const bool pc32 = st->precision == XCAL_PREC_FP32_PRECON;
    if (pc32 && st->solver == XCAL_SOLVER_CG && st->precon == XCAL_PRECON_DILU &&
        e->mesh->cone) {
      cone_factors32(e);
      S.pc32 = true;
      S.cfa32 = e->cfa32.p;
      S.cba32 = e->cba32.p;
      S.cbd32 = e->cbd32.p;
    }
    S.cg_batch = e->pred_slot(x, pc32 ? 2 : 0);
    S.bi_batch = e->pred_slot(x, pc32 ? 3 : 1);
    if (st->solver == XCAL_SOLVER_CG)
      rc = cg(S, st->precon, b, x, itmax, st->atol, st->rtol, niter, (double**)nullptr);
    else
      rc = bicgstab(S, st->precon, b, x, itmax, st->atol, st->rtol, niter, (double**)nullptr);
  }
  solve_epilogue(e, x);
  if (stats) {
    stats->niter = niter;
    stats->nouter = nouter;
    stats->residual = residual(e, x, b);
  }
  return rc;
}

static const double* cell_vec(const xcal_eqn* e, const xcal_field* f, int c, const char* what) {
  XCAL_CHECK(f && f->mesh == e->mesh, std::string(what) + ": field missing or on another mesh");
  XCAL_CHECK(f->kind == XCAL_CELL_SCALAR || f->kind == XCAL_CELL_VECTOR,
             std::string(what) + ": must be a cell field");
  XCAL_CHECK(c >= 0 && c < f->ncomp, std::string(what) + ": component out of range");
  return f->comp(c);
}

}  // namespace xcal

using namespace xcal;

extern "C" {

int xcal_precond_apply(xcal_eqn* e, const xcal_field* r, int32_t rc, xcal_field* z, int32_t zc) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    const double* rp = cell_vec(e, r, rc, "r");
    double* zp = const_cast<double*>(cell_vec(e, z, zc, "z"));
    XCAL_CHECK(rp != zp, "r and z must not alias");
    if (e->precon == XCAL_PRECON_DILU) {
      Solver<double> S = make_solver(e);
      dilu_apply(S, rp, zp, S.work, S.work + S.n);
    } else {
      apply_M_levels(e, rp, zp);
    }
    return 0;
  });
}

int xcal_kernel_timing_enable(xcal_ctx* c, int32_t kernel, int32_t enable) {
  return guarded([&] {
    XCAL_CHECK(c, "context is NULL");
    XCAL_CHECK(kernel >= 0 && kernel <= 2, "kernel id must be 0..2");
    XCAL_HIP(hipStreamSynchronize(c->stream));
    prof_resolve(c);
    c->prof_on[kernel] = enable != 0;
    c->prof_seq[kernel] = 0;
    c->prof_launches[kernel] = 0;
    c->prof_ms[kernel] = 0.0;
    return 0;
  });
}

int xcal_kernel_timing_read(xcal_ctx* c, int32_t kernel, int64_t* launches, double* total_ms) {
  return guarded([&] {
    XCAL_CHECK(c && launches && total_ms, "NULL argument");
    XCAL_CHECK(kernel >= 0 && kernel <= 2, "kernel id must be 0..2");
    XCAL_HIP(hipStreamSynchronize(c->stream));
    prof_resolve(c);
    *launches = c->prof_launches[kernel];
    *total_ms = c->prof_ms[kernel];
    return 0;
  });
}

int xcal_solve_system_batch(xcal_eqn* e, const xcal_solver_setup* s, xcal_field* phi,
                            int32_t comp, xcal_solve_stats* stats) {
  return guarded([&] {
    XCAL_CHECK(e && s && stats, "NULL argument");
    check_field(phi, e->mesh, e->vector ? XCAL_CELL_VECTOR : XCAL_CELL_SCALAR, "phi");
    XCAL_CHECK(comp >= 0 && comp < (e->vector ? 3 : 1), "component out of range");
    XCAL_CHECK(s->precon == e->precon,
               "preconditioner of the solver setup has not been updated on this equation");
    XCAL_CHECK(s->solver == XCAL_SOLVER_CG || s->solver == XCAL_SOLVER_BICGSTAB,
               "unknown Krylov solver");
    return solve_system_batch(e, s, phi->comp(comp), e->bc(e->vector ? comp : 0), stats);
  });
}

int xcal_solve_system(xcal_eqn* e, const xcal_solver_setup* s, xcal_field* phi, int32_t comp,
                      xcal_solve_stats* stats) {
  return guarded([&] {
    XCAL_CHECK(e && s, "NULL argument");
    check_field(phi, e->mesh, e->vector ? XCAL_CELL_VECTOR : XCAL_CELL_SCALAR, "phi");
    XCAL_CHECK(comp >= 0 && comp < (e->vector ? 3 : 1), "component out of range");
    xcal_solve_stats st{};
    int rc = solve_system(e, s, phi->comp(comp), e->bc(e->vector ? comp : 0), &st);
    if (stats) *stats = st;
    return rc;
  });
}

}  // extern "C"
