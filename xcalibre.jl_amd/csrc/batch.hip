// Batched Krylov solves: the independent systems of a batch mesh (one per
// case) solved together, with the same CG / BiCGSTAB iterations, Jacobi /
// DILU preconditioners and FP64 or FP32-inner-refinement modes as the
// single-mesh solver (solver.hip), in one launch per step for all cases.
//
//  * Every case has its own Krylov scalars (KS) and stops on its own; the
//    kernels of a stopped case skip their stores (sweeps skip the case).
//  * Workgroups never straddle two cases (a block table maps each workgroup
//    to a row range of one case), and each case's reductions are summed in a
//    fixed order over its own workgroups, so a case's iterates do not depend
//    on which other cases share the batch.
//  * DILU sweeps: one workgroup per case walks that case's levels with a
//    barrier between levels (cases of a batch are small; their levels are
//    many and short), all cases in one launch.
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <functional>
#include <map>
#include <mutex>
#include <vector>

#include "solver_common.h"
#include "sweep_pipe.h"

namespace xcal {

// workgroup b covers rows [row0[b], row1[b]) of case kcase[b]
struct BlkK {
  const int32_t* row0;
  const int32_t* row1;
  const int32_t* kcase;
};

// Sum of segment k ([seg[k], seg[k+1])) of the partial pairs by one
// wavefront: fixed per-lane strides, then a fixed shuffle tree (the result
// depends only on the segment).  A lane loads U of its slots together, then
// adds them in slot order, so a long segment takes few round trips.
__device__ inline void wave_sum_segment(const double* __restrict__ partial, int32_t j0, int32_t j1,
                                        double& t0, double& t1) {
  constexpr int U = 8;
  t0 = 0.0;
  t1 = 0.0;
  for (int32_t jb = j0 + (int32_t)threadIdx.x; jb < j1; jb += 64 * U) {
    double a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t j = jb + 64 * u < j1 ? jb + 64 * u : j0;
      a[u] = partial[2 * j];
      b[u] = partial[2 * j + 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (jb + 64 * u < j1) {
        t0 += a[u];
        t1 += b[u];
      }
  }
  for (int o = 32; o > 0; o >>= 1) {
    t0 += __shfl_down(t0, o, 64);
    t1 += __shfl_down(t1, o, 64);
  }
}

// one wavefront per case: sum its partials and run its scalar update
__global__ void __launch_bounds__(64)
    kb_finalize(const double* __restrict__ partial, const int32_t* __restrict__ seg, int ncase,
                KS* ks, int fin) {
  const int k = blockIdx.x;
  KS st;
  if (threadIdx.x == 0) st = ks[k];  // loaded while the partials load
  double t0, t1;
  wave_sum_segment(partial, seg[k], seg[k + 1], t0, t1);
  if (threadIdx.x == 0 && st.stop == STOP_NONE) {
    finalize(&st, fin, t0, t1);
    ks[k] = st;
  }
}

// The Krylov state of case k a kernel of an iteration works with.  Unfused
// (fin == FIN_NONE, out == nullptr) it is in[k], as kb_finalize left it.
// Fused, the kernel first does the pending finalize itself: every workgroup
// of the case sums the case's pending partials as kb_finalize does (one
// wavefront, the same order) and applies the same scalar update, so all
// hold the same state, and the case's first workgroup stores it to out[k]
// (never in[k], which the case's other workgroups are still reading).
struct BStep {
  const KS* in;
  KS* out;
  const double* pin;
  int fin;
};
inline BStep bstep_plain(const KS* ks) { return BStep{ks, nullptr, nullptr, FIN_NONE}; }
__device__ inline KS bstate(const BStep& K, int k, const int32_t* __restrict__ seg, double* lds) {
  KS st = K.in[k];
  if (K.fin != FIN_NONE) {  // uniform over the launch
    double t0 = 0.0, t1 = 0.0;
    if (threadIdx.x < 64) wave_sum_segment(K.pin, seg[k], seg[k + 1], t0, t1);
    if (threadIdx.x == 0) {
      lds[0] = t0;
      lds[1] = t1;
    }
    __syncthreads();
    t0 = lds[0];
    t1 = lds[1];
    if (st.stop == STOP_NONE) finalize(&st, K.fin, t0, t1);
  }
  if (K.out && (int32_t)blockIdx.x == seg[k] && threadIdx.x == 0) K.out[k] = st;
  return st;
}

// The pending finalize of a fused chain (fin, partial) applied to in[k] and
// stored to out[k] (in place allowed), before the host reads the states.
__global__ void __launch_bounds__(64)
    kb_flush(const double* __restrict__ partial, const int32_t* __restrict__ seg, const KS* in,
             KS* out, int fin) {
  const int k = blockIdx.x;
  KS st = in[k];
  double t0 = 0.0, t1 = 0.0;
  if (fin != FIN_NONE) wave_sum_segment(partial, seg[k], seg[k + 1], t0, t1);
  if (threadIdx.x == 0) {
    if (fin != FIN_NONE && st.stop == STOP_NONE) finalize(&st, fin, t0, t1);
    out[k] = st;
  }
}

// per-case sums of the partial pairs -> out[2k], out[2k+1]
__global__ void __launch_bounds__(64)
    kb_sum_cases(const double* __restrict__ partial, const int32_t* __restrict__ seg, int ncase,
                 double* __restrict__ out) {
  const int k = blockIdx.x;
  double t0, t1;
  wave_sum_segment(partial, seg[k], seg[k + 1], t0, t1);
  if (threadIdx.x == 0) {
    out[2 * k] = t0;
    out[2 * k + 1] = t1;
  }
}

// r = b - A x0 (r = b when x0 is NULL) ; x = 0 ; pold = 0
template <class T>
__global__ void __launch_bounds__(kBlock)
    kb_start(MatKT<T> A, const T* __restrict__ b, const T* __restrict__ x0, T* __restrict__ r,
             T* __restrict__ x, T* __restrict__ pold) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  if (x0) {
    const T acc = row_sum(A, i, [&](int j) { return x0[j]; });
    r[i] = b[i] - acc;
  } else {
    r[i] = b[i];
  }
  x[i] = 0;
  pold[i] = 0;
}

template <bool BICG, class T>
__global__ void __launch_bounds__(kBlock)
    kb_precond_diag_start(BlkK B, const T* __restrict__ S, T* __restrict__ r, T* __restrict__ z,
                          const T* __restrict__ c, double* partial) {
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  if (i < B.row1[blockIdx.x]) {
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

// p = z + beta p per case, over the rows each workgroup of kb_cg_spmv reads
// as its own
template <class T>
__global__ void __launch_bounds__(kBlock)
    kb_cg_p(BlkK B, const T* __restrict__ z, const T* __restrict__ pold, T* __restrict__ pnew,
            const KS* ks) {
  const KS* K = ks + B.kcase[blockIdx.x];
  const bool live = !stopped(K);
  const T beta = (T)K->beta;
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  if (i < B.row1[blockIdx.x] && live) pnew[i] = z[i] + beta * pold[i];
}

// q = A p ; p.q per case (one gather per entry; p formed by kb_cg_p)
template <int G, class T>
__global__ void __launch_bounds__(kBlock)
    kb_cg_spmv(BlkK B, MatKT<T> A, const T* __restrict__ p, T* __restrict__ q, double* partial,
               const KS* ks) {
  const KS* K = ks + B.kcase[blockIdx.x];
  const bool live = !stopped(K);
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  double s0 = 0.0;
  if (i < B.row1[blockIdx.x]) {
    const T pi = p[i];  // issued with the row's first loads
    const T acc = row_sum<G>(A, i, [&](int j) { return p[j]; });
    if (live) q[i] = acc;
    s0 = dd(pi, acc);
  }
  store_partial<kBlock>(s0, 0.0, partial, blockIdx.x);
}

// Fused form of kb_cg_p + kb_cg_spmv (fused chains): p = z + beta p formed
// for the row and for each neighbour it gathers, in the same arithmetic as
// kb_cg_p, so q and p.q equal the separate kernels' bit for bit.
template <int G, class T>
__global__ void __launch_bounds__(kBlock)
    kb_cg_pspmv(BlkK B, MatKT<T> A, const T* __restrict__ z, const T* __restrict__ pold,
                T* __restrict__ pnew, T* __restrict__ q, double* partial, BStep K,
                const int32_t* __restrict__ seg) {
  __shared__ double lds[2];
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  const bool in = i < B.row1[blockIdx.x];
  const int32_t ic = in ? i : B.row0[blockIdx.x];
  // the row's loads go out before the state is formed (its partials' round
  // trip overlaps the row's)
  const RowPair<G, T> R = row_pair_load<G>(A, ic, z, pold);
  const T zi = z[ic], poi = pold[ic];
  const KS st = bstate(K, B.kcase[blockIdx.x], seg, lds);
  const bool live = st.stop == STOP_NONE;
  const T beta = (T)st.beta;
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

template <class T>
__global__ void __launch_bounds__(kBlock)
    kb_cg_update_diag(BlkK B, const T* __restrict__ S, const T* __restrict__ p,
                      const T* __restrict__ q, T* __restrict__ x, T* __restrict__ r,
                      T* __restrict__ z, double* partial, BStep Kst,
                      const int32_t* __restrict__ seg) {
  __shared__ double lds[2];
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  const bool in = i < B.row1[blockIdx.x];
  const int32_t ic = in ? i : B.row0[blockIdx.x];
  const T x0 = x[ic], p0 = p[ic], r0 = r[ic], q0 = q[ic], S0 = S ? S[ic] : T(1);
  const KS st = bstate(Kst, B.kcase[blockIdx.x], seg, lds);
  const KS* K = &st;
  const bool live = !stopped(K);
  const T a = (T)K->alpha;
  double s0 = 0.0;
  if (in) {
    const T xi = x0 + a * p0;
    const T ri = r0 + (-a) * q0;
    const T zi = S ? S0 * ri : ri;
    if (live) {
      x[i] = xi;
      r[i] = ri;
      z[i] = zi;
    }
    s0 = dd(ri, zi);
  }
  store_partial<kBlock>(s0, 0.0, partial, blockIdx.x);
}

template <int G, bool VDIAG, class T>
__global__ void __launch_bounds__(kBlock)
    kb_bi_spmv_p(BlkK B, MatKT<T> A, const T* __restrict__ r, const T* __restrict__ pold,
                 T* __restrict__ pnew, const T* __restrict__ S, T* __restrict__ qv,
                 const T* __restrict__ c, double* partial, BStep Kst,
                 const int32_t* __restrict__ seg) {
  __shared__ double lds[2];
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  const bool in = i < B.row1[blockIdx.x];
  const int32_t ic = in ? i : B.row0[blockIdx.x];
  const RowPair<G, T> R = row_pair_load<G>(A, ic, r, pold);
  const T ri0 = r[ic], poi = pold[ic];
  const T Si = VDIAG && S ? S[ic] : T(1), ci = VDIAG ? c[ic] : T(0);
  const KS st = bstate(Kst, B.kcase[blockIdx.x], seg, lds);
  const KS* K = &st;
  const bool live = !stopped(K);
  const T beta = (T)K->beta;
  double s0 = 0.0;
  if (in) {
    const T acc = row_pair_sum(A, R, r, pold, [&](T u, T w) { return u + beta * w; });
    const T pi = ri0 + beta * poi;
    const T vi = VDIAG && S ? Si * acc : acc;
    if (live) {
      pnew[i] = pi;
      qv[i] = vi;
    }
    if (VDIAG) s0 = dd(ci, vi);
  }
  if (VDIAG) store_partial<kBlock>(s0, 0.0, partial, blockIdx.x);
}

template <int G, bool TDIAG, class T>
__global__ void __launch_bounds__(kBlock)
    kb_bi_spmv_s(BlkK B, MatKT<T> A, const T* __restrict__ r, const T* __restrict__ v,
                 const T* __restrict__ p, T* __restrict__ s, T* __restrict__ x,
                 const T* __restrict__ S, T* __restrict__ dt, double* partial, BStep Kst,
                 const int32_t* __restrict__ seg) {
  __shared__ double lds[2];
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  const bool in = i < B.row1[blockIdx.x];
  const int32_t ic = in ? i : B.row0[blockIdx.x];
  const RowPair<G, T> R = row_pair_load<G>(A, ic, r, v);
  const T ri0 = r[ic], vi0 = v[ic], xi0 = x[ic], pi0 = p[ic];
  const T Si = TDIAG && S ? S[ic] : T(1);
  const KS st = bstate(Kst, B.kcase[blockIdx.x], seg, lds);
  const KS* K = &st;
  const bool live = !stopped(K);
  const T a = (T)K->alpha;
  double s0 = 0.0, s1 = 0.0;
  if (in) {
    const T acc = row_pair_sum(A, R, r, v, [&](T u, T w) { return u + (-a) * w; });
    const T si = ri0 + (-a) * vi0;
    const T xi = xi0 + a * pi0;
    const T ti = TDIAG && S ? Si * acc : acc;
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

template <class T>
__global__ void __launch_bounds__(kBlock)
    kb_bi_tail(BlkK B, const T* __restrict__ s, const T* __restrict__ t, const T* __restrict__ v,
               T* __restrict__ x, T* __restrict__ r, T* __restrict__ p, const T* __restrict__ c,
               double* partial, BStep Kst, const int32_t* __restrict__ seg) {
  __shared__ double lds[2];
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  const bool in = i < B.row1[blockIdx.x];
  const int32_t ic = in ? i : B.row0[blockIdx.x];
  const T x0 = x[ic], s0v = s[ic], t0 = t[ic], p0 = p[ic], v0 = v[ic], c0 = c[ic];
  const KS st = bstate(Kst, B.kcase[blockIdx.x], seg, lds);
  const KS* K = &st;
  const bool live = !stopped(K);
  const T w = (T)K->omega;
  double s0 = 0.0, s1 = 0.0;
  if (in) {
    const T xi = x0 + w * s0v;
    const T ri = s0v + (-w) * t0;
    const T pi = p0 + (-w) * v0;
    if (live) {
      x[i] = xi;
      r[i] = ri;
      p[i] = pi;
    }
    s0 = dd(c0, ri);
    s1 = dd(ri, ri);
  }
  store_partial<kBlock>(s0, s1, partial, blockIdx.x);
}

// one workgroup per case: the case's forward levels in order
template <bool UPD, class T>
__global__ void __launch_bounds__(kRunBlock)
    kb_fwd(FwdK<T> W, const int32_t* __restrict__ lev_ptr, const int32_t* __restrict__ meta,
           const int32_t* __restrict__ case_lev, const KS* ks) {
  const int k = blockIdx.x;
  if (ks && stopped(ks + k)) return;  // uniform over the workgroup
  const T a = UPD ? (T)ks[k].alpha : T(0);
  for (int u = case_lev[k]; u < case_lev[k + 1]; ++u) {
    const int32_t r0 = lev_ptr[u], cnt = lev_ptr[u + 1] - r0, base = meta[2 * u], w = meta[2 * u + 1];
    for (int32_t q = threadIdx.x; q < cnt; q += blockDim.x)
      fwd_row<UPD, 0>(W, r0 + q, (int64_t)base + q, cnt, w, a, true);
    __syncthreads();
  }
}

// one workgroup per case: the case's backward levels; dots into slot k
template <int DOT, class T>
__global__ void __launch_bounds__(kRunBlock)
    kb_bwd(BwdK<T> W, const int32_t* __restrict__ lev_ptr, const int32_t* __restrict__ meta,
           const int32_t* __restrict__ case_lev, const KS* ks) {
  const int k = blockIdx.x;
  if (ks && stopped(ks + k)) return;
  double s0 = 0.0, s1 = 0.0;
  for (int u = case_lev[k]; u < case_lev[k + 1]; ++u) {
    const int32_t p0 = lev_ptr[u], cnt = lev_ptr[u + 1] - p0, base = meta[2 * u], w = meta[2 * u + 1];
    for (int32_t q = threadIdx.x; q < cnt; q += blockDim.x)
      bwd_row<DOT, 0>(W, p0 + q, (int64_t)base + q, cnt, w, true, s0, s1);
    __syncthreads();
  }
  if (DOT != 0) store_partial<kRunBlock>(s0, s1, W.partial, k);
}

// ------------------------------------------- pipelined per-case sweeps
// One workgroup per case walks the case's levels with the pipelined bodies of
// sweep_pipe.h (same results as kb_fwd/kb_bwd/kb_dilu_diag).  Within a case
// nothing depends on rows before it, so the ring starts empty.

// CG update of the pipelined path: x += a p ; r -= a q per case (the part of
// the fused forward row that does not depend on other rows)
template <class T>
__global__ void __launch_bounds__(kBlock)
    kb_cg_xr(BlkK B, T* __restrict__ x, const T* __restrict__ p, T* __restrict__ r,
             const T* __restrict__ q, const KS* ks) {
  const int k = B.kcase[blockIdx.x];
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  if (stopped(ks + k) || i >= B.row1[blockIdx.x]) return;
  const T a = (T)ks[k].alpha;
  const T xi = x[i] + a * p[i];
  const T ri = r[i] + (-a) * q[i];
  x[i] = xi;
  r[i] = ri;
}

// UPD: x and r were updated by kb_cg_xr; the sweep's input is u = r.  With
// FAR, y is also stored to W.Y (device order): a row gathers a dependency
// that lies before the LDS ring from there.
// RW: rows of a level per thread, kPipeBlock / RW threads per workgroup
// (the same rows, values and sums for every RW).
template <bool UPD, bool FAR, class T, int WK, int RW>
__global__ void __launch_bounds__(kPipeBlock / RW)
    kb_fwd_pipe(FwdK<T> W, const T* __restrict__ Lrd, const int32_t* __restrict__ lev_ptr,
                const int32_t* __restrict__ meta, const int32_t* __restrict__ case_lev,
                const KS* ks) {
  __shared__ PipeLDS<T> L;
  const int k = blockIdx.x;
  if (ks && stopped(ks + k)) return;  // uniform over the workgroup
  const int u0 = case_lev[k];
  fwd_pipe<FAR, FAR, T, WK, pipe_depth<WK, RW>(), RW>(W, UPD ? W.r : W.u, Lrd, lev_ptr, meta, u0,
                                                   case_lev[k + 1] - u0, false, L);
}

template <int DOT, bool FAR, class T, int WK, int RW>
__global__ void __launch_bounds__(kPipeBlock / RW)
    kb_bwd_pipe(BwdK<T> W, const int32_t* __restrict__ Upos, const int32_t* __restrict__ lev_ptr,
                const int32_t* __restrict__ meta, const int32_t* __restrict__ case_lev,
                const KS* ks) {
  __shared__ PipeLDS<T> L;
  const int k = blockIdx.x;
  if (ks && stopped(ks + k)) return;
  const int u0 = case_lev[k];
  double s0[RW], s1[RW];
#pragma unroll
  for (int j = 0; j < RW; ++j) s0[j] = s1[j] = 0.0;
  bwd_pipe<DOT, FAR, T, WK, pipe_depth<WK, RW>(), RW>(W, Upos, lev_ptr, meta, u0,
                                                   case_lev[k + 1] - u0, false, L, s0, s1);
  if (DOT != 0) store_partial_rows<kPipeBlock, RW>(s0, s1, W.partial, k);
}

template <bool FAR, int WK, int RW>
__global__ void __launch_bounds__(kPipeBlock / RW)
    kb_dilu_diag_pipe(double* D, const double* __restrict__ prod, const int32_t* __restrict__ Lcol,
                      const int32_t* __restrict__ lev_ptr, const int32_t* __restrict__ meta,
                      const int32_t* __restrict__ case_lev) {
  __shared__ PipeLDS<double> L;
  const int k = blockIdx.x;
  const int u0 = case_lev[k];
  diag_pipe<FAR, WK, pipe_depth<WK, RW>(), RW>(D, prod, Lcol, lev_ptr, meta, u0,
                                               case_lev[k + 1] - u0, false, L);
}

// Host: calls f(RW) for a rows-per-thread count of 1, 2 or 4.
template <class F>
inline void with_pipe_rows(int rw, F&& f) {
  if (rw == 4) f(std::integral_constant<int, 4>{});
  else if (rw == 2) f(std::integral_constant<int, 2>{});
  else f(std::integral_constant<int, 1>{});
}

// update_dilu_diagonal! per case: one workgroup walks the case's levels
__global__ void __launch_bounds__(kRunBlock)
    kb_dilu_diag(MatK A, const int32_t* __restrict__ trans, double* D,
                 const int32_t* __restrict__ lev_ptr, const int32_t* __restrict__ case_lev) {
  const int k = blockIdx.x;
  for (int u = case_lev[k]; u < case_lev[k + 1]; ++u) {
    const int32_t b = lev_ptr[u + 1];
    for (int32_t j = lev_ptr[u] + threadIdx.x; j < b; j += blockDim.x) {
      double Dj = D[j];
      for (int32_t s = A.rb[j]; s < A.diag[j]; s += kSlice) {
        const int i = A.colval[s];
        const int kt = trans[s];
        if (kt >= 0) Dj = Dj - (A.nz[kt] * A.nz[s]) / D[i];
      }
      D[j] = Dj;
    }
    __syncthreads();
  }
}

// per-workgroup partials of |b - A x|^2 and |b|^2; FP32 copy of b - A x
__global__ void __launch_bounds__(kBlock)
    kb_residual(BlkK B, MatK A, const double* __restrict__ x, const double* __restrict__ b,
                float* __restrict__ r32, double* partial) {
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  if (i < B.row1[blockIdx.x]) {
    const double acc = row_sum(A, (int64_t)i, [&](int j) { return x[j]; });
    const double r = b[i] - acc;
    if (r32) r32[i] = (float)r;
    s0 = r * r;
    s1 = b[i] * b[i];
  }
  store_partial<kBlock>(s0, s1, partial, blockIdx.x);
}

// x += d on the rows of active cases
__global__ void __launch_bounds__(kBlock)
    kb_ir_update(BlkK B, const int32_t* __restrict__ active, const float* __restrict__ d,
                 double* __restrict__ x) {
  if (!active[B.kcase[blockIdx.x]]) return;
  const int32_t i = B.row0[blockIdx.x] + threadIdx.x;
  if (i < B.row1[blockIdx.x]) x[i] = x[i] + (double)d[i];
}

__global__ void __launch_bounds__(kBlock)
    kb_add(int64_t n, const double* __restrict__ x, double* __restrict__ x0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x0[i] = x[i] + x0[i];
}

__global__ void __launch_bounds__(kBlock)
    kb_to_float(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ------------------------------------------------------------ host side
template <class T>
struct BSolver {
  xcal_eqn* e;
  const xcal_mesh* m;
  hipStream_t s;
  int64_t n;
  int ncase;
  MatKT<T> A;
  const T *Sj, *rD, *Lnz, *Unz, *Db;
  const T* Lrd;  // 1/D per packed lower entry (pipelined sweeps)
  T* work;
  KS* ks;
  double* partial;  // nblk slots (also used for the ncase sweep slots)
  KS* fks;          // fused chains: the second state array (ncase)
  double* fpart;    // fused chains: the second partial area (nblk slots)
  BlkK B;
  int nblk;
  int* pred;  // iterations of the last solve (first-batch size)
};

static void batch_workspace(xcal_eqn* e, bool fp32) {
  const xcal_mesh* m = e->mesh;
  const int64_t n = m->n;
  const int64_t slots = std::max<int64_t>(m->nblk, m->ncase) + 8;
  const int64_t ksb = (int64_t)sizeof(KS) * m->ncase;
  const bool grow = (e->work.n > 0 && e->work.n < 12 * n) ||
                    (e->kpart.n > 0 && e->kpart.n < 2 * slots) ||
                    (e->fpart.n > 0 && e->fpart.n < 2 * slots) ||
                    (e->fstate.n > 0 && e->fstate.n < ksb) ||
                    (e->kstate.n > 0 && e->kstate.n < ksb) ||
                    (fp32 && e->work32.n > 0 && e->work32.n < 12 * n);
  if (grow) XCAL_HIP(hipStreamSynchronize(m->ctx->stream));
  if (e->work.n < 12 * n) e->work.alloc(12 * n);
  if (e->kpart.n < 2 * slots) e->kpart.alloc(2 * slots);
  if (e->fpart.n < 2 * slots) e->fpart.alloc(2 * slots);
  if (e->fstate.n < ksb) e->fstate.alloc(ksb);
  if (e->kstate.n < ksb) {
    e->kstate.alloc(ksb);
    if (e->kstate_host) (void)hipHostFree(e->kstate_host);
    e->kstate_host = nullptr;
  }
  if (!e->kstate_host) XCAL_HIP(hipHostMalloc((void**)&e->kstate_host, ksb, 0));
  if (e->batch_pred.size() < 2 * 17) e->batch_pred.assign(2 * 17, 0);
  if (e->case_active.n < m->ncase) e->case_active.alloc(m->ncase);
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

static BlkK blkk(const xcal_mesh* m) { return BlkK{m->blk_row0.p, m->blk_row1.p, m->blk_case.p}; }

static BSolver<double> batch_solver(xcal_eqn* e) {
  batch_workspace(e, false);
  const xcal_mesh* m = e->mesh;
  BSolver<double> S;
  S.e = e;
  S.m = m;
  S.s = m->ctx->stream;
  S.n = m->n;
  S.ncase = m->ncase;
  S.A = matk(m, (const double*)e->nzval.p);
  S.Sj = e->pS.p;
  S.rD = e->prD.p;
  S.Lnz = e->Lnz.p;
  S.Unz = e->Unz.p;
  S.Db = e->Db.p;
  S.Lrd = e->Lrd.p;
  S.work = e->work.p;
  S.ks = (KS*)e->kstate.p;
  S.partial = e->kpart.p;
  S.fks = (KS*)e->fstate.p;
  S.fpart = e->fpart.p;
  S.B = blkk(m);
  S.nblk = (int)m->nblk;
  S.pred = &e->batch_pred[0];
  return S;
}

static BSolver<float> batch_solver32(xcal_eqn* e) {
  batch_workspace(e, true);
  const xcal_mesh* m = e->mesh;
  hipStream_t s = m->ctx->stream;
  const int64_t n = m->n;
  auto tofloat = [&](int64_t len, const double* a, float* b) {
    if (len > 0) hipLaunchKernelGGL(kb_to_float, dim3(grid_for(len)), dim3(kBlock), 0, s, len, a, b);
  };
  tofloat(m->nzlen, e->nzval.p, e->nz32.p);
  if (e->precon == XCAL_PRECON_DILU) {
    if (e->Lnz32.n != e->Lnz.n) e->Lnz32.alloc(e->Lnz.n);
    if (e->Unz32.n != e->Unz.n) e->Unz32.alloc(e->Unz.n);
    if (e->Db32.n != n) e->Db32.alloc(n);
    tofloat(n, e->prD.p, e->prD32.p);
    tofloat(n, e->Db.p, e->Db32.p);
    tofloat(e->Lnz.n, e->Lnz.p, e->Lnz32.p);
    tofloat(e->Unz.n, e->Unz.p, e->Unz32.p);
    if (m->pipe_wf > 0) {
      if (e->Lrd32.n != e->Lrd.n) e->Lrd32.alloc(e->Lrd.n);
      tofloat(e->Lrd.n, e->Lrd.p, e->Lrd32.p);
    }
  } else if (e->precon == XCAL_PRECON_JACOBI) {
    tofloat(n, e->pS.p, e->pS32.p);
  }
  XCAL_HIP(hipGetLastError());
  BSolver<float> S;
  S.e = e;
  S.m = m;
  S.s = s;
  S.n = n;
  S.ncase = m->ncase;
  S.A = matk(m, (const float*)e->nz32.p);
  S.Sj = e->pS32.p;
  S.rD = e->prD32.p;
  S.Lnz = e->Lnz32.p;
  S.Unz = e->Unz32.p;
  S.Db = e->Db32.p;
  S.Lrd = e->Lrd32.p;
  S.work = e->work32.p;
  S.ks = (KS*)e->kstate.p;
  S.partial = e->kpart.p;
  S.fks = (KS*)e->fstate.p;
  S.fpart = e->fpart.p;
  S.B = blkk(m);
  S.nblk = (int)m->nblk;
  S.pred = &e->batch_pred[2];
  return S;
}

template <class T>
static void fin_blocks(const BSolver<T>& S, int fin) {
  hipLaunchKernelGGL(kb_finalize, dim3(S.ncase), dim3(64), 0, S.s, S.partial, S.m->case_blk_d.p,
                     S.ncase, S.ks, fin);
}
template <class T>
static void fin_cases(const BSolver<T>& S, int fin) {
  hipLaunchKernelGGL(kb_finalize, dim3(S.ncase), dim3(64), 0, S.s, S.partial, S.m->case_iota_d.p,
                     S.ncase, S.ks, fin);
}

// Fused chains for batched CG and BiCGSTAB with a diagonal (or no)
// preconditioner: each kernel of an iteration does the pending finalize of
// the kernel before it (bstate), so an iteration takes 2 (CG) or 3
// (BiCGSTAB) launches instead of 5 or 6.  States alternate between S.ks and
// S.fks, partial areas between S.partial and S.fpart (a kernel never writes
// what another workgroup of it still reads); flush() applies the last
// pending finalize into S.ks before the host reads the states.  Results are
// those of the unfused chain bit for bit (the same sums in the same order,
// the same scalar updates).  Off with XCAL_NO_FUSE=1 at mesh creation, for
// DILU (its sweeps take the state from S.ks) and above kBatchFuseRows rows,
// where a second gather per entry costs more than a launch.
constexpr int64_t kBatchFuseRows = 1 << 20;

template <class T>
struct BFuse {
  const BSolver<T>& S;
  bool on;
  int cur = 0;         // the state array holding the current state: 0 S.ks, 1 S.fks
  int pend = -1;       // partial area of the pending finalize: 0 S.partial, 1 S.fpart
  int pfin = FIN_NONE;
  KS* ks(int a) const { return a ? S.fks : S.ks; }
  double* pa(int a) const { return a ? S.fpart : S.partial; }
  // a kernel that reads no state: its partials, then done(fin)
  double* start() {
    if (!on) return S.partial;
    pend = 0;
    return pa(0);
  }
  // a kernel of the chain: its state and, in `out`, its partial area
  BStep step(double*& out) {
    if (!on) {
      out = S.partial;
      return bstep_plain(S.ks);
    }
    const int po = pend == 0 ? 1 : 0;
    const BStep K{ks(cur), ks(cur ^ 1), pend >= 0 ? pa(pend) : nullptr,
                  pend >= 0 ? pfin : FIN_NONE};
    cur ^= 1;
    pend = po;
    pfin = FIN_NONE;
    out = pa(po);
    return K;
  }
  void done(int fin) {
    if (on) pfin = fin;
    else hipLaunchKernelGGL(kb_finalize, dim3(S.ncase), dim3(64), 0, S.s, S.partial,
                            S.m->case_blk_d.p, S.ncase, S.ks, fin);
  }
  void flush() {
    if (!on || (cur == 0 && pfin == FIN_NONE)) return;
    hipLaunchKernelGGL(kb_flush, dim3(S.ncase), dim3(64), 0, S.s,
                       (const double*)(pend >= 0 ? pa(pend) : S.partial), S.m->case_blk_d.p,
                       (const KS*)ks(cur), S.ks, pfin);
    cur = 0;
    pend = -1;
    pfin = FIN_NONE;
  }
};

template <class T>
static bool fuse_on(const BSolver<T>& S, int precon) {
  return precon != XCAL_PRECON_DILU && S.m->fuse && S.n <= kBatchFuseRows;
}

template <bool UPD, class T>
static void bsweep_fwd(const BSolver<T>& S, const FwdK<T>& W, const KS* ks, int rw = 0) {
  const xcal_mesh* m = S.m;
  if (rw <= 0) rw = m->pipe_rw_f > 0 ? m->pipe_rw_f : 1;
  auto pipe = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(S.ncase), dim3(kPipeBlock / rw), 0, S.s, W, S.Lrd,
                       m->lev_ptr_f_d.p, m->lmeta_f_d.p, m->case_lev_f_d.p, ks);
  };
  if (UPD && m->pipe_wf > 0)
    hipLaunchKernelGGL(kb_cg_xr<T>, dim3(S.nblk), dim3(kBlock), 0, S.s, S.B, W.x, W.p, W.r, W.q,
                       ks);
  if (m->pipe_wf > 0)
    with_pipe_form(m->pipe_wf, m->pipe_far_f, [&](auto WK, auto FAR) {
      with_pipe_rows(rw, [&](auto RW) {
        pipe(kb_fwd_pipe<UPD, decltype(FAR)::value, T, decltype(WK)::value, decltype(RW)::value>);
      });
    });
  else
    hipLaunchKernelGGL((kb_fwd<UPD, T>), dim3(S.ncase), dim3(kRunBlock), 0, S.s, W,
                       m->lev_ptr_f_d.p, m->lmeta_f_d.p, m->case_lev_f_d.p, ks);
}
template <int DOT, class T>
static void bsweep_bwd(const BSolver<T>& S, const BwdK<T>& W, const KS* ks, int rw = 0) {
  const xcal_mesh* m = S.m;
  if (rw <= 0) rw = m->pipe_rw_b > 0 ? m->pipe_rw_b : 1;
  auto pipe = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(S.ncase), dim3(kPipeBlock / rw), 0, S.s, W, m->Upos.p,
                       m->lev_ptr_b_d.p, m->lmeta_b_d.p, m->case_lev_b_d.p, ks);
  };
  if (m->pipe_wb > 0)
    with_pipe_form(m->pipe_wb, m->pipe_far_b, [&](auto WK, auto FAR) {
      with_pipe_rows(rw, [&](auto RW) {
        pipe(kb_bwd_pipe<DOT, decltype(FAR)::value, T, decltype(WK)::value, decltype(RW)::value>);
      });
    });
  else
    hipLaunchKernelGGL((kb_bwd<DOT, T>), dim3(S.ncase), dim3(kRunBlock), 0, S.s, W,
                       m->lev_ptr_b_d.p, m->lmeta_b_d.p, m->case_lev_b_d.p, ks);
}
template <class T>
static FwdK<T> bfwdk(const BSolver<T>& S, const T* u, T* Y, T* Yb) {
  return FwdK<T>{S.m->Lcol.p, S.Lnz, S.rD, u, Y, Yb, S.m->posb.p, nullptr, nullptr, nullptr,
                 nullptr, nullptr};
}
template <class T>
static BwdK<T> bbwdk(const BSolver<T>& S, const T* Yb, T* Z, const T* dv) {
  return BwdK<T>{S.m->Ucol.p, S.Unz, S.m->lev_rows_b.p, S.Db, Yb, Z, dv, S.partial, nullptr};
}

// KS of every case; `frozen` cases start stopped (they take no update)
template <class T>
static void binit(BSolver<T>& S, int itmax, double atol, double rtol, const int* frozen) {
  KS* h = (KS*)S.e->kstate_host;
  for (int k = 0; k < S.ncase; ++k) {
    KS q{};
    q.atol = atol;
    q.rtol = rtol;
    q.itmax = itmax;
    q.stop = frozen && frozen[k] ? STOP_SOLVED : STOP_NONE;
    h[k] = q;
  }
  XCAL_HIP(hipMemcpyAsync(S.ks, h, sizeof(KS) * S.ncase, hipMemcpyHostToDevice, S.s));
}

template <class T>
static bool bread(BSolver<T>& S) {
  XCAL_HIP(hipMemcpyAsync(S.e->kstate_host, S.ks, sizeof(KS) * S.ncase, hipMemcpyDeviceToHost, S.s));
  XCAL_HIP(hipStreamSynchronize(S.s));
  prof_resolve(S.m->ctx);
  const KS* h = (const KS*)S.e->kstate_host;
  for (int k = 0; k < S.ncase; ++k)
    if (h[k].stop == STOP_NONE) return false;
  return true;
}

// iterations in batches until every case has stopped; the first batch is
// sized from the previous solve (longest case)
// Launch signature of a batched solve's iteration: every pointer and size
// its kernels take, with the solver kind and value type.  Two solves with the
// same signature launch the same kernels with the same arguments.
template <class T>
static std::vector<uint64_t> graph_key(const BSolver<T>& S, int kind, int precon, const void* b) {
  auto u = [](const void* p) { return (uint64_t)(uintptr_t)p; };
  const xcal_mesh* m = S.m;
  return {(uint64_t)kind, sizeof(T), (uint64_t)precon, (uint64_t)S.n, (uint64_t)S.ncase,
          (uint64_t)S.nblk, u(S.s), u(S.work), u(S.ks), u(S.partial), u(S.A.nz), u(S.A.colval),
          u(S.A.rb), u(S.Sj), u(S.rD), u(S.Lnz), u(S.Unz), u(S.Db), u(S.Lrd), u(b),
          (uint64_t)m->pipe_rw_f, (uint64_t)m->pipe_rw_b, (uint64_t)m->pipe_wf,
          (uint64_t)m->pipe_wb};
}

// With XCAL_GRAPH=1 batched iterations are captured in HIP graphs of two
// iterations (the iteration's ping-pong buffers return to their roles after
// two), unless kernel timing is on.  A graph holds kernel launches only;
// every kernel reads the cases' Krylov state from device memory, so a replay
// does what the eager launches would.  Off by default: on the batches
// measured the device, not the launches, sets the time (replay 0.99x eager).
static bool graphs_enabled(const xcal_ctx* c) {
  const char* e = std::getenv("XCAL_GRAPH");
  if (!(e && *e == '1')) return false;
  for (int k = 0; k < xcal_ctx::kProfKernels; ++k)
    if (c->prof_on[k]) return false;
  return true;
}

// Runs iterations until every case has stopped: a first batch sized from the
// last solve (less a tenth), then `tail` at a time, reading the stop flags in
// between.  With a graph key, pairs of iterations are replayed from the
// equation's graph for the current parity (captured on first use).
template <class T>
static void brun(BSolver<T>& S, int tail, const std::function<void()>& iteration,
                 const std::vector<uint64_t>* gkey = nullptr, const int* parity = nullptr,
                 const std::function<void()>& flush = nullptr) {
  if (flush) flush();
  if (bread(S)) return;
  int batch = *S.pred > 0 ? std::max(1, (*S.pred * 9) / 10) : 8;
  auto run = [&](int count) {
    int done = 0;
    if (gkey && parity)
      for (; done + 2 <= count; done += 2) {
        hipGraphExec_t& g = S.e->graphs[*gkey][*parity];
        if (!g) {
          hipGraph_t graph = nullptr;
          XCAL_HIP(hipStreamBeginCapture(S.s, hipStreamCaptureModeThreadLocal));
          iteration();
          iteration();
          XCAL_HIP(hipStreamEndCapture(S.s, &graph));
          const hipError_t rc = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          XCAL_HIP(rc);
        }
        XCAL_HIP(hipGraphLaunch(g, S.s));
      }
    for (; done < count; ++done) iteration();
  };
  for (;;) {
    run(batch);
    XCAL_HIP(hipGetLastError());
    if (flush) flush();
    if (bread(S)) break;
    batch = tail;
  }
  const KS* h = (const KS*)S.e->kstate_host;
  int mx = 0;
  for (int k = 0; k < S.ncase; ++k) mx = std::max(mx, h[k].iter);
  *S.pred = mx;
}

// ------------------------------------------------- rows-per-thread tuner
// The pipelined batch sweeps run the same rows, values and sums whatever
// their rows per thread (RW), so RW is picked by time alone: on the first
// DILU solve on a batch mesh each direction is timed at RW = 1, 2 and 4 (one
// warm launch, then kTuneReps launches between HIP events, every case
// swept), and the fastest is kept for the mesh.  Meshes of the same shape
// (cells, entries, cases, levels, sweep form) reuse the choice and the
// timings.  XCAL_PIPE_ROWS=1|2|4 fixes RW instead.
constexpr int kTuneReps = 3;
constexpr int kTuneRows[3] = {1, 2, 4};

using PipeTune = xcal_mesh::PipeTuning;

static std::mutex g_tune_mu;
static std::map<std::vector<int64_t>, PipeTune> g_tune;

static std::vector<int64_t> tune_key(const xcal_mesh* m, int bytes) {
  return {m->n, m->nnz, m->ncase, (int64_t)m->lev_ptr_f.size(), (int64_t)m->lev_ptr_b.size(),
          m->pipe_wf, m->pipe_wb, m->pipe_far_f, m->pipe_far_b, bytes};
}

template <class T>
static void tune_pipes(const BSolver<T>& S, const T* u, T* Y, T* Yb, T* z) {
  const xcal_mesh* m = S.m;
  if (m->pipe_rw_f > 0 || m->pipe_wf <= 0 || m->pipe_wb <= 0) return;
  const char* env = std::getenv("XCAL_PIPE_ROWS");
  if (env && (*env == '1' || *env == '2' || *env == '4') && env[1] == 0) {
    m->pipe_rw_f = m->pipe_rw_b = *env - '0';
    return;
  }
  const auto key = tune_key(m, (int)sizeof(T));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      m->pipe_tune = it->second;
      m->pipe_rw_f = it->second.rw_f;
      m->pipe_rw_b = it->second.rw_b;
      return;
    }
  }
  struct Ev {
    hipEvent_t a = nullptr, b = nullptr;
    ~Ev() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  XCAL_HIP(hipEventCreate(&ev.a));
  XCAL_HIP(hipEventCreate(&ev.b));
  PipeTune t;
  for (int dir = 0; dir < 2; ++dir) {
    double best = 1e300;
    for (int c = 0; c < 3; ++c) {
      const int rw = kTuneRows[c];
      auto launch = [&] {
        if (dir == 0) bsweep_fwd<false>(S, bfwdk(S, u, Y, Yb), nullptr, rw);
        else bsweep_bwd<1>(S, bbwdk(S, (const T*)Yb, z, u), nullptr, rw);
      };
      launch();
      XCAL_HIP(hipEventRecord(ev.a, S.s));
      for (int r = 0; r < kTuneReps; ++r) launch();
      XCAL_HIP(hipEventRecord(ev.b, S.s));
      XCAL_HIP(hipEventSynchronize(ev.b));
      float ms = 0.f;
      XCAL_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
      const double us = 1e3 * ms / kTuneReps;
      t.us[dir][c] = us;
      if (us < best) {
        best = us;
        (dir == 0 ? t.rw_f : t.rw_b) = rw;
      }
    }
  }
  XCAL_HIP(hipGetLastError());
  m->pipe_tune = t;
  m->pipe_rw_f = t.rw_f;
  m->pipe_rw_b = t.rw_b;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.emplace(key, t);
}

// x0 == NULL: zero start, solution left in *xsol
template <class T>
static void bcg(BSolver<T>& S, int precon, const T* b, T* x0, int itmax, double atol, double rtol,
                const int* frozen, T** xsol) {
  const int64_t n = S.n;
  T* w = S.work;
  T *r = w, *x = w + n, *pa = w + 2 * n, *pb = w + 3 * n, *q = w + 4 * n, *zd = w + 5 * n,
    *Y = w + 6 * n, *Yb = w + 7 * n;
  binit(S, itmax, atol, rtol, frozen);
  hipLaunchKernelGGL(kb_start<T>, dim3(grid_for(n)), dim3(kBlock), 0, S.s, S.A, b, (const T*)x0, r,
                     x, pa);
  const bool dilu = precon == XCAL_PRECON_DILU;
  const T* Sd = precon == XCAL_PRECON_JACOBI ? S.Sj : nullptr;
  const KS* ks = S.ks;
  BFuse<T> F{S, fuse_on(S, precon)};
  const int32_t* seg = S.m->case_blk_d.p;
  if (dilu) {
    tune_pipes(S, (const T*)r, Y, Yb, zd);
    bsweep_fwd<false>(S, bfwdk(S, (const T*)r, Y, Yb), ks);
    bsweep_bwd<1>(S, bbwdk(S, (const T*)Yb, zd, (const T*)r), ks);
    fin_cases(S, FIN_CG_INIT);
  } else {
    hipLaunchKernelGGL((kb_precond_diag_start<false, T>), dim3(S.nblk), dim3(kBlock), 0, S.s, S.B,
                       Sd, r, zd, (const T*)nullptr, F.start());
    F.done(FIN_CG_INIT);
  }
  int parity = 0;
  xcal_ctx* c = S.m->ctx;
  auto iteration = [&]() {
    T* pold = parity ? pb : pa;
    T* pnew = parity ? pa : pb;
    parity ^= 1;
    if (F.on) {  // two launches: [p, q = A p, p.q], [x, r, z, r.z]
      double* P;
      const BStep K1 = F.step(P);
      with_row_group(S.A.rw, [&](auto G) {
        hipLaunchKernelGGL((kb_cg_pspmv<decltype(G)::value, T>), dim3(S.nblk), dim3(kBlock), 0,
                           S.s, S.B, S.A, (const T*)zd, (const T*)pold, pnew, q, P, K1, seg);
      });
      F.done(FIN_CG_PAP);
      const BStep K2 = F.step(P);
      hipLaunchKernelGGL(kb_cg_update_diag<T>, dim3(S.nblk), dim3(kBlock), 0, S.s, S.B, Sd,
                         (const T*)pnew, (const T*)q, x, r, zd, P, K2, seg);
      F.done(FIN_CG_ITER);
      return;
    }
    hipLaunchKernelGGL(kb_cg_p<T>, dim3(S.nblk), dim3(kBlock), 0, S.s, S.B, (const T*)zd,
                       (const T*)pold, pnew, ks);
    timed(c, 0, [&] {
      with_row_group(S.A.rw, [&](auto G) {
        hipLaunchKernelGGL((kb_cg_spmv<decltype(G)::value, T>), dim3(S.nblk), dim3(kBlock), 0, S.s,
                           S.B, S.A, (const T*)pnew, q, S.partial, ks);
      });
    });
    fin_blocks(S, FIN_CG_PAP);
    if (dilu) {
      FwdK<T> W = bfwdk(S, (const T*)nullptr, Y, Yb);
      W.p = pnew;
      W.q = q;
      W.x = x;
      W.r = r;
      timed(c, 1, [&] { bsweep_fwd<true>(S, W, ks); });
      timed(c, 2, [&] { bsweep_bwd<1>(S, bbwdk(S, (const T*)Yb, zd, (const T*)r), ks); });
      fin_cases(S, FIN_CG_ITER);
    } else {
      hipLaunchKernelGGL(kb_cg_update_diag<T>, dim3(S.nblk), dim3(kBlock), 0, S.s, S.B, Sd,
                         (const T*)pnew, (const T*)q, x, r, zd, S.partial, bstep_plain(ks), seg);
      fin_blocks(S, FIN_CG_ITER);
    }
  };
  const auto key = graph_key(S, 0, precon, nullptr);
  // fused chains carry their pending finalize between launches, so they are
  // not captured
  brun(S, 4, iteration, graphs_enabled(c) && !F.on ? &key : nullptr, &parity,
       [&] { F.flush(); });
  if (x0) hipLaunchKernelGGL(kb_add, dim3(grid_for(n)), dim3(kBlock), 0, S.s, n, (const double*)x,
                             (double*)x0);
  if (xsol) *xsol = x;
  XCAL_HIP(hipGetLastError());
}

template <class T>
static void bbicg(BSolver<T>& S, int precon, const T* b, T* x0, int itmax, double atol,
                  double rtol, const int* frozen, T** xsol) {
  const int64_t n = S.n;
  T* w = S.work;
  T *r = w, *x = w + n, *pa = w + 2 * n, *pb = w + 3 * n, *v = w + 4 * n, *sv = w + 5 * n,
    *t = w + 6 * n, *d = w + 7 * n, *Y = w + 8 * n, *Yb = w + 9 * n;
  binit(S, itmax, atol, rtol, frozen);
  hipLaunchKernelGGL(kb_start<T>, dim3(grid_for(n)), dim3(kBlock), 0, S.s, S.A, b, (const T*)x0, r,
                     x, pa);
  const bool dilu = precon == XCAL_PRECON_DILU;
  const T* Sd = precon == XCAL_PRECON_JACOBI ? S.Sj : nullptr;
  const KS* ks = S.ks;
  BFuse<T> F{S, fuse_on(S, precon)};
  const int32_t* seg = S.m->case_blk_d.p;
  if (dilu) {
    tune_pipes(S, (const T*)r, Y, Yb, t);
    bsweep_fwd<false>(S, bfwdk(S, (const T*)r, Y, Yb), ks);
    bsweep_bwd<4>(S, bbwdk(S, (const T*)Yb, r, b), ks);
    fin_cases(S, FIN_BI_INIT);
  } else {
    hipLaunchKernelGGL((kb_precond_diag_start<true, T>), dim3(S.nblk), dim3(kBlock), 0, S.s, S.B, Sd,
                       r, (T*)nullptr, b, F.start());
    F.done(FIN_BI_INIT);
  }
  int parity = 0;
  auto iteration = [&]() {
    T* pold = parity ? pb : pa;
    T* pnew = parity ? pa : pb;
    parity ^= 1;
    if (dilu) {
      with_row_group(S.A.rw, [&](auto G) {
        constexpr int g = decltype(G)::value;
        hipLaunchKernelGGL((kb_bi_spmv_p<g, false, T>), dim3(S.nblk), dim3(kBlock), 0,
                           S.s, S.B, S.A,
                           (const T*)r, (const T*)pold, pnew, (const T*)nullptr, d, b, S.partial,
                           bstep_plain(ks), seg);
      });
      bsweep_fwd<false>(S, bfwdk(S, (const T*)d, Y, Yb), ks);
      bsweep_bwd<2>(S, bbwdk(S, (const T*)Yb, v, b), ks);
      fin_cases(S, FIN_BI_ALPHA);
      with_row_group(S.A.rw, [&](auto G) {
        constexpr int g = decltype(G)::value;
        hipLaunchKernelGGL((kb_bi_spmv_s<g, false, T>), dim3(S.nblk), dim3(kBlock), 0,
                           S.s, S.B, S.A,
                           (const T*)r, (const T*)v, (const T*)pnew, sv, x, (const T*)nullptr, d,
                           S.partial, bstep_plain(ks), seg);
      });
      bsweep_fwd<false>(S, bfwdk(S, (const T*)d, Y, Yb), ks);
      bsweep_bwd<3>(S, bbwdk(S, (const T*)Yb, t, (const T*)sv), ks);
      fin_cases(S, FIN_BI_OMEGA);
    } else {  // fused: three launches per iteration
      double* P;
      const BStep Kp = F.step(P);
      with_row_group(S.A.rw, [&](auto G) {
        constexpr int g = decltype(G)::value;
        hipLaunchKernelGGL((kb_bi_spmv_p<g, true, T>), dim3(S.nblk), dim3(kBlock), 0,
                           S.s, S.B, S.A,
                           (const T*)r, (const T*)pold, pnew, Sd, v, b, P, Kp, seg);
      });
      F.done(FIN_BI_ALPHA);
      const BStep Ks = F.step(P);
      with_row_group(S.A.rw, [&](auto G) {
        constexpr int g = decltype(G)::value;
        hipLaunchKernelGGL((kb_bi_spmv_s<g, true, T>), dim3(S.nblk), dim3(kBlock), 0,
                           S.s, S.B, S.A,
                           (const T*)r, (const T*)v, (const T*)pnew, sv, x, Sd, t, P, Ks, seg);
      });
      F.done(FIN_BI_OMEGA);
    }
    double* P;
    const BStep Kt = F.step(P);
    hipLaunchKernelGGL(kb_bi_tail<T>, dim3(S.nblk), dim3(kBlock), 0, S.s, S.B, (const T*)sv,
                       (const T*)t, (const T*)v, x, r, pnew, b, P, Kt, seg);
    F.done(FIN_BI_END);
  };
  const auto key = graph_key(S, 1, precon, b);
  brun(S, 2, iteration, graphs_enabled(S.m->ctx) && !F.on ? &key : nullptr, &parity,
       [&] { F.flush(); });
  if (x0) hipLaunchKernelGGL(kb_add, dim3(grid_for(n)), dim3(kBlock), 0, S.s, n, (const double*)x,
                             (double*)x0);
  if (xsol) *xsol = x;
  XCAL_HIP(hipGetLastError());
}

// per-case |b - A x|^2 and |b|^2 (host arrays of ncase); r32 optional
static void bresidual_sums(BSolver<double>& S, const double* x, const double* b, float* r32,
                           std::vector<double>& rr, std::vector<double>& bb) {
  xcal_ctx* c = S.m->ctx;
  const int K = S.ncase;
  if (c->red_result.n < 2 * (int64_t)K) c->red_result.alloc(2 * (int64_t)K + 64);
  if (S.e->case_host.size() < 2 * (size_t)K) {
    S.e->case_host.resize(2 * (size_t)K);
  }
  hipLaunchKernelGGL(kb_residual, dim3(S.nblk), dim3(kBlock), 0, S.s, S.B, S.A, x, b, r32, S.partial);
  hipLaunchKernelGGL(kb_sum_cases, dim3(K), dim3(64), 0, S.s, S.partial, S.m->case_blk_d.p, K,
                     c->red_result.p);
  XCAL_HIP(hipGetLastError());
  XCAL_HIP(hipMemcpyAsync(S.e->case_host.data(), c->red_result.p, 2 * sizeof(double) * K,
                          hipMemcpyDeviceToHost, S.s));
  XCAL_HIP(hipStreamSynchronize(S.s));
  rr.resize(K);
  bb.resize(K);
  for (int k = 0; k < K; ++k) {
    rr[k] = S.e->case_host[2 * k];
    bb[k] = S.e->case_host[2 * k + 1];
  }
}

// residual() of Solve_1_api.jl:441-464 per case
static void bresidual(BSolver<double>& S, const double* x, const double* b, double* out) {
  std::vector<double> rr, bb;
  bresidual_sums(S, x, b, nullptr, rr, bb);
  for (int k = 0; k < S.ncase; ++k) {
    const double normb = std::sqrt(bb[k]);
    const double a = std::fabs(normb);
    const double eps_nb = std::nextafter(a, INFINITY) - a;
    out[k] = std::sqrt(rr[k]) / (normb > eps_nb ? normb : 1.0);
  }
}

// FP64 refinement around FP32 batched solves, per case as in ir_solve:
// a case whose true residual met its tolerance (or stalled) is frozen and
// takes no further update, so its result does not depend on the others.
static void bir_solve(xcal_eqn* e, const xcal_solver_setup* st, int itmax, double* x,
                      const double* b, std::vector<int>& niter, std::vector<int>& nouter,
                      std::vector<int>& rcs) {
  BSolver<double> S = batch_solver(e);
  BSolver<float> S32 = batch_solver32(e);
  const int K = S.ncase;
  float* r32 = e->work32.p + 11 * S.n;
  // an inner solve need not be tighter than the outer tolerance
  const double inner_rtol = std::max(st->ir_inner_rtol > 0 ? st->ir_inner_rtol : 1e-4, st->rtol);
  std::vector<double> rr, bb, res(K), prev(K, INFINITY), tol(K);
  bresidual_sums(S, x, b, r32, rr, bb);
  std::vector<int> frozen(K, 0);
  niter.assign(K, 0);
  nouter.assign(K, 0);
  rcs.assign(K, 0);
  for (int k = 0; k < K; ++k) {
    res[k] = std::sqrt(rr[k]);
    tol[k] = st->atol + st->rtol * res[k];
  }
  const bool single = st->rtol >= kIrSingleSweep;  // one FP32 solve (see ir_solve)
  for (int sweep = 0;; ++sweep) {
    bool any = false;
    for (int k = 0; k < K; ++k) {
      if (frozen[k]) continue;
      if (res[k] <= tol[k]) {
        frozen[k] = 1;
      } else if (single && sweep == 1) {
        frozen[k] = 1;
        rcs[k] = niter[k] >= itmax ? 1 : 0;
      } else if (niter[k] >= itmax || nouter[k] >= 100 || res[k] >= 0.5 * prev[k]) {
        frozen[k] = 1;
        rcs[k] = 1;
      } else {
        any = true;
      }
    }
    if (!any) break;
    float* d = nullptr;
    S32.pred = e->pred_slot(x, 16 + 2 * std::min(sweep, 12) + (st->solver == XCAL_SOLVER_CG ? 0 : 1));
    const double ia = single ? st->atol : 0.0;
    if (st->solver == XCAL_SOLVER_CG)
      bcg(S32, st->precon, (const float*)r32, (float*)nullptr, itmax, ia, inner_rtol, frozen.data(),
          &d);
    else
      bbicg(S32, st->precon, (const float*)r32, (float*)nullptr, itmax, ia, inner_rtol,
            frozen.data(), &d);
    const KS* h = (const KS*)e->kstate_host;
    std::vector<int32_t> active(K);
    for (int k = 0; k < K; ++k) {
      active[k] = !frozen[k];
      if (!frozen[k]) {
        niter[k] += h[k].iter;
        nouter[k] += 1;
      }
    }
    XCAL_HIP(hipMemcpy(e->case_active.p, active.data(), sizeof(int32_t) * K, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(kb_ir_update, dim3(S.nblk), dim3(kBlock), 0, S.s, S.B,
                       (const int32_t*)e->case_active.p, (const float*)d, x);
    for (int k = 0; k < K; ++k)
      if (!frozen[k]) prev[k] = res[k];
    bresidual_sums(S, x, b, r32, rr, bb);
    for (int k = 0; k < K; ++k)
      if (!frozen[k]) res[k] = std::sqrt(rr[k]);
  }
}

void precond_update_batch(xcal_eqn* e) {
  const xcal_mesh* m = e->mesh;
  hipStream_t s = m->ctx->stream;
  if (m->pipe_wf > 0) {
    const int64_t nl = m->Lslot.n;
    if (e->Lprod.n != nl) e->Lprod.alloc(nl);
    hipLaunchKernelGGL(k_lower_products, dim3(grid_for(nl)), dim3(kBlock), 0, s, nl, m->Lslot.p,
                       m->trans.p, e->nzval.p, e->Lprod.p);
    const int rw = m->pipe_rw_f > 0 ? m->pipe_rw_f : 1;  // tuned on the first solve
    auto pipe = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(m->ncase), dim3(kPipeBlock / rw), 0, s, e->pD.p,
                         (const double*)e->Lprod.p, m->Lcol.p, m->lev_ptr_f_d.p, m->lmeta_f_d.p,
                         m->case_lev_f_d.p);
    };
    with_pipe_form(m->pipe_wf, m->pipe_far_f, [&](auto WK, auto FAR) {
      with_pipe_rows(rw, [&](auto RW) {
        pipe(kb_dilu_diag_pipe<decltype(FAR)::value, decltype(WK)::value, decltype(RW)::value>);
      });
    });
    return;
  }
  hipLaunchKernelGGL(kb_dilu_diag, dim3(m->ncase), dim3(kRunBlock), 0, s,
                     matk(m, (const double*)e->nzval.p), m->trans.p, e->pD.p, m->lev_ptr_f_d.p,
                     m->case_lev_f_d.p);
}

// solve_system! over every case of a batch mesh; per-case statistics
int solve_system_batch(xcal_eqn* e, const xcal_solver_setup* st, double* x, const double* b,
                       xcal_solve_stats* stats) {
  const int K = e->mesh->ncase;
  const int itmax = st->itmax == 0 ? (int)(2 * e->mesh->n) : st->itmax;
  std::vector<int> niter(K, 0), nouter(K, 1), rcs(K, 0);
  solve_prologue(e, st, x, b);
  if (st->precision == XCAL_PREC_FP32_IR) {
    bir_solve(e, st, itmax, x, b, niter, nouter, rcs);
  } else {
    BSolver<double> S = batch_solver(e);
    S.pred = e->pred_slot(x, st->solver == XCAL_SOLVER_CG ? 0 : 1);
    if (st->solver == XCAL_SOLVER_CG)
      bcg(S, st->precon, b, x, itmax, st->atol, st->rtol, (const int*)nullptr, (double**)nullptr);
    else
      bbicg(S, st->precon, b, x, itmax, st->atol, st->rtol, (const int*)nullptr, (double**)nullptr);
    const KS* h = (const KS*)e->kstate_host;
    for (int k = 0; k < K; ++k) {
      niter[k] = h[k].iter;
      rcs[k] = (h[k].stop == STOP_ITMAX || h[k].iter >= itmax) ? 1 : 0;
    }
  }
  solve_epilogue(e, x);
  BSolver<double> S = batch_solver(e);
  std::vector<double> res(K);
  bresidual(S, x, b, res.data());
  int rc = 0;
  for (int k = 0; k < K; ++k) {
    if (stats) {
      stats[k].niter = niter[k];
      stats[k].nouter = nouter[k];
      stats[k].residual = res[k];
    }
    rc |= rcs[k];
  }
  return rc;
}

}  // namespace xcal
