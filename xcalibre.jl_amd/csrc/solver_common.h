// Krylov state and the per-row sweep functions shared by the single-mesh
// solver (solver.hip) and the batched solver (batch.hip).
#pragma once

#include "device_common.h"

namespace xcal {

// mixed precision: outer tolerances at least this loose take one FP32 solve
constexpr double kIrSingleSweep = 1e-3;

enum : int { STOP_NONE = 0, STOP_SOLVED = 1, STOP_ITMAX = 2, STOP_BREAKDOWN = 3,
             STOP_ZERO_RHS = 4 };

struct KS {
  double gamma, pAp, alpha, beta, rNorm, eps, pNorm2, rho, next_rho, omega, atol, rtol;
  int iter, itmax, stop, err;
};

enum Fin : int { FIN_NONE = 0, FIN_CG_INIT, FIN_CG_PAP, FIN_CG_ITER, FIN_BI_INIT, FIN_BI_ALPHA,
                 FIN_BI_OMEGA, FIN_BI_END };

__device__ inline void finalize(KS* ks, int fin, double d0, double d1) {
  const double epsm = 2.220446049250313e-16;
  switch (fin) {
    case FIN_CG_INIT: {
      ks->gamma = d0;
      ks->iter = 0;
      ks->beta = 0.0;
      if (d0 == 0.0) { ks->stop = STOP_ZERO_RHS; return; }
      ks->rNorm = sqrt(d0);
      ks->pNorm2 = d0;
      ks->eps = ks->atol + ks->rtol * ks->rNorm;
      if (ks->rNorm <= ks->eps) ks->stop = STOP_SOLVED;
      else if (ks->iter >= ks->itmax) ks->stop = STOP_ITMAX;
      return;
    }
    case FIN_CG_PAP: {
      ks->pAp = d0;
      if (d0 <= epsm * ks->pNorm2 && fabs(d0) <= epsm * ks->pNorm2) ks->stop = STOP_BREAKDOWN;
      else ks->alpha = ks->gamma / d0;
      return;
    }
    case FIN_CG_ITER: {
      const double gn = d0;
      ks->rNorm = sqrt(gn);
      const bool solved = ks->rNorm <= ks->eps || (ks->rNorm + 1.0 <= 1.0);
      if (!solved) {
        const double beta = gn / ks->gamma;
        ks->beta = beta;
        ks->pNorm2 = gn + beta * beta * ks->pNorm2;
        ks->gamma = gn;
      }
      ks->iter += 1;
      if (solved) ks->stop = STOP_SOLVED;
      else if (ks->iter >= ks->itmax) ks->stop = STOP_ITMAX;
      return;
    }
    case FIN_BI_INIT: {
      ks->iter = 0;
      ks->beta = 0.0;
      ks->rNorm = sqrt(d0);
      if (ks->rNorm == 0.0) { ks->stop = STOP_ZERO_RHS; return; }
      ks->eps = ks->atol + ks->rtol * ks->rNorm;
      ks->next_rho = d1;
      if (d1 == 0.0) { ks->stop = STOP_ZERO_RHS; return; }
      if (ks->rNorm <= ks->eps) ks->stop = STOP_SOLVED;
      else if (ks->iter >= ks->itmax) ks->stop = STOP_ITMAX;
      return;
    }
    case FIN_BI_ALPHA: {
      ks->iter += 1;
      ks->rho = ks->next_rho;
      ks->alpha = ks->rho / d0;
      return;
    }
    case FIN_BI_OMEGA: {
      ks->omega = d0 / d1;
      return;
    }
    case FIN_BI_END: {
      ks->next_rho = d0;
      ks->beta = (ks->next_rho / ks->rho) * (ks->alpha / ks->omega);
      ks->rNorm = sqrt(d1);
      const bool solved = ks->rNorm <= ks->eps || (ks->rNorm + 1.0 <= 1.0);
      const bool tired = ks->iter >= ks->itmax;
      const bool brk = ks->alpha == 0.0 || ks->alpha != ks->alpha;
      if (solved) ks->stop = STOP_SOLVED;
      else if (tired) ks->stop = STOP_ITMAX;
      else if (brk) ks->stop = STOP_BREAKDOWN;
      return;
    }
    default:
      return;
  }
}


// ---------------------------------------------------------- reductions
// The kernels below are templates on the value type T of the matrix and the
// vectors: double for the FP64 solve, float for the inner solves of the
// mixed-precision (FP32 inner, FP64 refinement) mode.  Dot products are
// accumulated in double in both cases; the Krylov scalars are double.
__device__ inline bool stopped(const KS* ks) { return ks->stop != STOP_NONE; }
// Kernels of an iteration read the stop flag early but use it only to
// predicate their stores, so their loads do not wait behind it.

// workgroup sums of s0, s1 -> partial[2 slot], partial[2 slot + 1]
template <int NT>
__device__ inline void store_partial(double s0, double s1, double* partial, int slot) {
  __shared__ double lds[NT / 64];
  s0 = block_sum(s0, lds);
  s1 = block_sum(s1, lds);
  if (threadIdx.x == 0) {
    partial[2 * slot] = s0;
    partial[2 * slot + 1] = s1;
  }
}


__device__ inline double dd(double a, double b) { return a * b; }  // products of dots in double

// ------------------------------------------------------------ DILU sweeps
// Both sweeps read the packed factors (level l stores entry t of its q-th
// row at base_l + t*cnt_l + q).  A row first loads all its entries (their
// addresses depend only on the row's position), then gathers the values
// they depend on, then sums in the reference's column order: two dependent
// memory round trips per level.
constexpr int kAhead = 8;  // entries in flight per row

// Forward sweep y = (I + L D^-1)^-1 u: row order and per-row summation order
// of forward_substitution!.  UPD fuses the CG update of the row's own x and
// r first (u = r after r -= a q).  y goes to Y (device order, gathered by
// later levels) and to Yb in backward-level order for the backward sweep.
template <class T>
struct FwdK {
  const int32_t* Lcol;
  const T* Lnz;
  const T* rD;
  const T* u;
  T* Y;
  T* Yb;
  const int32_t* posb;
  const T* p;
  const T* q;
  T* x;
  T* r;
  const KS* ks;  // NULL outside a Krylov solve
};

template <bool UPD, class T>
__device__ inline void fwd_row(const FwdK<T>& W, int32_t i, int64_t e0, int32_t cnt, int w, T a,
                               bool live) {
  T u;
  if (UPD) {
    const T xi = W.x[i] + a * W.p[i];
    u = W.r[i] + (-a) * W.q[i];
    if (live) {
      W.x[i] = xi;
      W.r[i] = u;
    }
  } else {
    u = W.u[i];
  }
  const int32_t pb = W.posb[i];
  T sum = 0;
  for (int t0 = 0; t0 < w; t0 += kAhead) {
    int col[kAhead];
    T av[kAhead], yv[kAhead], rv[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      col[k] = -1;
      av[k] = 0;
      if (t0 + k < w) {
        const int64_t e = e0 + (int64_t)(t0 + k) * cnt;
        col[k] = W.Lcol[e];
        av[k] = W.Lnz[e];
      }
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      yv[k] = 0;
      rv[k] = 0;
      if (col[k] >= 0) {
        yv[k] = W.Y[col[k]];
        rv[k] = W.rD[col[k]];
      }
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k)
      if (col[k] >= 0) sum += (av[k] * yv[k]) * rv[k];
  }
  const T y = u - sum;
  if (live) {
    W.Y[i] = y;
    W.Yb[pb] = y;
  }
}

// Backward sweep z = (D + U)^-1 y: order of backward_substitution!.  Works
// in backward-level order (position p, row rows[p]): y and D come packed in
// that order, z is written to Z in device order, and up to two dot products
// are folded into partials:
//   DOT 0: none, 1: r.z (CG), 2: c.z (BiCGSTAB v), 3: z.s and z.z (BiCGSTAB t),
//   4: z.z and c.z (BiCGSTAB start, z = M r0)
template <class T>
struct BwdK {
  const int32_t* Ucol;
  const T* Unz;
  const int32_t* rows;
  const T* Db;
  const T* Yb;
  T* Z;
  const T* dv;
  double* partial;
  const KS* ks;
};

template <int DOT, class T>
__device__ inline void bwd_row(const BwdK<T>& W, int32_t pos, int64_t e0, int32_t cnt, int w,
                               bool live, double& s0, double& s1) {
  const int32_t i = W.rows[pos];
  const T yb = W.Yb[pos], d = W.Db[pos];
  T sum = 0;
  for (int t0 = 0; t0 < w; t0 += kAhead) {
    int col[kAhead];
    T av[kAhead], zv[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      col[k] = -1;
      av[k] = 0;
      if (t0 + k < w) {
        const int64_t e = e0 + (int64_t)(t0 + k) * cnt;
        col[k] = W.Ucol[e];
        av[k] = W.Unz[e];
      }
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k) zv[k] = col[k] >= 0 ? W.Z[col[k]] : T(0);
#pragma unroll
    for (int k = 0; k < kAhead; ++k)
      if (col[k] >= 0) sum += av[k] * zv[k];
  }
  const T zi = (yb - sum) / d;
  if (live) W.Z[i] = zi;
  if (DOT == 1 || DOT == 2) s0 += dd(W.dv[i], zi);
  if (DOT == 3) { s0 += dd(zi, W.dv[i]); s1 += dd(zi, zi); }
  if (DOT == 4) { s0 += dd(zi, zi); s1 += dd(W.dv[i], zi); }
}

// ----------------------------------------------------- kernel timing hook
inline hipEvent_t prof_event(xcal_ctx* c) {
  if (c->prof_pool.empty()) {
    hipEvent_t ev;
    XCAL_HIP(hipEventCreate(&ev));
    return ev;
  }
  hipEvent_t ev = c->prof_pool.back();
  c->prof_pool.pop_back();
  return ev;
}

// Launch `fn` on the context stream, bracketed by events when kernel `k` is
// the one being timed.
template <class Fn>
inline void timed(xcal_ctx* c, int k, Fn&& fn) {
  if (c->prof_kernel != k) {
    fn();
    return;
  }
  hipEvent_t a = prof_event(c), b = prof_event(c);
  XCAL_HIP(hipEventRecord(a, c->stream));
  fn();
  XCAL_HIP(hipEventRecord(b, c->stream));
  c->prof_pending.emplace_back(a, b);
}

// after a stream synchronisation: fold completed event pairs into the totals
inline void prof_resolve(xcal_ctx* c) {
  for (auto& pr : c->prof_pending) {
    float ms = 0.0f;
    XCAL_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
    c->prof_ms += ms;
    c->prof_launches += 1;
    c->prof_pool.push_back(pr.first);
    c->prof_pool.push_back(pr.second);
  }
  c->prof_pending.clear();
}

}  // namespace xcal
