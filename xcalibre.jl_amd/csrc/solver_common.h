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
static_assert(sizeof(KS) <= 32 * sizeof(double), "the context's zeroed state holds one KS");

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

// Sum of `count` partial pairs in slot order, the same in every kernel that
// reduces them: thread t < kBlock adds slots t, t + kBlock, ... in turn,
// then block_sum; every thread receives both sums.  lds: >= kBlock / 64.
// U slots per thread are loaded together, then added in slot order: U does
// not change the sums, only how many round trips a long list takes (the
// one-workgroup finalize takes 16, one round trip for up to 4096 slots).
template <int U = 8>
__device__ inline void reduce_partials(const double* __restrict__ partial, int count, double& t0,
                                       double& t1, double* lds) {
  double a = 0.0, b = 0.0;
  if ((int)threadIdx.x < kBlock)
    for (int s0 = threadIdx.x; s0 < count; s0 += U * kBlock) {
      double pa[U], pb[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int s = s0 + u * kBlock < count ? s0 + u * kBlock : 0;
        pa[u] = partial[2 * s];
        pb[u] = partial[2 * s + 1];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (s0 + u * kBlock < count) {
          a += pa[u];
          b += pb[u];
        }
    }
  a = block_sum(a, lds);
  b = block_sum(b, lds);
  if (threadIdx.x == 0) {
    lds[0] = a;
    lds[1] = b;
  }
  __syncthreads();
  t0 = lds[0];
  t1 = lds[1];
  __syncthreads();
}

// The Krylov state a kernel of an iteration works with.  Unfused (fin ==
// FIN_NONE, out == nullptr) it is `in`, as a one-workgroup finalize left it.
// Fused, the kernel first does that finalize itself: every workgroup reduces
// the previous kernel's partials in the same order and applies the same
// scalar update, so all hold the same state, and workgroup 0 stores it to
// `out` (never `in`, which the other workgroups are still reading).
struct KStep {
  const KS* in;
  KS* out;
  const double* pin;
  int cin;
  int fin;
};
inline KStep kstep_plain(const KS* ks) { return KStep{ks, nullptr, nullptr, 0, FIN_NONE}; }
__device__ inline KS kstate(const KStep& K, double* lds) {
  KS k = *K.in;
  if (K.fin != FIN_NONE) {
    double t0, t1;
    reduce_partials(K.pin, K.cin, t0, t1, lds);
    if (k.stop == STOP_NONE) finalize(&k, K.fin, t0, t1);
  }
  if (K.out && blockIdx.x == 0 && threadIdx.x == 0) *K.out = k;
  return k;
}

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


// store_partial when each thread holds RW row slots, slot j of thread t
// standing for thread t + j blockDim.x of a workgroup RW times as wide: the
// same wavefront sums, added in the same wavefront order, so the partials are
// those of that wider workgroup bit for bit.
template <int RW>
__device__ inline double block_sum_rows(const double (&v)[RW], double* lds) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < RW; ++j) {
    double x = v[j];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) lds[(j * (int)blockDim.x + (int)threadIdx.x) >> 6] = x;
  }
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (RW * (int)blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) s += lds[w];
  }
  __syncthreads();
  return s;
}
template <int NV, int RW>
__device__ inline void store_partial_rows(const double (&s0)[RW], const double (&s1)[RW],
                                          double* partial, int slot) {
  __shared__ double lds[NV / 64];
  const double t0 = block_sum_rows<RW>(s0, lds);
  const double t1 = block_sum_rows<RW>(s1, lds);
  if (threadIdx.x == 0) {
    partial[2 * slot] = t0;
    partial[2 * slot + 1] = t1;
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
  // the cone form's UPD: the updated r goes here (r is read by other
  // workgroups' cones during the sweep, so it is not updated in place)
  T* rout = nullptr;
};

// Loads of one row's entries.  WK > 0: the level's entry count w satisfies
// 1 <= w <= WK and all WK loads are issued at once, branch-free (entries past
// w re-read entry w-1 and are masked with column -1).
template <int WK, class T>
__device__ inline void load_entries(const int32_t* __restrict__ col_of, const T* __restrict__ val,
                                    int64_t e0, int32_t cnt, int w, int (&col)[WK],
                                    T (&av)[WK]) {
#pragma unroll
  for (int k = 0; k < WK; ++k) {
    const int64_t e = e0 + (int64_t)(k < w ? k : w - 1) * cnt;
    const int c = col_of[e];
    col[k] = k < w ? c : -1;
    av[k] = val[e];
  }
}

// Sum over a row's entries for a level with more than 8 entries per row (or
// none): kAhead entries at a time.
template <class T, class G>
__device__ inline T entry_sum_loop(const int32_t* __restrict__ col_of, const T* __restrict__ val,
                                   int64_t e0, int32_t cnt, int w, G&& term) {
  T sum = 0;
  for (int t0 = 0; t0 < w; t0 += kAhead) {
    int col[kAhead];
    T av[kAhead], tv[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      col[k] = -1;
      av[k] = 0;
      if (t0 + k < w) {
        const int64_t e = e0 + (int64_t)(t0 + k) * cnt;
        col[k] = col_of[e];
        av[k] = val[e];
      }
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k) tv[k] = col[k] >= 0 ? term(av[k], col[k]) : T(0);
#pragma unroll
    for (int k = 0; k < kAhead; ++k)
      if (col[k] >= 0) sum += tv[k];
  }
  return sum;
}

// A row's memory work comes in two round trips: (1) everything addressed by
// the row's position (own values, packed entries), (2) the gathers addressed
// by the entries.  Scheduling barriers keep arithmetic that needs the stop
// flag or alpha from being hoisted between them, where it would add a third.
// READ: the stop flag and alpha are read from W.ks after the row's first
// round trip has been issued (a and live are then ignored).
template <bool UPD, int WK, class T, bool READ = false>
__device__ inline void fwd_row(const FwdK<T>& W, int32_t i, int64_t e0, int32_t cnt, int w, T a,
                               bool live) {
  T xo = 0, po = 0, ro = 0, qo = 0, uo = 0;
  if (UPD) {
    xo = W.x[i];
    po = W.p[i];
    ro = W.r[i];
    qo = W.q[i];
  } else {
    uo = W.u[i];
  }
  const int32_t pb = W.posb[i];
  T sum = 0;
  if constexpr (WK > 0) {
    int col[WK];
    T av[WK], yv[WK], rv[WK];
    load_entries<WK>(W.Lcol, W.Lnz, e0, cnt, w, col, av);
    __builtin_amdgcn_sched_barrier(0);
    if (READ) {
      live = !stopped(W.ks);
      a = UPD ? (T)W.ks->alpha : T(0);
    }
#pragma unroll
    for (int k = 0; k < WK; ++k) {
      const int c = col[k] >= 0 ? col[k] : 0;
      yv[k] = W.Y[c];
      rv[k] = W.rD[c];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < WK; ++k) {
      const T t = (av[k] * yv[k]) * rv[k];
      if (col[k] >= 0) sum += t;
    }
  } else {
    if (READ) {
      live = !stopped(W.ks);
      a = UPD ? (T)W.ks->alpha : T(0);
    }
    sum = entry_sum_loop(W.Lcol, W.Lnz, e0, cnt, w,
                         [&](T av, int c) { return (av * W.Y[c]) * W.rD[c]; });
  }
  // x and r are stored back unconditionally (their old values once the
  // solve has stopped) so their loads are not sunk under the predicate
  T u;
  if (UPD) {
    const T xi = xo + a * po;
    u = ro + (-a) * qo;
    W.x[i] = live ? xi : xo;
    W.r[i] = live ? u : ro;
  } else {
    u = uo;
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

template <int DOT, int WK, class T, bool READ = false>
__device__ inline void bwd_row(const BwdK<T>& W, int32_t pos, int64_t e0, int32_t cnt, int w,
                               bool live, double& s0, double& s1) {
  const int32_t i = W.rows[pos];
  const T yb = W.Yb[pos], d = W.Db[pos];
  T sum = 0, dvi = 0;
  if constexpr (WK > 0) {
    int col[WK];
    T av[WK], zv[WK];
    load_entries<WK>(W.Ucol, W.Unz, e0, cnt, w, col, av);
    __builtin_amdgcn_sched_barrier(0);
    if (READ) live = !stopped(W.ks);
    // the dot operand is read before Z is written (it may alias Z), in the
    // same round trip as the gathers
    if (DOT != 0) dvi = W.dv[i];
#pragma unroll
    for (int k = 0; k < WK; ++k) zv[k] = W.Z[col[k] >= 0 ? col[k] : 0];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < WK; ++k) {
      const T t = av[k] * zv[k];
      if (col[k] >= 0) sum += t;
    }
  } else {
    if (READ) live = !stopped(W.ks);
    if (DOT != 0) dvi = W.dv[i];
    sum = entry_sum_loop(W.Ucol, W.Unz, e0, cnt, w, [&](T av, int c) { return av * W.Z[c]; });
  }
  const T zi = (yb - sum) / d;
  if (live) W.Z[i] = zi;
  if (DOT == 1 || DOT == 2) s0 += dd(dvi, zi);
  if (DOT == 3) { s0 += dd(zi, dvi); s1 += dd(zi, zi); }
  if (DOT == 4) { s0 += dd(zi, zi); s1 += dd(dvi, zi); }
}


// ----------------------------------------------------- kernel timing hook
// Timing-only events: without the system-scope fence, recording one does not
// write back and invalidate the caches between the kernels it separates.
// Elapsed times are read only after the stream has been synchronised.
inline hipEvent_t prof_event(xcal_ctx* c) {
  if (c->prof_pool.empty()) {
    hipEvent_t ev;
    XCAL_HIP(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
    return ev;
  }
  hipEvent_t ev = c->prof_pool.back();
  c->prof_pool.pop_back();
  return ev;
}

// Launch `fn` on the context stream, bracketed by events when kernel `k` is
// the one being timed: every kProfEvery-th launch of it, starting with the
// first, so that the events cost the timed work little.
constexpr int64_t kProfEvery = 8;
template <class Fn>
inline void timed(xcal_ctx* c, int k, Fn&& fn) {
  if (k < 0 || k >= xcal_ctx::kProfKernels || !c->prof_on[k] ||
      c->prof_seq[k]++ % kProfEvery != 0) {
    fn();
    return;
  }
  hipEvent_t a = prof_event(c), b = prof_event(c);
  XCAL_HIP(hipEventRecord(a, c->stream));
  fn();
  XCAL_HIP(hipEventRecord(b, c->stream));
  c->prof_pending.push_back({k, a, b});
}

// after a stream synchronisation: fold completed event pairs into the totals
inline void prof_resolve(xcal_ctx* c) {
  for (auto& pr : c->prof_pending) {
    float ms = 0.0f;
    XCAL_HIP(hipEventElapsedTime(&ms, pr.a, pr.b));
    c->prof_ms[pr.kernel] += ms;
    c->prof_launches[pr.kernel] += 1;
    c->prof_pool.push_back(pr.a);
    c->prof_pool.push_back(pr.b);
  }
  c->prof_pending.clear();
}

}  // namespace xcal
