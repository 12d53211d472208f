// Fused, device-resident Krylov solves for solve_system!
// (Solve_1_api.jl:268-307): CG and BiCGSTAB (the iterations of Krylov.jl's
// cg!/bicgstab!, compat 0.10, restated) with Jacobi or DILU preconditioning.
//
// Design for gfx950:
//  * All Krylov scalars live in device memory (KS).  Each reduction ends in
//    its kernel's last-finishing block, which updates the scalars
//    (alpha, beta, convergence flags) in place; every kernel returns at entry
//    once the solve has stopped.  The host enqueues batches of iterations
//    and reads the state once per batch instead of once per dot product.
//  * BLAS-1 work is fused into the SpMV and preconditioner sweeps:
//      CG+DILU per iteration: [p = z + beta p ; q = A p ; p.q]
//                             [x += a p ; r -= a q ; forward DILU sweep]
//                             [backward DILU sweep ; r.z ; beta]
//  * DILU sweeps are single sync-free kernels over the level-ordered rows:
//    a row's result is published with one 64-bit atomic store into a buffer
//    pre-filled with a signalling-NaN sentinel, and dependants spin on the
//    value itself (no separate flags, no fences).  Chunks of 64 rows are taken
//    in level order from a ticket counter, so every chunk a wave waits on was
//    taken earlier by a resident wave and the oldest unfinished row can
//    always proceed; the spin is bounded and sets an error flag instead of
//    hanging.  The summation order inside every row is the reference's
//    (ascending column), so the sweeps are bit-identical to
//    forward_substitution!/backward_substitution!
//    (preconditioners_1_DILU.jl:84-114).
//  * Reductions are deterministic: fixed partial-sum slots per block or per
//    chunk, summed in a fixed order by the finishing block.
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <functional>
#include <tuple>
#include <utility>

#include "device_common.h"

namespace xcal {

constexpr uint64_t kSentinel = 0x7FF4DEADBEEF0001ull;  // never produced by arithmetic
constexpr int kArrive = 64;         // first-level arrival counters of a reduction
constexpr unsigned kSpinLimit = 1u << 22;
constexpr int kMaxSweepBlocks = 2048;  // partial-sum slots reserved for the sweep reductions

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

// ------------------------------------------------------- reduction tail
__device__ inline void st_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ld_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Arrival of one of `count` blocks at a reduction.  Same-address atomics
// serialise, so blocks first count in one of 64 first-level counters
// (slot % 64) and only the block completing a group touches the second
// level; returns true in exactly one block, the last to arrive.
__device__ inline bool arrive(unsigned* cnt, int slot, int count) {
  const int groups = count < kArrive ? count : kArrive;
  const int g = slot % kArrive;
  const unsigned gsize = (unsigned)((count - g + kArrive - 1) / kArrive);
  __threadfence();
  if (atomicAdd(&cnt[g], 1u) != gsize - 1u) return false;
  cnt[g] = 0u;  // every member of this group has arrived
  __threadfence();
  if (atomicAdd(&cnt[kArrive], 1u) != (unsigned)groups - 1u) return false;
  cnt[kArrive] = 0u;
  __threadfence();
  return true;
}

// Block `slot` publishes its two partial sums; the last of `count` blocks to
// arrive sums all slots in slot order and runs the scalar update.
__device__ void reduce_tail(double s0, double s1, int slot, int count, double* partial,
                            unsigned* done, KS* ks, int fin, double* lds) {
  __shared__ bool last;
  s0 = block_sum(s0, lds);
  s1 = block_sum(s1, lds);
  if (threadIdx.x == 0) {
    st_agent(&partial[2 * slot], s0);
    st_agent(&partial[2 * slot + 1], s1);
    last = arrive(done, slot, count);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double t0 = 0.0, t1 = 0.0;
  for (int b = threadIdx.x; b < count; b += blockDim.x) {
    t0 += ld_agent(&partial[2 * b]);
    t1 += ld_agent(&partial[2 * b + 1]);
  }
  t0 = block_sum(t0, lds);
  t1 = block_sum(t1, lds);
  if (threadIdx.x == 0) {
    finalize(ks, fin, t0, t1);
    __threadfence();
  }
}

__device__ inline bool stopped(const KS* ks) {
  return __hip_atomic_load(&ks->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != STOP_NONE;
}

__device__ inline double ld_flag(const double* p) {
  return __longlong_as_double(
      (long long)__hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline void st_flag(double* p, double v) {
  __hip_atomic_store((unsigned long long*)p, (unsigned long long)__double_as_longlong(v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool is_sentinel(double v) {
  return (unsigned long long)__double_as_longlong(v) == kSentinel;
}

struct MatK {
  int64_t n;
  const int32_t *rowptr, *colval, *diag;
  const double* nz;
};

// --------------------------------------------------------- static kernels
// r = b - A x0 ; x = 0 ; pold = 0 (CG / BiCGSTAB start)
__global__ void __launch_bounds__(kBlock)
    k_start(MatK A, const double* __restrict__ b, const double* __restrict__ x0,
            double* __restrict__ r, double* __restrict__ x, double* __restrict__ pold) {
  const int64_t i = xcd_row();
  if (i >= A.n) return;
  double acc = 0.0;
  for (int32_t k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) acc += A.nz[k] * x0[A.colval[k]];
  r[i] = b[i] - acc;
  x[i] = 0.0;
  pold[i] = 0.0;
}

// z = S r (Jacobi) or z = r; partials r.z  [CG start]   or
// r = S r0 ; partials r.r, c.r               [BiCGSTAB start]
template <bool BICG>
__global__ void __launch_bounds__(kBlock)
    k_precond_diag_start(int64_t n, const double* __restrict__ S, double* __restrict__ r,
                         double* __restrict__ z, const double* __restrict__ c, double* partial,
                         unsigned* done, KS* ks) {
  __shared__ double lds[kBlock / 64];
  const int64_t i = xcd_row();
  double s0 = 0.0, s1 = 0.0;
  if (i < n) {
    if (BICG) {
      const double ri = S ? S[i] * r[i] : r[i];
      r[i] = ri;
      s0 = ri * ri;
      s1 = c[i] * ri;
    } else {
      const double zi = S ? S[i] * r[i] : r[i];
      z[i] = zi;
      s0 = r[i] * zi;
    }
  }
  reduce_tail(s0, s1, blockIdx.x, gridDim.x, partial, done, ks, BICG ? FIN_BI_INIT : FIN_CG_INIT,
              lds);
}

// CG: pnew = z + beta pold (on the fly for neighbours) ; q = A pnew ; p.q
__global__ void __launch_bounds__(kBlock)
    k_cg_spmv(MatK A, const double* __restrict__ z, const double* __restrict__ pold,
              double* __restrict__ pnew, double* __restrict__ q, double* partial, unsigned* done,
              KS* ks) {
  __shared__ double lds[kBlock / 64];
  if (stopped(ks)) return;
  const double beta = ks->beta;
  const int64_t i = xcd_row();
  double s0 = 0.0;
  if (i < A.n) {
    double acc = 0.0;
    for (int32_t k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) {
      const int j = A.colval[k];
      acc += A.nz[k] * (z[j] + beta * pold[j]);
    }
    const double pi = z[i] + beta * pold[i];
    pnew[i] = pi;
    q[i] = acc;
    s0 = pi * acc;
  }
  reduce_tail(s0, 0.0, blockIdx.x, gridDim.x, partial, done, ks, FIN_CG_PAP, lds);
}

// CG with a diagonal preconditioner: x += a p ; r -= a q ; z = S r ; r.z
__global__ void __launch_bounds__(kBlock)
    k_cg_update_diag(int64_t n, const double* __restrict__ S, const double* __restrict__ p,
                     const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r,
                     double* __restrict__ z, double* partial, unsigned* done, KS* ks) {
  __shared__ double lds[kBlock / 64];
  if (stopped(ks)) return;
  const double a = ks->alpha;
  const int64_t i = xcd_row();
  double s0 = 0.0;
  if (i < n) {
    x[i] = x[i] + a * p[i];
    const double ri = r[i] + (-a) * q[i];
    r[i] = ri;
    const double zi = S ? S[i] * ri : ri;
    z[i] = zi;
    s0 = ri * zi;
  }
  reduce_tail(s0, 0.0, blockIdx.x, gridDim.x, partial, done, ks, FIN_CG_ITER, lds);
}

// BiCGSTAB: pnew = r + beta pold (on the fly) ; q = A pnew ; v = S q (diag
// preconditioner, fused when VDIAG) ; c.v
template <bool VDIAG>
__global__ void __launch_bounds__(kBlock)
    k_bi_spmv_p(MatK A, const double* __restrict__ r, const double* __restrict__ pold,
                double* __restrict__ pnew, const double* __restrict__ S, double* __restrict__ qv,
                const double* __restrict__ c, double* partial, unsigned* done, KS* ks) {
  __shared__ double lds[kBlock / 64];
  if (stopped(ks)) return;
  const double beta = ks->beta;
  const int64_t i = xcd_row();
  double s0 = 0.0;
  if (i < A.n) {
    double acc = 0.0;
    for (int32_t k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) {
      const int j = A.colval[k];
      acc += A.nz[k] * (r[j] + beta * pold[j]);
    }
    pnew[i] = r[i] + beta * pold[i];
    if (VDIAG) {
      const double vi = S ? S[i] * acc : acc;
      qv[i] = vi;
      s0 = c[i] * vi;
    } else {
      qv[i] = acc;
    }
  }
  if (VDIAG)
    reduce_tail(s0, 0.0, blockIdx.x, gridDim.x, partial, done, ks, FIN_BI_ALPHA, lds);
}

// BiCGSTAB: s = r - a v (on the fly) ; d = A s ; x += a p ;
// t = S d and the dots t.s, t.t when TDIAG
template <bool TDIAG>
__global__ void __launch_bounds__(kBlock)
    k_bi_spmv_s(MatK A, const double* __restrict__ r, const double* __restrict__ v,
                const double* __restrict__ p, double* __restrict__ s, double* __restrict__ x,
                const double* __restrict__ S, double* __restrict__ dt, double* partial,
                unsigned* done, KS* ks) {
  __shared__ double lds[kBlock / 64];
  if (stopped(ks)) return;
  const double a = ks->alpha;
  const int64_t i = xcd_row();
  double s0 = 0.0, s1 = 0.0;
  if (i < A.n) {
    double acc = 0.0;
    for (int32_t k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) {
      const int j = A.colval[k];
      acc += A.nz[k] * (r[j] + (-a) * v[j]);
    }
    const double si = r[i] + (-a) * v[i];
    s[i] = si;
    x[i] = x[i] + a * p[i];
    if (TDIAG) {
      const double ti = S ? S[i] * acc : acc;
      dt[i] = ti;
      s0 = ti * si;
      s1 = ti * ti;
    } else {
      dt[i] = acc;
    }
  }
  if (TDIAG)
    reduce_tail(s0, s1, blockIdx.x, gridDim.x, partial, done, ks, FIN_BI_OMEGA, lds);
}

// BiCGSTAB tail: x += w s ; r = s - w t ; p = p - w v ; c.r and r.r
__global__ void __launch_bounds__(kBlock)
    k_bi_tail(int64_t n, const double* __restrict__ s, const double* __restrict__ t,
              const double* __restrict__ v, double* __restrict__ x, double* __restrict__ r,
              double* __restrict__ p, const double* __restrict__ c, double* partial,
              unsigned* done, KS* ks) {
  __shared__ double lds[kBlock / 64];
  if (stopped(ks)) return;
  const double w = ks->omega;
  const int64_t i = xcd_row();
  double s0 = 0.0, s1 = 0.0;
  if (i < n) {
    x[i] = x[i] + w * s[i];
    const double ri = s[i] + (-w) * t[i];
    r[i] = ri;
    p[i] = p[i] + (-w) * v[i];
    s0 = c[i] * ri;
    s1 = ri * ri;
  }
  reduce_tail(s0, s1, blockIdx.x, gridDim.x, partial, done, ks, FIN_BI_END, lds);
}

// x0 = x0 + x (warm start, Krylov.jl adds the start vector last)
__global__ void __launch_bounds__(kBlock)
    k_add(int64_t n, const double* __restrict__ x, double* __restrict__ x0) {
  const int64_t i = xcd_row();
  if (i < n) x0[i] = x[i] + x0[i];
}

__global__ void __launch_bounds__(kBlock) k_fill_bits(int64_t n, double* p, uint64_t bits) {
  const int64_t i = xcd_row();
  if (i < n) p[i] = __longlong_as_double((long long)bits);
}

__global__ void __launch_bounds__(kBlock)
    k_copy_vec(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  const int64_t i = xcd_row();
  if (i < n) b[i] = a[i];
}

// ------------------------------------------------------ sync-free sweeps
// The DILU sweeps run as persistent, co-resident (cooperative) kernels.  The
// level-ordered rows are cut into 64-row chunks; wave w of W processes chunks
// w, w+W, w+2W, ... in order, so the waves work on a moving window of about
// W chunks that follows the dependency front.  Deadlock freedom: the smallest
// unfinished chunk belongs to a wave that has finished all its earlier
// (smaller) chunks and whose rows depend only on finished rows, so it can
// always proceed; co-residency is guaranteed by the cooperative launch.
// Polling backs off with s_sleep and is bounded (error flag, no hang).
struct SweepK {
  int64_t n;
  int nchunks;          // 64-row chunks
  const int32_t* rows;  // level-ordered rows
  unsigned* done;
  double* partial;
  KS* ks;
  int* err;
};

// Forward DILU sweep y = (I + L D^-1)^-1 u, row order and per-row summation
// order of forward_substitution!.  UPD fuses the CG update of the row's own
// x and r first (u = r after r -= a q).  `reset` (may be NULL) is the output
// buffer of the following backward sweep: its row entries are set back to
// the sentinel here.
template <bool UPD>
__global__ void __launch_bounds__(kBlock)
    k_fwd(SweepK W, MatK A, const double* __restrict__ rD, const double* u_in, double* Y,
          double* reset, const double* __restrict__ p, const double* __restrict__ q,
          double* __restrict__ x, double* __restrict__ r) {
  if (W.ks && stopped(W.ks)) return;
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
  const double a = UPD ? W.ks->alpha : 0.0;
  for (int c = wave; c < W.nchunks; c += nwaves) {
    const int64_t pos = (int64_t)c * 64 + lane;
    const bool active = pos < W.n;
    const int i = active ? W.rows[pos] : 0;
    double ui = 0.0;
    int32_t kp = 0, k1 = 0;
    if (active) {
      if (UPD) {
        x[i] = x[i] + a * p[i];
        ui = r[i] + (-a) * q[i];
        r[i] = ui;
      } else {
        ui = u_in[i];
      }
      if (reset) st_flag(&reset[i], __longlong_as_double((long long)kSentinel));
      kp = A.rowptr[i];
      k1 = A.diag[i];
    }
    double sum = 0.0;
    bool done = !active;
    unsigned spins = 0;
    while (true) {
      if (!done) {
        while (kp < k1) {
          const int j = A.colval[kp];
          const double yj = ld_flag(&Y[j]);
          if (is_sentinel(yj)) break;
          sum += (A.nz[kp] * yj) * rD[j];
          ++kp;
        }
        if (kp == k1) {
          st_flag(&Y[i], ui - sum);
          done = true;
        }
      }
      if (__all(done)) break;
      if (++spins > kSpinLimit) {
        atomicOr(W.err, 1);
        return;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
}

// Backward DILU sweep z = (D + U)^-1 y, order of backward_substitution!,
// same scheme as k_fwd.  Reads y from Y (and resets it), publishes z in Z,
// optionally also writes a plain copy to `out`, and accumulates up to two dot
// products:
//   DOT 0: none, 1: r.z (CG), 2: c.z (BiCGSTAB v), 3: z.s and z.z (BiCGSTAB t),
//   4: z.z and c.z (BiCGSTAB start, z = M r0)
template <int DOT>
__global__ void __launch_bounds__(kBlock)
    k_bwd(SweepK W, MatK A, const double* __restrict__ D, double* Y, double* Z,
          double* __restrict__ out, const double* __restrict__ dv, int fin) {
  __shared__ double lds[kBlock / 64];
  if (W.ks && stopped(W.ks)) return;
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
  double s0 = 0.0, s1 = 0.0;
  bool failed = false;
  for (int c = wave; c < W.nchunks && !failed; c += nwaves) {
    const int64_t pos = (int64_t)c * 64 + lane;
    const bool active = pos < W.n;
    const int i = active ? W.rows[pos] : 0;
    double yi = 0.0;
    int32_t kp = 0, k1 = 0;
    if (active) {
      yi = ld_flag(&Y[i]);
      st_flag(&Y[i], __longlong_as_double((long long)kSentinel));
      kp = A.diag[i] + 1;
      k1 = A.rowptr[i + 1];
    }
    double sum = 0.0, zi = 0.0;
    bool done = !active;
    unsigned spins = 0;
    while (true) {
      if (!done) {
        while (kp < k1) {
          const double zj = ld_flag(&Z[A.colval[kp]]);
          if (is_sentinel(zj)) break;
          sum += A.nz[kp] * zj;
          ++kp;
        }
        if (kp == k1) {
          zi = (yi - sum) / D[i];
          st_flag(&Z[i], zi);
          if (out) out[i] = zi;
          done = true;
        }
      }
      if (__all(done)) break;
      if (++spins > kSpinLimit) {
        atomicOr(W.err, 1);
        failed = true;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    if (DOT != 0 && active) {
      if (DOT == 1 || DOT == 2) s0 += dv[i] * zi;
      if (DOT == 3) { s0 += zi * dv[i]; s1 += zi * zi; }
      if (DOT == 4) { s0 += zi * zi; s1 += dv[i] * zi; }
    }
  }
  if (DOT == 0) return;
  reduce_tail(s0, s1, blockIdx.x, gridDim.x, W.partial, W.done, W.ks, fin, lds);
}

// ------------------------------------------------------------ host side
struct Solver;
template <class... P, class... Args>
static void coop(void (*kern)(P...), const Solver& S, Args&&... a);

struct Solver {
  xcal_eqn* e;
  const xcal_mesh* m;
  hipStream_t s;
  int64_t n;
  MatK A;
  SweepK Wf, Wb;
  KS* ks;
  double* partial;
  unsigned* done;
  int grid;
  int sgrid;  // blocks of the cooperative sweep kernels
};

// Launch a sweep kernel cooperatively (all blocks co-resident).  Arguments are
// converted to the kernel's exact parameter types before their addresses
// are handed to the runtime.
template <class... P, size_t... I>
static void coop_impl(void (*kern)(P...), const Solver& S, std::tuple<P...>& t,
                      std::index_sequence<I...>) {
  void* argv[] = {(void*)&std::get<I>(t)...};
  XCAL_HIP(hipLaunchCooperativeKernel((const void*)kern, dim3(S.sgrid), dim3(kBlock), argv, 0, S.s));
}
template <class... P, class... Args>
static void coop(void (*kern)(P...), const Solver& S, Args&&... a) {
  std::tuple<P...> t{static_cast<P>(std::forward<Args>(a))...};
  coop_impl(kern, S, t, std::index_sequence_for<P...>{});
}

// Co-resident block count for the sweeps: bounded by occupancy (required by
// the cooperative launch) and by about four chunks per wave, and capped at
// 512 blocks (2048 waves, a window of ~131k rows: about one DILU level of a
// 1M-cell mesh); XCAL_SWEEP_BLOCKS overrides.
static int sweep_grid(xcal_ctx* c, int nchunks) {
  if (c->coop_max_blocks == 0) {
    int cus = 0;
    XCAL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    int occ = 1 << 30, b = 0;
    const void* kerns[] = {(const void*)k_fwd<true>, (const void*)k_fwd<false>,
                           (const void*)k_bwd<0>, (const void*)k_bwd<1>, (const void*)k_bwd<2>,
                           (const void*)k_bwd<3>, (const void*)k_bwd<4>};
    for (const void* k : kerns) {
      XCAL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, kBlock, 0));
      occ = std::min(occ, b);
    }
    c->coop_max_blocks = std::max(1, occ) * std::max(1, cus);
  }
  int want = std::min(512, (nchunks + 15) / 16);
  if (const char* env = std::getenv("XCAL_SWEEP_BLOCKS")) want = std::min(kMaxSweepBlocks, std::atoi(env));
  return std::max(1, std::min(want, c->coop_max_blocks));
}

static void ensure_workspace(xcal_eqn* e) {
  const int64_t n = e->mesh->n;
  if (e->work.n < 12 * n) e->work.alloc(12 * n);
  const int64_t slots = std::max<int64_t>(grid_for(n), kMaxSweepBlocks) + 8;
  if (e->kpart.n < 2 * slots) e->kpart.alloc(2 * slots);
  if (e->kcount.n < 80) {
    e->kcount.alloc(80);
    XCAL_HIP(hipMemsetAsync(e->kcount.p, 0, e->kcount.bytes(), e->mesh->ctx->stream));
  }
  if (e->kstate.n < (int64_t)sizeof(KS)) e->kstate.alloc(sizeof(KS));
  if (!e->kstate_host) XCAL_HIP(hipHostMalloc((void**)&e->kstate_host, sizeof(KS), 0));
  if (e->yflag.n < n) {
    e->yflag.alloc(n);
    e->zflag.alloc(3 * n);
  }
}

static Solver make_solver(xcal_eqn* e) {
  ensure_workspace(e);
  Solver S;
  S.e = e;
  S.m = e->mesh;
  S.s = e->mesh->ctx->stream;
  S.n = e->mesh->n;
  S.A = MatK{S.n, S.m->rowptr.p, S.m->colval.p, S.m->diag.p, e->nzval.p};
  S.ks = (KS*)e->kstate.p;
  S.partial = e->kpart.p;
  S.done = e->kcount.p + 8;  // kArrive + 1 arrival counters
  S.grid = grid_for(S.n);
  const int nchunks = (int)((S.n + 63) / 64);
  S.sgrid = sweep_grid(S.m->ctx, nchunks);
  S.Wf = SweepK{S.n, nchunks, S.m->lev_rows_f.p, S.done, S.partial, S.ks,
                (int*)(e->kcount.p + 2)};
  S.Wb = S.Wf;
  S.Wb.rows = S.m->lev_rows_b.p;
  return S;
}

static void fill_sentinel(hipStream_t s, double* p, int64_t n) {
  if (n > 0) hipLaunchKernelGGL(k_fill_bits, dim3(grid_for(n)), dim3(kBlock), 0, s, n, p, kSentinel);
}

// z = M^-1 u through the sync-free sweeps (DILU), used by precond_apply
static void dilu_apply(Solver& S, const double* u, double* z) {
  xcal_eqn* e = S.e;
  double* Y = e->yflag.p;
  double* Z = e->zflag.p;
  fill_sentinel(S.s, Y, S.n);
  fill_sentinel(S.s, Z, S.n);
  SweepK wf = S.Wf, wb = S.Wb;
  wf.ks = nullptr;
  wb.ks = nullptr;
  coop(k_fwd<false>, S, wf, S.A, e->prD.p,
                     u, Y, nullptr, nullptr, nullptr, nullptr, nullptr);
  coop(k_bwd<0>, S, wb, S.A, e->pD.p, Y, Z,
                     z, nullptr, (int)FIN_NONE);
  XCAL_HIP(hipGetLastError());
}

// ----------------------------------------------------- kernel timing hook
static hipEvent_t prof_event(xcal_ctx* c) {
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
static void timed(xcal_ctx* c, int k, Fn&& fn) {
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
static void prof_resolve(xcal_ctx* c) {
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

static KS read_state(Solver& S) {
  XCAL_HIP(hipMemcpyAsync(S.e->kstate_host, S.ks, sizeof(KS), hipMemcpyDeviceToHost, S.s));
  XCAL_HIP(hipStreamSynchronize(S.s));
  prof_resolve(S.m->ctx);
  KS k;
  std::memcpy(&k, S.e->kstate_host, sizeof(KS));
  return k;
}

static void init_state(Solver& S, int itmax, double atol, double rtol) {
  KS k{};
  k.atol = atol;
  k.rtol = rtol;
  k.itmax = itmax;
  k.stop = STOP_NONE;
  std::memcpy(S.e->kstate_host, &k, sizeof(KS));
  XCAL_HIP(hipMemcpyAsync(S.ks, S.e->kstate_host, sizeof(KS), hipMemcpyHostToDevice, S.s));
}

static void check_sweep_error(Solver& S) {
  int err = 0;
  XCAL_HIP(hipMemcpyAsync(&err, S.e->kcount.p + 2, sizeof(int), hipMemcpyDeviceToHost, S.s));
  XCAL_HIP(hipStreamSynchronize(S.s));
  if (err) {
    XCAL_HIP(hipMemsetAsync(S.e->kcount.p, 0, S.e->kcount.bytes(), S.s));
    XCAL_HIP(hipStreamSynchronize(S.s));
    throw Error(XCAL_ERR_STATE, "DILU sweep exceeded its spin limit (dependency not published)");
  }
}

static int run_batches(Solver& S, int batch, const std::function<void()>& iteration, KS& out) {
  for (;;) {
    for (int b = 0; b < batch; ++b) iteration();
    XCAL_HIP(hipGetLastError());
    out = read_state(S);
    if (out.stop != STOP_NONE) return out.stop;
  }
}

static int cg(Solver& S, int precon, const double* b, double* x0, int itmax, double atol,
              double rtol, int& niter) {
  xcal_eqn* e = S.e;
  const int64_t n = S.n;
  double* w = e->work.p;
  double *r = w, *x = w + n, *pa = w + 2 * n, *pb = w + 3 * n, *q = w + 4 * n, *zd = w + 5 * n;
  init_state(S, itmax, atol, rtol);
  hipLaunchKernelGGL(k_start, dim3(S.grid), dim3(kBlock), 0, S.s, S.A, b, x0, r, x, pa);
  const bool dilu = precon == XCAL_PRECON_DILU;
  double* Y = e->yflag.p;
  double* Z = e->zflag.p;
  const double* Sd = precon == XCAL_PRECON_JACOBI ? e->pS.p : nullptr;
  if (dilu) {
    fill_sentinel(S.s, Y, n);
    fill_sentinel(S.s, Z, n);
    coop(k_fwd<false>, S, S.Wf, S.A,
                       e->prD.p, r, Y, nullptr, nullptr, nullptr, nullptr, nullptr);
    coop(k_bwd<1>, S, S.Wb, S.A, e->pD.p,
                       Y, Z, nullptr, r, (int)FIN_CG_INIT);
  } else {
    hipLaunchKernelGGL(k_precond_diag_start<false>, dim3(S.grid), dim3(kBlock), 0, S.s, n, Sd, r, zd,
                       nullptr, S.partial, S.done, S.ks);
  }
  const double* z = dilu ? Z : zd;
  int parity = 0;
  auto iteration = [&]() {
    double* pold = parity ? pb : pa;
    double* pnew = parity ? pa : pb;
    parity ^= 1;
    xcal_ctx* c = S.m->ctx;
    timed(c, 0, [&] {
      hipLaunchKernelGGL(k_cg_spmv, dim3(S.grid), dim3(kBlock), 0, S.s, S.A, z, pold, pnew, q,
                         S.partial, S.done, S.ks);
    });
    if (dilu) {
      timed(c, 1, [&] {
        coop(k_fwd<true>, S, S.Wf, S.A,
                           e->prD.p, nullptr, Y, Z, pnew, q, x, r);
      });
      timed(c, 2, [&] {
        coop(k_bwd<1>, S, S.Wb, S.A,
                           e->pD.p, Y, Z, nullptr, r, (int)FIN_CG_ITER);
      });
    } else {
      hipLaunchKernelGGL(k_cg_update_diag, dim3(S.grid), dim3(kBlock), 0, S.s, n, Sd, pnew, q, x, r,
                         zd, S.partial, S.done, S.ks);
    }
  };
  KS st;
  {
    st = read_state(S);
    if (st.stop == STOP_NONE) run_batches(S, e->cg_batch, iteration, st);
  }
  if (dilu) check_sweep_error(S);
  hipLaunchKernelGGL(k_add, dim3(S.grid), dim3(kBlock), 0, S.s, n, x, x0);
  XCAL_HIP(hipGetLastError());
  niter = st.iter;
  // next solve: batch sized to finish in about two reads
  e->cg_batch = std::max(4, std::min(64, niter / 2 + 1));
  return (st.stop == STOP_ITMAX || niter >= itmax) ? 1 : 0;
}

static int bicgstab(Solver& S, int precon, const double* b, double* x0, int itmax, double atol,
                    double rtol, int& niter) {
  xcal_eqn* e = S.e;
  const int64_t n = S.n;
  double* w = e->work.p;
  double *r = w, *x = w + n, *pa = w + 2 * n, *pb = w + 3 * n, *v = w + 4 * n, *sv = w + 5 * n,
         *t = w + 6 * n, *d = w + 7 * n;
  init_state(S, itmax, atol, rtol);
  hipLaunchKernelGGL(k_start, dim3(S.grid), dim3(kBlock), 0, S.s, S.A, b, x0, r, x, pa);
  const bool dilu = precon == XCAL_PRECON_DILU;
  const double* Sd = precon == XCAL_PRECON_JACOBI ? e->pS.p : nullptr;
  double* Y = e->yflag.p;
  double* Zv = e->zflag.p;       // v = M q
  double* Zt = e->zflag.p + n;   // t = M d
  if (dilu) {
    fill_sentinel(S.s, Y, n);
    fill_sentinel(S.s, Zv, 2 * n);
    // r = M r0 ; ||r||^2 and <b, r>
    coop(k_fwd<false>, S, S.Wf, S.A,
                       e->prD.p, r, Y, nullptr, nullptr, nullptr, nullptr, nullptr);
    coop(k_bwd<4>, S, S.Wb, S.A, e->pD.p,
                       Y, Zv, r, b, (int)FIN_BI_INIT);
  } else {
    hipLaunchKernelGGL(k_precond_diag_start<true>, dim3(S.grid), dim3(kBlock), 0, S.s, n, Sd, r,
                       nullptr, b, S.partial, S.done, S.ks);
  }
  int parity = 0;
  auto iteration = [&]() {
    double* pold = parity ? pb : pa;
    double* pnew = parity ? pa : pb;
    parity ^= 1;
    if (dilu) {
      hipLaunchKernelGGL(k_bi_spmv_p<false>, dim3(S.grid), dim3(kBlock), 0, S.s, S.A, r, pold, pnew,
                         nullptr, d, b, S.partial, S.done, S.ks);  // d holds q
      coop(k_fwd<false>, S, S.Wf, S.A,
                         e->prD.p, d, Y, Zv, nullptr, nullptr, nullptr, nullptr);
      coop(k_bwd<2>, S, S.Wb, S.A,
                         e->pD.p, Y, Zv, v, b, (int)FIN_BI_ALPHA);
      hipLaunchKernelGGL(k_bi_spmv_s<false>, dim3(S.grid), dim3(kBlock), 0, S.s, S.A, r, v, pnew, sv,
                         x, nullptr, d, S.partial, S.done, S.ks);
      coop(k_fwd<false>, S, S.Wf, S.A,
                         e->prD.p, d, Y, Zt, nullptr, nullptr, nullptr, nullptr);
      coop(k_bwd<3>, S, S.Wb, S.A,
                         e->pD.p, Y, Zt, t, sv, (int)FIN_BI_OMEGA);
    } else {
      hipLaunchKernelGGL(k_bi_spmv_p<true>, dim3(S.grid), dim3(kBlock), 0, S.s, S.A, r, pold, pnew,
                         Sd, v, b, S.partial, S.done, S.ks);
      hipLaunchKernelGGL(k_bi_spmv_s<true>, dim3(S.grid), dim3(kBlock), 0, S.s, S.A, r, v, pnew, sv,
                         x, Sd, t, S.partial, S.done, S.ks);
    }
    hipLaunchKernelGGL(k_bi_tail, dim3(S.grid), dim3(kBlock), 0, S.s, n, sv, t, v, x, r, pnew, b,
                       S.partial, S.done, S.ks);
  };
  KS st = read_state(S);
  if (st.stop == STOP_NONE) run_batches(S, e->bi_batch, iteration, st);
  if (dilu) check_sweep_error(S);
  hipLaunchKernelGGL(k_add, dim3(S.grid), dim3(kBlock), 0, S.s, n, x, x0);
  XCAL_HIP(hipGetLastError());
  niter = st.iter;
  e->bi_batch = std::max(1, std::min(16, niter + 1));
  return (st.stop == STOP_ITMAX || niter >= itmax) ? 1 : 0;
}

int solve_system(xcal_eqn* e, const xcal_solver_setup* st, double* x, const double* b,
                 xcal_solve_stats* stats) {
  XCAL_CHECK(st->precision == XCAL_PREC_FP64 || st->precision == XCAL_PREC_FP32_IR,
             "unknown precision mode");
  XCAL_CHECK(st->precon == e->precon,
             "preconditioner of the solver setup has not been updated on this equation");
  Solver S = make_solver(e);
  const int itmax = st->itmax == 0 ? (int)(2 * e->mesh->n) : st->itmax;
  int niter = 0, rc;
  if (st->solver == XCAL_SOLVER_CG) rc = cg(S, st->precon, b, x, itmax, st->atol, st->rtol, niter);
  else if (st->solver == XCAL_SOLVER_BICGSTAB)
    rc = bicgstab(S, st->precon, b, x, itmax, st->atol, st->rtol, niter);
  else throw Error(XCAL_ERR_UNSUPPORTED, "unknown Krylov solver");
  if (stats) {
    stats->niter = niter;
    stats->nouter = 1;
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
      Solver S = make_solver(e);
      dilu_apply(S, rp, zp);
      check_sweep_error(S);
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
    c->prof_kernel = enable ? kernel : -1;
    c->prof_launches = 0;
    c->prof_ms = 0.0;
    return 0;
  });
}

int xcal_kernel_timing_read(xcal_ctx* c, int64_t* launches, double* total_ms) {
  return guarded([&] {
    XCAL_CHECK(c && launches && total_ms, "NULL argument");
    XCAL_HIP(hipStreamSynchronize(c->stream));
    prof_resolve(c);
    *launches = c->prof_launches;
    *total_ms = c->prof_ms;
    return 0;
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
