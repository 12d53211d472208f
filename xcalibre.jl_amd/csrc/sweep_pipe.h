// Pipelined one-workgroup DILU sweeps over a run of small levels.
//
// The same sweeps as the level-stepping loops (one thread per row of a
// level, the same row-to-thread assignment and summation orders, hence the
// same results), with the memory latency taken off the level-to-level chain:
//  * the run's level table is staged in LDS once;
//  * each row's own loads -- its values, its packed entries and, per entry,
//    1/D of the column (forward) or the column's backward position
//    (backward) -- are issued P levels ahead into registers;
//  * the outputs of the last kPipeRing rows (positions) live in an LDS ring,
//    from which a row reads the values it depends on.  With `preload`, the
//    ring first receives the kPipeRing rows before the run from global
//    memory (a run that follows other launches of the same sweep).  With
//    FAR, a row also gathers its dependencies from global memory Q levels
//    before it is processed and uses them where they lie before the ring:
//    such a row is more than kPipeRing - kPipeBlock rows back, so at least
//    15 levels back, long since written (and, in a run after other launches,
//    written by them).
// The mesh build admits a run only if it fits: <= kPipeBlock rows per
// level, <= kPipeLev levels, <= WK entries per row; FAR is set when some
// dependency lies before the ring.  The code is branch-free apart from the
// level loop, so the
// compiler can count outstanding loads exactly and wait only for the level at
// hand: threads past a level's last row repeat that row (same inputs, same
// values stored, dot products masked), and entries past w load entry 0 of
// the packed arrays and are masked.
#pragma once

#include <type_traits>

#include "solver_common.h"

namespace xcal {

// LDS of one pipelined run: the ring and the staged level table
template <class T>
struct PipeLDS {
  T ring[kPipeRing];
  int32_t r0[kPipeLev + 1], base[kPipeLev], w[kPipeLev];
};

// A thread's records of one level when each thread works RW rows of it
// (kPipeBlock / RW threads): row slot j of thread t is the level's row
// t + j kPipeBlock / RW, the row thread t + j kPipeBlock / RW would have in
// a workgroup of kPipeBlock threads.  Rows, values and summation orders are
// the same for every RW; fewer threads take fewer barrier round trips per
// level but issue more per thread.
template <class Rec, int RW>
struct PipeRows {
  Rec r[RW];
};

// Stages levels [u0, u0 + nl) (row starts, packed bases, entry counts).
__device__ inline void stage_levels(const int32_t* __restrict__ lev_ptr,
                                    const int32_t* __restrict__ meta, int u0, int nl,
                                    int32_t* s_r0, int32_t* s_base, int32_t* s_w) {
  for (int j = threadIdx.x; j <= nl; j += blockDim.x) {
    s_r0[j] = lev_ptr[u0 + j];
    if (j < nl) {
      s_base[j] = meta[2 * (u0 + j)];
      s_w[j] = meta[2 * (u0 + j) + 1];
    }
  }
  __syncthreads();
}

// Drives P-deep prefetch: fetch(R, l) loads level l's record (a repeat of
// the last level past the end); gather(R) issues the record's dependent
// gathers Q levels before it is processed; process(R) works it; then a
// barrier.  Full groups of P levels run without conditionals; the last
// nl % P follow.
template <int P, class Rec, class Fetch, class Gather, class Process>
__device__ inline void pipe_levels(int nl, Rec (&buf)[P], Fetch&& fetch, Gather&& gather,
                                   Process&& process) {
  constexpr int Q = P / 2 > 0 ? P / 2 : 1;
#pragma unroll
  for (int s = 0; s < P; ++s) {
    fetch(buf[s], s);
    if (s < Q) gather(buf[s]);
  }
  const int nfull = nl - nl % P;
  for (int l0 = 0; l0 < nfull; l0 += P) {
#pragma unroll
    for (int s = 0; s < P; ++s) {
      if (Q < P) gather(buf[(s + Q) % P]);  // level l0 + s + Q (a repeat past the end)
      process(buf[s]);
      fetch(buf[s], l0 + s + P);
      __syncthreads();
    }
  }
#pragma unroll
  for (int s = 0; s < P; ++s)
    if (nfull + s < nl) {  // uniform
      if (s + Q < P && nfull + s + Q < nl) gather(buf[s + Q < P ? s + Q : 0]);
      process(buf[s]);
      __syncthreads();
    }
}

template <class T, int WK>
struct FRec {
  int32_t i, pb, lo;  // lo: first row still in the ring at this level
  T u;
  int32_t col[WK];
  T av[WK], rv[WK], yg[WK];
};

// Loads the ring for positions [first - kPipeRing, first) from global
// memory, through value(p) (a run after other launches of its sweep).
template <class T, class V>
__device__ inline void preload_ring(T* ring, int32_t first, V&& value) {
  for (int j = threadIdx.x; j < kPipeRing; j += blockDim.x) {
    const int32_t p = first - kPipeRing + j;
    if (p >= 0) ring[p & (kPipeRing - 1)] = value(p);
  }
  __syncthreads();
}

// Forward sweep y = (I + L D^-1)^-1 u over levels [u0, u0 + nl): y to the
// ring, to Yb (backward order) and, with STOREY, to Y (for later launches).
template <bool STOREY, bool FAR, class T, int WK, int P, int RW = 1>
__device__ inline void fwd_pipe(const FwdK<T>& W, const T* __restrict__ uin,
                                const T* __restrict__ Lrd, const int32_t* __restrict__ lev_ptr,
                                const int32_t* __restrict__ meta, int u0, int nl, bool preload,
                                PipeLDS<T>& L) {
  constexpr int NT = kPipeBlock / RW;
  stage_levels(lev_ptr, meta, u0, nl, L.r0, L.base, L.w);
  const int32_t first = L.r0[0];
  if (preload) preload_ring(L.ring, first, [&](int32_t p) { return W.Y[p]; });
  const int32_t ring_lo = preload ? first - kPipeRing : first;
  auto fetch1 = [&](FRec<T, WK>& R, int l, int32_t vt) {
    const int lc = l < nl ? l : nl - 1;
    const int32_t r0 = L.r0[lc], cnt = L.r0[lc + 1] - r0;
    const int32_t q = vt < cnt ? vt : cnt - 1;
    const int w = L.w[lc];
    const int64_t e0 = (int64_t)L.base[lc] + q;
    const int32_t i = r0 + q;
    R.i = i;
    R.lo = max(ring_lo, r0 + cnt - kPipeRing);
    R.pb = W.posb[i];
    R.u = uin[i];
#pragma unroll
    for (int t = 0; t < WK; ++t) {
      const int64_t e = t < w ? e0 + (int64_t)t * cnt : 0;
      const int c = W.Lcol[e];
      R.col[t] = t < w ? c : -1;
      R.av[t] = W.Lnz[e];
      R.rv[t] = Lrd[e];
    }
  };
  auto gather1 = [&](FRec<T, WK>& R) {
    if (FAR) {
#pragma unroll
      for (int t = 0; t < WK; ++t) R.yg[t] = W.Y[R.col[t] >= 0 ? R.col[t] : 0];
    }
  };
  auto process1 = [&](const FRec<T, WK>& R) {
    T sum = 0;
#pragma unroll
    for (int t = 0; t < WK; ++t) {
      const int c = R.col[t];
      T yv = L.ring[(c >= 0 ? c : 0) & (kPipeRing - 1)];
      if (FAR) yv = c >= R.lo ? yv : R.yg[t];
      const T term = (R.av[t] * yv) * R.rv[t];
      sum = c >= 0 ? sum + term : sum;
    }
    const T y = R.u - sum;
    L.ring[R.i & (kPipeRing - 1)] = y;
    if (STOREY) W.Y[R.i] = y;
    W.Yb[R.pb] = y;
  };
  using Rec = PipeRows<FRec<T, WK>, RW>;
  auto fetch = [&](Rec& R, int l) {
#pragma unroll
    for (int j = 0; j < RW; ++j) fetch1(R.r[j], l, (int32_t)threadIdx.x + j * NT);
  };
  auto gather = [&](Rec& R) {
#pragma unroll
    for (int j = 0; j < RW; ++j) gather1(R.r[j]);
  };
  auto process = [&](const Rec& R) {
#pragma unroll
    for (int j = 0; j < RW; ++j) process1(R.r[j]);
  };
  Rec buf[P];
  pipe_levels<P>(nl, buf, fetch, gather, process);
}

template <class T, int WK>
struct BRec {
  int32_t pos, i, lo;
  bool own;  // the thread's own row (not a repeat of the level's last row)
  T yb, d, dvi;
  int32_t up[WK], col[WK];
  T av[WK], zg[WK];
};

// Backward sweep z = (D + U)^-1 y over backward levels [u0, u0 + nl) with
// up to two dot products (DOT as bwd_row) into s0, s1.  The dot operand
// dv[i] is gathered with the record's dependencies, Q levels before its
// row is processed (its row index arrives with the record, P levels ahead).
template <int DOT, bool FAR, class T, int WK, int P, int RW = 1>
__device__ inline void bwd_pipe(const BwdK<T>& W, const int32_t* __restrict__ Upos,
                                const int32_t* __restrict__ lev_ptr,
                                const int32_t* __restrict__ meta, int u0, int nl, bool preload,
                                PipeLDS<T>& L, double (&s0)[RW], double (&s1)[RW]) {
  constexpr int NT = kPipeBlock / RW;
  stage_levels(lev_ptr, meta, u0, nl, L.r0, L.base, L.w);
  const int32_t first = L.r0[0];
  if (preload) preload_ring(L.ring, first, [&](int32_t p) { return W.Z[W.rows[p]]; });
  const int32_t ring_lo = preload ? first - kPipeRing : first;
  auto fetch1 = [&](BRec<T, WK>& R, int l, int32_t vt) {
    const int lc = l < nl ? l : nl - 1;
    const int32_t p0 = L.r0[lc], cnt = L.r0[lc + 1] - p0;
    R.own = l < nl && vt < cnt;
    const int32_t q = vt < cnt ? vt : cnt - 1;
    const int w = L.w[lc];
    const int64_t e0 = (int64_t)L.base[lc] + q;
    const int32_t pos = p0 + q;
    R.pos = pos;
    R.lo = max(ring_lo, p0 + cnt - kPipeRing);
    R.i = W.rows[pos];
    R.yb = W.Yb[pos];
    R.d = W.Db[pos];
#pragma unroll
    for (int t = 0; t < WK; ++t) {
      const int64_t e = t < w ? e0 + (int64_t)t * cnt : 0;
      const int up = Upos[e];  // -1 on padding
      R.up[t] = t < w ? up : -1;
      if (FAR) R.col[t] = W.Ucol[e];
      R.av[t] = W.Unz[e];
    }
  };
  auto gather1 = [&](BRec<T, WK>& R) {
    if (DOT != 0) R.dvi = W.dv[R.i];
    if (FAR) {
#pragma unroll
      for (int t = 0; t < WK; ++t) R.zg[t] = W.Z[R.up[t] >= 0 ? R.col[t] : 0];
    }
  };
  auto process1 = [&](const BRec<T, WK>& R, double& s0, double& s1) {
    T sum = 0;
#pragma unroll
    for (int t = 0; t < WK; ++t) {
      const int up = R.up[t];
      T zv = L.ring[(up >= 0 ? up : 0) & (kPipeRing - 1)];
      if (FAR) zv = up >= R.lo ? zv : R.zg[t];
      const T term = R.av[t] * zv;
      sum = up >= 0 ? sum + term : sum;
    }
    const T zi = (R.yb - sum) / R.d;
    L.ring[R.pos & (kPipeRing - 1)] = zi;
    W.Z[R.i] = zi;
    const T dvi = DOT != 0 ? R.dvi : T(0);
    double a0 = 0.0, a1 = 0.0;
    if (DOT == 1 || DOT == 2) a0 = dd(dvi, zi);
    if (DOT == 3) { a0 = dd(zi, dvi); a1 = dd(zi, zi); }
    if (DOT == 4) { a0 = dd(zi, zi); a1 = dd(dvi, zi); }
    s0 = R.own ? s0 + a0 : s0;
    s1 = R.own ? s1 + a1 : s1;
  };
  using Rec = PipeRows<BRec<T, WK>, RW>;
  auto fetch = [&](Rec& R, int l) {
#pragma unroll
    for (int j = 0; j < RW; ++j) fetch1(R.r[j], l, (int32_t)threadIdx.x + j * NT);
  };
  auto gather = [&](Rec& R) {
#pragma unroll
    for (int j = 0; j < RW; ++j) gather1(R.r[j]);
  };
  auto process = [&](const Rec& R) {
#pragma unroll
    for (int j = 0; j < RW; ++j) process1(R.r[j], s0[j], s1[j]);
  };
  Rec buf[P];
  pipe_levels<P>(nl, buf, fetch, gather, process);
}

// update_dilu_diagonal! (preconditioners_1_DILU.jl:43-75) over levels
// [u0, u0 + nl): D_j starts from a_jj and takes D_j -= (a_ij a_ji) / D_i over
// its lower entries in column order; prod holds a_ij a_ji per packed entry,
// D_i comes from the ring (or, with FAR, from D itself).
template <bool FAR, int WK, int P, int RW = 1>
__device__ inline void diag_pipe(double* D, const double* __restrict__ prod,
                                 const int32_t* __restrict__ Lcol,
                                 const int32_t* __restrict__ lev_ptr,
                                 const int32_t* __restrict__ meta, int u0, int nl, bool preload,
                                 PipeLDS<double>& L) {
  stage_levels(lev_ptr, meta, u0, nl, L.r0, L.base, L.w);
  const int32_t first = L.r0[0];
  if (preload) preload_ring(L.ring, first, [&](int32_t p) { return D[p]; });
  const int32_t ring_lo = preload ? first - kPipeRing : first;
  struct Rec {
    int32_t j, lo;
    double d0;
    int32_t col[WK];
    double pr[WK], dg[WK];
  };
  auto fetch1 = [&](Rec& R, int l, int32_t vt) {
    const int lc = l < nl ? l : nl - 1;
    const int32_t r0 = L.r0[lc], cnt = L.r0[lc + 1] - r0;
    const int32_t q = vt < cnt ? vt : cnt - 1;
    const int w = L.w[lc];
    const int64_t e0 = (int64_t)L.base[lc] + q;
    R.j = r0 + q;
    R.lo = max(ring_lo, r0 + cnt - kPipeRing);
    R.d0 = D[R.j];
#pragma unroll
    for (int t = 0; t < WK; ++t) {
      const int64_t e = t < w ? e0 + (int64_t)t * cnt : 0;
      const int c = Lcol[e];
      R.col[t] = t < w ? c : -1;
      R.pr[t] = prod[e];
    }
  };
  auto gather1 = [&](Rec& R) {
    if (FAR) {
#pragma unroll
      for (int t = 0; t < WK; ++t) R.dg[t] = D[R.col[t] >= 0 ? R.col[t] : 0];
    }
  };
  auto process1 = [&](const Rec& R) {
    double Dj = R.d0;
#pragma unroll
    for (int t = 0; t < WK; ++t) {
      const int c = R.col[t];
      double di = L.ring[(c >= 0 ? c : 0) & (kPipeRing - 1)];
      if (FAR) di = c >= R.lo ? di : R.dg[t];
      const double nd = Dj - R.pr[t] / di;
      Dj = c >= 0 ? nd : Dj;
    }
    L.ring[R.j & (kPipeRing - 1)] = Dj;
    D[R.j] = Dj;
  };
  constexpr int NT = kPipeBlock / RW;
  using Recs = PipeRows<Rec, RW>;
  auto fetch = [&](Recs& R, int l) {
#pragma unroll
    for (int j = 0; j < RW; ++j) fetch1(R.r[j], l, (int32_t)threadIdx.x + j * NT);
  };
  auto gather = [&](Recs& R) {
#pragma unroll
    for (int j = 0; j < RW; ++j) gather1(R.r[j]);
  };
  auto process = [&](const Recs& R) {
#pragma unroll
    for (int j = 0; j < RW; ++j) process1(R.r[j]);
  };
  Recs buf[P];
  pipe_levels<P>(nl, buf, fetch, gather, process);
}

// prod[e] = a_ij a_ji of packed lower entry e (nz[trans[slot]] nz[slot]; 0 on
// padding)
static __global__ void __launch_bounds__(kBlock)
    k_lower_products(int64_t total, const int32_t* __restrict__ slot,
                     const int32_t* __restrict__ trans, const double* __restrict__ nz,
                     double* __restrict__ prod) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int32_t k = slot[e];
  const int32_t kt = k >= 0 ? trans[k] : -1;
  prod[e] = kt >= 0 ? nz[kt] * nz[k] : 0.0;
}

// Host: calls f(WK, FAR) with the compile-time form (entry bound 2, 4 or 8,
// far dependencies or not) of a pipelined launch.
template <class F>
inline void with_pipe_form(int wk, bool far, F&& f) {
  auto go = [&](auto W) {
    if (far) f(W, std::true_type{});
    else f(W, std::false_type{});
  };
  switch (wk) {
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    default: go(std::integral_constant<int, 8>{});
  }
}

// Prefetch depth by entry bound: as deep as the 63 outstanding loads of a
// wavefront allow, with whole groups of levels per loop iteration.
template <int WK, int RW = 1>
constexpr int pipe_depth() {
  constexpr int d = WK <= 2 ? 8 : WK <= 4 ? 6 : 4;  // levels ahead at one row per thread
  return d / RW >= 2 ? d / RW : 2;                 // about as many records in flight
}

}  // namespace xcal
