// Cone form of the DILU sweeps on single meshes: one launch per band of
// levels (one band where the cones fit), no synchronisation between
// workgroups.
//
// The cells are cut into blocks of consecutive Morton keys (a compact region
// of space each).  The workgroup of a block sweeps the block's rows together
// with every row they depend on, directly or through other rows: the block's
// dependency cone.  It walks the cone in level order, one step (at most
// kConeBlock rows of one level) between barriers, and keeps every value it
// computes in LDS, where later rows of the cone gather them.  A cone row of
// another block is computed again, from the same inputs and in the same
// summation order as in its own block, so it has the same value bit for bit.
// Each row sums its entries in the reference's column order
// (forward_substitution!, backward_substitution!,
// preconditioners_1_DILU.jl:84-114), so the sweeps equal the sequential loops
// whatever the blocks.
//
// This pays where cones are small.  On a randomly numbered mesh a chain of
// ascending neighbours is short (18 levels on the 1M-cell box), and a block
// of 1024 cells has a cone of about 2000 rows: one band holds every level.
// On a lexicographically numbered mesh a chain crosses the domain, so the
// levels are cut into bands, one launch each: a block's cone stops at its
// band's first level, and the rows of earlier bands it reads (its halo) are
// loaded from global memory, where the earlier launches stored them.
//
// Memory latency is kept off the step-to-step chain as in the pipelined runs
// (sweep_pipe.h): each row's position-addressed loads (its row index, its
// entries' cone positions and values) are issued P steps ahead, and the loads
// addressed by its row index Q steps ahead, all branch-free, so the compiler
// waits only for the step at hand.  Entries are stored jagged-diagonal (rows
// of a step by entry count, descending), so a step carries no padding.  A
// row's outputs are stored, and the row-indexed inputs that only its own
// rows use (the backward sweep's dot operand) loaded, by its owner block
// only: the other cones hold equal copies, and their stores to the same
// lines from other XCDs cost more than the predicated access.
#pragma once

#include <type_traits>

#include "sweep_pipe.h"

namespace xcal {

// Static cone structure of one sweep direction (device copies of ConeHost).
struct ConeK {
  const int32_t* blk;    // first step of each block (nblk + 1)
  const int32_t* base;   // first position of each block (nblk + 1)
  const int32_t* steps;  // kConeTab ints per step: r0, cnt, eoff[8], ecnt[8], 2 unused
  const int32_t* crow;   // device row of each position (~row: another block's)
  const uint16_t* lidx;  // cone-local position of each entry's dependency
  const int32_t* own;    // forward blocks' own rows: block b owns own[ownp[b], ownp[b+1])
  const int32_t* ownp;
  const int32_t* hnum;   // halo positions of each block
  int32_t blk0;          // this launch's band: global index of its first block
};

struct alignas(16) ConeStep {
  int32_t r0, cnt, eoff[8], ecnt[8], unused[2];
};
static_assert(sizeof(ConeStep) == kConeTab * sizeof(int32_t), "step table row");

// Cone values in LDS; RD adds the 1/D of each position's row (the FP64
// forward sweep, whose entries read it there).
template <class T, bool RD = false>
struct ConeLDS {
  T v[kConeMax];
  T rd[RD ? kConeMax : 1];
  ConeStep tab[kConeSteps];
};

// Stages block b's step table; returns its step count and first position.
// The positions before the first step's are the block's halo: rows of
// earlier bands, whose values `src` holds (stored by earlier launches); they
// are loaded into the cone's first LDS slots (and, with rdl, their 1/D from
// the per-position rdp).
template <class T, class S>
__device__ inline int cone_stage(const ConeK& C, int b, ConeStep* tab, int32_t& base,
                                 const S* __restrict__ src, T* v,
                                 const T* __restrict__ rdp = nullptr, T* rdl = nullptr) {
  const int s0 = C.blk[b], s1 = C.blk[b + 1];
  const int4* srct = reinterpret_cast<const int4*>(C.steps + (int64_t)s0 * kConeTab);
  int4* dst = reinterpret_cast<int4*>(tab);
  for (int j = threadIdx.x; j < (s1 - s0) * (kConeTab / 4); j += blockDim.x) dst[j] = srct[j];
  base = C.base[b];
  const int32_t nh = C.hnum[b];  // halo positions
  for (int32_t p = threadIdx.x; p < nh; p += blockDim.x) {
    const int32_t c = C.crow[base + p];
    v[p] = (T)src[c >= 0 ? c : ~c];
    if (rdl) rdl[p] = rdp[base + p];
  }
  __syncthreads();
  return s1 - s0;
}

// Position-addressed part of a record: the row's position and, for each of
// its entries, the dependency's cone position (-1: none) and entry index.
template <int WK>
__device__ inline void cone_fetch_pos(const ConeK& C, const ConeStep& S, int32_t base,
                                      int32_t& c, int32_t& lp, bool& rep, int32_t (&li)[WK],
                                      int64_t (&e)[WK]) {
  const int32_t cnt = S.cnt;
  rep = (int32_t)threadIdx.x >= cnt;
  const int32_t q = rep ? cnt - 1 : (int32_t)threadIdx.x;
  const int32_t pos = S.r0 + q;
  c = C.crow[pos];
  lp = pos - base;
#pragma unroll
  for (int k = 0; k < WK; ++k) {
    const bool has = q < S.ecnt[k];
    e[k] = has ? (int64_t)S.eoff[k] + q : 0;
    const int l = C.lidx[e[k]];
    li[k] = has ? l : -1;
  }
}

template <class T, int WK, class V>
struct CFRec {
  int32_t c, lp;  // crow entry (row or ~row), cone-local position
  bool rep;       // a repeat of the step's last row
  V ro, qo;       // gathered: UPD r and q of the row; otherwise ro = u
  T rdi;          // 1/D of the row (unfolded factors)
  int32_t li[WK];
  T av[WK];
};

// Forward sweep y = (I + L D^-1)^-1 u (fwd_row's arithmetic): y to Yt in
// device order (the cone backward sweep reads it there); a holds a_ij per
// cone entry and rd 1/D per cone position: each row loads its own 1/D and
// keeps it in LDS beside y, where its dependents' entries read both.  The vectors of W have type V and the sweep
// computes in T: V = double with T = float is the FP32 preconditioner of an
// FP64 Krylov solve (u is rounded to float as the sweep reads it, y stays
// float); V = T otherwise, with Yt = W.Yb.  UPD fuses the CG update: u = r - alpha q, stored
// to W.rout (other cones read the old r of every row during the sweep), and
// x += alpha p over the block's own rows after the sweep (alpha: al_in).
// The caller returns before the sweep once the solve has stopped, so every
// store here is live.
// FP32 (the inner solves of the mixed-precision form, refined in FP64 to
// the FP64 tolerance) takes folded factors: a holds a_ij / D_j per forward
// entry (rd unused) and d holds 1 / D per backward position, so a forward
// entry loads one value and a backward row multiplies.  FP64 keeps the
// reference's (a_ij y_j) (1/D_j) and division.
template <class T>
constexpr bool cone_folded() {
  return std::is_same<T, float>::value;
}

template <bool UPD, class T, int WK, int P, class V = T>
__device__ inline void fwd_cone(const ConeK& C, const T* __restrict__ a, const T* __restrict__ rd,
                                const FwdK<V>& W, T* __restrict__ Yt,
                                ConeLDS<T, !cone_folded<T>()>& L, double al_in) {
  constexpr bool fold = cone_folded<T>();
  const int b = C.blk0 + (int)xcd_block();
  int32_t base;
  const int ns = cone_stage(C, b, L.tab, base, (const T*)Yt, L.v, fold ? nullptr : rd,
                            fold ? nullptr : L.rd);
  const V al = UPD ? (V)al_in : V(0);
  auto fetch = [&](CFRec<T, WK, V>& R, int s) {
    int64_t e[WK];
    cone_fetch_pos<WK>(C, L.tab[s < ns ? s : ns - 1], base, R.c, R.lp, R.rep, R.li, e);
    R.rdi = fold ? T(0) : rd[base + R.lp];
#pragma unroll
    for (int k = 0; k < WK; ++k) R.av[k] = a[e[k]];
  };
  auto gather = [&](CFRec<T, WK, V>& R) {
    const int32_t i = R.c >= 0 ? R.c : ~R.c;
    if (UPD) {
      R.ro = W.r[i];
      R.qo = W.q[i];
    } else {
      R.ro = W.u[i];
    }
  };
  auto process = [&](const CFRec<T, WK, V>& R) {
    T sum = 0;
#pragma unroll
    for (int k = 0; k < WK; ++k) {
      const int li = R.li[k];
      const int lj = li >= 0 ? li : 0;
      const T yv = L.v[lj];
      const T term = fold ? R.av[k] * yv : (R.av[k] * yv) * L.rd[lj];
      sum = li >= 0 ? sum + term : sum;
    }
    const V u = UPD ? R.ro + (-al) * R.qo : R.ro;
    const T y = (T)u - sum;
    L.v[R.lp] = y;
    if (!fold) L.rd[R.lp] = R.rdi;
    const int32_t i = R.c >= 0 ? R.c : ~R.c;
    if (R.c >= 0) {  // the row's owner stores it (other cones hold equal copies)
      Yt[i] = y;
      if (UPD) W.rout[i] = u;
    }
  };
  CFRec<T, WK, V> buf[P];
  pipe_levels<P>(ns, buf, fetch, gather, process);
  if (UPD) {  // x += alpha p over the own rows, U rows per thread at a time
    constexpr int U = 8;
    const int32_t o0 = C.ownp[b], o1 = C.ownp[b + 1];
    for (int32_t j0 = o0 + (int32_t)threadIdx.x; j0 < o1; j0 += U * (int32_t)blockDim.x) {
      int32_t ii[U];
      V xv[U], pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t j = j0 + u * (int32_t)blockDim.x;
        ii[u] = C.own[j < o1 ? j : o0];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xv[u] = W.x[ii[u]];
        pv[u] = W.p[ii[u]];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j0 + u * (int32_t)blockDim.x < o1) W.x[ii[u]] = xv[u] + al * pv[u];
    }
  }
}

template <class T, int WK, class V>
struct CBRec {
  int32_t c, lp;
  bool rep;
  T d, yb;
  V dvi;
  int32_t li[WK];
  T av[WK];
};

// Backward sweep z = (D + U)^-1 y (bwd_row's arithmetic) with up to two dot
// products over the block's own rows (DOT as bwd_row), stored as the
// block's partial pair.  y is read from Yt in device order; a holds a_ij
// per cone entry, d D per cone position.  Types as for the forward sweep: z
// is computed in T and stored to W.Z in V, and the dot products take the
// stored value.  dv must not alias Z (another
// workgroup may store a row's z before the row's owner reads its dv).  As
// for the forward sweep, the caller returns once the solve has stopped.
template <int DOT, class T, int WK, int P, class V = T>
__device__ inline void bwd_cone(const ConeK& C, const T* __restrict__ a, const T* __restrict__ d,
                                const BwdK<V>& W, const T* __restrict__ Yt, ConeLDS<T>& L) {
  const int b = C.blk0 + (int)xcd_block();
  int32_t base;
  const int ns = cone_stage(C, b, L.tab, base, (const V*)W.Z, L.v);
  auto fetch = [&](CBRec<T, WK, V>& R, int s) {
    const ConeStep& S = L.tab[s < ns ? s : ns - 1];
    int64_t e[WK];
    cone_fetch_pos<WK>(C, S, base, R.c, R.lp, R.rep, R.li, e);
    R.d = d[base + R.lp];
#pragma unroll
    for (int k = 0; k < WK; ++k) R.av[k] = a[e[k]];
  };
  auto gather = [&](CBRec<T, WK, V>& R) {
    const int32_t i = R.c >= 0 ? R.c : ~R.c;
    R.yb = Yt[i];
    R.dvi = DOT != 0 && R.c >= 0 ? W.dv[R.c] : V(0);  // own rows only enter the dots
  };
  double s0 = 0.0, s1 = 0.0;
  auto process = [&](const CBRec<T, WK, V>& R) {
    T sum = 0;
#pragma unroll
    for (int k = 0; k < WK; ++k) {
      const int li = R.li[k];
      const T zv = L.v[li >= 0 ? li : 0];
      const T term = R.av[k] * zv;
      sum = li >= 0 ? sum + term : sum;
    }
    const T zt = cone_folded<T>() ? (R.yb - sum) * R.d : (R.yb - sum) / R.d;
    L.v[R.lp] = zt;
    const int32_t i = R.c >= 0 ? R.c : ~R.c;
    const V zi = (V)zt;
    if (R.c >= 0) W.Z[i] = zi;
    const bool own = R.c >= 0 && !R.rep;
    double a0 = 0.0, a1 = 0.0;
    if (DOT == 1 || DOT == 2) a0 = dd(R.dvi, zi);
    if (DOT == 3) { a0 = dd(zi, R.dvi); a1 = dd(zi, zi); }
    if (DOT == 4) { a0 = dd(zi, zi); a1 = dd(R.dvi, zi); }
    s0 = own ? s0 + a0 : s0;
    s1 = own ? s1 + a1 : s1;
  };
  CBRec<T, WK, V> buf[P];
  pipe_levels<P>(ns, buf, fetch, gather, process);
  if (DOT != 0) store_partial<kConeBlock>(s0, s1, W.partial, b);
}

template <int WK>
struct CDRec {
  int32_t c, lp;
  bool rep;
  double d0;  // a_jj
  int32_t li[WK];
  double pr[WK];
};

// update_dilu_diagonal! over the forward cone (the sweep's dependencies):
// D_j = a_jj, then D_j -= (a_ji a_ij) / D_i over its lower entries in column
// order (dilu_diag_row's arithmetic).  prod holds a_ji a_ij per forward cone
// entry, ajj a_jj per forward cone position; D_j goes to D in device order,
// where later bands read their halo.
template <int WK, int P>
__device__ inline void diag_cone(const ConeK& C, const double* __restrict__ prod,
                                 const double* __restrict__ ajj, double* __restrict__ D,
                                 ConeLDS<double>& L) {
  const int b = C.blk0 + (int)xcd_block();
  int32_t base;
  const int ns = cone_stage(C, b, L.tab, base, (const double*)D, L.v);
  auto fetch = [&](CDRec<WK>& R, int s) {
    int64_t e[WK];
    cone_fetch_pos<WK>(C, L.tab[s < ns ? s : ns - 1], base, R.c, R.lp, R.rep, R.li, e);
    R.d0 = ajj[base + R.lp];
#pragma unroll
    for (int k = 0; k < WK; ++k) R.pr[k] = prod[e[k]];
  };
  auto gather = [&](CDRec<WK>&) {};
  auto process = [&](const CDRec<WK>& R) {
    double dj = R.d0;
#pragma unroll
    for (int k = 0; k < WK; ++k) {
      const int li = R.li[k];
      const double di = L.v[li >= 0 ? li : 0];
      dj = li >= 0 ? dj - R.pr[k] / di : dj;
    }
    L.v[R.lp] = dj;
    if (R.c >= 0) D[R.c] = dj;
  };
  CDRec<WK> buf[P];
  pipe_levels<P>(ns, buf, fetch, gather, process);
}

// Host: the kernel view of one sweep direction's cone, for the launch of
// the band whose first block is blk0.
inline ConeK cone_k(const xcal_mesh* m, const xcal_mesh::Cone& c, int32_t blk0) {
  return ConeK{c.blk.p,       c.base.p,        c.steps.p, c.crow.p, c.lidx.p,
               m->cone_own.p, m->cone_ownp.p, c.hnum.p,  blk0};
}

// Host: calls f(WK, P) with the cone sweeps' compile-time entry bound and
// prefetch depth (shallow: 3 steps; deep: 6 steps for up to 4 entries, 4
// above).  Shallow is the default: on the 1M-cell box its fewer registers
// (139 against 175 for the CG forward sweep) let more workgroups share a CU,
// which hid latency better than the deeper prefetch (39.3M against 38.5M
// cell-iters/s).
template <class F>
inline void with_cone_form(int wk, bool shallow, F&& f) {
  auto go = [&](auto W, auto Pdeep) {
    if (shallow) f(W, std::integral_constant<int, 3>{});
    else f(W, Pdeep);
  };
  switch (wk) {
    case 2: go(std::integral_constant<int, 2>{}, std::integral_constant<int, 6>{}); break;
    case 4: go(std::integral_constant<int, 4>{}, std::integral_constant<int, 6>{}); break;
    case 6: go(std::integral_constant<int, 6>{}, std::integral_constant<int, 4>{}); break;
    default: go(std::integral_constant<int, 8>{}, std::integral_constant<int, 4>{});
  }
}

}  // namespace xcal
