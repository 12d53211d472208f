// Device-side views and the per-face coefficient functions shared by the
// assembly kernels.  Every expression keeps the reference's evaluation order
// (the library is built with -ffp-contract=off) so that results match the
// oracle bit for bit where the reference's order is defined.
#pragma once

#include "xcal_internal.h"

namespace xcal {

struct MeshK {
  int64_t n, nf, nbf;
  const double* vol;
  const int32_t *cf_ptr, *cf_face, *cf_nbr, *cf_slot;
  const int8_t* cf_sign;
  const double *cf_w, *cf_lapEf, *cf_delta;  // face values per cf entry
  const double *cf_sx, *cf_sy, *cf_sz;       // nsign S_f per cf entry
  const int32_t* diag;
  const int32_t *own, *nei, *fslot;
  const double *area, *weight, *delta, *nx, *ny, *nz, *ex, *ey, *ez, *sfx, *sfy, *sfz, *lapEf;
  int64_t nbcell;
  const int32_t *bc_cells, *bc_ptr, *bc_face;
  const int8_t* bface_patch;
};

// Matrix view (sliced-ELL storage): the entries of row i sit at
// rb[i], rb[i] + kSlice, ... < re[i], in the reference's column order, and
// diag[i] is the slot of the diagonal entry.
template <class T>
struct MatKT {
  int64_t n;
  const int32_t *rb, *re, *colval, *diag;
  const T* nz;
  const int32_t* blk;  // row blocks in dispatch order (xcal_mesh::xblk)
  int32_t rw;          // most entries in one row (selects row_sum's group size)
  const int32_t* sinfo = nullptr;  // per slice: first slot, width (single meshes)
};
using MatK = MatKT<double>;
template <class T>
inline MatKT<T> matk(const xcal_mesh* m, const T* nz) {
  return MatKT<T>{m->n, m->rbeg.p, m->rend.p, m->colval.p, m->diag.p, nz, m->xblk_d.p, m->roww,
                  m->sinfo.p};
}
// Row of this work-item in a launch over A's row blocks (grid = nxblk);
// valid while row < end.
template <class T>
__device__ inline int64_t block_row(const MatKT<T>& A, int64_t& end) {
  const int32_t* b = A.blk + 2 * blockIdx.x;
  end = b[1];
  return (int64_t)b[0] + threadIdx.x;
}

inline MeshK meshk(const xcal_mesh* m) {
  MeshK k;
  k.n = m->n;
  k.nf = m->nf;
  k.nbf = m->nbf;
  k.vol = m->vol.p;
  k.cf_ptr = m->cf_ptr.p;
  k.cf_face = m->cf_face.p;
  k.cf_nbr = m->cf_nbr.p;
  k.cf_slot = m->cf_slot.p;
  k.cf_sign = m->cf_sign.p;
  k.cf_w = m->cf_w.p;
  k.cf_lapEf = m->cf_lapEf.p;
  k.cf_delta = m->cf_delta.p;
  k.cf_sx = m->cf_sx.p;
  k.cf_sy = m->cf_sy.p;
  k.cf_sz = m->cf_sz.p;
  k.diag = m->diag.p;
  k.own = m->own.p;
  k.nei = m->nei.p;
  k.fslot = m->fslot.p;
  k.area = m->area.p;
  k.weight = m->weight.p;
  k.delta = m->delta.p;
  k.nx = m->nx.p;
  k.ny = m->ny.p;
  k.nz = m->nz.p;
  k.ex = m->ex.p;
  k.ey = m->ey.p;
  k.ez = m->ez.p;
  k.sfx = m->sfx.p;
  k.sfy = m->sfy.p;
  k.sfz = m->sfz.p;
  k.lapEf = m->lapEf.p;
  k.nbcell = m->nbcell;
  k.bc_cells = m->bc_cells.p;
  k.bc_ptr = m->bc_ptr.p;
  k.bc_face = m->bc_face.p;
  k.bface_patch = m->bface_patch.p;
  return k;
}

// sum over row i of nz[k] * xv(colval[k]) in the row's column order, G
// entries at a time.  All loads of a group are issued unconditionally: slots
// past the row's last entry re-read that entry (every row holds at least its
// diagonal) and are left out of the sum.  With no branch between them, the
// compiler issues the group's column and value loads together, then all its
// gathers, and waits for each gathered value only where it is added; a
// conditional load per entry made it wait for every load after each entry.
template <int G, class T, class XV>
__device__ inline T row_sum_g(const MatKT<T>& A, int64_t i, XV&& xv) {
  const int32_t k1 = A.re[i];
  const int32_t klast = k1 - kSlice;
  T acc = 0;
  for (int32_t k0 = A.rb[i]; k0 < k1; k0 += G * kSlice) {
    int col[G];
    T a[G], x[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int32_t k = min(k0 + u * kSlice, klast);
      // streamed once: keep them out of the way of the gathers
      col[u] = __builtin_nontemporal_load(&A.colval[k]);
      a[u] = __builtin_nontemporal_load(&A.nz[k]);
    }
#pragma unroll
    for (int u = 0; u < G; ++u) x[u] = xv(col[u]);
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const T t = a[u] * x[u];
      acc = k0 + u * kSlice < k1 ? acc + t : acc;
    }
  }
  return acc;
}
// The group is as wide as the mesh's widest row, up to 8 (quads and tets 5,
// hexahedra 7), so a row's loads are issued in one group with few re-reads.
// A.rw is uniform over the launch: the choice is a scalar branch.
// G > 0 fixes the group at compile time (a kernel instantiated per group
// width keeps the registers, and so the occupancy, of its own width).
template <int G = 0, class T, class XV>
__device__ inline T row_sum(const MatKT<T>& A, int64_t i, XV&& xv) {
  if constexpr (G > 0) {
    return row_sum_g<G>(A, i, xv);
  } else {
    if (A.rw <= 5) return row_sum_g<5>(A, i, xv);
    if (A.rw <= 7) return row_sum_g<7>(A, i, xv);
    return row_sum_g<8>(A, i, xv);
  }
}
// row_sum_g over a whole slice width: row i's entries are found from its
// slice's first slot and width (one small table, cached, instead of the
// row's own begin/end), and the slice's padding -- column 0, value 0 --
// is summed like an entry.  a*x of a padding slot is +-0, so the sum is the
// same as row_sum_g's (up to the sign of an all-zero row's 0) for finite x.
// Needs the slice width <= G (rw <= G) and zeroed padding in nz.
template <int G, class T, class XV>
__device__ inline T row_sum_slice(const MatKT<T>& A, int64_t i, XV&& xv) {
  const int2 si = reinterpret_cast<const int2*>(A.sinfo)[i / kSlice];
  const int32_t k0 = si.x + (int32_t)(i % kSlice), w = si.y;
  int col[G];
  T a[G], x[G];
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const int32_t k = k0 + min(u, w - 1) * kSlice;
    col[u] = __builtin_nontemporal_load(&A.colval[k]);
    a[u] = __builtin_nontemporal_load(&A.nz[k]);
  }
#pragma unroll
  for (int u = 0; u < G; ++u) x[u] = xv(col[u]);
  T acc = 0;
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const T t = a[u] * x[u];
    acc = u < w ? acc + t : acc;
  }
  return acc;
}

// A row sum over entries nz * f(U[col], W[col]) in two phases: row_pair_load
// issues the row's first G entries and both gathers, before the caller has
// the scalars f needs (a Krylov state formed in the same kernel);
// row_pair_sum then adds in the row's order -- the same arithmetic as
// row_sum<G> with xv(j) = f(U[j], W[j]) -- and sums any entries past the
// first group itself.
template <int G, class T>
struct RowPair {
  int32_t k0, k1;
  T a[G], u[G], w[G];
};
template <int G, class T>
__device__ inline RowPair<G, T> row_pair_load(const MatKT<T>& A, int64_t i,
                                              const T* __restrict__ U, const T* __restrict__ W) {
  RowPair<G, T> R;
  R.k0 = A.rb[i];
  R.k1 = A.re[i];
  const int32_t klast = R.k1 - kSlice;
  int col[G];
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const int32_t k = min(R.k0 + u * kSlice, klast);
    col[u] = __builtin_nontemporal_load(&A.colval[k]);
    R.a[u] = __builtin_nontemporal_load(&A.nz[k]);
  }
#pragma unroll
  for (int u = 0; u < G; ++u) {
    R.u[u] = U[col[u]];
    R.w[u] = W[col[u]];
  }
  return R;
}
template <int G, class T, class F>
__device__ inline T row_pair_sum(const MatKT<T>& A, const RowPair<G, T>& R,
                                 const T* __restrict__ U, const T* __restrict__ W, F&& f) {
  T acc = 0;
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const T t = R.a[u] * f(R.u[u], R.w[u]);
    acc = R.k0 + u * kSlice < R.k1 ? acc + t : acc;
  }
  const int32_t klast = R.k1 - kSlice;
  for (int32_t k0 = R.k0 + G * kSlice; k0 < R.k1; k0 += G * kSlice) {  // rows wider than G
    int col[G];
    T a[G], x[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int32_t k = min(k0 + u * kSlice, klast);
      col[u] = A.colval[k];
      a[u] = A.nz[k];
    }
#pragma unroll
    for (int u = 0; u < G; ++u) x[u] = f(U[col[u]], W[col[u]]);
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const T t = a[u] * x[u];
      acc = k0 + u * kSlice < R.k1 ? acc + t : acc;
    }
  }
  return acc;
}

// Host side: calls f(std::integral_constant<int, G>) with row_sum's group
// width for rows of at most rw entries.
template <class F>
inline void with_row_group(int rw, F&& f) {
  if (rw <= 5) f(std::integral_constant<int, 5>{});
  else if (rw <= 7) f(std::integral_constant<int, 7>{});
  else f(std::integral_constant<int, 8>{});
}

// Julia's max for Float64: NaN-propagating, +0.0 preferred over -0.0
__device__ __host__ inline double jmax(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a == b) return __builtin_signbit(a) ? b : a;
  return a > b ? a : b;
}

__device__ inline double term_flux(const TermK& t, int64_t i) {
  return t.flux ? t.flux[i] : t.flux_const;
}

// Terms with a face contribution, whose flux is a face field
__device__ __host__ inline bool face_term(int type) {
  return type == XCAL_LAPLACIAN_LINEAR || type == XCAL_DIV_LINEAR || type == XCAL_DIV_UPWIND ||
         type == XCAL_DIV_LUST || type == XCAL_DIV_BOUNDED_UPWIND;
}

// scheme! (Discretise_1_schemes.jl) for one internal face visit, from the
// face's values: the term's flux J, the weight w, |E_f| (lapEf) and delta.
__device__ inline void scheme_vals(const TermK& t, double J, double ns, double w0, double lapEf,
                                   double delta, double& ac, double& an) {
  const double sign = (double)t.sign;
  switch (t.type) {
    case XCAL_LAPLACIAN_LINEAR: {  // :68-104
      double ap = (sign * (J * lapEf)) / delta;
      ac = -ap;
      an = ap;
      return;
    }
    case XCAL_DIV_LINEAR: {  // :114-129
      double w = 0.5 + ns * (w0 - 0.5);
      double ap = sign * (J * ns);
      ac = ap * w;
      an = ap * (1.0 - w);
      return;
    }
    case XCAL_DIV_UPWIND: {  // :137-147
      double ap = sign * (J * ns);
      ac = jmax(ap, 0.0);
      an = -jmax(-ap, 0.0);
      return;
    }
    case XCAL_DIV_LUST: {  // :155-175
      double w = w0;
      if (ns < 0.0) w = 1.0 - w;
      double ap = sign * (J * ns);
      double acL = ap * w, anL = ap * (1.0 - w);
      double acU = jmax(ap, 0.0), anU = -jmax(-ap, 0.0);
      ac = 0.75 * acL + 0.25 * acU;
      an = 0.75 * anL + 0.25 * anU;
      return;
    }
    case XCAL_DIV_BOUNDED_UPWIND: {  // :183-195
      double ap = sign * (J * ns);
      ac = jmax(-ap, 0.0);
      an = -jmax(-ap, 0.0);
      return;
    }
    default:  // Time, Si: no face contribution
      ac = 0.0;
      an = 0.0;
      return;
  }
}
__device__ inline void scheme(const MeshK& m, const TermK& t, int f, double ns, double& ac,
                              double& an) {
  scheme_vals(t, term_flux(t, f), ns, m.weight[f], m.lapEf[f], m.delta[f], ac, an);
}

// Visits a cell's internal faces k0 <= k < k1 (its cf list) kFaceGroup at a
// time: load(k) forms the face's record (loads addressed by k and by the
// face id) for the whole group, slots past k1 repeating the last face, before
// use(record, valid) runs on each in list order.  With no branch between the
// loads, the compiler issues a group's loads together and waits for each
// only where it is used; a loop over dependent loads paid two round trips
// per face.  Six faces cover a hexahedron.
constexpr int kFaceGroup = 6;
template <class Load, class Use>
__device__ inline void cell_faces(int32_t k0, int32_t k1, Load&& load, Use&& use) {
  for (int32_t kb = k0; kb < k1; kb += kFaceGroup) {
    using R = decltype(load(k0));
    R rec[kFaceGroup];
#pragma unroll
    for (int u = 0; u < kFaceGroup; ++u) rec[u] = load(min(kb + u, k1 - 1));
#pragma unroll
    for (int u = 0; u < kFaceGroup; ++u) use(rec[u], kb + u < k1);
  }
}

// scheme_source! (Discretise_1_schemes.jl) per cell
__device__ inline void scheme_source(const MeshK& m, const TermK& t, int64_t c,
                                     const double* prev, double dt, double& ac, double& b) {
  if (t.type == XCAL_TIME_EULER) {  // :30-42
    double vol_rdt = (term_flux(t, c) * m.vol[c]) / dt;
    ac = vol_rdt;
    b = prev[c] * vol_rdt;
  } else if (t.type == XCAL_SI) {  // :208-218
    ac = ((double)t.sign * term_flux(t, c)) * m.vol[c];
    b = 0.0;
  } else {
    ac = 0.0;
    b = 0.0;
  }
}

// Boundary-condition functors (files under Discretise/boundary_conditions/).
// comp < 0: scalar equation; 0..2: vector component.
__device__ inline void bc_coeffs(const MeshK& m, const BCK& bc, const TermK& t, int f, int c,
                                 int comp, const double* px, const double* py,
                                 const double* pz, double& ap_out, double& bp_out) {
  ap_out = 0.0;
  bp_out = 0.0;
  if (t.type == XCAL_TIME_STEADY || t.type == XCAL_TIME_EULER || t.type == XCAL_SI) return;
  const double sign = (double)t.sign;
  const bool vec = comp >= 0;
  const double value = vec ? bc.v[comp] : bc.v[0];
  const bool lap = t.type == XCAL_LAPLACIAN_LINEAR;
  const bool divBU = t.type == XCAL_DIV_BOUNDED_UPWIND;
  const bool divLUL = !lap && !divBU;
  const double J = term_flux(t, f);
  switch (bc.kind) {
    case XCAL_BC_DIRICHLET:  // dirichlet.jl:19-99
      if (lap) {
        double flux = (J * m.area[f]) / m.delta[f];
        double ap = sign * (-flux);
        ap_out = ap;
        bp_out = ap * value;
      } else if (divLUL) {
        double ap = sign * (-J);
        bp_out = ap * value;
      } else {
        double ap = sign * J;
        ap_out = jmax(-ap, 0.0);
        double an = -jmax(-ap, 0.0);
        bp_out = -an * value;
      }
      return;
    case XCAL_BC_NEUMANN:  // neumann.jl:23-63 (scalar)
      if (lap) {
        bp_out = (J * m.area[f]) * value;
      } else {
        double ap = sign * J;
        ap_out = ap;
        bp_out = ((-value) * ap) * m.delta[f];
      }
      return;
    case XCAL_BC_EMPTY:  // empty.jl:46-106: (0, 0) for every operator
      return;
    case XCAL_BC_ZEROGRADIENT:  // zerogradient.jl:25-76
    case XCAL_BC_KWALL:         // Discretise_3_boundary_conditions.jl:11-95
    case XCAL_BC_NUTWALL:
      if (divLUL) ap_out = sign * J;
      else if (divBU) ap_out = sign * J - J;
      return;
    case XCAL_BC_OMEGAWALL:       // Discretise_3_boundary_conditions.jl:27-123
    case XCAL_BC_EXTRAPOLATED: {  // extrapolated.jl:21-72
      if (lap) {
        double flux = ((-J) * m.area[f]) / m.delta[f];
        double ap = sign * flux;
        const double* ph = vec ? (comp == 0 ? px : (comp == 1 ? py : pz)) : px;
        ap_out = ap;
        bp_out = ap * ph[c];
      } else if (divLUL) {
        ap_out = sign * J;
      } else {
        ap_out = sign * J - J;
      }
      return;
    }
    case XCAL_BC_WALL:  // wall.jl:26-120
      if (vec) {
        if (lap) {
          double nx = m.nx[f], ny = m.ny[f], nz = m.nz[f];
          double flux = (J * m.area[f]) / m.delta[f];
          double ap = sign * (-flux);
          double vcn = px[c] * nx + py[c] * ny + pz[c] * nz;
          double vbn = bc.v[0] * nx + bc.v[1] * ny + bc.v[2] * nz;
          double nk = comp == 0 ? nx : (comp == 1 ? ny : nz);
          double vb_p = bc.v[comp] - vbn * nk;
          double vc_n = vcn * nk;
          ap_out = ap;
          bp_out = ap * (vb_p + vc_n);
        }
      } else {
        if (divLUL) {
          ap_out = sign * J;
        } else if (divBU) {
          double ap = sign * J;
          ap_out = jmax(-ap, 0.0);
          double an = -jmax(-ap, 0.0);
          bp_out = -an * value;
        }
      }
      return;
    case XCAL_BC_SYMMETRY:  // symmetry.jl:23-151
      if (vec) {
        double nx = m.nx[f], ny = m.ny[f], nz = m.nz[f];
        double vcn = px[c] * nx + py[c] * ny + pz[c] * nz;
        double nk = comp == 0 ? nx : (comp == 1 ? ny : nz);
        double vck = comp == 0 ? px[c] : (comp == 1 ? py[c] : pz[c]);
        double vn = vcn * nk;
        if (lap) {
          double flux = (J * m.area[f]) / m.delta[f];
          double ap = sign * (-flux);
          ap_out = ap;
          bp_out = ap * (vck - vn);
        } else {
          double ap = sign * J;
          ap_out = ap;
          bp_out = ap * vn;
        }
      } else {
        if (!lap) ap_out = sign * J;
      }
      return;
    default:
      return;
  }
}

// XCD-aware block order: the dispatcher deals workgroups round-robin to the
// 8 XCDs, so block b lands on XCD b%8.  Remapping gives each XCD one
// contiguous range of rows, so the cells it gathers stay in its own L2.
__device__ inline int64_t xcd_block() {
  const int64_t G = gridDim.x, b = blockIdx.x;
  const int64_t q = G >> 3, r = G & 7, x = b & 7, l = b >> 3;
  return x * q + (x < r ? x : r) + l;
}
__device__ inline int64_t xcd_row() { return xcd_block() * blockDim.x + threadIdx.x; }

// Position of this work-item in a level launch of cnt rows (grid =
// level_grid(cnt)): workgroups on XCD x (blockIdx % 8) take eighth x of the
// level, whose rows are one Morton range, so their gathers share that XCD's
// L2.  Valid while the position is below `end`.
__device__ inline int32_t level_pos(int32_t cnt, int32_t& end) {
  const int x = blockIdx.x & 7, l = blockIdx.x >> 3;
  const int64_t c0 = (int64_t)cnt * x / 8 / 64 * 64;
  end = (int32_t)(x == 7 ? cnt : (int64_t)cnt * (x + 1) / 8 / 64 * 64);
  return (int32_t)(c0 + (int64_t)l * blockDim.x + threadIdx.x);
}

// Deterministic block-level sum (wave shuffles + LDS), result in thread 0.
__device__ inline double block_sum(double v, double* lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) s += lds[w];
  }
  __syncthreads();
  return s;
}

}  // namespace xcal
