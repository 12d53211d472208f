// Mesh ingest on the device: the matrix connectivity of a mesh built by
// kernels instead of host loops.  The reference assembles the sparsity
// pattern on the host (sparse_matrix_connectivity + sparsecsr,
// ModelFramework_0_types.jl:139-140, 223-242, flagged there as a temporary
// solution that should be a kernel) and finds each face's matrix entry with
// spindex (:246-258) inside every assembly loop.  Here one pass per row
// builds the CSR pattern, the slot maps that replace spindex and the DILU
// level of every row, equal element for element to host_connectivity.
//
//   k_row_count   entries per row: 1 + distinct neighbours (no sort needed:
//                 an entry counts at its first occurrence in the list)
//   scan          row pointers (hipCUB exclusive sum)
//   k_row_fill    columns by rank among the row's distinct values, the
//                 diagonal slot and each cell-face pair's slot
//   k_trans       slot of each entry's transpose (binary search)
//   k_face_slot   owner-row slot of each internal face (binary search)
//   k_level_round one round of level propagation: a row whose lower (upper)
//                 neighbours all have their levels takes 1 + their maximum.
//                 A level is written once, from final values, so a round that
//                 reads a neighbour still unset merely defers the row.
//   k_level_tail  after 28 rounds, the rows left (meshes whose levels form
//                 long chains, such as lexicographically numbered ones) are
//                 drained by one workgroup from a worklist (Kahn's
//                 topological order), one barrier per level.
#include <hipcub/hipcub.hpp>

#include <chrono>

#include "xcal_internal.h"

namespace xcal {
namespace {

constexpr int kIngestBlock = 256;

inline unsigned grid_of(int64_t n) { return (unsigned)((n + kIngestBlock - 1) / kIngestBlock); }

__device__ inline int64_t gtid() { return blockIdx.x * (int64_t)blockDim.x + threadIdx.x; }

// L[k] is the first occurrence of its value in L[a, k) and not the row itself
__device__ inline bool first_of(const int32_t* __restrict__ L, int32_t a, int32_t k, int32_t c) {
  const int32_t v = L[k];
  if (v == c) return false;
  for (int32_t q = a; q < k; ++q)
    if (L[q] == v) return false;
  return true;
}

// rank of v among the distinct values of {c} U L[a, b)
__device__ inline int32_t rank_of(const int32_t* __restrict__ L, int32_t a, int32_t b, int32_t c,
                                  int32_t v) {
  int32_t r = c < v ? 1 : 0;
  for (int32_t k = a; k < b; ++k)
    if (L[k] < v && first_of(L, a, k, c)) ++r;
  return r;
}

__global__ void k_row_count(int64_t n, const int32_t* __restrict__ cf_ptr,
                            const int32_t* __restrict__ cf_nbr, int32_t* __restrict__ cnt) {
  const int64_t c = gtid();
  if (c > n) return;
  if (c == n) {  // the scan's last element: rowptr[n] = nnz
    cnt[n] = 0;
    return;
  }
  const int32_t a = cf_ptr[c], b = cf_ptr[c + 1];
  int32_t u = 1;
  for (int32_t k = a; k < b; ++k) u += first_of(cf_nbr, a, k, (int32_t)c) ? 1 : 0;
  cnt[c] = u;
}

__global__ void k_row_fill(int64_t n, const int32_t* __restrict__ cf_ptr,
                           const int32_t* __restrict__ cf_nbr, const int32_t* __restrict__ rowptr,
                           int32_t* __restrict__ colval, int32_t* __restrict__ diag,
                           int32_t* __restrict__ cf_slot) {
  const int64_t c = gtid();
  if (c >= n) return;
  const int32_t a = cf_ptr[c], b = cf_ptr[c + 1], base = rowptr[c], ci = (int32_t)c;
  const int32_t rd = base + rank_of(cf_nbr, a, b, ci, ci);
  colval[rd] = ci;
  diag[c] = rd;
  for (int32_t k = a; k < b; ++k) {
    const int32_t v = cf_nbr[k];
    const int32_t s = base + rank_of(cf_nbr, a, b, ci, v);
    cf_slot[k] = s;
    colval[s] = v;  // repeats of a value store the same column
  }
}

// entry of column j in row i, -1 if none (rows ascending)
__device__ inline int32_t find_entry(const int32_t* __restrict__ rowptr,
                                     const int32_t* __restrict__ colval, int32_t i, int32_t j) {
  int32_t lo = rowptr[i], hi = rowptr[i + 1];
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (colval[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  return lo < rowptr[i + 1] && colval[lo] == j ? lo : -1;
}

__global__ void k_trans(int64_t n, const int32_t* __restrict__ rowptr,
                        const int32_t* __restrict__ colval, int32_t* __restrict__ trans) {
  const int64_t c = gtid();
  if (c >= n) return;
  for (int32_t k = rowptr[c]; k < rowptr[c + 1]; ++k)
    trans[k] = find_entry(rowptr, colval, colval[k], (int32_t)c);
}

__global__ void k_face_slot(int64_t nf, int64_t nbf, const int32_t* __restrict__ own,
                            const int32_t* __restrict__ nei, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ colval, int32_t* __restrict__ fslot) {
  const int64_t f = gtid();
  if (f >= nf) return;
  fslot[f] = f < nbf ? -1 : find_entry(rowptr, colval, own[f], nei[f]);
}

__global__ void k_fill_i32(int64_t n, int32_t* __restrict__ p, int32_t v) {
  const int64_t i = gtid();
  if (i < n) p[i] = v;
}

// One round of level propagation (FWD: over lower neighbours, entries before
// the diagonal; otherwise upper).  *pending is raised (once per wavefront)
// while rows are left unset.
template <bool FWD>
__global__ void k_level_round(int64_t n, const int32_t* __restrict__ rowptr,
                              const int32_t* __restrict__ colval, const int32_t* __restrict__ diag,
                              int32_t* lev, int32_t* __restrict__ pending) {
  const int64_t i = gtid();
  bool left = false;
  if (i < n && __atomic_load_n(lev + i, __ATOMIC_RELAXED) < 0) {
    const int32_t k0 = FWD ? rowptr[i] : diag[i] + 1, k1 = FWD ? diag[i] : rowptr[i + 1];
    int32_t m = 0;
    bool ready = true;
    for (int32_t k = k0; k < k1; ++k) {
      const int32_t l = __atomic_load_n(lev + colval[k], __ATOMIC_RELAXED);
      ready = ready && l >= 0;
      m = max(m, l + 1);
    }
    if (ready) __atomic_store_n(lev + i, m, __ATOMIC_RELAXED);
    left = !ready;
  }
  if (__ballot(left) != 0 && (threadIdx.x & 63) == 0) atomicAdd(pending, 1);
}

// Dependencies of row i (FWD: lower neighbours) and dependents (upper).
template <bool FWD>
__device__ inline void dep_range(const int32_t* rowptr, const int32_t* diag, int32_t i, int32_t& a,
                                 int32_t& b) {
  a = FWD ? rowptr[i] : diag[i] + 1;
  b = FWD ? diag[i] : rowptr[i + 1];
}
template <bool FWD>
__device__ inline void succ_range(const int32_t* rowptr, const int32_t* diag, int32_t i,
                                  int32_t& a, int32_t& b) {
  dep_range<!FWD>(rowptr, diag, i, a, b);
}

// Worklist form for the rows the rounds left (long chains of levels): the
// unset dependencies of every unset row ...
template <bool FWD>
__global__ void k_level_indeg(int64_t n, const int32_t* __restrict__ rowptr,
                              const int32_t* __restrict__ colval, const int32_t* __restrict__ diag,
                              const int32_t* __restrict__ lev, int32_t* __restrict__ indeg) {
  const int64_t i = gtid();
  if (i >= n) return;
  int32_t u = 0;
  if (lev[i] < 0) {
    int32_t a, b;
    dep_range<FWD>(rowptr, diag, (int32_t)i, a, b);
    for (int32_t k = a; k < b; ++k) u += lev[colval[k]] < 0 ? 1 : 0;
  }
  indeg[i] = u;
}

// ... the unset rows with none: their levels, and the first worklist ...
template <bool FWD>
__global__ void k_level_seed(int64_t n, const int32_t* __restrict__ rowptr,
                             const int32_t* __restrict__ colval, const int32_t* __restrict__ diag,
                             int32_t* __restrict__ lev, const int32_t* __restrict__ indeg,
                             int32_t* __restrict__ work, int32_t* __restrict__ count) {
  const int64_t i = gtid();
  if (i >= n || lev[i] >= 0 || indeg[i] != 0) return;
  int32_t a, b, m = 0;
  dep_range<FWD>(rowptr, diag, (int32_t)i, a, b);
  for (int32_t k = a; k < b; ++k) m = max(m, lev[colval[k]] + 1);
  lev[i] = m;  // every dependency was set by the rounds
  work[atomicAdd(count, 1)] = (int32_t)i;
}

// ... and one workgroup that drains the worklist: a row whose last unset
// dependency is processed takes its level (all its dependencies are set
// then, by earlier iterations of this workgroup or earlier kernels) and
// joins the next list.  Each row joins once, so the loop ends.
constexpr int kTailBlock = 1024;
template <bool FWD>
__global__ void __launch_bounds__(kTailBlock)
    k_level_tail(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colval,
                 const int32_t* __restrict__ diag, int32_t* lev, int32_t* indeg, int32_t* wa,
                 int32_t* wb, const int32_t* __restrict__ count0) {
  __shared__ int32_t next;
  int32_t cnt = *count0;
  int32_t *cur = wa, *nxt = wb;
  while (cnt > 0) {
    if (threadIdx.x == 0) next = 0;
    __syncthreads();
    for (int32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
      const int32_t i = cur[t];
      int32_t a, b;
      succ_range<FWD>(rowptr, diag, i, a, b);
      for (int32_t k = a; k < b; ++k) {
        const int32_t j = colval[k];
        if (atomicSub(indeg + j, 1) != 1) continue;
        int32_t c0, c1, m = 0;
        dep_range<FWD>(rowptr, diag, j, c0, c1);
        for (int32_t q = c0; q < c1; ++q) m = max(m, lev[colval[q]] + 1);
        lev[j] = m;
        nxt[atomicAdd(&next, 1)] = j;
      }
    }
    __syncthreads();
    cnt = next;
    int32_t* w = cur;
    cur = nxt;
    nxt = w;
    __syncthreads();  // every thread has read `next` before it is cleared
  }
}

template <class T>
void upload_vec(DBuf<T>& d, const std::vector<T>& h, hipStream_t s) {
  d.alloc((int64_t)h.size());
  if (!h.empty())
    XCAL_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

template <class T>
void download_vec(std::vector<T>& h, const DBuf<T>& d, int64_t len, hipStream_t s) {
  h.resize((size_t)len);
  if (len > 0) XCAL_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(T) * len, hipMemcpyDeviceToHost, s));
}

}  // namespace

void device_connectivity(xcal_ctx* ctx, const CellLists& L, Connectivity& K) {
  XCAL_CHECK(ctx, "NULL context");
  XCAL_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = L.n, nf = L.nf, nbf = L.nbf, ncf = (int64_t)L.cf_nbr.size();
  XCAL_CHECK((int64_t)L.cf_ptr.size() == n + 1 && (int64_t)L.own.size() == nf &&
                 (int64_t)L.nei.size() == nf && L.cf_ptr[n] == ncf,
             "internal: cell lists do not match the mesh sizes");
  DBuf<int32_t> cf_ptr, cf_nbr, own, nei, cnt(n + 1), rowptr(n + 1), diag(n), cf_slot(ncf),
      fslot(nf), flev(n), blev(n), pend;
  upload_vec(cf_ptr, L.cf_ptr, s);
  upload_vec(cf_nbr, L.cf_nbr, s);
  upload_vec(own, L.own, s);
  upload_vec(nei, L.nei, s);
  struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  XCAL_HIP(hipEventCreate(&ev.a));
  XCAL_HIP(hipEventCreate(&ev.b));
  hipEvent_t e0 = ev.a, e1 = ev.b;
  XCAL_HIP(hipEventRecord(e0, s));
  hipLaunchKernelGGL(k_row_count, dim3(grid_of(n + 1)), dim3(kIngestBlock), 0, s, n, cf_ptr.p,
                     cf_nbr.p, cnt.p);
  size_t tmp_bytes = 0;
  XCAL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt.p, rowptr.p, (int)(n + 1), s));
  DBuf<uint8_t> tmp((int64_t)tmp_bytes);
  XCAL_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, cnt.p, rowptr.p, (int)(n + 1), s));
  int32_t nnz = 0;
  XCAL_HIP(hipMemcpyAsync(&nnz, rowptr.p + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  XCAL_HIP(hipStreamSynchronize(s));
  XCAL_CHECK(nnz >= n && (int64_t)nnz <= n + ncf, "internal: device row count out of range");
  DBuf<int32_t> colval(nnz), trans(nnz);
  hipLaunchKernelGGL(k_row_fill, dim3(grid_of(n)), dim3(kIngestBlock), 0, s, n, cf_ptr.p, cf_nbr.p,
                     rowptr.p, colval.p, diag.p, cf_slot.p);
  hipLaunchKernelGGL(k_trans, dim3(grid_of(n)), dim3(kIngestBlock), 0, s, n, rowptr.p, colval.p,
                     trans.p);
  hipLaunchKernelGGL(k_face_slot, dim3(grid_of(nf)), dim3(kIngestBlock), 0, s, nf, nbf, own.p,
                     nei.p, rowptr.p, colval.p, fslot.p);
  // levels: rounds in chunks of 4, 8 and 16, the last round's flag read
  // after each (a round over set rows is a cheap no-op); rows still unset
  // after 28 rounds (meshes with long chains of levels) go to the worklist.
  constexpr int kChunks[3] = {4, 8, 16};
  pend.alloc(16);
  DBuf<int32_t> indeg, wa, wb, wcount(1);
  int32_t* pinned = ctx->host_pinned ? reinterpret_cast<int32_t*>(ctx->host_pinned) : nullptr;
  XCAL_CHECK(pinned, "internal: context has no pinned staging");
  auto levels = [&](auto fwd, DBuf<int32_t>& lev) {
    constexpr bool F = decltype(fwd)::value;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_of(n)), dim3(kIngestBlock), 0, s, n, lev.p, -1);
    for (int chunk : kChunks) {
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kIngestBlock), 0, s, (int64_t)chunk, pend.p, 0);
      for (int r = 0; r < chunk; ++r)
        hipLaunchKernelGGL(k_level_round<F>, dim3(grid_of(n)), dim3(kIngestBlock), 0, s, n,
                           rowptr.p, colval.p, diag.p, lev.p, pend.p + r);
      XCAL_HIP(hipMemcpyAsync(pinned, pend.p + chunk - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      XCAL_HIP(hipStreamSynchronize(s));
      if (*pinned == 0) return;
    }
    if (!indeg.p) {
      indeg.alloc(n);
      wa.alloc(n);
      wb.alloc(n);
    }
    hipLaunchKernelGGL(k_level_indeg<F>, dim3(grid_of(n)), dim3(kIngestBlock), 0, s, n, rowptr.p,
                       colval.p, diag.p, lev.p, indeg.p);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kIngestBlock), 0, s, (int64_t)1, wcount.p, 0);
    hipLaunchKernelGGL(k_level_seed<F>, dim3(grid_of(n)), dim3(kIngestBlock), 0, s, n, rowptr.p,
                       colval.p, diag.p, lev.p, indeg.p, wa.p, wcount.p);
    hipLaunchKernelGGL(k_level_tail<F>, dim3(1), dim3(kTailBlock), 0, s, rowptr.p, colval.p,
                       diag.p, lev.p, indeg.p, wa.p, wb.p, wcount.p);
  };
  levels(std::true_type{}, flev);
  levels(std::false_type{}, blev);
  XCAL_HIP(hipEventRecord(e1, s));
  download_vec(K.rowptr, rowptr, n + 1, s);
  download_vec(K.colval, colval, nnz, s);
  download_vec(K.diag, diag, n, s);
  download_vec(K.cf_slot, cf_slot, ncf, s);
  download_vec(K.trans, trans, nnz, s);
  download_vec(K.fslot, fslot, nf, s);
  download_vec(K.flev, flev, n, s);
  download_vec(K.blev, blev, n, s);
  XCAL_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  XCAL_HIP(hipEventElapsedTime(&ms, e0, e1));
  K.device_ms = ms;
  int mf = 0, mb = 0, lo = 0;
  for (int64_t i = 0; i < n; ++i) {
    mf = std::max(mf, K.flev[i]);
    mb = std::max(mb, K.blev[i]);
    lo = std::min(lo, std::min(K.flev[i], K.blev[i]));
  }
  XCAL_CHECK(lo == 0, "internal: a row was left without a DILU level");
  K.nlev_f = mf + 1;
  K.nlev_b = mb + 1;
}

}  // namespace xcal

using namespace xcal;

// Handle of one connectivity build (the ABI's view for tests and timing).
struct xcal_connectivity {
  Connectivity K;
  int64_t ncf = 0;
  double wall_ms = 0.0;
};

extern "C" {

int xcal_connectivity_create(xcal_ctx* ctx, const xcal_mesh_host* mesh, int32_t on_device,
                             xcal_connectivity** out) {
  return guarded([&] {
    XCAL_CHECK(mesh && out, "NULL argument");
    XCAL_CHECK(!on_device || ctx, "a device build needs a context");
    auto g = std::make_unique<xcal_connectivity>();
    CellLists L;
    cell_lists(mesh, L);
    g->ncf = (int64_t)L.cf_nbr.size();
    const auto t0 = std::chrono::steady_clock::now();
    if (on_device) device_connectivity(ctx, L, g->K);
    else host_connectivity(L, g->K);
    g->wall_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = g.release();
    return 0;
  });
}

int xcal_connectivity_info(const xcal_connectivity* g, int64_t* nnz, int64_t* n_cell_faces,
                           int32_t* nlev_f, int32_t* nlev_b, double* wall_ms, double* device_ms) {
  return guarded([&] {
    XCAL_CHECK(g && nnz && n_cell_faces && nlev_f && nlev_b && wall_ms && device_ms,
               "NULL argument");
    *nnz = (int64_t)g->K.colval.size();
    *n_cell_faces = g->ncf;
    *nlev_f = g->K.nlev_f;
    *nlev_b = g->K.nlev_b;
    *wall_ms = g->wall_ms;
    *device_ms = g->K.device_ms;
    return 0;
  });
}

int xcal_connectivity_get(const xcal_connectivity* g, int32_t which, int32_t* host) {
  return guarded([&] {
    XCAL_CHECK(g && host, "NULL argument");
    const Connectivity& K = g->K;
    const std::vector<int32_t>* v[] = {&K.rowptr, &K.colval, &K.diag,  &K.cf_slot,
                                       &K.trans,  &K.fslot,  &K.flev, &K.blev};
    XCAL_CHECK(which >= 0 && which < 8, "which: 0..7");
    std::copy(v[which]->begin(), v[which]->end(), host);
    return 0;
  });
}

int xcal_connectivity_destroy(xcal_connectivity* g) {
  return guarded([&] {
    delete g;
    return 0;
  });
}

}  // extern "C"
