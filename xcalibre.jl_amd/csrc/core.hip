// Context, mesh upload and fields of libxcalhip.so.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>

#include "xcal_internal.h"

namespace xcal {

thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

template <class T>
static void upload(DBuf<T>& d, const std::vector<T>& h, int64_t& bytes) {
  d.alloc((int64_t)h.size());
  if (!h.empty()) XCAL_HIP(hipMemcpy(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  bytes += d.bytes();
}

TermsK make_terms(const xcal_term* terms, int n) {
  XCAL_CHECK(n >= 1 && n <= kMaxTerms, "number of operator terms must be 1..4");
  TermsK t{};
  t.n = n;
  for (int i = 0; i < n; ++i) {
    XCAL_CHECK((terms[i].type >= XCAL_TIME_STEADY && terms[i].type <= XCAL_SI) ||
                   terms[i].type == XCAL_TIME_CRANK_NICOLSON,
               "unknown operator type");
    XCAL_CHECK(terms[i].sign == 1 || terms[i].sign == -1, "operator sign must be +1 or -1");
    // Crank-Nicolson's coefficients are Euler's (Discretise_1_schemes.jl:44-62)
    const bool cn = terms[i].type == XCAL_TIME_CRANK_NICOLSON;
    if (i == 0) t.cn_first = cn;
    t.t[i].type = cn ? XCAL_TIME_EULER : terms[i].type;
    t.t[i].sign = terms[i].sign;
    t.t[i].flux = terms[i].flux ? terms[i].flux->comp(0) : nullptr;
    t.t[i].flux_const = terms[i].flux_const;
  }
  return t;
}

BCsK make_bcs(const xcal_mesh* m, const xcal_bc* bcs, int n) {
  XCAL_CHECK(m->n_patches <= kMaxPatches, "mesh has more than 32 patches");
  BCsK b{};
  b.npatch = m->n_patches;
  for (int p = 0; p < kMaxPatches; ++p) b.p[p].kind = -1;
  for (int i = 0; i < n; ++i) {
    XCAL_CHECK(bcs[i].patch >= 0 && bcs[i].patch < m->n_patches, "BC patch index out of range");
    XCAL_CHECK(bcs[i].kind >= XCAL_BC_DIRICHLET && bcs[i].kind <= XCAL_BC_EMPTY,
               "unknown boundary condition kind");
    // the reference applies the first BC whose range holds the face
    if (b.p[bcs[i].patch].kind != -1) continue;
    b.p[bcs[i].patch].kind = bcs[i].kind;
    for (int k = 0; k < 3; ++k) b.p[bcs[i].patch].v[k] = bcs[i].value[k];
  }
  return b;
}

xcal_field* check_field(const xcal_field* f, const xcal_mesh* m, int kind, const char* what) {
  XCAL_CHECK(f != nullptr, std::string(what) + ": field is NULL");
  XCAL_CHECK(f->mesh == m, std::string(what) + ": field belongs to another mesh");
  XCAL_CHECK(f->kind == kind, std::string(what) + ": wrong field kind");
  return const_cast<xcal_field*>(f);
}

__global__ void k_fill(double* p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Launch plans of the DILU sweeps.  A level with at least `small` rows gets a
// launch of its own over many workgroups; a run of consecutive smaller levels
// is done by one workgroup that steps through them with barriers, which costs
// far less than a launch per level.  Each launch owns a range of
// partial-sum slots (one per workgroup) for the dot products that the
// backward sweep folds in.
static std::vector<SweepStep> plan_levels(const std::vector<int32_t>& ptr, int nlev, int small) {
  std::vector<SweepStep> plan;
  int32_t slot = 0;
  for (int l = 0; l < nlev;) {
    const int32_t cnt = ptr[l + 1] - ptr[l];
    SweepStep st{};
    st.lev0 = l;
    st.row0 = ptr[l];
    if (cnt >= small) {
      st.lev1 = l + 1;
      st.grid = level_grid(cnt);
    } else {
      int e = l;
      while (e < nlev && e - l < kPipeLev && ptr[e + 1] - ptr[e] < small) ++e;
      st.lev1 = e;
      st.grid = 1;
    }
    st.row1 = ptr[st.lev1];
    st.slot = slot;
    slot += st.grid;
    plan.push_back(st);
    l = st.lev1;
  }
  return plan;
}

// Row blocks of kBlock rows for the gathering kernels: each forward level
// [a, b) of at least 8 kBlock rows is cut at 64-aligned eighths, and eighth x
// of every such level goes to XCD x (the dispatcher deals block b to XCD
// b%8).  A maximal stretch of consecutive smaller levels is cut the same way
// as one range: neighbouring small levels are neighbouring bands of space.
// Shorter XCD lists are padded with empty blocks.
std::vector<int32_t> xcd_row_blocks(const std::vector<int32_t>& lev_ptr, int nlev) {
  std::vector<std::vector<int32_t>> per(8);
  std::vector<std::pair<int64_t, int64_t>> ranges;
  for (int l = 0; l < nlev;) {
    int e = l + 1;
    if (lev_ptr[l + 1] - lev_ptr[l] < 8 * kBlock)
      while (e < nlev && lev_ptr[e + 1] - lev_ptr[e] < 8 * kBlock) ++e;
    ranges.push_back({lev_ptr[l], lev_ptr[e]});
    l = e;
  }
  for (const auto& rg : ranges) {
    const int64_t a = rg.first, cnt = rg.second - rg.first;
    for (int x = 0; x < 8; ++x) {
      const int64_t c0 = a + (cnt * x / 8) / 64 * 64;
      const int64_t c1 = x == 7 ? a + cnt : a + (cnt * (x + 1) / 8) / 64 * 64;
      for (int64_t r = c0; r < c1; r += kBlock) {
        per[x].push_back((int32_t)r);
        per[x].push_back((int32_t)std::min<int64_t>(r + kBlock, c1));
      }
    }
  }
  size_t nb = 1;
  for (const auto& v : per) nb = std::max(nb, v.size() / 2);
  std::vector<int32_t> out(16 * nb, 0);
  for (int x = 0; x < 8; ++x)
    for (size_t k = 0; k < per[x].size() / 2; ++k) {
      out[2 * (8 * k + x)] = per[x][2 * k];
      out[2 * (8 * k + x) + 1] = per[x][2 * k + 1];
    }
  return out;
}

void make_sweep_plans(xcal_mesh* m, int small) {
  if (small <= 0) {
    small = kSmallLevel;
    if (const char* env = std::getenv("XCAL_SMALL_LEVEL")) small = std::max(1, std::atoi(env));
  }
  m->small_level = small;
  m->plan_f = plan_levels(m->lev_ptr_f, m->nlev_f, small);
  m->plan_b = plan_levels(m->lev_ptr_b, m->nlev_b, small);
  m->sweep_slots = 0;
  for (const auto& st : m->plan_b) m->sweep_slots = std::max(m->sweep_slots, st.slot + st.grid);
  if (m->cone) m->sweep_slots = m->cone_b.nblk();  // the cone form: one partial per block
  int64_t bytes = 0;
  if (m->lev_ptr_f_d.n == 0) {
    upload(m->lev_ptr_f_d, m->lev_ptr_f, bytes);
    upload(m->lev_ptr_b_d, m->lev_ptr_b, bytes);
    m->xblk = xcd_row_blocks(m->lev_ptr_f, m->nlev_f);
    m->nxblk = (int)(m->xblk.size() / 2);
    upload(m->xblk_d, m->xblk, bytes);
    m->device_bytes += bytes;
  }
}

void fill(hipStream_t s, double* p, int64_t n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n, v);
}

// ---------------------------------------------------------------- meshes
// Host-side build of one case in device numbering: validation, CSR pattern,
// Julia's dot and norm of 3-vectors (StaticArrays): (a1 b1 + a2 b2) + a3 b3,
// and the square root of the same sum of squares
static inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double norm3(const double* a) { return std::sqrt(dot3(a, a)); }

// Cone form of one DILU sweep (host side; see sweep_cone.h).  The sweep's
// levels are cut into bands of consecutive levels, one launch each (one band
// holding every level where the cones fit, as on randomly numbered meshes).
// The rows of a band, in Morton order, are cut into blocks of S rows.  A
// block's cone is its rows and every row of the band they depend on,
// directly or through other rows; the rows of earlier bands that the cone
// reads (its halo) come first among its positions and are loaded from the
// sweep's output in global memory, where earlier launches stored them.
// Cone positions of block b are [base[b], base[b+1]): the halo, then the
// cone rows ordered by level and, inside a level, by entry count
// (descending), cut into steps of at most kConeBlock rows.  Entry t of the
// q-th row of a step is entry eoff[t] + q, present for q < ecnt[t]
// (jagged-diagonal storage: no padding).
struct ConeHost {
  int wk = 0;                      // entry bound (2, 4, 6 or 8) of the cone kernels
  std::vector<int32_t> blk, base;  // per block: first step, first position (nblk + 1)
  std::vector<int32_t> steps;      // per step: r0, cnt, eoff[8], ecnt[8] (kConeTab ints)
  std::vector<int32_t> crow;       // per position: device row (own) or ~row (another block's)
  std::vector<uint16_t> lidx;      // per entry: cone-local position of the dependency
  std::vector<int32_t> slot, dep;  // per entry: matrix slot, device row of the dependency
  std::vector<int32_t> own, ownp;  // block b owns rows own[ownp[b], ownp[b+1])
  std::vector<int32_t> band;       // first block of each band (nbands + 1)
  std::vector<int32_t> hnum;       // per block: halo positions (rows of earlier bands)
};

namespace {
// Work arrays of a cone build; `mark` holds build stamps, so a failed band
// leaves nothing a later one could mistake for its own.
struct ConeScratch {
  std::vector<int32_t> owner, mark, loc, stack, cone, halo;
  int32_t stamp = 0;
  explicit ConeScratch(int64_t n) : owner(n, -1), mark(n, -1), loc(n, 0) {}
};
}  // namespace

// Appends to C the blocks of the band of levels [lo, hi); false (C
// unchanged) when a block's cone or step count exceeds the kernels' limits,
// or (ratio > 0) the band's cone positions so far exceed ratio times its
// rows so far by more than one block's limit (checked after every block, so
// a band that will not fit stops early).
static bool build_band(int S, double ratio, bool lower, const std::vector<int32_t>& ord,
                       const std::vector<int32_t>& lev, int lo, int hi,
                       const std::vector<int32_t>& rbeg, const std::vector<int32_t>& rend,
                       const std::vector<int32_t>& diag, const std::vector<int32_t>& colval,
                       ConeHost& C, ConeScratch& W) {
  auto span = [&](int32_t c, int32_t& k0, int32_t& k1) {
    k0 = lower ? rbeg[c] : diag[c] + kSlice;
    k1 = lower ? diag[c] : rend[c];
  };
  auto nent = [&](int32_t c) {
    int32_t k0, k1;
    span(c, k0, k1);
    return (k1 - k0) / kSlice;
  };
  std::vector<int32_t> rows;  // the band's rows in Morton order
  for (int32_t c : ord)
    if (lev[c] >= lo && lev[c] < hi) rows.push_back(c);
  const int64_t nb = ((int64_t)rows.size() + S - 1) / S;
  const int32_t b0 = (int32_t)C.blk.size() - 1;  // global index of the band's first block
  for (size_t t = 0; t < rows.size(); ++t) W.owner[rows[t]] = b0 + (int32_t)(t / S);
  ConeHost B;  // the band's blocks, positions and entries from 0
  B.blk.push_back(0);
  B.base.push_back(0);
  B.ownp.push_back(0);
  int maxw = 0;
  std::vector<int32_t>&cone = W.cone, &halo = W.halo, &stack = W.stack, &mark = W.mark,
  &loc = W.loc;
  for (int64_t b = 0; b < nb; ++b) {
    const int32_t stamp = ++W.stamp;
    cone.clear();
    halo.clear();
    const size_t a0 = (size_t)b * S, a1 = std::min(rows.size(), a0 + S);
    for (size_t t = a0; t < a1; ++t) {
      mark[rows[t]] = stamp;
      cone.push_back(rows[t]);
      stack.push_back(rows[t]);
      B.own.push_back(rows[t]);
    }
    B.ownp.push_back((int32_t)B.own.size());
    while (!stack.empty()) {
      const int32_t c = stack.back();
      stack.pop_back();
      int32_t k0, k1;
      span(c, k0, k1);
      for (int32_t k = k0; k < k1; k += kSlice) {
        const int32_t d = colval[k];
        if (mark[d] == stamp) continue;
        mark[d] = stamp;
        if (lev[d] < lo) {  // an earlier band's row: read, not swept
          halo.push_back(d);
        } else {
          cone.push_back(d);
          stack.push_back(d);
        }
        if ((int64_t)(cone.size() + halo.size()) > kConeMax) {
          stack.clear();
          return false;
        }
      }
    }
    std::sort(halo.begin(), halo.end());
    std::sort(cone.begin(), cone.end(), [&](int32_t x, int32_t y) {
      if (lev[x] != lev[y]) return lev[x] < lev[y];
      const int ex = nent(x), ey = nent(y);
      if (ex != ey) return ex > ey;
      return x < y;
    });
    const int32_t base = B.base.back();
    const int32_t nh = (int32_t)halo.size();
    for (size_t j = 0; j < halo.size(); ++j) loc[halo[j]] = (int32_t)j;
    for (size_t j = 0; j < cone.size(); ++j) loc[cone[j]] = nh + (int32_t)j;
    int nsteps = 0;
    for (size_t j0 = 0; j0 < cone.size();) {
      size_t j1 = j0;
      while (j1 < cone.size() && lev[cone[j1]] == lev[cone[j0]] && j1 - j0 < (size_t)kConeBlock) ++j1;
      int32_t tab[kConeTab] = {};  // r0, cnt, eoff[8], ecnt[8], 2 unused
      tab[0] = base + nh + (int32_t)j0;
      tab[1] = (int32_t)(j1 - j0);
      for (int t = 0; t < 8; ++t) {
        tab[2 + t] = (int32_t)B.lidx.size();
        for (size_t j = j0; j < j1 && nent(cone[j]) > t; ++j) {
          int32_t k0, k1;
          span(cone[j], k0, k1);
          const int32_t k = k0 + t * kSlice, d = colval[k];
          XCAL_CHECK(loc[d] < nh + (int32_t)j0, "internal: cone dependency after its row");
          B.lidx.push_back((uint16_t)loc[d]);
          B.slot.push_back(k);
          B.dep.push_back(d);
          tab[10 + t]++;
        }
      }
      for (size_t j = j0; j < j1; ++j) maxw = std::max(maxw, nent(cone[j]));
      if (maxw > 8) return false;
      B.steps.insert(B.steps.end(), tab, tab + kConeTab);
      ++nsteps;
      j0 = j1;
    }
    if (nsteps > kConeSteps) return false;
    for (int32_t c : halo) B.crow.push_back(~c);
    B.hnum.push_back(nh);
    const int32_t gb = b0 + (int32_t)b;
    for (int32_t c : cone) B.crow.push_back(W.owner[c] == gb ? c : ~c);
    B.base.push_back(base + (int32_t)(halo.size() + cone.size()));
    B.blk.push_back(B.blk.back() + nsteps);
    if (ratio > 0 && B.base.back() > ratio * (double)B.own.size() + kConeMax) return false;
  }
  // append the band: positions, entry offsets, steps and own rows move up
  const int32_t p0 = C.base.back(), e0 = (int32_t)C.lidx.size(), s0 = C.blk.back(),
                o0 = C.ownp.back();
  for (size_t t = 0; t < B.steps.size(); t += kConeTab) {
    B.steps[t] += p0;
    for (int k = 0; k < 8; ++k) B.steps[t + 2 + k] += e0;
  }
  for (size_t b = 1; b < B.blk.size(); ++b) {
    C.blk.push_back(B.blk[b] + s0);
    C.base.push_back(B.base[b] + p0);
    C.ownp.push_back(B.ownp[b] + o0);
  }
  C.steps.insert(C.steps.end(), B.steps.begin(), B.steps.end());
  C.crow.insert(C.crow.end(), B.crow.begin(), B.crow.end());
  C.lidx.insert(C.lidx.end(), B.lidx.begin(), B.lidx.end());
  C.slot.insert(C.slot.end(), B.slot.begin(), B.slot.end());
  C.dep.insert(C.dep.end(), B.dep.begin(), B.dep.end());
  C.own.insert(C.own.end(), B.own.begin(), B.own.end());
  C.hnum.insert(C.hnum.end(), B.hnum.begin(), B.hnum.end());
  C.wk = std::max(C.wk, maxw);
  return true;
}

// Builds the cone form of the forward (lower) or backward (upper) sweep with
// blocks of S rows over nlev levels.  First every level as one band; where
// those cones do not fit, bands of consecutive levels, each launch then
// covering only the band's rows: a band's block size is the power of two
// (kConeMinBlock..S) that gives about kConeBandBlocks blocks, so a launch
// still fills the chip, doubled while its cone positions exceed
// kConeBandRatio times its rows (larger blocks recompute fewer rows).  A
// band starts at twice the previous band's width (the levels left, for the
// first) and halves when no block size fits.  False
// when a band of kConeMinBand levels does not fit, the sweep needs more than
// kConeMaxBands bands, or the cones hold more than kConeMaxRatio n
// positions in all.
static bool build_cone(int64_t n, int S, bool lower, const std::vector<int32_t>& ord,
                       const std::vector<int32_t>& lev, int nlev, const std::vector<int32_t>& rbeg,
                       const std::vector<int32_t>& rend, const std::vector<int32_t>& diag,
                       const std::vector<int32_t>& colval, ConeHost& C) {
  C = ConeHost();
  C.blk.push_back(0);
  C.base.push_back(0);
  C.ownp.push_back(0);
  C.band.push_back(0);
  ConeScratch W(n);
  std::vector<int64_t> lcnt(nlev + 1, 0);  // rows below each level
  for (int64_t c = 0; c < n; ++c) lcnt[lev[c] + 1]++;
  for (int l = 0; l < nlev; ++l) lcnt[l + 1] += lcnt[l];
  double bratio = kConeBandRatio;  // XCAL_CONE_BAND_RATIO at creation overrides it
  if (const char* env = std::getenv("XCAL_CONE_BAND_RATIO")) bratio = std::max(1.0, std::atof(env));
  const double maxratio = std::max(kConeMaxRatio, bratio + 0.5);
  int lo = 0, H = nlev;
  if (build_band(S, 0, lower, ord, lev, 0, nlev, rbeg, rend, diag, colval, C, W)) {
    C.band.push_back((int32_t)C.blk.size() - 1);
    lo = nlev;
    if ((double)C.crow.size() > kConeMaxRatio * (double)n) return false;
  }
  while (lo < nlev) {
    H = std::min(H, nlev - lo);
    for (;;) {
      const int64_t rows = lcnt[lo + H] - lcnt[lo];
      int Sb = kConeMinBlock;
      while (Sb < S && rows / (2 * Sb) >= kConeBandBlocks) Sb *= 2;
      bool ok = false;
      for (; Sb <= S && !ok; Sb *= 2)
        ok = build_band(Sb, bratio, lower, ord, lev, lo, lo + H, rbeg, rend, diag, colval, C, W);
      if (ok) break;
      if (H <= kConeMinBand) return false;
      H = std::max(kConeMinBand, H / 2);
    }
    C.band.push_back((int32_t)C.blk.size() - 1);
    if ((int)C.band.size() - 1 > kConeMaxBands) return false;
    if ((double)C.crow.size() > maxratio * (double)n) return false;
    lo += H;
    H *= 2;
  }
  if (C.wk == 0 || C.lidx.empty()) return false;  // masked loads read entry 0
  C.wk = C.wk <= 2 ? 2 : C.wk <= 4 ? 4 : C.wk <= 6 ? 6 : 8;
  return true;
}

// DILU levels, renumbering, sliced-ELL slots and packed factors.
struct HostMesh {
  bool cone = false;   // the cone form of both sweeps was built
  int cone_rows = 0;   // its block size
  ConeHost cone_f, cone_b;
  int dim = 3, n_patches = 0, nlev_f = 0, nlev_b = 0;
  int64_t n = 0, nf = 0, nbf = 0, nnz = 0, ncf = 0, nzlen = 0, nbcell = 0;
  std::vector<int64_t> patch_first, patch_last;
  std::vector<double> vol, area, weight, delta, nx, ny, nz, ex, ey, ez, sfx, sfy, sfz, lapEf, rix,
      riy, riz, ncw, ntx, nty, ntz;
  std::vector<int32_t> cf_ptr, cf_face, cf_nbr, cf_slot, diag, own, nei, fslot, bc_cells, bc_ptr,
      bc_face, rbeg, rend, sinfo, colval, trans, rows_b, posb, Lcol, Lslot, Ucol, Uslot, lev_ptr_f,
      lev_ptr_b, lmeta_f, lmeta_b, cperm, fperm, nzperm, h_rowptr, h_colval;
  std::vector<int8_t> cf_sign, bpatch;
};

// Row i of the matrix is {i} U neighbours(i), ascending and unique
// (sparse_matrix_connectivity + sparsecsr, ModelFramework_0_types.jl:223-242).
// The slot maps replace spindex (ModelFramework_0_types.jl:246-258): the
// entry of each cell-face pair, of each internal face's owner row, of each
// diagonal and of each entry's transpose.  DILU level schedules: forward
// level(i) = 1 + max level of its lower neighbours, backward the same over
// upper neighbours.  Rows inside a level are independent; level order
// preserves every i<j dependency of the sequential sweeps.
void host_connectivity(const CellLists& L, Connectivity& K) {
  const int64_t n = L.n, nf = L.nf, nbf = L.nbf;
  const std::vector<int32_t>&cf_ptr = L.cf_ptr, &cf_nbr = L.cf_nbr, &own = L.own, &nei = L.nei;
  std::vector<int32_t>&rowptr = K.rowptr, &colval = K.colval;
  rowptr.assign(n + 1, 0);
  colval.clear();
  colval.reserve(n + cf_nbr.size());
  std::vector<int32_t> row;
  for (int64_t c = 0; c < n; ++c) {
    row.clear();
    row.push_back((int32_t)c);
    for (int32_t k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) row.push_back(cf_nbr[k]);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    colval.insert(colval.end(), row.begin(), row.end());
    rowptr[c + 1] = (int32_t)colval.size();
  }
  const int64_t nnz = (int64_t)colval.size();
  auto find = [&](int64_t i, int32_t j) -> int32_t {
    const int32_t* b = colval.data() + rowptr[i];
    const int32_t* e = colval.data() + rowptr[i + 1];
    const int32_t* it = std::lower_bound(b, e, j);
    return (it != e && *it == j) ? (int32_t)(it - colval.data()) : -1;
  };
  K.diag.assign(n, 0);
  K.cf_slot.assign(cf_nbr.size(), 0);
  K.trans.assign(nnz, 0);
  for (int64_t c = 0; c < n; ++c) {
    K.diag[c] = find(c, (int32_t)c);
    for (int32_t k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) K.cf_slot[k] = find(c, cf_nbr[k]);
    for (int32_t k = rowptr[c]; k < rowptr[c + 1]; ++k) K.trans[k] = find(colval[k], (int32_t)c);
  }
  K.fslot.assign(nf, -1);
  for (int64_t f = nbf; f < nf; ++f) K.fslot[f] = find(own[f], nei[f]);
  K.flev.assign(n, 0);
  K.blev.assign(n, 0);
  int maxl = 0;
  for (int64_t i = 0; i < n; ++i) {
    int l = 0;
    for (int32_t k = rowptr[i]; k < K.diag[i]; ++k) l = std::max(l, K.flev[colval[k]] + 1);
    K.flev[i] = l;
    maxl = std::max(maxl, l);
  }
  K.nlev_f = maxl + 1;
  maxl = 0;
  for (int64_t i = n - 1; i >= 0; --i) {
    int l = 0;
    for (int32_t k = K.diag[i] + 1; k < rowptr[i + 1]; ++k) l = std::max(l, K.blev[colval[k]] + 1);
    K.blev[i] = l;
    maxl = std::max(maxl, l);
  }
  K.nlev_b = maxl + 1;
}

void cell_lists(const xcal_mesh_host* h, CellLists& L) {
  XCAL_CHECK(h->dim == 2 || h->dim == 3, "dim must be 2 or 3");
  XCAL_CHECK(h->n_cells > 0 && h->n_faces > 0, "empty mesh");
  XCAL_CHECK(h->n_cells < (1ll << 30) && h->n_faces < (1ll << 30), "mesh too large for int32 indices");
  XCAL_CHECK(h->n_bfaces >= 0 && h->n_bfaces <= h->n_faces, "bad n_bfaces");
  XCAL_CHECK(h->n_cell_faces >= 0 && h->n_cell_faces < (1ll << 31), "bad n_cell_faces");
  const int64_t n = h->n_cells, nf = h->n_faces, nbf = h->n_bfaces;
  L.n = n;
  L.nf = nf;
  L.nbf = nbf;
  // cell -> internal face lists in the reference's faces_range order
  L.cf_ptr.assign(n + 1, 0);
  L.cf_face.clear();
  L.cf_nbr.clear();
  L.cf_sign.clear();
  L.cf_face.reserve(h->n_cell_faces);
  L.cf_nbr.reserve(h->n_cell_faces);
  L.cf_sign.reserve(h->n_cell_faces);
  for (int64_t c = 0; c < n; ++c) {
    int64_t a = h->cell_faces_range[2 * c], b = h->cell_faces_range[2 * c + 1];
    for (int64_t k = a; k <= b; ++k) {
      XCAL_CHECK(k >= 1 && k <= h->n_cell_faces, "faces_range outside cell_faces");
      int64_t f = h->cell_faces[k - 1] - 1, nb = h->cell_neighbours[k - 1] - 1;
      XCAL_CHECK(f >= nbf && f < nf, "cell face is not an internal face");
      XCAL_CHECK(nb >= 0 && nb < n, "neighbour index out of range");
      XCAL_CHECK(h->cell_nsign[k - 1] == 1 || h->cell_nsign[k - 1] == -1, "nsign must be +-1");
      L.cf_face.push_back((int32_t)f);
      L.cf_nbr.push_back((int32_t)nb);
      L.cf_sign.push_back((int8_t)h->cell_nsign[k - 1]);
    }
    XCAL_CHECK(L.cf_face.size() < (size_t)(1ll << 31), "cell-face lists exceed int32 indices");
    L.cf_ptr[c + 1] = (int32_t)L.cf_face.size();
  }
  L.own.assign(nf, 0);
  L.nei.assign(nf, 0);
  for (int64_t f = 0; f < nf; ++f) {
    L.own[f] = (int32_t)(h->face_owner[2 * f] - 1);
    L.nei[f] = (int32_t)(h->face_owner[2 * f + 1] - 1);
    XCAL_CHECK(L.own[f] >= 0 && L.own[f] < n && L.nei[f] >= 0 && L.nei[f] < n,
               "ownerCells out of range");
  }
}

static void build_host_mesh(const xcal_mesh_host* h, HostMesh& M, bool cones, xcal_ctx* ingest) {
  CellLists L;
  cell_lists(h, L);
  M.dim = h->dim;
  const int64_t n = L.n, nf = L.nf, nbf = L.nbf;
  M.n = n;
  M.nf = nf;
  M.nbf = nbf;
  M.n_patches = h->n_patches;
  XCAL_CHECK(h->n_patches <= kMaxPatches, "mesh has more than 32 patches");
  for (int p = 0; p < h->n_patches; ++p) {
    int64_t a = h->patch_range[2 * p] - 1, b = h->patch_range[2 * p + 1] - 1;
    XCAL_CHECK(a >= 0 && b < nbf && a <= b + 1, "patch range outside boundary faces");
    M.patch_first.push_back(a);
    M.patch_last.push_back(b);
  }
  std::vector<int32_t>&cf_ptr = L.cf_ptr, &cf_face = L.cf_face, &cf_nbr = L.cf_nbr, &own = L.own,
                      &nei = L.nei;
  std::vector<int8_t>& cf_sign = L.cf_sign;
  M.ncf = (int64_t)cf_face.size();
  // CSR pattern, slot maps and DILU levels, on the device when a context is
  // given (the arrays are equal either way)
  Connectivity K;
  if (ingest) device_connectivity(ingest, L, K);
  else host_connectivity(L, K);
  std::vector<int32_t>&rowptr = K.rowptr, &colval = K.colval, &diag = K.diag, &cf_slot = K.cf_slot,
                      &trans = K.trans, &fslot = K.fslot, &flev = K.flev, &blev = K.blev;
  const int64_t nnz = (int64_t)colval.size();
  M.nnz = nnz;
  M.nlev_f = K.nlev_f;
  M.nlev_b = K.nlev_b;
  // faces
  std::vector<double> area(h->face_area, h->face_area + nf), weight(h->face_weight, h->face_weight + nf),
      delta(h->face_delta, h->face_delta + nf), nx(nf), ny(nf), nz(nf), ex(nf), ey(nf), ez(nf),
      sfx(nf), sfy(nf), sfz(nf), lapEf(nf), rix(nf, 0.0), riy(nf, 0.0), riz(nf, 0.0),
      ncw(nf, 0.0), ntx(nf, 0.0), nty(nf, 0.0), ntz(nf, 0.0);
  for (int64_t f = 0; f < nf; ++f) {
    nx[f] = h->face_normal[3 * f];
    ny[f] = h->face_normal[3 * f + 1];
    nz[f] = h->face_normal[3 * f + 2];
    ex[f] = h->face_e[3 * f];
    ey[f] = h->face_e[3 * f + 1];
    ez[f] = h->face_e[3 * f + 2];
    // Sf = area*normal (flux!, green_gauss!), same rounding as the reference
    sfx[f] = area[f] * nx[f];
    sfy[f] = area[f] * ny[f];
    sfz[f] = area[f] * nz[f];
    // Laplacian{Linear} |E_f| (Discretise_1_schemes.jl:74-88); independent
    // of the visiting side's nsign because every negation there is exact
    double s0 = area[f] * nx[f], s1 = area[f] * ny[f], s2 = area[f] * nz[f];
    double cfac = (s0 * s0 + s1 * s1 + s2 * s2) / (s0 * ex[f] + s1 * ey[f] + s2 * ez[f]);
    double e0 = cfac * ex[f], e1 = cfac * ey[f], e2 = cfac * ez[f];
    lapEf[f] = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    // correct_interpolation! (Calculate_0_gradient.jl:224-243): Ri = rf -
    // 0.5 (rP + rN) of an internal face
    if (f >= nbf) {
      const double* rf = h->face_centre + 3 * f;
      const double* rP = h->cell_centre + 3 * (int64_t)own[f];
      const double* rN = h->cell_centre + 3 * (int64_t)nei[f];
      rix[f] = rf[0] - 0.5 * (rP[0] + rN[0]);
      riy[f] = rf[1] - 0.5 * (rP[1] + rN[1]);
      riz[f] = rf[2] - 0.5 * (rP[2] + rN[2]);
      // nonorthogonal_face_correction (Solvers_1_SIMPLE.jl:285-336): the
      // correction weight and T_hat = Sf - Ef, Ef = dPN |n|^2 / (dPN.n) A
      const double* n3 = h->face_normal + 3 * f;
      double c1f[3], c1c2[3], fp[3], d2[3];
      for (int a = 0; a < 3; ++a) {
        c1f[a] = rf[a] - rP[a];
        c1c2[a] = rN[a] - rP[a];
      }
      const double q = dot3(c1f, c1c2) / dot3(c1c2, c1c2);
      for (int a = 0; a < 3; ++a) {
        fp[a] = rP[a] - q * (rP[a] - rN[a]);
        d2[a] = rN[a] - fp[a];
      }
      ncw[f] = norm3(d2) / norm3(c1c2);
      const double nn = norm3(n3);
      const double fac = (nn * nn) / dot3(c1c2, n3);
      ntx[f] = area[f] * n3[0] - (c1c2[0] * fac) * area[f];
      nty[f] = area[f] * n3[1] - (c1c2[1] * fac) * area[f];
      ntz[f] = area[f] * n3[2] - (c1c2[2] * fac) * area[f];
    }
  }
  // boundary faces grouped per cell, ascending face id
  std::vector<int8_t> bpatch(nbf, -1);
  for (int p = 0; p < h->n_patches; ++p)
    for (int64_t f = M.patch_first[p]; f <= M.patch_last[p]; ++f)
      if (bpatch[f] < 0) bpatch[f] = (int8_t)p;
  std::vector<int32_t> bcount(n, 0);
  for (int64_t f = 0; f < nbf; ++f) {
    int64_t c = h->boundary_cellsID[f] - 1;
    XCAL_CHECK(c >= 0 && c < n, "boundary_cellsID out of range");
    bcount[c]++;
  }
  std::vector<int32_t> bc_cells, bc_ptr(1, 0), bc_face;
  std::vector<int32_t> bstart(n, -1);
  for (int64_t c = 0; c < n; ++c)
    if (bcount[c] > 0) {
      bstart[c] = (int32_t)bc_cells.size();
      bc_cells.push_back((int32_t)c);
      bc_ptr.push_back(bc_ptr.back() + bcount[c]);
    }
  bc_face.assign(nbf, 0);
  {
    std::vector<int32_t> cur(bc_cells.size(), 0);
    for (int64_t f = 0; f < nbf; ++f) {
      int64_t c = h->boundary_cellsID[f] - 1;
      int32_t s = bstart[c];
      bc_face[bc_ptr[s] + cur[s]++] = (int32_t)f;
    }
  }
  M.nbcell = (int64_t)bc_cells.size();
  std::vector<double> vol(h->cell_volume, h->cell_volume + n);

  // ---- device numbering ------------------------------------------------
  // Cells are stored level-major: by forward DILU level, and in Morton
  // order of their centres inside a level.  Each forward level is then one
  // contiguous row range, and rows that are neighbours in a level are
  // neighbours in space, so neighbour gathers hit nearby cache lines.
  // Internal faces follow their renumbered cells.  Every matrix row keeps
  // its entries in the reference's (original) column order, so row sums,
  // the DILU lower/upper split and all per-row summation orders are
  // unchanged; fields are permuted only when they cross the C ABI.
  std::vector<int32_t> cperm(n), cinv(n), fperm(nf), finv(nf);
  std::vector<uint64_t> dmort(n);  // Morton key of each device row
  {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t c = 0; c < n; ++c)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], h->cell_centre[3 * c + a]);
        hi[a] = std::max(hi[a], h->cell_centre[3 * c + a]);
      }
    const int naxis = h->dim == 3 ? 3 : 2;
    const int bits = naxis == 3 ? 21 : 31;
    auto spread = [&](uint64_t v) {  // interleave with naxis-1 zero bits
      uint64_t r = 0;
      for (int b = 0; b < bits; ++b) r |= ((v >> b) & 1ull) << (naxis * b);
      return r;
    };
    std::vector<std::pair<std::pair<int32_t, uint64_t>, int32_t>> key(n);
    for (int64_t c = 0; c < n; ++c) {
      uint64_t k = 0;
      for (int a = 0; a < naxis; ++a) {
        double span = hi[a] - lo[a];
        double t = span > 0 ? (h->cell_centre[3 * c + a] - lo[a]) / span : 0.0;
        uint64_t q = (uint64_t)std::min(t * (double)((1ull << bits) - 1), (double)((1ull << bits) - 1));
        k |= spread(q) << a;
      }
      key[c] = {{flev[c], k}, (int32_t)c};
    }
    std::sort(key.begin(), key.end());
    for (int64_t c = 0; c < n; ++c) {
      cinv[c] = key[c].second;
      cperm[key[c].second] = (int32_t)c;
      dmort[c] = key[c].first.second;
    }
    for (int64_t f = 0; f < nbf; ++f) fperm[f] = finv[f] = (int32_t)f;  // boundary faces keep ids
    std::vector<std::pair<std::pair<int32_t, int32_t>, int32_t>> fk;
    fk.reserve(nf - nbf);
    for (int64_t f = nbf; f < nf; ++f) {
      int32_t a = cperm[own[f]], b = cperm[nei[f]];
      fk.push_back({{std::min(a, b), std::max(a, b)}, (int32_t)f});
    }
    std::sort(fk.begin(), fk.end());
    for (int64_t k = 0; k < nf - nbf; ++k) {
      finv[nbf + k] = fk[k].second;
      fperm[fk[k].second] = (int32_t)(nbf + k);
    }
  }
  // Matrix storage: sliced ELL with slices of kSlice consecutive rows, each
  // slice column-major and as wide as its longest row.  Entry t of device
  // row c sits at rbeg[c] + t*kSlice, so a wavefront reading entry t of 64
  // rows reads one contiguous segment.  Padding entries are never read.
  const int64_t nslice = (n + kSlice - 1) / kSlice;
  std::vector<int32_t> rbeg(n), rend(n), sinfo(2 * nslice);
  int64_t nzlen = 0;
  for (int64_t s = 0; s < nslice; ++s) {
    int32_t w = 0;
    const int64_t c1 = std::min<int64_t>(n, (s + 1) * kSlice);
    for (int64_t c = s * kSlice; c < c1; ++c) w = std::max(w, rowptr[cinv[c] + 1] - rowptr[cinv[c]]);
    for (int64_t c = s * kSlice; c < c1; ++c) {
      rbeg[c] = (int32_t)(nzlen + (c - s * kSlice));
      rend[c] = rbeg[c] + kSlice * (rowptr[cinv[c] + 1] - rowptr[cinv[c]]);
    }
    sinfo[2 * s] = (int32_t)nzlen;
    sinfo[2 * s + 1] = w;
    nzlen += (int64_t)w * kSlice;
    XCAL_CHECK(nzlen < (1ll << 31), "matrix storage exceeds int32 indices");
  }
  M.nzlen = nzlen;
  std::vector<int32_t> colval_n(nzlen), nzperm(nnz), diag_n(n), trans_n(nzlen, -1);
  for (int64_t oc = 0; oc < n; ++oc)
    for (int32_t k = rowptr[oc]; k < rowptr[oc + 1]; ++k)
      nzperm[k] = rbeg[cperm[oc]] + kSlice * (k - rowptr[oc]);
  for (int64_t k = 0; k < nnz; ++k) colval_n[nzperm[k]] = cperm[colval[k]];
  for (int64_t k = 0; k < nnz; ++k) trans_n[nzperm[k]] = trans[k] >= 0 ? nzperm[trans[k]] : -1;
  for (int64_t oc = 0; oc < n; ++oc) diag_n[cperm[oc]] = nzperm[diag[oc]];
  // level schedules in device numbering: forward levels are the contiguous
  // ranges of the level-major order; backward levels list their rows
  std::vector<int32_t> rows_b;
  {
    M.lev_ptr_f.assign(M.nlev_f + 1, 0);
    for (int64_t i = 0; i < n; ++i) M.lev_ptr_f[flev[i] + 1]++;
    for (int l = 0; l < M.nlev_f; ++l) M.lev_ptr_f[l + 1] += M.lev_ptr_f[l];
    for (int64_t c = 0; c < n; ++c)
      XCAL_CHECK(c >= M.lev_ptr_f[flev[cinv[c]]] && c < M.lev_ptr_f[flev[cinv[c]] + 1],
                 "internal: level-major numbering");
    M.lev_ptr_b.assign(M.nlev_b + 1, 0);
    for (int64_t i = 0; i < n; ++i) M.lev_ptr_b[blev[i] + 1]++;
    for (int l = 0; l < M.nlev_b; ++l) M.lev_ptr_b[l + 1] += M.lev_ptr_b[l];
    rows_b.assign(n, 0);
    std::vector<int32_t> cur(M.lev_ptr_b.begin(), M.lev_ptr_b.end() - 1);
    for (int64_t c = 0; c < n; ++c) rows_b[cur[blev[cinv[c]]]++] = (int32_t)c;
    // Morton order inside each backward level, so that an eighth of a level
    // is one region of space (as for the forward levels)
    for (int l = 0; l < M.nlev_b; ++l)
      std::stable_sort(rows_b.begin() + M.lev_ptr_b[l], rows_b.begin() + M.lev_ptr_b[l + 1],
                       [&](int32_t a, int32_t b) { return dmort[a] < dmort[b]; });
  }
  // Packed triangular factors for the sweeps.  Level l of a sweep stores
  // entry t of its q-th row at base[l] + t*cnt_l + q (column-major inside
  // the level, as wide as the level's longest row, padding marked by
  // column -1), in the row's column order.  The address depends only on
  // the row's position, so a sweep thread loads all its entries at once.
  // Backward rows are in lev_rows_b order; posb maps a device row to it.
  std::vector<int32_t> Lcol, Lslot, Ucol, Uslot, posb(n);
  for (int64_t p = 0; p < n; ++p) posb[rows_b[p]] = (int32_t)p;
  auto pack = [&](const std::vector<int32_t>& ptr, int nlev, const int32_t* rows, bool lower,
                  std::vector<int32_t>& meta, std::vector<int32_t>& col,
                  std::vector<int32_t>& slot) {
    meta.assign(2 * (size_t)nlev, 0);
    auto span = [&](int32_t c, int32_t& k0, int32_t& k1) {
      k0 = lower ? rbeg[c] : diag_n[c] + kSlice;
      k1 = lower ? diag_n[c] : rend[c];
    };
    int64_t total = 0;
    for (int l = 0; l < nlev; ++l) {
      int32_t w = 0;
      for (int32_t pos = ptr[l]; pos < ptr[l + 1]; ++pos) {
        int32_t k0, k1;
        span(rows ? rows[pos] : pos, k0, k1);
        w = std::max(w, (k1 - k0) / kSlice);
      }
      meta[2 * l] = (int32_t)total;
      meta[2 * l + 1] = w;
      total += (int64_t)w * (ptr[l + 1] - ptr[l]);
      XCAL_CHECK(total < (1ll << 31), "packed factor exceeds int32 indices");
    }
    col.assign(total, -1);
    slot.assign(total, -1);
    for (int l = 0; l < nlev; ++l) {
      const int32_t cnt = ptr[l + 1] - ptr[l];
      for (int32_t q = 0; q < cnt; ++q) {
        const int32_t c = rows ? rows[ptr[l] + q] : ptr[l] + q;
        int32_t k0, k1;
        span(c, k0, k1);
        for (int32_t k = k0, t = 0; k < k1; k += kSlice, ++t) {
          const int64_t e = meta[2 * l] + (int64_t)t * cnt + q;
          col[e] = colval_n[k];
          slot[e] = k;
        }
      }
    }
  };
  pack(M.lev_ptr_f, M.nlev_f, nullptr, true, M.lmeta_f, Lcol, Lslot);
  pack(M.lev_ptr_b, M.nlev_b, rows_b.data(), false, M.lmeta_b, Ucol, Uslot);
  // Cone form of both sweeps (single meshes of at least kConeMinRows rows):
  // blocks of `rows` consecutive Morton keys, halved down to 256 rows
  // until the cones fit.  XCAL_NO_CONE=1 turns it off; XCAL_CONE_ROWS sets
  // the first block size tried.  On the 1M-cell random box, blocks of 2048
  // rows measured 38.5M cell-iters/s, 1024 rows 33.5M and 512 rows 25.9M:
  // larger blocks recompute fewer rows (cones of 1.75 n against 1.98 n and
  // 2.29 n) and 489 workgroups fit the chip at once.
  const char* nocone = std::getenv("XCAL_NO_CONE");
  if (cones && n >= kConeMinRows && !(nocone && *nocone == '1')) {
    int rows = kConeRows;
    if (const char* env = std::getenv("XCAL_CONE_ROWS")) rows = std::max(64, std::atoi(env));
    std::vector<int32_t> ord(n), flev_d(n), blev_d(n);
    for (int64_t c = 0; c < n; ++c) {
      ord[c] = (int32_t)c;
      flev_d[c] = flev[cinv[c]];
      blev_d[c] = blev[cinv[c]];
    }
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int32_t a, int32_t b) { return dmort[a] < dmort[b]; });
    for (int S = rows; S >= 64 && !M.cone; S /= 2) {
      M.cone = build_cone(n, S, true, ord, flev_d, M.nlev_f, rbeg, rend, diag_n, colval_n,
                          M.cone_f) &&
               build_cone(n, S, false, ord, blev_d, M.nlev_b, rbeg, rend, diag_n, colval_n,
                          M.cone_b);
      if (M.cone) M.cone_rows = S;
      if (S <= 256) break;
    }
    if (!M.cone) M.cone_f = M.cone_b = ConeHost();
  }
  std::vector<int32_t> cf_ptr_n(n + 1, 0), cf_face_n(cf_face.size()), cf_nbr_n(cf_face.size()),
      cf_slot_n(cf_face.size());
  std::vector<int8_t> cf_sign_n(cf_face.size());
  for (int64_t c = 0; c < n; ++c) {
    const int32_t oc = cinv[c];
    int32_t w = cf_ptr_n[c];
    for (int32_t k = cf_ptr[oc]; k < cf_ptr[oc + 1]; ++k, ++w) {
      cf_face_n[w] = fperm[cf_face[k]];
      cf_nbr_n[w] = cperm[cf_nbr[k]];
      cf_slot_n[w] = nzperm[cf_slot[k]];
      cf_sign_n[w] = cf_sign[k];
    }
    cf_ptr_n[c + 1] = w;
  }
  auto permf = [&](std::vector<double>& a) {
    std::vector<double> t(nf);
    for (int64_t f = 0; f < nf; ++f) t[fperm[f]] = a[f];
    a.swap(t);
  };
  for (auto* a : {&area, &weight, &delta, &nx, &ny, &nz, &ex, &ey, &ez, &sfx, &sfy, &sfz, &lapEf,
                  &rix, &riy, &riz, &ncw, &ntx, &nty, &ntz})
    permf(*a);
  std::vector<int32_t> own_n(nf), nei_n(nf), fslot_n(nf, -1);
  for (int64_t f = 0; f < nf; ++f) {
    own_n[fperm[f]] = cperm[own[f]];
    nei_n[fperm[f]] = cperm[nei[f]];
    if (f >= nbf) fslot_n[fperm[f]] = nzperm[fslot[f]];
  }
  for (auto& c : bc_cells) c = cperm[c];
  {
    std::vector<double> t(n);
    for (int64_t c = 0; c < n; ++c) t[cperm[c]] = vol[c];
    vol.swap(t);
  }
  cf_ptr.swap(cf_ptr_n);
  cf_face.swap(cf_face_n);
  cf_nbr.swap(cf_nbr_n);
  cf_slot.swap(cf_slot_n);
  cf_sign.swap(cf_sign_n);
  own.swap(own_n);
  nei.swap(nei_n);
  fslot.swap(fslot_n);
  diag.swap(diag_n);
  trans.swap(trans_n);
  M.cperm = cperm;
  M.fperm = fperm;
  M.nzperm = nzperm;
  // original-numbering CSR stays on the host for xcal_mesh_csr_pattern
  M.h_rowptr = rowptr;
  M.h_colval = colval;
  colval.swap(colval_n);
  M.vol.swap(vol);
  M.area.swap(area);
  M.weight.swap(weight);
  M.delta.swap(delta);
  M.nx.swap(nx);
  M.ny.swap(ny);
  M.nz.swap(nz);
  M.ex.swap(ex);
  M.ey.swap(ey);
  M.ez.swap(ez);
  M.sfx.swap(sfx);
  M.sfy.swap(sfy);
  M.sfz.swap(sfz);
  M.lapEf.swap(lapEf);
  M.rix.swap(rix);
  M.riy.swap(riy);
  M.riz.swap(riz);
  M.ncw.swap(ncw);
  M.ntx.swap(ntx);
  M.nty.swap(nty);
  M.ntz.swap(ntz);
  M.cf_ptr.swap(cf_ptr);
  M.cf_face.swap(cf_face);
  M.cf_nbr.swap(cf_nbr);
  M.cf_slot.swap(cf_slot);
  M.cf_sign.swap(cf_sign);
  M.diag.swap(diag);
  M.own.swap(own);
  M.nei.swap(nei);
  M.fslot.swap(fslot);
  M.bc_cells.swap(bc_cells);
  M.bc_ptr.swap(bc_ptr);
  M.bc_face.swap(bc_face);
  M.bpatch.swap(bpatch);
  M.rbeg.swap(rbeg);
  M.rend.swap(rend);
  M.sinfo.swap(sinfo);
  M.colval.swap(colval);
  M.trans.swap(trans);
  M.rows_b.swap(rows_b);
  M.posb.swap(posb);
  M.Lcol.swap(Lcol);
  M.Lslot.swap(Lslot);
  M.Ucol.swap(Ucol);
  M.Uslot.swap(Uslot);
}

// Union of independent cases.  Cells, matrix slots and packed factors are
// concatenated case by case (each case keeps its own device numbering);
// faces are numbered boundary-first over all cases (every case's boundary
// faces, then every case's internal faces), as the face kernels require.
// The reference numbering of the union is the cases' reference numberings
// concatenated.  All cases must have the same patches.
static void union_cases(std::vector<HostMesh>& C, HostMesh& U, std::vector<int64_t>& case_cell,
                        std::vector<int64_t>& case_bface, std::vector<int64_t>& case_iface,
                        std::vector<int32_t>& case_lev_f, std::vector<int32_t>& case_lev_b) {
  const size_t K = C.size();
  if (K == 1) {
    U = std::move(C[0]);
    case_cell = {0, U.n};
    case_bface = {0, U.nbf};
    case_iface = {U.nbf, U.nf};
    case_lev_f = {0, U.nlev_f};
    case_lev_b = {0, U.nlev_b};
    return;
  }
  int64_t N = 0, NF = 0, NBF = 0, NNZ = 0, NZ = 0, NCF = 0, NBC = 0, NL = 0, NU = 0;
  for (const HostMesh& M : C) {
    XCAL_CHECK(M.dim == C[0].dim, "all cases of a batch must have the same dimension");
    XCAL_CHECK(M.n_patches == C[0].n_patches, "all cases of a batch must have the same patches");
    N += M.n; NF += M.nf; NBF += M.nbf; NNZ += M.nnz; NZ += M.nzlen; NCF += M.ncf;
    NBC += M.nbcell; NL += (int64_t)M.Lcol.size(); NU += (int64_t)M.Ucol.size();
  }
  XCAL_CHECK(N < (1ll << 30) && NF < (1ll << 30) && NZ < (1ll << 31) && NL < (1ll << 31) &&
                 NU < (1ll << 31),
             "batch too large for int32 indices");
  U = HostMesh();
  U.dim = C[0].dim;
  U.n_patches = C[0].n_patches;
  U.patch_first = C[0].patch_first;
  U.patch_last = C[0].patch_last;
  U.n = N; U.nf = NF; U.nbf = NBF; U.nnz = NNZ; U.nzlen = NZ; U.ncf = NCF; U.nbcell = NBC;
  for (auto* v : {&U.vol}) v->resize(N);
  for (auto* v : {&U.area, &U.weight, &U.delta, &U.nx, &U.ny, &U.nz, &U.ex, &U.ey, &U.ez, &U.sfx,
                  &U.sfy, &U.sfz, &U.lapEf, &U.rix, &U.riy, &U.riz, &U.ncw, &U.ntx, &U.nty,
                  &U.ntz})
    v->resize(NF);
  U.own.resize(NF);
  U.nei.resize(NF);
  U.fslot.assign(NF, -1);
  U.bpatch.assign(NBF, -1);
  U.cperm.resize(N);
  U.fperm.resize(NF);
  U.nzperm.resize(NNZ);
  U.cf_ptr.assign(1, 0);
  U.bc_ptr.assign(1, 0);
  U.h_rowptr.assign(1, 0);
  case_cell.assign(1, 0);
  case_bface.assign(1, 0);
  case_iface.assign(1, NBF);
  case_lev_f.assign(1, 0);
  case_lev_b.assign(1, 0);
  int64_t co = 0, bo = 0, io = 0, zo = 0, ro = 0, Lo = 0, Uo = 0, fo = 0;
  for (size_t k = 0; k < K; ++k) {
    HostMesh& M = C[k];
    auto face = [&](int32_t f) -> int32_t {  // case face (device) -> union face
      return f < M.nbf ? (int32_t)(bo + f) : (int32_t)(NBF + io + (f - M.nbf));
    };
    auto cell = [&](int32_t c) -> int32_t { return (int32_t)(co + c); };
    auto slot = [&](int32_t z) -> int32_t { return z < 0 ? -1 : (int32_t)(zo + z); };
    for (int64_t c = 0; c < M.n; ++c) {
      U.vol[co + c] = M.vol[c];
      U.cperm[co + c] = cell(M.cperm[c]);
      U.diag.push_back(slot(M.diag[c]));
      U.rbeg.push_back(slot(M.rbeg[c]));
      U.rend.push_back(slot(M.rend[c]));
      U.rows_b.push_back(cell(M.rows_b[c]));
      U.posb.push_back(cell(M.posb[c]));
      for (int32_t j = M.cf_ptr[c]; j < M.cf_ptr[c + 1]; ++j) {
        U.cf_face.push_back(face(M.cf_face[j]));
        U.cf_nbr.push_back(cell(M.cf_nbr[j]));
        U.cf_slot.push_back(slot(M.cf_slot[j]));
        U.cf_sign.push_back(M.cf_sign[j]);
      }
      U.cf_ptr.push_back((int32_t)U.cf_face.size());
    }
    for (int64_t f = 0; f < M.nf; ++f) {
      const int32_t g = face((int32_t)f);
      U.area[g] = M.area[f]; U.weight[g] = M.weight[f]; U.delta[g] = M.delta[f];
      U.nx[g] = M.nx[f]; U.ny[g] = M.ny[f]; U.nz[g] = M.nz[f];
      U.ex[g] = M.ex[f]; U.ey[g] = M.ey[f]; U.ez[g] = M.ez[f];
      U.sfx[g] = M.sfx[f]; U.sfy[g] = M.sfy[f]; U.sfz[g] = M.sfz[f]; U.lapEf[g] = M.lapEf[f];
      U.rix[g] = M.rix[f]; U.riy[g] = M.riy[f]; U.riz[g] = M.riz[f];
      U.ncw[g] = M.ncw[f]; U.ntx[g] = M.ntx[f]; U.nty[g] = M.nty[f]; U.ntz[g] = M.ntz[f];
      U.own[g] = cell(M.own[f]);
      U.nei[g] = cell(M.nei[f]);
      U.fslot[g] = slot(M.fslot[f]);
      if (f < M.nbf) U.bpatch[g] = M.bpatch[f];
      U.fperm[fo + f] = face(M.fperm[f]);  // reference face of the union -> device face
    }
    for (int64_t s = 0; s < M.nbcell; ++s) {
      U.bc_cells.push_back(cell(M.bc_cells[s]));
      for (int32_t j = M.bc_ptr[s]; j < M.bc_ptr[s + 1]; ++j) U.bc_face.push_back(face(M.bc_face[j]));
      U.bc_ptr.push_back((int32_t)U.bc_face.size());
    }
    for (int64_t z = 0; z < M.nzlen; ++z) {
      U.colval.push_back(cell(M.colval[z]));
      U.trans.push_back(slot(M.trans[z]));
    }
    for (int64_t z = 0; z < M.nnz; ++z) U.nzperm[ro + z] = slot(M.nzperm[z]);
    for (int64_t c = 0; c < M.n; ++c) U.h_rowptr.push_back((int32_t)(ro + M.h_rowptr[c + 1]));
    for (int32_t v : M.h_colval) U.h_colval.push_back(cell(v));
    for (size_t e = 0; e < M.Lcol.size(); ++e) {
      U.Lcol.push_back(M.Lcol[e] < 0 ? -1 : cell(M.Lcol[e]));
      U.Lslot.push_back(slot(M.Lslot[e]));
    }
    for (size_t e = 0; e < M.Ucol.size(); ++e) {
      U.Ucol.push_back(M.Ucol[e] < 0 ? -1 : cell(M.Ucol[e]));
      U.Uslot.push_back(slot(M.Uslot[e]));
    }
    for (int l = 0; l < M.nlev_f; ++l) {
      U.lev_ptr_f.push_back(cell(M.lev_ptr_f[l]));
      U.lmeta_f.push_back((int32_t)(Lo + M.lmeta_f[2 * l]));
      U.lmeta_f.push_back(M.lmeta_f[2 * l + 1]);
    }
    for (int l = 0; l < M.nlev_b; ++l) {
      U.lev_ptr_b.push_back(cell(M.lev_ptr_b[l]));
      U.lmeta_b.push_back((int32_t)(Uo + M.lmeta_b[2 * l]));
      U.lmeta_b.push_back(M.lmeta_b[2 * l + 1]);
    }
    U.nlev_f += M.nlev_f;
    U.nlev_b += M.nlev_b;
    co += M.n; bo += M.nbf; io += M.nf - M.nbf; zo += M.nzlen; ro += M.nnz; fo += M.nf;
    Lo += (int64_t)M.Lcol.size(); Uo += (int64_t)M.Ucol.size();
    case_cell.push_back(co);
    case_bface.push_back(bo);
    case_iface.push_back(NBF + io);
    case_lev_f.push_back(U.nlev_f);
    case_lev_b.push_back(U.nlev_b);
    HostMesh().cperm.swap(M.cperm);  // release the case's memory early
    M = HostMesh();
  }
  U.lev_ptr_f.push_back((int32_t)N);
  U.lev_ptr_b.push_back((int32_t)N);
}

// Face values per cf entry k (face f, nsign ns): weight, |E_f| and delta of
// f, and S_f ns (exact: ns is +-1)
__global__ void __launch_bounds__(kBlock)
    k_cf_face_values(int64_t ncf, const int32_t* __restrict__ cf_face,
                     const int8_t* __restrict__ cf_sign, const double* __restrict__ w,
                     const double* __restrict__ lapEf, const double* __restrict__ delta,
                     const double* __restrict__ sx, const double* __restrict__ sy,
                     const double* __restrict__ sz, double* __restrict__ ow,
                     double* __restrict__ olap, double* __restrict__ odel, double* __restrict__ osx,
                     double* __restrict__ osy, double* __restrict__ osz) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncf) return;
  const int32_t f = cf_face[k];
  const double ns = (double)cf_sign[k];
  ow[k] = w[f];
  olap[k] = lapEf[f];
  odel[k] = delta[f];
  osx[k] = sx[f] * ns;
  osy[k] = sy[f] * ns;
  osz[k] = sz[f] * ns;
}

static xcal_mesh* upload_mesh(xcal_ctx* ctx, HostMesh& M) {
  XCAL_HIP(hipSetDevice(ctx->device));
  auto m = std::make_unique<xcal_mesh>();
  m->ctx = ctx;
  m->dim = M.dim;
  m->n = M.n;
  m->nf = M.nf;
  m->nbf = M.nbf;
  m->nnz = M.nnz;
  m->ncf = M.ncf;
  m->nzlen = M.nzlen;
  for (size_t c = 0; c < M.rbeg.size(); ++c)
    m->roww = std::max(m->roww, (M.rend[c] - M.rbeg[c]) / kSlice);
  const char* nofuse = std::getenv("XCAL_NO_FUSE");
  m->fuse = !(nofuse && *nofuse == '1');
  m->nbcell = M.nbcell;
  m->n_patches = M.n_patches;
  m->patch_first = M.patch_first;
  m->patch_last = M.patch_last;
  m->nlev_f = M.nlev_f;
  m->nlev_b = M.nlev_b;
  m->lev_ptr_f = M.lev_ptr_f;
  m->lev_ptr_b = M.lev_ptr_b;
  m->lmeta_f = M.lmeta_f;
  m->lmeta_b = M.lmeta_b;
  m->cperm = M.cperm;
  m->fperm = M.fperm;
  m->nzperm = M.nzperm;
  m->h_rowptr = M.h_rowptr;
  m->h_colval = M.h_colval;
  int64_t bytes = 0;
  upload(m->vol, M.vol, bytes);
  upload(m->cf_ptr, M.cf_ptr, bytes);
  upload(m->cf_face, M.cf_face, bytes);
  upload(m->cf_nbr, M.cf_nbr, bytes);
  upload(m->cf_slot, M.cf_slot, bytes);
  upload(m->cf_sign, M.cf_sign, bytes);
  upload(m->diag, M.diag, bytes);
  upload(m->own, M.own, bytes);
  upload(m->nei, M.nei, bytes);
  upload(m->fslot, M.fslot, bytes);
  upload(m->area, M.area, bytes);
  upload(m->weight, M.weight, bytes);
  upload(m->delta, M.delta, bytes);
  upload(m->nx, M.nx, bytes);
  upload(m->ny, M.ny, bytes);
  upload(m->nz, M.nz, bytes);
  upload(m->ex, M.ex, bytes);
  upload(m->ey, M.ey, bytes);
  upload(m->ez, M.ez, bytes);
  upload(m->sfx, M.sfx, bytes);
  upload(m->sfy, M.sfy, bytes);
  upload(m->sfz, M.sfz, bytes);
  upload(m->lapEf, M.lapEf, bytes);
  upload(m->rix, M.rix, bytes);
  upload(m->riy, M.riy, bytes);
  upload(m->riz, M.riz, bytes);
  upload(m->ncw, M.ncw, bytes);
  upload(m->ntx, M.ntx, bytes);
  upload(m->nty, M.nty, bytes);
  upload(m->ntz, M.ntz, bytes);
  {  // cf-ordered copies of the face values assembly reads per cell face
    const int64_t ncf = (int64_t)M.cf_face.size();
    for (auto* b : {&m->cf_w, &m->cf_lapEf, &m->cf_delta, &m->cf_sx, &m->cf_sy, &m->cf_sz})
      b->alloc(ncf);
    bytes += 6 * 8 * ncf;
    if (ncf > 0)
      hipLaunchKernelGGL(k_cf_face_values, dim3(grid_for(ncf)), dim3(kBlock), 0, ctx->stream, ncf,
                         m->cf_face.p, m->cf_sign.p, m->weight.p, m->lapEf.p, m->delta.p,
                         m->sfx.p, m->sfy.p, m->sfz.p, m->cf_w.p, m->cf_lapEf.p, m->cf_delta.p,
                         m->cf_sx.p, m->cf_sy.p, m->cf_sz.p);
    XCAL_HIP(hipGetLastError());
  }
  upload(m->bc_cells, M.bc_cells, bytes);
  upload(m->bc_ptr, M.bc_ptr, bytes);
  upload(m->bc_face, M.bc_face, bytes);
  upload(m->bface_patch, M.bpatch, bytes);
  upload(m->rbeg, M.rbeg, bytes);
  upload(m->rend, M.rend, bytes);
  upload(m->sinfo, M.sinfo, bytes);  // empty for a batch union (batch kernels use rbeg/rend)
  upload(m->colval, M.colval, bytes);
  upload(m->trans, M.trans, bytes);
  upload(m->lev_rows_b, M.rows_b, bytes);
  upload(m->posb, M.posb, bytes);
  upload(m->Lcol, M.Lcol, bytes);
  upload(m->Lslot, M.Lslot, bytes);
  upload(m->Ucol, M.Ucol, bytes);
  upload(m->Uslot, M.Uslot, bytes);
  {
    std::vector<int32_t> upos(M.Ucol.size());
    for (size_t e = 0; e < upos.size(); ++e) upos[e] = M.Ucol[e] < 0 ? -1 : M.posb[M.Ucol[e]];
    upload(m->Upos, upos, bytes);
  }
  upload(m->lmeta_f_d, M.lmeta_f, bytes);
  upload(m->lmeta_b_d, M.lmeta_b, bytes);
  if (M.cone) {
    auto up = [&](xcal_mesh::Cone& D, const ConeHost& H) {
      D.wk = H.wk;
      D.rows = (int64_t)H.crow.size();
      D.entries = (int64_t)H.lidx.size();
      upload(D.blk, H.blk, bytes);
      upload(D.base, H.base, bytes);
      upload(D.steps, H.steps, bytes);
      upload(D.crow, H.crow, bytes);
      upload(D.slot, H.slot, bytes);
      upload(D.dep, H.dep, bytes);
      upload(D.lidx, H.lidx, bytes);
      upload(D.hnum, H.hnum, bytes);
      D.band = H.band;
    };
    up(m->cone_f, M.cone_f);
    up(m->cone_b, M.cone_b);
    upload(m->cone_own, M.cone_f.own, bytes);
    upload(m->cone_ownp, M.cone_f.ownp, bytes);
    m->cone = true;
    m->cone_nblk = m->cone_f.nblk();
    m->cone_rows = M.cone_rows;
    const char* deep = std::getenv("XCAL_CONE_DEEP");
    m->cone_deep = deep && *deep == '1';
  }
  m->device_bytes = bytes;
  return m.release();
}

// A mesh (or batch) built on the host, ready to upload.
struct PreparedMesh {
  HostMesh U;
  std::vector<int64_t> cc, cb, ci;
  std::vector<int32_t> lf, lb;
  int ncase = 1;
  bool batch = false;
};

bool host_ingest_forced() {
  const char* e = std::getenv("XCAL_HOST_INGEST");
  return e && *e == '1';
}

std::shared_ptr<PreparedMesh> prepare_mesh(const xcal_mesh_host* const* hosts, int ncase, bool batch,
                                           xcal_ctx* ingest, bool cones) {
  XCAL_CHECK(hosts && ncase >= 1, "NULL argument or empty batch");
  if (host_ingest_forced()) ingest = nullptr;
  std::vector<HostMesh> C(ncase);
  for (int k = 0; k < ncase; ++k) {
    XCAL_CHECK(hosts[k], "NULL mesh in batch");
    build_host_mesh(hosts[k], C[k], cones && !batch && ncase == 1, ingest);
  }
  auto P = std::make_shared<PreparedMesh>();
  union_cases(C, P->U, P->cc, P->cb, P->ci, P->lf, P->lb);
  P->ncase = ncase;
  P->batch = batch;
  return P;
}

// The batch of several prepared single cases (the serving engine's batch of
// cases in flight); the parts are copied, not consumed.
std::shared_ptr<PreparedMesh> prepare_union(const std::vector<const PreparedMesh*>& parts) {
  XCAL_CHECK(!parts.empty(), "empty batch");
  std::vector<HostMesh> C;
  C.reserve(parts.size());
  for (const PreparedMesh* q : parts) {
    XCAL_CHECK(q && q->ncase == 1, "batch parts must be single cases");
    C.push_back(q->U);
  }
  auto P = std::make_shared<PreparedMesh>();
  union_cases(C, P->U, P->cc, P->cb, P->ci, P->lf, P->lb);
  P->ncase = (int)parts.size();
  P->batch = true;
  return P;
}

int64_t prepared_cells(const PreparedMesh& P) { return P.U.n; }

// Device bytes of a prepared mesh plus a steady SIMPLE run on it (fields,
// the U and p equations with DILU, FP64 and FP32 Krylov workspaces).
int64_t prepared_footprint(const PreparedMesh& P, bool turb) {
  const HostMesh& M = P.U;
  int64_t mesh = 0;
  mesh += 8 * (int64_t)(M.vol.size() + 20 * M.area.size());
  mesh += 4 * (int64_t)(M.cf_ptr.size() + M.cf_face.size() * 3 + M.diag.size() + 3 * M.own.size() +
                        M.bc_cells.size() + M.bc_ptr.size() + M.bc_face.size() + 2 * M.rbeg.size() +
                        M.sinfo.size() +
                        2 * M.colval.size() + 2 * M.rows_b.size() + 2 * M.Lcol.size() +
                        3 * M.Ucol.size());
  const int64_t n = M.n, nf = M.nf, nz = M.nzlen;
  const int64_t fields = 8 * (22 * n + 7 * nf);
  const int64_t eqns = 8 * (3 * nz + 4 * n) + 8 * 2 * (nz + 3 * n);
  const int64_t krylov = 2 * 8 * 12 * n + 4 * (12 * n + nz + 3 * n);
  const int64_t turbulence = turb ? 8 * (8 * n + 3 * nf) : 0;
  return mesh + fields + eqns + krylov + turbulence;
}

xcal_mesh* upload_prepared(xcal_ctx* ctx, PreparedMesh& P) {
  if (P.batch) {  // batches run the batched solvers, which have no cone form
    P.U.cone = false;
    P.U.cone_f = P.U.cone_b = ConeHost();
  }
  xcal_mesh* m = upload_mesh(ctx, P.U);
  std::unique_ptr<xcal_mesh> guard(m);
  const int ncase = P.ncase;
  const bool batch = P.batch;
  const std::vector<int64_t>&cc = P.cc, &cb = P.cb, &ci = P.ci;
  const std::vector<int32_t>&lf = P.lf, &lb = P.lb;
  m->ncase = ncase;
  m->batch = batch;
  m->case_cell = cc;
  m->case_bface = cb;
  m->case_iface = ci;
  m->case_lev_f = lf;
  m->case_lev_b = lb;
  // workgroup table: blocks of kBlock rows that never straddle two cases
  std::vector<int32_t> b0, b1, bk, cblk(1, 0), iota(ncase + 1), fcase(m->nf, 0);
  for (int k = 0; k < ncase; ++k) {
    for (int64_t r = cc[k]; r < cc[k + 1]; r += kBlock) {
      b0.push_back((int32_t)r);
      b1.push_back((int32_t)std::min<int64_t>(r + kBlock, cc[k + 1]));
      bk.push_back(k);
    }
    cblk.push_back((int32_t)b0.size());
    for (int64_t f = cb[k]; f < cb[k + 1]; ++f) fcase[f] = k;
    for (int64_t f = ci[k]; f < ci[k + 1]; ++f) fcase[f] = k;
  }
  for (int k = 0; k <= ncase; ++k) iota[k] = k;
  m->nblk = (int64_t)b0.size();
  int64_t bytes = 0;
  upload(m->case_cell_d, cc, bytes);
  upload(m->case_lev_f_d, lf, bytes);
  upload(m->case_lev_b_d, lb, bytes);
  upload(m->blk_row0, b0, bytes);
  upload(m->blk_row1, b1, bytes);
  upload(m->blk_case, bk, bytes);
  upload(m->case_blk_d, cblk, bytes);
  upload(m->case_iota_d, iota, bytes);
  if (batch) upload(m->face_case, fcase, bytes);
  m->device_bytes += bytes;
  make_sweep_plans(m, 0);
  // pipelined batch sweeps: every case within their limits
  // dep(e): the row (forward) or position (backward) that packed entry e
  // depends on, -1 on padding; every one must lie within the LDS ring.
  auto pipe_bound = [&](const std::vector<int32_t>& ptr, const std::vector<int32_t>& meta,
                        const std::vector<int32_t>& cl, bool& far, auto&& dep) {
    int maxw = 0;
    far = false;
    for (int k = 0; k < ncase; ++k) {
      if (cl[k + 1] - cl[k] > kPipeLev) return -1;
      for (int l = cl[k]; l < cl[k + 1]; ++l) {
        const int32_t cnt = ptr[l + 1] - ptr[l], w = meta[2 * l + 1];
        if (cnt > kPipeBlock) return -1;
        maxw = std::max(maxw, w);
        for (int t = 0; t < w; ++t)
          for (int32_t q = 0; q < cnt; ++q) {
            const int32_t d = dep((int64_t)meta[2 * l] + (int64_t)t * cnt + q);
            far = far || (d >= 0 && d < ptr[l + 1] - kPipeRing);
          }
      }
    }
    const int wk = entry_bound(maxw);
    return wk > 0 ? wk : -1;
  };
  const char* nopipe = std::getenv("XCAL_NO_PIPE");
  if (!batch && !(nopipe && *nopipe == '1')) {
    // single meshes: the runs of small levels (dependencies before a run
    // reach its ring through the preload)
    const HostMesh& M = P.U;
    auto run_bound = [&](SweepStep& st, const std::vector<int32_t>& ptr,
                         const std::vector<int32_t>& meta, size_t npacked, auto&& dep) {
      st.pipe = 0;
      st.far = false;
      if (npacked == 0) return;  // masked entries read entry 0 of the packed arrays
      if (st.lev1 - st.lev0 < kPipeMinLevels) return;
      int maxw = 0;
      bool far = false;
      for (int l = st.lev0; l < st.lev1; ++l) {
        const int32_t cnt = ptr[l + 1] - ptr[l], w = meta[2 * l + 1];
        if (cnt > kPipeBlock) return;
        maxw = std::max(maxw, w);
        for (int t = 0; t < w; ++t)
          for (int32_t q = 0; q < cnt; ++q) {
            const int32_t d = dep((int64_t)meta[2 * l] + (int64_t)t * cnt + q);
            far = far || (d >= 0 && d < ptr[l + 1] - kPipeRing);
          }
      }
      st.pipe = entry_bound(std::max(maxw, 1));
      st.far = far;
    };
    for (SweepStep& st : m->plan_f)
      if (st.grid == 1) {
        run_bound(st, m->lev_ptr_f, m->lmeta_f, M.Lcol.size(),
                  [&](int64_t e) { return M.Lcol[e]; });
        m->pipe_runs = m->pipe_runs || st.pipe > 0;
      }
    for (SweepStep& st : m->plan_b)
      if (st.grid == 1)
        run_bound(st, m->lev_ptr_b, m->lmeta_b, M.Ucol.size(), [&](int64_t e) {
          return M.Ucol[e] < 0 ? -1 : M.posb[M.Ucol[e]];
        });
  }
  if (batch && !(nopipe && *nopipe == '1')) {
    const HostMesh& M = P.U;
    m->pipe_wf = pipe_bound(m->lev_ptr_f, m->lmeta_f, lf, m->pipe_far_f,
                            [&](int64_t e) { return M.Lcol[e]; });
    m->pipe_wb = pipe_bound(m->lev_ptr_b, m->lmeta_b, lb, m->pipe_far_b, [&](int64_t e) {
      return M.Ucol[e] < 0 ? -1 : M.posb[M.Ucol[e]];
    });
  }
  return guard.release();
}

static xcal_mesh* create_mesh(xcal_ctx* ctx, const xcal_mesh_host* const* hosts, int ncase,
                              bool batch) {
  XCAL_CHECK(ctx, "NULL context");
  auto P = prepare_mesh(hosts, ncase, batch, ctx);
  return upload_prepared(ctx, *P);
}

}  // namespace xcal

using namespace xcal;

extern "C" {

const char* xcal_last_error(void) { return g_last_error.c_str(); }
int32_t xcal_abi_version(void) { return 2; }

int xcal_ctx_create(int32_t device, xcal_ctx** out) {
  return guarded([&] {
    XCAL_CHECK(out != nullptr, "out is NULL");
    int ndev = 0;
    XCAL_HIP(hipGetDeviceCount(&ndev));
    XCAL_CHECK(device >= 0 && device < ndev, "device index out of range");
    XCAL_HIP(hipSetDevice(device));
    auto* c = new xcal_ctx();
    c->device = device;
    try {
      XCAL_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      c->red_result.alloc(64);
      c->nostop.alloc(32);
      XCAL_HIP(hipMemset(c->nostop.p, 0, sizeof(double) * 32));
      XCAL_HIP(hipHostMalloc((void**)&c->host_pinned, sizeof(double) * 64, hipHostMallocDefault));
      noalloc_init(device);
    } catch (...) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      if (c->host_pinned) (void)hipHostFree(c->host_pinned);
      delete c;
      throw;
    }
    *out = c;
    return 0;
  });
}

int xcal_ctx_destroy(xcal_ctx* c) {
  return guarded([&] {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& pr : c->prof_pending) c->prof_pool.insert(c->prof_pool.end(), {pr.a, pr.b});
    for (hipEvent_t ev : c->prof_pool) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(c->stream);
    if (c->host_pinned) (void)hipHostFree(c->host_pinned);
    delete c;
    return 0;
  });
}

int xcal_ctx_synchronize(xcal_ctx* c) {
  return guarded([&] {
    XCAL_HIP(hipStreamSynchronize(c->stream));
    return 0;
  });
}

int xcal_mesh_create(xcal_ctx* ctx, const xcal_mesh_host* h, xcal_mesh** out) {
  return guarded([&] {
    XCAL_CHECK(ctx && h && out, "NULL argument");
    *out = create_mesh(ctx, &h, 1, false);
    return 0;
  });
}

int xcal_mesh_create_batch(xcal_ctx* ctx, const xcal_mesh_host* const* hosts, int32_t ncase,
                           xcal_mesh** out) {
  return guarded([&] {
    XCAL_CHECK(ctx && hosts && out, "NULL argument");
    *out = create_mesh(ctx, hosts, ncase, true);
    return 0;
  });
}

int xcal_mesh_cases(const xcal_mesh* m, int32_t* ncase, int64_t* cell_offsets) {
  return guarded([&] {
    XCAL_CHECK(m, "mesh is NULL");
    if (ncase) *ncase = m->ncase;
    if (cell_offsets) std::copy(m->case_cell.begin(), m->case_cell.end(), cell_offsets);
    return 0;
  });
}

int xcal_mesh_destroy(xcal_mesh* m) {
  return guarded([&] {
    // Destroy touches only the handle's own memory (never its context), so
    // handles may be released in any order; hipFree waits for the device.
    if (m) {
      (void)hipDeviceSynchronize();
      delete m;
    }
    return 0;
  });
}

int xcal_mesh_info(const xcal_mesh* m, int64_t* n_cells, int64_t* nnz, int32_t* fl, int32_t* bl,
                   int64_t* bytes) {
  return guarded([&] {
    XCAL_CHECK(m, "mesh is NULL");
    if (n_cells) *n_cells = m->n;
    if (nnz) *nnz = m->nnz;
    if (fl) *fl = m->nlev_f;
    if (bl) *bl = m->nlev_b;
    if (bytes) *bytes = m->device_bytes;
    return 0;
  });
}

int xcal_mesh_cone_form(const xcal_mesh* m, int32_t* block_rows, int32_t* blocks,
                        int64_t* fwd_rows, int64_t* bwd_rows) {
  return guarded([&] {
    XCAL_CHECK(m, "mesh is NULL");
    if (block_rows) *block_rows = m->cone ? m->cone_rows : 0;
    if (blocks) *blocks = m->cone ? m->cone_nblk : 0;
    if (fwd_rows) *fwd_rows = m->cone ? m->cone_f.rows : 0;
    if (bwd_rows) *bwd_rows = m->cone ? m->cone_b.rows : 0;
    return 0;
  });
}

int xcal_mesh_cone_bands(const xcal_mesh* m, int32_t* fwd, int32_t* bwd) {
  return guarded([&] {
    XCAL_CHECK(m && fwd && bwd, "NULL argument");
    *fwd = m->cone ? (int32_t)m->cone_f.band.size() - 1 : 0;
    *bwd = m->cone ? (int32_t)m->cone_b.band.size() - 1 : 0;
    return 0;
  });
}

int xcal_mesh_sweep_far(const xcal_mesh* m, int32_t* fwd, int32_t* bwd) {
  return guarded([&] {
    XCAL_CHECK(m && fwd && bwd, "NULL argument");
    *fwd = m->pipe_wf > 0 && m->pipe_far_f ? 1 : 0;
    *bwd = m->pipe_wb > 0 && m->pipe_far_b ? 1 : 0;
    return 0;
  });
}

int xcal_mesh_pipe_tuning(const xcal_mesh* m, int32_t* fwd_rows, int32_t* bwd_rows, double* us) {
  return guarded([&] {
    XCAL_CHECK(m && fwd_rows && bwd_rows && us, "NULL argument");
    *fwd_rows = m->pipe_rw_f;
    *bwd_rows = m->pipe_rw_b;
    for (int d = 0; d < 2; ++d)
      for (int c = 0; c < 3; ++c) us[3 * d + c] = m->pipe_tune.us[d][c];
    return 0;
  });
}

int xcal_mesh_sweep_form(const xcal_mesh* m, int32_t* fwd, int32_t* bwd) {
  return guarded([&] {
    XCAL_CHECK(m && fwd && bwd, "NULL argument");
    *fwd = m->pipe_wf;
    *bwd = m->pipe_wb;
    if (!m->batch) {  // single meshes: the widest pipelined run, -1 if none
      *fwd = *bwd = -1;
      for (const SweepStep& st : m->plan_f)
        if (st.pipe > 0) *fwd = std::max(*fwd, st.pipe);
      for (const SweepStep& st : m->plan_b)
        if (st.pipe > 0) *bwd = std::max(*bwd, st.pipe);
    }
    return 0;
  });
}

int xcal_mesh_csr_pattern(const xcal_mesh* m, int32_t* rowptr, int32_t* colval) {
  return guarded([&] {
    XCAL_CHECK(m && rowptr && colval, "NULL argument");
    std::copy(m->h_rowptr.begin(), m->h_rowptr.end(), rowptr);
    std::copy(m->h_colval.begin(), m->h_colval.end(), colval);
    return 0;
  });
}

int xcal_field_create(xcal_mesh* m, int32_t kind, xcal_field** out) {
  return guarded([&] {
    XCAL_CHECK(m && out, "NULL argument");
    XCAL_HIP(hipSetDevice(m->ctx->device));
    auto f = std::make_unique<xcal_field>();
    f->mesh = m;
    f->kind = kind;
    switch (kind) {
      case XCAL_CELL_SCALAR: f->ncomp = 1; f->len = m->n; break;
      case XCAL_CELL_VECTOR: f->ncomp = 3; f->len = m->n; break;
      case XCAL_FACE_SCALAR: f->ncomp = 1; f->len = m->nf; break;
      case XCAL_FACE_VECTOR: f->ncomp = 3; f->len = m->nf; break;
      case XCAL_CELL_TENSOR: f->ncomp = 9; f->len = m->n; break;
      default: throw Error(XCAL_ERR_ARG, "unknown field kind");
    }
    f->data.alloc((int64_t)f->ncomp * f->len);
    XCAL_HIP(hipMemsetAsync(f->data.p, 0, f->data.bytes(), m->ctx->stream));
    XCAL_HIP(hipStreamSynchronize(m->ctx->stream));
    *out = f.release();
    return 0;
  });
}

int xcal_field_destroy(xcal_field* f) {
  return guarded([&] {
    if (f) {
      (void)hipDeviceSynchronize();
      delete f;
    }
    return 0;
  });
}

}  // extern "C"

namespace xcal {
// reference id -> device id map of a field's entities
const std::vector<int32_t>& field_perm(const xcal_field* f) {
  return (f->kind == XCAL_FACE_SCALAR || f->kind == XCAL_FACE_VECTOR) ? f->mesh->fperm
                                                                      : f->mesh->cperm;
}
void upload_permuted(const xcal_mesh* m, const std::vector<int32_t>& perm, double* dev,
                     const double* host, int64_t len) {
  std::vector<double> t(len);
  for (int64_t i = 0; i < len; ++i) t[perm[i]] = host[i];
  XCAL_HIP(hipMemcpyAsync(dev, t.data(), sizeof(double) * len, hipMemcpyHostToDevice, m->ctx->stream));
  XCAL_HIP(hipStreamSynchronize(m->ctx->stream));
}
void download_permuted(const xcal_mesh* m, const std::vector<int32_t>& perm, const double* dev,
                       double* host, int64_t len, int64_t dev_len) {
  if (dev_len < 0) dev_len = len;
  std::vector<double> t(dev_len);
  XCAL_HIP(hipMemcpyAsync(t.data(), dev, sizeof(double) * dev_len, hipMemcpyDeviceToHost,
                          m->ctx->stream));
  XCAL_HIP(hipStreamSynchronize(m->ctx->stream));
  for (int64_t i = 0; i < len; ++i) host[i] = t[perm[i]];
}
}  // namespace xcal

extern "C" {

int xcal_field_upload(xcal_field* f, int32_t c, const double* host) {
  return guarded([&] {
    XCAL_CHECK(f && host, "NULL argument");
    XCAL_CHECK(c >= 0 && c < f->ncomp, "component out of range");
    upload_permuted(f->mesh, field_perm(f), f->comp(c), host, f->len);
    return 0;
  });
}

int xcal_field_download(const xcal_field* f, int32_t c, double* host) {
  return guarded([&] {
    XCAL_CHECK(f && host, "NULL argument");
    XCAL_CHECK(c >= 0 && c < f->ncomp, "component out of range");
    download_permuted(f->mesh, field_perm(f), f->comp(c), host, f->len);
    return 0;
  });
}

int xcal_field_fill(xcal_field* f, int32_t c, double v) {
  return guarded([&] {
    XCAL_CHECK(f, "NULL argument");
    XCAL_CHECK(c >= 0 && c < f->ncomp, "component out of range");
    fill(f->mesh->ctx->stream, f->comp(c), f->len, v);
    XCAL_HIP(hipGetLastError());
    return 0;
  });
}

}  // extern "C"
