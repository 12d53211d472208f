// Internal declarations of libxcalhip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <memory>
#include <string>
#include <map>
#include <array>
#include <vector>

#include "xcal.h"

namespace xcal {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_error(const std::string& msg);

#define XCAL_HIP(call)                                                               \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::xcal::Error(XCAL_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define XCAL_CHECK(cond, msg)                                      \
  do {                                                             \
    if (!(cond)) throw ::xcal::Error(XCAL_ERR_ARG, (msg));         \
  } while (0)

// Runs fn, mapping exceptions to status codes (returns fn's int result).
template <class Fn>
int guarded(Fn&& fn) {
  try {
    return fn();
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return XCAL_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return XCAL_ERR_ARG;
  }
}

constexpr int kBlock = 256;
constexpr int kMaxTerms = 4;
constexpr int kMaxSources = 2;
constexpr int kMaxPatches = 32;
constexpr int kSlice = 64;        // rows per slice of the sliced-ELL matrix storage
constexpr int kRunBlock = 1024;   // workgroup size of the single-workgroup sweep runs
constexpr int kSmallLevel = 256;  // levels below this many rows join a single-workgroup run
// Meshes with at least this many rows form CG's p = z + beta p in a streamed
// pass before a single-gather SpMV; smaller ones form it on the fly inside the
// SpMV, saving a launch.
constexpr int64_t kSplitRows = 1 << 18;
// Krylov solves with a diagonal (or no) preconditioner on meshes whose
// launches have at most this many workgroups fold each one-workgroup
// finalize into the next kernel of the iteration.
constexpr int kFuseBlocks = 256;

// Compile-time entry bound for a level whose rows have w entries: 0 selects
// the general loop.
inline int entry_bound(int w) {
  return w <= 0 ? 0 : w <= 2 ? 2 : w <= 4 ? 4 : w <= 8 ? 8 : 0;
}
// Limits of the pipelined one-workgroup-per-case sweeps of batch meshes:
// rows per level (one per thread) and levels per case (staged in LDS).
constexpr int kPipeBlock = 256;
constexpr int kPipeLev = 2048;
constexpr int kPipeRing = 4096;  // LDS ring of recent sweep outputs
// Runs shorter than this keep the level-stepping form: the pipeline's set-up
// (level table, ring preload, P levels of prefetch) outweighs its gain.
constexpr int kPipeMinLevels = 16;

// Cone form of the DILU sweeps on single meshes: one workgroup per block of
// cells sweeps the block's rows and their dependency cone in LDS.
constexpr int kConeBlock = 256;       // threads per workgroup = most rows per step
constexpr int kConeMax = 4608;        // most cone rows of one block (LDS values)
constexpr int kConeSteps = 64;        // most steps of one block (LDS table)
constexpr int kConeTab = 20;          // ints per step: r0, cnt, eoff[8], ecnt[8], 2 unused
constexpr int kConeRows = 2048;       // block size tried first
constexpr int64_t kConeMinRows = 1 << 16;  // smaller meshes keep the level launches
constexpr double kConeMaxRatio = 3.0;      // most cone rows in all, per mesh row
constexpr int kConeMinBand = 4;            // narrowest band of levels
constexpr int kConeMaxBands = 128;         // most bands (launches) per sweep
constexpr int kConeMinBlock = 64;          // smallest block of a band
constexpr int kConeBandBlocks = 512;       // blocks a band's block size aims at
constexpr double kConeBandRatio = 4.0;     // most cone positions per row of a band

// Cut c_x (x = 0..8) of a level of cnt rows into eighths at 64-row
// boundaries; eighth x is worked by the workgroups on XCD x.
inline int64_t eighth_cut(int64_t cnt, int x) { return x >= 8 ? cnt : cnt * x / 8 / 64 * 64; }
// Workgroups of a level launch: 8 x (blocks of the longest eighth).
inline int level_grid(int64_t cnt, int block = 256) {
  int64_t m = 0;
  for (int x = 0; x < 8; ++x) m = std::max(m, eighth_cut(cnt, x + 1) - eighth_cut(cnt, x));
  return (int)(8 * ((m + block - 1) / block));
}
inline int grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : g);
}

// device buffer with RAII
template <class T>
struct DBuf {
  T* p = nullptr;
  int64_t n = 0;
  DBuf() = default;
  explicit DBuf(int64_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  void alloc(int64_t count) {
    release();
    n = count;
    if (count > 0) XCAL_HIP(hipMalloc(&p, sizeof(T) * (size_t)count));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
  int64_t bytes() const { return n * (int64_t)sizeof(T); }
};

}  // namespace xcal

// ----------------------------------------------------------------- handles
struct xcal_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  xcal::DBuf<double> red_result;    // 64 slots: results of device reductions
  xcal::DBuf<double> nostop;        // zeroed Krylov state (stop flag clear) for sweeps
                                    // that run outside a Krylov solve
  double* host_pinned = nullptr;    // 64 doubles, pinned
  // optional HIP-event timing of the pressure-solve kernels (0 SpMV, 1
  // forward sweep, 2 backward sweep), each on or off independently
  static constexpr int kProfKernels = 3;
  bool prof_on[kProfKernels] = {};
  int64_t prof_seq[kProfKernels] = {};  // launches of each kernel since its enable
  int64_t prof_launches[kProfKernels] = {};
  double prof_ms[kProfKernels] = {};
  std::vector<hipEvent_t> prof_pool;
  struct ProfPair {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<ProfPair> prof_pending;
};

// One launch of a DILU sweep over levels [lev0, lev1): one level over `grid`
// workgroups, or (grid == 1) a run of small levels in one workgroup.  Rows
// [row0, row1) are device rows (forward) or positions in lev_rows_b
// (backward); the launch writes partial sums to slots [slot, slot + grid).
struct SweepStep {
  int32_t lev0, lev1, row0, row1, grid, slot;
  int32_t pipe = 0;  // runs: entry bound of the pipelined form, 0 = level stepping
  bool far = false;  // some dependency of the run lies before its LDS ring
};

// Device mesh in SoA form, 0-based int32 indices, device numbering.
struct xcal_mesh {
  xcal_ctx* ctx = nullptr;
  int dim = 3;
  int64_t n = 0, nf = 0, nbf = 0, nnz = 0, ncf = 0;
  int n_patches = 0;
  std::vector<int64_t> patch_first, patch_last;  // 0-based inclusive
  // cells
  xcal::DBuf<double> vol;
  xcal::DBuf<int32_t> cf_ptr, cf_face, cf_nbr, cf_slot;
  // per cell-face entry, in cf order: the face's weight, |E_f| and delta,
  // and S_f times the entry's nsign (assembly and Green-Gauss stream them
  // instead of gathering them by face)
  xcal::DBuf<double> cf_w, cf_lapEf, cf_delta, cf_sx, cf_sy, cf_sz;
  xcal::DBuf<int8_t> cf_sign;
  xcal::DBuf<int32_t> diag;
  // faces
  xcal::DBuf<int32_t> own, nei, fslot;  // fslot: nz index of (owner, neighbour)
  xcal::DBuf<double> area, weight, delta, nx, ny, nz, ex, ey, ez, sfx, sfy, sfz, lapEf;
  xcal::DBuf<double> rix, riy, riz;  // Ri = rf - (rP + rN)/2 of internal faces (Midpoint)
  xcal::DBuf<double> ncw, ntx, nty, ntz;  // non-orthogonal correction: weight, T_hat
  // boundary faces grouped by cell (ascending face id within a cell)
  int64_t nbcell = 0;
  xcal::DBuf<int32_t> bc_cells, bc_ptr, bc_face;
  xcal::DBuf<int8_t> bface_patch;  // per boundary face
  // device numbering: cperm/fperm map reference cell/face ids to device ids,
  // nzperm maps reference CSR slots to device slots
  std::vector<int32_t> cperm, fperm, nzperm;
  // CSR pattern in the reference's numbering (host only)
  std::vector<int32_t> h_rowptr, h_colval;
  // device matrix storage (sliced ELL): entries of row c at
  // rbeg[c], rbeg[c] + kSlice, ... < rend[c], in the reference's column order
  int64_t nzlen = 0;  // storage length including slice padding
  int32_t roww = 0;   // most entries in one row
  bool fuse = true;   // fused Krylov finalizes allowed (XCAL_NO_FUSE=1 at creation: off)
  xcal::DBuf<int32_t> rbeg, rend, colval, trans;
  xcal::DBuf<int32_t> sinfo;  // per slice of kSlice rows: first slot, width (entries)
  // DILU level schedules: forward level l = device rows
  // [lev_ptr_f[l], lev_ptr_f[l+1]); backward level l = lev_rows_b over
  // [lev_ptr_b[l], lev_ptr_b[l+1])
  int nlev_f = 0, nlev_b = 0;
  std::vector<int32_t> lev_ptr_f, lev_ptr_b;
  // Row blocks of the gathering row kernels (SpMV and friends), in dispatch
  // order: block b = rows [xblk[2b], xblk[2b+1]) runs on XCD b%8, which gets
  // the (b%8)-th 64-aligned eighth of every forward level.  Rows are Morton
  // ordered within a level, so an XCD's rows -- and the neighbours they
  // gather -- sit in one eighth of the domain and share its L2.
  // Backward-level position of each packed upper entry's column (-1 on
  // padding): where the pipelined backward sweep finds it in its LDS ring.
  xcal::DBuf<int32_t> Upos;
  // Batch meshes: entry bound of the pipelined one-workgroup-per-case
  // sweeps (forward, backward), or -1 when a case exceeds their limits.
  int pipe_wf = -1, pipe_wb = -1;
  bool pipe_far_f = false, pipe_far_b = false;  // some dependency before the ring
  // rows of a level per thread of the pipelined batch sweeps (1, 2 or 4):
  // set by the tuner on the first DILU solve (or XCAL_PIPE_ROWS), -1 untuned;
  // the tuner's timings (µs per launch, [forward, backward][RW 1, 2, 4])
  mutable int pipe_rw_f = -1, pipe_rw_b = -1;
  struct PipeTuning {
    int rw_f = 1, rw_b = 1;
    double us[2][3] = {};
  };
  mutable PipeTuning pipe_tune;
  // Single meshes: some forward run is pipelined (needs Lrd and Lprod)
  bool pipe_runs = false;
  std::vector<int32_t> xblk;
  xcal::DBuf<int32_t> xblk_d;
  int nxblk = 0;
  xcal::DBuf<int32_t> lev_ptr_f_d, lev_ptr_b_d, lev_rows_b, posb;
  // packed lower (forward) / upper (backward) factors: per level l,
  // lmeta[2l] = base, lmeta[2l+1] = width; entry t of the level's q-th row at
  // base + t*cnt + q; *col = device column (-1: padding), *slot = nz slot
  std::vector<int32_t> lmeta_f, lmeta_b;
  xcal::DBuf<int32_t> lmeta_f_d, lmeta_b_d, Lcol, Lslot, Ucol, Uslot;
  // cases of a batch (one case for a plain mesh): cells [case_cell[k],
  // case_cell[k+1]), boundary faces [case_bface[k], case_bface[k+1]),
  // internal faces [case_iface[k], case_iface[k+1]), forward/backward levels
  // [case_lev_*[k], case_lev_*[k+1]) of the level arrays
  int ncase = 1;
  bool batch = false;  // created as a batch (of any size): batched solvers
  std::vector<int64_t> case_cell, case_bface, case_iface;
  std::vector<int32_t> case_lev_f, case_lev_b;
  xcal::DBuf<int64_t> case_cell_d;
  xcal::DBuf<int32_t> case_lev_f_d, case_lev_b_d;
  // batched solves: workgroup b covers rows [blk_row0[b], blk_row1[b]) of
  // case blk_case[b]; case k owns workgroups [case_blk[k], case_blk[k+1]);
  // case_iota = 0..ncase; face_case = case of each face (batches only)
  int64_t nblk = 0;
  xcal::DBuf<int32_t> blk_row0, blk_row1, blk_case, case_blk_d, case_iota_d, face_case;
  // Cone form of both DILU sweeps (single meshes; sweep_cone.h): when `cone`
  // is set, each sweep is one launch per band of levels, of the band's
  // blocks, and the backward sweep's partial sums fill one slot per block.
  struct Cone {
    int wk = 0;  // entry bound
    int64_t rows = 0, entries = 0;
    xcal::DBuf<int32_t> blk, base, steps, crow, slot, dep, hnum;
    xcal::DBuf<uint16_t> lidx;
    std::vector<int32_t> band;  // host: first block of each band (nbands + 1)
    int nblk() const { return band.empty() ? 0 : band.back(); }
  };
  // the forward sweep's block b owns rows cone_own[cone_ownp[b], cone_ownp[b+1])
  xcal::DBuf<int32_t> cone_own, cone_ownp;
  bool cone = false;
  int cone_nblk = 0, cone_rows = 0;  // forward blocks, rows per block
  bool cone_deep = false;            // deep prefetch (XCAL_CONE_DEEP=1 at creation)
  Cone cone_f, cone_b;
  int small_level = 0;
  std::vector<SweepStep> plan_f, plan_b;
  int sweep_slots = 0;  // partial-sum slots used by one backward sweep
  int64_t device_bytes = 0;
};

struct xcal_field {
  xcal_mesh* mesh = nullptr;
  int kind = 0;
  int ncomp = 1;
  int64_t len = 0;  // entries per component
  xcal::DBuf<double> data;
  double* comp(int c) const { return data.p + (int64_t)c * len; }
};

struct xcal_eqn {
  xcal_mesh* mesh = nullptr;
  bool vector = false;
  xcal::DBuf<double> nzval, nzval0;
  xcal::DBuf<double> b;  // 1 or 3 components of length n
  double* bc(int c) const { return b.p + (int64_t)c * mesh->n; }
  // preconditioner storage
  int precon = XCAL_PRECON_NONE;
  // the last discretise!'s first term was Time{CrankNicolson}: solves take
  // the step x = 2 x - x_old (Solve_1_api.jl:291-296)
  bool cn = false;
  xcal::DBuf<double> pS;   // Jacobi 1/|a_ii|
  xcal::DBuf<double> pD;   // DILU D
  xcal::DBuf<double> prD;  // DILU 1/D
  xcal::DBuf<double> Lnz, Unz, Db;  // packed factor values; D in backward-level order
  xcal::DBuf<double> Lrd;  // 1/D of each packed lower entry's column (0 on padding)
  xcal::DBuf<double> Lprod;  // a_ij a_ji of each packed lower entry (DILU diagonal update)
  // cone form: a_ij and 1/D_j per forward cone entry, a_ij per backward cone
  // entry, D per backward cone position
  xcal::DBuf<double> cfa, cfr, cfd, cba, cbd;
  // Krylov workspace (allocated lazily)
  xcal::DBuf<double> work;    // solver vectors, 12 x n
  xcal::DBuf<double> kpart;   // reduction partial sums
  xcal::DBuf<char> kstate;    // device-resident Krylov scalars
  xcal::DBuf<double> ncflux;  // non-orthogonal correction per internal face
  xcal::DBuf<double> fpart;   // fused finalizes: two partial areas
  xcal::DBuf<char> fstate;    // fused finalizes: two more Krylov states
  void* kstate_host = nullptr;  // pinned copy
  int cg_batch = 0, bi_batch = 0;   // iterations of the last solve (first-batch size)
  // mixed precision: FP32 copies of the matrix and preconditioner, workspace
  xcal::DBuf<float> work32, nz32, pD32, prD32, pS32, Lnz32, Unz32, Db32, Lrd32;
  xcal::DBuf<float> cfa32, cfr32, cba32, cbd32;
  // batched solves: first-batch sizes, per-case flags, host staging
  std::vector<int> batch_pred;
  xcal::DBuf<int32_t> case_active;
  std::vector<double> case_host;
  int ir_pred[2][16] = {};  // FP32 inner solves: iterations of the last solve per sweep index
  // Iterations of the last solve of each field, by (field pointer, solve
  // kind): the first batch of iterations a solve enqueues is sized from the
  // same field's last solve, not from another field sharing the equation (p,
  // k and omega share one), whose count can differ by 10x.
  std::map<std::pair<const void*, int>, int> preds;
  int* pred_slot(const void* field, int kind) { return &preds[{field, kind}]; }
  // batched solves: instantiated HIP graphs of two Krylov iterations, by the
  // solve's launch signature (its kernels' arguments) and the iteration
  // parity at the graph's start
  std::map<std::vector<uint64_t>, std::array<hipGraphExec_t, 2>> graphs;
  ~xcal_eqn() {
    for (auto& g : graphs)
      for (hipGraphExec_t x : g.second)
        if (x) (void)hipGraphExecDestroy(x);
    if (kstate_host) (void)hipHostFree(kstate_host);
  }
};

namespace xcal {

// kernel-side descriptors (passed by value)
struct TermK {
  int type;
  int sign;
  const double* flux;
  double flux_const;
};
struct TermsK {
  int n;
  int cn_first;  // terms[0] is Time{CrankNicolson} (assembled as Time{Euler})
  TermK t[kMaxTerms];
};
struct SourcesK {
  int n;
  int sign[kMaxSources];
  const double* x[kMaxSources];
  const double* y[kMaxSources];
  const double* z[kMaxSources];
};
struct BCK {
  int kind;
  double v[3];
};
struct BCsK {
  int npatch;
  BCK p[kMaxPatches];  // indexed by patch id; kind -1 = no BC on patch
};

void make_sweep_plans(xcal_mesh* m, int small);
// meshes built on the host first (server admission), uploaded later
struct PreparedMesh;
// Host build of one mesh or a batch.  With `ingest`, the CSR pattern, slot
// maps and DILU levels are built on that context's device (unless
// XCAL_HOST_INGEST=1); `cones` allows the cone form of single meshes.
std::shared_ptr<PreparedMesh> prepare_mesh(const xcal_mesh_host* const* hosts, int ncase, bool batch,
                                           xcal_ctx* ingest = nullptr, bool cones = true);
bool host_ingest_forced();
// Matrix connectivity of one mesh in the reference numbering: CSR pattern,
// slot maps (cell-face pair, internal face's owner row, diagonal, transpose)
// and the forward/backward DILU level of every row.
struct Connectivity {
  std::vector<int32_t> rowptr, colval, diag, cf_slot, trans, fslot, flev, blev;
  int nlev_f = 0, nlev_b = 0;
  double device_ms = 0.0;  // device build: kernel time (events on the stream)
};
// A mesh's cell-face lists (0-based, validated): each cell's internal faces,
// neighbours and nsign in faces_range order, and the two cells of every face
// (boundary faces first, nbf of them).
struct CellLists {
  int64_t n = 0, nf = 0, nbf = 0;
  std::vector<int32_t> cf_ptr, cf_face, cf_nbr, own, nei;
  std::vector<int8_t> cf_sign;
};
void cell_lists(const xcal_mesh_host* h, CellLists& L);
void host_connectivity(const CellLists& L, Connectivity& K);
void device_connectivity(xcal_ctx* ctx, const CellLists& L, Connectivity& K);
std::shared_ptr<PreparedMesh> prepare_union(const std::vector<const PreparedMesh*>& parts);
xcal_mesh* upload_prepared(xcal_ctx* ctx, PreparedMesh& P);
// Serving policy of a SIMPLE object: stop a run at the first non-finite
// residual (the reference's loop would carry on); simple_diverged reports it.
void simple_stop_nonfinite(xcal_simple* S, bool on);
bool simple_diverged(const xcal_simple* S);
// Serving engine: one iteration of every case (res: Ux, Uy, Uz, p, k, omega
// residuals per case), per-case outcome, initial calculations, state
// transfer between batches and per-case read-back.
void simple_step(xcal_simple* S, std::vector<double>& res);
bool simple_case_converged(const xcal_simple* S, int k);
bool simple_case_bad(const xcal_simple* S, int k);
void simple_initialise(xcal_simple* S);
void simple_copy_case(xcal_simple* dst, int kd, const xcal_simple* src, int ks);
void simple_case_field(const xcal_simple* S, int which, int k, double* host);

// device bytes a single-case SIMPLE run on P occupies (estimate); turb adds
// the k-omega fields
int64_t prepared_footprint(const PreparedMesh& P, bool turb = false);
int64_t prepared_cells(const PreparedMesh& P);
void upload_permuted(const xcal_mesh* m, const std::vector<int32_t>& perm, double* dev,
                     const double* host, int64_t len);
// host[i] = dev[perm[i]] for i < len; dev holds dev_len entries (default len)
void download_permuted(const xcal_mesh* m, const std::vector<int32_t>& perm, const double* dev,
                       double* host, int64_t len, int64_t dev_len = -1);
TermsK make_terms(const xcal_term* terms, int n);
BCsK make_bcs(const xcal_mesh* m, const xcal_bc* bcs, int n);
xcal_field* check_field(const xcal_field* f, const xcal_mesh* m, int kind, const char* what);

// ---------------------------------------------------------------- kernels
// assembly
void discretise(xcal_eqn* e, const TermsK& t, const SourcesK& s, const xcal_field* prev, double dt);
void apply_bcs(xcal_eqn* e, const TermsK& t, const BCsK& b, const xcal_field* phi, int comp);
void relax_diagdom(xcal_eqn* e, const double* field, double* b, double alpha);
void set_reference(xcal_eqn* e, double pref, int64_t cell);
void copy_nz(xcal_mesh* m, double* dst, const double* src);
// calculate
void interpolate(const xcal_mesh* m, const double* phi, double* phif);
void correct_boundaries(const xcal_mesh* m, const BCsK& b, const double* const* phi,
                        double* const* phif, int ncomp);
void green_gauss(const xcal_mesh* m, const double* phif, double* gx, double* gy, double* gz);
void grad_midpoint(const xcal_mesh* m, const BCsK& b, const double* const* phi,
                   double* const* phif, int ncomp, double* grad);
void green_gauss_tensor(const xcal_mesh* m, const double* fx, const double* fy, const double* fz,
                        double* t0);
void divergence(const xcal_mesh* m, const double* psif, double* div);
void flux(const xcal_mesh* m, const double* fx, const double* fy, const double* fz, double* mdot);
void relax_explicit(const xcal_mesh* m, int64_t len, double* phi, const double* phi0, double alpha);
void fill(hipStream_t s, double* p, int64_t n, double v);
void inverse_diagonal(const xcal_eqn* e, double* rD);
void remove_pressure_source(xcal_eqn* e, const double* gx, const double* gy, const double* gz,
                            int sign);
void Hop(const xcal_eqn* e, const double* Ux, const double* Uy, const double* Uz, double* Hx,
         double* Hy, double* Hz);
void nonorthogonal_correction(xcal_eqn* pe, const double* gx, const double* gy, const double* gz,
                              const double* p, const double* flux);
void correct_mass_flux(const xcal_eqn* pe, const double* p, const double* pflux, int psign,
                       const BCsK& b, double* mdot);
void correct_velocity(const xcal_mesh* m, double* Ux, double* Uy, double* Uz, const double* Hx,
                      const double* Hy, const double* Hz, const double* gx, const double* gy,
                      const double* gz, const double* rD);
// k-omega model
void turb_production(const xcal_mesh* m, const double* gradU, const double* nut, double alpha1,
                     double* Pk, double* Pw);
void turb_wall_production(const xcal_mesh* m, const BCsK& bk, double nu, const double* Ux,
                          const double* Uy, const double* Uz, const double* k, double* Pk);
void turb_omega_constrain(xcal_eqn* e, const BCsK& bw, double nu, const double* k);
void turb_coeffs(const xcal_mesh* m, const double* omega, const double* nutf, double nu,
                 double beta1, double beta_star, double sigma_w, double sigma_k, double* Dw,
                 double* Dk, double* mw, double* mk);
void turb_bound(const xcal_mesh* m, double* v, double* tmp);
void turb_nut(const xcal_mesh* m, const double* k, const double* omega, double* nut);
void turb_nut_wall(const xcal_mesh* m, const BCsK& bn, double nu, const double* k, double* nutf);
void turb_nueff(const xcal_mesh* m, double nu, const double* nutf, double* nueff);
// linear algebra
void spmv(const xcal_eqn* e, const double* x, double* y);
void apply_M_levels(const xcal_eqn* e, const double* r, double* z);
void precond_update(xcal_eqn* e, int precon);
void precond_apply_staged(xcal_eqn* e, const double* r, double* z, double* Y, double* Yb);
// Resolves the current device's address of the no-allocation scratch array
// (once per device; xcal_ctx_create calls it with the device current).
void noalloc_init(int device);
void solve_prologue(xcal_eqn* e, const xcal_solver_setup* st, double* x, const double* b);
void solve_epilogue(xcal_eqn* e, double* x);
int solve_system(xcal_eqn* e, const xcal_solver_setup* s, double* x, const double* b,
                 xcal_solve_stats* st);
// batch meshes: every case solved independently; st has one entry per case
int solve_system_batch(xcal_eqn* e, const xcal_solver_setup* s, double* x, const double* b,
                       xcal_solve_stats* st);
void precond_update_batch(xcal_eqn* e);
double residual(xcal_eqn* e, const double* x, const double* b);

}  // namespace xcal
