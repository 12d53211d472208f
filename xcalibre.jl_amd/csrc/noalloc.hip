// The preconditioner apply for callers that cannot allocate device memory:
// code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.  xcal_precond_apply stages the DILU sweeps through two n-vectors
// of the equation's Krylov workspace, which it allocates on first use; this
// variant stages them in a file-scope device array instead and calls no
// allocating routine.
#include <mutex>

#include "xcal_internal.h"

namespace xcal {

// Most rows of a mesh the no-allocation apply accepts (the 1M-cell channel
// fits).
constexpr int64_t kNoAllocRows = 1 << 20;

// Two staging vectors of kNoAllocRows doubles, 16 MiB.  Every GPU the library
// runs on holds its own copy until the process ends.  Host code never uses
// this name as an address: it uses the per-device pointer from
// hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[2 * kNoAllocRows];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device
}  // namespace

void noalloc_init(int device) {
  std::lock_guard<std::mutex> lock(g_mu);
  XCAL_CHECK(device >= 0 && device < kMaxDevices, "device index beyond the scratch table");
  if (g_scratch[device]) return;
  void* p = nullptr;
  XCAL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  XCAL_CHECK(p != nullptr, "no device address for the scratch array");
  g_scratch[device] = static_cast<double*>(p);
}

static double* scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_mu);
  double* p = device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
  XCAL_CHECK(p != nullptr, "scratch array not resolved for this device (create a context first)");
  return p;
}

static const double* cell_component(const xcal_eqn* e, const xcal_field* f, int c,
                                    const char* what) {
  XCAL_CHECK(f && f->mesh == e->mesh, std::string(what) + ": field missing or on another mesh");
  XCAL_CHECK(f->kind == XCAL_CELL_SCALAR || f->kind == XCAL_CELL_VECTOR,
             std::string(what) + ": must be a cell field");
  XCAL_CHECK(c >= 0 && c < f->ncomp, std::string(what) + ": component out of range");
  return f->comp(c);
}

}  // namespace xcal

using namespace xcal;

extern "C" {

int xcal_precond_apply_noalloc(xcal_eqn* e, const xcal_field* r, int32_t rc, xcal_field* z,
                               int32_t zc) {
  return guarded([&] {
    XCAL_CHECK(e, "equation is NULL");
    const double* rp = cell_component(e, r, rc, "r");
    double* zp = const_cast<double*>(cell_component(e, z, zc, "z"));
    XCAL_CHECK(rp != zp, "r and z must not alias");
    const int64_t n = e->mesh->n;
    XCAL_CHECK(n <= kNoAllocRows,
               "mesh has " + std::to_string(n) + " rows; xcal_precond_apply_noalloc takes at most " +
                   std::to_string(kNoAllocRows));
    double* Y = scratch_of(e->mesh->ctx->device);
    XCAL_HIP(hipSetDevice(e->mesh->ctx->device));
    precond_apply_staged(e, rp, zp, Y, Y + kNoAllocRows);
    return 0;
  });
}

}  // extern "C"
