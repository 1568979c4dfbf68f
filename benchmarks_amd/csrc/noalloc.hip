// noalloc.hip -- mfx_cg_solve_noalloc: the conjugate-gradient solve of
// mfx_cg_solve for callers that cannot allocate device or pinned host memory
// (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc).  Its work buffers live in one file-scope device array,
// `scratch`, of fixed size; each GPU the library runs on holds its own copy
// until the process ends.  Host code never uses the array's name as a
// device address: the address is looked up once per device with
// hipGetSymbolAddress, with that device current, and kept per device.
//
// Layout of scratch (doubles): r, p, Ap (kNoAllocMaxDofs each), then
// kNoAllocMaxPartials partial sums, then the CG state.  Every call on one
// GPU shares these buffers, and they keep the last call's contents; the
// vectors and partial sums are written before they are read, and the state
// is reset by a host-to-device copy at the start of every solve.  Calls that
// overlap on one GPU (two host threads, or two streams) race.
#include <mutex>

#include "op.hpp"

namespace mfx {
namespace {

constexpr uint64_t kNoAllocMaxDofs = 1ull << 21;        // 2,097,152 DoFs
constexpr uint64_t kNoAllocMaxPartials = 1ull << 17;    // partial sums of one reduction
constexpr uint64_t kStateDoubles = (sizeof(CgState) + 7) / 8;
constexpr uint64_t kScratchDoubles = 3 * kNoAllocMaxDofs + kNoAllocMaxPartials + kStateDoubles;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];

// The device address of `scratch` on the current device (looked up on the
// first call per device).
double *scratch_on_current_device() {
  int dev = -1;
  MFX_HIP(hipGetDevice(&dev));
  MFX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  hipError_t err = hipSuccess;
  std::call_once(g_scratch_once[dev], [&] {
    void *ptr = nullptr;
    err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err == hipSuccess) g_scratch_ptr[dev] = static_cast<double *>(ptr);
  });
  MFX_HIP(err);
  MFX_REQUIRE(g_scratch_ptr[dev], "no device address for the solver scratch");
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace mfx

extern "C" {

uint64_t mfx_cg_noalloc_max_dofs(void) { return mfx::kNoAllocMaxDofs; }

mfx_status mfx_cg_solve_noalloc(mfx_op *op, double *x, const double *b, const double *inv_diag,
                                int max_steps, double abs_tol, double rel_tol,
                                mfx_cg_result *res, void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && x && b && res, "null pointer argument");
    MFX_REQUIRE(max_steps >= 0, "max_steps must be >= 0");
    MFX_REQUIRE(op->N <= kNoAllocMaxDofs,
                "mfx_cg_solve_noalloc: " + std::to_string(op->N) + " DoFs exceed its limit of " +
                    std::to_string(kNoAllocMaxDofs) + " (use mfx_cg_solve)");
    const size_t npart = cg_partial_count(*op);
    MFX_REQUIRE(npart <= kNoAllocMaxPartials,
                "mfx_cg_solve_noalloc: the operator needs " + std::to_string(npart) +
                    " partial sums, more than its limit of " +
                    std::to_string(kNoAllocMaxPartials) + " (use mfx_cg_solve)");
    double *base = scratch_on_current_device();
    const uint64_t n = kNoAllocMaxDofs;
    CgWork w{base, base + n, base + 2 * n, base + 3 * n,
             reinterpret_cast<CgState *>(base + 3 * n + kNoAllocMaxPartials), nullptr};
    cg_run(*op, x, b, inv_diag, max_steps, abs_tol, rel_tol, w, res, as_stream(stream));
  });
}

}  // extern "C"
