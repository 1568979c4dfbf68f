// op.hpp -- internal definition of the operator handle and the helpers the
// C-ABI translation units share (capi.cpp, mg.cpp).
#pragma once

#include <atomic>
#include <memory>
#include <new>
#include <string>

#include "cg_state.hpp"
#include "mfx_internal.hpp"
#include "patch.hpp"

namespace mfx {

// Runs f, converting exceptions to a status and the thread-local message.
template <typename F>
inline mfx_status guard(F &&f) {
  try {
    f();
    return MFX_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.status;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return MFX_ERR_OOM;
  } catch (const std::exception &e) {
    set_error(e.what());
    return MFX_ERR_HIP;
  }
}

// Owning device buffer.
template <typename T>
struct DBuf {
  T *ptr = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : ptr(o.ptr), n(o.n) {
    o.ptr = nullptr;
    o.n = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      reset();
      ptr = o.ptr;
      n = o.n;
      o.ptr = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    reset();
    if (count == 0) return;
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess)
      throw Error{e == hipErrorOutOfMemory ? MFX_ERR_OOM : MFX_ERR_HIP,
                  "hipMalloc(" + std::to_string(count * sizeof(T)) + " B): " +
                      hipGetErrorString(e)};
    ptr = static_cast<T *>(p);
    n = count;
  }
  void upload(const T *host, size_t count) {
    alloc(count);
    if (count) MFX_HIP(hipMemcpy(ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
  }
};

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Process-wide creation count of operators.  Caches that belong to one
// operator key on this serial, never on its address: a destroyed
// operator's heap address can be reused by the next one created.
inline uint64_t next_op_serial() {
  static std::atomic<uint64_t> count{0};
  return ++count;
}

}  // namespace mfx

using namespace mfx;

struct mfx_op {
  const uint64_t serial = mfx::next_op_serial();
  // set for operators made by mfx_op_create_box: the mesh they were built on
  bool is_box = false;
  mfx::BoxGeom box{};
  int p = 0, nq = 0, quad = 0, geometry = 0;
  bool colloc = false;
  uint32_t E = 0;
  uint64_t N = 0;
  uint64_t n_cons = 0;
  DBuf<uint32_t> map, cons;
  DBuf<double> G, B, D, xq, wq, support, cellX;
  DBuf<float> G32, B32, D32, xq32, wq32;
  // Gauss quadrature: the even-odd tables of B, B^T, D and D^T (eo_tab
  // layout, in that order) for the line-form contractions
  DBuf<double> eo;
  DBuf<float> eo32;
  // patch-of-cells kernels (box meshes, GLL collocation)
  bool patch = false;
  PatchLayout lay{};
  PatchShape shape{};
  PatchGrid grid64{}, grid32{};
  DBuf<uint32_t> pmap;
  DBuf<double> Gp;   // stored geometry: G in the patch kernel's order
  DBuf<float> Gp32;
  // deterministic scatter (mfx_op_set_deterministic): per-patch sums
  bool det = false;
  DBuf<double> pbuf64;
  DBuf<float> pbuf32;
  // CG scratch, allocated on first use
  DBuf<double> r, pdir, ap, part;
  DBuf<CgState> st;
  int *h_done = nullptr;  // pinned
  ~mfx_op() {
    if (h_done) (void)hipHostFree(h_done);
  }

  PatchOp<double> patch_op64() const {
    PatchOp<double> o;
    o.p = p;
    o.shape = shape;
    o.lay = lay;
    o.grid = grid64;
    o.D = D.ptr;
    o.w = wq.ptr;
    o.xq = xq.ptr;
    o.cellX = cellX.ptr;
    o.G = geometry == MFX_GEOM_STORED ? Gp.ptr : nullptr;
    o.pmap = pmap.ptr;
    o.pbuf = det ? pbuf64.ptr : nullptr;
    return o;
  }
  PatchOp<float> patch_op32() const {
    PatchOp<float> o;
    o.p = p;
    o.shape = shape;
    o.lay = lay;
    o.grid = grid32;
    o.D = D32.ptr;
    o.w = wq32.ptr;
    o.xq = xq32.ptr;
    o.cellX = cellX.ptr;
    o.G = geometry == MFX_GEOM_STORED ? Gp32.ptr : nullptr;
    o.pmap = pmap.ptr;
    o.pbuf = det ? pbuf32.ptr : nullptr;
    return o;
  }
  // Gauss applies on a box: x-rows of cells (their shared faces are summed
  // before the scatter); 0 for other meshes and collocation
  uint32_t row_cells() const { return is_box && !colloc ? box.n[0] : 0; }
  // an undeformed box: every cell's Jacobian columns lie along x, y and z
  bool axis_cells() const { return is_box && !colloc && box.deform == 0.0; }
  uint32_t apply_blocks() const {
    return patch ? (det ? grid64.det_dot : grid64.dot)
                 : apply_n_blocks(p, nq, colloc, E, row_cells());
  }
  // dst += A_free src (no zeroing, no constrained rows)
  void apply64(const double *src, double *dst, hipStream_t s, double *partials = nullptr,
               const int *skip = nullptr) const {
    if (patch)
      launch_patch<double>(patch_op64(), src, dst, s, partials, skip);
    else if (geometry == MFX_GEOM_ON_THE_FLY)
      launch_apply_otf<double>(p, nq, B.ptr, D.ptr, eo.ptr, xq.ptr, wq.ptr, cellX.ptr, src, dst,
                               map.ptr, E, s, partials, skip, row_cells(), axis_cells());
    else
      launch_apply<double>(p, nq, colloc, B.ptr, D.ptr, G.ptr, src, dst, map.ptr, E, s, partials,
                           skip, row_cells());
  }
  void apply32(const float *src, float *dst, hipStream_t s) const {
    if (patch)
      launch_patch<float>(patch_op32(), src, dst, s);
    else if (geometry == MFX_GEOM_ON_THE_FLY)
      launch_apply_otf<float>(p, nq, B32.ptr, D32.ptr, eo32.ptr, xq32.ptr, wq32.ptr, cellX.ptr,
                              src, dst, map.ptr, E, s, nullptr, nullptr, row_cells(),
                              axis_cells());
    else
      launch_apply<float>(p, nq, colloc, B32.ptr, D32.ptr, G32.ptr, src, dst, map.ptr, E, s,
                          nullptr, nullptr, row_cells());
  }
};

namespace mfx {

// Work buffers of one CG solve: three N-vectors, cg_partial_count(op)
// partial sums, the device state, and a host word for the `done` flag
// (pinned, or nullptr to read it into pageable host memory).
struct CgWork {
  double *r, *p, *ap, *part;
  CgState *st;
  int *h_done;
};
size_t cg_partial_count(const mfx_op &op);
// The CG iteration of mfx_cg_solve on the given work buffers (capi.cpp).
void cg_run(mfx_op &op, double *x, const double *b, const double *inv_diag, int max_steps,
            double abs_tol, double rel_tol, const CgWork &w, mfx_cg_result *res, hipStream_t s);

}  // namespace mfx
