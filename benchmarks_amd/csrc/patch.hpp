// patch.hpp -- data of the patch-of-cells operator kernels (patch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfx {

struct PatchShape {
  int bx, by, bz;  // cells per patch and direction
};

struct PatchLayout {
  uint32_t n[3];   // cells per direction
  uint32_t np[3];  // patches per direction
};

// Workgroups of the persistent patch kernel: as many as the device holds at
// once (resident workgroups per CU x CUs), at most one per patch.  `dot` is
// the grid of the variant that also returns per-workgroup p.Ap partials; the
// det_ grids are those of the deterministic-scatter variants.
struct PatchGrid {
  uint32_t plain = 0, dot = 0, det_plain = 0, det_dot = 0;
};

// Device pointers and parameters one patch-kernel launch needs.
template <typename T>
struct PatchOp {
  int p = 0;
  PatchShape shape{};            // cells per patch the map was built for
  PatchLayout lay{};
  PatchGrid grid{};
  const T *D = nullptr;          // collocation derivative [nq][nq]
  const T *w = nullptr;          // 1D quadrature weights [nq]
  const T *xq = nullptr;         // 1D quadrature points [nq]
  const double *cellX = nullptr; // cell vertices [n_cells][8][3], lexicographic cells
  const T *G = nullptr;          // stored metric terms in patch order (nullptr: from cellX)
  const uint32_t *pmap = nullptr;
  // deterministic scatter: per-patch sums [n_patches][P3], added per node in
  // patch order by a second kernel (nullptr: atomic scatter)
  T *pbuf = nullptr;
};

// Patches of many box operators of one degree in one launch (batched
// solves): the patch maps of the instances concatenated in patch order, with
// the instances' DoF offsets applied; per patch its instance and patch
// coordinates; per instance its cell counts and first cell in the
// concatenated vertex array.  Patches of instances whose `done` flag is set
// are skipped.
struct PatchBatchOp {
  int p = 0;
  uint32_t n_patches = 0;
  const uint32_t *slot = nullptr;  // per instance: its p.Ap slot among this degree's instances
  uint32_t grid = 0, grid_dot = 0;  // without / with the p.Ap partials
  const uint32_t *pmap = nullptr;  // [n_patches][P3]
  const uint4 *meta = nullptr;     // per patch: instance, px, py, pz
  const uint4 *inst = nullptr;     // per instance: n0, n1, n2, first cell
  const double *cellX = nullptr;   // [cells][8][3]
  const double *w = nullptr, *xq = nullptr;
};
uint32_t patch_batch_grid(int p, uint32_t n_patches, bool dot);
// waves per workgroup of the batch kernel: p.Ap partials per patch
uint32_t patch_batch_waves(int p);
// dst += A src over the batch's patches (FP64, no zeroing).  With `partials`
// (one slot of grid_dot x patch_batch_waves(p) doubles per instance of this
// degree, slot[instance] its index; zero on entry) every workgroup also
// stores, per instance and wave, its sum of src_i (A src)_i over the DoFs of
// its patches of that instance, and leaves the slots of instances it has no
// patch of untouched: each slot, summed, is src.(A src) of that instance's
// free rows.
void launch_patch_batch(const PatchBatchOp &o, const double *src, double *dst, const int *done,
                        hipStream_t s, double *partials = nullptr);

PatchShape patch_shape(int p);
uint32_t patch_count(const PatchLayout &lay);
// grids of the kernels with the metric from the vertices or (stored) from G
template <typename T>
PatchGrid patch_grid(int p, const PatchLayout &lay, bool stored);
template <typename T>
void launch_patch(const PatchOp<T> &o, const T *src, T *dst, hipStream_t s,
                  double *partials = nullptr, const int *skip = nullptr);
void launch_patch_map(const PatchLayout &lay, int p, PatchShape sh, uint32_t *pmap, hipStream_t s);
// The stored metric terms G [cell][6][nq^3] permuted into the patch order the
// kernel with stored geometry reads (patch_geometry_size elements).
uint64_t patch_geometry_size(int p, const PatchLayout &lay);
template <typename T>
void launch_patch_geometry(int p, const PatchLayout &lay, const T *G, T *Gp, hipStream_t s);

}  // namespace mfx
