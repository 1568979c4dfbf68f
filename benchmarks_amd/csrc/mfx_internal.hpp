// mfx_internal.hpp -- shared definitions of the gfx950 library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/mfx.h"

namespace mfx {

// Thread-local last error (mfx_last_error).
void set_error(const std::string &msg);

struct Error {
  mfx_status status;
  std::string msg;
};

#define MFX_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw ::mfx::Error{e_ == hipErrorOutOfMemory ? MFX_ERR_OOM : MFX_ERR_HIP, \
                         std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define MFX_REQUIRE(cond, msg)                                              \
  do {                                                                      \
    if (!(cond)) throw ::mfx::Error{MFX_ERR_INVALID_ARG, (msg)};            \
  } while (0)

constexpr uint32_t kInvalid = MFX_INVALID_DOF;

// Launch wrappers implemented in the .hip translation units.  Each checks
// its arguments against what the kernel assumes and throws mfx::Error.
template <typename T>
void launch_apply(int p, int nq, bool collocation, const T *B, const T *D,
                  const T *G, const T *src, T *dst, const uint32_t *map,
                  uint32_t n_cells, hipStream_t s, double *dot_partials = nullptr,
                  const int *skip = nullptr, uint32_t row_cells = 0);
// Non-collocated (Gauss) apply with the metric terms of the trilinear cells
// computed per quadrature point from their vertices (cellX [E][8][3]), xq and
// wq the 1D quadrature points and weights on [0,1].  row_cells > 0 (both
// Gauss applies): the cells form x-rows of row_cells cells numbered x
// fastest, and the map gives neighbours along x their shared face nodes (a
// box mesh); the scatter then sums shared x-faces in LDS first.
// eo: the even-odd tables of B, B^T, D, D^T (tables.hpp eo_tables) for the
// line-form contractions.
// axes: the cells are those of an undeformed box (Jacobian columns along x,
// y and z; row_cells > 0), whose diagonal metric is formed once per column.
template <typename T>
void launch_apply_otf(int p, int nq, const T *B, const T *D, const T *eo, const T *xq,
                      const T *wq, const double *cellX, const T *src, T *dst,
                      const uint32_t *map, uint32_t n_cells, hipStream_t s,
                      double *dot_partials = nullptr, const int *skip = nullptr,
                      uint32_t row_cells = 0, bool axes = false);
template <typename T>
void launch_evec(int p, int nq, bool collocation, const T *B, const T *D,
                 const T *G, const T *in, T *out, uint32_t n_cells,
                 hipStream_t s);
// Tile variants of the FP64 collocation apply (degrees 2..6) for the
// autotuner: variant v in [0, colloc_variant_count(p)); dst += A src over the
// cells; with inst_done set, cells of instances flagged done are skipped.
// epb / prefetch (optional) receive the variant's tile; src == nullptr only
// describes the variant.
int colloc_variant_count(int p);
int colloc_default_variant(int p);
void launch_colloc_variant(int p, int v, const double *D, const double *G, const double *src,
                           double *dst, const uint32_t *map, uint32_t n_cells,
                           const uint16_t *cell_inst, const int *inst_done, hipStream_t s,
                           int *epb = nullptr, int *prefetch = nullptr);
// number of partial sums written by launch_apply(..., dot_partials)
uint32_t apply_n_blocks(int p, int nq, bool collocation, uint32_t n_cells,
                        uint32_t row_cells = 0);

// set-up kernels (setup.hip)
struct BoxGeom {
  uint32_t n[3];
  double lo[3], hi[3];
  double deform;
};
void launch_box_cell_vertices(const BoxGeom &box, double *cellX, hipStream_t s);
void launch_box_map(const BoxGeom &box, int p, uint32_t *map, hipStream_t s);
void launch_geometry(uint32_t E, const double *cellX, int nq, const double *xq,
                     const double *wq, double *G, hipStream_t s);
void launch_rhs(uint32_t E, int p, int nq, const double *B, const double *xq,
                const double *wq, const double *cellX, const uint32_t *map,
                int f_kind, double *b, hipStream_t s);
void launch_diagonal(uint32_t E, int p, int nq, const double *B,
                     const double *D, const double *G, const uint32_t *map,
                     double *diag, hipStream_t s);
void launch_dof_coordinates(uint32_t E, int p, const double *support,
                            const double *cellX, const uint32_t *map,
                            double *coords, hipStream_t s);
void launch_to_float(const double *in, float *out, uint64_t n, hipStream_t s);

// vector kernels (vector.hip)
void launch_fill(double *x, double v, uint64_t n, hipStream_t s);
void launch_fill_f32(float *x, float v, uint64_t n, hipStream_t s);
template <typename T>
void launch_copy_listed(T *dst, const T *src, const uint32_t *list, uint64_t n,
                        hipStream_t s);
void launch_set_listed(double *dst, double v, const uint32_t *list, uint64_t n,
                       hipStream_t s);
void launch_gather(const uint32_t *map, uint64_t n, const double *src, double *evec,
                   hipStream_t s);
void launch_scatter_add(const uint32_t *map, uint64_t n, const double *evec, double *dst,
                        hipStream_t s);
void launch_invert_diagonal(double *d, uint64_t n, hipStream_t s);
void launch_random(double *x, uint64_t n, uint64_t seed, hipStream_t s);
// sum of up to kMax sine modes at n points (mfx_vec_sine_modes), passed by value
struct SineModes {
  static constexpr int kMax = 64;
  int n;
  int k[3 * kMax];
  double a[kMax];
};
void launch_sine_modes(const double *xyz, uint64_t n, const SineModes &m, double *out,
                       hipStream_t s);
// deterministic two-stage dot: partials[nb] then final sum into *result
uint32_t dot_n_blocks(uint64_t n);
void launch_dot(const double *x, const double *y, uint64_t n, double *partials,
                double *result, hipStream_t s);

// 1D rules and basis tables on [0,1] (basis.cpp)
void gauss_rule(int n, double *x, double *w);
void gll_rule(int n, double *x, double *w);
void basis_tables(int p, int nq, int quad, double *B, double *D, double *xq,
                  double *wq, double *support);
// values at x in [0,1] of the p+1 nodal basis functions on the Gauss-Lobatto
// support points (the FE_Q basis of degree p)
void lagrange_values(int p, double x, double *values);

}  // namespace mfx
