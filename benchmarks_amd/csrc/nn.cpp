// nn.cpp -- the learned preconditioner (mfx_nn_* in include/mfx.h): a 4-level
// 3D U-Net on the nodal grid of a box operator that maps a residual to an
// approximation of the error A^-1 r, trained on the GPU by this library's
// own forward (BF16 MFMA) and backward (FP32) kernels, and used inside
// flexible CG.
//
//   e0 = relu(conv_a(x))                 1 -> 16, level 0
//   e1 = relu(conv_b(e0, stride 2))     16 -> 32, level 1
//   e2 = relu(conv_c(e1, stride 2))     32 -> 64, level 2
//   e3 = relu(conv_d(e2, stride 2))     64 -> 64, level 3
//   d2 = relu(conv_e(up(e3)) + e2)      64 -> 64, level 2
//   d1 = relu(conv_f(up(d2)) + e1)      64 -> 32, level 1
//   d0 = relu(conv_g(up(d1)) + e0)      32 -> 16, level 0
//   y  = conv_h(d0)                     16 -> 1,  level 0
//
// 3^3 kernels, no biases (the network is positively homogeneous), levels of
// ceil(n/2) nodes, nearest up-sampling.  Master weights are FP32 [o][t][c];
// the forward pass uses their BF16 rounding, the backward pass the FP32
// value of that rounding, so the gradient is the gradient of the network
// that ran.
//
// As a preconditioner (FP32 inside, FP64 in and out, constrained rows passed
// through), with s the output scale fixed at training time:
//   smoother degree 0:  z = ||r|| s y(r / ||r||);
//   smoother degree k:  a two-level cycle with the network as the coarse
//     correction -- k Chebyshev(Jacobi) steps on A z = r from zero, the
//     network applied to the remaining residual rho, c = ||rho|| s
//     y(rho / ||rho||), added as z += omega c with omega = c.rho / c.Ac (the
//     energy-optimal step along c: the correction never increases the
//     energy norm of the error), then k Chebyshev steps from z.  Without the
//     network the same cycle is a plain Chebyshev polynomial preconditioner.
//     With a coarse operator set (mfx_nn_set_coarse) the network runs on its
//     node grid: rho is restricted (R = P^T of the nodal interpolation
//     between the two box operators), the network's output prolongated.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "mg.hpp"
#include "nn.hpp"
#include "op.hpp"
#include "transfer.hpp"

namespace mfx {
namespace {

struct LayerDesc {
  int cin, cout;
};
constexpr LayerDesc kLayers[8] = {{1, 16}, {16, 32}, {32, 64}, {64, 64},
                                  {64, 64}, {64, 32}, {32, 16}, {16, 1}};

uint64_t param_offset(int l) {
  uint64_t o = 0;
  for (int k = 0; k < l; ++k) o += 27ull * kLayers[k].cin * kLayers[k].cout;
  return o;
}
uint64_t n_params_total() { return param_offset(8); }
int kpad(int l) { return (27 * kLayers[l].cin + 31) / 32 * 32; }

void coarse(const uint32_t *f, uint32_t *c) {
  for (int k = 0; k < 3; ++k) c[k] = (f[k] + 1) / 2;
}
uint64_t count(const uint32_t *d) { return (uint64_t)d[0] * d[1] * d[2]; }

double dot64(const double *x, const double *y, uint64_t n, DBuf<double> &part, hipStream_t s) {
  const uint32_t nb = dot_n_blocks(n);
  if (part.n < nb + 1) part.alloc(nb + 1);
  launch_dot(x, y, n, part.ptr, part.ptr + nb, s);
  double h = 0.0;
  MFX_HIP(hipMemcpyAsync(&h, part.ptr + nb, sizeof(double), hipMemcpyDeviceToHost, s));
  MFX_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace
}  // namespace mfx

using namespace mfx;

struct mfx_nn {
  DBuf<float> w, m, v, grad, wr;
  DBuf<uint16_t> wb;
  uint64_t wb_off[9] = {};
  int step = 0;
  double out_scale = 1.0;
  bool packed = false;
  // workspace for one grid size
  uint32_t lv[4][3] = {};
  DBuf<uint16_t> e0, e1, e2, e3, d2, d1, d0;
  DBuf<float> x32, y32, dy;
  DBuf<float> gd0, gu0, gd1, gu1, gd2, gu2, ge3, t2, ge2, t1, ge1, t0, ge0, part;
  DBuf<double> lpart, dpart;
  DBuf<double> fr, fp, fq, fz;  // flexible CG
  // two-level mode: Chebyshev smoothing around the network
  int smooth_degree = 0;
  bool use_net = true;
  uint64_t sm_serial = 0;  // the operator the smoother is set up for (0: none)
  uint64_t sm_n = 0;
  double lmax = 0.0, lmin = 0.0;
  DBuf<float> dinv32, r32, z32, t32, c32, n32, an32;
  DBuf<double> tmp64, tmp64b, tmp64c;  // sized by setup_smoother only, for its operator
  DBuf<double> en64, en64b, en64c;     // energy-loss training, sized for its operator
  double last_omega = 0.0;
  // coarse correction: the network on another box operator's node grid
  mfx_op *coarse_op = nullptr;
  // the operator pair the transfer and its scratch were built for (0: none)
  uint64_t tr_fine_serial = 0, tr_coarse_serial = 0;
  Transfer tr;
  DBuf<float> rc32, ec32, ta32, tb32;
  DBuf<double> rc64;
  // fallback of the flexible CG: one V-cycle of this multigrid per step once
  // the learned cycle stagnates (the caller keeps it alive)
  mfx_mg *fb_mg = nullptr;
  int fb_window = 8;
  double fb_min_reduction = 2.0;
  int fb_step = -1;  // step at which the last solve switched to it (-1: it did not)

  void setup_smoother(mfx_op &op, hipStream_t s) {
    if (sm_serial == op.serial && sm_n == op.N && tmp64.n == op.N && tmp64c.n == op.N) return;
    MFX_REQUIRE(op.D32.ptr, "the smoothed preconditioner needs an operator made with with_fp32");
    const uint64_t N = op.N;
    tmp64.alloc(N);
    const mfx_status st = mfx_op_inverse_diagonal(&op, tmp64.ptr, s);
    if (st != MFX_OK) throw Error{st, mfx_last_error()};
    lmax = 1.2 * estimate_lambda_max(op, tmp64.ptr, 12, s);
    lmin = lmax / 15.0;
    dinv32.alloc(N);
    launch_to_float(tmp64.ptr, dinv32.ptr, N, s);
    r32.alloc(N);
    z32.alloc(N);
    t32.alloc(N);
    c32.alloc(N);
    n32.alloc(N);
    an32.alloc(N);
    tmp64b.alloc(N);
    tmp64c.alloc(N);
    sm_serial = op.serial;
    sm_n = N;
  }

  void pack(hipStream_t s) {
    if (packed) return;
    for (int l = 0; l < 8; ++l)
      nn_pack_weights(w.ptr + param_offset(l), wb.ptr + wb_off[l], wr.ptr + param_offset(l),
                      kLayers[l].cin, kLayers[l].cout, kpad(l), s);
    packed = true;
  }

  void ensure(const uint32_t *d, bool training) {
    const bool same = lv[0][0] == d[0] && lv[0][1] == d[1] && lv[0][2] == d[2];
    if (!same) {
      for (int k = 0; k < 3; ++k) lv[0][k] = d[k];
      for (int l = 1; l < 4; ++l) coarse(lv[l - 1], lv[l]);
      const uint64_t P0 = count(lv[0]), P1 = count(lv[1]), P2 = count(lv[2]), P3 = count(lv[3]);
      e0.alloc(P0 * 16);
      e1.alloc(P1 * 32);
      e2.alloc(P2 * 64);
      e3.alloc(P3 * 64);
      d2.alloc(P2 * 64);
      d1.alloc(P1 * 32);
      d0.alloc(P0 * 16);
      x32.alloc(P0);
      y32.alloc(P0);
      gd0.reset();  // training buffers follow the grid
    }
    if (training && !gd0.ptr) {
      const uint64_t P0 = count(lv[0]), P1 = count(lv[1]), P2 = count(lv[2]), P3 = count(lv[3]);
      dy.alloc(P0);
      gd0.alloc(P0 * 16);
      gu0.alloc(P0 * 32);
      gd1.alloc(P1 * 32);
      gu1.alloc(P1 * 64);
      gd2.alloc(P2 * 64);
      gu2.alloc(P2 * 64);
      ge3.alloc(P3 * 64);
      t2.alloc(P2 * 64);
      ge2.alloc(P2 * 64);
      t1.alloc(P1 * 32);
      ge1.alloc(P1 * 32);
      t0.alloc(P0 * 16);
      ge0.alloc(P0 * 16);
      if (!part.ptr) part.alloc(64ull * 27 * 64 * 64);
    }
  }

  const uint16_t *wbl(int l) const { return wb.ptr + wb_off[l]; }
  const float *wrl(int l) const { return wr.ptr + param_offset(l); }
  float *gl(int l) { return grad.ptr + param_offset(l); }

  // y32 = network(x32) on the current grid
  void forward(hipStream_t s) {
    pack(s);
    nn_conv_forward(0, nullptr, x32.ptr, nullptr, wrl(0), nullptr, e0.ptr, nullptr, lv[0], lv[0], s);
    nn_conv_forward(1, e0.ptr, nullptr, wbl(1), nullptr, nullptr, e1.ptr, nullptr, lv[1], lv[0], s);
    nn_conv_forward(2, e1.ptr, nullptr, wbl(2), nullptr, nullptr, e2.ptr, nullptr, lv[2], lv[1], s);
    nn_conv_forward(3, e2.ptr, nullptr, wbl(3), nullptr, nullptr, e3.ptr, nullptr, lv[3], lv[2], s);
    nn_conv_forward(4, e3.ptr, nullptr, wbl(4), nullptr, e2.ptr, d2.ptr, nullptr, lv[2], lv[3], s);
    nn_conv_forward(5, d2.ptr, nullptr, wbl(5), nullptr, e1.ptr, d1.ptr, nullptr, lv[1], lv[2], s);
    nn_conv_forward(6, d1.ptr, nullptr, wbl(6), nullptr, e0.ptr, d0.ptr, nullptr, lv[0], lv[1], s);
    nn_conv_forward(7, d0.ptr, nullptr, nullptr, wrl(7), nullptr, nullptr, y32.ptr, lv[0], lv[0], s);
  }

  static uint32_t parts_for(const uint32_t *d) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(64, count(d) / 2048));
  }

  // grad += dL/dW for dy = dL/dy (after forward on the same x32)
  void backward(hipStream_t s) {
    const uint64_t P0 = count(lv[0]), P1 = count(lv[1]), P2 = count(lv[2]);
    nn_conv_wgrad(1, false, d0.ptr, nullptr, dy.ptr, part.ptr, gl(7), lv[0], lv[0], 16, 1,
                  parts_for(lv[0]), s);
    nn_head_backward(dy.ptr, wrl(7), d0.ptr, gd0.ptr, lv[0], s);
    // layer g: conv over up(d1) on level 0
    nn_conv_wgrad(1, true, d1.ptr, nullptr, gd0.ptr, part.ptr, gl(6), lv[0], lv[1], 32, 16,
                  parts_for(lv[0]), s);
    nn_conv_adjoint(1, gd0.ptr, wrl(6), gu0.ptr, lv[0], lv[0], 32, 16, s);
    nn_up_adjoint_mask(gu0.ptr, lv[0], nullptr, d1.ptr, gd1.ptr, lv[1], 32, s);
    // layer f: conv over up(d2) on level 1
    nn_conv_wgrad(1, true, d2.ptr, nullptr, gd1.ptr, part.ptr, gl(5), lv[1], lv[2], 64, 32,
                  parts_for(lv[1]), s);
    nn_conv_adjoint(1, gd1.ptr, wrl(5), gu1.ptr, lv[1], lv[1], 64, 32, s);
    nn_up_adjoint_mask(gu1.ptr, lv[1], nullptr, d2.ptr, gd2.ptr, lv[2], 64, s);
    // layer e: conv over up(e3) on level 2
    nn_conv_wgrad(1, true, e3.ptr, nullptr, gd2.ptr, part.ptr, gl(4), lv[2], lv[3], 64, 64,
                  parts_for(lv[2]), s);
    nn_conv_adjoint(1, gd2.ptr, wrl(4), gu2.ptr, lv[2], lv[2], 64, 64, s);
    nn_up_adjoint_mask(gu2.ptr, lv[2], nullptr, e3.ptr, ge3.ptr, lv[3], 64, s);
    // layer d: stride-2 conv of e2
    nn_conv_wgrad(2, false, e2.ptr, nullptr, ge3.ptr, part.ptr, gl(3), lv[3], lv[2], 64, 64,
                  parts_for(lv[3]), s);
    nn_conv_adjoint(2, ge3.ptr, wrl(3), t2.ptr, lv[3], lv[2], 64, 64, s);
    nn_add_mask(gd2.ptr, t2.ptr, e2.ptr, ge2.ptr, P2 * 64, s);
    // layer c: stride-2 conv of e1
    nn_conv_wgrad(2, false, e1.ptr, nullptr, ge2.ptr, part.ptr, gl(2), lv[2], lv[1], 32, 64,
                  parts_for(lv[2]), s);
    nn_conv_adjoint(2, ge2.ptr, wrl(2), t1.ptr, lv[2], lv[1], 32, 64, s);
    nn_add_mask(gd1.ptr, t1.ptr, e1.ptr, ge1.ptr, P1 * 32, s);
    // layer b: stride-2 conv of e0
    nn_conv_wgrad(2, false, e0.ptr, nullptr, ge1.ptr, part.ptr, gl(1), lv[1], lv[0], 16, 32,
                  parts_for(lv[1]), s);
    nn_conv_adjoint(2, ge1.ptr, wrl(1), t0.ptr, lv[1], lv[0], 16, 32, s);
    nn_add_mask(gd0.ptr, t0.ptr, e0.ptr, ge0.ptr, P0 * 16, s);
    // layer a: the FP32 input
    nn_conv_wgrad(1, false, nullptr, x32.ptr, ge0.ptr, part.ptr, gl(0), lv[0], lv[0], 1, 16,
                  parts_for(lv[0]), s);
  }

  // z = ||r|| s y(r / ||r||) on a box operator's node grid; constrained rows
  // pass through
  void precondition(mfx_op &op, double *z, const double *r, hipStream_t s) {
    if (smooth_degree > 0) return precondition_two_level(op, z, r, s);
    uint32_t d[3];
    for (int k = 0; k < 3; ++k) d[k] = op.box.n[k] * op.p + 1;
    ensure(d, false);
    const uint64_t N = op.N;
    const double rn = std::sqrt(dot64(r, r, N, dpart, s));
    if (!(rn > 0.0)) {
      MFX_HIP(hipMemcpyAsync(z, r, N * sizeof(double), hipMemcpyDeviceToDevice, s));
      return;
    }
    nn_scale_to_f32(r, x32.ptr, 1.0 / rn, N, s);
    forward(s);
    nn_scale_to_f64(y32.ptr, z, rn * out_scale, N, s);
    launch_copy_listed<double>(z, r, op.cons.ptr, op.n_cons, s);
  }

  // z32 += omega P c with c = ||rho_c|| s net(rho_c / ||rho_c||) on the
  // coarse node grid, rho_c = R rho (rho in t32), omega energy-optimal
  // rc32 = R t32 with the coarse constrained rows zeroed, rc64 its FP64
  // copy: the residual the network sees on the coarse grid (transfer and
  // scratch built for the operator pair on first use)
  void restrict_residual(mfx_op &op, hipStream_t s) {
    mfx_op &C = *coarse_op;
    const uint64_t N = op.N, Nc = C.N;
    if (tr_fine_serial != op.serial || tr_coarse_serial != C.serial || ta32.n != N ||
        tb32.n != N || rc32.n != Nc) {
      build_transfer(op, C, tr, true, false);
      rc32.alloc(Nc);
      ec32.alloc(Nc);
      rc64.alloc(Nc);
      ta32.alloc(N);
      tb32.alloc(N);
      tr_fine_serial = op.serial;
      tr_coarse_serial = C.serial;
    }
    // the smoother's vectors (t32, n32, an32) were set up for this operator
    MFX_REQUIRE(sm_serial == op.serial && t32.n == N && n32.n == N && an32.n == N,
                "coarse correction: smoother scratch does not match the fine operator");
    MFX_REQUIRE(count(tr.fine) == N && count(tr.coarse) == Nc,
                "coarse correction: transfer does not match the operators");
    restrict_<float>(tr.R32, tr.fine, tr.coarse, t32.ptr, ta32.ptr, tb32.ptr, rc32.ptr, s);
    launch_set_listed_f32(rc32.ptr, 0.0f, C.cons.ptr, C.n_cons, s);
    launch_to_double(rc32.ptr, rc64.ptr, Nc, s);
  }

  // r32 = r (constrained rows 0), z32 = the pre-smoothing from zero, and
  // t32 = r32 - A z32: the first half of the two-level cycle
  void presmooth_residual(mfx_op &op, const double *r, hipStream_t s) {
    const uint64_t N = op.N;
    launch_to_float(r, r32.ptr, N, s);
    launch_set_listed_f32(r32.ptr, 0.0f, op.cons.ptr, op.n_cons, s);
    chebyshev(op, z32.ptr, r32.ptr, dinv32.ptr, t32.ptr, c32.ptr, smooth_degree, lmin, lmax, true, s);
    vmult32(op, t32.ptr, z32.ptr, s);
    launch_residual_f32(t32.ptr, r32.ptr, t32.ptr, N, s);
  }

  // z32 += omega P c with c = ||rho_c|| s net(rho_c / ||rho_c||) on the
  // coarse node grid, rho_c = R rho (rho in t32), omega energy-optimal
  void coarse_correction(mfx_op &op, hipStream_t s) {
    mfx_op &C = *coarse_op;
    const uint64_t N = op.N, Nc = C.N;
    restrict_residual(op, s);
    const double rn = std::sqrt(dot64(rc64.ptr, rc64.ptr, Nc, dpart, s));
    if (!(rn > 0.0)) return;
    uint32_t d[3];
    for (int k = 0; k < 3; ++k) d[k] = C.box.n[k] * C.p + 1;
    ensure(d, false);
    nn_scale_to_f32(rc64.ptr, x32.ptr, 1.0 / rn, Nc, s);
    forward(s);
    launch_fill_f32(ec32.ptr, 0.0f, Nc, s);
    nn_axpy_f32(ec32.ptr, (float)(rn * out_scale), y32.ptr, Nc, s);
    launch_set_listed_f32(ec32.ptr, 0.0f, C.cons.ptr, C.n_cons, s);
    prolongate<float>(tr.P32, tr.fine, tr.coarse, ec32.ptr, ta32.ptr, tb32.ptr, n32.ptr, false, s);
    launch_set_listed_f32(n32.ptr, 0.0f, op.cons.ptr, op.n_cons, s);
    vmult32(op, an32.ptr, n32.ptr, s);
    launch_to_double(t32.ptr, tmp64.ptr, N, s);
    launch_to_double(n32.ptr, tmp64b.ptr, N, s);
    launch_to_double(an32.ptr, tmp64c.ptr, N, s);
    const double crho = dot64(tmp64b.ptr, tmp64.ptr, N, dpart, s);
    const double cac = dot64(tmp64b.ptr, tmp64c.ptr, N, dpart, s);
    last_omega = cac > 0.0 ? crho / cac : 0.0;
    if (last_omega != 0.0) nn_axpy_f32(z32.ptr, (float)last_omega, n32.ptr, N, s);
  }

  void precondition_two_level(mfx_op &op, double *z, const double *r, hipStream_t s) {
    setup_smoother(op, s);
    last_omega = 0.0;  // stays 0 when the network adds nothing to this application
    const uint64_t N = op.N;
    if (use_net && coarse_op) {
      presmooth_residual(op, r, s);
      coarse_correction(op, s);
    } else if (!use_net) {
      launch_to_float(r, r32.ptr, N, s);
      launch_set_listed_f32(r32.ptr, 0.0f, op.cons.ptr, op.n_cons, s);
      chebyshev(op, z32.ptr, r32.ptr, dinv32.ptr, t32.ptr, c32.ptr, smooth_degree, lmin, lmax, true,
                s);
    } else {
      presmooth_residual(op, r, s);
      uint32_t d[3];
      for (int k = 0; k < 3; ++k) d[k] = op.box.n[k] * op.p + 1;
      ensure(d, false);
      launch_to_double(t32.ptr, tmp64.ptr, N, s);
      const double rn = std::sqrt(dot64(tmp64.ptr, tmp64.ptr, N, dpart, s));
      if (rn > 0.0) {
        // correction c = ||rho|| s net(rho / ||rho||), added with the weight
        // omega = c.rho / c.Ac that minimises the energy norm of the error
        // along c, so it never increases it
        nn_scale_to_f32(tmp64.ptr, x32.ptr, 1.0 / rn, N, s);
        forward(s);
        launch_fill_f32(n32.ptr, 0.0f, N, s);
        nn_axpy_f32(n32.ptr, (float)(rn * out_scale), y32.ptr, N, s);
        launch_set_listed_f32(n32.ptr, 0.0f, op.cons.ptr, op.n_cons, s);
        vmult32(op, an32.ptr, n32.ptr, s);
        launch_to_double(n32.ptr, tmp64b.ptr, N, s);
        launch_to_double(an32.ptr, tmp64c.ptr, N, s);
        const double crho = dot64(tmp64b.ptr, tmp64.ptr, N, dpart, s);
        const double cac = dot64(tmp64b.ptr, tmp64c.ptr, N, dpart, s);
        last_omega = cac > 0.0 ? crho / cac : 0.0;
        if (last_omega != 0.0) nn_axpy_f32(z32.ptr, (float)last_omega, n32.ptr, N, s);
      }
    }
    chebyshev(op, z32.ptr, r32.ptr, dinv32.ptr, t32.ptr, c32.ptr, smooth_degree, lmin, lmax, false,
              s);
    launch_to_double(z32.ptr, z, N, s);
    launch_copy_listed<double>(z, r, op.cons.ptr, op.n_cons, s);
  }
};

namespace {

void check_dims(const uint32_t *d) {
  MFX_REQUIRE(d, "null dims");
  for (int k = 0; k < 3; ++k) MFX_REQUIRE(d[k] >= 1 && d[k] <= 4096, "grid size out of range");
}

}  // namespace

extern "C" {

uint64_t mfx_nn_n_params(void) { return n_params_total(); }

mfx_status mfx_nn_create(uint64_t seed, mfx_nn **out) {
  return guard([&] {
    MFX_REQUIRE(out, "null pointer argument");
    *out = nullptr;
    auto nn = std::make_unique<mfx_nn>();
    const uint64_t n = n_params_total();
    // He-normal weights: Box-Muller on the SplitMix64 stream of `seed`
    std::vector<float> h(n);
    uint64_t k = 0;
    auto u01 = [&]() {
      ++k;
      uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      return (double)(z >> 11) * 0x1.0p-53;
    };
    uint64_t i = 0;
    for (int l = 0; l < 8; ++l) {
      const double sd = std::sqrt(2.0 / (27.0 * kLayers[l].cin));
      const uint64_t m = 27ull * kLayers[l].cin * kLayers[l].cout;
      for (uint64_t j = 0; j < m; ++j, ++i) {
        const double a = std::max(u01(), 0x1.0p-53), b = u01();
        h[i] = (float)(std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * M_PI * b) * sd);
      }
    }
    nn->w.upload(h.data(), n);
    std::vector<float> zero(n, 0.0f);
    nn->m.upload(zero.data(), n);
    nn->v.upload(zero.data(), n);
    nn->grad.upload(zero.data(), n);
    nn->wr.alloc(n);
    uint64_t off = 0;
    for (int l = 0; l < 8; ++l) {
      nn->wb_off[l] = off;
      off += (uint64_t)kLayers[l].cout * kpad(l);
    }
    nn->wb_off[8] = off;
    nn->wb.alloc(off);
    MFX_HIP(hipDeviceSynchronize());
    *out = nn.release();
  });
}

mfx_status mfx_nn_destroy(mfx_nn *nn) {
  return guard([&] {
    if (nn) {
      (void)hipDeviceSynchronize();
      delete nn;
    }
  });
}

mfx_status mfx_nn_get_params(const mfx_nn *nn, float *host) {
  return guard([&] {
    MFX_REQUIRE(nn && host, "null pointer argument");
    MFX_HIP(hipMemcpy(host, nn->w.ptr, n_params_total() * sizeof(float), hipMemcpyDeviceToHost));
  });
}

mfx_status mfx_nn_set_params(mfx_nn *nn, const float *host) {
  return guard([&] {
    MFX_REQUIRE(nn && host, "null pointer argument");
    MFX_HIP(hipMemcpy(nn->w.ptr, host, n_params_total() * sizeof(float), hipMemcpyHostToDevice));
    nn->packed = false;
  });
}

mfx_status mfx_nn_set_output_scale(mfx_nn *nn, double scale) {
  return guard([&] {
    MFX_REQUIRE(nn && scale > 0.0, "output scale must be positive");
    nn->out_scale = scale;
  });
}
double mfx_nn_output_scale(const mfx_nn *nn) { return nn ? nn->out_scale : 0.0; }

mfx_status mfx_nn_set_coarse(mfx_nn *nn, mfx_op *coarse) {
  return guard([&] {
    MFX_REQUIRE(nn, "null pointer argument");
    MFX_REQUIRE(!coarse || coarse->is_box, "the coarse operator must come from mfx_op_create_box");
    nn->coarse_op = coarse;
    nn->tr_fine_serial = 0;
    nn->tr_coarse_serial = 0;
  });
}

mfx_status mfx_nn_smoother_info(mfx_nn *nn, mfx_op *op, double *lambda_min, double *lambda_max,
                                double *last_omega, void *stream) {
  return guard([&] {
    MFX_REQUIRE(nn && op && lambda_min && lambda_max && last_omega, "null pointer argument");
    MFX_REQUIRE(op->is_box, "the preconditioner needs an operator from mfx_op_create_box");
    nn->setup_smoother(*op, as_stream(stream));
    *lambda_min = nn->lmin;
    *lambda_max = nn->lmax;
    *last_omega = nn->last_omega;
  });
}

mfx_status mfx_nn_set_mode(mfx_nn *nn, int smoother_degree, int use_network) {
  return guard([&] {
    MFX_REQUIRE(nn, "null pointer argument");
    MFX_REQUIRE(smoother_degree >= 0 && smoother_degree <= 32, "smoother degree must be in 0..32");
    MFX_REQUIRE(smoother_degree > 0 || use_network, "degree 0 without the network is no preconditioner");
    nn->smooth_degree = smoother_degree;
    nn->use_net = use_network != 0;
  });
}

mfx_status mfx_nn_forward(mfx_nn *nn, const uint32_t dims[3], const float *x, float *y,
                          void *stream) {
  return guard([&] {
    MFX_REQUIRE(nn && x && y, "null pointer argument");
    check_dims(dims);
    hipStream_t s = as_stream(stream);
    nn->ensure(dims, false);
    const uint64_t P = count(dims);
    MFX_HIP(hipMemcpyAsync(nn->x32.ptr, x, P * 4, hipMemcpyDeviceToDevice, s));
    nn->forward(s);
    MFX_HIP(hipMemcpyAsync(y, nn->y32.ptr, P * 4, hipMemcpyDeviceToDevice, s));
  });
}

mfx_status mfx_nn_gradient(mfx_nn *nn, const uint32_t dims[3], const float *x, const float *dy,
                           float *grad, void *stream) {
  return guard([&] {
    MFX_REQUIRE(nn && x && dy && grad, "null pointer argument");
    check_dims(dims);
    hipStream_t s = as_stream(stream);
    nn->ensure(dims, true);
    const uint64_t P = count(dims);
    MFX_HIP(hipMemcpyAsync(nn->x32.ptr, x, P * 4, hipMemcpyDeviceToDevice, s));
    MFX_HIP(hipMemcpyAsync(nn->dy.ptr, dy, P * 4, hipMemcpyDeviceToDevice, s));
    launch_fill_f32(nn->grad.ptr, 0.0f, n_params_total(), s);
    nn->forward(s);
    nn->backward(s);
    MFX_HIP(hipMemcpyAsync(grad, nn->grad.ptr, n_params_total() * 4, hipMemcpyDeviceToDevice, s));
  });
}

mfx_status mfx_nn_train_step(mfx_nn *nn, const uint32_t dims[3], int batch, const float *x,
                             const float *target, double lr, double *loss, void *stream) {
  return guard([&] {
    MFX_REQUIRE(nn && x && target && loss, "null pointer argument");
    MFX_REQUIRE(batch >= 1, "batch must be >= 1");
    check_dims(dims);
    hipStream_t s = as_stream(stream);
    nn->ensure(dims, true);
    const uint64_t P = count(dims);
    const uint32_t nb = nn_mse_blocks(P);
    if (nn->lpart.n < (uint64_t)nb * batch) nn->lpart.alloc((uint64_t)nb * batch);
    launch_fill_f32(nn->grad.ptr, 0.0f, n_params_total(), s);
    // mean over the batch of the mean squared error per node
    const float scale = (float)(2.0 / ((double)P * batch));
    for (int b = 0; b < batch; ++b) {
      MFX_HIP(hipMemcpyAsync(nn->x32.ptr, x + (uint64_t)b * P, P * 4, hipMemcpyDeviceToDevice, s));
      nn->forward(s);
      nn_mse(nn->y32.ptr, target + (uint64_t)b * P, nullptr, scale, nn->dy.ptr,
             nn->lpart.ptr + (uint64_t)b * nb, P, s);
      nn->backward(s);
    }
    nn->step += 1;
    nn_adam(nn->w.ptr, nn->m.ptr, nn->v.ptr, nn->grad.ptr, (uint32_t)n_params_total(), (float)lr,
            0.9f, 0.999f, 1e-8f, nn->step, s);
    nn->packed = false;
    std::vector<double> h((size_t)nb * batch);
    MFX_HIP(hipMemcpyAsync(h.data(), nn->lpart.ptr, h.size() * 8, hipMemcpyDeviceToHost, s));
    MFX_HIP(hipStreamSynchronize(s));
    double sum = 0.0;
    for (double v : h) sum += v;
    *loss = sum / ((double)P * batch);
  });
}

mfx_status mfx_nn_vmult(mfx_nn *nn, mfx_op *op, double *z, const double *r, void *stream) {
  return guard([&] {
    MFX_REQUIRE(nn && op && z && r, "null pointer argument");
    MFX_REQUIRE(op->is_box, "the network runs on the node grid of a box operator");
    nn->precondition(*op, z, r, as_stream(stream));
  });
}

mfx_status mfx_nn_restricted_residual(mfx_nn *nn, mfx_op *op, const double *r, double *rho_c,
                                      void *stream) {
  return guard([&] {
    MFX_REQUIRE(nn && op && r && rho_c, "null pointer argument");
    MFX_REQUIRE(op->is_box, "the network runs on the node grid of a box operator");
    MFX_REQUIRE(nn->smooth_degree > 0 && nn->coarse_op,
                "needs the two-level mode (smoother degree > 0) and a coarse operator");
    hipStream_t s = as_stream(stream);
    nn->setup_smoother(*op, s);
    nn->presmooth_residual(*op, r, s);
    nn->restrict_residual(*op, s);
    MFX_HIP(hipMemcpyAsync(rho_c, nn->rc64.ptr, nn->coarse_op->N * sizeof(double),
                           hipMemcpyDeviceToDevice, s));
  });
}

// Flexible CG with up to kFcgDirs stored directions (Notay's FCG(m)): the
// new direction is the preconditioned residual made A-orthogonal to the last
// m directions, p = z - sum_j (z.Ap_j / p_j.Ap_j) p_j, then
// alpha = p.r / p.Ap.  With a linear SPD preconditioner it is PCG; with the
// network it keeps the Krylov acceleration the single-direction form loses.
// Stopping rule and result fields of mfx_cg_solve.
//
// Fallback (mfx_nn_set_fallback): the learned cycle is judged stagnant when
// the residual falls by less than fb_min_reduction over the last fb_window
// steps, when the network's correction was rejected (omega <= 0: it does not
// lower the energy norm of the error) in fb_window applications in a row, or
// when a direction breaks down (p.Ap <= 0, e.g. a zero preconditioned
// residual).  The solve then continues with one V-cycle of the fallback
// multigrid per step; flexible CG needs no restart for the change of
// preconditioner.  This mirrors the only failure path of the reference's
// solver, which raises NoConvergence (bp5_kokkos/benchmark.cc:366-374),
// with a working method in place of the exception.
mfx_status mfx_nn_fcg_solve(mfx_nn *nn, mfx_op *op, double *x, const double *b, int max_steps,
                            double abs_tol, double rel_tol, mfx_cg_result *res, void *stream) {
  return guard([&] {
    MFX_REQUIRE(nn && op && x && b && res, "null pointer argument");
    MFX_REQUIRE(op->is_box, "the network runs on the node grid of a box operator");
    MFX_REQUIRE(max_steps >= 0, "max_steps must be >= 0");
    if (nn->fb_mg) {
      mfx_mg_level_info li{};
      const mfx_status st = mfx_mg_get_level(nn->fb_mg, 0, &li);
      if (st != MFX_OK) throw Error{st, mfx_last_error()};
      MFX_REQUIRE(li.op && li.op->serial == op->serial,
                  "the fallback multigrid's fine operator must be the solve's operator");
    }
    constexpr int kFcgDirs = 4;
    hipStream_t s = as_stream(stream);
    const uint64_t N = op->N;
    if (nn->fr.n != N) {
      nn->fr.alloc(N);
      nn->fz.alloc(N);
      nn->fp.alloc(N * kFcgDirs);
      nn->fq.alloc(N * kFcgDirs);
    }
    double *r = nn->fr.ptr, *z = nn->fz.ptr;
    auto P = [&](int j) { return nn->fp.ptr + (uint64_t)j * N; };
    auto Q = [&](int j) { return nn->fq.ptr + (uint64_t)j * N; };
    double pq[kFcgDirs] = {};
    const auto t0 = std::chrono::steady_clock::now();
    vmult64(*op, Q(0), x, s);
    launch_residual_f64(r, b, Q(0), N, s);
    double rn = std::sqrt(dot64(r, r, N, nn->dpart, s));
    res->initial_value = rn;
    const double tol = std::max(abs_tol, rel_tol * rn);
    std::vector<double> hist{rn};
    nn->fb_step = -1;
    const bool net_judged = nn->use_net && nn->smooth_degree > 0;
    int rejected_run = 0;  // consecutive applications whose correction was rejected
    int steps = 0;
    bool converged = rn <= tol, breakdown = false;
    auto engage = [&]() { nn->fb_step = steps; };
    while (!converged && steps < max_steps) {
      if (nn->fb_step >= 0) {
        const mfx_status st = mfx_mg_vmult(nn->fb_mg, z, r, stream);
        if (st != MFX_OK) throw Error{st, mfx_last_error()};
      } else {
        nn->precondition(*op, z, r, s);
        rejected_run = (net_judged && !(nn->last_omega > 0.0)) ? rejected_run + 1 : 0;
      }
      const int slot = steps % kFcgDirs, stored = std::min(steps, kFcgDirs);
      double beta[kFcgDirs] = {};
      for (int j = 0; j < stored; ++j)
        if (j != slot) beta[j] = -dot64(z, Q(j), N, nn->dpart, s) / pq[j];
      MFX_HIP(hipMemcpyAsync(P(slot), z, N * 8, hipMemcpyDeviceToDevice, s));
      for (int j = 0; j < stored; ++j)
        if (j != slot) launch_axpy_f64(P(slot), beta[j], P(j), N, s);
      vmult64(*op, Q(slot), P(slot), s);
      pq[slot] = dot64(P(slot), Q(slot), N, nn->dpart, s);
      if (!(pq[slot] > 0.0)) {
        if (nn->fb_mg && nn->fb_step < 0) {
          engage();  // redo this step with the multigrid
          continue;
        }
        breakdown = true;
        break;
      }
      const double alpha = dot64(P(slot), r, N, nn->dpart, s) / pq[slot];
      launch_pcg_update(x, r, P(slot), Q(slot), alpha, N, s);
      ++steps;
      rn = std::sqrt(dot64(r, r, N, nn->dpart, s));
      hist.push_back(rn);
      converged = rn <= tol;
      if (!converged && nn->fb_mg && nn->fb_step < 0) {
        const int w = nn->fb_window;
        const bool slow = steps >= w && rn * nn->fb_min_reduction > hist[steps - w];
        if (slow || rejected_run >= w) engage();
      }
    }
    MFX_HIP(hipStreamSynchronize(s));
    res->steps = steps;
    res->converged = converged ? 1 : 0;
    res->last_value = rn;
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!converged)
      throw Error{breakdown ? MFX_ERR_BREAKDOWN : MFX_ERR_NOT_CONVERGED,
                  "flexible CG did not converge: steps=" + std::to_string(steps) +
                      " residual=" + std::to_string(rn)};
  });
}

mfx_status mfx_nn_set_fallback(mfx_nn *nn, mfx_mg *mg, int window, double min_reduction) {
  return guard([&] {
    MFX_REQUIRE(nn, "null pointer argument");
    MFX_REQUIRE(window >= 0 && window <= 1000, "window must be in 0..1000 (0: 8)");
    MFX_REQUIRE(min_reduction == 0.0 || min_reduction > 1.0,
                "min_reduction must exceed 1 (0: 2)");
    nn->fb_mg = mg;
    nn->fb_window = window ? window : 8;
    nn->fb_min_reduction = min_reduction > 0.0 ? min_reduction : 2.0;
  });
}

int mfx_nn_fallback_step(const mfx_nn *nn) { return nn ? nn->fb_step : -1; }

// One Adam step on the energy-norm loss of whole fields of a box operator:
//   L = (1/B) sum_b (pred_b - e_b)^T A (pred_b - e_b) / e_b^T A e_b,
//   pred_b = ||r_b|| s net(r_b / ||r_b||) (constrained rows 0),
// with e_b^T A e_b = e_b . r_b (A e_b = r_b); dL/dpred_b = 2 A (pred_b - e_b) /
// (B e_b.r_b).  r, e: batch consecutive FP64 vectors of the operator's DoFs.
mfx_status mfx_nn_train_energy(mfx_nn *nn, mfx_op *op, int batch, const double *r,
                               const double *e, double lr, double *loss, void *stream) {
  return guard([&] {
    MFX_REQUIRE(nn && op && r && e && loss, "null pointer argument");
    MFX_REQUIRE(op->is_box, "the network runs on the node grid of a box operator");
    MFX_REQUIRE(batch >= 1, "batch must be >= 1");
    hipStream_t s = as_stream(stream);
    const uint64_t N = op->N;
    uint32_t d[3];
    for (int k = 0; k < 3; ++k) d[k] = op->box.n[k] * op->p + 1;
    nn->ensure(d, true);
    // its own scratch: the smoother's vectors stay sized for the smoothed operator
    if (nn->en64.n != N) {
      nn->en64.alloc(N);
      nn->en64b.alloc(N);
      nn->en64c.alloc(N);
    }
    launch_fill_f32(nn->grad.ptr, 0.0f, n_params_total(), s);
    double total = 0.0;
    for (int bi = 0; bi < batch; ++bi) {
      const double *rb = r + (uint64_t)bi * N, *eb = e + (uint64_t)bi * N;
      const double rn = std::sqrt(dot64(rb, rb, N, nn->dpart, s));
      const double eae = dot64(eb, rb, N, nn->dpart, s);
      MFX_REQUIRE(rn > 0.0 && eae > 0.0, "training pair with zero residual or energy");
      nn_scale_to_f32(rb, nn->x32.ptr, 1.0 / rn, N, s);
      nn->forward(s);
      nn_scale_to_f64(nn->y32.ptr, nn->en64.ptr, rn * nn->out_scale, N, s);
      launch_set_listed(nn->en64.ptr, 0.0, op->cons.ptr, op->n_cons, s);
      launch_residual_f64(nn->en64b.ptr, nn->en64.ptr, eb, N, s);  // pred - e
      vmult64(*op, nn->en64c.ptr, nn->en64b.ptr, s);
      total += dot64(nn->en64b.ptr, nn->en64c.ptr, N, nn->dpart, s) / eae;
      launch_set_listed(nn->en64c.ptr, 0.0, op->cons.ptr, op->n_cons, s);
      nn_scale_to_f32(nn->en64c.ptr, nn->dy.ptr, 2.0 * rn * nn->out_scale / (eae * batch), N, s);
      nn->backward(s);
    }
    nn->step += 1;
    nn_adam(nn->w.ptr, nn->m.ptr, nn->v.ptr, nn->grad.ptr, (uint32_t)n_params_total(), (float)lr,
            0.9f, 0.999f, 1e-8f, nn->step, s);
    nn->packed = false;
    MFX_HIP(hipStreamSynchronize(s));
    *loss = total / batch;
  });
}

}  // extern "C"
