// mg.cpp -- mixed-precision hybrid p/h-multigrid preconditioner and the FP64
// preconditioned CG around it (mfx_mg_* in include/mfx.h).
//
// The reference stops at identity/Jacobi CG; its Chebyshev/multigrid include
// is unused (CEED_bp/src/bp3.cc:27).  This module is the mixed-precision
// solver of the same BP problems:
//   levels   fine Q_p -> Q_{p/2} -> ... -> Q_1 on the same cells (p-levels),
//            then Q_1 on n/2, n/4, ... cells per direction (h-levels) until
//            the free DoFs fit a dense direct solve; under a Gauss fine
//            operator the coarser levels use Gauss q = p+1;
//   smoother Chebyshev iteration with the Jacobi preconditioner, eigenvalue
//            interval [lambda_max / range, lambda_max] with lambda_max from a
//            few CG (Lanczos) steps times a safety factor (the semantics of
//            deal.II's PreconditionChebyshev); on request the fine level
//            smooths with its own operator of fewer Gauss points (same
//            mesh, degree and DoFs), while the outer CG keeps the fine one;
//   transfer tensor products of 1D nodal interpolation matrices (mg_kernels.hip);
//   coarse   FP64 explicit inverse of the coarsest operator on its free DoFs;
//   cycle    V-cycle in FP32 (operators, smoother, transfers), zero initial
//            guess, pre- and post-smoothing;
//   outer    CG in FP64: residual update, dot products and convergence test
//            in FP64, stopping rule of deal.II's ReductionControl.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "mg.hpp"
#include "op.hpp"
#include "transfer.hpp"

namespace mfx {
namespace {

struct OpDeleter {
  void operator()(mfx_op *o) const { (void)mfx_op_destroy(o); }
};
using OpPtr = std::unique_ptr<mfx_op, OpDeleter>;

double dot(const double *x, const double *y, uint64_t n, DBuf<double> &part, hipStream_t s) {
  const uint32_t nb = dot_n_blocks(n);
  if (part.n < nb + 1) part.alloc(nb + 1);
  launch_dot(x, y, n, part.ptr, part.ptr + nb, s);
  double h = 0.0;
  MFX_HIP(hipMemcpyAsync(&h, part.ptr + nb, sizeof(double), hipMemcpyDeviceToHost, s));
  MFX_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace

void vmult64(mfx_op &op, double *dst, const double *src, hipStream_t s) {
  launch_fill(dst, 0.0, op.N, s);
  op.apply64(src, dst, s);
  launch_copy_listed<double>(dst, src, op.cons.ptr, op.n_cons, s);
}
void vmult32(mfx_op &op, float *dst, const float *src, hipStream_t s) {
  launch_fill_f32(dst, 0.0f, op.N, s);
  op.apply32(src, dst, s);
  launch_copy_listed<float>(dst, src, op.cons.ptr, op.n_cons, s);
}

// Chebyshev iteration for A x = b with the Jacobi preconditioner on the
// eigenvalue interval [lmin, lmax] of D^-1 A (theta, delta its centre and
// half-width): d_0 = D^-1 r_0 / theta, x_1 = x_0 + d_0, then
//   rho_k = 1 / (2 theta / delta - rho_{k-1}),  rho_0 = delta / theta,
//   d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k / delta) D^-1 r_k,  x_{k+1} = x_k + d_k,
// `degree` updates in all, so the error is multiplied by the scaled
// Chebyshev polynomial T_degree((theta - lambda) / delta) / T_degree(theta / delta).
// A is applied degree - 1 times from a zero x, degree times otherwise.
// t and d are scratch vectors of the operator's size.
void chebyshev(mfx_op &op, float *x, const float *b, const float *dinv, float *t, float *d,
               int degree, double lmin, double lmax, bool zero_initial, hipStream_t s) {
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const uint64_t N = op.N;
  if (zero_initial) {
    launch_cheb_step(x, d, b, nullptr, dinv, 0.0f, (float)(1.0 / theta), true, N, s);
  } else {
    vmult32(op, t, x, s);
    launch_cheb_step(x, d, b, t, dinv, 0.0f, (float)(1.0 / theta), false, N, s);
  }
  double rho_old = delta / theta;
  for (int k = 1; k < degree; ++k) {
    vmult32(op, t, x, s);
    const double rho = 1.0 / (2.0 * theta / delta - rho_old);
    launch_cheb_step(x, d, b, t, dinv, (float)(rho * rho_old), (float)(2.0 * rho / delta), false,
                     N, s);
    rho_old = rho;
  }
}

namespace {

// largest eigenvalue of a symmetric tridiagonal matrix (cyclic Jacobi on the
// dense k x k matrix; k is a few tens)
double tridiag_max_eig(const std::vector<double> &diag, const std::vector<double> &off) {
  const int k = (int)diag.size();
  std::vector<double> A((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) {
    A[i * k + i] = diag[i];
    if (i + 1 < k) A[i * k + i + 1] = A[(i + 1) * k + i] = off[i];
  }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double offn = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = i + 1; j < k; ++j) offn += A[i * k + j] * A[i * k + j];
    if (offn < 1e-30) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = A[p * k + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (A[q * k + q] - A[p * k + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int r = 0; r < k; ++r) {  // rotate columns p, q
          const double arp = A[r * k + p], arq = A[r * k + q];
          A[r * k + p] = c * arp - sn * arq;
          A[r * k + q] = sn * arp + c * arq;
        }
        for (int r = 0; r < k; ++r) {  // rotate rows p, q
          const double apr = A[p * k + r], aqr = A[q * k + r];
          A[p * k + r] = c * apr - sn * aqr;
          A[q * k + r] = sn * apr + c * aqr;
        }
      }
  }
  double m = A[0];
  for (int i = 1; i < k; ++i) m = std::max(m, A[i * k + i]);
  return m;
}

}  // namespace

// Largest eigenvalue of D^-1 A from `iters` Jacobi-preconditioned CG steps on
// a random right-hand side: the CG coefficients give the Lanczos tridiagonal
// matrix T (T_ii = 1/a_i + b_{i-1}/a_{i-1}, T_{i,i+1} = sqrt(b_i)/a_i).
double estimate_lambda_max(mfx_op &op, const double *inv_diag, int iters, hipStream_t s) {
  const uint64_t N = op.N;
  DBuf<double> r, z, p, ap, x, part;
  r.alloc(N);
  z.alloc(N);
  p.alloc(N);
  ap.alloc(N);
  x.alloc(N);
  launch_random(r.ptr, N, 0x5eedull, s);
  launch_set_listed(r.ptr, 0.0, op.cons.ptr, op.n_cons, s);
  launch_fill(x.ptr, 0.0, N, s);
  std::vector<double> alpha, beta;
  launch_scale(z.ptr, r.ptr, inv_diag, N, s);  // z = D^-1 r
  MFX_HIP(hipMemcpyAsync(p.ptr, z.ptr, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  double gamma = dot(r.ptr, z.ptr, N, part, s);
  const double gamma0 = gamma;
  for (int i = 0; i < iters && gamma > 1e-28 * gamma0; ++i) {
    vmult64(op, ap.ptr, p.ptr, s);
    const double pap = dot(p.ptr, ap.ptr, N, part, s);
    if (!(pap > 0.0)) break;
    const double a = gamma / pap;
    launch_pcg_update(x.ptr, r.ptr, p.ptr, ap.ptr, a, N, s);
    launch_scale(z.ptr, r.ptr, inv_diag, N, s);
    const double gn = dot(r.ptr, z.ptr, N, part, s);
    const double b = gn / gamma;
    launch_pcg_direction(p.ptr, z.ptr, b, N, s);
    gamma = gn;
    alpha.push_back(a);
    beta.push_back(b);
  }
  MFX_REQUIRE(!alpha.empty(), "eigenvalue estimate: operator is not positive definite");
  const int k = (int)alpha.size();
  std::vector<double> dg(k), od(k > 1 ? k - 1 : 0);
  for (int i = 0; i < k; ++i) {
    dg[i] = 1.0 / alpha[i] + (i > 0 ? beta[i - 1] / alpha[i - 1] : 0.0);
    if (i + 1 < k) od[i] = std::sqrt(beta[i]) / alpha[i];
  }
  return tridiag_max_eig(dg, od);
}

namespace {

// In-place inverse of an SPD matrix through its Cholesky factor.
void spd_inverse(std::vector<double> &A, int n) {
  std::vector<double> L((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    MFX_REQUIRE(d > 0.0, "coarse operator is not positive definite");
    const double ljj = std::sqrt(d);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double v = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) v -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = v / ljj;
    }
  }
  // columns of A^-1: solve L L^T x = e_c (y_i = 0 for i < c in the forward
  // solve); the columns are independent and split over host threads
  auto columns = [&](int c0, int c1) {
    std::vector<double> y(n);
    for (int c = c0; c < c1; ++c) {
      std::fill(y.begin(), y.begin() + c, 0.0);
      for (int i = c; i < n; ++i) {
        double v = (i == c) ? 1.0 : 0.0;
        for (int k = c; k < i; ++k) v -= L[(size_t)i * n + k] * y[k];
        y[i] = v / L[(size_t)i * n + i];
      }
      for (int i = n - 1; i >= 0; --i) {
        double v = y[i];
        for (int k = i + 1; k < n; ++k) v -= L[(size_t)k * n + i] * y[k];
        y[i] = v / L[(size_t)i * n + i];
      }
      for (int i = 0; i < n; ++i) A[(size_t)i * n + c] = y[i];
    }
  };
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 256 || nt == 1) {
    columns(0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t) pool.emplace_back(columns, (int)((int64_t)n * t / nt), (int)((int64_t)n * (t + 1) / nt));
  for (auto &th : pool) th.join();
}

}  // namespace
}  // namespace mfx

using namespace mfx;

struct mfx_mg {
  struct Level {
    OpPtr op;
    OpPtr sop;  // smoothing operator when it differs from op (fine level only)
    uint32_t dims[3] = {0, 0, 0};
    DBuf<float> dinv, x, b, t, d;
    double lmax = 0.0, lmin = 0.0;
    Transfer to_coarse;  // empty on the coarsest level
    // operator of the level's smoothing and residual (FP32)
    mfx_op &sm() const { return sop ? *sop : *op; }
  };
  std::vector<Level> lv;
  int degree = 5;
  // coarse direct solve
  DBuf<double> Ainv;
  DBuf<uint32_t> free_idx;
  uint32_t n_free = 0;
  // outer CG scratch (FP64)
  DBuf<double> r, z, p, ap, part;

  void smooth(Level &L, bool zero_initial, hipStream_t s) {
    chebyshev(L.sm(), L.x.ptr, L.b.ptr, L.dinv.ptr, L.t.ptr, L.d.ptr, degree, L.lmin, L.lmax,
              zero_initial, s);
  }

  // x_l = V-cycle(b_l), zero initial guess
  void vcycle(size_t l, hipStream_t s) {
    Level &L = lv[l];
    if (l + 1 == lv.size()) {
      launch_fill_f32(L.x.ptr, 0.0f, L.op->N, s);
      launch_dense_solve(Ainv.ptr, free_idx.ptr, n_free, L.b.ptr, L.x.ptr, s);
      return;
    }
    Level &Cl = lv[l + 1];
    smooth(L, true, s);
    vmult32(L.sm(), L.t.ptr, L.x.ptr, s);
    launch_residual_f32(L.t.ptr, L.b.ptr, L.t.ptr, L.op->N, s);
    restrict_<float>(L.to_coarse.R32, L.dims, Cl.dims, L.t.ptr, L.d.ptr, L.t.ptr, Cl.b.ptr, s);
    launch_set_listed_f32(Cl.b.ptr, 0.0f, Cl.op->cons.ptr, Cl.op->n_cons, s);
    vcycle(l + 1, s);
    prolongate<float>(L.to_coarse.P32, L.dims, Cl.dims, Cl.x.ptr, L.t.ptr, L.d.ptr, L.x.ptr, true, s);
    smooth(L, false, s);
  }

  // z = M r in FP64: constrained rows pass through, the rest is one FP32 V-cycle
  void precondition(double *z64, const double *r64, hipStream_t s) {
    Level &L = lv[0];
    launch_to_float(r64, L.b.ptr, L.op->N, s);
    launch_set_listed_f32(L.b.ptr, 0.0f, L.op->cons.ptr, L.op->n_cons, s);
    vcycle(0, s);
    launch_to_double(L.x.ptr, z64, L.op->N, s);
    launch_copy_listed<double>(z64, r64, L.op->cons.ptr, L.op->n_cons, s);
  }
};

namespace {

mfx_op *make_level_op(const mfx_box_desc &d) {
  mfx_op *op = nullptr;
  const mfx_status st = mfx_op_create_box(&d, &op);
  if (st != MFX_OK) throw Error{st, std::string("multigrid level: ") + mfx_last_error()};
  return op;
}

void setup_coarse(mfx_mg &mg, uint32_t max_free, hipStream_t s) {
  mfx_mg::Level &L = mg.lv.back();
  mfx_op &op = *L.op;
  std::vector<uint32_t> cons(op.n_cons);
  if (op.n_cons)
    MFX_HIP(hipMemcpy(cons.data(), op.cons.ptr, op.n_cons * sizeof(uint32_t), hipMemcpyDeviceToHost));
  std::vector<char> is_cons(op.N, 0);
  for (uint32_t c : cons) is_cons[c] = 1;
  std::vector<uint32_t> fr;
  for (uint64_t i = 0; i < op.N; ++i)
    if (!is_cons[i]) fr.push_back((uint32_t)i);
  MFX_REQUIRE(fr.size() <= max_free,
              "multigrid: the coarsest level has " + std::to_string(fr.size()) +
                  " free DoFs, more than the direct-solve limit " + std::to_string(max_free) +
                  " (cells per direction need more factors of 2)");
  const int n = (int)fr.size();
  std::vector<double> A((size_t)n * n), col(op.N);
  DBuf<double> e, y;
  e.alloc(op.N);
  y.alloc(op.N);
  const double one = 1.0;
  for (int j = 0; j < n; ++j) {
    launch_fill(e.ptr, 0.0, op.N, s);
    MFX_HIP(hipMemcpyAsync(e.ptr + fr[j], &one, sizeof(double), hipMemcpyHostToDevice, s));
    vmult64(op, y.ptr, e.ptr, s);
    MFX_HIP(hipMemcpyAsync(col.data(), y.ptr, op.N * sizeof(double), hipMemcpyDeviceToHost, s));
    MFX_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) A[(size_t)i * n + j] = col[fr[i]];
  }
  // symmetrize the rounding of the assembled columns before factorizing
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double m = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
      A[(size_t)i * n + j] = A[(size_t)j * n + i] = m;
    }
  if (n) spd_inverse(A, n);
  mg.n_free = (uint32_t)n;
  mg.free_idx.upload(fr.data(), fr.size());
  mg.Ainv.upload(A.data(), A.size());
}

}  // namespace

extern "C" {

mfx_status mfx_mg_create(const mfx_mg_desc *d, mfx_mg **out) {
  return guard([&] {
    MFX_REQUIRE(d && out, "null pointer argument");
    *out = nullptr;
    const int degree = d->smoother_degree > 0 ? d->smoother_degree : 5;
    const double range = d->smoothing_range > 1.0 ? d->smoothing_range : 15.0;
    const int eig_its = d->eig_cg_n_iterations > 0 ? d->eig_cg_n_iterations : 12;
    const double safety = d->eig_safety > 1.0 ? d->eig_safety : 1.2;
    const uint32_t max_free = d->coarse_max_dofs > 0 ? d->coarse_max_dofs : 2000;
    MFX_REQUIRE(degree <= 32, "smoother degree must be <= 32");
    MFX_REQUIRE(max_free <= 8000, "coarse_max_dofs must be <= 8000");
    auto mg = std::make_unique<mfx_mg>();
    mg->degree = degree;
    hipStream_t s = nullptr;

    // level descriptors: degrees p, p/2, ..., 1 on the fine cells, then Q1 on
    // halved meshes while the free DoFs exceed the direct-solve limit
    std::vector<mfx_box_desc> descs;
    mfx_box_desc cur = d->fine;
    const int q_smooth = d->fine_smoother_n_q_points_1d;
    const bool own_smoother = q_smooth > 0 && q_smooth != cur.n_q_points_1d;
    MFX_REQUIRE(q_smooth >= 0, "fine_smoother_n_q_points_1d must be >= 0");
    MFX_REQUIRE(!own_smoother || (cur.quadrature == MFX_QUAD_GAUSS && q_smooth > cur.degree),
                "a separate fine smoothing operator needs Gauss quadrature with q > p");
    // the fine operator serves the outer FP64 CG; it smooths in FP32 unless
    // a separate smoothing operator is asked for
    mfx_box_desc smooth_desc = cur;
    smooth_desc.n_q_points_1d = q_smooth;
    smooth_desc.with_fp32 = 1;
    cur.with_fp32 = own_smoother ? 0 : 1;
    descs.push_back(cur);
    cur.with_fp32 = 1;
    // levels below the fine one: Gauss q = p+1 under a Gauss fine operator
    // (exact for the stiffness of affine cells; the fine level keeps its own
    // rule), otherwise the fine operator's q - p
    const int q_coarse = cur.quadrature == MFX_QUAD_GAUSS ? 1 : cur.n_q_points_1d - cur.degree;
    while (cur.degree > 1) {
      cur.degree = std::max(1, cur.degree / 2);
      cur.n_q_points_1d = cur.degree + q_coarse;
      descs.push_back(cur);
    }
    cur.n_q_points_1d = cur.degree + q_coarse;
    auto free_dofs = [](const mfx_box_desc &b) {
      uint64_t f = 1;
      for (int k = 0; k < 3; ++k) f *= (uint64_t)b.n_cells[k] * b.degree - 1;
      return f;
    };
    while (free_dofs(cur) > max_free && cur.n_cells[0] % 2 == 0 && cur.n_cells[1] % 2 == 0 &&
           cur.n_cells[2] % 2 == 0) {
      for (int k = 0; k < 3; ++k) cur.n_cells[k] /= 2;
      descs.push_back(cur);
    }
    for (const auto &ld : descs) {
      mfx_mg::Level L;
      L.op.reset(make_level_op(ld));
      if (own_smoother && mg->lv.empty()) L.sop.reset(make_level_op(smooth_desc));
      for (int k = 0; k < 3; ++k) L.dims[k] = ld.n_cells[k] * ld.degree + 1;
      mg->lv.push_back(std::move(L));
    }
    for (size_t l = 0; l + 1 < mg->lv.size(); ++l)
      build_transfer(*mg->lv[l].op, *mg->lv[l + 1].op, mg->lv[l].to_coarse, true, false);
    for (auto &L : mg->lv) {
      const uint64_t N = L.op->N;
      DBuf<double> inv;
      inv.alloc(N);
      const mfx_status st = mfx_op_inverse_diagonal(&L.sm(), inv.ptr, s);
      if (st != MFX_OK) throw Error{st, mfx_last_error()};
      L.dinv.alloc(N);
      launch_to_float(inv.ptr, L.dinv.ptr, N, s);
      L.x.alloc(N);
      L.b.alloc(N);
      L.t.alloc(N);
      L.d.alloc(N);
      if (&L != &mg->lv.back()) {
        L.lmax = safety * estimate_lambda_max(L.sm(), inv.ptr, eig_its, s);
        L.lmin = L.lmax / range;
      }
      MFX_HIP(hipStreamSynchronize(s));
    }
    setup_coarse(*mg, max_free, s);
    MFX_HIP(hipStreamSynchronize(s));
    *out = mg.release();
  });
}

mfx_status mfx_mg_destroy(mfx_mg *mg) {
  return guard([&] {
    if (mg) {
      (void)hipDeviceSynchronize();
      delete mg;
    }
  });
}

int mfx_mg_n_levels(const mfx_mg *mg) { return mg ? (int)mg->lv.size() : 0; }

mfx_status mfx_mg_get_level(const mfx_mg *mg, int level, mfx_mg_level_info *info) {
  return guard([&] {
    MFX_REQUIRE(mg && info, "null pointer argument");
    MFX_REQUIRE(level >= 0 && level < (int)mg->lv.size(), "level out of range");
    const auto &L = mg->lv[level];
    info->degree = L.op->p;
    info->n_q_points_1d = L.op->nq;
    for (int k = 0; k < 3; ++k) info->n_cells[k] = L.op->box.n[k];
    info->n_dofs = L.op->N;
    info->lambda_max = L.lmax;
    info->lambda_min = L.lmin;
    info->op = L.op.get();
  });
}

mfx_status mfx_mg_vmult(mfx_mg *mg, double *dst, const double *src, void *stream) {
  return guard([&] {
    MFX_REQUIRE(mg && dst && src, "null pointer argument");
    MFX_REQUIRE(dst != src, "src and dst must not alias");
    mg->precondition(dst, src, as_stream(stream));
  });
}

mfx_status mfx_mg_solve(mfx_mg *mg, double *x, const double *b, int max_steps, double abs_tol,
                        double rel_tol, mfx_cg_result *res, void *stream) {
  return guard([&] {
    MFX_REQUIRE(mg && x && b && res, "null pointer argument");
    MFX_REQUIRE(max_steps >= 0, "max_steps must be >= 0");
    hipStream_t s = as_stream(stream);
    mfx_op &A = *mg->lv[0].op;
    const uint64_t N = A.N;
    if (mg->r.n != N) {
      mg->r.alloc(N);
      mg->z.alloc(N);
      mg->p.alloc(N);
      mg->ap.alloc(N);
    }
    const auto t0 = std::chrono::steady_clock::now();
    // r = b - A x
    vmult64(A, mg->ap.ptr, x, s);
    launch_residual_f64(mg->r.ptr, b, mg->ap.ptr, N, s);
    double res_norm = std::sqrt(dot(mg->r.ptr, mg->r.ptr, N, mg->part, s));
    const double tol = std::max(abs_tol, rel_tol * res_norm);
    res->initial_value = res_norm;
    int steps = 0;
    bool converged = res_norm <= tol;
    bool breakdown = false;
    if (!converged && max_steps > 0) {
      mg->precondition(mg->z.ptr, mg->r.ptr, s);
      MFX_HIP(hipMemcpyAsync(mg->p.ptr, mg->z.ptr, N * sizeof(double), hipMemcpyDeviceToDevice, s));
      double gamma = dot(mg->r.ptr, mg->z.ptr, N, mg->part, s);
      while (steps < max_steps) {
        vmult64(A, mg->ap.ptr, mg->p.ptr, s);
        const double pap = dot(mg->p.ptr, mg->ap.ptr, N, mg->part, s);
        if (!(pap > 0.0) || !(gamma > 0.0)) {
          breakdown = true;
          break;
        }
        const double alpha = gamma / pap;
        launch_pcg_update(x, mg->r.ptr, mg->p.ptr, mg->ap.ptr, alpha, N, s);
        ++steps;
        res_norm = std::sqrt(dot(mg->r.ptr, mg->r.ptr, N, mg->part, s));
        if (res_norm <= tol) {
          converged = true;
          break;
        }
        mg->precondition(mg->z.ptr, mg->r.ptr, s);
        const double gn = dot(mg->r.ptr, mg->z.ptr, N, mg->part, s);
        launch_pcg_direction(mg->p.ptr, mg->z.ptr, gn / gamma, N, s);
        gamma = gn;
      }
    }
    MFX_HIP(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    res->steps = steps;
    res->converged = converged ? 1 : 0;
    res->last_value = res_norm;
    res->seconds = std::chrono::duration<double>(t1 - t0).count();
    if (!converged)
      throw Error{breakdown ? MFX_ERR_BREAKDOWN : MFX_ERR_NOT_CONVERGED,
                  "multigrid CG did not converge: steps=" + std::to_string(steps) +
                      " residual=" + std::to_string(res_norm)};
  });
}

mfx_status mfx_transfer_prolongate_f64(const mfx_op *fine, const mfx_op *coarse, double *dst,
                                       const double *src, void *stream) {
  return guard([&] {
    MFX_REQUIRE(fine && coarse && dst && src, "null pointer argument");
    hipStream_t s = as_stream(stream);
    Transfer t;
    build_transfer(*fine, *coarse, t, false, true);
    DBuf<double> t1, t2;
    t1.alloc((uint64_t)t.fine[0] * t.coarse[1] * t.coarse[2]);
    t2.alloc((uint64_t)t.fine[0] * t.fine[1] * t.coarse[2]);
    prolongate<double>(t.P64, t.fine, t.coarse, src, t1.ptr, t2.ptr, dst, false, s);
    MFX_HIP(hipStreamSynchronize(s));
  });
}

mfx_status mfx_transfer_restrict_f64(const mfx_op *fine, const mfx_op *coarse, double *dst,
                                     const double *src, void *stream) {
  return guard([&] {
    MFX_REQUIRE(fine && coarse && dst && src, "null pointer argument");
    hipStream_t s = as_stream(stream);
    Transfer t;
    build_transfer(*fine, *coarse, t, false, true);
    DBuf<double> t1, t2;
    t1.alloc((uint64_t)t.fine[0] * t.fine[1] * t.coarse[2]);
    t2.alloc((uint64_t)t.fine[0] * t.coarse[1] * t.coarse[2]);
    restrict_<double>(t.R64, t.fine, t.coarse, src, t1.ptr, t2.ptr, dst, s);
    MFX_HIP(hipStreamSynchronize(s));
  });
}

mfx_status mfx_chebyshev_f32(mfx_op *op, const float *inv_diag, float *x, const float *b,
                             int degree, double lambda_min, double lambda_max, int x_is_zero,
                             void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && inv_diag && x && b, "null pointer argument");
    MFX_REQUIRE(op->D32.ptr, "operator created without with_fp32");
    MFX_REQUIRE(degree >= 1 && degree <= 32, "degree must be in 1..32");
    MFX_REQUIRE(lambda_max > lambda_min && lambda_min > 0.0, "need 0 < lambda_min < lambda_max");
    hipStream_t s = as_stream(stream);
    DBuf<float> t, dd;
    t.alloc(op->N);
    dd.alloc(op->N);
    chebyshev(*op, x, b, inv_diag, t.ptr, dd.ptr, degree, lambda_min, lambda_max, x_is_zero != 0, s);
    MFX_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
