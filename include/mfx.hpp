// mfx.hpp -- C++ host API over the C-ABI of include/mfx.h.
//
// Mirrors the pieces of deal.II the reference drivers use on this path, with
// the same member names and meaning:
//   Portable::LaplaceOperator<dim,fe_degree,nq,number>
//       (CEED_bp/include/portable_laplace_operator.h:17-96):
//       vmult, Tvmult, m, n, initialize_dof_vector;
//   LinearAlgebra::distributed::Vector<number, MemorySpace::Default>
//       (device vector: size, get_values, operator=(0), l2_norm, import);
//   SolverCG, ReductionControl, PreconditionIdentity, DiagonalMatrix,
//   SolverControl::NoConvergence (CEED_bp/src/bp3.cc:264-288,
//   bp5_kokkos/benchmark.cc:336-378);
//   PreconditionMG: the mixed-precision p/h-multigrid of mfx_mg_* (no
//   counterpart in the reference; deal.II's PreconditionChebyshev parameters);
//   BatchedLaplace: many independent problems solved together (mfx_batch_*);
//   LearnedPreconditioner: the network preconditioner of mfx_nn_*, with
//   SolverCG::solve running flexible CG for it.
// None of these three has a counterpart in the reference.
// Errors become exceptions (mfx::Error / NoConvergence), as deal.II's do.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mfx.h"

namespace mfx {

class Error : public std::runtime_error {
 public:
  Error(mfx_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
  mfx_status status;
};

inline void check(mfx_status s) {
  if (s != MFX_OK) throw Error(s, mfx_last_error());
}
inline void check_hip(hipError_t e) {
  if (e != hipSuccess) throw Error(MFX_ERR_HIP, hipGetErrorString(e));
}

// Device vector of owned size n (no ghosts on one GPU).
template <typename Number>
class Vector {
 public:
  Vector() = default;
  explicit Vector(std::size_t n) { reinit(n); }
  Vector(const Vector &o) { *this = o; }
  Vector(Vector &&o) noexcept { swap(o); }
  ~Vector() {
    if (values_) (void)hipFree(values_);
  }
  Vector &operator=(const Vector &o) {
    if (this != &o) {
      reinit(o.size_, true);
      if (size_) check_hip(hipMemcpy(values_, o.values_, bytes(), hipMemcpyDeviceToDevice));
    }
    return *this;
  }
  Vector &operator=(Vector &&o) noexcept {
    swap(o);
    return *this;
  }
  void swap(Vector &o) noexcept {
    std::swap(values_, o.values_);
    std::swap(size_, o.size_);
  }
  void reinit(std::size_t n, bool omit_zeroing = false) {
    if (n != size_) {
      if (values_) check_hip(hipFree(values_));
      values_ = nullptr;
      size_ = n;
      if (n) check_hip(hipMalloc(reinterpret_cast<void **>(&values_), bytes()));
    }
    if (!omit_zeroing && n) check_hip(hipMemset(values_, 0, bytes()));
  }
  // deal.II's `v = 0.`
  Vector &operator=(Number s) {
    if (s != Number(0)) throw Error(MFX_ERR_INVALID_ARG, "only v = 0 is supported");
    if (size_) check_hip(hipMemset(values_, 0, bytes()));
    return *this;
  }
  std::size_t size() const { return size_; }
  std::size_t locally_owned_size() const { return size_; }
  Number *get_values() { return values_; }
  const Number *get_values() const { return values_; }
  void import_elements(const std::vector<Number> &host) {
    if (host.size() != size_) throw Error(MFX_ERR_INVALID_ARG, "import: size mismatch");
    check_hip(hipMemcpy(values_, host.data(), bytes(), hipMemcpyHostToDevice));
  }
  std::vector<Number> to_host() const {
    std::vector<Number> h(size_);
    if (size_) check_hip(hipMemcpy(h.data(), values_, bytes(), hipMemcpyDeviceToHost));
    return h;
  }
  double operator*(const Vector &o) const {
    static_assert(sizeof(Number) == 8, "dot products on FP64 vectors");
    double r = 0.0;
    check(mfx_vec_dot(values_, o.values_, size_, &r, nullptr));
    return r;
  }
  double l2_norm() const;

 private:
  std::size_t bytes() const { return size_ * sizeof(Number); }
  Number *values_ = nullptr;
  std::size_t size_ = 0;
};

template <typename Number>
double Vector<Number>::l2_norm() const {
  double r = (*this) * (*this);
  return r > 0 ? __builtin_sqrt(r) : 0.0;
}

// Structured box mesh of the reference drivers
// (GridGenerator::subdivided_hyper_rectangle + refine_global,
// CEED_bp/src/bp3.cc:452-488), optionally deformed on the vertices.
struct BoxMesh {
  unsigned n_cells[3] = {1, 1, 1};
  double p1[3] = {0, 0, 0};
  double p2[3] = {1, 1, 1};
  double deform = 0.0;
};

enum class Quadrature { gauss = MFX_QUAD_GAUSS, gauss_lobatto = MFX_QUAD_GLL };
// metric terms stored per quadrature point, or computed on the fly from the
// cell vertices (GLL collocation only)
enum class Geometry { stored = MFX_GEOM_STORED, on_the_fly = MFX_GEOM_ON_THE_FLY };

template <typename Number = double>
class LaplaceOperator {
  static_assert(sizeof(Number) == 8 || sizeof(Number) == 4, "double or float");

 public:
  using VectorType = Vector<Number>;

  LaplaceOperator(const BoxMesh &mesh, unsigned fe_degree, unsigned n_q_points_1d,
                  Quadrature quad = Quadrature::gauss, bool with_fp32 = false,
                  Geometry geometry = Geometry::stored) {
    mfx_box_desc d{};
    d.degree = static_cast<int>(fe_degree);
    d.n_q_points_1d = static_cast<int>(n_q_points_1d);
    d.quadrature = static_cast<int>(quad);
    d.geometry = static_cast<int>(geometry);
    for (int k = 0; k < 3; ++k) {
      d.n_cells[k] = mesh.n_cells[k];
      d.lo[k] = mesh.p1[k];
      d.hi[k] = mesh.p2[k];
    }
    d.deform = mesh.deform;
    d.with_fp32 = (with_fp32 || sizeof(Number) == 4) ? 1 : 0;
    check(mfx_op_create_box(&d, &op_));
  }
  LaplaceOperator(const LaplaceOperator &) = delete;
  LaplaceOperator &operator=(const LaplaceOperator &) = delete;
  ~LaplaceOperator() {
    if (op_ && owned_) mfx_op_destroy(op_);
  }
  // view of an operator owned elsewhere (a multigrid level)
  static LaplaceOperator view(mfx_op *op) { return LaplaceOperator(op); }
  LaplaceOperator(LaplaceOperator &&o) noexcept : op_(o.op_), owned_(o.owned_) {
    o.op_ = nullptr;
  }

  // dst = A src; constrained rows are the identity (copy_constrained_values).
  void vmult(VectorType &dst, const VectorType &src) const {
    if constexpr (sizeof(Number) == 8)
      check(mfx_op_vmult(op_, dst.get_values(), src.get_values(), nullptr));
    else
      check(mfx_op_vmult_f32(op_, dst.get_values(), src.get_values(), nullptr));
  }
  void Tvmult(VectorType &dst, const VectorType &src) const { vmult(dst, src); }
  std::uint64_t m() const { return mfx_op_n_dofs(op_); }
  std::uint64_t n() const { return m(); }
  std::uint32_t n_cells() const { return mfx_op_n_cells(op_); }
  void initialize_dof_vector(VectorType &v) const { v.reinit(m()); }

  // inverse diagonal for Jacobi (MatrixFreeTools::compute_diagonal + the
  // inversion of bp5_kokkos/benchmark.cc:231-250)
  void compute_inverse_diagonal(Vector<double> &inv) const {
    inv.reinit(m(), true);
    check(mfx_op_inverse_diagonal(op_, inv.get_values(), nullptr));
  }
  // b_i = (phi_i, f); f_kind 0: f = 1, 1: f = 3 pi^2 prod sin(pi x_d)
  void compute_rhs(Vector<double> &b, int f_kind) const {
    b.reinit(m(), true);
    check(mfx_op_rhs(op_, f_kind, b.get_values(), nullptr));
  }
  mfx_op *handle() const { return op_; }

 private:
  explicit LaplaceOperator(mfx_op *op) : op_(op), owned_(false) {}
  mfx_op *op_ = nullptr;
  bool owned_ = true;
};

// Multigrid preconditioner: an FP32 V-cycle (Chebyshev smoothing, nodal
// transfers, FP64 coarse solve) over the levels Q_p ... Q_1 and h-coarsened
// Q_1 meshes.  It builds the fine operator itself; solve with
// SolverCG::solve(mg.get_operator(), x, b, mg).
struct MGAdditionalData {
  unsigned smoother_degree = 5;         // Chebyshev steps per smoothing
  double smoothing_range = 15.0;        // lambda_min = lambda_max / range
  unsigned eig_cg_n_iterations = 12;    // CG (Lanczos) steps of the estimate
  double eig_safety = 1.2;              // lambda_max = safety x estimate
  std::uint32_t coarse_max_dofs = 2000; // largest directly solved level
  unsigned fine_smoother_n_q_points_1d = 0;  // fine smoothing rule (0: the fine operator's)
};

class PreconditionMG {
 public:
  using AdditionalData = MGAdditionalData;
  PreconditionMG(const BoxMesh &mesh, unsigned fe_degree, unsigned n_q_points_1d,
                 Quadrature quad = Quadrature::gauss, const AdditionalData &data = {},
                 int geometry = MFX_GEOM_STORED) {
    mfx_mg_desc d{};
    d.fine.degree = static_cast<int>(fe_degree);
    d.fine.n_q_points_1d = static_cast<int>(n_q_points_1d);
    d.fine.quadrature = static_cast<int>(quad);
    d.fine.geometry = geometry;
    for (int k = 0; k < 3; ++k) {
      d.fine.n_cells[k] = mesh.n_cells[k];
      d.fine.lo[k] = mesh.p1[k];
      d.fine.hi[k] = mesh.p2[k];
    }
    d.fine.deform = mesh.deform;
    d.smoother_degree = static_cast<int>(data.smoother_degree);
    d.smoothing_range = data.smoothing_range;
    d.eig_cg_n_iterations = static_cast<int>(data.eig_cg_n_iterations);
    d.eig_safety = data.eig_safety;
    d.coarse_max_dofs = data.coarse_max_dofs;
    d.fine_smoother_n_q_points_1d = static_cast<int>(data.fine_smoother_n_q_points_1d);
    check(mfx_mg_create(&d, &mg_));
    mfx_mg_level_info info{};
    check(mfx_mg_get_level(mg_, 0, &info));
    fine_.push_back(LaplaceOperator<double>::view(info.op));
  }
  PreconditionMG(const PreconditionMG &) = delete;
  PreconditionMG &operator=(const PreconditionMG &) = delete;
  ~PreconditionMG() {
    fine_.clear();
    if (mg_) mfx_mg_destroy(mg_);
  }
  const LaplaceOperator<double> &get_operator() const { return fine_.front(); }
  unsigned n_levels() const { return static_cast<unsigned>(mfx_mg_n_levels(mg_)); }
  mfx_mg_level_info level(unsigned l) const {
    mfx_mg_level_info info{};
    check(mfx_mg_get_level(mg_, static_cast<int>(l), &info));
    return info;
  }
  // dst = V-cycle(src)
  void vmult(Vector<double> &dst, const Vector<double> &src) const {
    check(mfx_mg_vmult(mg_, dst.get_values(), src.get_values(), nullptr));
  }
  mfx_mg *handle() const { return mg_; }

 private:
  mfx_mg *mg_ = nullptr;
  std::vector<LaplaceOperator<double>> fine_;
};

// Independent BP5 problems (GLL collocation, degree 2..6) on one
// concatenated vector; per-instance Jacobi or identity CG in one batch.
struct BatchInstance {
  unsigned degree = 2;
  BoxMesh mesh;
};

class BatchedLaplace {
 public:
  explicit BatchedLaplace(const std::vector<BatchInstance> &instances, bool autotune = true) {
    std::vector<mfx_batch_instance> d(instances.size());
    for (std::size_t i = 0; i < instances.size(); ++i) {
      d[i].degree = static_cast<int>(instances[i].degree);
      for (int k = 0; k < 3; ++k) {
        d[i].n_cells[k] = instances[i].mesh.n_cells[k];
        d[i].lo[k] = instances[i].mesh.p1[k];
        d[i].hi[k] = instances[i].mesh.p2[k];
      }
      d[i].deform = instances[i].mesh.deform;
    }
    check(mfx_batch_create(d.data(), static_cast<int>(d.size()), autotune ? 1 : 0, &b_));
  }
  BatchedLaplace(const BatchedLaplace &) = delete;
  BatchedLaplace &operator=(const BatchedLaplace &) = delete;
  ~BatchedLaplace() {
    if (b_) mfx_batch_destroy(b_);
  }
  unsigned n_instances() const { return static_cast<unsigned>(mfx_batch_n_instances(b_)); }
  std::uint64_t m() const { return mfx_batch_n_dofs(b_); }
  // DoF range [offsets[i], offsets[i + 1]) of instance i
  std::vector<std::uint64_t> offsets() const {
    std::vector<std::uint64_t> o(n_instances() + 1);
    check(mfx_batch_offsets(b_, o.data()));
    return o;
  }
  void initialize_dof_vector(Vector<double> &v) const { v.reinit(m()); }
  void vmult(Vector<double> &dst, const Vector<double> &src) const {
    check(mfx_batch_vmult(b_, dst.get_values(), src.get_values(), nullptr));
  }
  void compute_inverse_diagonal(Vector<double> &inv) const {
    inv.reinit(m(), true);
    check(mfx_batch_inverse_diagonal(b_, inv.get_values(), nullptr));
  }
  mfx_batch_tuning tuning(unsigned degree) const {
    mfx_batch_tuning t{};
    check(mfx_batch_get_tuning(b_, static_cast<int>(degree), &t));
    return t;
  }
  void set_variant(unsigned degree, int variant) {
    check(mfx_batch_set_variant(b_, static_cast<int>(degree), variant));
  }
  // per-instance CG (Jacobi when inv_diag is given) with the stopping rule
  // of ReductionControl(max_steps, tolerance, reduction); one result per
  // instance; MFX_ERR_NOT_CONVERGED if any instance fails
  std::vector<mfx_cg_result> solve(Vector<double> &x, const Vector<double> &b,
                                   const Vector<double> *inv_diag, unsigned max_steps,
                                   double tolerance, double reduction,
                                   bool use_graph = true) const {
    std::vector<mfx_cg_result> r(n_instances());
    check(mfx_batch_cg_solve(b_, x.get_values(), b.get_values(),
                             inv_diag ? inv_diag->get_values() : nullptr,
                             static_cast<int>(max_steps), tolerance, reduction,
                             use_graph ? 1 : 0, r.data(), nullptr));
    return r;
  }
  mfx_batch *handle() const { return b_; }

 private:
  mfx_batch *b_ = nullptr;
};

// The learned preconditioner: with smoother degree k > 0, k Chebyshev steps,
// the network's correction (on a coarse operator's grid when one is set)
// with the energy-optimal weight, k Chebyshev steps; nonlinear, so solve
// with SolverCG::solve(A, x, b, P), which runs flexible CG.
class LearnedPreconditioner {
 public:
  explicit LearnedPreconditioner(std::uint64_t seed = 0) { check(mfx_nn_create(seed, &nn_)); }
  LearnedPreconditioner(const LearnedPreconditioner &) = delete;
  LearnedPreconditioner &operator=(const LearnedPreconditioner &) = delete;
  ~LearnedPreconditioner() {
    if (nn_) mfx_nn_destroy(nn_);
  }
  static std::uint64_t n_params() { return mfx_nn_n_params(); }
  std::vector<float> get_params() const {
    std::vector<float> w(n_params());
    check(mfx_nn_get_params(nn_, w.data()));
    return w;
  }
  void set_params(const std::vector<float> &w) {
    if (w.size() != n_params()) throw Error(MFX_ERR_INVALID_ARG, "wrong number of parameters");
    check(mfx_nn_set_params(nn_, w.data()));
  }
  void set_output_scale(double s) { check(mfx_nn_set_output_scale(nn_, s)); }
  double output_scale() const { return mfx_nn_output_scale(nn_); }
  void set_mode(int smoother_degree, bool use_network = true) {
    check(mfx_nn_set_mode(nn_, smoother_degree, use_network ? 1 : 0));
  }
  // the operator whose node grid the network runs on; nullptr: the fine grid
  void set_coarse(const LaplaceOperator<double> *coarse) {
    check(mfx_nn_set_coarse(nn_, coarse ? coarse->handle() : nullptr));
  }
  // one Adam step on the energy-norm loss over `batch` consecutive fields
  // (r, e: batch x A.m(), A e = r); returns the loss before the step
  double train_energy_step(const LaplaceOperator<double> &A, const Vector<double> &r,
                           const Vector<double> &e, int batch, double lr) {
    double loss = 0.0;
    check(mfx_nn_train_energy(nn_, A.handle(), batch, r.get_values(), e.get_values(), lr, &loss,
                              nullptr));
    return loss;
  }
  void vmult(const LaplaceOperator<double> &A, Vector<double> &z, const Vector<double> &r) const {
    check(mfx_nn_vmult(nn_, A.handle(), z.get_values(), r.get_values(), nullptr));
  }
  // the multigrid the flexible CG switches to when the learned cycle
  // stagnates (mfx_nn_set_fallback); nullptr: none
  void set_fallback(const PreconditionMG *mg, int window = 8, double min_reduction = 2.0) {
    check(mfx_nn_set_fallback(nn_, mg ? mg->handle() : nullptr, window, min_reduction));
  }
  // step at which the last solve switched to the fallback, -1 if it did not
  int fallback_step() const { return mfx_nn_fallback_step(nn_); }
  mfx_nn *handle() const { return nn_; }

 private:
  mfx_nn *nn_ = nullptr;
};

class SolverControl {
 public:
  struct NoConvergence : public std::runtime_error {
    NoConvergence(unsigned step, double value)
        : std::runtime_error("Iterative method reported convergence failure in step " +
                             std::to_string(step)),
          last_step(step), last_residual(value) {}
    unsigned last_step;
    double last_residual;
  };
  SolverControl(unsigned max_steps, double tolerance)
      : max_steps_(max_steps), tolerance_(tolerance) {}
  virtual ~SolverControl() = default;
  unsigned max_steps() const { return max_steps_; }
  double tolerance() const { return tolerance_; }
  virtual double reduction() const { return 0.0; }
  unsigned last_step() const { return last_step_; }
  double initial_value() const { return initial_; }
  double last_value() const { return last_; }
  double seconds() const { return seconds_; }
  void record(const mfx_cg_result &r) {
    last_step_ = static_cast<unsigned>(r.steps);
    initial_ = r.initial_value;
    last_ = r.last_value;
    seconds_ = r.seconds;
  }

 private:
  unsigned max_steps_;
  double tolerance_;
  unsigned last_step_ = 0;
  double initial_ = 0, last_ = 0, seconds_ = 0;
};

// Stops at ||r_k|| <= max(tolerance, reduction * ||r_0||).
class ReductionControl : public SolverControl {
 public:
  ReductionControl(unsigned max_steps, double tolerance, double reduction)
      : SolverControl(max_steps, tolerance), reduction_(reduction) {}
  double reduction() const override { return reduction_; }

 private:
  double reduction_;
};

struct PreconditionIdentity {};

template <typename VectorType>
class DiagonalMatrix {
 public:
  VectorType &get_vector() { return v_; }
  const VectorType &get_vector() const { return v_; }

 private:
  VectorType v_;
};

template <typename VectorType>
class SolverCG {
 public:
  explicit SolverCG(SolverControl &control) : control_(control) {}

  template <typename MatrixType>
  void solve(const MatrixType &A, VectorType &x, const VectorType &b,
             const PreconditionIdentity &) {
    run(A, x, b, nullptr);
  }
  template <typename MatrixType>
  void solve(const MatrixType &A, VectorType &x, const VectorType &b,
             const DiagonalMatrix<VectorType> &P) {
    run(A, x, b, P.get_vector().get_values());
  }
  // flexible CG (4 stored directions) preconditioned by the network
  void solve(const LaplaceOperator<double> &A, VectorType &x, const VectorType &b,
             const LearnedPreconditioner &P) {
    mfx_cg_result r{};
    const mfx_status s = mfx_nn_fcg_solve(P.handle(), A.handle(), x.get_values(), b.get_values(),
                                          static_cast<int>(control_.max_steps()),
                                          control_.tolerance(), control_.reduction(), &r, nullptr);
    control_.record(r);
    if (s == MFX_ERR_NOT_CONVERGED || s == MFX_ERR_BREAKDOWN)
      throw SolverControl::NoConvergence(control_.last_step(), control_.last_value());
    check(s);
  }
  // FP64 CG preconditioned by one multigrid V-cycle per step
  template <typename MatrixType>
  void solve(const MatrixType &A, VectorType &x, const VectorType &b, const PreconditionMG &P) {
    if (A.handle() != P.get_operator().handle())
      throw Error(MFX_ERR_INVALID_ARG, "A must be the multigrid's fine operator (get_operator())");
    mfx_cg_result r{};
    const mfx_status s = mfx_mg_solve(P.handle(), x.get_values(), b.get_values(),
                                      static_cast<int>(control_.max_steps()),
                                      control_.tolerance(), control_.reduction(), &r, nullptr);
    control_.record(r);
    if (s == MFX_ERR_NOT_CONVERGED || s == MFX_ERR_BREAKDOWN)
      throw SolverControl::NoConvergence(control_.last_step(), control_.last_value());
    check(s);
  }

 private:
  template <typename MatrixType>
  void run(const MatrixType &A, VectorType &x, const VectorType &b, const double *inv) {
    mfx_cg_result r{};
    const mfx_status s =
        mfx_cg_solve(A.handle(), x.get_values(), b.get_values(), inv,
                     static_cast<int>(control_.max_steps()), control_.tolerance(),
                     control_.reduction(), &r, nullptr);
    control_.record(r);
    if (s == MFX_ERR_NOT_CONVERGED || s == MFX_ERR_BREAKDOWN)
      throw SolverControl::NoConvergence(control_.last_step(), control_.last_value());
    check(s);
  }
  SolverControl &control_;
};

}  // namespace mfx
