/*
 * mfx.h -- C-ABI of the MI355X (gfx950) matrix-free Laplacian library
 * (benchmarks_amd/libmfx.so).
 *
 * Plain pointers and sizes only.  All vector arguments are DEVICE pointers
 * (HBM-resident, like deal.II's LinearAlgebra::distributed::Vector<number,
 * MemorySpace::Default>::get_values(), CEED_bp/include/portable_laplace_operator.h:135-136);
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 * Calls are stream-ordered and asynchronous unless stated otherwise.  One
 * operator handle must not be used from two host threads at once.
 *
 * Conventions: reference cell [0,1]^3; tensor index x fastest
 * (l = a + n (b + n c)); DoF map uint32 [n_cells][(p+1)^3] with
 * MFX_INVALID_DOF (0xFFFFFFFF) for Dirichlet-constrained DoFs; geometry
 * G[cell][6][q^3], components (xx, xy, xz, yy, yz, zz) of JxW * J^-1 J^-T.
 */
#ifndef MFX_H
#define MFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFX_INVALID_DOF 0xFFFFFFFFu

/* Status codes.  Replaces deal.II's Assert/ExcMessage and the exceptions
 * caught in main (CEED_bp/src/bp3.cc:600-624) and Kokkos' allocation
 * failures (bakeoff_problems_dealii/tests/scaling_bp35_kokkos_kernel/results/1_GPU.out:678). */
typedef enum {
  MFX_OK = 0,
  MFX_ERR_INVALID_ARG = 1,
  MFX_ERR_OOM = 2,
  MFX_ERR_HIP = 3,
  MFX_ERR_NOT_CONVERGED = 4, /* SolverControl::NoConvergence */
  MFX_ERR_UNSUPPORTED = 5,
  MFX_ERR_BREAKDOWN = 6 /* NaN / zero curvature inside CG */
} mfx_status;

typedef enum { MFX_QUAD_GAUSS = 0, MFX_QUAD_GLL = 1 } mfx_quadrature;
typedef enum { MFX_GEOM_STORED = 0, MFX_GEOM_ON_THE_FLY = 1 } mfx_geometry_mode;

/* Host-only: 1D tables on [0,1] for degree p and nq quadrature points
 * (quadrature = mfx_quadrature): B[q*(p+1)+m] (FE_Q basis on Gauss-Lobatto
 * support points at the quadrature points), D[q*nq+k] (collocation
 * derivative), quadrature points xq[nq] and weights wq[nq].  The shape data
 * deal.II's MatrixFree hands the reference kernel (shape_values,
 * co_shape_gradients: CEED_bp/include/portable_laplace_operator.h:147). */
mfx_status mfx_basis_tables(int p, int nq, int quadrature, double *B, double *D,
                            double *xq, double *wq);

/* Thread-local message for the last failing call on this host thread. */
const char *mfx_last_error(void);

/* Library version (major*10000 + minor*100 + patch). */
int mfx_version(void);

/* ----------------------------------------------------------------------
 * Kernel-level plug point.
 * Replaces BK3::Parallel::KokkosKernel<dim,nm,nq,number>(shape_values,
 * co_shape_gradients, G, src, dst, dof_indices, n_cells, numBlocks,
 * threadsPerBlock) (CEED_bp/include/bk3_kokkos_kernel.h:27-38, called once
 * per colour at CEED_bp/include/portable_laplace_operator.h:154-164) and
 * KokkosKernel_1D_Block (bakeoff_problems_dealii/include/bk3_kokkos_kernel.h:32-434).
 * dst += sum_cells P^T B^T D^T G D B P src over the given cells, skipping
 * MFX_INVALID_DOF entries of `map` (gather gives 0, scatter skips).
 * B[q*(p+1)+m]: nodal basis m at quadrature point q; D[q*nq+k]: collocation
 * derivative on the quadrature points.  Degrees p = 1..8, nq = p+1 or p+2.
 * ---------------------------------------------------------------------- */
mfx_status mfx_bk3_apply_f64(int p, int nq, const double *B, const double *D,
                             const double *G, const double *src, double *dst,
                             const uint32_t *map, uint32_t n_cells,
                             void *stream);
mfx_status mfx_bk3_apply_f32(int p, int nq, const float *B, const float *D,
                             const float *G, const float *src, float *dst,
                             const uint32_t *map, uint32_t n_cells,
                             void *stream);
/* GLL collocation (BK5, B = I, nq = p+1): the L-vector form of
 * BK5::Parallel::LaplaceOperator (CEED_BK/include/kernels/BK5/templated_cuda_kernels.cuh:11-126). */
mfx_status mfx_bk5_apply_f64(int p, const double *D, const double *G,
                             const double *src, double *dst,
                             const uint32_t *map, uint32_t n_cells,
                             void *stream);
mfx_status mfx_bk5_apply_f32(int p, const float *D, const float *G,
                             const float *src, float *dst, const uint32_t *map,
                             uint32_t n_cells, void *stream);

/* E-vector (no gather/scatter) element kernels: out[e] = K_e in[e],
 * the standalone kernels of CEED_BK/include/kernels/BK{3,5}/templated_cuda_kernels.cuh
 * (BK5 :11-126, BK3 :11-290).  in/out[e][(p+1)^3], x fastest. */
mfx_status mfx_bk3_evec_f64(int p, int nq, const double *B, const double *D,
                            const double *G, const double *in, double *out,
                            uint32_t n_cells, void *stream);
mfx_status mfx_bk5_evec_f64(int p, const double *D, const double *G,
                            const double *in, double *out, uint32_t n_cells,
                            void *stream);

/* Gather / scatter halves of the element kernels (L-vector <-> E-vector),
 * steps 1 and 12 of CEED_bp/include/bk3_kokkos_kernel.h:113-143, 356-382:
 * evec[e][l] = map[e][l] == MFX_INVALID_DOF ? 0 : src[map[e][l]];
 * dst[map[e][l]] += evec[e][l] for valid entries. */
mfx_status mfx_gather_f64(const uint32_t *map, uint32_t n_cells, int nodes_per_cell,
                          const double *src, double *evec, void *stream);
mfx_status mfx_scatter_add_f64(const uint32_t *map, uint32_t n_cells, int nodes_per_cell,
                               const double *evec, double *dst, void *stream);

/* ----------------------------------------------------------------------
 * Operator handle: Portable::LaplaceOperator<dim,fe_degree,nq,number>
 * (CEED_bp/include/portable_laplace_operator.h:17-96) and
 * LaplaceOperatorBK35 (bakeoff_problems_dealii/include/portable_laplace_operator_bk35.h:16-443).
 * The handle owns the DoF map, geometry, basis tables and scratch; the
 * caller owns every vector.
 * ---------------------------------------------------------------------- */
typedef struct mfx_op mfx_op;

/* Structured box mesh [lo,hi] with n[d] cells per direction, optionally
 * deformed on the vertices by x_d += deform * sin(2 pi x) sin(2 pi y)
 * sin(2 pi z); homogeneous Dirichlet on the whole boundary.  This replaces
 * the set-up sequence GridGenerator::subdivided_hyper_rectangle +
 * refine_global + distribute_dofs + interpolate_boundary_values +
 * LaplaceOperator ctor (CEED_bp/src/bp3.cc:129-182, 452-488,
 * CEED_bp/include/portable_laplace_operator.h:97-122); mesh, map and
 * geometry are generated on the GPU. */
typedef struct {
  int degree;        /* p = 1..8 */
  int n_q_points_1d; /* nq = p+1 or p+2 */
  int quadrature;    /* mfx_quadrature */
  int geometry;      /* mfx_geometry_mode */
  uint32_t n_cells[3];
  double lo[3];
  double hi[3];
  double deform;
  int with_fp32; /* also build the float copies used by mfx_op_vmult_f32 */
} mfx_box_desc;

mfx_status mfx_op_create_box(const mfx_box_desc *desc, mfx_op **out);

/* General (unstructured) creation from host arrays: the data an adapter
 * takes from deal.II's MatrixFree (dof_indices of
 * setup_dirichlet_boundary_dofs_masks, CEED_bp/include/portable_laplace_operator.h:304-394)
 * plus the 8 vertices of every cell (lexicographic vertex order) and the
 * constrained DoF list (AffineConstraints). */
typedef struct {
  int degree;
  int n_q_points_1d;
  int quadrature;
  int geometry;
  uint32_t n_cells;
  uint64_t n_dofs;
  const uint32_t *map;          /* host [n_cells][(p+1)^3] */
  const double *cell_vertices;  /* host [n_cells][8][3] */
  const uint32_t *constrained;  /* host list of constrained DoFs */
  uint64_t n_constrained;
  int with_fp32;
} mfx_op_desc;

mfx_status mfx_op_create(const mfx_op_desc *desc, mfx_op **out);
mfx_status mfx_op_destroy(mfx_op *op);

/* m() / n() (CEED_bp/include/portable_laplace_operator.h:432-443). */
uint64_t mfx_op_n_dofs(const mfx_op *op);
uint32_t mfx_op_n_cells(const mfx_op *op);

/* Device pointers owned by the handle (for kernel-level calls and tests). */
typedef struct {
  int degree, n_q_points_1d, quadrature, geometry;
  uint32_t n_cells;
  uint64_t n_dofs;
  uint64_t n_constrained;
  const uint32_t *map;         /* [n_cells][(p+1)^3] */
  const double *G;             /* [n_cells][6][nq^3] (stored mode) */
  const double *B, *D;         /* basis tables */
  const double *cell_vertices; /* [n_cells][8][3] */
  const uint32_t *constrained; /* [n_constrained] */
  const double *xq, *wq;       /* 1D quadrature points/weights on [0,1] */
  /* box meshes with GLL collocation run the patch-of-cells kernels: */
  uint32_t n_patches;           /* 0 when the per-cell map kernel is used */
  int patch_cells[3];           /* cells per patch and direction */
  const uint32_t *patch_map;    /* [n_patches][(bx p+1)(by p+1)(bz p+1)], bit 31 = shared */
} mfx_op_info;
mfx_status mfx_op_get_info(const mfx_op *op, mfx_op_info *info);

/* vmult (CEED_bp/include/portable_laplace_operator.h:124-172): dst = 0;
 * dst += A_free src; dst_c = src_c on constrained DoFs
 * (copy_constrained_values).  Tvmult is the same call (A is symmetric,
 * :397-411).  src and dst must not alias. */
mfx_status mfx_op_vmult(mfx_op *op, double *dst, const double *src,
                        void *stream);
/* dst += A_free src over the unconstrained rows only (no zeroing, no
 * constrained copy): deal.II's vmult_add without the identity rows; the
 * operator's element kernel on its own. */
mfx_status mfx_op_vmult_add(mfx_op *op, double *dst, const double *src,
                            void *stream);
/* The same in FP32 (geometry rounded to float at set-up; needs with_fp32). */
mfx_status mfx_op_vmult_f32(mfx_op *op, float *dst, const float *src,
                            void *stream);

/* Deterministic scatter (on != 0): instead of adding its sums to dst with
 * atomics, each patch of cells stores them, and a second pass adds the
 * (up to 8) patch sums of every DoF in a fixed order, so applies -- and the
 * CG iterations on them -- are bitwise reproducible from run to run.  The
 * reference's scatter is atomic (CEED_bp/include/bk3_kokkos_kernel.h:356-382)
 * and not reproducible.  Operators on the patch kernel only (box meshes,
 * GLL collocation, either geometry mode); holds one value per patch DoF
 * (FP64, and FP32 with with_fp32) while on.  Synchronizes the device. */
mfx_status mfx_op_set_deterministic(mfx_op *op, int on);

/* Diagonal of the assembled operator (constrained rows: 1), the quantity
 * MatrixFreeTools::compute_diagonal fills at bp5_kokkos/benchmark.cc:231-236;
 * and its inverse per bp5_kokkos/benchmark.cc:240-250 (1/d if d > 0, else 1). */
mfx_status mfx_op_diagonal(mfx_op *op, double *diag, void *stream);
mfx_status mfx_op_inverse_diagonal(mfx_op *op, double *inv_diag, void *stream);

/* Right-hand side b_i = sum_q phi_i f JxW, constrained entries 0
 * (compute_rhs, CEED_bp/src/bp3.cc:184-239).  f_kind 0: f = 1;
 * f_kind 1: f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z). */
mfx_status mfx_op_rhs(mfx_op *op, int f_kind, double *b, void *stream);

/* Physical coordinates of every DoF, [n_dofs][3]. */
mfx_status mfx_op_dof_coordinates(mfx_op *op, double *coords, void *stream);

/* ----------------------------------------------------------------------
 * Conjugate gradients: deal.II SolverCG<Vector>::solve(A, x, b, P) with
 * ReductionControl(max_steps, abs_tol, rel_tol) (CEED_bp/src/bp3.cc:264-288,
 * PreconditionIdentity; bp5_kokkos/benchmark.cc:355-368, DiagonalMatrix).
 * inv_diag == NULL: identity preconditioner; otherwise Jacobi with the
 * given inverse diagonal.  x holds the start vector and receives the
 * solution.  Synchronous: returns after convergence or failure.
 * Stops at ||r_k||_2 <= max(abs_tol, rel_tol ||r_0||_2).
 * ---------------------------------------------------------------------- */
typedef struct {
  int steps;            /* SolverControl::last_step() */
  int converged;        /* 1 success, 0 failure */
  double initial_value; /* ||r_0||_2 */
  double last_value;    /* ||r_last||_2 */
  double seconds;       /* wall time of the solve */
} mfx_cg_result;

mfx_status mfx_cg_solve(mfx_op *op, double *x, const double *b,
                        const double *inv_diag, int max_steps, double abs_tol,
                        double rel_tol, mfx_cg_result *result, void *stream);

/* The same solve without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device or pinned host memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * The work vectors, partial sums and solver state live in a fixed device
 * array of the library (one per GPU, held until the process ends), so the
 * operator may have at most mfx_cg_noalloc_max_dofs() DoFs (2,097,152);
 * larger operators are rejected with MFX_ERR_INVALID_ARG before any work.
 * Calls that overlap on one GPU share that array and race.  Synchronous;
 * same arguments, stopping rule and results as mfx_cg_solve. */
mfx_status mfx_cg_solve_noalloc(mfx_op *op, double *x, const double *b,
                                const double *inv_diag, int max_steps, double abs_tol,
                                double rel_tol, mfx_cg_result *result, void *stream);
uint64_t mfx_cg_noalloc_max_dofs(void);

/* ----------------------------------------------------------------------
 * Vector helpers (deal.II LA::distributed::Vector device operations).
 * ---------------------------------------------------------------------- */
/* Synthetic input: uniform[-1,1) from SplitMix64 with the given seed. */
mfx_status mfx_vec_random(double *x, uint64_t n, uint64_t seed, void *stream);
/* Smooth synthetic fields: out_i = sum_m a_m sin(pi k_m0 x_i) sin(pi k_m1 y_i)
 * sin(pi k_m2 z_i) at points coords[n][3] (device), modes summed in index
 * order; wave_numbers[3 n_modes] and amplitudes[n_modes] are host arrays,
 * n_modes <= 64.  The right-hand sides of the learned preconditioner's
 * training pairs. */
mfx_status mfx_vec_sine_modes(const double *coords, uint64_t n, int n_modes,
                              const int32_t *wave_numbers, const double *amplitudes, double *out,
                              void *stream);
/* Set the listed entries to zero (constrained rows of a right-hand side). */
mfx_status mfx_vec_zero_constrained(mfx_op *op, double *x, void *stream);
/* Deterministic dot product; result written to *host_result (synchronous). */
mfx_status mfx_vec_dot(const double *x, const double *y, uint64_t n,
                       double *host_result, void *stream);

/* ----------------------------------------------------------------------
 * Mixed-precision multigrid.  No counterpart in the reference, which solves
 * with identity/Jacobi CG only (CEED_bp/src/bp3.cc:264-288; its multigrid
 * include at CEED_bp/src/bp3.cc:27 is unused).  The parameters follow
 * deal.II's PreconditionChebyshev::AdditionalData and the p/h-multigrid
 * preconditioners of deal.II's matrix-free solvers.
 *
 * Levels: Q_p, Q_{p/2}, ..., Q_1 on the fine cells, then Q_1 on n/2, n/4, ...
 * cells per direction while the free DoFs exceed coarse_max_dofs (every
 * halving needs even cell counts).  Under a Gauss fine operator the levels
 * below it use Gauss q = p+1; a GLL hierarchy keeps the fine q - p.  The coarsest level is solved directly in
 * FP64; every other level smooths with Chebyshev(Jacobi) in FP32.  Transfers
 * are nodal interpolation P and its transpose R = P^T.
 * ---------------------------------------------------------------------- */
typedef struct mfx_mg mfx_mg;

typedef struct {
  mfx_box_desc fine;        /* mesh, degree, quadrature, geometry of the fine level */
  int smoother_degree;      /* Chebyshev steps per smoothing; 0: 5 */
  double smoothing_range;   /* lambda_min = lambda_max / range; 0: 15 */
  int eig_cg_n_iterations;  /* CG (Lanczos) steps of the lambda_max estimate; 0: 12 */
  double eig_safety;        /* lambda_max = safety x estimate; 0: 1.2 */
  uint32_t coarse_max_dofs; /* largest free-DoF count solved directly; 0: 2000 */
  /* Gauss points per direction of the fine level's FP32 smoothing operator
   * (Chebyshev and the residual restricted to the next level); 0: the fine
   * operator's own rule.  Under a Gauss q = p+2 fine operator, q = p+1
   * integrates the stiffness of affine cells exactly at 58% of the points;
   * the outer FP64 CG always applies the fine operator itself. */
  int fine_smoother_n_q_points_1d;
} mfx_mg_desc;

typedef struct {
  int degree, n_q_points_1d;
  uint32_t n_cells[3];
  uint64_t n_dofs;
  double lambda_max, lambda_min; /* Chebyshev interval (0 on the coarsest level) */
  mfx_op *op;                    /* level operator, owned by the multigrid object */
} mfx_mg_level_info;

mfx_status mfx_mg_create(const mfx_mg_desc *desc, mfx_mg **out);
mfx_status mfx_mg_destroy(mfx_mg *mg);
int mfx_mg_n_levels(const mfx_mg *mg);
/* level 0 is the fine level; its operator is the FP64 operator of the solve */
mfx_status mfx_mg_get_level(const mfx_mg *mg, int level, mfx_mg_level_info *info);
/* Preconditioner application (deal.II PreconditionMG::vmult): dst = one
 * FP32 V-cycle from a zero guess applied to src; constrained rows pass
 * through.  FP64 in and out. */
mfx_status mfx_mg_vmult(mfx_mg *mg, double *dst, const double *src, void *stream);
/* FP64 CG preconditioned by one V-cycle per step, on the fine operator; same
 * stopping rule and result fields as mfx_cg_solve.  Synchronous. */
mfx_status mfx_mg_solve(mfx_mg *mg, double *x, const double *b, int max_steps,
                        double abs_tol, double rel_tol, mfx_cg_result *result,
                        void *stream);

/* Transfers between two operators made by mfx_op_create_box on the same box:
 * same cells and lower degree (p-transfer), or same degree and half the
 * cells per direction (h-transfer).  Vectors are FP64 on the device. */
mfx_status mfx_transfer_prolongate_f64(const mfx_op *fine, const mfx_op *coarse,
                                       double *dst_fine, const double *src_coarse,
                                       void *stream);
mfx_status mfx_transfer_restrict_f64(const mfx_op *fine, const mfx_op *coarse,
                                     double *dst_coarse, const double *src_fine,
                                     void *stream);
/* Chebyshev smoother steps on an FP32 operator (created with with_fp32): x
 * receives `degree` Chebyshev(Jacobi) updates towards A x = b on the
 * interval [lambda_min, lambda_max]; x_is_zero skips the first operator
 * application.  Synchronous. */
mfx_status mfx_chebyshev_f32(mfx_op *op, const float *inv_diag, float *x, const float *b,
                             int degree, double lambda_min, double lambda_max,
                             int x_is_zero, void *stream);

/* ----------------------------------------------------------------------
 * Batched solves.  No counterpart in the reference, which runs one problem
 * per process (CEED_bp/src/bp3.cc:423-534).  Many independent BP5 problems
 * (GLL collocation q = p+1, degree 2..6, stored geometry) are concatenated
 * into one DoF vector, instance i occupying [offsets[i], offsets[i+1]).
 * mfx_batch_vmult applies every instance's operator (vmult semantics per
 * instance); mfx_batch_cg_solve runs an independent Jacobi-CG per instance
 * (mfx_cg_solve's stopping rule and result fields, per instance) as a
 * hipGraph of CG steps replayed until all have converged.
 * ---------------------------------------------------------------------- */
typedef struct mfx_batch mfx_batch;

typedef struct {
  int degree;           /* 2..6 */
  uint32_t n_cells[3];
  double lo[3], hi[3];
  double deform;
} mfx_batch_instance;

typedef struct {
  int degree;
  uint32_t n_cells;            /* cells of this degree over all instances */
  int n_variants;              /* tile variants of the collocation kernel */
  int variant;                 /* the one in use */
  int default_variant;         /* the library's default tile */
  int cells_per_workgroup[8];  /* per variant (cells per patch for the patch kernel) */
  int prefetch[8];             /* per variant: metric terms loaded ahead */
  int geometry[8];             /* per variant: MFX_GEOM_STORED (collocation kernel tiles) or
                                  MFX_GEOM_ON_THE_FLY (the persistent patch kernel, last) */
  double ms[8];                /* per variant: measured apply time (0 if not tuned) */
} mfx_batch_tuning;

/* autotune != 0: time every variant per degree (the collocation kernel's
 * tiles with stored metric terms, and the patch kernel with metric terms
 * from the cell vertices) on the batch's cells and keep the fastest;
 * otherwise the default tile */
mfx_status mfx_batch_create(const mfx_batch_instance *instances, int n_instances,
                            int autotune, mfx_batch **out);
mfx_status mfx_batch_destroy(mfx_batch *batch);
int mfx_batch_n_instances(const mfx_batch *batch);
uint64_t mfx_batch_n_dofs(const mfx_batch *batch);
mfx_status mfx_batch_offsets(const mfx_batch *batch, uint64_t *offsets /* n + 1 */);
/* the instance's own operator (owned by the batch) */
mfx_op *mfx_batch_instance_operator(mfx_batch *batch, int instance);
mfx_status mfx_batch_get_tuning(const mfx_batch *batch, int degree, mfx_batch_tuning *tuning);
/* use variant `variant` (0 .. n_variants - 1 of the tuning record) for the
 * instances of `degree` */
mfx_status mfx_batch_set_variant(mfx_batch *batch, int degree, int variant);
mfx_status mfx_batch_vmult(mfx_batch *batch, double *dst, const double *src, void *stream);
mfx_status mfx_batch_inverse_diagonal(mfx_batch *batch, double *inv, void *stream);
/* inv_diag NULL: identity preconditioner.  results: one per instance.
 * use_graph = 0 issues the same launches directly. */
mfx_status mfx_batch_cg_solve(mfx_batch *batch, double *x, const double *b,
                              const double *inv_diag, int max_steps, double abs_tol,
                              double rel_tol, int use_graph, mfx_cg_result *results,
                              void *stream);

/* ----------------------------------------------------------------------
 * Learned preconditioner.  No counterpart in the reference.  A 4-level 3D
 * U-Net (1 -> 16 -> 32 -> 64 channels, 3^3 kernels, additive skips, ReLU,
 * no biases; 360288 FP32 master weights in layer order, each [cout][27][cin])
 * on the node grid of a box operator, mapping a residual r to an error
 * estimate: z = ||r|| s net(r / ||r||).  Forward on BF16 MFMA with FP32
 * accumulation, backward in FP32, training with Adam (b1 0.9, b2 0.999),
 * all on the GPU.  Grids are node counts {nx, ny, nz}; fields are FP32
 * device arrays, x fastest.
 * ---------------------------------------------------------------------- */
typedef struct mfx_nn mfx_nn;

uint64_t mfx_nn_n_params(void);
/* He-normal weights from the SplitMix64 stream of seed */
mfx_status mfx_nn_create(uint64_t seed, mfx_nn **out);
mfx_status mfx_nn_destroy(mfx_nn *nn);
mfx_status mfx_nn_get_params(const mfx_nn *nn, float *host_params);
mfx_status mfx_nn_set_params(mfx_nn *nn, const float *host_params);
mfx_status mfx_nn_set_output_scale(mfx_nn *nn, double scale);
double mfx_nn_output_scale(const mfx_nn *nn);
/* Preconditioner mode: smoother_degree 0 applies the network alone; k > 0
 * runs k Chebyshev(Jacobi) steps (FP32, interval [lambda_max/15, lambda_max],
 * lambda_max = 1.2 x a 12-step Lanczos estimate; the operator needs
 * with_fp32), then the network on the remaining residual as a coarse
 * correction, then k Chebyshev steps again.  use_network = 0 keeps only the
 * Chebyshev steps (the method without the network). */
mfx_status mfx_nn_set_mode(mfx_nn *nn, int smoother_degree, int use_network);
/* Coarse correction on another box operator of the same box (cells of the
 * fine operator an integer refinement of its cells, and/or a lower degree,
 * e.g. Q4 on n^3 cells over Q1 on (n/2)^3): with smoother_degree > 0 the
 * network then runs on that operator's node grid, on the restricted
 * residual, and its output is prolongated (nodal interpolation and its
 * transpose) before the energy-optimal weighting.  The caller keeps
 * `coarse` alive; NULL returns to the fine grid. */
mfx_status mfx_nn_set_coarse(mfx_nn *nn, mfx_op *coarse);
/* The smoother's Chebyshev interval on `op` (set up on first use) and the
 * energy-optimal weight of the network's correction in the last
 * application (0 when the network did not contribute) */
mfx_status mfx_nn_smoother_info(mfx_nn *nn, mfx_op *op, double *lambda_min, double *lambda_max,
                                double *last_omega, void *stream);
/* y = net(x) on one grid */
mfx_status mfx_nn_forward(mfx_nn *nn, const uint32_t dims[3], const float *x, float *y,
                          void *stream);
/* grad (device, n_params floats) = d(sum dy . net(x)) / d(weights) */
mfx_status mfx_nn_gradient(mfx_nn *nn, const uint32_t dims[3], const float *x, const float *dy,
                           float *grad, void *stream);
/* one Adam step on the mean over `batch` grids of the mean squared error of
 * net(x_b) against target_b (x, target: batch consecutive grids); *loss
 * receives that mean before the step */
mfx_status mfx_nn_train_step(mfx_nn *nn, const uint32_t dims[3], int batch, const float *x,
                             const float *target, double lr, double *loss, void *stream);
/* one Adam step on the energy-norm loss over `batch` whole fields of a box
 * operator: mean of (pred - e)^T A (pred - e) / e^T A e with
 * pred = ||r|| s net(r / ||r||); r, e: batch consecutive FP64 DoF vectors
 * with A e = r; *loss receives that mean before the step */
mfx_status mfx_nn_train_energy(mfx_nn *nn, mfx_op *op, int batch, const double *r,
                               const double *e, double lr, double *loss, void *stream);
/* preconditioner application on a box operator's node grid (FP64 in/out,
 * constrained rows pass through) */
mfx_status mfx_nn_vmult(mfx_nn *nn, mfx_op *op, double *z, const double *r, void *stream);
/* The residual the two-level cycle hands the network, on the coarse
 * operator's DoFs (FP64 out, rho_c of mfx_op_n_dofs(coarse) entries):
 * rho_c = R (r - A z), z the cycle's FP32 Chebyshev pre-smoothing of A z = r
 * from zero (constrained rows of r taken as 0, those of rho_c set to 0), R
 * the transpose of the nodal interpolation.  Needs smoother degree > 0 and a
 * coarse operator (mfx_nn_set_coarse).  Training pairs (rho_c,
 * A_c^-1 rho_c) made from it are what the network meets in the solve. */
mfx_status mfx_nn_restricted_residual(mfx_nn *nn, mfx_op *op, const double *r, double *rho_c,
                                      void *stream);
/* flexible CG (up to 4 stored A-orthogonal directions) on op preconditioned
 * by the network;
 * stopping rule and result fields of mfx_cg_solve */
mfx_status mfx_nn_fcg_solve(mfx_nn *nn, mfx_op *op, double *x, const double *b, int max_steps,
                            double abs_tol, double rel_tol, mfx_cg_result *result,
                            void *stream);
/* Fallback of mfx_nn_fcg_solve, the counterpart of the reference solver's
 * only failure path (NoConvergence, bp5_kokkos/benchmark.cc:366-374): when
 * the residual falls by less than a factor min_reduction over the last
 * `window` steps, or the network's correction has been rejected (energy-
 * optimal weight <= 0) in `window` applications in a row, or a direction
 * breaks down, the solve continues with one V-cycle of `mg` per step (its
 * fine operator must be the solve's operator; the caller keeps it alive).
 * mg == NULL: no fallback.  window 0: 8; min_reduction 0: 2. */
mfx_status mfx_nn_set_fallback(mfx_nn *nn, mfx_mg *mg, int window, double min_reduction);
/* Step at which the last mfx_nn_fcg_solve switched to the fallback
 * (-1: it did not). */
int mfx_nn_fallback_step(const mfx_nn *nn);

#ifdef __cplusplus
}
#endif
#endif /* MFX_H */
