/*
 * mfx_oracle.h -- CPU oracle for the matrix-free Laplacian hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline).  The product library
 * (benchmarks_amd/libmfx.so) never links or calls it.
 *
 * Conventions (build-native, shared with the product):
 *   - reference cell [0,1]^3; tensor index x fastest: l = a + n*(b + n*c);
 *   - geometry G[e][c][q], c = (xx, xy, xz, yy, yz, zz), axis 0 = x;
 *     G = w_q * det(J) * A * A^T with A = J^{-1};
 *   - DoF map uint32 [E][(p+1)^3], constrained DoFs = 0xFFFFFFFF;
 *   - B[q*nm + m] = l_m(x_q) (nodal basis m at quadrature point q);
 *     D[q*nq + k] = l'_k(x_q) (collocation derivative on quadrature points).
 *
 * Parity status of every function is stated in its comment in mfx_oracle.c.
 */
#ifndef MFX_ORACLE_H
#define MFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_INVALID 0xFFFFFFFFu
#define ORC_GAUSS 0
#define ORC_GLL 1

/* ---- 1D rules and Lagrange basis ---- */
int orc_gauss_rule(int n, double *x, double *w);
int orc_gll_rule(int n, double *x, double *w);
double orc_lagrange_value(const double *nodes, int n, int j, double x);
double orc_lagrange_derivative(const double *nodes, int n, int j, double x);
int orc_basis_tables(int p, int nq, int quad_kind, double *B, double *D,
                     double *xq, double *wq);

/* ---- reference-layout E-vector kernels (bitwise restatements) ---- */
double orc_bk3_serial_ref(unsigned nq, unsigned nelmt, const double *basis,
                          const double *dbasis, const double *G,
                          const double *in, double *out);
double orc_bk5_serial_ref(unsigned nq, unsigned nelmt, const double *dbasis,
                          const double *G, const double *in, double *out);

/* ---- native-layout E-vector kernels ---- */
void orc_bk3_evec(int nm, int nq, uint64_t nelmt, const double *B,
                  const double *D, const double *G, const double *in,
                  double *out);
void orc_bk5_evec(int nq, uint64_t nelmt, const double *D, const double *G,
                  const double *in, double *out);

/* ---- structured box mesh ---- */
typedef struct {
  uint32_t n[3];     /* cells per direction */
  double lo[3];      /* box corners */
  double hi[3];
  double deform;     /* amplitude a of x_d += a sin(2 pi x) sin(2 pi y) sin(2 pi z) */
} orc_box;

uint64_t orc_box_n_cells(const orc_box *box);
uint64_t orc_box_n_dofs(const orc_box *box, int p);
void orc_box_vertices(const orc_box *box, double *X);
void orc_box_dof_map(const orc_box *box, int p, uint32_t *map,
                     uint8_t *constrained);
void orc_box_cell_vertices(const orc_box *box, const double *X, double *cellX);
void orc_geometry(uint64_t E, const double *cellX, int nq, const double *xq,
                  const double *wq, double *G);
void orc_dof_coordinates(const orc_box *box, int p, const double *cellX,
                         const uint32_t *map, double *coords);

/* ---- L-vector operator ---- */
void orc_vmult(int nm, int nq, const double *B, const double *D,
               const double *G, const uint32_t *map, uint64_t E, uint64_t N,
               const uint8_t *constrained, const double *src, double *dst,
               int threads);
void orc_gather(int nm, const uint32_t *map, uint64_t E, const double *src,
                double *evec);
void orc_diagonal(int nm, int nq, const double *B, const double *D,
                  const double *G, const uint32_t *map, uint64_t E,
                  uint64_t N, const uint8_t *constrained, double *diag);
void orc_invert_diagonal(uint64_t N, double *diag);

/* ---- right-hand side: f_kind 0: f = 1; 1: f = 3 pi^2 prod sin(pi x_d) ---- */
void orc_rhs(int nm, int nq, const double *B, const double *xq,
             const double *wq, const double *cellX, const uint32_t *map,
             uint64_t E, uint64_t N, int f_kind, double *b);

/* ---- conjugate gradients, deal.II SolverCG + ReductionControl semantics ---- */
typedef struct {
  int steps;            /* last_step(): iterations performed */
  int converged;        /* 1 success, 0 failure (max steps / NaN) */
  double initial_value; /* ||r_0||_2 */
  double last_value;    /* ||r_last||_2 */
} orc_cg_result;

orc_cg_result orc_cg(int nm, int nq, const double *B, const double *D,
                     const double *G, const uint32_t *map, uint64_t E,
                     uint64_t N, const uint8_t *constrained,
                     const double *inv_diag, const double *b, double *x,
                     int max_steps, double abs_tol, double rel_tol,
                     double *res_history, int threads);

/* L2 error of the nodal field against sin(pi x)sin(pi y)sin(pi z) */
double orc_l2_error_sin(int nm, int nq, const double *B, const double *xq,
                        const double *wq, const double *cellX,
                        const uint32_t *map, uint64_t E, const double *u);

#ifdef __cplusplus
}
#endif
#endif
