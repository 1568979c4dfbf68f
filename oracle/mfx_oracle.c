/*
 * mfx_oracle.c -- CPU restatement of the reference's matrix-free Laplacian
 * path.  TEST INFRASTRUCTURE ONLY: the product library never calls this code;
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it as
 * the checker.  Built with -ffp-contract=off so that the reference-layout
 * kernels reproduce the reference serial kernels bit for bit.
 *
 * Parity status per function (how each is pinned):
 *   orc_gauss_rule / orc_gll_rule      polynomial exactness, known GLL table,
 *                                      symmetry (tests/test_oracle_basis.py)
 *   orc_lagrange_value / _derivative   bitwise vs the reference
 *                                      LagrangeBasis built in oracle/_ref, and
 *                                      partition of unity (reference KAT)
 *   orc_bk3_serial_ref / bk5_...       bitwise vs the reference serial kernels
 *                                      (oracle/_ref) and committed golden files
 *   orc_bk3_evec / orc_bk5_evec        bitwise vs the reference-layout kernels
 *                                      through the layout adapter; dense
 *                                      element-matrix brute force for p <= 3
 *   orc_box_* / orc_geometry           DoF-count KAT, node-coordinate brute
 *                                      force, affine closed form of G,
 *                                      finite-difference Jacobian
 *   orc_vmult                          assembled-matrix brute force, identity
 *                                      rows, K*1 = 0, symmetry
 *   orc_diagonal                       diagonal of the brute-force matrix
 *   orc_rhs                            f = 1 integrates to the domain volume;
 *                                      manufactured-solution convergence
 *   orc_cg                             CG iteration-count KAT from the
 *                                      reference logs; h^(p+1) convergence
 */
#include "mfx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* 1D quadrature rules on [0,1]                                        */
/* ------------------------------------------------------------------ */

/* Legendre P_n(x) and P_{n-1}(x) by the three-term recurrence. */
static void legendre_pair(int n, long double x, long double *pn,
                          long double *pn1) {
  long double p0 = 1.0L, p1 = x;
  if (n == 0) {
    *pn = 1.0L;
    *pn1 = 0.0L;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *pn1 = p0;
}

/* Gauss-Legendre points: roots of P_n by Newton's method from the
 * asymptotic guess cos(pi (i + 3/4) / (n + 1/2)); weights
 * 2 / ((1 - x^2) P_n'(x)^2); mapped to [0,1] (deal.II's QGauss on the unit
 * interval).  Parity unpinned bitwise against deal.II (not installed); pinned
 * by exactness for polynomials of degree 2n-1 and by symmetry. */
int orc_gauss_rule(int n, double *x, double *w) {
  if (n < 1) return -1;
  long double xs[64], ws[64];
  if (n > 64) return -1;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double z = cosl(3.14159265358979323846264338327950288L *
                         (i + 0.75L) / (n + 0.5L));
    long double dp = 1.0L;
    for (int it = 0; it < 100; ++it) {
      long double pn, pn1;
      legendre_pair(n, z, &pn, &pn1);
      dp = n * (z * pn - pn1) / (z * z - 1.0L);
      long double dz = pn / dp;
      z -= dz;
      if (fabsl(dz) < 1e-30L) break;
    }
    {
      long double pn, pn1;
      legendre_pair(n, z, &pn, &pn1);
      dp = n * (z * pn - pn1) / (z * z - 1.0L);
    }
    long double wt = 2.0L / ((1.0L - z * z) * dp * dp);
    /* z is the i-th largest root; mirror for exact symmetry */
    xs[n - 1 - i] = z;
    xs[i] = -z;
    ws[n - 1 - i] = wt;
    ws[i] = wt;
  }
  if (n % 2 == 1) xs[n / 2] = 0.0L;
  for (int i = 0; i < n; ++i) {
    x[i] = (double)((xs[i] + 1.0L) * 0.5L);
    w[i] = (double)(ws[i] * 0.5L);
  }
  return 0;
}

/* Gauss-Lobatto-Legendre points: the end points and the roots of P_m',
 * m = n - 1, by Newton's method on P_m' with P_m'' from Legendre's equation;
 * weights 2 / (m (m + 1) P_m(x)^2); mapped to [0,1].  deal.II's FE_Q support
 * points and QGaussLobatto.  Pinned by the known p = 2 table, exactness for
 * degree 2n-3 and symmetry. */
int orc_gll_rule(int n, double *x, double *w) {
  if (n < 2 || n > 64) return -1;
  const int m = n - 1;
  long double xs[64], ws[64];
  const long double pi = 3.14159265358979323846264338327950288L;
  xs[0] = -1.0L;
  xs[n - 1] = 1.0L;
  for (int i = 1; i <= (n - 1) / 2; ++i) {
    /* i-th interior root counted from the right */
    long double z = cosl(pi * i / m);
    for (int it = 0; it < 100; ++it) {
      long double pm, pm1;
      legendre_pair(m, z, &pm, &pm1);
      long double d1 = m * (z * pm - pm1) / (z * z - 1.0L);
      long double d2 = (2.0L * z * d1 - m * (m + 1.0L) * pm) / (1.0L - z * z);
      long double dz = d1 / d2;
      z -= dz;
      if (fabsl(dz) < 1e-30L) break;
    }
    xs[n - 1 - i] = z;
    xs[i] = -z;
  }
  if (n % 2 == 1) xs[n / 2] = 0.0L;
  for (int i = 0; i < n; ++i) {
    long double pm, pm1;
    legendre_pair(m, xs[i], &pm, &pm1);
    ws[i] = 2.0L / (m * (m + 1.0L) * pm * pm);
  }
  for (int i = 0; i < n / 2; ++i) {
    long double avg = 0.5L * (ws[i] + ws[n - 1 - i]);
    ws[i] = avg;
    ws[n - 1 - i] = avg;
  }
  for (int i = 0; i < n; ++i) {
    x[i] = (double)((xs[i] + 1.0L) * 0.5L);
    w[i] = (double)(ws[i] * 0.5L);
  }
  x[0] = 0.0;
  x[n - 1] = 1.0;
  return 0;
}

/* Restates LagrangeBasis::LagrangeBasis + operator()
 * (sum_factorization/include/math/lagrange_basis.hpp:10-42) in the same
 * operation order.  Pinned bitwise against the reference header. */
double orc_lagrange_value(const double *c, int n, int j, double x) {
  double denominator = 1;
  for (int m = 0; m < n; ++m)
    if (j != m) denominator *= c[j] - c[m];
  double result = 1;
  for (int m = 0; m < n; m++)
    if (m != j) result *= (x - c[m]);
  result = result / denominator;
  return result;
}

/* Restates LagrangeBasis::derivative
 * (sum_factorization/include/math/lagrange_basis.hpp:44-65). */
double orc_lagrange_derivative(const double *c, int n, int j, double x) {
  double sum = 0;
  double product;
  for (int i = 0; i < n; ++i) {
    product = 1;
    if (i != j) {
      for (int m = 0; m < n; ++m)
        if (m != i && m != j) product *= (x - c[m]) / (c[j] - c[m]);
      sum += (1 / (c[j] - c[i])) * product;
    }
  }
  return sum;
}

/* Shape data as deal.II's MatrixFree hands it to the reference kernel
 * (shape_values / co_shape_gradients, CEED_bp/include/portable_laplace_operator.h:147,
 * used at CEED_bp/include/bk3_kokkos_kernel.h:156-161, 243-245): nodal FE_Q
 * basis on p+1 GLL support points evaluated at the nq quadrature points, and
 * the collocation derivative of the Lagrange basis through the quadrature
 * points.  Layout here: B[q*nm+m], D[q*nq+k]. */
int orc_basis_tables(int p, int nq, int quad_kind, double *B, double *D,
                     double *xq, double *wq) {
  const int nm = p + 1;
  double sup[64], sw[64];
  if (p < 1 || nq < 1 || nq > 64) return -1;
  if (orc_gll_rule(nm, sup, sw)) return -1;
  if (quad_kind == ORC_GAUSS) {
    if (orc_gauss_rule(nq, xq, wq)) return -1;
  } else {
    if (orc_gll_rule(nq, xq, wq)) return -1;
  }
  for (int q = 0; q < nq; ++q)
    for (int m = 0; m < nm; ++m)
      B[q * nm + m] = orc_lagrange_value(sup, nm, m, xq[q]);
  for (int q = 0; q < nq; ++q)
    for (int k = 0; k < nq; ++k)
      D[q * nq + k] = orc_lagrange_derivative(xq, nq, k, xq[q]);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Reference-layout serial kernels                                     */
/* ------------------------------------------------------------------ */

/* Restates BK3::Serial::SumFactorization
 * (CEED_BK/include/kernels/BK3/serial_kernels.hpp:9-197) for an isotropic
 * element, nm = nq - 1: same layouts (in/out[e][i][j][k], basis[p*nm+i],
 * dbasis[p*nq+n], G[e][p][q][c][r]), same loop nests, same summation order.
 * Returns sum(out^2) summed in index order (the reference reduces with
 * std::transform_reduce, whose association is unspecified). */
double orc_bk3_serial_ref(unsigned nq, unsigned nelmt, const double *basis,
                          const double *dbasis, const double *G,
                          const double *in, double *out) {
  const unsigned nm = nq - 1;
  const unsigned nq3 = nq * nq * nq;
  double *wsp0 = (double *)malloc(sizeof(double) * nq3);
  double *wsp1 = (double *)malloc(sizeof(double) * nq3);
  double *rqr = (double *)malloc(sizeof(double) * nq3);
  double *rqs = (double *)malloc(sizeof(double) * nq3);
  double *rqt = (double *)malloc(sizeof(double) * nq3);
  for (unsigned e = 0; e < nelmt; ++e) {
    const double *ine = in + (size_t)e * nm * nm * nm;
    /* interpolation, direction 0 (slowest index i) */
    for (unsigned p = 0; p < nq; p++)
      for (unsigned k = 0; k < nm; k++)
        for (unsigned j = 0; j < nm; j++) {
          double acc = 0;
          for (unsigned i = 0; i < nm; i++)
            acc += ine[i * nm * nm + j * nm + k] * basis[p * nm + i];
          wsp1[p * nm * nm + j * nm + k] = acc;
        }
    /* direction 1 -> layout [q][p][k] */
    for (unsigned q = 0; q < nq; q++)
      for (unsigned p = 0; p < nq; p++)
        for (unsigned k = 0; k < nm; k++) {
          double acc = 0;
          for (unsigned j = 0; j < nm; j++)
            acc += wsp1[p * nm * nm + j * nm + k] * basis[q * nm + j];
          wsp0[q * nq * nm + p * nm + k] = acc;
        }
    /* direction 2 -> layout [p][q][r] */
    for (unsigned r = 0; r < nq; r++)
      for (unsigned q = 0; q < nq; q++)
        for (unsigned p = 0; p < nq; p++) {
          double acc = 0;
          for (unsigned k = 0; k < nm; k++)
            acc += wsp0[q * nq * nm + p * nm + k] * basis[r * nm + k];
          wsp1[p * nq * nq + q * nq + r] = acc;
        }
    for (unsigned p = 0; p < nq; ++p)
      for (unsigned q = 0; q < nq; ++q)
        for (unsigned r = 0; r < nq; ++r) {
          const size_t gb = (size_t)e * nq * nq * 6 * nq + p * nq * 6 * nq + q * 6 * nq + r;
          const double Grr = G[gb + 0 * nq], Grs = G[gb + 1 * nq], Grt = G[gb + 2 * nq];
          const double Gss = G[gb + 3 * nq], Gst = G[gb + 4 * nq], Gtt = G[gb + 5 * nq];
          double qr = 0.0, qs = 0.0, qt = 0.0;
          for (unsigned n = 0; n < nq; ++n) qr += wsp1[n * nq * nq + q * nq + r] * dbasis[p * nq + n];
          for (unsigned n = 0; n < nq; ++n) qs += wsp1[p * nq * nq + n * nq + r] * dbasis[q * nq + n];
          for (unsigned n = 0; n < nq; ++n) qt += wsp1[p * nq * nq + q * nq + n] * dbasis[r * nq + n];
          rqr[p * nq * nq + q * nq + r] = Grr * qr + Grs * qs + Grt * qt;
          rqs[p * nq * nq + q * nq + r] = Grs * qr + Gss * qs + Gst * qt;
          rqt[p * nq * nq + q * nq + r] = Grt * qr + Gst * qs + Gtt * qt;
        }
    for (unsigned p = 0; p < nq; ++p)
      for (unsigned q = 0; q < nq; ++q)
        for (unsigned r = 0; r < nq; ++r) {
          double tmp0 = 0;
          for (unsigned n = 0; n < nq; ++n) tmp0 += rqr[n * nq * nq + q * nq + r] * dbasis[n * nq + p];
          for (unsigned n = 0; n < nq; ++n) tmp0 += rqs[p * nq * nq + n * nq + r] * dbasis[n * nq + q];
          for (unsigned n = 0; n < nq; ++n) tmp0 += rqt[p * nq * nq + q * nq + n] * dbasis[n * nq + r];
          wsp1[p * nq * nq + q * nq + r] = tmp0;
        }
    /* transposed interpolation, direction 2 -> [q][p][k] */
    for (unsigned k = 0; k < nm; k++)
      for (unsigned q = 0; q < nq; q++)
        for (unsigned p = 0; p < nq; p++) {
          double acc = 0;
          for (unsigned r = 0; r < nq; r++)
            acc += wsp1[p * nq * nq + q * nq + r] * basis[r * nm + k];
          wsp0[q * nq * nm + p * nm + k] = acc;
        }
    /* direction 1 -> [p][j][k] */
    for (unsigned j = 0; j < nm; j++)
      for (unsigned k = 0; k < nm; k++)
        for (unsigned p = 0; p < nq; p++) {
          double acc = 0;
          for (unsigned q = 0; q < nq; q++)
            acc += wsp0[q * nq * nm + p * nm + k] * basis[q * nm + j];
          wsp1[p * nm * nm + j * nm + k] = acc;
        }
    /* direction 0 -> out[i][j][k] */
    double *oute = out + (size_t)e * nm * nm * nm;
    for (unsigned i = 0; i < nm; i++)
      for (unsigned j = 0; j < nm; j++)
        for (unsigned k = 0; k < nm; k++) {
          double acc = 0;
          for (unsigned p = 0; p < nq; p++)
            acc += wsp1[p * nm * nm + j * nm + k] * basis[p * nm + i];
          oute[i * nm * nm + j * nm + k] = acc;
        }
  }
  free(wsp0);
  free(wsp1);
  free(rqr);
  free(rqs);
  free(rqt);
  double s = 0;
  for (size_t i = 0; i < (size_t)nelmt * nm * nm * nm; ++i) s += out[i] * out[i];
  return s;
}

/* Restates BK5::Serial::SumFactorization
 * (sum_factorization/include/kernels/BK5/serial_kernels.hpp:9-87; the copy
 * whose G loads read all six components, unlike
 * CEED_BK/include/kernels/BK5/serial_kernels.hpp:28-33).  Layouts
 * in/out[e][i][j][k], dbasis[i*nq+n], G[e][i][j][c][k]. */
double orc_bk5_serial_ref(unsigned nq, unsigned nelmt, const double *dbasis,
                          const double *G, const double *in, double *out) {
  const unsigned nq3 = nq * nq * nq;
  double *rqr = (double *)malloc(sizeof(double) * nq3);
  double *rqs = (double *)malloc(sizeof(double) * nq3);
  double *rqt = (double *)malloc(sizeof(double) * nq3);
  for (unsigned e = 0; e < nelmt; ++e) {
    const double *ine = in + (size_t)e * nq3;
    for (unsigned i = 0; i < nq; ++i)
      for (unsigned j = 0; j < nq; ++j)
        for (unsigned k = 0; k < nq; ++k) {
          const size_t gb = (size_t)e * nq * nq * 6 * nq + i * nq * 6 * nq + j * 6 * nq + k;
          const double Grr = G[gb + 0 * nq], Grs = G[gb + 1 * nq], Grt = G[gb + 2 * nq];
          const double Gss = G[gb + 3 * nq], Gst = G[gb + 4 * nq], Gtt = G[gb + 5 * nq];
          double qr = 0.0, qs = 0.0, qt = 0.0;
          for (unsigned n = 0; n < nq; ++n) qr += ine[n * nq * nq + j * nq + k] * dbasis[i * nq + n];
          for (unsigned n = 0; n < nq; ++n) qs += ine[i * nq * nq + n * nq + k] * dbasis[j * nq + n];
          for (unsigned n = 0; n < nq; ++n) qt += ine[i * nq * nq + j * nq + n] * dbasis[k * nq + n];
          rqr[i * nq * nq + j * nq + k] = Grr * qr + Grs * qs + Grt * qt;
          rqs[i * nq * nq + j * nq + k] = Grs * qr + Gss * qs + Gst * qt;
          rqt[i * nq * nq + j * nq + k] = Grt * qr + Gst * qs + Gtt * qt;
        }
    double *oute = out + (size_t)e * nq3;
    for (unsigned i = 0; i < nq; ++i)
      for (unsigned j = 0; j < nq; ++j)
        for (unsigned k = 0; k < nq; ++k) {
          double tmp0 = 0;
          for (unsigned n = 0; n < nq; ++n) tmp0 += rqr[n * nq * nq + j * nq + k] * dbasis[n * nq + i];
          for (unsigned n = 0; n < nq; ++n) tmp0 += rqs[i * nq * nq + n * nq + k] * dbasis[n * nq + j];
          for (unsigned n = 0; n < nq; ++n) tmp0 += rqt[i * nq * nq + j * nq + n] * dbasis[n * nq + k];
          oute[i * nq * nq + j * nq + k] = tmp0;
        }
  }
  free(rqr);
  free(rqs);
  free(rqt);
  double s = 0;
  for (size_t i = 0; i < (size_t)nelmt * nq3; ++i) s += out[i] * out[i];
  return s;
}

/* ------------------------------------------------------------------ */
/* Native-layout E-vector kernels                                      */
/* ------------------------------------------------------------------ */
/* The element cube u[c][b][a] (x index a fastest) is the same memory as the
 * reference's in[i][j][k] with i <-> z, j <-> y, k <-> x; the reference's
 * metric components (rr, rs, rt, ss, st, tt) are therefore this build's
 * (zz, yz, xz, yy, xy, xx).  Only the G addressing differs from the
 * reference-layout kernels, so both produce identical bits. */
enum { GXX = 0, GXY = 1, GXZ = 2, GYY = 3, GYZ = 4, GZZ = 5 };

static void bk3_evec_one(int nm, int nq, const double *basis,
                         const double *dbasis, const double *Ge,
                         const double *ine, double *oute, double *wsp0,
                         double *wsp1, double *rqr, double *rqs, double *rqt) {
  const int nq3 = nq * nq * nq;
  for (int p = 0; p < nq; p++)
    for (int k = 0; k < nm; k++)
      for (int j = 0; j < nm; j++) {
        double acc = 0;
        for (int i = 0; i < nm; i++) acc += ine[i * nm * nm + j * nm + k] * basis[p * nm + i];
        wsp1[p * nm * nm + j * nm + k] = acc;
      }
  for (int q = 0; q < nq; q++)
    for (int p = 0; p < nq; p++)
      for (int k = 0; k < nm; k++) {
        double acc = 0;
        for (int j = 0; j < nm; j++) acc += wsp1[p * nm * nm + j * nm + k] * basis[q * nm + j];
        wsp0[q * nq * nm + p * nm + k] = acc;
      }
  for (int r = 0; r < nq; r++)
    for (int q = 0; q < nq; q++)
      for (int p = 0; p < nq; p++) {
        double acc = 0;
        for (int k = 0; k < nm; k++) acc += wsp0[q * nq * nm + p * nm + k] * basis[r * nm + k];
        wsp1[p * nq * nq + q * nq + r] = acc;
      }
  for (int p = 0; p < nq; ++p)
    for (int q = 0; q < nq; ++q)
      for (int r = 0; r < nq; ++r) {
        const int pt = p * nq * nq + q * nq + r;
        const double Grr = Ge[GZZ * nq3 + pt], Grs = Ge[GYZ * nq3 + pt], Grt = Ge[GXZ * nq3 + pt];
        const double Gss = Ge[GYY * nq3 + pt], Gst = Ge[GXY * nq3 + pt], Gtt = Ge[GXX * nq3 + pt];
        double qr = 0.0, qs = 0.0, qt = 0.0;
        for (int n = 0; n < nq; ++n) qr += wsp1[n * nq * nq + q * nq + r] * dbasis[p * nq + n];
        for (int n = 0; n < nq; ++n) qs += wsp1[p * nq * nq + n * nq + r] * dbasis[q * nq + n];
        for (int n = 0; n < nq; ++n) qt += wsp1[p * nq * nq + q * nq + n] * dbasis[r * nq + n];
        rqr[pt] = Grr * qr + Grs * qs + Grt * qt;
        rqs[pt] = Grs * qr + Gss * qs + Gst * qt;
        rqt[pt] = Grt * qr + Gst * qs + Gtt * qt;
      }
  for (int p = 0; p < nq; ++p)
    for (int q = 0; q < nq; ++q)
      for (int r = 0; r < nq; ++r) {
        double tmp0 = 0;
        for (int n = 0; n < nq; ++n) tmp0 += rqr[n * nq * nq + q * nq + r] * dbasis[n * nq + p];
        for (int n = 0; n < nq; ++n) tmp0 += rqs[p * nq * nq + n * nq + r] * dbasis[n * nq + q];
        for (int n = 0; n < nq; ++n) tmp0 += rqt[p * nq * nq + q * nq + n] * dbasis[n * nq + r];
        wsp1[p * nq * nq + q * nq + r] = tmp0;
      }
  for (int k = 0; k < nm; k++)
    for (int q = 0; q < nq; q++)
      for (int p = 0; p < nq; p++) {
        double acc = 0;
        for (int r = 0; r < nq; r++) acc += wsp1[p * nq * nq + q * nq + r] * basis[r * nm + k];
        wsp0[q * nq * nm + p * nm + k] = acc;
      }
  for (int j = 0; j < nm; j++)
    for (int k = 0; k < nm; k++)
      for (int p = 0; p < nq; p++) {
        double acc = 0;
        for (int q = 0; q < nq; q++) acc += wsp0[q * nq * nm + p * nm + k] * basis[q * nm + j];
        wsp1[p * nm * nm + j * nm + k] = acc;
      }
  for (int i = 0; i < nm; i++)
    for (int j = 0; j < nm; j++)
      for (int k = 0; k < nm; k++) {
        double acc = 0;
        for (int p = 0; p < nq; p++) acc += wsp1[p * nm * nm + j * nm + k] * basis[p * nm + i];
        oute[i * nm * nm + j * nm + k] = acc;
      }
}

static void bk5_evec_one(int nq, const double *dbasis, const double *Ge,
                         const double *ine, double *oute, double *rqr,
                         double *rqs, double *rqt) {
  const int nq3 = nq * nq * nq;
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < nq; ++j)
      for (int k = 0; k < nq; ++k) {
        const int pt = i * nq * nq + j * nq + k;
        const double Grr = Ge[GZZ * nq3 + pt], Grs = Ge[GYZ * nq3 + pt], Grt = Ge[GXZ * nq3 + pt];
        const double Gss = Ge[GYY * nq3 + pt], Gst = Ge[GXY * nq3 + pt], Gtt = Ge[GXX * nq3 + pt];
        double qr = 0.0, qs = 0.0, qt = 0.0;
        for (int n = 0; n < nq; ++n) qr += ine[n * nq * nq + j * nq + k] * dbasis[i * nq + n];
        for (int n = 0; n < nq; ++n) qs += ine[i * nq * nq + n * nq + k] * dbasis[j * nq + n];
        for (int n = 0; n < nq; ++n) qt += ine[i * nq * nq + j * nq + n] * dbasis[k * nq + n];
        rqr[pt] = Grr * qr + Grs * qs + Grt * qt;
        rqs[pt] = Grs * qr + Gss * qs + Gst * qt;
        rqt[pt] = Grt * qr + Gst * qs + Gtt * qt;
      }
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < nq; ++j)
      for (int k = 0; k < nq; ++k) {
        double tmp0 = 0;
        for (int n = 0; n < nq; ++n) tmp0 += rqr[n * nq * nq + j * nq + k] * dbasis[n * nq + i];
        for (int n = 0; n < nq; ++n) tmp0 += rqs[i * nq * nq + n * nq + k] * dbasis[n * nq + j];
        for (int n = 0; n < nq; ++n) tmp0 += rqt[i * nq * nq + j * nq + n] * dbasis[n * nq + k];
        oute[i * nq * nq + j * nq + k] = tmp0;
      }
}

static void evec_range(int nm, int nq, const double *B, const double *D,
                       const double *G, const double *in, double *out,
                       uint64_t e0, uint64_t e1) {
  const size_t nq3 = (size_t)nq * nq * nq, nm3 = (size_t)nm * nm * nm;
  double *w = (double *)malloc(sizeof(double) * 5 * nq3);
  for (uint64_t e = e0; e < e1; ++e) {
    if (B)
      bk3_evec_one(nm, nq, B, D, G + e * 6 * nq3, in + e * nm3, out + e * nm3,
                   w, w + nq3, w + 2 * nq3, w + 3 * nq3, w + 4 * nq3);
    else
      bk5_evec_one(nq, D, G + e * 6 * nq3, in + e * nm3, out + e * nm3, w,
                   w + nq3, w + 2 * nq3);
  }
  free(w);
}

void orc_bk3_evec(int nm, int nq, uint64_t nelmt, const double *B,
                  const double *D, const double *G, const double *in,
                  double *out) {
  evec_range(nm, nq, B, D, G, in, out, 0, nelmt);
}

void orc_bk5_evec(int nq, uint64_t nelmt, const double *D, const double *G,
                  const double *in, double *out) {
  evec_range(nq, nq, NULL, D, G, in, out, 0, nelmt);
}

/* ------------------------------------------------------------------ */
/* Structured box mesh                                                 */
/* ------------------------------------------------------------------ */
/* Box [lo, hi] with n[d] cells per direction, the mesh of
 * GridGenerator::subdivided_hyper_rectangle + refine_global in
 * CEED_bp/src/bp3.cc:452-488 and bp5_kokkos/create_triangulation.h:17-42,
 * optionally deformed on the vertices (trilinear cells). */
uint64_t orc_box_n_cells(const orc_box *box) {
  return (uint64_t)box->n[0] * box->n[1] * box->n[2];
}

uint64_t orc_box_n_dofs(const orc_box *box, int p) {
  return (uint64_t)(box->n[0] * (uint64_t)p + 1) * (box->n[1] * (uint64_t)p + 1) *
         (box->n[2] * (uint64_t)p + 1);
}

void orc_box_vertices(const orc_box *box, double *X) {
  const uint64_t vx = box->n[0] + 1, vy = box->n[1] + 1, vz = box->n[2] + 1;
  const double two_pi = 6.283185307179586476925286766559;
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = (box->hi[d] - box->lo[d]) / box->n[d];
  for (uint64_t k = 0; k < vz; ++k)
    for (uint64_t j = 0; j < vy; ++j)
      for (uint64_t i = 0; i < vx; ++i) {
        const uint64_t v = i + vx * (j + vy * k);
        double x = box->lo[0] + (double)i * h[0];
        double y = box->lo[1] + (double)j * h[1];
        double z = box->lo[2] + (double)k * h[2];
        if (i == box->n[0]) x = box->hi[0];
        if (j == box->n[1]) y = box->hi[1];
        if (k == box->n[2]) z = box->hi[2];
        double s = 0.0;
        if (box->deform != 0.0)
          s = box->deform * sin(two_pi * x) * sin(two_pi * y) * sin(two_pi * z);
        X[3 * v + 0] = x + s;
        X[3 * v + 1] = y + s;
        X[3 * v + 2] = z + s;
      }
}

/* Element -> DoF index map in lexicographic (x fastest) local order, the
 * content of dirichlet_boundary_dofs_masks built by
 * setup_dirichlet_boundary_dofs_masks (CEED_bp/include/portable_laplace_operator.h:304-394):
 * map[e][l] = global DoF of the l-th lexicographic node, or 0xFFFFFFFF when
 * that DoF carries a homogeneous Dirichlet constraint (boundary id 0 = whole
 * boundary, CEED_bp/src/bp3.cc:147-151).  Global numbering: lexicographic
 * over the structured node grid (deal.II's own numbering cannot be
 * reproduced without deal.II). */
void orc_box_dof_map(const orc_box *box, int p, uint32_t *map,
                     uint8_t *constrained) {
  const uint64_t Nx = (uint64_t)box->n[0] * p + 1, Ny = (uint64_t)box->n[1] * p + 1,
                 Nz = (uint64_t)box->n[2] * p + 1;
  const int n = p + 1;
  if (constrained) {
    for (uint64_t gz = 0; gz < Nz; ++gz)
      for (uint64_t gy = 0; gy < Ny; ++gy)
        for (uint64_t gx = 0; gx < Nx; ++gx)
          constrained[gx + Nx * (gy + Ny * gz)] =
              (gx == 0 || gy == 0 || gz == 0 || gx == Nx - 1 || gy == Ny - 1 ||
               gz == Nz - 1);
  }
  for (uint64_t ez = 0; ez < box->n[2]; ++ez)
    for (uint64_t ey = 0; ey < box->n[1]; ++ey)
      for (uint64_t ex = 0; ex < box->n[0]; ++ex) {
        const uint64_t e = ex + box->n[0] * (ey + box->n[1] * ez);
        for (int c = 0; c < n; ++c)
          for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
              const uint64_t gx = ex * p + a, gy = ey * p + b, gz = ez * p + c;
              const int bnd = (gx == 0 || gy == 0 || gz == 0 || gx == Nx - 1 ||
                               gy == Ny - 1 || gz == Nz - 1);
              map[e * n * n * n + a + n * (b + n * c)] =
                  bnd ? ORC_INVALID : (uint32_t)(gx + Nx * (gy + Ny * gz));
            }
      }
}

void orc_box_cell_vertices(const orc_box *box, const double *X, double *cellX) {
  const uint64_t vx = box->n[0] + 1, vy = box->n[1] + 1;
  for (uint64_t ez = 0; ez < box->n[2]; ++ez)
    for (uint64_t ey = 0; ey < box->n[1]; ++ey)
      for (uint64_t ex = 0; ex < box->n[0]; ++ex) {
        const uint64_t e = ex + box->n[0] * (ey + box->n[1] * ez);
        for (int v = 0; v < 8; ++v) {
          const uint64_t gv = (ex + (v & 1)) + vx * ((ey + ((v >> 1) & 1)) + vy * (ez + (v >> 2)));
          for (int d = 0; d < 3; ++d) cellX[(e * 8 + v) * 3 + d] = X[gv * 3 + d];
        }
      }
}

/* Trilinear map x(xi) = sum_v X_v phi_v(xi), phi_v = prod_d (bit_d(v) ? xi_d : 1 - xi_d);
 * J[i][j] = dx_i/dxi_j. */
static void trilinear(const double *Xe, const double xi[3], double x[3],
                      double J[3][3]) {
  for (int i = 0; i < 3; ++i) {
    x[i] = 0;
    for (int j = 0; j < 3; ++j) J[i][j] = 0;
  }
  for (int v = 0; v < 8; ++v) {
    double f[3], df[3];
    for (int d = 0; d < 3; ++d) {
      const int bit = (v >> d) & 1;
      f[d] = bit ? xi[d] : 1.0 - xi[d];
      df[d] = bit ? 1.0 : -1.0;
    }
    const double phi = f[0] * f[1] * f[2];
    const double dphi[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
    for (int i = 0; i < 3; ++i) {
      x[i] += Xe[v * 3 + i] * phi;
      for (int j = 0; j < 3; ++j) J[i][j] += Xe[v * 3 + i] * dphi[j];
    }
  }
}

/* G = JxW * A A^T with A = J^{-1}, A(a,b) = d xi_a / d x_b: the metric the
 * deal.II cell operator applies (bakeoff_problems_dealii/include/portable_laplace_operator.h:229-255).
 * The reference's compute_G_tensors forms A^T A instead
 * (CEED_bp/include/portable_laplace_operator.h:280-289), which agrees only
 * on axis-aligned cells; this follows the mathematics.  Layout
 * G[e][c][q], q = qx + nq (qy + nq qz), c = (xx, xy, xz, yy, yz, zz) like
 * CEED_bp/include/portable_laplace_operator.h:294-295. */
void orc_geometry(uint64_t E, const double *cellX, int nq, const double *xq,
                  const double *wq, double *G) {
  const int nq3 = nq * nq * nq;
  for (uint64_t e = 0; e < E; ++e) {
    const double *Xe = cellX + e * 24;
    double *Ge = G + e * 6 * nq3;
    for (int qz = 0; qz < nq; ++qz)
      for (int qy = 0; qy < nq; ++qy)
        for (int qx = 0; qx < nq; ++qx) {
          const int q = qx + nq * (qy + nq * qz);
          const double xi[3] = {xq[qx], xq[qy], xq[qz]};
          double x[3], J[3][3];
          trilinear(Xe, xi, x, J);
          const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
          const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
          const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
          const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
          const double id = 1.0 / det;
          double A[3][3];
          A[0][0] = c00 * id;
          A[1][0] = c01 * id;
          A[2][0] = c02 * id;
          A[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
          A[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
          A[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
          A[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
          A[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
          A[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
          const double jxw = wq[qx] * wq[qy] * wq[qz] * det;
          const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
          for (int c = 0; c < 6; ++c) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += A[ca[c]][k] * A[cb[c]][k];
            Ge[c * nq3 + q] = jxw * s;
          }
        }
  }
}

void orc_dof_coordinates(const orc_box *box, int p, const double *cellX,
                         const uint32_t *map, double *coords) {
  const int n = p + 1;
  double s[64], w[64];
  orc_gll_rule(n, s, w);
  const uint64_t Nx = (uint64_t)box->n[0] * p + 1, Ny = (uint64_t)box->n[1] * p + 1;
  (void)map;
  for (uint64_t ez = 0; ez < box->n[2]; ++ez)
    for (uint64_t ey = 0; ey < box->n[1]; ++ey)
      for (uint64_t ex = 0; ex < box->n[0]; ++ex) {
        const uint64_t e = ex + box->n[0] * (ey + box->n[1] * ez);
        for (int c = 0; c < n; ++c)
          for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
              const double xi[3] = {s[a], s[b], s[c]};
              double x[3], J[3][3];
              trilinear(cellX + e * 24, xi, x, J);
              const uint64_t g = (ex * p + a) + Nx * ((ey * p + b) + Ny * (ez * p + c));
              for (int d = 0; d < 3; ++d) coords[g * 3 + d] = x[d];
            }
      }
}

/* ------------------------------------------------------------------ */
/* L-vector operator                                                    */
/* ------------------------------------------------------------------ */

/* Gather with masking, CEED_bp/include/bk3_kokkos_kernel.h:113-143. */
void orc_gather(int nm, const uint32_t *map, uint64_t E, const double *src,
                double *evec) {
  const uint64_t nm3 = (uint64_t)nm * nm * nm;
  for (uint64_t i = 0; i < E * nm3; ++i)
    evec[i] = (map[i] == ORC_INVALID) ? 0.0 : src[map[i]];
}

/* LaplaceOperator::vmult (CEED_bp/include/portable_laplace_operator.h:124-172):
 * dst = 0; per element gather -> sum factorization -> scatter-add of the
 * unconstrained entries (bk3_kokkos_kernel.h:356-382); then
 * copy_constrained_values(src, dst), i.e. identity rows on constrained
 * DoFs.  B == NULL selects GLL collocation (BK5, B = I).  The element loop
 * runs on `threads` OpenMP threads; the scatter is serial in element order,
 * so the result is deterministic. */
void orc_vmult(int nm, int nq, const double *B, const double *D,
               const double *G, const uint32_t *map, uint64_t E, uint64_t N,
               const uint8_t *constrained, const double *src, double *dst,
               int threads) {
  const uint64_t nm3 = (uint64_t)nm * nm * nm;
  double *ein = (double *)malloc(sizeof(double) * E * nm3);
  double *eout = (double *)malloc(sizeof(double) * E * nm3);
  orc_gather(nm, map, E, src, ein);
  if (threads < 1) threads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const uint64_t e0 = E * t / nt, e1 = E * (t + 1) / nt;
    evec_range(nm, nq, B, D, G, ein, eout, e0, e1);
  }
#else
  evec_range(nm, nq, B, D, G, ein, eout, 0, E);
#endif
  memset(dst, 0, sizeof(double) * N);
  for (uint64_t i = 0; i < E * nm3; ++i)
    if (map[i] != ORC_INVALID) dst[map[i]] += eout[i];
  for (uint64_t i = 0; i < N; ++i)
    if (constrained[i]) dst[i] = src[i];
  free(ein);
  free(eout);
}

/* Diagonal of the assembled operator, the quantity
 * MatrixFreeTools::compute_diagonal returns for bp5_kokkos/benchmark.cc:231-236
 * (deal.II, not installed: parity unpinned against it; pinned against the
 * diagonal of the brute-force assembled matrix).  Element diagonal
 * sum_q grad(phi_l)^T G_q grad(phi_l) with grad(phi_l) from
 * Dhat = D B (derivative of the nodal basis at the quadrature points);
 * constrained rows get 1 (identity rows of vmult). */
void orc_diagonal(int nm, int nq, const double *B, const double *D,
                  const double *G, const uint32_t *map, uint64_t E,
                  uint64_t N, const uint8_t *constrained, double *diag) {
  const int nq3 = nq * nq * nq, nm3 = nm * nm * nm;
  double *Bm = (double *)malloc(sizeof(double) * nq * nm);
  double *Dh = (double *)malloc(sizeof(double) * nq * nm);
  for (int q = 0; q < nq; ++q)
    for (int m = 0; m < nm; ++m) {
      Bm[q * nm + m] = B ? B[q * nm + m] : (q == m ? 1.0 : 0.0);
      double s = 0;
      for (int k = 0; k < nq; ++k) s += D[q * nq + k] * (B ? B[k * nm + m] : (k == m ? 1.0 : 0.0));
      Dh[q * nm + m] = s;
    }
  memset(diag, 0, sizeof(double) * N);
  for (uint64_t e = 0; e < E; ++e) {
    const double *Ge = G + e * 6 * (uint64_t)nq3;
    for (int c = 0; c < nm; ++c)
      for (int b = 0; b < nm; ++b)
        for (int a = 0; a < nm; ++a) {
          double s = 0;
          for (int qz = 0; qz < nq; ++qz)
            for (int qy = 0; qy < nq; ++qy)
              for (int qx = 0; qx < nq; ++qx) {
                const int q = qx + nq * (qy + nq * qz);
                const double gx = Dh[qx * nm + a] * Bm[qy * nm + b] * Bm[qz * nm + c];
                const double gy = Bm[qx * nm + a] * Dh[qy * nm + b] * Bm[qz * nm + c];
                const double gz = Bm[qx * nm + a] * Bm[qy * nm + b] * Dh[qz * nm + c];
                s += Ge[GXX * nq3 + q] * gx * gx + Ge[GYY * nq3 + q] * gy * gy +
                     Ge[GZZ * nq3 + q] * gz * gz +
                     2.0 * (Ge[GXY * nq3 + q] * gx * gy + Ge[GXZ * nq3 + q] * gx * gz +
                            Ge[GYZ * nq3 + q] * gy * gz);
              }
          const uint32_t g = map[e * nm3 + a + nm * (b + nm * c)];
          if (g != ORC_INVALID) diag[g] += s;
        }
  }
  for (uint64_t i = 0; i < N; ++i)
    if (constrained[i]) diag[i] = 1.0;
  free(Bm);
  free(Dh);
}

/* Inversion of bp5_kokkos/benchmark.cc:240-250: 1/d if d > 0, else 1. */
void orc_invert_diagonal(uint64_t N, double *diag) {
  for (uint64_t i = 0; i < N; ++i) diag[i] = diag[i] > 0 ? 1.0 / diag[i] : 1.0;
}

static double rhs_f(int f_kind, const double x[3]) {
  if (f_kind == 0) return 1.0;
  const double pi = 3.14159265358979323846;
  return 3.0 * pi * pi * sin(pi * x[0]) * sin(pi * x[1]) * sin(pi * x[2]);
}

/* compute_rhs (CEED_bp/src/bp3.cc:184-239): b_i = sum_cells sum_q
 * phi_i(x_q) f(x_q) JxW_q, constrained entries left zero
 * (AffineConstraints::distribute_local_to_global). */
void orc_rhs(int nm, int nq, const double *B, const double *xq,
             const double *wq, const double *cellX, const uint32_t *map,
             uint64_t E, uint64_t N, int f_kind, double *b) {
  const int nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  double *fj = (double *)malloc(sizeof(double) * nq3);
  memset(b, 0, sizeof(double) * N);
  for (uint64_t e = 0; e < E; ++e) {
    for (int qz = 0; qz < nq; ++qz)
      for (int qy = 0; qy < nq; ++qy)
        for (int qx = 0; qx < nq; ++qx) {
          const double xi[3] = {xq[qx], xq[qy], xq[qz]};
          double x[3], J[3][3];
          trilinear(cellX + e * 24, xi, x, J);
          const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                             J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          fj[qx + nq * (qy + nq * qz)] = rhs_f(f_kind, x) * wq[qx] * wq[qy] * wq[qz] * det;
        }
    for (int c = 0; c < nm; ++c)
      for (int bb = 0; bb < nm; ++bb)
        for (int a = 0; a < nm; ++a) {
          const uint32_t g = map[e * nm3 + a + nm * (bb + nm * c)];
          if (g == ORC_INVALID) continue;
          double s = 0;
          for (int qz = 0; qz < nq; ++qz)
            for (int qy = 0; qy < nq; ++qy)
              for (int qx = 0; qx < nq; ++qx)
                s += B[qx * nm + a] * B[qy * nm + bb] * B[qz * nm + c] *
                     fj[qx + nq * (qy + nq * qz)];
          b[g] += s;
        }
  }
  free(fj);
}

static double dot(uint64_t N, const double *a, const double *b) {
  double s = 0;
  for (uint64_t i = 0; i < N; ++i) s += a[i] * b[i];
  return s;
}

/* deal.II SolverCG::solve with ReductionControl(max_steps, abs_tol, rel_tol)
 * as called at CEED_bp/src/bp3.cc:264-288 (PreconditionIdentity) and
 * bp5_kokkos/benchmark.cc:355-368 (DiagonalMatrix = Jacobi):
 *   r = b - A x, tol = max(abs_tol, rel_tol ||r_0||); stop when
 *   ||r_k||_2 <= tol (success) or k >= max_steps / NaN (failure).
 *   Each step: v = P r; gamma = r.v; beta = gamma/gamma_old (k > 1);
 *   p = v + beta p; v = A p; alpha = gamma / p.v; x += alpha p;
 *   r -= alpha v.
 * steps = last_step(), the number of operator applications after the
 * initial residual. */
orc_cg_result orc_cg(int nm, int nq, const double *B, const double *D,
                     const double *G, const uint32_t *map, uint64_t E,
                     uint64_t N, const uint8_t *constrained,
                     const double *inv_diag, const double *b, double *x,
                     int max_steps, double abs_tol, double rel_tol,
                     double *res_history, int threads) {
  orc_cg_result res = {0, 0, 0.0, 0.0};
  double *r = (double *)malloc(sizeof(double) * N);
  double *p = (double *)malloc(sizeof(double) * N);
  double *v = (double *)malloc(sizeof(double) * N);
  orc_vmult(nm, nq, B, D, G, map, E, N, constrained, x, r, threads);
  for (uint64_t i = 0; i < N; ++i) r[i] = b[i] - r[i];
  double rn = sqrt(dot(N, r, r));
  res.initial_value = rn;
  res.last_value = rn;
  if (res_history) res_history[0] = rn;
  const double tol = (abs_tol > rel_tol * rn) ? abs_tol : rel_tol * rn;
  if (rn <= tol) {
    res.converged = 1;
    goto done;
  }
  if (max_steps <= 0 || isnan(rn)) { /* SolverControl::check fails at step 0 */
    res.converged = 0;
    goto done;
  }
  double gamma_old = 0.0;
  for (int it = 1;; ++it) {
    for (uint64_t i = 0; i < N; ++i) v[i] = inv_diag ? inv_diag[i] * r[i] : r[i];
    const double gamma = dot(N, r, v);
    if (it == 1)
      for (uint64_t i = 0; i < N; ++i) p[i] = v[i];
    else {
      const double beta = gamma / gamma_old;
      for (uint64_t i = 0; i < N; ++i) p[i] = v[i] + beta * p[i];
    }
    orc_vmult(nm, nq, B, D, G, map, E, N, constrained, p, v, threads);
    const double alpha = gamma / dot(N, p, v);
    for (uint64_t i = 0; i < N; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * v[i];
    }
    rn = sqrt(dot(N, r, r));
    res.steps = it;
    res.last_value = rn;
    if (res_history) res_history[it] = rn;
    if (rn <= tol) {
      res.converged = 1;
      break;
    }
    if (it >= max_steps || isnan(rn)) {
      res.converged = 0;
      break;
    }
    gamma_old = gamma;
  }
done:
  free(r);
  free(p);
  free(v);
  return res;
}

double orc_l2_error_sin(int nm, int nq, const double *B, const double *xq,
                        const double *wq, const double *cellX,
                        const uint32_t *map, uint64_t E, const double *u) {
  const int nm3 = nm * nm * nm;
  const double pi = 3.14159265358979323846;
  double err = 0;
  for (uint64_t e = 0; e < E; ++e)
    for (int qz = 0; qz < nq; ++qz)
      for (int qy = 0; qy < nq; ++qy)
        for (int qx = 0; qx < nq; ++qx) {
          const double xi[3] = {xq[qx], xq[qy], xq[qz]};
          double x[3], J[3][3];
          trilinear(cellX + e * 24, xi, x, J);
          const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                             J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          double uh = 0;
          for (int c = 0; c < nm; ++c)
            for (int b = 0; b < nm; ++b)
              for (int a = 0; a < nm; ++a) {
                const uint32_t g = map[e * nm3 + a + nm * (b + nm * c)];
                const double val = (g == ORC_INVALID) ? 0.0 : u[g];
                uh += B[qx * nm + a] * B[qy * nm + b] * B[qz * nm + c] * val;
              }
          const double ex = sin(pi * x[0]) * sin(pi * x[1]) * sin(pi * x[2]);
          err += (uh - ex) * (uh - ex) * wq[qx] * wq[qy] * wq[qz] * det;
        }
  return sqrt(err);
}
