// setup.hip -- GPU-side set-up: structured mesh, DoF map, geometric factors,
// right-hand side, operator diagonal.  The reference builds all of these on
// the host (108.6 s + 16 s at 67.6M DoFs, CEED_bp/results/1xGH200_P4.txt:608-609);
// here every step is an embarrassingly parallel kernel.
#include "mfx_internal.hpp"

namespace mfx {
namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kPi = 3.14159265358979323846264338327950288;

__device__ void box_vertex(const BoxGeom &b, uint32_t i, uint32_t j, uint32_t k,
                           double *X) {
  double c[3];
  const uint32_t ijk[3] = {i, j, k};
  for (int d = 0; d < 3; ++d) {
    const double h = (b.hi[d] - b.lo[d]) / b.n[d];
    c[d] = (ijk[d] == b.n[d]) ? b.hi[d] : b.lo[d] + (double)ijk[d] * h;
  }
  double s = 0.0;
  if (b.deform != 0.0)
    s = b.deform * sin(kTwoPi * c[0]) * sin(kTwoPi * c[1]) * sin(kTwoPi * c[2]);
  X[0] = c[0] + s;
  X[1] = c[1] + s;
  X[2] = c[2] + s;
}

// cellX[e][v][d], v = a + 2b + 4c lexicographic vertex order.
__global__ void k_box_cell_vertices(BoxGeom b, double *cellX) {
  const uint64_t E = (uint64_t)b.n[0] * b.n[1] * b.n[2];
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * 8) return;
  const uint64_t e = t >> 3;
  const int v = (int)(t & 7);
  const uint32_t ex = (uint32_t)(e % b.n[0]);
  const uint32_t ey = (uint32_t)((e / b.n[0]) % b.n[1]);
  const uint32_t ez = (uint32_t)(e / ((uint64_t)b.n[0] * b.n[1]));
  double X[3];
  box_vertex(b, ex + (v & 1), ey + ((v >> 1) & 1), ez + (v >> 2), X);
  for (int d = 0; d < 3; ++d) cellX[t * 3 + d] = X[d];
}

// map[e][a + n(b + n c)] = lexicographic global node, or invalid on the
// boundary (homogeneous Dirichlet on the whole boundary).
__global__ void k_box_map(BoxGeom b, int p, uint32_t *map) {
  const int n = p + 1, n3 = n * n * n;
  const uint64_t E = (uint64_t)b.n[0] * b.n[1] * b.n[2];
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * n3) return;
  const uint64_t e = t / n3;
  const int l = (int)(t - e * n3);
  const int a = l % n, bb = (l / n) % n, c = l / (n * n);
  const uint64_t ex = e % b.n[0], ey = (e / b.n[0]) % b.n[1], ez = e / ((uint64_t)b.n[0] * b.n[1]);
  const uint64_t Nx = (uint64_t)b.n[0] * p + 1, Ny = (uint64_t)b.n[1] * p + 1,
                 Nz = (uint64_t)b.n[2] * p + 1;
  const uint64_t gx = ex * p + a, gy = ey * p + bb, gz = ez * p + c;
  const bool bnd = gx == 0 || gy == 0 || gz == 0 || gx == Nx - 1 || gy == Ny - 1 || gz == Nz - 1;
  map[t] = bnd ? kInvalid : (uint32_t)(gx + Nx * (gy + Ny * gz));
}

// Trilinear map of one cell: position and Jacobian J[i][j] = dx_i/dxi_j.
__device__ void trilinear(const double *Xe, double x0, double x1, double x2,
                          double x[3], double J[3][3]) {
  const double xi[3] = {x0, x1, x2};
  for (int i = 0; i < 3; ++i) {
    x[i] = 0.0;
    for (int j = 0; j < 3; ++j) J[i][j] = 0.0;
  }
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    double f[3], df[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const bool bit = (v >> d) & 1;
      f[d] = bit ? xi[d] : 1.0 - xi[d];
      df[d] = bit ? 1.0 : -1.0;
    }
    const double phi = f[0] * f[1] * f[2];
    const double dphi[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double Xi = Xe[v * 3 + i];
      x[i] += Xi * phi;
#pragma unroll
      for (int j = 0; j < 3; ++j) J[i][j] += Xi * dphi[j];
    }
  }
}

__device__ double det3(const double J[3][3]) {
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
         J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// G[e][c][q] = JxW (A A^T)_c, A = J^{-1}: the metric of the deal.II cell
// operator (bakeoff_problems_dealii/include/portable_laplace_operator.h:229-255),
// replacing compute_G_tensors (CEED_bp/include/portable_laplace_operator.h:239-302),
// whose A^T A product is correct only on axis-aligned cells.
__global__ void k_geometry(uint32_t E, const double *__restrict__ cellX, int nq,
                           const double *__restrict__ xq, const double *__restrict__ wq,
                           double *__restrict__ G) {
  const int nq3 = nq * nq * nq;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)E * nq3) return;
  const uint64_t e = t / nq3;
  const int q = (int)(t - e * nq3);
  const int qx = q % nq, qy = (q / nq) % nq, qz = q / (nq * nq);
  double Xe[24];
  for (int i = 0; i < 24; ++i) Xe[i] = cellX[e * 24 + i];
  double x[3], J[3][3];
  trilinear(Xe, xq[qx], xq[qy], xq[qz], x, J);
  const double det = det3(J);
  // adjugate / det = inverse, A[a][b] = d xi_a / d x_b
  double A[3][3];
  const double id = 1.0 / det;
  A[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  A[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  A[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  A[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  A[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  A[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  A[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  A[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  A[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  const double jxw = wq[qx] * wq[qy] * wq[qz] * det;
  double *Ge = G + e * 6 * (uint64_t)nq3 + q;
  const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double s = A[ca[c]][0] * A[cb[c]][0] + A[ca[c]][1] * A[cb[c]][1] +
                     A[ca[c]][2] * A[cb[c]][2];
    Ge[c * nq3] = jxw * s;
  }
}

// One workgroup per cell: f JxW at the quadrature points into LDS, then
// b_l = sum_q B(qx,a) B(qy,b) B(qz,c) fJ(q), scatter-added.
__global__ void k_rhs(int p, int nq, const double *__restrict__ B,
                      const double *__restrict__ xq, const double *__restrict__ wq,
                      const double *__restrict__ cellX, const uint32_t *__restrict__ map,
                      int f_kind, double *__restrict__ b) {
  extern __shared__ double sh[];
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  double *fj = sh;
  double *sB = sh + nq3;
  const uint64_t e = blockIdx.x;
  for (int i = threadIdx.x; i < nq * nm; i += blockDim.x) sB[i] = B[i];
  double Xe[24];
  for (int i = 0; i < 24; ++i) Xe[i] = cellX[e * 24 + i];
  for (int q = threadIdx.x; q < nq3; q += blockDim.x) {
    const int qx = q % nq, qy = (q / nq) % nq, qz = q / (nq * nq);
    double x[3], J[3][3];
    trilinear(Xe, xq[qx], xq[qy], xq[qz], x, J);
    const double f = (f_kind == 0) ? 1.0
                                   : 3.0 * kPi * kPi * sin(kPi * x[0]) * sin(kPi * x[1]) *
                                         sin(kPi * x[2]);
    fj[q] = f * wq[qx] * wq[qy] * wq[qz] * det3(J);
  }
  __syncthreads();
  for (int l = threadIdx.x; l < nm3; l += blockDim.x) {
    const uint32_t g = map[e * nm3 + l];
    if (g == kInvalid) continue;
    const int a = l % nm, bb = (l / nm) % nm, c = l / (nm * nm);
    double s = 0.0;
    for (int qz = 0; qz < nq; ++qz)
      for (int qy = 0; qy < nq; ++qy) {
        const double byz = sB[qy * nm + bb] * sB[qz * nm + c];
        for (int qx = 0; qx < nq; ++qx) s += sB[qx * nm + a] * byz * fj[qx + nq * (qy + nq * qz)];
      }
    unsafeAtomicAdd(b + g, s);
  }
}

// One workgroup per cell: element diagonal sum_q grad(phi_l)^T G grad(phi_l)
// with grad(phi_l) from Dhat = D B, scatter-added.
__global__ void k_diagonal(int p, int nq, const double *__restrict__ B,
                           const double *__restrict__ D, const double *__restrict__ G,
                           const uint32_t *__restrict__ map, double *__restrict__ diag) {
  extern __shared__ double sh[];
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  double *sG = sh;               // 6 nq3
  double *sB = sG + 6 * nq3;     // nq nm
  double *sDh = sB + nq * nm;    // nq nm
  const uint64_t e = blockIdx.x;
  for (int i = threadIdx.x; i < 6 * nq3; i += blockDim.x) sG[i] = G[e * 6 * (uint64_t)nq3 + i];
  for (int i = threadIdx.x; i < nq * nm; i += blockDim.x) {
    const int q = i / nm, m = i % nm;
    sB[i] = B[i];
    double s = 0.0;
    for (int k = 0; k < nq; ++k) s += D[q * nq + k] * B[k * nm + m];
    sDh[i] = s;
  }
  __syncthreads();
  for (int l = threadIdx.x; l < nm3; l += blockDim.x) {
    const uint32_t g = map[e * nm3 + l];
    if (g == kInvalid) continue;
    const int a = l % nm, bb = (l / nm) % nm, c = l / (nm * nm);
    double s = 0.0;
    for (int qz = 0; qz < nq; ++qz)
      for (int qy = 0; qy < nq; ++qy)
        for (int qx = 0; qx < nq; ++qx) {
          const int q = qx + nq * (qy + nq * qz);
          const double gx = sDh[qx * nm + a] * sB[qy * nm + bb] * sB[qz * nm + c];
          const double gy = sB[qx * nm + a] * sDh[qy * nm + bb] * sB[qz * nm + c];
          const double gz = sB[qx * nm + a] * sB[qy * nm + bb] * sDh[qz * nm + c];
          s += sG[q] * gx * gx + sG[3 * nq3 + q] * gy * gy + sG[5 * nq3 + q] * gz * gz +
               2.0 * (sG[nq3 + q] * gx * gy + sG[2 * nq3 + q] * gx * gz + sG[4 * nq3 + q] * gy * gz);
        }
    unsafeAtomicAdd(diag + g, s);
  }
}

__global__ void k_dof_coordinates(uint32_t E, int p, const double *__restrict__ support,
                                  const double *__restrict__ cellX,
                                  const uint32_t *__restrict__ map, double *coords,
                                  const uint32_t *__restrict__ unused) {
  (void)unused;
  const int n = p + 1, n3 = n * n * n;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)E * n3) return;
  const uint64_t e = t / n3;
  const int l = (int)(t - e * n3);
  const uint32_t g = map[t];
  if (g == kInvalid) return;
  double Xe[24];
  for (int i = 0; i < 24; ++i) Xe[i] = cellX[e * 24 + i];
  double x[3], J[3][3];
  trilinear(Xe, support[l % n], support[(l / n) % n], support[l / (n * n)], x, J);
  for (int d = 0; d < 3; ++d) coords[(uint64_t)g * 3 + d] = x[d];
}

__global__ void k_to_float(const double *__restrict__ in, float *__restrict__ out, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

inline dim3 grid_for(uint64_t n, int block) {
  return dim3((unsigned)((n + block - 1) / block));
}

}  // namespace

void launch_box_cell_vertices(const BoxGeom &box, double *cellX, hipStream_t s) {
  const uint64_t E = (uint64_t)box.n[0] * box.n[1] * box.n[2];
  hipLaunchKernelGGL(k_box_cell_vertices, grid_for(E * 8, 256), dim3(256), 0, s, box, cellX);
  MFX_HIP(hipGetLastError());
}

void launch_box_map(const BoxGeom &box, int p, uint32_t *map, hipStream_t s) {
  const uint64_t E = (uint64_t)box.n[0] * box.n[1] * box.n[2];
  const uint64_t n3 = (uint64_t)(p + 1) * (p + 1) * (p + 1);
  hipLaunchKernelGGL(k_box_map, grid_for(E * n3, 256), dim3(256), 0, s, box, p, map);
  MFX_HIP(hipGetLastError());
}

void launch_geometry(uint32_t E, const double *cellX, int nq, const double *xq,
                     const double *wq, double *G, hipStream_t s) {
  const uint64_t n = (uint64_t)E * nq * nq * nq;
  hipLaunchKernelGGL(k_geometry, grid_for(n, 256), dim3(256), 0, s, E, cellX, nq, xq, wq, G);
  MFX_HIP(hipGetLastError());
}

void launch_rhs(uint32_t E, int p, int nq, const double *B, const double *xq,
                const double *wq, const double *cellX, const uint32_t *map,
                int f_kind, double *b, hipStream_t s) {
  const size_t lds = sizeof(double) * ((size_t)nq * nq * nq + (size_t)nq * (p + 1));
  hipLaunchKernelGGL(k_rhs, dim3(E), dim3(256), lds, s, p, nq, B, xq, wq, cellX, map, f_kind, b);
  MFX_HIP(hipGetLastError());
}

void launch_diagonal(uint32_t E, int p, int nq, const double *B, const double *D,
                     const double *G, const uint32_t *map, double *diag, hipStream_t s) {
  const size_t lds = sizeof(double) * (6 * (size_t)nq * nq * nq + 2 * (size_t)nq * (p + 1));
  hipLaunchKernelGGL(k_diagonal, dim3(E), dim3(256), lds, s, p, nq, B, D, G, map, diag);
  MFX_HIP(hipGetLastError());
}

void launch_dof_coordinates(uint32_t E, int p, const double *support, const double *cellX,
                            const uint32_t *map, double *coords, hipStream_t s) {
  const uint64_t n = (uint64_t)E * (p + 1) * (p + 1) * (p + 1);
  hipLaunchKernelGGL(k_dof_coordinates, grid_for(n, 256), dim3(256), 0, s, E, p, support,
                     cellX, map, coords, nullptr);
  MFX_HIP(hipGetLastError());
}

void launch_to_float(const double *in, float *out, uint64_t n, hipStream_t s) {
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_to_float, dim3((unsigned)std::max<uint64_t>(blocks, 1)), dim3(256), 0, s,
                     in, out, n);
  MFX_HIP(hipGetLastError());
}

}  // namespace mfx
