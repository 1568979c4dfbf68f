// apply.hip -- sum-factorized scalar-Laplacian element kernels for gfx950.
//
// One kernel family covers the reference's BK3 (Gauss, interpolation B) and
// BK5 (GLL collocation, B = I) element operators
//   v_e = B^T D^T G_e D B u_e
// (CEED_bp/include/bk3_kokkos_kernel.h:113-382,
//  CEED_BK/include/kernels/BK5/templated_cuda_kernels.cuh:37-125), with the
// DoF-map gather (masked entries read as 0) and the scatter-add (masked
// entries skipped) of CEED_bp/include/bk3_kokkos_kernel.h:113-143, 356-382.
//
// Work decomposition (CDNA4, 64-wide wavefronts):
//   * a workgroup holds EPB elements, one thread per (x,y) quadrature column
//     (EPB * nq^2 <= 256 threads); each thread keeps its z-column in
//     registers, so z contractions are register-local and only the x and y
//     contractions go through LDS (3 LDS cubes per element);
//   * the map and G are read with the column index fastest, so every load
//     instruction of a wavefront touches contiguous 8-byte runs (G layout
//     [cell][6][q^3], x fastest);
//   * collocation kernel: the 6 q-columns of G a thread needs are loaded into
//     registers before the first barrier, so the geometry stream is in flight
//     while the dependent map -> src gather completes;
//   * scatter: hardware FP64/FP32 atomic add in L2 (global_atomic_add_f64).
// Cells per workgroup and the prefetch switch were chosen per nq by timing
// the p = 2/4/6/8 BP5 operators at ~32M DoFs (a remap of workgroups onto
// XCD-contiguous cell ranges was measured too and did not pay).
#include <type_traits>

#include "mfx_internal.hpp"
#include "tables.hpp"

namespace mfx {
namespace {

// Cells per workgroup (EPB) and G prefetch for the collocation kernel;
// the interpolation kernel keeps EPB = floor(256 / nq^2).
template <int NQ>
struct Tile {
  static constexpr int EPB = (256 / (NQ * NQ)) > 0 ? 256 / (NQ * NQ) : 1;
  static constexpr int THREADS = EPB * NQ * NQ;
};
template <int NQ>
struct CollocTile {
  static constexpr int EPB = NQ <= 6 ? 128 / (NQ * NQ) : (NQ == 7 ? 1 : (NQ == 8 ? 2 : 3));
  static constexpr int THREADS = EPB * NQ * NQ;
  static constexpr bool PREFETCH = NQ <= 7;
};

__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

// Sum over the workgroup; result valid in thread 0.
template <int THREADS>
__device__ __forceinline__ double block_sum(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  constexpr int NW = (THREADS + 63) / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NW; ++i) s += red[i];
  }
  return s;
}

// ---------------------------------------------------------------------------
// GLL collocation (BK5): NM == NQ, B = I.
// ---------------------------------------------------------------------------
// EPB cells per workgroup; PREF: G columns loaded into registers before the
// first barrier.  cell_inst / inst_done (optional): cells whose instance is
// flagged done are skipped (batched solves).
template <typename T, int NQ, int EPB, bool PREF, bool GATHER, bool DOT>
__global__ void __launch_bounds__(EPB * NQ * NQ)
k_apply_colloc(const T *__restrict__ gD, const T *__restrict__ G,
               const T *__restrict__ src, T *__restrict__ dst,
               const uint32_t *__restrict__ map, uint32_t n_cells,
               double *__restrict__ partials, const int *__restrict__ skip,
               const uint16_t *__restrict__ cell_inst, const int *__restrict__ inst_done) {
  constexpr int THREADS = EPB * NQ * NQ;
  constexpr int NQ2 = NQ * NQ, NQ3 = NQ2 * NQ;
  __shared__ T sD[NQ * NQ];
  __shared__ T su[EPB][NQ3];
  __shared__ T sfx[EPB][NQ3];
  __shared__ T sfy[EPB][NQ3];
  __shared__ double red[(THREADS + 63) / 64];

  if (skip && *skip) return;  // converged CG: nothing left to do
  const int tid = threadIdx.x;
  const int el = tid / NQ2;
  const int col = tid - el * NQ2;
  const int qx = col % NQ, qy = col / NQ;
  const uint32_t e = blockIdx.x * EPB + el;
  bool active = e < n_cells;
  if (active && inst_done) active = inst_done[cell_inst[e]] == 0;

  for (int i = tid; i < NQ * NQ; i += THREADS) sD[i] = gD[i];

  // geometry columns first: independent of the gather, so in flight early
  const T *Ge = G + (size_t)(active ? e : 0) * 6 * NQ3 + col;
  T g[PREF ? 6 : 1][PREF ? NQ : 1];
  if (PREF && active) {
#pragma unroll
    for (int z = 0; z < NQ; ++z)
#pragma unroll
      for (int c = 0; c < 6; ++c) g[c][z] = __builtin_nontemporal_load(Ge + c * NQ3 + z * NQ2);
  }

  uint32_t idx[NQ];
  T u[NQ];
  if (active) {
    if (GATHER) {
      const uint32_t *me = map + (size_t)e * NQ3 + col;
#pragma unroll
      for (int z = 0; z < NQ; ++z) idx[z] = __builtin_nontemporal_load(me + z * NQ2);
#pragma unroll
      for (int z = 0; z < NQ; ++z) u[z] = (idx[z] != kInvalid) ? src[idx[z]] : T(0);
    } else {
      const T *se = src + (size_t)e * NQ3 + col;
#pragma unroll
      for (int z = 0; z < NQ; ++z) u[z] = se[z * NQ2];
    }
#pragma unroll
    for (int z = 0; z < NQ; ++z) su[el][z * NQ2 + col] = u[z];
  }
  __syncthreads();

  T fz[NQ];
  if (active) {
#pragma unroll
    for (int z = 0; z < NQ; ++z) {
      T gxx, gxy, gxz, gyy, gyz, gzz;
      if constexpr (PREF) {
        gxx = g[0][z]; gxy = g[1][z]; gxz = g[2][z];
        gyy = g[3][z]; gyz = g[4][z]; gzz = g[5][z];
      } else {
        gxx = __builtin_nontemporal_load(Ge + 0 * NQ3 + z * NQ2);
        gxy = __builtin_nontemporal_load(Ge + 1 * NQ3 + z * NQ2);
        gxz = __builtin_nontemporal_load(Ge + 2 * NQ3 + z * NQ2);
        gyy = __builtin_nontemporal_load(Ge + 3 * NQ3 + z * NQ2);
        gyz = __builtin_nontemporal_load(Ge + 4 * NQ3 + z * NQ2);
        gzz = __builtin_nontemporal_load(Ge + 5 * NQ3 + z * NQ2);
      }
      T dx = 0, dy = 0, dz = 0;
#pragma unroll
      for (int a = 0; a < NQ; ++a) {
        dx += sD[qx * NQ + a] * su[el][z * NQ2 + qy * NQ + a];
        dy += sD[qy * NQ + a] * su[el][z * NQ2 + a * NQ + qx];
        dz += sD[z * NQ + a] * u[a];
      }
      sfx[el][z * NQ2 + col] = gxx * dx + gxy * dy + gxz * dz;
      sfy[el][z * NQ2 + col] = gxy * dx + gyy * dy + gyz * dz;
      fz[z] = gxz * dx + gyz * dy + gzz * dz;
    }
  }
  __syncthreads();

  double dot = 0.0;
  if (active) {
#pragma unroll
    for (int z = 0; z < NQ; ++z) {
      T v = 0;
#pragma unroll
      for (int a = 0; a < NQ; ++a) {
        v += sD[a * NQ + qx] * sfx[el][z * NQ2 + qy * NQ + a];
        v += sD[a * NQ + qy] * sfy[el][z * NQ2 + a * NQ + qx];
        v += sD[a * NQ + z] * fz[a];
      }
      if (DOT) dot += (double)u[z] * (double)v;
      if (GATHER) {
        if (idx[z] != kInvalid) atomic_add(dst + idx[z], v);
      } else {
        dst[(size_t)e * NQ3 + z * NQ2 + col] = v;
      }
    }
  }
  if (DOT) {
    const double s = block_sum<THREADS>(dot, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// General tensor-product element (BK3): NM = p+1 nodes, NQ >= NM quadrature
// points, interpolation B[q][m] and collocation derivative D[q][k].
// ---------------------------------------------------------------------------
//
// OTF: the metric terms are computed per quadrature point from the 8
// vertices of the trilinear cell (cellX [E][8][3], Gauss points gxq and
// weights gwq on [0,1]) instead of being read from G: per cell seven
// coefficient vectors of the Jacobian (LDS), per (x,y) column its Jacobian
// columns as functions of zeta, per point the cofactor columns
// C0 = J1 x J2, C1 = J2 x J0, C2 = J0 x J1 (rows of det J J^-1), det J =
// J2 . C2 and (G grad u)_a = (w / det J) C_a . sum_b (du/dxi_b) C_b, with
// G = w det J J^-1 J^-T the stored form.  24 doubles per cell replace
// 48 nq^3 bytes of G.
//
// ROW (box meshes, cells numbered x fastest, `row_cells` cells per x-row):
// each workgroup takes up to EPB consecutive cells of one x-row.  The scatter
// then first sums the cells' shared x-faces in LDS (the right face of cell
// c and the left face of cell c+1, in that order) and adds each node of the
// resulting row segment once, with consecutive lanes on consecutive nodes
// of an x-line: ~(p+1)^2 (EPB p + 1) atomic adds in few wide requests
// instead of EPB (p+1)^3 in requests of about p+1 lanes.  The device-scope
// FP atomics execute beyond the per-XCD L2s, and their request count bounds
// the per-cell scatter (measured: plain stores in their place made the FP32
// q = p+1 apply at 135M DoFs 2.1x faster).
template <typename T, int NM, int NQ, bool GATHER, bool DOT, bool OTF = false, bool ROW = false,
          bool AX = false>
__global__ void __launch_bounds__(Tile<NQ>::THREADS)
k_apply_interp(const T *__restrict__ gB, const T *__restrict__ gD,
               const T *__restrict__ G, const T *__restrict__ src,
               T *__restrict__ dst, const uint32_t *__restrict__ map,
               uint32_t n_cells, double *__restrict__ partials,
               const int *__restrict__ skip, const double *__restrict__ cellX,
               const T *__restrict__ gxq, const T *__restrict__ gwq, uint32_t row_cells,
               const T *__restrict__ geo) {
  constexpr int EPB = Tile<NQ>::EPB, NQ2 = NQ * NQ, NQ3 = NQ2 * NQ;
  static_assert(!AX || (OTF && ROW), "axis-aligned cells: a box-mesh form of the vertex metric");
  constexpr int NM2 = NM * NM, NM3 = NM2 * NM;
  constexpr int THREADS = Tile<NQ>::THREADS;
  __shared__ T sB[NQ * NM];
  __shared__ T sD[NQ * NQ];
  __shared__ T s0[EPB][NQ3];
  __shared__ T s1[EPB][NQ3];
  __shared__ T s2[EPB][NQ3];
  __shared__ T sXv[OTF ? EPB * 24 : 1];  // cell vertices, then
  __shared__ T sKc[OTF ? EPB * 21 : 1];  // per cell 7 Jacobian coefficient vectors
  __shared__ double red[(Tile<NQ>::THREADS + 63) / 64];

  if (skip && *skip) return;  // converged CG: nothing left to do
  const int tid = threadIdx.x;
  const int el = tid / NQ2;
  const int col = tid - el * NQ2;
  const int qx = col % NQ, qy = col / NQ;
  // the workgroup's cells: e0 .. e0 + m - 1
  uint32_t e0, m;
  if constexpr (ROW) {
    const uint32_t wpr = (row_cells + EPB - 1) / EPB;  // workgroups per x-row
    const uint32_t row = blockIdx.x / wpr, xs = (blockIdx.x - row * wpr) * EPB;
    e0 = row * row_cells + xs;
    m = min((uint32_t)EPB, row_cells - xs);
  } else {
    e0 = blockIdx.x * EPB;
    m = min((uint32_t)EPB, n_cells - e0);
  }
  const uint32_t e = e0 + el;
  const bool active = (uint32_t)el < m;
  const bool node_col = active && qx < NM && qy < NM;

  for (int i = tid; i < NQ * NM; i += Tile<NQ>::THREADS) sB[i] = gB[i];
  for (int i = tid; i < NQ * NQ; i += Tile<NQ>::THREADS) sD[i] = gD[i];
  // wave-uniform entries (the z contractions): through the scalar cache up
  // to nq = 7, from LDS above (their scalar registers would spill)
  auto uB = [&](int i, int j) -> T {
    if constexpr (NQ <= 7) return gB[i * NM + j]; else return sB[i * NM + j];
  };
  auto uD = [&](int i, int j) -> T {
    if constexpr (NQ <= 7) return gD[i * NQ + j]; else return sD[i * NQ + j];
  };

  // nq >= 9 (two or three cells per workgroup): the metric terms are loaded into
  // registers after the x interpolation, so their stream overlaps the y/z
  // interpolation (3-20% faster at p = 7, 8 with q = 9, 10; slower at nq = 7
  // in FP64, no gain below).  Smaller nq loads them in the gradient phase.
  constexpr bool PF = NQ >= 9 && !OTF;
  // LINE: every x and y contraction (interpolation, derivative and their
  // transposes) is taken along whole lines, one line per thread, instead of
  // reading a line per output point, with the metric from the vertices; the
  // line contractions read the even-odd tables of B and D (below) through
  // the scalar cache.  Measured against the per-point form: 25-45% faster in
  // FP32 and 10-40% in FP64 at nq = 5..10 (the gain grows with nq), 1-5% at
  // nq = 3 and in FP32 at nq = 4; FP64 at nq = 4 was 3% slower.  With stored
  // G (memory-bound) level, so not used there.
  constexpr bool LINE =
      OTF && (NQ >= 5 || NQ == 3 || (NQ == 4 && std::is_same<T, float>::value));
  // the line-form contractions apply B, B^T, D, D^T through their even-odd
  // tables (half the multiply-adds; geo holds the four in that order)
  constexpr int EO_BT = eo_table_size(NQ, NM), EO_D = EO_BT + eo_table_size(NM, NQ),
                EO_DT = EO_D + eo_table_size(NQ, NQ);
  const T *Ge = OTF ? nullptr : G + (size_t)(active ? e : 0) * 6 * NQ3 + col;
  T g[PF ? 6 : 1][PF ? NQ : 1];
  if constexpr (OTF) {
    for (int i = tid; i < EPB * 24; i += THREADS) {
      const uint32_t c = (uint32_t)(i / 24);
      sXv[i] = c < m ? (T)cellX[(size_t)(e0 + c) * 24 + i % 24] : T(0);
    }
  }

  // 1. gather the (x,y) node column
  uint32_t idx[NM];
  T uin[NM];
  if (node_col) {
    const int l0 = qy * NM + qx;
    if (GATHER) {
      const uint32_t *me = map + (size_t)e * NM3 + l0;
#pragma unroll
      for (int z = 0; z < NM; ++z) idx[z] = __builtin_nontemporal_load(me + z * NM2);
#pragma unroll
      for (int z = 0; z < NM; ++z) uin[z] = (idx[z] != kInvalid) ? src[idx[z]] : T(0);
    } else {
      const T *se = src + (size_t)e * NM3 + l0;
#pragma unroll
      for (int z = 0; z < NM; ++z) uin[z] = se[z * NM2];
    }
#pragma unroll
    for (int z = 0; z < NM; ++z) s0[el][z * NM2 + l0] = uin[z];
  }
  __syncthreads();

  // per cell the Jacobian coefficient vectors (vertex v = a + 2b + 4c from
  // the bits of (xi, eta, zeta)): A0 = x1-x0, A1 = x3-x2-x1+x0, G0 = x2-x0,
  // K0 = x4-x0, K1 = x5-x1-x4+x0, K2 = x6-x2-x4+x0, K3 = x7-x6-x5+x4-x3+x2+x1-x0
  if constexpr (OTF) {
    for (int j = tid; j < EPB * 3; j += THREADS) {  // j = 3 cell + coordinate
      const T *X = sXv + (j / 3) * 24 + j % 3;
      const T x0 = X[0], x1 = X[3], x2 = X[6], x3 = X[9];
      const T x4 = X[12], x5 = X[15], x6 = X[18], x7 = X[21];
      T *Kd = sKc + (j / 3) * 21 + 7 * (j % 3);
      const T a0 = x1 - x0, a1 = (x3 - x2) - a0, k0 = x4 - x0, k1 = (x5 - x1) - k0;
      Kd[0] = a0;
      Kd[1] = a1;
      Kd[2] = x2 - x0;
      Kd[3] = k0;
      Kd[4] = k1;
      Kd[5] = (x6 - x2) - k0;
      Kd[6] = ((x7 - x6) - (x5 - x4)) - a1;
    }
  }
  // 2. interpolate along x: s1[z][y][qx], z,y < NM (LINE: thread (qx, qy)
  // interpolates the node x-line y = qx, z = qy whole)
  if constexpr (LINE) {
    if (active && qx < NM && qy < NM) {
      T ln[NM];
#pragma unroll
      for (int a = 0; a < NM; ++a) ln[a] = s0[el][qy * NM2 + qx * NM + a];
      T o[NQ];
      eo_tab<NQ, NM, 1>(geo, ln, o);
#pragma unroll
      for (int i = 0; i < NQ; ++i) s1[el][qy * NM * NQ + qx * NQ + i] = o[i];
    }
  } else if (active && qy < NM) {
    T bx[NM];  // this lane's row of B, read from LDS once
#pragma unroll
    for (int a = 0; a < NM; ++a) bx[a] = sB[qx * NM + a];
#pragma unroll
    for (int z = 0; z < NM; ++z) {
      T t = 0;
#pragma unroll
      for (int a = 0; a < NM; ++a) t += bx[a] * s0[el][z * NM2 + qy * NM + a];
      s1[el][z * NM * NQ + qy * NQ + qx] = t;
    }
  }
  __syncthreads();

  // 3./4. interpolate along y (LDS) and z (registers), then 5. gradient
  // OTF: the column's Jacobian columns J0 = a0 + zeta da, J1 = b0 + zeta db,
  // J2 = jz at (xi, eta) = (gxq[qx], gxq[qy])
  T ca0[3], cda[3], cb0[3], cdb[3], cjz[3];
  if (OTF && active) {
    const T xi = gxq[qx], eta = gxq[qy];
    const T *K = sKc + el * 21;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T A0 = K[7 * d], A1 = K[7 * d + 1], G0 = K[7 * d + 2], K0 = K[7 * d + 3];
      const T K1 = K[7 * d + 4], K2 = K[7 * d + 5], K3 = K[7 * d + 6];
      ca0[d] = A0 + eta * A1;
      cda[d] = K1 + eta * K3;
      cb0[d] = G0 + xi * A1;
      cdb[d] = K2 + xi * K3;
      cjz[d] = (K0 + xi * K1) + eta * cdb[d];
    }
  }
  T uq[NQ];
  if (PF && active) {
#pragma unroll
    for (int z = 0; z < NQ; ++z)
#pragma unroll
      for (int c = 0; c < 6; ++c) g[c][z] = __builtin_nontemporal_load(Ge + c * NQ3 + z * NQ2);
  }
  if constexpr (LINE) {
    // thread (qx, qy) interpolates the y-line x = qx, z = qy (< NM) into s0
    if (active && qy < NM) {
      T ln[NM];
#pragma unroll
      for (int c = 0; c < NM; ++c) ln[c] = s1[el][qy * NM * NQ + c * NQ + qx];
      T o[NQ];
      eo_tab<NQ, NM, 1>(geo, ln, o);
#pragma unroll
      for (int yq = 0; yq < NQ; ++yq) s0[el][qy * NQ2 + yq * NQ + qx] = o[yq];
    }
    __syncthreads();
  }
  if (active) {
    T t2[NM], by[LINE ? 1 : NM];
#pragma unroll
    for (int b = 0; b < (LINE ? 0 : NM); ++b) by[b] = sB[qy * NM + b];
#pragma unroll
    for (int z = 0; z < NM; ++z) {
      T t = 0;
      if constexpr (LINE) {
        t = s0[el][z * NQ2 + col];
      } else {
#pragma unroll
        for (int b = 0; b < NM; ++b) t += by[b] * s1[el][z * NM * NQ + b * NQ + qx];
      }
      t2[z] = t;
    }
    if constexpr (LINE) {
      eo_tab<NQ, NM, 1>(geo, t2, uq);
#pragma unroll
      for (int zq = 0; zq < NQ; ++zq) s2[el][zq * NQ2 + col] = uq[zq];
    } else {
#pragma unroll
      for (int zq = 0; zq < NQ; ++zq) {
        T t = 0;
#pragma unroll
        for (int c = 0; c < NM; ++c) t += uB(zq, c) * t2[c];
        uq[zq] = t;
        s2[el][zq * NQ2 + col] = t;
      }
    }
  }
  __syncthreads();

  // p.Ap (DOT): the sum over the cells of u_c^T A_c u_c, formed at the
  // quadrature points as grad u . (G grad u) while both are in registers
  // (constrained entries are gathered as zero, so this is the sum over the
  // free rows of p_i (A p)_i)
  // LINE: the x and y derivatives (and their transposes in 6.) are taken
  // along whole lines first -- thread (l0, l1) = (qx, qy) of a cell
  // differentiates the x-line (y = l0, z = l1) and the y-line (x = l0,
  // z = l1) into s0 / s1, which are free after the interpolation -- so each
  // point's value is read once per line instead of NQ times per point.
  if constexpr (LINE) {
    if (active) {
      T ln[NQ];
#pragma unroll
      for (int a = 0; a < NQ; ++a) ln[a] = s2[el][qy * NQ2 + qx * NQ + a];
      T o[NQ];
      eo_tab<NQ, NQ, -1>(geo + EO_D, ln, o);
#pragma unroll
      for (int i = 0; i < NQ; ++i) s0[el][qy * NQ2 + qx * NQ + i] = o[i];
#pragma unroll
      for (int a = 0; a < NQ; ++a) ln[a] = s2[el][qy * NQ2 + a * NQ + qx];
      eo_tab<NQ, NQ, -1>(geo + EO_D, ln, o);
#pragma unroll
      for (int i = 0; i < NQ; ++i) s1[el][qy * NQ2 + i * NQ + qx] = o[i];
    }
    __syncthreads();
  }
  double dot = 0.0;
  T fz[NQ];
  if (active) {
    T dxr[LINE ? 1 : NQ], dyr[LINE ? 1 : NQ];  // this lane's rows of D
#pragma unroll
    for (int a = 0; a < (LINE ? 0 : NQ); ++a) {
      dxr[a] = sD[qx * NQ + a];
      dyr[a] = sD[qy * NQ + a];
    }
    const T wxy = OTF ? gwq[qx] * gwq[qy] : T(0);
    T dzl[LINE ? NQ : 1];  // LINE: the column's z derivatives
    if constexpr (LINE) eo_tab<NQ, NQ, -1>(geo + EO_D, uq, dzl);
    // the column's points.  AX (cells of an undeformed box, whose Jacobian
    // columns lie along x, y and z): the cofactor columns and det J are the
    // same at every point of the column and diagonal, so they are formed once
    // and the products with their zero entries are left out
    [[maybe_unused]] T H0[3], H1[3], H2[3], hr = T(0);
    if constexpr (OTF && AX) {
      H0[0] = cb0[1] * cjz[2] - cb0[2] * cjz[1];
      H0[1] = cb0[2] * cjz[0] - cb0[0] * cjz[2];
      H0[2] = cb0[0] * cjz[1] - cb0[1] * cjz[0];
      H1[0] = cjz[1] * ca0[2] - cjz[2] * ca0[1];
      H1[1] = cjz[2] * ca0[0] - cjz[0] * ca0[2];
      H1[2] = cjz[0] * ca0[1] - cjz[1] * ca0[0];
      H2[0] = ca0[1] * cb0[2] - ca0[2] * cb0[1];
      H2[1] = ca0[2] * cb0[0] - ca0[0] * cb0[2];
      H2[2] = ca0[0] * cb0[1] - ca0[1] * cb0[0];
      hr = recip(cjz[0] * H2[0] + cjz[1] * H2[1] + cjz[2] * H2[2]);
    }
#pragma unroll
    for (int z = 0; z < NQ; ++z) {
      T dx = 0, dy = 0, dz = 0;
      if constexpr (LINE) {
        dx = s0[el][z * NQ2 + col];
        dy = s1[el][z * NQ2 + col];
        dz = dzl[z];
      } else {
#pragma unroll
        for (int a = 0; a < NQ; ++a) {
          dx += dxr[a] * s2[el][z * NQ2 + qy * NQ + a];
          dy += dyr[a] * s2[el][z * NQ2 + a * NQ + qx];
          dz += uD(z, a) * uq[a];
        }
      }
      if constexpr (OTF) {
        T fx, fy, fzv;
        if constexpr (AX) {
          const T sc = wxy * gwq[z] * hr;
          fx = sc * (H0[0] * (dx * H0[0]));
          fy = sc * (H1[1] * (dy * H1[1]));
          fzv = sc * (H2[2] * (dz * H2[2]));
        } else {
          const T zeta = gxq[z];
          T J0[3], J1[3], C0[3], C1[3], C2[3], W[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            J0[d] = ca0[d] + zeta * cda[d];
            J1[d] = cb0[d] + zeta * cdb[d];
          }
          C0[0] = J1[1] * cjz[2] - J1[2] * cjz[1];
          C0[1] = J1[2] * cjz[0] - J1[0] * cjz[2];
          C0[2] = J1[0] * cjz[1] - J1[1] * cjz[0];
          C1[0] = cjz[1] * J0[2] - cjz[2] * J0[1];
          C1[1] = cjz[2] * J0[0] - cjz[0] * J0[2];
          C1[2] = cjz[0] * J0[1] - cjz[1] * J0[0];
          C2[0] = J0[1] * J1[2] - J0[2] * J1[1];
          C2[1] = J0[2] * J1[0] - J0[0] * J1[2];
          C2[2] = J0[0] * J1[1] - J0[1] * J1[0];
          const T det = cjz[0] * C2[0] + cjz[1] * C2[1] + cjz[2] * C2[2];
          const T sc = wxy * gwq[z] * recip(det);
#pragma unroll
          for (int d = 0; d < 3; ++d) W[d] = dx * C0[d] + dy * C1[d] + dz * C2[d];
          fx = sc * (C0[0] * W[0] + C0[1] * W[1] + C0[2] * W[2]);
          fy = sc * (C1[0] * W[0] + C1[1] * W[1] + C1[2] * W[2]);
          fzv = sc * (C2[0] * W[0] + C2[1] * W[1] + C2[2] * W[2]);
        }
        s0[el][z * NQ2 + col] = fx;
        s1[el][z * NQ2 + col] = fy;
        fz[z] = fzv;
        if (DOT) dot += (double)dx * (double)fx + (double)dy * (double)fy + (double)dz * (double)fz[z];
      } else {
        T gxx, gxy, gxz, gyy, gyz, gzz;
        if constexpr (PF) {
          gxx = g[0][z]; gxy = g[1][z]; gxz = g[2][z];
          gyy = g[3][z]; gyz = g[4][z]; gzz = g[5][z];
        } else {
          gxx = __builtin_nontemporal_load(Ge + 0 * NQ3 + z * NQ2);
          gxy = __builtin_nontemporal_load(Ge + 1 * NQ3 + z * NQ2);
          gxz = __builtin_nontemporal_load(Ge + 2 * NQ3 + z * NQ2);
          gyy = __builtin_nontemporal_load(Ge + 3 * NQ3 + z * NQ2);
          gyz = __builtin_nontemporal_load(Ge + 4 * NQ3 + z * NQ2);
          gzz = __builtin_nontemporal_load(Ge + 5 * NQ3 + z * NQ2);
        }
        const T fx = gxx * dx + gxy * dy + gxz * dz;
        const T fy = gxy * dx + gyy * dy + gyz * dz;
        s0[el][z * NQ2 + col] = fx;
        s1[el][z * NQ2 + col] = fy;
        fz[z] = gxz * dx + gyz * dy + gzz * dz;
        if (DOT) dot += (double)dx * (double)fx + (double)dy * (double)fy + (double)dz * (double)fz[z];
      }
    }
  }
  __syncthreads();

  // 6. transposed gradient, 7. transposed interpolation along z (registers)
  if constexpr (LINE) {
    if (active) {
      T ln[NQ];
#pragma unroll
      for (int i = 0; i < NQ; ++i) ln[i] = s0[el][qy * NQ2 + qx * NQ + i];
      T o[NQ];
      eo_tab<NQ, NQ, -1>(geo + EO_DT, ln, o);
#pragma unroll
      for (int k = 0; k < NQ; ++k) s0[el][qy * NQ2 + qx * NQ + k] = o[k];
#pragma unroll
      for (int i = 0; i < NQ; ++i) ln[i] = s1[el][qy * NQ2 + i * NQ + qx];
      eo_tab<NQ, NQ, -1>(geo + EO_DT, ln, o);
#pragma unroll
      for (int k = 0; k < NQ; ++k) s1[el][qy * NQ2 + k * NQ + qx] = o[k];
    }
    __syncthreads();
  }
  if (active) {
    T w[NQ], dxt[LINE ? 1 : NQ], dyt[LINE ? 1 : NQ];  // this lane's columns of D
#pragma unroll
    for (int a = 0; a < (LINE ? 0 : NQ); ++a) {
      dxt[a] = sD[a * NQ + qx];
      dyt[a] = sD[a * NQ + qy];
    }
    T dzt[LINE ? NQ : 1];  // LINE: D^T of the column's z fluxes
    if constexpr (LINE) eo_tab<NQ, NQ, -1>(geo + EO_DT, fz, dzt);
#pragma unroll
    for (int z = 0; z < NQ; ++z) {
      T v = 0;
      if constexpr (LINE) {
        v = s0[el][z * NQ2 + col] + s1[el][z * NQ2 + col] + dzt[z];
      } else {
#pragma unroll
        for (int a = 0; a < NQ; ++a) {
          v += dxt[a] * s0[el][z * NQ2 + qy * NQ + a];
          v += dyt[a] * s1[el][z * NQ2 + a * NQ + qx];
          v += uD(a, z) * fz[a];
        }
      }
      w[z] = v;
    }
    if constexpr (LINE) {
      T o[NM];
      eo_tab<NM, NQ, 1>(geo + EO_BT, w, o);
#pragma unroll
      for (int zn = 0; zn < NM; ++zn) s2[el][zn * NQ2 + col] = o[zn];
    } else {
#pragma unroll
      for (int zn = 0; zn < NM; ++zn) {
        T t = 0;
#pragma unroll
        for (int c = 0; c < NQ; ++c) t += uB(c, zn) * w[c];
        s2[el][zn * NQ2 + col] = t;
      }
    }
  }
  __syncthreads();

  // 8. transposed interpolation along y: s0[zn][yn][qx] (LINE: thread
  // (qx, qy) takes the y-line x = qx, zn = qy whole)
  if constexpr (LINE) {
    if (active && qy < NM) {
      T ln[NQ];
#pragma unroll
      for (int b = 0; b < NQ; ++b) ln[b] = s2[el][qy * NQ2 + b * NQ + qx];
      T o[NM];
      eo_tab<NM, NQ, 1>(geo + EO_BT, ln, o);
#pragma unroll
      for (int yn = 0; yn < NM; ++yn) s0[el][qy * NM * NQ + yn * NQ + qx] = o[yn];
    }
  } else if (active && qy < NM) {
    T byt[NQ];
#pragma unroll
    for (int b = 0; b < NQ; ++b) byt[b] = sB[b * NM + qy];
#pragma unroll
    for (int zn = 0; zn < NM; ++zn) {
      T t = 0;
#pragma unroll
      for (int b = 0; b < NQ; ++b) t += byt[b] * s2[el][zn * NQ2 + b * NQ + qx];
      s0[el][zn * NM * NQ + qy * NQ + qx] = t;
    }
  }
  __syncthreads();

  // 9. transposed interpolation along x and scatter
  constexpr int P = NM - 1, XW = EPB * P + 1;  // ROW: nodes of a full segment row
  T *acc = &s1[0][0];  // ROW: [zn][yn][X], X = c P + xn, the left faces
  T *rgt = &s2[0][0];  // ROW: [zn][yn][c], right faces of all but the last cell
  static_assert(!ROW || NM * NM * XW <= EPB * NQ3, "row segment does not fit s1");
  static_assert(!ROW || NM * NM * EPB <= EPB * NQ3, "right faces do not fit s2");
  if constexpr (LINE && ROW) {
    // thread (qx, qy) takes the x-line yn = qx, zn = qy whole
    if (node_col) {
      T ln[NQ];
#pragma unroll
      for (int a = 0; a < NQ; ++a) ln[a] = s0[el][qy * NM * NQ + qx * NQ + a];
      T o[NM];
      eo_tab<NM, NQ, 1>(geo + EO_BT, ln, o);
#pragma unroll
      for (int xn = 0; xn < NM; ++xn) {
        const T v = o[xn];
        if (xn < P || (uint32_t)el + 1 == m)
          acc[(qy * NM + qx) * XW + el * P + xn] = v;
        else
          rgt[(qy * NM + qx) * EPB + el] = v;
      }
    }
  } else if (node_col) {
    T bxt[NQ];
#pragma unroll
    for (int a = 0; a < NQ; ++a) bxt[a] = sB[a * NM + qx];
#pragma unroll
    for (int zn = 0; zn < NM; ++zn) {
      T v = 0;
#pragma unroll
      for (int a = 0; a < NQ; ++a) v += bxt[a] * s0[el][zn * NM * NQ + qy * NQ + a];
      if constexpr (ROW) {
        if (qx < P || (uint32_t)el + 1 == m)
          acc[(zn * NM + qy) * XW + el * P + qx] = v;
        else
          rgt[(zn * NM + qy) * EPB + el] = v;
      } else if (GATHER) {
        if (idx[zn] != kInvalid) atomic_add(dst + idx[zn], v);
      } else {
        dst[(size_t)e * NM3 + zn * NM2 + qy * NM + qx] = v;
      }
    }
  }
  if constexpr (ROW) {
    __syncthreads();
    // node X of row r = zn NM + yn of the segment: its left-face value plus,
    // on an inner face, the right face of the cell before it; the map entry
    // of the cell that holds X as its left face or interior node (the last
    // cell for the segment's end)
    const int nx = (int)m * P + 1;
    for (int i = tid; i < NM * NM * nx; i += THREADS) {
      const int r = i / nx, X = i - r * nx;
      const int c = X / P;
      T v = acc[r * XW + X];
      if (X == c * P && c >= 1 && c < (int)m) v += rgt[r * EPB + c - 1];
      const int co = min(c, (int)m - 1);
      const uint32_t gi = map[(size_t)(e0 + co) * NM3 + r * NM + (X - co * P)];
      if (gi != kInvalid) atomic_add(dst + gi, v);
    }
  }
  if (DOT) {
    const double s = block_sum<Tile<NQ>::THREADS>(dot, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
}

// Workgroups of an apply: cells in batches of EPB, or (Gauss kernel on a box
// mesh, row_cells > 0) EPB-cell segments of each x-row.
template <int NQ, bool COLLOC>
uint32_t n_blocks_for(uint32_t n_cells, uint32_t row_cells = 0) {
  constexpr int EPB = COLLOC ? CollocTile<NQ>::EPB : Tile<NQ>::EPB;
  if (!COLLOC && row_cells)
    return (n_cells / row_cells) * ((row_cells + EPB - 1) / EPB);
  return (n_cells + EPB - 1) / EPB;
}

template <typename T, int NM, int NQ, bool COLLOC, bool GATHER>
void launch_t(const T *B, const T *D, const T *G, const T *src, T *dst,
              const uint32_t *map, uint32_t n_cells, hipStream_t s,
              double *partials, const int *skip, uint32_t row_cells) {
  if (n_cells == 0) return;
  const bool row = !COLLOC && GATHER && row_cells > 0;
  MFX_REQUIRE(!row || n_cells % row_cells == 0, "row segments need whole x-rows of cells");
  const dim3 grid(n_blocks_for<NQ, COLLOC>(n_cells, row ? row_cells : 0));
  if constexpr (COLLOC) {
    static_assert(NM == NQ, "collocation needs nm == nq");
    const dim3 block(CollocTile<NQ>::THREADS);
    if (partials) {
      if constexpr (std::is_same<T, double>::value && GATHER)
        hipLaunchKernelGGL((k_apply_colloc<T, NQ, CollocTile<NQ>::EPB, CollocTile<NQ>::PREFETCH,
                                           GATHER, true>),
                           grid, block, 0, s, D, G, src, dst, map, n_cells, partials, skip,
                           nullptr, nullptr);
    } else {
      hipLaunchKernelGGL((k_apply_colloc<T, NQ, CollocTile<NQ>::EPB, CollocTile<NQ>::PREFETCH,
                                         GATHER, false>),
                         grid, block, 0, s, D, G, src, dst, map, n_cells, nullptr, nullptr,
                         nullptr, nullptr);
    }
  } else {
    const dim3 block(Tile<NQ>::THREADS);
    if (partials) {
      if constexpr (std::is_same<T, double>::value && GATHER) {
        if (row)
          hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, GATHER, true, false, true>), grid, block,
                             0, s, B, D, G, src, dst, map, n_cells, partials, skip, nullptr,
                             nullptr, nullptr, row_cells, (const T *)nullptr);
        else
          hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, GATHER, true>), grid, block, 0, s, B, D,
                             G, src, dst, map, n_cells, partials, skip, nullptr, nullptr, nullptr,
                             0u, (const T *)nullptr);
      }
    } else {
      if constexpr (GATHER) {
        if (row) {
          hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, GATHER, false, false, true>), grid,
                             block, 0, s, B, D, G, src, dst, map, n_cells, nullptr, nullptr,
                             nullptr, nullptr, nullptr, row_cells, (const T *)nullptr);
          MFX_HIP(hipGetLastError());
          return;
        }
      }
      hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, GATHER, false>), grid, block, 0, s,
                         B, D, G, src, dst, map, n_cells, nullptr, nullptr, nullptr, nullptr,
                         nullptr, 0u, (const T *)nullptr);
    }
  }
  MFX_HIP(hipGetLastError());
}

// Runtime (p, nq, collocation) -> template instance.  Supported: p = 1..8;
// collocation nq = p+1; interpolation nq = p+1 or p+2.
template <typename T, bool GATHER>
void dispatch(int p, int nq, bool colloc, const T *B, const T *D, const T *G,
              const T *src, T *dst, const uint32_t *map, uint32_t n,
              hipStream_t s, double *partials, const int *skip, uint32_t rc) {
#define MFX_CASE_C(P)                                                         \
  case P:                                                                     \
    return launch_t<T, P + 1, P + 1, true, GATHER>(B, D, G, src, dst, map, n, s, partials, skip, rc);
#define MFX_CASE_I(P)                                                         \
  case P:                                                                     \
    if (nq == P + 1)                                                          \
      return launch_t<T, P + 1, P + 1, false, GATHER>(B, D, G, src, dst, map, n, s, partials, skip, rc); \
    if (nq == P + 2)                                                          \
      return launch_t<T, P + 1, P + 2, false, GATHER>(B, D, G, src, dst, map, n, s, partials, skip, rc); \
    break;
  if (colloc) {
    MFX_REQUIRE(nq == p + 1, "collocation requires nq == p+1");
    switch (p) { MFX_CASE_C(1) MFX_CASE_C(2) MFX_CASE_C(3) MFX_CASE_C(4)
                 MFX_CASE_C(5) MFX_CASE_C(6) MFX_CASE_C(7) MFX_CASE_C(8) default: break; }
  } else {
    switch (p) { MFX_CASE_I(1) MFX_CASE_I(2) MFX_CASE_I(3) MFX_CASE_I(4)
                 MFX_CASE_I(5) MFX_CASE_I(6) MFX_CASE_I(7) MFX_CASE_I(8) default: break; }
  }
#undef MFX_CASE_C
#undef MFX_CASE_I
  throw Error{MFX_ERR_UNSUPPORTED,
              "unsupported (degree, n_q_points_1d) = (" + std::to_string(p) + ", " +
                  std::to_string(nq) + ")"};
}

// Tile variants of the collocation kernel the autotuner chooses from: three
// cell counts per workgroup, each with and without the G prefetch.
template <int NQ, int EPB, bool PREF>
void launch_cv(const double *D, const double *G, const double *src, double *dst,
               const uint32_t *map, uint32_t n_cells, const uint16_t *cell_inst,
               const int *inst_done, hipStream_t s) {
  const dim3 grid((n_cells + EPB - 1) / EPB), block(EPB * NQ * NQ);
  hipLaunchKernelGGL((k_apply_colloc<double, NQ, EPB, PREF, true, false>), grid, block, 0, s, D, G,
                     src, dst, map, n_cells, nullptr, nullptr, cell_inst, inst_done);
  MFX_HIP(hipGetLastError());
}
template <int NQ, int E0, int E1, int E2>
void launch_cv_sel(int v, int *epb, int *pref, const double *D, const double *G, const double *src,
                   double *dst, const uint32_t *map, uint32_t n, const uint16_t *ci,
                   const int *done, hipStream_t s) {
  const int e[3] = {E0, E1, E2};
  MFX_REQUIRE(v >= 0 && v < 6, "tile variant out of range");
  *epb = e[v / 2];
  *pref = (v % 2 == 0);
  if (!src) return;  // description only
  if (n == 0) return;
  switch (v) {
    case 0: return launch_cv<NQ, E0, true>(D, G, src, dst, map, n, ci, done, s);
    case 1: return launch_cv<NQ, E0, false>(D, G, src, dst, map, n, ci, done, s);
    case 2: return launch_cv<NQ, E1, true>(D, G, src, dst, map, n, ci, done, s);
    case 3: return launch_cv<NQ, E1, false>(D, G, src, dst, map, n, ci, done, s);
    case 4: return launch_cv<NQ, E2, true>(D, G, src, dst, map, n, ci, done, s);
    default: return launch_cv<NQ, E2, false>(D, G, src, dst, map, n, ci, done, s);
  }
}

}  // namespace

int colloc_variant_count(int p) { return (p >= 2 && p <= 6) ? 6 : 0; }

int colloc_default_variant(int p) {
  // the variant equal to the default tile (CollocTile) of launch_apply
  switch (p) {
    case 2: return 2;  // 14 cells, prefetch
    case 3: return 2;  // 8 cells, prefetch
    case 4: return 2;  // 5 cells, prefetch
    case 5: return 2;  // 3 cells, prefetch
    case 6: return 0;  // 1 cell, prefetch
    default: return 0;
  }
}

void launch_colloc_variant(int p, int v, const double *D, const double *G, const double *src,
                           double *dst, const uint32_t *map, uint32_t n_cells,
                           const uint16_t *cell_inst, const int *inst_done, hipStream_t s,
                           int *epb, int *prefetch) {
  int e = 0, pf = 0;
  switch (p) {
    case 2: launch_cv_sel<3, 7, 14, 28>(v, &e, &pf, D, G, src, dst, map, n_cells, cell_inst, inst_done, s); break;
    case 3: launch_cv_sel<4, 4, 8, 16>(v, &e, &pf, D, G, src, dst, map, n_cells, cell_inst, inst_done, s); break;
    case 4: launch_cv_sel<5, 2, 5, 10>(v, &e, &pf, D, G, src, dst, map, n_cells, cell_inst, inst_done, s); break;
    case 5: launch_cv_sel<6, 1, 3, 7>(v, &e, &pf, D, G, src, dst, map, n_cells, cell_inst, inst_done, s); break;
    case 6: launch_cv_sel<7, 1, 2, 5>(v, &e, &pf, D, G, src, dst, map, n_cells, cell_inst, inst_done, s); break;
    default: throw Error{MFX_ERR_UNSUPPORTED, "tile variants exist for degrees 2..6"};
  }
  if (epb) *epb = e;
  if (prefetch) *prefetch = pf;
}

template <typename T>
void launch_apply(int p, int nq, bool colloc, const T *B, const T *D, const T *G,
                  const T *src, T *dst, const uint32_t *map, uint32_t n_cells,
                  hipStream_t s, double *dot_partials, const int *skip, uint32_t row_cells) {
  dispatch<T, true>(p, nq, colloc, B, D, G, src, dst, map, n_cells, s, dot_partials, skip,
                    row_cells);
}
template <typename T>
void launch_evec(int p, int nq, bool colloc, const T *B, const T *D, const T *G,
                 const T *in, T *out, uint32_t n_cells, hipStream_t s) {
  dispatch<T, false>(p, nq, colloc, B, D, G, in, out, nullptr, n_cells, s, nullptr, nullptr, 0u);
}

// Gauss (non-collocated) apply with the metric terms computed from the cell
// vertices (k_apply_interp<..., OTF>).
template <typename T, int NM, int NQ>
void launch_interp_otf(const T *B, const T *D, const T *eo, const T *xq, const T *wq,
                       const double *cellX, const T *src, T *dst, const uint32_t *map,
                       uint32_t n_cells, hipStream_t s, double *partials, const int *skip,
                       uint32_t row_cells, bool axes) {
  MFX_REQUIRE(!axes || row_cells, "axis-aligned cells are a box-mesh form");
  if (n_cells == 0) return;
  MFX_REQUIRE(!row_cells || n_cells % row_cells == 0, "row segments need whole x-rows of cells");
  const dim3 grid(n_blocks_for<NQ, false>(n_cells, row_cells)), block(Tile<NQ>::THREADS);
  if (partials) {
    if constexpr (std::is_same<T, double>::value) {
      if (axes)
        hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, true, true, true, true, true>), grid, block,
                           0, s, B, D, nullptr, src, dst, map, n_cells, partials, skip, cellX, xq,
                           wq, row_cells, eo);
      else if (row_cells)
        hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, true, true, true, true>), grid, block, 0, s,
                           B, D, nullptr, src, dst, map, n_cells, partials, skip, cellX, xq, wq,
                           row_cells, eo);
      else
        hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, true, true, true>), grid, block, 0, s, B, D,
                           nullptr, src, dst, map, n_cells, partials, skip, cellX, xq, wq, 0u,
                           eo);
    } else {
      throw Error{MFX_ERR_UNSUPPORTED, "p.Ap partials are computed in FP64 only"};
    }
  } else if (axes) {
    hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, true, false, true, true, true>), grid, block, 0,
                       s, B, D, nullptr, src, dst, map, n_cells, nullptr, nullptr, cellX, xq, wq,
                       row_cells, eo);
  } else if (row_cells) {
    hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, true, false, true, true>), grid, block, 0, s, B,
                       D, nullptr, src, dst, map, n_cells, nullptr, nullptr, cellX, xq, wq,
                       row_cells, eo);
  } else {
    hipLaunchKernelGGL((k_apply_interp<T, NM, NQ, true, false, true>), grid, block, 0, s, B, D,
                       nullptr, src, dst, map, n_cells, nullptr, nullptr, cellX, xq, wq, 0u,
                       eo);
  }
  MFX_HIP(hipGetLastError());
}

template <typename T>
void launch_apply_otf(int p, int nq, const T *B, const T *D, const T *eo, const T *xq,
                      const T *wq, const double *cellX, const T *src, T *dst,
                      const uint32_t *map, uint32_t n_cells, hipStream_t s, double *partials,
                      const int *skip, uint32_t row_cells, bool axes) {
  MFX_REQUIRE(B && D && eo && xq && wq && cellX, "on-the-fly apply: missing tables or vertices");
#define MFX_CASE_O(P)                                                                         \
  case P:                                                                                    \
    if (nq == P + 1)                                                                         \
      return launch_interp_otf<T, P + 1, P + 1>(B, D, eo, xq, wq, cellX, src, dst, map,      \
                                                n_cells, s, partials, skip, row_cells, axes); \
    if (nq == P + 2)                                                                         \
      return launch_interp_otf<T, P + 1, P + 2>(B, D, eo, xq, wq, cellX, src, dst, map,      \
                                                n_cells, s, partials, skip, row_cells, axes); \
    break;
  switch (p) {
    MFX_CASE_O(1) MFX_CASE_O(2) MFX_CASE_O(3) MFX_CASE_O(4)
    MFX_CASE_O(5) MFX_CASE_O(6) MFX_CASE_O(7) MFX_CASE_O(8) default: break;
  }
#undef MFX_CASE_O
  throw Error{MFX_ERR_UNSUPPORTED,
              "unsupported (degree, n_q_points_1d) = (" + std::to_string(p) + ", " +
                  std::to_string(nq) + ")"};
}

template void launch_apply_otf<double>(int, int, const double *, const double *, const double *,
                                       const double *, const double *, const double *,
                                       const double *, double *, const uint32_t *, uint32_t,
                                       hipStream_t, double *, const int *, uint32_t, bool);
template void launch_apply_otf<float>(int, int, const float *, const float *, const float *,
                                      const float *, const float *, const double *, const float *,
                                      float *, const uint32_t *, uint32_t, hipStream_t, double *,
                                      const int *, uint32_t, bool);

template void launch_apply<double>(int, int, bool, const double *, const double *,
                                   const double *, const double *, double *,
                                   const uint32_t *, uint32_t, hipStream_t, double *,
                                   const int *, uint32_t);
template void launch_apply<float>(int, int, bool, const float *, const float *,
                                  const float *, const float *, float *,
                                  const uint32_t *, uint32_t, hipStream_t, double *,
                                  const int *, uint32_t);
template void launch_evec<double>(int, int, bool, const double *, const double *,
                                  const double *, const double *, double *, uint32_t,
                                  hipStream_t);

uint32_t apply_n_blocks(int p, int nq, bool colloc, uint32_t n_cells, uint32_t row_cells) {
  (void)p;
#define MFX_NB(Q)                                                                    \
  case Q:                                                                           \
    return colloc ? n_blocks_for<Q, true>(n_cells) : n_blocks_for<Q, false>(n_cells, row_cells);
  switch (nq) {
    MFX_NB(2) MFX_NB(3) MFX_NB(4) MFX_NB(5) MFX_NB(6) MFX_NB(7) MFX_NB(8) MFX_NB(9) MFX_NB(10)
    default: return 0;
  }
#undef MFX_NB
}

}  // namespace mfx
