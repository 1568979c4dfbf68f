// transfer.hpp -- grid transfers between two box operators of one box:
// tensor products of 1D nodal-interpolation matrices (CSR, on the device),
// applied by the axis passes of mg_kernels.hip.  Used by the multigrid
// (mg.cpp) and by the learned preconditioner's coarse correction (nn.cpp).
#pragma once

#include <algorithm>
#include <vector>

#include "mg.hpp"
#include "op.hpp"

namespace mfx {

template <typename T>
struct DevAxis {
  DBuf<int> rowptr, col;
  DBuf<T> val;
  int rows = 0, cols = 0;
  AxisMatrix<T> view() const { return {rows, cols, rowptr.ptr, col.ptr, val.ptr}; }
};

struct HostCsr {
  int rows = 0, cols = 0;
  std::vector<int> rowptr, col;
  std::vector<double> val;
};

// 1D prolongation from the coarse node line (nc cells of degree pc) to the
// fine one (nf cells of degree pf, nf = r nc with r >= 1): the fine node at
// reference coordinate xi_f of fine cell e_f takes the coarse degree-pc
// interpolant at the same point, coarse cell e_f / r at xi_c = (e_f mod r +
// xi_f) / r.  The multigrid uses p-steps (r = 1, pf > pc) and h-steps (r =
// 2, pf == pc); both at once (e.g. Q4 -> Q1 on half the cells) also work.
inline HostCsr prolongation_1d(int nc, int pc, int nf, int pf) {
  MFX_REQUIRE(nc >= 1 && pc >= 1 && nf % nc == 0 && (nf > nc || pf > pc),
              "transfer: the fine level must refine the coarse one (cells by an integer factor, or degree)");
  const int r = nf / nc;
  std::vector<double> sup(pf + 1), sw(pf + 1), w(pc + 1);
  gll_rule(pf + 1, sup.data(), sw.data());
  HostCsr P;
  P.rows = nf * pf + 1;
  P.cols = nc * pc + 1;
  P.rowptr.push_back(0);
  for (int j = 0; j < P.rows; ++j) {
    const int ef = std::min(j / pf, nf - 1), l = j - ef * pf;
    const int ec = ef / r;
    const double xi = ((ef % r) + sup[l]) / r;
    lagrange_values(pc, xi, w.data());
    for (int m = 0; m <= pc; ++m)
      if (w[m] != 0.0) {
        P.col.push_back(ec * pc + m);
        P.val.push_back(w[m]);
      }
    P.rowptr.push_back((int)P.col.size());
  }
  return P;
}

inline HostCsr transpose(const HostCsr &A) {
  HostCsr T;
  T.rows = A.cols;
  T.cols = A.rows;
  T.rowptr.assign(T.rows + 1, 0);
  for (int c : A.col) ++T.rowptr[c + 1];
  for (int i = 0; i < T.rows; ++i) T.rowptr[i + 1] += T.rowptr[i];
  T.col.resize(A.col.size());
  T.val.resize(A.val.size());
  std::vector<int> pos(T.rowptr.begin(), T.rowptr.end() - 1);
  for (int r = 0; r < A.rows; ++r)
    for (int k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k) {
      const int at = pos[A.col[k]]++;
      T.col[at] = r;
      T.val[at] = A.val[k];
    }
  return T;
}

template <typename T>
void upload(const HostCsr &h, DevAxis<T> &d) {
  d.rows = h.rows;
  d.cols = h.cols;
  d.rowptr.upload(h.rowptr.data(), h.rowptr.size());
  d.col.upload(h.col.data(), h.col.size());
  std::vector<T> v(h.val.begin(), h.val.end());
  d.val.upload(v.data(), v.size());
}

// 1D matrices of one fine -> coarse step, per axis
struct Transfer {
  DevAxis<float> P32[3], R32[3];
  DevAxis<double> P64[3], R64[3];
  uint32_t fine[3] = {0, 0, 0}, coarse[3] = {0, 0, 0};
};

inline void build_transfer(const mfx_op &fine, const mfx_op &coarse, Transfer &t, bool fp32, bool fp64) {
  MFX_REQUIRE(fine.is_box && coarse.is_box, "transfer: both operators must come from mfx_op_create_box");
  for (int a = 0; a < 3; ++a) {
    MFX_REQUIRE(fine.box.lo[a] == coarse.box.lo[a] && fine.box.hi[a] == coarse.box.hi[a],
                "transfer: operators on different boxes");
    const HostCsr P = prolongation_1d((int)coarse.box.n[a], coarse.p, (int)fine.box.n[a], fine.p);
    const HostCsr R = transpose(P);
    if (fp32) {
      upload(P, t.P32[a]);
      upload(R, t.R32[a]);
    }
    if (fp64) {
      upload(P, t.P64[a]);
      upload(R, t.R64[a]);
    }
    t.fine[a] = (uint32_t)P.rows;
    t.coarse[a] = (uint32_t)P.cols;
  }
  MFX_REQUIRE(fine.box.deform == coarse.box.deform, "transfer: operators on differently deformed meshes");
}

// prolongation: coarse (c0,c1,c2) -> x-pass (f0,c1,c2) -> y-pass (f0,f1,c2) -> z-pass (f0,f1,f2)
template <typename T>
void prolongate(const DevAxis<T> *P, const uint32_t f[3], const uint32_t c[3], const T *src, T *tmp1,
                T *tmp2, T *dst, bool add, hipStream_t s) {
  const uint32_t d1[3] = {f[0], c[1], c[2]}, d2[3] = {f[0], f[1], c[2]};
  launch_axis_pass<T>(P[0].view(), 0, d1, src, tmp1, false, s);
  launch_axis_pass<T>(P[1].view(), 1, d2, tmp1, tmp2, false, s);
  launch_axis_pass<T>(P[2].view(), 2, f, tmp2, dst, add, s);
}
// restriction: fine (f0,f1,f2) -> z-pass (f0,f1,c2) -> y-pass (f0,c1,c2) -> x-pass (c0,c1,c2)
template <typename T>
void restrict_(const DevAxis<T> *R, const uint32_t f[3], const uint32_t c[3], const T *src, T *tmp1,
               T *tmp2, T *dst, hipStream_t s) {
  const uint32_t d1[3] = {f[0], f[1], c[2]}, d2[3] = {f[0], c[1], c[2]};
  launch_axis_pass<T>(R[2].view(), 2, d1, src, tmp1, false, s);
  launch_axis_pass<T>(R[1].view(), 1, d2, tmp1, tmp2, false, s);
  launch_axis_pass<T>(R[0].view(), 0, c, tmp2, dst, false, s);
}

}  // namespace mfx
