// capi.cpp -- extern "C" entry points of libmfx (declared in include/mfx.h).
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "op.hpp"
#include "tables.hpp"

namespace mfx {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

}  // namespace mfx

using namespace mfx;

static void finish_setup(mfx_op *op, int with_fp32, hipStream_t s) {
  const uint64_t nq3 = (uint64_t)op->nq * op->nq * op->nq;
  if (op->geometry == MFX_GEOM_STORED) {
    op->G.alloc((uint64_t)op->E * 6 * nq3);
    launch_geometry(op->E, op->cellX.ptr, op->nq, op->xq.ptr, op->wq.ptr, op->G.ptr, s);
  }
  if (op->patch) {
    const uint64_t npatch = patch_count(op->lay);
    const PatchShape sh = op->shape;
    const uint64_t P3 = (uint64_t)(sh.bx * op->p + 1) * (sh.by * op->p + 1) * (sh.bz * op->p + 1);
    op->pmap.alloc(npatch * P3);
    launch_patch_map(op->lay, op->p, sh, op->pmap.ptr, s);
    const bool stored = op->geometry == MFX_GEOM_STORED;
    if (stored) {
      op->Gp.alloc(patch_geometry_size(op->p, op->lay));
      launch_patch_geometry<double>(op->p, op->lay, op->G.ptr, op->Gp.ptr, s);
    }
    op->grid64 = patch_grid<double>(op->p, op->lay, stored);
    if (with_fp32) op->grid32 = patch_grid<float>(op->p, op->lay, stored);
  }
  if (with_fp32) {
    if (op->geometry == MFX_GEOM_STORED) {
      op->G32.alloc(op->G.n);
      launch_to_float(op->G.ptr, op->G32.ptr, op->G.n, s);
      if (op->patch) {
        op->Gp32.alloc(op->Gp.n);
        launch_patch_geometry<float>(op->p, op->lay, op->G32.ptr, op->Gp32.ptr, s);
      }
    }
    op->B32.alloc(op->B.n);
    launch_to_float(op->B.ptr, op->B32.ptr, op->B.n, s);
    op->D32.alloc(op->D.n);
    launch_to_float(op->D.ptr, op->D32.ptr, op->D.n, s);
    if (op->eo.n) {
      op->eo32.alloc(op->eo.n);
      launch_to_float(op->eo.ptr, op->eo32.ptr, op->eo.n, s);
    }
    op->xq32.alloc(op->xq.n);
    launch_to_float(op->xq.ptr, op->xq32.ptr, op->xq.n, s);
    op->wq32.alloc(op->wq.n);
    launch_to_float(op->wq.ptr, op->wq32.ptr, op->wq.n, s);
  }
  MFX_HIP(hipStreamSynchronize(s));
}

static void common_init(mfx_op *op, int p, int nq, int quad, int geometry) {
  MFX_REQUIRE(p >= 1 && p <= 8, "degree must be in 1..8");
  MFX_REQUIRE(quad == MFX_QUAD_GAUSS || quad == MFX_QUAD_GLL, "unknown quadrature");
  MFX_REQUIRE(nq == p + 1 || nq == p + 2, "n_q_points_1d must be degree+1 or degree+2");
  MFX_REQUIRE(geometry == MFX_GEOM_STORED || geometry == MFX_GEOM_ON_THE_FLY,
              "unknown geometry mode");
  op->p = p;
  op->nq = nq;
  op->quad = quad;
  op->geometry = geometry;
  op->colloc = (quad == MFX_QUAD_GLL && nq == p + 1);
  std::vector<double> B((size_t)nq * (p + 1)), D((size_t)nq * nq), xq(nq), wq(nq), sup(p + 1);
  basis_tables(p, nq, quad, B.data(), D.data(), xq.data(), wq.data(), sup.data());
  if (op->colloc) {
    // the patch kernel's compile-time derivative table must be this matrix
    double mx = 1.0;
    for (double v : D) mx = std::max(mx, std::fabs(v));
    MFX_REQUIRE(gll_d_table_error(nq, D.data()) <= 1e-13 * mx,
                "compile-time GLL derivative table disagrees with the basis tables");
  }
  op->B.upload(B.data(), B.size());
  op->D.upload(D.data(), D.size());
  if (!op->colloc) {
    // even-odd tables of the line-form Gauss kernel: B and D on symmetric
    // points are centro-symmetric and centro-antisymmetric
    const int nm = p + 1;
    std::vector<double> Bt((size_t)nm * nq), Dt((size_t)nq * nq), eo;
    for (int q = 0; q < nq; ++q)
      for (int m = 0; m < nm; ++m) Bt[(size_t)m * nq + q] = B[(size_t)q * nm + m];
    for (int i = 0; i < nq; ++i)
      for (int j = 0; j < nq; ++j) Dt[(size_t)j * nq + i] = D[(size_t)i * nq + j];
    double dev = eo_tables(B.data(), nq, nm, 1, eo);
    dev = std::max(dev, eo_tables(Bt.data(), nm, nq, 1, eo));
    dev = std::max(dev, eo_tables(D.data(), nq, nq, -1, eo));
    dev = std::max(dev, eo_tables(Dt.data(), nq, nq, -1, eo));
    MFX_REQUIRE(dev <= 1e-13, "basis tables are not centro-symmetric");
    op->eo.upload(eo.data(), eo.size());
  }
  op->xq.upload(xq.data(), xq.size());
  op->wq.upload(wq.data(), wq.size());
  op->support.upload(sup.data(), sup.size());
}

// Metric terms of the lexicographic cell order: the stored array, or a
// temporary one for on-the-fly operators (set-up kernels only).
struct GeometryView {
  const double *G = nullptr;
  DBuf<double> tmp;
};
static void lexicographic_G(const mfx_op *op, GeometryView &v, hipStream_t s) {
  if (op->G.ptr) {
    v.G = op->G.ptr;
    return;
  }
  v.tmp.alloc((uint64_t)op->E * 6 * op->nq * op->nq * op->nq);
  launch_geometry(op->E, op->cellX.ptr, op->nq, op->xq.ptr, op->wq.ptr, v.tmp.ptr, s);
  v.G = v.tmp.ptr;
}

namespace mfx {

size_t cg_partial_count(const mfx_op &op) {
  const uint32_t nb_apply = op.apply_blocks();
  const uint32_t nb_cons = cg_cons_blocks(op.n_cons);
  const uint32_t nb_vec = cg_vec_blocks(op.N);
  const size_t na = (size_t)nb_apply + nb_cons;
  return std::max<size_t>(na + cg_alpha_extra((uint32_t)na), 2 * (size_t)nb_vec);
}

// deal.II SolverCG + ReductionControl on the device (cg_state.hpp): the
// step kernels are enqueued in batches of 4, 8, ... 64 and the host reads
// the `done` flag once per batch.
void cg_run(mfx_op &op, double *x, const double *b, const double *inv_diag, int max_steps,
            double abs_tol, double rel_tol, const CgWork &w, mfx_cg_result *res, hipStream_t s) {
  const uint64_t N = op.N;
  const uint32_t nb_apply = op.apply_blocks();
  const uint32_t nb_cons = cg_cons_blocks(op.n_cons);
  const uint32_t nb_vec = cg_vec_blocks(N);
  CgState init{};
  MFX_HIP(hipMemcpyAsync(w.st, &init, sizeof(CgState), hipMemcpyHostToDevice, s));

  const auto t0 = std::chrono::steady_clock::now();
  // r = A x
  launch_fill(w.r, 0.0, N, s);
  op.apply64(x, w.r, s);
  launch_copy_listed<double>(w.r, x, op.cons.ptr, op.n_cons, s);
  cg_init(w.st, b, w.r, inv_diag, N, w.part, abs_tol, rel_tol, max_steps, s);
  const int *done = &w.st->done;
  int h_local = 0;
  int *h_done = w.h_done ? w.h_done : &h_local;
  int batch = 4;
  for (int it = 0;;) {
    for (int k = 0; k < batch; ++k, ++it) {
      cg_direction(w.st, w.r, inv_diag, w.p, w.ap, x, N, s);
      op.apply64(w.p, w.ap, s, w.part, done);
      cg_constrained(w.st, op.cons.ptr, op.n_cons, w.p, w.ap, w.part + nb_apply, s);
      cg_alpha(w.st, w.part, nb_apply + nb_cons, s);
      cg_update(w.st, w.r, w.ap, inv_diag, N, w.part, s);
      cg_check(w.st, w.part, nb_vec, s);
    }
    MFX_HIP(hipGetLastError());
    MFX_HIP(hipMemcpyAsync(h_done, done, sizeof(int), hipMemcpyDeviceToHost, s));
    MFX_HIP(hipStreamSynchronize(s));
    if (*h_done) break;
    if (it > max_steps + 8) break;  // device enforces max_steps; this bounds the host loop
    batch = std::min(batch * 2, 64);
  }
  cg_finish(w.st, x, w.p, N, s);
  CgState fin{};
  MFX_HIP(hipMemcpyAsync(&fin, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
  MFX_HIP(hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  res->steps = fin.steps;
  res->converged = fin.done == 1;
  res->initial_value = fin.initial;
  res->last_value = fin.last;
  res->seconds = std::chrono::duration<double>(t1 - t0).count();
  if (fin.done != 1)
    throw Error{fin.breakdown ? MFX_ERR_BREAKDOWN : MFX_ERR_NOT_CONVERGED,
                "CG did not converge: steps=" + std::to_string(fin.steps) +
                    " residual=" + std::to_string(fin.last)};
}

}  // namespace mfx

extern "C" {

const char *mfx_last_error(void) { return g_last_error.c_str(); }
int mfx_version(void) { return 100; }

mfx_status mfx_basis_tables(int p, int nq, int quadrature, double *B, double *D, double *xq,
                            double *wq) {
  return guard([&] {
    MFX_REQUIRE(B && D && xq && wq, "null pointer argument");
    MFX_REQUIRE(quadrature == MFX_QUAD_GAUSS || quadrature == MFX_QUAD_GLL, "unknown quadrature");
    std::vector<double> sup(p + 1 > 0 ? p + 1 : 1);
    basis_tables(p, nq, quadrature, B, D, xq, wq, sup.data());
  });
}

// ---------------------------------------------------------------- kernels
mfx_status mfx_bk3_apply_f64(int p, int nq, const double *B, const double *D, const double *G,
                             const double *src, double *dst, const uint32_t *map,
                             uint32_t n_cells, void *stream) {
  return guard([&] {
    MFX_REQUIRE(B && D && G && src && dst && map, "null pointer argument");
    launch_apply<double>(p, nq, false, B, D, G, src, dst, map, n_cells, as_stream(stream));
  });
}
mfx_status mfx_bk3_apply_f32(int p, int nq, const float *B, const float *D, const float *G,
                             const float *src, float *dst, const uint32_t *map,
                             uint32_t n_cells, void *stream) {
  return guard([&] {
    MFX_REQUIRE(B && D && G && src && dst && map, "null pointer argument");
    launch_apply<float>(p, nq, false, B, D, G, src, dst, map, n_cells, as_stream(stream));
  });
}
mfx_status mfx_bk5_apply_f64(int p, const double *D, const double *G, const double *src,
                             double *dst, const uint32_t *map, uint32_t n_cells, void *stream) {
  return guard([&] {
    MFX_REQUIRE(D && G && src && dst && map, "null pointer argument");
    launch_apply<double>(p, p + 1, true, nullptr, D, G, src, dst, map, n_cells,
                         as_stream(stream));
  });
}
mfx_status mfx_bk5_apply_f32(int p, const float *D, const float *G, const float *src,
                             float *dst, const uint32_t *map, uint32_t n_cells, void *stream) {
  return guard([&] {
    MFX_REQUIRE(D && G && src && dst && map, "null pointer argument");
    launch_apply<float>(p, p + 1, true, nullptr, D, G, src, dst, map, n_cells,
                        as_stream(stream));
  });
}
mfx_status mfx_bk3_evec_f64(int p, int nq, const double *B, const double *D, const double *G,
                            const double *in, double *out, uint32_t n_cells, void *stream) {
  return guard([&] {
    MFX_REQUIRE(B && D && G && in && out, "null pointer argument");
    launch_evec<double>(p, nq, false, B, D, G, in, out, n_cells, as_stream(stream));
  });
}
mfx_status mfx_bk5_evec_f64(int p, const double *D, const double *G, const double *in,
                            double *out, uint32_t n_cells, void *stream) {
  return guard([&] {
    MFX_REQUIRE(D && G && in && out, "null pointer argument");
    launch_evec<double>(p, p + 1, true, nullptr, D, G, in, out, n_cells, as_stream(stream));
  });
}

mfx_status mfx_gather_f64(const uint32_t *map, uint32_t n_cells, int nodes_per_cell,
                          const double *src, double *evec, void *stream) {
  return guard([&] {
    MFX_REQUIRE(map && src && evec, "null pointer argument");
    MFX_REQUIRE(nodes_per_cell > 0, "nodes_per_cell must be positive");
    launch_gather(map, (uint64_t)n_cells * nodes_per_cell, src, evec, as_stream(stream));
  });
}
mfx_status mfx_scatter_add_f64(const uint32_t *map, uint32_t n_cells, int nodes_per_cell,
                               const double *evec, double *dst, void *stream) {
  return guard([&] {
    MFX_REQUIRE(map && evec && dst, "null pointer argument");
    MFX_REQUIRE(nodes_per_cell > 0, "nodes_per_cell must be positive");
    launch_scatter_add(map, (uint64_t)n_cells * nodes_per_cell, evec, dst, as_stream(stream));
  });
}

// ---------------------------------------------------------------- operator
mfx_status mfx_op_create_box(const mfx_box_desc *d, mfx_op **out) {
  return guard([&] {
    MFX_REQUIRE(d && out, "null pointer argument");
    *out = nullptr;
    auto op = std::make_unique<mfx_op>();
    common_init(op.get(), d->degree, d->n_q_points_1d, d->quadrature, d->geometry);
    const int p = d->degree;
    uint64_t E = 1, N = 1, Ninner = 1;
    for (int k = 0; k < 3; ++k) {
      MFX_REQUIRE(d->n_cells[k] >= 1, "n_cells must be >= 1");
      MFX_REQUIRE(d->hi[k] > d->lo[k], "box must have hi > lo");
      E *= d->n_cells[k];
      const uint64_t nk = (uint64_t)d->n_cells[k] * p + 1;
      N *= nk;
      Ninner *= nk - 2;
    }
    MFX_REQUIRE(E <= 0x7FFFFFFFull, "too many cells");
    MFX_REQUIRE(N < 0x7FFFFFFFull, "too many DoFs for 31-bit DoF indices");
    op->E = (uint32_t)E;
    op->N = N;
    BoxGeom box;
    for (int k = 0; k < 3; ++k) {
      box.n[k] = d->n_cells[k];
      box.lo[k] = d->lo[k];
      box.hi[k] = d->hi[k];
    }
    box.deform = d->deform;
    op->is_box = true;
    op->box = box;
    // constrained DoFs: the boundary nodes of the structured node grid
    const uint64_t Nx = (uint64_t)d->n_cells[0] * p + 1, Ny = (uint64_t)d->n_cells[1] * p + 1,
                   Nz = (uint64_t)d->n_cells[2] * p + 1;
    std::vector<uint32_t> cons;
    cons.reserve(N - Ninner);
    for (uint64_t z = 0; z < Nz; ++z)
      for (uint64_t y = 0; y < Ny; ++y) {
        const bool face = (z == 0 || z == Nz - 1 || y == 0 || y == Ny - 1);
        const uint64_t row = Nx * (y + Ny * z);
        if (face) {
          for (uint64_t x = 0; x < Nx; ++x) cons.push_back((uint32_t)(row + x));
        } else {
          cons.push_back((uint32_t)row);
          cons.push_back((uint32_t)(row + Nx - 1));
        }
      }
    op->n_cons = cons.size();
    op->cons.upload(cons.data(), cons.size());
    // GLL collocation runs the patch kernel, with the metric terms from the
    // cell vertices or from G; Gauss quadrature runs the per-cell kernels
    if (op->colloc) {
      op->patch = true;
      op->shape = patch_shape(p);
      const int b[3] = {op->shape.bx, op->shape.by, op->shape.bz};
      for (int k = 0; k < 3; ++k) {
        op->lay.n[k] = d->n_cells[k];
        op->lay.np[k] = (d->n_cells[k] + b[k] - 1) / b[k];
      }
    }
    hipStream_t s = nullptr;
    op->cellX.alloc(E * 24);
    launch_box_cell_vertices(box, op->cellX.ptr, s);
    op->map.alloc(E * (uint64_t)(p + 1) * (p + 1) * (p + 1));
    launch_box_map(box, p, op->map.ptr, s);
    finish_setup(op.get(), d->with_fp32, s);
    *out = op.release();
  });
}

mfx_status mfx_op_create(const mfx_op_desc *d, mfx_op **out) {
  return guard([&] {
    MFX_REQUIRE(d && out, "null pointer argument");
    *out = nullptr;
    MFX_REQUIRE(d->map && d->cell_vertices, "map and cell_vertices are required");
    MFX_REQUIRE(d->n_constrained == 0 || d->constrained, "constrained list missing");
    MFX_REQUIRE(d->n_dofs < 0x7FFFFFFFull, "too many DoFs for 31-bit DoF indices");
    auto op = std::make_unique<mfx_op>();
    common_init(op.get(), d->degree, d->n_q_points_1d, d->quadrature, d->geometry);
    MFX_REQUIRE(d->geometry == MFX_GEOM_STORED || !op->colloc,
                "general meshes with GLL collocation use stored geometry (the on-the-fly "
                "collocation kernel works on patches of a box mesh)");
    const uint64_t nm3 = (uint64_t)(d->degree + 1) * (d->degree + 1) * (d->degree + 1);
    const uint64_t nmap = (uint64_t)d->n_cells * nm3;
    for (uint64_t i = 0; i < nmap; ++i)
      MFX_REQUIRE(d->map[i] == MFX_INVALID_DOF || d->map[i] < d->n_dofs,
                  "map entry out of range at " + std::to_string(i));
    for (uint64_t i = 0; i < d->n_constrained; ++i)
      MFX_REQUIRE(d->constrained[i] < d->n_dofs, "constrained index out of range");
    op->E = d->n_cells;
    op->N = d->n_dofs;
    op->n_cons = d->n_constrained;
    op->map.upload(d->map, nmap);
    op->cellX.upload(d->cell_vertices, (uint64_t)d->n_cells * 24);
    op->cons.upload(d->constrained, d->n_constrained);
    finish_setup(op.get(), d->with_fp32, nullptr);
    *out = op.release();
  });
}

mfx_status mfx_op_destroy(mfx_op *op) {
  return guard([&] {
    if (op) {
      (void)hipDeviceSynchronize();
      delete op;
    }
  });
}

uint64_t mfx_op_n_dofs(const mfx_op *op) { return op ? op->N : 0; }
uint32_t mfx_op_n_cells(const mfx_op *op) { return op ? op->E : 0; }

mfx_status mfx_op_get_info(const mfx_op *op, mfx_op_info *info) {
  return guard([&] {
    MFX_REQUIRE(op && info, "null pointer argument");
    info->degree = op->p;
    info->n_q_points_1d = op->nq;
    info->quadrature = op->quad;
    info->geometry = op->geometry;
    info->n_cells = op->E;
    info->n_dofs = op->N;
    info->n_constrained = op->n_cons;
    info->map = op->map.ptr;
    info->G = op->G.ptr;
    info->B = op->B.ptr;
    info->D = op->D.ptr;
    info->cell_vertices = op->cellX.ptr;
    info->constrained = op->cons.ptr;
    info->xq = op->xq.ptr;
    info->wq = op->wq.ptr;
    info->n_patches = op->patch ? patch_count(op->lay) : 0;
    info->patch_cells[0] = op->patch ? op->shape.bx : 0;
    info->patch_cells[1] = op->patch ? op->shape.by : 0;
    info->patch_cells[2] = op->patch ? op->shape.bz : 0;
    info->patch_map = op->pmap.ptr;
  });
}

mfx_status mfx_op_set_deterministic(mfx_op *op, int on) {
  return guard([&] {
    MFX_REQUIRE(op, "null pointer argument");
    MFX_REQUIRE(!on || op->patch,
                "deterministic scatter is implemented for the patch kernel (box operators with "
                "GLL collocation)");
    MFX_HIP(hipDeviceSynchronize());  // no apply of this operator in flight
    if (on) {
      const PatchShape sh = op->shape;
      const uint64_t n = (uint64_t)patch_count(op->lay) * (sh.bx * op->p + 1) *
                         (sh.by * op->p + 1) * (sh.bz * op->p + 1);
      if (op->pbuf64.n != n) op->pbuf64.alloc(n);
      if (op->D32.ptr && op->pbuf32.n != n) op->pbuf32.alloc(n);
    } else {
      op->pbuf64.reset();
      op->pbuf32.reset();
    }
    op->det = on != 0;
  });
}

mfx_status mfx_op_vmult(mfx_op *op, double *dst, const double *src, void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && dst && src, "null pointer argument");
    MFX_REQUIRE(dst != src, "vmult: src and dst must not alias");
    hipStream_t s = as_stream(stream);
    launch_fill(dst, 0.0, op->N, s);
    op->apply64(src, dst, s);
    launch_copy_listed<double>(dst, src, op->cons.ptr, op->n_cons, s);
  });
}

mfx_status mfx_op_vmult_add(mfx_op *op, double *dst, const double *src, void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && dst && src, "null pointer argument");
    MFX_REQUIRE(dst != src, "vmult_add: src and dst must not alias");
    op->apply64(src, dst, as_stream(stream));
  });
}

mfx_status mfx_op_vmult_f32(mfx_op *op, float *dst, const float *src, void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && dst && src, "null pointer argument");
    MFX_REQUIRE(dst != src, "vmult: src and dst must not alias");
    MFX_REQUIRE(op->D32.ptr, "operator created without with_fp32");
    hipStream_t s = as_stream(stream);
    launch_fill_f32(dst, 0.0f, op->N, s);
    op->apply32(src, dst, s);
    launch_copy_listed<float>(dst, src, op->cons.ptr, op->n_cons, s);
  });
}

mfx_status mfx_op_diagonal(mfx_op *op, double *diag, void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && diag, "null pointer argument");
    hipStream_t s = as_stream(stream);
    GeometryView gv;
    lexicographic_G(op, gv, s);
    launch_fill(diag, 0.0, op->N, s);
    launch_diagonal(op->E, op->p, op->nq, op->B.ptr, op->D.ptr, gv.G, op->map.ptr, diag, s);
    launch_set_listed(diag, 1.0, op->cons.ptr, op->n_cons, s);
    if (gv.tmp.ptr) MFX_HIP(hipStreamSynchronize(s));  // temporary geometry freed on return
  });
}

mfx_status mfx_op_inverse_diagonal(mfx_op *op, double *inv, void *stream) {
  mfx_status st = mfx_op_diagonal(op, inv, stream);
  if (st != MFX_OK) return st;
  return guard([&] { launch_invert_diagonal(inv, op->N, as_stream(stream)); });
}

mfx_status mfx_op_rhs(mfx_op *op, int f_kind, double *b, void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && b, "null pointer argument");
    MFX_REQUIRE(f_kind == 0 || f_kind == 1, "f_kind must be 0 or 1");
    hipStream_t s = as_stream(stream);
    launch_fill(b, 0.0, op->N, s);
    launch_rhs(op->E, op->p, op->nq, op->B.ptr, op->xq.ptr, op->wq.ptr, op->cellX.ptr,
               op->map.ptr, f_kind, b, s);
  });
}

mfx_status mfx_op_dof_coordinates(mfx_op *op, double *coords, void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && coords, "null pointer argument");
    launch_dof_coordinates(op->E, op->p, op->support.ptr, op->cellX.ptr, op->map.ptr, coords,
                           as_stream(stream));
  });
}

// ---------------------------------------------------------------- CG
mfx_status mfx_cg_solve(mfx_op *op, double *x, const double *b, const double *inv_diag,
                        int max_steps, double abs_tol, double rel_tol, mfx_cg_result *res,
                        void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && x && b && res, "null pointer argument");
    MFX_REQUIRE(max_steps >= 0, "max_steps must be >= 0");
    const uint64_t N = op->N;
    const size_t npart = cg_partial_count(*op);
    if (op->r.n != N) {
      op->r.alloc(N);
      op->pdir.alloc(N);
      op->ap.alloc(N);
    }
    if (op->part.n < npart) op->part.alloc(npart);
    if (!op->st.ptr) op->st.alloc(1);
    if (!op->h_done) MFX_HIP(hipHostMalloc(reinterpret_cast<void **>(&op->h_done), sizeof(int)));
    const CgWork w{op->r.ptr, op->pdir.ptr, op->ap.ptr, op->part.ptr, op->st.ptr, op->h_done};
    cg_run(*op, x, b, inv_diag, max_steps, abs_tol, rel_tol, w, res, as_stream(stream));
  });
}

// ---------------------------------------------------------------- vectors
mfx_status mfx_vec_random(double *x, uint64_t n, uint64_t seed, void *stream) {
  return guard([&] {
    MFX_REQUIRE(x || n == 0, "null pointer argument");
    launch_random(x, n, seed, as_stream(stream));
  });
}

mfx_status mfx_vec_sine_modes(const double *coords, uint64_t n, int n_modes,
                              const int32_t *wave_numbers, const double *amplitudes, double *out,
                              void *stream) {
  return guard([&] {
    MFX_REQUIRE((coords && out) || n == 0, "null pointer argument");
    MFX_REQUIRE(n_modes >= 0 && n_modes <= SineModes::kMax, "n_modes must be in 0..64");
    MFX_REQUIRE(n_modes == 0 || (wave_numbers && amplitudes), "null mode arrays");
    SineModes m{};
    m.n = n_modes;
    for (int i = 0; i < n_modes; ++i) {
      for (int d = 0; d < 3; ++d) m.k[3 * i + d] = wave_numbers[3 * i + d];
      m.a[i] = amplitudes[i];
    }
    launch_sine_modes(coords, n, m, out, as_stream(stream));
  });
}

mfx_status mfx_vec_zero_constrained(mfx_op *op, double *x, void *stream) {
  return guard([&] {
    MFX_REQUIRE(op && x, "null pointer argument");
    launch_set_listed(x, 0.0, op->cons.ptr, op->n_cons, as_stream(stream));
  });
}

mfx_status mfx_vec_dot(const double *x, const double *y, uint64_t n, double *host_result,
                       void *stream) {
  return guard([&] {
    MFX_REQUIRE(x && y && host_result, "null pointer argument");
    hipStream_t s = as_stream(stream);
    DBuf<double> part;
    part.alloc(dot_n_blocks(n) + 1);
    launch_dot(x, y, n, part.ptr, part.ptr + dot_n_blocks(n), s);
    MFX_HIP(hipMemcpyAsync(host_result, part.ptr + dot_n_blocks(n), sizeof(double),
                           hipMemcpyDeviceToHost, s));
    MFX_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
