// batch.cpp -- many independent BP5 problems solved together
// (mfx_batch_* in include/mfx.h).
//
// Instances of uneven size and degree are concatenated into one DoF vector.
// Per degree, the DoF maps (offset into the concatenated vector), the metric
// terms and a cell -> instance table of all instances of that degree are
// concatenated, so one launch of the collocation kernel applies the operator
// of every instance of that degree.  Each instance keeps its own Jacobi-CG
// scalars and convergence state (batch_kernels.hip); cells and segments of
// converged instances are skipped.  A solve runs as a hipGraph of K CG steps
// replayed until every instance has converged (or as the same launches
// issued directly).  Per degree there is a second form of the apply: the
// persistent patch kernel with the metric terms computed from the cell
// vertices (patch.hip), over the concatenated patches of all instances of
// that degree.  At creation an autotuner may time the tile variants of the
// collocation kernel (cells per workgroup, G prefetch) and the patch kernel
// per degree on the batch's own cells and keep the fastest.
#include <algorithm>
#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"
#include "op.hpp"
#include "patch.hpp"

namespace mfx {
namespace {

constexpr uint64_t kSegment = 8192;  // DoFs per vector-kernel workgroup
constexpr int kGraphSteps = 8;       // CG steps per graph launch

struct OpDeleter {
  void operator()(mfx_op *o) const { (void)mfx_op_destroy(o); }
};
using OpPtr = std::unique_ptr<mfx_op, OpDeleter>;

}  // namespace
}  // namespace mfx

using namespace mfx;

struct mfx_batch {
  struct Group {
    int p = 0;
    uint32_t E = 0;
    DBuf<uint32_t> map;
    DBuf<double> G, D;
    DBuf<uint16_t> cell_inst;
    // the on-the-fly form: concatenated patch maps, patch -> (instance,
    // patch coordinates), instance -> (cells per direction, first cell),
    // concatenated cell vertices
    PatchBatchOp patches;
    DBuf<double> ppart;  // p.Ap partials of the patch kernel, [slot][workgroup][wave]
    DBuf<uint32_t> slot;  // per instance of the batch: its slot (instances of this degree)
    std::vector<uint32_t> slot_h;
    uint32_t waves = 0;
    DBuf<uint32_t> pmap;
    DBuf<uint4> meta, inst;
    DBuf<double> cellX, w, xq;
    int variant = 0, default_variant = 0;
    std::vector<double> ms;  // per variant, when tuned
    int n_variants() const { return colloc_variant_count(p) + 1; }
    bool on_the_fly(int v) const { return v == colloc_variant_count(p); }
  };
  std::vector<OpPtr> ops;
  std::vector<uint64_t> off;  // n + 1
  uint64_t N = 0;
  std::vector<Group> groups;
  DBuf<uint32_t> cons;
  uint64_t n_cons = 0;
  DBuf<Seg> segs;
  uint32_t n_segs = 0;
  DBuf<uint32_t> inst_seg;
  DBuf<uint8_t> cmask;  // 1 on constrained rows
  // fused p.Ap: per instance its range of the patch kernel's partials
  DBuf<PapRange> pap_ranges;
  DBuf<double> part_c;  // constrained-row share of p.Ap per segment
  // CG scratch
  DBuf<BatchCg> st;
  DBuf<double> r, p, ap, t, part, part2, ones;
  DBuf<int> flags, n_active;
  int *h_active = nullptr;
  hipStream_t cs = nullptr;

  ~mfx_batch() {
    if (h_active) (void)hipHostFree(h_active);
    if (cs) (void)hipStreamDestroy(cs);
  }

  static void apply_group(const Group &g, int v, const double *src, double *dst, const int *done,
                          hipStream_t s) {
    if (g.on_the_fly(v))
      launch_patch_batch(g.patches, src, dst, done, s);
    else
      launch_colloc_variant(g.p, v, g.D.ptr, g.G.ptr, src, dst, g.map.ptr, g.E,
                            done ? g.cell_inst.ptr : nullptr, done, s);
  }
  void apply(const double *src, double *dst, const int *done, hipStream_t s) {
    for (auto &g : groups) apply_group(g, g.variant, src, dst, done, s);
  }
  void vmult(double *dst, const double *src, hipStream_t s) {
    launch_fill(dst, 0.0, N, s);
    apply(src, dst, nullptr, s);
    launch_copy_listed<double>(dst, src, cons.ptr, n_cons, s);
  }
  // p.Ap comes from the patch kernel when every degree uses it
  bool fused_pap() const {
    for (auto &g : groups)
      if (!g.on_the_fly(g.variant)) return false;
    return true;
  }
  void cg_step(double *x, const double *dinv, hipStream_t s) {
    // ap leaves the direction pass with the identity rows set
    launch_b_direction(segs.ptr, n_segs, st.ptr, r.ptr, dinv, cmask.ptr, p.ptr, ap.ptr, x,
                       part_c.ptr, s);
    if (fused_pap()) {
      for (auto &g : groups) launch_patch_batch(g.patches, p.ptr, ap.ptr, flags.ptr, s, g.ppart.ptr);
      launch_b_alpha_fused(st.ptr, pap_ranges.ptr, inst_seg.ptr, (uint32_t)ops.size(), part_c.ptr,
                           s);
    } else {
      apply(p.ptr, ap.ptr, flags.ptr, s);
      launch_b_pap(segs.ptr, n_segs, st.ptr, p.ptr, ap.ptr, part.ptr, s);
      launch_b_alpha(st.ptr, inst_seg.ptr, (uint32_t)ops.size(), part.ptr, s);
    }
    launch_b_update(segs.ptr, n_segs, st.ptr, r.ptr, ap.ptr, dinv, part.ptr, part2.ptr, s);
    launch_b_check(st.ptr, inst_seg.ptr, (uint32_t)ops.size(), part.ptr, part2.ptr, flags.ptr,
                   n_active.ptr, s);
  }
};

namespace {

void autotune(mfx_batch &b, hipStream_t s) {
  DBuf<double> src, dst;
  src.alloc(b.N);
  dst.alloc(b.N);
  launch_random(src.ptr, b.N, 7, s);
  launch_fill(dst.ptr, 0.0, b.N, s);
  struct Ev {  // destroys what was created, also when the second create throws
    hipEvent_t a = nullptr, b = nullptr;
    ~Ev() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  MFX_HIP(hipEventCreate(&ev.a));
  MFX_HIP(hipEventCreate(&ev.b));
  const hipEvent_t e0 = ev.a, e1 = ev.b;
  for (auto &g : b.groups) {
    const int nv = g.n_variants();
    g.ms.assign(nv, 0.0);
    double best = 1e30;
    for (int v = 0; v < nv; ++v) {
      for (int w = 0; w < 2; ++w) mfx_batch::apply_group(g, v, src.ptr, dst.ptr, nullptr, s);
      MFX_HIP(hipEventRecord(e0, s));
      constexpr int reps = 5;
      for (int w = 0; w < reps; ++w) mfx_batch::apply_group(g, v, src.ptr, dst.ptr, nullptr, s);
      MFX_HIP(hipEventRecord(e1, s));
      MFX_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      MFX_HIP(hipEventElapsedTime(&ms, e0, e1));
      g.ms[v] = ms / reps;
      if (g.ms[v] < best) {
        best = g.ms[v];
        g.variant = v;
      }
    }
  }
}

// The on-the-fly form of a degree group: every instance's patch map (the
// same kernel shape for all), offset to its DoF range, concatenated in
// instance order, with the matching per-patch and per-instance tables.
void build_patches(mfx_batch &b, mfx_batch::Group &g, int n, hipStream_t s) {
  const int p = g.p;
  const PatchShape sh = patch_shape(p);
  const uint64_t P3 = (uint64_t)(sh.bx * p + 1) * (sh.by * p + 1) * (sh.bz * p + 1);
  uint64_t np_total = 0;
  std::vector<uint4> meta, inst(n, uint4{0, 0, 0, 0});
  g.slot_h.assign(n, 0);
  uint32_t n_local = 0;
  uint64_t cells = 0;
  for (int i = 0; i < n; ++i) {
    const mfx_op &op = *b.ops[i];
    if (op.p != p) continue;
    const uint32_t n0 = op.box.n[0], n1 = op.box.n[1], n2 = op.box.n[2];
    const uint32_t q0 = (n0 + sh.bx - 1) / sh.bx, q1 = (n1 + sh.by - 1) / sh.by,
                   q2 = (n2 + sh.bz - 1) / sh.bz;
    inst[i] = uint4{n0, n1, n2, (uint32_t)cells};
    g.slot_h[i] = n_local++;
    for (uint32_t z = 0; z < q2; ++z)
      for (uint32_t y = 0; y < q1; ++y)
        for (uint32_t x = 0; x < q0; ++x) meta.push_back(uint4{(uint32_t)i, x, y, z});
    np_total += (uint64_t)q0 * q1 * q2;
    cells += op.E;
  }
  MFX_REQUIRE(np_total < 0xFFFFFFFFull && np_total * P3 < (1ull << 40) && cells < 0xFFFFFFFFull,
              "batch too large for the patch tables");
  g.pmap.alloc(np_total * P3);
  g.cellX.alloc(cells * 24);
  DBuf<uint32_t> tmp;
  uint64_t pt0 = 0;
  for (int i = 0; i < n; ++i) {
    const mfx_op &op = *b.ops[i];
    if (op.p != p) continue;
    PatchLayout lay{};
    const int bs[3] = {sh.bx, sh.by, sh.bz};
    for (int k = 0; k < 3; ++k) {
      lay.n[k] = op.box.n[k];
      lay.np[k] = (op.box.n[k] + bs[k] - 1) / bs[k];
    }
    const uint64_t npi = patch_count(lay);
    tmp.alloc(npi * P3);
    launch_patch_map(lay, p, sh, tmp.ptr, s);
    launch_offset_map(tmp.ptr, g.pmap.ptr + pt0 * P3, npi * P3, (uint32_t)b.off[i], s);
    MFX_HIP(hipMemcpyAsync(g.cellX.ptr + (uint64_t)inst[i].w * 24, op.cellX.ptr,
                           (uint64_t)op.E * 24 * sizeof(double), hipMemcpyDeviceToDevice, s));
    MFX_HIP(hipStreamSynchronize(s));  // tmp is reused
    pt0 += npi;
    if (!g.w.ptr) {
      g.w.alloc(op.wq.n);
      g.xq.alloc(op.xq.n);
      MFX_HIP(hipMemcpyAsync(g.w.ptr, op.wq.ptr, op.wq.n * 8, hipMemcpyDeviceToDevice, s));
      MFX_HIP(hipMemcpyAsync(g.xq.ptr, op.xq.ptr, op.xq.n * 8, hipMemcpyDeviceToDevice, s));
    }
  }
  g.meta.upload(meta.data(), meta.size());
  g.inst.upload(inst.data(), inst.size());
  g.patches.p = p;
  g.patches.n_patches = (uint32_t)np_total;
  g.patches.grid = patch_batch_grid(p, (uint32_t)np_total, false);
  g.patches.grid_dot = patch_batch_grid(p, (uint32_t)np_total, true);
  g.slot.upload(g.slot_h.data(), g.slot_h.size());
  g.patches.slot = g.slot.ptr;
  g.waves = patch_batch_waves(p);
  g.ppart.alloc((uint64_t)n_local * g.patches.grid_dot * g.waves);
  launch_fill(g.ppart.ptr, 0.0, g.ppart.n, s);
  g.patches.pmap = g.pmap.ptr;
  g.patches.meta = g.meta.ptr;
  g.patches.inst = g.inst.ptr;
  g.patches.cellX = g.cellX.ptr;
  g.patches.w = g.w.ptr;
  g.patches.xq = g.xq.ptr;
}

}  // namespace

extern "C" {

mfx_status mfx_batch_create(const mfx_batch_instance *inst, int n, int tune, mfx_batch **out) {
  return guard([&] {
    MFX_REQUIRE(inst && out, "null pointer argument");
    MFX_REQUIRE(n >= 1 && n <= 65535, "number of instances must be in 1..65535");
    *out = nullptr;
    auto b = std::make_unique<mfx_batch>();
    MFX_HIP(hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking));
    hipStream_t s = b->cs;
    b->off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
      const auto &in = inst[i];
      MFX_REQUIRE(colloc_variant_count(in.degree) > 0, "batched degrees must be in 2..6");
      mfx_box_desc d{};
      d.degree = in.degree;
      d.n_q_points_1d = in.degree + 1;
      d.quadrature = MFX_QUAD_GLL;
      d.geometry = MFX_GEOM_STORED;
      for (int k = 0; k < 3; ++k) {
        d.n_cells[k] = in.n_cells[k];
        d.lo[k] = in.lo[k];
        d.hi[k] = in.hi[k];
      }
      d.deform = in.deform;
      mfx_op *op = nullptr;
      const mfx_status st = mfx_op_create_box(&d, &op);
      if (st != MFX_OK)
        throw Error{st, "instance " + std::to_string(i) + ": " + mfx_last_error()};
      b->ops.emplace_back(op);
      b->off[i + 1] = b->off[i] + op->N;
    }
    b->N = b->off[n];
    MFX_REQUIRE(b->N < 0x7FFFFFFFull, "too many DoFs in the batch for 31-bit DoF indices");
    // per-degree concatenation of maps, metric terms and cell -> instance ids
    for (int p = 2; p <= 6; ++p) {
      uint64_t E = 0;
      for (auto &op : b->ops)
        if (op->p == p) E += op->E;
      if (E == 0) continue;
      MFX_REQUIRE(E <= 0xFFFFFFFFull, "too many cells of one degree");
      mfx_batch::Group g;
      g.p = p;
      g.E = (uint32_t)E;
      const uint64_t nm3 = (uint64_t)(p + 1) * (p + 1) * (p + 1);
      g.map.alloc(E * nm3);
      g.G.alloc(E * 6 * nm3);
      g.cell_inst.alloc(E);
      uint64_t e0 = 0;
      for (int i = 0; i < n; ++i) {
        mfx_op &op = *b->ops[i];
        if (op.p != p) continue;
        launch_offset_map(op.map.ptr, g.map.ptr + e0 * nm3, (uint64_t)op.E * nm3,
                          (uint32_t)b->off[i], s);
        MFX_HIP(hipMemcpyAsync(g.G.ptr + e0 * 6 * nm3, op.G.ptr, (uint64_t)op.E * 6 * nm3 * 8,
                               hipMemcpyDeviceToDevice, s));
        launch_fill_u16(g.cell_inst.ptr + e0, (uint16_t)i, op.E, s);
        if (!g.D.ptr) {
          g.D.alloc(op.D.n);
          MFX_HIP(hipMemcpyAsync(g.D.ptr, op.D.ptr, op.D.n * 8, hipMemcpyDeviceToDevice, s));
        }
        e0 += op.E;
      }
      build_patches(*b, g, n, s);
      g.default_variant = colloc_default_variant(p);
      g.variant = g.default_variant;
      b->groups.push_back(std::move(g));
    }
    // constrained rows of every instance, offset
    for (auto &op : b->ops) b->n_cons += op->n_cons;
    b->cons.alloc(b->n_cons);
    uint64_t c0 = 0;
    for (int i = 0; i < n; ++i) {
      launch_offset_map(b->ops[i]->cons.ptr, b->cons.ptr + c0, b->ops[i]->n_cons, (uint32_t)b->off[i],
                        s);
      c0 += b->ops[i]->n_cons;
    }
    // fused p.Ap: each instance's range of patch partials in its group, and
    // segments of the constrained list (per instance, at most kSegment rows)
    {
      std::vector<PapRange> rg(n);
      for (int i = 0; i < n; ++i)
        for (auto &g : b->groups)
          if (g.p == b->ops[i]->p) {
            const uint64_t slot_len = (uint64_t)g.patches.grid_dot * g.waves;
            rg[i] = PapRange{g.ppart.ptr + (uint64_t)g.slot_h[i] * slot_len, slot_len};
          }
      b->pap_ranges.upload(rg.data(), rg.size());
    }
    b->cmask.alloc(b->N);
    MFX_HIP(hipMemsetAsync(b->cmask.ptr, 0, b->N, s));
    launch_mark_u8(b->cons.ptr, b->n_cons, b->cmask.ptr, s);
    // segments: contiguous DoF ranges of one instance
    std::vector<Seg> segs;
    std::vector<uint32_t> inst_seg(n + 1, 0);
    for (int i = 0; i < n; ++i) {
      inst_seg[i] = (uint32_t)segs.size();
      for (uint64_t a = b->off[i]; a < b->off[i + 1]; a += kSegment)
        segs.push_back({a, std::min(a + kSegment, b->off[i + 1]), (uint32_t)i, 0});
    }
    inst_seg[n] = (uint32_t)segs.size();
    b->n_segs = (uint32_t)segs.size();
    b->segs.upload(segs.data(), segs.size());
    b->part_c.alloc(b->n_segs);
    b->inst_seg.upload(inst_seg.data(), inst_seg.size());
    MFX_HIP(hipStreamSynchronize(s));
    if (tune) autotune(*b, s);
    MFX_HIP(hipStreamSynchronize(s));
    *out = b.release();
  });
}

mfx_status mfx_batch_destroy(mfx_batch *b) {
  return guard([&] {
    if (b) {
      (void)hipDeviceSynchronize();
      delete b;
    }
  });
}

int mfx_batch_n_instances(const mfx_batch *b) { return b ? (int)b->ops.size() : 0; }
uint64_t mfx_batch_n_dofs(const mfx_batch *b) { return b ? b->N : 0; }

mfx_status mfx_batch_offsets(const mfx_batch *b, uint64_t *offsets) {
  return guard([&] {
    MFX_REQUIRE(b && offsets, "null pointer argument");
    std::copy(b->off.begin(), b->off.end(), offsets);
  });
}

mfx_op *mfx_batch_instance_operator(mfx_batch *b, int i) {
  if (!b || i < 0 || i >= (int)b->ops.size()) return nullptr;
  return b->ops[i].get();
}

mfx_status mfx_batch_get_tuning(const mfx_batch *b, int degree, mfx_batch_tuning *t) {
  return guard([&] {
    MFX_REQUIRE(b && t, "null pointer argument");
    *t = mfx_batch_tuning{};
    t->degree = degree;
    for (auto &g : b->groups) {
      if (g.p != degree) continue;
      t->n_cells = g.E;
      t->n_variants = g.n_variants();
      t->variant = g.variant;
      t->default_variant = g.default_variant;
      for (int v = 0; v < t->n_variants && v < 8; ++v) {
        if (g.on_the_fly(v)) {
          const PatchShape sh = patch_shape(degree);
          t->cells_per_workgroup[v] = sh.bx * sh.by * sh.bz;
          t->prefetch[v] = 1;
          t->geometry[v] = MFX_GEOM_ON_THE_FLY;
        } else {
          launch_colloc_variant(degree, v, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                nullptr, nullptr, nullptr, &t->cells_per_workgroup[v],
                                &t->prefetch[v]);
          t->geometry[v] = MFX_GEOM_STORED;
        }
        t->ms[v] = v < (int)g.ms.size() ? g.ms[v] : 0.0;
      }
      return;
    }
    throw Error{MFX_ERR_INVALID_ARG, "no instance of degree " + std::to_string(degree)};
  });
}

mfx_status mfx_batch_set_variant(mfx_batch *b, int degree, int variant) {
  return guard([&] {
    MFX_REQUIRE(b, "null pointer argument");
    for (auto &g : b->groups) {
      if (g.p != degree) continue;
      MFX_REQUIRE(variant >= 0 && variant < g.n_variants(), "variant out of range");
      g.variant = variant;
      return;
    }
    throw Error{MFX_ERR_INVALID_ARG, "no instance of degree " + std::to_string(degree)};
  });
}

mfx_status mfx_batch_vmult(mfx_batch *b, double *dst, const double *src, void *stream) {
  return guard([&] {
    MFX_REQUIRE(b && dst && src, "null pointer argument");
    MFX_REQUIRE(dst != src, "src and dst must not alias");
    b->vmult(dst, src, as_stream(stream));
  });
}

mfx_status mfx_batch_inverse_diagonal(mfx_batch *b, double *inv, void *stream) {
  return guard([&] {
    MFX_REQUIRE(b && inv, "null pointer argument");
    for (size_t i = 0; i < b->ops.size(); ++i) {
      const mfx_status st = mfx_op_inverse_diagonal(b->ops[i].get(), inv + b->off[i], stream);
      if (st != MFX_OK) throw Error{st, mfx_last_error()};
    }
  });
}

mfx_status mfx_batch_cg_solve(mfx_batch *b, double *x, const double *rhs, const double *inv_diag,
                              int max_steps, double abs_tol, double rel_tol, int use_graph,
                              mfx_cg_result *results, void *stream) {
  return guard([&] {
    MFX_REQUIRE(b && x && rhs && results, "null pointer argument");
    MFX_REQUIRE(max_steps >= 0, "max_steps must be >= 0");
    const uint32_t n = (uint32_t)b->ops.size();
    // the solve runs on the batch's own stream (graph capture needs one);
    // the caller's stream is drained first so that x and rhs are ready
    MFX_HIP(hipStreamSynchronize(as_stream(stream)));
    hipStream_t s = b->cs;
    if (b->r.n != b->N) {
      b->r.alloc(b->N);
      b->p.alloc(b->N);
      b->ap.alloc(b->N);
      b->t.alloc(b->N);
    }
    if (b->part.n < b->n_segs) {
      b->part.alloc(b->n_segs);
      b->part2.alloc(b->n_segs);
    }
    if (!b->st.ptr) {
      b->st.alloc(n);
      b->flags.alloc(n);
      b->n_active.alloc(1);
    }
    if (!b->h_active) MFX_HIP(hipHostMalloc(reinterpret_cast<void **>(&b->h_active), sizeof(int)));
    const double *dinv = inv_diag;
    if (!dinv) {
      if (b->ones.n != b->N) {
        b->ones.alloc(b->N);
        launch_fill(b->ones.ptr, 1.0, b->N, s);
      }
      dinv = b->ones.ptr;
    }
    const auto t0 = std::chrono::steady_clock::now();
    // the fused p.Ap keeps its partial slots zero between steps; start clean
    for (auto &g : b->groups) launch_fill(g.ppart.ptr, 0.0, g.ppart.n, s);
    b->vmult(b->t.ptr, x, s);
    launch_b_residual(b->segs.ptr, b->n_segs, b->r.ptr, rhs, b->t.ptr, dinv, b->part.ptr,
                      b->part2.ptr, s);
    launch_b_init(b->st.ptr, b->inst_seg.ptr, n, b->part.ptr, b->part2.ptr, abs_tol, rel_tol,
                  max_steps, b->flags.ptr, b->n_active.ptr, s);
    MFX_HIP(hipMemcpyAsync(b->h_active, b->n_active.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    MFX_HIP(hipStreamSynchronize(s));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    struct GraphGuard {
      hipGraph_t &g;
      hipGraphExec_t &e;
      ~GraphGuard() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
      }
    } gg{graph, exec};
    if (use_graph && *b->h_active > 0) {
      MFX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < kGraphSteps; ++k) b->cg_step(x, dinv, s);
      MFX_HIP(hipStreamEndCapture(s, &graph));
      MFX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    const int max_rounds = max_steps / kGraphSteps + 2;
    for (int round = 0; *b->h_active > 0 && round < max_rounds; ++round) {
      if (exec) {
        MFX_HIP(hipGraphLaunch(exec, s));
      } else {
        for (int k = 0; k < kGraphSteps; ++k) b->cg_step(x, dinv, s);
      }
      MFX_HIP(hipMemcpyAsync(b->h_active, b->n_active.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
      MFX_HIP(hipStreamSynchronize(s));
    }
    launch_b_finish(b->segs.ptr, b->n_segs, b->st.ptr, x, b->p.ptr, s);
    std::vector<BatchCg> fin(n);
    MFX_HIP(hipMemcpyAsync(fin.data(), b->st.ptr, n * sizeof(BatchCg), hipMemcpyDeviceToHost, s));
    MFX_HIP(hipStreamSynchronize(s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int failed = 0;
    for (uint32_t i = 0; i < n; ++i) {
      results[i].steps = fin[i].steps;
      results[i].converged = fin[i].converged;
      results[i].initial_value = fin[i].initial;
      results[i].last_value = fin[i].last;
      results[i].seconds = secs;
      failed += fin[i].converged ? 0 : 1;
    }
    if (failed)
      throw Error{MFX_ERR_NOT_CONVERGED,
                  std::to_string(failed) + " of " + std::to_string(n) + " instances did not converge"};
  });
}

}  // extern "C"
