"""BP5 matrix-free operator benchmark on one MI355X.

Workload (one "step"): one operator apply (vmult, deal.II semantics: dst = 0,
gather/apply/scatter, constrained rows copied) for each of p = 4, 6, 8 of the
BP5 plain method: scalar Laplacian, Q_p with GLL collocation (q = p+1), unit
cube deformed by x += 0.05 sin(2 pi x) sin(2 pi y) sin(2 pi z), n^3 cells with
n = 79 / 53 / 40 (about 32M DoFs each), FP64, input uniform[-1,1] from
SplitMix64 seed 1.  Value = aggregate GDoF/s = sum_p N_p / sum_p t_p.

Each degree is measured with the metric terms stored (G, 48 q^3 B per cell,
per-cell kernel) and computed on the fly from the cell vertices (patch
kernel); the faster mode per degree enters the aggregate.  Also reported
per degree: the apply kernel's time (HIP events on the launch stream), its
algorithmic bytes and FLOPs, and Jacobi-CG time-to-1e-8 relative residual.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
FP64_PEAK = 78.6e12
CELLS = {2: 158, 4: 79, 6: 53, 8: 40}
STORED, ON_THE_FLY = 0, 1


def algorithmic_bytes(p, N, E, geometry, patch_cells=None, n_patches=0):
    """Bytes an apply must move at least once: src read + dst write (8 B per
    DoF each) + the DoF map + the geometry.  Stored: uint32 map (4 (p+1)^3
    per cell) and G (6 doubles per quadrature point).  On the fly: the patch
    DoF map (4 B per patch DoF, patches of patch_cells cells) and 8 vertices
    per cell (192 B)."""
    q = p + 1
    if geometry == STORED:
        return 16 * N + 4 * (p + 1) ** 3 * E + 48 * q ** 3 * E
    bx, by, bz = patch_cells
    return 16 * N + 4 * (bx * p + 1) * (by * p + 1) * (bz * p + 1) * n_patches + 192 * E


def flops_per_cell(p, geometry=STORED):
    """FP64 operations of one cell's apply (an FMA counts 2).  Stored metric:
    the sum-factorized BK5 count 12 q^4 + 15 q^3 (rooflineBK5.py:23-35:
    6 q^4 gradient + 6 q^4 transposed gradient + 15 per point for G grad u).
    On the fly (patch kernel): per point, 73 in place of the 15 (cofactor
    polynomials 24, det 5, reciprocal 9, weight 2, W 15, three projections
    18) and per (x,y) column 180 for the trilinear coefficients."""
    q = p + 1
    if geometry == STORED:
        return 12 * q ** 4 + 15 * q ** 3
    return 12 * q ** 4 + 73 * q ** 3 + 180 * q ** 2


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_mode(p, geometry, steps, warmup):
    import torch

    from benchmarks_amd.laplace import LaplaceOperator, random_vector
    n = CELLS[p]
    op = LaplaceOperator(p, p + 1, 1, n_cells=n, deform=0.05, geometry=geometry)
    N, E = op.m(), op.n_cells()
    x = random_vector(N, 1)
    y = op.initialize_dof_vector()
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        op.vmult(y, x)
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record(stream)
    for _ in range(steps):
        op.vmult(y, x)
    e1.record(stream)
    torch.cuda.synchronize()
    t_vmult = e0.elapsed_time(e1) * 1e-3 / steps
    # the element kernel alone (vmult_add: no zero-fill, no constrained copy)
    scratch = torch.zeros(N, dtype=torch.float64, device="cuda")
    for _ in range(3):
        op.vmult_add(scratch, x)
    torch.cuda.synchronize()
    k0, k1 = _events()
    k0.record(stream)
    for _ in range(steps):
        op.vmult_add(scratch, x)
    k1.record(stream)
    torch.cuda.synchronize()
    t_kernel = k0.elapsed_time(k1) * 1e-3 / steps
    # known-size probes for hardware-counter scaling: the deterministic dot
    # reads exactly 16 N bytes, vmult's zero-fill writes exactly 8 N bytes
    from benchmarks_amd._lib import check, lib
    dot = C.c_double()
    check(lib().mfx_vec_dot(C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), N,
                            C.byref(dot), C.c_void_p(stream.cuda_stream)))
    info = op.info
    b_alg = algorithmic_bytes(p, N, E, geometry, tuple(info.patch_cells), info.n_patches)
    f_alg = flops_per_cell(p, geometry) * E
    res = {"geometry": "stored" if geometry == STORED else "on_the_fly", "ms_vmult": t_vmult * 1e3,
           "gdofs_vmult": N / t_vmult * 1e-9, "ms_kernel": t_kernel * 1e3, "bytes_alg": b_alg,
           "flops_alg": f_alg, "kernel_GBs": b_alg / t_kernel * 1e-9,
           "kernel_TFLOPs": f_alg / t_kernel * 1e-12}
    return op, x, y, res


def bench_degree(p, steps, warmup, do_cg, modes):
    import torch

    from benchmarks_amd.laplace import (DiagonalMatrix, ReductionControl, SolverCG,
                                        random_vector)
    out = {"p": p, "n_cells_1d": CELLS[p], "modes": {}}
    best = None
    for g in modes:
        op, x, y, r = bench_mode(p, g, steps, warmup)
        out["n_dofs"], out["n_cells"] = op.m(), op.n_cells()
        out["modes"][r["geometry"]] = r
        if best is None or r["ms_vmult"] < best[1]["ms_vmult"]:
            best = (g, r)
        del op, x, y
        torch.cuda.empty_cache()
    g, r = best
    out.update({"geometry": r["geometry"], "ms_vmult": r["ms_vmult"],
                "gdofs_vmult": r["gdofs_vmult"], "ms_kernel": r["ms_kernel"],
                "bytes_alg": r["bytes_alg"], "flops_alg": r["flops_alg"]})
    if do_cg:
        from benchmarks_amd.laplace import LaplaceOperator
        op = LaplaceOperator(p, p + 1, 1, n_cells=CELLS[p], deform=0.05, geometry=g)
        N = op.m()
        b = random_vector(N, 1)
        op.zero_constrained(b)
        inv = op.compute_inverse_diagonal()
        sol = op.initialize_dof_vector()
        ctl = ReductionControl(100000, 1e-300, 1e-8)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        SolverCG(ctl).solve(op, sol, b, DiagonalMatrix(inv))
        torch.cuda.synchronize()
        out.update({"cg_jacobi_its": ctl.last_step(), "cg_jacobi_s": time.perf_counter() - t0,
                    "cg_final_rel_res": ctl.last_value() / ctl.initial_value()})
        del op, b, inv, sol
        torch.cuda.empty_cache()
    return out


def stream_copy_GBs(n=1 << 28, reps=10):
    """Measured HBM ceiling: device-to-device copy of 2 GiB (read + write bytes)."""
    import torch
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = _events()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    del a, b
    torch.cuda.empty_cache()
    return 16.0 * n / t * 1e-9


def _cpu_worker(args):
    """One host process: the reference kernel on its own cell slice for about
    `seconds`; returns (cells processed, elapsed).  A process per slice
    because the reference kernel returns before its delete[] (it leaks its
    three q^3 scratch arrays per cell); the memory goes with the process."""
    p, cells, seconds, use_ref = args
    import numpy as np

    from oracle import oracle as O
    nq = p + 1
    _, D, _, _ = O.basis_tables(p, nq, O.GLL)
    rng = np.random.default_rng(p)
    G = rng.standard_normal(cells * 6 * nq ** 3)
    u = rng.standard_normal(cells * nq ** 3)
    out = np.zeros(cells * nq ** 3)
    fn = O.ref_lib(fast=True).ref_bk5_serial if use_ref else None
    done, t0 = 0, time.perf_counter()
    while True:
        if fn is not None:
            fn(nq, cells, D, G, u, out)
        else:
            out[:] = O.bk5_evec(nq, cells, D, G, u)
        done += cells
        el = time.perf_counter() - t0
        if el >= seconds:
            return done, el


def cpu_baseline(degrees, budget_s=6.0):
    """The reference's own serial BK5 kernel
    (sum_factorization/include/kernels/BK5/serial_kernels.hpp:9-87, built by
    oracle/Makefile into oracle/_ref/) on T host processes, each on its own
    slice of cells (E-vector form: the reference CPU path has no gather or
    scatter); the per-cell time is extrapolated to each degree's full apply."""
    import multiprocessing as mp

    from oracle import oracle as O
    use_ref = True
    try:
        O.ref_lib(fast=True)
    except (FileNotFoundError, OSError):
        use_ref = False
    procs = max(1, min(16, os.cpu_count() or 1))
    per_deg = budget_s / len(degrees)
    t_tot, n_tot, samples = 0.0, 0, []
    ctx = mp.get_context("fork")
    for p in degrees:
        nq = p + 1
        n = CELLS[p]
        E, N = n ** 3, (n * p + 1) ** 3
        cells = max(32, int(512 * (5 / nq) ** 3))
        with ctx.Pool(procs) as pool:
            res = pool.map(_cpu_worker, [(p, cells, per_deg, use_ref)] * procs)
        done = sum(r[0] for r in res)
        wall = max(r[1] for r in res)
        t_tot += wall / done * E
        n_tot += N
        samples.append(f"p={p}: {done} cells in {wall:.1f}s")
    return {"value": n_tot / t_tot * 1e-9, "unit": "GDoF/s", "cores": procs,
            "kind": "reference" if use_ref else "port",
            "sample": f"{'reference' if use_ref else 'oracle'} BK5 serial kernel (E-vector, no "
                      f"gather/scatter), {procs} processes on disjoint cell slices, per-cell time "
                      "extrapolated to the full p=4,6,8 applies; " + "; ".join(samples)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--degrees", default="4,6,8")
    ap.add_argument("--geometry", choices=["both", "stored", "otf"], default="both")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the multi-GPU path (slab decomposition with halo "
              "exchange) is designed but not built; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    degrees = [int(d) for d in args.degrees.split(",")]
    modes = {"both": [STORED, ON_THE_FLY], "stored": [STORED], "otf": [ON_THE_FLY]}[args.geometry]
    # host baseline first, before anything initialises the GPU in this process
    cpu = None if args.no_cpu_baseline else cpu_baseline(degrees)
    import torch
    torch.cuda.synchronize()
    stream_copy = stream_copy_GBs()
    per = [bench_degree(p, args.steps, args.warmup, not args.no_cg, modes) for p in degrees]
    t_sum = sum(r["ms_vmult"] for r in per) * 1e-3
    n_sum = sum(r["n_dofs"] for r in per)
    dom = max(per, key=lambda r: r["ms_kernel"])
    t_k = dom["ms_kernel"] * 1e-3
    bw, flops = dom["bytes_alg"] / t_k, dom["flops_alg"] / t_k
    # the roof that bounds the kernel: FP64 compute when its arithmetic
    # intensity exceeds the machine balance (78.6 TFLOP/s / 8 TB/s), else HBM
    intensity = dom["flops_alg"] / dom["bytes_alg"]
    if intensity > FP64_PEAK / HBM_PEAK:
        roof = {"bound": "mfma", "achieved": flops * 1e-12, "peak": FP64_PEAK * 1e-12,
                "unit": "TFLOP/s", "frac": flops / FP64_PEAK}
    else:
        roof = {"bound": "hbm", "achieved": bw * 1e-9, "peak": HBM_PEAK * 1e-9,
                "unit": "GB/s", "frac": bw / HBM_PEAK}
    roof.update({"traffic": None, "kernel": f"BK5 apply p={dom['p']} ({dom['geometry']})",
                 "ms_per_launch": dom["ms_kernel"], "flops_per_launch": dom["flops_alg"],
                 "bytes_per_launch": dom["bytes_alg"], "flop_per_byte": intensity,
                 "hbm_GBs": bw * 1e-9, "hbm_frac": bw / HBM_PEAK,
                 "fp64_TFLOPs": flops * 1e-12, "fp64_frac": flops / FP64_PEAK,
                 "note": "FP64 arithmetic on the vector ALUs; MI355X's FP64 vector and "
                         "matrix peaks are both 78.6 TFLOP/s"})
    out = {
        "metric": "BP5 matvec GDoF/s at p=4-8 (aggregate over p=4,6,8); CG time-to-1e-8 per degree",
        "value": n_sum / t_sum * 1e-9,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_sum * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (deformed unit-cube mesh, SplitMix64 seed-1 input)",
        "config": {"workload": "BP5 GLL q=p+1, deformed unit cube, p=4/6/8 on 79^3/53^3/40^3 "
                               "cells (~32M DoFs each), FP64 vmult; metric terms stored or "
                               "on the fly, faster per degree",
                   "degrees": degrees, "geometry": {r["p"]: r["geometry"] for r in per},
                   "parallelism": "single GPU"},
        "roofline": roof,
        "per_degree": per,
        "stream_copy_GBs": stream_copy,
    }
    traffic = measured_traffic(dom)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


def measured_traffic(dom):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/), when one exists for
    this degree and geometry mode; None otherwise."""
    path = os.path.join(ROOT, "profiles", f"bk5_apply_bp5_p{dom['p']}_{dom['geometry']}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("APPLY_TRAFFIC_BYTES"):
            return float(line.split()[-1])
    return None


if __name__ == "__main__":
    main()
