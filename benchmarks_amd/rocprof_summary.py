"""Summaries of rocprofv3 CSV output as short markdown (for profiles/).

  python -m benchmarks_amd.rocprof_summary stats <kernel_stats.csv> [title]
      per-kernel calls / average / share from `rocprofv3 --kernel-trace --stats`
  python -m benchmarks_amd.rocprof_summary traffic <fetch counter_collection.csv>
      <write counter_collection.csv> <n_dofs> [title]
      HBM bytes per dispatch from two `rocprofv3 --pmc FETCH_SIZE` /
      `--pmc WRITE_SIZE` runs of `bench.py --degrees <p>`, each counter scaled
      by what it reads for a kernel of known traffic in the same run: the
      deterministic dot (k_dot_partials, reads 16 N bytes) for FETCH_SIZE and
      the zero-fill (k_fill, writes 8 N bytes) for WRITE_SIZE.
  python -m benchmarks_amd.rocprof_summary counters <kernel name part> <points>
      <pmc counter_collection.csv> [...]
      per-dispatch averages of the SQ counters of one kernel over one or more
      `rocprofv3 --pmc` runs, and the ratios they give (wait share, FP64
      instructions per quadrature point, LDS bank conflicts).
"""
import csv
import re
import sys
from collections import defaultdict


def short(name: str) -> str:
    name = name.replace("mfx::(anonymous namespace)::", "").replace("void ", "")
    name = re.sub(r"\(.*$", "", name)
    return name[:90]


def stats(path, title="kernel trace"):
    rows = list(csv.DictReader(open(path)))
    out = [f"### {title}", "", "| kernel | calls | avg (us) | total (ms) | share |",
           "|---|---:|---:|---:|---:|"]
    for r in rows[:14]:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.1f}% |")
    return "\n".join(out)


def _per_kernel(path, counter):
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name", "").startswith(counter):
            acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def traffic(fetch_csv, write_csv, n_dofs, title="HBM traffic"):
    n_dofs = int(n_dofs)
    fetch = _per_kernel(fetch_csv, "FETCH_SIZE")
    write = _per_kernel(write_csv, "WRITE_SIZE")
    f_probe = next(v for k, v in fetch.items() if "k_dot_partials" in k)
    w_probe = next(v for k, v in write.items() if "k_fill(" in k)
    f_scale = 16.0 * n_dofs / (f_probe * 1024.0)
    w_scale = 8.0 * n_dofs / (w_probe * 1024.0)
    out = [f"### {title}", "",
           f"FETCH_SIZE scale {f_scale:.4f} (k_dot_partials reads 16 N = {16 * n_dofs} B; "
           f"counter {f_probe:.0f} KiB); WRITE_SIZE scale {w_scale:.4f} (k_fill writes "
           f"8 N = {8 * n_dofs} B; counter {w_probe:.0f} KiB).", "",
           "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes / dispatch (scaled) |",
           "|---|---:|---:|---:|"]
    for k in sorted(set(fetch) | set(write), key=lambda k: -fetch.get(k, 0)):
        f, w = fetch.get(k, 0.0), write.get(k, 0.0)
        tot = f * 1024 * f_scale + w * 1024 * w_scale
        out.append(f"| `{short(k)}` | {f:.0f} | {w:.0f} | {tot:.4e} |")
        if "k_apply" in k or "k_patch<" in k:
            out.append("")
            out.append(f"APPLY_TRAFFIC_BYTES {short(k)} {tot:.6e}")
            out.append("")
    return "\n".join(out)


def counters(pmc_csvs, pattern, points=None, title="SQ counters"):
    """Per-dispatch averages of every counter of the kernels whose name
    contains `pattern`, over one or more `rocprofv3 --pmc` runs, with the
    ratios that follow from those present: share of wave-cycles spent
    waiting (SQ_WAIT_ANY / SQ_WAVE_CYCLES), FP64 VALU instructions per
    quadrature point (wave instructions x 64 lanes / points), executed FP64
    FLOPs (FMA = 2), LDS bank-conflict cycles per CU and the kernel's
    register and LDS footprint."""
    acc, meta = defaultdict(list), {}
    for path in pmc_csvs:
        for r in csv.DictReader(open(path)):
            if pattern in r["Kernel_Name"]:
                acc[r["Counter_Name"]].append(float(r["Counter_Value"]))
                meta = {k: r[k] for k in ("Kernel_Name", "VGPR_Count", "Accum_VGPR_Count",
                                          "SGPR_Count", "LDS_Block_Size", "Scratch_Size",
                                          "Workgroup_Size", "Grid_Size") if k in r}
    avg = {k: sum(v) / len(v) for k, v in acc.items()}
    out = [f"### {title}", "", f"Kernel: `{short(meta.get('Kernel_Name', pattern))}`", "",
           "| resource | value |", "|---|---:|"]
    for k in ("VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "LDS_Block_Size", "Scratch_Size",
              "Workgroup_Size", "Grid_Size"):
        if k in meta:
            out.append(f"| {k} | {meta[k]} |")
    out += ["", "| counter | dispatches | per dispatch |", "|---|---:|---:|"]
    for k in sorted(avg):
        out.append(f"| {k} | {len(acc[k])} | {avg[k]:.6g} |")
    out += ["", "| derived | value |", "|---|---:|"]
    if "SQ_WAIT_ANY" in avg and avg.get("SQ_WAVE_CYCLES"):
        out.append(f"| wave-cycles waiting (SQ_WAIT_ANY / SQ_WAVE_CYCLES) | "
                   f"{avg['SQ_WAIT_ANY'] / avg['SQ_WAVE_CYCLES']:.3f} |")
    f64 = [avg.get(f"SQ_INSTS_VALU_{op}_F64") for op in ("ADD", "MUL", "FMA")]
    if all(v is not None for v in f64):
        lanes = 64.0 * sum(f64)
        flops = 64.0 * (f64[0] + f64[1] + 2.0 * f64[2])
        out.append(f"| executed FP64 FLOPs per dispatch (FMA = 2) | {flops:.4e} |")
        if points:
            out.append(f"| FP64 lane instructions per quadrature point ({int(points)} points) | "
                       f"{lanes / float(points):.1f} |")
    if "SQ_LDS_BANK_CONFLICT" in avg:
        out.append(f"| LDS bank-conflict cycles per CU (256 CUs) | "
                   f"{avg['SQ_LDS_BANK_CONFLICT'] / 256:.4g} |")
    if "SQ_LDS_BANK_CONFLICT" in avg and "SQ_LDS_IDX_ACTIVE" in avg:
        out.append(f"| LDS bank-conflict share of LDS-active cycles | "
                   f"{avg['SQ_LDS_BANK_CONFLICT'] / avg['SQ_LDS_IDX_ACTIVE']:.3f} |")
    if "GRBM_GUI_ACTIVE" in avg:
        # GRBM_GUI_ACTIVE comes summed over the 8 XCDs (8x the kernel's cycles)
        cyc = avg["GRBM_GUI_ACTIVE"] / 8.0
        out.append(f"| kernel cycles (GRBM_GUI_ACTIVE / 8 XCDs) | {cyc:.4g} |")
        if "SQ_ACTIVE_INST_VALU" in avg:
            out.append(f"| VALU busy (4 SQ_ACTIVE_INST_VALU / 1024 SIMDs / cycles) | "
                       f"{4.0 * avg['SQ_ACTIVE_INST_VALU'] / 1024 / cyc:.3f} |")
        if "SQ_LDS_IDX_ACTIVE" in avg:
            out.append(f"| LDS busy (SQ_LDS_IDX_ACTIVE / 256 CUs / cycles) | "
                       f"{avg['SQ_LDS_IDX_ACTIVE'] / 256 / cyc:.3f} |")
        if "SQ_LDS_BANK_CONFLICT" in avg:
            out.append(f"| LDS bank-conflict share of kernel cycles | "
                       f"{avg['SQ_LDS_BANK_CONFLICT'] / 256 / cyc:.3f} |")
    if "SQ_WAIT_INST_LDS" in avg and avg.get("SQ_WAVE_CYCLES"):
        out.append(f"| wave-cycles waiting to issue LDS (SQ_WAIT_INST_LDS / SQ_WAVE_CYCLES) | "
                   f"{avg['SQ_WAIT_INST_LDS'] / avg['SQ_WAVE_CYCLES']:.3f} |")
    if "SQ_INSTS_LDS" in avg and points:
        out.append(f"| LDS instructions per quadrature point (x 64 lanes) | "
                   f"{64.0 * avg['SQ_INSTS_LDS'] / float(points):.2f} |")
    return "\n".join(out)


if __name__ == "__main__":
    mode = sys.argv[1]
    if mode == "stats":
        print(stats(sys.argv[2], *(sys.argv[3:4] or ["kernel trace"])))
    elif mode == "traffic":
        print(traffic(sys.argv[2], sys.argv[3], sys.argv[4], *(sys.argv[5:6] or ["HBM traffic"])))
    elif mode == "counters":
        # counters <pattern> <points> <pmc csv> [<pmc csv> ...]
        print(counters(sys.argv[4:], sys.argv[2], float(sys.argv[3]) or None))
    else:
        raise SystemExit(__doc__)
