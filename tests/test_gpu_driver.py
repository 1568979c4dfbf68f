"""The C++ host API (include/mfx.hpp) through the reference-style driver
benchmarks_amd/apps/bp_driver, read like the reference's own drivers: CG
iteration counts of CEED_bp/src/bp3.cc's mesh cycles, and the verification
fixture norms of the reference's serial_verification programs."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "benchmarks_amd", "apps", "bp_driver")

# CEED_bp/results/1xGH200_P4.txt:635-639
KAT1 = {256: 92, 512: 99, 1024: 153, 2048: 170, 4096: 180}


def run(*args, timeout=120):
    assert os.path.exists(DRIVER), "bp_driver not built"
    print("run", DRIVER, *args, flush=True)
    out = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True,
                         timeout=timeout)
    assert out.returncode == 0, out.stderr + out.stdout
    return out.stdout


def test_driver_bp3_cycles(gpu):
    text = run(4, 10000, 300000)
    rows = re.findall(r"RESULT cycle=(\d+) cells=(\d+) dofs=(\d+) .* cg_its=(\d+)", text)
    seen = {int(c): int(its) for _, c, _, its in rows}
    assert set(KAT1) <= set(seen), text
    for cells, its in KAT1.items():
        assert seen[cells] == its, (cells, seen[cells], its)


@pytest.mark.parametrize("kind,nq", [("bk3", q) for q in range(3, 8)] +
                         [("bk5", q) for q in range(2, 10)])
def test_driver_verification_fixture(gpu, golden, kind, nq):
    text = run("--verify", kind, nq, 64)
    norm = float(re.search(r"norm=(\S+)", text).group(1))
    ref = float(golden[f"kat3_{kind}_nq{nq}_norm"])
    assert abs(norm - ref) <= 1e-12 * abs(ref), (norm, ref)


def test_driver_bp5_runs(gpu):
    text = run("--bp5", 4, 8, 1)
    m = re.search(r"RESULT bp5 p=4 cells=512 dofs=35937 .* cg_its=(\d+)", text)
    assert m and int(m.group(1)) > 10


def test_driver_mixed_precision_multigrid(gpu):
    """PreconditionMG through the C++ API: FP64 CG to 1e-10 with the FP32
    V-cycle takes a handful of steps where Jacobi-CG takes hundreds, and the
    two solutions agree to the Jacobi solve's 1e-8 tolerance."""
    text = run("--mg", 4, 16)
    m = re.search(r"RESULT mg p=4 cells=4096 dofs=274625 .* mg_its=(\d+) .* jacobi_its=(\d+) "
                  r".* rel_diff=(\S+)", text)
    assert m, text
    mg_its, j_its, diff = int(m.group(1)), int(m.group(2)), float(m.group(3))
    assert mg_its <= 12 and j_its >= 10 * mg_its
    assert diff <= 1e-6
    assert "levels: Q4/16^3 Q2/16^3 Q1/16^3 Q1/8^3" in text


def test_driver_batch_matches_python_instances(gpu):
    """BatchedLaplace through the C++ API on the batched-solve instance
    generator (C++ SplitMix64, the stream of laplace.batch_instances): the
    same DoF count as the Python list, and every instance converges."""
    from benchmarks_amd.laplace import batch_instances
    inst = batch_instances(12, 2)
    dofs = sum((n * p + 1) ** 3 for p, n in inst)
    text = run("--batch", 12, 2)
    m = re.search(r"RESULT batch instances=12 dofs=(\d+) .* steps_max=(\d+) .* converged=(\d+)",
                  text)
    assert m, text
    assert int(m.group(1)) == dofs and int(m.group(3)) == 12 and int(m.group(2)) > 10


def test_driver_flexible_cg(gpu):
    """LearnedPreconditioner and SolverCG's flexible CG through the C++ API
    (network off: the Chebyshev cycle alone) beat Jacobi-CG in steps."""
    text = run("--fcg", 3, 6, 2)
    m = re.search(r"RESULT fcg p=3 cells=216 .* fcg_its=(\d+) jacobi_its=(\d+)", text)
    assert m, text
    assert 0 < int(m.group(1)) < int(m.group(2))
