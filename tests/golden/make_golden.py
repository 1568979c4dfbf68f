"""Writes the golden vectors under tests/golden/ by running the reference's own
serial kernels, compiled in place by oracle/Makefile into oracle/_ref/:

  BK3::Serial::SumFactorization  CEED_BK/include/kernels/BK3/serial_kernels.hpp:9-197
  BK5::Serial::SumFactorization  sum_factorization/include/kernels/BK5/serial_kernels.hpp:9-87
  LagrangeBasis                  sum_factorization/include/math/lagrange_basis.hpp:8-70

Cases:
  * random, non-symmetric G / input / basis tables (the reference's own fixture,
    G = 2 and in = 3, cannot tell a transposed layout apart);
  * the reference's verification fixture KAT-3 (G = 2, in = 3,
    basis[p*nm+i] = cos(p*nm+i), dbasis[i*nq+p] = cos(i*nq+p);
    CEED_BK/src/BK3/serial_verification.cc:25-74,
    sum_factorization/tests/BK5/serial_verification.cc:16-42);
  * Lagrange values/derivatives on the node sets of
    sum_factorization/tests/test_lagrange.cc:5-27 and on GLL nodes.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def bk3_case(nq, nelmt, rng):
    nm = nq - 1
    basis = rng.standard_normal(nq * nm)
    dbasis = rng.standard_normal(nq * nq)
    G = rng.standard_normal(nelmt * 6 * nq ** 3)
    inp = rng.standard_normal(nelmt * nm ** 3)
    out = np.zeros(nelmt * nm ** 3)
    s = O.ref_lib().ref_bk3_serial(nq, nelmt, basis, dbasis, G, inp.copy(), out)
    return dict(basis=basis, dbasis=dbasis, G=G, inp=inp, out=out, sum_sq=np.array(s))


def bk5_case(nq, nelmt, rng):
    dbasis = rng.standard_normal(nq * nq)
    G = rng.standard_normal(nelmt * 6 * nq ** 3)
    inp = rng.standard_normal(nelmt * nq ** 3)
    out = np.zeros(nelmt * nq ** 3)
    s = O.ref_lib().ref_bk5_serial(nq, nelmt, dbasis, G, inp, out)
    return dict(dbasis=dbasis, G=G, inp=inp, out=out, sum_sq=np.array(s))


def kat3(nq, nelmt, bk):
    """The reference's own fixture; returns sqrt(sum out^2) as the drivers print it."""
    if bk == 3:
        nm = nq - 1
        basis = np.array([np.cos(float(p * nm + i)) for p in range(nq) for i in range(nm)])
        dbasis = np.array([np.cos(float(i * nq + p)) for i in range(nq) for p in range(nq)])
        G = np.full(nelmt * 6 * nq ** 3, 2.0)
        inp = np.full(nelmt * nm ** 3, 3.0)
        out = np.zeros(nelmt * nm ** 3)
        s = O.ref_lib().ref_bk3_serial(nq, nelmt, basis, dbasis, G, inp, out)
    else:
        dbasis = np.array([np.cos(float(i * nq + a)) for i in range(nq) for a in range(nq)])
        G = np.full(nelmt * 6 * nq ** 3, 2.0)
        inp = np.full(nelmt * nq ** 3, 3.0)
        out = np.zeros(nelmt * nq ** 3)
        s = O.ref_lib().ref_bk5_serial(nq, nelmt, dbasis, G, inp, out)
    return np.sqrt(s)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref not built: run `make -C oracle` where the reference is present")
    rng = np.random.default_rng(20240611)
    arrays = {}
    for nq in range(3, 8):          # BK3: nm = nq - 1, p = 1..6 (q = p + 2)
        for k, v in bk3_case(nq, 2, rng).items():
            arrays[f"bk3_nq{nq}_{k}"] = v
    for nq in range(2, 10):         # BK5: p = 1..8
        for k, v in bk5_case(nq, 2, rng).items():
            arrays[f"bk5_nq{nq}_{k}"] = v
    for nq in range(3, 8):
        arrays[f"kat3_bk3_nq{nq}_norm"] = np.array(kat3(nq, 64, 3))
    for nq in range(2, 10):
        arrays[f"kat3_bk5_nq{nq}_norm"] = np.array(kat3(nq, 64, 5))
    # Lagrange basis on the reference test's node sets and on GLL nodes
    node_sets = {"x2": [-2.0, 0.5], "y3": [-2.0, 0.0, 1.0], "z4": [-3.0, 0.5, 1.0, 1.5]}
    for n in range(2, 10):
        node_sets[f"gll{n}"] = list(O.gll_rule(n)[0])
    xs = np.linspace(-2.0, 1.5, 29)
    for name, nodes in node_sets.items():
        nd = np.array(nodes, dtype=np.float64)
        val = np.array([[O.ref_lib().ref_lagrange_value(nd, len(nd), j, x) for x in xs]
                        for j in range(len(nd))])
        der = np.array([[O.ref_lib().ref_lagrange_derivative(nd, len(nd), j, x) for x in xs]
                        for j in range(len(nd))])
        arrays[f"lag_{name}_nodes"] = nd
        arrays[f"lag_{name}_value"] = val
        arrays[f"lag_{name}_deriv"] = der
    arrays["lag_x"] = xs
    path = os.path.join(HERE, "reference_serial.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path}: {len(arrays)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
