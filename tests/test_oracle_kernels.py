"""Pins the oracle's element kernels: bitwise against the reference's own
serial kernels (golden vectors written by tests/golden/make_golden.py from the
reference compiled in place), and by a dense element matrix assembled
independently in numpy."""
import numpy as np
import pytest

from oracle import oracle as O

# reference G component order (rr, rs, rt, ss, st, tt) -> build-native index
# (xx=0, xy=1, xz=2, yy=3, yz=4, zz=5) with r <-> z, s <-> y, t <-> x
REF_TO_NATIVE = [5, 4, 2, 3, 1, 0]


@pytest.mark.parametrize("nq", range(3, 8))
def test_bk3_serial_bitwise(golden, nq):
    g = {k: golden[f"bk3_nq{nq}_{k}"] for k in ["basis", "dbasis", "G", "inp", "out", "sum_sq"]}
    out, s = O.bk3_serial_ref(nq, 2, g["basis"], g["dbasis"], g["G"], g["inp"])
    assert np.array_equal(out, g["out"])
    assert abs(s - float(g["sum_sq"])) <= 1e-13 * float(g["sum_sq"])


@pytest.mark.parametrize("nq", range(2, 10))
def test_bk5_serial_bitwise(golden, nq):
    g = {k: golden[f"bk5_nq{nq}_{k}"] for k in ["dbasis", "G", "inp", "out", "sum_sq"]}
    out, s = O.bk5_serial_ref(nq, 2, g["dbasis"], g["G"], g["inp"])
    assert np.array_equal(out, g["out"])


def _native_G_from_ref_bk5(G, nq, nelmt):
    # reference BK5 layout G[e][i][j][c][k] -> native G[e][c][i][j][k]
    Gr = G.reshape(nelmt, nq, nq, 6, nq)
    Gn = np.empty((nelmt, 6, nq, nq, nq))
    for c in range(6):
        Gn[:, REF_TO_NATIVE[c]] = Gr[:, :, :, c, :]
    return Gn.ravel()


def _native_G_from_ref_bk3(G, nq, nelmt):
    # reference BK3 layout G[e][p][q][c][r] -> native G[e][c][p][q][r]
    return _native_G_from_ref_bk5(G, nq, nelmt)


@pytest.mark.parametrize("nq", range(2, 10))
def test_bk5_native_layout_bitwise(golden, nq):
    g = {k: golden[f"bk5_nq{nq}_{k}"] for k in ["dbasis", "G", "inp", "out"]}
    Gn = _native_G_from_ref_bk5(g["G"], nq, 2)
    out = O.bk5_evec(nq, 2, g["dbasis"], Gn, g["inp"])
    assert np.array_equal(out, g["out"])


@pytest.mark.parametrize("nq", range(3, 8))
def test_bk3_native_layout_bitwise(golden, nq):
    g = {k: golden[f"bk3_nq{nq}_{k}"] for k in ["basis", "dbasis", "G", "inp", "out"]}
    Gn = _native_G_from_ref_bk3(g["G"], nq, 2)
    out = O.bk3_evec(nq - 1, nq, 2, g["basis"], g["dbasis"], Gn, g["inp"])
    assert np.array_equal(out, g["out"])


@pytest.mark.parametrize("kind,nq", [("bk3", q) for q in range(3, 8)] + [("bk5", q) for q in range(2, 10)])
def test_reference_fixture_norms(golden, kind, nq):
    """KAT-3: G = 2, in = 3, cos() tables -> sqrt(sum out^2) of the reference."""
    nelmt = 64
    if kind == "bk5":
        db = np.array([np.cos(float(i * nq + a)) for i in range(nq) for a in range(nq)])
        out = O.bk5_evec(nq, nelmt, db, np.full(nelmt * 6 * nq ** 3, 2.0),
                         np.full(nelmt * nq ** 3, 3.0))
    else:
        nm = nq - 1
        b = np.array([np.cos(float(p * nm + i)) for p in range(nq) for i in range(nm)])
        db = np.array([np.cos(float(i * nq + p)) for i in range(nq) for p in range(nq)])
        out = O.bk3_evec(nm, nq, nelmt, b, db, np.full(nelmt * 6 * nq ** 3, 2.0),
                         np.full(nelmt * nm ** 3, 3.0))
    ref = float(golden[f"kat3_{kind}_nq{nq}_norm"])
    assert abs(np.linalg.norm(out) - ref) <= 1e-13 * ref


def dense_element_matrix(p, nq, B, D, Ge):
    """K[l, m] = sum_q grad(phi_l)(q)^T G(q) grad(phi_m)(q), built from 1D
    tables with numpy (x fastest): an independent statement of the element
    operator."""
    nm = p + 1
    Bm, Dm = B.reshape(nq, nm), D.reshape(nq, nq)
    Dh = Dm @ Bm
    gx = np.einsum("xa,yb,zc->zyxcba", Dh, Bm, Bm).reshape(nq ** 3, nm ** 3)
    gy = np.einsum("xa,yb,zc->zyxcba", Bm, Dh, Bm).reshape(nq ** 3, nm ** 3)
    gz = np.einsum("xa,yb,zc->zyxcba", Bm, Bm, Dh).reshape(nq ** 3, nm ** 3)
    G = Ge.reshape(6, nq ** 3)
    grads = [gx, gy, gz]
    comp = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    K = np.zeros((nm ** 3, nm ** 3))
    for a in range(3):
        for b in range(3):
            c = comp[(min(a, b), max(a, b))]
            K += grads[a].T @ (G[c][:, None] * grads[b])
    return K


@pytest.mark.parametrize("p,nq,quad", [(1, 2, O.GLL), (2, 3, O.GLL), (3, 4, O.GLL),
                                       (1, 3, O.GAUSS), (2, 4, O.GAUSS), (3, 5, O.GAUSS),
                                       (2, 3, O.GAUSS)])
def test_element_kernel_matches_dense_matrix(p, nq, quad):
    rng = np.random.default_rng(p * 10 + nq)
    B, D, _, _ = O.basis_tables(p, nq, quad)
    nm = p + 1
    Ge = rng.standard_normal(6 * nq ** 3)  # arbitrary (non-symmetric-looking) metric field
    K = dense_element_matrix(p, nq, B, D, Ge)
    u = rng.standard_normal(nm ** 3)
    if quad == O.GLL and nq == p + 1:
        v = O.bk5_evec(nq, 1, D, Ge, u)
    else:
        v = O.bk3_evec(nm, nq, 1, B, D, Ge, u)
    assert np.allclose(v, K @ u, rtol=1e-12, atol=1e-12 * np.abs(K @ u).max())
    assert np.allclose(K, K.T, atol=1e-12 * np.abs(K).max())          # symmetric
    assert np.allclose(K @ np.ones(nm ** 3), 0, atol=1e-11 * np.abs(K).max())  # K 1 = 0


def test_planted_transposition_is_caught(golden):
    """A plausible mistake -- the transposed derivative in the forward pass --
    changes the output on the random golden fixture (it would not on the
    reference's symmetric G = 2, in = 3 fixture)."""
    nq = 4
    g = {k: golden[f"bk5_nq{nq}_{k}"] for k in ["dbasis", "G", "inp", "out"]}
    Gn = _native_G_from_ref_bk5(g["G"], nq, 2)
    Dt = g["dbasis"].reshape(nq, nq).T.copy().ravel()
    out = O.bk5_evec(nq, 2, Dt, Gn, g["inp"])
    assert not np.allclose(out, g["out"])
