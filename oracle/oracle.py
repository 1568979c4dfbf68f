"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the
reference's own serial kernels (oracle/_ref/libref_serial*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INVALID = 0xFFFFFFFF
GAUSS, GLL = 0, 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class OrcBox(C.Structure):
    _fields_ = [("n", C.c_uint32 * 3), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3), ("deform", C.c_double)]


class OrcCgResult(C.Structure):
    _fields_ = [("steps", C.c_int), ("converged", C.c_int),
                ("initial_value", C.c_double), ("last_value", C.c_double)]


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built: run `make -C oracle`")
    return C.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _load(os.path.join(HERE, "liboracle.so"))
        L.orc_gauss_rule.argtypes = [C.c_int, _dp, _dp]
        L.orc_gll_rule.argtypes = [C.c_int, _dp, _dp]
        L.orc_lagrange_value.argtypes = [_dp, C.c_int, C.c_int, C.c_double]
        L.orc_lagrange_value.restype = C.c_double
        L.orc_lagrange_derivative.argtypes = [_dp, C.c_int, C.c_int, C.c_double]
        L.orc_lagrange_derivative.restype = C.c_double
        L.orc_basis_tables.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_bk3_serial_ref.argtypes = [C.c_uint, C.c_uint, _dp, _dp, _dp, _dp, _dp]
        L.orc_bk3_serial_ref.restype = C.c_double
        L.orc_bk5_serial_ref.argtypes = [C.c_uint, C.c_uint, _dp, _dp, _dp, _dp]
        L.orc_bk5_serial_ref.restype = C.c_double
        L.orc_bk3_evec.argtypes = [C.c_int, C.c_int, C.c_uint64, _dp, _dp, _dp, _dp, _dp]
        L.orc_bk5_evec.argtypes = [C.c_int, C.c_uint64, _dp, _dp, _dp, _dp]
        L.orc_box_n_cells.argtypes = [C.POINTER(OrcBox)]
        L.orc_box_n_cells.restype = C.c_uint64
        L.orc_box_n_dofs.argtypes = [C.POINTER(OrcBox), C.c_int]
        L.orc_box_n_dofs.restype = C.c_uint64
        L.orc_box_vertices.argtypes = [C.POINTER(OrcBox), _dp]
        L.orc_box_dof_map.argtypes = [C.POINTER(OrcBox), C.c_int, _u32p, _u8p]
        L.orc_box_cell_vertices.argtypes = [C.POINTER(OrcBox), _dp, _dp]
        L.orc_geometry.argtypes = [C.c_uint64, _dp, C.c_int, _dp, _dp, _dp]
        L.orc_dof_coordinates.argtypes = [C.POINTER(OrcBox), C.c_int, _dp, _u32p, _dp]
        L.orc_vmult.argtypes = [C.c_int, C.c_int, C.c_void_p, _dp, _dp, _u32p,
                                C.c_uint64, C.c_uint64, _u8p, _dp, _dp, C.c_int]
        L.orc_gather.argtypes = [C.c_int, _u32p, C.c_uint64, _dp, _dp]
        L.orc_diagonal.argtypes = [C.c_int, C.c_int, C.c_void_p, _dp, _dp, _u32p,
                                   C.c_uint64, C.c_uint64, _u8p, _dp]
        L.orc_invert_diagonal.argtypes = [C.c_uint64, _dp]
        L.orc_rhs.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _u32p,
                              C.c_uint64, C.c_uint64, C.c_int, _dp]
        L.orc_cg.argtypes = [C.c_int, C.c_int, C.c_void_p, _dp, _dp, _u32p,
                             C.c_uint64, C.c_uint64, _u8p, C.c_void_p, _dp, _dp,
                             C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_int]
        L.orc_cg.restype = OrcCgResult
        L.orc_l2_error_sin.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _u32p,
                                       C.c_uint64, _dp]
        L.orc_l2_error_sin.restype = C.c_double
        _lib = L
    return _lib


_ref = {}


def ref_lib(fast: bool = False):
    """The reference's own serial kernels, compiled in place (oracle/_ref)."""
    name = "libref_serial_fast.so" if fast else "libref_serial.so"
    if name not in _ref:
        L = _load(os.path.join(HERE, "_ref", name))
        L.ref_bk3_serial.argtypes = [C.c_uint, C.c_uint, _dp, _dp, _dp, _dp, _dp]
        L.ref_bk3_serial.restype = C.c_double
        L.ref_bk5_serial.argtypes = [C.c_uint, C.c_uint, _dp, _dp, _dp, _dp]
        L.ref_bk5_serial.restype = C.c_double
        L.ref_lagrange_value.argtypes = [_dp, C.c_int, C.c_int, C.c_double]
        L.ref_lagrange_value.restype = C.c_double
        L.ref_lagrange_derivative.argtypes = [_dp, C.c_int, C.c_int, C.c_double]
        L.ref_lagrange_derivative.restype = C.c_double
        L.ref_lagrange_value_f32.argtypes = [
            np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS"),
            C.c_int, C.c_int, C.c_float]
        L.ref_lagrange_value_f32.restype = C.c_float
        _ref[name] = L
    return _ref[name]


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libref_serial.so"))


# ---------------------------------------------------------------------------
def gauss_rule(n):
    x = np.zeros(n); w = np.zeros(n)
    assert lib().orc_gauss_rule(n, x, w) == 0
    return x, w


def gll_rule(n):
    x = np.zeros(n); w = np.zeros(n)
    assert lib().orc_gll_rule(n, x, w) == 0
    return x, w


def basis_tables(p, nq, quad=GAUSS):
    nm = p + 1
    B = np.zeros(nq * nm); D = np.zeros(nq * nq); xq = np.zeros(nq); wq = np.zeros(nq)
    assert lib().orc_basis_tables(p, nq, quad, B, D, xq, wq) == 0
    return B, D, xq, wq


def bk3_serial_ref(nq, nelmt, basis, dbasis, G, inp):
    out = np.zeros(nelmt * (nq - 1) ** 3)
    s = lib().orc_bk3_serial_ref(nq, nelmt, basis, dbasis, G, inp, out)
    return out, s


def bk5_serial_ref(nq, nelmt, dbasis, G, inp):
    out = np.zeros(nelmt * nq ** 3)
    s = lib().orc_bk5_serial_ref(nq, nelmt, dbasis, G, inp, out)
    return out, s


def bk3_evec(nm, nq, nelmt, B, D, G, inp):
    out = np.zeros(nelmt * nm ** 3)
    lib().orc_bk3_evec(nm, nq, nelmt, B, D, G, inp, out)
    return out


def bk5_evec(nq, nelmt, D, G, inp):
    out = np.zeros(nelmt * nq ** 3)
    lib().orc_bk5_evec(nq, nelmt, D, G, inp, out)
    return out


@dataclass
class Problem:
    """A structured box problem assembled by the oracle."""
    p: int
    nq: int
    quad: int
    box: OrcBox
    E: int
    N: int
    B: np.ndarray
    D: np.ndarray
    xq: np.ndarray
    wq: np.ndarray
    X: np.ndarray
    cellX: np.ndarray
    map: np.ndarray
    constrained: np.ndarray
    G: np.ndarray

    @property
    def nm(self):
        return self.p + 1

    @property
    def collocation(self):
        return self.quad == GLL and self.nq == self.p + 1

    def _Bptr(self):
        return None if self.collocation else self.B.ctypes.data

    def vmult(self, src, threads=8):
        dst = np.zeros(self.N)
        lib().orc_vmult(self.nm, self.nq, self._Bptr(), self.D, self.G, self.map,
                        self.E, self.N, self.constrained,
                        np.ascontiguousarray(src, dtype=np.float64), dst, threads)
        return dst

    def diagonal(self):
        d = np.zeros(self.N)
        lib().orc_diagonal(self.nm, self.nq, self._Bptr(), self.D, self.G, self.map,
                           self.E, self.N, self.constrained, d)
        return d

    def inverse_diagonal(self):
        d = self.diagonal()
        lib().orc_invert_diagonal(self.N, d)
        return d

    def rhs(self, f_kind):
        b = np.zeros(self.N)
        lib().orc_rhs(self.nm, self.nq, self.B, self.xq, self.wq, self.cellX, self.map,
                      self.E, self.N, f_kind, b)
        return b

    def cg(self, b, inv_diag=None, max_steps=100000, abs_tol=1e-16, rel_tol=1e-9,
           x0=None, threads=8):
        x = np.zeros(self.N) if x0 is None else np.array(x0, dtype=np.float64)
        hist = np.zeros(max_steps + 2)
        idp = None if inv_diag is None else np.ascontiguousarray(inv_diag).ctypes.data
        keep = inv_diag  # keep alive
        r = lib().orc_cg(self.nm, self.nq, self._Bptr(), self.D, self.G, self.map,
                         self.E, self.N, self.constrained, idp,
                         np.ascontiguousarray(b, dtype=np.float64), x, max_steps,
                         abs_tol, rel_tol, hist.ctypes.data, threads)
        del keep
        return x, r, hist[: r.steps + 1]

    def gather(self, src):
        ev = np.zeros(self.E * self.nm ** 3)
        lib().orc_gather(self.nm, self.map, self.E, np.ascontiguousarray(src, dtype=np.float64), ev)
        return ev

    def dof_coordinates(self):
        c = np.zeros(self.N * 3)
        lib().orc_dof_coordinates(C.byref(self.box), self.p, self.cellX, self.map, c)
        return c.reshape(self.N, 3)

    def l2_error_sin(self, u):
        return lib().orc_l2_error_sin(self.nm, self.nq, self.B, self.xq, self.wq,
                                      self.cellX, self.map, self.E,
                                      np.ascontiguousarray(u, dtype=np.float64))


def make_box(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), deform=0.0):
    if isinstance(n, int):
        n = (n, n, n)
    b = OrcBox()
    for d in range(3):
        b.n[d] = n[d]; b.lo[d] = lo[d]; b.hi[d] = hi[d]
    b.deform = deform
    return b


def make_problem(p, nq, quad, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), deform=0.0):
    L = lib()
    box = make_box(n, lo, hi, deform)
    E = L.orc_box_n_cells(C.byref(box))
    N = L.orc_box_n_dofs(C.byref(box), p)
    B, D, xq, wq = basis_tables(p, nq, quad)
    V = (box.n[0] + 1) * (box.n[1] + 1) * (box.n[2] + 1)
    X = np.zeros(V * 3)
    L.orc_box_vertices(C.byref(box), X)
    cellX = np.zeros(E * 24)
    L.orc_box_cell_vertices(C.byref(box), X, cellX)
    m = np.zeros(E * (p + 1) ** 3, dtype=np.uint32)
    cons = np.zeros(N, dtype=np.uint8)
    L.orc_box_dof_map(C.byref(box), p, m, cons)
    G = np.zeros(E * 6 * nq ** 3)
    L.orc_geometry(E, cellX, nq, xq, wq, G)
    return Problem(p, nq, quad, box, int(E), int(N), B, D, xq, wq, X, cellX, m, cons, G)


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """Uniform[-1,1) doubles from SplitMix64 (x += 0x9E3779B97F4A7C15, standard
    mix), u = (x >> 11) * 2^-53 * 2 - 1 -- the synthetic-input generator shared
    with the product library."""
    M = (1 << 64) - 1
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)) & np.uint64(M)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0
