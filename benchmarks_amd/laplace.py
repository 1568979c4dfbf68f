"""Host-side mirror of the reference's operator interface on top of libmfx.

Names and semantics follow the matrix concept deal.II's SolverCG consumes
(CEED_bp/include/portable_laplace_operator.h:17-96) and the solver pieces the
reference drivers use (SolverCG, ReductionControl, PreconditionIdentity,
DiagonalMatrix; CEED_bp/src/bp3.cc:264-288, bp5_kokkos/benchmark.cc:336-378).
Vectors are torch tensors on the GPU (torch is the device-memory and stream
plumbing only); every operation runs in libmfx's HIP kernels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import check, lib

GAUSS, GLL = _lib.QUAD_GAUSS, _lib.QUAD_GLL


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


F64, F32, I32 = torch.float64, torch.float32, torch.int32


def _ptr(t: torch.Tensor, dtype=None, n: int = None):
    """Device pointer of a contiguous tensor, after checking its element
    type (the C entry points take typed pointers: a tensor of another type
    would be read past its end) and, when given, that it holds n elements."""
    if not t.is_cuda:
        raise ValueError("libmfx works on device vectors; got a host tensor")
    if not t.is_contiguous():
        raise ValueError("vectors must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"expected a {dtype} tensor, got {t.dtype}")
    if n is not None and t.numel() != n:
        raise ValueError(f"expected {n} elements, got {t.numel()}")
    return C.c_void_p(t.data_ptr())


class NoConvergence(RuntimeError):
    """deal.II SolverControl::NoConvergence."""

    def __init__(self, last_step: int, last_residual: float):
        super().__init__(f"Iterative method reported convergence failure in step {last_step}. "
                         f"The residual in the last step was {last_residual}.")
        self.last_step = last_step
        self.last_residual = last_residual


class LaplaceOperator:
    """Matrix-free scalar Laplacian on a (deformed) structured hex mesh.

    Mirrors Portable::LaplaceOperator<dim, fe_degree, nq, number>
    (CEED_bp/include/portable_laplace_operator.h:17-96): vmult, Tvmult, m, n,
    initialize_dof_vector; set-up by degree, quadrature and mesh as in the
    LaplaceProblem drivers (CEED_bp/src/bp3.cc:129-182)."""

    def __init__(self, degree: int, n_q_points_1d: int, quadrature: int = GAUSS,
                 n_cells=(8, 8, 8), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                 deform: float = 0.0, with_fp32: bool = False, geometry: int = 0,
                 _handle=None, _owner=None):
        self._h = None
        # a handle owned by another object (a multigrid level) is not destroyed here
        self._owner = _owner
        # cells per direction of a box operator (None for general meshes)
        self.box_cells = None
        if _handle is not None:
            self._h = _handle
        else:
            if isinstance(n_cells, int):
                n_cells = (n_cells,) * 3
            d = _lib.BoxDesc()
            d.degree, d.n_q_points_1d, d.quadrature = degree, n_q_points_1d, quadrature
            d.geometry = geometry
            for k in range(3):
                d.n_cells[k] = n_cells[k]
                d.lo[k] = lo[k]
                d.hi[k] = hi[k]
            d.deform = deform
            d.with_fp32 = int(with_fp32)
            h = C.c_void_p()
            check(lib().mfx_op_create_box(C.byref(d), C.byref(h)))
            self._h = h
            self.box_cells = tuple(int(c) for c in n_cells)
        info = _lib.OpInfo()
        check(lib().mfx_op_get_info(self._h, C.byref(info)))
        self.info = info
        self.degree = info.degree
        self.n_q_points_1d = info.n_q_points_1d
        self.quadrature = info.quadrature

    @classmethod
    def from_arrays(cls, degree, n_q_points_1d, quadrature, map_, cell_vertices, constrained,
                    n_dofs, with_fp32=False, geometry: int = 0):
        """General (unstructured) creation from host numpy arrays: the DoF map
        [n_cells][(p+1)^3] (0xFFFFFFFF = constrained), the 8 vertices of each
        cell and the list of constrained DoFs; geometry 1 computes the metric
        terms from the vertices (Gauss quadrature only)."""
        import numpy as np
        m = np.ascontiguousarray(map_, dtype=np.uint32)
        cv = np.ascontiguousarray(cell_vertices, dtype=np.float64)
        cons = np.ascontiguousarray(constrained, dtype=np.uint32)
        nm3 = (degree + 1) ** 3
        d = _lib.OpDesc()
        d.degree, d.n_q_points_1d, d.quadrature = degree, n_q_points_1d, quadrature
        d.geometry = geometry
        d.n_cells = m.size // nm3
        d.n_dofs = n_dofs
        d.map = m.ctypes.data
        d.cell_vertices = cv.ctypes.data
        d.constrained = cons.ctypes.data if cons.size else None
        d.n_constrained = cons.size
        d.with_fp32 = int(with_fp32)
        h = C.c_void_p()
        check(lib().mfx_op_create(C.byref(d), C.byref(h)))
        return cls(degree, n_q_points_1d, quadrature, _handle=h)

    def __del__(self):
        if getattr(self, "_owner", None) is not None:
            return
        if getattr(self, "_h", None) is not None and lib is not None:
            try:
                lib().mfx_op_destroy(self._h)
            except Exception:
                pass
            self._h = None

    # --- matrix concept -------------------------------------------------
    def m(self) -> int:
        return int(lib().mfx_op_n_dofs(self._h))

    def n(self) -> int:
        return self.m()

    def n_cells(self) -> int:
        return int(lib().mfx_op_n_cells(self._h))

    def initialize_dof_vector(self, dtype=torch.float64) -> torch.Tensor:
        return torch.zeros(self.m(), dtype=dtype, device="cuda")

    def vmult(self, dst: torch.Tensor, src: torch.Tensor) -> None:
        """dst = A src (constrained rows: identity),
        CEED_bp/include/portable_laplace_operator.h:124-172."""
        self._check_vec(dst)
        self._check_vec(src)
        if dst.dtype == torch.float64:
            check(lib().mfx_op_vmult(self._h, _ptr(dst, F64), _ptr(src, F64), _stream()))
        elif dst.dtype == torch.float32:
            check(lib().mfx_op_vmult_f32(self._h, _ptr(dst, F32), _ptr(src, F32), _stream()))
        else:
            raise TypeError("vmult supports float64 and float32 vectors")

    def vmult_add(self, dst: torch.Tensor, src: torch.Tensor) -> None:
        """dst += A_free src on the unconstrained rows (the element kernel alone)."""
        self._check_vec(dst)
        self._check_vec(src)
        check(lib().mfx_op_vmult_add(self._h, _ptr(dst, F64), _ptr(src, F64), _stream()))

    def Tvmult(self, dst, src) -> None:
        self.vmult(dst, src)

    def set_deterministic(self, on: bool = True) -> None:
        """Bitwise-reproducible applies: per-patch sums added per DoF in a
        fixed order instead of atomically (patch-kernel operators only)."""
        check(lib().mfx_op_set_deterministic(self._h, int(on)))

    def _check_vec(self, v: torch.Tensor):
        if v.numel() != self.m():
            raise ValueError(f"vector size {v.numel()} != m() = {self.m()}")

    # --- set-up helpers -------------------------------------------------
    def compute_diagonal(self) -> torch.Tensor:
        d = self.initialize_dof_vector()
        check(lib().mfx_op_diagonal(self._h, _ptr(d, F64, self.m()), _stream()))
        return d

    def compute_inverse_diagonal(self) -> torch.Tensor:
        d = self.initialize_dof_vector()
        check(lib().mfx_op_inverse_diagonal(self._h, _ptr(d, F64, self.m()), _stream()))
        return d

    def rhs(self, f_kind: int = 0) -> torch.Tensor:
        b = self.initialize_dof_vector()
        check(lib().mfx_op_rhs(self._h, f_kind, _ptr(b, F64, self.m()), _stream()))
        return b

    def zero_constrained(self, v: torch.Tensor) -> None:
        check(lib().mfx_vec_zero_constrained(self._h, _ptr(v, F64, self.m()), _stream()))

    def dof_coordinates(self) -> torch.Tensor:
        c = torch.zeros(self.m() * 3, dtype=torch.float64, device="cuda")
        check(lib().mfx_op_dof_coordinates(self._h, _ptr(c, F64, 3 * self.m()), _stream()))
        return c.view(-1, 3)

    def device_array(self, name: str, dtype, count: int) -> torch.Tensor:
        """Copy of an operator-owned device array (map, G, B, D, cell_vertices,
        constrained, xq, wq) as a new tensor."""
        return _device_view(getattr(self.info, name), dtype, count).clone()


_TYPESTR = {torch.float64: "<f8", torch.float32: "<f4", torch.int32: "<i4"}


class _DevArray:
    def __init__(self, ptr, typestr, count):
        self.__cuda_array_interface__ = {"shape": (count,), "typestr": typestr,
                                         "data": (ptr, False), "version": 2}


def _device_view(ptr: int, dtype, count: int) -> torch.Tensor:
    """Zero-copy tensor view of device memory owned by libmfx (uint32 arrays
    are viewed as int32)."""
    return torch.as_tensor(_DevArray(ptr, _TYPESTR[dtype], count), device="cuda")


@dataclass
class ReductionControl:
    """deal.II ReductionControl(max_steps, tolerance, reduction)."""
    max_steps: int = 100
    tolerance: float = 1e-10
    reduction: float = 1e-2
    _last_step: int = 0
    _initial: float = 0.0
    _last: float = 0.0
    seconds: float = 0.0

    def last_step(self) -> int:
        return self._last_step

    def initial_value(self) -> float:
        return self._initial

    def last_value(self) -> float:
        return self._last


class PreconditionIdentity:
    pass


class DiagonalMatrix:
    """deal.II DiagonalMatrix holding the inverse diagonal (Jacobi)."""

    def __init__(self, inverse_diagonal: torch.Tensor):
        self.vector = inverse_diagonal

    def get_vector(self) -> torch.Tensor:
        return self.vector


class PreconditionMG:
    """Mixed-precision p/h-multigrid preconditioner (mfx_mg_*): FP32 V-cycle
    with Chebyshev(Jacobi) smoothing, nodal-interpolation transfers and an
    FP64 direct solve on the coarsest level.  No counterpart in the
    reference, whose multigrid include is unused (CEED_bp/src/bp3.cc:27);
    the parameters are those of deal.II's PreconditionChebyshev.  The fine
    operator (level 0) is built here: operator() returns it."""

    def __init__(self, degree: int, n_q_points_1d: int, quadrature: int = GAUSS,
                 n_cells=(8, 8, 8), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), deform: float = 0.0,
                 geometry: int = 0, smoother_degree: int = 5, smoothing_range: float = 15.0,
                 eig_cg_n_iterations: int = 12, eig_safety: float = 1.2,
                 coarse_max_dofs: int = 2000, fine_smoother_n_q_points_1d: int = 0):
        if isinstance(n_cells, int):
            n_cells = (n_cells,) * 3
        d = _lib.MgDesc()
        f = d.fine
        f.degree, f.n_q_points_1d = degree, n_q_points_1d
        f.quadrature, f.geometry = quadrature, geometry
        for k in range(3):
            f.n_cells[k], f.lo[k], f.hi[k] = n_cells[k], lo[k], hi[k]
        f.deform = deform
        f.with_fp32 = 1
        d.smoother_degree, d.smoothing_range = smoother_degree, smoothing_range
        d.eig_cg_n_iterations, d.eig_safety = eig_cg_n_iterations, eig_safety
        d.coarse_max_dofs = coarse_max_dofs
        d.fine_smoother_n_q_points_1d = fine_smoother_n_q_points_1d
        self._h = None
        h = C.c_void_p()
        check(lib().mfx_mg_create(C.byref(d), C.byref(h)))
        self._h = h
        self._levels = [self._level(i) for i in range(self.n_levels())]

    def __del__(self):
        if getattr(self, "_h", None) is not None and lib is not None:
            try:
                lib().mfx_mg_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def n_levels(self) -> int:
        return int(lib().mfx_mg_n_levels(self._h))

    def _level(self, i):
        info = _lib.MgLevelInfo()
        check(lib().mfx_mg_get_level(self._h, i, C.byref(info)))
        op = LaplaceOperator(0, 0, _handle=C.c_void_p(info.op), _owner=self)
        op.box_cells = tuple(int(c) for c in info.n_cells)
        return {"degree": info.degree, "n_q_points_1d": info.n_q_points_1d,
                "n_cells": tuple(info.n_cells), "n_dofs": int(info.n_dofs),
                "lambda_max": info.lambda_max, "lambda_min": info.lambda_min, "operator": op}

    def level(self, i: int) -> dict:
        return self._levels[i]

    def operator(self) -> LaplaceOperator:
        return self._levels[0]["operator"]

    def vmult(self, dst: torch.Tensor, src: torch.Tensor) -> None:
        """dst = one V-cycle applied to src (FP64 in and out)."""
        n = self.operator().m()
        check(lib().mfx_mg_vmult(self._h, _ptr(dst, F64, n), _ptr(src, F64, n), _stream()))


def transfer_prolongate(fine: LaplaceOperator, coarse: LaplaceOperator, src: torch.Tensor):
    dst = fine.initialize_dof_vector()
    check(lib().mfx_transfer_prolongate_f64(fine._h, coarse._h, _ptr(dst, F64, fine.m()),
                                             _ptr(src, F64, coarse.m()), _stream()))
    return dst


def transfer_restrict(fine: LaplaceOperator, coarse: LaplaceOperator, src: torch.Tensor):
    dst = coarse.initialize_dof_vector()
    check(lib().mfx_transfer_restrict_f64(fine._h, coarse._h, _ptr(dst, F64, coarse.m()),
                                           _ptr(src, F64, fine.m()), _stream()))
    return dst


def chebyshev_f32(op: LaplaceOperator, inv_diag: torch.Tensor, x: torch.Tensor, b: torch.Tensor,
                  degree: int, lambda_min: float, lambda_max: float, x_is_zero: bool) -> None:
    """x receives `degree` Chebyshev(Jacobi) steps towards A x = b (FP32)."""
    n = op.m()
    check(lib().mfx_chebyshev_f32(op._h, _ptr(inv_diag, F32, n), _ptr(x, F32, n), _ptr(b, F32, n),
                                  degree, lambda_min,
                                  lambda_max, int(x_is_zero), _stream()))


class SolverCG:
    """deal.II SolverCG<Vector>(control).solve(A, x, b, preconditioner).  With
    identity or Jacobi preconditioning the whole iteration runs on the GPU
    (mfx_cg_solve); with PreconditionMG it is the FP64 CG of mfx_mg_solve."""

    def __init__(self, control: ReductionControl):
        self.control = control

    def solve(self, A: LaplaceOperator, x: torch.Tensor, b: torch.Tensor, preconditioner) -> None:
        A._check_vec(x)
        A._check_vec(b)
        res = _lib.CgResult()
        if isinstance(preconditioner, LearnedPreconditioner):
            st = lib().mfx_nn_fcg_solve(preconditioner._h, A._h, _ptr(x, F64, A.m()),
                                        _ptr(b, F64, A.m()),
                                        self.control.max_steps, self.control.tolerance,
                                        self.control.reduction, C.byref(res), _stream())
            self._record(res, st)
            return
        if isinstance(preconditioner, PreconditionMG):
            if A._h.value != preconditioner.operator()._h.value:
                raise ValueError("A must be the multigrid's fine operator "
                                 "(PreconditionMG.operator())")
            st = lib().mfx_mg_solve(preconditioner._h, _ptr(x, F64, A.m()), _ptr(b, F64, A.m()),
                                    self.control.max_steps,
                                    self.control.tolerance, self.control.reduction, C.byref(res),
                                    _stream())
            self._record(res, st)
            return
        if isinstance(preconditioner, DiagonalMatrix):
            inv = _ptr(preconditioner.vector, F64, A.m())
        elif isinstance(preconditioner, PreconditionIdentity) or preconditioner is None:
            inv = None
        else:
            raise TypeError("preconditioner must be PreconditionIdentity, DiagonalMatrix or "
                            "PreconditionMG")
        st = lib().mfx_cg_solve(A._h, _ptr(x, F64, A.m()), _ptr(b, F64, A.m()), inv,
                                self.control.max_steps,
                                self.control.tolerance, self.control.reduction, C.byref(res),
                                _stream())
        self._record(res, st)

    def _record(self, res, st):
        c = self.control
        c._last_step, c._initial, c._last, c.seconds = (res.steps, res.initial_value,
                                                        res.last_value, res.seconds)
        if st == 4 or st == 6:
            raise NoConvergence(res.steps, res.last_value)
        check(st)


class BatchedLaplace:
    """Many independent BP5 problems (GLL collocation, degree 2..6, deformed
    boxes) on one concatenated DoF vector (mfx_batch_*): instance i owns
    [offsets[i], offsets[i+1]).  No counterpart in the reference, which runs
    one problem per process.  instances: iterable of (degree, n_cells) or
    (degree, n_cells, deform)."""

    def __init__(self, instances, autotune: bool = True, deform: float = 0.05):
        inst = list(instances)
        arr = (_lib.BatchInstance * len(inst))()
        for i, it in enumerate(inst):
            p, n = it[0], it[1]
            d = it[2] if len(it) > 2 else deform
            n = (n,) * 3 if isinstance(n, int) else n
            arr[i].degree = p
            for k in range(3):
                arr[i].n_cells[k], arr[i].lo[k], arr[i].hi[k] = n[k], 0.0, 1.0
            arr[i].deform = d
        self._h = None
        h = C.c_void_p()
        check(lib().mfx_batch_create(arr, len(inst), int(autotune), C.byref(h)))
        self._h = h
        self.n_instances = len(inst)
        off = (C.c_uint64 * (len(inst) + 1))()
        check(lib().mfx_batch_offsets(self._h, off))
        self.offsets = [int(v) for v in off]

    def __del__(self):
        if getattr(self, "_h", None) is not None and lib is not None:
            try:
                lib().mfx_batch_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def m(self) -> int:
        return int(lib().mfx_batch_n_dofs(self._h))

    def initialize_dof_vector(self) -> torch.Tensor:
        return torch.zeros(self.m(), dtype=torch.float64, device="cuda")

    def instance_operator(self, i: int) -> LaplaceOperator:
        h = lib().mfx_batch_instance_operator(self._h, i)
        return LaplaceOperator(0, 0, _handle=C.c_void_p(h), _owner=self)

    def vmult(self, dst: torch.Tensor, src: torch.Tensor) -> None:
        check(lib().mfx_batch_vmult(self._h, _ptr(dst, F64, self.m()), _ptr(src, F64, self.m()),
                                    _stream()))

    def compute_inverse_diagonal(self) -> torch.Tensor:
        d = self.initialize_dof_vector()
        check(lib().mfx_batch_inverse_diagonal(self._h, _ptr(d, F64, self.m()), _stream()))
        return d

    def zero_constrained(self, v: torch.Tensor) -> None:
        for i in range(self.n_instances):
            self.instance_operator(i).zero_constrained(v[self.offsets[i]:self.offsets[i + 1]])

    def tuning(self, degree: int) -> dict:
        t = _lib.BatchTuning()
        check(lib().mfx_batch_get_tuning(self._h, degree, C.byref(t)))
        nv = t.n_variants
        geo = {_lib.GEOM_STORED: "stored", _lib.GEOM_ON_THE_FLY: "on_the_fly"}
        return {"degree": t.degree, "n_cells": t.n_cells, "variant": t.variant,
                "default_variant": t.default_variant,
                "cells_per_workgroup": list(t.cells_per_workgroup)[:nv],
                "prefetch": list(t.prefetch)[:nv],
                "geometry": [geo[g] for g in list(t.geometry)[:nv]], "ms": list(t.ms)[:nv]}

    def set_variant(self, degree: int, variant: int) -> None:
        """Use tuning variant `variant` for the instances of `degree`."""
        check(lib().mfx_batch_set_variant(self._h, degree, variant))

    def solve(self, x: torch.Tensor, b: torch.Tensor, inverse_diagonal=None, max_steps=100000,
              abs_tol=1e-300, rel_tol=1e-8, use_graph=True):
        """Independent Jacobi-CG (identity if inverse_diagonal is None) per
        instance; returns one ReductionControl per instance.  Raises
        NoConvergence if any instance fails."""
        res = (_lib.CgResult * self.n_instances)()
        inv = None if inverse_diagonal is None else _ptr(inverse_diagonal, F64, self.m())
        st = lib().mfx_batch_cg_solve(self._h, _ptr(x, F64, self.m()), _ptr(b, F64, self.m()), inv,
                                      max_steps, abs_tol, rel_tol,
                                      int(use_graph), res, _stream())
        out = []
        for r in res:
            c = ReductionControl(max_steps, abs_tol, rel_tol)
            c._last_step, c._initial, c._last, c.seconds = (r.steps, r.initial_value,
                                                            r.last_value, r.seconds)
            out.append(c)
        if st == 4:
            raise NoConvergence(max(r.steps for r in res), max(r.last_value for r in res))
        check(st)
        return out


class LearnedPreconditioner:
    """The learned preconditioner (mfx_nn_*): a 4-level 3D U-Net on the node
    grid of a box operator mapping a residual r to z = ||r|| s net(r/||r||),
    trained on the GPU by the library's own BF16-MFMA forward and FP32
    backward kernels.  No counterpart in the reference.  With SolverCG it
    runs flexible CG (mfx_nn_fcg_solve)."""

    def __init__(self, seed: int = 0):
        self._h = None
        h = C.c_void_p()
        check(lib().mfx_nn_create(seed, C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) is not None and lib is not None:
            try:
                lib().mfx_nn_destroy(self._h)
            except Exception:
                pass
            self._h = None

    @staticmethod
    def n_params() -> int:
        return int(lib().mfx_nn_n_params())

    def get_params(self):
        import numpy as np
        out = np.zeros(self.n_params(), dtype=np.float32)
        check(lib().mfx_nn_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, params) -> None:
        import numpy as np
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.n_params()
        check(lib().mfx_nn_set_params(self._h, p.ctypes.data))

    @property
    def output_scale(self) -> float:
        return float(lib().mfx_nn_output_scale(self._h))

    @output_scale.setter
    def output_scale(self, s: float) -> None:
        check(lib().mfx_nn_set_output_scale(self._h, s))

    def set_mode(self, smoother_degree: int, use_network: bool = True) -> None:
        """0: the network alone; k > 0: k Chebyshev(Jacobi) steps, the network
        on the remaining residual, k steps again (needs an operator built
        with with_fp32); use_network=False keeps only the Chebyshev steps."""
        check(lib().mfx_nn_set_mode(self._h, smoother_degree, int(use_network)))

    def set_coarse(self, coarse) -> None:
        """Run the network on a coarser box operator's node grid (restricted
        residual in, prolongated correction out); None: the fine grid."""
        self._coarse = coarse  # keep the operator alive while in use
        check(lib().mfx_nn_set_coarse(self._h, None if coarse is None else coarse._h))

    def smoother_info(self, op: LaplaceOperator) -> dict:
        """Chebyshev interval on `op` and the weight of the network's
        correction in the last application."""
        lo, hi, om = C.c_double(), C.c_double(), C.c_double()
        check(lib().mfx_nn_smoother_info(self._h, op._h, C.byref(lo), C.byref(hi), C.byref(om),
                                         _stream()))
        return {"lambda_min": lo.value, "lambda_max": hi.value, "omega": om.value}

    @staticmethod
    def _dims(dims):
        return (C.c_uint32 * 3)(*[int(d) for d in dims])

    def forward(self, dims, x: torch.Tensor) -> torch.Tensor:
        """net(x) on a grid of dims = (nx, ny, nz) nodes (FP32 fields)."""
        y = torch.empty_like(x)
        n = dims[0] * dims[1] * dims[2]
        check(lib().mfx_nn_forward(self._h, self._dims(dims), _ptr(x, F32, n), _ptr(y, F32, n),
                                   _stream()))
        return y

    def gradient(self, dims, x: torch.Tensor, dy: torch.Tensor) -> torch.Tensor:
        g = torch.empty(self.n_params(), dtype=torch.float32, device="cuda")
        n = dims[0] * dims[1] * dims[2]
        check(lib().mfx_nn_gradient(self._h, self._dims(dims), _ptr(x, F32, n), _ptr(dy, F32, n),
                                    _ptr(g, F32),
                                    _stream()))
        return g

    def train_step(self, dims, x: torch.Tensor, target: torch.Tensor, lr: float) -> float:
        """One Adam step on the mean squared error over a batch of grids
        (x, target: [batch, nz, ny, nx] FP32); returns the loss before it."""
        batch = x.shape[0]
        loss = C.c_double()
        n = batch * dims[0] * dims[1] * dims[2]
        check(lib().mfx_nn_train_step(self._h, self._dims(dims), batch, _ptr(x, F32, n),
                                      _ptr(target, F32, n),
                                      lr, C.byref(loss), _stream()))
        return loss.value

    def train_energy_step(self, op: LaplaceOperator, r: torch.Tensor, e: torch.Tensor,
                          lr: float) -> float:
        """One Adam step on the energy-norm loss of whole fields (r, e:
        [batch, N] FP64, A e = r); returns the loss before it."""
        loss = C.c_double()
        n = r.shape[0] * op.m()
        check(lib().mfx_nn_train_energy(self._h, op._h, r.shape[0], _ptr(r, F64, n),
                                        _ptr(e, F64, n), lr,
                                        C.byref(loss), _stream()))
        return loss.value

    def vmult(self, op: LaplaceOperator, z: torch.Tensor, r: torch.Tensor) -> None:
        check(lib().mfx_nn_vmult(self._h, op._h, _ptr(z, F64, op.m()), _ptr(r, F64, op.m()),
                                 _stream()))

    def restricted_residual(self, op: LaplaceOperator, r: torch.Tensor, rho_c: torch.Tensor):
        """rho_c = R (r - A z), z the cycle's pre-smoothing of A z = r: the
        residual the two-level cycle hands the network, on the coarse
        operator's DoFs (needs set_mode(k > 0) and set_coarse)."""
        n_c = self._coarse.m() if getattr(self, "_coarse", None) is not None else rho_c.numel()
        check(lib().mfx_nn_restricted_residual(self._h, op._h, _ptr(r, F64, op.m()),
                                               _ptr(rho_c, F64, n_c), _stream()))

    def set_fallback(self, mg, window: int = 8, min_reduction: float = 2.0) -> None:
        """The multigrid (a PreconditionMG whose operator() is the solve's
        operator) flexible CG switches to when the learned cycle stagnates:
        the residual falls by less than min_reduction over `window` steps,
        the network's correction is rejected `window` times in a row, or a
        direction breaks down.  None: no fallback."""
        self._fallback = mg  # keep the multigrid alive while in use
        check(lib().mfx_nn_set_fallback(self._h, None if mg is None else mg._h, window,
                                        min_reduction))

    @property
    def fallback_step(self) -> int:
        """Step at which the last solve switched to the fallback (-1: it did not)."""
        return int(lib().mfx_nn_fallback_step(self._h))


def batch_instances(count: int = 64, seed: int = 2, n_min: int = 4, n_max: int = 32,
                    degrees=(2, 3, 4, 5, 6)):
    """Instance list of the batched-solve configuration: cells per direction
    log-uniform in [n_min, n_max] (rounded), degree uniform over `degrees`,
    from the SplitMix64 stream of `seed` (u = (z >> 11) 2^-53 in [0,1))."""
    import math
    M = (1 << 64) - 1
    out, k = [], 0

    def u01():
        nonlocal k
        k += 1
        z = (seed + k * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        return (z >> 11) * 2.0 ** -53
    for _ in range(count):
        n = int(round(math.exp(math.log(n_min) + u01() * (math.log(n_max) - math.log(n_min)))))
        p = degrees[min(int(u01() * len(degrees)), len(degrees) - 1)]
        out.append((p, n))
    return out


# --- kernel-level plug point -------------------------------------------
def bk5_apply(degree, D, G, src, dst, map_, n_cells):
    """dst += sum_cells P^T D^T G D P src (GLL collocation), no zeroing."""
    fn = lib().mfx_bk5_apply_f64 if src.dtype == torch.float64 else lib().mfx_bk5_apply_f32
    t, nm3 = src.dtype, (degree + 1) ** 3
    check(fn(degree, _ptr(D, t), _ptr(G, t, 6 * nm3 * n_cells), _ptr(src, t),
             _ptr(dst, t, src.numel()),
             _ptr(map_, I32, nm3 * n_cells), n_cells, _stream()))


def bk3_apply(degree, nq, B, D, G, src, dst, map_, n_cells):
    fn = lib().mfx_bk3_apply_f64 if src.dtype == torch.float64 else lib().mfx_bk3_apply_f32
    t = src.dtype
    check(fn(degree, nq, _ptr(B, t, nq * (degree + 1)), _ptr(D, t, nq * nq),
             _ptr(G, t, 6 * nq ** 3 * n_cells),
             _ptr(src, t), _ptr(dst, t, src.numel()), _ptr(map_, I32, (degree + 1) ** 3 * n_cells),
             n_cells, _stream()))


def evec_apply(degree, nq, collocation, B, D, G, inp, out, n_cells):
    if collocation:
        check(lib().mfx_bk5_evec_f64(degree, _ptr(D, F64), _ptr(G, F64), _ptr(inp, F64),
                                     _ptr(out, F64), n_cells,
                                     _stream()))
    else:
        check(lib().mfx_bk3_evec_f64(degree, nq, _ptr(B, F64), _ptr(D, F64), _ptr(G, F64),
                                     _ptr(inp, F64), _ptr(out, F64),
                                     n_cells, _stream()))


def gather(map_, n_cells, nodes_per_cell, src):
    """E-vector of src through the DoF map (invalid entries -> 0)."""
    ev = torch.empty(n_cells * nodes_per_cell, dtype=torch.float64, device="cuda")
    check(lib().mfx_gather_f64(_ptr(map_, I32, n_cells * nodes_per_cell), n_cells, nodes_per_cell,
                               _ptr(src, F64),
                                _ptr(ev, F64, n_cells * nodes_per_cell), _stream()))
    return ev


def scatter_add(map_, n_cells, nodes_per_cell, evec, dst):
    check(lib().mfx_scatter_add_f64(_ptr(map_, I32, n_cells * nodes_per_cell), n_cells,
                                    nodes_per_cell,
                                    _ptr(evec, F64, n_cells * nodes_per_cell), _ptr(dst, F64),
                                    _stream()))


def sine_modes(coords: torch.Tensor, wave_numbers, amplitudes, out: torch.Tensor) -> None:
    """out_i = sum_m a_m sin(pi k_m0 x_i) sin(pi k_m1 y_i) sin(pi k_m2 z_i) at
    the points coords [n, 3] (mfx_vec_sine_modes); wave_numbers: 3 integers
    per mode, amplitudes: one per mode (at most 64 modes)."""
    import numpy as np
    k = np.ascontiguousarray(wave_numbers, dtype=np.int32)
    a = np.ascontiguousarray(amplitudes, dtype=np.float64)
    if k.size != 3 * a.size:
        raise ValueError("three wave numbers per amplitude")
    n = out.numel()
    check(lib().mfx_vec_sine_modes(_ptr(coords, F64, 3 * n), n, a.size, k.ctypes.data,
                                   a.ctypes.data, _ptr(out, F64, n), _stream()))


def random_vector(n: int, seed: int) -> torch.Tensor:
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    check(lib().mfx_vec_random(_ptr(v, F64, n), n, seed, _stream()))
    return v
