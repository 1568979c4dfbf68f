"""GPU parity of the operator apply against the CPU oracle.

Tolerances: DoF maps and gathered values bit-exact; FP64 applies <= 1e-12
relative (2-norm) to the oracle; FP32 applies <= 1e-5 relative."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _gpu_problem(p, nq, quad, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), deform=0.0,
                 with_fp32=False):
    from benchmarks_amd.laplace import LaplaceOperator
    return LaplaceOperator(p, nq, quad, n_cells=n, lo=lo, hi=hi, deform=deform,
                           with_fp32=with_fp32)


def _map_of(op):
    E = op.n_cells()
    nm3 = (op.degree + 1) ** 3
    return op.device_array("map", torch.int32, E * nm3).cpu().numpy().view(np.uint32)


def _G_of(op):
    E = op.n_cells()
    return op.device_array("G", torch.float64, E * 6 * op.n_q_points_1d ** 3).cpu().numpy()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = ([(p, p + 1, O.GLL) for p in range(1, 9)] +
         [(p, p + 2, O.GAUSS) for p in range(1, 9)] +
         [(p, p + 1, O.GAUSS) for p in (1, 2, 4, 6, 8)])


@pytest.mark.parametrize("p,nq,quad", CASES)
def test_vmult_deformed_matches_oracle(gpu, p, nq, quad):
    n = (3, 4, 5)  # ragged: the last workgroup of cells is partial
    print(f"case p={p} nq={nq} quad={quad} n={n}", flush=True)
    op = _gpu_problem(p, nq, quad, n, deform=0.05)
    P = O.make_problem(p, nq, quad, n, deform=0.05)
    assert op.m() == P.N and op.n_cells() == P.E
    # DoF map: bit-exact
    assert np.array_equal(_map_of(op), P.map)
    # geometry from GPU-generated vertices vs the oracle's
    G = _G_of(op)
    assert _rel(G, P.G) < 1e-13
    x = O.splitmix64_uniform(1, P.N)
    xt = torch.from_numpy(x).cuda()
    y = op.initialize_dof_vector()
    op.vmult(y, xt)
    yg = y.cpu().numpy()
    ref = P.vmult(x)
    assert _rel(yg, ref) <= TOL64, _rel(yg, ref)
    # identity rows on constrained DoFs: bit-exact copy
    cons = P.constrained.astype(bool)
    assert np.array_equal(yg[cons], x[cons])
    # same geometry array in both: isolates the kernel arithmetic
    P.G = G
    assert _rel(yg, P.vmult(x)) <= 1e-13


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("n", [(3, 4, 5), (1, 1, 1), (5, 2, 3)])
def test_vmult_on_the_fly_geometry_matches_oracle(gpu, p, n):
    """Patch kernel with the metric computed from the cell vertices: same
    operator as the oracle's stored-G one; meshes whose cell counts are not
    multiples of the patch size leave partial patches."""
    from benchmarks_amd.laplace import LaplaceOperator
    print(f"case p={p} n={n} on-the-fly", flush=True)
    op = LaplaceOperator(p, p + 1, O.GLL, n_cells=n, deform=0.05, geometry=1)
    P = O.make_problem(p, p + 1, O.GLL, n, deform=0.05)
    assert op.info.n_patches > 0 and op.info.G is None
    x = O.splitmix64_uniform(1, P.N)
    y = op.initialize_dof_vector()
    op.vmult(y, torch.from_numpy(x).cuda())
    yg = y.cpu().numpy()
    assert _rel(yg, P.vmult(x)) <= TOL64
    cons = P.constrained.astype(bool)
    assert np.array_equal(yg[cons], x[cons])


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("n", [(3, 4, 5), (1, 1, 1)])
@pytest.mark.parametrize("geometry", [1, 0])
def test_deterministic_scatter_matches_oracle_bit_for_bit_on_repeat(gpu, p, n, geometry):
    """Deterministic scatter (mfx_op_set_deterministic): repeated applies are
    bitwise identical (FP64 and FP32); they agree with the oracle to 1e-12
    and with the atomic scatter to 1e-13 (only the summation order differs);
    vmult_add adds to dst; switching the mode off restores the atomic form.
    Meshes whose cell counts are not multiples of the patch size leave
    partial patches, whose nodes the gather must still find.  Both geometry
    modes (metric from the vertices, stored metric terms)."""
    from benchmarks_amd.laplace import LaplaceOperator
    print(f"case p={p} n={n} geometry={geometry} deterministic", flush=True)
    op = LaplaceOperator(p, p + 1, O.GLL, n_cells=n, deform=0.05, geometry=geometry,
                         with_fp32=True)
    P = O.make_problem(p, p + 1, O.GLL, n, deform=0.05)
    x = torch.from_numpy(O.splitmix64_uniform(1, P.N)).cuda()
    ya = op.initialize_dof_vector()
    op.vmult(ya, x)
    op.set_deterministic(True)
    y1, y2 = op.initialize_dof_vector(), op.initialize_dof_vector()
    op.vmult(y1, x)
    op.vmult(y2, x)
    assert torch.equal(y1, y2)
    ref = P.vmult(x.cpu().numpy())
    assert _rel(y1.cpu().numpy(), ref) <= TOL64
    assert (torch.linalg.norm(y1 - ya) / torch.linalg.norm(ya)).item() <= 1e-13
    z = torch.ones_like(x)
    op.vmult_add(z, x)
    cons = torch.from_numpy(P.constrained.astype(bool)).cuda()
    assert torch.equal(z[cons], torch.ones_like(z[cons]))
    if (~cons).any():  # a single cell has no unconstrained DoF
        assert ((z - 1.0 - y1)[~cons].abs().max() / y1.abs().max()).item() <= 1e-15
    f1, f2 = (torch.zeros(op.m(), dtype=torch.float32, device="cuda") for _ in range(2))
    op.vmult(f1, x.float())
    op.vmult(f2, x.float())
    assert torch.equal(f1, f2)
    assert _rel(f1.double().cpu().numpy(), ref) <= 1e-5
    op.set_deterministic(False)
    y3 = op.initialize_dof_vector()
    op.vmult(y3, x)
    assert (torch.linalg.norm(y3 - y1) / torch.linalg.norm(y1)).item() <= 1e-13


@pytest.mark.parametrize("p,n", [(1, 64), (2, 48), (3, 40), (4, 40), (5, 28), (6, 24), (7, 18),
                                 (8, 16)])
def test_patch_kernel_over_many_patches_per_workgroup(gpu, p, n):
    """Meshes with several times more patches than the persistent patch
    kernel has workgroups, so that every workgroup walks many patches
    through its software pipeline (the small meshes of the other oracle
    tests fit one patch per workgroup): the apply with the metric from the
    vertices and the one with stored metric terms both agree with the oracle
    to 1e-12, and two applies agree to 1e-13 (FP64) and 1e-5 (FP32 against
    FP64)."""
    from benchmarks_amd.laplace import LaplaceOperator, random_vector
    print(f"case p={p} n={n} many patches", flush=True)
    x = None
    for geometry in (1, 0):
        op = LaplaceOperator(p, p + 1, O.GLL, n_cells=n, deform=0.05, geometry=geometry,
                             with_fp32=True)
        assert op.info.n_patches >= 3 * 256
        if x is None:
            x = random_vector(op.m(), 4)
            P = O.make_problem(p, p + 1, O.GLL, n, deform=0.05)
            ref = P.vmult(x.cpu().numpy())
        y1, y2 = op.initialize_dof_vector(), op.initialize_dof_vector()
        op.vmult(y1, x)
        op.vmult(y2, x)
        assert (torch.linalg.norm(y2 - y1) / torch.linalg.norm(y1)).item() <= 1e-13
        assert _rel(y1.cpu().numpy(), ref) <= TOL64
        f = torch.zeros(op.m(), dtype=torch.float32, device="cuda")
        op.vmult(f, x.float())
        assert (torch.linalg.norm(f.double() - y1) / torch.linalg.norm(y1)).item() <= 1e-5
        del op


def test_deterministic_scatter_needs_the_patch_kernel(gpu):
    from benchmarks_amd._lib import MfxError
    from benchmarks_amd.laplace import LaplaceOperator
    op = LaplaceOperator(3, 5, O.GAUSS, n_cells=(2, 2, 2))
    with pytest.raises(MfxError):
        op.set_deterministic(True)


@pytest.mark.parametrize("p,nq", [(p, p + 2) for p in range(1, 9)] +
                         [(p, p + 1) for p in (1, 2, 4, 6, 8)])
def test_gauss_on_the_fly_geometry_matches_oracle(gpu, p, nq):
    """Gauss (BK3) kernel with the metric terms computed per quadrature point
    from the cell vertices: the oracle's stored-G operator to 1e-12 (FP64)
    and 1e-5 (FP32) on a deformed, ragged mesh; identity rows bit-exact; the
    same operator built from host arrays (mfx_op_create) agrees."""
    from benchmarks_amd.laplace import LaplaceOperator
    n = (3, 4, 5)
    print(f"case p={p} nq={nq} Gauss on-the-fly", flush=True)
    op = LaplaceOperator(p, nq, O.GAUSS, n_cells=n, deform=0.05, geometry=1, with_fp32=True)
    P = O.make_problem(p, nq, O.GAUSS, n, deform=0.05)
    assert op.info.n_patches == 0 and op.info.G is None
    x = O.splitmix64_uniform(1, P.N)
    xt = torch.from_numpy(x).cuda()
    y = op.initialize_dof_vector()
    op.vmult(y, xt)
    yg, ref = y.cpu().numpy(), P.vmult(x)
    assert _rel(yg, ref) <= TOL64, _rel(yg, ref)
    cons = P.constrained.astype(bool)
    assert np.array_equal(yg[cons], x[cons])
    y32 = torch.zeros(op.m(), dtype=torch.float32, device="cuda")
    op.vmult(y32, xt.float())
    assert _rel(y32.double().cpu().numpy(), ref) <= 1e-5
    ga = LaplaceOperator.from_arrays(p, nq, O.GAUSS, P.map, P.cellX,
                                     np.nonzero(P.constrained)[0].astype(np.uint32), P.N,
                                     geometry=1)
    y2 = ga.initialize_dof_vector()
    ga.vmult(y2, xt)
    assert _rel(y2.cpu().numpy(), ref) <= TOL64


@pytest.mark.parametrize("p,nq", [(p, p + 2) for p in range(1, 9)] +
                         [(p, p + 1) for p in (2, 3, 4, 6, 8)])
def test_gauss_on_the_fly_over_full_row_segments(gpu, p, nq):
    """x-rows of 17 cells: every row holds full segments of the cells a
    workgroup takes (7 at nq = 6, 2 at nq = 10) and a partial last one, so
    the row-segment sums of shared x-faces and the line-form contractions run
    on whole segments, which the small ragged mesh above does not reach; FP64
    to 1e-12 and FP32 to 1e-5 of the oracle, identity rows bit-exact."""
    from benchmarks_amd.laplace import LaplaceOperator
    n = (17, 3, 2)
    print(f"case p={p} nq={nq} Gauss on-the-fly, rows of 17 cells", flush=True)
    op = LaplaceOperator(p, nq, O.GAUSS, n_cells=n, deform=0.05, geometry=1, with_fp32=True)
    P = O.make_problem(p, nq, O.GAUSS, n, deform=0.05)
    x = O.splitmix64_uniform(3, P.N)
    xt = torch.from_numpy(x).cuda()
    y = op.initialize_dof_vector()
    op.vmult(y, xt)
    yg, ref = y.cpu().numpy(), P.vmult(x)
    assert _rel(yg, ref) <= TOL64, _rel(yg, ref)
    cons = P.constrained.astype(bool)
    assert np.array_equal(yg[cons], x[cons])
    y32 = torch.zeros(op.m(), dtype=torch.float32, device="cuda")
    op.vmult(y32, xt.float())
    assert _rel(y32.double().cpu().numpy(), ref) <= 1e-5


@pytest.mark.parametrize("p,nq", [(1, 3), (2, 4), (4, 5), (4, 6), (6, 8), (8, 10)])
def test_gauss_on_the_fly_undeformed_sheared_and_mixed_cells(gpu, p, nq):
    """On an undeformed box the on-the-fly Gauss kernel forms the diagonal
    cofactors and det J once per column (the cells' Jacobian columns lie
    along x, y and z); every other mesh forms the metric per point.  The box
    (row segments), a sheared box from host arrays (affine cells, none
    axis-aligned; dyadic coordinates and shear, so the cells are affine to
    the last bit) and a mesh from host arrays in which every third box cell
    has a moved vertex each match the oracle: FP64 to 1e-12, FP32 to 1e-5."""
    import dataclasses
    from benchmarks_amd.laplace import LaplaceOperator
    n = (6, 4, 3)
    print(f"case p={p} nq={nq} Gauss on-the-fly, undeformed, sheared and mixed cells", flush=True)
    P = O.make_problem(p, nq, O.GAUSS, n, lo=(-1.0, -0.5, 0.0), hi=(2.8, 1.5, 0.9))
    x = O.splitmix64_uniform(4, P.N)
    xt = torch.from_numpy(x).cuda()
    box = LaplaceOperator(p, nq, O.GAUSS, n_cells=n, lo=(-1.0, -0.5, 0.0), hi=(2.8, 1.5, 0.9),
                          geometry=1, with_fp32=True)
    cellX = P.cellX.copy().reshape(P.E, 8, 3)
    cellX[::3, 7] += np.array([0.03, -0.05, 0.02])
    cellX = cellX.reshape(-1)
    G = np.zeros_like(P.G)
    O.lib().orc_geometry(P.E, cellX, nq, P.xq, P.wq, G)
    Pm = dataclasses.replace(P, cellX=cellX, G=G)
    cons = np.nonzero(P.constrained)[0].astype(np.uint32)
    mixed = LaplaceOperator.from_arrays(p, nq, O.GAUSS, P.map, cellX, cons, P.N,
                                        with_fp32=True, geometry=1)
    for op, ref in ((box, P.vmult(x)), (mixed, Pm.vmult(x))):
        y = op.initialize_dof_vector()
        op.vmult(y, xt)
        assert _rel(y.cpu().numpy(), ref) <= TOL64, _rel(y.cpu().numpy(), ref)
        y32 = torch.zeros(op.m(), dtype=torch.float32, device="cuda")
        op.vmult(y32, xt.float())
        assert _rel(y32.double().cpu().numpy(), ref) <= 1e-5
    assert _rel(Pm.vmult(x), P.vmult(x)) > 1e-4  # the moved vertices change the operator
    # sheared box: X -> S X + t on the dyadic unit-cube grid of 4 x 4 x 2 cells
    Q = O.make_problem(p, nq, O.GAUSS, (4, 4, 2))
    S = np.array([[1.0, 0.25, 0.125], [0.125, 1.0, 0.25], [0.25, -0.125, 1.0]])
    V = Q.cellX.reshape(Q.E, 8, 3) @ S.T + np.array([0.5, -0.25, 1.0])
    a1 = V[:, 3] - V[:, 2] - V[:, 1] + V[:, 0]
    k1 = V[:, 5] - V[:, 1] - V[:, 4] + V[:, 0]
    k2 = V[:, 6] - V[:, 2] - V[:, 4] + V[:, 0]
    assert not a1.any() and not k1.any() and not k2.any()  # affine to the last bit
    assert (V[:, 1] - V[:, 0])[:, 1:].all()  # and not axis-aligned
    cellXs = V.reshape(-1)
    Gs = np.zeros_like(Q.G)
    O.lib().orc_geometry(Q.E, cellXs, nq, Q.xq, Q.wq, Gs)
    Qs = dataclasses.replace(Q, cellX=cellXs, G=Gs)
    sheared = LaplaceOperator.from_arrays(p, nq, O.GAUSS, Q.map, cellXs,
                                          np.nonzero(Q.constrained)[0].astype(np.uint32), Q.N,
                                          with_fp32=True, geometry=1)
    xs = O.splitmix64_uniform(5, Q.N)
    ref = Qs.vmult(xs)
    assert _rel(ref, Q.vmult(xs)) > 1e-3  # the shear changes the operator
    y = sheared.initialize_dof_vector()
    sheared.vmult(y, torch.from_numpy(xs).cuda())
    assert _rel(y.cpu().numpy(), ref) <= TOL64, _rel(y.cpu().numpy(), ref)
    y32 = torch.zeros(sheared.m(), dtype=torch.float32, device="cuda")
    sheared.vmult(y32, torch.from_numpy(xs).cuda().float())
    assert _rel(y32.double().cpu().numpy(), ref) <= 1e-5


@pytest.mark.parametrize("p", [2, 4, 6, 8])
def test_patch_map_covers_every_dof(gpu, p):
    """Patch DoF map: every unconstrained DoF appears, constrained ones never;
    bit 31 marks exactly the DoFs seen by more than one patch."""
    from benchmarks_amd.laplace import LaplaceOperator
    op = LaplaceOperator(p, p + 1, O.GLL, n_cells=(5, 4, 3), geometry=1)
    P = O.make_problem(p, p + 1, O.GLL, (5, 4, 3))
    bx, by, bz = op.info.patch_cells
    P3 = (bx * p + 1) * (by * p + 1) * (bz * p + 1)
    from benchmarks_amd.laplace import _device_view
    pm = _device_view(op.info.patch_map, torch.int32, op.info.n_patches * P3).cpu().numpy()
    pm = pm.view(np.uint32)
    valid = pm[pm != O.INVALID]
    idx = valid & 0x7FFFFFFF
    count = np.bincount(idx, minlength=P.N)
    assert np.array_equal(count > 0, P.constrained == 0)
    shared = np.zeros(P.N, bool)
    shared[idx[(valid & 0x80000000) != 0]] = True
    assert np.array_equal(shared, count > 1)


@pytest.mark.parametrize("p,nq,quad", [(2, 3, O.GLL), (4, 5, O.GLL), (4, 6, O.GAUSS),
                                       (8, 9, O.GLL)])
def test_gather_scatter_bit_exact(gpu, p, nq, quad):
    """Gather through the GPU's DoF map is bit-exact against the oracle's
    gather; scatter-add of an integer-valued E-vector is exact too."""
    from benchmarks_amd.laplace import gather, scatter_add
    op = _gpu_problem(p, nq, quad, (2, 3, 2))
    P = O.make_problem(p, nq, quad, (2, 3, 2))
    mp = op.device_array("map", torch.int32, P.E * (p + 1) ** 3)
    x = O.splitmix64_uniform(7, P.N)
    ev = gather(mp, P.E, (p + 1) ** 3, torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(ev, P.gather(x))
    ones = torch.ones(P.E * (p + 1) ** 3, dtype=torch.float64, device="cuda")
    cnt = torch.zeros(P.N, dtype=torch.float64, device="cuda")
    scatter_add(mp, P.E, (p + 1) ** 3, ones, cnt)
    valid = P.map[P.map != O.INVALID]
    assert np.array_equal(cnt.cpu().numpy(), np.bincount(valid, minlength=P.N).astype(float))


@pytest.mark.parametrize("p", range(1, 9))
def test_bk5_evec_random_geometry(gpu, p):
    from benchmarks_amd.laplace import evec_apply
    nq = p + 1
    rng = np.random.default_rng(p)
    E = 37
    B, D, xq, wq = O.basis_tables(p, nq, O.GLL)
    G = rng.standard_normal(E * 6 * nq ** 3)
    u = rng.standard_normal(E * nq ** 3)
    ref = O.bk5_evec(nq, E, D, G, u)
    out = torch.zeros(E * nq ** 3, dtype=torch.float64, device="cuda")
    evec_apply(p, nq, True, None, torch.from_numpy(D).cuda(), torch.from_numpy(G).cuda(),
               torch.from_numpy(u).cuda(), out, E)
    assert _rel(out.cpu().numpy(), ref) <= TOL64


@pytest.mark.parametrize("p,nq", [(p, p + 2) for p in range(1, 9)] + [(p, p + 1) for p in range(1, 9)])
def test_bk3_evec_random_geometry(gpu, p, nq):
    from benchmarks_amd.laplace import evec_apply
    rng = np.random.default_rng(100 + p)
    E = 23
    nm = p + 1
    B, D, xq, wq = O.basis_tables(p, nq, O.GAUSS)
    G = rng.standard_normal(E * 6 * nq ** 3)
    u = rng.standard_normal(E * nm ** 3)
    ref = O.bk3_evec(nm, nq, E, B, D, G, u)
    out = torch.zeros(E * nm ** 3, dtype=torch.float64, device="cuda")
    evec_apply(p, nq, False, torch.from_numpy(B).cuda(), torch.from_numpy(D).cuda(),
               torch.from_numpy(G).cuda(), torch.from_numpy(u).cuda(), out, E)
    assert _rel(out.cpu().numpy(), ref) <= TOL64


def test_kernel_level_accumulates(gpu):
    """mfx_bk5_apply_f64 adds into dst (no zeroing): applying twice doubles."""
    from benchmarks_amd.laplace import bk5_apply
    p = 3
    op = _gpu_problem(p, p + 1, O.GLL, (3, 3, 3), deform=0.05)
    N, E = op.m(), op.n_cells()
    info = op.info
    D = op.device_array("D", torch.float64, (p + 1) ** 2)
    G = op.device_array("G", torch.float64, E * 6 * (p + 1) ** 3)
    mp = op.device_array("map", torch.int32, E * (p + 1) ** 3)
    x = torch.from_numpy(O.splitmix64_uniform(3, N)).cuda()
    d1 = torch.zeros(N, dtype=torch.float64, device="cuda")
    bk5_apply(p, D, G, x, d1, mp, E)
    d2 = d1.clone()
    bk5_apply(p, D, G, x, d2, mp, E)
    torch.cuda.synchronize()
    # atomics add in a different order each time: equal to rounding
    assert torch.allclose(d2, 2 * d1, rtol=1e-13, atol=1e-13 * float(d1.abs().max()))
    assert info.n_constrained > 0


def test_single_cell_and_all_constrained(gpu):
    """Edge cases: one cell (every DoF constrained for p=1), p=2 single cell
    (one free DoF)."""
    for p, nfree in [(1, 0), (2, 1)]:
        op = _gpu_problem(p, p + 1, O.GLL, (1, 1, 1))
        P = O.make_problem(p, p + 1, O.GLL, (1, 1, 1))
        x = O.splitmix64_uniform(5, P.N)
        y = op.initialize_dof_vector()
        op.vmult(y, torch.from_numpy(x).cuda())
        yg = y.cpu().numpy()
        assert int((P.constrained == 0).sum()) == nfree
        assert _rel(yg, P.vmult(x)) <= TOL64


@pytest.mark.parametrize("p,nq,quad,geom", [(2, 3, O.GLL, 0), (4, 5, O.GLL, 0), (4, 6, O.GAUSS, 0),
                                            (4, 5, O.GLL, 1), (8, 9, O.GLL, 1)])
def test_vmult_f32(gpu, p, nq, quad, geom):
    from benchmarks_amd.laplace import LaplaceOperator
    op = LaplaceOperator(p, nq, quad, n_cells=(4, 3, 5), deform=0.05, with_fp32=True,
                         geometry=geom)
    P = O.make_problem(p, nq, quad, (4, 3, 5), deform=0.05)
    x = O.splitmix64_uniform(11, P.N)
    y = op.initialize_dof_vector(torch.float32)
    op.vmult(y, torch.from_numpy(x).float().cuda())
    ref = P.vmult(x.astype(np.float32).astype(np.float64))
    assert _rel(y.cpu().numpy().astype(np.float64), ref) <= 1e-5


@pytest.mark.parametrize("p,nq,quad", [(2, 3, O.GLL), (4, 5, O.GLL), (4, 6, O.GAUSS),
                                       (8, 9, O.GLL)])
def test_diagonal_matches_oracle(gpu, p, nq, quad):
    op = _gpu_problem(p, nq, quad, (3, 2, 4), deform=0.05)
    P = O.make_problem(p, nq, quad, (3, 2, 4), deform=0.05)
    d = op.compute_diagonal().cpu().numpy()
    assert _rel(d, P.diagonal()) <= TOL64
    inv = op.compute_inverse_diagonal().cpu().numpy()
    assert _rel(inv, P.inverse_diagonal()) <= TOL64


@pytest.mark.parametrize("f_kind", [0, 1])
@pytest.mark.parametrize("p,nq,quad", [(2, 3, O.GLL), (4, 6, O.GAUSS)])
def test_rhs_matches_oracle(gpu, p, nq, quad, f_kind):
    op = _gpu_problem(p, nq, quad, (3, 4, 2), deform=0.05)
    P = O.make_problem(p, nq, quad, (3, 4, 2), deform=0.05)
    b = op.rhs(f_kind).cpu().numpy()
    assert _rel(b, P.rhs(f_kind)) <= TOL64


def test_random_vector_matches_generator(gpu):
    from benchmarks_amd.laplace import random_vector
    v = random_vector(10007, 1).cpu().numpy()
    assert np.array_equal(v, O.splitmix64_uniform(1, 10007))


def test_unstructured_create_matches_box(gpu):
    """mfx_op_create from host arrays (the general adapter path) gives the
    same operator as the box generator."""
    from benchmarks_amd.laplace import LaplaceOperator
    p, nq = 3, 4
    P = O.make_problem(p, nq, O.GLL, (2, 3, 2), deform=0.05)
    cons = np.nonzero(P.constrained)[0].astype(np.uint32)
    op = LaplaceOperator.from_arrays(p, nq, O.GLL, P.map, P.cellX, cons, P.N)
    x = O.splitmix64_uniform(2, P.N)
    y = op.initialize_dof_vector()
    op.vmult(y, torch.from_numpy(x).cuda())
    assert _rel(y.cpu().numpy(), P.vmult(x)) <= TOL64


def test_vector_type_mismatch_raises_before_launch(gpu):
    """The C entry points take typed pointers: a vector of the wrong element
    type (same length) would be read and written past its end, so the Python
    layer refuses it before any launch."""
    from benchmarks_amd.laplace import LaplaceOperator
    op = LaplaceOperator(2, 4, O.GAUSS, (3, 2, 2), with_fp32=True)
    x32 = torch.zeros(op.m(), dtype=torch.float32, device="cuda")
    x64 = torch.zeros(op.m(), dtype=torch.float64, device="cuda")
    with pytest.raises(TypeError):
        op.vmult_add(x32, x32)
    with pytest.raises(TypeError):
        op.vmult(x64, x32)
    with pytest.raises(TypeError):
        op.vmult(x32, x64)
    with pytest.raises(ValueError):
        op.zero_constrained(x64[:-1])


def test_invalid_arguments_raise(gpu):
    from benchmarks_amd._lib import MfxError
    from benchmarks_amd.laplace import LaplaceOperator
    with pytest.raises(MfxError):
        LaplaceOperator(9, 10, O.GLL, n_cells=2)
    with pytest.raises(MfxError):
        LaplaceOperator(4, 7, O.GAUSS, n_cells=2)
    # size limits, checked before anything is allocated: DoF indices are
    # 31-bit (1291^3 Q1 nodes > 2^31 - 1) and cell counts 32-bit
    with pytest.raises(MfxError, match="31-bit"):
        LaplaceOperator(1, 2, O.GLL, n_cells=1290)
    with pytest.raises(MfxError, match="too many cells"):
        LaplaceOperator(1, 2, O.GLL, n_cells=(2048, 2048, 512))
    with pytest.raises(MfxError, match="n_cells"):
        LaplaceOperator(1, 2, O.GLL, n_cells=(0, 4, 4))
    op = LaplaceOperator(2, 3, O.GLL, n_cells=2)
    x = op.initialize_dof_vector()
    with pytest.raises(MfxError):
        op.vmult(x, x)  # aliasing is rejected
    with pytest.raises(ValueError):
        op.vmult(torch.zeros(3, dtype=torch.float64, device="cuda"), x)
