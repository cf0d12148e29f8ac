"""Multigrid-preconditioned batched CG (TFM_SOLVE_PCG64_MG) on SIMP grid
systems, against the oracle's direct solve of every design and right-hand
side, and the SIMP loop driven by it against the oracle's loop.

The V-cycle only changes how fast CG converges, so the bar is the plain FP64
solver's: 1e-9 relative error against the direct solution for a 1e-12
residual.  Grid sizes cover a single level (the coarse inverse alone solves
the system), even 2:1 coarsening and odd element counts (one coarse element
spanning a single fine element), like the reference's own 20 x 10 and 26 x 13
grids."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def simp_batch(ctx, P, nx, ny, seed, emin=1e-9, rho_lo=0.01):
    import tptfem as T

    m = T.Mesh.grid_batch(ctx, P, nx, ny, 1.0)
    ne = (nx - 1) * (ny - 1)
    rng = np.random.default_rng(seed)
    rho = rng.uniform(rho_lo, 1.0, P * ne)
    nd1 = 2 * nx * ny
    bcs = [(2 * (p * nx * ny + j * nx) + c, 0.0) for p in range(P) for j in range(ny) for c in (0, 1)]
    op = T.Operator.simp(rho, nu=0.3, penal=3.0, e0=1.0, emin=emin, h=1.0)
    s = T.System.create(m, op, 2, bcs, None, flags=T.SYS_FP64)
    rows = rng.integers(0, ny, size=(P, 8))
    f = np.zeros((8, P * nd1))
    for p in range(P):
        for k in range(8):
            f[k, 2 * (p * nx * ny + rows[p, k] * nx + nx - 1) + 1] = -1.0
            f[k, 2 * (p * nx * ny + rng.integers(0, nx * ny))] += 0.5 * (k + 1)
    return m, s, rho, f, np.arange(P + 1) * nd1


def oracle_solution(nx, ny, rho, emin, f):
    coords, conn = O.grid(nx, ny, 1.0)
    coo = O.assemble(coords, conn, O.make_op(O.QUAD, 2, 2, rho=rho, penal=3.0, e0=1.0, emin=emin,
                                             ke0=O.simp_ke_unit(0.3, 1.0)), 2)
    lb = [(2 * j * nx + c, 0.0) for j in range(ny) for c in (0, 1)]
    return O.solve_with_dirichlet(coo, f, lb)


@pytest.mark.parametrize("grid", [(9, 5), (17, 9), (20, 10), (26, 13), (33, 17)])
@pytest.mark.parametrize("nrhs", [1, 8])
def test_mg_batch_matches_direct(ctx, grid, nrhs):
    import tptfem as T

    nx, ny = grid
    P = 3
    m, s, rho, f, off = simp_batch(ctx, P, nx, ny, seed=nx * 31 + ny)
    ne, nd1 = (nx - 1) * (ny - 1), 2 * nx * ny
    u = s.solve_batch(off, rhs=f[:nrhs] if nrhs > 1 else f[0], nrhs=nrhs, mode=T.SOLVE_PCG64_MG, rtol=1e-12,
                      max_iter=5000)
    st = s.batch_stats
    # stopped on the recurrence residual; the reported true residual drifts a little above it
    assert all(x[2] <= 1e-11 for x in st), st
    for p in range(P):
        for k in range(nrhs):
            fp = f[k, p * nd1:(p + 1) * nd1]
            ur = oracle_solution(nx, ny, rho[p * ne:(p + 1) * ne], 1e-9, fp)
            up = u[k, p * nd1:(p + 1) * nd1]
            assert np.linalg.norm(up - ur) <= 1e-9 * np.linalg.norm(ur), (p, k)
    s.close()
    m.close()


def test_mg_cuts_iterations(ctx):
    """Jacobi PCG iterations grow with the grid; the V-cycle keeps them at a
    few tens on a 65 x 33 grid with a 1e6 stiffness contrast."""
    import tptfem as T

    m, s, rho, f, off = simp_batch(ctx, 2, 65, 33, seed=5)
    uj = s.solve_batch(off, rhs=f[0], nrhs=1, mode=T.SOLVE_PCG64, rtol=1e-10, max_iter=20000)
    itj = max(x[0] for x in s.batch_stats)
    um = s.solve_batch(off, rhs=f[0], nrhs=1, mode=T.SOLVE_PCG64_MG, rtol=1e-10, max_iter=2000)
    itm = max(x[0] for x in s.batch_stats)
    assert itm <= 100 and itm * 8 < itj, (itm, itj)
    assert np.linalg.norm(um - uj) <= 1e-8 * np.linalg.norm(uj)
    s.close()
    m.close()


def test_mg_rejects_other_systems(ctx):
    """Only SIMP systems on a grid_batch mesh, one problem per design."""
    import tptfem as T

    m, noff, eoff = T.Mesh.box_batch(ctx, [(2, 2, 2), (3, 2, 2)])
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, [(0, 0.0)], T.source_vector(m, 1.0), flags=T.SYS_FP64)
    with pytest.raises(T.TfmError) as ei:
        s.solve_batch(noff, mode=T.SOLVE_PCG64_MG)
    assert ei.value.code == T.E_UNSUPPORTED
    s.close()
    m.close()
    m, s, rho, f, off = simp_batch(ctx, 2, 9, 5, seed=1)
    with pytest.raises(T.TfmError) as ei:  # two designs as one problem
        s.solve_batch(np.array([0, off[-1]]), rhs=f[0], mode=T.SOLVE_PCG64_MG)
    assert ei.value.code == T.E_UNSUPPORTED
    s.close()
    m.close()


@pytest.mark.parametrize("case", [(20, 10, 0.5, 30), (26, 13, 0.4, 20)])
def test_simp_loop_with_mg_matches_oracle(ctx, case):
    """The SIMP loop with multigrid state solves gives the oracle's densities
    and compliance history (same bar as the Jacobi-PCG loop)."""
    import tptfem as T

    nx, ny, vf, iters = case
    f, bcs = O.cantilever_load(nx, ny, 1.0)
    rr, cr, _ = O.topopt_simp(nx, ny, 1.0, 1.0, 0.3, vf, 3.0, 2.0, iters, 0.2, f, bcs)
    g = T.topopt_simp(ctx, nx, ny, 1.0, 1.0, 0.3, vf, 3.0, 2.0, iters, 0.2, f, bcs, mode=T.SOLVE_PCG64_MG)
    assert np.abs(g.densities - rr).max() <= 1e-6
    np.testing.assert_allclose(g.compliance, cr, rtol=1e-7, atol=0)


def test_single_system_solve_with_mg(ctx):
    """System.solve with the multigrid mode on one design (a grid_batch mesh
    of one grid) equals the direct solution, Dirichlet values in place."""
    import tptfem as T

    m, s, rho, f, off = simp_batch(ctx, 1, 26, 13, seed=3)
    ur = oracle_solution(26, 13, rho, 1e-9, f[0])
    u = s.solve_batch(off, rhs=f[0], nrhs=1, mode=T.SOLVE_PCG64_MG, rtol=1e-12)[0]
    assert np.linalg.norm(u - ur) <= 1e-9 * np.linalg.norm(ur)
    s.close()
    m.close()
    # the system's own right-hand side through System.solve
    m = T.Mesh.grid_batch(ctx, 1, 26, 13, 1.0)
    fo, bcs = O.cantilever_load(26, 13, 1.0)
    op = T.Operator.simp(rho, nu=0.3, penal=3.0, e0=1.0, emin=1e-9, h=1.0)
    s = T.System.create(m, op, 2, bcs, fo, flags=T.SYS_FP64)
    u = s.solve(mode=T.SOLVE_PCG64_MG, rtol=1e-12)
    ur = oracle_solution(26, 13, rho, 1e-9, fo)
    assert np.linalg.norm(u - ur) <= 1e-9 * np.linalg.norm(ur) and s.stats.iterations < 200
    s.close()
    m.close()


# ---- box meshes (3-D multigrid) ---------------------------------------------

def box_system(ctx, n, dpn, seed, jitter=0.2, faces=((0, 0.0),), flags=None):
    """A jittered box_mesh system: Poisson (dpn 1, kappa_e = exp(N(0,1)),
    f = 1) or elasticity (dpn 3, E_e log-uniform [0.1, 10], nu 0.3, traction
    (0, 0, -1e-3) on x = 1), DOFs on the listed (axis, coord) faces fixed;
    also returns the oracle's system for the direct solve."""
    import tptfem as T

    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    if jitter:
        m.jitter(n, jitter, seed)
    coords, conn = m.download()
    rng = np.random.default_rng(seed)
    fixed = sorted({int(v) for axis, val in faces for v in O.nodes_on_plane(coords, axis, val, 1e-9)})
    bcs = [(dpn * v + c, 0.0) for v in fixed for c in range(dpn)]
    if dpn == 1:
        coef = np.exp(rng.standard_normal(len(conn)))
        op = T.Operator.poisson(1.0, 2, kappa_e=coef)
        rhs = O.poisson_rhs(coords, O.TET, conn, 1.0)
        oop = O.make_op(O.TET, 3, 0, coef_e=coef)
    else:
        coef = 10.0 ** rng.uniform(-1, 1, len(conn))
        op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=coef)
        rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
        oop = O.make_op(O.TET, 3, 1, coef_e=coef, nu=0.3)
    s = T.System.create(m, op, dpn, bcs, rhs, flags=T.SYS_FP64 if flags is None else flags)
    return m, s, (coords, conn, oop, rhs, bcs)


@pytest.mark.parametrize("case", [((2, 2, 2), 3), ((6, 5, 4), 1), ((9, 7, 4), 1), ((6, 6, 6), 3), ((7, 5, 9), 3),
                                  ((12, 1, 10), 1), ((10, 9, 1), 3)])
def test_mg_box_matches_direct(ctx, case):
    """One level (the coarse inverse solves the system), even and odd
    coarsening on all three axes, an axis of one cell (never coarsened; with
    two nodes on a fast axis the stencil conversion decodes neighbours by
    division), Poisson and elasticity: the multigrid CG solution equals the
    oracle's dense-LU solve to 1e-9."""
    import tptfem as T

    n, dpn = case
    m, s, (coords, conn, oop, rhs, bcs) = box_system(ctx, n, dpn, seed=sum(n) * dpn)
    u = s.solve(T.SOLVE_PCG64_MG, rtol=1e-12, max_iter=2000)
    ur = O.solve_with_dirichlet(O.assemble(coords, conn, oop, dpn), rhs, bcs)
    assert np.linalg.norm(u - ur) <= 1e-9 * np.linalg.norm(ur), n
    assert s.stats.rel_residual <= 1e-11
    s.close()
    m.close()


def test_mg_box_iterations_do_not_grow(ctx):
    """Jacobi PCG needs ~2x the iterations per refinement; the V-cycle keeps
    the count flat (elasticity, 8^3 and 16^3 jittered boxes)."""
    import tptfem as T

    its = []
    for n in (8, 16):
        m, s, _ = box_system(ctx, (n, n, n), 3, seed=n)
        uj = s.solve(T.SOLVE_PCG64, rtol=1e-10, max_iter=20000)
        itj = s.stats.iterations
        um = s.solve(T.SOLVE_PCG64_MG, rtol=1e-10, max_iter=2000)
        itm = s.stats.iterations
        assert np.linalg.norm(um - uj) <= 1e-8 * np.linalg.norm(uj)
        its.append((itm, itj))
        s.close()
        m.close()
    (m8, j8), (m16, j16) = its
    assert m16 <= 60 and m16 <= m8 + 16 and 4 * m16 < j16, its


def test_mg_box_needs_box_mesh(ctx):
    """A box system built from host arrays carries no node grid: refused."""
    import tptfem as T

    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (3, 3, 3))
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    bcs = [(int(v), 0.0) for v in O.nodes_on_plane(coords, 0, 0.0, 1e-9)]
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, bcs, O.poisson_rhs(coords, O.TET, conn, 1.0),
                        flags=T.SYS_FP64)
    with pytest.raises(T.TfmError) as ei:
        s.solve(T.SOLVE_PCG64_MG)
    assert ei.value.code == T.E_UNSUPPORTED
    s.close()
    m.close()


def test_solve_drivers_with_mg(ctx):
    """solve_poisson and solve_elasticity on a box mesh with the multigrid
    mode: the reference's drivers' results (oracle dense LU) to 1e-10."""
    import tptfem as T

    n = (6, 5, 7)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    coords, conn = m.download()
    faces = [(a, v) for a in range(3) for v in (0.0, 1.0)]
    bcs = [(int(i), 0.0) for a, v in faces for i in O.nodes_on_plane(coords, a, v, 1e-9)]
    u = T.solve_poisson(m, 1.0, 2, 1.0, bcs, rtol=1e-13, mode=T.SOLVE_PCG64_MG)
    ur = O.solve_poisson(coords, O.TET, conn, 1.0, 2, 1.0, bcs)
    assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()
    clamp = [(int(i) * 3 + c, 0.0) for i in O.nodes_on_plane(coords, 0, 0.0, 1e-9) for c in range(3)]
    u = T.solve_elasticity(m, T.CONTINUUM3D, 1.0, 0.3, 2, [0.0, 0.0, -1.0], clamp, rtol=1e-13,
                           mode=T.SOLVE_PCG64_MG)
    bq = np.tile([0.0, 0.0, -1.0], (len(conn), 4, 1))
    ur = O.solve_elasticity(coords, O.TET, conn, O.CONT3D, 1.0, 0.3, 2, bq, clamp)
    assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()
    m.close()


@pytest.mark.parametrize("case", [((6, 6, 6), 3), ((9, 7, 4), 1)])
def test_mg_box_with_fp32_values(ctx, case):
    """A system that also holds FP32 values runs the V-cycle's fine passes on
    them; the CG stays FP64, so the solution is still the direct one to 1e-9."""
    import tptfem as T

    n, dpn = case
    m, s, (coords, conn, oop, rhs, bcs) = box_system(ctx, n, dpn, seed=5 * dpn, flags=T.SYS_FP64 | T.SYS_FP32)
    u = s.solve(T.SOLVE_PCG64_MG, rtol=1e-12, max_iter=2000)
    ur = O.solve_with_dirichlet(O.assemble(coords, conn, oop, dpn), rhs, bcs)
    assert np.linalg.norm(u - ur) <= 1e-9 * np.linalg.norm(ur)
    assert s.stats.iterations <= 60
    s.close()
    m.close()
