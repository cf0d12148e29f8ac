"""Batched solves and the SIMP loop run as sequences of captured graphs
(refresh / pre-stage, initial state, chunks of iterations with the precision
form's refinement decided on the device, certificate / post-stage).  Every
replay must give exactly what the same kernels give launched one by one
(TFM_NO_GRAPH=1), with the program's own eager work and synchronisations
between the replays, and the results still meet the oracle's."""
import contextlib
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N3 = [(3, 4, 5), (6, 6, 6), (2, 3, 2), (5, 4, 3), (4, 4, 4), (7, 3, 4)]


@contextlib.contextmanager
def no_graph():
    os.environ["TFM_NO_GRAPH"] = "1"
    try:
        yield
    finally:
        del os.environ["TFM_NO_GRAPH"]


def boundary_nodes(n):
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    on = (i == 0) | (i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz)
    return ((i * (ny + 1) + j) * (nz + 1) + k)[on]


def poisson_batch(T, ctx, flags):
    m, noff, eoff = T.Mesh.box_batch(ctx, N3)
    kappa = np.exp(np.random.default_rng(11).standard_normal(m.n_elem))
    bcs = [(int(noff[p] + i), 0.0) for p, n in enumerate(N3) for i in boundary_nodes(n)]
    rhs = T.source_vector(m, 1.0)
    s = T.System.create(m, T.Operator.poisson(1.0, 2, kappa_e=kappa), 1, bcs, rhs, flags=flags)
    return m, noff, eoff, kappa, s


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_batched_solve_replays_equal_direct_launches(ctx, mode):
    import tptfem as T

    flags = T.SYS_FP64 | (T.SYS_FP32 if mode == "mixed" else 0)
    m, noff, eoff, kappa, s = poisson_batch(T, ctx, flags)
    sm = T.SOLVE_PCG64 if mode == "plain" else T.SOLVE_MIXED
    with no_graph():
        u0 = s.solve_batch(noff, mode=sm, rtol=1e-11)
        st0 = [x[:3] for x in s.batch_stats]
    for rep in range(3):
        u = s.solve_batch(noff, mode=sm, rtol=1e-11)
        assert np.array_equal(u, u0), rep
        assert [x[:3] for x in s.batch_stats] == st0, rep
        s.download()  # eager work and a synchronisation between the replays
    # and against the oracle's direct solves
    for p, n in enumerate(N3):
        c, e = O.box_mesh([0, 0, 0], [1, 1, 1], n)
        coo = O.assemble(c, e, O.make_op(O.TET, 3, 0, coef_e=kappa[eoff[p]:eoff[p + 1]]), 1)
        ur = O.solve_with_dirichlet(coo, O.poisson_rhs(c, O.TET, e, 1.0), [(int(i), 0.0) for i in boundary_nodes(n)])
        err = np.linalg.norm(u0[0, noff[p]:noff[p + 1]] - ur) / np.linalg.norm(ur)
        assert err <= (1e-9 if mode == "plain" else 1e-8), (p, err)


def test_batched_solve_after_reassembly_replays_refresh(ctx):
    """Stage A copies the system's current values: a re-assembly with new
    coefficients between two solves changes the answer exactly as a fresh
    system with those coefficients does."""
    import tptfem as T

    m, noff, eoff, kappa, s = poisson_batch(T, ctx, T.SYS_FP64 | T.SYS_FP32)
    s.solve_batch(noff, mode=T.SOLVE_MIXED, rtol=1e-11)
    k2 = kappa * np.exp(np.random.default_rng(12).uniform(-1, 1, m.n_elem))
    s.assemble(T.Operator.poisson(1.0, 2, kappa_e=k2), T.SYS_FP64 | T.SYS_FP32)
    u = s.solve_batch(noff, mode=T.SOLVE_MIXED, rtol=1e-11)
    bcs = [(int(noff[p] + i), 0.0) for p, n in enumerate(N3) for i in boundary_nodes(n)]
    s2 = T.System.create(m, T.Operator.poisson(1.0, 2, kappa_e=k2), 1, bcs, T.source_vector(m, 1.0),
                         flags=T.SYS_FP64 | T.SYS_FP32)
    u2 = s2.solve_batch(noff, mode=T.SOLVE_MIXED, rtol=1e-11)
    assert np.array_equal(u, u2)


@pytest.mark.parametrize("mode", ["plain", "mixed", "multigrid"])
def test_simp_loop_replays_equal_direct_launches(ctx, mode):
    """One OC iteration = assembly + refresh, solve, certificate + OC step as
    graph stages: the loop's densities and compliance history are bitwise
    those of the kernels launched one by one, and meet the oracle's."""
    import tptfem as T

    sm = {"plain": T.SOLVE_PCG64, "mixed": T.SOLVE_MIXED, "multigrid": T.SOLVE_PCG64_MG}[mode]
    nx, ny, iters = 18, 9, 8
    loads, vfs = [0, 4, 8], [0.5, 0.3, 0.6]
    F = np.zeros((3, 2 * nx * ny))
    for p, j in enumerate(loads):
        F[p, 2 * (j * nx + nx - 1) + 1] = -1.0
    _, bcs = O.cantilever_load(nx, ny, 1.0)
    args = (ctx, nx, ny, 1.0, 1.0, 0.3, vfs, 3.0, 1.5, iters, 0.2, F, bcs)
    with no_graph():
        g0 = T.topopt_simp(*args, mode=sm, rtol=1e-12)
    for rep in range(2):
        g = T.topopt_simp(*args, mode=sm, rtol=1e-12)
        assert np.array_equal(g.densities, g0.densities), rep
        assert np.array_equal(g.compliance, g0.compliance), rep
        assert [x[:3] for x in g.stats] == [x[:3] for x in g0.stats], rep
    for p in range(3):
        rr, cr, _ = O.topopt_simp(nx, ny, 1.0, 1.0, 0.3, vfs[p], 3.0, 1.5, iters, 0.2, F[p], bcs)
        assert np.abs(g0.densities[p] - rr).max() <= 1e-6, p
        np.testing.assert_allclose(g0.compliance[p], cr, rtol=1e-7, atol=0)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_assemble_solve_batch_equals_assemble_then_solve(ctx, mode):
    """tfm_system_assemble_solve_batch (the numeric pass in the solve's first
    graph stage) gives bitwise what tfm_system_assemble + solve_batch give,
    call after call with new coefficients (replayed graphs), and the values
    it leaves equal a fresh assembly."""
    import tptfem as T

    flags = T.SYS_FP64 | T.SYS_FP32
    sm = T.SOLVE_PCG64 if mode == "plain" else T.SOLVE_MIXED
    m, noff, eoff, kappa, s = poisson_batch(T, ctx, flags)
    _, _, _, _, r = poisson_batch(T, ctx, flags)
    rng = np.random.default_rng(21)
    for rep in range(3):
        k = kappa * np.exp(rng.uniform(-1, 1, m.n_elem))
        op = T.Operator.poisson(1.0, 2, kappa_e=k)
        u = s.assemble_solve_batch(op, noff, mode=sm, rtol=1e-11)
        st = [x[:3] for x in s.batch_stats]
        r.assemble(op, flags)
        ur = r.solve_batch(noff, mode=sm, rtol=1e-11)
        assert np.array_equal(u, ur), rep
        assert st == [x[:3] for x in r.batch_stats], rep
        Ks, Kr = s.download()[0], r.download()[0]
        assert np.array_equal(Ks.values, Kr.values), rep
    # the current coefficients again (op None) and the eager path agree
    u1 = s.assemble_solve_batch(None, noff, mode=sm, rtol=1e-11)
    with no_graph():
        u2 = s.assemble_solve_batch(None, noff, mode=sm, rtol=1e-11)
    assert np.array_equal(u1, u2) and np.array_equal(u1, u)
