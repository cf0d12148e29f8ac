"""The reference's direct solve on the GPU (TFM_SOLVE_DIRECT: dense
partial-pivot LU of the reduced system) and the general-solve contract around
it: bit-identical to the oracle's dense LU (the same loop run sequentially),
the fallback of the iterative modes on systems CG cannot solve (not SPD),
singular matrices reported as the reference's numeric error, and any number
of right-hand sides (solve_multi, crates/tpt-fem-sparse/src/lib.rs:244-278)."""
import numpy as np
import pytest

import oracle as O
from test_gpu_parity import face_bcs, tp

pytestmark = pytest.mark.gpu


def _elastic_box(ctx, n, seed):
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    coords = O.jitter_interior(coords, n, 0.2, seed)
    ye = 10.0 ** np.random.default_rng(seed).uniform(-1, 1, len(conn))
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    bcs = face_bcs(coords, 3, [(0, 0.0)])
    rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [0.0, 0.3, -1.0])
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef_e=ye, nu=0.3), 3)
    return m, op, bcs, rhs, coo


def test_direct_solve_is_bitexact_to_the_oracle_lu(ctx):
    T = tp()
    m, op, bcs, rhs, coo = _elastic_box(ctx, (4, 3, 3), 3)
    s = T.System.create(m, op, 3, bcs, rhs)
    u = s.solve(T.SOLVE_DIRECT)
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert np.array_equal(u, ur)
    assert s.stats.rel_residual < 1e-13 and s.stats.iterations == 0
    # non-homogeneous Dirichlet values (general lifting) as well
    bcs2 = [(d, 0.01 * (i % 5)) for i, (d, _) in enumerate(bcs)]
    s2 = T.System.create(m, op, 3, bcs2, rhs)
    assert np.array_equal(s2.solve(T.SOLVE_DIRECT), O.solve_with_dirichlet(coo, rhs, bcs2))


def _indefinite(n, seed):
    # a non-symmetric, indefinite sparse matrix (diagonal +4 on the first half
    # of the unknowns, -4 on the second, couplings summing below 4 in
    # magnitude, so it is diagonally dominant and well conditioned): CG cannot
    # solve it, the reference's LU can
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        rows += [i]; cols += [i]; vals += [(4.0 if i < n // 2 else -4.0) + 0.1 * rng.random()]
        if i > 0:
            rows += [i]; cols += [i - 1]; vals += [-1.3]
        if i + 1 < n:
            rows += [i]; cols += [i + 1]; vals += [-0.7]
        if i + 7 < n:
            rows += [i]; cols += [i + 7]; vals += [0.2 * rng.standard_normal()]
    return O.Coo(np.array(rows, np.int64), np.array(cols, np.int64), np.array(vals))


@pytest.mark.parametrize("mode", ["pcg64", "direct"])
def test_non_spd_system_falls_back_to_the_direct_solve(ctx, mode):
    T = tp()
    coo = _indefinite(300, 4)
    K = T.Matrix.from_csr(ctx, coo.to_csr())
    rhs = np.sin(np.arange(300.0))
    bcs = [(0, 0.25), (150, -1.0)]
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    s = T.System.from_matrix(K, rhs, bcs)
    md = {"pcg64": T.SOLVE_PCG64, "direct": T.SOLVE_DIRECT}[mode]
    u = s.solve(md, rtol=1e-12, max_iter=2000)
    assert np.array_equal(u, ur)


def test_singular_system_is_a_numeric_error_in_every_mode(ctx):
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (2, 2, 2))
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, [], np.ones(len(coords)))
    for mode in (T.SOLVE_DIRECT, T.SOLVE_PCG64):
        with pytest.raises(T.SparseError, match="singular"):
            s.solve(mode, rtol=1e-14, max_iter=500)


@pytest.mark.parametrize("nrhs", [3, 11])
def test_any_number_of_right_hand_sides(ctx, nrhs):
    T = tp()
    m, op, bcs, rhs, coo = _elastic_box(ctx, (4, 3, 3), 7)
    s = T.System.create(m, op, 3, bcs, rhs)
    F = np.stack([rhs * (k + 1) + 0.01 * np.cos(np.arange(len(rhs)) * (k + 1)) for k in range(nrhs)])
    off = [0, s.n_dof]
    ud = s.solve_batch(off, rhs=F, nrhs=nrhs, mode=T.SOLVE_DIRECT)
    ui = s.solve_batch(off, rhs=F, nrhs=nrhs, mode=T.SOLVE_PCG64, rtol=1e-12, max_iter=20000)
    assert len(s.batch_stats) == nrhs
    for k in range(nrhs):
        ur = O.solve_with_dirichlet(coo, F[k], bcs)
        assert np.array_equal(ud[k], ur), k
        assert np.abs(ui[k] - ur).max() <= 1e-10 * np.abs(ur).max(), k


def test_direct_solve_size_limit(ctx):
    T = tp()
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], (21, 21, 21))  # 10,648 nodes, no DOF fixed
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, [(0, 0.0)], np.ones(m.n_nodes))
    with pytest.raises(T.TfmError) as ei:
        s.solve(T.SOLVE_DIRECT)
    assert ei.value.code == T.E_UNSUPPORTED


def test_direct_solve_of_a_few_thousand_unknowns(ctx):
    T = tp()
    n = (14, 14, 14)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    coords, conn = m.download()
    bcs = [(int(i), 0.0) for a in range(3) for v in (0.0, 1.0) for i in O.nodes_on_plane(coords, a, v, 1e-9)]
    rhs = O.poisson_rhs(coords, O.TET, conn, 1.0)
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, bcs, rhs)
    assert s.n_free == 13 ** 3
    u = s.solve(T.SOLVE_DIRECT)
    ui = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    assert np.abs(u - ui).max() <= 1e-11 * np.abs(ui).max()
