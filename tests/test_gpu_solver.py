"""GPU solver forms against the oracle's dense-LU direct solve.

The iteration runs on the node-blocked sliced copy of the system matrix
(block size = DOFs per node: 1 for Poisson and matrices given as CSR, 2 for
plane elasticity, 3 for 3-D elasticity), in FP64 (plain) and FP32 with FP64
residual replacement (precision form).  Bars: plain solutions within 1e-10
relative of the direct solve; precision-form solutions within 1e-8 relative
of the FP64 solution; repeated solves bit-identical (deterministic
reductions).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def tp():
    import tptfem

    return tptfem


def rel_max(u, ur):
    return np.abs(u - ur).max() / np.abs(ur).max()


def plane_bcs(coords, dpn, axis, val, comps):
    return [(int(n) * dpn + c, 0.0) for n in O.nodes_on_plane(coords, axis, val, 1e-9) for c in comps]


def test_plane_strain_quad_solve_block2(ctx):
    """Block size 2: plane-strain Q4 cantilever, clamped at x = 0."""
    T = tp()
    coords, conn = O.grid(13, 7, 1.0 / 12)
    bcs = plane_bcs(coords, 2, 0, 0.0, (0, 1))
    inner = coords[:, 0] > 1e-9  # the clamped column stays on x = 0
    coords[inner] += np.random.default_rng(5).uniform(-0.01, 0.01, (inner.sum(), 2))
    m = T.Mesh.from_arrays(ctx, T.QUAD, coords, conn)
    rhs = np.zeros(len(coords) * 2)
    rhs[1::2] = -1e-3
    op = T.Operator.elasticity(T.PLANE_STRAIN, 1.0, 0.3, 2)
    s = T.System.create(m, op, 2, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    u = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    coo = O.assemble(coords, conn, O.make_op(O.QUAD, 2, 1, coef=1.0, nu=0.3, model=T.PLANE_STRAIN), 2)
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert rel_max(u, ur) <= 1e-10
    um = s.solve(T.SOLVE_MIXED, rtol=1e-12, inner_rtol=0.1)
    assert np.linalg.norm(um - u) <= 1e-8 * np.linalg.norm(u)


def test_symmetry_planes_tet_solve_partial_blocks(ctx):
    """Block size 3 with partially fixed nodes: u_x = 0 on x = 0, u_y = 0 on
    y = 0, u_z = 0 on z = 0 (blocks with one or two empty rows)."""
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (4, 3, 3))
    bcs = (plane_bcs(coords, 3, 0, 0.0, (0,)) + plane_bcs(coords, 3, 1, 0.0, (1,)) +
           plane_bcs(coords, 3, 2, 0.0, (2,)))
    rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [1e-3, 0.0, 0.0])
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    op = T.Operator.elasticity(T.CONTINUUM3D, 2.0, 0.25, 2)
    s = T.System.create(m, op, 3, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    u = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef=2.0, nu=0.25), 3)
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert rel_max(u, ur) <= 1e-10
    um = s.solve(T.SOLVE_MIXED, rtol=1e-12, inner_rtol=0.1)
    assert np.linalg.norm(um - u) <= 1e-8 * np.linalg.norm(u)


def test_matrix_given_as_csr_solve_block1(ctx):
    """Block size 1 on a system built from an uploaded CSR (an elasticity
    matrix, so rows are longer than a Poisson row)."""
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (3, 3, 2))
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef=1.0, nu=0.3), 3)
    csr = coo.to_csr()
    bcs = plane_bcs(coords, 3, 0, 0.0, (0, 1, 2))
    rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [0.0, -1e-3, 0.0])
    k = T.Matrix.from_csr(ctx, T.Csr(csr.nrows, csr.ncols, csr.row_ptrs, csr.col_ind, csr.values))
    s = T.System.from_matrix(k, rhs, bcs)
    u = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert rel_max(u, ur) <= 1e-10


def test_poisson_mixed_block1_and_determinism(ctx):
    T = tp()
    n = (6, 5, 4)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    m.jitter(n, 0.2, 3)
    coords, conn = m.download()
    kap = 10.0 ** np.random.default_rng(9).uniform(-1, 1, m.n_elem)
    bcs = plane_bcs(coords, 1, 0, 0.0, (0,)) + plane_bcs(coords, 1, 0, 1.0, (0,))
    rhs = T.source_vector(m, 1.0, 2)
    op = T.Operator.poisson(1.0, 2, kappa_e=kap)
    s = T.System.create(m, op, 1, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    u1 = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    u2 = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    assert np.array_equal(u1, u2)
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 0, coef_e=kap), 1)
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert rel_max(u1, ur) <= 1e-10
    um = s.solve(T.SOLVE_MIXED, rtol=1e-12, inner_rtol=0.1)
    assert np.linalg.norm(um - u1) <= 1e-8 * np.linalg.norm(u1)


def test_elasticity_resolve_after_reassembly(ctx):
    """New moduli through System.assemble: the solver's copy follows."""
    T = tp()
    n = (4, 4, 3)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    coords, conn = m.download()
    bcs = plane_bcs(coords, 3, 0, 0.0, (0, 1, 2))
    rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    s = T.System.create(m, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2), 3, bcs, rhs)
    u1 = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    ye = np.full(m.n_elem, 4.0)
    s.assemble(T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye))
    u4 = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    # linear in 1/E
    assert np.abs(u4 * 4.0 - u1).max() <= 1e-9 * np.abs(u1).max()


@pytest.mark.parametrize("n", [(2, 3, 2), (3, 3, 3)])
def test_mixed_tiny_system_converges(ctx, n):
    """A system with 2 (or 8) free DOFs: its FP32 iteration converges (and
    breaks down) inside one 16-iteration chunk; the precision form must still
    end at the FP64 tolerance."""
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    bnd = [i for i, c in enumerate(coords) if np.any(np.isclose(c, 0.0)) or np.any(np.isclose(c, 1.0))]
    bcs = [(int(i), 0.0) for i in bnd]
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    rhs = T.source_vector(m, 1.0, 2)
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    um = s.solve(T.SOLVE_MIXED, rtol=1e-12, max_iter=2000)
    assert s.stats.rel_residual <= 1e-12
    ur = O.solve_with_dirichlet(O.assemble(coords, conn, O.make_op(O.TET, 3, 0, coef=1.0), 1),
                                O.poisson_rhs(coords, O.TET, conn, 1.0), bcs)
    assert rel_max(um, ur) <= 1e-10


def test_mixed_small_replacement_factor_still_converges(ctx):
    """inner_rtol far below what FP32 can reach: the stagnation guard makes
    the replacements, and the solve still meets rtol."""
    T = tp()
    n = (6, 5, 4)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    m.jitter(n, 0.2, 5)
    coords, conn = m.download()
    bcs = plane_bcs(coords, 3, 0, 0.0, (0, 1, 2))
    rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    s = T.System.create(m, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2), 3, bcs, rhs,
                        flags=T.SYS_FP64 | T.SYS_FP32)
    u64 = s.solve(T.SOLVE_PCG64, rtol=1e-12)
    um = s.solve(T.SOLVE_MIXED, rtol=1e-11, inner_rtol=1e-4, max_iter=100000)
    assert s.stats.rel_residual <= 1e-11
    assert np.linalg.norm(um - u64) <= 1e-8 * np.linalg.norm(u64)
