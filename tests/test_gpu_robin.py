"""Robin terms on the GPU (apply_robin appended to assemble()'s triplets,
crates/tpt-fem-assembly/src/lib.rs:674-753) and solve_poisson with Neumann
and Robin (crates/tpt-fem-thermal/src/lib.rs:206-240), against the oracle:
patterns and FP64 values bit-identical, solutions within 1e-10."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def tp():
    import tptfem

    return tptfem


def csr_equal(g, r):
    assert (g.nrows, g.ncols) == (r.nrows, r.ncols)
    np.testing.assert_array_equal(g.row_ptrs, r.row_ptrs)
    np.testing.assert_array_equal(g.col_ind, r.col_ind)
    assert np.array_equal(g.values, r.values)


@pytest.mark.parametrize("band", [False, True])
def test_robin_matrices_bitexact(ctx, band):
    """Jittered tet Poisson and elasticity, Q4 plane stress and Tri3 Poisson,
    h constant or limited to a plane band (zero entries still created)."""
    T = tp()
    ax = (0, 1.0, 1e-9) if band else (-1, 0.0, 0.0)
    n3 = (3, 2, 3)
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], n3)
    c = O.jitter_interior(c, n3, 0.2, 4)
    m = T.Mesh.from_arrays(ctx, T.TET, c, e)
    cases = [(m, c, e, O.TET, 1, T.Operator.poisson(1.5, 2), O.make_op(O.TET, 3, 0, coef=1.5)),
             (m, c, e, O.TET, 3, T.Operator.elasticity(T.CONTINUUM3D, 2.0, 0.3, 2),
              O.make_op(O.TET, 3, 1, coef=2.0, nu=0.3))]
    qc, qe = O.grid(5, 4, 0.25)
    mq = T.Mesh.from_arrays(ctx, T.QUAD, qc, qe)
    cases.append((mq, qc, qe, O.QUAD, 2, T.Operator.elasticity(T.PLANE_STRESS, 1.0, 0.25, 2),
                  O.make_op(O.QUAD, 2, 1, model=O.PSTRESS, coef=1.0, nu=0.25)))
    tc = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [0.45, 0.55]], float)
    te = np.array([[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]])
    mt = T.Mesh.from_arrays(ctx, T.TRI, tc, te)
    cases.append((mt, tc, te, O.TRI, 1, T.Operator.poisson(0.7, 2), O.make_op(O.TRI, 2, 0, coef=0.7)))
    for mesh, cc, ee, cell, dpn, op, oop in cases:
        g = T.assemble_robin(mesh, dpn, op, 3.5, *ax).to_csr()
        r = O.assemble_robin(cc, cell, ee, oop, dpn, 3.5, *ax).to_csr()
        csr_equal(g, r)


def test_solve_poisson_with_neumann_and_robin(ctx):
    """u = 0 on x = 0, flux 0.3 and convection h = 2 on the rest of the
    boundary, source 1: the GPU driver against the oracle's dense solve."""
    T = tp()
    n3 = (4, 3, 3)
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], n3)
    c = O.jitter_interior(c, n3, 0.15, 8)
    m = T.Mesh.from_arrays(ctx, T.TET, c, e)
    bcs = [(int(i), 0.0) for i in O.nodes_on_plane(c, 0, 0.0, 1e-9)]
    u = T.solve_poisson_nr(m, 1.0, 2, 1.0, bcs, neumann=0.3, robin=2.0)
    coo = O.assemble_robin(c, O.TET, e, O.make_op(O.TET, 3, 0, coef=1.0), 1, 2.0)
    rhs = O.apply_neumann(c, O.TET, e, 1, 0.3, O.poisson_rhs(c, O.TET, e, 1.0))
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()
    # without Robin/Neumann terms solve_poisson_nr reduces to solve_poisson
    u0 = T.solve_poisson_nr(m, 1.0, 2, 1.0, bcs)
    ur0 = O.solve_with_dirichlet(O.assemble(c, e, O.make_op(O.TET, 3, 0, coef=1.0), 1),
                                 O.poisson_rhs(c, O.TET, e, 1.0), bcs)
    assert np.abs(u0 - ur0).max() <= 1e-10 * np.abs(ur0).max()
