"""GPU edge cases against the oracle: isolated nodes, an element-free mesh,
systems with every DOF fixed, and batches of one-cell problems."""
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


def with_isolated_nodes(coords, conn, where, extra_at_end):
    """Insert nodes used by no element at index `where` and at the end."""
    k = len(where)
    n = len(coords)
    new_id = np.arange(n)
    for w in sorted(where):
        new_id[new_id >= w] += 1
    out = np.zeros((n + k + extra_at_end, 3))
    out[new_id] = coords
    rng = np.random.default_rng(1)
    iso = np.setdiff1d(np.arange(len(out)), new_id)
    out[iso] = rng.uniform(0, 1, (len(iso), 3))
    return out, new_id[conn]


@pytest.mark.parametrize("kind", ["poisson", "elasticity"])
def test_isolated_nodes_tet_assembly_bitexact(ctx, kind, monkeypatch):
    """Nodes that belong to no element (inside the numbering and at its end)
    give empty rows; the fast path and the generic kernel agree with the
    oracle bit for bit."""
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (3, 3, 2))
    coords, conn = with_isolated_nodes(coords, conn, [5, 17, 18, 40], 3)
    dpn = 1 if kind == "poisson" else 3
    op = T.Operator.poisson(1.5, 2) if kind == "poisson" else T.Operator.elasticity(T.CONTINUUM3D, 2.0, 0.3, 2)
    ref = O.make_op(O.TET, 3, 0, coef=1.5) if kind == "poisson" else O.make_op(O.TET, 3, 1, coef=2.0, nu=0.3)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    fast = T.assemble(m, dpn, op).to_csr()
    kr = O.assemble(coords, conn, ref, dpn).to_csr()
    csr_equal(fast, kr)
    monkeypatch.setenv("TFM_GENERIC_ASSEMBLY", "1")
    m2 = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    csr_equal(T.assemble(m2, dpn, op).to_csr(), kr)


def test_element_free_mesh_is_the_empty_csr(ctx):
    """assemble() over no elements: Coo::to_csr of no triplets is 1 x 1 with
    no entries."""
    T = tp()
    coords = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    conn = np.zeros((0, 4), dtype=np.int64)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    g = T.assemble(m, 1, T.Operator.poisson(1.0, 2)).to_csr()
    r = O.assemble(coords, conn, O.make_op(O.TET, 3, 0, coef=1.0), 1).to_csr()
    csr_equal(g, r)
    assert (g.nrows, g.ncols, g.nnz()) == (1, 1, 0)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_every_dof_fixed_fails_like_the_reference(ctx, mode):
    """No free DOF: the reference's solve gets the 1 x 1 CSR of no triplets
    against a length-0 rhs and reports a numeric error; so does the GPU."""
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (1, 1, 1))
    bcs = [(i, float(i) + 0.5) for i in range(len(coords))] + [(3, -7.0)]
    rhs = np.ones(len(coords))
    with pytest.raises(O.SolveError, match="does not match matrix dimension"):
        O.solve_with_dirichlet(O.assemble(coords, conn, O.make_op(O.TET, 3, 0, coef=1.0), 1), rhs, bcs)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    assert s.n_free == 0
    with pytest.raises(T.SparseError, match="does not match matrix dimension"):
        s.solve(T.SOLVE_PCG64 if mode == "plain" else T.SOLVE_MIXED, rtol=1e-12)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_batch_of_one_cell_problems(ctx, mode):
    """Ragged batch with problems that have no free DOF (one cell, all faces
    fixed) next to ones that have a few."""
    T = tp()
    N3 = [(1, 1, 1), (2, 2, 2), (1, 2, 1), (3, 2, 2), (1, 1, 1)]
    m, noff, eoff = T.Mesh.box_batch(ctx, N3)
    bcs = []
    for p, n in enumerate(N3):
        c, _ = O.box_mesh([0, 0, 0], [1, 1, 1], n)
        for i, x in enumerate(c):
            if np.any(np.isclose(x, 0.0)) or np.any(np.isclose(x, 1.0)):
                bcs.append((int(noff[p] + i), 0.0))
    rhs = T.source_vector(m, 1.0)
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    u = s.solve_batch(noff, mode=T.SOLVE_PCG64 if mode == "plain" else T.SOLVE_MIXED, rtol=1e-12)[0]
    for p, n in enumerate(N3):
        c, e = O.box_mesh([0, 0, 0], [1, 1, 1], n)
        lb = [(b - int(noff[p]), v) for b, v in bcs if noff[p] <= b < noff[p + 1]]
        up = u[noff[p]:noff[p + 1]]
        if len(lb) == len(c):  # no free DOF: the Dirichlet values (the oracle's solve refuses)
            assert np.array_equal(up, np.zeros(len(c))), p
            continue
        ur = O.solve_with_dirichlet(O.assemble(c, e, O.make_op(O.TET, 3, 0, coef=1.0), 1),
                                    O.poisson_rhs(c, O.TET, e, 1.0), lb)
        scale = max(np.abs(ur).max(), 1e-300)
        assert np.abs(up - ur).max() <= 1e-10 * scale, p
