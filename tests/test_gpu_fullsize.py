"""GPU parity at the full size of the headline workload: box_mesh 128^3 cells
(12,582,912 tets), interior nodes jittered by +-0.2h (seed 2024), E_e =
10^U(-1,1), nu = 0.3, x = 0 clamped, traction (0, 0, -1e-3) on x = 1.

The oracle cannot assemble 12.6M tets within a test, so the checks are the
ones that do not depend on size:
* the mesh and the free-DOF map, in full;
* the reduced right-hand side, in full;
* reduced stiffness rows of 1,500 random nodes plus the nodes next to the
  clamp, on the loaded face and at the corners, bit for bit: row n of the
  assembled matrix only receives the elements that contain node n, summed in
  ascending element order, so the oracle's assemble + reduce_system over just
  those elements (kept in ascending order) reproduces those rows exactly;
* the symmetry of the sampled rows (B^T D B evaluated in the reference's
  order is symmetric only to rounding: 1e-14 of the row's largest entry);
* the multigrid-preconditioned CG at full size, by its true FP64 residual.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N = 128
SEED = 2024
TRACTION = [0.0, 0.0, -1e-3]


def node_id(i, j, k):
    return (i * (N + 1) + j) * (N + 1) + k


@pytest.fixture(scope="module")
def c3(ctx):
    import tptfem as T

    mesh = T.Mesh.box(ctx, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], (N, N, N))
    mesh.jitter((N, N, N), 0.2, SEED)
    ye = 10.0 ** np.random.default_rng(SEED).uniform(-1.0, 1.0, mesh.n_elem)
    clamp = mesh.nodes_on_plane(0, 0.0, 1e-9)
    bcs = [(int(i) * 3 + k, 0.0) for i in clamp for k in range(3)]
    rhs = T.plane_traction(mesh, 3, 0, 1.0, 1e-9, TRACTION)
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)
    s = T.System.create(mesh, op, 3, bcs, rhs, flags=T.SYS_FP64)
    K, f, free = s.download()
    rc, re = O.box_mesh([0, 0, 0], [1, 1, 1], (N, N, N))
    rc = O.jitter_interior(rc, (N, N, N), 0.2, SEED)
    yield dict(T=T, mesh=mesh, s=s, ye=ye, bcs=bcs, K=K, f=f, free=free, rc=rc, re=re)
    s.close()
    mesh.close()


def test_full_size_mesh_and_free_map(c3):
    c, e = c3["mesh"].download()
    assert c.shape == (129 ** 3, 3) and e.shape == (12_582_912, 4)
    assert np.array_equal(c, c3["rc"]) and np.array_equal(e, c3["re"])
    fixed = np.unique([d for d, _ in c3["bcs"]])
    free = c3["free"]
    assert free.size == 6_390_144 and c3["K"].nrows == free.size
    np.testing.assert_array_equal(free, np.setdiff1d(np.arange(3 * 129 ** 3), fixed))


def oracle_rhs(rc, re):
    """plane_traction on x = 1 over the elements with a node on that plane
    (ascending): a face on x = 1 lies on the boundary and in one element,
    which has nodes on the plane, so the load equals the whole mesh's."""
    on = np.zeros(len(rc), bool)
    on[O.nodes_on_plane(rc, 0, 1.0, 1e-9)] = True
    sub = np.flatnonzero(on[re].any(1))
    return O.plane_traction(rc, O.TET, re[sub], 3, 0, 1.0, 1e-9, TRACTION)


def sample_nodes():
    rng = np.random.default_rng(3)
    nodes = set(rng.choice(129 ** 3, 1500, replace=False).tolist())
    for j in (0, 1, 64, 127, 128):
        for k in (0, 1, 64, 127, 128):
            nodes.update([node_id(1, j, k), node_id(2, j, k), node_id(N, j, k), node_id(N - 1, j, k)])
    nodes.update([node_id(N, N, N), node_id(N, 0, 0), node_id(64, 64, 64)])
    return np.array(sorted(nodes))


def sampled_rows(free, nodes):
    """Reduced row indices of the sampled nodes' free DOFs."""
    dofs = (3 * nodes[:, None] + np.arange(3)).ravel()
    i = np.searchsorted(free, dofs)
    ok = (i < free.size) & (free[np.minimum(i, free.size - 1)] == dofs)
    return i[ok]


def test_full_size_reduced_rows_and_rhs_bitexact(c3):
    rc, re, ye, bcs = c3["rc"], c3["re"], c3["ye"], c3["bcs"]
    nodes = sample_nodes()
    sel = np.zeros(len(rc), bool)
    sel[nodes] = True
    patch = np.flatnonzero(sel[re].any(1))
    coo = O.assemble(rc, re[patch], O.make_op(O.TET, 3, 1, coef_e=ye[patch], nu=0.3), 3)
    # one explicit 0.0 at (last DOF, a clamped DOF) sizes the matrix to every
    # DOF; it only meets the lifting of a zero Dirichlet value (f - 0 K = f)
    ndof = 3 * len(rc)
    coo = O.Coo(np.append(coo.rows, ndof - 1), np.append(coo.cols, 0), np.append(coo.vals, 0.0))
    kr, fr, free = O.reduce_system(coo, oracle_rhs(rc, re), bcs)
    np.testing.assert_array_equal(free, c3["free"])
    assert np.array_equal(c3["f"].view(np.uint64), fr.view(np.uint64))
    kc = kr.to_csr()
    K = c3["K"]
    rows = sampled_rows(free, nodes)
    assert rows.size >= 4000
    for i in rows:
        g0, g1 = K.row_ptrs[i], K.row_ptrs[i + 1]
        r0, r1 = kc.row_ptrs[i], kc.row_ptrs[i + 1]
        np.testing.assert_array_equal(K.col_ind[g0:g1], kc.col_ind[r0:r1])
        assert np.array_equal(K.values[g0:g1].view(np.uint64), kc.values[r0:r1].view(np.uint64)), f"row {i}"


def test_full_size_sampled_rows_symmetric(c3):
    K = c3["K"]
    worst = 0.0
    for i in sampled_rows(c3["free"], sample_nodes())[::2]:
        g0, g1 = K.row_ptrs[i], K.row_ptrs[i + 1]
        cols, vals = K.col_ind[g0:g1], K.values[g0:g1]
        scale = np.abs(vals).max()
        for j, v in zip(cols, vals):
            h0, h1 = K.row_ptrs[j], K.row_ptrs[j + 1]
            q = h0 + np.searchsorted(K.col_ind[h0:h1], i)
            assert q < h1 and K.col_ind[q] == i
            worst = max(worst, abs(K.values[q] - v) / scale)
    assert worst <= 1e-14


def test_full_size_multigrid_solve_true_residual(c3):
    import scipy.sparse as sp

    T, s, K, f, free = c3["T"], c3["s"], c3["K"], c3["f"], c3["free"]
    u = s.solve(T.SOLVE_PCG64_MG, rtol=1e-10)
    A = sp.csr_matrix((K.values, K.col_ind, K.row_ptrs), shape=(K.nrows, K.ncols))
    r = f - A @ u[free]
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(f)
    assert np.all(u[: 3 * 129 * 129] == 0.0)  # the clamped x = 0 face (node ids below 129^2)
