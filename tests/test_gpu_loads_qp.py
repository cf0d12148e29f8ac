"""GPU parity of the closure-valued loads and drivers: the points handed out
(element quadrature points, boundary points with outward normals) and the
loads integrated from per-point values, bit-identical to the oracle; the
drivers' solutions within 1e-10 of the oracle's dense-LU solve; and the
reference's 3-D manufactured-solution rate test run through the GPU path
(crates/tpt-fem-thermal/tests/mms_convergence.rs:336-384)."""
import numpy as np
import pytest

import oracle as O
from test_gpu_parity import hex_box, perturb, tp

pytestmark = pytest.mark.gpu


def _meshes(ctx):
    T = tp()
    n = (4, 3, 3)
    tc, te = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    tc = O.jitter_interior(tc, n, 0.2, 21)
    qc, qe = O.grid(6, 5, 0.2)
    qc = perturb(qc, 0.03, 4)
    hc, he = hex_box((3, 2, 3))
    hc = perturb(hc, 0.04, 6)
    return [(O.TET, tc, te, T.Mesh.from_arrays(ctx, T.TET, tc, te)),
            (O.QUAD, qc, qe, T.Mesh.from_arrays(ctx, T.QUAD, qc, qe)),
            (O.HEX, hc, he, T.Mesh.from_arrays(ctx, T.HEX, hc, he))]


def _b(x):
    # a body force / source with every component depending on position
    return np.stack([np.sin(3 * x[:, 0]) + x[:, 1], x[:, 0] * x[:, -1] - 0.5, np.cos(x[:, 1]) + 2 * x[:, 0]], axis=1)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_quad_points_and_element_loads_bitexact(ctx, order):
    T = tp()
    for cell, c, e, m in _meshes(ctx):
        xq = m.quad_points(order)
        ref = O.mesh_qp_coords(c, cell, e, order)
        assert xq.shape == ref.shape and np.array_equal(xq, ref), cell
        dim = O.DIMS[cell]
        bq = _b(xq.reshape(-1, xq.shape[2]))[:, :dim].reshape(len(e), -1, dim)
        assert np.array_equal(T.body_vector_qp(m, bq, order), O.body_rhs(c, cell, e, bq, order)), cell
        sq = np.ascontiguousarray(bq[:, :, 0])
        assert np.array_equal(T.source_vector_qp(m, sq, order), O.source_rhs_qp(c, cell, e, sq, order)), cell


@pytest.mark.parametrize("order", [2, 3])
def test_boundary_points_neumann_and_robin_bitexact(ctx, order):
    T = tp()
    for cell, c, e, m in _meshes(ctx):
        x, nrm = m.boundary_points(order)
        rx, rn = O.boundary_points(c, cell, e, order)
        assert np.array_equal(x, rx) and np.array_equal(nrm, rn), cell
        g = np.sin(2 * x[:, 0]) * nrm[:, 0] + x[:, 1] * nrm[:, 1] + 0.25
        dpn = O.DIMS[cell]
        got = T.apply_neumann(m, dpn, lambda X, N: np.sin(2 * X[:, 0]) * N[:, 0] + X[:, 1] * N[:, 1] + 0.25,
                              np.zeros(len(c) * dpn), order=order, vectorized=True)
        assert np.array_equal(got, O.apply_neumann_qp(c, cell, e, dpn, g, np.zeros(len(c) * dpn), order)), cell
        h = 1.0 + x[:, 0] ** 2
        op = T.Operator.poisson(1.3, 2)
        K = T.assemble_robin(m, 1, op, lambda X, N: 1.0 + X[:, 0] ** 2, order=order, vectorized=True).to_csr()
        a = O.assemble(c, e, O.make_op(cell, c.shape[1], 0, order=2, coef=1.3), 1)
        b = O.apply_robin_qp(c, cell, e, 1, h, order)
        R = O.Coo(np.concatenate([a.rows, b.rows]), np.concatenate([a.cols, b.cols]),
                  np.concatenate([a.vals, b.vals])).to_csr()
        np.testing.assert_array_equal(K.row_ptrs, R.row_ptrs)
        np.testing.assert_array_equal(K.col_ind, R.col_ind)
        assert np.array_equal(K.values, R.values), cell


def test_per_point_closure_calls_match_vectorized(ctx):
    # the binding's per-point calls (the reference's calling pattern) and one
    # vectorized call give the same values
    T = tp()
    cell, c, e, m = _meshes(ctx)[0]
    f1 = T.apply_neumann(m, 1, lambda x, n: float(x[0] * n[0] + 0.5), np.zeros(len(c)))
    f2 = T.apply_neumann(m, 1, lambda X, N: X[:, 0] * N[:, 0] + 0.5, np.zeros(len(c)), vectorized=True)
    assert np.array_equal(f1, f2)


def test_solve_elasticity_with_body_force_matches_dense_lu(ctx):
    T = tp()
    for cell, c, e, m in _meshes(ctx):
        dim = O.DIMS[cell]
        model = T.CONTINUUM3D if dim == 3 else T.PLANE_STRAIN
        # clamp the first column of nodes (the Q4 and Hex8 meshes move every node)
        bcs = [(int(n) * dim + a, 0.0) for n in np.flatnonzero(c[:, 0] < 0.1) for a in range(dim)]
        const = np.array([0.3, -1.0, 0.7])[:dim]
        nq = m.quad_points(2).shape[1]
        u = T.solve_elasticity(m, model, 50.0, 0.3, 2, const, bcs)
        ur = O.solve_elasticity(c, cell, e, model, 50.0, 0.3, 2, np.tile(const, (len(e), nq, 1)), bcs)
        assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max(), cell
        fn = lambda X: _b(X)[:, :dim]
        u = T.solve_elasticity(m, model, 50.0, 0.3, 2, fn, bcs, vectorized=True)
        bq = fn(m.quad_points(2).reshape(-1, c.shape[1])).reshape(len(e), nq, dim)
        ur = O.solve_elasticity(c, cell, e, model, 50.0, 0.3, 2, bq, bcs)
        assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max(), cell


def test_solve_poisson_with_closures_matches_dense_lu(ctx):
    T = tp()
    cell, c, e, m = _meshes(ctx)[0]
    src = lambda X: 1.0 + X[:, 0] * X[:, 1]
    flux = lambda X, N: 0.3 * N[:, 2] + X[:, 0]
    rob = lambda X, N: 2.0 + X[:, 1]
    bcs = [(int(n), 0.5) for n in O.nodes_on_plane(c, 0, 0.0, 1e-9)]
    u = T.solve_poisson(m, 1.7, 2, src, bcs, neumann=flux, robin=rob, vectorized=True)
    xq = m.quad_points(2)
    sq = src(xq.reshape(-1, 3)).reshape(len(e), -1)
    xb, nb = m.boundary_points(2)
    rhs = O.source_rhs_qp(c, cell, e, sq, 2)
    rhs = O.apply_neumann_qp(c, cell, e, 1, flux(xb, nb), rhs, 2)
    a = O.assemble(c, e, O.make_op(cell, 3, 0, order=2, coef=1.7), 1)
    b = O.apply_robin_qp(c, cell, e, 1, rob(xb, nb), 2)
    coo = O.Coo(np.concatenate([a.rows, b.rows]), np.concatenate([a.cols, b.cols]), np.concatenate([a.vals, b.vals]))
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()
    # the source alone through the closure path equals the constant path
    u1 = T.solve_poisson(m, 1.0, 2, lambda x: 2.0, bcs)
    u2 = T.solve_poisson(m, 1.0, 2, 2.0, bcs)
    assert np.abs(u1 - u2).max() <= 1e-12 * np.abs(u2).max()


def _tet_rule_deg5():
    # collapsed-cube (Duffy) product of 4-point Gauss-Legendre rules on [0, 1]:
    # x = u, y = v (1 - u), z = w (1 - u)(1 - v), |J| = (1 - u)^2 (1 - v)
    g, w = np.polynomial.legendre.leggauss(4)
    g, w = (g + 1) / 2, w / 2
    pts, wts = [], []
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            for c, wc in zip(g, w):
                pts.append([a, b * (1 - a), c * (1 - a) * (1 - b)])
                wts.append(wa * wb * wc * (1 - a) ** 2 * (1 - b))
    return np.array(pts), np.array(wts)


def _l2_h1_errors(c, e, u, uex, gex):
    # the reference's l2_error / h1_error (mms_convergence.rs:84-262) in numpy,
    # with an independent degree-5 tet rule in place of Keast3
    p, w = _tet_rule_deg5()
    N = np.stack([1 - p.sum(axis=1), p[:, 0], p[:, 1], p[:, 2]], axis=1)
    X = c[e]                                           # (ne, 4, 3)
    J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    det = np.abs(np.linalg.det(J))
    xq = np.einsum("qk,ekd->eqd", N, X)
    uh = np.einsum("qk,ek->eq", N, u[e])
    l2 = np.sqrt(np.sum(w[None] * det[:, None] * (uh - uex(xq.reshape(-1, 3)).reshape(uh.shape)) ** 2))
    G = np.linalg.inv(J)                               # rows: gradients of N1..N3
    gr = np.concatenate([-G.sum(axis=1, keepdims=True), G], axis=1)  # (ne, 4, 3)
    guh = np.einsum("ek,ekd->ed", u[e], gr)
    ge = gex(xq.reshape(-1, 3)).reshape(len(e), len(w), 3)
    h1 = np.sqrt(np.sum(w[None] * det[:, None] * np.sum((guh[:, None] - ge) ** 2, axis=2)))
    return l2, h1


def test_mms_3d_rates_through_the_gpu_path(ctx):
    # the reference's mms_3d_converges (mms_convergence.rs:336-384): u = sin(pi x)
    # sin(pi y) sin(pi z), f = 3 pi^2 u, u = u_exact on the boundary, box_mesh
    # n^3, quad order 4; L2 ratio > 4 * 0.85 and H1 ratio > 2 * 0.92 -- here on
    # n = 4, 8, 16 instead of the reference's 4, 8
    T = tp()
    pi = np.pi
    uex = lambda X: np.prod(np.sin(pi * X), axis=1)
    fsrc = lambda X: 3 * pi * pi * uex(X)
    gex = lambda X: pi * np.stack([np.cos(pi * X[:, 0]) * np.sin(pi * X[:, 1]) * np.sin(pi * X[:, 2]),
                                   np.sin(pi * X[:, 0]) * np.cos(pi * X[:, 1]) * np.sin(pi * X[:, 2]),
                                   np.sin(pi * X[:, 0]) * np.sin(pi * X[:, 1]) * np.cos(pi * X[:, 2])], axis=1)
    l2s, h1s = [], []
    for n in (4, 8, 16):
        m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], (n, n, n))
        c, e = m.download()
        onb = np.any((np.abs(c) < 1e-9) | (np.abs(c - 1) < 1e-9), axis=1)
        bcs = [(int(i), float(uex(c[i:i + 1])[0])) for i in np.flatnonzero(onb)]
        u = T.solve_poisson(m, 1.0, 4, fsrc, bcs, vectorized=True)
        l2, h1 = _l2_h1_errors(c, e, u, uex, gex)
        l2s.append(l2)
        h1s.append(h1)
    for a, b in zip(l2s, l2s[1:]):
        assert a / b > 4.0 * 0.85, l2s
    for a, b in zip(h1s, h1s[1:]):
        assert a / b > 2.0 * 0.92, h1s
