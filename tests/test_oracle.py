"""CPU pins of the oracle (no GPU): every oracle function is checked against
the reference's own known answers (tests/golden/kats.json), closed forms or
textbook routines, chosen so that a plausible slip (dropped term, transposed
operand, wrong sign or index) fails at least one of them."""
import json
import os

import numpy as np
import pytest
from scipy import linalg

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(GOLD, "kats.json")))
CELL = {"line": O.LINE, "tri": O.TRI, "tet": O.TET}


# ---- quadrature --------------------------------------------------------------
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_gauss_legendre_exact_to_degree(order):
    p, w = O.gauss_legendre(order)
    for deg in range(2 * order):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, p ** deg) - exact) < 1e-9


def test_tet_and_triangle_rules():
    p, w = O.cell_quad(O.TET, 7)  # order is ignored for tets
    assert len(w) == 4 and abs(w.sum() - 1 / 6) < 1e-15
    x, y, z = p.T
    assert abs(np.dot(w, x * x) - 1 / 60) < 1e-12 and abs(np.dot(w, x * y) - 1 / 120) < 1e-12
    p, w = O.cell_quad(O.TRI, 5)
    assert abs(w.sum() - 0.5) < 1e-15 and abs(np.dot(w, p[:, 0] ** 2) - 1 / 12) < 1e-12
    p, w = O.cell_quad(O.HEX, 2)
    assert len(w) == 8 and abs(w.sum() - 8.0) < 1e-14 and abs(np.dot(w, p[:, 0] ** 2 * p[:, 2] ** 2) - 8 / 9) < 1e-12


# ---- shape functions ---------------------------------------------------------
@pytest.mark.parametrize("cell", [O.LINE, O.TRI, O.QUAD, O.TET, O.HEX])
def test_shape_partition_kronecker_and_gradient(cell):
    ref = {O.LINE: [[-1], [1]], O.TRI: [[0, 0], [1, 0], [0, 1]], O.QUAD: [[-1, -1], [1, -1], [1, 1], [-1, 1]],
           O.TET: [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
           O.HEX: [[x, y, z] for z in (-1, 1) for (x, y) in ((-1, -1), (1, -1), (1, 1), (-1, 1))]}[cell]
    for i, xi in enumerate(ref):
        np.testing.assert_allclose(O.shape(cell, xi), np.eye(len(ref))[i], atol=1e-15)
    xi = np.array([0.13, 0.21, 0.17])[: O.DIMS[cell]]
    assert abs(O.shape(cell, xi).sum() - 1.0) < 1e-14
    g = O.grad(cell, xi)
    for d in range(O.DIMS[cell]):
        e = np.zeros(len(xi))
        e[d] = 1e-6
        fd = (O.shape(cell, xi + e) - O.shape(cell, xi - e)) / 2e-6
        np.testing.assert_allclose(g[:, d], fd, atol=1e-9)


# ---- Map ---------------------------------------------------------------------
def test_jacobian_known_answers():
    k = KATS["jacobian_dets"]
    for c in k["cases"]:
        cell = CELL[c["cell"]]
        J, det, inv = O.map_jacobian(np.array(c["nodes"], float), O.grad(cell, c["xi"]))
        assert abs(det - c["det"]) < k["tol"]
        if "jac" in c:
            np.testing.assert_allclose(J.ravel(), c["jac"], atol=k["tol"])
        if "pgrad" in c:
            np.testing.assert_allclose(np.array(c["pgrad_in"]) @ inv, c["pgrad"], atol=k["tol"])


def test_inverse_is_the_inverse_not_its_transpose():
    m = np.array([[2.0, 0.7, -0.3], [0.1, 3.0, 0.5], [-0.4, 0.2, 1.5]])
    assert abs(O.mat_det(m) - np.linalg.det(m)) < 1e-13
    np.testing.assert_allclose(O.mat_inv(m), np.linalg.inv(m), rtol=1e-14)
    np.testing.assert_allclose(O.mat_inv(m[:2, :2]), np.linalg.inv(m[:2, :2]), rtol=1e-14)


# ---- element operators -------------------------------------------------------
def test_tri3_conductivity_known_answer():
    k = KATS["tri3_conductivity"]
    K = O.poisson_element_matrix(O.TRI, np.array(k["nodes"], float))
    np.testing.assert_allclose(K, k["k"], atol=k["tol"])


def test_tet_poisson_matrix_closed_form():
    rng = np.random.default_rng(0)
    X = rng.uniform(0, 1, (4, 3))
    K = O.poisson_element_matrix(O.TET, X, kappa=2.5)
    A = np.hstack([np.ones((4, 1)), X])
    G = np.linalg.inv(A)[1:].T  # rows: gradients of the barycentric functions
    vol = abs(np.linalg.det(A)) / 6
    np.testing.assert_allclose(K, 2.5 * vol * G @ G.T, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(O.poisson_source_vector(O.TET, X, 3.0), np.full(4, 3.0 * vol / 4), rtol=1e-14)


def test_elasticity_tet_rigid_modes_symmetry_and_closed_form():
    rng = np.random.default_rng(1)
    X = rng.uniform(0, 1, (4, 3))
    E, nu = 2.0, 0.3
    K = O.elasticity_element_matrix(O.TET, X, O.CONT3D, E, nu)
    np.testing.assert_allclose(K, K.T, rtol=1e-12, atol=1e-14)
    modes = []
    for t in np.eye(3):
        modes.append(np.tile(t, 4))
    for w in np.eye(3):
        modes.append(np.cross(w, X).ravel())
    for mvec in modes:
        assert np.abs(K @ mvec).max() < 1e-12 * np.abs(K).max()
    # textbook B^T D B V with independent B (Voigt xx,yy,zz,xy,yz,zx)
    A = np.hstack([np.ones((4, 1)), X])
    G = np.linalg.inv(A)[1:].T
    vol = abs(np.linalg.det(A)) / 6
    B = np.zeros((6, 12))
    for a in range(4):
        gx, gy, gz = G[a]
        B[:, 3 * a:3 * a + 3] = [[gx, 0, 0], [0, gy, 0], [0, 0, gz], [gy, gx, 0], [0, gz, gy], [gz, 0, gx]]
    c = E / ((1 + nu) * (1 - 2 * nu))
    D = np.zeros((6, 6))
    D[:3, :3] = c * nu
    np.fill_diagonal(D[:3, :3], c * (1 - nu))
    D[3:, 3:] = np.eye(3) * c * (1 - 2 * nu) / 2
    np.testing.assert_allclose(K, vol * B.T @ D @ B, rtol=1e-11, atol=1e-13)


def test_four_equal_contributions_sum_is_one_multiplication():
    """The GPU tet kernels write 4 c where the reference adds a P1 tet's four
    equal quadrature contributions, ((c + c) + c) + c: bitwise equal in binary
    round-to-nearest-even.  Checked on every half-precision value (the
    argument does not depend on the format; numpy's float16 sums here are
    exact in float32, then rounded once), on random doubles and on the doubles
    where fl(3c) is a tie (significand = 2 mod 4, c's significand >= 4/3)."""
    h = np.arange(1 << 16, dtype=np.uint16).view(np.float16)
    h = h[np.isfinite(h)]
    with np.errstate(over="ignore"):
        s = ((h + h) + h) + h
        np.testing.assert_array_equal(s.view(np.uint16), (h * np.float16(4)).view(np.uint16))
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 1 << 63, 2_000_000, dtype=np.int64).astype(np.uint64)
    sig = rng.integers(1 << 50, 1 << 52, 500_000, dtype=np.int64).astype(np.uint64)
    sig = (sig & ~np.uint64(3)) | np.uint64(2)  # significand = 2 mod 4
    sig = sig[sig >= np.uint64((1 << 52) // 3)]  # c's significand >= 4/3
    ties = sig | (rng.integers(1, 2046, sig.size, dtype=np.int64).astype(np.uint64) << np.uint64(52))
    c = np.concatenate([bits, ties, bits | np.uint64(1 << 63)]).view(np.float64)
    c = c[np.isfinite(c)]
    with np.errstate(over="ignore"):
        s = ((c + c) + c) + c
        np.testing.assert_array_equal(s.view(np.uint64), (4.0 * c).view(np.uint64))
    t = ties.view(np.float64)  # fl(3 t) really is a tie: 3 M needs two bits more
    M = (sig & np.uint64((1 << 52) - 1)) | np.uint64(1 << 52)
    assert np.all((np.uint64(3) * M) % np.uint64(4) == np.uint64(2)) and np.all(np.uint64(3) * M >= np.uint64(1 << 54))
    assert t.size > 100_000


def test_q4_simp_unit_stiffness_matches_textbook():
    nu = 0.3
    k = np.array([1 / 2 - nu / 6, 1 / 8 + nu / 8, -1 / 4 - nu / 12, -1 / 8 + 3 * nu / 8, -1 / 4 + nu / 12,
                  -1 / 8 - nu / 8, nu / 6, 1 / 8 - 3 * nu / 8])
    idx = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2], [2, 7, 0, 5, 6, 3, 4, 1], [3, 6, 5, 0, 7, 2, 1, 4],
           [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6], [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]]
    KE = k[np.array(idx)] / (1 - nu * nu)
    np.testing.assert_allclose(O.simp_ke_unit(nu, 1.0), KE, atol=1e-15)
    np.testing.assert_allclose(O.simp_ke_unit(nu, 0.37), KE, atol=1e-14)  # 2-D: independent of h
    # and equals the generic plane-stress Q4 operator on a unit square
    sq = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    np.testing.assert_allclose(O.elasticity_element_matrix(O.QUAD, sq, O.PSTRESS, 1.0, nu), KE, atol=1e-15)


def test_hex_elasticity_rigid_modes():
    X = np.array([[x, y, z] for z in (0, 1) for (x, y) in ((0, 0), (1, 0), (1, 1), (0, 1))], float)
    X += np.random.default_rng(2).uniform(-0.05, 0.05, X.shape)
    K = O.elasticity_element_matrix(O.HEX, X, O.CONT3D, 1.0, 0.3)
    for w in np.eye(3):
        assert np.abs(K @ np.cross(w, X).ravel()).max() < 1e-12
    assert np.abs(K @ np.tile([1.0, 0, 0], 8)).max() < 1e-13


# ---- sparse ------------------------------------------------------------------
@pytest.mark.parametrize("name", ["coo_to_csr", "coo_doctest"])
def test_to_csr_known_answers(name):
    k = KATS[name]
    csr = O.to_csr(O.Coo(k["rows"], k["cols"], k["vals"]))
    np.testing.assert_array_equal(csr.row_ptrs, k["row_ptrs"])
    np.testing.assert_array_equal(csr.values, k["values"])
    if "col_ind" in k:
        np.testing.assert_array_equal(csr.col_ind, k["col_ind"])


def test_to_csr_stable_insertion_order_sum():
    # (1e16 + 1) - 1e16 depends on the order: insertion order must be kept
    csr = O.to_csr(O.Coo([0, 0, 0], [0, 0, 0], [1e16, 1.0, -1e16]))
    assert csr.values[0] == (1e16 + 1.0) - 1e16
    csr = O.to_csr(O.Coo([0, 0, 0], [0, 0, 0], [1.0, 1e16, -1e16]))
    assert csr.values[0] == 0.0
    e = O.to_csr(O.Coo())
    assert (e.nrows, e.ncols, list(e.row_ptrs)) == (1, 1, [0, 0])


@pytest.mark.parametrize("name", ["solve_2x2", "solve_3x3"])
def test_dense_solve_known_answers(name):
    k = KATS[name]
    x = O.solve(O.Coo(k["rows"], k["cols"], k["vals"]), k["rhs"])
    np.testing.assert_allclose(x, k["x"], atol=k["tol"])


def test_dense_solve_matches_lapack():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 40)) + 40 * np.eye(40)
    b = rng.standard_normal(40)
    r, c = np.nonzero(A)
    x = O.solve(O.Coo(r, c, A[r, c]), b)
    np.testing.assert_allclose(x, linalg.solve(A, b), rtol=1e-12)
    with pytest.raises(O.SolveError):
        O.solve(O.Coo([0, 1], [0, 1], [1.0, 0.0]), [1.0, 1.0])


def test_matvec_known_answer():
    k = KATS["matvec"]
    y = O.to_csr(O.Coo(k["rows"], k["cols"], k["vals"])).matvec(k["x"])
    np.testing.assert_allclose(y, k["y"], atol=k["tol"])
    np.testing.assert_array_equal(O.to_csr(O.Coo()).matvec([1.0, 2.0, 3.0]), [0.0, 0.0, 0.0])


# ---- assembly / reduction ------------------------------------------------------
def test_assemble_skips_exact_zeros_and_dirichlet_known_answer():
    k = KATS["dirichlet_line"]
    coords, conn = np.array(k["coords"]), np.array(k["conn"])
    coo = O.assemble(coords, conn, O.make_op(O.LINE, 1, 0), 1)
    u = O.solve_with_dirichlet(coo, np.zeros(3), [tuple(b) for b in k["bcs"]])
    np.testing.assert_allclose(u, k["u"], atol=k["tol"])
    # Kuhn tets: 6 of 16 Poisson entries per tet are exact zeros
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (2, 2, 2))
    assert len(O.assemble(c, e, O.make_op(O.TET, 3, 0), 1).rows) == 10 * len(e)


def test_reduce_system_lifting_first_match_and_scatter_last():
    coo = O.Coo([0, 0, 1, 1, 1, 2, 2], [0, 1, 0, 1, 2, 1, 2], [2, -1, -1, 2, -1, -1, 2])
    bcs = [(0, 1.0), (2, 3.0), (0, 5.0)]
    kred, fred, free = O.reduce_system(coo, [0.0, 0.5, 0.0], bcs)
    np.testing.assert_array_equal(free, [1])
    assert fred[0] == 0.5 + 1.0 + 3.0  # first value of DOF 0 lifts
    u = O.solve_with_dirichlet(coo, [0.0, 0.5, 0.0], bcs)
    assert u[0] == 5.0 and abs(u[1] - 2.25) < 1e-15


def test_neumann_total_load_and_boundary_faces():
    k = KATS["neumann_total"]
    c, e = np.array(k["coords"], float), np.array(k["conn"])
    rhs = O.apply_neumann(c, O.TRI, e, 1, 1.0, np.zeros(3))
    assert abs(rhs.sum() - k["total"]) < k["tol"]
    bc, be = O.box_mesh([0, 0, 0], [1, 1, 1], (3, 2, 2))
    assert len(O.boundary_faces(O.TET, be)) == 2 * 2 * (3 * 2 + 3 * 2 + 2 * 2)
    t = O.plane_traction(bc, O.TET, be, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    assert abs(t[2::3].sum() + 1e-3) < 1e-15 and not t[0::3].any()


def test_box_mesh_counts_orientation_and_numbering():
    k = KATS["box_mesh_2"]
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (2, 2, 2))
    assert len(c) == k["nodes"] and len(e) == k["tets"]
    vols = [np.linalg.det(np.array([c[t[1]] - c[t[0]], c[t[2]] - c[t[0]], c[t[3]] - c[t[0]]])) for t in e]
    assert min(vols) > 0
    c, e = O.box_mesh([0, 0, 0], [2, 3, 4], (2, 3, 4))
    i, j, kk = 1, 2, 3
    np.testing.assert_array_equal(c[(i * 4 + j) * 5 + kk], [1.0, 2.0, 3.0])


# ---- drivers against closed forms / reference KATs -------------------------------
def test_c1_poisson_equals_dst_closed_form():
    g = json.load(open(os.path.join(GOLD, "c1_fd_solution.json")))
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (8, 8, 8))
    bcs = [(int(n), 0.0) for a in range(3) for v in (0.0, 1.0) for n in O.nodes_on_plane(c, a, v, 1e-9)]
    u = O.solve_poisson(c, O.TET, e, 1.0, 2, 1.0, bcs)
    interior = u.reshape(9, 9, 9)[1:-1, 1:-1, 1:-1].ravel()
    np.testing.assert_allclose(interior, g["interior_u"], rtol=0, atol=1e-15)
    # the reduced matrix is h*(6I - neighbours): diag 0.75, off-diagonal
    # -0.125 in exact arithmetic; the reference's IEEE order (w = 1/24 is not
    # representable) lands within a few ulp of those values
    coo = O.assemble(c, e, O.make_op(O.TET, 3, 0), 1)
    kr, fr, free = O.reduce_system(coo, O.poisson_rhs(c, O.TET, e), bcs)
    k = kr.to_csr()
    assert k.nnz() == 2107
    diag = k.values[k.col_ind == np.repeat(np.arange(343), np.diff(k.row_ptrs))]
    off = k.values[k.col_ind != np.repeat(np.arange(343), np.diff(k.row_ptrs))]
    assert len(diag) == 343 and np.abs(diag - 0.75).max() <= 4e-16
    assert np.abs(off + 0.125).max() <= 1e-16
    np.testing.assert_allclose(fr, np.full(343, 1 / 512), rtol=1e-15)


def test_stale_reference_vtk_is_not_reproduced():
    s = json.load(open(os.path.join(GOLD, "stale_vtk.json")))
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (6, 6, 6))
    bcs = [(int(n), 0.0) for a in range(3) for v in (0.0, 1.0) for n in O.nodes_on_plane(c, a, v, 1e-9)]
    u = O.solve_poisson(c, O.TET, e, 1.0, 2, 1.0, bcs)
    assert len(c) == s["points"] and len(e) == s["cells"]
    assert abs(u.max() - s["temperature_max"]) > 1e-3


def test_kuhn_elasticity_pattern_counts():
    g = json.load(open(os.path.join(GOLD, "kuhn_counts.json")))
    for n in (2, 4):
        c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (n, n, n))
        coo = O.assemble(c, e, O.make_op(O.TET, 3, 1, nu=0.3), 3)
        assert coo.to_csr().nnz() == g[str(n)]["full"]
        bcs = [(int(i) * 3 + k, 0.0) for i in O.nodes_on_plane(c, 0, 0.0, 1e-9) for k in range(3)]
        kr, _, _ = O.reduce_system(coo, np.zeros(len(c) * 3), bcs)
        assert kr.to_csr().nnz() == g[str(n)]["reduced_x0_clamped"]


def test_reference_solver_known_answers():
    k = KATS["poisson_2d_linear"]
    u = O.solve_poisson(np.array(k["coords"], float), O.TRI, np.array(k["conn"]), 1.0, 2, 0.0,
                        [tuple(b) for b in k["bcs"]])
    assert abs(u[4] - k["u_mid"]) < k["tol"]
    k = KATS["bar_axial"]
    c, e = np.array(k["coords"]), np.array(k["conn"])
    coo = O.assemble(c, e, O.make_op(O.LINE, 1, 1, model=O.BAR, coef=1.0, nu=0.0, order=1), 1)
    u = O.solve_with_dirichlet(coo, np.zeros(3), [tuple(b) for b in k["bcs"]])
    np.testing.assert_allclose(u, k["u"], atol=k["tol"])
    k = KATS["plane_stress_patch"]
    c, e = np.array(k["coords"], float), np.array(k["conn"])
    coo = O.assemble(c, e, O.make_op(O.TRI, 2, 1, model=O.PSTRESS, nu=0.3), 2)
    u = O.solve_with_dirichlet(coo, np.zeros(10), [tuple(b) for b in k["bcs"]])
    np.testing.assert_allclose(u[8:10], k["u_mid"], atol=k["tol"])
    k = KATS["poisson_1d"]
    n = k["nx"]
    c = np.linspace(0, 1, n + 1)[:, None]
    e = np.array([[i, i + 1] for i in range(n)])
    u = O.solve_poisson(c, O.LINE, e, 1.0, 2, 1.0, [(0, 0.0), (n, 0.0)])
    assert abs(u[n // 2] - k["u_mid"]) < k["tol"]


def test_planted_transposed_inverse_is_caught():
    """A transposed J^-1 (the reference's old bug) must fail the closed-form
    element check; restated here in numpy so the oracle itself is untouched."""
    rng = np.random.default_rng(9)
    X = rng.uniform(0, 1, (4, 3))
    J = (X[1:] - X[0]).T
    lg = O.grad(O.TET, [0.2, 0.2, 0.2])
    good = lg @ np.linalg.inv(J)
    bad = lg @ np.linalg.inv(J).T
    A = np.hstack([np.ones((4, 1)), X])
    G = np.linalg.inv(A)[1:].T
    assert np.allclose(good, G) and not np.allclose(bad, G)


# ---- SIMP outer loop (crates/tpt-fem-topopt/src/lib.rs:241-500) ------------------
def _simp(nx, ny, vf, iters):
    f, bcs = O.cantilever_load(nx, ny, 1.0)
    return O.topopt_simp(nx, ny, 1.0, 1.0, 0.3, vf, 3.0, 2.0, iters, 0.2, f, bcs)


def test_cantilever_load_layout():
    """cantilever_load (:398-416): both DOFs of every left-edge node fixed to
    zero, -load on the y DOF of node nx-1 (bottom right), nothing else."""
    f, bcs = O.cantilever_load(5, 3, 2.5)
    assert bcs == [(0, 0.0), (1, 0.0), (10, 0.0), (11, 0.0), (20, 0.0), (21, 0.0)]
    assert f[9] == -2.5 and np.count_nonzero(f) == 1 and len(f) == 30


def test_simp_reference_properties():
    """The reference's three tests (:421-495): compliance falls from the
    uniform start, never rises by more than 1e-9, the volume holds to 1e-6;
    full volume stays solid; a 40 % optimized design beats the uniform one."""
    rho, c, it = _simp(20, 10, 0.5, 40)
    assert it == 40 and len(rho) == 19 * 9 and len(c) == 41
    assert c[-1] < c[0] and np.all(np.diff(c) <= 1e-9)
    assert abs(rho.sum() - 0.5 * len(rho)) < 1e-6
    # index 0 and 1 solve the same uniform design
    assert c[0] == c[1]
    rho, c, _ = _simp(16, 8, 1.0, 10)
    assert abs(rho.sum() - len(rho)) < 1e-6 and np.all(np.abs(rho - 1.0) < 1e-6)
    _, c, _ = _simp(24, 12, 0.4, 30)
    assert c[-1] < c[0]


def test_simp_quad_form_matches_numpy():
    """element_quad_form (:241-254): u_e^T K0 u_e with u_e = (u[2n], u[2n+1])
    per node, against numpy's product (other summation order)."""
    rng = np.random.default_rng(3)
    ke0 = O.simp_ke_unit(0.3, 1.0)
    u = rng.standard_normal(2 * 12)
    nodes = np.array([1, 2, 6, 5])
    ue = np.ravel([[u[2 * n], u[2 * n + 1]] for n in nodes])
    assert abs(O.topopt_quad_form(ke0, u, nodes) - ue @ ke0 @ ue) <= 1e-13 * abs(ue @ ke0 @ ue)
    # a rigid translation carries no strain energy
    t = np.tile([0.7, -0.2], 12)
    assert abs(O.topopt_quad_form(ke0, t, nodes)) < 1e-15


def test_simp_filter_closed_forms():
    """filter_sensitivity (:259-279) on a 9 x 7 node grid (8 x 6 elements):
    radius 0.5 < h keeps only the element itself (exactly); uniform input is
    a fixed point; an interior element at radius 2 weighs itself 2, its four
    edge neighbours 1, its diagonals 2 - sqrt(2) and the distance-2 ones 0."""
    rng = np.random.default_rng(5)
    raw = rng.standard_normal(8 * 6)
    assert np.array_equal(O.topopt_filter(9, 7, 1.0, raw, 0.5), raw)
    np.testing.assert_allclose(O.topopt_filter(9, 7, 1.0, np.full(48, -3.25), 2.0), -3.25, rtol=1e-15)
    out = O.topopt_filter(9, 7, 1.0, raw, 2.0)
    R = raw.reshape(6, 8)
    i, j = 3, 2  # element (ei, ej) = (3, 2)
    d = 2.0 - np.sqrt(2.0)
    num = 2 * R[j, i] + R[j, i - 1] + R[j, i + 1] + R[j - 1, i] + R[j + 1, i] + d * (
        R[j - 1, i - 1] + R[j - 1, i + 1] + R[j + 1, i - 1] + R[j + 1, i + 1])
    assert abs(out[j * 8 + i] - num / (6 + 4 * d)) < 1e-14
    # corner element (0, 0): itself, two edge neighbours, one diagonal
    num = 2 * R[0, 0] + R[0, 1] + R[1, 0] + d * R[1, 1]
    assert abs(out[0] - num / (4 + d)) < 1e-14


def test_simp_oc_closed_form_and_bounds():
    """oc_update (:284-331): with no bound active the update is
    rho_e sqrt(B_e / lambda) with lambda fixing the volume, i.e.
    target rho_e sqrt(B_e) / sum(rho sqrt(B)); otherwise the move limit and
    [1e-3, 1] hold and the volume meets the target to bisection accuracy."""
    rng = np.random.default_rng(11)
    n = 200
    rho = np.full(n, 0.5)
    b = rng.uniform(0.9, 1.1, n)
    out = O.topopt_oc(rho, -b, 0.5, 0.2)
    w = rho * np.sqrt(b)
    np.testing.assert_allclose(out, 0.5 * n * w / w.sum(), rtol=1e-8)
    rho = rng.uniform(0.01, 1.0, n)
    s = -np.exp(rng.standard_normal(n) * 2)
    s[:7] = 0.0  # no gain: ratio 0, pushed to the lower bound
    out = O.topopt_oc(rho, s, 0.3, 0.2)
    assert abs(out.sum() - 0.3 * n) < 1e-6
    assert np.all(out >= np.maximum(rho - 0.2, 1e-3) - 1e-15) and np.all(out <= np.minimum(rho + 0.2, 1.0) + 1e-15)
    np.testing.assert_array_equal(out[:7], np.maximum(rho[:7] - 0.2, 1e-3))



# ---- consistent / lumped mass (crates/tpt-fem-elasticity/src/lib.rs:386-457) ------
def test_mass_element_closed_forms():
    """Exact-integration closed forms: P1 tet rho V/20 (1 + d_ij), Tri3
    rho A/12 (1 + d_ij), Line2 rho L/6 (1 + d_ij), and the tensor products
    rho V/6^d * 2^(#shared coordinates) for axis-aligned Q4/Hex8 at order 2;
    each copied on the component diagonal only (nothing across components)."""
    rng = np.random.default_rng(21)
    X = np.array([[0, 0, 0], [1.1, 0.1, 0], [0.2, 0.9, 0.1], [0.3, 0.2, 1.2]]) + rng.uniform(0, 0.1, (4, 3))
    V = abs(np.linalg.det((X[1:] - X[0]).T)) / 6
    k = O.mass_element_matrix(O.TET, X, 2.5, 2)
    np.testing.assert_allclose(k[0::3, 0::3], 2.5 * V / 20 * (1 + np.eye(4)), rtol=1e-14)
    assert np.count_nonzero(k) == 48 and np.array_equal(k[0::3, 0::3], k[2::3, 2::3])
    T3 = np.array([[0.1, 0.2], [1.3, 0.1], [0.4, 0.8]])
    A = abs(np.linalg.det(T3[1:] - T3[0])) / 2
    np.testing.assert_allclose(O.mass_element_matrix(O.TRI, T3, 3.0, 2)[0::2, 0::2], 3.0 * A / 12 * (1 + np.eye(3)),
                               rtol=1e-14)
    np.testing.assert_allclose(O.mass_element_matrix(O.LINE, np.array([[0.5], [2.0]]), 2.0, 2),
                               2.0 * 1.5 / 6 * (1 + np.eye(2)), rtol=1e-15)
    a, b, c = 0.7, 1.3, 0.4
    Q = np.array([[0, 0], [a, 0], [a, b], [0, b]])
    rq = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])
    shared = (rq[:, None, :] == rq[None, :, :]).sum(-1)
    np.testing.assert_allclose(O.mass_element_matrix(O.QUAD, Q, 1.5, 2)[0::2, 0::2], 1.5 * a * b / 36 * 2.0 ** shared,
                               rtol=1e-14)
    rh = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    H = rh * [a, b, c]
    shared = (rh[:, None, :] == rh[None, :, :]).sum(-1)
    np.testing.assert_allclose(O.mass_element_matrix(O.HEX, H, 1.5, 2)[0::3, 0::3],
                               1.5 * a * b * c / 216 * 2.0 ** shared, rtol=1e-14)


def test_mass_matrix_total_and_lumped():
    """Sum of the consistent mass = rho |Omega| per component; the lumped
    diagonal holds each node's row sums over dim (restated in numpy)."""
    c, e = O.box_mesh([0, 0, 0], [1.0, 2.0, 0.5], (3, 2, 2))
    c = O.jitter_interior(c, (3, 2, 2), 0.2, 5)
    coo = O.mass_matrix(c, O.TET, e, 3.0)
    assert abs(coo.vals.sum() - 3.0 * 1.0 * 3) < 1e-13
    csr = coo.to_csr()
    lump = O.lumped_mass(c, O.TET, e, 3.0)
    rows = np.repeat(np.arange(csr.nrows), np.diff(csr.row_ptrs))
    per_node = np.bincount(rows // 3, weights=csr.values)
    np.testing.assert_allclose(lump.vals, np.repeat(per_node / 3, 3), rtol=1e-14)
    assert np.array_equal(lump.rows, np.arange(len(c) * 3))


# ---- Lanczos eigenpairs (crates/tpt-fem-eigen/src/lib.rs) -------------------------
def _laplace_1d(n):
    r, c, v = [], [], []
    for i in range(n):
        r.append(i); c.append(i); v.append(2.0)
        if i + 1 < n:
            r += [i, i + 1]; c += [i + 1, i]; v += [-1.0, -1.0]
    return O.Coo(r, c, v).to_csr()


def test_lanczos_reference_kat_and_closed_form():
    """lanczos_1d_laplacian (:595-623): the extreme eigenvalues of the n = 10
    discrete Laplacian, 2 - 2 cos(k pi / 11), to 1e-6 (and in fact 1e-13);
    with a full basis every eigenvalue and unit eigenvector is recovered."""
    n = 10
    a = _laplace_1d(n)
    lo = O.lanczos_eigs(a, 1, O.SMALLEST, n)[0][0]
    hi = O.lanczos_eigs(a, 1, O.LARGEST, n)[0][0]
    assert abs(lo - (2 - 2 * np.cos(np.pi / 11))) < 1e-13
    assert abs(hi - (2 - 2 * np.cos(np.pi * 10 / 11))) < 1e-13
    pairs = O.lanczos_eigs(a, n, O.SMALLEST, n)
    k = np.arange(1, n + 1)
    np.testing.assert_allclose([p[0] for p in pairs], 2 - 2 * np.cos(k * np.pi / 11), atol=1e-13)
    dense = np.diag(np.full(n, 2.0)) - np.eye(n, k=1) - np.eye(n, k=-1)
    for lam, vec in pairs:
        assert abs(np.linalg.norm(vec) - 1) < 1e-14 and np.linalg.norm(dense @ vec - lam * vec) < 1e-12


def test_jacobi_matches_lapack():
    """jacobi (:170-218) against LAPACK's symmetric eigensolver; A Q = Q diag."""
    rng = np.random.default_rng(17)
    for m in (1, 2, 5, 12):
        b = rng.standard_normal((m, m))
        a = b + b.T
        eig, q = O.eig_jacobi(a)
        np.testing.assert_allclose(np.sort(eig), np.linalg.eigvalsh(a), atol=1e-12)
        np.testing.assert_allclose(a @ q, q * eig, atol=1e-11)


def test_lanczos_partial_basis_and_invariant_subspace():
    """A partial basis converges the ends of the spectrum first (Ritz values
    bracketed by the true ones); a diagonal matrix is an invariant subspace of
    e_0, so the basis stops after one vector and one pair comes back."""
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (4, 4, 4))
    a = O.assemble(c, e, O.make_op(O.TET, 3, 0, coef=1.0), 1).to_csr()
    dense = np.zeros((a.nrows, a.nrows))
    for r in range(a.nrows):
        dense[r, a.col_ind[a.row_ptrs[r]:a.row_ptrs[r + 1]]] = a.values[a.row_ptrs[r]:a.row_ptrs[r + 1]]
    # the pure Neumann Laplacian is singular: its smallest eigenvalue is 0
    w = np.linalg.eigvalsh(dense)
    hi = O.lanczos_eigs(a, 3, O.LARGEST, 40)
    assert abs(hi[0][0] - w[-1]) < 1e-9 * w[-1] and hi[0][0] <= w[-1] * (1 + 1e-14)
    assert hi[0][0] >= hi[1][0] >= hi[2][0]
    d = O.Coo([0, 1, 2], [0, 1, 2], [3.0, 1.0, 2.0]).to_csr()
    pairs = O.lanczos_eigs(d, 3, O.SMALLEST, 3)
    assert len(pairs) == 1 and pairs[0][0] == 3.0


# ---- time integration (crates/tpt-fem-dynamic/src/lib.rs) -------------------------
def _sdof(mass, stiff):
    return (O.Coo([0], [0], [mass]).to_csr(), O.Coo([], [], []).to_csr(), O.Coo([0], [0], [stiff]).to_csr())


def test_dynamic_reference_tests():
    """The reference's three tests (:507-550) on the oracle."""
    m, c, k = _sdof(1.0, 4.0)
    h = O.newmark(m, c, k, [1.0], [0.0], [0.0], 0.005, 0.25, 0.5, 400)
    assert abs(h[400][0] - np.cos(4.0)) < 5e-3
    with pytest.raises(O.CflViolation):
        O.central_difference(m, c, k, [1.0], [0.0], [0.0], 2.0, 10)
    m, c, k = _sdof(2.0, 8.0)
    h = O.newmark(m, c, k, [1.0], [0.0], [0.0], 0.002, 0.25, 0.5, 500)
    for w in range(1, len(h) - 1):
        vm = (h[w + 1][0] - h[w - 1][0]) / (2 * 0.002)
        assert abs(0.5 * 2 * vm * vm + 0.5 * 8 * h[w][0] ** 2 - 4.0) < 1e-2


def test_dynamic_discrete_closed_forms():
    """Exact discrete solutions of the undamped oscillator from u0 = 1, v0 = 0:
    central difference u_k = cos(k t) with cos t = 1 - w^2 dt^2 / 2, and the
    average-acceleration Newmark (the trapezoidal rule) u_k = cos(k t') with
    tan(t'/2) = w dt / 2; and for two coupled DOFs the same per mode of the
    generalized eigenproblem (scipy)."""
    w, dt, N = 3.0, 0.05, 80
    m, c, k = _sdof(2.0, 2.0 * w * w)
    kk = np.arange(N + 1)
    cd = O.central_difference(m, c, k, [1.0], [0.0], [0.0], dt, N)[:, 0]
    np.testing.assert_allclose(cd, np.cos(kk * np.arccos(1 - w * w * dt * dt / 2)), atol=1e-12)
    nm = O.newmark(m, c, k, [1.0], [0.0], [0.0], dt, 0.25, 0.5, N)[:, 0]
    np.testing.assert_allclose(nm, np.cos(kk * 2 * np.arctan(w * dt / 2)), atol=1e-12)
    Md = np.array([[2.0, 0.5], [0.5, 1.0]])
    Kd = np.array([[6.0, -2.0], [-2.0, 4.0]])
    tocsr = lambda A: O.Coo(*np.nonzero(A), A[np.nonzero(A)]).to_csr()
    u0 = np.array([1.0, -0.5])
    lam, phi = linalg.eigh(Kd, Md)  # phi^T M phi = I
    q0 = phi.T @ Md @ u0
    th = 2 * np.arctan(np.sqrt(lam) * dt / 2)
    want = np.array([phi @ (q0 * np.cos(j * th)) for j in kk])
    got = O.newmark(tocsr(Md), O.Coo([], [], []).to_csr(), tocsr(Kd), u0, [0.0, 0.0], [0.0, 0.0], dt, 0.25, 0.5, N)
    np.testing.assert_allclose(got, want, atol=1e-12)


# ---- Robin terms (crates/tpt-fem-assembly/src/lib.rs:684-753) ----------------------
def test_robin_reference_kat_and_face_mass():
    """robin_conserves_total_measure (:834-851): h = 1 on the unit right
    triangle sums to its perimeter 2 + sqrt(2).  On a tet box each boundary
    face contributes the exact P1 face mass h A / 12 (1 + d_kl) (the Degree2
    rule integrates it exactly), the whole boundary h * area, and a band on
    x = 1 only that face's share."""
    tc, te = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]), np.array([[0, 1, 2]])
    assert abs(O.apply_robin(tc, O.TRI, te, 1, 1.0).vals.sum() - (2 + np.sqrt(2))) < 1e-12
    c, e = O.box_mesh([0, 0, 0], [1.0, 2.0, 1.5], (2, 2, 3))
    r = O.apply_robin(c, O.TET, e, 1, 2.0)
    area = 2 * (1.0 * 2.0 + 1.0 * 1.5 + 2.0 * 1.5)
    assert abs(r.vals.sum() - 2.0 * area) < 1e-12
    # the first boundary face: 3 points x 9 node pairs, consecutive
    f0 = r.vals[:27].reshape(3, 3, 3).sum(0)
    rows0 = r.rows[:9].reshape(3, 3)[:, 0]
    X = c[rows0]
    A = 0.5 * np.linalg.norm(np.cross(X[1] - X[0], X[2] - X[0]))
    np.testing.assert_allclose(f0, 2.0 * A / 12 * (1 + np.eye(3)), rtol=1e-14)
    band = O.apply_robin(c, O.TET, e, 1, 2.0, axis=0, coord=1.0, tol=1e-9)
    assert len(band.vals) == len(r.vals) and abs(band.vals.sum() - 2.0 * 2.0 * 1.5) < 1e-12


# ---- legacy binary VTK writer (crates/tpt-fem-io-vtk/src/lib.rs:117-170) -------------
VTK_DIR = os.path.join(GOLD, "vtk")
from vtk_util import read_legacy_vtk  # noqa: E402
@pytest.mark.parametrize("name", ["mesh_gen_box_structured.vtk", "mesh_gen_box_delaunay.vtk", "modal_analysis.vtk",
                                  "elasticity_frame.vtk", "abaqus_import.vtk", "thermal_solve.vtk"])
def test_vtk_writer_reproduces_reference_files(name):
    """The six files the reference's examples wrote (tests/golden/vtk, copied
    from its repository root) come back byte for byte from their own points,
    cells and fields: header lines, blank lines, big-endian blocks."""
    d = open(os.path.join(VTK_DIR, name), "rb").read()
    title, pts, conn, cell, fields = read_legacy_vtk(d)
    assert O.vtk_legacy_bytes(pts, cell, conn, fields, title) == d


def test_vtk_writer_of_box_mesh_is_the_reference_example():
    """mesh_gen_box.rs writes write_vtk(box_mesh([0,0,0],[1,1,1],[4,4,4])):
    the oracle's box mesh through the oracle's writer is that file."""
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (4, 4, 4))
    assert O.vtk_legacy_bytes(c, O.TET, e) == open(os.path.join(VTK_DIR, "mesh_gen_box_structured.vtk"), "rb").read()


# ---- closure-valued loads: quadrature points, source and body vectors ----------
def _tet_vol(x):
    return abs(np.linalg.det((x[1:] - x[0]).T)) / 6.0


def test_qp_coords_are_the_affine_images_of_the_rule_points():
    rng = np.random.default_rng(5)
    x = rng.random((4, 3))
    xq = O.qp_coords(O.TET, x, 2)
    p, _ = O.cell_quad(O.TET, 2)
    # independent: x0 + J xi with J = [x1-x0, x2-x0, x3-x0]
    ref = x[0] + p @ (x[1:] - x[0])
    np.testing.assert_allclose(xq, ref, atol=1e-15)
    np.testing.assert_allclose(xq.mean(axis=0), x.mean(axis=0), atol=1e-15)
    # Q4 at order 2: the four Gauss points of a rectangle [0,2]x[0,1]
    xr = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 1.0], [0.0, 1.0]])
    g = 1 / np.sqrt(3)
    got = sorted(map(tuple, np.round(O.qp_coords(O.QUAD, xr, 2), 14)))
    want = sorted((round(1 + s * g, 14), round(0.5 + t * g / 2, 14)) for s in (-1, 1) for t in (-1, 1))
    np.testing.assert_allclose(np.array(got), np.array(want), atol=1e-14)


def test_body_vector_constant_force_closed_forms():
    rng = np.random.default_rng(11)
    b = np.array([0.3, -1.7, 2.9])
    # tet: every node gets b * vol / 4
    x = rng.random((4, 3))
    f = O.elasticity_body_vector(O.TET, x, np.tile(b, (4, 1)), 2).reshape(4, 3)
    np.testing.assert_allclose(f, np.tile(b * _tet_vol(x) / 4, (4, 1)), rtol=1e-13)
    # affine Q4 (parallelogram): b * area / 4 per node
    a0, e1, e2 = rng.random(2), np.array([1.3, 0.2]), np.array([-0.4, 0.9])
    xq = np.array([a0, a0 + e1, a0 + e1 + e2, a0 + e2])
    area = abs(e1[0] * e2[1] - e1[1] * e2[0])
    f = O.elasticity_body_vector(O.QUAD, xq, np.tile(b[:2], (4, 1)), 2).reshape(4, 2)
    np.testing.assert_allclose(f, np.tile(b[:2] * area / 4, (4, 1)), rtol=1e-13)
    # affine Hex8 (parallelepiped): b * vol / 8 per node
    E = np.array([[1.1, 0.1, 0.0], [0.2, 0.9, 0.3], [-0.1, 0.2, 1.4]])
    cube = np.array([[x, y, z] for z in (0, 1) for (x, y) in ((0, 0), (1, 0), (1, 1), (0, 1))], float)
    xh = cube @ E.T + 0.5
    f = O.elasticity_body_vector(O.HEX, xh, np.tile(b, (8, 1)), 2).reshape(8, 3)
    np.testing.assert_allclose(f, np.tile(b * abs(np.linalg.det(E)) / 8, (8, 1)), rtol=1e-13)
    # Q4 is reached through the same code as Tri3 with dim 2 -- and a wrong
    # node index or dropped det fails the per-node equality above


def test_body_vector_non_affine_hex_total_is_force_times_volume():
    rng = np.random.default_rng(2)
    cube = np.array([[x, y, z] for z in (0, 1) for (x, y) in ((0, 0), (1, 0), (1, 1), (0, 1))], float)
    xh = cube + 0.15 * rng.standard_normal((8, 3))
    # volume of the trilinear map by a 6-point Gauss rule per axis (numpy), exact
    g, w = np.polynomial.legendre.leggauss(6)
    S = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    vol = 0.0
    for i, a in enumerate(g):
        for j, bq in enumerate(g):
            for k, c in enumerate(g):
                dN = np.stack([S[:, 0] * (1 + S[:, 1] * bq) * (1 + S[:, 2] * c),
                               (1 + S[:, 0] * a) * S[:, 1] * (1 + S[:, 2] * c),
                               (1 + S[:, 0] * a) * (1 + S[:, 1] * bq) * S[:, 2]], axis=1) / 8
                vol += w[i] * w[j] * w[k] * abs(np.linalg.det(xh.T @ dN))
    b = np.array([1.0, -2.0, 0.5])
    f = O.elasticity_body_vector(O.HEX, xh, np.tile(b, (8, 1)), 2).reshape(8, 3)
    np.testing.assert_allclose(f.sum(axis=0), b * vol, rtol=1e-12)


def test_linear_closures_integrate_exactly_on_tets():
    # int N_i (c . x) dV = (c . (sum_j x_j + x_i)) * V / 20 (degree-2 exact rule)
    rng = np.random.default_rng(7)
    x = rng.random((4, 3))
    V = _tet_vol(x)
    cs = np.array([[0.7, -0.2, 1.1], [0.1, 0.4, -0.9], [-1.3, 0.6, 0.2]])
    xq = O.qp_coords(O.TET, x, 2)
    bq = xq @ cs.T  # b_a(x) = cs[a] . x
    f = O.elasticity_body_vector(O.TET, x, bq, 2).reshape(4, 3)
    want = np.array([[cs[a] @ (x.sum(axis=0) + x[i]) * V / 20 for a in range(3)] for i in range(4)])
    np.testing.assert_allclose(f, want, rtol=1e-13)
    sq = xq @ cs[0]
    fs = O.poisson_source_vector_qp(O.TET, x, sq, 2)
    np.testing.assert_allclose(fs, want[:, 0], rtol=1e-13)
    # a constant closure reproduces the constant-source function bit for bit
    np.testing.assert_array_equal(O.poisson_source_vector_qp(O.TET, x, np.full(4, 2.5), 2),
                                  O.poisson_source_vector(O.TET, x, 2.5, 2))


def test_body_rhs_scatter_and_solve_elasticity_against_closed_form():
    # uniaxial bar: u_x(x) = -rho g (L x - x^2 / 2) / E under body force -rho g along x,
    # clamped at x = 0, free elsewhere, nu = 0 (no lateral coupling): P1 tets
    # reproduce the quadratic only approximately, so check the total load and
    # global equilibrium: sum of rhs = b * volume, K u = f at free DOFs.
    coords, conn = O.box_mesh([0, 0, 0], [1, 0.5, 0.25], [4, 2, 2])
    nq = 4
    b = np.array([0.0, 0.0, -9.81])
    bq = np.tile(b, (len(conn), nq, 1))
    rhs = O.body_rhs(coords, O.TET, conn, bq, 2)
    np.testing.assert_allclose(rhs.reshape(-1, 3).sum(axis=0), b * 0.125, rtol=1e-13)
    fixed = O.nodes_on_plane(coords, 0, 0.0, 1e-9)
    bcs = [(int(n) * 3 + a, 0.0) for n in fixed for a in range(3)]
    u = O.solve_elasticity(coords, O.TET, conn, O.CONT3D, 200.0, 0.3, 2, bq, bcs)
    op = O.make_op(O.TET, 3, 1, order=2, coef=200.0, nu=0.3)
    K = O.to_csr(O.assemble(coords, conn, op, 3))
    Ku = np.zeros(len(u))
    for r in range(K.nrows):
        Ku[r] = np.dot(K.values[K.row_ptrs[r]:K.row_ptrs[r + 1]], u[K.col_ind[K.row_ptrs[r]:K.row_ptrs[r + 1]]])
    free = np.setdiff1d(np.arange(len(u)), [d for d, _ in bcs])
    np.testing.assert_allclose(Ku[free], rhs[free], atol=1e-12 * np.abs(rhs).max())
    assert np.all(u[[d for d, _ in bcs]] == 0.0) and u[2::3].min() < 0


def test_boundary_points_normals_and_divergence_theorem():
    # on a box every boundary point sits on a face with the axis normal
    # pointing out; int_boundary x.n dA = 3 * volume (divergence theorem), which
    # an inward-oriented or unnormalised normal fails
    for cell, (coords, conn) in (
            (O.TET, O.box_mesh([0, 0, 0], [1.0, 0.5, 2.0], [3, 2, 4])),
            (O.QUAD, O.grid(5, 4, 0.5))):
        x, nrm = O.boundary_points(coords, cell, conn, 2)
        pdim = coords.shape[1]
        lo, hi = coords.min(axis=0), coords.max(axis=0)
        np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-15)
        for p, nv in zip(x, nrm):
            ax = int(np.argmax(np.abs(nv)))
            assert abs(abs(nv[ax]) - 1) < 1e-15
            side = hi[ax] if nv[ax] > 0 else lo[ax]
            assert abs(p[ax] - side) < 1e-14
        vol = np.prod(hi - lo)
        gq = np.einsum("ij,ij->i", x, nrm)
        rhs = O.apply_neumann_qp(coords, cell, conn, 1, gq, np.zeros(len(coords)), 2)
        np.testing.assert_allclose(rhs.sum(), pdim * vol, rtol=1e-13)
        # constant values reproduce the constant-flux pass bit for bit
        np.testing.assert_array_equal(
            O.apply_neumann_qp(coords, cell, conn, 2, np.full(len(x), 0.7), np.zeros(2 * len(coords)), 2),
            O.apply_neumann(coords, cell, conn, 2, 0.7, np.zeros(2 * len(coords)), 2))
        # Robin: sum over k, l of N_k N_l is 1, so the triplets total int h dA
        coo = O.apply_robin_qp(coords, cell, conn, 1, gq, 2)
        np.testing.assert_allclose(coo.vals.sum(), pdim * vol, rtol=1e-13)
        ref = O.apply_robin(coords, cell, conn, 1, 0.3)
        got = O.apply_robin_qp(coords, cell, conn, 1, np.full(len(x), 0.3), 2)
        np.testing.assert_array_equal(ref.rows, got.rows)
        np.testing.assert_array_equal(ref.vals, got.vals)


def test_boundary_points_on_a_jittered_mesh_and_higher_order():
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], [4, 4, 4])
    coords = O.jitter_interior(coords, [4, 4, 4], 0.2, 9)
    for order in (1, 3):
        x, nrm = O.boundary_points(coords, O.TET, conn, order)
        gq = np.einsum("ij,ij->i", x, nrm)
        rhs = O.apply_neumann_qp(coords, O.TET, conn, 1, gq, np.zeros(len(coords)), order)
        np.testing.assert_allclose(rhs.sum(), 3.0, rtol=1e-13)
        # a quadratic flux x_0^2 n_0: int = int_V 2 x_0 dV = 1 on the unit cube (order >= 2 exact on flat faces)
        if order == 3:
            rhs = O.apply_neumann_qp(coords, O.TET, conn, 1, x[:, 0] ** 2 * nrm[:, 0], np.zeros(len(coords)), order)
            np.testing.assert_allclose(rhs.sum(), 1.0, rtol=1e-13)


def _fma(a, b, c):
    """a * b + c rounded once (what a fused multiply-add or a matrix core's
    accumulate step does), exactly via rationals."""
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _hex_elasticity_ke(phys, young, nu, order, fused):
    """elasticity_element_matrix for a Hex8 in the reference's operation
    order (crates/tpt-fem-elasticity/src/lib.rs:262-309: D, then per point
    physical gradients, B, and s = sum_{s,t} B_i[s][a] D[s][t] B_j[t][b]
    from 0.0, k += w det s), the inner accumulation either separately
    rounded (the reference) or fused."""
    c = young / ((1.0 + nu) * (1.0 - 2.0 * nu))
    m, dd = (1.0 - nu) * c, (1.0 - 2.0 * nu) / 2.0 * c
    d = [[0.0] * 6 for _ in range(6)]
    for i in range(3):
        for j in range(3):
            d[i][j] = m if i == j else c * nu
    d[3][3] = d[4][4] = d[5][5] = dd
    qp, qw = O.cell_quad(O.HEX, order)
    k = [[0.0] * 24 for _ in range(24)]
    for q in range(len(qw)):
        lg = O.grad(O.HEX, qp[q])
        _, det, inv = O.map_jacobian(phys, lg)
        det = abs(det)
        b = []
        for a in range(8):
            g = []
            for i in range(3):
                s = 0.0
                for r in range(3):
                    s += float(lg[a][r]) * float(inv[r][i])
                g.append(s)
            bm = [[0.0] * 3 for _ in range(6)]
            bm[0][0], bm[1][1], bm[2][2] = g[0], g[1], g[2]
            bm[3][0], bm[3][1], bm[4][1], bm[4][2], bm[5][0], bm[5][2] = g[1], g[0], g[2], g[1], g[2], g[0]
            b.append(bm)
        for i in range(8):
            for j in range(8):
                for a in range(3):
                    for bb in range(3):
                        s = 0.0
                        for s_ in range(6):
                            for t in range(6):
                                x = b[i][s_][a] * d[s_][t]
                                s = _fma(x, b[j][t][bb], s) if fused else s + x * b[j][t][bb]
                        k[i * 3 + a][j * 3 + bb] += float(qw[q]) * det * s
    return np.array(k)


def test_fused_accumulation_cannot_reproduce_the_reference_bits():
    """Why the bit-exact FP64 path stays off the matrix cores (and off FMA):
    the matrix cores' multiply-accumulate rounds a * b + c once, the
    reference rounds the product and the sum separately.  The same
    Hex8 element matrix in the same order, with only that accumulation
    fused, differs from the reference in most entries (by ~1e-16 of the
    largest), while the separately rounded restatement reproduces the
    oracle's matrix bit for bit."""
    rng = np.random.default_rng(11)
    phys = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    phys = phys * [1.0, 0.8, 1.3] + rng.uniform(-0.15, 0.15, phys.shape)
    ref = O.elasticity_element_matrix(O.HEX, phys, O.CONT3D, 2.5, 0.3, 2)
    sep = _hex_elasticity_ke(phys, 2.5, 0.3, 2, fused=False)
    assert np.array_equal(sep.view(np.uint64), ref.view(np.uint64))
    fus = _hex_elasticity_ke(phys, 2.5, 0.3, 2, fused=True)
    differ = np.count_nonzero(fus != ref)
    assert differ > 24 * 24 // 4
    assert np.abs(fus - ref).max() <= 1e-14 * np.abs(ref).max()
