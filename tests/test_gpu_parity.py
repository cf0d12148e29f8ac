"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: sparsity patterns, DOF numbering and free maps identical; FP64 values
bit-identical (up to the sign of an exact zero) wherever the reference's
arithmetic is deterministic IEEE (every element operator except SIMP, whose
rho^p uses the device pow: 1e-14 relative); solutions within 1e-10 relative
of the dense-LU direct solve.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

def tp():
    import tptfem

    return tptfem


def csr_equal(g, r, exact=True, rtol=0.0):
    assert (g.nrows, g.ncols) == (r.nrows, r.ncols)
    np.testing.assert_array_equal(g.row_ptrs, r.row_ptrs)
    np.testing.assert_array_equal(g.col_ind, r.col_ind)
    if exact:
        bad = np.flatnonzero(g.values != r.values)
        assert bad.size == 0, f"{bad.size} values differ, first {bad[:5]}: {g.values[bad[:5]]} vs {r.values[bad[:5]]}"
    else:
        scale = np.abs(r.values).max() if r.values.size else 1.0
        assert np.abs(g.values - r.values).max() <= rtol * scale


def hex_box(n):
    nx, ny, nz = n
    idx = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    coords = np.array([[i / nx, j / ny, k / nz] for i in range(nx + 1) for j in range(ny + 1)
                       for k in range(nz + 1)])
    conn = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                conn.append([idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
                             idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1),
                             idx(i, j + 1, k + 1)])
    return coords, np.array(conn, dtype=np.int64)


def tri_grid(nx, ny):
    coords, quads = O.grid(nx, ny, 1.0 / (nx - 1))
    tris = []
    for q in quads:
        tris.append([q[0], q[1], q[2]])
        tris.append([q[0], q[2], q[3]])
    return coords, np.array(tris, dtype=np.int64)


def perturb(coords, scale, seed):
    rng = np.random.default_rng(seed)
    return coords + rng.uniform(-scale, scale, coords.shape)


def face_bcs(coords, dpn, faces):
    bcs = []
    for axis, val in faces:
        for n in O.nodes_on_plane(coords, axis, val, 1e-9):
            for c in range(dpn):
                bcs.append((int(n) * dpn + c, 0.0))
    return bcs


# ---- generators -----------------------------------------------------------
def test_box_mesh_matches_oracle(ctx):
    T = tp()
    mn, mx, n = [-1.0, 0.5, 2.0], [0.3, 1.7, 2.9], (3, 4, 5)
    m = T.Mesh.box(ctx, mn, mx, n)
    c, e = m.download()
    rc, re = O.box_mesh(mn, mx, n)
    assert np.array_equal(c, rc) and np.array_equal(e, re)


def test_grid_and_jitter_match_oracle(ctx):
    T = tp()
    m = T.Mesh.grid(ctx, 9, 5, 0.25)
    c, e = m.download()
    rc, re = O.grid(9, 5, 0.25)
    assert np.array_equal(c, rc) and np.array_equal(e, re)
    b = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], (6, 5, 4))
    b.jitter((6, 5, 4), 0.2, 1234)
    c, _ = b.download()
    rc, _ = O.box_mesh([0, 0, 0], [1, 1, 1], (6, 5, 4))
    assert np.array_equal(c, O.jitter_interior(rc, (6, 5, 4), 0.2, 1234))
    assert not np.array_equal(c, rc)


def test_nodes_on_plane(ctx):
    T = tp()
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], (4, 4, 4))
    c, _ = m.download()
    for axis in range(3):
        np.testing.assert_array_equal(m.nodes_on_plane(axis, 1.0, 1e-9), O.nodes_on_plane(c, axis, 1.0, 1e-9))


# ---- assembly parity --------------------------------------------------------
def _assemble_pair(ctx, cell, coords, conn, op_gpu, op_orc, dpn):
    T = tp()
    m = T.Mesh.from_arrays(ctx, cell, coords, conn)
    g = T.assemble(m, dpn, op_gpu).to_csr()
    r = O.assemble(coords, conn, op_orc, dpn).to_csr()
    return g, r


def test_c1_poisson_assembly_bitexact(ctx):
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (8, 8, 8))
    g, r = _assemble_pair(ctx, O.TET, coords, conn, T.Operator.poisson(1.0, 2), O.make_op(O.TET, 3, 0, coef=1.0), 1)
    assert r.nnz() == 4617
    csr_equal(g, r)


def test_c1_poisson_solution_matches_direct_and_closed_form(ctx):
    T = tp()
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], (8, 8, 8))
    coords, conn = m.download()
    bcs = face_bcs(coords, 1, [(a, v) for a in range(3) for v in (0.0, 1.0)])
    u = T.solve_poisson(m, 1.0, 2, 1.0, bcs)
    ur = O.solve_poisson(coords, O.TET, conn, 1.0, 2, 1.0, bcs)
    assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()
    # 7-point FD solution of -lap u = 1 (DST-I closed form) at the centre node
    assert abs(u.max() - 0.05491766911624052) <= 1e-12


@pytest.mark.parametrize("n,jit", [((4, 4, 4), 0.0), ((5, 4, 3), 0.2)])
def test_elasticity_tet_assembly_bitexact(ctx, n, jit):
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    if jit:
        coords = O.jitter_interior(coords, n, jit, 99)
    rng = np.random.default_rng(5)
    ye = 10.0 ** rng.uniform(-1, 1, len(conn))
    g, r = _assemble_pair(ctx, O.TET, coords, conn,
                          T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye),
                          O.make_op(O.TET, 3, 1, coef_e=ye, nu=0.3), 3)
    if not jit:
        nn = n[0]
        assert r.nnz() == 111 * nn ** 3 + 171 * nn ** 2 + 45 * nn - 3
    csr_equal(g, r)


def test_elasticity_reduced_system_bitexact(ctx):
    T = tp()
    n = (4, 3, 3)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    bcs = face_bcs(coords, 3, [(0, 0.0)])
    rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2)
    s = T.System.create(m, op, 3, bcs, rhs)
    kg, fg, frg = s.download()
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef=1.0, nu=0.3), 3)
    kr, fr, frr = O.reduce_system(coo, rhs, bcs)
    np.testing.assert_array_equal(frg, frr)
    assert np.array_equal(fg, fr)
    csr_equal(kg, kr.to_csr())


def test_elasticity_solution_matches_direct(ctx):
    T = tp()
    n = (4, 3, 3)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    bcs = face_bcs(coords, 3, [(0, 0.0)])
    rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    s = T.System.create(m, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2), 3, bcs, rhs)
    u = s.solve(T.SOLVE_PCG64, rtol=1e-14)
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef=1.0, nu=0.3), 3)
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()


@pytest.mark.parametrize("order", [2, 3])
def test_quad_poisson_and_plane_models_bitexact(ctx, order):
    T = tp()
    coords, conn = O.grid(7, 5, 0.3)
    coords = perturb(coords, 0.05, 3)
    g, r = _assemble_pair(ctx, O.QUAD, coords, conn, T.Operator.poisson(2.5, order),
                          O.make_op(O.QUAD, 2, 0, coef=2.5, order=order), 1)
    csr_equal(g, r)
    for model in (T.PLANE_STRESS, T.PLANE_STRAIN):
        g, r = _assemble_pair(ctx, O.QUAD, coords, conn, T.Operator.elasticity(model, 3.0, 0.25, order),
                              O.make_op(O.QUAD, 2, 1, coef=3.0, nu=0.25, model=model, order=order), 2)
        csr_equal(g, r)


def test_hex_elasticity_and_poisson_bitexact(ctx):
    T = tp()
    coords, conn = hex_box((3, 2, 2))
    coords = perturb(coords, 0.03, 11)
    g, r = _assemble_pair(ctx, O.HEX, coords, conn, T.Operator.elasticity(T.CONTINUUM3D, 2.0, 0.3, 2),
                          O.make_op(O.HEX, 3, 1, coef=2.0, nu=0.3, order=2), 3)
    csr_equal(g, r)
    g, r = _assemble_pair(ctx, O.HEX, coords, conn, T.Operator.poisson(1.0, 3),
                          O.make_op(O.HEX, 3, 0, coef=1.0, order=3), 1)
    csr_equal(g, r)


def test_tri_poisson_and_plane_stress_bitexact(ctx):
    T = tp()
    coords, conn = tri_grid(6, 5)
    coords = perturb(coords, 0.02, 7)
    g, r = _assemble_pair(ctx, O.TRI, coords, conn, T.Operator.poisson(1.0, 2), O.make_op(O.TRI, 2, 0), 1)
    csr_equal(g, r)
    g, r = _assemble_pair(ctx, O.TRI, coords, conn, T.Operator.elasticity(T.PLANE_STRESS, 1.0, 0.3, 2),
                          O.make_op(O.TRI, 2, 1, model=O.PSTRESS, nu=0.3), 2)
    csr_equal(g, r)


def test_simp_q4_pattern_exact_values_close(ctx):
    T = tp()
    coords, conn = O.grid(9, 5, 1.0)
    rho = np.random.default_rng(1).uniform(0.01, 1.0, len(conn))
    ke0 = O.simp_ke_unit(0.3, 1.0)
    g, r = _assemble_pair(ctx, O.QUAD, coords, conn, T.Operator.simp(rho, 0.3, 3.0, 1.0, 1e-9, 1.0),
                          O.make_op(O.QUAD, 2, 2, rho=rho, penal=3.0, e0=1.0, emin=1e-9, ke0=ke0), 2)
    csr_equal(g, r, exact=False, rtol=1e-14)


@pytest.mark.parametrize("fixed_left", [False, True])
def test_simp_grid_rows_equal_tile_kernel(ctx, monkeypatch, fixed_left):
    """SIMP systems on Mesh.grid_batch meshes form their rows directly from
    the four elements around each node; the values (FP64 and FP32) equal
    the generic tile kernel's bit for bit, before and after new densities,
    and the pattern is the same.  Several designs, odd grid sizes."""
    T = tp()
    P, nx, ny = 3, 9, 6
    m = T.Mesh.grid_batch(ctx, P, nx, ny, 1.0)
    rho = np.random.default_rng(8).uniform(0.01, 1.0, m.n_elem)
    op = T.Operator.simp(rho, nu=0.3, penal=3.0, e0=1.0, emin=1e-9, h=1.0)
    bcs = []
    if fixed_left:
        left = (np.arange(P)[:, None] * nx * ny + np.arange(ny)[None, :] * nx).ravel()
        bcs = [(int(2 * q + c), 0.0) for q in left for c in (0, 1)]
    rho2 = np.random.default_rng(9).uniform(0.01, 1.0, m.n_elem)
    op2 = T.Operator.simp(rho2, nu=0.3, penal=3.0, e0=1.0, emin=1e-9, h=1.0)
    got = []
    for generic in (False, True):
        if generic:
            monkeypatch.setenv("TFM_GENERIC_ASSEMBLY", "1")
        s = T.System.create(m, op, 2, bcs, None, flags=T.SYS_FP64 | T.SYS_FP32)
        assert s.assembly_info()["kernel"] == ("generic" if generic else "simp_grid")
        k64 = s.download()[0]
        s.assemble(op2, 0)
        k64b = s.download()[0]
        s32 = T.System.create(m, op, 2, bcs, None, flags=T.SYS_FP32)
        got.append((k64, k64b, s32.download()[0]))
    for a, b in zip(got[0], got[1]):
        csr_equal(a, b)


def test_reassembly_with_new_coefficients(ctx):
    T = tp()
    n = (3, 3, 3)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    coords = O.jitter_interior(coords, n, 0.2, 3)
    bcs = face_bcs(coords, 3, [(0, 0.0)])
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    y1 = np.full(len(conn), 1.0)
    y2 = np.random.default_rng(2).uniform(0.5, 2.0, len(conn))
    s = T.System.create(m, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=y1), 3, bcs)
    s.assemble(T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=y2))
    kg, _, _ = s.download()
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef_e=y2, nu=0.3), 3)
    kr, _, _ = O.reduce_system(coo, np.zeros(len(coords) * 3), bcs)
    csr_equal(kg, kr.to_csr())


# ---- loads ------------------------------------------------------------------
def test_source_and_body_vectors_bitexact(ctx):
    T = tp()
    n = (4, 3, 5)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    coords = O.jitter_interior(coords, n, 0.2, 5)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    assert np.array_equal(T.source_vector(m, 1.7), O.poisson_rhs(coords, O.TET, conn, 1.7))
    qc, qe = O.grid(6, 4, 0.5)
    mq = T.Mesh.from_arrays(ctx, T.QUAD, qc, qe)
    assert np.array_equal(T.source_vector(mq, 1.0, 3), O.poisson_rhs(qc, O.QUAD, qe, 1.0, 3))
    bvec = np.array([0.5, -1.0, 2.0])
    b = T.body_vector(m, bvec)
    assert np.array_equal(b, O.body_rhs(coords, O.TET, conn, np.tile(bvec, (len(conn), 4, 1))))
    # Q4 (dim 2) and Hex8 (8 points at order 2, 27 at order 3)
    qcj = perturb(qc, 0.05, 3)
    mq = T.Mesh.from_arrays(ctx, T.QUAD, qcj, qe)
    assert np.array_equal(T.body_vector(mq, bvec[:2], 2), O.body_rhs(qcj, O.QUAD, qe, np.tile(bvec[:2], (len(qe), 4, 1)), 2))
    hc, he = hex_box((3, 2, 2))
    hc = perturb(hc, 0.04, 8)
    mh = T.Mesh.from_arrays(ctx, T.HEX, hc, he)
    for order, nq in ((2, 8), (3, 27)):
        assert np.array_equal(T.body_vector(mh, bvec, order),
                              O.body_rhs(hc, O.HEX, he, np.tile(bvec, (len(he), nq, 1)), order))


def test_plane_traction_and_neumann_match_oracle(ctx):
    T = tp()
    n = (4, 3, 3)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    t = [0.0, 0.3, -1e-3]
    g = T.plane_traction(m, 3, 0, 1.0, 1e-9, t)
    r = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, t)
    assert np.array_equal(g, r)
    assert abs(r[2::3].sum() + 1e-3) < 1e-15
    # constant flux over the whole boundary tiled over components
    g = T.apply_neumann(m, 1, 1.0, np.zeros(len(coords)))
    r = O.apply_neumann(coords, O.TET, conn, 1, 1.0, np.zeros(len(coords)))
    assert np.array_equal(g, r)
    assert abs(r.sum() - 6.0) < 1e-12
    # 2-D: Tri3 boundary (reference neumann_conserves_total_load)
    tc = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    mt = T.Mesh.from_arrays(ctx, T.TRI, tc, np.array([[0, 1, 2]]))
    gt = T.apply_neumann(mt, 1, 1.0, np.zeros(3))
    assert abs(gt.sum() - (2.0 + np.sqrt(2.0))) < 1e-9


@pytest.mark.parametrize("axis,coord,tol", [(0, 1.0, 1e-9), (1, 0.0, 0.3), (2, 0.62, 0.25), (0, 1.0, 0.45)])
def test_plane_traction_wide_bands_on_jittered_mesh(ctx, axis, coord, tol):
    """Traction bands that cut through elements of a jittered box: faces with
    some nodes inside the band and some outside, quadrature points on either
    side, and planes with no boundary face at all -- bit-identical to the
    oracle."""
    T = tp()
    n = (5, 4, 4)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    coords = O.jitter_interior(coords, n, 0.2, 11)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    t = [0.25, -1.5, 1e-3]
    g = T.plane_traction(m, 3, axis, coord, tol, t)
    r = O.plane_traction(coords, O.TET, conn, 3, axis, coord, tol, t)
    assert np.array_equal(g, r)
    qc, qe = O.grid(6, 5, 0.2)
    mq = T.Mesh.from_arrays(ctx, T.QUAD, qc, qe)
    g = T.plane_traction(mq, 2, axis % 2, coord * 0.8, tol, t[:2])
    r = O.plane_traction(qc, O.QUAD, qe, 2, axis % 2, coord * 0.8, tol, t[:2])
    assert np.array_equal(g, r)


def _mass_meshes():
    c, e = O.box_mesh([0, 0, 0], [1.0, 0.7, 1.3], (3, 2, 3))
    yield "tet", O.TET, O.jitter_interior(c, (3, 2, 3), 0.2, 3), e
    hc, he = hex_box((2, 3, 2))
    yield "hex", O.HEX, perturb(hc, 0.05, 4), he
    tc, te = tri_grid(5, 4)
    yield "tri", O.TRI, perturb(tc, 0.03, 5), te
    qc, qe = O.grid(5, 4, 0.3)
    yield "quad", O.QUAD, perturb(qc, 0.03, 6), qe
    lc = np.linspace(0.0, 2.0, 7)[:, None] ** 1.3
    yield "line", O.LINE, lc, np.array([[i, i + 1] for i in range(6)], dtype=np.int64)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_mass_matrices_bitexact(ctx, order):
    """Consistent mass (elasticity_mass_matrix) for every cell type and the
    lumped diagonal built from it, bit-identical to the oracle; per-element
    densities through the operator's coefficient array."""
    T = tp()
    for name, cell, c, e in _mass_meshes():
        m = T.Mesh.from_arrays(ctx, cell, c, e)
        g = T.mass_matrix(m, 2.5, order).to_csr()
        r = O.mass_matrix(c, cell, e, 2.5, order).to_csr()
        csr_equal(g, r)
        gl = T.lumped_mass(m, 2.5, order).to_csr()
        rl = O.lumped_mass(c, cell, e, 2.5, order).to_csr()
        csr_equal(gl, rl)
        rho_e = np.random.default_rng(order).uniform(0.5, 2.0, len(e))
        dim = O.DIMS[cell]
        g = T.assemble(m, dim, T.Operator.mass(1.0, order, density_e=rho_e)).to_csr()
        r = O.assemble(c, e, O.make_op(cell, c.shape[1], 3, order=order, coef_e=rho_e), dim).to_csr()
        csr_equal(g, r)


def test_mass_contract(ctx):
    """The mass operator has one DOF per spatial dimension."""
    T = tp()
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (1, 1, 1))
    m = T.Mesh.from_arrays(ctx, T.TET, c, e)
    with pytest.raises(T.TfmError):
        T.assemble(m, 1, T.Operator.mass(1.0))


# ---- sparse API -------------------------------------------------------------
def test_coo_to_csr_known_answer_and_matvec(ctx):
    T = tp()
    # crates/tpt-fem-sparse/src/lib.rs:378-391 and :500-524
    coo = O.Coo([1, 0, 0, 0, 1], [1, 1, 0, 0, 0], [1.0, 1.0, 1.0, 1.0, 1.0])
    csr = O.to_csr(coo)
    np.testing.assert_array_equal(csr.row_ptrs, [0, 2, 4])
    np.testing.assert_array_equal(csr.col_ind, [0, 1, 0, 1])
    np.testing.assert_array_equal(csr.values, [2.0, 1.0, 1.0, 1.0])
    M = T.Matrix.from_csr(ctx, T.Csr(csr.nrows, csr.ncols, csr.row_ptrs, csr.col_ind, csr.values))
    back = M.to_csr()
    csr_equal(back, csr)
    a = O.to_csr(O.Coo([0, 0, 1], [0, 1, 1], [2.0, 1.0, 3.0]))
    Ma = T.Matrix.from_csr(ctx, T.Csr(a.nrows, a.ncols, a.row_ptrs, a.col_ind, a.values))
    y = Ma.matvec([1.0, -2.0])
    np.testing.assert_array_equal(y, [0.0, -6.0])
    e = O.to_csr(O.Coo())
    Me = T.Matrix.from_csr(ctx, T.Csr(e.nrows, e.ncols, e.row_ptrs, e.col_ind, e.values))
    np.testing.assert_array_equal(Me.matvec([1.0, 2.0, 3.0]), [0.0, 0.0, 0.0])


def test_matvec_bitexact_on_assembled_matrix(ctx):
    T = tp()
    n = (5, 4, 4)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    coords = O.jitter_interior(coords, n, 0.2, 8)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    K = T.assemble(m, 3, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2))
    x = np.random.default_rng(4).standard_normal(K.ncols)
    r = O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef=1.0, nu=0.3), 3).to_csr()
    assert np.array_equal(K.matvec(x), r.matvec(x))


def test_dirichlet_line_known_answer(ctx):
    T = tp()
    # crates/tpt-fem-assembly/src/lib.rs:795-811: 2 Line2 elements, k=1,
    # element matrix 2*[[1,-1],[-1,1]], u(0)=0, u(1)=1 -> u(0.5) = 0.5
    m = T.Mesh.from_arrays(ctx, T.LINE, np.array([[0.0], [0.5], [1.0]]), np.array([[0, 1], [1, 2]]))
    K = T.assemble(m, 1, T.Operator.poisson(1.0, 2))
    np.testing.assert_allclose(K.to_csr().values, [2.0, -2.0, -2.0, 4.0, -2.0, -2.0, 2.0], rtol=0, atol=1e-15)
    u = T.solve_with_dirichlet(K, [0.0, 0.0, 0.0], [(0, 0.0), (2, 1.0)])
    np.testing.assert_allclose(u, [0.0, 0.5, 1.0], atol=1e-12)


def test_nonhomogeneous_dirichlet_lifting_matches_oracle(ctx):
    T = tp()
    n = (4, 4, 3)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    bcs = [(int(i), 1.0 + 0.1 * k) for k, i in enumerate(O.nodes_on_plane(coords, 0, 0.0, 1e-9))]
    bcs += [(int(i), -0.5) for i in O.nodes_on_plane(coords, 0, 1.0, 1e-9)]
    bcs.append((bcs[0][0], 9.0))  # duplicate: lifting uses the first, the scatter the last
    rhs = O.poisson_rhs(coords, O.TET, conn, 1.0)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, bcs, rhs)
    kg, fg, frg = s.download()
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 0), 1)
    kr, fr, frr = O.reduce_system(coo, rhs, bcs)
    np.testing.assert_array_equal(frg, frr)
    assert np.array_equal(fg, fr)
    csr_equal(kg, kr.to_csr())
    u = s.solve(T.SOLVE_PCG64, rtol=1e-14)
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()
    assert u[bcs[0][0]] == 9.0


# ---- precision form ---------------------------------------------------------
def test_mixed_precision_refinement_reaches_1e8(ctx):
    T = tp()
    n = (10, 10, 10)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    m.jitter(n, 0.2, 17)
    coords, conn = m.download()
    ye = 10.0 ** np.random.default_rng(3).uniform(-1, 1, m.n_elem)
    bcs = face_bcs(coords, 3, [(0, 0.0)])
    rhs = T.plane_traction(m, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)
    s = T.System.create(m, op, 3, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    u64 = s.solve(T.SOLVE_PCG64, rtol=1e-13)
    umx = s.solve(T.SOLVE_MIXED, rtol=1e-11, inner_rtol=0.1)
    err = np.linalg.norm(umx - u64) / np.linalg.norm(u64)
    assert err <= 1e-8, err
    assert s.stats.outer >= 2


# ---- error paths --------------------------------------------------------------
def test_model_dimension_mismatch_raises(ctx):
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (1, 1, 1))
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    with pytest.raises(T.ElasticityError):
        T.assemble(m, 3, T.Operator.elasticity(T.PLANE_STRESS, 1.0, 0.3, 2))


def test_singular_system_reports_numeric_error(ctx):
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (2, 2, 2))
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    s = T.System.create(m, T.Operator.poisson(1.0, 2), 1, [], np.ones(len(coords)))
    with pytest.raises((T.SparseError, T.TfmError)):
        s.solve(T.SOLVE_PCG64, rtol=1e-14, max_iter=500)


def test_tet_fast_path_equals_generic_kernel(ctx, monkeypatch):
    T = tp()
    n = (6, 5, 4)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    m.jitter(n, 0.2, 21)
    coords, conn = m.download()
    ye = 10.0 ** np.random.default_rng(6).uniform(-1, 1, m.n_elem)
    bcs = face_bcs(coords, 3, [(0, 0.0)])
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)
    fast = T.System.create(m, op, 3, bcs, flags=T.SYS_FP64 | T.SYS_FP32)
    kf, _, _ = fast.download()
    monkeypatch.setenv("TFM_GENERIC_ASSEMBLY", "1")
    gen = T.System.create(m, op, 3, bcs, flags=T.SYS_FP64 | T.SYS_FP32)
    kg, _, _ = gen.download()
    csr_equal(kf, kg)
    coo = O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef_e=ye, nu=0.3), 3)
    kr, _, _ = O.reduce_system(coo, np.zeros(len(coords) * 3), bcs)
    csr_equal(kf, kr.to_csr())


@pytest.mark.parametrize("dpn", [1, 3])
def test_tet_tile_sizes_bitexact_and_tuner_record(ctx, monkeypatch, dpn):
    """A node's pairs never span two P1-tet tiles, so both tile sizes sum
    every entry in the same element order: FP64 and FP32 matrices are
    bit-identical across them.  The autotuner's record names the forced
    size, and an unforced system either timed both candidates (and kept
    the faster) or reused the choice of an earlier system of its class."""
    T = tp()
    n = (7, 6, 5)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    m.jitter(n, 0.2, 5)
    coords, _ = m.download()
    cf = 10.0 ** np.random.default_rng(4).uniform(-1, 1, m.n_elem)
    if dpn == 1:
        op = T.Operator.poisson(1.0, 2, kappa_e=cf)
    else:
        op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=cf)
    bcs = face_bcs(coords, dpn, [(0, 0.0)])
    got = {}
    for pairs in (192, 256):
        monkeypatch.setenv("TFM_TET_TILE_PAIRS", str(pairs))
        for flags in (T.SYS_FP64, T.SYS_FP32):
            s = T.System.create(m, op, dpn, bcs, flags=flags)
            info = s.assembly_info()
            assert (info["kernel"], info["tile_pairs"], info["source"]) == ("p1_tet", pairs, "env")
            got[(pairs, flags)] = s.download()[0]
    for flags in (T.SYS_FP64, T.SYS_FP32):
        csr_equal(got[(192, flags)], got[(256, flags)])
    monkeypatch.delenv("TFM_TET_TILE_PAIRS")
    first = T.System.create(m, op, dpn, bcs).assembly_info()
    assert first["kernel"] == "p1_tet" and first["tile_pairs"] in (192, 256)
    assert first["source"] in ("timed", "cached")
    if first["source"] == "timed":
        t256, t192 = first["ms"]["256"], first["ms"]["192"]
        assert t256 > 0 and t192 > 0
        assert first["tile_pairs"] == (256 if t256 <= t192 else 192)
    again = T.System.create(m, op, dpn, bcs).assembly_info()
    assert (again["source"], again["tile_pairs"]) == ("cached", first["tile_pairs"])


@pytest.mark.parametrize("kind", ["poisson", "elasticity"])
def test_tet_fp32_values_close_to_fp64(ctx, kind):
    """The precision form's FP32 matrix: same pattern, values within FP32
    rounding of the exact FP64 ones (1e-5 of the largest entry)."""
    T = tp()
    n = (5, 4, 4)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    m.jitter(n, 0.2, 8)
    coords, conn = m.download()
    cf = 10.0 ** np.random.default_rng(2).uniform(-1, 1, m.n_elem)
    dpn = 1 if kind == "poisson" else 3
    if kind == "poisson":
        op = T.Operator.poisson(1.0, 2, kappa_e=cf)
        ref = O.make_op(O.TET, 3, 0, coef_e=cf)
    else:
        op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=cf)
        ref = O.make_op(O.TET, 3, 1, coef_e=cf, nu=0.3)
    bcs = face_bcs(coords, dpn, [(0, 0.0)])
    s = T.System.create(m, op, dpn, bcs, flags=T.SYS_FP32)
    k32, _, _ = s.download()
    kr, _, _ = O.reduce_system(O.assemble(coords, conn, ref, dpn), np.zeros(len(coords) * dpn), bcs)
    csr_equal(k32, kr.to_csr(), exact=False, rtol=1e-5)


@pytest.mark.parametrize("kind", ["poisson", "elasticity"])
def test_tet_both_arrays_fp32_is_rounded_fp64(ctx, kind):
    """A system holding both value arrays fills them in one pass: the FP32
    values are the exact FP64 values rounded to nearest, at creation and
    after a re-assembly with new coefficients; a re-assembly of the FP32
    array alone takes the FP32 pass (within 1e-5 of each row's maximum)."""
    T = tp()
    n = (6, 4, 5)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    m.jitter(n, 0.2, 12)
    coords, conn = m.download()
    rng = np.random.default_rng(9)
    dpn = 1 if kind == "poisson" else 3

    def make(cf):
        if kind == "poisson":
            return T.Operator.poisson(1.0, 2, kappa_e=cf), O.make_op(O.TET, 3, 0, coef_e=cf)
        return (T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=cf),
                O.make_op(O.TET, 3, 1, coef_e=cf, nu=0.3))

    def oracle_csr(ref):
        kr, _, _ = O.reduce_system(O.assemble(coords, conn, ref, dpn), np.zeros(len(coords) * dpn), bcs)
        return kr.to_csr()

    bcs = face_bcs(coords, dpn, [(0, 0.0)])
    op, ref = make(10.0 ** rng.uniform(-1, 1, m.n_elem))
    s = T.System.create(m, op, dpn, bcs, flags=T.SYS_FP64 | T.SYS_FP32)
    assert s.assembly_info()["kernel"] == "p1_tet"
    for step in range(2):
        k64, _, _ = s.download()
        csr_equal(k64, oracle_csr(ref))
        np.testing.assert_array_equal(s.values32(), k64.values.astype(np.float32).astype(np.float64))
        op, ref = make(10.0 ** rng.uniform(-1, 1, m.n_elem))
        s.assemble(op, T.SYS_FP64 | T.SYS_FP32)
    op, ref = make(10.0 ** rng.uniform(-1, 1, m.n_elem))
    s.assemble(op, T.SYS_FP32)
    kc = oracle_csr(ref)
    v32 = s.values32()
    for i in range(len(kc.row_ptrs) - 1):
        r0, r1 = kc.row_ptrs[i], kc.row_ptrs[i + 1]
        if r1 > r0:
            assert np.abs(v32[r0:r1] - kc.values[r0:r1]).max() <= 1e-5 * np.abs(kc.values[r0:r1]).max(), i
    s.close()
    m.close()
