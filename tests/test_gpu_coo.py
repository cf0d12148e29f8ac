"""Meshes beyond the tile kernels' valence limits: the sorted-triplet path.

A Delaunay tetrahedralisation of random points has nodes in more than 64
elements (the reference builds such meshes in crates/tpt-fem-mesh-gen/src/
lib.rs:163-409 and imports them from Gmsh); those go through the
sorted-triplet assembly, which restates assemble's push order and
Coo::to_csr's insertion-order sums.  Checked bit for bit against the oracle:
full matrices and reduced systems, Poisson and 3-D elasticity with
per-element coefficients, on a 20,000-point random Delaunay mesh (generated
here with scipy's Qhull from a seeded point set) and on the reference's own
Delaunay box mesh (`tests/golden/vtk/mesh_gen_box_delaunay.vtk`); the path
forced on a box mesh equals the tile kernels; a solve on it matches the
oracle's dense LU.  The triplet export (tfm_assemble_coo) is assemble's Coo
itself: the oracle's triplet list, element by element in push order.
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtk")


def csr_equal(g, r):
    assert (g.nrows, g.ncols) == (r.nrows, r.ncols)
    np.testing.assert_array_equal(g.row_ptrs, r.row_ptrs)
    np.testing.assert_array_equal(g.col_ind, r.col_ind)
    bad = np.flatnonzero(g.values != r.values)
    assert bad.size == 0, f"{bad.size} values differ"


def delaunay(n, seed):
    import scipy.spatial as sp

    pts = np.random.default_rng(seed).uniform(0.0, 1.0, (n, 3))
    return pts, sp.Delaunay(pts).simplices.astype(np.int64)


def ops(T, kind, cf):
    if kind == "poisson":
        return 1, T.Operator.poisson(1.0, 2, kappa_e=cf), O.make_op(O.TET, 3, 0, coef_e=cf)
    return 3, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=cf), O.make_op(O.TET, 3, 1, coef_e=cf, nu=0.3)


@pytest.mark.parametrize("kind", ["poisson", "elasticity"])
def test_high_valence_delaunay_bitexact(ctx, kind):
    import tptfem as T

    coords, conn = delaunay(20000, 5)
    valence = np.bincount(conn.ravel(), minlength=len(coords))
    assert valence.max() > 64
    cf = 10.0 ** np.random.default_rng(1).uniform(-1, 1, len(conn))
    dpn, op, ref = ops(T, kind, cf)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    K = T.assemble(m, dpn, op).to_csr()
    coo = O.assemble(coords, conn, ref, dpn)
    csr_equal(K, coo.to_csr())
    bcs = [(int(q) * dpn + c, 0.0) for q in np.flatnonzero(coords[:, 0] < 0.05) for c in range(dpn)]
    s = T.System.create(m, op, dpn, bcs, np.ones(len(coords) * dpn))
    assert s.assembly_info()["kernel"] == "triplets"
    kg, fg, free = s.download()
    kr, fr, frr = O.reduce_system(coo, np.ones(len(coords) * dpn), bcs)
    np.testing.assert_array_equal(free, frr)
    csr_equal(kg, kr.to_csr())
    # re-assembly with new coefficients on the kept pattern
    cf2 = cf * np.random.default_rng(2).uniform(0.5, 2.0, len(conn))
    dpn, op2, ref2 = ops(T, kind, cf2)
    s.assemble(op2, T.SYS_FP64)
    kg2, _, _ = s.download()
    kr2, _, _ = O.reduce_system(O.assemble(coords, conn, ref2, dpn), np.ones(len(coords) * dpn), bcs)
    csr_equal(kg2, kr2.to_csr())
    s.close()
    m.close()


def test_reference_delaunay_box_mesh_bitexact(ctx):
    import tptfem as T
    from vtk_util import read_legacy_vtk

    _, pts, conn, cell, _ = read_legacy_vtk(open(os.path.join(GOLD, "mesh_gen_box_delaunay.vtk"), "rb").read())
    assert cell == O.TET
    cf = 10.0 ** np.random.default_rng(3).uniform(-1, 1, len(conn))
    for kind in ("poisson", "elasticity"):
        dpn, op, ref = ops(T, kind, cf)
        m = T.Mesh.from_arrays(ctx, T.TET, pts, conn)
        csr_equal(T.assemble(m, dpn, op).to_csr(), O.assemble(pts, conn, ref, dpn).to_csr())
        m.close()


@pytest.mark.parametrize("kind", ["poisson", "elasticity"])
def test_forced_triplet_path_equals_tile_kernels(ctx, monkeypatch, kind):
    import tptfem as T

    n = (5, 4, 6)
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], n)
    m.jitter(n, 0.2, 3)
    coords, conn = m.download()
    cf = 10.0 ** np.random.default_rng(4).uniform(-1, 1, len(conn))
    dpn, op, ref = ops(T, kind, cf)
    bcs = [(int(q) * dpn + c, 0.0) for q in O.nodes_on_plane(coords, 0, 0.0, 1e-9) for c in range(dpn)]
    rhs = np.random.default_rng(5).standard_normal(len(coords) * dpn)
    tile = T.System.create(m, op, dpn, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    monkeypatch.setenv("TFM_COO_ASSEMBLY", "1")
    trip = T.System.create(m, op, dpn, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    assert trip.assembly_info()["kernel"] == "triplets"
    csr_equal(trip.download()[0], tile.download()[0])
    u = trip.solve(T.SOLVE_PCG64, rtol=1e-12)
    u_ref = O.solve_with_dirichlet(O.assemble(coords, conn, ref, dpn), rhs, bcs)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) <= 1e-10
    m.close()


def _coo_equal(g, r, exact=True):
    rows, cols, vals = g
    np.testing.assert_array_equal(rows, r.rows)
    np.testing.assert_array_equal(cols, r.cols)
    if exact:
        assert np.array_equal(vals, r.vals)
    else:
        np.testing.assert_allclose(vals, r.vals, rtol=1e-14, atol=0)


def test_triplet_export_is_the_reference_coo(ctx):
    """tfm_assemble_coo against the oracle's assemble (rows, cols and values in
    push order): the C1 box (30,720 triplets: 49,152 element entries less
    18,432 exact zeros), jittered P1-tet elasticity with per-element moduli,
    Q4 plane stress, Hex8 elasticity, SIMP Q4 (every entry pushed, values to
    the device pow's 1e-14), and an element-free mesh."""
    import tptfem as T

    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], (8, 8, 8))
    c, e = m.download()
    g = T.assemble_coo(m, 1, T.Operator.poisson(1.0, 2))
    assert len(g[0]) == 30720
    _coo_equal(g, O.assemble(c, e, O.make_op(O.TET, 3, 0), 1))
    m.close()

    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (3, 4, 2))
    c = O.jitter_interior(c, (3, 4, 2), 0.2, 3)
    ye = 10.0 ** np.random.default_rng(3).uniform(-1, 1, len(e))
    m = T.Mesh.from_arrays(ctx, T.TET, c, e)
    _coo_equal(T.assemble_coo(m, 3, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)),
               O.assemble(c, e, O.make_op(O.TET, 3, 1, coef_e=ye, nu=0.3), 3))
    m.close()

    qc, qe = O.grid(7, 5, 0.3)
    qc = qc + np.random.default_rng(4).uniform(-0.05, 0.05, qc.shape)
    m = T.Mesh.from_arrays(ctx, T.QUAD, qc, qe)
    _coo_equal(T.assemble_coo(m, 2, T.Operator.elasticity(T.PLANE_STRESS, 3.0, 0.25, 2)),
               O.assemble(qc, qe, O.make_op(O.QUAD, 2, 1, coef=3.0, nu=0.25, model=T.PLANE_STRESS), 2))
    rho = np.random.default_rng(1).uniform(0.01, 1.0, len(qe))
    _coo_equal(T.assemble_coo(m, 2, T.Operator.simp(rho, 0.3, 3.0, 1.0, 1e-9, 1.0)),
               O.assemble(qc, qe, O.make_op(O.QUAD, 2, 2, rho=rho, penal=3.0, e0=1.0, emin=1e-9,
                                            ke0=O.simp_ke_unit(0.3, 1.0)), 2), exact=False)
    m.close()

    from test_gpu_hex_mfma import hex_mesh

    hc, he = hex_mesh((3, 2, 2), 0.2, 6)
    ye = 10.0 ** np.random.default_rng(6).uniform(-1, 1, len(he))
    m = T.Mesh.from_arrays(ctx, T.HEX, hc, he)
    _coo_equal(T.assemble_coo(m, 3, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)),
               O.assemble(hc, he, O.make_op(O.HEX, 3, 1, coef_e=ye, nu=0.3), 3))
    m.close()

    m = T.Mesh.from_arrays(ctx, T.TET, c, np.zeros((0, 4), np.int64))
    rows, cols, vals = T.assemble_coo(m, 1, T.Operator.poisson(1.0, 2))
    assert len(rows) == len(cols) == len(vals) == 0
    m.close()
