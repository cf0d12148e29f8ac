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
oracle's dense LU.
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
