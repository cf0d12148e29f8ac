"""GPU parity at the full sizes of the two batching workloads.

C4: 256 Poisson boxes, cells per axis U{8..48} (drawn with seed 2024),
kappa_e = exp(N(0,1)), f = 1, u = 0 on the six faces (about 32M tets).
C5: 64 SIMP designs on Grid(257, 129, h = 1), rho_e ~ U[0.01, 1], p = 3,
E0 = 1, Emin = 1e-9, nu = 0.3, left edge clamped, 8 unit point loads (-y) per
design at seeded right-edge nodes.

A block-diagonal system's block is its problem's own reduced system, numbered
from the block's offset, so the oracle's assemble + reduce_system of single
problems checks whole blocks at full size: C4's smallest and largest boxes
bit for bit, C5's first and last designs to 1e-14 (rho^p uses the device
pow).  Every segment's true FP64 residual after the batched solve closes the
check.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def box_face_nodes(n):
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    on = (i == 0) | (i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz)
    return ((i * (ny + 1) + j) * (nz + 1) + k)[on]


def block(K, free, lo, hi):
    """Rows and columns of the reduced CSR K whose DOFs lie in [lo, hi), as
    (row_ptrs, col_ind, values) numbered from the block's first free DOF."""
    a, b = np.searchsorted(free, lo), np.searchsorted(free, hi)
    rp = K.row_ptrs[a:b + 1] - K.row_ptrs[a]
    sl = slice(K.row_ptrs[a], K.row_ptrs[b])
    return rp, K.col_ind[sl] - a, K.values[sl], a, b


def segment_residuals(K, f, free, u, off):
    import scipy.sparse as sp

    A = sp.csr_matrix((K.values, K.col_ind, K.row_ptrs), shape=(K.nrows, K.ncols))
    r = f - A @ u[free]
    out = []
    for p in range(len(off) - 1):
        a, b = np.searchsorted(free, off[p]), np.searchsorted(free, off[p + 1])
        out.append(np.linalg.norm(r[a:b]) / np.linalg.norm(f[a:b]))
    return np.array(out)


def test_c4_full_size_blocks_bitexact_and_residuals(ctx):
    import tptfem as T

    rng = np.random.default_rng(2024)
    dims = [tuple(int(v) for v in d) for d in rng.integers(8, 49, size=(256, 3))]
    mesh, noff, eoff = T.Mesh.box_batch(ctx, dims)
    kappa = np.exp(rng.standard_normal(mesh.n_elem))
    ids = np.concatenate([noff[p] + box_face_nodes(d) for p, d in enumerate(dims)])
    bcs = np.stack([ids.astype(np.float64), np.zeros(len(ids))], axis=1)
    rhs = T.source_vector(mesh, 1.0)
    s = T.System.create(mesh, T.Operator.poisson(1.0, 2, kappa_e=kappa), 1, bcs, rhs)
    K, f, free = s.download()
    assert mesh.n_elem == int(sum(6 * a * b * c for a, b, c in dims))
    sizes = [a * b * c for a, b, c in dims]
    for p in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        c, e = O.box_mesh([0, 0, 0], [1, 1, 1], dims[p])
        coo = O.assemble(c, e, O.make_op(O.TET, 3, 0, coef_e=kappa[eoff[p]:eoff[p + 1]]), 1)
        bp = [(int(i), 0.0) for i in box_face_nodes(dims[p])]
        kr, fr, frr = O.reduce_system(coo, O.poisson_rhs(c, O.TET, e, 1.0), bp)
        kc = kr.to_csr()
        rp, ci, v, a, b = block(K, free, noff[p], noff[p + 1])
        np.testing.assert_array_equal(free[a:b] - noff[p], frr)
        np.testing.assert_array_equal(rp, kc.row_ptrs)
        np.testing.assert_array_equal(ci, kc.col_ind)
        assert np.array_equal(v.view(np.uint64), kc.values.view(np.uint64)), p
        assert np.array_equal(f[a:b].view(np.uint64), fr.view(np.uint64)), p
    u = s.solve_batch(noff, nrhs=1, mode=T.SOLVE_PCG64, rtol=1e-12, max_iter=200000)[0]
    assert all(st[2] <= 1e-11 for st in s.batch_stats)  # the stop tests the recurrence residual
    assert segment_residuals(K, f, free, u, noff).max() <= 1e-10
    s.close()
    mesh.close()


def test_c5_full_size_design_blocks_and_multigrid_residuals(ctx):
    import tptfem as T

    P, nx, ny = 64, 257, 129
    nd1, ne = 2 * nx * ny, (nx - 1) * (ny - 1)
    rng = np.random.default_rng(2024)
    mesh = T.Mesh.grid_batch(ctx, P, nx, ny, 1.0)
    rho = rng.uniform(0.01, 1.0, mesh.n_elem)
    left = (np.arange(P)[:, None] * nx * ny + np.arange(ny)[None, :] * nx).ravel()
    ids = np.concatenate([2 * left, 2 * left + 1])
    bcs = np.stack([ids.astype(np.float64), np.zeros(len(ids))], axis=1)
    F = np.zeros((8, P * nd1))
    rows = rng.integers(0, ny, size=(P, 8))
    for p in range(P):
        for k in range(8):
            F[k, 2 * (p * nx * ny + rows[p, k] * nx + nx - 1) + 1] = -1.0
    s = T.System.create(mesh, T.Operator.simp(rho, nu=0.3, penal=3.0, e0=1.0, emin=1e-9, h=1.0), 2, bcs, None)
    K, _, free = s.download()
    off = np.arange(P + 1) * nd1
    c, q = O.grid(nx, ny, 1.0)
    ke0 = O.simp_ke_unit(0.3, 1.0)
    b1 = [(2 * j * nx + cc, 0.0) for j in range(ny) for cc in (0, 1)]
    for p in (0, P - 1):
        op = O.make_op(O.QUAD, 2, 2, rho=rho[p * ne:(p + 1) * ne], penal=3.0, e0=1.0, emin=1e-9, ke0=ke0)
        kr, _, frr = O.reduce_system(O.assemble(c, q, op, 2), np.zeros(nd1), b1)
        kc = kr.to_csr()
        rp, ci, v, a, b = block(K, free, off[p], off[p + 1])
        np.testing.assert_array_equal(free[a:b] - off[p], frr)
        np.testing.assert_array_equal(rp, kc.row_ptrs)
        np.testing.assert_array_equal(ci, kc.col_ind)
        assert np.abs(v - kc.values).max() <= 1e-14 * np.abs(kc.values).max()
    u = s.solve_batch(off, rhs=F, nrhs=8, mode=T.SOLVE_PCG64_MG, rtol=1e-10, max_iter=20000)
    assert all(st[2] <= 1e-9 for st in s.batch_stats)
    worst = max(segment_residuals(K, F[k][free], free, u[k], off).max() for k in range(8))
    assert worst <= 1e-9
    s.close()
    mesh.close()
