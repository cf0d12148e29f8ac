"""The solver legs of the workloads at full size: the plain FP64 Jacobi PCG
by its true FP64 residual, the precision form's FP32 values against the
oracle's FP64 rows, and the precision form's solutions (FP32 PCG with FP64
residual replacement) against an FP64 solve of the same system.

Workloads (as the digests and full-size tests build them):
* C2: box_mesh 64^3, E = 1, nu = 0.3, x = 0 clamped, traction (0, 0, -1e-3)
  on x = 1; 811,200 free DOFs;
* C3: box_mesh 128^3 jittered +-0.2h (seed 2024), E_e = 10^U(-1,1), same
  loads; 6,390,144 free DOFs;
* C4: 256 Poisson boxes U{8..48}^3, kappa_e = exp(N(0,1)), f = 1, u = 0 on
  the faces;
* C5: 64 SIMP designs on Grid(257, 129), rho_e ~ U[0.01, 1], p = 3, left edge
  clamped, 8 point loads per design.

The FP64 reference solutions are multigrid CG (C3, C5) or Jacobi PCG (C4)
solves whose recurrence residual reaches 1e-12; the precision form's error is
||u - u_ref|| / ||u_ref|| per problem and right-hand side.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ERR = 1e-8  # relative error of the precision form vs the FP64 solution


def csr_matrix(K):
    import scipy.sparse as sp

    return sp.csr_matrix((K.values, K.col_ind, K.row_ptrs), shape=(K.nrows, K.ncols))


def box_face_nodes(n):
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    on = (i == 0) | (i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz)
    return ((i * (ny + 1) + j) * (nz + 1) + k)[on]


def c3_system(T, ctx, flags):
    n = (128, 128, 128)
    mesh = T.Mesh.box(ctx, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], n)
    mesh.jitter(n, 0.2, 2024)
    ye = 10.0 ** np.random.default_rng(2024).uniform(-1.0, 1.0, mesh.n_elem)
    bcs = [(int(i) * 3 + k, 0.0) for i in mesh.nodes_on_plane(0, 0.0, 1e-9) for k in range(3)]
    rhs = T.plane_traction(mesh, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    s = T.System.create(mesh, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye), 3, bcs, rhs,
                        flags=flags)
    return mesh, s, ye, bcs


def test_c2_jacobi_pcg_to_1e12_true_residual(ctx):
    import tptfem as T

    mesh = T.Mesh.box(ctx, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], (64, 64, 64))
    bcs = [(int(i) * 3 + k, 0.0) for i in mesh.nodes_on_plane(0, 0.0, 1e-9) for k in range(3)]
    rhs = T.plane_traction(mesh, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    s = T.System.create(mesh, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2), 3, bcs, rhs)
    # The recurrence residual reaches 1e-12; the true FP64 residual of the
    # iterate then sits at the floor of evaluating f - K u in FP64 (about
    # 1e-12 here: the solver restarts from it while that still gains and
    # accepts up to 10 rtol), checked again on the host from the download.
    u = s.solve(T.SOLVE_PCG64, rtol=1e-12, max_iter=200000)
    assert s.stats.rel_residual <= 1e-11
    K, f, free = s.download()
    res = np.linalg.norm(f - csr_matrix(K) @ u[free]) / np.linalg.norm(f)
    print(f"C2 Jacobi PCG: {s.stats.iterations} iterations, {s.stats.outer} restarts, device certificate "
          f"{s.stats.rel_residual:.2e}, host true rel. residual {res:.2e}")
    assert res <= 1e-11
    s.close()
    mesh.close()


def test_c3_fp32_values_on_sampled_rows(ctx):
    """FP32 reduced rows of the sampled nodes against the oracle's FP64 rows
    (same pattern; each value within 1e-6 of its row's largest entry)."""
    import tptfem as T
    from test_gpu_fullsize import sample_nodes, sampled_rows

    mesh, s, ye, bcs = c3_system(T, ctx, T.SYS_FP32)
    K, _, free = s.download()
    s.close()
    mesh.close()
    rc, re = O.box_mesh([0, 0, 0], [1, 1, 1], (128, 128, 128))
    rc = O.jitter_interior(rc, (128, 128, 128), 0.2, 2024)
    nodes = sample_nodes()
    sel = np.zeros(len(rc), bool)
    sel[nodes] = True
    patch = np.flatnonzero(sel[re].any(1))
    coo = O.assemble(rc, re[patch], O.make_op(O.TET, 3, 1, coef_e=ye[patch], nu=0.3), 3)
    ndof = 3 * len(rc)
    coo = O.Coo(np.append(coo.rows, ndof - 1), np.append(coo.cols, 0), np.append(coo.vals, 0.0))
    kr, _, ofree = O.reduce_system(coo, np.zeros(ndof), bcs)
    np.testing.assert_array_equal(ofree, free)
    kc = kr.to_csr()
    worst = 0.0
    rows = sampled_rows(free, nodes)
    for i in rows:
        g0, g1 = K.row_ptrs[i], K.row_ptrs[i + 1]
        r0, r1 = kc.row_ptrs[i], kc.row_ptrs[i + 1]
        np.testing.assert_array_equal(K.col_ind[g0:g1], kc.col_ind[r0:r1])
        ref = kc.values[r0:r1]
        worst = max(worst, np.abs(K.values[g0:g1] - ref).max() / np.abs(ref).max())
    print(f"C3 FP32 rows: {rows.size} rows, worst |v32 - v64| / row max {worst:.2e}")
    assert worst <= 1e-6


def test_c3_precision_form_error_vs_fp64(ctx):
    import tptfem as T

    mesh, s, _, _ = c3_system(T, ctx, T.SYS_FP64 | T.SYS_FP32)
    # the FP64 reference: its recurrence stops at 1e-12; the true FP64
    # residual floor of this system (E contrast 100, jittered) is ~1e-11
    u_ref = s.solve(T.SOLVE_PCG64_MG, rtol=1e-12, max_iter=2000)
    print(f"C3 FP64 reference: {s.stats.iterations} iterations, true rel. residual {s.stats.rel_residual:.2e}")
    assert s.stats.rel_residual <= 1e-10
    u = s.solve(T.SOLVE_MIXED, rtol=1e-9, max_iter=200000)
    err = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    print(f"C3 precision form: {s.stats.iterations} iterations, {s.stats.outer} replacements, "
          f"rel. residual {s.stats.rel_residual:.2e}, error vs FP64 {err:.2e}")
    assert s.stats.rel_residual <= 1e-9
    assert err <= ERR
    s.close()
    mesh.close()


def segment_errors(u, u_ref, off):
    out = []
    for p in range(len(off) - 1):
        a, b = off[p], off[p + 1]
        nr = np.linalg.norm(u_ref[a:b])
        out.append(np.linalg.norm(u[a:b] - u_ref[a:b]) / nr if nr > 0 else np.linalg.norm(u[a:b]))
    return np.array(out)


def test_c4_precision_form_error_per_box(ctx):
    import tptfem as T

    rng = np.random.default_rng(2024)
    dims = [tuple(int(v) for v in d) for d in rng.integers(8, 49, size=(256, 3))]
    mesh, noff, _ = T.Mesh.box_batch(ctx, dims)
    kappa = np.exp(rng.standard_normal(mesh.n_elem))
    ids = np.concatenate([noff[p] + box_face_nodes(d) for p, d in enumerate(dims)])
    bcs = np.stack([ids.astype(np.float64), np.zeros(len(ids))], axis=1)
    s = T.System.create(mesh, T.Operator.poisson(1.0, 2, kappa_e=kappa), 1, bcs, T.source_vector(mesh, 1.0),
                        flags=T.SYS_FP64 | T.SYS_FP32)
    u_ref = s.solve_batch(noff, nrhs=1, mode=T.SOLVE_PCG64, rtol=1e-12, max_iter=200000)[0]
    u = s.solve_batch(noff, nrhs=1, mode=T.SOLVE_MIXED, rtol=1e-10, max_iter=200000)[0]
    assert all(st[2] <= 1e-10 for st in s.batch_stats)
    err = segment_errors(u, u_ref, noff)
    print(f"C4 precision form: max iterations {max(st[0] for st in s.batch_stats)}, worst box error {err.max():.2e}")
    assert err.max() <= ERR
    s.close()
    mesh.close()


def test_c5_jacobi_and_precision_form_error_per_design(ctx):
    import tptfem as T

    P, nx, ny = 64, 257, 129
    nd1 = 2 * nx * ny
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
    op = T.Operator.simp(rho, nu=0.3, penal=3.0, e0=1.0, emin=1e-9, h=1.0)
    s = T.System.create(mesh, op, 2, bcs, None, flags=T.SYS_FP64 | T.SYS_FP32)
    off = np.arange(P + 1) * nd1
    # the FP64 reference (stiffness contrast 1e6: the true residual floor of
    # these systems lies above the 1e-12 the recurrence stops on)
    u_ref = s.solve_batch(off, rhs=F, nrhs=8, mode=T.SOLVE_PCG64_MG, rtol=1e-12, max_iter=20000)
    print(f"C5 FP64 reference: worst true rel. residual {max(st[2] for st in s.batch_stats):.2e}")
    assert all(st[2] <= 1e-10 for st in s.batch_stats)
    for mode, name, rtol in ((T.SOLVE_PCG64, "Jacobi PCG", 1e-10), (T.SOLVE_MIXED, "precision form", 1e-10)):
        u = s.solve_batch(off, rhs=F, nrhs=8, mode=mode, rtol=rtol, max_iter=200000)
        res = max(st[2] for st in s.batch_stats)
        err = max(segment_errors(u[k], u_ref[k], off).max() for k in range(8))
        print(f"C5 {name}: max iterations {max(st[0] for st in s.batch_stats)}, worst true rel. residual "
              f"{res:.2e}, worst (design, load) error {err:.2e}")
        assert res <= 10 * rtol
        assert err <= ERR
    s.close()
    mesh.close()
