"""FP32 Hex8 elasticity element matrices on the matrix cores (precision form).

The FP32 numeric pass of a 3-D isotropic elasticity system on Hex8 cells
forms each element matrix once as E (lambda P + mu P' + mu delta tr P) with
P = H H^T on v_mfma_f32_32x32x2f32 (H: scaled gradients, 24 x quadrature
points), then sums the entries into the CSR in ascending element order.
Checked against the oracle's FP64 reduced matrix (identical pattern, values
within 2e-6 of each row's largest entry) for Gauss-Legendre orders 1, 2 and
3 (1, 8 and 27 points: the padded inner dimension 2, 8 and 28), against the
per-pair FP32 kernel it replaces (TFM_NO_MFMA), and through a precision-form
solve against the oracle's dense LU.

Hex8 elasticity is assembled from per-element tables by the sorted-triplet
sum: FP64 element matrices formed once per element by a wavefront in the
reference's operation order, FP32 ones on the matrix cores.  Its FP64
values are the oracle's bit for bit and those of the tile kernels
(TFM_HEX_TILE), before and after re-assembly.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def hex_mesh(n, jit, seed):
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    coords = np.stack([i / nx, j / ny * 0.8, k / nz * 1.3], -1).reshape(-1, 3).astype(np.float64)
    inner = ((i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)).ravel()
    rng = np.random.default_rng(seed)
    coords[inner] += rng.uniform(-jit / max(n), jit / max(n), (int(inner.sum()), 3))
    idx = lambda a, b, c: (a * (ny + 1) + b) * (nz + 1) + c
    I, J, K = [v.ravel() for v in np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")]
    conn = np.stack([idx(I, J, K), idx(I + 1, J, K), idx(I + 1, J + 1, K), idx(I, J + 1, K),
                     idx(I, J, K + 1), idx(I + 1, J, K + 1), idx(I + 1, J + 1, K + 1), idx(I, J + 1, K + 1)], -1)
    return coords, conn.astype(np.int64)


def setup(ctx, n, order, seed, flags):
    import tptfem as T

    coords, conn = hex_mesh(n, 0.2, seed)
    ye = 10.0 ** np.random.default_rng(seed).uniform(-1, 1, len(conn))
    m = T.Mesh.from_arrays(ctx, T.HEX, coords, conn)
    bcs = [(int(q) * 3 + c, 0.0) for q in O.nodes_on_plane(coords, 0, 0.0, 1e-9) for c in range(3)]
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, order, young_e=ye)
    rhs = np.random.default_rng(seed + 1).standard_normal(len(coords) * 3)
    s = T.System.create(m, op, 3, bcs, rhs, flags=flags)
    ref = O.make_op(O.HEX, 3, 1, order=order, coef_e=ye, nu=0.3)
    return T, m, s, coords, conn, ref, bcs, rhs


def row_close(K, kc, tol):
    np.testing.assert_array_equal(K.row_ptrs, kc.row_ptrs)
    np.testing.assert_array_equal(K.col_ind, kc.col_ind)
    worst = 0.0
    for i in range(K.nrows):
        a, b = kc.row_ptrs[i], kc.row_ptrs[i + 1]
        worst = max(worst, np.abs(K.values[a:b] - kc.values[a:b]).max() / np.abs(kc.values[a:b]).max())
    assert worst <= tol, worst
    return worst


@pytest.mark.parametrize("order", [1, 2, 3])
def test_hex_fp32_mfma_values_vs_oracle(ctx, order):
    T, m, s, coords, conn, ref, bcs, rhs = setup(ctx, (4, 3, 5), order, 7 + order, 2)  # TFM_SYS_FP32
    K, _, _ = s.download()
    kr, _, _ = O.reduce_system(O.assemble(coords, conn, ref, 3), np.zeros(len(coords) * 3), bcs)
    row_close(K, kr.to_csr(), 2e-6)
    m.close()


def test_hex_fp32_mfma_matches_per_pair_kernel(ctx, monkeypatch):
    T, m, s, coords, conn, ref, bcs, rhs = setup(ctx, (5, 4, 3), 2, 3, 2)
    K, _, _ = s.download()
    monkeypatch.setenv("TFM_NO_MFMA", "1")
    s2 = T.System.create(m, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2,
                                                  young_e=10.0 ** np.random.default_rng(3).uniform(-1, 1, len(conn))),
                         3, bcs, rhs, flags=T.SYS_FP32)
    K2, _, _ = s2.download()
    row_close(K, K2, 2e-6)
    m.close()


def test_hex_precision_form_solve_vs_dense_lu(ctx):
    T, m, s, coords, conn, ref, bcs, rhs = setup(ctx, (4, 4, 4), 2, 5, 3)  # FP64 | FP32
    u = s.solve(T.SOLVE_MIXED, rtol=1e-12, max_iter=20000)
    u_ref = O.solve_with_dirichlet(O.assemble(coords, conn, ref, 3), rhs, bcs)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) <= 1e-10
    m.close()


@pytest.mark.parametrize("order", [1, 2, 3])
def test_hex_table_route_fp64_bitexact(ctx, order, monkeypatch):
    T, m, s, coords, conn, ref, bcs, rhs = setup(ctx, (4, 3, 5), order, 11 + order, 3)  # FP64 | FP32
    assert s.assembly_info()["kernel"] == "triplets"
    K, f, _ = s.download()
    kr, fr, _ = O.reduce_system(O.assemble(coords, conn, ref, 3), rhs, bcs)
    kr = kr.to_csr()
    np.testing.assert_array_equal(K.row_ptrs, kr.row_ptrs)
    np.testing.assert_array_equal(K.col_ind, kr.col_ind)
    assert np.array_equal(K.values, kr.values)
    # both arrays in one sum: the FP32 values are the FP64 values rounded
    np.testing.assert_array_equal(s.values32(), K.values.astype(np.float32).astype(np.float64))
    # the tile kernels give the same FP64 values, and the same FP32 sums of the
    # same FP32 element tables
    monkeypatch.setenv("TFM_HEX_TILE", "1")
    ye = 10.0 ** np.random.default_rng(11 + order).uniform(-1, 1, len(conn))
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, order, young_e=ye)
    t64 = T.System.create(m, op, 3, bcs, rhs, flags=T.SYS_FP64)
    assert t64.assembly_info()["kernel"] != "triplets"
    assert np.array_equal(t64.download()[0].values, K.values)
    t32 = T.System.create(m, op, 3, bcs, rhs, flags=T.SYS_FP32)
    s32 = None
    monkeypatch.delenv("TFM_HEX_TILE")
    s32 = T.System.create(m, op, 3, bcs, rhs, flags=T.SYS_FP32)
    assert np.array_equal(t32.download()[0].values, s32.download()[0].values)
    # re-assembly with new moduli
    ye2 = ye * np.exp(np.random.default_rng(5).uniform(-1, 1, len(conn)))
    s.assemble(T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, order, young_e=ye2), T.SYS_FP64)
    kr2, _, _ = O.reduce_system(O.assemble(coords, conn, O.make_op(O.HEX, 3, 1, order=order, coef_e=ye2, nu=0.3), 3),
                                rhs, bcs)
    assert np.array_equal(s.download()[0].values, kr2.to_csr().values)
    for x in (s, t64, t32, s32):
        x.close()
    m.close()


def test_hex_table_route_one_shot_assemble(ctx):
    """tfm_assemble (assemble(...).to_csr(), no reduction) on the same route."""
    import tptfem as T

    coords, conn = hex_mesh((3, 3, 2), 0.25, 4)
    ye = 10.0 ** np.random.default_rng(4).uniform(-1, 1, len(conn))
    m = T.Mesh.from_arrays(ctx, T.HEX, coords, conn)
    K = T.assemble(m, 3, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)).to_csr()
    R = O.assemble(coords, conn, O.make_op(O.HEX, 3, 1, order=2, coef_e=ye, nu=0.3), 3).to_csr()
    np.testing.assert_array_equal(K.row_ptrs, R.row_ptrs)
    np.testing.assert_array_equal(K.col_ind, R.col_ind)
    assert np.array_equal(K.values, R.values)
    m.close()
