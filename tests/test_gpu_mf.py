"""The precision form's matrix-free FP32 operator on box meshes of P1 tets.

On a Mesh.box system, TFM_SOLVE_MIXED runs its FP32 CG on an operator that
recomputes every element's contribution from the coordinates, the element
coefficients and the vector (FP64 determinants, FP32 otherwise) instead of
reading stored FP32 values.  Checked here:
* K x against the oracle's assembled FP64 reduced matrix, for elasticity and
  Poisson on jittered boxes with per-element coefficients: every free row
  within 2e-6 of max |K x| (FP32 rounding of the element arithmetic), fixed
  rows exactly 0, and two products bit-identical (fixed summation order);
* the mixed solve against the oracle's dense-LU solution to 1e-10 relative,
  with the same solve on the assembled FP32 values (TFM_SOLVE_MIXED_CSR)
  beside it;
* refusal when the FP64 values predate the last coefficients, and that
  systems not built on a box mesh have no matrix-free operator.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def tp():
    import tptfem

    return tptfem


def face_bcs(coords, dpn, faces):
    bcs = []
    for axis, val in faces:
        for n in O.nodes_on_plane(coords, axis, val, 1e-9):
            for c in range(dpn):
                bcs.append((int(n) * dpn + c, 0.0))
    return bcs


def box_system(ctx, n, dpn, seed, faces=((0, 0.0),), rhs=None):
    T = tp()
    m = T.Mesh.box(ctx, [0, 0, 0], [1.0, 0.8, 1.3], n)
    m.jitter(n, 0.2, seed)
    coords, conn = m.download()
    cf = 10.0 ** np.random.default_rng(seed).uniform(-1, 1, m.n_elem)
    if dpn == 1:
        op = T.Operator.poisson(1.0, 2, kappa_e=cf)
        ref = O.make_op(O.TET, 3, 0, coef_e=cf)
    else:
        op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=cf)
        ref = O.make_op(O.TET, 3, 1, coef_e=cf, nu=0.3)
    bcs = face_bcs(coords, dpn, faces)
    s = T.System.create(m, op, dpn, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    return T, m, s, coords, conn, op, ref, bcs, cf


@pytest.mark.parametrize("dpn,n", [(3, (6, 5, 4)), (3, (9, 3, 17)), (1, (7, 6, 5)), (1, (1, 1, 1))])
def test_matrix_free_product_vs_oracle_matrix(ctx, dpn, n):
    T, m, s, coords, conn, op, ref, bcs, _ = box_system(ctx, n, dpn, 11 + n[0], faces=((0, 0.0), (2, 1.3)))
    ndof = len(coords) * dpn
    kr, _, free = O.reduce_system(O.assemble(coords, conn, ref, dpn), np.zeros(ndof), bcs)
    kc = kr.to_csr()
    rng = np.random.default_rng(5)
    x = np.zeros(ndof)
    x[free] = rng.standard_normal(free.size)
    y = s.mf_apply(x)
    y2 = s.mf_apply(x)
    assert np.array_equal(y.view(np.uint64), y2.view(np.uint64))
    fixed = np.setdiff1d(np.arange(ndof), free)
    assert np.all(y[fixed] == 0.0)
    yr = np.zeros(free.size)
    for i in range(free.size):
        a, b = kc.row_ptrs[i], kc.row_ptrs[i + 1]
        yr[i] = kc.values[a:b] @ x[free][kc.col_ind[a:b]]
    err = np.abs(y[free] - yr).max() / np.abs(yr).max()
    assert err <= 2e-6, err
    m.close()


@pytest.mark.parametrize("dpn", [1, 3])
def test_matrix_free_mixed_solve_vs_dense_lu(ctx, dpn):
    n = (5, 4, 6)
    T = tp()
    m0 = T.Mesh.box(ctx, [0, 0, 0], [1.0, 0.8, 1.3], n)
    rhs = np.random.default_rng(9).standard_normal(m0.n_nodes * dpn)
    m0.close()
    T, m, s, coords, conn, op, ref, bcs, _ = box_system(ctx, n, dpn, 3, rhs=rhs)
    u_ref = O.solve_with_dirichlet(O.assemble(coords, conn, ref, dpn), rhs, bcs)
    nr = np.linalg.norm(u_ref)
    for mode in (T.SOLVE_MIXED, T.SOLVE_MIXED_CSR):
        u = s.solve(mode, rtol=1e-12, max_iter=20000)
        assert s.stats.rel_residual <= 1e-12
        assert np.linalg.norm(u - u_ref) / nr <= 1e-10, mode
    m.close()


def test_matrix_free_refuses_stale_fp64_values(ctx):
    T, m, s, coords, conn, op, ref, bcs, cf = box_system(ctx, (4, 4, 4), 3, 2,
                                                         rhs=np.ones(5 ** 3 * 3))
    s.solve(T.SOLVE_MIXED, rtol=1e-8)
    op2 = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=cf * 2.0)
    s.assemble(op2, T.SYS_FP32)  # new coefficients reach the FP32 values only
    with pytest.raises(T.TfmError):
        s.solve(T.SOLVE_MIXED, rtol=1e-8)
    s.assemble(None, T.SYS_FP64)
    s.solve(T.SOLVE_MIXED, rtol=1e-8)
    m.close()


def test_no_matrix_free_operator_off_box_meshes(ctx):
    T = tp()
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (3, 3, 3))
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    bcs = face_bcs(coords, 3, [(0, 0.0)])
    rhs = np.ones(len(coords) * 3)
    s = T.System.create(m, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2), 3, bcs, rhs,
                        flags=T.SYS_FP64 | T.SYS_FP32)
    with pytest.raises(T.TfmError):
        s.mf_apply(np.zeros(len(coords) * 3))
    u = s.solve(T.SOLVE_MIXED, rtol=1e-10)  # the assembled FP32 operator
    u_ref = O.solve_with_dirichlet(O.assemble(coords, conn, O.make_op(O.TET, 3, 1, nu=0.3), 3), rhs, bcs)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) <= 1e-8
    m.close()
