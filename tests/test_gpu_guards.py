"""Guards of the library's contracts on the GPU: the int32 range guard of
every pattern build's prefix sum, FP32 values that predate an FP64-only
re-assembly, re-assembly that would need pattern entries the first
assembly skipped, and a system that outlives its mesh's handle."""
import gc

import numpy as np
import pytest

import oracle as O
from test_gpu_mg import box_system

pytestmark = pytest.mark.gpu


def test_scan_total_guard(ctx):
    import tptfem as T

    assert ctx.scan_total([3, 0, 7, 1]) == 11
    assert ctx.scan_total([]) == 0
    big = 2_000_000_000
    # totals in [2^31, 2^32) (an int32 sum wraps negative), at 2^32 and above
    # (wraps to a small positive value) and exactly 2^31 - 1: all refused
    for counts in ([big, big], [2_147_483_000, 2_147_483_000, 1296], [2_147_483_647], [big] * 5,
                   [1] * 16 + [2_147_483_640]):
        with pytest.raises(T.TfmError) as ei:
            ctx.scan_total(counts)
        assert ei.value.code == T.E_UNSUPPORTED, counts
    assert ctx.scan_total([2_147_483_646]) == 2_147_483_646
    with pytest.raises(T.TfmError) as ei:
        ctx.scan_total([5, -1])
    assert ei.value.code == T.E_INVALID


def test_stale_fp32_values_are_not_used(ctx):
    """Create with FP64 and FP32 values, re-assemble the FP64 values alone with
    new moduli: the precision form on the assembled FP32 values refuses the
    stale copy, the matrix-free precision form (which reads the current
    coefficients) and the multigrid CG (whose fine passes would read the
    copy) solve the new system -- against the oracle's dense LU of the new
    coefficients."""
    import tptfem as T

    n = (6, 5, 6)
    m, s, (coords, conn, oop, rhs, bcs) = box_system(ctx, n, 3, seed=13, flags=T.SYS_FP64 | T.SYS_FP32)
    u0 = s.solve(T.SOLVE_PCG64_MG, rtol=1e-12, max_iter=2000)  # builds the solver with the FP32 copy
    coef2 = 4.0 * 10.0 ** np.random.default_rng(3).uniform(-1, 1, len(conn))
    s.assemble(T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=coef2), which=T.SYS_FP64)
    with pytest.raises(T.TfmError) as ei:
        s.solve(T.SOLVE_MIXED_CSR, rtol=1e-10)
    assert ei.value.code == T.E_INVALID
    u = s.solve(T.SOLVE_PCG64_MG, rtol=1e-12, max_iter=2000)
    ur = O.solve_with_dirichlet(O.assemble(coords, conn, O.make_op(O.TET, 3, 1, coef_e=coef2, nu=0.3), 3), rhs, bcs)
    assert np.linalg.norm(u - ur) <= 1e-9 * np.linalg.norm(ur)
    assert np.linalg.norm(u - u0) > 0.1 * np.linalg.norm(ur)  # the solution did change
    umf = s.solve(T.SOLVE_MIXED, rtol=1e-10, max_iter=20000)
    assert np.linalg.norm(umf - ur) <= 1e-8 * np.linalg.norm(ur)
    # an FP32 pass with the resident coefficients makes both forms available again
    s.assemble(None, which=T.SYS_FP32)
    um = s.solve(T.SOLVE_MIXED_CSR, rtol=1e-10, max_iter=20000)
    assert np.linalg.norm(um - ur) <= 1e-8 * np.linalg.norm(ur)
    s.close()
    m.close()


def test_reassembly_cannot_revive_zero_coefficient_elements(ctx):
    """Elements with coefficient 0 at creation own no pattern entries (all
    their entries were skipped): giving them a non-zero coefficient later is
    refused; the other direction keeps explicit zeros and equal values."""
    import tptfem as T

    n = (3, 3, 3)
    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    coords = O.jitter_interior(coords, n, 0.2, 2)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    bcs = [(int(v), 0.0) for v in O.nodes_on_plane(coords, 0, 0.0, 1e-9)]
    k0 = np.ones(len(conn))
    k0[::7] = 0.0
    s = T.System.create(m, T.Operator.poisson(1.0, 2, kappa_e=k0), 1, bcs, np.ones(len(coords)))
    with pytest.raises(T.TfmError) as ei:
        s.assemble(T.Operator.poisson(1.0, 2, kappa_e=np.ones(len(conn))))
    assert ei.value.code == T.E_INVALID
    with pytest.raises(T.TfmError):
        s.assemble(T.Operator.poisson(2.0, 2))  # scalar coefficient: every element non-zero
    # zero -> zero and non-zero -> other values are fine, bit-identical to the oracle on the pattern
    k1 = np.where(k0 == 0.0, 0.0, 3.0)
    s.assemble(T.Operator.poisson(1.0, 2, kappa_e=k1))
    kg, _, _ = s.download()
    kr, _, _ = O.reduce_system(O.assemble(coords, conn, O.make_op(O.TET, 3, 0, coef_e=k1), 1), np.ones(len(coords)), bcs)
    kr = kr.to_csr()
    np.testing.assert_array_equal(kg.col_ind, kr.col_ind)
    assert np.array_equal(kg.values, kr.values)
    s.close()
    # a scalar coefficient of 0 at creation: every element is void
    s = T.System.create(m, T.Operator.poisson(0.0, 2), 1, bcs, np.ones(len(coords)))
    with pytest.raises(T.TfmError):
        s.assemble(T.Operator.poisson(1.0, 2))
    s.close()
    m.close()


@pytest.mark.parametrize("dpn", [1, 3])
def test_system_outlives_its_mesh_handle(ctx, dpn):
    """A system holds its mesh: after the caller destroys the mesh handle (and
    new allocations take any freed memory), re-assembly with new
    coefficients, the batched solve and the single solve still run on the
    mesh the system was built on -- bit-identical values and the oracle's
    dense-LU solution."""
    import tptfem as T

    n = (4, 3, 5)
    m, s, (coords, conn, oop, rhs, bcs) = box_system(ctx, n, dpn, seed=5, flags=T.SYS_FP64 | T.SYS_FP32)
    m.close()
    del m
    gc.collect()
    # same-sized allocations that would reuse the mesh's memory had it been freed
    others = [T.Mesh.box(ctx, [2, 2, 2], [3, 3, 3], n) for _ in range(3)]
    coef2 = 2.0 + np.random.default_rng(8).uniform(0, 1, len(conn))
    op2 = (T.Operator.poisson(1.0, 2, kappa_e=coef2) if dpn == 1
           else T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=coef2))
    oop2 = O.make_op(O.TET, 3, 0 if dpn == 1 else 1, coef_e=coef2, nu=0.3)
    s.assemble(op2, T.SYS_FP64 | T.SYS_FP32)
    coo = O.assemble(coords, conn, oop2, dpn)
    kg, fg, _ = s.download()
    kr, fr, _ = O.reduce_system(coo, rhs, bcs)
    kr = kr.to_csr()
    np.testing.assert_array_equal(kg.row_ptrs, kr.row_ptrs)
    np.testing.assert_array_equal(kg.col_ind, kr.col_ind)
    assert np.array_equal(kg.values, kr.values)
    ur = O.solve_with_dirichlet(coo, rhs, bcs)
    for mode in (T.SOLVE_PCG64, T.SOLVE_MIXED):
        ub = s.solve_batch([0, s.n_dof], mode=mode, rtol=1e-11)[0]
        assert np.linalg.norm(ub - ur) <= 1e-8 * np.linalg.norm(ur), mode
    u = s.solve(T.SOLVE_PCG64, rtol=1e-12)
    assert np.linalg.norm(u - ur) <= 1e-10 * np.linalg.norm(ur)
    s.close()
    for o in others:
        o.close()
