"""tfm_system_solve_noalloc (the FP64 Jacobi PCG with its work space in a
fixed device array) against tfm_system_solve and the oracle's dense LU, on
two calls in a row -- the second on a smaller system, so stale contents of
the shared array would show -- plus its refusals."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _system(ctx, n, dpn, seed):
    import tptfem as T

    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    coords = O.jitter_interior(coords, n, 0.2, seed)
    m = T.Mesh.from_arrays(ctx, T.TET, coords, conn)
    rng = np.random.default_rng(seed)
    if dpn == 3:
        ye = 10.0 ** rng.uniform(-1, 1, len(conn))
        op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)
        oop = O.make_op(O.TET, 3, 1, coef_e=ye, nu=0.3)
        rhs = O.plane_traction(coords, O.TET, conn, 3, 0, 1.0, 1e-9, [0.0, 0.2, -1.0])
        bcs = [(int(v) * 3 + c, 0.0) for v in O.nodes_on_plane(coords, 0, 0.0, 1e-9) for c in range(3)]
    else:
        ke = np.exp(rng.standard_normal(len(conn)))
        op = T.Operator.poisson(1.0, 2, kappa_e=ke)
        oop = O.make_op(O.TET, 3, 0, coef_e=ke)
        rhs = O.poisson_rhs(coords, O.TET, conn, 1.0)
        bcs = [(int(v), 0.0) for a in range(3) for v in O.nodes_on_plane(coords, a, 0.0, 1e-9)]
    s = T.System.create(m, op, dpn, bcs, rhs)
    return m, s, O.solve_with_dirichlet(O.assemble(coords, conn, oop, dpn), rhs, bcs)


def test_noalloc_solve_matches_the_allocating_solve_twice(ctx):
    results = []
    for n, dpn, seed in (((9, 8, 7), 3, 1), ((5, 4, 6), 1, 2)):
        m, s, ur = _system(ctx, n, dpn, seed)
        u_ref = s.solve(rtol=1e-12)
        before = _free_device_bytes()
        u = s.solve_noalloc(rtol=1e-12)
        after = _free_device_bytes()
        st = s.stats
        assert after == before, (before, after)
        assert st.rel_residual <= 1e-11 and st.iterations > 0
        assert np.abs(u - u_ref).max() <= 1e-10 * np.abs(u_ref).max()
        assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()
        # a repeat on the same system is bit-identical (deterministic reductions)
        assert np.array_equal(s.solve_noalloc(rtol=1e-12), u)
        results.append(u)
        s.close()
        m.close()
    assert len(results[1]) < len(results[0])


def test_noalloc_refusals(ctx):
    import tptfem as T

    m, s, _ = _system(ctx, (3, 3, 3), 1, 3)
    with pytest.raises(T.TfmError) as ei:
        s.ctx.check(T._lib.tfm_system_solve_noalloc(
            s.h, ctypes.byref(T.SolveOpts(mode=T.SOLVE_MIXED, rtol=1e-10, max_iter=100, inner_rtol=0.3)),
            np.zeros(s.n_dof).ctypes.data, None))
    assert ei.value.code == T.E_UNSUPPORTED
    s.close()
    m.close()
    # 108^3 nodes: more DOFs than the fixed work space holds
    big = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], (107, 107, 107))
    sb = T.System.create(big, T.Operator.poisson(1.0, 2), 1, [(0, 0.0)], np.ones(big.n_nodes))
    with pytest.raises(T.TfmError) as ei:
        sb.solve_noalloc()
    assert ei.value.code == T.E_UNSUPPORTED and "work space" in str(ei.value)
    sb.close()
    big.close()
