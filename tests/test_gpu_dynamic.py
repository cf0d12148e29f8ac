"""Time integration on the GPU (crates/tpt-fem-dynamic/src/lib.rs) against the
oracle's restatement: central_difference bit for bit (elementwise updates and
CSR row sums in the reference's order), newmark within the PCG tolerance (the
reference solves each step by dense LU)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def tp():
    import tptfem

    return tptfem


def sdof(mass, stiff):
    return (O.Coo([0], [0], [mass]).to_csr(), O.Coo([], [], []).to_csr(), O.Coo([0], [0], [stiff]).to_csr())


def dev(ctx, *csrs):
    T = tp()
    return [T.Matrix.from_csr(ctx, c) for c in csrs]


def test_reference_sdof_tests(ctx):
    """newmark_matches_closed_form_sdof, central_difference_rejects_over_cfl_step
    and newmark_conserves_energy_undamped (:507-550)."""
    T = tp()
    m, c, k = dev(ctx, *sdof(1.0, 4.0))
    h = T.newmark(m, c, k, [1.0], [0.0], [0.0], 0.005, 0.25, 0.5, 400)
    assert abs(h[400][0] - np.cos(4.0)) < 5e-3
    with pytest.raises(T.CflViolation) as ex:
        T.central_difference(m, c, k, [1.0], [0.0], [0.0], 2.0, 10)
    assert ex.value.critical == 1.0 and "exceeds the critical (CFL) limit" in str(ex.value)
    m2, c2, k2 = dev(ctx, *sdof(2.0, 8.0))
    h = T.newmark(m2, c2, k2, [1.0], [0.0], [0.0], 0.002, 0.25, 0.5, 500)
    for w in range(1, len(h) - 1):
        vm = (h[w + 1][0] - h[w - 1][0]) / (2 * 0.002)
        assert abs(0.5 * 2 * vm * vm + 0.5 * 8 * h[w][0] ** 2 - 4.0) < 1e-2


def elastic_bar():
    """Jittered P1-tet elasticity, x = 0 clamped (reduced), consistent mass and
    Rayleigh damping C = 0.05 M + 0.001 K, a traction-like load on x = 1."""
    n3 = (4, 2, 2)
    c, e = O.box_mesh([0, 0, 0], [2.0, 0.5, 0.5], n3)
    c = O.jitter_interior(c, n3, 0.2, 12)
    bcs = [(int(i) * 3 + a, 0.0) for i in O.nodes_on_plane(c, 0, 0.0, 1e-9) for a in range(3)]
    kc = O.assemble(c, e, O.make_op(O.TET, 3, 1, coef=100.0, nu=0.3), 3)
    mc = O.mass_matrix(c, O.TET, e, 1.0)
    f = O.plane_traction(c, O.TET, e, 3, 0, 2.0, 1e-9, [0.0, 0.0, -1.0])
    kr, fr, free = O.reduce_system(kc, f, bcs)
    mr, _, _ = O.reduce_system(mc, np.zeros(len(f)), bcs)
    K, M = kr.to_csr(), mr.to_csr()
    C = O.Coo(np.concatenate([mr.rows, kr.rows]), np.concatenate([mr.cols, kr.cols]),
              np.concatenate([mr.vals * 0.05, kr.vals * 0.001])).to_csr()
    return M, C, K, fr


def test_central_difference_bitexact_on_elasticity(ctx):
    T = tp()
    M, C, K, f = elastic_bar()
    n = M.nrows
    rng = np.random.default_rng(3)
    u0, v0 = rng.uniform(-1e-3, 1e-3, n), rng.uniform(-1e-2, 1e-2, n)
    mh = O.Coo(*_lumped(M)).to_csr()  # a lumped (diagonal) mass, as explicit codes use
    crit = None
    try:
        O.central_difference(mh, C, K, u0, v0, f, 1e9, 1)
    except O.CflViolation as ex:
        crit = ex.critical
    dt = 0.5 * crit
    fs = np.sin(np.arange(61) * 0.1)
    r = O.central_difference(mh, C, K, u0, v0, f, dt, 60, fscale=fs)
    g = T.central_difference(*dev(ctx, mh, C, K), u0, v0, f, dt, 60, fscale=fs)
    assert np.array_equal(g, r)
    with pytest.raises(T.CflViolation) as ex:
        T.central_difference(*dev(ctx, mh, C, K), u0, v0, f, 1.01 * crit, 5)
    assert ex.value.critical == crit
    # consistent mass: its row sums are the lumped masses the scheme uses
    g = T.central_difference(*dev(ctx, M, C, K), u0, v0, f, dt, 20)
    r = O.central_difference(M, C, K, u0, v0, f, dt, 20)
    assert np.array_equal(g, r)


def _lumped(M):
    rows = np.arange(M.nrows)
    sums = np.array([M.values[M.row_ptrs[i]:M.row_ptrs[i + 1]].sum() for i in rows])
    return rows, rows, sums


def test_newmark_on_elasticity_matches_dense_lu(ctx):
    """40 Newmark steps of the damped bar under a ramped load: every step's
    displacements within 1e-9 of the oracle's dense-LU history (relative to
    its largest displacement)."""
    T = tp()
    M, C, K, f = elastic_bar()
    n = M.nrows
    u0, v0 = np.zeros(n), np.zeros(n)
    fs = np.minimum(np.arange(41) / 10.0, 1.0)
    r = O.newmark(M, C, K, u0, v0, f, 0.01, 0.25, 0.5, 40, fscale=fs)
    m, c, k = dev(ctx, M, C, K)
    g = T.newmark(m, c, k, u0, v0, f, 0.01, 0.25, 0.5, 40, fscale=fs)
    scale = np.abs(r).max()
    assert scale > 0 and np.abs(g - r).max() <= 1e-9 * scale
    assert m.stats.outer == 40 and m.stats.rel_residual <= 1e-12
