"""SIMP topology optimization on the GPU (topopt_simp,
crates/tpt-fem-topopt/src/lib.rs:342-385) against the oracle's loop.

The reference solves every state equation by dense LU; the GPU solves it by
PCG to a relative residual of 1e-12, so results agree to a tolerance, not
bit for bit.  Perturbing the oracle's displacements by a relative 1e-9 moves
its densities by at most 7e-8 and its compliances by 1e-8 relative on these
cases, so the bar is 1e-6 on densities and 1e-7 relative on compliance."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RHO_TOL, COMP_RTOL = 1e-6, 1e-7


def tp():
    import tptfem

    return tptfem


def run_both(ctx, nx, ny, vf, iters, **kw):
    T = tp()
    f, bcs = O.cantilever_load(nx, ny, 1.0)
    ref = O.topopt_simp(nx, ny, 1.0, 1.0, 0.3, vf, 3.0, 2.0, iters, 0.2, f, bcs)
    g = T.topopt_simp(ctx, nx, ny, 1.0, 1.0, 0.3, vf, 3.0, 2.0, iters, 0.2, f, bcs, **kw)
    return ref, g


@pytest.mark.parametrize("case", [(20, 10, 0.5, 40), (16, 8, 1.0, 10), (24, 12, 0.4, 30)])
def test_reference_cases_match_oracle(ctx, case):
    """The reference's three tests (:421-495): same densities and compliance
    history as the oracle, and the reference's own assertions hold."""
    nx, ny, vf, iters = case
    (rr, cr, ir), g = run_both(ctx, nx, ny, vf, iters)
    assert g.iterations == ir == iters
    assert np.abs(g.densities - rr).max() <= RHO_TOL
    np.testing.assert_allclose(g.compliance, cr, rtol=COMP_RTOL, atol=0)
    c = g.compliance
    assert np.all(np.diff(c) <= 1e-9) and abs(g.densities.sum() - vf * len(rr)) < 1e-6
    if vf < 1.0:
        assert c[-1] < c[0]
    else:  # full volume stays solid
        assert np.all(np.abs(g.densities - 1.0) < 1e-6)
    assert c[0] == c[1]  # iteration 1 solves the uniform start again


def test_batch_of_designs_matches_oracle_per_design(ctx):
    """Four designs in one batch, each with its own point load on the right
    edge and its own volume fraction, each equal to its own oracle run."""
    T = tp()
    nx, ny, iters = 18, 9, 12
    loads, vfs = [0, 4, 8, 2], [0.5, 0.3, 0.6, 0.45]
    F = np.zeros((4, 2 * nx * ny))
    for p, j in enumerate(loads):
        F[p, 2 * (j * nx + nx - 1) + 1] = -1.0
    _, bcs = O.cantilever_load(nx, ny, 1.0)
    g = T.topopt_simp(ctx, nx, ny, 1.0, 1.0, 0.3, vfs, 3.0, 1.5, iters, 0.2, F, bcs)
    assert g.densities.shape == (4, (nx - 1) * (ny - 1)) and g.compliance.shape == (4, iters + 1)
    for p in range(4):
        rr, cr, _ = O.topopt_simp(nx, ny, 1.0, 1.0, 0.3, vfs[p], 3.0, 1.5, iters, 0.2, F[p], bcs)
        assert np.abs(g.densities[p] - rr).max() <= RHO_TOL, p
        np.testing.assert_allclose(g.compliance[p], cr, rtol=COMP_RTOL, atol=0)


def test_mixed_precision_solves_match_oracle(ctx):
    """FP32 PCG with FP64 residual replacement as the state solver."""
    T = tp()
    (rr, cr, _), g = run_both(ctx, 20, 10, 0.5, 20, mode=T.SOLVE_MIXED, rtol=1e-12)
    assert np.abs(g.densities - rr).max() <= RHO_TOL
    np.testing.assert_allclose(g.compliance, cr, rtol=COMP_RTOL, atol=0)


def test_zero_iterations_and_wide_filter(ctx):
    """max_iter 0 returns the uniform design and its compliance; a filter
    radius spanning the whole grid visits every element, like the reference."""
    T = tp()
    f, bcs = O.cantilever_load(10, 6, 1.0)
    rr, cr, _ = O.topopt_simp(10, 6, 1.0, 1.0, 0.3, 0.5, 3.0, 2.0, 0, 0.2, f, bcs)
    g = T.topopt_simp(ctx, 10, 6, 1.0, 1.0, 0.3, 0.5, 3.0, 2.0, 0, 0.2, f, bcs)
    assert g.iterations == 0 and np.array_equal(g.densities, rr) and len(g.compliance) == 1
    np.testing.assert_allclose(g.compliance, cr, rtol=COMP_RTOL, atol=0)
    rr, cr, _ = O.topopt_simp(10, 6, 0.5, 2.0, 0.25, 0.4, 3.0, 40.0, 6, 0.1, f, bcs)
    g = T.topopt_simp(ctx, 10, 6, 0.5, 2.0, 0.25, 0.4, 3.0, 40.0, 6, 0.1, f, bcs)
    assert np.abs(g.densities - rr).max() <= RHO_TOL
    np.testing.assert_allclose(g.compliance, cr, rtol=COMP_RTOL, atol=0)


def test_contract_errors(ctx):
    """Non-zero Dirichlet values are outside the batched solver; a negative
    move limit is the reference's clamp panic."""
    T = tp()
    f, bcs = O.cantilever_load(6, 4, 1.0)
    with pytest.raises(T.TfmError):
        T.topopt_simp(ctx, 6, 4, 1.0, 1.0, 0.3, 0.5, 3.0, 2.0, 2, 0.2, f, [(0, 0.1)] + bcs[1:])
    with pytest.raises(T.TfmError):
        T.topopt_simp(ctx, 6, 4, 1.0, 1.0, 0.3, 0.5, 3.0, 2.0, 2, -0.2, f, bcs)
