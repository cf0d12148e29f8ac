"""Batched CG on problems spread over many blocks forms each iteration's stop
test, beta and alpha once per problem (one block per problem) instead of in
every block from all of its problem's partial sums.  Both forms add the same
partials in the same order, so the solutions, iteration counts and residuals
must be bitwise equal.  TFM_SCALAR_BLOCKS moves the block count at which the
solver switches; 0 puts every solve here in the scalar form, a large value
keeps every one in the per-block form."""
import contextlib
import os

import numpy as np
import pytest

from test_gpu_graph_stages import poisson_batch
from test_gpu_mg import box_system, simp_batch

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def scalar_blocks(n):
    os.environ["TFM_SCALAR_BLOCKS"] = str(n)
    try:
        yield
    finally:
        del os.environ["TFM_SCALAR_BLOCKS"]


def both_forms(run):
    with scalar_blocks(1 << 30):
        a = run()
    with scalar_blocks(0):
        b = run()
    return a, b


@pytest.mark.parametrize("rtol", [1e-11, 1e-3])
def test_scalar_form_equals_per_block_form_jacobi(ctx, rtol):
    import tptfem as T

    m, noff, eoff, kappa, s = poisson_batch(T, ctx, T.SYS_FP64)

    def run():
        u = s.solve_batch(noff, mode=T.SOLVE_PCG64, rtol=rtol)
        return u, [tuple(x[:3]) for x in s.batch_stats]

    (u0, st0), (u1, st1) = both_forms(run)
    assert np.array_equal(u0, u1)
    assert st0 == st1
    s.close()
    m.close()


@pytest.mark.parametrize("nrhs", [1, 8])
def test_scalar_form_equals_per_block_form_mg_grid(ctx, nrhs):
    import tptfem as T

    m, s, rho, f, off = simp_batch(ctx, 3, 26, 13, seed=5)

    def run():
        u = s.solve_batch(off, rhs=f[:nrhs] if nrhs > 1 else f[0], nrhs=nrhs, mode=T.SOLVE_PCG64_MG, rtol=1e-12,
                          max_iter=5000)
        return u, [tuple(x[:3]) for x in s.batch_stats]

    (u0, st0), (u1, st1) = both_forms(run)
    assert np.array_equal(u0, u1)
    assert st0 == st1
    s.close()
    m.close()


@pytest.mark.parametrize("dpn", [1, 3])
def test_scalar_form_equals_per_block_form_mg_box(ctx, dpn):
    import tptfem as T

    m, s, _ = box_system(ctx, (9, 7, 6), dpn, seed=17 + dpn)

    def run():
        u = s.solve(T.SOLVE_PCG64_MG, rtol=1e-12, max_iter=2000)
        st = s.stats
        return u, (st.iterations, st.rel_residual)

    (u0, st0), (u1, st1) = both_forms(run)
    assert np.array_equal(u0, u1)
    assert st0 == st1
    s.close()
    m.close()
