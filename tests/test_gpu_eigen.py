"""Lanczos eigenpairs on the GPU (lanczos_eigs, crates/tpt-fem-eigen/src/lib.rs:
223-268) against the oracle's restatement.

The device dot products sum in a different (fixed) order than the reference's
sequential folds, so the Lanczos scalars differ in the last bits: Ritz values
that have converged agree to 1e-10 of the spectral radius, their Ritz vectors
to 1e-8 in angle; a full basis recovers every eigenvalue."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def tp():
    import tptfem

    return tptfem


def laplace_1d(n):
    r, c, v = [], [], []
    for i in range(n):
        r.append(i); c.append(i); v.append(2.0)
        if i + 1 < n:
            r += [i, i + 1]; c += [i + 1, i]; v += [-1.0, -1.0]
    return O.Coo(r, c, v).to_csr()


def test_reference_kat_1d_laplacian(ctx):
    """lanczos_1d_laplacian (:595-623), both ends, and the full spectrum."""
    T = tp()
    n = 10
    a = laplace_1d(n)
    m = T.Matrix.from_csr(ctx, a)
    lo = T.lanczos_eigs(m, 1, T.EIG_SMALLEST, n)[0][0]
    hi = T.lanczos_eigs(m, 1, T.EIG_LARGEST, n)[0][0]
    assert abs(lo - (2 - 2 * np.cos(np.pi / 11))) < 1e-13
    assert abs(hi - (2 - 2 * np.cos(np.pi * 10 / 11))) < 1e-13
    g = T.lanczos_eigs(m, n, T.EIG_SMALLEST, n)
    r = O.lanczos_eigs(a, n, O.SMALLEST, n)
    np.testing.assert_allclose([p[0] for p in g], [p[0] for p in r], atol=1e-13)
    for (lg, vg), (lr, vr) in zip(g, r):
        assert abs(abs(vg @ vr) - 1.0) < 1e-10


@pytest.mark.parametrize("which", ["smallest", "largest"])
def test_partial_basis_on_assembled_matrices(ctx, which):
    """Jittered P1-tet stiffness (reduced, SPD) and consistent mass with a
    40-vector basis: the converged end of the spectrum matches the oracle."""
    T = tp()
    n3 = (5, 4, 4)
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], n3)
    c = O.jitter_interior(c, n3, 0.2, 9)
    mesh = T.Mesh.from_arrays(ctx, T.TET, c, e)
    bcs = [(int(i) * 3 + k, 0.0) for i in O.nodes_on_plane(c, 0, 0.0, 1e-9) for k in range(3)]
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2)
    s = T.System.create(mesh, op, 3, bcs, None)
    kred, _, _ = s.download()
    mats = [kred, T.mass_matrix(mesh, 2.0).to_csr()]
    w = T.EIG_SMALLEST if which == "smallest" else T.EIG_LARGEST
    ow = O.SMALLEST if which == "smallest" else O.LARGEST
    for csr in mats:
        g = T.lanczos_eigs(T.Matrix.from_csr(ctx, csr), 2, w, 40)
        r = O.lanczos_eigs(csr, 2, ow, 40)
        scale = max(abs(p[0]) for p in O.lanczos_eigs(csr, 1, O.LARGEST, 40))
        assert len(g) == len(r) == 2
        if which == "largest":  # the top of the spectrum converges within 40 vectors
            for (lg, vg), (lr, vr) in zip(g, r):
                assert abs(lg - lr) <= 1e-10 * scale
                assert abs(abs(vg @ vr) - 1.0) < 1e-8
        else:  # the bottom has not converged: the same Krylov space to rounding
            for (lg, _), (lr, _) in zip(g, r):
                assert abs(lg - lr) <= 1e-8 * scale


def test_invariant_subspace_and_determinism(ctx):
    """A diagonal matrix stops the basis after e_0 (one pair); repeated runs
    are bit-identical."""
    T = tp()
    d = T.Matrix.from_csr(ctx, O.Coo([0, 1, 2], [0, 1, 2], [3.0, 1.0, 2.0]).to_csr())
    pairs = T.lanczos_eigs(d, 3, T.EIG_SMALLEST, 3)
    assert len(pairs) == 1 and pairs[0][0] == 3.0 and np.array_equal(pairs[0][1], [1.0, 0.0, 0.0])
    a = laplace_1d(300)
    m = T.Matrix.from_csr(ctx, a)
    x = T.lanczos_eigs(m, 4, T.EIG_LARGEST, 60)
    y = T.lanczos_eigs(m, 4, T.EIG_LARGEST, 60)
    assert all(p[0] == q[0] and np.array_equal(p[1], q[1]) for p, q in zip(x, y))
    with pytest.raises(T.TfmError):
        T.lanczos_eigs(m, 1, T.EIG_SMALLEST, 0)
