"""GPU parity of the batching form: many problems through one assembly pass
and one segmented PCG, each compared with its own oracle direct solve."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def boundary_nodes(n):
    nx, ny, nz = n
    ids = []
    for i in range(nx + 1):
        for j in range(ny + 1):
            for k in range(nz + 1):
                if i in (0, nx) or j in (0, ny) or k in (0, nz):
                    ids.append((i * (ny + 1) + j) * (nz + 1) + k)
    return np.array(ids, dtype=np.int64)


N3 = [(3, 4, 5), (6, 6, 6), (2, 3, 2), (5, 4, 3), (4, 4, 4)]


def test_box_batch_mesh_is_concatenated_box_meshes(ctx):
    import tptfem as T

    m, noff, eoff = T.Mesh.box_batch(ctx, N3)
    c, e = m.download()
    for p, n in enumerate(N3):
        rc, re = O.box_mesh([0, 0, 0], [1, 1, 1], n)
        assert np.array_equal(c[noff[p]:noff[p + 1]], rc)
        assert np.array_equal(e[eoff[p]:eoff[p + 1]] - noff[p], re)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_batched_poisson_matches_per_problem_direct_solves(ctx, mode):
    import tptfem as T

    m, noff, eoff = T.Mesh.box_batch(ctx, N3)
    kappa = np.exp(np.random.default_rng(4).standard_normal(m.n_elem))
    bcs = [(int(noff[p] + i), 0.0) for p, n in enumerate(N3) for i in boundary_nodes(n)]
    rhs = T.source_vector(m, 1.0)
    flags = T.SYS_FP64 | (T.SYS_FP32 if mode == "mixed" else 0)
    s = T.System.create(m, T.Operator.poisson(1.0, 2, kappa_e=kappa), 1, bcs, rhs, flags=flags)
    u = s.solve_batch(noff, mode=T.SOLVE_PCG64 if mode == "plain" else T.SOLVE_MIXED, rtol=1e-12)[0]
    for p, n in enumerate(N3):
        c, e = O.box_mesh([0, 0, 0], [1, 1, 1], n)
        kp = kappa[eoff[p]:eoff[p + 1]]
        coo = O.assemble(c, e, O.make_op(O.TET, 3, 0, coef_e=kp), 1)
        rp = O.poisson_rhs(c, O.TET, e, 1.0)
        ur = O.solve_with_dirichlet(coo, rp, [(int(i), 0.0) for i in boundary_nodes(n)])
        up = u[noff[p]:noff[p + 1]]
        err = np.linalg.norm(up - ur) / np.linalg.norm(ur)
        assert err <= (1e-10 if mode == "plain" else 1e-8), (p, err)
    assert all(st[2] <= 1e-12 for st in s.batch_stats)


def test_batched_simp_multi_rhs_matches_direct(ctx):
    import tptfem as T

    P, nx, ny = 3, 9, 5
    m = T.Mesh.grid_batch(ctx, P, nx, ny, 1.0)
    ne = (nx - 1) * (ny - 1)
    rng = np.random.default_rng(8)
    rho = rng.uniform(0.01, 1.0, P * ne)
    ndof1 = 2 * nx * ny
    bcs = []
    for p in range(P):
        for j in range(ny):
            node = p * nx * ny + j * nx
            bcs += [(2 * node, 0.0), (2 * node + 1, 0.0)]
    loads = rng.integers(0, ny, size=(P, 8))
    f = np.zeros((8, P * ndof1))
    for p in range(P):
        for k in range(8):
            node = p * nx * ny + loads[p, k] * nx + (nx - 1)
            f[k, 2 * node + 1] = -1.0
    op = T.Operator.simp(rho, nu=0.3, penal=3.0, e0=1.0, emin=1e-9, h=1.0)
    s = T.System.create(m, op, 2, bcs, None, flags=T.SYS_FP64 | T.SYS_FP32)
    assert s.assembly_info()["kernel"] == "simp_grid"
    # both arrays in one pass: the FP32 values are the FP64 values rounded
    np.testing.assert_array_equal(s.values32(), s.download()[0].values.astype(np.float32).astype(np.float64))
    off = np.arange(P + 1) * ndof1
    u = s.solve_batch(off, rhs=f, nrhs=8, mode=T.SOLVE_PCG64, rtol=1e-12)
    um = s.solve_batch(off, rhs=f, nrhs=8, mode=T.SOLVE_MIXED, rtol=1e-10)
    coords, conn = O.grid(nx, ny, 1.0)
    ke0 = O.simp_ke_unit(0.3, 1.0)
    lb = [(2 * j * nx + c, 0.0) for j in range(ny) for c in (0, 1)]
    for p in range(P):
        coo = O.assemble(coords, conn, O.make_op(O.QUAD, 2, 2, rho=rho[p * ne:(p + 1) * ne], penal=3.0, e0=1.0,
                                                 emin=1e-9, ke0=ke0), 2)
        for k in range(8):
            fp = f[k, p * ndof1:(p + 1) * ndof1]
            ur = O.solve_with_dirichlet(coo, fp, lb)
            up = u[k, p * ndof1:(p + 1) * ndof1]
            assert np.linalg.norm(up - ur) <= 1e-9 * np.linalg.norm(ur), (p, k)
            upm = um[k, p * ndof1:(p + 1) * ndof1]
            assert np.linalg.norm(upm - ur) <= 1e-8 * np.linalg.norm(ur), (p, k)
            # compliance f.u
            assert abs(fp @ up - fp @ ur) <= 1e-9 * abs(fp @ ur)
