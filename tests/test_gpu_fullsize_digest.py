"""Whole reduced systems at full workload size, bit for bit: the GPU's
pattern, values, right-hand side and free map hash to the digests of the
oracle's (tests/golden/fullsize_digests.json, written by
tests/golden/make_fullsize_digests.py, which also states the workloads and
the canonical byte form: int64 indices, float64 values with -0.0 folded
into +0.0)."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize_digests.json")))


def check(s, want):
    K, f, free = s.download()
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    got = {"n_free": int(len(free)), "nnz": int(len(K.col_ind)),
           "row_ptrs": h(np.asarray(K.row_ptrs, "<i8")), "col_ind": h(np.asarray(K.col_ind, "<i8")),
           "values": h(np.asarray(K.values, "<f8") + 0.0), "rhs": h(np.asarray(f, "<f8") + 0.0),
           "free": h(np.asarray(free, "<i8"))}
    assert got == want


def box_face_nodes(n):
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    on = (i == 0) | (i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz)
    return ((i * (ny + 1) + j) * (nz + 1) + k)[on]


def test_c2_elasticity_64cubed_reduced_system(ctx):
    """box_mesh 64^3, E = 1, nu = 0.3, x = 0 clamped, traction on x = 1."""
    import tptfem as T

    mesh = T.Mesh.box(ctx, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], (64, 64, 64))
    bcs = [(int(i) * 3 + k, 0.0) for i in mesh.nodes_on_plane(0, 0.0, 1e-9) for k in range(3)]
    rhs = T.plane_traction(mesh, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    s = T.System.create(mesh, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2), 3, bcs, rhs)
    assert (s.n_free, s.nnz) == (811_200, 29_340_082)
    check(s, GOLD["c2"])
    s.close()
    mesh.close()


def test_c4_batched_poisson_reduced_system(ctx):
    """256 Poisson boxes U{8..48}^3, kappa_e = exp(N(0,1)), f = 1, u = 0 on faces."""
    import tptfem as T

    rng = np.random.default_rng(2024)
    dims = [tuple(int(v) for v in d) for d in rng.integers(8, 49, size=(256, 3))]
    mesh, noff, _ = T.Mesh.box_batch(ctx, dims)
    kappa = np.exp(rng.standard_normal(mesh.n_elem))
    ids = np.concatenate([noff[p] + box_face_nodes(d) for p, d in enumerate(dims)])
    bcs = np.stack([ids.astype(np.float64), np.zeros(len(ids))], axis=1)
    s = T.System.create(mesh, T.Operator.poisson(1.0, 2, kappa_e=kappa), 1, bcs, T.source_vector(mesh, 1.0))
    check(s, GOLD["c4"])
    s.close()
    mesh.close()


def test_c3_jittered_elasticity_128cubed_reduced_system(ctx):
    """The headline workload: box_mesh 128^3 jittered +-0.2h (seed 2024),
    E_e = 10^U(-1,1) (seed 2024), nu = 0.3, x = 0 clamped, traction on x = 1;
    6,390,144 free DOFs, 283,988,358 reduced non-zeros."""
    import tptfem as T

    n = (128, 128, 128)
    mesh = T.Mesh.box(ctx, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], n)
    mesh.jitter(n, 0.2, 2024)
    ye = 10.0 ** np.random.default_rng(2024).uniform(-1.0, 1.0, mesh.n_elem)
    bcs = [(int(i) * 3 + k, 0.0) for i in mesh.nodes_on_plane(0, 0.0, 1e-9) for k in range(3)]
    rhs = T.plane_traction(mesh, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    s = T.System.create(mesh, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye), 3, bcs, rhs)
    check(s, GOLD["c3"])
    s.close()
    mesh.close()
