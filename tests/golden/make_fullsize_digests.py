"""Writes fullsize_digests.json: SHA-256 digests of the oracle's reduced
systems at two full workload sizes, for tests/test_gpu_fullsize_digest.py.

  python tests/golden/make_fullsize_digests.py      (about two minutes, ~20 GB RAM)

c2: box_mesh 64^3 (1,572,864 tets), E = 1, nu = 0.3, every DOF on x = 0
    fixed, traction (0, 0, -1e-3) on x = 1.
c4: 256 Poisson boxes, cells per axis U{8..48} drawn with seed 2024,
    kappa_e = exp(N(0,1)) from the same generator, f = 1, u = 0 on the six
    faces; the block-diagonal system is the problems' reduced systems in
    order (each numbered from its offset), built here problem by problem.

Each digest covers one array in a canonical byte form: row_ptrs and col_ind
as little-endian int64, values and the reduced right-hand side as float64
with -0.0 folded into +0.0, the free map as int64.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def digests(row_ptrs, col_ind, values, fred, free):
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    return {"n_free": int(len(free)), "nnz": int(len(col_ind)),
            "row_ptrs": h(np.asarray(row_ptrs, "<i8")), "col_ind": h(np.asarray(col_ind, "<i8")),
            "values": h(np.asarray(values, "<f8") + 0.0), "rhs": h(np.asarray(fred, "<f8") + 0.0),
            "free": h(np.asarray(free, "<i8"))}


def box_face_nodes(n):
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    on = (i == 0) | (i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz)
    return ((i * (ny + 1) + j) * (nz + 1) + k)[on]


def c2():
    n = (64, 64, 64)
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], n)
    coo = O.assemble(c, e, O.make_op(O.TET, 3, 1, coef=1.0, nu=0.3), 3)
    rhs = O.plane_traction(c, O.TET, e, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    bcs = [(int(i) * 3 + k, 0.0) for i in O.nodes_on_plane(c, 0, 0.0, 1e-9) for k in range(3)]
    kr, fr, free = O.reduce_system(coo, rhs, bcs)
    del coo
    kc = kr.to_csr()
    return digests(kc.row_ptrs, kc.col_ind, kc.values, fr, free)


def c4():
    rng = np.random.default_rng(2024)
    dims = [tuple(int(v) for v in d) for d in rng.integers(8, 49, size=(256, 3))]
    ne = sum(6 * a * b * c for a, b, c in dims)
    kappa = np.exp(rng.standard_normal(ne))
    rps, cis, vals, frs, frees = [np.zeros(1, np.int64)], [], [], [], []
    eoff = noff = roff = nzoff = 0
    for d in dims:
        c, e = O.box_mesh([0, 0, 0], [1, 1, 1], d)
        coo = O.assemble(c, e, O.make_op(O.TET, 3, 0, coef_e=kappa[eoff:eoff + len(e)]), 1)
        bp = [(int(i), 0.0) for i in box_face_nodes(d)]
        kr, fr, free = O.reduce_system(coo, O.poisson_rhs(c, O.TET, e, 1.0), bp)
        kc = kr.to_csr()
        # rows of a problem without free DOFs would be absent; every box here has interior nodes
        assert kc.nrows == len(free)
        rps.append(kc.row_ptrs[1:] + nzoff)
        cis.append(kc.col_ind + roff)
        vals.append(kc.values)
        frs.append(fr)
        frees.append(free + noff)
        eoff += len(e)
        noff += len(c)
        roff += len(free)
        nzoff += len(kc.col_ind)
    return digests(np.concatenate(rps), np.concatenate(cis), np.concatenate(vals), np.concatenate(frs),
                   np.concatenate(frees))


if __name__ == "__main__":
    out = {"c2": c2(), "c4": c4()}
    with open(os.path.join(HERE, "fullsize_digests.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))
