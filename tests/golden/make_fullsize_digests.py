"""Writes fullsize_digests.json: SHA-256 digests of the oracle's reduced
systems at three full workload sizes, for tests/test_gpu_fullsize_digest.py.

  python tests/golden/make_fullsize_digests.py [c2 c4 c3]
  (c2 + c4: about two minutes and 20 GB of memory; c3: about ten minutes)

c2: box_mesh 64^3 (1,572,864 tets), E = 1, nu = 0.3, every DOF on x = 0
    fixed, traction (0, 0, -1e-3) on x = 1.
c4: 256 Poisson boxes, cells per axis U{8..48} drawn with seed 2024,
    kappa_e = exp(N(0,1)) from the same generator, f = 1, u = 0 on the six
    faces; the block-diagonal system is the problems' reduced systems in
    order (each numbered from its offset), built here problem by problem.
c3: box_mesh 128^3 (12,582,912 tets), interior nodes jittered by +-0.2h
    (seed 2024), E_e = 10^U(-1,1) (seed 2024), nu = 0.3, x = 0 clamped,
    traction (0, 0, -1e-3) on x = 1.  The whole triplet list would need
    ~43 GB, so the rows are built in slabs of node layers: a row only
    receives the elements containing its node, in ascending element order, so
    assemble + reduce_system over the elements touching a slab (ascending)
    gives that slab's rows exactly.  The traction is assembled over the
    elements with a node on x = 1: its faces on that plane are boundary faces
    of one element each, so the load equals the whole mesh's.

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


def c3(slab=4):
    N, seed = 128, 2024
    c, e = O.box_mesh([0, 0, 0], [1, 1, 1], (N, N, N))
    c = O.jitter_interior(c, (N, N, N), 0.2, seed)
    ye = 10.0 ** np.random.default_rng(seed).uniform(-1.0, 1.0, len(e))
    bcs = [(int(i) * 3 + k, 0.0) for i in O.nodes_on_plane(c, 0, 0.0, 1e-9) for k in range(3)]
    on = np.zeros(len(c), bool)
    on[O.nodes_on_plane(c, 0, 1.0, 1e-9)] = True
    sub = np.flatnonzero(on[e].any(1))
    rhs = O.plane_traction(c, O.TET, e[sub], 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    layer = (N + 1) * (N + 1)  # nodes per x layer; node ids are x-major
    ndof = 3 * len(c)
    rps, cis, vals, frs = [np.zeros(1, np.int64)], [], [], []
    free_all, nz = None, 0
    for i0 in range(0, N + 1, slab):
        lo, hi = i0 * layer, min(i0 + slab, N + 1) * layer
        patch = np.flatnonzero(((e >= lo) & (e < hi)).any(1))
        coo = O.assemble(c, e[patch], O.make_op(O.TET, 3, 1, coef_e=ye[patch], nu=0.3), 3)
        # an explicit 0.0 at (last DOF, a clamped DOF) sizes the matrix to every DOF;
        # it only meets the lifting of a zero Dirichlet value
        coo = O.Coo(np.append(coo.rows, ndof - 1), np.append(coo.cols, 0), np.append(coo.vals, 0.0))
        kr, fr, free = O.reduce_system(coo, rhs, bcs)
        del coo
        kc = kr.to_csr()
        a, b = np.searchsorted(free, 3 * lo), np.searchsorted(free, 3 * hi)
        rp = kc.row_ptrs[a:b + 1]
        rps.append(rp[1:] - rp[0] + nz)
        cis.append(kc.col_ind[rp[0]:rp[-1]].copy())
        vals.append(kc.values[rp[0]:rp[-1]].copy())
        frs.append(fr[a:b].copy())
        nz += int(rp[-1] - rp[0])
        free_all = free
        print(f"c3 slab {i0}: rows {a}..{b}, nnz so far {nz}", flush=True)
    return digests(np.concatenate(rps), np.concatenate(cis), np.concatenate(vals), np.concatenate(frs), free_all)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c2", "c4", "c3"]
    path = os.path.join(HERE, "fullsize_digests.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for w in which:
        out[w] = {"c2": c2, "c4": c4, "c3": c3}[w]()
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))
