"""Writes the golden fixtures under tests/golden/ (run from the repo root:
python tests/golden/make_golden.py).

Sources, none of them the oracle:
  * kats.json -- known answers transcribed from the reference's own unit tests
    and examples (file:line recorded per case);
  * c1_fd_solution.json -- config C1 (Poisson -lap u = 1 on the unit cube,
    box_mesh 8^3 Kuhn tets, u = 0 on the boundary).  On that mesh the reduced
    P1 stiffness is exactly h*(6I - neighbours) and the interior load exactly
    h^3, so the FE solution is the 7-point finite-difference solution, which a
    DST-I diagonalises: u = DST^-1[ DST[h^2 f] / eigenvalues ] (scipy.fft);
  * kuhn_counts.json -- nnz of the assembled elasticity matrix on box_mesh
    n^3 Kuhn meshes, closed form 111n^3 + 171n^2 + 45n - 3 (and the reduced
    x = 0 clamped count 111n^3 + 60n^2 - 57n - 14);
  * stale_vtk.json -- the temperature maximum stored in the reference's
    committed thermal_solve.vtk (6^3 box, k = 1, f = 1), produced by an older
    build that applied J^-1 transposed; the current algorithm must NOT match it.
"""
import json
import os

import numpy as np
from scipy.fft import dstn, idstn

HERE = os.path.dirname(os.path.abspath(__file__))


def c1_solution(n=8):
    h = 1.0 / n
    m = n - 1
    k = np.arange(1, n)
    lam1 = 2.0 - 2.0 * np.cos(np.pi * k / n)
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    f = np.full((m, m, m), h * h)  # h^2 * f with f = 1
    u = idstn(dstn(f, type=1) / lam, type=1)
    return u  # interior nodes (i, j, k), k fastest, i, j, k in 1..n-1


def main():
    kats = {
        "coo_to_csr": {"src": "crates/tpt-fem-sparse/src/lib.rs:378-391",
                       "rows": [1, 0, 0, 0, 1], "cols": [1, 1, 0, 0, 0], "vals": [1, 1, 1, 1, 1],
                       "row_ptrs": [0, 2, 4], "col_ind": [0, 1, 0, 1], "values": [2, 1, 1, 1]},
        "coo_doctest": {"src": "crates/tpt-fem-sparse/src/lib.rs:16-31",
                        "rows": [0, 0, 0, 1, 1, 1], "cols": [0, 0, 1, 0, 1, 1],
                        "vals": [1.0, 1.0, 1.0, 1.0, 1.5, 1.5],
                        "row_ptrs": [0, 2, 4], "values": [2.0, 1.0, 1.0, 3.0]},
        "solve_2x2": {"src": "crates/tpt-fem-sparse/src/lib.rs:394-404",
                      "rows": [0, 0, 1, 1], "cols": [0, 1, 0, 1], "vals": [2, 1, 1, 3],
                      "rhs": [3, 5], "x": [0.8, 1.4], "tol": 1e-10},
        "solve_3x3": {"src": "crates/tpt-fem-sparse/src/lib.rs:406-423",
                      "rows": [0, 0, 1, 1, 1, 2, 2], "cols": [0, 1, 0, 1, 2, 1, 2],
                      "vals": [4, -1, -1, 4, -1, -1, 4], "rhs": [3, 2, 3], "x": [1, 1, 1], "tol": 1e-9},
        "matvec": {"src": "crates/tpt-fem-sparse/src/lib.rs:500-512",
                   "rows": [0, 0, 1], "cols": [0, 1, 1], "vals": [2, 1, 3], "x": [1, -2], "y": [0, -6],
                   "tol": 1e-14},
        "dirichlet_line": {"src": "crates/tpt-fem-assembly/src/lib.rs:795-811",
                           "coords": [[0.0], [0.5], [1.0]], "conn": [[0, 1], [1, 2]],
                           "bcs": [[0, 0.0], [2, 1.0]], "u": [0.0, 0.5, 1.0], "tol": 1e-12},
        "neumann_total": {"src": "crates/tpt-fem-assembly/src/lib.rs:814-832",
                          "coords": [[0, 0], [1, 0], [0, 1]], "conn": [[0, 1, 2]], "total": 2 + 2 ** 0.5,
                          "tol": 1e-9},
        "assemble_identity_tris": {"src": "crates/tpt-fem-assembly/src/lib.rs:775-792",
                                   "nrows": 4, "ncols": 4, "nnz": 4},
        "jacobian_dets": {"src": "crates/tpt-fem-element/src/lib.rs:817-907",
                          "cases": [
                              {"cell": "tri", "nodes": [[0, 0], [1, 0], [0, 1]], "xi": [0.3, 0.2], "det": 1.0,
                               "jac": [1, 0, 0, 1]},
                              {"cell": "line", "nodes": [[2.0], [5.0]], "xi": [0.0], "det": 1.5},
                              {"cell": "tri", "nodes": [[1, 1], [3, 1], [1, 4]], "xi": [0.2, 0.3], "det": 6.0,
                               "jac": [2, 0, 0, 3], "pgrad_in": [1.0, 0.0], "pgrad": [0.5, 0.0]},
                              {"cell": "tet", "nodes": [[1, 1, 1], [3, 1, 1], [1, 4, 1], [1, 1, 5]],
                               "xi": [0.2, 0.2, 0.2], "det": 24.0, "jac": [2, 0, 0, 0, 3, 0, 0, 0, 4]}],
                          "tol": 1e-12},
        "tri3_conductivity": {"src": "crates/tpt-fem-thermal/examples/tri3_conductivity.rs:26",
                              "nodes": [[0, 0], [1, 0], [0, 1]],
                              "k": [[1, -0.5, -0.5], [-0.5, 0.5, 0], [-0.5, 0, 0.5]], "tol": 1e-12},
        "poisson_2d_linear": {"src": "crates/tpt-fem-thermal/src/lib.rs:347-371",
                              "coords": [[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0.5]],
                              "conn": [[0, 1, 4], [1, 3, 4], [3, 2, 4], [2, 0, 4]],
                              "bcs": [[0, 0.0], [1, 1.0], [2, 1.0], [3, 2.0]], "u_mid": 1.0, "tol": 1e-10},
        "bar_axial": {"src": "crates/tpt-fem-elasticity/src/lib.rs:716-739",
                      "coords": [[0.0], [0.5], [1.0]], "conn": [[0, 1], [1, 2]],
                      "bcs": [[0, 0.0], [2, 1.0]], "u": [0.0, 0.5, 1.0], "order": 1, "tol": 1e-12},
        "plane_stress_patch": {"src": "crates/tpt-fem-elasticity/src/lib.rs:742-784",
                               "coords": [[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0.5]],
                               "conn": [[0, 1, 4], [1, 3, 4], [3, 2, 4], [2, 0, 4]],
                               "bcs": [[0, 0.0], [1, 0.0], [2, 0.1], [3, 0.0], [4, 0.0], [5, 0.2],
                                       [6, 0.1], [7, 0.2]],
                               "u_mid": [0.05, 0.1], "tol": 1e-9},
        "box_mesh_2": {"src": "crates/tpt-fem-mesh-gen/src/lib.rs:830-835", "nodes": 27, "tets": 48},
        "poisson_1d": {"src": "crates/tpt-fem-thermal/src/lib.rs:247-279", "nx": 16,
                       "u_mid": 0.125, "tol": 2e-3},
    }
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(kats, f, indent=1)

    u = c1_solution(8)
    with open(os.path.join(HERE, "c1_fd_solution.json"), "w") as f:
        json.dump({"n": 8, "interior_u": u.ravel().tolist(), "u_max": float(u.max()),
                   "note": "7-point FD = P1 Kuhn-tet FE solution of -lap u = 1, u = 0 on the boundary"}, f)

    counts = {str(n): {"full": 111 * n ** 3 + 171 * n ** 2 + 45 * n - 3,
                       "reduced_x0_clamped": 111 * n ** 3 + 60 * n ** 2 - 57 * n - 14} for n in (2, 4, 8, 16, 64)}
    with open(os.path.join(HERE, "kuhn_counts.json"), "w") as f:
        json.dump(counts, f, indent=1)

    with open(os.path.join(HERE, "stale_vtk.json"), "w") as f:
        json.dump({"src": "thermal_solve.vtk (crates/tpt-fem/examples/thermal_solve.rs:20-45)", "n": 6,
                   "points": 343, "cells": 1296, "temperature_max": 0.038521449390516764}, f, indent=1)


if __name__ == "__main__":
    main()
