"""ctypes front-end of the CPU oracle (oracle/tfm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
(tpt-fem_amd/) never imports this module.

Each wrapper names the reference function it restates; the C source carries
the file:line citations.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

LINE, TRI, QUAD, TET, HEX = 0, 1, 2, 3, 4
BAR, PSTRESS, PSTRAIN, CONT3D = 0, 1, 2, 3
NODES = {LINE: 2, TRI: 3, QUAD: 4, TET: 4, HEX: 8}
DIMS = {LINE: 1, TRI: 2, QUAD: 2, TET: 3, HEX: 3}

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, strict IEEE)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
I64 = C.c_int64
D = C.c_double
INT = C.c_int


class OrcOp(C.Structure):
    _fields_ = [
        ("cell", INT), ("pdim", INT), ("kind", INT), ("model", INT), ("order", INT),
        ("coef", D), ("coef_e", P), ("nu", D), ("rho", P), ("penal", D), ("e0", D),
        ("emin", D), ("ke0", P),
    ]


class OrcCsr(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("rp", C.c_void_p), ("ci", C.c_void_p), ("v", C.c_void_p)]


def _declare(L):
    sig = {
        "orc_gauss_legendre": (INT, [INT, P, P]),
        "orc_cell_quad": (INT, [INT, INT, P, P]),
        "orc_shape": (None, [INT, P, P]),
        "orc_grad": (None, [INT, P, P]),
        "orc_mat_det": (D, [INT, P]),
        "orc_mat_inv": (None, [INT, P, P]),
        "orc_map": (D, [INT, INT, P, INT, P, P, P]),
        "orc_poisson_element_matrix": (None, [INT, INT, P, D, INT, P]),
        "orc_poisson_source_vector": (None, [INT, INT, P, D, INT, P]),
        "orc_elasticity_element_matrix": (INT, [INT, INT, P, INT, D, D, INT, P]),
        "orc_simp_ke_unit": (None, [D, D, P]),
        "orc_box_mesh": (None, [P, P, I64, I64, I64, P, P]),
        "orc_grid": (None, [I64, I64, D, P, P]),
        "orc_nodes_on_plane": (I64, [I64, INT, P, INT, D, D, P]),
        "orc_assemble": (I64, [I64, P, I64, P, C.POINTER(OrcOp), INT, P, P, P]),
        "orc_coo_dims": (I64, [I64, P, P, P, P]),
        "orc_to_csr": (I64, [I64, P, P, P, I64, P, P, P]),
        "orc_csr_matvec": (None, [I64, P, P, P, I64, P, P]),
        "orc_reduce_system": (I64, [I64, P, P, P, I64, P, I64, P, P, P, P, P, P, P, P]),
        "orc_dense_solve": (INT, [I64, P, P]),
        "orc_solve_coo": (INT, [I64, P, P, P, I64, P]),
        "orc_solve_with_dirichlet": (INT, [I64, P, P, P, I64, P, I64, P, P, P]),
        "orc_mass_element_matrix": (None, [INT, INT, P, D, INT, P]),
        "orc_lumped_mass": (I64, [I64, P, P, INT, P]),
        "orc_eig_jacobi": (None, [INT, P, INT, P, P]),
        "orc_lanczos_eigs": (INT, [I64, P, P, P, INT, INT, INT, P, P]),
        "orc_central_difference": (INT, [I64, OrcCsr, OrcCsr, OrcCsr, P, P, P, P, D, I64, P, P]),
        "orc_newmark": (INT, [I64, OrcCsr, OrcCsr, OrcCsr, P, P, P, P, D, D, D, I64, P]),
        "orc_apply_robin": (I64, [I64, INT, P, I64, INT, P, INT, D, INT, D, D, INT, P, P, P]),
        "orc_topopt_quad_form": (D, [P, P, P]),
        "orc_topopt_filter": (None, [I64, I64, D, P, D, P]),
        "orc_topopt_oc": (None, [I64, P, P, D, D, P]),
        "orc_topopt_simp": (INT, [I64, I64, D, D, D, D, D, D, INT, D, P, I64, P, P, P, P, C.POINTER(INT)]),
        "orc_boundary_faces": (I64, [I64, INT, P, P, P]),
        "orc_apply_neumann": (None, [I64, INT, P, I64, INT, P, INT, D, INT, P]),
        "orc_plane_traction": (None, [I64, INT, P, I64, INT, P, INT, INT, D, D, P, INT, P]),
        "orc_poisson_rhs": (None, [I64, INT, P, I64, INT, P, D, INT, P]),
        "orc_qp_coords": (INT, [INT, INT, P, INT, P]),
        "orc_boundary_points": (I64, [I64, INT, P, I64, INT, P, INT, P, P]),
        "orc_apply_neumann_qp": (None, [I64, INT, P, I64, INT, P, INT, P, INT, P]),
        "orc_apply_robin_qp": (I64, [I64, INT, P, I64, INT, P, INT, P, INT, P, P, P]),
        "orc_poisson_source_vector_qp": (None, [INT, INT, P, P, INT, P]),
        "orc_elasticity_body_vector": (None, [INT, INT, P, P, INT, P]),
        "orc_body_rhs": (None, [I64, INT, P, I64, INT, P, P, INT, P]),
        "orc_source_rhs_qp": (None, [I64, INT, P, I64, INT, P, P, INT, P]),
        "orc_solve_elasticity": (INT, [I64, INT, P, I64, INT, P, INT, D, D, INT, P, I64, P, P, P]),
        "orc_solve_poisson": (INT, [I64, INT, P, I64, INT, P, D, INT, D, I64, P, P, P]),
        "orc_jitter_interior": (None, [I64, I64, I64, D, C.c_uint64, P]),
        "orc_loguniform": (None, [I64, D, D, C.c_uint64, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# ---- quadrature / elements ------------------------------------------------
def gauss_legendre(order):
    p, w = np.zeros(5), np.zeros(5)
    n = lib().orc_gauss_legendre(order, _p(p), _p(w))
    return p[:n], w[:n]


def cell_quad(cell, order):
    p, w = np.zeros(27 * 3), np.zeros(27)
    n = lib().orc_cell_quad(cell, order, _p(p), _p(w))
    return p[: 3 * n].reshape(n, 3), w[:n]


def shape(cell, xi):
    x = np.zeros(3)
    x[: len(xi)] = xi
    n = np.zeros(NODES[cell])
    lib().orc_shape(cell, _p(x), _p(n))
    return n


def grad(cell, xi):
    x = np.zeros(3)
    x[: len(xi)] = xi
    g = np.zeros(NODES[cell] * DIMS[cell])
    lib().orc_grad(cell, _p(x), _p(g))
    return g.reshape(NODES[cell], DIMS[cell])


def mat_det(m):
    m = f64(m)
    n = int(round(np.sqrt(m.size)))
    return lib().orc_mat_det(n, _p(m.ravel()))


def mat_inv(m):
    m = f64(m)
    n = int(round(np.sqrt(m.size)))
    o = np.zeros(n * n)
    lib().orc_mat_inv(n, _p(m.ravel()), _p(o))
    return o.reshape(n, n)


def map_jacobian(phys, local_grad):
    """Map::from_nodes_and_grad -> (J, signed det, J^-1)."""
    phys, lg = f64(phys), f64(local_grad)
    nn, pdim = phys.shape
    n = lg.shape[1]
    jac, inv = np.zeros(n * n), np.zeros(n * n)
    det = lib().orc_map(nn, pdim, _p(phys), n, _p(lg), _p(jac), _p(inv))
    return jac.reshape(n, n), det, inv.reshape(n, n)


def poisson_element_matrix(cell, phys, kappa=1.0, order=2):
    phys = f64(phys)
    nn = NODES[cell]
    k = np.zeros(nn * nn)
    lib().orc_poisson_element_matrix(cell, phys.shape[1], _p(phys), kappa, order, _p(k))
    return k.reshape(nn, nn)


def poisson_source_vector(cell, phys, f_const=1.0, order=2):
    phys = f64(phys)
    f = np.zeros(NODES[cell])
    lib().orc_poisson_source_vector(cell, phys.shape[1], _p(phys), f_const, order, _p(f))
    return f


def qp_coords(cell, phys, order=2):
    """Quadrature-point positions a source/body closure is called at (nq x pdim)."""
    phys = f64(phys)
    x = np.zeros((125, phys.shape[1]))
    nq = lib().orc_qp_coords(cell, phys.shape[1], _p(phys), order, _p(x))
    return x[:nq].copy()


def mesh_qp_coords(coords, cell, conn, order=2):
    """qp_coords of every element: (nelem, nq, pdim)."""
    coords, conn = f64(coords), i64(conn)
    return np.array([qp_coords(cell, coords[c], order) for c in conn]).reshape(len(conn), -1, coords.shape[1])


def poisson_source_vector_qp(cell, phys, sq, order=2):
    """poisson_source_vector with the source closure's values at the quadrature points."""
    phys, sq = f64(phys), f64(sq)
    f = np.zeros(NODES[cell])
    lib().orc_poisson_source_vector_qp(cell, phys.shape[1], _p(phys), _p(sq), order, _p(f))
    return f


def elasticity_body_vector(cell, phys, bq, order=2):
    """elasticity_body_vector with the body-force closure's values bq (nq x dim)."""
    phys, bq = f64(phys), f64(bq)
    f = np.zeros(NODES[cell] * DIMS[cell])
    lib().orc_elasticity_body_vector(cell, phys.shape[1], _p(phys), _p(bq), order, _p(f))
    return f


def body_rhs(coords, cell, conn, bq, order=2):
    """solve_elasticity's rhs: body vectors scattered in element order; bq (nelem, nq, dim)."""
    coords, conn, bq = f64(coords), i64(conn), f64(bq)
    rhs = np.zeros(coords.shape[0] * DIMS[cell])
    lib().orc_body_rhs(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell, _p(conn), _p(bq),
                       order, _p(rhs))
    return rhs


def source_rhs_qp(coords, cell, conn, sq, order=2):
    """solve_poisson's rhs with per-point source values sq (nelem, nq)."""
    coords, conn, sq = f64(coords), i64(conn), f64(sq)
    rhs = np.zeros(coords.shape[0])
    lib().orc_source_rhs_qp(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell, _p(conn), _p(sq),
                            order, _p(rhs))
    return rhs


def elasticity_element_matrix(cell, phys, model, young, nu, order=2):
    phys = f64(phys)
    nd = NODES[cell] * DIMS[cell]
    k = np.zeros(nd * nd)
    rc = lib().orc_elasticity_element_matrix(cell, phys.shape[1], _p(phys), model, young, nu,
                                             order, _p(k))
    if rc != 0:
        raise ValueError("model/dimension mismatch")
    return k.reshape(nd, nd)


def mass_element_matrix(cell, phys, density=1.0, order=2):
    """elasticity_mass_matrix's element block (crates/tpt-fem-elasticity/src/lib.rs:392-424)."""
    phys = f64(phys)
    nd = NODES[cell] * DIMS[cell]
    k = np.zeros(nd * nd)
    lib().orc_mass_element_matrix(cell, phys.shape[1], _p(phys), density, order, _p(k))
    return k.reshape(nd, nd)


def mass_matrix(coords, cell, conn, density=1.0, order=2):
    """elasticity_mass_matrix (crates/tpt-fem-elasticity/src/lib.rs:386-428) -> Coo."""
    return assemble(coords, conn, make_op(cell, f64(coords).shape[1], 3, order=order, coef=density), DIMS[cell])


def lumped_mass(coords, cell, conn, density=1.0, order=2):
    """elasticity_lumped_mass (crates/tpt-fem-elasticity/src/lib.rs:432-457) -> Coo (diagonal)."""
    csr = mass_matrix(coords, cell, conn, density, order).to_csr()
    d = np.zeros(max(csr.nrows, 1))
    m = lib().orc_lumped_mass(csr.nrows, _p(csr.row_ptrs), _p(csr.values), DIMS[cell], _p(d))
    idx = np.arange(m, dtype=np.int64)
    return Coo(idx, idx, d[:m].copy())


def simp_ke_unit(nu, h=1.0):
    ke = np.zeros(64)
    lib().orc_simp_ke_unit(nu, h, _p(ke))
    return ke.reshape(8, 8)


# ---- meshes ---------------------------------------------------------------
def box_mesh(mn, mx, n):
    nx, ny, nz = n
    coords = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    conn = np.zeros((nx * ny * nz * 6, 4), dtype=np.int64)
    lo, hi = f64(mn), f64(mx)  # kept alive across the call
    lib().orc_box_mesh(_p(lo), _p(hi), nx, ny, nz, _p(coords), _p(conn))
    return coords, conn


def grid(nx, ny, h=1.0):
    coords = np.zeros((nx * ny, 2))
    conn = np.zeros(((nx - 1) * (ny - 1), 4), dtype=np.int64)
    lib().orc_grid(nx, ny, h, _p(coords), _p(conn))
    return coords, conn


def nodes_on_plane(coords, axis, coord, tol):
    coords = f64(coords)
    out = np.zeros(coords.shape[0], dtype=np.int64)
    m = lib().orc_nodes_on_plane(coords.shape[0], coords.shape[1], _p(coords), axis, coord, tol,
                                 _p(out))
    return out[:m]


def jitter_interior(coords, n, h_frac, seed):
    coords = f64(coords).copy()
    lib().orc_jitter_interior(n[0], n[1], n[2], h_frac, seed, _p(coords))
    return coords


# ---- assembly / sparse ----------------------------------------------------
class Coo:
    """Mirror of tpt_fem_sparse::Coo (crates/tpt-fem-sparse/src/lib.rs:39-47)."""

    def __init__(self, rows=(), cols=(), vals=()):
        self.rows = i64(rows)
        self.cols = i64(cols)
        self.vals = f64(vals)

    def to_csr(self):
        return to_csr(self)


class Csr:
    def __init__(self, nrows, ncols, row_ptrs, col_ind, values):
        self.nrows, self.ncols = nrows, ncols
        self.row_ptrs, self.col_ind, self.values = row_ptrs, col_ind, values

    def nnz(self):
        return len(self.values)

    def matvec(self, x):
        x = f64(x)
        y = np.zeros(max(self.nrows, len(x)))
        lib().orc_csr_matvec(self.nrows, _p(self.row_ptrs), _p(self.col_ind), _p(self.values),
                             len(x), _p(x), _p(y))
        return y


def make_op(cell, pdim, kind, order=2, coef=1.0, coef_e=None, model=CONT3D, nu=0.3,
            rho=None, penal=3.0, e0=1.0, emin=1e-9, ke0=None):
    keep = []
    op = OrcOp()
    op.cell, op.pdim, op.kind, op.model, op.order = cell, pdim, kind, model, order
    op.coef, op.nu, op.penal, op.e0, op.emin = coef, nu, penal, e0, emin
    for name, arr in (("coef_e", coef_e), ("rho", rho), ("ke0", ke0)):
        if arr is not None:
            a = f64(arr).ravel()
            keep.append(a)
            setattr(op, name, _p(a))
    op._keep = keep
    return op


def assemble(coords, conn, op, dpn):
    """assemble / try_assemble (crates/tpt-fem-assembly/src/lib.rs:428-480)."""
    coords, conn = f64(coords), i64(conn)
    nn = conn.shape[1]
    nd = 8 if op.kind == 2 else nn * dpn
    cap = max(1, conn.shape[0] * nd * nd)
    r, c, v = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(cap)
    t = lib().orc_assemble(coords.shape[0], _p(coords), conn.shape[0], _p(conn), C.byref(op),
                           dpn, _p(r), _p(c), _p(v))
    if t < 0:
        raise ValueError("element operator failed")
    return Coo(r[:t], c[:t], v[:t])


def to_csr(coo):
    n = len(coo.rows)
    nr, nc = C.c_int64(), C.c_int64()
    lib().orc_coo_dims(n, _p(coo.rows), _p(coo.cols), C.byref(nr), C.byref(nc))
    nrows = nr.value
    rp = np.zeros(nrows + 1, np.int64)
    ci = np.zeros(max(n, 1), np.int64)
    v = np.zeros(max(n, 1))
    nnz = lib().orc_to_csr(n, _p(coo.rows), _p(coo.cols), _p(coo.vals), nrows, _p(rp), _p(ci),
                           _p(v))
    return Csr(nrows, nc.value, rp, ci[:nnz].copy(), v[:nnz].copy())


def reduce_system(coo, rhs, bcs):
    """reduce_system (crates/tpt-fem-assembly/src/lib.rs:504-535) -> (kred Coo, fred, free)."""
    csr = to_csr(coo)
    rhs = f64(rhs)
    n = len(rhs)
    bd = i64([b[0] for b in bcs]) if bcs else np.zeros(1, np.int64)
    bv = f64([b[1] for b in bcs]) if bcs else np.zeros(1)
    nnz = max(1, csr.nnz())
    kr, kc, kv = np.zeros(nnz, np.int64), np.zeros(nnz, np.int64), np.zeros(nnz)
    fred, free = np.zeros(max(n, 1)), np.zeros(max(n, 1), np.int64)
    nf = C.c_int64()
    t = lib().orc_reduce_system(csr.nrows, _p(csr.row_ptrs), _p(csr.col_ind), _p(csr.values), n,
                                _p(rhs), len(bcs), _p(bd), _p(bv), _p(kr), _p(kc), _p(kv),
                                _p(fred), _p(free), C.byref(nf))
    if t < 0:
        raise IndexError("reduce_system: DOF outside the assembled matrix")
    m = nf.value
    return Coo(kr[:t], kc[:t], kv[:t]), fred[:m].copy(), free[:m].copy()


class SolveError(RuntimeError):
    pass


_SOLVE_MSG = {
    -1: "sparse numeric factorization/solve failed: singular matrix",
    -2: "sparse numeric factorization/solve failed: solve requires a square matrix",
    -3: "sparse numeric factorization/solve failed: rhs length does not match matrix dimension",
    -4: "reduce_system: DOF outside the assembled matrix",
}


def solve(coo, rhs):
    """solve (crates/tpt-fem-sparse/src/lib.rs:225-231): dense partial-pivot LU."""
    b = f64(rhs).copy()
    rc = lib().orc_solve_coo(len(coo.rows), _p(coo.rows), _p(coo.cols), _p(coo.vals), len(b),
                             _p(b))
    if rc != 0:
        raise SolveError(_SOLVE_MSG[rc])
    return b


def solve_with_dirichlet(coo, rhs, bcs):
    rhs = f64(rhs)
    n = len(rhs)
    bd = i64([b[0] for b in bcs]) if bcs else np.zeros(1, np.int64)
    bv = f64([b[1] for b in bcs]) if bcs else np.zeros(1)
    u = np.zeros(max(n, 1))
    rc = lib().orc_solve_with_dirichlet(len(coo.rows), _p(coo.rows), _p(coo.cols), _p(coo.vals),
                                        n, _p(rhs), len(bcs), _p(bd), _p(bv), _p(u))
    if rc != 0:
        raise SolveError(_SOLVE_MSG[rc])
    return u[:n]


def boundary_faces(cell, conn):
    conn = i64(conn)
    nf = {LINE: 2, TRI: 3, QUAD: 4, TET: 4, HEX: 6}[cell]
    cap = max(1, conn.shape[0] * nf)
    e, f = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
    m = lib().orc_boundary_faces(conn.shape[0], cell, _p(conn), _p(e), _p(f))
    return list(zip(e[:m].tolist(), f[:m].tolist()))


def apply_neumann(coords, cell, conn, dpn, g, rhs, order=2):
    coords, conn = f64(coords), i64(conn)
    rhs = f64(rhs).copy()
    lib().orc_apply_neumann(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell,
                            _p(conn), dpn, g, order, _p(rhs))
    return rhs


def plane_traction(coords, cell, conn, dpn, axis, coord, tol, t, rhs=None, order=2):
    coords, conn = f64(coords), i64(conn)
    rhs = np.zeros(coords.shape[0] * dpn) if rhs is None else f64(rhs).copy()
    tt = f64(t)
    lib().orc_plane_traction(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell,
                             _p(conn), dpn, axis, coord, tol, _p(tt), order, _p(rhs))
    return rhs


def poisson_rhs(coords, cell, conn, f_const=1.0, order=2):
    coords, conn = f64(coords), i64(conn)
    rhs = np.zeros(coords.shape[0])
    lib().orc_poisson_rhs(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell,
                          _p(conn), f_const, order, _p(rhs))
    return rhs


def solve_poisson(coords, cell, conn, kappa, order, f_const, bcs):
    coords, conn = f64(coords), i64(conn)
    n = coords.shape[0]
    bd = i64([b[0] for b in bcs]) if bcs else np.zeros(1, np.int64)
    bv = f64([b[1] for b in bcs]) if bcs else np.zeros(1)
    u = np.zeros(n)
    rc = lib().orc_solve_poisson(n, coords.shape[1], _p(coords), conn.shape[0], cell, _p(conn),
                                 kappa, order, f_const, len(bcs), _p(bd), _p(bv), _p(u))
    if rc != 0:
        raise SolveError(_SOLVE_MSG[rc])
    return u


def solve_elasticity(coords, cell, conn, model, young, nu, order, bq, bcs):
    """solve_elasticity (crates/tpt-fem-elasticity/src/lib.rs:355-381); bq (nelem, nq, dim)
    holds the body-force closure's values at every element quadrature point."""
    coords, conn, bq = f64(coords), i64(conn), f64(bq)
    n = coords.shape[0] * DIMS[cell]
    bd = i64([b[0] for b in bcs]) if bcs else np.zeros(1, np.int64)
    bv = f64([b[1] for b in bcs]) if bcs else np.zeros(1)
    u = np.zeros(n)
    rc = lib().orc_solve_elasticity(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell, _p(conn),
                                    model, young, nu, order, _p(bq), len(bcs), _p(bd), _p(bv), _p(u))
    if rc == -5:
        raise ValueError("model/dimension mismatch")
    if rc != 0:
        raise SolveError(_SOLVE_MSG[rc])
    return u


# ---- SIMP topology optimization (crates/tpt-fem-topopt/src/lib.rs) -------------
def topopt_quad_form(ke0, u, nodes):
    """element_quad_form (:241-254)."""
    ke0, u, nodes = f64(ke0).ravel(), f64(u), i64(nodes)
    return lib().orc_topopt_quad_form(_p(ke0), _p(u), _p(nodes))


def topopt_filter(nx, ny, h, raw, radius):
    """filter_sensitivity (:259-279) on Grid::new(nx, ny, h)."""
    raw = f64(raw)
    out = np.zeros(len(raw))
    lib().orc_topopt_filter(nx, ny, h, _p(raw), radius, _p(out))
    return out


def topopt_oc(rho, sens, vol_frac, move_limit):
    """oc_update (:284-331)."""
    rho, sens = f64(rho), f64(sens)
    out = np.zeros(len(rho))
    lib().orc_topopt_oc(len(rho), _p(rho), _p(sens), vol_frac, move_limit, _p(out))
    return out


def cantilever_load(nx, ny, load):
    """cantilever_load (:398-416): left edge clamped, -load in y at the
    bottom-right node."""
    f = np.zeros(2 * nx * ny)
    bcs = []
    for j in range(ny):
        n = j * nx
        bcs += [(2 * n, 0.0), (2 * n + 1, 0.0)]
    f[2 * (nx - 1) + 1] = -load
    return f, bcs


def topopt_simp(nx, ny, h, e0, nu, vol_frac, penal, radius, max_iter, move_limit, f, bcs):
    """topopt_simp (:342-385) -> (densities, compliance history, iterations)."""
    f = f64(f)
    bd = i64([b[0] for b in bcs]) if bcs else np.zeros(1, np.int64)
    bv = f64([b[1] for b in bcs]) if bcs else np.zeros(1)
    rho = np.zeros((nx - 1) * (ny - 1))
    comp = np.zeros(max_iter + 1)
    it = INT(0)
    rc = lib().orc_topopt_simp(nx, ny, h, e0, nu, vol_frac, penal, radius, max_iter, move_limit, _p(f),
                               len(bcs), _p(bd), _p(bv), _p(rho), _p(comp), C.byref(it))
    if rc != 0:
        raise SolveError(_SOLVE_MSG[rc])
    return rho, comp, it.value


# ---- Lanczos eigenpairs (crates/tpt-fem-eigen/src/lib.rs) ---------------------
SMALLEST, LARGEST = 0, 1


def eig_jacobi(a, max_sweeps=200):
    """jacobi (:170-218) -> (eigenvalues, eigenvectors as columns)."""
    a = f64(a).copy()
    m = a.shape[0]
    eig, q = np.zeros(m), np.zeros((m, m))
    lib().orc_eig_jacobi(m, _p(a), max_sweeps, _p(eig), _p(q))
    return eig, q


def lanczos_eigs(csr, num, which, lanczos_dim):
    """lanczos_eigs(coo, num, which, lanczos_dim) (:223-268) on coo.to_csr() -> [(lambda, vec)]."""
    n = csr.nrows
    ev, vec = np.zeros(max(num, 1)), np.zeros((max(num, 1), n))
    k = lib().orc_lanczos_eigs(n, _p(csr.row_ptrs), _p(csr.col_ind), _p(csr.values), num, which, lanczos_dim,
                               _p(ev), _p(vec))
    if k < 0:
        raise ValueError("lanczos_dim must be >= 1")
    return [(ev[i], vec[i].copy()) for i in range(k)]


# ---- time integration (crates/tpt-fem-dynamic/src/lib.rs) ----------------------
class CflViolation(RuntimeError):
    """DynamicError::CflViolation (dynamic/src/lib.rs:46, 58-61)."""

    def __init__(self, dt, critical):
        super().__init__(f"central-difference timestep {dt:e} exceeds the critical (CFL) limit {critical:e}")
        self.dt, self.critical = dt, critical


def _orc_csr(csr, keep):
    rp, ci, v = i64(csr.row_ptrs), i64(csr.col_ind), f64(csr.values)
    keep += [rp, ci, v]
    return OrcCsr(csr.nrows, rp.ctypes.data, ci.ctypes.data, v.ctypes.data)


def central_difference(m, c, k, u0, v0, F, dt, nsteps, fscale=None):
    """central_difference (:289-377) on CSR operators with f(t_k) = fscale[k] * F -> history [nsteps+1, n]."""
    keep = []
    u0, v0, F = f64(u0), f64(v0), f64(F)
    fs = None if fscale is None else f64(fscale)
    n = len(u0)
    hist = np.zeros((nsteps + 1, n))
    crit = D(0.0)
    rc = lib().orc_central_difference(n, _orc_csr(m, keep), _orc_csr(c, keep), _orc_csr(k, keep), _p(u0), _p(v0),
                                      _p(F), _p(fs), dt, nsteps, _p(hist), C.byref(crit))
    if rc == 1:
        raise CflViolation(dt, crit.value)
    return hist


def newmark(m, c, k, u0, v0, F, dt, beta, gamma, nsteps, fscale=None):
    """newmark (:194-263) with dense LU solves -> history [nsteps+1, n]."""
    keep = []
    u0, v0, F = f64(u0), f64(v0), f64(F)
    fs = None if fscale is None else f64(fscale)
    n = len(u0)
    hist = np.zeros((nsteps + 1, n))
    rc = lib().orc_newmark(n, _orc_csr(m, keep), _orc_csr(c, keep), _orc_csr(k, keep), _p(u0), _p(v0), _p(F),
                           _p(fs), dt, beta, gamma, nsteps, _p(hist))
    if rc != 0:
        raise SolveError("sparse numeric factorization/solve failed: singular matrix")
    return hist


def apply_robin(coords, cell, conn, dpn, h, axis=-1, coord=0.0, tol=0.0, order=2):
    """apply_robin_order (crates/tpt-fem-assembly/src/lib.rs:684-753) -> the appended triplets as a Coo."""
    coords, conn = f64(coords), i64(conn)
    fn = {LINE: 1, TRI: 2, QUAD: 2, TET: 3, HEX: 4}[cell]
    cap = max(1, conn.shape[0] * 6 * 9 * fn * fn * dpn)
    r, c, v = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(cap)
    t = lib().orc_apply_robin(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell, _p(conn), dpn, h,
                              axis, coord, tol, order, _p(r), _p(c), _p(v))
    return Coo(r[:t], c[:t], v[:t])


def boundary_points(coords, cell, conn, order=2):
    """Every boundary-face quadrature point's position and outward unit normal --
    the (x, n) apply_neumann_order / apply_robin_order pass to their closure --
    in boundary_faces order, points in rule order: two (count, pdim) arrays."""
    coords, conn = f64(coords), i64(conn)
    pdim = coords.shape[1]
    cap = max(1, conn.shape[0] * 6 * 25)
    x, nrm = np.zeros((cap, pdim)), np.zeros((cap, pdim))
    m = lib().orc_boundary_points(coords.shape[0], pdim, _p(coords), conn.shape[0], cell, _p(conn), order, _p(x),
                                  _p(nrm))
    return x[:m].copy(), nrm[:m].copy()


def apply_neumann_qp(coords, cell, conn, dpn, gq, rhs, order=2):
    """apply_neumann_order with the flux closure's values at boundary_points."""
    coords, conn, gq = f64(coords), i64(conn), f64(gq)
    rhs = f64(rhs).copy()
    lib().orc_apply_neumann_qp(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell, _p(conn), dpn,
                               _p(gq), order, _p(rhs))
    return rhs


def apply_robin_qp(coords, cell, conn, dpn, hq, order=2):
    """apply_robin_order with the coefficient closure's values at boundary_points -> Coo."""
    coords, conn, hq = f64(coords), i64(conn), f64(hq)
    fn = {LINE: 1, TRI: 2, QUAD: 2, TET: 3, HEX: 4}[cell]
    cap = max(1, conn.shape[0] * 6 * 25 * fn * fn * dpn)
    r, c, v = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(cap)
    t = lib().orc_apply_robin_qp(coords.shape[0], coords.shape[1], _p(coords), conn.shape[0], cell, _p(conn), dpn,
                                 _p(hq), order, _p(r), _p(c), _p(v))
    return Coo(r[:t], c[:t], v[:t])


def assemble_robin(coords, cell, conn, op, dpn, h, axis=-1, coord=0.0, tol=0.0):
    """assemble(...) then apply_robin(...) on the same Coo (thermal/src/lib.rs:219-237 composition)."""
    a = assemble(coords, conn, op, dpn)
    b = apply_robin(coords, cell, conn, dpn, h, axis, coord, tol)
    return Coo(np.concatenate([a.rows, b.rows]), np.concatenate([a.cols, b.cols]), np.concatenate([a.vals, b.vals]))


# ---- legacy binary VTK (crates/tpt-fem-io-vtk/src/lib.rs:117-170) --------------
VTK_TYPE = {LINE: 3, TRI: 5, QUAD: 9, TET: 10, HEX: 12}  # vtk_cell_type (:101-114), P1 cells


def vtk_legacy_bytes(coords, cell, conn, fields=(), title="tpt-fem mesh"):
    """write_vtk_with_data's file (mesh_to_vtk :117-170 exported by vtkio as
    legacy BINARY, big endian): points padded to 3 components as f64, cells as
    u32 (count then node ids), types as i32, one SCALARS block per field."""
    coords, conn = f64(coords), i64(conn)
    n, ne = coords.shape[0], conn.shape[0]
    pts = np.zeros((n, 3))
    pts[:, :coords.shape[1]] = coords
    nn = conn.shape[1]
    verts = np.concatenate([np.full((ne, 1), nn), conn], axis=1).astype('>u4').ravel()
    out = [f"# vtk DataFile Version 4.1\n{title}\nBINARY\n\nDATASET UNSTRUCTURED_GRID\nPOINTS {n} double\n".encode(),
           pts.astype('>f8').tobytes(),
           f"\n\nCELLS {ne} {ne * (nn + 1)}\n".encode(), verts.tobytes(),
           f"\n\nCELL_TYPES {ne}\n".encode(), np.full(ne, VTK_TYPE[cell], '>i4').tobytes(),
           f"\n\nPOINT_DATA {n}\n".encode()]
    for name, vals in fields:
        out += [f"\nSCALARS {name} double 1\nLOOKUP_TABLE default\n".encode(), f64(vals).astype('>f8').tobytes(),
                b"\n"]
    out.append(f"\nCELL_DATA {ne}\n\n".encode())
    return b"".join(out)
