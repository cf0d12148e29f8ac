"""Python binding of libtptfem.so (the MI355X assembly/solve library).

Mirrors the reference crates' API on this path -- tpt_fem_sparse (Coo, Csr,
solve), tpt_fem_assembly (assemble, reduce_system, solve_with_dirichlet,
apply_neumann), tpt_fem_thermal (solve_poisson), tpt_fem_elasticity
(solve_elasticity), tpt_fem_mesh_gen (box_mesh) -- over the C ABI in
include/tptfem.h.  There is no CPU fallback: if the HIP library is not built,
importing this module raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtptfem.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C tpt-fem_amd)")

_lib = C.CDLL(LIB_PATH)

OK, E_INVALID, E_NUMERIC, E_NOT_CONVERGED, E_BREAKDOWN, E_HIP, E_UNSUPPORTED, E_MODEL, E_CFL = (
    0, -1, -2, -3, -4, -5, -6, -7, -8)
LINE, TRI, QUAD, TET, HEX = 0, 1, 2, 3, 4
BAR, PLANE_STRESS, PLANE_STRAIN, CONTINUUM3D = 0, 1, 2, 3
OP_POISSON, OP_ELASTICITY, OP_SIMP, OP_MASS = 0, 1, 2, 3
SYS_FP64, SYS_FP32 = 1, 2
SOLVE_PCG64, SOLVE_MIXED = 0, 1
# FP64 CG with a geometric-multigrid V-cycle: one system on a Mesh.box mesh
# (one rhs) or SIMP systems on a Mesh.grid_batch mesh (one problem per design)
SOLVE_PCG64_MG = 2
# the reference's dense partial-pivot LU on the GPU (at most 8000 free DOFs)
SOLVE_DIRECT = 3
# the precision form on the assembled FP32 values (SOLVE_MIXED is matrix-free
# on box meshes of P1 tets)
SOLVE_MIXED_CSR = 4
NODES = {LINE: 2, TRI: 3, QUAD: 4, TET: 4, HEX: 8}
DIMS = {LINE: 1, TRI: 2, QUAD: 2, TET: 3, HEX: 3}

P, I64, D, INT = C.c_void_p, C.c_int64, C.c_double, C.c_int
PP = C.POINTER(C.c_void_p)


class Operator(C.Structure):
    """tfm_operator: the element operator an assemble() closure would apply."""

    _fields_ = [("kind", INT), ("model", INT), ("quad_order", INT), ("coef", D), ("coef_e", P),
                ("nu", D), ("rho", P), ("penal", D), ("e0", D), ("emin", D), ("h", D)]

    @staticmethod
    def poisson(kappa=1.0, quad_order=2, kappa_e=None):
        op = Operator(kind=OP_POISSON, quad_order=quad_order, coef=kappa)
        op._attach("coef_e", kappa_e)
        return op

    @staticmethod
    def elasticity(model, young, nu, quad_order=2, young_e=None):
        op = Operator(kind=OP_ELASTICITY, model=model, quad_order=quad_order, coef=young, nu=nu)
        op._attach("coef_e", young_e)
        return op

    @staticmethod
    def mass(density, quad_order=2, density_e=None):
        """elasticity_mass_matrix's operator (crates/tpt-fem-elasticity/src/lib.rs:386-428):
        assemble with dofs_per_node equal to the cell's dimension."""
        op = Operator(kind=OP_MASS, quad_order=quad_order, coef=density)
        op._attach("coef_e", density_e)
        return op

    @staticmethod
    def simp(rho, nu=0.3, penal=3.0, e0=1.0, emin=1e-9, h=1.0):
        op = Operator(kind=OP_SIMP, quad_order=2, nu=nu, penal=penal, e0=e0, emin=emin, h=h)
        op._attach("rho", rho)
        return op

    def _attach(self, name, arr):
        if not hasattr(self, "_keep"):
            self._keep = {}
        if arr is None:
            setattr(self, name, None)
            return
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._keep[name] = a
        setattr(self, name, a.ctypes.data)


class SolveOpts(C.Structure):
    _fields_ = [("mode", INT), ("rtol", D), ("max_iter", INT), ("inner_rtol", D)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", INT), ("outer", INT), ("rel_residual", D), ("seconds", D)]


class TopOptParams(C.Structure):
    """TopOptParams (crates/tpt-fem-topopt/src/lib.rs:119-137), the grid as Grid::new(nx, ny, h)."""
    _fields_ = [("nx", I64), ("ny", I64), ("h", D), ("e0", D), ("nu", D), ("vol_frac", D), ("penal", D),
                ("filter_radius", D), ("max_iter", INT), ("move_limit", D)]


_SIGS = {
    "tfm_ctx_create": (INT, [INT, PP]),
    "tfm_ctx_destroy": (None, [P]),
    "tfm_last_error": (C.c_char_p, [P]),
    "tfm_ctx_synchronize": (INT, [P]),
    "tfm_ctx_timer_begin": (INT, [P]),
    "tfm_ctx_timer_end": (INT, [P, C.POINTER(D)]),
    "tfm_debug_scan_total": (INT, [P, P, I64, P]),
    "tfm_mesh_create": (INT, [P, INT, I64, INT, P, I64, P, PP]),
    "tfm_mesh_box": (INT, [P, P, P, I64, I64, I64, PP]),
    "tfm_mesh_grid": (INT, [P, I64, I64, D, PP]),
    "tfm_mesh_jitter": (INT, [P, I64, I64, I64, D, C.c_uint64]),
    "tfm_mesh_info": (INT, [P, P, P, P, P]),
    "tfm_mesh_download": (INT, [P, P, P]),
    "tfm_mesh_nodes_on_plane": (INT, [P, INT, D, D, P, P]),
    "tfm_mesh_destroy": (None, [P]),
    "tfm_assemble": (INT, [P, P, C.POINTER(Operator), INT, PP]),
    "tfm_assemble_coo": (INT, [P, P, C.POINTER(Operator), INT, C.POINTER(I64), P, P, P]),
    "tfm_lumped_mass": (INT, [P, P, D, INT, PP]),
    "tfm_matrix_create": (INT, [P, I64, I64, P, P, P, PP]),
    "tfm_matrix_info": (INT, [P, P, P, P]),
    "tfm_matrix_download": (INT, [P, P, P, P]),
    "tfm_matrix_matvec": (INT, [P, P, I64, P]),
    "tfm_matrix_destroy": (None, [P]),
    "tfm_load_source": (INT, [P, P, INT, D, P]),
    "tfm_load_body": (INT, [P, P, INT, P, P]),
    "tfm_load_neumann": (INT, [P, P, INT, D, P]),
    "tfm_mesh_quad_points": (INT, [P, INT, P, P]),
    "tfm_load_source_qp": (INT, [P, P, INT, P, P]),
    "tfm_load_body_qp": (INT, [P, P, INT, P, P]),
    "tfm_mesh_boundary_points": (INT, [P, INT, P, P, P]),
    "tfm_load_neumann_qp": (INT, [P, P, INT, INT, P, P]),
    "tfm_assemble_robin_qp": (INT, [P, P, C.POINTER(Operator), INT, INT, P, PP]),
    "tfm_solve_poisson_qp": (INT, [P, P, D, INT, P, P, P, I64, P, P, C.POINTER(SolveOpts), P]),
    "tfm_solve_elasticity_qp": (INT, [P, P, INT, D, D, INT, P, I64, P, P, C.POINTER(SolveOpts), P]),
    "tfm_load_plane_traction": (INT, [P, P, INT, INT, D, D, P, P]),
    "tfm_system_create": (INT, [P, P, C.POINTER(Operator), INT, I64, P, P, P, INT, PP]),
    "tfm_system_from_matrix": (INT, [P, P, I64, P, I64, P, P, PP]),
    "tfm_system_assemble": (INT, [P, C.POINTER(Operator), INT]),
    "tfm_system_info": (INT, [P, P, P, P]),
    "tfm_system_assembly_info": (INT, [P, P, P, P, P]),
    "tfm_debug_mf_apply": (INT, [P, P, P]),
    "tfm_debug_system_values32": (INT, [P, P]),
    "tfm_system_download": (INT, [P, P, P, P, P, P]),
    "tfm_system_solve": (INT, [P, C.POINTER(SolveOpts), P, C.POINTER(SolveStats)]),
    "tfm_system_solve_noalloc": (INT, [P, C.POINTER(SolveOpts), P, C.POINTER(SolveStats)]),
    "tfm_system_destroy": (None, [P]),
    "tfm_mesh_box_batch": (INT, [P, I64, P, PP, P, P]),
    "tfm_mesh_grid_batch": (INT, [P, I64, I64, I64, D, PP]),
    "tfm_system_solve_batch": (INT, [P, I64, P, INT, P, C.POINTER(SolveOpts), P, P]),
    "tfm_system_assemble_solve_batch": (INT, [P, P, I64, P, INT, P, C.POINTER(SolveOpts), P, P]),
    "tfm_topopt_simp": (INT, [P, C.POINTER(TopOptParams), P, I64, P, P, C.POINTER(SolveOpts), P, P, P, P]),
    "tfm_topopt_simp_batch": (INT, [P, I64, C.POINTER(TopOptParams), P, P, I64, P, P, C.POINTER(SolveOpts), P, P,
                                    P, P]),
    "tfm_lanczos_eigs": (INT, [P, INT, INT, INT, P, P, P]),
    "tfm_write_vtk": (INT, [P, INT, P, P, C.c_char_p]),
    "tfm_central_difference": (INT, [P, P, P, I64, P, P, P, P, D, INT, P, P]),
    "tfm_newmark": (INT, [P, P, P, I64, P, P, P, P, D, D, D, INT, C.POINTER(SolveOpts), P, P]),
    "tfm_solve_poisson": (INT, [P, P, D, INT, D, I64, P, P, C.POINTER(SolveOpts), P]),
    "tfm_assemble_robin": (INT, [P, P, C.POINTER(Operator), INT, D, INT, D, D, PP]),
    "tfm_solve_poisson_nr": (INT, [P, P, D, INT, D, P, P, I64, P, P, C.POINTER(SolveOpts), P]),
    "tfm_solve_elasticity": (INT, [P, P, INT, D, D, INT, P, I64, P, P, C.POINTER(SolveOpts), P]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_SIGS)


class SparseError(RuntimeError):
    """Mirrors tpt_fem_sparse::SparseError (crates/tpt-fem-sparse/src/lib.rs:66-84)."""


class ElasticityError(RuntimeError):
    """Mirrors tpt_fem_elasticity::ElasticityError::ModelDimMismatch."""


class TfmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class Context:
    """One tfm_ctx (device + stream).  Not shared across host threads."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = _lib.tfm_ctx_create(device, C.byref(h))
        if rc != OK:
            raise TfmError(rc, "tfm_ctx_create failed (no usable GPU?)")
        self.h = h

    def check(self, rc):
        if rc == OK:
            return
        msg = (_lib.tfm_last_error(self.h) or b"").decode()
        if rc == E_NUMERIC:
            raise SparseError(msg)
        if rc == E_MODEL:
            raise ElasticityError(msg)
        raise TfmError(rc, msg)

    def close(self):
        if self.h:
            _lib.tfm_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        self.check(_lib.tfm_ctx_synchronize(self.h))

    def timer_begin(self):
        self.check(_lib.tfm_ctx_timer_begin(self.h))

    def timer_end(self):
        """Device milliseconds since timer_begin on this context's stream."""
        ms = D()
        self.check(_lib.tfm_ctx_timer_end(self.h, C.byref(ms)))
        return ms.value

    def scan_total(self, counts):
        """The library's device prefix sum of int32 counts (its int32 range guard)."""
        c = np.ascontiguousarray(counts, dtype=np.int32)
        t = I64()
        self.check(_lib.tfm_debug_scan_total(self.h, _ptr(c), len(c), C.byref(t)))
        return t.value


class Mesh:
    """Device-resident mesh (crates/tpt-fem-mesh/src/lib.rs:39-148)."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle
        cell, nn, pdim, ne = INT(), I64(), INT(), I64()
        _lib.tfm_mesh_info(handle, C.byref(cell), C.byref(nn), C.byref(pdim), C.byref(ne))
        self.cell, self.n_nodes, self.pdim, self.n_elem = cell.value, nn.value, pdim.value, ne.value
        self.dim = DIMS[self.cell]

    @staticmethod
    def from_arrays(ctx, cell, coords, conn):
        coords, conn = _f64(coords), _i64(conn)
        h = C.c_void_p()
        ctx.check(_lib.tfm_mesh_create(ctx.h, cell, coords.shape[0], coords.shape[1], _ptr(coords),
                                       conn.shape[0], _ptr(conn), C.byref(h)))
        return Mesh(ctx, h)

    @staticmethod
    def box(ctx, mn, mx, n):
        """box_mesh(min, max, [nx, ny, nz]) (crates/tpt-fem-mesh-gen/src/lib.rs:422-495)."""
        h = C.c_void_p()
        lo, hi = _f64(mn), _f64(mx)  # kept alive across the call
        ctx.check(_lib.tfm_mesh_box(ctx.h, _ptr(lo), _ptr(hi), n[0], n[1], n[2], C.byref(h)))
        return Mesh(ctx, h)

    @staticmethod
    def grid(ctx, nx, ny, h=1.0):
        """topopt Grid::new (crates/tpt-fem-topopt/src/lib.rs:89-116) as a Q4 mesh."""
        hh = C.c_void_p()
        ctx.check(_lib.tfm_mesh_grid(ctx.h, nx, ny, h, C.byref(hh)))
        return Mesh(ctx, hh)

    @staticmethod
    def box_batch(ctx, n3):
        """P unit-cube box meshes, concatenated -> (mesh, node_off, elem_off)."""
        n3 = _i64(n3).reshape(-1, 3)
        P = n3.shape[0]
        noff, eoff = np.zeros(P + 1, np.int64), np.zeros(P + 1, np.int64)
        h = C.c_void_p()
        ctx.check(_lib.tfm_mesh_box_batch(ctx.h, P, _ptr(n3), C.byref(h), _ptr(noff), _ptr(eoff)))
        return Mesh(ctx, h), noff, eoff

    @staticmethod
    def grid_batch(ctx, P, nx, ny, h=1.0):
        """P copies of the topopt Grid (one per SIMP design), concatenated."""
        hh = C.c_void_p()
        ctx.check(_lib.tfm_mesh_grid_batch(ctx.h, P, nx, ny, h, C.byref(hh)))
        return Mesh(ctx, hh)

    def jitter(self, n, frac, seed):
        self.ctx.check(_lib.tfm_mesh_jitter(self.h, n[0], n[1], n[2], frac, seed))

    def download(self):
        coords = np.zeros((self.n_nodes, self.pdim))
        conn = np.zeros((self.n_elem, NODES[self.cell]), dtype=np.int64)
        self.ctx.check(_lib.tfm_mesh_download(self.h, _ptr(coords), _ptr(conn)))
        return coords, conn

    def quad_points(self, quad_order=2):
        """Positions of every element quadrature point, (n_elem, nq, pdim): the
        points the reference calls a source / body-force closure at."""
        nq = I64()
        self.ctx.check(_lib.tfm_mesh_quad_points(self.h, quad_order, None, C.byref(nq)))
        x = np.zeros((max(1, self.n_elem), nq.value, self.pdim))
        self.ctx.check(_lib.tfm_mesh_quad_points(self.h, quad_order, _ptr(x), C.byref(nq)))
        return x[: self.n_elem]

    def boundary_points(self, order=2):
        """Every boundary-face quadrature point's position and outward unit normal,
        two (count, pdim) arrays: the (x, n) a Neumann / Robin closure is called with."""
        cnt = I64()
        self.ctx.check(_lib.tfm_mesh_boundary_points(self.h, order, None, None, C.byref(cnt)))
        x = np.zeros((max(1, cnt.value), self.pdim))
        n = np.zeros((max(1, cnt.value), self.pdim))
        self.ctx.check(_lib.tfm_mesh_boundary_points(self.h, order, _ptr(x), _ptr(n), C.byref(cnt)))
        return x[: cnt.value], n[: cnt.value]

    def nodes_on_plane(self, axis, coord, tol):
        out = np.zeros(max(1, self.n_nodes), dtype=np.int64)
        cnt = I64()
        self.ctx.check(_lib.tfm_mesh_nodes_on_plane(self.h, axis, coord, tol, _ptr(out), C.byref(cnt)))
        return out[: cnt.value].copy()

    def close(self):
        if self.h:
            _lib.tfm_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Csr:
    """Host copy of a tpt_fem_sparse::Csr (crates/tpt-fem-sparse/src/lib.rs:50-63)."""

    def __init__(self, nrows, ncols, row_ptrs, col_ind, values):
        self.nrows, self.ncols = nrows, ncols
        self.row_ptrs, self.col_ind, self.values = row_ptrs, col_ind, values

    def nnz(self):
        return len(self.values)


class Matrix:
    """Device CSR (tfm_matrix)."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle
        nr, nc, nz = I64(), I64(), I64()
        _lib.tfm_matrix_info(handle, C.byref(nr), C.byref(nc), C.byref(nz))
        self.nrows, self.ncols, self.nnz = nr.value, nc.value, nz.value

    @staticmethod
    def from_csr(ctx, csr):
        h = C.c_void_p()
        rp, ci, v = _i64(csr.row_ptrs), _i64(csr.col_ind), _f64(csr.values)
        ctx.check(_lib.tfm_matrix_create(ctx.h, csr.nrows, csr.ncols, _ptr(rp), _ptr(ci), _ptr(v), C.byref(h)))
        return Matrix(ctx, h)

    def to_csr(self):
        rp = np.zeros(self.nrows + 1, dtype=np.int64)
        ci = np.zeros(max(1, self.nnz), dtype=np.int64)
        v = np.zeros(max(1, self.nnz))
        self.ctx.check(_lib.tfm_matrix_download(self.h, _ptr(rp), _ptr(ci), _ptr(v)))
        return Csr(self.nrows, self.ncols, rp, ci[: self.nnz].copy(), v[: self.nnz].copy())

    def matvec(self, x):
        """Csr::matvec (crates/tpt-fem-sparse/src/lib.rs:195-205)."""
        x = _f64(x)
        y = np.zeros(max(self.nrows, len(x)))
        self.ctx.check(_lib.tfm_matrix_matvec(self.h, _ptr(x), len(x), _ptr(y)))
        return y

    def close(self):
        if self.h:
            _lib.tfm_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def assemble(mesh, dofs_per_node, op):
    """assemble(mesh, dpn, |eid, m| element_matrix) -> device CSR equal to Coo::to_csr()."""
    h = C.c_void_p()
    mesh.ctx.check(_lib.tfm_assemble(mesh.ctx.h, mesh.h, C.byref(op), dofs_per_node, C.byref(h)))
    return Matrix(mesh.ctx, h)


def assemble_coo(mesh, dofs_per_node, op):
    """assemble(mesh, dpn, op) as the reference's Coo (crates/tpt-fem-assembly/src/lib.rs:458-480):
    (rows, cols, vals) in push order."""
    n = I64(0)
    mesh.ctx.check(_lib.tfm_assemble_coo(mesh.ctx.h, mesh.h, C.byref(op), dofs_per_node, C.byref(n), None, None,
                                         None))
    rows, cols, vals = np.zeros(max(1, n.value), np.int64), np.zeros(max(1, n.value), np.int64), np.zeros(max(1, n.value))
    mesh.ctx.check(_lib.tfm_assemble_coo(mesh.ctx.h, mesh.h, C.byref(op), dofs_per_node, C.byref(n), _ptr(rows),
                                         _ptr(cols), _ptr(vals)))
    k = n.value
    return rows[:k], cols[:k], vals[:k]


# ---- loads ------------------------------------------------------------------
def mass_matrix(mesh, density, quad_order=2):
    """elasticity_mass_matrix(mesh, model, density, quad_order).to_csr()
    (crates/tpt-fem-elasticity/src/lib.rs:386-428) as a device Matrix."""
    return assemble(mesh, mesh.dim, Operator.mass(density, quad_order))


def lumped_mass(mesh, density, quad_order=2):
    """elasticity_lumped_mass(...).to_csr() (crates/tpt-fem-elasticity/src/lib.rs:432-457)."""
    h = C.c_void_p()
    mesh.ctx.check(_lib.tfm_lumped_mass(mesh.ctx.h, mesh.h, density, quad_order, C.byref(h)))
    return Matrix(mesh.ctx, h)


def source_vector(mesh, f=1.0, quad_order=2, rhs=None):
    rhs = np.zeros(mesh.n_nodes) if rhs is None else _f64(rhs).copy()
    mesh.ctx.check(_lib.tfm_load_source(mesh.ctx.h, mesh.h, quad_order, f, _ptr(rhs)))
    return rhs


def body_vector(mesh, b, quad_order=2, rhs=None):
    rhs = np.zeros(mesh.n_nodes * mesh.dim) if rhs is None else _f64(rhs).copy()
    bb = _f64(list(b) + [0.0] * (3 - len(b)))
    mesh.ctx.check(_lib.tfm_load_body(mesh.ctx.h, mesh.h, quad_order, _ptr(bb), _ptr(rhs)))
    return rhs


def _values_at(fn, pts, ncomp, vectorized, normals=None):
    """A closure's values at points (n, pdim) [and normals]: one call per point
    as the reference makes them, or one call on the whole array (vectorized)."""
    if vectorized:
        v = fn(pts) if normals is None else fn(pts, normals)
        return np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (len(pts),) + ((ncomp,) if ncomp > 1 else ())))
    if normals is None:
        vals = [fn(x) for x in pts]
    else:
        vals = [fn(x, n) for x, n in zip(pts, normals)]
    out = np.asarray(vals, dtype=np.float64).reshape(len(pts), -1)
    return np.ascontiguousarray(out[:, :ncomp] if ncomp > 1 else out[:, 0])


def source_vector_qp(mesh, sq, quad_order=2, rhs=None):
    """rhs += poisson_source_vector scattered, from the source's values sq (n_elem, nq)."""
    rhs = np.zeros(mesh.n_nodes) if rhs is None else _f64(rhs).copy()
    sq = _f64(sq)
    mesh.ctx.check(_lib.tfm_load_source_qp(mesh.ctx.h, mesh.h, quad_order, _ptr(sq), _ptr(rhs)))
    return rhs


def body_vector_qp(mesh, bq, quad_order=2, rhs=None):
    """rhs += elasticity_body_vector scattered, from the body force's values bq (n_elem, nq, dim)."""
    rhs = np.zeros(mesh.n_nodes * mesh.dim) if rhs is None else _f64(rhs).copy()
    bq = _f64(bq)
    mesh.ctx.check(_lib.tfm_load_body_qp(mesh.ctx.h, mesh.h, quad_order, _ptr(bq), _ptr(rhs)))
    return rhs


def apply_neumann(mesh, dofs_per_node, g, rhs, order=2, vectorized=False):
    """apply_neumann / apply_neumann_order (crates/tpt-fem-assembly/src/lib.rs:591-667):
    g a constant flux, or a closure g(x, n) evaluated at the boundary points."""
    rhs = _f64(rhs).copy()
    if not callable(g) and order == 2:
        mesh.ctx.check(_lib.tfm_load_neumann(mesh.ctx.h, mesh.h, dofs_per_node, g, _ptr(rhs)))
        return rhs
    x, n = mesh.boundary_points(order)
    gq = _values_at(g, x, 1, vectorized, n) if callable(g) else np.full(len(x), float(g))
    mesh.ctx.check(_lib.tfm_load_neumann_qp(mesh.ctx.h, mesh.h, dofs_per_node, order, _ptr(gq), _ptr(rhs)))
    return rhs


def plane_traction(mesh, dofs_per_node, axis, coord, tol, t, rhs=None):
    rhs = np.zeros(mesh.n_nodes * dofs_per_node) if rhs is None else _f64(rhs).copy()
    tt = _f64(list(t) + [0.0] * (3 - len(t)))
    mesh.ctx.check(_lib.tfm_load_plane_traction(mesh.ctx.h, mesh.h, dofs_per_node, axis, coord, tol,
                                                _ptr(tt), _ptr(rhs)))
    return rhs


# ---- reduced systems ----------------------------------------------------------
def _bc_arrays(bcs):
    """Dirichlet list as (dof, value) pairs, or a (k, 2) array of them."""
    if isinstance(bcs, np.ndarray):
        if bcs.size == 0:
            return 0, np.zeros(1, np.int64), np.zeros(1)
        b = bcs.reshape(-1, 2)
        return len(b), _i64(b[:, 0]), _f64(b[:, 1])
    if not bcs:
        return 0, np.zeros(1, np.int64), np.zeros(1)
    return len(bcs), _i64([b[0] for b in bcs]), _f64([b[1] for b in bcs])


class System:
    """Reduced system on the device (ReducedSystem + its solver)."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle
        nf, nz, nd = I64(), I64(), I64()
        _lib.tfm_system_info(handle, C.byref(nf), C.byref(nz), C.byref(nd))
        self.n_free, self.nnz, self.n_dof = nf.value, nz.value, nd.value

    def mf_apply(self, x):
        """K x by the precision form's matrix-free FP32 operator (test hook)."""
        xx = _f64(x)
        y = np.zeros(self.n_dof)
        self.ctx.check(_lib.tfm_debug_mf_apply(self.h, _ptr(xx), _ptr(y)))
        return y

    def assembly_info(self):
        """Numeric assembly kernel and how its tile size was chosen."""
        path, tp, src = C.c_int(), C.c_int(), C.c_int()
        ms = (C.c_double * 2)()
        self.ctx.check(_lib.tfm_system_assembly_info(self.h, C.byref(path), C.byref(tp), C.byref(src), ms))
        return {"kernel": ("generic", "p1_tet", "triplets", "simp_grid")[path.value], "tile_pairs": tp.value,
                "source": ("fixed", "timed", "cached", "env")[src.value], "ms": {"256": ms[0], "192": ms[1]}}

    @staticmethod
    def create(mesh, op, dofs_per_node, bcs, rhs=None, flags=SYS_FP64):
        nb, bd, bv = _bc_arrays(bcs)
        r = None if rhs is None else _f64(rhs)
        h = C.c_void_p()
        mesh.ctx.check(_lib.tfm_system_create(mesh.ctx.h, mesh.h, C.byref(op), dofs_per_node, nb, _ptr(bd),
                                              _ptr(bv), _ptr(r), flags, C.byref(h)))
        return System(mesh.ctx, h)

    @staticmethod
    def from_matrix(matrix, rhs, bcs):
        nb, bd, bv = _bc_arrays(bcs)
        rhs = _f64(rhs)
        h = C.c_void_p()
        matrix.ctx.check(_lib.tfm_system_from_matrix(matrix.ctx.h, matrix.h, len(rhs), _ptr(rhs), nb, _ptr(bd),
                                                     _ptr(bv), C.byref(h)))
        return System(matrix.ctx, h)

    def assemble(self, op=None, which=0):
        """Numeric re-assembly (values only) on the existing pattern."""
        self.ctx.check(_lib.tfm_system_assemble(self.h, C.byref(op) if op is not None else None, which))

    def download(self):
        """(kred Csr, fred, free) like reduce_system's ReducedSystem."""
        rp = np.zeros(self.n_free + 1, np.int64)
        ci = np.zeros(max(1, self.nnz), np.int64)
        v = np.zeros(max(1, self.nnz))
        fred = np.zeros(max(1, self.n_free))
        free = np.zeros(max(1, self.n_free), np.int64)
        self.ctx.check(_lib.tfm_system_download(self.h, _ptr(rp), _ptr(ci), _ptr(v), _ptr(fred), _ptr(free)))
        ncols = int(ci[: self.nnz].max()) + 1 if self.nnz else 1
        nrows = self.n_free
        return (Csr(nrows, ncols, rp, ci[: self.nnz].copy(), v[: self.nnz].copy()),
                fred[: self.n_free].copy(), free[: self.n_free].copy())

    def values32(self):
        """The FP32 values (widened), in download()'s order."""
        v = np.zeros(max(1, self.nnz))
        self.ctx.check(_lib.tfm_debug_system_values32(self.h, _ptr(v)))
        return v[: self.nnz].copy()

    def solve(self, mode=SOLVE_PCG64, rtol=1e-12, max_iter=100000, inner_rtol=0.3):
        """Single-system solve.  Mixed mode: inner_rtol is the FP32 residual
        reduction between FP64 residual replacements."""
        u = np.zeros(max(1, self.n_dof))
        o = SolveOpts(mode=mode, rtol=rtol, max_iter=max_iter, inner_rtol=inner_rtol)
        st = SolveStats()
        self.ctx.check(_lib.tfm_system_solve(self.h, C.byref(o), _ptr(u), C.byref(st)))
        self.stats = st
        return u[: self.n_dof]

    def solve_noalloc(self, rtol=1e-12, max_iter=100000):
        """tfm_system_solve_noalloc: the FP64 Jacobi PCG without device
        allocations (work space in a fixed device array; at most ~1.197M DOFs)."""
        u = np.zeros(max(1, self.n_dof))
        o = SolveOpts(mode=SOLVE_PCG64, rtol=rtol, max_iter=max_iter, inner_rtol=0.3)
        st = SolveStats()
        self.ctx.check(_lib.tfm_system_solve_noalloc(self.h, C.byref(o), _ptr(u), C.byref(st)))
        self.stats = st
        return u[: self.n_dof]

    def solve_batch(self, dof_off, rhs=None, nrhs=1, mode=SOLVE_PCG64, rtol=1e-10, max_iter=20000,
                    inner_rtol=0.3):
        """Block-diagonal batched solve: returns u [nrhs, n_dof] and per-segment stats."""
        off = _i64(dof_off)
        P = len(off) - 1
        r = None if rhs is None else _f64(rhs).reshape(nrhs, self.n_dof)
        u = np.zeros((nrhs, self.n_dof))
        st = (SolveStats * (P * nrhs))()
        o = SolveOpts(mode=mode, rtol=rtol, max_iter=max_iter, inner_rtol=inner_rtol)
        self.ctx.check(_lib.tfm_system_solve_batch(self.h, P, _ptr(off), nrhs, _ptr(r), C.byref(o), _ptr(u),
                                                   C.cast(st, C.c_void_p)))
        self.batch_stats = [(s.iterations, s.outer, s.rel_residual, s.seconds) for s in st]
        return u

    def assemble_solve_batch(self, op, dof_off, rhs=None, nrhs=1, mode=SOLVE_PCG64, rtol=1e-10, max_iter=20000,
                             inner_rtol=0.3):
        """Re-assembly with op's coefficients (None: the current ones) and the
        batched solve as one sequence of captured graph launches; returns u
        [nrhs, n_dof] like solve_batch."""
        off = _i64(dof_off)
        P = len(off) - 1
        r = None if rhs is None else _f64(rhs).reshape(nrhs, self.n_dof)
        u = np.zeros((nrhs, self.n_dof))
        st = (SolveStats * (P * nrhs))()
        o = SolveOpts(mode=mode, rtol=rtol, max_iter=max_iter, inner_rtol=inner_rtol)
        self.ctx.check(_lib.tfm_system_assemble_solve_batch(
            self.h, C.byref(op) if op is not None else None, P, _ptr(off), nrhs, _ptr(r), C.byref(o), _ptr(u),
            C.cast(st, C.c_void_p)))
        self.batch_stats = [(s.iterations, s.outer, s.rel_residual, s.seconds) for s in st]
        return u

    def close(self):
        if self.h:
            _lib.tfm_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def reduce_system(matrix, rhs, bcs):
    """reduce_system (crates/tpt-fem-assembly/src/lib.rs:504-535) -> (kred, fred, free)."""
    s = System.from_matrix(matrix, rhs, bcs)
    try:
        return s.download()
    finally:
        s.close()


def solve_with_dirichlet(matrix, rhs, bcs, rtol=1e-13):
    """solve_with_dirichlet (crates/tpt-fem-assembly/src/lib.rs:543-559), solved by FP64 PCG."""
    s = System.from_matrix(matrix, rhs, bcs)
    try:
        return s.solve(SOLVE_PCG64, rtol=rtol)
    finally:
        s.close()


def solve_poisson(mesh, kappa, quad_order, f, bcs, rtol=1e-13, mode=SOLVE_PCG64, neumann=None, robin=None,
                  vectorized=False):
    """solve_poisson (crates/tpt-fem-thermal/src/lib.rs:206-240).

    f is a constant source or a closure s(x); neumann and robin are None or
    closures g(x, n) / h(x, n), evaluated at the points the reference calls
    them at (Mesh.quad_points, Mesh.boundary_points) -- one call per point, or
    one call on the whole (m, pdim) arrays with vectorized=True.  mode
    SOLVE_PCG64_MG solves with the multigrid CG (meshes from Mesh.box,
    constant source, no Neumann/Robin terms)."""
    nb, bd, bv = _bc_arrays(bcs)
    u = np.zeros(mesh.n_nodes)
    o = SolveOpts(mode=mode, rtol=rtol, max_iter=200000, inner_rtol=1e-5)
    if not callable(f) and neumann is None and robin is None:
        mesh.ctx.check(_lib.tfm_solve_poisson(mesh.ctx.h, mesh.h, kappa, quad_order, f, nb, _ptr(bd), _ptr(bv),
                                              C.byref(o), _ptr(u)))
        return u
    xq = mesh.quad_points(quad_order)
    nq = xq.shape[1]
    if callable(f):
        sq = _values_at(f, xq.reshape(-1, mesh.pdim), 1, vectorized).reshape(mesh.n_elem, nq)
    else:
        sq = np.full((mesh.n_elem, nq), float(f))
    gq = hq = None
    if neumann is not None or robin is not None:
        xb, nrm = mesh.boundary_points(2)
        if neumann is not None:
            gq = _values_at(neumann, xb, 1, vectorized, nrm)
        if robin is not None:
            hq = _values_at(robin, xb, 1, vectorized, nrm)
    mesh.ctx.check(_lib.tfm_solve_poisson_qp(mesh.ctx.h, mesh.h, kappa, quad_order, _ptr(sq), _ptr(gq), _ptr(hq),
                                             nb, _ptr(bd), _ptr(bv), C.byref(o), _ptr(u)))
    return u


def solve_elasticity(mesh, model, young, nu, quad_order, body, bcs, rtol=1e-13, mode=SOLVE_PCG64,
                     vectorized=False):
    """solve_elasticity (crates/tpt-fem-elasticity/src/lib.rs:355-381).

    body is a constant vector or a closure b(x) returning dim components,
    evaluated at Mesh.quad_points.  mode SOLVE_PCG64_MG solves with the
    multigrid CG (meshes from Mesh.box, constant body force)."""
    nb, bd, bv = _bc_arrays(bcs)
    u = np.zeros(mesh.n_nodes * mesh.dim)
    o = SolveOpts(mode=mode, rtol=rtol, max_iter=200000, inner_rtol=1e-5)
    if callable(body):
        xq = mesh.quad_points(quad_order)
        bq = _values_at(body, xq.reshape(-1, mesh.pdim), mesh.dim, vectorized)
        bq = np.ascontiguousarray(bq.reshape(mesh.n_elem, xq.shape[1], mesh.dim))
        mesh.ctx.check(_lib.tfm_solve_elasticity_qp(mesh.ctx.h, mesh.h, model, young, nu, quad_order, _ptr(bq), nb,
                                                    _ptr(bd), _ptr(bv), C.byref(o), _ptr(u)))
        return u
    b = _f64(list(body) + [0.0] * (3 - len(body)))
    mesh.ctx.check(_lib.tfm_solve_elasticity(mesh.ctx.h, mesh.h, model, young, nu, quad_order, _ptr(b), nb,
                                             _ptr(bd), _ptr(bv), C.byref(o), _ptr(u)))
    return u


class TopOptResult:
    """TopOptResult (crates/tpt-fem-topopt/src/lib.rs:140-148) plus the solver statistics."""

    def __init__(self, densities, compliance, iterations, stats):
        self.densities, self.compliance, self.iterations, self.stats = densities, compliance, iterations, stats


def topopt_simp(ctx, nx, ny, h, e0, nu, vol_frac, penal, filter_radius, max_iter, move_limit, f, bcs,
                mode=SOLVE_PCG64_MG, rtol=1e-12, pcg_max_iter=100000, inner_rtol=0.3):
    """topopt_simp(params, f, bcs) (crates/tpt-fem-topopt/src/lib.rs:342-385) on the GPU.

    The state equations are solved by FP64 CG with the multigrid V-cycle by
    default (mode SOLVE_PCG64 / SOLVE_MIXED select Jacobi PCG / the precision
    form).

    With f of shape (P, 2*nx*ny) (and vol_frac a scalar or P values) the P
    designs run as one batch; the result's arrays then gain a leading P axis."""
    prm = TopOptParams(nx=nx, ny=ny, h=h, e0=e0, nu=nu, vol_frac=float(np.ravel(vol_frac)[0]), penal=penal,
                       filter_radius=filter_radius, max_iter=max_iter, move_limit=move_limit)
    ff = _f64(f)
    batched = ff.ndim == 2
    ff = ff.reshape(-1, 2 * nx * ny)
    Pn = ff.shape[0]
    vf = _f64(np.broadcast_to(np.ravel(vol_frac), (Pn,)))
    nb, bd, bv = _bc_arrays(bcs)
    ne = (nx - 1) * (ny - 1)
    rho = np.zeros((Pn, ne))
    comp = np.zeros((Pn, max_iter + 1))
    its = np.zeros(Pn, dtype=np.int32)
    st = (SolveStats * Pn)()
    o = SolveOpts(mode=mode, rtol=rtol, max_iter=pcg_max_iter, inner_rtol=inner_rtol)
    ctx.check(_lib.tfm_topopt_simp_batch(ctx.h, Pn, C.byref(prm), _ptr(ff), _ptr(vf), nb, _ptr(bd), _ptr(bv),
                                         C.byref(o), _ptr(rho), _ptr(comp), its.ctypes.data,
                                         C.cast(st, C.c_void_p)))
    stats = [(s.iterations, s.outer, s.rel_residual, s.seconds) for s in st]
    if batched:
        return TopOptResult(rho, comp, its.astype(np.int64), stats)
    return TopOptResult(rho[0], comp[0], int(its[0]), stats[0])


EIG_SMALLEST, EIG_LARGEST = 0, 1


def lanczos_eigs(matrix, num, which, lanczos_dim, vectors=True):
    """lanczos_eigs(coo, num, which, lanczos_dim) (crates/tpt-fem-eigen/src/lib.rs:223-268)
    on a device Matrix -> [(lambda, vec)] (vec None when vectors is False)."""
    n = matrix.nrows
    ev = np.zeros(max(num, 1))
    vec = np.zeros((max(num, 1), n)) if vectors else None
    cnt = INT(0)
    matrix.ctx.check(_lib.tfm_lanczos_eigs(matrix.h, num, which, lanczos_dim, _ptr(ev), _ptr(vec),
                                           C.byref(cnt)))
    return [(float(ev[i]), vec[i].copy() if vectors else None) for i in range(cnt.value)]


class CflViolation(RuntimeError):
    """DynamicError::CflViolation (crates/tpt-fem-dynamic/src/lib.rs:46, 58-61)."""

    def __init__(self, msg, critical):
        super().__init__(msg)
        self.critical = critical


def central_difference(m, c, k, u0, v0, f, dt, nsteps, fscale=None):
    """central_difference (crates/tpt-fem-dynamic/src/lib.rs:289-377) on device matrices,
    f(t_k) = fscale[k] * f -> history [nsteps+1, n]."""
    u0, v0, f = _f64(u0), _f64(v0), _f64(f)
    fs = None if fscale is None else _f64(fscale)
    n = len(u0)
    hist = np.zeros((nsteps + 1, n))
    crit = D(0.0)
    rc = _lib.tfm_central_difference(m.h, c.h, k.h, n, _ptr(u0), _ptr(v0), _ptr(f), _ptr(fs), dt, nsteps,
                                     _ptr(hist), C.byref(crit))
    if rc == E_CFL:
        raise CflViolation((_lib.tfm_last_error(m.ctx.h) or b"").decode(), crit.value)
    m.ctx.check(rc)
    return hist


def newmark(m, c, k, u0, v0, f, dt, beta, gamma, nsteps, fscale=None, rtol=1e-12, max_iter=100000):
    """newmark (crates/tpt-fem-dynamic/src/lib.rs:194-263), PCG per step -> history [nsteps+1, n]."""
    u0, v0, f = _f64(u0), _f64(v0), _f64(f)
    fs = None if fscale is None else _f64(fscale)
    n = len(u0)
    hist = np.zeros((nsteps + 1, n))
    o = SolveOpts(mode=SOLVE_PCG64, rtol=rtol, max_iter=max_iter, inner_rtol=0.3)
    st = SolveStats()
    m.ctx.check(_lib.tfm_newmark(m.h, c.h, k.h, n, _ptr(u0), _ptr(v0), _ptr(f), _ptr(fs), dt, beta, gamma, nsteps,
                                 C.byref(o), _ptr(hist), C.byref(st)))
    m.stats = st
    return hist


def assemble_robin(mesh, dofs_per_node, op, h, axis=-1, coord=0.0, tol=0.0, order=2, vectorized=False):
    """to_csr() of assemble(...) followed by apply_robin(mesh, dpn, h, coo)
    (crates/tpt-fem-assembly/src/lib.rs:674-753): h constant (or on a plane
    band), or a closure h(x, n) evaluated at the boundary points of `order`."""
    out = C.c_void_p()
    if callable(h) or order != 2:
        x, n = mesh.boundary_points(order)
        hq = _values_at(h, x, 1, vectorized, n) if callable(h) else np.full(len(x), float(h))
        mesh.ctx.check(_lib.tfm_assemble_robin_qp(mesh.ctx.h, mesh.h, C.byref(op), dofs_per_node, order, _ptr(hq),
                                                  C.byref(out)))
        return Matrix(mesh.ctx, out)
    mesh.ctx.check(_lib.tfm_assemble_robin(mesh.ctx.h, mesh.h, C.byref(op), dofs_per_node, h, axis, coord, tol,
                                           C.byref(out)))
    return Matrix(mesh.ctx, out)


def solve_poisson_nr(mesh, kappa, quad_order, f, bcs, neumann=None, robin=None, rtol=1e-13):
    """solve_poisson (crates/tpt-fem-thermal/src/lib.rs:206-240) with constant Neumann flux and Robin coefficient."""
    nb, bd, bv = _bc_arrays(bcs)
    u = np.zeros(mesh.n_nodes)
    g = None if neumann is None else _f64([neumann])
    hh = None if robin is None else _f64([robin])
    o = SolveOpts(mode=SOLVE_PCG64, rtol=rtol, max_iter=200000, inner_rtol=1e-5)
    mesh.ctx.check(_lib.tfm_solve_poisson_nr(mesh.ctx.h, mesh.h, kappa, quad_order, f, _ptr(g), _ptr(hh), nb,
                                             _ptr(bd), _ptr(bv), C.byref(o), _ptr(u)))
    return u


def write_vtk(mesh, path, fields=()):
    """write_vtk / write_vtk_with_data (crates/tpt-fem-io-vtk/src/lib.rs:117-183):
    fields is a sequence of (name, per-node values)."""
    names = (C.c_char_p * max(1, len(fields)))(*[f[0].encode() for f in fields])
    arrs = [_f64(f[1]) for f in fields]
    for a in arrs:
        if len(a) != mesh.n_nodes:
            raise ValueError("a field needs one value per node")
    vals = (C.c_void_p * max(1, len(fields)))(*[a.ctypes.data for a in arrs])
    mesh.ctx.check(_lib.tfm_write_vtk(mesh.h, len(fields), C.cast(names, C.c_void_p), C.cast(vals, C.c_void_p),
                                      str(path).encode()))
