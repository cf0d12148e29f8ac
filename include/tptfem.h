/*
 * tptfem.h -- C ABI of the MI355X (gfx950) element-assembly and reduced-system
 * solve library.  Plain C99: opaque handles, plain pointers and sizes, int
 * status codes, no exceptions across the boundary.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the tpt-solutions/tpt-fem repository).
 *
 * Threading: one tfm_ctx per host thread; calls on one context are not
 * concurrent.  Contexts are independent (the reference has no global state).
 * Host buffers are caller-owned; device memory is owned by the handles.
 */
#ifndef TPTFEM_H
#define TPTFEM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TFM_OK 0
#define TFM_E_INVALID (-1)       /* contract violation; the reference panics  */
#define TFM_E_NUMERIC (-2)       /* SparseError::Numeric (sparse/src/lib.rs:66-84) */
#define TFM_E_NOT_CONVERGED (-3) /* PCG hit max_iter (no reference counterpart) */
#define TFM_E_BREAKDOWN (-4)     /* PCG p'Ap <= 0: operator not SPD           */
#define TFM_E_HIP (-5)           /* HIP runtime error                          */
#define TFM_E_UNSUPPORTED (-6)   /* input outside what is built: keep the CPU path */
#define TFM_E_MODEL (-7)         /* ElasticityError::ModelDimMismatch (elasticity/src/lib.rs:28-35) */
#define TFM_E_CFL (-8)           /* DynamicError::CflViolation (dynamic/src/lib.rs:46, 58-61) */

/* CellType order: crates/tpt-fem-mesh/src/lib.rs:39-52 (P1 cells). */
enum { TFM_LINE = 0, TFM_TRI = 1, TFM_QUAD = 2, TFM_TET = 3, TFM_HEX = 4 };
/* ElasticModel order: crates/tpt-fem-elasticity/src/lib.rs:58-67. */
enum { TFM_BAR = 0, TFM_PLANE_STRESS = 1, TFM_PLANE_STRAIN = 2, TFM_CONTINUUM3D = 3 };
/* Element operators the library evaluates on the GPU. */
enum {
  TFM_OP_POISSON = 0,    /* poisson_element_matrix, thermal/src/lib.rs:132-160      */
  TFM_OP_ELASTICITY = 1, /* elasticity_element_matrix, elasticity/src/lib.rs:262-309 */
  TFM_OP_SIMP = 2,       /* topopt assemble, topopt/src/lib.rs:166-238 (Q4 only)      */
  TFM_OP_MASS = 3        /* elasticity_mass_matrix, elasticity/src/lib.rs:386-428:
                            coef = density, dofs_per_node = the cell's dimension  */
};
enum { TFM_FP64 = 0, TFM_FP32 = 1 };

typedef struct tfm_ctx tfm_ctx;
typedef struct tfm_mesh tfm_mesh;
typedef struct tfm_matrix tfm_matrix;
typedef struct tfm_system tfm_system;

/* Element operator: which closure the reference's assemble() would be given.
 * coef is kappa (Poisson) or Young's modulus (elasticity) when coef_e is NULL;
 * coef_e (host, n_elem) gives a per-element value instead, as an
 * `|eid, m| ...element_matrix(m, eid, coef[eid], ...)` closure would.
 * SIMP: E(e) = emin + rho[e]^penal * (e0 - emin), Ke = E(e) * Ke_unit(nu, h). */
typedef struct {
  int kind;
  int model;
  int quad_order;
  double coef;
  const double *coef_e;
  double nu;
  const double *rho;
  double penal, e0, emin, h;
} tfm_operator;

/* ---- context ------------------------------------------------------------ */
int tfm_ctx_create(int device, tfm_ctx **out);
void tfm_ctx_destroy(tfm_ctx *ctx);
/* Text of the last error on this context; mirrors SparseError's Display
 * strings (crates/tpt-fem-sparse/src/lib.rs:76-84) where one applies. */
const char *tfm_last_error(const tfm_ctx *ctx);
int tfm_ctx_synchronize(tfm_ctx *ctx);
/* Device-time bracket on the context's stream (HIP events): begin records the
 * start, end records the stop, synchronises and returns the elapsed ms.  No
 * reference counterpart (measurement utility). */
int tfm_ctx_timer_begin(tfm_ctx *ctx);
int tfm_ctx_timer_end(tfm_ctx *ctx, double *ms);
/* The library's device prefix sum of n non-negative int32 counts (host
 * buffer), as every pattern build runs it: *total = their sum, or
 * TFM_E_UNSUPPORTED when the sum reaches 2^31 - 1 (row pointers, slot and
 * block indices are int32).  No reference counterpart (diagnostic of the
 * range guard). */
int tfm_debug_scan_total(tfm_ctx *ctx, const int32_t *counts, int64_t n, int64_t *total);

/* ---- meshes (crates/tpt-fem-mesh/src/lib.rs:39-148) ---------------------- */
/* Mesh with one cell type; coords node-major [n_nodes x pdim], conn
 * element-major [n_elem x nodes_per_cell].  Node and element order are kept,
 * so DOF numbering equals Mesh::number_dofs (mesh/src/lib.rs:242-250). */
int tfm_mesh_create(tfm_ctx *ctx, int cell, int64_t n_nodes, int pdim, const double *coords,
                    int64_t n_elem, const int64_t *conn, tfm_mesh **out);
/* box_mesh (crates/tpt-fem-mesh-gen/src/lib.rs:422-495), generated on the GPU. */
int tfm_mesh_box(tfm_ctx *ctx, const double mn[3], const double mx[3], int64_t nx,
                 int64_t ny, int64_t nz, tfm_mesh **out);
/* topopt Grid::new (crates/tpt-fem-topopt/src/lib.rs:89-116), generated on the GPU. */
int tfm_mesh_grid(tfm_ctx *ctx, int64_t nx, int64_t ny, double h, tfm_mesh **out);
/* Move every interior node of an (nx,ny,nz) box_mesh by (2u-1)*frac*h per axis,
 * u = counter hash(seed, 3*node+axis).  No reference counterpart (input generator). */
int tfm_mesh_jitter(tfm_mesh *mesh, int64_t nx, int64_t ny, int64_t nz, double frac,
                    uint64_t seed);
int tfm_mesh_info(const tfm_mesh *mesh, int *cell, int64_t *n_nodes, int *pdim,
                  int64_t *n_elem);
int tfm_mesh_download(const tfm_mesh *mesh, double *coords, int64_t *conn);
/* Mesh::nodes_on_plane (mesh/src/lib.rs:425-431); out holds n_nodes entries. */
int tfm_mesh_nodes_on_plane(const tfm_mesh *mesh, int axis, double coord, double tol,
                            int64_t *out, int64_t *count);
/* Releases the caller's handle.  A system built on the mesh holds it too: the
 * device memory is freed when the handle and every such system are gone, so
 * a system may outlive the handle (as a Rust owner would drop it). */
void tfm_mesh_destroy(tfm_mesh *mesh);

/* ---- assembly (crates/tpt-fem-assembly/src/lib.rs:428-480) -------------- */
/* K = assemble(mesh, dpn, op).to_csr() (sparse/src/lib.rs:122-175): exact-zero
 * element entries skipped, columns ascending, nrows = max row + 1. */
int tfm_assemble(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op,
                 int dofs_per_node, tfm_matrix **out);
/* assemble(mesh, dpn, op) itself (assembly/src/lib.rs:458-480): the Coo's
 * triplets in push order -- element order, then local row, then local
 * column, exact-zero entries skipped (every SIMP entry kept) -- as host rows,
 * cols, vals.  Two calls: with rows, cols and vals NULL, *count receives the
 * number of triplets; then *count must hold the buffers' capacity and
 * receives the number written. */
int tfm_assemble_coo(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op,
                     int dofs_per_node, int64_t *count, int64_t *rows, int64_t *cols,
                     double *vals);
/* elasticity_lumped_mass(mesh, model, density, quad_order).to_csr()
 * (elasticity/src/lib.rs:432-457): the diagonal CSR whose entry on each DOF
 * of a node is the sum of the node's consistent-mass rows over dim.  */
int tfm_lumped_mass(tfm_ctx *ctx, const tfm_mesh *mesh, double density, int quad_order,
                    tfm_matrix **out);
/* A CSR matrix from host arrays (a Csr built by the caller). */
int tfm_matrix_create(tfm_ctx *ctx, int64_t nrows, int64_t ncols, const int64_t *row_ptrs,
                      const int64_t *col_ind, const double *values, tfm_matrix **out);
int tfm_matrix_info(const tfm_matrix *m, int64_t *nrows, int64_t *ncols, int64_t *nnz);
int tfm_matrix_download(const tfm_matrix *m, int64_t *row_ptrs, int64_t *col_ind,
                        double *values);
/* Csr::matvec (sparse/src/lib.rs:195-205): y has max(nrows, nx) entries. */
int tfm_matrix_matvec(const tfm_matrix *m, const double *x, int64_t nx, double *y);
void tfm_matrix_destroy(tfm_matrix *m);

/* ---- loads -------------------------------------------------------------- */
/* rhs[node] += poisson_source_vector with s(x) = f (thermal/src/lib.rs:163-196,
 * scattered as in solve_poisson :223-230).  rhs: host, n_nodes. */
int tfm_load_source(tfm_ctx *ctx, const tfm_mesh *mesh, int quad_order, double f,
                    double *rhs);
/* rhs[node*dim+a] += elasticity_body_vector with constant b (elasticity/src/lib.rs:312-348). */
int tfm_load_body(tfm_ctx *ctx, const tfm_mesh *mesh, int quad_order, const double *b,
                  double *rhs);
/* Closure-valued loads.  The reference's loads take closures: the source
 * s(x) of poisson_source_vector (thermal/src/lib.rs:163-196), the body force
 * b(x) of elasticity_body_vector (elasticity/src/lib.rs:312-348), the flux
 * g(x, n) of apply_neumann_order and the coefficient h(x, n) of
 * apply_robin_order (assembly/src/lib.rs:601-753).  Across the C ABI a
 * closure becomes its values at the points the reference calls it at: the
 * library hands out those points, the caller evaluates, the library
 * integrates in the reference's order (bit-identical to a closure returning
 * the same values).
 * Element points: xq [n_elem][nq][pdim] (xq NULL: *nq only), positions formed
 * as the reference forms them (x[d] = sum_k N_k x_k[d], node order). */
int tfm_mesh_quad_points(const tfm_mesh *mesh, int quad_order, double *xq, int64_t *nq);
/* rhs[node] += poisson_source_vector from per-point values sq [n_elem][nq]. */
int tfm_load_source_qp(tfm_ctx *ctx, const tfm_mesh *mesh, int quad_order, const double *sq,
                       double *rhs);
/* rhs[node*dim+a] += elasticity_body_vector from per-point values bq [n_elem][nq][dim]. */
int tfm_load_body_qp(tfm_ctx *ctx, const tfm_mesh *mesh, int quad_order, const double *bq,
                     double *rhs);
/* Boundary points: every boundary-face quadrature point (boundary_faces order,
 * assembly/src/lib.rs:563-583, then the face rule of `order`) with its
 * position and outward unit normal, x and normal [count][pdim] (either NULL:
 * not written).  Point faces of Line2 cells have no points. */
int tfm_mesh_boundary_points(const tfm_mesh *mesh, int order, double *x, double *normal,
                             int64_t *count);
/* apply_neumann_order(mesh, dpn, flux, rhs, order) from per-point values gq [count]. */
int tfm_load_neumann_qp(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, int order, const double *gq,
                        double *rhs);
/* to_csr() of assemble(mesh, dpn, op) with apply_robin_order(mesh, dpn, coeff,
 * coo, order)'s triplets from per-point values hq [count]. */
int tfm_assemble_robin_qp(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn,
                          int order, const double *hq, tfm_matrix **out);

/* apply_neumann (assembly/src/lib.rs:591-667) with constant flux g tiled over dpn. */
int tfm_load_neumann(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, double g, double *rhs);
/* Traction t[dpn] on the plane x[axis] == coord (|.| <= tol): one apply_neumann
 * pass per nonzero component into a scalar buffer, interleaved into rhs. */
int tfm_load_plane_traction(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, int axis,
                            double coord, double tol, const double *t, double *rhs);

/* to_csr() of assemble(mesh, dpn, op) with apply_robin(mesh, dpn, h, coo)'s
 * triplets appended (assembly/src/lib.rs:674-753): every boundary face adds
 * w * measure * h * N_k * N_l on the component diagonal of its node pairs,
 * summed after the element contributions.  h is constant, or -- axis >= 0 --
 * h on the band |x[axis] - coord| <= tol and 0 elsewhere (zero entries are
 * still created, as the reference pushes them). */
int tfm_assemble_robin(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dofs_per_node,
                       double h, int axis, double coord, double tol, tfm_matrix **out);

/* ---- reduced systems (assembly/src/lib.rs:489-559) ----------------------- */
enum { TFM_SYS_FP64 = 1, TFM_SYS_FP32 = 2 };
/* reduce_system(assemble(mesh, dpn, op), rhs, bcs) without forming the full K:
 * the reduced CSR pattern is built on the GPU (symbolic phase) and filled
 * (numeric phase) in the values named by flags.  rhs: host, n_nodes*dpn. */
int tfm_system_create(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn,
                      int64_t n_bc, const int64_t *bc_dof, const double *bc_val,
                      const double *rhs, int flags, tfm_system **out);
/* reduce_system(K, rhs, bcs) of an assembled matrix (general lifting). */
int tfm_system_from_matrix(tfm_ctx *ctx, const tfm_matrix *k, int64_t n, const double *rhs,
                           int64_t n_bc, const int64_t *bc_dof, const double *bc_val,
                           tfm_system **out);
/* Numeric re-assembly on the existing pattern (the hot path).  op may carry
 * new coefficients (coef_e, rho), or be NULL to reuse the device-resident
 * ones.  which: TFM_SYS_FP64 and/or TFM_SYS_FP32 value arrays to refill (0 =
 * every array the system holds).  Refilling both on the P1-tet kernel is one
 * pass whose FP32 values are the FP64 values rounded to nearest; the FP32
 * array alone takes the FP32 pass (FP64 geometry, FP32 element arithmetic).
 * The pattern was built from the creation coefficients: an element whose
 * coefficient was exactly 0 then has no entries in it, so giving it a
 * non-zero coefficient returns TFM_E_INVALID (create the system again); a
 * coefficient that becomes 0 leaves explicit zero entries.
 * FP32 values last filled with older coefficients than the FP64 values are
 * stale: TFM_SOLVE_MIXED refuses them (TFM_E_INVALID) and the multigrid CG
 * runs its fine passes on the FP64 values instead; an FP32 pass with op NULL
 * brings them up to date. */
int tfm_system_assemble(tfm_system *sys, const tfm_operator *op, int which);
int tfm_system_info(const tfm_system *sys, int64_t *n_free, int64_t *nnz, int64_t *n_dof);
/* No counterpart in the reference (diagnostic): which numeric assembly
 * kernel the system refills with and how its launch shape was chosen.
 * path: 0 = generic tile kernel, 1 = P1-tet kernel, 2 = sorted triplets (node
 * valence beyond the tile kernels), 3 = SIMP-grid rows (tfm_mesh_grid_batch
 * meshes); tile_pairs: element-node
 * pairs (= threads) per P1-tet tile (192 or 256); source: 1 = timed at
 * creation, 2 = reused from an earlier system of the same dofs per node,
 * value arrays and size class (floor(log2(elements))), 3 = forced by
 * TFM_TET_TILE_PAIRS; ms[2]: the timed pass of each candidate (256, 192)
 * when source is 1, else 0.  Any pointer may be NULL. */
int tfm_system_assembly_info(const tfm_system *sys, int *path, int *tile_pairs, int *source, double *ms);
/* No counterpart in the reference (test hook): y = K x with the precision
 * form's matrix-free FP32 operator (x, y: host, n_dof; x rounded to FP32;
 * rows of fixed DOFs give 0 and x must be 0 there).  TFM_E_UNSUPPORTED when
 * the system has no such operator (see TFM_SOLVE_MIXED). */
int tfm_debug_mf_apply(tfm_system *sys, const double *x, double *y);
/* ReducedSystem {kred (as its to_csr()), fred, free}. */
int tfm_system_download(const tfm_system *sys, int64_t *row_ptrs, int64_t *col_ind,
                        double *values, double *fred, int64_t *free_dofs);
/* The FP32 values of a system that holds them, in tfm_system_download's
 * order (nnz entries, widened to double); TFM_E_INVALID without them.  No
 * reference counterpart (diagnostic of the precision form's matrix). */
int tfm_debug_system_values32(const tfm_system *sys, double *values);

/* TFM_SOLVE_PCG64: Jacobi-preconditioned FP64 CG.  TFM_SOLVE_MIXED: FP32 CG
   with FP64 residual refinement (r = f - K u on the FP64 values); its FP32
   operator is matrix-free on systems built from a tfm_mesh_box mesh of P1
   tets (Poisson, 3-D elasticity): element contributions recomputed in FP32
   from the coordinates and coefficients on every product, determinants in
   FP64 -- no FP32 values needed, but the FP64 values must be current with
   the last coefficients (TFM_E_INVALID otherwise); elsewhere it is the
   assembled FP32 values.  TFM_SOLVE_MIXED_CSR: the same refinement on the
   assembled FP32 values always (single systems; the batched solve takes
   TFM_SOLVE_MIXED).  TFM_SOLVE_PCG64_MG: FP64 CG preconditioned
   by a geometric multigrid V-cycle, for systems on a tfm_mesh_box mesh (one
   system, one rhs) or on a tfm_mesh_grid_batch mesh (SIMP designs, one
   problem per design); TFM_E_UNSUPPORTED elsewhere.  TFM_SOLVE_DIRECT: the
   reference's dense partial-pivot LU (sparse/src/lib.rs:244-278) on the GPU,
   bit-identical to that loop run sequentially, for at most 8000 free DOFs
   (TFM_E_UNSUPPORTED above).  PCG64 and MIXED fall back to DIRECT on such
   small systems when CG breaks down (the reduced matrix is not SPD) or
   stalls, as the reference solves every system directly (MIXED_CSR too). */
enum { TFM_SOLVE_PCG64 = 0, TFM_SOLVE_MIXED = 1, TFM_SOLVE_PCG64_MG = 2, TFM_SOLVE_DIRECT = 3,
       TFM_SOLVE_MIXED_CSR = 4 };
typedef struct {
  int mode;           /* TFM_SOLVE_*                                          */
  double rtol;        /* stop when ||f - K u|| <= rtol * ||f|| (FP64 residual) */
  int max_iter;       /* total inner iterations                               */
  double inner_rtol;  /* mixed: FP32 residual reduction between FP64 residual
                         replacements (outside (0,1): 0.3), per segment when
                         batched; refinement cycles run until rtol is met or
                         max_iter inner iterations are spent                  */
} tfm_solve_opts;
typedef struct {
  int iterations;     /* inner PCG iterations in total          */
  int outer;          /* refinement cycles (1 for PCG64)        */
  double rel_residual;/* final ||f - K u|| / ||f|| in FP64      */
  double seconds;     /* device time of the solve               */
} tfm_solve_stats;
/* solve_with_dirichlet (assembly/src/lib.rs:543-559): solve the reduced system
 * by Jacobi-PCG (FP64) or FP32 PCG with FP64 refinement, scatter into u
 * (host, n_dof) with u[bc] = value in bcs order (last duplicate wins). */
int tfm_system_solve(tfm_system *sys, const tfm_solve_opts *opts, double *u,
                     tfm_solve_stats *stats);
/* tfm_system_solve in the TFM_SOLVE_PCG64 mode without any device
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate (serving loops, harnesses that forbid hipMalloc).  The PCG
 * scalars, block partial sums, flags and work vectors live in one file-scope
 * device array of 64 MiB per GPU (held from the first tfm_ctx_create on that
 * GPU until the process ends), so the system may have at most about
 * 1,197,000 DOFs (n_dof, free and fixed) -- larger systems return
 * TFM_E_UNSUPPORTED -- and the iterations run the CSR-stream SpMV, launched
 * directly.  Unlike tfm_system_solve there is no dense-LU fallback: a system
 * CG cannot solve returns TFM_E_BREAKDOWN / TFM_E_NOT_CONVERGED.  Every call
 * on one GPU uses the same array: calls that overlap on one GPU (two host
 * threads, two contexts) race and must be serialised by the caller. */
int tfm_system_solve_noalloc(tfm_system *sys, const tfm_solve_opts *opts, double *u,
                             tfm_solve_stats *stats);
void tfm_system_destroy(tfm_system *sys);

/* ---- batching form (no reference counterpart: the reference solves one
 * problem per call; these run many through one assembly and one PCG) ------ */
/* P unit-cube box_mesh problems with per-problem cells n3[3p..3p+2],
 * concatenated (node_off / elem_off: P+1 entries, may be NULL). */
int tfm_mesh_box_batch(tfm_ctx *ctx, int64_t P, const int64_t *n3, tfm_mesh **out,
                       int64_t *node_off, int64_t *elem_off);
/* P copies of the topopt Grid (one per SIMP design), concatenated. */
int tfm_mesh_grid_batch(tfm_ctx *ctx, int64_t P, int64_t nx, int64_t ny, double h,
                        tfm_mesh **out);
/* Solve a block-diagonal system (P problems, DOF ranges dof_off[p]..dof_off[p+1])
 * for nrhs >= 1 right-hand sides each, every (problem, rhs) segment with its
 * own PCG scalars and convergence (the kernels interleave 8 right-hand
 * sides: other counts run as chunks of 8 and then one at a time).  rhs: host
 * [nrhs][n_dof] or NULL (the system's own rhs, nrhs = 1); u: host
 * [nrhs][n_dof]; stats: [P*nrhs] (segment p*nrhs + k).  Homogeneous
 * Dirichlet values only.  With TFM_SOLVE_DIRECT (P = 1 or any block-diagonal
 * system up to 8000 free DOFs) this is solve_multi (sparse/src/lib.rs:
 * 244-278): one dense LU, every right-hand side solved against it. */
int tfm_system_solve_batch(tfm_system *sys, int64_t P, const int64_t *dof_off, int nrhs,
                           const double *rhs, const tfm_solve_opts *opts, double *u,
                           tfm_solve_stats *stats);
/* Batching form, re-assembly and solve together (no reference counterpart:
 * the reference assembles then solves one problem per call): op's
 * coefficients (NULL: the current ones; kind, model and rule unchanged) are
 * assembled into every value array the system holds and the batched solve of
 * tfm_system_solve_batch follows, as one sequence of captured hipGraph
 * launches -- numeric assembly and the solver's refresh, initial state,
 * chunks of iterations until every segment is done, residual certificate.
 * The solver and its graphs stay with the system, so a later call with the
 * same P, DOF ranges, nrhs and mode replays them.  Results equal
 * tfm_system_assemble followed by tfm_system_solve_batch.  TFM_SOLVE_DIRECT
 * and rhs counts other than 1 and 8 run those two calls. */
int tfm_system_assemble_solve_batch(tfm_system *sys, const tfm_operator *op, int64_t P,
                                    const int64_t *dof_off, int nrhs, const double *rhs,
                                    const tfm_solve_opts *opts, double *u,
                                    tfm_solve_stats *stats);

/* ---- SIMP topology optimization (crates/tpt-fem-topopt/src/lib.rs) ------ */
/* TopOptParams (topopt/src/lib.rs:119-137), the Grid given by Grid::new's
 * arguments (:89-116): nx x ny nodes, (nx-1)(ny-1) Q4 elements of side h. */
typedef struct {
  int64_t nx, ny;
  double h;
  double e0, nu, vol_frac, penal, filter_radius;
  int max_iter;
  double move_limit;
} tfm_topopt_params;
/* topopt_simp(params, f, bcs) (topopt/src/lib.rs:342-385) -> TopOptResult
 * {densities, compliance, iterations} (:140-148).  f: host, 2*nx*ny; bcs as
 * (dof, value) pairs with value 0 (other values: TFM_E_UNSUPPORTED).  Each
 * state equation K(rho) u = f is solved on the GPU by the PCG of opts (NULL:
 * FP64 CG with the multigrid V-cycle, TFM_SOLVE_PCG64_MG, to rtol 1e-12)
 * instead of the dense LU, warm-started from the
 * previous iteration.  densities: (nx-1)(ny-1); compliance: max_iter + 1
 * (index 0 the uniform start); iterations: max_iter; stats (may be NULL):
 * PCG iterations and device seconds summed over the solves, worst residual. */
int tfm_topopt_simp(tfm_ctx *ctx, const tfm_topopt_params *prm, const double *f, int64_t n_bc,
                    const int64_t *bc_dof, const double *bc_val, const tfm_solve_opts *opts,
                    double *densities, double *compliance, int *iterations,
                    tfm_solve_stats *stats);
/* Batching form of topopt_simp (no reference counterpart): P designs on the
 * same grid and supports, design p with its own load f[p*2*nx*ny ...] and
 * volume fraction vol_frac[p] (NULL: prm->vol_frac), one assembly, one
 * batched PCG and one filter/OC pass per iteration for all of them.
 * densities [P][(nx-1)(ny-1)], compliance [P][max_iter+1], iterations and
 * stats [P]. */
int tfm_topopt_simp_batch(tfm_ctx *ctx, int64_t P, const tfm_topopt_params *prm, const double *f,
                          const double *vol_frac, int64_t n_bc, const int64_t *bc_dof,
                          const double *bc_val, const tfm_solve_opts *opts, double *densities,
                          double *compliance, int *iterations, tfm_solve_stats *stats);

/* ---- eigenpairs (crates/tpt-fem-eigen/src/lib.rs) ------------------------ */
enum { TFM_EIG_SMALLEST = 0, TFM_EIG_LARGEST = 1 }; /* EigWhich (eigen/src/lib.rs:125-131) */
/* lanczos_eigs(coo, num, which, lanczos_dim) (eigen/src/lib.rs:223-268) on the
 * square CSR A (n = nrows): a Lanczos basis of dimension lanczos_dim from e_0
 * with full re-orthogonalisation, built with device SpMV and dot products,
 * the tridiagonal solved by the reference's cyclic Jacobi.  evals [num],
 * evecs [num][n] (unit Ritz vectors; may be NULL); *count = min(num, basis
 * size), the basis stopping early at an invariant subspace (beta < 1e-12). */
int tfm_lanczos_eigs(const tfm_matrix *A, int num, int which, int lanczos_dim, double *evals,
                     double *evecs, int *count);

/* ---- time integration (crates/tpt-fem-dynamic/src/lib.rs) --------------- */
/* CachedSystem (dynamic/src/lib.rs:133-166) is the three device matrices M, C,
 * K (C may be the empty 1 x 1 CSR); every column must index a DOF (< n).
 * The load is f(t_k) = fscale[k] * f (fscale NULL: 1), k = 0..nsteps;
 * history: host [nsteps + 1][n] displacements, row 0 = u0. */
/* central_difference (dynamic/src/lib.rs:289-377): lumped (row-sum) mass, the
 * CFL guard 2 / sqrt(max_i sum_j |K_ij| / m_i) -- TFM_E_CFL with *critical
 * set when dt exceeds it -- then the explicit steps, bit-identical. */
int tfm_central_difference(const tfm_matrix *M, const tfm_matrix *C, const tfm_matrix *K, int64_t n,
                           const double *u0, const double *v0, const double *f,
                           const double *fscale, double dt, int nsteps, double *history,
                           double *critical);
/* newmark (dynamic/src/lib.rs:194-263): a0 from M a0 = f0 - C v0 - K u0 and
 * every step's K_hat u = rhs solved by FP64 PCG (opts; NULL: rtol 1e-12),
 * warm-started from the previous step, where the reference factors densely.
 * stats (may be NULL): PCG iterations and device seconds over the steps. */
int tfm_newmark(const tfm_matrix *M, const tfm_matrix *C, const tfm_matrix *K, int64_t n,
                const double *u0, const double *v0, const double *f, const double *fscale,
                double dt, double beta, double gamma, int nsteps, const tfm_solve_opts *opts,
                double *history, tfm_solve_stats *stats);

/* ---- results ------------------------------------------------------------- */
/* write_vtk / write_vtk_with_data (io-vtk/src/lib.rs:117-183): the mesh and
 * n_fields per-node scalar fields (values[f]: host, n_nodes; names without
 * blanks) as the legacy BINARY (big-endian) unstructured-grid file vtkio
 * exports; the byte-swapped blocks are built on the device.  I/O failures
 * return TFM_E_INVALID with VtkError::Io's text. */
int tfm_write_vtk(const tfm_mesh *mesh, int n_fields, const char *const *names,
                  const double *const *values, const char *path);

/* ---- drivers ------------------------------------------------------------ */
/* solve_poisson (thermal/src/lib.rs:206-240) with constant source f, no
 * Neumann/Robin terms. */
int tfm_solve_poisson(tfm_ctx *ctx, const tfm_mesh *mesh, double kappa, int quad_order,
                      double f, int64_t n_bc, const int64_t *bc_dof, const double *bc_val,
                      const tfm_solve_opts *opts, double *u);
/* solve_poisson with Neumann and Robin terms (thermal/src/lib.rs:206-240):
 * constant flux *neumann_g and coefficient *robin_h (NULL: None). */
int tfm_solve_poisson_nr(tfm_ctx *ctx, const tfm_mesh *mesh, double kappa, int quad_order, double f,
                         const double *neumann_g, const double *robin_h, int64_t n_bc,
                         const int64_t *bc_dof, const double *bc_val, const tfm_solve_opts *opts,
                         double *u);
/* solve_poisson (thermal/src/lib.rs:206-240) with closure-valued source,
 * Neumann flux and Robin coefficient given at their points: sq [n_elem][nq]
 * (tfm_mesh_quad_points), neumann_gq / robin_hq [count] at
 * tfm_mesh_boundary_points(mesh, 2, ...) (FACE_QUAD_ORDER), NULL: None. */
int tfm_solve_poisson_qp(tfm_ctx *ctx, const tfm_mesh *mesh, double kappa, int quad_order,
                         const double *sq, const double *neumann_gq, const double *robin_hq,
                         int64_t n_bc, const int64_t *bc_dof, const double *bc_val,
                         const tfm_solve_opts *opts, double *u);
/* solve_elasticity (elasticity/src/lib.rs:355-381) with the body-force
 * closure's values bq [n_elem][nq][dim] at tfm_mesh_quad_points. */
int tfm_solve_elasticity_qp(tfm_ctx *ctx, const tfm_mesh *mesh, int model, double young, double nu,
                            int quad_order, const double *bq, int64_t n_bc, const int64_t *bc_dof,
                            const double *bc_val, const tfm_solve_opts *opts, double *u);
/* solve_elasticity (elasticity/src/lib.rs:355-381) with constant body force. */
int tfm_solve_elasticity(tfm_ctx *ctx, const tfm_mesh *mesh, int model, double young,
                         double nu, int quad_order, const double *body, int64_t n_bc,
                         const int64_t *bc_dof, const double *bc_val,
                         const tfm_solve_opts *opts, double *u);

#ifdef __cplusplus
}
#endif
#endif
