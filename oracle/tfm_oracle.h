/*
 * tfm_oracle.h -- CPU oracle for the tpt-fem element-assembly and
 * reduced-system solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, strict-IEEE restatement of
 * the reference algorithms (built with -ffp-contract=off, no fast-math, so
 * every expression rounds exactly as the reference's Rust f64 code does).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product library never links it.
 *
 * Conventions: indices are int64_t (the reference uses usize); matrices are
 * row-major; element connectivity is element-major.  Cell codes follow the
 * reference CellType order (crates/tpt-fem-mesh/src/lib.rs:39-52):
 *   0 Line2, 1 Tri3, 2 Quad4, 3 Tet4, 4 Hex8.
 * Elastic models follow crates/tpt-fem-elasticity/src/lib.rs:58-67:
 *   0 BarAxial, 1 PlaneStress, 2 PlaneStrain, 3 Continuum3D.
 */
#ifndef TFM_ORACLE_H
#define TFM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LINE = 0, ORC_TRI = 1, ORC_QUAD = 2, ORC_TET = 3, ORC_HEX = 4 };
enum { ORC_BAR = 0, ORC_PSTRESS = 1, ORC_PSTRAIN = 2, ORC_3D = 3 };

/* ---- quadrature (crates/tpt-fem-quadrature/src/lib.rs) ---- */
int orc_gauss_legendre(int order, double *pts, double *wts);
int orc_cell_quad(int cell, int order, double *pts /* n x 3 */, double *wts);

/* ---- reference elements + Map (crates/tpt-fem-element/src/lib.rs) ---- */
int orc_cell_nodes(int cell);
int orc_cell_dim(int cell);
void orc_shape(int cell, const double *xi, double *n);
void orc_grad(int cell, const double *xi, double *dn /* nn x dim */);
double orc_mat_det(int n, const double *m);
void orc_mat_inv(int n, const double *m, double *out);
double orc_map(int nn, int pdim, const double *phys, int n, const double *lgrad,
               double *jac, double *inv);

/* ---- physics element operators ---- */
void orc_poisson_element_matrix(int cell, int pdim, const double *phys,
                                double kappa, int order, double *k);
void orc_poisson_source_vector(int cell, int pdim, const double *phys,
                               double f_const, int order, double *f);
/* Closure-valued loads: the quadrature-point positions a closure is called
 * at, and the element vectors from the closure's per-point values. */
int orc_qp_coords(int cell, int pdim, const double *phys, int order, double *x);
void orc_poisson_source_vector_qp(int cell, int pdim, const double *phys, const double *sq, int order,
                                  double *f);
void orc_elasticity_body_vector(int cell, int pdim, const double *phys, const double *bq, int order,
                                double *f);
int orc_elasticity_element_matrix(int cell, int pdim, const double *phys,
                                  int model, double young, double nu,
                                  int order, double *k);
void orc_simp_ke_unit(double nu, double h, double *ke /* 8x8 */);

/* ---- mesh generation and selectors ---- */
void orc_box_mesh(const double *mn, const double *mx, int64_t nx, int64_t ny,
                  int64_t nz, double *coords, int64_t *conn);
void orc_grid(int64_t nx, int64_t ny, double h, double *coords, int64_t *conn);
int64_t orc_nodes_on_plane(int64_t nnodes, int pdim, const double *coords,
                           int axis, double coord, double tol, int64_t *out);

/* ---- assembly / sparse ---- */
/* Generic element operator description used by the oracle's assemble. */
typedef struct {
  int cell;        /* ORC_* */
  int pdim;        /* coordinate components per node */
  int kind;        /* 0 poisson, 1 elasticity, 2 simp-q4, 3 consistent mass (density in coef) */
  int model;       /* elastic model */
  int order;       /* quadrature order */
  double coef;     /* kappa or young when coef_e == NULL */
  const double *coef_e; /* per-element kappa / young, or NULL */
  double nu;
  /* SIMP: rho per element, penal, e0, emin, and unit ke (8x8) */
  const double *rho;
  double penal, e0, emin;
  const double *ke0;
} orc_op;

void orc_mass_element_matrix(int cell, int pdim, const double *phys, double density, int order,
                             double *k);
int64_t orc_lumped_mass(int64_t nrows, const int64_t *rp, const double *vals, int dim, double *diag);
int64_t orc_assemble(int64_t nnodes, const double *coords, int64_t nelem,
                     const int64_t *conn, const orc_op *op, int dpn,
                     int64_t *rows, int64_t *cols, double *vals);
int64_t orc_coo_dims(int64_t n, const int64_t *rows, const int64_t *cols,
                     int64_t *nrows, int64_t *ncols);
int64_t orc_to_csr(int64_t n, const int64_t *rows, const int64_t *cols,
                   const double *vals, int64_t nrows, int64_t *row_ptrs,
                   int64_t *col_ind, double *values);
void orc_csr_matvec(int64_t nrows, const int64_t *row_ptrs,
                    const int64_t *col_ind, const double *values,
                    int64_t nx, const double *x, double *y);
int64_t orc_reduce_system(int64_t nrows_csr, const int64_t *row_ptrs,
                          const int64_t *col_ind, const double *values,
                          int64_t n, const double *rhs, int64_t nbc,
                          const int64_t *bc_dof, const double *bc_val,
                          int64_t *kr, int64_t *kc, double *kv, double *fred,
                          int64_t *free_out, int64_t *n_free);
int orc_dense_solve(int64_t n, double *a_colmajor, double *b);
int orc_solve_coo(int64_t ntrip, const int64_t *rows, const int64_t *cols,
                  const double *vals, int64_t nrhs_len, double *b_inout);
int orc_solve_with_dirichlet(int64_t ntrip, const int64_t *rows,
                             const int64_t *cols, const double *vals,
                             int64_t n, const double *rhs, int64_t nbc,
                             const int64_t *bc_dof, const double *bc_val,
                             double *u);

/* ---- boundary loads ---- */
int64_t orc_boundary_faces(int64_t nelem, int cell, const int64_t *conn,
                           int64_t *out_eid, int64_t *out_face);
/* Scalar constant flux tiled across dpn components (apply_neumann_order). */
void orc_apply_neumann(int64_t nnodes, int pdim, const double *coords,
                       int64_t nelem, int cell, const int64_t *conn, int dpn,
                       double g, int order, double *rhs);
/* Per-component traction on the plane x[axis] == coord (within tol):
 * one scalar pass per nonzero component, interleaved into rhs. */
void orc_plane_traction(int64_t nnodes, int pdim, const double *coords,
                        int64_t nelem, int cell, const int64_t *conn, int dpn,
                        int axis, double coord, double tol, const double *t,
                        int order, double *rhs);

/* Closure-valued boundary terms: the points (position, outward unit normal)
 * a flux / Robin closure is called at, and the passes from its values. */
int64_t orc_boundary_points(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                            const int64_t *conn, int order, double *x, double *nrm);
void orc_apply_neumann_qp(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                          const int64_t *conn, int dpn, const double *gq, int order, double *rhs);
int64_t orc_apply_robin_qp(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                           const int64_t *conn, int dpn, const double *hq, int order, int64_t *rows,
                           int64_t *cols, double *vals);

/* ---- drivers ---- */
void orc_poisson_rhs(int64_t nnodes, int pdim, const double *coords,
                     int64_t nelem, int cell, const int64_t *conn,
                     double f_const, int order, double *rhs);
void orc_body_rhs(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                  const int64_t *conn, const double *bq, int order, double *rhs);
void orc_source_rhs_qp(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                       const int64_t *conn, const double *sq, int order, double *rhs);
int orc_solve_elasticity(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                         const int64_t *conn, int model, double young, double nu, int order,
                         const double *bq, int64_t nbc, const int64_t *bc_dof, const double *bc_val,
                         double *u);
int orc_solve_poisson(int64_t nnodes, int pdim, const double *coords,
                      int64_t nelem, int cell, const int64_t *conn,
                      double kappa, int order, double f_const, int64_t nbc,
                      const int64_t *bc_dof, const double *bc_val, double *u);

/* ---- SIMP topology optimization (crates/tpt-fem-topopt/src/lib.rs) ---- */
double orc_topopt_quad_form(const double *ke0, const double *u, const int64_t *nodes);
void orc_topopt_filter(int64_t nx, int64_t ny, double h, const double *raw, double radius,
                       double *out);
void orc_topopt_oc(int64_t n, const double *rho, const double *sens, double vol_frac,
                   double move_limit, double *out);
int orc_topopt_simp(int64_t nx, int64_t ny, double h, double e0, double nu, double vol_frac,
                    double penal, double radius, int max_iter, double move_limit,
                    const double *f, int64_t nbc, const int64_t *bc_dof, const double *bc_val,
                    double *rho_out, double *compliance_out, int *iterations);

/* Deterministic input generators shared with the product's generators. */
uint64_t orc_splitmix64(uint64_t *state);
double orc_uniform01(uint64_t *state);
void orc_jitter_interior(int64_t nx, int64_t ny, int64_t nz, double h_frac,
                         uint64_t seed, double *coords);
void orc_loguniform(int64_t n, double lo_exp10, double hi_exp10, uint64_t seed,
                    double *out);

#ifdef __cplusplus
}
#endif

/* ---- Lanczos eigenpairs (crates/tpt-fem-eigen/src/lib.rs) ---- */
void orc_eig_matvec(int64_t n, const int64_t *rp, const int64_t *ci, const double *v,
                    const double *x, double *y);
void orc_eig_jacobi(int m, double *a, int max_sweeps, double *eig, double *q);
int orc_lanczos_eigs(int64_t n, const int64_t *rp, const int64_t *ci, const double *v, int num,
                     int which, int lanczos_dim, double *evals, double *evecs);

/* ---- time integration (crates/tpt-fem-dynamic/src/lib.rs) ---- */
typedef struct { int64_t nrows; const int64_t *rp, *ci; const double *v; } orc_csr;
int orc_central_difference(int64_t n, orc_csr M, orc_csr D, orc_csr K, const double *u0,
                           const double *v0, const double *F, const double *fscale, double dt,
                           int64_t nsteps, double *history, double *critical);
int orc_newmark(int64_t n, orc_csr M, orc_csr D, orc_csr K, const double *u0, const double *v0,
                const double *F, const double *fscale, double dt, double beta, double gamma,
                int64_t nsteps, double *history);

/* ---- Robin terms (crates/tpt-fem-assembly/src/lib.rs:684-753) ---- */
int64_t orc_apply_robin(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                        const int64_t *conn, int dpn, double hconst, int axis, double coord,
                        double tol, int order, int64_t *rows, int64_t *cols, double *vals);
#endif
