/*
 * tfm_oracle.c -- CPU oracle (test infrastructure, never shipped) restating
 * the reference tpt-fem assembly and solve path in strict-IEEE C.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (oracle/Makefile).  Rust
 * f64 arithmetic never fuses multiply-add and evaluates left to right, so
 * every expression below keeps the reference's operation order; that is what
 * makes the element matrices (and the exact-zero skip decisions) bit-identical
 * to the reference's.
 *
 * Parity pins (see tests/test_oracle.py): every function here is checked
 * against the reference's own unit-test known answers, closed forms or
 * independent library routines; see the per-function comments.
 */
#include "tfm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Quadrature: crates/tpt-fem-quadrature/src/lib.rs:118-162 (Gauss-Legendre),
 * :180-207 (tensor rules), :223-260 (triangle), :279-289 (tet Degree2).     */

int orc_gauss_legendre(int order, double *p, double *w) {
  switch (order) {
  case 1:
    p[0] = 0.0; w[0] = 2.0;
    return 1;
  case 2:
    p[0] = -1.0 / sqrt(3.0); p[1] = 1.0 / sqrt(3.0);
    w[0] = 1.0; w[1] = 1.0;
    return 2;
  case 3:
    p[0] = 0.0; p[1] = -sqrt(3.0 / 5.0); p[2] = sqrt(3.0 / 5.0);
    w[0] = 8.0 / 9.0; w[1] = 5.0 / 9.0; w[2] = 5.0 / 9.0;
    return 3;
  case 4:
    p[0] = -0.3399810435848563; p[1] = 0.3399810435848563;
    p[2] = -0.8611363115940526; p[3] = 0.8611363115940526;
    w[0] = 0.6521451548625461; w[1] = 0.6521451548625461;
    w[2] = 0.3478548451374539; w[3] = 0.3478548451374539;
    return 4;
  case 5:
    p[0] = 0.0; p[1] = -0.5384693101056831; p[2] = 0.5384693101056831;
    p[3] = -0.9061798459386640; p[4] = 0.9061798459386640;
    w[0] = 0.5688888888888889; w[1] = 0.4786286704993665;
    w[2] = 0.4786286704993665; w[3] = 0.2369268850561891;
    w[4] = 0.2369268850561891;
    return 5;
  default:
    return 0;
  }
}

/* cell_quad: crates/tpt-fem-thermal/src/lib.rs:29-92 (identical selection in
 * crates/tpt-fem-elasticity/src/lib.rs:117-183).  Tets always use Degree2 and
 * triangles Degree2, whatever `order` says.  Points are stored with stride 3. */
int orc_cell_quad(int cell, int order, double *pts, double *wts) {
  double gp[5], gw[5];
  int n, q = 0;
  switch (cell) {
  case ORC_LINE:
    n = orc_gauss_legendre(order, gp, gw);
    for (int i = 0; i < n; i++) {
      pts[3 * i] = gp[i]; pts[3 * i + 1] = 0.0; pts[3 * i + 2] = 0.0;
      wts[i] = gw[i];
    }
    return n;
  case ORC_TRI: {
    const double a = 2.0 / 3.0, b = 1.0 / 6.0;
    const double P[3][2] = {{a, b}, {b, a}, {b, b}};
    for (int i = 0; i < 3; i++) {
      pts[3 * i] = P[i][0]; pts[3 * i + 1] = P[i][1]; pts[3 * i + 2] = 0.0;
      wts[i] = 1.0 / 6.0;
    }
    return 3;
  }
  case ORC_QUAD:
    n = orc_gauss_legendre(order, gp, gw);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        pts[3 * q] = gp[i]; pts[3 * q + 1] = gp[j]; pts[3 * q + 2] = 0.0;
        wts[q] = gw[i] * gw[j];
        q++;
      }
    return q;
  case ORC_TET: {
    const double a = 0.5854101966249685, b = 0.1381966011250105;
    const double P[4][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
    for (int i = 0; i < 4; i++) {
      pts[3 * i] = P[i][0]; pts[3 * i + 1] = P[i][1]; pts[3 * i + 2] = P[i][2];
      wts[i] = 1.0 / 24.0;
    }
    return 4;
  }
  case ORC_HEX:
    n = orc_gauss_legendre(order, gp, gw);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
          pts[3 * q] = gp[i]; pts[3 * q + 1] = gp[j]; pts[3 * q + 2] = gp[k];
          wts[q] = gw[i] * gw[j] * gw[k];
          q++;
        }
    return q;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Reference elements: crates/tpt-fem-element/src/lib.rs:90-221.            */

int orc_cell_nodes(int cell) {
  static const int nn[5] = {2, 3, 4, 4, 8};
  return (cell >= 0 && cell < 5) ? nn[cell] : 0;
}
int orc_cell_dim(int cell) {
  static const int dd[5] = {1, 2, 2, 3, 3};
  return (cell >= 0 && cell < 5) ? dd[cell] : 0;
}

static const double QS[4] = {-1.0, 1.0, 1.0, -1.0};
static const double QT[4] = {-1.0, -1.0, 1.0, 1.0};
static const double HS[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
static const double HT[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
static const double HU[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};

void orc_shape(int cell, const double *xi, double *n) {
  double x = xi[0], y = xi[1], z = xi[2];
  switch (cell) {
  case ORC_LINE:
    n[0] = (1.0 - x) / 2.0; n[1] = (1.0 + x) / 2.0;
    break;
  case ORC_TRI:
    n[0] = 1.0 - x - y; n[1] = x; n[2] = y;
    break;
  case ORC_QUAD:
    for (int i = 0; i < 4; i++)
      n[i] = 0.25 * (1.0 + QS[i] * x) * (1.0 + QT[i] * y);
    break;
  case ORC_TET:
    n[0] = 1.0 - x - y - z; n[1] = x; n[2] = y; n[3] = z;
    break;
  case ORC_HEX:
    for (int i = 0; i < 8; i++)
      n[i] = 0.125 * (1.0 + HS[i] * x) * (1.0 + HT[i] * y) * (1.0 + HU[i] * z);
    break;
  }
}

void orc_grad(int cell, const double *xi, double *g) {
  double x = xi[0], y = xi[1], z = xi[2];
  switch (cell) {
  case ORC_LINE:
    g[0] = -0.5; g[1] = 0.5;
    break;
  case ORC_TRI:
    g[0] = -1.0; g[1] = -1.0; g[2] = 1.0; g[3] = 0.0; g[4] = 0.0; g[5] = 1.0;
    break;
  case ORC_QUAD:
    for (int i = 0; i < 4; i++) {
      g[2 * i] = 0.25 * QS[i] * (1.0 + QT[i] * y);
      g[2 * i + 1] = 0.25 * (1.0 + QS[i] * x) * QT[i];
    }
    break;
  case ORC_TET: {
    static const double G[12] = {-1.0, -1.0, -1.0, 1.0, 0.0, 0.0,
                                 0.0,  1.0,  0.0,  0.0, 0.0, 1.0};
    memcpy(g, G, sizeof G);
    break;
  }
  case ORC_HEX:
    for (int i = 0; i < 8; i++) {
      g[3 * i] = 0.125 * HS[i] * (1.0 + HT[i] * y) * (1.0 + HU[i] * z);
      g[3 * i + 1] = 0.125 * (1.0 + HS[i] * x) * HT[i] * (1.0 + HU[i] * z);
      g[3 * i + 2] = 0.125 * (1.0 + HS[i] * x) * (1.0 + HT[i] * y) * HU[i];
    }
    break;
  }
}

/* mat_det / mat_inv: crates/tpt-fem-element/src/lib.rs:624-680.            */
double orc_mat_det(int n, const double *m) {
  switch (n) {
  case 1: return m[0];
  case 2: return m[0] * m[3] - m[1] * m[2];
  case 3:
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
  }
  abort();
}

void orc_mat_inv(int n, const double *m, double *o) {
  if (n == 1) {
    o[0] = 1.0 / m[0];
  } else if (n == 2) {
    double det = m[0] * m[3] - m[1] * m[2];
    double inv = 1.0 / det;
    o[0] = m[3] * inv; o[1] = -m[1] * inv; o[2] = -m[2] * inv; o[3] = m[0] * inv;
  } else if (n == 3) {
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                 m[2] * (m[3] * m[7] - m[4] * m[6]);
    double inv = 1.0 / det;
    o[0] = (m[4] * m[8] - m[5] * m[7]) * inv;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
    o[3] = (m[5] * m[6] - m[3] * m[8]) * inv;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
    o[6] = (m[3] * m[7] - m[4] * m[6]) * inv;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
  } else {
    abort();
  }
}

/* Map::from_nodes_and_grad: crates/tpt-fem-element/src/lib.rs:702-724.
 * J[p][r] accumulated node by node from 0.0; returns the SIGNED det.       */
double orc_map(int nn, int pdim, const double *phys, int n, const double *lg,
               double *jac, double *inv) {
  for (int i = 0; i < n * n; i++) jac[i] = 0.0;
  for (int a = 0; a < nn; a++)
    for (int p = 0; p < n; p++)
      for (int r = 0; r < n; r++)
        jac[p * n + r] += phys[a * pdim + p] * lg[a * n + r];
  double det = orc_mat_det(n, jac);
  orc_mat_inv(n, jac, inv);
  return det;
}

/* Map::physical_grad: crates/tpt-fem-element/src/lib.rs:729-740.          */
static void phys_grad(int n, const double *inv, const double *loc, double *out) {
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int r = 0; r < n; r++) s += loc[r] * inv[r * n + i];
    out[i] = s;
  }
}

/* Rust's f64 Iterator::sum folds from -0.0 (crates/tpt-fem-thermal/src/lib.rs:125). */
static double dotn(int n, const double *a, const double *b) {
  double s = -0.0;
  for (int i = 0; i < n; i++) s = s + a[i] * b[i];
  return s;
}

/* ------------------------------------------------------------------------ */
/* poisson_element_matrix: crates/tpt-fem-thermal/src/lib.rs:132-160.      */
void orc_poisson_element_matrix(int cell, int pdim, const double *phys,
                                double kappa, int order, double *k) {
  double qp[125 * 3], qw[125], lg[8 * 3], jac[9], inv[9], g[8 * 3];
  int nn = orc_cell_nodes(cell), dim = orc_cell_dim(cell);
  int nq = orc_cell_quad(cell, order, qp, qw);
  for (int i = 0; i < nn * nn; i++) k[i] = 0.0;
  for (int q = 0; q < nq; q++) {
    orc_grad(cell, qp + 3 * q, lg);
    double det = fabs(orc_map(nn, pdim, phys, dim, lg, jac, inv));
    for (int a = 0; a < nn; a++) phys_grad(dim, inv, lg + a * dim, g + a * dim);
    for (int i = 0; i < nn; i++)
      for (int j = 0; j < nn; j++)
        k[i * nn + j] += qw[q] * kappa * dotn(dim, g + i * dim, g + j * dim) * det;
  }
}

/* poisson_source_vector: crates/tpt-fem-thermal/src/lib.rs:163-196 with a
 * constant source s(x) = f_const (the quadrature-point position is still
 * formed, as in the reference, but a constant source ignores it).          */
void orc_poisson_source_vector(int cell, int pdim, const double *phys,
                               double f_const, int order, double *f) {
  double qp[125 * 3], qw[125], ns[8], lg[8 * 3], jac[9], inv[9];
  int nn = orc_cell_nodes(cell), dim = orc_cell_dim(cell);
  int nq = orc_cell_quad(cell, order, qp, qw);
  for (int i = 0; i < nn; i++) f[i] = 0.0;
  for (int q = 0; q < nq; q++) {
    orc_shape(cell, qp + 3 * q, ns);
    orc_grad(cell, qp + 3 * q, lg);
    double det = fabs(orc_map(nn, pdim, phys, dim, lg, jac, inv));
    double s = f_const;
    for (int i = 0; i < nn; i++) f[i] += qw[q] * s * ns[i] * det;
  }
}

/* Physical quadrature-point positions of one element, formed as
 * poisson_source_vector and elasticity_body_vector form the argument of the
 * closure (crates/tpt-fem-thermal/src/lib.rs:180-185,
 * crates/tpt-fem-elasticity/src/lib.rs:331-336): x[d] = sum_k ns[k]*phys[k][d]
 * accumulated from 0.0 in node order.  x: nq x pdim.  Returns nq.          */
int orc_qp_coords(int cell, int pdim, const double *phys, int order, double *x) {
  double qp[125 * 3], qw[125], ns[8];
  int nn = orc_cell_nodes(cell);
  int nq = orc_cell_quad(cell, order, qp, qw);
  for (int q = 0; q < nq; q++) {
    orc_shape(cell, qp + 3 * q, ns);
    for (int d = 0; d < pdim; d++) x[q * pdim + d] = 0.0;
    for (int k = 0; k < nn; k++)
      for (int d = 0; d < pdim; d++) x[q * pdim + d] += ns[k] * phys[k * pdim + d];
  }
  return nq;
}

/* poisson_source_vector (crates/tpt-fem-thermal/src/lib.rs:163-196) with the
 * closure's value s(x_q) given per quadrature point: f[i] += w*s*ns[i]*det. */
void orc_poisson_source_vector_qp(int cell, int pdim, const double *phys, const double *sq, int order,
                                  double *f) {
  double qp[125 * 3], qw[125], ns[8], lg[8 * 3], jac[9], inv[9];
  int nn = orc_cell_nodes(cell), dim = orc_cell_dim(cell);
  int nq = orc_cell_quad(cell, order, qp, qw);
  for (int i = 0; i < nn; i++) f[i] = 0.0;
  for (int q = 0; q < nq; q++) {
    orc_shape(cell, qp + 3 * q, ns);
    orc_grad(cell, qp + 3 * q, lg);
    double det = fabs(orc_map(nn, pdim, phys, dim, lg, jac, inv));
    for (int i = 0; i < nn; i++) f[i] += qw[q] * sq[q] * ns[i] * det;
  }
}

/* elasticity_body_vector (crates/tpt-fem-elasticity/src/lib.rs:312-348) with
 * the closure's value b(x_q) (dim components) given per quadrature point:
 * f[i*dim + a] += w * ns[i] * b[a] * det, points in rule order.             */
void orc_elasticity_body_vector(int cell, int pdim, const double *phys, const double *bq, int order,
                                double *f) {
  double qp[125 * 3], qw[125], ns[8], lg[8 * 3], jac[9], inv[9];
  int nn = orc_cell_nodes(cell), dim = orc_cell_dim(cell);
  int nq = orc_cell_quad(cell, order, qp, qw);
  for (int i = 0; i < nn * dim; i++) f[i] = 0.0;
  for (int q = 0; q < nq; q++) {
    orc_shape(cell, qp + 3 * q, ns);
    orc_grad(cell, qp + 3 * q, lg);
    double det = fabs(orc_map(nn, pdim, phys, dim, lg, jac, inv));
    for (int i = 0; i < nn; i++)
      for (int a = 0; a < dim; a++) f[i * dim + a] += qw[q] * ns[i] * bq[q * dim + a] * det;
  }
}

/* constitutive: crates/tpt-fem-elasticity/src/lib.rs:197-238.             */
static int constitutive(int model, double e, double nu, int dim, double d[6][6]) {
  memset(d, 0, sizeof(double) * 36);
  if (model == ORC_BAR && dim == 1) {
    d[0][0] = e;
  } else if (model == ORC_PSTRESS && dim == 2) {
    double c = e / (1.0 - nu * nu);
    d[0][0] = c; d[0][1] = c * nu; d[1][0] = c * nu; d[1][1] = c;
    d[2][2] = c * (1.0 - nu) / 2.0;
  } else if (model == ORC_PSTRAIN && dim == 2) {
    double c = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
    d[0][0] = c * (1.0 - nu); d[0][1] = c * nu;
    d[1][0] = c * nu; d[1][1] = c * (1.0 - nu);
    d[2][2] = c * (1.0 - 2.0 * nu) / 2.0;
  } else if (model == ORC_3D && dim == 3) {
    double c = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double m = (1.0 - nu) * c;
    double dd = (1.0 - 2.0 * nu) / 2.0 * c;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) d[i][j] = (i == j) ? m : c * nu;
    d[3][3] = dd; d[4][4] = dd; d[5][5] = dd;
  } else {
    return -1;
  }
  return 0;
}

/* b_matrix (Voigt xx,yy,zz,xy,yz,zx): crates/tpt-fem-elasticity/src/lib.rs:242-258. */
static void b_matrix(const double *g, int dim, double b[6][3]) {
  memset(b, 0, sizeof(double) * 18);
  if (dim == 1) {
    b[0][0] = g[0];
  } else if (dim == 2) {
    b[0][0] = g[0]; b[1][1] = g[1]; b[2][0] = g[1]; b[2][1] = g[0];
  } else {
    b[0][0] = g[0]; b[1][1] = g[1]; b[2][2] = g[2];
    b[3][0] = g[1]; b[3][1] = g[0];
    b[4][1] = g[2]; b[4][2] = g[1];
    b[5][0] = g[2]; b[5][2] = g[0];
  }
}

/* elasticity_element_matrix: crates/tpt-fem-elasticity/src/lib.rs:262-309. */
int orc_elasticity_element_matrix(int cell, int pdim, const double *phys,
                                  int model, double young, double nu,
                                  int order, double *k) {
  double qp[125 * 3], qw[125], lg[8 * 3], jac[9], inv[9], g[8 * 3];
  double d[6][6], b[8][6][3];
  int nn = orc_cell_nodes(cell), dim = orc_cell_dim(cell);
  int nstr = dim == 1 ? 1 : (dim == 2 ? 3 : 6);
  if (constitutive(model, young, nu, dim, d) != 0) return -1;
  int nq = orc_cell_quad(cell, order, qp, qw);
  int nd = nn * dim;
  for (int i = 0; i < nd * nd; i++) k[i] = 0.0;
  for (int q = 0; q < nq; q++) {
    orc_grad(cell, qp + 3 * q, lg);
    double det = fabs(orc_map(nn, pdim, phys, dim, lg, jac, inv));
    for (int a = 0; a < nn; a++) {
      phys_grad(dim, inv, lg + a * dim, g + a * dim);
      b_matrix(g + a * dim, dim, b[a]);
    }
    for (int i = 0; i < nn; i++)
      for (int j = 0; j < nn; j++)
        for (int a = 0; a < dim; a++)
          for (int bb = 0; bb < dim; bb++) {
            double s = 0.0;
            for (int s_ = 0; s_ < nstr; s_++)
              for (int t = 0; t < nstr; t++) s += b[i][s_][a] * d[s_][t] * b[j][t][bb];
            k[(i * dim + a) * nd + j * dim + bb] += qw[q] * det * s;
          }
  }
  return 0;
}

/* element_stiffness_unit: crates/tpt-fem-topopt/src/lib.rs:166-217.       */
void orc_simp_ke_unit(double nu, double h, double *ke) {
  double c = 1.0 / (1.0 - nu * nu);
  double d[3][3] = {{c, c * nu, 0.0}, {c * nu, c, 0.0}, {0.0, 0.0, c * (1.0 - nu) / 2.0}};
  double gp[2], gw[2], lg[8], bm[3][8], db[3][8];
  orc_gauss_legendre(2, gp, gw);
  double jdet = h * h / 4.0;
  double scale = 2.0 / h;
  for (int i = 0; i < 64; i++) ke[i] = 0.0;
  for (int a = 0; a < 2; a++)
    for (int bq = 0; bq < 2; bq++) {
      double xi[3] = {gp[a], gp[bq], 0.0};
      double w = gw[a] * gw[bq];
      orc_grad(ORC_QUAD, xi, lg);
      memset(bm, 0, sizeof bm);
      for (int n = 0; n < 4; n++) {
        double nx_ = scale * lg[2 * n], ny_ = scale * lg[2 * n + 1];
        bm[0][2 * n] = nx_; bm[1][2 * n + 1] = ny_;
        bm[2][2 * n] = ny_; bm[2][2 * n + 1] = nx_;
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 8; j++) {
          double s = 0.0;
          for (int kk = 0; kk < 3; kk++) s += d[i][kk] * bm[kk][j];
          db[i][j] = s;
        }
      for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
          double s = 0.0;
          for (int kk = 0; kk < 3; kk++) s += bm[kk][i] * db[kk][j];
          ke[i * 8 + j] += w * jdet * s;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* box_mesh + orient3: crates/tpt-fem-mesh-gen/src/lib.rs:61-67, 422-495.   */
static double orient3(const double *a, const double *b, const double *c, const double *d) {
  double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double ad[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  return ab[0] * (ac[1] * ad[2] - ac[2] * ad[1]) - ab[1] * (ac[0] * ad[2] - ac[2] * ad[0]) +
         ab[2] * (ac[0] * ad[1] - ac[1] * ad[0]);
}

static void box_point(const double *mn, const double *mx, int64_t nx, int64_t ny,
                      int64_t nz, int64_t i, int64_t j, int64_t k, double *o) {
  double t = nx > 0 ? (double)i / (double)nx : 0.0;
  double u = ny > 0 ? (double)j / (double)ny : 0.0;
  double w = nz > 0 ? (double)k / (double)nz : 0.0;
  o[0] = mn[0] + t * (mx[0] - mn[0]);
  o[1] = mn[1] + u * (mx[1] - mn[1]);
  o[2] = mn[2] + w * (mx[2] - mn[2]);
}

void orc_box_mesh(const double *mn, const double *mx, int64_t nx, int64_t ny,
                  int64_t nz, double *coords, int64_t *conn) {
  int64_t id = 0;
  for (int64_t i = 0; i <= nx; i++)
    for (int64_t j = 0; j <= ny; j++)
      for (int64_t k = 0; k <= nz; k++) box_point(mn, mx, nx, ny, nz, i, j, k, coords + 3 * id++);
  static const int TETS[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                                 {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
  static const int OFF[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  int64_t e = 0;
  for (int64_t i = 0; i < nx; i++)
    for (int64_t j = 0; j < ny; j++)
      for (int64_t k = 0; k < nz; k++) {
        int64_t ids[8];
        double co[8][3];
        for (int c = 0; c < 8; c++) {
          int64_t ii = i + OFF[c][0], jj = j + OFF[c][1], kk = k + OFF[c][2];
          ids[c] = (ii * (ny + 1) + jj) * (nz + 1) + kk;
          /* corner(): the offset is evaluated at the brick-local 0/1 index,
           * exactly as the reference's closure does (it ignores i, j, k). */
          box_point(mn, mx, nx, ny, nz, OFF[c][0], OFF[c][1], OFF[c][2], co[c]);
        }
        for (int t = 0; t < 6; t++) {
          int64_t tv[4] = {ids[TETS[t][0]], ids[TETS[t][1]], ids[TETS[t][2]], ids[TETS[t][3]]};
          if (orient3(co[TETS[t][0]], co[TETS[t][1]], co[TETS[t][2]], co[TETS[t][3]]) < 0.0) {
            int64_t tmp = tv[1]; tv[1] = tv[2]; tv[2] = tmp;
          }
          memcpy(conn + 4 * e, tv, sizeof tv);
          e++;
        }
      }
}

/* Grid::new: crates/tpt-fem-topopt/src/lib.rs:89-116 (2 coords per node). */
void orc_grid(int64_t nx, int64_t ny, double h, double *coords, int64_t *conn) {
  for (int64_t j = 0; j < ny; j++)
    for (int64_t i = 0; i < nx; i++) {
      coords[2 * (j * nx + i)] = (double)i * h;
      coords[2 * (j * nx + i) + 1] = (double)j * h;
    }
  int64_t e = 0;
  for (int64_t ej = 0; ej < ny - 1; ej++)
    for (int64_t ei = 0; ei < nx - 1; ei++) {
      conn[4 * e] = ej * nx + ei;
      conn[4 * e + 1] = ej * nx + ei + 1;
      conn[4 * e + 2] = (ej + 1) * nx + ei + 1;
      conn[4 * e + 3] = (ej + 1) * nx + ei;
      e++;
    }
}

/* nodes_on_plane: crates/tpt-fem-mesh/src/lib.rs:425-431.                 */
int64_t orc_nodes_on_plane(int64_t nnodes, int pdim, const double *coords, int axis,
                           double coord, double tol, int64_t *out) {
  int64_t m = 0;
  for (int64_t n = 0; n < nnodes; n++) {
    double c = axis < pdim ? coords[n * pdim + axis] : 0.0;
    if (fabs(c - coord) <= tol) out[m++] = n;
  }
  return m;
}

/* elasticity_mass_matrix's element block: crates/tpt-fem-elasticity/src/lib.rs
 * :392-414.  m_ij = sum_q w * det * rho * N_i * N_j (left to right, from 0.0),
 * det = |det J| of Map::from_nodes_and_grad; k is the (nn*dim)^2 matrix whose
 * only non-zeros are the component-diagonal copies m_ij (:418-424 pushes
 * (n_i*dim + a, n_j*dim + a) for a < dim, skipping m_ij == 0).              */
void orc_mass_element_matrix(int cell, int pdim, const double *phys, double density, int order,
                             double *k) {
  double qp[125 * 3], qw[125], ns[8], lg[8 * 3], jac[9], inv[9], m[64];
  int nn = orc_cell_nodes(cell), dim = orc_cell_dim(cell), nd = nn * dim;
  int nq = orc_cell_quad(cell, order, qp, qw);
  for (int i = 0; i < nn * nn; i++) m[i] = 0.0;
  for (int q = 0; q < nq; q++) {
    orc_grad(cell, qp + 3 * q, lg);
    double det = fabs(orc_map(nn, pdim, phys, dim, lg, jac, inv));
    orc_shape(cell, qp + 3 * q, ns);
    for (int i = 0; i < nn; i++)
      for (int j = 0; j < nn; j++) m[i * nn + j] += qw[q] * det * density * ns[i] * ns[j];
  }
  for (int i = 0; i < nd * nd; i++) k[i] = 0.0;
  for (int i = 0; i < nn; i++)
    for (int j = 0; j < nn; j++)
      for (int a = 0; a < dim; a++) k[(i * dim + a) * nd + j * dim + a] = m[i * nn + j];
}

/* elasticity_lumped_mass: crates/tpt-fem-elasticity/src/lib.rs:432-457 on the
 * consistent mass CSR: per node the sum of its dim rows (row order, then CSR
 * order, from 0.0) divided by dim, on each of its dim diagonal entries.
 * Returns the number of diagonal entries, (nrows / dim) * dim.              */
int64_t orc_lumped_mass(int64_t nrows, const int64_t *rp, const double *vals, int dim, double *diag) {
  int64_t nodes = nrows / dim;
  for (int64_t node = 0; node < nodes; node++) {
    double total = 0.0;
    for (int a = 0; a < dim; a++) {
      int64_t row = node * dim + a;
      for (int64_t c = rp[row]; c < rp[row + 1]; c++) total += vals[c];
    }
    double lump = total / (double)dim;
    for (int a = 0; a < dim; a++) diag[node * dim + a] = lump;
  }
  return nodes * dim;
}

/* ------------------------------------------------------------------------ */
/* Element operator dispatch used by orc_assemble.                           */
static int elem_matrix(const orc_op *op, const double *phys, int64_t e, double *k) {
  switch (op->kind) {
  case 0:
    orc_poisson_element_matrix(op->cell, op->pdim, phys,
                               op->coef_e ? op->coef_e[e] : op->coef, op->order, k);
    return orc_cell_nodes(op->cell);
  case 1: {
    double young = op->coef_e ? op->coef_e[e] : op->coef;
    if (orc_elasticity_element_matrix(op->cell, op->pdim, phys, op->model, young, op->nu,
                                      op->order, k) != 0)
      return -1;
    return orc_cell_nodes(op->cell) * orc_cell_dim(op->cell);
  }
  case 2: {
    /* SIMP Young's modulus: crates/tpt-fem-topopt/src/lib.rs:224, with
     * Emin a parameter instead of the hard-coded E_MIN_FACTOR * E0. */
    double e_rho = op->emin + pow(op->rho[e], op->penal) * (op->e0 - op->emin);
    for (int i = 0; i < 64; i++) k[i] = e_rho * op->ke0[i];
    return 8;
  }
  case 3: /* consistent mass, density in coef (or per element in coef_e) */
    orc_mass_element_matrix(op->cell, op->pdim, phys, op->coef_e ? op->coef_e[e] : op->coef, op->order, k);
    return orc_cell_nodes(op->cell) * orc_cell_dim(op->cell);
  }
  return -1;
}

/* assemble / try_assemble: crates/tpt-fem-assembly/src/lib.rs:428-480.
 * Exact-zero entries are skipped, except for the SIMP operator whose own
 * assemble (crates/tpt-fem-topopt/src/lib.rs:220-238) pushes all 64.
 * Returns the triplet count, or -1 if an element operator fails.           */
int64_t orc_assemble(int64_t nnodes, const double *coords, int64_t nelem,
                     const int64_t *conn, const orc_op *op, int dpn,
                     int64_t *rows, int64_t *cols, double *vals) {
  (void)nnodes;
  int nn = orc_cell_nodes(op->cell);
  double phys[8 * 3];
  double *k = (double *)malloc(sizeof(double) * 24 * 24);  /* Hex8 x 3 DOFs */
  int64_t t = 0;
  for (int64_t e = 0; e < nelem; e++) {
    for (int a = 0; a < nn; a++)
      for (int p = 0; p < op->pdim; p++) phys[a * op->pdim + p] = coords[conn[e * nn + a] * op->pdim + p];
    int nd = elem_matrix(op, phys, e, k);
    if (nd < 0) { free(k); return -1; }
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++) {
        double kij = k[i * nd + j];
        if (op->kind != 2 && kij == 0.0) continue;
        rows[t] = conn[e * nn + i / dpn] * dpn + i % dpn;
        cols[t] = conn[e * nn + j / dpn] * dpn + j % dpn;
        vals[t] = kij;
        t++;
      }
  }
  free(k);
  return t;
}

int64_t orc_coo_dims(int64_t n, const int64_t *rows, const int64_t *cols,
                     int64_t *nrows, int64_t *ncols) {
  int64_t mr = 0, mc = 0;
  for (int64_t i = 0; i < n; i++) {
    if (rows[i] > mr) mr = rows[i];
    if (cols[i] > mc) mc = cols[i];
  }
  *nrows = n ? mr + 1 : 1; /* max over rows chained with 0, plus one */
  *ncols = n ? mc + 1 : 1;
  return n;
}

static const int64_t *g_sr, *g_sc;
static int cmp_rc(const void *pa, const void *pb) {
  int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
  if (g_sr[a] != g_sr[b]) return g_sr[a] < g_sr[b] ? -1 : 1;
  if (g_sc[a] != g_sc[b]) return g_sc[a] < g_sc[b] ? -1 : 1;
  return a < b ? -1 : (a > b); /* tie-break on position: a stable sort */
}

/* Coo::to_csr: crates/tpt-fem-sparse/src/lib.rs:122-175.  Stable sort by
 * (row, col), duplicates summed in insertion order starting from the first
 * value, zero sums kept.  Returns nnz.  row_ptrs has nrows+1 entries.       */
int64_t orc_to_csr(int64_t n, const int64_t *rows, const int64_t *cols,
                   const double *vals, int64_t nrows, int64_t *row_ptrs,
                   int64_t *col_ind, double *values) {
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t i = 0; i < n; i++) idx[i] = i;
  g_sr = rows; g_sc = cols;
  qsort(idx, (size_t)n, sizeof(int64_t), cmp_rc);
  for (int64_t r = 0; r <= nrows; r++) row_ptrs[r] = 0;
  int64_t k = 0, nnz = 0;
  while (k < n) {
    int64_t r = rows[idx[k]], c = cols[idx[k]];
    double sum = vals[idx[k]];
    int64_t kk = k + 1;
    while (kk < n && rows[idx[kk]] == r && cols[idx[kk]] == c) {
      sum += vals[idx[kk]];
      kk++;
    }
    col_ind[nnz] = c;
    values[nnz] = sum;
    nnz++;
    row_ptrs[r + 1] += 1;
    k = kk;
  }
  for (int64_t r = 0; r < nrows; r++) row_ptrs[r + 1] += row_ptrs[r];
  free(idx);
  return nnz;
}

/* Csr::matvec: crates/tpt-fem-sparse/src/lib.rs:195-205.  y has
 * max(nrows, nx) entries; rows beyond nrows are zero.                        */
void orc_csr_matvec(int64_t nrows, const int64_t *rp, const int64_t *ci,
                    const double *v, int64_t nx, const double *x, double *y) {
  int64_t ny = nrows > nx ? nrows : nx;
  for (int64_t i = 0; i < ny; i++) y[i] = 0.0;
  for (int64_t r = 0; r < nrows; r++) {
    double s = 0.0;
    for (int64_t c = rp[r]; c < rp[r + 1]; c++) s += v[c] * x[ci[c]];
    y[r] = s;
  }
}

/* reduce_system: crates/tpt-fem-assembly/src/lib.rs:504-535.  The lifting
 * value of a fixed column is the FIRST matching (dof, value) pair in bcs.
 * Returns the kred triplet count, or -1 when the CSR has fewer rows than a
 * free DOF needs (the reference panics on that index).                      */
int64_t orc_reduce_system(int64_t nrows_csr, const int64_t *rp, const int64_t *ci,
                          const double *v, int64_t n, const double *rhs, int64_t nbc,
                          const int64_t *bc_dof, const double *bc_val, int64_t *kr,
                          int64_t *kc, double *kv, double *fred, int64_t *free_out,
                          int64_t *n_free) {
  int64_t ncol_max = n;
  for (int64_t r = 0; r < nrows_csr; r++)
    for (int64_t c = rp[r]; c < rp[r + 1]; c++)
      if (ci[c] + 1 > ncol_max) ncol_max = ci[c] + 1;
  for (int64_t b = 0; b < nbc; b++)
    if (bc_dof[b] + 1 > ncol_max) ncol_max = bc_dof[b] + 1;
  char *fixed = (char *)calloc((size_t)ncol_max, 1);
  double *first = (double *)malloc(sizeof(double) * (size_t)ncol_max);
  int64_t *fidx = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncol_max);
  for (int64_t b = 0; b < nbc; b++) {
    if (!fixed[bc_dof[b]]) first[bc_dof[b]] = bc_val[b];
    fixed[bc_dof[b]] = 1;
  }
  int64_t nf = 0;
  for (int64_t i = 0; i < n; i++) {
    fidx[i] = -1;
    if (!fixed[i]) { fidx[i] = nf; free_out[nf++] = i; }
  }
  for (int64_t i = n; i < ncol_max; i++) fidx[i] = -1;
  int64_t t = 0;
  for (int64_t q = 0; q < nf; q++) {
    int64_t f = free_out[q];
    if (f >= nrows_csr) { t = -1; break; }
    double r = rhs[f];
    for (int64_t c = rp[f]; c < rp[f + 1]; c++) {
      int64_t col = ci[c];
      if (fixed[col]) {
        r -= v[c] * first[col];
      } else {
        if (fidx[col] < 0) { t = -1; break; } /* reference: free_idx[&col] panics */
        kr[t] = q; kc[t] = fidx[col]; kv[t] = v[c]; t++;
      }
    }
    if (t < 0) break;
    fred[q] = r;
  }
  *n_free = nf;
  free(fixed); free(first); free(fidx);
  return t;
}

/* Dense partial-pivot LU solve (the external tpt-math-linalg-dense
 * DMatrix::solve called at crates/tpt-fem-sparse/src/lib.rs:268-274): the
 * textbook Doolittle elimination with row partial pivoting (largest |a|,
 * first on ties).  Returns 0, or -1 on an exactly zero pivot.               */
int orc_dense_solve(int64_t n, double *a, double *b) {
#define A(i, j) a[(i) + (j) * n]
  for (int64_t k = 0; k < n; k++) {
    int64_t p = k;
    double best = fabs(A(k, k));
    for (int64_t i = k + 1; i < n; i++)
      if (fabs(A(i, k)) > best) { best = fabs(A(i, k)); p = i; }
    if (best == 0.0) return -1;
    if (p != k) {
      for (int64_t j = 0; j < n; j++) { double t = A(k, j); A(k, j) = A(p, j); A(p, j) = t; }
      double t = b[k]; b[k] = b[p]; b[p] = t;
    }
    double piv = A(k, k);
    for (int64_t i = k + 1; i < n; i++) {
      double l = A(i, k) / piv;
      A(i, k) = l;
      if (l == 0.0) continue;
      for (int64_t j = k + 1; j < n; j++) A(i, j) -= l * A(k, j);
      b[i] -= l * b[k];
    }
  }
  for (int64_t i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int64_t j = i + 1; j < n; j++) s -= A(i, j) * b[j];
    b[i] = s / A(i, i);
  }
#undef A
  return 0;
}

/* solve / solve_multi: crates/tpt-fem-sparse/src/lib.rs:225-278.  Returns 0,
 * -2 for a non-square matrix, -3 for a rhs length mismatch, -1 singular.    */
int orc_solve_coo(int64_t ntrip, const int64_t *rows, const int64_t *cols,
                  const double *vals, int64_t nb, double *b) {
  int64_t nrows, ncols;
  orc_coo_dims(ntrip, rows, cols, &nrows, &ncols);
  if (nrows != ncols) return -2;
  if (nb != nrows) return -3;
  int64_t n = nrows;
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  int64_t *ci = (int64_t *)malloc(sizeof(int64_t) * (ntrip ? ntrip : 1));
  double *v = (double *)malloc(sizeof(double) * (ntrip ? ntrip : 1));
  orc_to_csr(ntrip, rows, cols, vals, n, rp, ci, v);
  double *a = (double *)calloc((size_t)(n > 0 ? n * n : 1), sizeof(double));
  for (int64_t r = 0; r < n; r++)
    for (int64_t c = rp[r]; c < rp[r + 1]; c++) a[r + ci[c] * n] += v[c];
  int rc = orc_dense_solve(n, a, b);
  free(rp); free(ci); free(v); free(a);
  return rc;
}

/* solve_with_dirichlet: crates/tpt-fem-assembly/src/lib.rs:543-559.        */
int orc_solve_with_dirichlet(int64_t ntrip, const int64_t *rows, const int64_t *cols,
                             const double *vals, int64_t n, const double *rhs, int64_t nbc,
                             const int64_t *bc_dof, const double *bc_val, double *u) {
  int64_t nrows, ncols;
  orc_coo_dims(ntrip, rows, cols, &nrows, &ncols);
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (nrows + 1));
  int64_t *ci = (int64_t *)malloc(sizeof(int64_t) * (ntrip ? ntrip : 1));
  double *v = (double *)malloc(sizeof(double) * (ntrip ? ntrip : 1));
  int64_t nnz = orc_to_csr(ntrip, rows, cols, vals, nrows, rp, ci, v);
  int64_t *kr = (int64_t *)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  int64_t *kc = (int64_t *)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  double *kv = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
  double *fred = (double *)malloc(sizeof(double) * (n ? n : 1));
  int64_t *fr = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
  int64_t nf = 0;
  int64_t nt = orc_reduce_system(nrows, rp, ci, v, n, rhs, nbc, bc_dof, bc_val, kr, kc, kv,
                                 fred, fr, &nf);
  int rc = nt < 0 ? -4 : orc_solve_coo(nt, kr, kc, kv, nf, fred);
  if (rc == 0) {
    for (int64_t i = 0; i < n; i++) u[i] = 0.0;
    for (int64_t i = 0; i < nf; i++) u[fr[i]] = fred[i];
    for (int64_t b = 0; b < nbc; b++) u[bc_dof[b]] = bc_val[b];
  }
  free(rp); free(ci); free(v); free(kr); free(kc); free(kv); free(fred); free(fr);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Boundary faces and Neumann loads: crates/tpt-fem-assembly/src/lib.rs
 * :73-117 (face tables), :336-358 (face_quad), :394-420 (surface),
 * :563-583 (boundary_faces), :601-667 (apply_neumann_order).               */
typedef struct { int nn; int fcell; int idx[4]; } face_def;
static const face_def TRI_F[3] = {{2, ORC_LINE, {1, 2}}, {2, ORC_LINE, {2, 0}}, {2, ORC_LINE, {0, 1}}};
static const face_def QUAD_F[4] = {{2, ORC_LINE, {0, 1}}, {2, ORC_LINE, {1, 2}},
                                   {2, ORC_LINE, {2, 3}}, {2, ORC_LINE, {3, 0}}};
static const face_def TET_F[4] = {{3, ORC_TRI, {1, 2, 3}}, {3, ORC_TRI, {0, 2, 3}},
                                  {3, ORC_TRI, {0, 1, 3}}, {3, ORC_TRI, {0, 1, 2}}};
static const face_def HEX_F[6] = {{4, ORC_QUAD, {0, 1, 2, 3}}, {4, ORC_QUAD, {4, 5, 6, 7}},
                                  {4, ORC_QUAD, {0, 1, 5, 4}}, {4, ORC_QUAD, {3, 2, 6, 7}},
                                  {4, ORC_QUAD, {0, 4, 7, 3}}, {4, ORC_QUAD, {1, 5, 6, 2}}};
static const face_def LINE_F[2] = {{1, ORC_LINE, {0}}, {1, ORC_LINE, {1}}};

static const face_def *faces_of(int cell, int *nf) {
  switch (cell) {
  case ORC_LINE: *nf = 2; return LINE_F;
  case ORC_TRI: *nf = 3; return TRI_F;
  case ORC_QUAD: *nf = 4; return QUAD_F;
  case ORC_TET: *nf = 4; return TET_F;
  case ORC_HEX: *nf = 6; return HEX_F;
  }
  *nf = 0;
  return NULL;
}

typedef struct { int64_t k[4]; int64_t pos; } face_key;
static int cmp_key(const void *pa, const void *pb) {
  const face_key *a = (const face_key *)pa, *b = (const face_key *)pb;
  for (int i = 0; i < 4; i++)
    if (a->k[i] != b->k[i]) return a->k[i] < b->k[i] ? -1 : 1;
  return 0;
}

int64_t orc_boundary_faces(int64_t nelem, int cell, const int64_t *conn, int64_t *out_eid,
                           int64_t *out_face) {
  int nf, nn = orc_cell_nodes(cell);
  const face_def *F = faces_of(cell, &nf);
  int64_t nk = nelem * nf;
  face_key *keys = (face_key *)malloc(sizeof(face_key) * (nk ? nk : 1));
  for (int64_t e = 0; e < nelem; e++)
    for (int f = 0; f < nf; f++) {
      face_key *fk = &keys[e * nf + f];
      for (int i = 0; i < 4; i++) fk->k[i] = -1;
      for (int i = 0; i < F[f].nn; i++) fk->k[i] = conn[e * nn + F[f].idx[i]];
      /* sort_unstable of the node ids */
      for (int i = 1; i < F[f].nn; i++)
        for (int j = i; j > 0 && fk->k[j - 1] > fk->k[j]; j--) {
          int64_t t = fk->k[j]; fk->k[j] = fk->k[j - 1]; fk->k[j - 1] = t;
        }
      fk->pos = e * nf + f;
    }
  face_key *sorted = (face_key *)malloc(sizeof(face_key) * (nk ? nk : 1));
  memcpy(sorted, keys, sizeof(face_key) * nk);
  qsort(sorted, (size_t)nk, sizeof(face_key), cmp_key);
  char *once = (char *)calloc((size_t)(nk ? nk : 1), 1);
  for (int64_t i = 0; i < nk;) {
    int64_t j = i + 1;
    while (j < nk && cmp_key(&sorted[i], &sorted[j]) == 0) j++;
    if (j - i == 1) once[sorted[i].pos] = 1;
    i = j;
  }
  int64_t m = 0;
  for (int64_t p = 0; p < nk; p++)
    if (once[p]) { out_eid[m] = p / nf; out_face[m] = p % nf; m++; }
  free(keys); free(sorted); free(once);
  return m;
}

static void face_quad(int fcell, int order, double *pts, double *wts, int *nq) {
  double gp[5], gw[5];
  if (fcell == ORC_LINE) {
    int n = orc_gauss_legendre(order, gp, gw);
    for (int i = 0; i < n; i++) { pts[3 * i] = gp[i]; pts[3 * i + 1] = 0; pts[3 * i + 2] = 0; wts[i] = gw[i]; }
    *nq = n;
  } else if (fcell == ORC_TRI) {
    *nq = orc_cell_quad(ORC_TRI, order, pts, wts);
  } else {
    *nq = orc_cell_quad(ORC_QUAD, order, pts, wts);
  }
}

/* One scalar Neumann pass with flux g(x, n) = gconst if (axis < 0 or
 * |x[axis] - coord| <= tol) else 0.  comp < 0 tiles the load over all dpn
 * components (the reference's behaviour); comp >= 0 writes one component. */
static void neumann_pass_ex(int64_t nnodes, int pdim, const double *coords, int64_t nelem,
                            int cell, const int64_t *conn, int dpn, int comp, double gconst,
                            int axis, double coord, double tol, int order, double *rhs,
                            const double *gq, double *xout, double *nout, int64_t *npts);
static void neumann_pass(int64_t nnodes, int pdim, const double *coords, int64_t nelem,
                         int cell, const int64_t *conn, int dpn, int comp, double gconst,
                         int axis, double coord, double tol, int order, double *rhs) {
  neumann_pass_ex(nnodes, pdim, coords, nelem, cell, conn, dpn, comp, gconst, axis, coord, tol, order, rhs,
                  NULL, NULL, NULL, NULL);
}

/* The same pass with the closure's value at every boundary-face quadrature
 * point given in gq (boundary_faces order, then points in rule order) when
 * gq != NULL; xout / nout (may be NULL) receive each point's position and
 * outward unit normal -- the arguments apply_neumann_order passes to the
 * closure (crates/tpt-fem-assembly/src/lib.rs:620-656) -- and *npts their
 * count.  rhs may be NULL (points only).                                    */
static void neumann_pass_ex(int64_t nnodes, int pdim, const double *coords, int64_t nelem,
                            int cell, const int64_t *conn, int dpn, int comp, double gconst,
                            int axis, double coord, double tol, int order, double *rhs,
                            const double *gq, double *xout, double *nout, int64_t *npts) {
  (void)nnodes;
  int64_t ip = 0;
  int nf, nn = orc_cell_nodes(cell);
  const face_def *F = faces_of(cell, &nf);
  int64_t *be = (int64_t *)malloc(sizeof(int64_t) * (nelem * nf + 1));
  int64_t *bf = (int64_t *)malloc(sizeof(int64_t) * (nelem * nf + 1));
  int64_t nb = orc_boundary_faces(nelem, cell, conn, be, bf);
  int dim = pdim;
  for (int64_t q = 0; q < nb; q++) {
    int64_t e = be[q];
    const face_def *fd = &F[bf[q]];
    double phys[8][3], ecen[3] = {0, 0, 0}, fcen[3] = {0, 0, 0};
    for (int a = 0; a < nn; a++)
      for (int d = 0; d < dim; d++) phys[a][d] = coords[conn[e * nn + a] * pdim + d];
    /* mean(): sum in node order then divide by the count */
    for (int a = 0; a < nn; a++)
      for (int d = 0; d < dim; d++) ecen[d] += phys[a][d];
    for (int d = 0; d < dim; d++) ecen[d] /= (double)nn;
    for (int a = 0; a < fd->nn; a++)
      for (int d = 0; d < dim; d++) fcen[d] += phys[fd->idx[a]][d];
    for (int d = 0; d < dim; d++) fcen[d] /= (double)fd->nn;
    double qp[125 * 3], qw[125];
    int nq = 0;
    if (fd->fcell == ORC_LINE && fd->nn == 1) {
      /* point face of a Line2 element: GL rule of the given order on a 0-d
       * face is not defined in the reference either; skip. */
      continue;
    }
    face_quad(fd->fcell, order, qp, qw, &nq);
    for (int iq = 0; iq < nq; iq++) {
      double ns[8], ng[8 * 3];
      orc_shape(fd->fcell, qp + 3 * iq, ns);
      orc_grad(fd->fcell, qp + 3 * iq, ng);
      int fdim = orc_cell_dim(fd->fcell);
      double xp[3] = {0, 0, 0};
      for (int k = 0; k < fd->nn; k++)
        for (int d = 0; d < dim; d++) xp[d] += ns[k] * phys[fd->idx[k]][d];
      double tang[2][3] = {{0, 0, 0}, {0, 0, 0}};
      for (int m = 0; m < fdim; m++)
        for (int k = 0; k < fd->nn; k++)
          for (int d = 0; d < dim; d++) tang[m][d] += ng[k * fdim + m] * phys[fd->idx[k]][d];
      double measure, nrm[3] = {0, 0, 0};
      if (fdim == 1) {
        double s = -0.0;
        for (int d = 0; d < dim; d++) s = s + tang[0][d] * tang[0][d];
        measure = sqrt(s);
        nrm[0] = -tang[0][1];
        nrm[1] = tang[0][0];
        double s2 = -0.0;
        for (int d = 0; d < dim; d++) s2 = s2 + nrm[d] * nrm[d];
        double ln = sqrt(s2);
        for (int d = 0; d < dim; d++) nrm[d] /= ln;
      } else {
        double c[3] = {tang[0][1] * tang[1][2] - tang[0][2] * tang[1][1],
                       tang[0][2] * tang[1][0] - tang[0][0] * tang[1][2],
                       tang[0][0] * tang[1][1] - tang[0][1] * tang[1][0]};
        double s = -0.0;
        for (int d = 0; d < 3; d++) s = s + c[d] * c[d];
        measure = sqrt(s);
        double ln = measure;
        for (int d = 0; d < 3; d++) nrm[d] = c[d] / ln;
      }
      /* outward: flip when the normal points against face centroid minus
       * element centroid (:642-651) */
      double dd = -0.0;
      for (int d = 0; d < dim; d++) dd = dd + nrm[d] * (fcen[d] - ecen[d]);
      if (dd < 0.0)
        for (int d = 0; d < dim; d++) nrm[d] = -nrm[d];
      if (xout)
        for (int d = 0; d < dim; d++) xout[ip * dim + d] = xp[d];
      if (nout)
        for (int d = 0; d < dim; d++) nout[ip * dim + d] = nrm[d];
      double g = gq ? gq[ip] : gconst;
      ip++;
      if (!gq && axis >= 0 && !(fabs(xp[axis] - coord) <= tol)) g = 0.0;
      if (!rhs) continue;
      for (int k = 0; k < fd->nn; k++) {
        int64_t base = conn[e * nn + fd->idx[k]] * dpn;
        if (comp < 0) {
          for (int c = 0; c < dpn; c++) rhs[base + c] += qw[iq] * measure * g * ns[k];
        } else {
          rhs[base + comp] += qw[iq] * measure * g * ns[k];
        }
      }
    }
  }
  if (npts) *npts = ip;
  free(be); free(bf);
}

/* Positions and outward unit normals of every boundary-face quadrature point
 * (the closure arguments of apply_neumann_order / apply_robin_order); x and
 * nrm: count x pdim, either may be NULL.  Returns the count.              */
int64_t orc_boundary_points(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                            const int64_t *conn, int order, double *x, double *nrm) {
  int64_t np = 0;
  neumann_pass_ex(nnodes, pdim, coords, nelem, cell, conn, 1, -1, 0.0, -1, 0.0, 0.0, order, NULL, NULL, x, nrm,
                  &np);
  return np;
}

/* apply_neumann_order (crates/tpt-fem-assembly/src/lib.rs:601-667) with the
 * flux closure's value g(x_q, n_q) at every boundary point (gq, in
 * orc_boundary_points order), tiled over dpn components.                   */
void orc_apply_neumann_qp(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                          const int64_t *conn, int dpn, const double *gq, int order, double *rhs) {
  neumann_pass_ex(nnodes, pdim, coords, nelem, cell, conn, dpn, -1, 0.0, -1, 0.0, 0.0, order, rhs, gq, NULL, NULL,
                  NULL);
}

/* apply_robin_order: crates/tpt-fem-assembly/src/lib.rs:684-753.  Appends,
 * for every boundary face (boundary_faces order), quadrature point, face node
 * pair (k, l) and component c, the triplet (n_k*dpn + c, n_l*dpn + c,
 * w * measure * h * N_k * N_l) -- no zero skip -- with h the constant hconst,
 * or hconst on |x[axis] - coord| <= tol and 0 elsewhere when axis >= 0.
 * Returns the number of triplets written.                                   */
static int64_t robin_pass(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                          const int64_t *conn, int dpn, double hconst, int axis, double coord, double tol,
                          int order, const double *hq, int64_t *rows, int64_t *cols, double *vals);
int64_t orc_apply_robin(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                        const int64_t *conn, int dpn, double hconst, int axis, double coord,
                        double tol, int order, int64_t *rows, int64_t *cols, double *vals) {
  return robin_pass(nnodes, pdim, coords, nelem, cell, conn, dpn, hconst, axis, coord, tol, order, NULL, rows, cols,
                    vals);
}

/* apply_robin_order with the coefficient closure's value h(x_q, n_q) at every
 * boundary point (hq, in orc_boundary_points order).                        */
int64_t orc_apply_robin_qp(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                           const int64_t *conn, int dpn, const double *hq, int order, int64_t *rows,
                           int64_t *cols, double *vals) {
  return robin_pass(nnodes, pdim, coords, nelem, cell, conn, dpn, 0.0, -1, 0.0, 0.0, order, hq, rows, cols, vals);
}

static int64_t robin_pass(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                          const int64_t *conn, int dpn, double hconst, int axis, double coord, double tol,
                          int order, const double *hq, int64_t *rows, int64_t *cols, double *vals) {
  (void)nnodes;
  int64_t ip = 0;
  int nf, nn = orc_cell_nodes(cell);
  const face_def *F = faces_of(cell, &nf);
  int64_t *be = (int64_t *)malloc(sizeof(int64_t) * (nelem * nf + 1));
  int64_t *bf = (int64_t *)malloc(sizeof(int64_t) * (nelem * nf + 1));
  int64_t nb = orc_boundary_faces(nelem, cell, conn, be, bf);
  int64_t t = 0;
  for (int64_t q = 0; q < nb; q++) {
    const int64_t e = be[q];
    const face_def *fd = &F[bf[q]];
    if (fd->nn == 1) continue; /* point faces of Line2 cells, as in the Neumann pass */
    double phys[8][3];
    for (int a = 0; a < nn; a++)
      for (int d = 0; d < pdim; d++) phys[a][d] = coords[conn[e * nn + a] * pdim + d];
    double qp[125 * 3], qw[125];
    int nq = 0;
    face_quad(fd->fcell, order, qp, qw, &nq);
    const int fdim = orc_cell_dim(fd->fcell);
    for (int iq = 0; iq < nq; iq++) {
      double ns[8], ng[8 * 3], xp[3] = {0, 0, 0}, tang[2][3] = {{0, 0, 0}, {0, 0, 0}};
      orc_shape(fd->fcell, qp + 3 * iq, ns);
      orc_grad(fd->fcell, qp + 3 * iq, ng);
      for (int k = 0; k < fd->nn; k++)
        for (int d = 0; d < pdim; d++) xp[d] += ns[k] * phys[fd->idx[k]][d];
      for (int m = 0; m < fdim; m++)
        for (int k = 0; k < fd->nn; k++)
          for (int d = 0; d < pdim; d++) tang[m][d] += ng[k * fdim + m] * phys[fd->idx[k]][d];
      double measure;
      if (fdim == 1) {
        double s2 = -0.0;
        for (int d = 0; d < pdim; d++) s2 = s2 + tang[0][d] * tang[0][d];
        measure = sqrt(s2);
      } else {
        const double c0 = tang[0][1] * tang[1][2] - tang[0][2] * tang[1][1];
        const double c1 = tang[0][2] * tang[1][0] - tang[0][0] * tang[1][2];
        const double c2 = tang[0][0] * tang[1][1] - tang[0][1] * tang[1][0];
        double s2 = -0.0;
        s2 = s2 + c0 * c0;
        s2 = s2 + c1 * c1;
        s2 = s2 + c2 * c2;
        measure = sqrt(s2);
      }
      double h = hq ? hq[ip] : hconst;
      ip++;
      if (!hq && axis >= 0 && !(fabs(xp[axis] - coord) <= tol)) h = 0.0;
      for (int k = 0; k < fd->nn; k++)
        for (int l = 0; l < fd->nn; l++) {
          const int64_t bi = conn[e * nn + fd->idx[k]] * dpn, bj = conn[e * nn + fd->idx[l]] * dpn;
          for (int c = 0; c < dpn; c++) {
            rows[t] = bi + c;
            cols[t] = bj + c;
            vals[t] = qw[iq] * measure * h * ns[k] * ns[l];
            t++;
          }
        }
    }
  }
  free(be); free(bf);
  return t;
}

void orc_apply_neumann(int64_t nnodes, int pdim, const double *coords, int64_t nelem,
                       int cell, const int64_t *conn, int dpn, double g, int order,
                       double *rhs) {
  neumann_pass(nnodes, pdim, coords, nelem, cell, conn, dpn, -1, g, -1, 0.0, 0.0, order, rhs);
}

/* Traction t on a plane, composed from the reference's scalar Neumann pass:
 * for each nonzero component c, buf = apply_neumann(mesh, 1, g_c, buf) with
 * g_c(x) = t[c] on the plane and 0 elsewhere, then rhs[n*dpn + c] += buf[n]. */
void orc_plane_traction(int64_t nnodes, int pdim, const double *coords, int64_t nelem,
                        int cell, const int64_t *conn, int dpn, int axis, double coord,
                        double tol, const double *t, int order, double *rhs) {
  double *buf = (double *)malloc(sizeof(double) * (nnodes ? nnodes : 1));
  for (int c = 0; c < dpn; c++) {
    if (t[c] == 0.0) continue;
    for (int64_t n = 0; n < nnodes; n++) buf[n] = 0.0;
    neumann_pass(nnodes, pdim, coords, nelem, cell, conn, 1, 0, t[c], axis, coord, tol, order, buf);
    for (int64_t n = 0; n < nnodes; n++) rhs[n * dpn + c] += buf[n];
  }
  free(buf);
}

/* ------------------------------------------------------------------------ */
/* Drivers: solve_poisson (crates/tpt-fem-thermal/src/lib.rs:206-240).      */
void orc_poisson_rhs(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                     const int64_t *conn, double f_const, int order, double *rhs) {
  int nn = orc_cell_nodes(cell);
  double phys[8 * 3], f[8];
  for (int64_t i = 0; i < nnodes; i++) rhs[i] = 0.0;
  for (int64_t e = 0; e < nelem; e++) {
    for (int a = 0; a < nn; a++)
      for (int p = 0; p < pdim; p++) phys[a * pdim + p] = coords[conn[e * nn + a] * pdim + p];
    orc_poisson_source_vector(cell, pdim, phys, f_const, order, f);
    for (int i = 0; i < nn; i++) rhs[conn[e * nn + i]] += f[i];
  }
}

/* solve_elasticity's body-force scatter (crates/tpt-fem-elasticity/src/lib.rs:370-379):
 * rhs[node*dim + a] += f[i*dim + a], elements in order.  bq: nelem x nq x dim. */
void orc_body_rhs(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                  const int64_t *conn, const double *bq, int order, double *rhs) {
  double qp[125 * 3], qw[125];
  int nn = orc_cell_nodes(cell), dim = orc_cell_dim(cell);
  int nq = orc_cell_quad(cell, order, qp, qw);
  double phys[8 * 3], f[8 * 3];
  for (int64_t i = 0; i < nnodes * dim; i++) rhs[i] = 0.0;
  for (int64_t e = 0; e < nelem; e++) {
    for (int a = 0; a < nn; a++)
      for (int p = 0; p < pdim; p++) phys[a * pdim + p] = coords[conn[e * nn + a] * pdim + p];
    orc_elasticity_body_vector(cell, pdim, phys, bq + e * nq * dim, order, f);
    for (int i = 0; i < nn; i++)
      for (int a = 0; a < dim; a++) rhs[conn[e * nn + i] * dim + a] += f[i * dim + a];
  }
}

/* solve_poisson's source scatter (crates/tpt-fem-thermal/src/lib.rs:223-230)
 * with per-point source values sq: nelem x nq.                              */
void orc_source_rhs_qp(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                       const int64_t *conn, const double *sq, int order, double *rhs) {
  double qp[125 * 3], qw[125];
  int nn = orc_cell_nodes(cell);
  int nq = orc_cell_quad(cell, order, qp, qw);
  double phys[8 * 3], f[8];
  for (int64_t i = 0; i < nnodes; i++) rhs[i] = 0.0;
  for (int64_t e = 0; e < nelem; e++) {
    for (int a = 0; a < nn; a++)
      for (int p = 0; p < pdim; p++) phys[a * pdim + p] = coords[conn[e * nn + a] * pdim + p];
    orc_poisson_source_vector_qp(cell, pdim, phys, sq + e * nq, order, f);
    for (int i = 0; i < nn; i++) rhs[conn[e * nn + i]] += f[i];
  }
}

/* solve_elasticity (crates/tpt-fem-elasticity/src/lib.rs:355-381): try_assemble
 * of elasticity_element_matrix, the body-force rhs, solve_with_dirichlet.
 * Returns 0, -5 on a model/dimension mismatch, or the solver's status.      */
int orc_solve_elasticity(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                         const int64_t *conn, int model, double young, double nu, int order,
                         const double *bq, int64_t nbc, const int64_t *bc_dof, const double *bc_val,
                         double *u) {
  int nn = orc_cell_nodes(cell), dim = orc_cell_dim(cell);
  int64_t cap = nelem * nn * nn * dim * dim;
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * (cap ? cap : 1));
  int64_t *c = (int64_t *)malloc(sizeof(int64_t) * (cap ? cap : 1));
  double *v = (double *)malloc(sizeof(double) * (cap ? cap : 1));
  double *rhs = (double *)malloc(sizeof(double) * (nnodes * dim > 0 ? nnodes * dim : 1));
  orc_op op;
  memset(&op, 0, sizeof op);
  op.cell = cell; op.pdim = pdim; op.kind = 1; op.model = model; op.order = order;
  op.coef = young; op.nu = nu;
  int64_t t = orc_assemble(nnodes, coords, nelem, conn, &op, dim, r, c, v);
  int rc = -5;
  if (t >= 0) {
    orc_body_rhs(nnodes, pdim, coords, nelem, cell, conn, bq, order, rhs);
    rc = orc_solve_with_dirichlet(t, r, c, v, nnodes * dim, rhs, nbc, bc_dof, bc_val, u);
  }
  free(r); free(c); free(v); free(rhs);
  return rc;
}

int orc_solve_poisson(int64_t nnodes, int pdim, const double *coords, int64_t nelem, int cell,
                      const int64_t *conn, double kappa, int order, double f_const, int64_t nbc,
                      const int64_t *bc_dof, const double *bc_val, double *u) {
  int nn = orc_cell_nodes(cell);
  int64_t cap = nelem * nn * nn;
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * (cap ? cap : 1));
  int64_t *c = (int64_t *)malloc(sizeof(int64_t) * (cap ? cap : 1));
  double *v = (double *)malloc(sizeof(double) * (cap ? cap : 1));
  orc_op op;
  memset(&op, 0, sizeof op);
  op.cell = cell; op.pdim = pdim; op.kind = 0; op.order = order; op.coef = kappa;
  int64_t t = orc_assemble(nnodes, coords, nelem, conn, &op, 1, r, c, v);
  double *rhs = (double *)malloc(sizeof(double) * (nnodes ? nnodes : 1));
  orc_poisson_rhs(nnodes, pdim, coords, nelem, cell, conn, f_const, order, rhs);
  int rc = orc_solve_with_dirichlet(t, r, c, v, nnodes, rhs, nbc, bc_dof, bc_val, u);
  free(r); free(c); free(v); free(rhs);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Counter-based generators (inputs only; no reference counterpart).        */
uint64_t orc_splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double hash_u01(uint64_t seed, uint64_t ctr) {
  uint64_t s = seed ^ (ctr * 0xD1B54A32D192ED03ull);
  return (double)(orc_splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0);
}

double orc_uniform01(uint64_t *state) {
  return (double)(orc_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* Interior nodes of a box_mesh grid (not on any face) move by
 * (2u - 1) * h_frac * h_axis per axis, u = hash(seed, 3*node + axis).       */
void orc_jitter_interior(int64_t nx, int64_t ny, int64_t nz, double h_frac, uint64_t seed,
                         double *coords) {
  double h[3] = {1.0 / (double)nx, 1.0 / (double)ny, 1.0 / (double)nz};
  for (int64_t i = 1; i < nx; i++)
    for (int64_t j = 1; j < ny; j++)
      for (int64_t k = 1; k < nz; k++) {
        int64_t id = (i * (ny + 1) + j) * (nz + 1) + k;
        for (int a = 0; a < 3; a++) {
          double u = hash_u01(seed, (uint64_t)(3 * id + a));
          coords[3 * id + a] += (2.0 * u - 1.0) * (h_frac * h[a]);
        }
      }
}

void orc_loguniform(int64_t n, double lo, double hi, uint64_t seed, double *out) {
  for (int64_t i = 0; i < n; i++) out[i] = pow(10.0, lo + (hi - lo) * hash_u01(seed, (uint64_t)i));
}


/* ======================================================================== *
 * SIMP minimum-compliance loop: crates/tpt-fem-topopt/src/lib.rs.           *
 * ======================================================================== */
#define TOPOPT_E_MIN_FACTOR 1e-3 /* :53 */
#define TOPOPT_RHO_MIN 1e-3      /* :57 */
#define TOPOPT_RHO_MAX 1.0       /* :59 */
#define TOPOPT_OC_DAMP 0.5       /* :61 */

/* element_quad_form (:241-254): u_e^T K0 u_e, i then j, from +0.0.          */
double orc_topopt_quad_form(const double *ke0, const double *u, const int64_t *nodes) {
  double ue[8];
  for (int k = 0; k < 4; k++) {
    ue[2 * k] = u[2 * nodes[k]];
    ue[2 * k + 1] = u[2 * nodes[k] + 1];
  }
  double s = 0.0;
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) s += ue[i] * ke0[i * 8 + j] * ue[j];
  return s;
}

/* filter_sensitivity (:259-279): every element f, ascending, with
 * |c_e - c_f|^2 <= R^2 weighted by R - |c_e - c_f|; centres from Grid::new
 * (:94-107).                                                                */
void orc_topopt_filter(int64_t nx, int64_t ny, double h, const double *raw, double radius,
                       double *out) {
  const int64_t nex = nx - 1, ne = (nx - 1) * (ny - 1);
  const double r2 = radius * radius;
  for (int64_t e = 0; e < ne; e++) {
    const double cex = ((double)(e % nex) + 0.5) * h, cey = ((double)(e / nex) + 0.5) * h;
    double num = 0.0, den = 0.0;
    for (int64_t f = 0; f < ne; f++) {
      const double dx = cex - ((double)(f % nex) + 0.5) * h;
      const double dy = cey - ((double)(f / nex) + 0.5) * h;
      const double d2 = dx * dx + dy * dy;
      if (d2 <= r2) {
        const double w = radius - sqrt(d2);
        num += w * raw[f];
        den += w;
      }
    }
    out[e] = den > 0.0 ? num / den : raw[e];
  }
}

static double topopt_clamp(double x, double lo, double hi) {
  if (x < lo) x = lo; /* f64::clamp */
  if (x > hi) x = hi;
  return x;
}

/* oc_update (:284-331): 60 bisection steps on lambda in [1e-9, 1e9], then
 * the damped, move-limited update at the last midpoint.                    */
void orc_topopt_oc(int64_t n, const double *rho, const double *sens, double vol_frac,
                   double move_limit, double *out) {
  const double target = vol_frac * (double)n;
  double lo = 1e-9, hi = 1e9, lambda = 1.0;
  for (int it = 0; it < 60; it++) {
    lambda = 0.5 * (lo + hi);
    double sum = 0.0;
    for (int64_t e = 0; e < n; e++) {
      const double b = -sens[e];
      const double ratio = b > 0.0 ? pow(b / lambda, TOPOPT_OC_DAMP) : 0.0;
      const double cand = rho[e] * ratio;
      sum += topopt_clamp(cand, fmax(rho[e] - move_limit, TOPOPT_RHO_MIN),
                          fmin(rho[e] + move_limit, TOPOPT_RHO_MAX));
    }
    if (sum > target) lo = lambda;
    else hi = lambda;
  }
  for (int64_t e = 0; e < n; e++) {
    const double b = -sens[e];
    const double ratio = b > 0.0 ? pow(b / lambda, TOPOPT_OC_DAMP) : 0.0;
    out[e] = topopt_clamp(rho[e] * ratio, fmax(rho[e] - move_limit, TOPOPT_RHO_MIN),
                          fmin(rho[e] + move_limit, TOPOPT_RHO_MAX));
  }
}

/* dot (:388-390): Iterator::sum folds from -0.0.                           */
static double topopt_dot(int64_t n, const double *a, const double *b) {
  double s = -0.0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* topopt_simp (:342-385) on Grid::new(nx, ny, h): compliance_out gets
 * max_iter + 1 values (the uniform start, then one per OC iteration, before
 * its update), rho_out the final densities.  Returns the solve's status.  */
int orc_topopt_simp(int64_t nx, int64_t ny, double h, double e0, double nu, double vol_frac,
                    double penal, double radius, int max_iter, double move_limit,
                    const double *f, int64_t nbc, const int64_t *bc_dof, const double *bc_val,
                    double *rho_out, double *compliance_out, int *iterations) {
  const int64_t nn = nx * ny, ne = (nx - 1) * (ny - 1), ndof = 2 * nn;
  const double emin = TOPOPT_E_MIN_FACTOR * e0;
  double ke0[64];
  orc_simp_ke_unit(nu, h, ke0);
  double *coords = (double *)malloc(sizeof(double) * 2 * nn);
  int64_t *conn = (int64_t *)malloc(sizeof(int64_t) * 4 * ne);
  orc_grid(nx, ny, h, coords, conn);
  double *rho = (double *)malloc(sizeof(double) * ne);
  double *raw = (double *)malloc(sizeof(double) * ne);
  double *sens = (double *)malloc(sizeof(double) * ne);
  double *u = (double *)malloc(sizeof(double) * ndof);
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * 64 * ne);
  int64_t *c = (int64_t *)malloc(sizeof(int64_t) * 64 * ne);
  double *v = (double *)malloc(sizeof(double) * 64 * ne);
  for (int64_t e = 0; e < ne; e++) rho[e] = vol_frac;
  orc_op op;
  memset(&op, 0, sizeof op);
  op.cell = ORC_QUAD;
  op.pdim = 2;
  op.kind = 2;
  op.rho = rho;
  op.penal = penal;
  op.e0 = e0;
  op.emin = emin;
  op.ke0 = ke0;
  int rc = 0, it = 0;
  for (int k = 0; k <= max_iter && !rc; k++) {
    const int64_t t = orc_assemble(nn, coords, ne, conn, &op, 2, r, c, v);
    rc = orc_solve_with_dirichlet(t, r, c, v, ndof, f, nbc, bc_dof, bc_val, u);
    if (rc) break;
    compliance_out[k] = topopt_dot(ndof, f, u);
    if (k == 0) continue; /* the uniform start is solved once more in iteration 1 */
    it++;
    for (int64_t e = 0; e < ne; e++)
      raw[e] = -penal * pow(rho[e], penal - 1.0) * (e0 - emin) * orc_topopt_quad_form(ke0, u, conn + 4 * e);
    orc_topopt_filter(nx, ny, h, raw, radius, sens);
    orc_topopt_oc(ne, rho, sens, vol_frac, move_limit, rho);
  }
  memcpy(rho_out, rho, sizeof(double) * ne);
  *iterations = it;
  free(coords); free(conn); free(rho); free(raw); free(sens); free(u); free(r); free(c); free(v);
  return rc;
}


/* ======================================================================== *
 * Lanczos eigenpairs: crates/tpt-fem-eigen/src/lib.rs.                      *
 * ======================================================================== */
/* dot / norm (:15-21): Iterator::sum folds from -0.0.                      */
static double eig_dot(int64_t n, const double *a, const double *b) {
  double s = -0.0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* matvec (:24-35): row sums from 0.0 in CSR order; y has nrows entries.    */
void orc_eig_matvec(int64_t n, const int64_t *rp, const int64_t *ci, const double *v,
                    const double *x, double *y) {
  for (int64_t r = 0; r < n; r++) {
    double s = 0.0;
    for (int64_t c = rp[r]; c < rp[r + 1]; c++) s += v[c] * x[ci[c]];
    y[r] = s;
  }
}

/* jacobi (:170-218): cyclic Jacobi on a dense symmetric m x m matrix a
 * (row-major, overwritten), eigenvectors as the columns of q.              */
void orc_eig_jacobi(int m, double *a, int max_sweeps, double *eig, double *q) {
  for (int i = 0; i < m * m; i++) q[i] = 0.0;
  for (int i = 0; i < m; i++) q[i * m + i] = 1.0;
  for (int sw = 0; sw < max_sweeps; sw++) {
    int changed = 0;
    for (int p = 0; p < m; p++)
      for (int r = p + 1; r < m; r++) {
        if (!(fabs(a[p * m + r]) > 1e-12)) continue;
        changed = 1;
        const double theta = (a[r * m + r] - a[p * m + p]) / (2.0 * a[p * m + r]);
        const double t = theta >= 0.0 ? 1.0 / (theta + sqrt(1.0 + theta * theta))
                                      : 1.0 / (theta - sqrt(1.0 + theta * theta));
        const double c = 1.0 / sqrt(1.0 + t * t);
        const double s = t * c;
        for (int i = 0; i < m; i++) {
          const double aip = a[i * m + p], aiq = a[i * m + r];
          a[i * m + p] = c * aip - s * aiq;
          a[i * m + r] = s * aip + c * aiq;
        }
        for (int i = 0; i < m; i++) {
          const double api = a[p * m + i], aqi = a[r * m + i];
          a[p * m + i] = c * api - s * aqi;
          a[r * m + i] = s * api + c * aqi;
        }
        for (int i = 0; i < m; i++) {
          const double vip = q[i * m + p], viq = q[i * m + r];
          q[i * m + p] = c * vip - s * viq;
          q[i * m + r] = s * vip + c * viq;
        }
      }
    if (!changed) break;
  }
  for (int i = 0; i < m; i++) eig[i] = a[i * m + i];
}

static const double *g_eig;
static int cmp_eig(const void *pa, const void *pb) {
  const int a = *(const int *)pa, b = *(const int *)pb;
  const double x = g_eig[a], y = g_eig[b];
  if (x < y) return -1;
  if (x > y) return 1;
  return a < b ? -1 : (a > b); /* stable; NaN compares Equal */
}

/* lanczos (:135-167) + lanczos_eigs (:223-268) on a square CSR of n rows.
 * which 0 = Smallest, 1 = Largest.  Returns the number of pairs (<= num);
 * evecs gets them row by row (n values each).                              */
int orc_lanczos_eigs(int64_t n, const int64_t *rp, const int64_t *ci, const double *v, int num,
                     int which, int lanczos_dim, double *evals, double *evecs) {
  const int md = lanczos_dim;
  if (md < 1 || n < 1) return -1;
  double *basis = (double *)calloc((size_t)md * (size_t)n, sizeof(double));
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  double *alpha = (double *)malloc(sizeof(double) * md), *beta = (double *)malloc(sizeof(double) * md);
  basis[0] = 1.0; /* e_0, normalised by its norm 1 */
  int m = 1, na = 0;
  for (int j = 0; j < md; j++) {
    const double *vj = basis + (size_t)j * n;
    orc_eig_matvec(n, rp, ci, v, vj, w);
    const double a = eig_dot(n, w, vj);
    for (int k = 0; k <= j; k++) {
      const double *vk = basis + (size_t)k * n;
      const double p = eig_dot(n, w, vk);
      for (int64_t d = 0; d < n; d++) w[d] -= p * vk[d];
    }
    alpha[na++] = a;
    if (j + 1 == md) break;
    const double b = sqrt(eig_dot(n, w, w));
    if (b < 1e-12) break;
    beta[j] = b;
    double *vn = basis + (size_t)(j + 1) * n;
    for (int64_t d = 0; d < n; d++) vn[d] = w[d] / b;
    m++;
  }
  double *t = (double *)calloc((size_t)m * m, sizeof(double)), *q = (double *)malloc(sizeof(double) * m * m);
  double *eig = (double *)malloc(sizeof(double) * m);
  int *idx = (int *)malloc(sizeof(int) * m);
  for (int i = 0; i < m; i++) t[i * m + i] = alpha[i];
  for (int i = 0; i + 1 < m; i++) t[i * m + i + 1] = t[(i + 1) * m + i] = beta[i];
  orc_eig_jacobi(m, t, 200, eig, q);
  for (int i = 0; i < m; i++) idx[i] = i;
  g_eig = eig;
  qsort(idx, (size_t)m, sizeof(int), cmp_eig);
  const int take = num < m ? num : m;
  for (int s = 0; s < take; s++) {
    const int i = which == 0 ? idx[s] : idx[m - 1 - s];
    double *vec = evecs + (size_t)s * n;
    for (int64_t d = 0; d < n; d++) vec[d] = 0.0;
    for (int r = 0; r < m; r++) {
      const double c = q[r * m + i];
      for (int64_t d = 0; d < n; d++) vec[d] += c * basis[(size_t)r * n + d];
    }
    const double nv = sqrt(eig_dot(n, vec, vec));
    for (int64_t d = 0; d < n; d++) vec[d] /= nv;
    evals[s] = eig[i];
  }
  free(basis); free(w); free(alpha); free(beta); free(t); free(q); free(eig); free(idx);
  return take;
}


/* ======================================================================== *
 * Time integration: crates/tpt-fem-dynamic/src/lib.rs.  Operators are given *
 * as CSR (CachedSystem::new's to_csr of the Coo inputs, :145-151); the load *
 * is f(t_k) = fscale[k] * F (fscale NULL: 1).                               *
 * ======================================================================== */
/* Csr::matvec (crates/tpt-fem-sparse/src/lib.rs:195-205) into y[0..n).     */
static void dyn_matvec(const orc_csr *a, int64_t n, const double *x, double *y) {
  for (int64_t r = 0; r < n; r++) {
    double s = 0.0;
    if (r < a->nrows)
      for (int64_t c = a->rp[r]; c < a->rp[r + 1]; c++) s += a->v[c] * x[a->ci[c]];
    y[r] = s;
  }
}

static double dyn_scale(const double *fscale, int64_t k) { return fscale ? fscale[k] : 1.0; }

/* central_difference (:289-377).  Returns 0, or 1 for a CFL violation with
 * *critical set; history gets (nsteps + 1) rows of n.                      */
int orc_central_difference(int64_t n, orc_csr M, orc_csr D, orc_csr K, const double *u0,
                           const double *v0, const double *F, const double *fscale, double dt,
                           int64_t nsteps, double *history, double *critical) {
  double *md = (double *)malloc(sizeof(double) * (size_t)n), *dv = (double *)malloc(sizeof(double) * (size_t)n);
  double *ku = (double *)malloc(sizeof(double) * (size_t)n), *vh = (double *)malloc(sizeof(double) * (size_t)n);
  double *u = (double *)malloc(sizeof(double) * (size_t)n), *a = (double *)malloc(sizeof(double) * (size_t)n);
  for (int64_t r = 0; r < n; r++) {
    double s = 0.0;
    for (int64_t c = M.rp[r]; c < M.rp[r + 1]; c++) s += M.v[c];
    md[r] = s;
  }
  double rho = 0.0;
  int has_mass = 0;
  for (int64_t i = 0; i < n; i++)
    if (md[i] > 0.0) {
      double rs = 0.0;
      for (int64_t c = K.rp[i]; c < K.rp[i + 1]; c++) rs += fabs(K.v[c]);
      rho = fmax(rho, rs / md[i]);
      has_mass = 1;
    }
  const double safe = has_mass && rho > 0.0 ? 2.0 / sqrt(rho) : INFINITY;
  *critical = safe;
  int rc = 0;
  if (dt > safe) {
    rc = 1;
  } else {
    dyn_matvec(&D, n, v0, dv);
    dyn_matvec(&K, n, u0, ku);
    const double s0 = dyn_scale(fscale, 0);
    for (int64_t i = 0; i < n; i++) a[i] = (s0 * F[i] - dv[i] - ku[i]) / md[i];
    for (int64_t i = 0; i < n; i++) {
      u[i] = u0[i];
      vh[i] = v0[i] - 0.5 * dt * a[i];
      history[i] = u[i];
    }
    for (int64_t step = 1; step <= nsteps; step++) {
      const double sk = dyn_scale(fscale, step);
      dyn_matvec(&D, n, vh, dv);
      dyn_matvec(&K, n, u, ku);
      for (int64_t i = 0; i < n; i++) a[i] = (sk * F[i] - dv[i] - ku[i]) / md[i];
      for (int64_t i = 0; i < n; i++) vh[i] += dt * a[i];
      for (int64_t i = 0; i < n; i++) u[i] += dt * vh[i];
      memcpy(history + (size_t)step * n, u, sizeof(double) * (size_t)n);
    }
  }
  free(md); free(dv); free(ku); free(vh); free(u); free(a);
  return rc;
}

/* Dense copy of a CSR's first n rows / columns (column-major, summed).      */
static void dyn_dense(const orc_csr *a, int64_t n, double s, double *out) {
  for (int64_t r = 0; r < n && r < a->nrows; r++)
    for (int64_t c = a->rp[r]; c < a->rp[r + 1]; c++)
      if (a->ci[c] < n) out[r + a->ci[c] * n] += a->v[c] * s;
}

/* newmark (:194-263) with dense LU solves, as the reference.  K_hat is the
 * Coo sum K + M/(b dt^2) + C g/(b dt) collapsed in that insertion order
 * (:221-229).  Returns 0, or -1 when a solve hits a zero pivot.            */
int orc_newmark(int64_t n, orc_csr M, orc_csr D, orc_csr K, const double *u0, const double *v0,
                const double *F, const double *fscale, double dt, double beta, double gamma,
                int64_t nsteps, double *history) {
  const double b = beta, g = gamma;
  const size_t nn = (size_t)n;
  double *dv = (double *)malloc(sizeof(double) * nn), *ku = (double *)malloc(sizeof(double) * nn);
  double *a = (double *)malloc(sizeof(double) * nn), *u = (double *)malloc(sizeof(double) * nn);
  double *v = (double *)malloc(sizeof(double) * nn), *pm = (double *)malloc(sizeof(double) * nn);
  double *pc = (double *)malloc(sizeof(double) * nn), *rhs = (double *)malloc(sizeof(double) * nn);
  double *mp = (double *)malloc(sizeof(double) * nn), *dp = (double *)malloc(sizeof(double) * nn);
  double *dense = (double *)malloc(sizeof(double) * nn * nn), *kh = (double *)calloc(nn * nn, sizeof(double));
  int rc = 0;
  dyn_matvec(&D, n, v0, dv);
  dyn_matvec(&K, n, u0, ku);
  const double s0 = dyn_scale(fscale, 0);
  for (int64_t i = 0; i < n; i++) a[i] = s0 * F[i] - dv[i] - ku[i];
  memset(dense, 0, sizeof(double) * nn * nn);
  dyn_dense(&M, n, 1.0, dense);
  if (orc_dense_solve(n, dense, a)) rc = -1;
  /* K_hat: each entry sums K's, then M's scaled, then C's scaled values */
  dyn_dense(&K, n, 1.0, kh);
  dyn_dense(&M, n, 1.0 / (b * dt * dt), kh);
  dyn_dense(&D, n, g / (b * dt), kh);
  for (int64_t i = 0; i < n; i++) {
    u[i] = u0[i];
    v[i] = v0[i];
    history[i] = u[i];
  }
  for (int64_t step = 1; step <= nsteps && !rc; step++) {
    const double sk = dyn_scale(fscale, step);
    for (int64_t i = 0; i < n; i++) {
      pm[i] = u[i] / (b * dt * dt) + v[i] / (b * dt) + (0.5 / b - 1.0) * a[i];
      pc[i] = g / (b * dt) * u[i] + (g / b - 1.0) * v[i] + dt * (g / (2.0 * b) - 1.0) * a[i];
    }
    dyn_matvec(&M, n, pm, mp);
    dyn_matvec(&D, n, pc, dp);
    for (int64_t i = 0; i < n; i++) rhs[i] = sk * F[i] + mp[i] + dp[i];
    memcpy(dense, kh, sizeof(double) * nn * nn);
    if (orc_dense_solve(n, dense, rhs)) { rc = -1; break; }
    for (int64_t i = 0; i < n; i++) {
      const double an = (rhs[i] - u[i]) / (b * dt * dt) - v[i] / (b * dt) - (0.5 / b - 1.0) * a[i];
      v[i] = v[i] + dt * ((1.0 - g) * a[i] + g * an);
      a[i] = an;
      u[i] = rhs[i];
    }
    memcpy(history + (size_t)step * n, u, sizeof(double) * nn);
  }
  free(dv); free(ku); free(a); free(u); free(v); free(pm); free(pc); free(rhs); free(mp); free(dp);
  free(dense); free(kh);
  return rc;
}
