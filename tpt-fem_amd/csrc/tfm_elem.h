// tfm_elem.h -- per-element operators evaluated on the GPU, one "row block"
// at a time: the DPN rows of element-local node `ln` against all NN*DPN
// element columns.
//
// Exactness.  The FP64 path reproduces the reference's element arithmetic
// operation for operation (this file must be compiled without FMA
// contraction): Map::from_nodes_and_grad, mat_det, mat_inv, physical_grad
// (crates/tpt-fem-element/src/lib.rs:624-745), poisson_element_matrix
// (crates/tpt-fem-thermal/src/lib.rs:132-160) and elasticity_element_matrix
// (crates/tpt-fem-elasticity/src/lib.rs:262-309), summed over quadrature
// points in rule order.  Two identities keep it cheap without changing a
// single non-zero result:
//   * a product with a structural 0 of the reference gradient/B/D tables is
//     +-0, and x + (+-0) == x for every x that is not itself a zero, so those
//     terms are dropped (only the sign of an exactly-zero sum could differ,
//     and exactly-zero entries are skipped by assemble() either way);
//   * multiplications by +-1 are exact.
// Results are therefore bit-identical to the reference, which the parity
// tests check against the CPU oracle entry by entry.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/tptfem.h"

namespace tfm {

// Quadrature rule + reference tables, built on the host and uploaded once
// per plan.  Layout per point q: w, xi[3], ns[NN], lg[NN][DIM].
constexpr int kMaxQp = 125;  // Hex8 with Gauss-Legendre order 5
constexpr int kRuleStride = 1 + 3 + 8 + 8 * 3;

struct OpParams {
  int kind, model, nq;
  double coef, nu, penal, e0, emin;
  const double *coef_e;  // device, per element, or nullptr
  const double *rho;     // device, per element (SIMP), or nullptr
  const double *rule;    // device [nq][kRuleStride]
  const double *ke0;     // device [64] (SIMP unit stiffness)
  const float *ke32;     // device [n_elem][W][W] FP32 element matrices, or nullptr
};

// MI355X dispatches the workgroups of a launch round-robin over its 8 XCDs,
// each with a private L2.  Work items that share data (consecutive node tiles,
// consecutive row chunks) should land on the same XCD: launch a grid of
// 8 * ceil(nwork / 8) blocks and map block b to work item
// (b % 8) * ceil(nwork / 8) + b / 8 -- a bijection; items >= nwork exit.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item(int b, int nwork) {
  const int q = (nwork + kXcds - 1) / kXcds;
  return (b % kXcds) * q + b / kXcds;
}
inline int xcd_grid(int nwork) { return kXcds * ((nwork + kXcds - 1) / kXcds); }

template <int CELL> struct Cell;
template <> struct Cell<TFM_LINE> { static constexpr int NN = 2, DIM = 1; };
template <> struct Cell<TFM_TRI> { static constexpr int NN = 3, DIM = 2; };
template <> struct Cell<TFM_QUAD> { static constexpr int NN = 4, DIM = 2; };
template <> struct Cell<TFM_TET> { static constexpr int NN = 4, DIM = 3; };
template <> struct Cell<TFM_HEX> { static constexpr int NN = 8, DIM = 3; };

// Constitutive constants (crates/tpt-fem-elasticity/src/lib.rs:197-238).
struct DConst {
  double d00, d01, d10, d11, d22;  // 2-D: [[d00 d01 0][d10 d11 0][0 0 d22]]
  double m, cnu, dd;               // 3-D: diag m, off-diag c*nu, shear dd
  double e;                        // 1-D bar
};

__device__ __forceinline__ DConst make_d(int model, double e, double nu) {
  DConst D{};
  if (model == TFM_BAR) {
    D.e = e;
  } else if (model == TFM_PLANE_STRESS) {
    double c = e / (1.0 - nu * nu);
    D.d00 = c; D.d01 = c * nu; D.d10 = c * nu; D.d11 = c;
    D.d22 = c * (1.0 - nu) / 2.0;
  } else if (model == TFM_PLANE_STRAIN) {
    double c = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
    D.d00 = c * (1.0 - nu); D.d01 = c * nu; D.d10 = c * nu; D.d11 = c * (1.0 - nu);
    D.d22 = c * (1.0 - 2.0 * nu) / 2.0;
  } else {
    double c = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
    D.m = (1.0 - nu) * c;
    D.cnu = c * nu;
    D.dd = (1.0 - 2.0 * nu) / 2.0 * c;
  }
  return D;
}

// The reference's sum of four equal quadrature contributions, ((c + c) + c)
// + c from 0.0, as one multiplication: in binary round-to-nearest-even the
// two are bitwise equal.  c + c = 2c and fl(2c + c) = fl(3c) = 4c - d; 4c is
// representable and |d| <= ulp(3c) / 2, which is <= ulp(4c) / 4 when 3c lies
// below 4c's binade and <= ulp(4c) / 2 otherwise, so fl(4c - d) = 4c except
// at a tie, which needs c's integer significand M = 2 (mod 4) and then
// rounds to 4c, whose significand M is even.  (+0.0 + c only differs for c =
// -0.0, an entry assemble() skips either way.)
__device__ __forceinline__ double quad4(double c) { return 4.0 * c; }

// s(a,b) = sum_s sum_t B_i[s][a] D[s][t] B_j[t][b], non-zero terms in the
// reference loop order (s outer, t inner), B/D of b_matrix/constitutive.
__device__ __forceinline__ double bdb3(const double *gi, const double *gj, int a, int b,
                                       const DConst &D) {
  switch (a * 3 + b) {
  case 0: return gi[0] * D.m * gj[0] + gi[1] * D.dd * gj[1] + gi[2] * D.dd * gj[2];
  case 1: return gi[0] * D.cnu * gj[1] + gi[1] * D.dd * gj[0];
  case 2: return gi[0] * D.cnu * gj[2] + gi[2] * D.dd * gj[0];
  case 3: return gi[1] * D.cnu * gj[0] + gi[0] * D.dd * gj[1];
  case 4: return gi[1] * D.m * gj[1] + gi[0] * D.dd * gj[0] + gi[2] * D.dd * gj[2];
  case 5: return gi[1] * D.cnu * gj[2] + gi[2] * D.dd * gj[1];
  case 6: return gi[2] * D.cnu * gj[0] + gi[0] * D.dd * gj[2];
  case 7: return gi[2] * D.cnu * gj[1] + gi[1] * D.dd * gj[2];
  default: return gi[2] * D.m * gj[2] + gi[1] * D.dd * gj[1] + gi[0] * D.dd * gj[0];
  }
}

__device__ __forceinline__ double bdb2(const double *gi, const double *gj, int a, int b,
                                       const DConst &D) {
  switch (a * 2 + b) {
  case 0: return gi[0] * D.d00 * gj[0] + gi[1] * D.d22 * gj[1];
  case 1: return gi[0] * D.d01 * gj[1] + gi[1] * D.d22 * gj[0];
  case 2: return gi[1] * D.d10 * gj[0] + gi[0] * D.d22 * gj[1];
  default: return gi[1] * D.d11 * gj[1] + gi[0] * D.d22 * gj[0];
  }
}

// Map for one quadrature point: J from node coords and reference gradients,
// signed det, inverse, physical gradients of all nodes.
template <int NN, int DIM>
__device__ __forceinline__ double map_point(const double (&X)[NN][3], const double *lg,
                                            double (&g)[NN][3]) {
  double J[DIM * DIM];
#pragma unroll
  for (int i = 0; i < DIM * DIM; i++) J[i] = 0.0;
#pragma unroll
  for (int a = 0; a < NN; a++)
#pragma unroll
    for (int p = 0; p < DIM; p++)
#pragma unroll
      for (int r = 0; r < DIM; r++) J[p * DIM + r] += X[a][p] * lg[a * DIM + r];
  double det, inv[DIM * DIM];
  if constexpr (DIM == 1) {
    det = J[0];
    inv[0] = 1.0 / J[0];
  } else if constexpr (DIM == 2) {
    det = J[0] * J[3] - J[1] * J[2];
    double iv = 1.0 / det;
    inv[0] = J[3] * iv; inv[1] = -J[1] * iv; inv[2] = -J[2] * iv; inv[3] = J[0] * iv;
  } else {
    const double *m = J;
    det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
          m[2] * (m[3] * m[7] - m[4] * m[6]);
    double iv = 1.0 / det;
    inv[0] = (m[4] * m[8] - m[5] * m[7]) * iv;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) * iv;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * iv;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) * iv;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * iv;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * iv;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) * iv;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * iv;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * iv;
  }
#pragma unroll
  for (int a = 0; a < NN; a++)
#pragma unroll
    for (int i = 0; i < DIM; i++) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < DIM; r++) s += lg[a * DIM + r] * inv[r * DIM + i];
      g[a][i] = s;
    }
  return det;
}

// Tet4 specialisation: reference gradients are the constants
// (-1,-1,-1),(1,0,0),(0,1,0),(0,0,1), so J[p][r] = x_{r+1,p} - x_{0,p} and the
// physical gradients are rows of J^-1 (structural zero terms dropped).
__device__ __forceinline__ double map_tet(const double (&X)[4][3], double (&g)[4][3]) {
  double m[9];
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int r = 0; r < 3; r++) m[p * 3 + r] = -X[0][p] + X[r + 1][p];
  double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
               m[2] * (m[3] * m[7] - m[4] * m[6]);
  double iv = 1.0 / det;
  double inv[9];
  inv[0] = (m[4] * m[8] - m[5] * m[7]) * iv;
  inv[1] = (m[2] * m[7] - m[1] * m[8]) * iv;
  inv[2] = (m[1] * m[5] - m[2] * m[4]) * iv;
  inv[3] = (m[5] * m[6] - m[3] * m[8]) * iv;
  inv[4] = (m[0] * m[8] - m[2] * m[6]) * iv;
  inv[5] = (m[2] * m[3] - m[0] * m[5]) * iv;
  inv[6] = (m[3] * m[7] - m[4] * m[6]) * iv;
  inv[7] = (m[1] * m[6] - m[0] * m[7]) * iv;
  inv[8] = (m[0] * m[4] - m[1] * m[3]) * iv;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    g[0][i] = -inv[i] - inv[3 + i] - inv[6 + i];
    g[1][i] = inv[i];
    g[2][i] = inv[3 + i];
    g[3][i] = inv[6 + i];
  }
  return det;
}

// Row block of element-local node ln: out[a * (NN*DPN) + j*DPN + b] =
// K_e[ln*DPN + a][j*DPN + b], FP64, exactly the reference's value.
template <int CELL, int DPN>
__device__ __forceinline__ void rowblock_f64(const double (&X)[Cell<CELL>::NN][3], int ln,
                                             double coef, double rho, const OpParams &P,
                                             double *out) {
  constexpr int NN = Cell<CELL>::NN, DIM = Cell<CELL>::DIM, W = NN * DPN;
  if (P.kind == TFM_OP_SIMP) {
    // crates/tpt-fem-topopt/src/lib.rs:224, 233-235
    double e_rho = P.emin + pow(rho, P.penal) * (P.e0 - P.emin);
#pragma unroll
    for (int a = 0; a < DPN; a++)
#pragma unroll
      for (int j = 0; j < W; j++) out[a * W + j] = e_rho * P.ke0[(ln * DPN + a) * 8 + j];
    return;
  }
#pragma unroll
  for (int i = 0; i < DPN * W; i++) out[i] = 0.0;
  if (P.kind == TFM_OP_MASS) {
    // elasticity_mass_matrix (crates/tpt-fem-elasticity/src/lib.rs:403-424):
    // m_ij = sum_q w det rho N_i N_j on every component of the diagonal
    // block, nothing across components.
    for (int q = 0; q < P.nq; q++) {
      const double *R = P.rule + q * kRuleStride;
      const double w = R[0];
      const double *ns = R + 4;
      const double *lg = R + 4 + 8;
      double g[NN][3];
      const double det = fabs(map_point<NN, DIM>(X, lg, g));
      const double wdr = w * det * coef * ns[ln];
#pragma unroll
      for (int j = 0; j < NN; j++) {
        const double v = wdr * ns[j];
#pragma unroll
        for (int a = 0; a < DPN; a++) out[a * W + j * DPN + a] += v;
      }
    }
    return;
  }
  DConst D{};
  if (P.kind == TFM_OP_ELASTICITY) D = make_d(P.model, coef, P.nu);

  if constexpr (CELL == TFM_TET) {
    // Four identical quadrature contributions (constant gradients, equal
    // weights): the reference adds c four times starting from 0.0 (quad4).
    double g[4][3];
    double det = fabs(map_tet(X, g));
    const double w = 1.0 / 24.0;
    if (P.kind == TFM_OP_POISSON) {
      double wk = w * coef;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        double dot = g[ln][0] * g[j][0] + g[ln][1] * g[j][1] + g[ln][2] * g[j][2];
        double c = wk * dot * det;
        out[j] = quad4(c);
      }
    } else {
      double wd = w * det;
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int b = 0; b < 3; b++) {
            double c = wd * bdb3(g[ln], g[j], a, b, D);
            out[a * W + j * 3 + b] = quad4(c);
          }
    }
    return;
  }

  for (int q = 0; q < P.nq; q++) {
    const double *R = P.rule + q * kRuleStride;
    double w = R[0];
    const double *lg = R + 4 + 8;
    double g[NN][3];
    double det = fabs(map_point<NN, DIM>(X, lg, g));
    if (P.kind == TFM_OP_POISSON) {
      double wk = w * coef;
#pragma unroll
      for (int j = 0; j < NN; j++) {
        double dot = g[ln][0] * g[j][0];
        if constexpr (DIM >= 2) dot = dot + g[ln][1] * g[j][1];
        if constexpr (DIM >= 3) dot = dot + g[ln][2] * g[j][2];
        out[j] += wk * dot * det;
      }
    } else {
      double wd = w * det;
#pragma unroll
      for (int a = 0; a < DPN; a++)
#pragma unroll
        for (int j = 0; j < NN; j++)
#pragma unroll
          for (int b = 0; b < DPN; b++) {
            double s;
            if constexpr (DIM == 3) s = bdb3(g[ln], g[j], a, b, D);
            else if constexpr (DIM == 2) s = bdb2(g[ln], g[j], a, b, D);
            else s = g[ln][0] * D.e * g[j][0];
            out[a * W + j * DPN + b] += wd * s;
          }
    }
  }
}

// FP32 row block for the precision form: geometry (J, det, J^-1) in FP64,
// gradients rounded to FP32, B^T D B and the quadrature sum in FP32.
template <int CELL, int DPN>
__device__ __forceinline__ void rowblock_f32(const double (&X)[Cell<CELL>::NN][3], int ln,
                                             double coef, double rho, const OpParams &P,
                                             float *out) {
  constexpr int NN = Cell<CELL>::NN, W = NN * DPN;
  if constexpr (CELL == TFM_TET) {
    if (P.kind == TFM_OP_POISSON || P.kind == TFM_OP_ELASTICITY) {
      double gd[4][3];
      double det = fabs(map_tet(X, gd));
      float g[4][3];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) g[a][i] = (float)gd[a][i];
      // the four equal-weight points sum to vol = det/6
      float vol = (float)(det * (1.0 / 6.0));
      if (P.kind == TFM_OP_POISSON) {
        float wk = vol * (float)coef;
#pragma unroll
        for (int j = 0; j < 4; j++)
          out[j] = wk * (g[ln][0] * g[j][0] + g[ln][1] * g[j][1] + g[ln][2] * g[j][2]);
      } else {
        double c = coef / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
        float m = (float)((1.0 - P.nu) * c), cnu = (float)(c * P.nu),
              dd = (float)((1.0 - 2.0 * P.nu) / 2.0 * c);
        const float *gi = g[ln];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const float *gj = g[j];
          float *o = out + j * 3;
          o[0] = vol * (gi[0] * m * gj[0] + dd * (gi[1] * gj[1] + gi[2] * gj[2]));
          o[1] = vol * (gi[0] * cnu * gj[1] + gi[1] * dd * gj[0]);
          o[2] = vol * (gi[0] * cnu * gj[2] + gi[2] * dd * gj[0]);
          o[W + 0] = vol * (gi[1] * cnu * gj[0] + gi[0] * dd * gj[1]);
          o[W + 1] = vol * (gi[1] * m * gj[1] + dd * (gi[0] * gj[0] + gi[2] * gj[2]));
          o[W + 2] = vol * (gi[1] * cnu * gj[2] + gi[2] * dd * gj[1]);
          o[2 * W + 0] = vol * (gi[2] * cnu * gj[0] + gi[0] * dd * gj[2]);
          o[2 * W + 1] = vol * (gi[2] * cnu * gj[1] + gi[1] * dd * gj[2]);
          o[2 * W + 2] = vol * (gi[2] * m * gj[2] + dd * (gi[1] * gj[1] + gi[0] * gj[0]));
        }
      }
      return;
    }
  }
  // Other cells: evaluate in FP64 and round (their cost is not on the
  // precision-form workloads).
  double tmp[DPN * W];
  rowblock_f64<CELL, DPN>(X, ln, coef, rho, P, tmp);
#pragma unroll
  for (int i = 0; i < DPN * W; i++) out[i] = (float)tmp[i];
}

}  // namespace tfm
