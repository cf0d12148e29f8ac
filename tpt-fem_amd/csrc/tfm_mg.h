// tfm_mg.h -- device side of the grid operator shared by the multigrid
// V-cycle (tfm_mg.hip) and the batched CG's matrix-free product
// (tfm_batch.hip).
//
// Designs are P concatenated topopt grids (Grid::new(nx, ny, h),
// crates/tpt-fem-topopt/src/lib.rs:89-116): node n = j nx + i of design p
// owns DOFs p ndof + 2 n + {0, 1}, element e = ej (nx - 1) + ei has nodes
// (bottom-left, bottom-right, top-right, top-left).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tfm {

constexpr double kMgOmega = 0.6;   // Jacobi weight; lambda_max(D^-1 K) ~ 2.8 for Q4 plane stress
constexpr double kMg3Omega = 0.5;  // box meshes; lambda_max(D^-1 K) ~ 3.1-3.3 for P1-tet elasticity

// Operator of one level.
struct MgOp {
  int nx, ny;
  int64_t ndof, ne;             // per design
  const double *E, *ke0;        // level 0: element moduli and unit stiffness
  const unsigned char *fixed;   // level 0 (may be null)
  const double *K;              // levels >= 1: element matrices (null at level 0)
};

// Local index of node (i, j) in element (ei, ej), and the inverse map.
__device__ __forceinline__ int elem_local(int i, int j, int ei, int ej) {
  const int di = i - ei, dj = j - ej;
  return dj == 0 ? di : 3 - di;
}
__device__ __forceinline__ int local_di(int l) { return l == 1 || l == 2; }
__device__ __forceinline__ int local_dj(int l) { return l >= 2; }

// (A x) at node (i, j) of design p, both components; ld(d) returns x at DOF
// d (already zero at fixed DOFs, so the level-0 operator needs no column
// mask).  FINE: level 0 (E_e Ke_unit, Ke_unit read at compile-time offsets);
// else the level's element matrices.  The 3 x 3 neighbourhood is loaded once
// (18 values; nodes off the grid read as 0), then the node's four elements,
// slots q (ei = i - 1 + (q & 1), ej = j - 1 + (q >> 1)), add their rows: the
// node's local index in slot q and every neighbour's place are constants.
template <bool FINE, typename LD>
__device__ __forceinline__ void apply_node(const MgOp &o, int p, int i, int j, const LD &ld, double (&y)[2]) {
  const int nex = o.nx - 1, ney = o.ny - 1;
  const int g0 = p * (int)o.ndof + 2 * (j * o.nx + i);
  double xn[3][3][2];
#pragma unroll
  for (int dy = -1; dy <= 1; dy++)
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
      const bool in = i + dx >= 0 && i + dx < o.nx && j + dy >= 0 && j + dy < o.ny;
      const int d = g0 + 2 * (dy * o.nx + dx);
      xn[dy + 1][dx + 1][0] = in ? ld(d) : 0.0;
      xn[dy + 1][dx + 1][1] = in ? ld(d + 1) : 0.0;
    }
  y[0] = y[1] = 0.0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int dxe = q & 1, dye = q >> 1;
    const int ei = i - 1 + dxe, ej = j - 1 + dye;
    if (ei < 0 || ei >= nex || ej < 0 || ej >= ney) continue;
    const int ln = dye == 1 ? 1 - dxe : 2 + dxe;  // local index of (1 - dxe, 1 - dye)
    const int e = p * (int)o.ne + ej * nex + ei;
    const double *Ke = FINE ? o.ke0 : o.K + (int64_t)e * 64;
    double t0 = 0.0, t1 = 0.0;
#pragma unroll
    for (int lm = 0; lm < 4; lm++) {
      const int oy = dye + local_dj(lm), ox = dxe + local_di(lm);  // neighbour (ox - 1, oy - 1)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const double xv = xn[oy][ox][b];
        t0 += Ke[(2 * ln) * 8 + 2 * lm + b] * xv;
        t1 += Ke[(2 * ln + 1) * 8 + 2 * lm + b] * xv;
      }
    }
    const double s = FINE ? o.E[e] : 1.0;
    y[0] += s * t0;
    y[1] += s * t1;
  }
}


// Segments (design p, rhs k) a V-cycle skips: done[(p * NR + k) * stride]
// non-zero (the batched CG's per-segment stop flags); done null skips none.
struct MgAct {
  const int *done;
  int stride;
};
template <int NR>
__device__ __forceinline__ bool skipped(const MgAct &a, int p, int k) {
  return a.done && a.done[(p * NR + k) * a.stride];
}

// One axis of the interpolation.  A coarsened axis of n fine elements has
// coarse nodes at fine positions X(I) = min(2 I, n), I = 0..ceil(n / 2); a
// fine node either coincides with a coarse node or is the midpoint of two.
__device__ __forceinline__ int coarse_pos(int I, int n, int co) { return co ? min(2 * I, n) : I; }

__device__ __forceinline__ void interp1(int i, int n, int co, int &I0, int &I1, double &w1) {
  if (!co) {
    I0 = I1 = i;
    w1 = 0.0;
  } else if (i == n) {
    I0 = I1 = (n + 1) / 2;
    w1 = 0.0;
  } else if ((i & 1) == 0) {
    I0 = I1 = i / 2;
    w1 = 0.0;
  } else {
    I0 = (i - 1) / 2;
    I1 = I0 + 1;
    w1 = 0.5;
  }
}

// Weight of coarse node I in the interpolation at fine node i.
__device__ __forceinline__ double hat1(int i, int I, int n, int co) {
  int I0, I1;
  double w1;
  interp1(i, n, co, I0, I1, w1);
  if (I0 == I1) return I0 == I ? 1.0 : 0.0;
  return I0 == I ? 1.0 - w1 : (I1 == I ? w1 : 0.0);
}

// In-place inverse of the SPD nc x nc matrix A in LDS by the Gauss-Jordan
// sweep (no pivoting), the whole block cooperating; d0 holds A's diagonal.
// A DOF whose pivot vanishes relative to its diagonal -- one no coarse basis
// function reaches -- gets a zero row and column.  Ends with a barrier.
__device__ __forceinline__ void gj_invert_lds(double *A, const double *d0, int nc) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < nc; k++) {
    __syncthreads();
    const double piv = A[k * nc + k];
    const bool bad = !(d0[k] > 0.0) || !(piv > 1e-13 * d0[k]);
    __syncthreads();
    if (bad) {
      for (int q = tid; q < nc; q += nt) {
        A[k * nc + q] = 0.0;
        A[q * nc + k] = 0.0;
      }
      continue;
    }
    for (int q = tid; q < nc; q += nt)
      if (q != k) A[k * nc + q] /= piv;
    __syncthreads();
    for (int q = tid; q < nc * nc; q += nt) {
      const int r = q / nc, c = q % nc;
      if (r != k && c != k) A[q] -= A[r * nc + k] * A[k * nc + c];
    }
    __syncthreads();
    for (int q = tid; q < nc; q += nt)
      if (q != k) A[q * nc + k] = -A[q * nc + k] / piv;
    if (tid == 0) A[k * nc + k] = 1.0 / piv;
  }
  __syncthreads();
}

// Stores the symmetrised inverse (A + A^T) / 2.
__device__ __forceinline__ void gj_store_sym(const double *A, int nc, double *__restrict__ out) {
  for (int q = threadIdx.x; q < nc * nc; q += blockDim.x) {
    const int r = q / nc, c = q % nc;
    out[q] = 0.5 * (A[q] + A[c * nc + r]);
  }
}

}  // namespace tfm
