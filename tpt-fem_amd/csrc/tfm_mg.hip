// tfm_mg.hip -- geometric multigrid V-cycle for batched SIMP grids.
//
// Preconditioner of the batched PCG for P concatenated topopt grids
// (Grid::new(nx, ny, h), crates/tpt-fem-topopt/src/lib.rs:89-116; Q4 plane
// stress, 2 DOFs per node n = j nx + i, element matrices E_e Ke_unit with
// E_e = Emin + rho_e^p (E0 - Emin), :166-238 and :346).  The reference solves
// these systems directly; Jacobi PCG needs thousands of iterations on them
// (stiffness contrast 1e3 to 1e9), CG preconditioned by this V-cycle a few
// tens, independent of the grid size.
//
// Levels: every axis with at least two elements is coarsened 2:1 (an odd
// element count leaves one coarse element spanning a single fine element),
// until a design has at most kMgMaxC DOFs.  The fine operator is applied
// matrix-free (E_e Ke_unit with the fixed DOFs' rows and columns removed);
// each coarse operator is the Galerkin product P^T A P of the level below,
// formed element by element (K_E = sum over E's children of P_c^T K_c P_c,
// P_c the bilinear interpolation inside E), so soft and stiff regions keep
// their coarse stiffness.  Smoothing: kMgNu weighted-Jacobi sweeps before and
// after the coarse correction, which keeps the V-cycle symmetric positive
// definite; coarsest level: the dense inverse per design, formed on the GPU
// by an in-place Gauss-Jordan sweep in LDS.
//
// Vectors follow the batched solver's layout: [P * 2 * nodes][NR], NR right-
// hand sides interleaved; one thread per (node, rhs) with the rhs fastest.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tfm_internal.h"

namespace tfm {

constexpr double kMgOmega = 0.6;  // Jacobi weight; lambda_max(D^-1 K) ~ 2.8 for Q4 plane stress
constexpr int kMgNu = 1;          // sweeps before and after the coarse correction
constexpr int kMgB = 256;
constexpr int kMgMaxC = 90;       // coarsest DOFs per design: the dense matrix fills 64 KB of LDS

struct MgLevel {
  int nx = 0, ny = 0;       // nodes per axis
  int cox = 0, coy = 0;     // coarsened from the level above along x / y
  int64_t ne = 0, ndof = 0;  // per design
  double *K = nullptr;      // [P][ne][64] element matrices (levels >= 1)
  double *dinv = nullptr;   // [P * ndof]
  double *b = nullptr, *x = nullptr, *x2 = nullptr, *r = nullptr;  // [P * ndof][NR]
};

struct MgData {
  tfm_ctx *ctx = nullptr;
  int P = 0, NR = 1, L = 0;  // levels 0..L, L the coarsest
  std::vector<MgLevel> lv;
  double *d_ke0 = nullptr;   // [64] unit element stiffness
  double *d_E = nullptr;     // [P][ne0] element Young's moduli
  const unsigned char *d_fixed = nullptr;  // [P * ndof0] (borrowed; null: nothing fixed)
  double *d_inv = nullptr;   // [P][nc][nc] inverse of the coarsest operator
  int nc = 0;
};

// Operator of one level as the kernels see it.
struct MgOp {
  int nx, ny;
  int64_t ndof, ne;
  const double *E, *ke0;          // level 0
  const unsigned char *fixed;     // level 0 (may be null)
  const double *K;                // levels >= 1 (null at level 0)
};

// Local index of node (i, j) in element (ei, ej) (Grid::new order: bottom-
// left, bottom-right, top-right, top-left) and the inverse map.
__device__ __forceinline__ int elem_local(int i, int j, int ei, int ej) {
  const int di = i - ei, dj = j - ej;
  return dj == 0 ? di : 3 - di;
}
__device__ __forceinline__ int local_di(int l) { return l == 1 || l == 2; }
__device__ __forceinline__ int local_dj(int l) { return l >= 2; }

__device__ __forceinline__ bool is_fixed(const MgOp &o, int64_t g) { return o.fixed && o.fixed[g]; }

// Entry (r8, c8) of element (ei, ej) of design p.  Level 0: E_e Ke_unit with
// the rows and columns of fixed DOFs zero.
__device__ __forceinline__ double ke_entry(const MgOp &o, int p, int ei, int ej, int r8, int c8) {
  const int64_t e = (int64_t)p * o.ne + (int64_t)ej * (o.nx - 1) + ei;
  if (o.K) return o.K[e * 64 + r8 * 8 + c8];
  const int64_t base = (int64_t)p * o.ndof;
  const int lr = r8 >> 1, lc = c8 >> 1;
  const int64_t gr = base + 2 * ((int64_t)(ej + local_dj(lr)) * o.nx + ei + local_di(lr)) + (r8 & 1);
  const int64_t gc = base + 2 * ((int64_t)(ej + local_dj(lc)) * o.nx + ei + local_di(lc)) + (c8 & 1);
  if (is_fixed(o, gr) || is_fixed(o, gc)) return 0.0;
  return o.E[e] * o.ke0[r8 * 8 + c8];
}

// (A x) at node (i, j) of design p, both components, right-hand side k.
template <int NR>
__device__ __forceinline__ void apply_node(const MgOp &o, int p, int i, int j, int k, const double *__restrict__ x,
                                           double (&y)[2]) {
  const int nex = o.nx - 1, ney = o.ny - 1;
  y[0] = y[1] = 0.0;
  const int64_t base = (int64_t)p * o.ndof;
  for (int ej = j - 1; ej <= j; ej++) {
    if (ej < 0 || ej >= ney) continue;
    for (int ei = i - 1; ei <= i; ei++) {
      if (ei < 0 || ei >= nex) continue;
      const int ln = elem_local(i, j, ei, ej);
      const int64_t e = (int64_t)p * o.ne + (int64_t)ej * nex + ei;
      const double *Ke = o.K ? o.K + e * 64 : o.ke0;
      const double s = o.K ? 1.0 : o.E[e];
#pragma unroll
      for (int lm = 0; lm < 4; lm++) {
        const int64_t nd = base + 2 * ((int64_t)(ej + local_dj(lm)) * o.nx + ei + local_di(lm));
#pragma unroll
        for (int b = 0; b < 2; b++) {
          if (is_fixed(o, nd + b)) continue;
          const double xv = s * x[(nd + b) * NR + k];
          y[0] += Ke[(2 * ln) * 8 + 2 * lm + b] * xv;
          y[1] += Ke[(2 * ln + 1) * 8 + 2 * lm + b] * xv;
        }
      }
    }
  }
}

__global__ void km_young(int64_t n, const double *__restrict__ rho, double penal, double e0, double emin,
                         double *__restrict__ E) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) E[t] = emin + pow(rho[t], penal) * (e0 - emin);
}

// Jacobi diagonal inverse of a level (0 at fixed or zero-diagonal DOFs).
__global__ void km_dinv(int P, MgOp o, double *__restrict__ dinv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)o.nx * o.ny;
  if (t >= P * nn) return;
  const int p = (int)(t / nn), n = (int)(t % nn);
  const int i = n % o.nx, j = n / o.nx;
  double d[2] = {0.0, 0.0};
  for (int ej = j - 1; ej <= j; ej++) {
    if (ej < 0 || ej >= o.ny - 1) continue;
    for (int ei = i - 1; ei <= i; ei++) {
      if (ei < 0 || ei >= o.nx - 1) continue;
      const int ln = elem_local(i, j, ei, ej);
      for (int a = 0; a < 2; a++) d[a] += ke_entry(o, p, ei, ej, 2 * ln + a, 2 * ln + a);
    }
  }
  for (int a = 0; a < 2; a++) {
    const int64_t g = (int64_t)p * o.ndof + 2 * n + a;
    dinv[g] = (!is_fixed(o, g) && d[a] > 0.0) ? 1.0 / d[a] : 0.0;
  }
}

// One weighted-Jacobi sweep xo = xi + w Dinv (b - A xi); xi == nullptr means
// xi = 0 (no product is formed).
template <int NR>
__global__ __launch_bounds__(kMgB) void km_smooth(int P, MgOp o, const double *__restrict__ dinv,
                                                  const double *__restrict__ b, const double *__restrict__ xi,
                                                  double *__restrict__ xo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)o.nx * o.ny;
  if (t >= P * nn * NR) return;
  const int k = (int)(t % NR);
  const int64_t pn = t / NR;
  const int p = (int)(pn / nn), n = (int)(pn % nn);
  double y[2] = {0.0, 0.0};
  if (xi) apply_node<NR>(o, p, n % o.nx, n / o.nx, k, xi, y);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int64_t g = (int64_t)p * o.ndof + 2 * n + a;
    const double x0 = xi ? xi[g * NR + k] : 0.0;
    xo[g * NR + k] = x0 + kMgOmega * dinv[g] * (b[g * NR + k] - y[a]);
  }
}

// r = b - A x (0 at fixed DOFs).
template <int NR>
__global__ __launch_bounds__(kMgB) void km_resid(int P, MgOp o, const double *__restrict__ b,
                                                 const double *__restrict__ x, double *__restrict__ r) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)o.nx * o.ny;
  if (t >= P * nn * NR) return;
  const int k = (int)(t % NR);
  const int64_t pn = t / NR;
  const int p = (int)(pn / nn), n = (int)(pn % nn);
  double y[2];
  apply_node<NR>(o, p, n % o.nx, n / o.nx, k, x, y);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int64_t g = (int64_t)p * o.ndof + 2 * n + a;
    r[g * NR + k] = is_fixed(o, g) ? 0.0 : b[g * NR + k] - y[a];
  }
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

// Restriction (transpose of the interpolation): bc = P^T rf.
template <int NR>
__global__ __launch_bounds__(kMgB) void km_restrict(int P, int nxf, int nyf, int64_t ndf, int nxc, int nyc,
                                                    int64_t ndc, int cox, int coy, const double *__restrict__ rf,
                                                    double *__restrict__ bc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)nxc * nyc;
  if (t >= P * nn * NR) return;
  const int k = (int)(t % NR);
  const int64_t pn = t / NR;
  const int p = (int)(pn / nn), n = (int)(pn % nn);
  const int I = n % nxc, J = n / nxc;
  const int X = coarse_pos(I, nxf - 1, cox), Y = coarse_pos(J, nyf - 1, coy);
  double acc[2] = {0.0, 0.0};
  for (int dj = -coy; dj <= coy; dj++) {
    const int j = Y + dj;
    if (j < 0 || j >= nyf) continue;
    const double wy = hat1(j, J, nyf - 1, coy);
    if (wy == 0.0) continue;
    for (int di = -cox; di <= cox; di++) {
      const int i = X + di;
      if (i < 0 || i >= nxf) continue;
      const double w = wy * hat1(i, I, nxf - 1, cox);
      if (w == 0.0) continue;
      const int64_t g = (int64_t)p * ndf + 2 * ((int64_t)j * nxf + i);
      acc[0] += w * rf[g * NR + k];
      acc[1] += w * rf[(g + 1) * NR + k];
    }
  }
  const int64_t g = (int64_t)p * ndc + 2 * n;
  bc[g * NR + k] = acc[0];
  bc[(g + 1) * NR + k] = acc[1];
}

// xf += P xc; fixed fine DOFs are left alone (they stay 0).
template <int NR>
__global__ __launch_bounds__(kMgB) void km_prolong(int P, int nxf, int nyf, int64_t ndf, int nxc, int64_t ndc,
                                                   int cox, int coy, const unsigned char *__restrict__ fixed,
                                                   const double *__restrict__ xc, double *__restrict__ xf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)nxf * nyf;
  if (t >= P * nn * NR) return;
  const int k = (int)(t % NR);
  const int64_t pn = t / NR;
  const int p = (int)(pn / nn), n = (int)(pn % nn);
  const int i = n % nxf, j = n / nxf;
  int I0, I1, J0, J1;
  double wx1, wy1;
  interp1(i, nxf - 1, cox, I0, I1, wx1);
  interp1(j, nyf - 1, coy, J0, J1, wy1);
  const double wx0 = 1.0 - wx1, wy0 = 1.0 - wy1;
  const int64_t cb = (int64_t)p * ndc;
  const int64_t c00 = cb + 2 * ((int64_t)J0 * nxc + I0), c10 = cb + 2 * ((int64_t)J0 * nxc + I1);
  const int64_t c01 = cb + 2 * ((int64_t)J1 * nxc + I0), c11 = cb + 2 * ((int64_t)J1 * nxc + I1);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int64_t g = (int64_t)p * ndf + 2 * n + a;
    if (fixed && fixed[g]) continue;
    const double v = wy0 * (wx0 * xc[(c00 + a) * NR + k] + wx1 * xc[(c10 + a) * NR + k]) +
                     wy1 * (wx0 * xc[(c01 + a) * NR + k] + wx1 * xc[(c11 + a) * NR + k]);
    xf[g * NR + k] += v;
  }
}

// Galerkin coarse element matrices, one thread per (design, coarse element,
// entry): K_E[r][c] = sum over E's children ch and child nodes l, m of
// W[l][mr] K_ch[2l + a][2m + b] W[m][mc], r = 2 mr + a, c = 2 mc + b, with
// W[l][M] the bilinear weight of E's node M at child node l.
__global__ void km_galerkin(int P, MgOp f, int nxc, int nyc, int cox, int coy, double *__restrict__ Kc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nexc = nxc - 1, neyc = nyc - 1;
  const int64_t nec = (int64_t)nexc * neyc;
  if (t >= P * nec * 64) return;
  const int rc = (int)(t % 64);
  const int64_t pe = t / 64;
  const int p = (int)(pe / nec), E0 = (int)(pe % nec);
  const int EI = E0 % nexc, EJ = E0 / nexc;
  const int r8 = rc >> 3, c8 = rc & 7, mr = r8 >> 1, a = r8 & 1, mc = c8 >> 1, b = c8 & 1;
  const int x0 = coarse_pos(EI, f.nx - 1, cox), x1 = coarse_pos(EI + 1, f.nx - 1, cox);
  const int y0 = coarse_pos(EJ, f.ny - 1, coy), y1 = coarse_pos(EJ + 1, f.ny - 1, coy);
  const double sx = 1.0 / (double)(x1 - x0), sy = 1.0 / (double)(y1 - y0);
  double acc = 0.0;
  for (int ej = y0; ej < y1; ej++)
    for (int ei = x0; ei < x1; ei++) {
      double wr[4], wc[4];  // weight of E's nodes mr / mc at the child's nodes
      for (int l = 0; l < 4; l++) {
        const double xs = (double)(ei + local_di(l) - x0) * sx, ys = (double)(ej + local_dj(l) - y0) * sy;
        const double w[4] = {(1.0 - xs) * (1.0 - ys), xs * (1.0 - ys), xs * ys, (1.0 - xs) * ys};
        wr[l] = w[mr];
        wc[l] = w[mc];
      }
      for (int l = 0; l < 4; l++) {
        if (wr[l] == 0.0) continue;
        double s = 0.0;
        for (int m = 0; m < 4; m++)
          if (wc[m] != 0.0) s += ke_entry(f, p, ei, ej, 2 * l + a, 2 * m + b) * wc[m];
        acc += wr[l] * s;
      }
    }
  Kc[pe * 64 + rc] = acc;
}

// Dense inverse of the coarsest operator per design (one block each): the
// matrix is gathered into LDS from its element matrices, inverted in place by
// the Gauss-Jordan sweep (SPD, no pivoting; a DOF whose pivot vanishes
// relative to its diagonal -- one no coarse basis function reaches -- gets a
// zero row and column) and symmetrised.
__global__ __launch_bounds__(256) void km_coarse_inv(MgOp o, int nc, double *__restrict__ inv) {
  __shared__ double A[kMgMaxC * kMgMaxC];
  __shared__ double d0[kMgMaxC];
  const int p = blockIdx.x, tid = threadIdx.x;
  for (int q = tid; q < nc * nc; q += 256) {
    const int r = q / nc, c = q % nc;
    const int nr = r >> 1, ncn = c >> 1;
    const int i = nr % o.nx, j = nr / o.nx, ic = ncn % o.nx, jc = ncn / o.nx;
    double s = 0.0;
    for (int ej = j - 1; ej <= j; ej++) {
      if (ej < 0 || ej >= o.ny - 1 || jc - ej < 0 || jc - ej > 1) continue;
      for (int ei = i - 1; ei <= i; ei++) {
        if (ei < 0 || ei >= o.nx - 1 || ic - ei < 0 || ic - ei > 1) continue;
        s += ke_entry(o, p, ei, ej, 2 * elem_local(i, j, ei, ej) + (r & 1), 2 * elem_local(ic, jc, ei, ej) + (c & 1));
      }
    }
    A[q] = s;
    if (r == c) d0[r] = s;
  }
  for (int k = 0; k < nc; k++) {
    __syncthreads();
    const double piv = A[k * nc + k];
    const bool bad = !(d0[k] > 0.0) || !(piv > 1e-13 * d0[k]);
    __syncthreads();
    if (bad) {
      for (int q = tid; q < nc; q += 256) {
        A[k * nc + q] = 0.0;
        A[q * nc + k] = 0.0;
      }
      continue;
    }
    for (int q = tid; q < nc; q += 256)
      if (q != k) A[k * nc + q] /= piv;
    __syncthreads();
    for (int q = tid; q < nc * nc; q += 256) {
      const int r = q / nc, c = q % nc;
      if (r != k && c != k) A[q] -= A[r * nc + k] * A[k * nc + c];
    }
    __syncthreads();
    for (int q = tid; q < nc; q += 256)
      if (q != k) A[q * nc + k] = -A[q * nc + k] / piv;
    if (tid == 0) A[k * nc + k] = 1.0 / piv;
  }
  __syncthreads();
  double *out = inv + (int64_t)p * nc * nc;
  for (int q = tid; q < nc * nc; q += 256) {
    const int r = q / nc, c = q % nc;
    out[q] = 0.5 * (A[q] + A[c * nc + r]);
  }
}

// Coarsest solve x = A^-1 b, one thread per (design, row, rhs); the inverse
// is symmetric, so row r is read as column r (coalesced).
template <int NR>
__global__ __launch_bounds__(kMgB) void km_coarse_solve(int P, int nc, const double *__restrict__ inv,
                                                        const double *__restrict__ b, double *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)P * nc * NR) return;
  const int k = (int)(t % NR);
  const int64_t pr = t / NR;
  const int p = (int)(pr / nc), r = (int)(pr % nc);
  const double *Ai = inv + (int64_t)p * nc * nc;
  const double *bp = b + (int64_t)p * nc * NR;
  double s = 0.0;
  for (int c = 0; c < nc; c++) s += Ai[(int64_t)c * nc + r] * bp[(int64_t)c * NR + k];
  x[pr * NR + k] = s;
}

static unsigned mg_grid(int64_t n) { return (unsigned)((n + kMgB - 1) / kMgB); }

static MgOp level_op(const MgData *mg, int l) {
  const MgLevel &v = mg->lv[l];
  MgOp o;
  o.nx = v.nx;
  o.ny = v.ny;
  o.ndof = v.ndof;
  o.ne = v.ne;
  o.E = mg->d_E;
  o.ke0 = mg->d_ke0;
  o.fixed = l == 0 ? mg->d_fixed : nullptr;
  o.K = l == 0 ? nullptr : v.K;
  return o;
}

void mg_destroy(MgData *mg) {
  if (!mg) return;
  for (auto &l : mg->lv) {
    dfree(l.K); dfree(l.dinv); dfree(l.b); dfree(l.x); dfree(l.x2); dfree(l.r);
  }
  dfree(mg->d_ke0);
  dfree(mg->d_E);
  dfree(mg->d_inv);
  delete mg;
}

// Level dimensions {nx, ny, cox, coy} per level; returns the coarsest index.
int mg_levels(int nx, int ny, std::vector<int> *dims) {
  int cx = nx, cy = ny, L = 0;
  if (dims) dims->assign({cx, cy, 0, 0});
  while (2 * cx * cy > kMgMaxC && (cx >= 3 || cy >= 3)) {
    const int ox = cx >= 3, oy = cy >= 3;
    if (ox) cx = cx / 2 + 1;  // ceil((cx - 1) / 2) elements
    if (oy) cy = cy / 2 + 1;
    if (dims) dims->insert(dims->end(), {cx, cy, ox, oy});
    L++;
  }
  return L;
}

int mg_create(tfm_ctx *ctx, int P, int nx, int ny, int NR, const double *ke0, MgData **out) {
  *out = nullptr;
  if (P < 1 || nx < 2 || ny < 2 || (NR != 1 && NR != 8)) return fail(ctx, TFM_E_UNSUPPORTED, "multigrid: bad grid");
  auto *mg = new MgData();
  mg->ctx = ctx;
  mg->P = P;
  mg->NR = NR;
  std::vector<int> dims;
  mg->L = mg_levels(nx, ny, &dims);
  for (int l = 0; l <= mg->L; l++) {
    MgLevel v;
    v.nx = dims[4 * l];
    v.ny = dims[4 * l + 1];
    v.cox = dims[4 * l + 2];
    v.coy = dims[4 * l + 3];
    v.ne = (int64_t)(v.nx - 1) * (v.ny - 1);
    v.ndof = 2 * (int64_t)v.nx * v.ny;
    mg->lv.push_back(v);
  }
  mg->nc = (int)mg->lv.back().ndof;
  int rc = TFM_OK;
  for (int l = 0; l <= mg->L && !rc; l++) {
    MgLevel &v = mg->lv[l];
    const size_t nv = (size_t)P * v.ndof * NR;
    // level 0 works in the caller's r (as b) and z (as x)
    if (l > 0 && ((rc = dalloc(ctx, &v.b, nv)) || (rc = dalloc(ctx, &v.x, nv)) ||
                  (rc = dalloc(ctx, &v.K, (size_t)P * v.ne * 64))))
      break;
    if (l < mg->L && ((rc = dalloc(ctx, &v.dinv, (size_t)P * v.ndof)) || (rc = dalloc(ctx, &v.x2, nv)) ||
                      (rc = dalloc(ctx, &v.r, nv))))
      break;
  }
  if (!rc && !(rc = dalloc(ctx, &mg->d_ke0, 64)) && !(rc = dalloc(ctx, &mg->d_E, (size_t)P * mg->lv[0].ne)) &&
      !(rc = dalloc(ctx, &mg->d_inv, (size_t)P * mg->nc * mg->nc)))
    if (hipMemcpy(mg->d_ke0, ke0, sizeof(double) * 64, hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(ctx, TFM_E_HIP, "multigrid: upload failed");
  if (rc) {
    mg_destroy(mg);
    return rc;
  }
  *out = mg;
  return TFM_OK;
}

// New element moduli (after a re-assembly): every level's operator, Jacobi
// diagonal and the coarsest inverse.
int mg_update(MgData *mg, const double *d_rho, double penal, double e0, double emin, const unsigned char *d_fixed) {
  tfm_ctx *ctx = mg->ctx;
  hipStream_t st = ctx->stream;
  const int P = mg->P;
  mg->d_fixed = d_fixed;
  const int64_t ne0 = (int64_t)P * mg->lv[0].ne;
  hipLaunchKernelGGL(km_young, dim3(mg_grid(ne0)), dim3(kMgB), 0, st, ne0, d_rho, penal, e0, emin, mg->d_E);
  for (int l = 0; l <= mg->L; l++) {
    MgLevel &v = mg->lv[l];
    if (l >= 1) {
      const MgOp f = level_op(mg, l - 1);
      hipLaunchKernelGGL(km_galerkin, dim3(mg_grid((int64_t)P * v.ne * 64)), dim3(kMgB), 0, st, P, f, v.nx, v.ny,
                         v.cox, v.coy, v.K);
    }
    if (l < mg->L)
      hipLaunchKernelGGL(km_dinv, dim3(mg_grid((int64_t)P * v.nx * v.ny)), dim3(kMgB), 0, st, P, level_op(mg, l),
                         v.dinv);
  }
  hipLaunchKernelGGL(km_coarse_inv, dim3((unsigned)P), dim3(256), 0, st, level_op(mg, mg->L), mg->nc, mg->d_inv);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

// z = M r, one V-cycle.  Per level, the 2 kMgNu sweeps alternate between two
// buffers so the last lands in the level's output (z at level 0, x2 below).
template <int NR>
static void mg_vcycle(const MgData *mg, const double *r0, double *z0, hipStream_t st) {
  const int P = mg->P, L = mg->L;
  auto bufs = [&](int l, double *(&bf)[2]) {
    const MgLevel &v = mg->lv[l];
    bf[0] = l == 0 ? v.x2 : v.x;
    bf[1] = l == 0 ? z0 : v.x2;
  };
  auto rhs = [&](int l) -> const double * { return l == 0 ? r0 : mg->lv[l].b; };
  for (int l = 0; l < L; l++) {
    const MgLevel &v = mg->lv[l], &c = mg->lv[l + 1];
    const MgOp o = level_op(mg, l);
    double *bf[2];
    bufs(l, bf);
    const unsigned g = mg_grid((int64_t)P * v.nx * v.ny * NR);
    hipLaunchKernelGGL((km_smooth<NR>), dim3(g), dim3(kMgB), 0, st, P, o, (const double *)v.dinv, rhs(l),
                       (const double *)nullptr, bf[0]);
    for (int s = 1; s < kMgNu; s++)
      hipLaunchKernelGGL((km_smooth<NR>), dim3(g), dim3(kMgB), 0, st, P, o, (const double *)v.dinv, rhs(l),
                         (const double *)bf[(s - 1) & 1], bf[s & 1]);
    hipLaunchKernelGGL((km_resid<NR>), dim3(g), dim3(kMgB), 0, st, P, o, rhs(l),
                       (const double *)bf[(kMgNu - 1) & 1], v.r);
    hipLaunchKernelGGL((km_restrict<NR>), dim3(mg_grid((int64_t)P * c.nx * c.ny * NR)), dim3(kMgB), 0, st, P, v.nx,
                       v.ny, v.ndof, c.nx, c.ny, c.ndof, c.cox, c.coy, (const double *)v.r, c.b);
  }
  const MgLevel &cl = mg->lv[L];
  hipLaunchKernelGGL((km_coarse_solve<NR>), dim3(mg_grid((int64_t)P * mg->nc * NR)), dim3(kMgB), 0, st, P, mg->nc,
                     (const double *)mg->d_inv, rhs(L), L == 0 ? z0 : cl.x);
  for (int l = L - 1; l >= 0; l--) {
    const MgLevel &v = mg->lv[l], &c = mg->lv[l + 1];
    const MgOp o = level_op(mg, l);
    double *bf[2];
    bufs(l, bf);
    const double *xc = l + 1 == L ? c.x : c.x2;
    const unsigned g = mg_grid((int64_t)P * v.nx * v.ny * NR);
    hipLaunchKernelGGL((km_prolong<NR>), dim3(g), dim3(kMgB), 0, st, P, v.nx, v.ny, v.ndof, c.nx, c.ndof, c.cox,
                       c.coy, o.fixed, xc, bf[(kMgNu - 1) & 1]);
    for (int s = kMgNu; s < 2 * kMgNu; s++)
      hipLaunchKernelGGL((km_smooth<NR>), dim3(g), dim3(kMgB), 0, st, P, o, (const double *)v.dinv, rhs(l),
                         (const double *)bf[(s - 1) & 1], bf[s & 1]);
  }
}

void mg_apply(const MgData *mg, const double *r, double *z, hipStream_t st) {
  if (mg->NR == 1) mg_vcycle<1>(mg, r, z, st);
  else mg_vcycle<8>(mg, r, z, st);
}

}  // namespace tfm
