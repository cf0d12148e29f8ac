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
// their coarse stiffness.  Smoothing: one weighted-Jacobi sweep before and
// one after the coarse correction, which keeps the V-cycle symmetric positive
// definite (the pre-sweep from zero is fused into the residual); coarsest
// level: the dense inverse per design, formed on the GPU by an in-place
// Gauss-Jordan sweep in LDS.
//
// Vectors follow the batched solver's layout: [P * 2 * nodes][NR], NR right-
// hand sides interleaved; one thread per (node, rhs) with the rhs fastest.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tfm_internal.h"
#include "tfm_mg.h"

namespace tfm {

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
  double *d_wd0 = nullptr;   // [P * ndof0] w Dinv of level 0 (the first sweep's weights)
  int nc = 0;
};

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

// Thread -> (design, node, rhs) of a level, rhs fastest; false past the end.
template <int NR>
__device__ __forceinline__ bool node_thread(int P, int nn, int &p, int &n, int &k) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * nn * NR) return false;
  k = t % NR;
  const int pn = t / NR;
  p = pn / nn;
  n = pn - p * nn;
  return true;
}

__global__ void km_young(int64_t n, const double *__restrict__ rho, double penal, double e0, double emin,
                         double *__restrict__ E) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) E[t] = emin + pow(rho[t], penal) * (e0 - emin);
}

// Jacobi diagonal inverse of a level (0 at fixed or zero-diagonal DOFs) and,
// when wd is given, the first sweep's weights w Dinv.
__global__ void km_dinv(int P, MgOp o, double *__restrict__ dinv, double *__restrict__ wd) {
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
    const double di = (!is_fixed(o, g) && d[a] > 0.0) ? 1.0 / d[a] : 0.0;
    dinv[g] = di;
    if (wd) wd[g] = kMgOmega * di;
  }
}

// Descent with one pre-smoothing sweep from zero, x1 = w Dinv b (formed by
// whoever wrote b): r = b - A x1.  Rows with a zero Jacobi weight (fixed
// DOFs, or DOFs no basis function reaches) get r = 0.
template <int NR, bool FINE>
__global__ __launch_bounds__(kMgB) void km_resid0(int P, MgOp o, const double *__restrict__ dinv,
                                                  const double *__restrict__ b, const double *__restrict__ x1,
                                                  double *__restrict__ r, MgAct act) {
  int p, n, k;
  if (!node_thread<NR>(P, o.nx * o.ny, p, n, k) || skipped<NR>(act, p, k)) return;
  double y[2];
  apply_node<FINE>(o, p, n % o.nx, n / o.nx, [&](int d) { return x1[d * NR + k]; }, y);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int g = p * (int)o.ndof + 2 * n + a;
    r[g * NR + k] = dinv[g] == 0.0 ? 0.0 : b[g * NR + k] - y[a];
  }
}

// Post-smoothing sweep: out = y + w Dinv (b - A y).
template <int NR, bool FINE>
__global__ __launch_bounds__(kMgB) void km_smooth(int P, MgOp o, const double *__restrict__ dinv,
                                                  const double *__restrict__ b, const double *__restrict__ yv,
                                                  double *__restrict__ out, MgAct act) {
  int p, n, k;
  if (!node_thread<NR>(P, o.nx * o.ny, p, n, k) || skipped<NR>(act, p, k)) return;
  double y[2];
  apply_node<FINE>(o, p, n % o.nx, n / o.nx, [&](int d) { return yv[d * NR + k]; }, y);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int g = p * (int)o.ndof + 2 * n + a;
    out[g * NR + k] = yv[g * NR + k] + kMgOmega * dinv[g] * (b[g * NR + k] - y[a]);
  }
}

// Restriction (transpose of the interpolation): bc = P^T rf, and x1c = w
// Dinv_c bc when the coarse level smooths (x1c non-null).
template <int NR>
__global__ __launch_bounds__(kMgB) void km_restrict(int P, int nxf, int nyf, int ndf, int nxc, int nyc, int ndc,
                                                    int cox, int coy, const double *__restrict__ rf,
                                                    double *__restrict__ bc, const double *__restrict__ dinvc,
                                                    double *__restrict__ x1c, MgAct act) {
  int p, n, k;
  if (!node_thread<NR>(P, nxc * nyc, p, n, k) || skipped<NR>(act, p, k)) return;
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
      const int g = p * ndf + 2 * (j * nxf + i);
      acc[0] += w * rf[g * NR + k];
      acc[1] += w * rf[(g + 1) * NR + k];
    }
  }
  const int g = p * ndc + 2 * n;
  bc[g * NR + k] = acc[0];
  bc[(g + 1) * NR + k] = acc[1];
  if (x1c) {  // the coarse level's pre-smoothed start w Dinv b
    x1c[g * NR + k] = kMgOmega * dinvc[g] * acc[0];
    x1c[(g + 1) * NR + k] = kMgOmega * dinvc[g + 1] * acc[1];
  }
}

// Coarse correction on top of the pre-smoothed x1: y = x1 + P xc, 0 where
// the Jacobi weight is 0 (fixed DOFs).
template <int NR>
__global__ __launch_bounds__(kMgB) void km_prolong0(int P, int nxf, int nyf, int ndf, int nxc, int ndc, int cox,
                                                    int coy, const double *__restrict__ dinv,
                                                    const double *__restrict__ x1, const double *__restrict__ xc,
                                                    double *__restrict__ y, MgAct act) {
  int p, n, k;
  if (!node_thread<NR>(P, nxf * nyf, p, n, k) || skipped<NR>(act, p, k)) return;
  const int i = n % nxf, j = n / nxf;
  int I0, I1, J0, J1;
  double wx1, wy1;
  interp1(i, nxf - 1, cox, I0, I1, wx1);
  interp1(j, nyf - 1, coy, J0, J1, wy1);
  const double wx0 = 1.0 - wx1, wy0 = 1.0 - wy1;
  const int cb = p * ndc;
  const int c00 = cb + 2 * (J0 * nxc + I0), c10 = cb + 2 * (J0 * nxc + I1);
  const int c01 = cb + 2 * (J1 * nxc + I0), c11 = cb + 2 * (J1 * nxc + I1);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int g = p * ndf + 2 * n + a;
    const double d = dinv[g];
    const double v = wy0 * (wx0 * xc[(c00 + a) * NR + k] + wx1 * xc[(c10 + a) * NR + k]) +
                     wy1 * (wx0 * xc[(c01 + a) * NR + k] + wx1 * xc[(c11 + a) * NR + k]);
    y[g * NR + k] = d == 0.0 ? 0.0 : x1[g * NR + k] + v;
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
  const bool fine = f.K == nullptr;
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
      const int64_t e = (int64_t)p * f.ne + (int64_t)ej * (f.nx - 1) + ei;
      // a level-0 child is E_e Ke_unit with its fixed rows and columns zero:
      // its modulus and the fixed flags of its row (a) and column (b) DOFs
      double sc = 1.0;
      bool fr[4] = {false, false, false, false}, fc[4] = {false, false, false, false};
      if (fine) {
        sc = f.E[e];
        const int64_t base = (int64_t)p * f.ndof;
        for (int l = 0; l < 4; l++) {
          const int64_t g = base + 2 * ((int64_t)(ej + local_dj(l)) * f.nx + ei + local_di(l));
          fr[l] = is_fixed(f, g + a);
          fc[l] = is_fixed(f, g + b);
        }
      }
      const double *Kk = fine ? f.ke0 : f.K + e * 64;
      for (int l = 0; l < 4; l++) {
        if (wr[l] == 0.0 || fr[l]) continue;
        double row = 0.0;
        for (int m = 0; m < 4; m++)
          if (wc[m] != 0.0 && !fc[m]) row += Kk[(2 * l + a) * 8 + 2 * m + b] * wc[m];
        acc += wr[l] * sc * row;
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
  gj_invert_lds(A, d0, nc);
  gj_store_sym(A, nc, inv + (int64_t)p * nc * nc);
}

// Coarsest solve x = A^-1 b, one thread per (design, row, rhs); the inverse
// is symmetric, so row r is read as column r (coalesced).
template <int NR>
__global__ __launch_bounds__(kMgB) void km_coarse_solve(int P, int nc, const double *__restrict__ inv,
                                                        const double *__restrict__ b, double *__restrict__ x,
                                                        MgAct act) {
  int p, r, k;
  if (!node_thread<NR>(P, nc, p, r, k) || skipped<NR>(act, p, k)) return;
  const double *Ai = inv + (int64_t)p * nc * nc;
  const double *bp = b + p * nc * NR;
  double s = 0.0;
  for (int c = 0; c < nc; c++) s += Ai[c * nc + r] * bp[c * NR + k];
  x[(p * nc + r) * NR + k] = s;
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
  dfree(mg->d_wd0);
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
  if ((int64_t)P * mg->lv[0].ndof * NR >= INT32_MAX || (int64_t)P * mg->lv[0].ne * 64 >= INT32_MAX) {
    delete mg;
    return fail(ctx, TFM_E_UNSUPPORTED, "multigrid: batch too large for int32 indexing");
  }
  int rc = TFM_OK;
  for (int l = 0; l <= mg->L && !rc; l++) {
    MgLevel &v = mg->lv[l];
    const size_t nv = (size_t)P * v.ndof * NR;
    // level 0 works in the caller's r (as b) and z (as x)
    if (l > 0 && ((rc = dalloc(ctx, &v.b, nv)) || (rc = dalloc(ctx, &v.x, nv)) ||
                  (rc = dalloc(ctx, &v.K, (size_t)P * v.ne * 64))))
      break;
    // the Jacobi diagonal of level 0 also masks the batched CG's product
    if ((l < mg->L || l == 0) && (rc = dalloc(ctx, &v.dinv, (size_t)P * v.ndof))) break;
    if (l < mg->L && ((rc = dalloc(ctx, &v.x2, nv)) || (rc = dalloc(ctx, &v.r, nv)))) break;
  }
  if (!rc && !(rc = dalloc(ctx, &mg->d_ke0, 64)) && !(rc = dalloc(ctx, &mg->d_E, (size_t)P * mg->lv[0].ne)) &&
      !(rc = dalloc(ctx, &mg->d_wd0, (size_t)P * mg->lv[0].ndof)) &&
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
    if (l < mg->L || l == 0)
      hipLaunchKernelGGL(km_dinv, dim3(mg_grid((int64_t)P * v.nx * v.ny)), dim3(kMgB), 0, st, P, level_op(mg, l),
                         v.dinv, l == 0 ? mg->d_wd0 : nullptr);
  }
  hipLaunchKernelGGL(km_coarse_inv, dim3((unsigned)P), dim3(256), 0, st, level_op(mg, mg->L), mg->nc, mg->d_inv);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

// z = M r, one V-cycle with one Jacobi sweep before and one after the coarse
// correction.  On entry z holds w D_0^-1 r (the caller forms it with the
// residual; see mg_sweep_weights).  Level l (b_0 = r, x1_l = w D_l^-1 b_l):
// r_l = b_l - A_l x1_l, b_{l+1} = P^T r_l (with x1_{l+1}, in x); coarsest
// x_L = A_L^-1 b_L; back up, y_l = x1_l + P x_{l+1} (in x2) and x_l = y_l +
// w D_l^-1 (b_l - A_l y_l) (z at level 0, x below).
template <int NR>
static void mg_vcycle(const MgData *mg, const double *r0, double *z0, hipStream_t st, MgAct act, bool last) {
  const int P = mg->P, L = mg->L;
  auto rhs = [&](int l) -> const double * { return l == 0 ? r0 : mg->lv[l].b; };
  auto sol = [&](int l) -> double * { return l == 0 ? z0 : mg->lv[l].x; };
  for (int l = 0; l < L; l++) {
    const MgLevel &v = mg->lv[l], &c = mg->lv[l + 1];
    const unsigned g = mg_grid((int64_t)P * v.nx * v.ny * NR);
    if (l == 0)
      hipLaunchKernelGGL((km_resid0<NR, true>), dim3(g), dim3(kMgB), 0, st, P, level_op(mg, l),
                         (const double *)v.dinv, rhs(l), (const double *)sol(l), v.r, act);
    else
      hipLaunchKernelGGL((km_resid0<NR, false>), dim3(g), dim3(kMgB), 0, st, P, level_op(mg, l),
                         (const double *)v.dinv, rhs(l), (const double *)sol(l), v.r, act);
    const bool smooths = l + 1 < L;
    hipLaunchKernelGGL((km_restrict<NR>), dim3(mg_grid((int64_t)P * c.nx * c.ny * NR)), dim3(kMgB), 0, st, P, v.nx,
                       v.ny, (int)v.ndof, c.nx, c.ny, (int)c.ndof, c.cox, c.coy, (const double *)v.r, c.b,
                       (const double *)(smooths ? c.dinv : nullptr), smooths ? c.x : nullptr, act);
  }
  hipLaunchKernelGGL((km_coarse_solve<NR>), dim3(mg_grid((int64_t)P * mg->nc * NR)), dim3(kMgB), 0, st, P, mg->nc,
                     (const double *)mg->d_inv, rhs(L), sol(L), act);
  for (int l = L - 1; l >= 0; l--) {
    const MgLevel &v = mg->lv[l], &c = mg->lv[l + 1];
    const unsigned g = mg_grid((int64_t)P * v.nx * v.ny * NR);
    hipLaunchKernelGGL((km_prolong0<NR>), dim3(g), dim3(kMgB), 0, st, P, v.nx, v.ny, (int)v.ndof, c.nx, (int)c.ndof,
                       c.cox, c.coy, (const double *)v.dinv, (const double *)sol(l), (const double *)c.x, v.x2, act);
    if (l == 0) {
      if (last)
        hipLaunchKernelGGL((km_smooth<NR, true>), dim3(g), dim3(kMgB), 0, st, P, level_op(mg, l),
                           (const double *)v.dinv, rhs(l), (const double *)v.x2, sol(l), act);
    } else
      hipLaunchKernelGGL((km_smooth<NR, false>), dim3(g), dim3(kMgB), 0, st, P, level_op(mg, l),
                         (const double *)v.dinv, rhs(l), (const double *)v.x2, sol(l), act);
  }
}

const double *mg_apply(const MgData *mg, const double *r, double *z, hipStream_t st, const int *done, int stride,
                       bool last) {
  const MgAct act{done, stride};
  if (mg->NR == 1) mg_vcycle<1>(mg, r, z, st, act, last);
  else mg_vcycle<8>(mg, r, z, st, act, last);
  return (!last && mg->L > 0) ? mg->lv[0].x2 : nullptr;
}

const double *mg_sweep_weights(const MgData *mg) { return mg->L > 0 ? mg->d_wd0 : nullptr; }

MgOp mg_fine_op(const MgData *mg) { return level_op(mg, 0); }

const double *mg_fine_dinv(const MgData *mg) { return mg->lv[0].dinv; }

}  // namespace tfm
