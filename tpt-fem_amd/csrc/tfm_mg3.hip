// tfm_mg3.hip -- geometric multigrid V-cycle for systems on box meshes.
//
// Preconditioner of the FP64 CG for one reduced system assembled on a
// box_mesh (crates/tpt-fem-mesh-gen/src/lib.rs:422-470; node (i, j, k) is
// (i (ny + 1) + j) (nz + 1) + k, DOF dpn node + c) -- Poisson (dpn 1) or
// elasticity (dpn 3), any coefficients, jittered or not: the hierarchy only
// uses the node grid.  Jacobi PCG needs thousands of iterations on the
// 128^3 elasticity problem and the count doubles with every refinement; CG
// preconditioned by this V-cycle needs a few tens at every size.
//
// Levels: every axis with at least two elements is coarsened 2:1 (an odd
// count leaves one coarse cell spanning a single fine cell), until at most
// kMg3MaxC DOFs remain.  Level 0 is the assembled matrix itself (its products
// run on the batched solver's sliced copy); each refresh also writes its CSR
// rows into the 27-point stencil form of the coarser levels, and every level
// l >= 1 is the Galerkin product P^T A P of level l - 1 with trilinear P.
// With trilinear interpolation each coarse operator is a 27-point stencil of
// dpn x dpn blocks, stored per node.
// Smoothing: one weighted-Jacobi sweep before and one after the coarse
// correction (the V-cycle stays symmetric positive definite); coarsest: the
// dense inverse formed in LDS.  One right-hand side.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tfm_internal.h"
#include "tfm_mg.h"

namespace tfm {

constexpr int kMg3B = 256;
constexpr int kMg3MaxC = 90;  // coarsest DOFs: the dense matrix fills 64 KB of LDS
constexpr int kMg3GB = 256;   // Galerkin: (coarse node, row, column component) items per block

struct Mg3Level {
  int n[3] = {0, 0, 0};   // nodes per axis; node id (i n[1] + j) n[2] + k
  int co[3] = {0, 0, 0};  // coarsened from the level above along each axis
  int nn = 0;
  double *S = nullptr;     // [nn][27][dpn][dpn], offset d = (dx + 1) 9 + (dy + 1) 3 + dz + 1 (level 0: when L > 0)
  double *dinv = nullptr;  // [nn dpn]
  double *b = nullptr, *x = nullptr, *x2 = nullptr, *r = nullptr;  // [nn dpn]
};

struct Mg3Data {
  tfm_ctx *ctx = nullptr;
  int dpn = 1, L = 0, nc = 0;
  std::vector<Mg3Level> lv;
  double *d_wd0 = nullptr;  // w Dinv of level 0 (the first sweep's weights)
  double *d_inv = nullptr;  // [nc][nc] inverse of the coarsest operator
};

struct G3 {
  int n0, n1, n2;
};

__device__ __forceinline__ void node3(int id, const G3 &g, int &i, int &j, int &k) {
  k = id % g.n2;
  const int t = id / g.n2;
  j = t % g.n1;
  i = t / g.n1;
}

__device__ __forceinline__ int thread_id() { return blockIdx.x * blockDim.x + threadIdx.x; }

__device__ __forceinline__ bool mg3_skip(const MgAct &a) { return a.done && a.done[0]; }

// The first sweep's weights w Dinv of level 0.
__global__ void km3_weights(int n, const double *__restrict__ dinv, double *__restrict__ wd) {
  const int g = thread_id();
  if (g < n) wd[g] = kMg3Omega * dinv[g];
}

// Jacobi diagonal inverse of a stencil level.
template <int DPN>
__global__ void km3_dinv(int nn, const double *__restrict__ S, double *__restrict__ dinv) {
  const int t = thread_id();
  if (t >= nn * DPN) return;
  const int n = t / DPN, a = t % DPN;
  const double d = S[((int64_t)n * 27 + 13) * DPN * DPN + a * DPN + a];
  dinv[t] = d > 0.0 ? 1.0 / d : 0.0;
}

// Interpolation parents of the fine positions i - 1, i, i + 1 along one axis,
// as stencil offsets from the coarse node I (times the axis stride, 9, 3 or
// 1) and weights.  The fine node i is at most one fine step from I's
// position, so every parent is at most one coarse step from I; a single
// parent appears twice, the second time with weight 0, so all eight terms of
// a coupling are added without branching.
struct GalAxis {
  int o[3][2];
  double w[3][2];
};
__device__ __forceinline__ void gal_axis(int i, int nel, int co, int I, int stride, GalAxis &p) {
#pragma unroll
  for (int q = 0; q < 3; q++) {
    int A, B;
    double x;
    interp1(i + q - 1, nel, co, A, B, x);
    p.o[q][0] = (A - I + 1) * stride;
    p.o[q][1] = (B - I + 1) * stride;
    p.w[q][0] = 1.0 - x;
    p.w[q][1] = x;
  }
}

// Level 0's CSR rows in the stencil form: entry (a, b) of the block at the
// column node's offset d from the row node (absent couplings and fixed DOFs'
// empty rows give 0).  One wave per node: the lanes stride over the node's
// DPN contiguous rows (coalesced reads), scatter into the node's 27 DPN^2
// values in LDS, then store them contiguously.  A box mesh couples a node
// only to nodes at most one step away per axis.
constexpr int kMg3CsN = 4;  // nodes (waves) per block
template <int DPN>
__global__ __launch_bounds__(64 * kMg3CsN) void km3_csr_stencil(G3 f, const int *__restrict__ rp,
                                                                const int *__restrict__ ci,
                                                                const double *__restrict__ v, double *__restrict__ S) {
  constexpr int NV = 27 * DPN * DPN;
  __shared__ double s_blk[kMg3CsN][NV];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nn = f.n0 * f.n1 * f.n2;
  const int fn = blockIdx.x * kMg3CsN + w;
  double *blk = s_blk[w];
  for (int q = lane; q < NV; q += 64) blk[q] = 0.0;
  __syncthreads();
  if (fn < nn) {
    // with >= 3 nodes on the two fastest axes the offset dx n1 n2 + dy n2 + dz
    // of a neighbour decodes without division
    const bool fast = f.n1 >= 3 && f.n2 >= 3;
    int i0, i1, i2;
    node3(fn, f, i0, i1, i2);
    int rb[DPN + 1];
#pragma unroll
    for (int a = 0; a <= DPN; a++) rb[a] = rp[fn * DPN + a];
    for (int e = rb[0] + lane; e < rb[DPN]; e += 64) {
      int a = 0;
#pragma unroll
      for (int q = 1; q < DPN; q++)
        if (e >= rb[q]) a = q;
      const int col = ci[e], gn = col / DPN, b = col - gn * DPN;
      int d0, d1, d2;
      if (fast) {
        int dd = gn - fn;
        d0 = dd > f.n2 + 1 ? 1 : (dd < -(f.n2 + 1) ? -1 : 0);
        dd -= d0 * f.n1 * f.n2;
        d1 = dd > 1 ? 1 : (dd < -1 ? -1 : 0);
        d2 = dd - d1 * f.n2;
      } else {
        int g0, g1, g2;
        node3(gn, f, g0, g1, g2);
        d0 = g0 - i0;
        d1 = g1 - i1;
        d2 = g2 - i2;
      }
      if (d0 < -1 || d0 > 1 || d1 < -1 || d1 > 1 || d2 < -1 || d2 > 1) continue;
      blk[((d0 + 1) * 9 + (d1 + 1) * 3 + d2 + 1) * DPN * DPN + a * DPN + b] = v[e];
    }
  }
  __syncthreads();
  if (fn < nn)
    for (int q = lane; q < NV; q += 64) S[(int64_t)fn * NV + q] = blk[q];
}

// Galerkin coarse stencil, one thread per (coarse node I, row a, column
// component b): entry (a, b) of every block of I's 27-point stencil, sum over
// the fine nodes i of I's support (weight w_iI) of w_iI A_(i,a),(:,b) P,
// accumulated in the thread's LDS slice; Sf is the fine level's stencil.
template <int DPN>
__global__ __launch_bounds__(kMg3GB) void km3_galerkin(G3 f, G3 c, int co0, int co1, int co2,
                                                       const double *__restrict__ Sf, double *__restrict__ Sc) {
  __shared__ double s_acc[kMg3GB][27];
  double *acc = s_acc[threadIdx.x];
  const int item = thread_id();
  const int ncn = c.n0 * c.n1 * c.n2;
  if (item >= ncn * DPN * DPN) return;  // no barrier below: each thread owns its slice
  const int b = item % DPN, ia = item / DPN;
  const int I = ia / DPN, a = ia % DPN;
  for (int q = 0; q < 27; q++) acc[q] = 0.0;
  int I0, I1, I2;
  node3(I, c, I0, I1, I2);
  const int X0 = coarse_pos(I0, f.n0 - 1, co0), X1 = coarse_pos(I1, f.n1 - 1, co1),
            X2 = coarse_pos(I2, f.n2 - 1, co2);
  for (int e0 = -co0; e0 <= co0; e0++) {
    const int i0 = X0 + e0;
    if (i0 < 0 || i0 >= f.n0) continue;
    const double w0 = hat1(i0, I0, f.n0 - 1, co0);
    if (w0 == 0.0) continue;
    for (int e1 = -co1; e1 <= co1; e1++) {
      const int i1 = X1 + e1;
      if (i1 < 0 || i1 >= f.n1) continue;
      const double w1 = hat1(i1, I1, f.n1 - 1, co1);
      if (w1 == 0.0) continue;
      for (int e2 = -co2; e2 <= co2; e2++) {
        const int i2 = X2 + e2;
        if (i2 < 0 || i2 >= f.n2) continue;
        const double w2 = hat1(i2, I2, f.n2 - 1, co2);
        if (w2 == 0.0) continue;
        const double wi = w0 * w1 * w2;
        const int fn = (i0 * f.n1 + i1) * f.n2 + i2;
        GalAxis p0, p1, p2;  // parents of the neighbours i + d - 1 on each axis
        gal_axis(i0, f.n0 - 1, co0, I0, 9, p0);
        gal_axis(i1, f.n1 - 1, co1, I1, 3, p1);
        gal_axis(i2, f.n2 - 1, co2, I2, 1, p2);
#pragma unroll
        for (int d = 0; d < 27; d++) {
          const int q0 = d / 9, q1 = (d / 3) % 3, q2 = d % 3;
          const int g0 = i0 + q0 - 1, g1 = i1 + q1 - 1, g2 = i2 + q2 - 1;
          if (g0 < 0 || g0 >= f.n0 || g1 < 0 || g1 >= f.n1 || g2 < 0 || g2 >= f.n2) continue;
          const double val = Sf[((int64_t)fn * 27 + d) * DPN * DPN + a * DPN + b];
          if (val == 0.0) continue;
          const double w = wi * val;
#pragma unroll
          for (int u = 0; u < 2; u++)
#pragma unroll
            for (int t = 0; t < 2; t++)
#pragma unroll
              for (int q = 0; q < 2; q++)
                acc[p0.o[q0][u] + p1.o[q1][t] + p2.o[q2][q]] += w * p0.w[q0][u] * p1.w[q1][t] * p2.w[q2][q];
        }
      }
    }
  }
  for (int d = 0; d < 27; d++) Sc[((int64_t)I * 27 + d) * DPN * DPN + a * DPN + b] = acc[d];
}

// y = S x at node n (all components); neighbours off the grid have zero blocks.
template <int DPN>
__device__ __forceinline__ void st_apply(const G3 &g, const double *__restrict__ S, int n,
                                         const double *__restrict__ x, double (&y)[DPN]) {
  int i, j, k;
  node3(n, g, i, j, k);
  for (int a = 0; a < DPN; a++) y[a] = 0.0;
  const double *Sn = S + (int64_t)n * 27 * DPN * DPN;
#pragma unroll
  for (int d = 0; d < 27; d++) {
    const int i2 = i + d / 9 - 1, j2 = j + (d / 3) % 3 - 1, k2 = k + d % 3 - 1;
    if (i2 < 0 || i2 >= g.n0 || j2 < 0 || j2 >= g.n1 || k2 < 0 || k2 >= g.n2) continue;
    const int m = (i2 * g.n1 + j2) * g.n2 + k2;
    double xm[DPN];
#pragma unroll
    for (int b = 0; b < DPN; b++) xm[b] = x[m * DPN + b];
#pragma unroll
    for (int a = 0; a < DPN; a++)
#pragma unroll
      for (int b = 0; b < DPN; b++) y[a] += Sn[d * DPN * DPN + a * DPN + b] * xm[b];
  }
}

// Descent on a stencil level: r = b - S x1 (x1 = w Dinv b from the
// restriction); rows with a zero Jacobi weight get r = 0.
template <int DPN>
__global__ __launch_bounds__(kMg3B) void km3_resid0(G3 g, const double *__restrict__ S,
                                                    const double *__restrict__ dinv, const double *__restrict__ b,
                                                    const double *__restrict__ x1, double *__restrict__ r,
                                                    MgAct act) {
  const int n = thread_id();
  if (n >= g.n0 * g.n1 * g.n2 || mg3_skip(act)) return;
  double y[DPN];
  st_apply<DPN>(g, S, n, x1, y);
  for (int a = 0; a < DPN; a++) {
    const int q = n * DPN + a;
    r[q] = dinv[q] == 0.0 ? 0.0 : b[q] - y[a];
  }
}

// Post-smoothing on a stencil level: out = y + w Dinv (b - S y).
template <int DPN>
__global__ __launch_bounds__(kMg3B) void km3_smooth(G3 g, const double *__restrict__ S,
                                                    const double *__restrict__ dinv, const double *__restrict__ b,
                                                    const double *__restrict__ yv, double *__restrict__ out,
                                                    MgAct act) {
  const int n = thread_id();
  if (n >= g.n0 * g.n1 * g.n2 || mg3_skip(act)) return;
  double y[DPN];
  st_apply<DPN>(g, S, n, yv, y);
  for (int a = 0; a < DPN; a++) {
    const int q = n * DPN + a;
    out[q] = yv[q] + kMg3Omega * dinv[q] * (b[q] - y[a]);
  }
}

// Restriction bc = P^T rf (trilinear), and x1c = w Dinv_c bc when the coarse
// level smooths (x1c non-null).
template <int DPN>
__global__ __launch_bounds__(kMg3B) void km3_restrict(G3 f, G3 c, int co0, int co1, int co2,
                                                      const double *__restrict__ rf, double *__restrict__ bc,
                                                      const double *__restrict__ dinvc, double *__restrict__ x1c,
                                                      MgAct act) {
  const int I = thread_id();
  if (I >= c.n0 * c.n1 * c.n2 || mg3_skip(act)) return;
  int I0, I1, I2;
  node3(I, c, I0, I1, I2);
  const int X0 = coarse_pos(I0, f.n0 - 1, co0), X1 = coarse_pos(I1, f.n1 - 1, co1),
            X2 = coarse_pos(I2, f.n2 - 1, co2);
  double acc[DPN];
  for (int a = 0; a < DPN; a++) acc[a] = 0.0;
  for (int e0 = -co0; e0 <= co0; e0++) {
    const int i0 = X0 + e0;
    if (i0 < 0 || i0 >= f.n0) continue;
    const double w0 = hat1(i0, I0, f.n0 - 1, co0);
    if (w0 == 0.0) continue;
    for (int e1 = -co1; e1 <= co1; e1++) {
      const int i1 = X1 + e1;
      if (i1 < 0 || i1 >= f.n1) continue;
      const double w1 = hat1(i1, I1, f.n1 - 1, co1);
      if (w1 == 0.0) continue;
      for (int e2 = -co2; e2 <= co2; e2++) {
        const int i2 = X2 + e2;
        if (i2 < 0 || i2 >= f.n2) continue;
        const double w = w0 * w1 * hat1(i2, I2, f.n2 - 1, co2);
        if (w == 0.0) continue;
        const int fn = (i0 * f.n1 + i1) * f.n2 + i2;
        for (int a = 0; a < DPN; a++) acc[a] += w * rf[fn * DPN + a];
      }
    }
  }
  for (int a = 0; a < DPN; a++) {
    bc[I * DPN + a] = acc[a];
    if (x1c) x1c[I * DPN + a] = kMg3Omega * dinvc[I * DPN + a] * acc[a];
  }
}

// Coarse correction on top of the pre-smoothed x1: y = x1 + P xc, 0 where
// the Jacobi weight is 0 (fixed DOFs).
template <int DPN>
__global__ __launch_bounds__(kMg3B) void km3_prolong0(G3 f, G3 c, int co0, int co1, int co2,
                                                      const double *__restrict__ dinv, const double *__restrict__ x1,
                                                      const double *__restrict__ xc, double *__restrict__ y,
                                                      MgAct act) {
  const int n = thread_id();
  if (n >= f.n0 * f.n1 * f.n2 || mg3_skip(act)) return;
  int i0, i1, i2;
  node3(n, f, i0, i1, i2);
  int P0[2], P1[2], P2[2];
  double W0[2], W1[2], W2[2], w;
  interp1(i0, f.n0 - 1, co0, P0[0], P0[1], w);
  W0[0] = 1.0 - w;
  W0[1] = w;
  interp1(i1, f.n1 - 1, co1, P1[0], P1[1], w);
  W1[0] = 1.0 - w;
  W1[1] = w;
  interp1(i2, f.n2 - 1, co2, P2[0], P2[1], w);
  W2[0] = 1.0 - w;
  W2[1] = w;
  double v[DPN];
  for (int a = 0; a < DPN; a++) v[a] = 0.0;
  for (int u = 0; u < 2; u++) {
    if (W0[u] == 0.0) continue;
    for (int t = 0; t < 2; t++) {
      if (W1[t] == 0.0) continue;
      for (int s = 0; s < 2; s++) {
        if (W2[s] == 0.0) continue;
        const int cn = (P0[u] * c.n1 + P1[t]) * c.n2 + P2[s];
        const double ww = W0[u] * W1[t] * W2[s];
        for (int a = 0; a < DPN; a++) v[a] += ww * xc[cn * DPN + a];
      }
    }
  }
  for (int a = 0; a < DPN; a++) {
    const int q = n * DPN + a;
    y[q] = dinv[q] == 0.0 ? 0.0 : x1[q] + v[a];
  }
}

// Dense inverse of the coarsest operator (one block), gathered from its
// stencil (S non-null) or, when the system is itself that small, from the
// CSR rows.
template <int DPN>
__global__ __launch_bounds__(256) void km3_coarse_inv(G3 g, const double *__restrict__ S, const int *__restrict__ rp,
                                                      const int *__restrict__ ci, const double *__restrict__ v,
                                                      int nc, double *__restrict__ inv) {
  __shared__ double A[kMg3MaxC * kMg3MaxC];
  __shared__ double d0[kMg3MaxC];
  for (int q = threadIdx.x; q < nc * nc; q += blockDim.x) A[q] = 0.0;
  __syncthreads();
  if (S) {
    for (int q = threadIdx.x; q < nc * nc; q += blockDim.x) {
      const int r = q / nc, c = q % nc;
      const int rn = r / DPN, a = r % DPN, cn = c / DPN, b = c % DPN;
      int i, j, k, i2, j2, k2;
      node3(rn, g, i, j, k);
      node3(cn, g, i2, j2, k2);
      const int d0i = i2 - i, d1 = j2 - j, d2 = k2 - k;
      if (d0i < -1 || d0i > 1 || d1 < -1 || d1 > 1 || d2 < -1 || d2 > 1) continue;
      const int d = (d0i + 1) * 9 + (d1 + 1) * 3 + d2 + 1;
      A[q] = S[((int64_t)rn * 27 + d) * DPN * DPN + a * DPN + b];
    }
  } else {
    for (int r = threadIdx.x; r < nc; r += blockDim.x)
      for (int e = rp[r]; e < rp[r + 1]; e++) A[r * nc + ci[e]] = v[e];
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nc; r += blockDim.x) d0[r] = A[r * nc + r];
  gj_invert_lds(A, d0, nc);  // starts with a barrier
  gj_store_sym(A, nc, inv);
}

// Coarsest solve x = A^-1 b (the inverse is symmetric: row r read as column r).
__global__ __launch_bounds__(kMg3B) void km3_coarse_solve(int nc, const double *__restrict__ inv,
                                                          const double *__restrict__ b, double *__restrict__ x,
                                                          MgAct act) {
  const int r = thread_id();
  if (r >= nc || mg3_skip(act)) return;
  double s = 0.0;
  for (int c = 0; c < nc; c++) s += inv[c * nc + r] * b[c];
  x[r] = s;
}

static unsigned g3grid(int64_t n, int b = kMg3B) { return (unsigned)((n + b - 1) / b); }

static G3 g3(const Mg3Level &v) { return G3{v.n[0], v.n[1], v.n[2]}; }

void mg3_destroy(Mg3Data *mg) {
  if (!mg) return;
  for (auto &l : mg->lv) {
    dfree(l.S); dfree(l.dinv); dfree(l.b); dfree(l.x); dfree(l.x2); dfree(l.r);
  }
  dfree(mg->d_wd0);
  dfree(mg->d_inv);
  delete mg;
}

// Level dimensions {n0, n1, n2, co0, co1, co2} per level; returns the coarsest index.
int mg3_levels(const int n[3], int dpn, std::vector<int> *dims) {
  int c[3] = {n[0], n[1], n[2]}, L = 0;
  if (dims) dims->assign({c[0], c[1], c[2], 0, 0, 0});
  while ((int64_t)dpn * c[0] * c[1] * c[2] > kMg3MaxC && (c[0] >= 3 || c[1] >= 3 || c[2] >= 3)) {
    int co[3];
    for (int a = 0; a < 3; a++) {
      co[a] = c[a] >= 3;
      if (co[a]) c[a] = c[a] / 2 + 1;  // ceil((c - 1) / 2) cells
    }
    if (dims) dims->insert(dims->end(), {c[0], c[1], c[2], co[0], co[1], co[2]});
    L++;
  }
  return L;
}

int mg3_create(tfm_ctx *ctx, const int n[3], int dpn, Mg3Data **out) {
  *out = nullptr;
  if ((dpn != 1 && dpn != 3) || n[0] < 1 || n[1] < 1 || n[2] < 1)
    return fail(ctx, TFM_E_UNSUPPORTED, "multigrid: box systems with 1 or 3 DOFs per node");
  if ((int64_t)n[0] * n[1] * n[2] * 27 * dpn >= INT32_MAX)
    return fail(ctx, TFM_E_UNSUPPORTED, "multigrid: box too large for int32 indexing");
  auto *mg = new Mg3Data();
  mg->ctx = ctx;
  mg->dpn = dpn;
  std::vector<int> dims;
  mg->L = mg3_levels(n, dpn, &dims);
  for (int l = 0; l <= mg->L; l++) {
    Mg3Level v;
    for (int a = 0; a < 3; a++) {
      v.n[a] = dims[6 * l + a];
      v.co[a] = dims[6 * l + 3 + a];
    }
    v.nn = v.n[0] * v.n[1] * v.n[2];
    mg->lv.push_back(v);
  }
  mg->nc = mg->lv.back().nn * dpn;
  int rc = TFM_OK;
  for (int l = 0; l <= mg->L && !rc; l++) {
    Mg3Level &v = mg->lv[l];
    const size_t nv = (size_t)v.nn * dpn;
    if ((l > 0 || mg->L > 0) && (rc = dalloc(ctx, &v.S, nv * 27 * dpn))) break;
    if (l > 0 && ((rc = dalloc(ctx, &v.b, nv)) || (rc = dalloc(ctx, &v.x, nv)))) break;
    if ((l < mg->L || l == 0) && (rc = dalloc(ctx, &v.dinv, nv))) break;
    if (l < mg->L && ((rc = dalloc(ctx, &v.x2, nv)) || (rc = dalloc(ctx, &v.r, nv)))) break;
  }
  if (!rc && !(rc = dalloc(ctx, &mg->d_wd0, (size_t)mg->lv[0].nn * dpn)))
    rc = dalloc(ctx, &mg->d_inv, (size_t)mg->nc * mg->nc);
  if (rc) {
    mg3_destroy(mg);
    return rc;
  }
  *out = mg;
  return TFM_OK;
}

template <int DPN>
static int mg3_update_t(Mg3Data *mg, const int *rp, const int *ci, const double *v, hipStream_t st) {
  const int nrows = mg->lv[0].nn * DPN;
  hipLaunchKernelGGL(km3_weights, dim3(g3grid(nrows)), dim3(kMg3B), 0, st, nrows, (const double *)mg->lv[0].dinv,
                     mg->d_wd0);
  if (mg->L > 0) {
    const Mg3Level &l0 = mg->lv[0];
    hipLaunchKernelGGL((km3_csr_stencil<DPN>), dim3((unsigned)((l0.nn + kMg3CsN - 1) / kMg3CsN)),
                       dim3(64 * kMg3CsN), 0, st, g3(l0), rp, ci, v, l0.S);
  }
  for (int l = 1; l <= mg->L; l++) {
    const Mg3Level &f = mg->lv[l - 1], &c = mg->lv[l];
    const unsigned gg = g3grid((int64_t)c.nn * DPN * DPN, kMg3GB);
    hipLaunchKernelGGL((km3_galerkin<DPN>), dim3(gg), dim3(kMg3GB), 0, st, g3(f), g3(c), c.co[0], c.co[1], c.co[2],
                       (const double *)f.S, c.S);
    if (l < mg->L)
      hipLaunchKernelGGL((km3_dinv<DPN>), dim3(g3grid((int64_t)c.nn * DPN)), dim3(kMg3B), 0, st, c.nn,
                         (const double *)c.S, c.dinv);
  }
  const Mg3Level &cl = mg->lv[mg->L];
  hipLaunchKernelGGL((km3_coarse_inv<DPN>), dim3(1), dim3(256), 0, st, g3(cl),
                     (const double *)(mg->L > 0 ? cl.S : nullptr), rp, ci, v, mg->nc, mg->d_inv);
  return TFM_OK;
}

// New matrix values (level 0's CSR, fixed rows empty, and its Jacobi
// diagonal inverse, 0 at fixed DOFs): every level's operator, Jacobi
// diagonal and the coarsest inverse.
int mg3_update(Mg3Data *mg, const int *rp, const int *ci, const double *v, const double *dinv0) {
  hipStream_t st = mg->ctx->stream;
  TFM_CHECK_HIP(mg->ctx, hipMemcpyAsync(mg->lv[0].dinv, dinv0, sizeof(double) * mg->lv[0].nn * mg->dpn,
                                        hipMemcpyDeviceToDevice, st));
  const int rc = mg->dpn == 3 ? mg3_update_t<3>(mg, rp, ci, v, st) : mg3_update_t<1>(mg, rp, ci, v, st);
  if (rc) return rc;
  TFM_CHECK_HIP(mg->ctx, hipGetLastError());
  return TFM_OK;
}

// The V-cycle below level 0.  On entry mg3_fine_r holds level 0's r0 = b0 -
// A x1 and z holds x1 = w D_0^-1 b0; returns level 0's coarse-corrected y =
// x1 + P x_1, leaving the closing fine sweep to the caller.
template <int DPN>
static const double *mg3_cycle_t(const Mg3Data *mg, const double *z, hipStream_t st, MgAct act) {
  const int L = mg->L;
  for (int l = 0; l < L; l++) {
    const Mg3Level &v = mg->lv[l], &c = mg->lv[l + 1];
    if (l > 0)
      hipLaunchKernelGGL((km3_resid0<DPN>), dim3(g3grid(v.nn)), dim3(kMg3B), 0, st, g3(v), (const double *)v.S,
                         (const double *)v.dinv, (const double *)v.b, (const double *)v.x, v.r, act);
    const bool smooths = l + 1 < L;
    hipLaunchKernelGGL((km3_restrict<DPN>), dim3(g3grid(c.nn)), dim3(kMg3B), 0, st, g3(v), g3(c), c.co[0], c.co[1],
                       c.co[2], (const double *)v.r, c.b, (const double *)(smooths ? c.dinv : nullptr),
                       smooths ? c.x : nullptr, act);
  }
  const Mg3Level &cl = mg->lv[L];
  hipLaunchKernelGGL(km3_coarse_solve, dim3(g3grid(mg->nc)), dim3(kMg3B), 0, st, mg->nc, (const double *)mg->d_inv,
                     (const double *)cl.b, cl.x, act);
  for (int l = L - 1; l >= 0; l--) {
    const Mg3Level &v = mg->lv[l], &c = mg->lv[l + 1];
    const double *x1 = l == 0 ? z : v.x;
    hipLaunchKernelGGL((km3_prolong0<DPN>), dim3(g3grid(v.nn)), dim3(kMg3B), 0, st, g3(v), g3(c), c.co[0], c.co[1],
                       c.co[2], (const double *)v.dinv, x1, (const double *)c.x, v.x2, act);
    if (l > 0)
      hipLaunchKernelGGL((km3_smooth<DPN>), dim3(g3grid(v.nn)), dim3(kMg3B), 0, st, g3(v), (const double *)v.S,
                         (const double *)v.dinv, (const double *)v.b, (const double *)v.x2, v.x, act);
  }
  return mg->lv[0].x2;
}

const double *mg3_cycle(const Mg3Data *mg, const double *z, hipStream_t st, const int *done) {
  const MgAct act{done, 0};
  if (mg->L == 0) return nullptr;
  return mg->dpn == 3 ? mg3_cycle_t<3>(mg, z, st, act) : mg3_cycle_t<1>(mg, z, st, act);
}

// A system no larger than the coarsest level: z = A^-1 r directly.
void mg3_direct(const Mg3Data *mg, const double *r, double *z, hipStream_t st, const int *done) {
  const MgAct act{done, 0};
  hipLaunchKernelGGL(km3_coarse_solve, dim3(g3grid(mg->nc)), dim3(kMg3B), 0, st, mg->nc, (const double *)mg->d_inv,
                     r, z, act);
}

int mg3_depth(const Mg3Data *mg) { return mg->L; }
double *mg3_fine_r(const Mg3Data *mg) { return mg->lv[0].r; }
const double *mg3_fine_dinv(const Mg3Data *mg) { return mg->lv[0].dinv; }
const double *mg3_sweep_weights(const Mg3Data *mg) { return mg->L > 0 ? mg->d_wd0 : nullptr; }

}  // namespace tfm
