// tfm_assemble.hip -- GPU assembly of the reference's CSR matrix.
//
// What is computed.  K = Coo::to_csr(assemble(mesh, dpn, op))
// (crates/tpt-fem-assembly/src/lib.rs:428-480, crates/tpt-fem-sparse/src/
// lib.rs:122-175): element entries that are exactly 0.0 are never pushed,
// duplicates are summed in insertion order (element id, then local i, j),
// columns ascend, zero sums are kept.  Optionally restricted to free rows and
// columns, which is reduce_system's kred (assembly/src/lib.rs:504-535).
//
// How.  Row-block gather, no atomics on values:
//   * rows are grouped into tiles of consecutive nodes whose node->element
//     adjacency lists (sorted by element id) hold at most TP (node, element)
//     pairs in total;
//   * phase 1: one lane per pair evaluates the element's row block for that
//     node (tfm_elem.h) into LDS, and registers the pair in a per-node LDS
//     hash of neighbour node -> bitmask of contributing pairs;
//   * phase 2: one lane per CSR slot of the tile (a contiguous slot range, so
//     col_ind reads and value writes are coalesced) adds its contributions in
//     pair order, which is element-id order -- the reference's summation
//     order, so FP64 values are bit-identical.
// The symbolic pass runs the same phase 1 in exact FP64 to decide which
// entries are non-zero (the reference's exact-zero skip), counts the row
// lengths, scans them and fills col_ind.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "tfm_elem.h"
#include "tfm_internal.h"

#pragma clang fp contract(off)

namespace tfm {

// M_F32T: FP32 values summed from a table of element matrices (P.ke32)
enum { M_COUNT = 0, M_FILL = 1, M_F64 = 2, M_F32 = 3, M_F32T = 4 };
constexpr int kTileNodes = 16;  // nodes per tile (hash tables in LDS)
constexpr int kHash = 64;       // hash slots per node (> distinct neighbours)

__device__ __forceinline__ unsigned hslot(int key) {
  return ((unsigned)key * 2654435761u) >> (32 - 6);
}

template <int CELL, int DPN>
struct TileCfg {
  static constexpr int NN = Cell<CELL>::NN;
  static constexpr int W = NN * DPN;
  static constexpr int ENT = DPN * W;
  static constexpr int TP = ENT <= 16 ? 256 : (ENT <= 48 ? 128 : 64);
};

template <int CELL, int DPN, int MODE>
__global__ __launch_bounds__((TileCfg<CELL, DPN>::TP)) void k_asm_tile(
    const double *__restrict__ coords, int pdim, const int *__restrict__ conn,
    const int *__restrict__ n2e_ptr, const int *__restrict__ n2e, const int *__restrict__ tiles,
    OpParams P, const unsigned char *__restrict__ fixed, const int *__restrict__ rp,
    const int *__restrict__ ci, int *__restrict__ row_cnt, int *__restrict__ ci_out,
    int *__restrict__ dims, double *__restrict__ v64, float *__restrict__ v32,
    int *__restrict__ err) {
  using C = TileCfg<CELL, DPN>;
  constexpr int NN = C::NN, W = C::W, ENT = C::ENT, TP = C::TP;
  constexpr int DIM = Cell<CELL>::DIM;
  using VT = typename std::conditional<MODE == M_F32 || MODE == M_F32T, float, double>::type;

  // Row blocks of the tile's pairs (the table mode reads them from P.ke32).
  // The symbolic passes only need which entries are non-zero: one bit per
  // entry when the block has at most 64 (the values stay in registers), so
  // their LDS is a fraction of the numeric pass's and more tiles run per CU.
  constexpr bool kBits = MODE <= M_FILL && ENT <= 64;
  __shared__ VT s_tab[(MODE == M_F32T || kBits) ? 1 : TP][ENT];
  __shared__ unsigned long long s_nzb[kBits ? TP : 1];
  __shared__ int s_en[TP][NN];
  __shared__ int s_pe[(MODE == M_F32T) ? TP : 1];  // table mode: element << 3 | local node
  __shared__ int s_nptr[kTileNodes + 1];
  __shared__ int s_key[kTileNodes][kHash];
  __shared__ unsigned long long s_mask[kTileNodes][kHash];
  __shared__ int s_rp[kTileNodes * DPN + 1];
  __shared__ unsigned short s_nz[(MODE <= M_FILL) ? kTileNodes : 1][(MODE <= M_FILL) ? kHash : 1];
  __shared__ unsigned char s_ord[(MODE <= M_FILL) ? kTileNodes : 1][(MODE <= M_FILL) ? kHash : 1];
  __shared__ int s_nkeys[kTileNodes];

  const int tid = threadIdx.x;
  const int n0 = tiles[blockIdx.x], n1 = tiles[blockIdx.x + 1];
  const int nt = n1 - n0;
  const int pa = n2e_ptr[n0];
  const int np = n2e_ptr[n1] - pa;

  for (int i = tid; i <= nt; i += TP) s_nptr[i] = n2e_ptr[n0 + i] - pa;
  for (int i = tid; i < kTileNodes * kHash; i += TP) {
    s_key[i / kHash][i % kHash] = -1;
    s_mask[i / kHash][i % kHash] = 0ull;
  }
  if (tid < kTileNodes) s_nkeys[tid] = 0;
  if (MODE != M_COUNT)
    for (int i = tid; i <= nt * DPN; i += TP) s_rp[i] = rp[n0 * DPN + i];
  __syncthreads();

  // ---- phase 1: one lane per (node, element) pair ------------------------
  if (tid < np) {
    int v = n2e[pa + tid];
    int e = v >> 3, ln = v & 7;
    int tn = 0;
    while (tn + 1 < nt && s_nptr[tn + 1] <= tid) tn++;
    int k = tid - s_nptr[tn];
    double X[NN][3];
    int en[NN];
#pragma unroll
    for (int a = 0; a < NN; a++) {
      en[a] = conn[(int64_t)e * NN + a];
      s_en[tid][a] = en[a];
#pragma unroll
      for (int p = 0; p < 3; p++) X[a][p] = p < DIM ? coords[(int64_t)en[a] * pdim + p] : 0.0;
    }
    double coef = P.coef_e ? P.coef_e[e] : P.coef;
    double rho = P.rho ? P.rho[e] : 0.0;
    if constexpr (MODE == M_F32T) {
      s_pe[tid] = v;  // phase 2 reads the element's matrix rows from the table
    } else if constexpr (MODE == M_F32) {
      float tmp[ENT];
      rowblock_f32<CELL, DPN>(X, ln, coef, rho, P, tmp);
#pragma unroll
      for (int i = 0; i < ENT; i++) s_tab[tid][i] = tmp[i];
    } else if constexpr (kBits) {
      double tmp[ENT];
      rowblock_f64<CELL, DPN>(X, ln, coef, rho, P, tmp);
      unsigned long long nzb = 0ull;
#pragma unroll
      for (int i = 0; i < ENT; i++)
        if (tmp[i] != 0.0) nzb |= 1ull << i;
      s_nzb[tid] = nzb;
    } else {
      // accumulated in place in LDS (no register copy of the row block)
      rowblock_f64<CELL, DPN>(X, ln, coef, rho, P, s_tab[tid]);
    }
#pragma unroll
    for (int lm = 0; lm < NN; lm++) {
      int key = en[lm];
      unsigned h = hslot(key) & (kHash - 1);
      int probe = 0;
      for (; probe < kHash; probe++) {
        int old = atomicCAS(&s_key[tn][h], -1, key);
        if (old == -1 || old == key) {
          atomicOr(&s_mask[tn][h], 1ull << k);
          if (old == -1) atomicAdd(&s_nkeys[tn], 1);
          break;
        }
        h = (h + 1) & (kHash - 1);
      }
      if (probe == kHash) atomicExch(err, 1);
    }
  }
  __syncthreads();

  if constexpr (MODE == M_F64 || MODE == M_F32 || MODE == M_F32T) {
    // ---- phase 2: one lane per CSR slot, contributions in element order ---
    const int sb = s_rp[0], se = s_rp[nt * DPN];
    for (int s = sb + tid; s < se; s += TP) {
      int lo = 0, hi = nt * DPN;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s_rp[mid] <= s) lo = mid; else hi = mid;
      }
      int tn = lo / DPN, a = lo % DPN;
      int col = ci[s];
      int m = col / DPN, b = col - m * DPN;
      unsigned h = hslot(m) & (kHash - 1);
      unsigned long long mask = 0ull;
      for (int probe = 0; probe < kHash; probe++) {
        int kk = s_key[tn][h];
        if (kk == m) { mask = s_mask[tn][h]; break; }
        if (kk == -1) break;
        h = (h + 1) & (kHash - 1);
      }
      VT acc = 0;
      const int base = s_nptr[tn];
      while (mask) {
        int k = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        int i = base + k;
        int lm = 0;
#pragma unroll
        for (int q = 1; q < NN; q++)
          if (s_en[i][q] == m) lm = q;
        VT val;
        if constexpr (MODE == M_F32T) {
          const int pv = s_pe[i];
          val = P.ke32[(int64_t)(pv >> 3) * (W * W) + ((pv & 7) * DPN + a) * W + lm * DPN + b];
        } else {
          val = s_tab[i][a * W + lm * DPN + b];
        }
        if (val != (VT)0) acc += val;
      }
      if constexpr (MODE == M_F64) v64[s] = acc;
      else v32[s] = acc;
    }
  } else {
    // ---- symbolic: non-zero mask per (node, neighbour) block -------------
    for (int idx = tid; idx < nt * kHash; idx += TP) {
      int tn = idx / kHash, h = idx % kHash;
      int key = s_key[tn][h];
      if (key < 0) continue;
      unsigned long long mask = s_mask[tn][h];
      unsigned nz = 0;
      const int base = s_nptr[tn];
      while (mask) {
        int k = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        int i = base + k;
        int lm = 0;
#pragma unroll
        for (int q = 1; q < NN; q++)
          if (s_en[i][q] == key) lm = q;
#pragma unroll
        for (int a = 0; a < DPN; a++)
#pragma unroll
          for (int b = 0; b < DPN; b++)
          {
            bool keep;
            if constexpr (kBits) keep = (s_nzb[i] >> (a * W + lm * DPN + b)) & 1ull;
            else keep = s_tab[i][a * W + lm * DPN + b] != 0.0;
            if (P.kind == TFM_OP_SIMP || keep) nz |= 1u << (a * DPN + b);
          }
      }
      s_nz[tn][h] = (unsigned short)nz;
      int rank = 0;
      for (int h2 = 0; h2 < kHash; h2++) {
        int k2 = s_key[tn][h2];
        rank += (k2 >= 0 && k2 < key) ? 1 : 0;
      }
      s_ord[tn][rank] = (unsigned char)h;
    }
    __syncthreads();
    int dmax = 0;  // row count (count pass) / column count (fill pass) seen by this thread
    for (int r = tid; r < nt * DPN; r += TP) {
      int tn = r / DPN, a = r % DPN;
      int row = (n0 + tn) * DPN + a;
      if (fixed && fixed[row]) {
        if (MODE == M_COUNT) row_cnt[row] = 0;
        continue;
      }
      int c = 0, maxcol = -1;
      int out0 = (MODE == M_FILL) ? s_rp[r] : 0;
      for (int q = 0; q < s_nkeys[tn]; q++) {
        int h = s_ord[tn][q];
        int m = s_key[tn][h];
        unsigned nz = s_nz[tn][h];
        for (int b = 0; b < DPN; b++) {
          int col = m * DPN + b;
          if (!((nz >> (a * DPN + b)) & 1u)) continue;
          if (fixed && fixed[col]) continue;
          if (MODE == M_FILL) ci_out[out0 + c] = col;
          maxcol = col;
          c++;
        }
      }
      if (MODE == M_COUNT) {
        row_cnt[row] = c;
        if (c > 0) dmax = max(dmax, row + 1);
      } else if (maxcol >= 0) {
        dmax = max(dmax, maxcol + 1);
      }
    }
    block_max(dmax, MODE == M_COUNT ? &dims[0] : &dims[1]);
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static void host_gl(int order, double *p, double *w, int *n) {
  switch (order) {
  case 1: p[0] = 0.0; w[0] = 2.0; *n = 1; return;
  case 2: p[0] = -1.0 / std::sqrt(3.0); p[1] = 1.0 / std::sqrt(3.0); w[0] = 1.0; w[1] = 1.0; *n = 2; return;
  case 3:
    p[0] = 0.0; p[1] = -std::sqrt(3.0 / 5.0); p[2] = std::sqrt(3.0 / 5.0);
    w[0] = 8.0 / 9.0; w[1] = 5.0 / 9.0; w[2] = 5.0 / 9.0; *n = 3; return;
  case 4:
    p[0] = -0.3399810435848563; p[1] = 0.3399810435848563; p[2] = -0.8611363115940526; p[3] = 0.8611363115940526;
    w[0] = 0.6521451548625461; w[1] = 0.6521451548625461; w[2] = 0.3478548451374539; w[3] = 0.3478548451374539;
    *n = 4; return;
  case 5:
    p[0] = 0.0; p[1] = -0.5384693101056831; p[2] = 0.5384693101056831; p[3] = -0.9061798459386640;
    p[4] = 0.9061798459386640;
    w[0] = 0.5688888888888889; w[1] = 0.4786286704993665; w[2] = 0.4786286704993665;
    w[3] = 0.2369268850561891; w[4] = 0.2369268850561891; *n = 5; return;
  }
  *n = 0;
}

// Reference shape functions and gradients (crates/tpt-fem-element/src/lib.rs:90-221).
static void host_shape_grad(int cell, const double *xi, double *ns, double *lg) {
  double x = xi[0], y = xi[1], z = xi[2];
  static const double QS[4] = {-1.0, 1.0, 1.0, -1.0}, QT[4] = {-1.0, -1.0, 1.0, 1.0};
  static const double HS[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
  static const double HT[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
  static const double HU[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};
  switch (cell) {
  case TFM_LINE:
    ns[0] = (1.0 - x) / 2.0; ns[1] = (1.0 + x) / 2.0; lg[0] = -0.5; lg[1] = 0.5; break;
  case TFM_TRI:
    ns[0] = 1.0 - x - y; ns[1] = x; ns[2] = y;
    lg[0] = -1.0; lg[1] = -1.0; lg[2] = 1.0; lg[3] = 0.0; lg[4] = 0.0; lg[5] = 1.0; break;
  case TFM_QUAD:
    for (int i = 0; i < 4; i++) {
      ns[i] = 0.25 * (1.0 + QS[i] * x) * (1.0 + QT[i] * y);
      lg[2 * i] = 0.25 * QS[i] * (1.0 + QT[i] * y);
      lg[2 * i + 1] = 0.25 * (1.0 + QS[i] * x) * QT[i];
    }
    break;
  case TFM_TET: {
    ns[0] = 1.0 - x - y - z; ns[1] = x; ns[2] = y; ns[3] = z;
    const double G[12] = {-1.0, -1.0, -1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
    for (int i = 0; i < 12; i++) lg[i] = G[i];
    break;
  }
  case TFM_HEX:
    for (int i = 0; i < 8; i++) {
      ns[i] = 0.125 * (1.0 + HS[i] * x) * (1.0 + HT[i] * y) * (1.0 + HU[i] * z);
      lg[3 * i] = 0.125 * HS[i] * (1.0 + HT[i] * y) * (1.0 + HU[i] * z);
      lg[3 * i + 1] = 0.125 * (1.0 + HS[i] * x) * HT[i] * (1.0 + HU[i] * z);
      lg[3 * i + 2] = 0.125 * (1.0 + HS[i] * x) * (1.0 + HT[i] * y) * HU[i];
    }
    break;
  }
}

// cell_quad (crates/tpt-fem-thermal/src/lib.rs:29-92): tets and triangles use
// their Degree2 rules whatever the order; lines/quads/hexes Gauss-Legendre.
int build_rule(int cell, int order, std::vector<double> &rule) {
  std::vector<double> pts, wts;
  double gp[5], gw[5];
  int n = 0;
  auto push = [&](double a, double b, double c, double w) {
    pts.push_back(a); pts.push_back(b); pts.push_back(c); wts.push_back(w);
  };
  switch (cell) {
  case TFM_LINE:
    host_gl(order, gp, gw, &n);
    for (int i = 0; i < n; i++) push(gp[i], 0, 0, gw[i]);
    break;
  case TFM_TRI: {
    const double a = 2.0 / 3.0, b = 1.0 / 6.0;
    push(a, b, 0, 1.0 / 6.0); push(b, a, 0, 1.0 / 6.0); push(b, b, 0, 1.0 / 6.0);
    break;
  }
  case TFM_QUAD:
    host_gl(order, gp, gw, &n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) push(gp[i], gp[j], 0, gw[i] * gw[j]);
    break;
  case TFM_TET: {
    const double a = 0.5854101966249685, b = 0.1381966011250105;
    push(a, b, b, 1.0 / 24.0); push(b, a, b, 1.0 / 24.0); push(b, b, a, 1.0 / 24.0); push(b, b, b, 1.0 / 24.0);
    break;
  }
  case TFM_HEX:
    host_gl(order, gp, gw, &n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) push(gp[i], gp[j], gp[k], gw[i] * gw[j] * gw[k]);
    break;
  }
  int nq = (int)wts.size();
  if (nq == 0) return -1;
  rule.assign((size_t)nq * kRuleStride, 0.0);
  for (int q = 0; q < nq; q++) {
    double *R = rule.data() + q * kRuleStride;
    R[0] = wts[q];
    R[1] = pts[3 * q]; R[2] = pts[3 * q + 1]; R[3] = pts[3 * q + 2];
    host_shape_grad(cell, R + 1, R + 4, R + 12);
  }
  return nq;
}

// element_stiffness_unit (crates/tpt-fem-topopt/src/lib.rs:166-217).
void simp_ke_unit(double nu, double h, double *ke) {
  double c = 1.0 / (1.0 - nu * nu);
  double d[3][3] = {{c, c * nu, 0.0}, {c * nu, c, 0.0}, {0.0, 0.0, c * (1.0 - nu) / 2.0}};
  double gp[5], gw[5];
  int n;
  host_gl(2, gp, gw, &n);
  double jdet = h * h / 4.0, scale = 2.0 / h;
  for (int i = 0; i < 64; i++) ke[i] = 0.0;
  for (int a = 0; a < 2; a++)
    for (int bq = 0; bq < 2; bq++) {
      double xi[3] = {gp[a], gp[bq], 0.0}, ns[8], lg[8];
      double w = gw[a] * gw[bq];
      host_shape_grad(TFM_QUAD, xi, ns, lg);
      double bm[3][8] = {{0}}, db[3][8];
      for (int k = 0; k < 4; k++) {
        double nx_ = scale * lg[2 * k], ny_ = scale * lg[2 * k + 1];
        bm[0][2 * k] = nx_; bm[1][2 * k + 1] = ny_; bm[2][2 * k] = ny_; bm[2][2 * k + 1] = nx_;
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 8; j++) {
          double s = 0.0;
          for (int k = 0; k < 3; k++) s += d[i][k] * bm[k][j];
          db[i][j] = s;
        }
      for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
          double s = 0.0;
          for (int k = 0; k < 3; k++) s += bm[k][i] * db[k][j];
          ke[i * 8 + j] += w * jdet * s;
        }
    }
}

int validate_operator(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn) {
  if (!op) return fail(ctx, TFM_E_INVALID, "null operator");
  int dim = mesh->dim;
  if (op->kind == TFM_OP_POISSON) {
    if (dpn != 1) return fail(ctx, TFM_E_INVALID, "Poisson operator needs dofs_per_node == 1");
  } else if (op->kind == TFM_OP_ELASTICITY) {
    bool ok = (op->model == TFM_BAR && dim == 1) ||
              ((op->model == TFM_PLANE_STRESS || op->model == TFM_PLANE_STRAIN) && dim == 2) ||
              (op->model == TFM_CONTINUUM3D && dim == 3);
    if (!ok) return fail(ctx, TFM_E_MODEL, "model is incompatible with spatial dimension " + std::to_string(dim));
    if (dpn != dim) return fail(ctx, TFM_E_INVALID, "elasticity needs dofs_per_node == spatial dimension");
  } else if (op->kind == TFM_OP_SIMP) {
    if (mesh->cell != TFM_QUAD || dpn != 2) return fail(ctx, TFM_E_INVALID, "SIMP operator is Q4 with 2 DOFs per node");
    if (!op->rho) return fail(ctx, TFM_E_INVALID, "SIMP operator needs densities");
  } else if (op->kind == TFM_OP_MASS) {
    if (dpn != dim) return fail(ctx, TFM_E_INVALID, "mass matrix has one DOF per spatial dimension");
  } else {
    return fail(ctx, TFM_E_INVALID, "unknown operator kind");
  }
  if (mesh->cell != TFM_TET && mesh->cell != TFM_TRI && (op->quad_order < 1 || op->quad_order > 5))
    return fail(ctx, TFM_E_INVALID, "gauss_legendre: order must be in 1..=5");
  return TFM_OK;
}

template <int CELL, int DPN, int MODE>
static void launch_tile(tfm_ctx *ctx, const AsmPlan *pl, const OpParams &P, int *row_cnt, int *ci_out,
                        int *dims, double *v64, float *v32, int *err) {
  const tfm_mesh *m = pl->mesh;
  constexpr int TP = TileCfg<CELL, DPN>::TP;
  hipLaunchKernelGGL((k_asm_tile<CELL, DPN, MODE>), dim3(pl->ntiles), dim3(TP), 0, ctx->stream, m->d_coords,
                     m->pdim, m->d_conn, m->d_n2e_ptr, m->d_n2e, pl->d_tiles, P, pl->d_fixed, pl->d_rp, pl->d_ci,
                     row_cnt, ci_out, dims, v64, v32, err);
}

template <int MODE>
static int dispatch(tfm_ctx *ctx, const AsmPlan *pl, const OpParams &P, int *row_cnt, int *ci_out, int *dims,
                    double *v64, float *v32, int *err) {
  int cell = pl->mesh->cell, dpn = pl->dpn;
#define TFM_CASE(C, D)                                                                   \
  if (cell == C && dpn == D) {                                                           \
    launch_tile<C, D, MODE>(ctx, pl, P, row_cnt, ci_out, dims, v64, v32, err);           \
    return TFM_OK;                                                                       \
  }
  TFM_CASE(TFM_TET, 1)
  TFM_CASE(TFM_TET, 3)
  TFM_CASE(TFM_QUAD, 1)
  TFM_CASE(TFM_QUAD, 2)
  TFM_CASE(TFM_HEX, 1)
  TFM_CASE(TFM_HEX, 3)
  TFM_CASE(TFM_TRI, 1)
  TFM_CASE(TFM_TRI, 2)
  TFM_CASE(TFM_LINE, 1)
#undef TFM_CASE
  return fail(ctx, TFM_E_UNSUPPORTED, "cell / dofs-per-node combination not built");
}

static int tile_pairs_for(int cell, int dpn) {
  int nn = cell_nodes(cell), ent = dpn * nn * dpn;
  return ent <= 16 ? 256 : (ent <= 48 ? 128 : 64);
}

static OpParams make_params(const AsmPlan *pl, const double *d_rule, int nq, const double *d_ke0) {
  OpParams P{};
  P.kind = pl->op.kind;
  P.model = pl->op.model;
  P.nq = nq;
  P.coef = pl->op.coef;
  P.nu = pl->op.nu;
  P.penal = pl->op.penal;
  P.e0 = pl->op.e0;
  P.emin = pl->op.emin;
  P.coef_e = pl->d_coef_e;
  P.rho = pl->d_rho;
  P.rule = d_rule;
  P.ke0 = d_ke0;
  P.ke32 = nullptr;
  return P;
}

int plan_set_coefficients(tfm_ctx *ctx, AsmPlan *pl, const tfm_operator *op) {
  const tfm_mesh *m = pl->mesh;
  int rc;
  // The pattern holds no entries of an element whose coefficient was 0 when
  // it was built (assemble() skipped them all): a re-assembly that gives such
  // an element a non-zero coefficient would need entries the pattern lacks.
  if (pl->d_rp && op->kind != TFM_OP_SIMP) {
    auto coef = [&](int64_t e) { return op->coef_e ? op->coef_e[e] : op->coef; };
    bool bad = false;
    if (pl->zero_all)
      for (int64_t e = 0; e < m->n_elem && !bad; e++) bad = coef(e) != 0.0;
    else
      for (int64_t e : pl->zero_elems)
        if ((bad = coef(e) != 0.0)) break;
    if (bad)
      return fail(ctx, TFM_E_INVALID,
                  "re-assembly gives a non-zero coefficient to an element that had coefficient 0 when the "
                  "pattern was built (its entries are not in the pattern); create the system again");
  }
  pl->op = *op;
  if (op->coef_e) {
    if (!pl->d_coef_e && (rc = dalloc(ctx, &pl->d_coef_e, m->n_elem))) return rc;
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(pl->d_coef_e, op->coef_e, sizeof(double) * m->n_elem,
                                      hipMemcpyHostToDevice, ctx->stream));
  } else {
    dfree(pl->d_coef_e);
  }
  if (op->rho) {
    if (!pl->d_rho && (rc = dalloc(ctx, &pl->d_rho, m->n_elem))) return rc;
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(pl->d_rho, op->rho, sizeof(double) * m->n_elem, hipMemcpyHostToDevice,
                                      ctx->stream));
  } else {
    dfree(pl->d_rho);
  }
  if (op->kind == TFM_OP_SIMP) simp_ke_unit(op->nu, op->h, pl->ke0);
  pl->op.coef_e = nullptr;
  pl->op.rho = nullptr;
  return TFM_OK;
}

static int upload_tables(tfm_ctx *ctx, AsmPlan *pl, double **d_tab, int *nq) {
  std::vector<double> rule;
  int q = build_rule(pl->mesh->cell, pl->op.quad_order, rule);
  if (q <= 0) return fail(ctx, TFM_E_INVALID, "gauss_legendre: order must be in 1..=5");
  if (q > kMaxQp) return fail(ctx, TFM_E_UNSUPPORTED, "quadrature rule too large");
  std::vector<double> buf(rule);
  buf.insert(buf.end(), pl->ke0, pl->ke0 + 64);
  int rc = dalloc(ctx, d_tab, buf.size());
  if (rc) return rc;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(*d_tab, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice,
                                    ctx->stream));
  *nq = q;
  return TFM_OK;
}

// A plan's operator tables without a pattern: what the triplet export needs.
int plan_tables(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn, AsmPlan *pl, OpParams *P) {
  int rc;
  if ((rc = validate_operator(ctx, mesh, op, dpn))) return rc;
  pl->mesh = mesh;
  mesh_retain(mesh);  // released by plan_destroy
  pl->dpn = dpn;
  pl->ndof = mesh->n_nodes * dpn;
  if (pl->ndof >= INT32_MAX) return fail(ctx, TFM_E_UNSUPPORTED, "too many DOFs for int32 indexing");
  if ((rc = plan_set_coefficients(ctx, pl, op))) return rc;
  if ((rc = upload_tables(ctx, pl, &pl->d_tab, &pl->nq))) return rc;
  *P = make_params(pl, pl->d_tab, pl->nq, pl->d_tab + (size_t)pl->nq * kRuleStride);
  return TFM_OK;
}

// ---------------------------------------------------------------------------
// SIMP grids (tfm_mesh_grid_batch: P copies of an nx x ny node grid, Q4
// element (ei, ej) of design p = p (nx-1)(ny-1) + ej (nx-1) + ei with nodes
// (ei, ej), (ei+1, ej), (ei+1, ej+1), (ei, ej+1)): every element matrix is
// E(rho_e) Ke_unit (crates/tpt-fem-topopt/src/lib.rs:224, 233-235), so a row
// of the stiffness is formed directly from the four elements around its
// node, in ascending element order, each column's contributions summed from
// +0.0 as the tile kernel sums them -- the same values, bit for bit, at a
// fraction of the work.
// ---------------------------------------------------------------------------
__global__ void k_simp_erho(int64_t ne, const double *__restrict__ rho, double emin, double e0, double penal,
                            double *__restrict__ er) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < ne) er[e] = emin + pow(rho[e], penal) * (e0 - emin);
}

// One thread per DOF row (node n, component a), 256 rows per block: each
// row's entries (at most 9 neighbours x 2) go to LDS at their CSR offsets,
// then the block's contiguous CSR range is written with coalesced stores.
// With CHECK, only compares each row's entry count with the pattern's.
// X32 (FP64 only): also stores every value rounded to FP32 into v32, the
// FP32 array of a system that holds both.
constexpr int kSimpRows = 256;
template <typename VT, bool CHECK, bool X32 = false>
__global__ __launch_bounds__(kSimpRows) void k_simp_rows(int64_t nrows, int nx, int ny, const double *__restrict__ er,
                                                          const double *__restrict__ ke0,
                                                          const unsigned char *__restrict__ fixed,
                                                          const int *__restrict__ rp, VT *__restrict__ vals,
                                                          int *__restrict__ bad, float *__restrict__ v32) {
  __shared__ VT s_v[CHECK ? 1 : kSimpRows * 18];
  const int64_t r0 = (int64_t)blockIdx.x * kSimpRows;
  const int64_t r = r0 + threadIdx.x;
  const int64_t r1 = r0 + kSimpRows < nrows ? r0 + kSimpRows : nrows;
  const int base = rp[r0];
  if (r < nrows && !(fixed && fixed[r])) {
    const int a = (int)(r & 1);
    const int64_t n = r >> 1;
    const int64_t per = (int64_t)nx * ny;
    const int64_t p = n / per;
    const int rem = (int)(n - p * per), j = rem / nx, i = rem - j * nx;
    const int64_t ebase = p * (int64_t)(nx - 1) * (ny - 1);
    // the node's elements in ascending id: (i-1, j-1), (i, j-1), (i-1, j), (i, j)
    double E[4];
    bool has[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int ei = i - 1 + (k & 1), ej = j - 1 + (k >> 1);
      has[k] = ei >= 0 && ei < nx - 1 && ej >= 0 && ej < ny - 1;
      E[k] = has[k] ? er[ebase + (int64_t)ej * (nx - 1) + ei] : 0.0;
    }
    int q = rp[r];
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        const int mi = i + di, mj = j + dj;
        if (mi < 0 || mi >= nx || mj < 0 || mj >= ny) continue;
        bool share = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int ei = i - 1 + (k & 1), ej = j - 1 + (k >> 1);
          share |= has[k] && mi >= ei && mi <= ei + 1 && mj >= ej && mj <= ej + 1;
        }
        if (!share) continue;
        const int64_t m = p * per + (int64_t)mj * nx + mi;
#pragma unroll
        for (int b = 0; b < 2; b++) {
          if (fixed && fixed[m * 2 + b]) continue;
          if (!CHECK) {
            VT acc = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const int ei = i - 1 + (k & 1), ej = j - 1 + (k >> 1);
              if (!has[k] || mi < ei || mi > ei + 1 || mj < ej || mj > ej + 1) continue;
              // local index of (x, y) in element (ei, ej): (0,0) 0, (1,0) 1, (1,1) 2, (0,1) 3
              const int ln = (j - ej) ? ((i - ei) ? 2 : 3) : (i - ei);
              const int lm = (mj - ej) ? ((mi - ei) ? 2 : 3) : (mi - ei);
              VT v;
              if constexpr (sizeof(VT) == 8) {
                v = E[k] * ke0[(ln * 2 + a) * 8 + lm * 2 + b];
              } else {
                v = (VT)(E[k] * ke0[(ln * 2 + a) * 8 + lm * 2 + b]);  // the FP64 product rounded
              }
              if (v != (VT)0) acc += v;
            }
            s_v[q - base] = acc;
          }
          q++;
        }
      }
    if (CHECK && q != rp[r + 1]) atomicExch(bad, 1);
  } else if (CHECK && r < nrows && rp[r + 1] != rp[r]) {
    atomicExch(bad, 1);
  }
  if constexpr (!CHECK) {
    __syncthreads();
    const int len = rp[r1] - base;
    for (int k = threadIdx.x; k < len; k += kSimpRows) {
      vals[base + k] = s_v[k];
      if constexpr (X32) v32[base + k] = (float)s_v[k];
    }
  }
}

bool simp_grid_eligible(const AsmPlan *pl) {
  const tfm_mesh *m = pl->mesh;
  return pl->op.kind == TFM_OP_SIMP && pl->dpn == 2 && m->cell == TFM_QUAD && m->grid_P > 0 && m->grid_nx >= 2 &&
         m->grid_ny >= 2 && m->n_nodes == m->grid_P * m->grid_nx * m->grid_ny &&
         m->n_elem == m->grid_P * (m->grid_nx - 1) * (m->grid_ny - 1) && pl->d_rho && !getenv("TFM_GENERIC_ASSEMBLY");
}

static int simp_grid_numeric(tfm_ctx *ctx, const AsmPlan *pl, double *v64, float *v32) {
  const tfm_mesh *m = pl->mesh;
  hipStream_t st = ctx->stream;
  int rc;
  if (!pl->d_erho && (rc = dalloc(ctx, &pl->d_erho, m->n_elem))) return rc;
  hipLaunchKernelGGL(k_simp_erho, dim3((unsigned)((m->n_elem + 255) / 256)), dim3(256), 0, st, m->n_elem, pl->d_rho,
                     pl->op.emin, pl->op.e0, pl->op.penal, pl->d_erho);
  const int64_t nrows = pl->ndof;
  const dim3 g((unsigned)((nrows + kSimpRows - 1) / kSimpRows)), b(kSimpRows);
  const double *ke0 = pl->d_tab + (size_t)pl->nq * kRuleStride;
  if (v64 && v32) {
    // both arrays in one pass: the FP32 values are the FP64 sums rounded
    hipLaunchKernelGGL((k_simp_rows<double, false, true>), g, b, 0, st, nrows, (int)m->grid_nx, (int)m->grid_ny,
                       pl->d_erho, ke0, pl->d_fixed, pl->d_rp, v64, nullptr, v32);
  } else if (v64) {
    hipLaunchKernelGGL((k_simp_rows<double, false>), g, b, 0, st, nrows, (int)m->grid_nx, (int)m->grid_ny,
                       pl->d_erho, ke0, pl->d_fixed, pl->d_rp, v64, nullptr, nullptr);
  } else if (v32) {
    hipLaunchKernelGGL((k_simp_rows<float, false>), g, b, 0, st, nrows, (int)m->grid_nx, (int)m->grid_ny,
                       pl->d_erho, ke0, pl->d_fixed, pl->d_rp, v32, nullptr, nullptr);
  }
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

// Once per plan: the direct row kernel applies when it reproduces the
// pattern's row lengths exactly.
static int simp_grid_check(tfm_ctx *ctx, AsmPlan *pl) {
  const tfm_mesh *m = pl->mesh;
  int *bad = nullptr, hb = 0, rc;
  if ((rc = dalloc(ctx, &bad, 1))) return rc;
  TFM_CHECK_HIP(ctx, hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
  const int64_t nrows = pl->ndof;
  hipLaunchKernelGGL((k_simp_rows<double, true>), dim3((unsigned)((nrows + kSimpRows - 1) / kSimpRows)),
                     dim3(kSimpRows), 0, ctx->stream,
                     nrows, (int)m->grid_nx, (int)m->grid_ny, pl->d_rho, pl->d_rho, pl->d_fixed, pl->d_rp,
                     (double *)nullptr, bad, (float *)nullptr);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dfree(bad);
  pl->use_simp_grid = hb == 0;
  return TFM_OK;
}

int plan_create(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn,
                const unsigned char *h_fixed, int vals, AsmPlan *pl) {
  int rc;
  if ((rc = validate_operator(ctx, mesh, op, dpn))) return rc;
  if ((rc = ensure_adjacency(const_cast<tfm_mesh *>(mesh)))) return rc;
  pl->mesh = mesh;
  mesh_retain(mesh);  // released by plan_destroy: a system may outlive the caller's handle
  pl->dpn = dpn;
  pl->ndof = mesh->n_nodes * dpn;
  if (pl->ndof >= INT32_MAX) return fail(ctx, TFM_E_UNSUPPORTED, "too many DOFs for int32 indexing");
  if ((rc = plan_set_coefficients(ctx, pl, op))) return rc;
  if (op->kind != TFM_OP_SIMP) {
    if (op->coef_e) {
      for (int64_t e = 0; e < mesh->n_elem; e++)
        if (op->coef_e[e] == 0.0) pl->zero_elems.push_back(e);
    } else {
      pl->zero_all = op->coef == 0.0 && mesh->n_elem > 0;
    }
  }
  if (h_fixed) {
    if ((rc = dalloc(ctx, &pl->d_fixed, pl->ndof))) return rc;
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(pl->d_fixed, h_fixed, pl->ndof, hipMemcpyHostToDevice, ctx->stream));
  }
  // A node in more elements than a tile holds (or than its pair mask has
  // bits): the sorted-triplet path assembles the whole mesh.
  int TP = tile_pairs_for(mesh->cell, dpn);
  if (mesh->max_adj > 64 || mesh->max_adj > TP || getenv("TFM_COO_ASSEMBLY") || hex_table_route(mesh, dpn, op)) {
    if ((rc = upload_tables(ctx, pl, &pl->d_tab, &pl->nq))) return rc;
    OpParams P = make_params(pl, pl->d_tab, pl->nq, pl->d_tab + (size_t)pl->nq * kRuleStride);
    return coo_plan_create(ctx, pl, P);
  }
  // Tile schedule: consecutive nodes, <= TP pairs and <= kTileNodes nodes.
  pl->tile_pairs = TP;
  std::vector<int> tiles;
  tiles.push_back(0);
  const std::vector<int> &p = mesh->h_n2e_ptr;
  int start = 0;
  for (int n = 0; n < (int)mesh->n_nodes; n++) {
    int pairs = p[n + 1] - p[start];
    int nodes = n - start + 1;
    if (pairs > TP || nodes > kTileNodes) {
      tiles.push_back(n);
      start = n;
    }
  }
  if (mesh->n_nodes > 0) tiles.push_back((int)mesh->n_nodes);
  pl->ntiles = (int)tiles.size() - 1;
  if ((rc = dalloc(ctx, &pl->d_tiles, tiles.size()))) return rc;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(pl->d_tiles, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice,
                                    ctx->stream));
  if ((rc = upload_tables(ctx, pl, &pl->d_tab, &pl->nq))) return rc;
  OpParams P = make_params(pl, pl->d_tab, pl->nq, pl->d_tab + (size_t)pl->nq * kRuleStride);

  // Symbolic pass 1: row counts.
  int *cnt = nullptr, *dims = nullptr, *err = nullptr;
  if ((rc = dalloc(ctx, &cnt, pl->ndof + 1)) || (rc = dalloc(ctx, &dims, 2)) || (rc = dalloc(ctx, &err, 1)))
    return rc;
  TFM_CHECK_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int) * (pl->ndof + 1), ctx->stream));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(dims, 0, sizeof(int) * 2, ctx->stream));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
  if (pl->ntiles > 0 && (rc = dispatch<M_COUNT>(ctx, pl, P, cnt, nullptr, dims, nullptr, nullptr, err))) return rc;
  TFM_CHECK_HIP(ctx, hipGetLastError());
  if ((rc = dalloc(ctx, &pl->d_rp, pl->ndof + 1))) return rc;
  int64_t nnz = 0;
  if ((rc = exclusive_scan_i32(ctx, cnt, pl->d_rp, pl->ndof, &nnz))) return rc;
  pl->nnz = nnz;
  if ((rc = dalloc(ctx, &pl->d_ci, nnz))) return rc;
  if (pl->ntiles > 0 && (rc = dispatch<M_FILL>(ctx, pl, P, nullptr, pl->d_ci, dims, nullptr, nullptr, err)))
    return rc;
  TFM_CHECK_HIP(ctx, hipGetLastError());
  int hd[2], herr = 0;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hd, dims, sizeof hd, hipMemcpyDeviceToHost, ctx->stream));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, ctx->stream));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dfree(cnt);
  dfree(dims);
  dfree(err);
  if (herr) {
    // more distinct neighbours than a node's hash holds: sorted triplets
    dfree(pl->d_rp);
    dfree(pl->d_ci);
    pl->d_rp = pl->d_ci = nullptr;
    return coo_plan_create(ctx, pl, P);
  }
  // Coo::to_csr: nrows = max row + 1, ncols = max col + 1, and an empty
  // triplet list still yields a 1 x 1 shape.
  pl->nrows = nnz ? hd[0] : 1;
  pl->ncols = nnz ? hd[1] : 1;
  if (simp_grid_eligible(pl) && (rc = simp_grid_check(ctx, pl))) return rc;
  // P1-tet Poisson / 3-D elasticity: the balanced fast path, when every node
  // fits its tile limits (otherwise the generic kernel keeps the job).
  const bool tet_op = mesh->cell == TFM_TET && mesh->n_elem > 0 && ((op->kind == TFM_OP_POISSON && dpn == 1) ||
                                                (op->kind == TFM_OP_ELASTICITY && dpn == 3));
  if (tet_op && mesh->max_adj <= 32 && !getenv("TFM_GENERIC_ASSEMBLY")) {
    if (tet_plan_tuned(ctx, mesh, dpn, pl->d_rp, pl->d_ci, nnz, P, vals ? vals : TFM_SYS_FP64, &pl->tet,
                       &pl->tune) == TFM_OK) {
      pl->use_tet = true;
    } else {
      tet_plan_destroy(&pl->tet);
      ctx->err.clear();
    }
  }
  return TFM_OK;
}

int plan_numeric(tfm_ctx *ctx, const AsmPlan *pl, double *d_v64, float *d_v32) {
  if (!pl->d_tab) return fail(ctx, TFM_E_INVALID, "assembly plan has no tables");
  OpParams P = make_params(pl, pl->d_tab, pl->nq, pl->d_tab + (size_t)pl->nq * kRuleStride);
  int rc;
  if (pl->use_coo) return coo_numeric(ctx, pl, P, d_v64, d_v32);
  if (pl->use_simp_grid) return simp_grid_numeric(ctx, pl, d_v64, d_v32);
  if (pl->ntiles == 0) return TFM_OK;
  if (pl->use_tet) return tet_numeric(ctx, pl->mesh, pl->dpn, pl->tet, pl->d_rp, P, d_v64, d_v32);
  if (d_v64 && (rc = dispatch<M_F64>(ctx, pl, P, nullptr, nullptr, nullptr, d_v64, nullptr, nullptr))) return rc;
  if (d_v32) {
    // FP32 Hex8 elasticity: element matrices on the matrix cores, once per
    // element, then the tile pass sums them
    if (hex_mfma_eligible(pl)) {
      if (!pl->d_ke32 && (rc = dalloc(ctx, &pl->d_ke32, pl->mesh->n_elem * 576))) return rc;
      if ((rc = hex_ke_mfma(ctx, pl, pl->d_ke32))) return rc;
      OpParams P32 = P;
      P32.ke32 = pl->d_ke32;
      if ((rc = dispatch<M_F32T>(ctx, pl, P32, nullptr, nullptr, nullptr, nullptr, d_v32, nullptr))) return rc;
    } else if ((rc = dispatch<M_F32>(ctx, pl, P, nullptr, nullptr, nullptr, nullptr, d_v32, nullptr))) {
      return rc;
    }
  }
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

void plan_destroy(AsmPlan *pl) {
  tet_plan_destroy(&pl->tet);
  dfree(pl->d_erho);
  coo_plan_destroy(pl);
  dfree(pl->d_ke32);
  dfree(pl->d_tab);
  dfree(pl->d_coef_e);
  dfree(pl->d_rho);
  dfree(pl->d_fixed);
  dfree(pl->d_tiles);
  dfree(pl->d_rp);
  dfree(pl->d_ci);
  if (pl->mesh) mesh_release(pl->mesh);
  pl->mesh = nullptr;
}

}  // namespace tfm
