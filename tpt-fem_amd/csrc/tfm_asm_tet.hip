// tfm_asm_tet.hip -- numeric assembly fast path for P1 tetrahedra.
//
// Same values as the generic tile kernel (k_asm_tile; bit-identical FP64),
// split into a plan built once per sparsity pattern and a lean numeric pass.
//
// Plan (on the GPU, once): for every node n its row blocks (n, m), m running
// over the column nodes its CSR rows reach, in column order, with
//   bmask  u32  which of n's (node, element) pairs -- the node -> element
//               adjacency, ascending element id -- contribute to (n, m)
//   blmb   u64  the 2-bit element-local index of m in each of those pairs
// and for every row DPN*n + a and block, the CSR position of the block's first
// present entry and a presence mask of its DPN columns (columns of fixed DOFs
// and entries assemble() skipped as exact zeros are absent).
//
// The plan stores each tile's row blocks as items in the order the numeric
// pass takes them (the tile's diagonal blocks first), so a tile reads its
// items straight from its first-item index: no per-node tables, prefix sums
// or barriers before the element math starts.
//
// Numeric pass: tiles of consecutive nodes (<= kTetTP pairs).  One lane per
// (node, element) pair computes the element map once; then, one DOF row a at
// a time, each pair writes its 4 x DPN entries of row a into LDS and one lane
// per row block adds the contributions of the block's pairs in ascending
// element order -- the reference's duplicate-summation order (Coo::to_csr)
// -- and writes them at the planned positions.  No hashing, searching or
// atomics.
//
// Entry arithmetic is the reference's (tfm_elem.h: bdb3, and quad4 for the
// sum of the four equal quadrature contributions); this translation unit is
// built without FMA contraction.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>

#include "tfm_elem.h"
#include "tfm_internal.h"

#pragma clang fp contract(off)

namespace tfm {

constexpr int kTetTP = 192;    // pairs (and threads) per tile
constexpr int kTetNodes = 32;  // nodes per tile

// FP32 entry of the precision form: vol * B_i^T D B_j (structural zeros dropped).
__device__ __forceinline__ float bdb3f(const float *gi, const float *gj, int a, int b, float m, float cnu,
                                       float dd) {
  switch (a * 3 + b) {
  case 0: return gi[0] * m * gj[0] + dd * (gi[1] * gj[1] + gi[2] * gj[2]);
  case 1: return gi[0] * cnu * gj[1] + gi[1] * dd * gj[0];
  case 2: return gi[0] * cnu * gj[2] + gi[2] * dd * gj[0];
  case 3: return gi[1] * cnu * gj[0] + gi[0] * dd * gj[1];
  case 4: return gi[1] * m * gj[1] + dd * (gi[0] * gj[0] + gi[2] * gj[2]);
  case 5: return gi[1] * cnu * gj[2] + gi[2] * dd * gj[1];
  case 6: return gi[2] * cnu * gj[0] + gi[0] * dd * gj[2];
  case 7: return gi[2] * cnu * gj[1] + gi[1] * dd * gj[2];
  default: return gi[2] * m * gj[2] + dd * (gi[1] * gj[1] + gi[0] * gj[0]);
  }
}

// Element quantities of one (node, element) pair, kept in registers while
// its row block is emitted one DOF row at a time.
template <int DPN, typename VT>
struct TetPair {
  double g[4][3];  // physical gradients (FP64 path)
  double gl[3];    // the pair's own node's gradient, g[ln]
  float gf[4][3];  // the same in FP32 (precision form)
  float gfl[3];
  double k0, k1;   // FP64 Poisson: w * coef, |det|; FP64 elasticity: w * |det|
  DConst D;        // FP64 elasticity
  float f0, m, cnu, dd;  // FP32: vol (* coef for Poisson) and the D entries
};

template <int DPN, typename VT>
__device__ __forceinline__ void tet_setup(const double (&X)[4][3], int ln, double coef, const OpParams &P,
                                          TetPair<DPN, VT> &pr) {
  const double det = fabs(map_tet(X, pr.g));
  // g[ln] by selects (a dynamic index would put the array in scratch memory)
#pragma unroll
  for (int p = 0; p < 3; p++) {
    pr.gl[p] = pr.g[0][p];
#pragma unroll
    for (int i = 1; i < 4; i++)
      if (ln == i) pr.gl[p] = pr.g[i][p];
  }
  const double w = 1.0 / 24.0;
  if constexpr (sizeof(VT) == 8) {
    if constexpr (DPN == 1) {
      pr.k0 = w * coef;
      pr.k1 = det;
    } else {
      pr.D = make_d(TFM_CONTINUUM3D, coef, P.nu);
      pr.k0 = w * det;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int p = 0; p < 3; p++) pr.gf[i][p] = (float)pr.g[i][p];
#pragma unroll
    for (int p = 0; p < 3; p++) pr.gfl[p] = (float)pr.gl[p];
    const double vol = det * (1.0 / 6.0);  // the four equal-weight points
    if constexpr (DPN == 1) {
      pr.f0 = (float)(vol * coef);
    } else {
      const double c = coef / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
      pr.f0 = (float)vol;
      pr.m = (float)((1.0 - P.nu) * c);
      pr.cnu = (float)(c * P.nu);
      pr.dd = (float)((1.0 - 2.0 * P.nu) / 2.0 * c);
    }
  }
}

// Row a of element-local node ln's row block: K_e[ln*DPN + a][lm*DPN + b]
// into out[(lm * DPN + b) * stride].
template <int DPN, typename VT>
__device__ __forceinline__ void tet_row(const TetPair<DPN, VT> &pr, int a, VT *out, int stride) {
  if constexpr (sizeof(VT) == 8) {
    if constexpr (DPN == 1) {
#pragma unroll
      for (int lm = 0; lm < 4; lm++) {
        const double dot = pr.gl[0] * pr.g[lm][0] + pr.gl[1] * pr.g[lm][1] + pr.gl[2] * pr.g[lm][2];
        const double c = pr.k0 * dot * pr.k1;
        out[lm * stride] = quad4(c);
      }
    } else {
#pragma unroll
      for (int lm = 0; lm < 4; lm++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
          const double c = pr.k0 * bdb3(pr.gl, pr.g[lm], a, b, pr.D);
          out[(lm * 3 + b) * stride] = quad4(c);
        }
    }
  } else {
    if constexpr (DPN == 1) {
#pragma unroll
      for (int lm = 0; lm < 4; lm++)
        out[lm * stride] =
            pr.f0 * (pr.gfl[0] * pr.gf[lm][0] + pr.gfl[1] * pr.gf[lm][1] + pr.gfl[2] * pr.gf[lm][2]);
    } else {
#pragma unroll
      for (int lm = 0; lm < 4; lm++)
#pragma unroll
        for (int b = 0; b < 3; b++)
          out[(lm * 3 + b) * stride] = pr.f0 * bdb3f(pr.gfl, pr.gf[lm], a, b, pr.m, pr.cnu, pr.dd);
    }
  }
}

// Plan entries of one item (row block), for all DPN rows.
struct TetBlk {
  unsigned mask;
  unsigned long long lmb;
  int kb;             // first pair of the block's node within the tile
  int pos[3];         // per row: CSR position of the first present entry
  unsigned pm;        // per row a: presence mask of the DPN columns in bits 4a..4a+3
};

template <int DPN>
__device__ __forceinline__ TetBlk tet_blk(int64_t item, int64_t nitems, const unsigned *__restrict__ mask,
                                          const unsigned long long *__restrict__ lmb, const int *__restrict__ pos,
                                          const unsigned *__restrict__ pmk) {
  TetBlk B;
  B.mask = mask[item];
  B.lmb = lmb[item];
  const unsigned w = pmk[item];
  B.kb = (int)(w >> 16);
  B.pm = w & 0xfffu;
#pragma unroll
  for (int a = 0; a < DPN; a++) B.pos[a] = pos[a * nitems + item];
  return B;
}

// Sums one row of a block from the pairs' rows in LDS, in ascending pair (=
// element) order.  LDS reads are issued two contributions at a time; missing
// ones read as +0.0, which leaves the sum unchanged (it is never -0.0: it
// starts at +0.0 and exact cancellation rounds to +0.0).
template <int DPN, typename VT>
__device__ __forceinline__ void tet_sum(const VT *s_c, const TetBlk &B, VT (&acc)[DPN]) {
#pragma unroll
  for (int b = 0; b < DPN; b++) acc[b] = 0;
  unsigned mask = B.mask;
  while (mask) {
    VT c[2][DPN];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      if (mask) {
        const int j = __ffs((int)mask) - 1;
        mask &= mask - 1;
        const int lm = (int)((B.lmb >> (2 * j)) & 3ull);
        const VT *src = s_c + lm * DPN * kTetTP + B.kb + j;
#pragma unroll
        for (int b = 0; b < DPN; b++) c[u][b] = src[b * kTetTP];
      } else {
#pragma unroll
        for (int b = 0; b < DPN; b++) c[u][b] = 0;
      }
    }
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int b = 0; b < DPN; b++) acc[b] += c[u][b];
  }
}

template <int DPN, typename VT>
__device__ __forceinline__ void tet_write(const TetBlk &B, int a, const VT (&acc)[DPN], VT *__restrict__ vals) {
  int q = B.pos[0];
#pragma unroll
  for (int i = 1; i < DPN; i++)  // selects: a dynamic index would use scratch memory
    if (a == i) q = B.pos[i];
  const unsigned pm = (B.pm >> (4 * a)) & 0xfu;
#pragma unroll
  for (int b = 0; b < DPN; b++)
    if ((pm >> b) & 1u) vals[q++] = acc[b];
}

template <int DPN, typename VT>
__global__ __launch_bounds__(kTetTP, 5) void k_asm_tet(const double *__restrict__ coords, const int *__restrict__ conn,
                                                     const int *__restrict__ n2e, const int *__restrict__ tpa,
                                                     const int *__restrict__ tib, int ntiles, OpParams P,
                                                     const unsigned *__restrict__ imask,
                                                     const unsigned long long *__restrict__ ilmb,
                                                     const int *__restrict__ ipos, const unsigned *__restrict__ ipmk,
                                                     int64_t nitems, VT *__restrict__ vals) {
  __shared__ VT s_c[4 * DPN * kTetTP];  // one DOF row of every pair: s_c[(lm * DPN + b) * kTetTP + pair]

  const int tid = threadIdx.x;
  const int tile = xcd_item(blockIdx.x, ntiles);  // neighbouring tiles share an XCD's L2
  if (tile >= ntiles) return;
  const int pa = tpa[tile], np = tpa[tile + 1] - pa;
  const int ib = tib[tile], nblk = tib[tile + 1] - ib;

  // element loads, one (node, element) pair per lane (unconditional: inactive
  // lanes read pair 0 of the tile, or of the mesh when the tile's nodes have
  // no elements)
  const bool pair = tid < np;
  const int v = n2e[np > 0 ? pa + (pair ? tid : 0) : 0];
  const int e = v >> 3, ln = v & 7;
  double X[4][3];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int nd = conn[(int64_t)e * 4 + a];
#pragma unroll
    for (int p = 0; p < 3; p++) X[a][p] = coords[(int64_t)nd * 3 + p];
  }
  const double coef = P.coef_e ? P.coef_e[e] : P.coef;

  // the first round's items, fetched while the element math runs
  TetBlk B0{};
  if (tid < nblk) B0 = tet_blk<DPN>(ib + tid, nitems, imask, ilmb, ipos, ipmk);

  TetPair<DPN, VT> pr;
  tet_setup<DPN, VT>(X, ln, coef, P, pr);

#pragma unroll
  for (int a = 0; a < DPN; a++) {
    if (a > 0) __syncthreads();  // the previous row's sums are done with s_c
    if (pair) tet_row<DPN, VT>(pr, a, s_c + tid, kTetTP);
    __syncthreads();
    VT acc[DPN];
    if (tid < nblk) {
      tet_sum<DPN, VT>(s_c, B0, acc);
      tet_write<DPN, VT>(B0, a, acc, vals);
    }
    for (int it = tid + kTetTP; it < nblk; it += kTetTP) {
      const TetBlk B = tet_blk<DPN>(ib + it, nitems, imask, ilmb, ipos, ipmk);
      tet_sum<DPN, VT>(s_c, B, acc);
      tet_write<DPN, VT>(B, a, acc, vals);
    }
  }
}

// ---- plan -------------------------------------------------------------------
template <int DPN>
__device__ __forceinline__ int tet_merge_min(const int *__restrict__ ci, const int *j, const int *e) {
  int m = INT_MAX;
#pragma unroll
  for (int a = 0; a < DPN; a++)
    if (j[a] < e[a]) m = min(m, ci[j[a]] / DPN);
  return m;
}

// Row blocks of node n: distinct column nodes over its DPN rows; and whether
// one of them is the node itself (its diagonal block).
template <int DPN>
__global__ void k_tet_plan_count(int N, const int *__restrict__ rp, const int *__restrict__ ci, int *__restrict__ cnt,
                                 unsigned char *__restrict__ hasdiag) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  int j[DPN], e[DPN];
#pragma unroll
  for (int a = 0; a < DPN; a++) {
    j[a] = rp[n * DPN + a];
    e[a] = rp[n * DPN + a + 1];
  }
  int k = 0;
  unsigned char hd = 0;
  for (int m; (m = tet_merge_min<DPN>(ci, j, e)) != INT_MAX; k++) {
    hd |= (unsigned char)(m == n);
#pragma unroll
    for (int a = 0; a < DPN; a++)
      while (j[a] < e[a] && ci[j[a]] / DPN == m) j[a]++;
  }
  cnt[n] = k;
  hasdiag[n] = hd;
}

template <int DPN>
__global__ void k_tet_plan_fill(int N, const int *__restrict__ rp, const int *__restrict__ ci,
                                const int *__restrict__ conn, const int *__restrict__ n2e_ptr,
                                const int *__restrict__ n2e, const int *__restrict__ bptr,
                                const unsigned char *__restrict__ hasdiag, unsigned *__restrict__ bmask,
                                unsigned long long *__restrict__ blmb, int *__restrict__ ipos,
                                unsigned char *__restrict__ ipm, int *__restrict__ err) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  int j[DPN], e[DPN];
#pragma unroll
  for (int a = 0; a < DPN; a++) {
    j[a] = rp[n * DPN + a];
    e[a] = rp[n * DPN + a + 1];
  }
  const int pa = n2e_ptr[n], np = n2e_ptr[n + 1] - pa;
  if (np > 32) {
    atomicExch(err, 1);
    return;
  }
  const int bb = bptr[n], nb = bptr[n + 1] - bb;
  // the diagonal block is stored first, the others follow in column order
  const int hd = hasdiag[n];
  bool seen = false;
  for (int m, kc = 0; (m = tet_merge_min<DPN>(ci, j, e)) != INT_MAX; kc++) {
    const int k = (m == n) ? 0 : kc + (hd && !seen ? 1 : 0);
    seen |= m == n;
    unsigned mask = 0u;
    unsigned long long lmb = 0ull;
    for (int p = 0; p < np; p++) {
      const int el = n2e[pa + p] >> 3;
#pragma unroll
      for (int lm = 0; lm < 4; lm++)
        if (conn[(int64_t)el * 4 + lm] == m) {
          mask |= 1u << p;
          lmb |= (unsigned long long)lm << (2 * p);
        }
    }
    if (mask == 0u) atomicExch(err, 1);  // an entry no element produces
    bmask[bb + k] = mask;
    blmb[bb + k] = lmb;
#pragma unroll
    for (int a = 0; a < DPN; a++) {
      const int64_t gi = (int64_t)DPN * bb + a * nb + k;
      const int first = j[a];
      unsigned bits = 0u;
      for (; j[a] < e[a] && ci[j[a]] / DPN == m; j[a]++) bits |= 1u << (ci[j[a]] % DPN);
      ipos[gi] = first;
      ipm[gi] = (unsigned char)bits;
    }
  }
}

// Item records in the numeric pass's order, one thread per tile: the tile's
// diagonal blocks (first in each node's block list) in node order, then the
// off-diagonal blocks in node and column order.  A tile's items occupy the
// block range of its nodes, [bptr[n0], bptr[n1]).
template <int DPN>
__global__ void k_tet_plan_items(int ntiles, const int *__restrict__ tiles, const int *__restrict__ tpa,
                                 const int *__restrict__ n2e_ptr, const int *__restrict__ bptr,
                                 const unsigned char *__restrict__ hasdiag, const unsigned *__restrict__ bmask,
                                 const unsigned long long *__restrict__ blmb, const int *__restrict__ bpos,
                                 const unsigned char *__restrict__ bpm, int64_t nitems, int *__restrict__ tib,
                                 unsigned *__restrict__ imask, unsigned long long *__restrict__ ilmb,
                                 int *__restrict__ ipos, unsigned *__restrict__ ipmk) {
  const int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= ntiles) return;
  const int n0 = tiles[tile], n1 = tiles[tile + 1];
  const int b0 = bptr[n0];
  tib[tile] = b0;
  if (tile == ntiles - 1) tib[ntiles] = bptr[n1];
  int nd = 0;
  for (int n = n0; n < n1; n++) nd += hasdiag[n];
  int d = 0, o = 0;
  for (int n = n0; n < n1; n++) {
    const int bb = bptr[n], nb = bptr[n + 1] - bb, hd = hasdiag[n];
    const unsigned kb = (unsigned)(n2e_ptr[n] - tpa[tile]);
    for (int k = 0; k < nb; k++) {
      const int64_t item = b0 + (hd && k == 0 ? d : nd + o + (k - hd));
      imask[item] = bmask[bb + k];
      ilmb[item] = blmb[bb + k];
      unsigned w = kb << 16;
#pragma unroll
      for (int a = 0; a < DPN; a++) {
        const int64_t gi = (int64_t)DPN * bb + a * nb + k;
        ipos[a * nitems + item] = bpos[gi];
        w |= (unsigned)bpm[gi] << (4 * a);
      }
      ipmk[item] = w;
    }
    d += hd;
    o += nb - hd;
  }
}

// Tile schedule: consecutive nodes with <= kTetTP pairs and <= kTetNodes nodes.
static int tet_tiles(tfm_ctx *ctx, const tfm_mesh *mesh, TetPlan *tp) {
  const int64_t N = mesh->n_nodes;
  const std::vector<int> &p = mesh->h_n2e_ptr;
  std::vector<int> t;
  t.push_back(0);
  int start = 0;
  for (int n = 0; n < (int)N; n++) {
    if (p[n + 1] - p[n] > 32) return fail(ctx, TFM_E_UNSUPPORTED, "node too large for the tet tile");
    if (p[n + 1] - p[start] > kTetTP || n - start + 1 > kTetNodes) {
      t.push_back(n);
      start = n;
    }
  }
  if (N > 0) t.push_back((int)N);
  tp->ntiles = (int)t.size() - 1;
  std::vector<int> pa(t.size());
  for (size_t i = 0; i < t.size(); i++) pa[i] = p[t[i]];
  int rc;
  if ((rc = dalloc(ctx, &tp->d_tiles, t.size())) || (rc = dalloc(ctx, &tp->d_tpa, t.size()))) return rc;
  TFM_CHECK_HIP(ctx, hipMemcpy(tp->d_tiles, t.data(), sizeof(int) * t.size(), hipMemcpyHostToDevice));
  TFM_CHECK_HIP(ctx, hipMemcpy(tp->d_tpa, pa.data(), sizeof(int) * t.size(), hipMemcpyHostToDevice));
  return TFM_OK;
}

int tet_plan_create(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, const int *d_rp, const int *d_ci, TetPlan *tp) {
  const int64_t N = mesh->n_nodes;
  hipStream_t st = ctx->stream;
  int rc;
  if ((rc = tet_tiles(ctx, mesh, tp))) return rc;
  // per-node block lists (temporaries), then the item records
  int *cnt = nullptr, *err = nullptr, *bptr = nullptr, *bpos = nullptr;
  unsigned char *hasdiag = nullptr, *bpm = nullptr;
  unsigned *bmask = nullptr;
  unsigned long long *blmb = nullptr;
  auto cleanup = [&]() {
    dfree(cnt);
    dfree(err);
    dfree(bptr);
    dfree(hasdiag);
    dfree(bmask);
    dfree(blmb);
    dfree(bpos);
    dfree(bpm);
  };
  if ((rc = dalloc(ctx, &cnt, N)) || (rc = dalloc(ctx, &bptr, N + 1)) || (rc = dalloc(ctx, &err, 1)) ||
      (rc = dalloc(ctx, &hasdiag, N))) {
    cleanup();
    return rc;
  }
  const unsigned g = (unsigned)((N + 255) / 256);
  if (N > 0) {
    if (dpn == 1)
      hipLaunchKernelGGL(k_tet_plan_count<1>, dim3(g), dim3(256), 0, st, (int)N, d_rp, d_ci, cnt, hasdiag);
    else
      hipLaunchKernelGGL(k_tet_plan_count<3>, dim3(g), dim3(256), 0, st, (int)N, d_rp, d_ci, cnt, hasdiag);
  }
  int64_t nblk = 0;
  if ((rc = exclusive_scan_i32(ctx, cnt, bptr, N, &nblk))) {
    cleanup();
    return rc;
  }
  tp->nitems = nblk;
  if ((rc = dalloc(ctx, &bmask, nblk)) || (rc = dalloc(ctx, &blmb, nblk)) || (rc = dalloc(ctx, &bpos, nblk * dpn)) ||
      (rc = dalloc(ctx, &bpm, nblk * dpn)) || (rc = dalloc(ctx, &tp->d_tib, (int64_t)tp->ntiles + 1)) ||
      (rc = dalloc(ctx, &tp->d_mask, nblk)) || (rc = dalloc(ctx, &tp->d_lmb, nblk)) ||
      (rc = dalloc(ctx, &tp->d_pos, nblk * dpn)) || (rc = dalloc(ctx, &tp->d_pmk, nblk))) {
    cleanup();
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipMemsetAsync(err, 0, sizeof(int), st));
  if (N > 0) {
    const unsigned gt = (unsigned)((tp->ntiles + 255) / 256);
    if (dpn == 1) {
      hipLaunchKernelGGL(k_tet_plan_fill<1>, dim3(g), dim3(256), 0, st, (int)N, d_rp, d_ci, mesh->d_conn,
                         mesh->d_n2e_ptr, mesh->d_n2e, bptr, hasdiag, bmask, blmb, bpos, bpm, err);
      hipLaunchKernelGGL(k_tet_plan_items<1>, dim3(gt), dim3(256), 0, st, tp->ntiles, tp->d_tiles, tp->d_tpa,
                         mesh->d_n2e_ptr, bptr, hasdiag, bmask, blmb, bpos, bpm, nblk, tp->d_tib, tp->d_mask,
                         tp->d_lmb, tp->d_pos, tp->d_pmk);
    } else {
      hipLaunchKernelGGL(k_tet_plan_fill<3>, dim3(g), dim3(256), 0, st, (int)N, d_rp, d_ci, mesh->d_conn,
                         mesh->d_n2e_ptr, mesh->d_n2e, bptr, hasdiag, bmask, blmb, bpos, bpm, err);
      hipLaunchKernelGGL(k_tet_plan_items<3>, dim3(gt), dim3(256), 0, st, tp->ntiles, tp->d_tiles, tp->d_tpa,
                         mesh->d_n2e_ptr, bptr, hasdiag, bmask, blmb, bpos, bpm, nblk, tp->d_tib, tp->d_mask,
                         tp->d_lmb, tp->d_pos, tp->d_pmk);
    }
  }
  int herr = 0;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  cleanup();
  if (herr) return fail(ctx, TFM_E_UNSUPPORTED, "pattern not covered by the tet plan");
  return TFM_OK;
}

void tet_plan_destroy(TetPlan *tp) {
  dfree(tp->d_tiles);
  dfree(tp->d_tpa);
  dfree(tp->d_tib);
  dfree(tp->d_mask);
  dfree(tp->d_lmb);
  dfree(tp->d_pos);
  dfree(tp->d_pmk);
  *tp = TetPlan{};
}

int tet_numeric(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, const TetPlan &tp, const OpParams &P, double *v64,
                float *v32) {
  if (tp.ntiles == 0) return TFM_OK;
  hipStream_t st = ctx->stream;
  dim3 g(xcd_grid(tp.ntiles)), t(kTetTP);
#define TFM_TET(D, VT, V)                                                                                        \
  hipLaunchKernelGGL((k_asm_tet<D, VT>), g, t, 0, st, mesh->d_coords, mesh->d_conn, mesh->d_n2e, tp.d_tpa, tp.d_tib, \
                     tp.ntiles, P, tp.d_mask, tp.d_lmb, tp.d_pos, tp.d_pmk, tp.nitems, V)
  if (dpn == 1) {
    if (v64) TFM_TET(1, double, v64);
    if (v32) TFM_TET(1, float, v32);
  } else {
    if (v64) TFM_TET(3, double, v64);
    if (v32) TFM_TET(3, float, v32);
  }
#undef TFM_TET
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

}  // namespace tfm
