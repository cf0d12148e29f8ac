// tfm_asm_tet.hip -- numeric assembly fast path for P1 tetrahedra.
//
// Same result as the generic tile kernel (k_asm_tile, bit-identical FP64
// values), organised for occupancy and balance:
//   * phase 0: the tile's col_ind range is staged in LDS with coalesced loads;
//   * phase 1: one lane per (node, element) pair computes the element's
//     Map (J, det, J^-1, the four physical gradients) once and stores a
//     16-value geometry record -- not the 36 entries of the row block -- and
//     registers the pair with each off-diagonal neighbour in a per-node LDS
//     hash (neighbour -> bitmask of contributing pairs, 2-bit local index);
//   * phase 2: work items are the 9 entries of each diagonal block (24
//     contributions each, one lane per entry) and the off-diagonal blocks (one
//     lane per 3x3 block, ~5 contributions of 9 entries), so lanes carry
//     similar work.  Each item evaluates its entries from the geometry records
//     in element-id order -- the reference's summation order -- and writes them
//     at the column positions found by binary search in the staged col_ind.
// Entry arithmetic is the reference's (tfm_elem.h: bdb3, four equal
// quadrature contributions); this translation unit is built without FMA
// contraction.
#include <hip/hip_runtime.h>

#include <cstring>

#include "tfm_elem.h"
#include "tfm_internal.h"

#pragma clang fp contract(off)

namespace tfm {

constexpr int kTetTP = 128;     // pairs (and threads) per tile
constexpr int kTetNodes = 16;   // nodes per tile
constexpr int kTetHash = 32;    // off-diagonal neighbours per node (hash slots)
constexpr int kTetSlots = 1024; // col_ind entries staged per tile
constexpr int kTetPairsPerNode = 32;  // pair bitmask width

template <typename VT>
struct TetRec {
  // g[4][3] physical gradients, then (elasticity) wd = w*|det|, m, c*nu, dd
  // or (Poisson) wk = w*kappa, |det|
  VT v[16];
};

__device__ __forceinline__ unsigned thash(int key) { return ((unsigned)key * 2654435761u) >> (32 - 5); }

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

// Entry (a, b) of element row block (ln, lm) from a geometry record.
template <int DPN, typename VT>
__device__ __forceinline__ VT tet_entry(const VT *R, int ln, int lm, int a, int b) {
  if constexpr (sizeof(VT) == 8) {
    const double *gi = R + 3 * ln, *gj = R + 3 * lm;
    if constexpr (DPN == 1) {
      double dot = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
      double c = R[12] * dot * R[13];
      return c + c + c + c;
    } else {
      DConst D{};
      D.m = R[13];
      D.cnu = R[14];
      D.dd = R[15];
      double c = R[12] * bdb3(gi, gj, a, b, D);
      return c + c + c + c;
    }
  } else {
    const float *gi = R + 3 * ln, *gj = R + 3 * lm;
    if constexpr (DPN == 1) {
      return R[12] * (gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]);
    } else {
      return R[12] * bdb3f(gi, gj, a, b, R[13], R[14], R[15]);
    }
  }
}

// First position >= key in the sorted array s[lo, hi).
__device__ __forceinline__ int lower_bound_lds(const int *s, int lo, int hi, int key) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (s[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <int DPN, typename VT>
__global__ __launch_bounds__(kTetTP) void k_asm_tet(const double *__restrict__ coords, const int *__restrict__ conn,
                                                     const int *__restrict__ n2e_ptr, const int *__restrict__ n2e,
                                                     const int *__restrict__ tiles, int ntiles, OpParams P,
                                                     const int *__restrict__ rp, const int *__restrict__ ci,
                                                     VT *__restrict__ vals, int *__restrict__ err) {
  __shared__ TetRec<VT> s_rec[kTetTP];
  __shared__ int s_en[kTetTP][4];
  __shared__ unsigned char s_ln[kTetTP];
  __shared__ int s_nptr[kTetNodes + 1];
  __shared__ int s_key[kTetNodes][kTetHash];
  __shared__ unsigned int s_mask[kTetNodes][kTetHash];
  __shared__ unsigned long long s_lmb[kTetNodes][kTetHash];
  __shared__ unsigned char s_list[kTetNodes][kTetHash];
  __shared__ int s_nk[kTetNodes];
  __shared__ int s_ipre[kTetNodes + 1];
  __shared__ int s_rp[kTetNodes * DPN + 1];
  __shared__ int s_ci[kTetSlots];

  const int tid = threadIdx.x;
  const int tile = xcd_item(blockIdx.x, ntiles);  // neighbouring tiles share an XCD's L2
  if (tile >= ntiles) return;
  const int n0 = tiles[tile], n1 = tiles[tile + 1];
  const int nt = n1 - n0;
  const int pa = n2e_ptr[n0];
  const int np = n2e_ptr[n1] - pa;

  for (int i = tid; i <= nt; i += kTetTP) s_nptr[i] = n2e_ptr[n0 + i] - pa;
  for (int i = tid; i <= nt * DPN; i += kTetTP) s_rp[i] = rp[n0 * DPN + i];
  for (int i = tid; i < kTetNodes * kTetHash; i += kTetTP) {
    s_key[i / kTetHash][i % kTetHash] = -1;
    s_mask[i / kTetHash][i % kTetHash] = 0u;
    s_lmb[i / kTetHash][i % kTetHash] = 0ull;
  }
  if (tid < kTetNodes) s_nk[tid] = 0;
  __syncthreads();
  const int sb = s_rp[0], se = s_rp[nt * DPN];
  for (int s = sb + tid; s < se; s += kTetTP) s_ci[s - sb] = ci[s];

  // ---- phase 1: geometry record per pair, neighbour hash ------------------
  if (tid < np) {
    const int v = n2e[pa + tid];
    const int e = v >> 3, ln = v & 7;
    int tn = 0;
    while (tn + 1 < nt && s_nptr[tn + 1] <= tid) tn++;
    const int k = tid - s_nptr[tn];
    if (k >= kTetPairsPerNode) atomicExch(err, 1);
    double X[4][3];
    int en[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      en[a] = conn[(int64_t)e * 4 + a];
      s_en[tid][a] = en[a];
#pragma unroll
      for (int p = 0; p < 3; p++) X[a][p] = coords[(int64_t)en[a] * 3 + p];
    }
    s_ln[tid] = (unsigned char)ln;
    const double coef = P.coef_e ? P.coef_e[e] : P.coef;
    double g[4][3];
    const double det = fabs(map_tet(X, g));
    VT *R = s_rec[tid].v;
    if constexpr (sizeof(VT) == 8) {
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int p = 0; p < 3; p++) R[3 * a + p] = g[a][p];
      const double w = 1.0 / 24.0;
      if constexpr (DPN == 1) {
        R[12] = w * coef;
        R[13] = det;
      } else {
        DConst D = make_d(TFM_CONTINUUM3D, coef, P.nu);
        R[12] = w * det;
        R[13] = D.m;
        R[14] = D.cnu;
        R[15] = D.dd;
      }
    } else {
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int p = 0; p < 3; p++) R[3 * a + p] = (float)g[a][p];
      const double vol = det * (1.0 / 6.0);  // the four equal-weight points
      if constexpr (DPN == 1) {
        R[12] = (float)(vol * coef);
      } else {
        const double c = coef / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
        R[12] = (float)vol;
        R[13] = (float)((1.0 - P.nu) * c);
        R[14] = (float)(c * P.nu);
        R[15] = (float)((1.0 - 2.0 * P.nu) / 2.0 * c);
      }
    }
#pragma unroll
    for (int lm = 0; lm < 4; lm++) {
      if (lm == ln) continue;
      const int key = en[lm];
      unsigned h = thash(key) & (kTetHash - 1);
      int probe = 0;
      for (; probe < kTetHash; probe++) {
        const int old = atomicCAS(&s_key[tn][h], -1, key);
        if (old == -1 || old == key) {
          atomicOr(&s_mask[tn][h], 1u << (k & 31));
          atomicOr(&s_lmb[tn][h], (unsigned long long)lm << (2 * (k & 31)));
          if (old == -1) s_list[tn][atomicAdd(&s_nk[tn], 1)] = (unsigned char)h;
          break;
        }
        h = (h + 1) & (kTetHash - 1);
      }
      if (probe == kTetHash) atomicExch(err, 1);
    }
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int t = 0; t < nt; t++) {
      s_ipre[t] = acc;
      acc += DPN * DPN + s_nk[t];
    }
    s_ipre[nt] = acc;
  }
  __syncthreads();

  // ---- phase 2: diagonal entries and off-diagonal blocks ------------------
  const int nitems = s_ipre[nt];
  for (int it = tid; it < nitems; it += kTetTP) {
    int tn = 0;
    while (tn + 1 < nt && s_ipre[tn + 1] <= it) tn++;
    const int loc = it - s_ipre[tn];
    const int node = n0 + tn;
    const int kb = s_nptr[tn], ke = s_nptr[tn + 1];
    if (loc < DPN * DPN) {
      const int a = loc / DPN, b = loc % DPN;
      const int r = tn * DPN + a;
      const int lo = s_rp[r] - sb, hi = s_rp[r + 1] - sb;
      const int pos = lower_bound_lds(s_ci, lo, hi, node * DPN + b);
      if (pos >= hi || s_ci[pos] != node * DPN + b) continue;  // fixed column or absent
      VT acc = 0;
      for (int i = kb; i < ke; i++) {
        const int ln = s_ln[i];
        const VT x = tet_entry<DPN, VT>(s_rec[i].v, ln, ln, a, b);
        if (x != (VT)0) acc += x;
      }
      vals[sb + pos] = acc;
    } else {
      const int h = s_list[tn][loc - DPN * DPN];
      const int m = s_key[tn][h];
      unsigned int mask = s_mask[tn][h];
      const unsigned long long lb = s_lmb[tn][h];
      VT acc[DPN * DPN];
#pragma unroll
      for (int q = 0; q < DPN * DPN; q++) acc[q] = 0;
      while (mask) {
        const int k = __ffs((int)mask) - 1;
        mask &= mask - 1;
        const int i = kb + k;
        const int lm = (int)((lb >> (2 * k)) & 3ull);
        const int ln = s_ln[i];
#pragma unroll
        for (int a = 0; a < DPN; a++)
#pragma unroll
          for (int b = 0; b < DPN; b++) {
            const VT x = tet_entry<DPN, VT>(s_rec[i].v, ln, lm, a, b);
            if (x != (VT)0) acc[a * DPN + b] += x;
          }
      }
#pragma unroll
      for (int a = 0; a < DPN; a++) {
        const int r = tn * DPN + a;
        const int lo = s_rp[r] - sb, hi = s_rp[r + 1] - sb;
        int pos = lower_bound_lds(s_ci, lo, hi, m * DPN);
#pragma unroll
        for (int b = 0; b < DPN; b++)
          if (pos < hi && s_ci[pos] == m * DPN + b) {
            vals[sb + pos] = acc[a * DPN + b];
            pos++;
          }
      }
    }
  }
}

// Tile schedule for this kernel: consecutive nodes with <= kTetTP pairs,
// <= kTetNodes nodes and <= kTetSlots CSR slots.
int tet_tiles(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, const int *d_rp, int **d_tiles, int *ntiles) {
  const int64_t N = mesh->n_nodes;
  std::vector<int> rp((size_t)(N * dpn + 1));
  TFM_CHECK_HIP(ctx, hipMemcpy(rp.data(), d_rp, sizeof(int) * rp.size(), hipMemcpyDeviceToHost));
  const std::vector<int> &p = mesh->h_n2e_ptr;
  std::vector<int> t;
  t.push_back(0);
  int start = 0;
  for (int n = 0; n < (int)N; n++) {
    const int pairs = p[n + 1] - p[start];
    const int nodes = n - start + 1;
    const int slots = rp[(size_t)(n + 1) * dpn] - rp[(size_t)start * dpn];
    if (pairs > kTetTP || nodes > kTetNodes || slots > kTetSlots) {
      t.push_back(n);
      start = n;
    }
    const int own = rp[(size_t)(n + 1) * dpn] - rp[(size_t)n * dpn];
    if (p[n + 1] - p[n] > kTetTP || own > kTetSlots) return fail(ctx, TFM_E_UNSUPPORTED, "node too large for the tet tile");
  }
  if (N > 0) t.push_back((int)N);
  *ntiles = (int)t.size() - 1;
  int rc = dalloc(ctx, d_tiles, t.size());
  if (rc) return rc;
  TFM_CHECK_HIP(ctx, hipMemcpy(*d_tiles, t.data(), sizeof(int) * t.size(), hipMemcpyHostToDevice));
  return TFM_OK;
}

int tet_numeric(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, const int *d_tiles, int ntiles, const OpParams &P,
                const int *d_rp, const int *d_ci, double *v64, float *v32, int *d_err) {
  if (ntiles == 0) return TFM_OK;
  hipStream_t st = ctx->stream;
  dim3 g(xcd_grid(ntiles)), t(kTetTP);
  if (dpn == 1) {
    if (v64) hipLaunchKernelGGL((k_asm_tet<1, double>), g, t, 0, st, mesh->d_coords, mesh->d_conn, mesh->d_n2e_ptr,
                                mesh->d_n2e, d_tiles, ntiles, P, d_rp, d_ci, v64, d_err);
    if (v32) hipLaunchKernelGGL((k_asm_tet<1, float>), g, t, 0, st, mesh->d_coords, mesh->d_conn, mesh->d_n2e_ptr,
                                mesh->d_n2e, d_tiles, ntiles, P, d_rp, d_ci, v32, d_err);
  } else {
    if (v64) hipLaunchKernelGGL((k_asm_tet<3, double>), g, t, 0, st, mesh->d_coords, mesh->d_conn, mesh->d_n2e_ptr,
                                mesh->d_n2e, d_tiles, ntiles, P, d_rp, d_ci, v64, d_err);
    if (v32) hipLaunchKernelGGL((k_asm_tet<3, float>), g, t, 0, st, mesh->d_coords, mesh->d_conn, mesh->d_n2e_ptr,
                                mesh->d_n2e, d_tiles, ntiles, P, d_rp, d_ci, v32, d_err);
  }
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

}  // namespace tfm
