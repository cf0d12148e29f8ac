// tfm_asm_coo.hip -- assembly by sorted triplets, for meshes whose node
// valence exceeds the tile kernels' limits (a node in more than 64 elements,
// or with more distinct neighbours than a tile's per-node hash holds: the
// Delaunay meshes crates/tpt-fem-mesh-gen/src/lib.rs:163-409 builds and the
// Gmsh meshes crates/tpt-fem-mesh/src/lib.rs:260-388 imports).
//
// This is the reference's own route, restated for the GPU: assemble pushes
// every non-zero element entry as a triplet in element order, local row,
// local column (crates/tpt-fem-assembly/src/lib.rs:428-480), and
// Coo::to_csr sums duplicates in insertion order
// (crates/tpt-fem-sparse/src/lib.rs:39-130).  Here one thread per
// (element, local node) pair evaluates its row block and writes its kept
// triplets at its scanned offset (key = row << 32 | column, payload = the
// entry's position in element-major order), a stable radix sort by key
// brings every CSR entry's contributions together in that insertion order,
// and the numeric pass sums each entry's contributions in that order from a
// dense table of element values.  Results equal the tile kernels' (and the
// reference's) bit for bit; it costs a sort at plan creation and a table of
// all element entries per numeric pass.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "tfm_elem.h"
#include "tfm_internal.h"

#pragma clang fp contract(off)

namespace tfm {

enum { C_COUNT = 0, C_FILL = 1, C_F64 = 2, C_F32 = 3 };

// One (element, local node) pair per thread, t = e * NN + ln (element-major).
template <int CELL, int DPN, int WHAT>
__global__ __launch_bounds__(128) void k_coo_pairs(int64_t npairs, const double *__restrict__ coords, int pdim,
                                                    const int *__restrict__ conn, OpParams P,
                                                    const unsigned char *__restrict__ fixed, int *__restrict__ cnt,
                                                    const int *__restrict__ off, unsigned long long *__restrict__ keys,
                                                    int64_t *__restrict__ idx, double *__restrict__ v64,
                                                    float *__restrict__ v32) {
  constexpr int NN = Cell<CELL>::NN, DIM = Cell<CELL>::DIM, W = NN * DPN, ND = NN * DPN;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  const int64_t e = t / NN;
  const int ln = (int)(t - e * NN);
  int en[NN];
  double X[NN][3];
#pragma unroll
  for (int a = 0; a < NN; a++) {
    en[a] = conn[e * NN + a];
#pragma unroll
    for (int p = 0; p < 3; p++) X[a][p] = p < DIM ? coords[(int64_t)en[a] * pdim + p] : 0.0;
  }
  const double coef = P.coef_e ? P.coef_e[e] : P.coef;
  const double rho = P.rho ? P.rho[e] : 0.0;
  const int64_t base = e * (ND * ND) + (int64_t)ln * DPN * ND;  // first entry of the pair's rows
  if constexpr (WHAT == C_F32) {
    float tmp[DPN * W];
    rowblock_f32<CELL, DPN>(X, ln, coef, rho, P, tmp);
#pragma unroll
    for (int i = 0; i < DPN * W; i++) v32[base + i] = tmp[i];
  } else {
    double tmp[DPN * W];
    rowblock_f64<CELL, DPN>(X, ln, coef, rho, P, tmp);
    if constexpr (WHAT == C_F64) {
#pragma unroll
      for (int i = 0; i < DPN * W; i++) v64[base + i] = tmp[i];
    } else {
      int c = 0, o = WHAT == C_FILL ? off[t] : 0;
#pragma unroll
      for (int a = 0; a < DPN; a++) {
        const int row = en[ln] * DPN + a;
        if (fixed && fixed[row]) continue;
#pragma unroll
        for (int lm = 0; lm < NN; lm++)
#pragma unroll
          for (int b = 0; b < DPN; b++) {
            const int col = en[lm] * DPN + b;
            // the reference pushes non-zero entries (every SIMP entry)
            if ((P.kind != TFM_OP_SIMP && tmp[a * W + lm * DPN + b] == 0.0) || (fixed && fixed[col])) continue;
            if constexpr (WHAT == C_FILL) {
              keys[o + c] = ((unsigned long long)row << 32) | (unsigned)col;
              idx[o + c] = base + a * ND + lm * DPN + b;
            }
            c++;
          }
      }
      if constexpr (WHAT == C_COUNT) cnt[t] = c;
    }
  }
}

// First triplet of each CSR entry: flag[k] = key[k] != key[k - 1].
__global__ void k_coo_heads(int64_t n, const unsigned long long *__restrict__ keys, int *__restrict__ flag) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) flag[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}

// Per CSR entry q (its head k): column, first triplet, and the row count.
// Keys are sorted, so a wavefront's heads come in runs of equal rows: one
// atomic per run.  The largest row and column go through a block maximum
// into one of kSpread slot pairs (atomics on a single word from every block
// would serialise), which k_coo_dims folds.
constexpr int kSpread = 256;
__global__ __launch_bounds__(256) void k_coo_entries(int64_t n, const unsigned long long *__restrict__ keys,
                                                     const int *__restrict__ flag, const int *__restrict__ eid,
                                                     int *__restrict__ ci, int *__restrict__ seg,
                                                     int *__restrict__ rcnt, int *__restrict__ spread) {
  __shared__ int s_m[2][4];
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool head = k < n && flag[k];
  int row = -1, col = -1;
  if (head) {
    const int q = eid[k];
    row = (int)(keys[k] >> 32);
    col = (int)(keys[k] & 0xffffffffull);
    ci[q] = col;
    seg[q] = (int)k;
  }
  const int lane = threadIdx.x & 63;
  unsigned long long left = __ballot(head);
  while (left) {
    const int lead = __ffsll((long long)left) - 1;
    const int r = __shfl(row, lead, 64);
    const unsigned long long run = __ballot(head && row == r);
    if (lane == lead) atomicAdd(&rcnt[r], __popcll(run));
    left &= ~run;
  }
  int mr = row, mc = col;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mr = max(mr, __shfl_xor(mr, o, 64));
    mc = max(mc, __shfl_xor(mc, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if (lane == 0) {
    s_m[0][w] = mr;
    s_m[1][w] = mc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    mr = max(max(s_m[0][0], s_m[0][1]), max(s_m[0][2], s_m[0][3]));
    mc = max(max(s_m[1][0], s_m[1][1]), max(s_m[1][2], s_m[1][3]));
    if (mr >= 0) {
      int *sl = spread + 2 * (blockIdx.x % kSpread);
      atomicMax(&sl[0], mr + 1);
      atomicMax(&sl[1], mc + 1);
    }
  }
}

// dims = the slot pairs' maxima (one block of kSpread threads).
__global__ void k_coo_dims(const int *__restrict__ spread, int *__restrict__ dims) {
  __shared__ int s[2][kSpread];
  const int t = threadIdx.x;
  s[0][t] = spread[2 * t];
  s[1][t] = spread[2 * t + 1];
  __syncthreads();
  for (int o = kSpread / 2; o > 0; o >>= 1) {
    if (t < o) {
      s[0][t] = max(s[0][t], s[0][t + o]);
      s[1][t] = max(s[1][t], s[1][t + o]);
    }
    __syncthreads();
  }
  if (t == 0) {
    dims[0] = s[0][0];
    dims[1] = s[1][0];
  }
}

// One CSR entry per thread: its contributions in insertion order from +0.0.
template <typename VT, typename IT>
__global__ void k_coo_sum(int64_t nnz, const int *__restrict__ seg, const IT *__restrict__ perm,
                          const VT *__restrict__ dense, VT *__restrict__ out) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  VT acc = 0;
  for (int k = seg[q]; k < seg[q + 1]; k++) {
    const VT v = dense[perm[k]];
    if (v != (VT)0) acc += v;
  }
  out[q] = acc;
}

// Both value arrays: the FP64 sum and its rounding to FP32.
template <typename IT>
__global__ void k_coo_sum_dual(int64_t nnz, const int *__restrict__ seg, const IT *__restrict__ perm,
                               const double *__restrict__ dense, double *__restrict__ out,
                               float *__restrict__ out32) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  double acc = 0;
  for (int k = seg[q]; k < seg[q + 1]; k++) {
    const double v = dense[perm[k]];
    if (v != 0.0) acc += v;
  }
  out[q] = acc;
  out32[q] = (float)acc;
}

// FP64 element matrices of 3-D elasticity on Hex8 cells, the reference's
// values bit for bit (crates/tpt-fem-elasticity/src/lib.rs:284-307), into the
// dense table ke[e][24][24] (row (i, a), column (j, b)) the sorted-triplet sum
// reads.  One wavefront per element: lane q < nq forms quadrature point q's
// map (map_point, as every FP64 pass does) into LDS; then lane (i, j) adds
// the nine entries K[(i,a),(j,b)] += (w |det|) * bdb3(g_i, g_j, a, b) over the
// points in rule order, from 0.0 -- the per-pair row block's operations and
// order -- with (a, b) the same across the wavefront; the 576 values leave
// through LDS as one contiguous block.  The element is formed once, where a
// per-pair pass forms it once per element node.
constexpr int kKeWaves = 4;
constexpr int kKeMaxQ = 27;  // Gauss-Legendre orders 1-3

__device__ __forceinline__ void ke_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * kKeWaves) void k_hex_ke64(int64_t ne, const double *__restrict__ coords,
                                                             const int *__restrict__ conn, OpParams P,
                                                             double *__restrict__ ke) {
  __shared__ double s_x[kKeWaves][8][3];
  __shared__ double s_g[kKeWaves][kKeMaxQ][8][3];
  __shared__ double s_wd[kKeWaves][kKeMaxQ];
  __shared__ double s_k[kKeWaves][576];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * kKeWaves;
  const int nq = P.nq;
  for (int64_t e = (int64_t)blockIdx.x * kKeWaves + w; e < ne; e += nwaves) {
    if (l < 24) s_x[w][l / 3][l % 3] = coords[(int64_t)conn[e * 8 + l / 3] * 3 + l % 3];
    const double coef = P.coef_e ? P.coef_e[e] : P.coef;
    const DConst D = make_d(P.model, coef, P.nu);
    ke_wave_sync();
    if (l < nq) {
      double X[8][3];
#pragma unroll
      for (int a = 0; a < 8; a++)
#pragma unroll
        for (int p = 0; p < 3; p++) X[a][p] = s_x[w][a][p];
      const double *R = P.rule + l * kRuleStride;
      double g[8][3];
      const double det = fabs(map_point<8, 3>(X, R + 12, g));
      s_wd[w][l] = R[0] * det;
#pragma unroll
      for (int a = 0; a < 8; a++)
#pragma unroll
        for (int p = 0; p < 3; p++) s_g[w][l][a][p] = g[a][p];
    }
    ke_wave_sync();
    const int i = l >> 3, j = l & 7;
    double acc[9];
#pragma unroll
    for (int c = 0; c < 9; c++) acc[c] = 0.0;
    for (int q = 0; q < nq; q++) {
      const double wd = s_wd[w][q];
      double gi[3], gj[3];
#pragma unroll
      for (int p = 0; p < 3; p++) {
        gi[p] = s_g[w][q][i][p];
        gj[p] = s_g[w][q][j][p];
      }
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) acc[a * 3 + b] += wd * bdb3(gi, gj, a, b, D);
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) s_k[w][(i * 3 + a) * 24 + j * 3 + b] = acc[a * 3 + b];
    ke_wave_sync();
#pragma unroll
    for (int t = 0; t < 9; t++) ke[e * 576 + t * 64 + l] = s_k[w][t * 64 + l];
    ke_wave_sync();  // s_x, s_g and s_k are rewritten for the next element
  }
}

bool hex_table_route(const tfm_mesh *m, int dpn, const tfm_operator *op) {
  return m->cell == TFM_HEX && dpn == 3 && op->kind == TFM_OP_ELASTICITY && op->model == TFM_CONTINUUM3D &&
         op->quad_order >= 1 && op->quad_order <= 3 && !getenv("TFM_HEX_TILE");
}

template <int WHAT>
static int coo_launch(tfm_ctx *ctx, const AsmPlan *pl, const OpParams &P, int *cnt, const int *off,
                      unsigned long long *keys, int64_t *idx, double *v64, float *v32) {
  const tfm_mesh *m = pl->mesh;
  const int64_t npairs = m->n_elem * m->nn;
  if (npairs == 0) return TFM_OK;
  const dim3 g((unsigned)((npairs + 127) / 128)), b(128);
#define TFM_COO(C, D)                                                                                           \
  if (m->cell == C && pl->dpn == D) {                                                                           \
    hipLaunchKernelGGL((k_coo_pairs<C, D, WHAT>), g, b, 0, ctx->stream, npairs, m->d_coords, m->pdim, m->d_conn, \
                       P, pl->d_fixed, cnt, off, keys, idx, v64, v32);                                          \
    TFM_CHECK_HIP(ctx, hipGetLastError());                                                                      \
    return TFM_OK;                                                                                              \
  }
  TFM_COO(TFM_TET, 1)
  TFM_COO(TFM_TET, 3)
  TFM_COO(TFM_QUAD, 1)
  TFM_COO(TFM_QUAD, 2)
  TFM_COO(TFM_HEX, 1)
  TFM_COO(TFM_HEX, 3)
  TFM_COO(TFM_TRI, 1)
  TFM_COO(TFM_TRI, 2)
  TFM_COO(TFM_LINE, 1)
#undef TFM_COO
  return fail(ctx, TFM_E_UNSUPPORTED, "cell / dofs-per-node combination not built");
}

__global__ void k_narrow(int64_t n, const int64_t *__restrict__ a, int *__restrict__ b) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    b[k] = (int)a[k];
}

int coo_plan_create(tfm_ctx *ctx, AsmPlan *pl, const OpParams &P) {
  const tfm_mesh *m = pl->mesh;
  hipStream_t st = ctx->stream;
  const int64_t npairs = m->n_elem * m->nn;
  const int nd = m->nn * pl->dpn;
  if (m->n_elem * nd * nd >= ((int64_t)1 << 40)) return fail(ctx, TFM_E_UNSUPPORTED, "mesh too large for triplets");
  int *cnt = nullptr, *off = nullptr, *flag = nullptr, *eid = nullptr, *rcnt = nullptr, *dims = nullptr;
  unsigned long long *keys = nullptr, *keys2 = nullptr;
  int64_t *idx = nullptr;
  void *tmp = nullptr;
  auto cleanup = [&]() {
    dfree(cnt); dfree(off); dfree(flag); dfree(eid); dfree(rcnt); dfree(dims); dfree(keys); dfree(keys2); dfree(idx);
    dfree(tmp);
  };
  int rc;
  if ((rc = dalloc(ctx, &cnt, std::max<int64_t>(npairs, 1))) || (rc = dalloc(ctx, &off, npairs + 1))) {
    cleanup();
    return rc;
  }
  if ((rc = coo_launch<C_COUNT>(ctx, pl, P, cnt, nullptr, nullptr, nullptr, nullptr, nullptr))) {
    cleanup();
    return rc;
  }
  int64_t ntrip = 0;
  if ((rc = exclusive_scan_i32(ctx, cnt, off, npairs, &ntrip))) {  // fails beyond 2^31 triplets
    cleanup();
    return rc;
  }
  if ((rc = dalloc(ctx, &keys, std::max<int64_t>(ntrip, 1))) || (rc = dalloc(ctx, &keys2, std::max<int64_t>(ntrip, 1))) ||
      (rc = dalloc(ctx, &idx, std::max<int64_t>(ntrip, 1))) ||
      (rc = dalloc(ctx, &pl->d_coo_perm, std::max<int64_t>(ntrip, 1)))) {
    cleanup();
    return rc;
  }
  if ((rc = coo_launch<C_FILL>(ctx, pl, P, nullptr, off, keys, idx, nullptr, nullptr))) {
    cleanup();
    return rc;
  }
  // stable LSD radix sort of (key, element-major position): equal keys keep
  // their insertion order
  if (ntrip > 0) {
    size_t bytes = 0;
    TFM_CHECK_HIP(ctx, rocprim::radix_sort_pairs(nullptr, bytes, keys, keys2, idx, pl->d_coo_perm, (size_t)ntrip, 0,
                                                 64, st));
    if ((rc = dalloc(ctx, (char **)&tmp, bytes))) {
      cleanup();
      return rc;
    }
    TFM_CHECK_HIP(ctx, rocprim::radix_sort_pairs(tmp, bytes, keys, keys2, idx, pl->d_coo_perm, (size_t)ntrip, 0, 64,
                                                 st));
  }
  if ((rc = dalloc(ctx, &flag, std::max<int64_t>(ntrip, 1))) || (rc = dalloc(ctx, &eid, ntrip + 1))) {
    cleanup();
    return rc;
  }
  const unsigned gk = (unsigned)((ntrip + 255) / 256);
  if (ntrip > 0) hipLaunchKernelGGL(k_coo_heads, dim3(gk), dim3(256), 0, st, ntrip, keys2, flag);
  int64_t nnz = 0;
  if ((rc = exclusive_scan_i32(ctx, flag, eid, ntrip, &nnz))) {
    cleanup();
    return rc;
  }
  if ((rc = dalloc(ctx, &pl->d_ci, std::max<int64_t>(nnz, 1))) || (rc = dalloc(ctx, &pl->d_coo_seg, nnz + 1)) ||
      (rc = dalloc(ctx, &rcnt, pl->ndof + 1)) || (rc = dalloc(ctx, &dims, 2 + 2 * kSpread)) ||
      (rc = dalloc(ctx, &pl->d_rp, pl->ndof + 1))) {
    cleanup();
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipMemsetAsync(rcnt, 0, sizeof(int) * (pl->ndof + 1), st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(dims, 0, sizeof(int) * (2 + 2 * kSpread), st));
  if (ntrip > 0) {
    hipLaunchKernelGGL(k_coo_entries, dim3(gk), dim3(256), 0, st, ntrip, keys2, flag, eid, pl->d_ci, pl->d_coo_seg,
                       rcnt, dims + 2);
    hipLaunchKernelGGL(k_coo_dims, dim3(1), dim3(kSpread), 0, st, (const int *)(dims + 2), dims);
  }
  const int segend = (int)ntrip;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(pl->d_coo_seg + nnz, &segend, sizeof(int), hipMemcpyHostToDevice, st));
  int64_t nnz2 = 0;
  if ((rc = exclusive_scan_i32(ctx, rcnt, pl->d_rp, pl->ndof, &nnz2))) {
    cleanup();
    return rc;
  }
  int hd[2];
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hd, dims, sizeof hd, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  cleanup();
  // positions fit 32 bits for tables below 2^31 entries: the sum pass then
  // reads half the index bytes
  if (m->n_elem * nd * nd < INT32_MAX) {
    if ((rc = dalloc(ctx, &pl->d_coo_perm32, std::max<int64_t>(ntrip, 1)))) return rc;
    if (ntrip > 0)
      hipLaunchKernelGGL(k_narrow, dim3((unsigned)std::min<int64_t>((ntrip + 255) / 256, 65536)), dim3(256), 0, st,
                         ntrip, (const int64_t *)pl->d_coo_perm, pl->d_coo_perm32);
    TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
    dfree(pl->d_coo_perm);
    pl->d_coo_perm = nullptr;
  }
  pl->use_coo = true;
  pl->coo_ntrip = ntrip;
  pl->nnz = nnz;
  // Coo::to_csr: nrows = max row + 1, ncols = max col + 1, and an empty
  // triplet list still yields a 1 x 1 shape.
  pl->nrows = nnz ? hd[0] : 1;
  pl->ncols = nnz ? hd[1] : 1;
  return TFM_OK;
}

template <typename VT>
static void coo_sum(tfm_ctx *ctx, const AsmPlan *pl, const VT *dense, VT *out) {
  if (!pl->nnz) return;
  const dim3 g((unsigned)((pl->nnz + 255) / 256)), b(256);
  if (pl->d_coo_perm32)
    hipLaunchKernelGGL((k_coo_sum<VT, int>), g, b, 0, ctx->stream, pl->nnz, pl->d_coo_seg,
                       (const int *)pl->d_coo_perm32, dense, out);
  else
    hipLaunchKernelGGL((k_coo_sum<VT, int64_t>), g, b, 0, ctx->stream, pl->nnz, pl->d_coo_seg,
                       (const int64_t *)pl->d_coo_perm, dense, out);
}

static void coo_sum_dual(tfm_ctx *ctx, const AsmPlan *pl, const double *dense, double *out, float *out32) {
  if (!pl->nnz) return;
  const dim3 g((unsigned)((pl->nnz + 255) / 256)), b(256);
  if (pl->d_coo_perm32)
    hipLaunchKernelGGL((k_coo_sum_dual<int>), g, b, 0, ctx->stream, pl->nnz, pl->d_coo_seg,
                       (const int *)pl->d_coo_perm32, dense, out, out32);
  else
    hipLaunchKernelGGL((k_coo_sum_dual<int64_t>), g, b, 0, ctx->stream, pl->nnz, pl->d_coo_seg,
                       (const int64_t *)pl->d_coo_perm, dense, out, out32);
}

int coo_numeric(tfm_ctx *ctx, const AsmPlan *pl, const OpParams &P, double *v64, float *v32) {
  const tfm_mesh *m = pl->mesh;
  const int nd = m->nn * pl->dpn;
  const int64_t dense = m->n_elem * nd * nd;
  const bool hex = hex_table_route(m, pl->dpn, &pl->op) && P.nq <= kKeMaxQ;
  int rc;
  if (v64) {
    if (!pl->d_coo_v64 && (rc = dalloc(ctx, &pl->d_coo_v64, std::max<int64_t>(dense, 1)))) return rc;
    if (hex) {
      // each element once, by a wavefront (the table is the per-pair layout)
      if (m->n_elem > 0)
        hipLaunchKernelGGL(k_hex_ke64, dim3((unsigned)std::min<int64_t>((m->n_elem + kKeWaves - 1) / kKeWaves, 16384)),
                           dim3(64 * kKeWaves), 0, ctx->stream, m->n_elem, m->d_coords, m->d_conn, P, pl->d_coo_v64);
    } else if ((rc = coo_launch<C_F64>(ctx, pl, P, nullptr, nullptr, nullptr, nullptr, pl->d_coo_v64, nullptr))) {
      return rc;
    }
    if (v32) {
      // both arrays: the FP32 values are the FP64 sums rounded to nearest
      coo_sum_dual(ctx, pl, pl->d_coo_v64, v64, v32);
      TFM_CHECK_HIP(ctx, hipGetLastError());
      return TFM_OK;
    }
    coo_sum<double>(ctx, pl, pl->d_coo_v64, v64);
  }
  if (v32) {
    if (!pl->d_coo_v32 && (rc = dalloc(ctx, &pl->d_coo_v32, std::max<int64_t>(dense, 1)))) return rc;
    if (hex_mfma_eligible(pl)) {
      // the matrix cores' element table has the same [e][24][24] layout
      if ((rc = hex_ke_mfma(ctx, pl, pl->d_coo_v32))) return rc;
    } else if ((rc = coo_launch<C_F32>(ctx, pl, P, nullptr, nullptr, nullptr, nullptr, nullptr, pl->d_coo_v32))) {
      return rc;
    }
    coo_sum<float>(ctx, pl, pl->d_coo_v32, v32);
  }
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

void coo_plan_destroy(AsmPlan *pl) {
  dfree(pl->d_coo_perm);
  dfree(pl->d_coo_perm32);
  dfree(pl->d_coo_seg);
  dfree(pl->d_coo_v64);
  dfree(pl->d_coo_v32);
  pl->use_coo = false;
}

// Triplet t of the export: its row, column and value.
__global__ void k_coo_export(int64_t n, const unsigned long long *__restrict__ keys, const int64_t *__restrict__ idx,
                             const double *__restrict__ dense, int64_t *__restrict__ rows, int64_t *__restrict__ cols,
                             double *__restrict__ vals) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = keys[t];
    rows[t] = (int64_t)(k >> 32);
    cols[t] = (int64_t)(k & 0xffffffffull);
    vals[t] = dense[idx[t]];
  }
}

}  // namespace tfm

using namespace tfm;

// assemble(mesh, dpn, op) as the reference's Coo: the count pass, its scan
// (each (element, local node) pair's first triplet), the fill pass (keys and
// element-major positions in push order) and the FP64 element table.
extern "C" int tfm_assemble_coo(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn,
                                int64_t *count, int64_t *rows, int64_t *cols, double *vals) {
  if (!ctx || !mesh || !op || !count) return TFM_E_INVALID;
  const bool want = rows || cols || vals;
  if (want && !(rows && cols && vals)) return fail(ctx, TFM_E_INVALID, "rows, cols and vals go together");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  AsmPlan pl;
  OpParams P{};
  int *cnt = nullptr, *off = nullptr;
  unsigned long long *keys = nullptr;
  int64_t *idx = nullptr, *d_rows = nullptr, *d_cols = nullptr;
  double *dense = nullptr, *d_vals = nullptr;
  auto done = [&](int code) {
    (void)hipStreamSynchronize(st);
    dfree(cnt); dfree(off); dfree(keys); dfree(idx); dfree(d_rows); dfree(d_cols); dfree(dense); dfree(d_vals);
    plan_destroy(&pl);
    return code;
  };
  if ((rc = plan_tables(ctx, mesh, op, dpn, &pl, &P))) return done(rc);
  const int64_t npairs = mesh->n_elem * mesh->nn;
  const int nd = mesh->nn * dpn;
  if ((rc = dalloc(ctx, &cnt, std::max<int64_t>(npairs, 1))) || (rc = dalloc(ctx, &off, npairs + 1))) return done(rc);
  if ((rc = coo_launch<C_COUNT>(ctx, &pl, P, cnt, nullptr, nullptr, nullptr, nullptr, nullptr))) return done(rc);
  int64_t ntrip = 0;
  if ((rc = exclusive_scan_i32(ctx, cnt, off, npairs, &ntrip))) return done(rc);
  if (!want) {
    *count = ntrip;
    return done(TFM_OK);
  }
  if (*count < ntrip) return done(fail(ctx, TFM_E_INVALID, "buffers hold fewer entries than the triplets"));
  *count = ntrip;
  if (ntrip == 0) return done(TFM_OK);
  if ((rc = dalloc(ctx, &keys, ntrip)) || (rc = dalloc(ctx, &idx, ntrip)) ||
      (rc = dalloc(ctx, &dense, std::max<int64_t>(mesh->n_elem * nd * nd, 1))) || (rc = dalloc(ctx, &d_rows, ntrip)) ||
      (rc = dalloc(ctx, &d_cols, ntrip)) || (rc = dalloc(ctx, &d_vals, ntrip)))
    return done(rc);
  if ((rc = coo_launch<C_FILL>(ctx, &pl, P, nullptr, off, keys, idx, nullptr, nullptr)) ||
      (rc = coo_launch<C_F64>(ctx, &pl, P, nullptr, nullptr, nullptr, nullptr, dense, nullptr)))
    return done(rc);
  hipLaunchKernelGGL(k_coo_export, dim3((unsigned)std::min<int64_t>((ntrip + 255) / 256, 65536)), dim3(256), 0, st,
                     ntrip, (const unsigned long long *)keys, (const int64_t *)idx, (const double *)dense, d_rows,
                     d_cols, d_vals);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  if (hipMemcpyAsync(rows, d_rows, sizeof(int64_t) * ntrip, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(cols, d_cols, sizeof(int64_t) * ntrip, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(vals, d_vals, sizeof(double) * ntrip, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return done(fail(ctx, TFM_E_HIP, "triplet download failed"));
  return done(TFM_OK);
}

