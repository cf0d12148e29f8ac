// tfm_solve.hip -- CSR SpMV and Jacobi-preconditioned CG on the GPU.
//
// The reference solves the reduced system by densifying it and calling a
// partial-pivot LU (crates/tpt-fem-sparse/src/lib.rs:244-278); that cannot
// scale past a few thousand DOFs, so this library solves the same reduced
// system with Jacobi PCG:
//   * plain form: FP64 matrix, vectors and reductions;
//   * precision form: FP32 matrix, vectors and SpMV accumulation, FP64 dot
//     products, wrapped in FP64 iterative refinement (r = f - K u with the FP64
//     matrix, u += d) until the FP64 residual meets the tolerance.
// Iterations run in batches captured in a hipGraph.  Reductions need no
// fences or atomics: each block writes its partial sum, and every block of
// the next kernel re-sums the partials in the same fixed order, so all blocks
// agree on alpha, beta and convergence and results are deterministic.
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "tfm_bsell.h"
#include "tfm_elem.h"
#include "tfm_internal.h"
#include "tfm_pcg.h"

namespace tfm {

constexpr double kFloor2 = 1e-12;  // precision form: FP32 residual floor (squared, relative to the last replacement)

// Wave-level CSR stream: each wavefront independently takes R consecutive
// rows at a time, loads their whole (contiguous) non-zero range with
// coalesced reads into its own LDS slice as products, then 64/R lanes per row
// add the row's products in a fixed order.  All loads of a chunk are
// independent and waves never wait for each other, so many loads stay in
// flight per CU.
constexpr int kWaveCap = 512;  // products per wave chunk
template <int R>
struct Stream {
  static constexpr int T = 64 / R;  // lanes per row
};

// Orders a wave's LDS writes before its LDS reads (no block barrier: waves
// of a block run different chunk counts).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Iteration i (parity e): p_i = z_i + beta_i p_{i-1} on the fly, q = K p_i,
// partial p_i.q.  beta_i = rz_i / rz_{i-1}; rz_i, rr_i come from the previous
// update kernel (slot parity 1-e), rz_{i-1} from the one before (parity e).
// Convergence of r_i is decided here, identically in every block.
template <typename VT, int R>
__global__ __launch_bounds__(kBlock) void k_pcg_dir_spmv(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                          const VT *__restrict__ v, const VT *__restrict__ z,
                                                          const VT *__restrict__ pold, VT *__restrict__ pnew,
                                                          VT *__restrict__ q, Scal *sc, double *part, int e) {
  extern __shared__ unsigned char dyn_lds[];
  VT *prod = (VT *)dyn_lds;
  constexpr int T = Stream<R>::T;
  double beta;
  if (!dir_prologue(sc, part, e, &beta)) return;  // uniform per block
  const VT b = (VT)beta;
  const int lane = threadIdx.x & 63, lr = lane / T, sub = lane % T;
  VT *wprod = prod + (threadIdx.x >> 6) * kWaveCap;
  const int nchunks = (n + R - 1) / R;
  const int nwaves = gridDim.x * (kBlock / 64);
  double dot = 0.0;
  for (int ch = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); ch < nchunks; ch += nwaves) {
    const int r0 = ch * R, r1 = min(n, r0 + R);
    const int j0 = rp[r0], j1 = rp[r1];
    for (int j = j0 + lane; j < j1; j += 64) {
      const int c = ci[j];
      wprod[j - j0] = v[j] * (z[c] + b * pold[c]);
    }
    wave_lds_sync();
    const int row = r0 + lr;
    VT acc = 0;
    if (row < r1)
      for (int j = rp[row] + sub; j < rp[row + 1]; j += T) acc += wprod[j - j0];
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, T);
    if (row < r1 && sub == 0) {
      const VT pn = z[row] + b * pold[row];
      pnew[row] = pn;
      q[row] = acc;
      dot += (double)pn * (double)acc;
    }
    wave_lds_sync();
  }
  put_part(part, P_PQ, dot);
}

// The same iteration step on the node-blocked sliced ELL copy (BSell): one
// lane per block row (node), 64 block rows per slice, each wave a contiguous
// run of slices and each XCD a contiguous run of waves (neighbouring nodes'
// vector entries stay in one L2).  Per block: one column index, BS*BS
// coalesced values, BS gathered entries of z and p_{i-1}.  Padding blocks
// carry zeros against column 0, so the loop has no branch.
template <typename VT, int BS>
__global__ __launch_bounds__(kBlock) void k_pcg_dir_bsell(int n, int nslices, const int *__restrict__ soff,
                                                           const int *__restrict__ bcol, const VT *__restrict__ bv,
                                                           const VT *__restrict__ z, const VT *__restrict__ pold,
                                                           VT *__restrict__ pnew, VT *__restrict__ q, Scal *sc,
                                                           double *part, int e) {
  double beta;
  if (!dir_prologue(sc, part, e, &beta)) return;  // uniform per block
  const VT b = (VT)beta;
  constexpr int BB = BS * BS;
  const int lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * (kBlock / 64);
  const int gw = xcd_item(blockIdx.x, gridDim.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int per = (nslices + nwaves - 1) / nwaves;
  const int s1 = min(nslices, (gw + 1) * per);
  double dot = 0.0;
  for (int s = gw * per; s < s1; s++) {
    const int o = soff[s], w = (soff[s + 1] - o) >> 6;
    const int *cp = bcol + o + lane;
    const VT *vp = bv + (int64_t)o * BB + lane;
    VT acc[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) acc[i] = 0;
#pragma unroll 2
    for (int k = 0; k < w; k++) {
      const int c = cp[k * 64] * BS;
      VT xv[BS];
#pragma unroll
      for (int j = 0; j < BS; j++) xv[j] = z[c + j] + b * pold[c + j];
      const VT *vk = vp + (int64_t)k * (64 * BB);
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int j = 0; j < BS; j++) acc[i] += vk[(i * BS + j) * 64] * xv[j];
    }
    const int r0 = (s * 64 + lane) * BS;
    if (r0 < n) {
#pragma unroll
      for (int i = 0; i < BS; i++) {
        const VT pn = z[r0 + i] + b * pold[r0 + i];
        pnew[r0 + i] = pn;
        q[r0 + i] = acc[i];
        dot += (double)pn * (double)acc[i];
      }
    }
  }
  put_part(part, P_PQ, dot);
}

// alpha = rz_i / p.q; x += alpha p, r -= alpha q, z = Dinv r; partial rz, rr.
template <typename VT, typename XT>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, XT *__restrict__ x, VT *__restrict__ r,
                                                        VT *__restrict__ z, const VT *__restrict__ p,
                                                        const VT *__restrict__ q, const VT *__restrict__ dinv,
                                                        Scal *sc, double *part, int e) {
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = sc->done;
  __syncthreads();
  if (s_done) return;  // uniform per block (barriers follow)
  const int nb = gridDim.x;
  const double pq = sum_part(part, P_PQ, nb);
  if (!(pq > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->breakdown = 1;
      sc->done = 1;
    }
    return;
  }
  const double a = sum_part(part, e ? P_RZ0 : P_RZ1, nb) / pq;
  const VT af = (VT)a;
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += (XT)(a * (double)p[i]);
    const VT ri = r[i] - af * q[i];
    r[i] = ri;
    const VT zi = dinv[i] * ri;
    z[i] = zi;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  put_part(part, e ? P_RZ1 : P_RZ0, rz);
  put_part(part, e ? P_RR1 : P_RR0, rr);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->it = sc->it + 1;
}

// z = Dinv r, p buffers zeroed, partial rz / rr into the parity-1 slots
// (they play "iteration -1").
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const VT *__restrict__ r, VT *__restrict__ z,
                                                      VT *__restrict__ p0, VT *__restrict__ p1,
                                                      const VT *__restrict__ dinv, double *part) {
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const VT zi = dinv[i] * r[i];
    z[i] = zi;
    p0[i] = 0;
    p1[i] = 0;
    rz += (double)r[i] * (double)zi;
    rr += (double)r[i] * (double)r[i];
  }
  put_part(part, P_RZ1, rz);
  put_part(part, P_RR1, rr);
}

// One block: scalars after init / after a residual evaluation.
__global__ void k_pcg_start(Scal *sc, const double *part, int nb, double rtol2, int max_it) {
  double rr0 = sum_part(part, P_RR1, nb);
  if (threadIdx.x == 0) {
    sc->rr0 = rr0;
    sc->rr = rr0;
    sc->tol2 = rtol2 * rr0;
    sc->it = 0;
    sc->max_it = max_it;
    sc->breakdown = 0;
    sc->done = (rr0 == 0.0 || rr0 <= sc->tol2) ? 1 : 0;
  }
}

// r = f - K u (FP64, CSR-stream), optional FP32 copy; partial ||r||^2, ||f||^2.
template <int R>
__global__ __launch_bounds__(kBlock) void k_residual64(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                        const double *__restrict__ v, const double *__restrict__ u,
                                                        const double *__restrict__ f, double *__restrict__ r,
                                                        float *__restrict__ r32, double *part) {
  extern __shared__ unsigned char dyn_lds[];
  double *prod = (double *)dyn_lds;
  constexpr int T = Stream<R>::T;
  const int lane = threadIdx.x & 63, lr = lane / T, sub = lane % T;
  double *wprod = prod + (threadIdx.x >> 6) * kWaveCap;
  const int nchunks = (n + R - 1) / R;
  const int nwaves = gridDim.x * (kBlock / 64);
  double rr = 0.0, ff = 0.0;
  for (int ch = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); ch < nchunks; ch += nwaves) {
    const int r0 = ch * R, r1 = min(n, r0 + R);
    const int j0 = rp[r0], j1 = rp[r1];
    for (int j = j0 + lane; j < j1; j += 64) wprod[j - j0] = v[j] * u[ci[j]];
    wave_lds_sync();
    const int row = r0 + lr;
    double acc = 0.0;
    if (row < r1)
      for (int j = rp[row] + sub; j < rp[row + 1]; j += T) acc += wprod[j - j0];
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, T);
    if (row < r1 && sub == 0) {
      const double ri = (rp[row] == rp[row + 1]) ? 0.0 : f[row] - acc;
      r[row] = ri;
      if (r32) r32[row] = (float)ri;
      rr += ri * ri;
      ff += f[row] * f[row];
    }
    wave_lds_sync();
  }
  put_part(part, P_RES, rr);
  put_part(part, P_FF, ff);
}

// r = f - K u on the node-blocked sliced copy (the iteration's traversal);
// rows of fixed DOFs (empty CSR rows) give 0; optional FP32 copy; partial
// ||r||^2, ||f||^2.
template <int BS>
__global__ __launch_bounds__(kBlock) void k_resid_bsell(int n, int nslices, const int *__restrict__ soff,
                                                         const int *__restrict__ bcol, const double *__restrict__ bv,
                                                         const int *__restrict__ rp, const double *__restrict__ u,
                                                         const double *__restrict__ f, double *__restrict__ r,
                                                         float *__restrict__ r32, double *part) {
  constexpr int BB = BS * BS;
  const int lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * (kBlock / 64);
  const int gw = xcd_item(blockIdx.x, gridDim.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int per = (nslices + nwaves - 1) / nwaves;
  const int s1 = min(nslices, (gw + 1) * per);
  double rr = 0.0, ff = 0.0;
  for (int s = gw * per; s < s1; s++) {
    const int o = soff[s], w = (soff[s + 1] - o) >> 6;
    const int *cp = bcol + o + lane;
    const double *vp = bv + (int64_t)o * BB + lane;
    double acc[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) acc[i] = 0.0;
#pragma unroll 2
    for (int k = 0; k < w; k++) {
      const int c = cp[k * 64] * BS;
      double xv[BS];
#pragma unroll
      for (int j = 0; j < BS; j++) xv[j] = u[c + j];
      const double *vk = vp + (int64_t)k * (64 * BB);
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int j = 0; j < BS; j++) acc[i] += vk[(i * BS + j) * 64] * xv[j];
    }
    const int r0 = (s * 64 + lane) * BS;
    if (r0 < n) {
#pragma unroll
      for (int i = 0; i < BS; i++) {
        const int row = r0 + i;
        const double ri = (rp[row] == rp[row + 1]) ? 0.0 : f[row] - acc[i];
        r[row] = ri;
        if (r32) r32[row] = (float)ri;
        rr += ri * ri;
        ff += f[row] * f[row];
      }
    }
  }
  put_part(part, P_RES, rr);
  put_part(part, P_FF, ff);
}

__global__ void k_residual_finish(Scal *sc, const double *part, int nb) {
  double rr = sum_part(part, P_RES, nb), ff = sum_part(part, P_FF, nb);
  if (threadIdx.x == 0) {
    sc->rr = rr;
    sc->rr0 = ff;
  }
}

// Jacobi: dinv[i] = 1 / K[i][i] on non-empty rows; 0 on empty (fixed) rows.
// A free row without a positive diagonal is flagged (singular / not SPD).
template <typename VT>
__global__ void k_diag_inv(int n, const int *__restrict__ rp, const int *__restrict__ ci, const VT *__restrict__ v,
                           const unsigned char *__restrict__ fixed, VT *__restrict__ dinv, int *bad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int rb = rp[i], re = rp[i + 1];
  VT d = 0;
  int lo = rb, hi = re;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ci[mid] < i) lo = mid + 1; else hi = mid;
  }
  if (lo < re && ci[lo] == i) d = v[lo];
  bool isfixed = fixed && fixed[i];
  if (isfixed || rb == re) {
    dinv[i] = 0;
    if (!isfixed) atomicExch(bad, 1);
  } else if (!(d > (VT)0)) {
    dinv[i] = 0;
    atomicExch(bad, 1);
  } else {
    dinv[i] = (VT)1 / d;
  }
}

// FP32 copy of an FP64 vector (the matrix-free operator's Jacobi weights).
__global__ void k_round_f32(int n, const double *__restrict__ a, float *__restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_add_f32_to_f64(int n, double *u, const float *d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] += (double)d[i];
}

// Csr::matvec in the reference's order (sparse/src/lib.rs:195-205): one lane
// per row, left-to-right sum from 0.0, no fused multiply-add.
__global__ void k_matvec_exact(int64_t nrows, const int *__restrict__ rp, const int *__restrict__ ci,
                               const double *__restrict__ v, const double *__restrict__ x, double *__restrict__ y) {
#pragma clang fp contract(off)
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double s = 0.0;
  for (int j = rp[r]; j < rp[r + 1]; j++) s += v[j] * x[ci[j]];
  y[r] = s;
}

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

static int nblocks_for(int64_t work, int per_block) {
  int64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (int)b;
}

// Lanes per row for the CSR SpMV from the average row length.
static int lanes_for(double avg) {
  if (avg <= 6) return 2;
  if (avg <= 12) return 4;
  if (avg <= 24) return 8;
  if (avg <= 48) return 16;
  return 32;
}

template <typename VT>
struct PcgBufs {
  VT *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr, *p2 = nullptr;
};

// Rows per wave chunk from the longest row: a chunk's products must fit the
// wave's LDS slice (kWaveCap entries); a block holds kBlock/64 slices.
constexpr int kStreamCap = kWaveCap * (kBlock / 64);
static int rows_per_chunk(int max_row_nnz) {
  int R = 32;
  while (R > 1 && R * max_row_nnz > kWaveCap) R >>= 1;
  return R;
}
// Blocks for n rows in chunks of R rows, one chunk per wave at a time.
static int stream_blocks(int n, int R) {
  int chunks = (n + R - 1) / R;
  int b = (chunks + (kBlock / 64) - 1) / (kBlock / 64);
  return std::max(1, std::min(kMaxBlocks, b));
}

template <typename VT, int R>
static void launch_dir(hipStream_t st, int nb, int n, const int *rp, const int *ci, const VT *v, const VT *z,
                       const VT *pold, VT *pnew, VT *q, Scal *sc, double *part, int e) {
  hipLaunchKernelGGL((k_pcg_dir_spmv<VT, R>), dim3(nb), dim3(kBlock), sizeof(VT) * kStreamCap, st, n, rp, ci, v, z,
                     pold, pnew, q, sc, part, e);
}

template <typename VT>
static void dir_dispatch(int R, hipStream_t st, int nb, int n, const int *rp, const int *ci, const VT *v, const VT *z,
                         const VT *pold, VT *pnew, VT *q, Scal *sc, double *part, int e) {
  switch (R) {
  case 32: launch_dir<VT, 32>(st, nb, n, rp, ci, v, z, pold, pnew, q, sc, part, e); break;
  case 16: launch_dir<VT, 16>(st, nb, n, rp, ci, v, z, pold, pnew, q, sc, part, e); break;
  case 8: launch_dir<VT, 8>(st, nb, n, rp, ci, v, z, pold, pnew, q, sc, part, e); break;
  case 4: launch_dir<VT, 4>(st, nb, n, rp, ci, v, z, pold, pnew, q, sc, part, e); break;
  case 2: launch_dir<VT, 2>(st, nb, n, rp, ci, v, z, pold, pnew, q, sc, part, e); break;
  default: launch_dir<VT, 1>(st, nb, n, rp, ci, v, z, pold, pnew, q, sc, part, e); break;
  }
}

// Residual r = f - K u into r (and r32), then sc->rr = ||r||^2, sc->rr0 = ||f||^2.
static void residual_dispatch(int R, hipStream_t st, int nb, int n, const int *rp, const int *ci, const double *v,
                              const double *u, const double *f, double *r, float *r32, Scal *sc, double *part) {
  const size_t sh = sizeof(double) * kStreamCap;
#define TFM_RES(RR) hipLaunchKernelGGL((k_residual64<RR>), dim3(nb), dim3(kBlock), sh, st, n, rp, ci, v, u, f, r, r32, part)
  switch (R) {
  case 32: TFM_RES(32); break;
  case 16: TFM_RES(16); break;
  case 8: TFM_RES(8); break;
  case 4: TFM_RES(4); break;
  case 2: TFM_RES(2); break;
  default: TFM_RES(1); break;
  }
#undef TFM_RES
  hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(kBlock), 0, st, sc, part, nb);
}

// The matrix operand of the PCG iteration: the CSR stream form, or the
// node-blocked sliced copy when one was built (bs > 0).  nb is the grid of
// every kernel of the iteration (their partial sums are indexed by block).
template <typename VT>
struct SpOp {
  int n = 0, nb = 1;
  const int *rp = nullptr, *ci = nullptr;
  const VT *v = nullptr;
  int R = 1;
  int bs = 0, nslices = 0;
  const int *soff = nullptr, *bcol = nullptr;
  const VT *bv = nullptr;
  bool mf = false;  // FP32 only: the matrix-free box-mesh operator
  MfBox box;
};

template <typename VT>
static SpOp<VT> csr_op(int n, const int *rp, const int *ci, const VT *v, int R) {
  SpOp<VT> o;
  o.n = n;
  o.rp = rp;
  o.ci = ci;
  o.v = v;
  o.R = R;
  o.nb = stream_blocks(n, R);
  return o;
}

template <typename VT>
static SpOp<VT> bsell_op(int n, const BSell &B, const VT *bv) {
  SpOp<VT> o;
  o.n = n;
  o.bs = B.bs;
  o.nslices = B.nslices;
  o.soff = B.d_soff;
  o.bcol = B.d_bcol;
  o.bv = bv;
  // four waves per block; a multiple of 8 blocks so that the XCD mapping is
  // a permutation of the grid
  int nb = (B.nslices + 3) / 4;
  nb = std::min(kMaxBlocks, std::max(8, (nb + 7) / 8 * 8));
  o.nb = nb;
  return o;
}

template <typename VT>
static void dir_launch(const SpOp<VT> &A, hipStream_t st, const VT *z, const VT *pold, VT *pnew, VT *q, Scal *sc,
                       double *part, int e) {
  if constexpr (sizeof(VT) == 4) {
    if (A.mf) {
      mf_dir_launch(A.box, st, A.nb, z, pold, pnew, q, sc, part, e);
      return;
    }
  }
  if (A.bs == 0) {
    dir_dispatch<VT>(A.R, st, A.nb, A.n, A.rp, A.ci, A.v, z, pold, pnew, q, sc, part, e);
    return;
  }
#define TFM_BS(BSZ)                                                                                          \
  hipLaunchKernelGGL((k_pcg_dir_bsell<VT, BSZ>), dim3(A.nb), dim3(kBlock), 0, st, A.n, A.nslices, A.soff, A.bcol, \
                     A.bv, z, pold, pnew, q, sc, part, e)
  switch (A.bs) {
  case 3: TFM_BS(3); break;
  case 2: TFM_BS(2); break;
  default: TFM_BS(1); break;
  }
#undef TFM_BS
}

__global__ __launch_bounds__(256) void k_max_row_nnz(int n, const int *rp, int *out) {
  int m = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = max(m, rp[i + 1] - rp[i]);
  block_max(m, out);
}

// Runs PCG on K x = r0 (x accumulates; r holds r0 on entry) until
// ||r|| <= rtol ||r0|| or max_iter.  Two kernels per iteration (direction +
// SpMV + p.q, then the update + r.z, r.r); 16 iterations per graph replay.
template <typename VT, typename XT>
static int pcg_run(tfm_ctx *ctx, const SpOp<VT> &A, PcgBufs<VT> &B, XT *x, double rtol, int max_iter, Scal *sc,
                   double *part, int *iters, int *breakdown, double *rr_out, bool allow_graph = true) {
  hipStream_t st = ctx->stream;
  const int n = A.n, nb = A.nb;
  hipLaunchKernelGGL((k_pcg_init<VT>), dim3(nb), dim3(kBlock), 0, st, n, B.r, B.z, B.p, B.p2, B.dinv, part);
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(kBlock), 0, st, sc, part, nb, rtol * rtol, max_iter);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  const int CH = 16;  // iterations per replay (even: buffer parity is static)
  VT *pb[2] = {B.p, B.p2};
  auto chunk = [&]() {
    for (int k = 0; k < CH; k++) {
      const int e = k & 1;
      dir_launch<VT>(A, st, B.z, pb[1 - e], pb[e], B.q, sc, part, e);
      hipLaunchKernelGGL((k_pcg_update<VT, XT>), dim3(nb), dim3(kBlock), 0, st, n, x, B.r, B.z, pb[e], B.q, B.dinv,
                         sc, part, e);
    }
  };
  // TFM_NO_GRAPH=1 launches the iterations directly (for tools that do not
  // follow graph launches); results are identical.
  const bool use_graph = allow_graph && !getenv("TFM_NO_GRAPH");
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    TFM_CHECK_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    chunk();
    TFM_CHECK_HIP(ctx, hipStreamEndCapture(st, &graph));
    TFM_CHECK_HIP(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  Scal h{};
  int rc = TFM_OK;
  for (;;) {
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
    TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
    if (h.done) break;
    if (!use_graph) {
      chunk();
    } else if (hipGraphLaunch(exec, st) != hipSuccess) {
      rc = fail(ctx, TFM_E_HIP, "hipGraphLaunch failed");
      break;
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  *iters = h.it;
  *breakdown = h.breakdown;
  *rr_out = h.rr0 > 0 ? std::sqrt(h.rr / h.rr0) : 0.0;
  return rc;
}

// After an FP64 residual replacement (r already holds the new residual):
// z = Dinv r and the r.z / r.r partials of the coming iteration (parity-1
// slots); the search directions are kept.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_replace(int n, const VT *__restrict__ r, VT *__restrict__ z,
                                                         const VT *__restrict__ dinv, double *part) {
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const VT zi = dinv[i] * r[i];
    z[i] = zi;
    rz += (double)r[i] * (double)zi;
    rr += (double)r[i] * (double)r[i];
  }
  put_part(part, P_RZ1, rz);
  put_part(part, P_RR1, rr);
}

// After a residual replacement: the FP32 iteration runs on, freezing itself
// only when its residual falls by sqrt(d2) below the replaced one (sc->rr).
__global__ void k_pcg_rearm(Scal *sc, double d2) {
  sc->tol2 = d2 * sc->rr;
  sc->breakdown = 0;
  sc->done = 0;
}

// x += d (FP64), d = 0.
__global__ void k_fold(int n, double *__restrict__ x, float *__restrict__ d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] += (double)d[i];
    d[i] = 0.0f;
  }
}

// Residual r = f - K u through the operator: the sliced copy when it holds
// the FP64 values, else the CSR stream; then sc->rr, sc->rr0.
static void residual_op(const SpOp<double> &K, hipStream_t st, const double *u, const double *f, double *r,
                        float *r32, Scal *sc, double *part) {
  if (K.bs == 0) {
    residual_dispatch(K.R, st, K.nb, K.n, K.rp, K.ci, K.v, u, f, r, r32, sc, part);
    return;
  }
#define TFM_BS(BSZ)                                                                                             \
  hipLaunchKernelGGL((k_resid_bsell<BSZ>), dim3(K.nb), dim3(kBlock), 0, st, K.n, K.nslices, K.soff, K.bcol, K.bv, \
                     K.rp, u, f, r, r32, part)
  switch (K.bs) {
  case 3: TFM_BS(3); break;
  case 2: TFM_BS(2); break;
  default: TFM_BS(1); break;
  }
#undef TFM_BS
  hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(kBlock), 0, st, sc, part, K.nb);
}

// Precision form: FP32 PCG on the correction with reliable updates.  The
// FP32 iteration runs on; whenever its recurrence residual has dropped by
// delta since the last replacement, the FP32 correction is folded into the
// FP64 solution x and the residual is replaced by the true FP64 r = f - K x
// (FP64 matrix), keeping the search direction.  Stops when the FP64 residual
// meets rtol.  Keeping the Krylov direction (instead of restarting per
// refinement cycle) keeps the iteration count near plain FP64 PCG's.
static int mixed_run(tfm_ctx *ctx, const SpOp<double> &K64, const SpOp<float> &A, const double *f,
                     PcgBufs<float> &B, double *r64, float *x32, double *x, double rtol, double delta, int max_iter,
                     Scal *sc, double *part, int *iters, int *repl, int *breakdown, double *relres) {
  hipStream_t st = ctx->stream;
  const int n = A.n, nb = A.nb;
  const unsigned g1 = (unsigned)((n + 255) / 256);
  Scal h{};
  *iters = 0;
  *repl = 0;
  *breakdown = 0;
  residual_op(K64, st, x, f, r64, B.r, sc, part);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  *relres = h.rr0 > 0 ? std::sqrt(h.rr / h.rr0) : 0.0;
  if (*relres <= rtol) return TFM_OK;
  double rr_ref = h.rr;
  const double target2 = rtol * rtol * h.rr0;  // ||f||^2 rtol^2
  TFM_CHECK_HIP(ctx, hipMemsetAsync(x32, 0, sizeof(float) * n, st));
  hipLaunchKernelGGL((k_pcg_init<float>), dim3(nb), dim3(kBlock), 0, st, n, B.r, B.z, B.p, B.p2, B.dinv, part);
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(kBlock), 0, st, sc, part, nb, kFloor2, max_iter);
  const int CH = 16;
  float *pb[2] = {B.p, B.p2};
  auto chunk = [&]() {
    for (int k = 0; k < CH; k++) {
      const int e = k & 1;
      dir_launch<float>(A, st, B.z, pb[1 - e], pb[e], B.q, sc, part, e);
      hipLaunchKernelGGL((k_pcg_update<float, float>), dim3(nb), dim3(kBlock), 0, st, n, x32, B.r, B.z, pb[e], B.q,
                         B.dinv, sc, part, e);
    }
  };
  const bool use_graph = !getenv("TFM_NO_GRAPH");
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    TFM_CHECK_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    chunk();
    TFM_CHECK_HIP(ctx, hipStreamEndCapture(st, &graph));
    TFM_CHECK_HIP(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  int rc = TFM_OK;
  constexpr int kStall = 64;  // chunks of 16 iterations
  double best = rr_ref;
  int since = 0, nbrk = 0;
  for (;;) {
    if (!use_graph) {
      chunk();
    } else if (hipGraphLaunch(exec, st) != hipSuccess) {
      rc = fail(ctx, TFM_E_HIP, "hipGraphLaunch failed");
      break;
    }
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
    TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
    *iters = h.it;
    // Replacements keeping the direction happen at chunk boundaries, when the
    // residual has dropped by delta (the buffer and partial-slot parity is
    // then consistent).  Inside a chunk the FP32 iteration freezes itself
    // (done) only at its accuracy floor -- 1e-6 below the last replacement --
    // or on a breakdown (p.q <= 0), and then restarts its direction.  The
    // stagnation guard also replaces when the smallest recurrence residual
    // has not improved by 10% for kStall chunks.  A recurrence residual below
    // the final target triggers the check at once, so the last cycle stops
    // near rtol rather than at the next multiple of delta.
    if (h.rr < 0.81 * best) {
      best = h.rr;
      since = 0;
    } else {
      since++;
    }
    if (h.rr <= delta * delta * rr_ref || h.rr <= target2 || h.done || since >= kStall) {
      hipLaunchKernelGGL(k_fold, dim3(g1), dim3(256), 0, st, n, x, x32);
      residual_op(K64, st, x, f, r64, B.r, sc, part);
      Scal h2{};
      TFM_CHECK_HIP(ctx, hipMemcpyAsync(&h2, sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
      TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
      *repl += 1;
      *relres = h2.rr0 > 0 ? std::sqrt(h2.rr / h2.rr0) : 0.0;
      if (*relres <= rtol || h.it >= max_iter) break;
      if (h.breakdown && ++nbrk > 4) {  // a breakdown that keeps coming back fails the solve
        *breakdown = 1;
        break;
      }
      hipLaunchKernelGGL((k_pcg_replace<float>), dim3(nb), dim3(kBlock), 0, st, n, B.r, B.z, B.dinv, part);
      if (h.done) {  // frozen inside the chunk: restart the direction
        TFM_CHECK_HIP(ctx, hipMemsetAsync(B.p, 0, sizeof(float) * n, st));
        TFM_CHECK_HIP(ctx, hipMemsetAsync(B.p2, 0, sizeof(float) * n, st));
      }
      hipLaunchKernelGGL(k_pcg_rearm, dim3(1), dim3(1), 0, st, sc, kFloor2);
      rr_ref = h2.rr;
      best = rr_ref;
      since = 0;
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return rc;
}

// Builds the node-blocked sliced copy's structure from the system's CSR
// (once per system; the pattern never changes).  Leaves state -1 when the
// value index would not fit int32.
static int bsell_build(tfm_system *s) {
  BSell &B = s->bsell;
  if (B.state != 0) return TFM_OK;
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int bs = (s->dpn >= 1 && s->dpn <= 3 && s->ndof % s->dpn == 0) ? s->dpn : 1;
  const int64_t nbr = s->ndof / bs;
  const int64_t nsl = (nbr + 63) / 64;
  if (nbr == 0 || s->nnz == 0 || nsl * 64 >= INT32_MAX / 16) {
    B.state = -1;
    return TFM_OK;
  }
  int *cnt = nullptr, *w = nullptr;
  int rc;
  if ((rc = dalloc(ctx, &cnt, nbr)) || (rc = dalloc(ctx, &w, nsl)) || (rc = dalloc(ctx, &B.d_soff, nsl + 1))) {
    dfree(cnt);
    dfree(w);
    bsell_destroy(&B);
    return rc;
  }
  const unsigned g = (unsigned)((nbr + 255) / 256);
  switch (bs) {
  case 3: hipLaunchKernelGGL(k_bs_count<3>, dim3(g), dim3(256), 0, st, (int)nbr, s->d_rp, s->d_ci, cnt); break;
  case 2: hipLaunchKernelGGL(k_bs_count<2>, dim3(g), dim3(256), 0, st, (int)nbr, s->d_rp, s->d_ci, cnt); break;
  default: hipLaunchKernelGGL(k_bs_count<1>, dim3(g), dim3(256), 0, st, (int)nbr, s->d_rp, s->d_ci, cnt); break;
  }
  hipLaunchKernelGGL(k_bs_width, dim3((unsigned)((nsl + 255) / 256)), dim3(256), 0, st, (int)nbr, (int)nsl,
                     (const int *)nullptr, cnt, w);
  int64_t nblk = 0;
  rc = exclusive_scan_i32(ctx, w, B.d_soff, nsl, &nblk);
  dfree(cnt);
  dfree(w);
  if (rc) {
    bsell_destroy(&B);
    return rc;
  }
  if (nblk * bs * bs >= INT32_MAX) {
    bsell_destroy(&B);
    B.state = -1;
    return TFM_OK;
  }
  if ((rc = dalloc(ctx, &B.d_bcol, nblk)) || (rc = dalloc(ctx, &B.d_map, s->nnz))) {
    bsell_destroy(&B);
    return rc;
  }
  const unsigned g2 = (unsigned)((nsl * 64 + 255) / 256);
  switch (bs) {
  case 3:
    hipLaunchKernelGGL(k_bs_fill<3>, dim3(g2), dim3(256), 0, st, (int)nbr, (int)nsl, (const int *)nullptr,
                       s->d_rp, s->d_ci, B.d_soff, B.d_bcol, B.d_map);
    break;
  case 2:
    hipLaunchKernelGGL(k_bs_fill<2>, dim3(g2), dim3(256), 0, st, (int)nbr, (int)nsl, (const int *)nullptr,
                       s->d_rp, s->d_ci, B.d_soff, B.d_bcol, B.d_map);
    break;
  default:
    hipLaunchKernelGGL(k_bs_fill<1>, dim3(g2), dim3(256), 0, st, (int)nbr, (int)nsl, (const int *)nullptr,
                       s->d_rp, s->d_ci, B.d_soff, B.d_bcol, B.d_map);
    break;
  }
  TFM_CHECK_HIP(ctx, hipGetLastError());
  if ((rc = bs_invert_map(ctx, &B.d_map, s->nnz, nblk * bs * bs))) {
    bsell_destroy(&B);
    return rc;
  }
  B.bs = bs;
  B.nbr = (int)nbr;
  B.nslices = (int)nsl;
  B.nblk = nblk;
  B.nnz = s->nnz;
  B.state = 1;
  return TFM_OK;
}

// Copies the system's current FP64 or FP32 values into the sliced copy
// (entries absent from the CSR stay zero).
template <typename VT>
static int bsell_values(tfm_system *s, const VT *src, VT **dst) {
  BSell &B = s->bsell;
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t nv = B.nblk * B.bs * B.bs;
  if (!*dst) {
    int rc = dalloc(ctx, dst, nv);
    if (rc) return rc;
  }
  hipLaunchKernelGGL((k_bs_values<VT, VT>), dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, nv, B.d_map, src,
                     *dst);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

void bsell_destroy(BSell *b) {
  dfree(b->d_soff);
  dfree(b->d_bcol);
  dfree(b->d_map);
  dfree(b->d_v64);
  dfree(b->d_v32);
  *b = BSell{};
}

}  // namespace tfm

using namespace tfm;

// ---------------------------------------------------------------------------
// Matrices
// ---------------------------------------------------------------------------
extern "C" int tfm_matrix_create(tfm_ctx *ctx, int64_t nrows, int64_t ncols, const int64_t *row_ptrs,
                                 const int64_t *col_ind, const double *values, tfm_matrix **out) {
  if (!ctx || !out || nrows < 0 || ncols < 0 || !row_ptrs) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  int64_t nnz = row_ptrs[nrows];
  if (nnz >= INT32_MAX || nrows >= INT32_MAX) return fail(ctx, TFM_E_UNSUPPORTED, "matrix too large for int32 indexing");
  std::vector<int> rp(nrows + 1), ci(nnz);
  for (int64_t i = 0; i <= nrows; i++) rp[i] = (int)row_ptrs[i];
  for (int64_t i = 0; i < nnz; i++) {
    if (col_ind[i] < 0 || col_ind[i] >= ncols) return fail(ctx, TFM_E_INVALID, "column index out of range");
    ci[i] = (int)col_ind[i];
  }
  tfm_matrix *m = new tfm_matrix();
  m->ctx = ctx;
  m->nrows = nrows;
  m->ncols = ncols;
  m->nnz = nnz;
  if ((rc = dalloc(ctx, &m->d_rp, nrows + 1)) || (rc = dalloc(ctx, &m->d_ci, nnz)) || (rc = dalloc(ctx, &m->d_v, nnz))) {
    tfm_matrix_destroy(m);
    return rc;
  }
  if (hipMemcpy(m->d_rp, rp.data(), sizeof(int) * (nrows + 1), hipMemcpyHostToDevice) != hipSuccess ||
      (nnz && hipMemcpy(m->d_ci, ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice) != hipSuccess) ||
      (nnz && hipMemcpy(m->d_v, values, sizeof(double) * nnz, hipMemcpyHostToDevice) != hipSuccess)) {
    tfm_matrix_destroy(m);
    return fail(ctx, TFM_E_HIP, "matrix upload failed");
  }
  *out = m;
  return TFM_OK;
}

extern "C" int tfm_assemble(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn, tfm_matrix **out) {
  if (!ctx || !mesh || !op || !out) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  AsmPlan pl;
  if ((rc = plan_create(ctx, mesh, op, dpn, nullptr, TFM_SYS_FP64, &pl))) {
    plan_destroy(&pl);
    return rc;
  }
  tfm_matrix *m = new tfm_matrix();
  m->ctx = ctx;
  m->nrows = pl.nrows;
  m->ncols = pl.ncols;
  m->nnz = pl.nnz;
  if ((rc = dalloc(ctx, &m->d_v, pl.nnz)) || (rc = plan_numeric(ctx, &pl, m->d_v, nullptr))) {
    plan_destroy(&pl);
    tfm_matrix_destroy(m);
    return rc;
  }
  // keep the pattern; the plan's row pointers cover all ndof rows, and rows
  // past nrows are empty, so the first nrows+1 entries are the CSR's.
  m->d_rp = pl.d_rp;
  m->d_ci = pl.d_ci;
  pl.d_rp = nullptr;
  pl.d_ci = nullptr;
  if (pl.nnz == 0) {
    // an empty assembly is the 1 x 1 empty CSR of Coo::to_csr
    TFM_CHECK_HIP(ctx, hipMemsetAsync(m->d_rp, 0, sizeof(int) * 2, ctx->stream));
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  plan_destroy(&pl);
  if (e != hipSuccess) {
    tfm_matrix_destroy(m);
    return fail(ctx, TFM_E_HIP, "assembly failed");
  }
  *out = m;
  return TFM_OK;
}

// elasticity_lumped_mass (crates/tpt-fem-elasticity/src/lib.rs:436-456): one
// thread per node sums its dim consistent-mass rows (row order, CSR order,
// from 0.0), divides by dim and writes the node's dim diagonal entries.
__global__ void k_lumped(int64_t nodes, int dim, const int *__restrict__ rp, const double *__restrict__ v,
                         int *__restrict__ orp, int *__restrict__ oci, double *__restrict__ ov) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nodes) return;
  double total = 0.0;
  for (int a = 0; a < dim; a++) {
    const int64_t row = node * dim + a;
    for (int c = rp[row]; c < rp[row + 1]; c++) total += v[c];
  }
  const double lump = total / (double)dim;
  for (int a = 0; a < dim; a++) {
    const int64_t i = node * dim + a;
    orp[i] = (int)i;
    oci[i] = (int)i;
    ov[i] = lump;
  }
  if (node == nodes - 1) orp[nodes * dim] = (int)(nodes * dim);
}

extern "C" int tfm_lumped_mass(tfm_ctx *ctx, const tfm_mesh *mesh, double density, int quad_order,
                               tfm_matrix **out) {
  if (!ctx || !mesh || !out) return TFM_E_INVALID;
  tfm_operator op{};
  op.kind = TFM_OP_MASS;
  op.quad_order = quad_order;
  op.coef = density;
  tfm_matrix *M = nullptr;
  int rc = tfm_assemble(ctx, mesh, &op, mesh->dim, &M);
  if (rc) return rc;
  const int dim = mesh->dim;
  const int64_t nodes = M->nrows / dim, n = nodes * dim;
  tfm_matrix *L = new tfm_matrix();
  L->ctx = ctx;
  L->nrows = n > 0 ? n : 1;  // Coo::to_csr of no triplets is 1 x 1
  L->ncols = L->nrows;
  L->nnz = n;
  if ((rc = dalloc(ctx, &L->d_rp, L->nrows + 1)) || (rc = dalloc(ctx, &L->d_ci, n)) ||
      (rc = dalloc(ctx, &L->d_v, n))) {
    tfm_matrix_destroy(M);
    tfm_matrix_destroy(L);
    return rc;
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_lumped, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, ctx->stream, nodes, dim,
                       (const int *)M->d_rp, (const double *)M->d_v, L->d_rp, L->d_ci, L->d_v);
  } else {
    TFM_CHECK_HIP(ctx, hipMemsetAsync(L->d_rp, 0, sizeof(int) * 2, ctx->stream));
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  tfm_matrix_destroy(M);
  if (e != hipSuccess) {
    tfm_matrix_destroy(L);
    return fail(ctx, TFM_E_HIP, "lumped-mass kernel failed");
  }
  *out = L;
  return TFM_OK;
}

extern "C" int tfm_matrix_info(const tfm_matrix *m, int64_t *nrows, int64_t *ncols, int64_t *nnz) {
  if (!m) return TFM_E_INVALID;
  if (nrows) *nrows = m->nrows;
  if (ncols) *ncols = m->ncols;
  if (nnz) *nnz = m->nnz;
  return TFM_OK;
}

extern "C" int tfm_matrix_download(const tfm_matrix *m, int64_t *row_ptrs, int64_t *col_ind, double *values) {
  if (!m) return TFM_E_INVALID;
  tfm_ctx *ctx = m->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  std::vector<int> rp(m->nrows + 1), ci(m->nnz);
  TFM_CHECK_HIP(ctx, hipMemcpy(rp.data(), m->d_rp, sizeof(int) * (m->nrows + 1), hipMemcpyDeviceToHost));
  if (m->nnz) TFM_CHECK_HIP(ctx, hipMemcpy(ci.data(), m->d_ci, sizeof(int) * m->nnz, hipMemcpyDeviceToHost));
  if (row_ptrs)
    for (int64_t i = 0; i <= m->nrows; i++) row_ptrs[i] = rp[i];
  if (col_ind)
    for (int64_t i = 0; i < m->nnz; i++) col_ind[i] = ci[i];
  if (values && m->nnz) TFM_CHECK_HIP(ctx, hipMemcpy(values, m->d_v, sizeof(double) * m->nnz, hipMemcpyDeviceToHost));
  return TFM_OK;
}

extern "C" int tfm_matrix_matvec(const tfm_matrix *m, const double *x, int64_t nx, double *y) {
  if (!m || (!x && nx) || !y) return TFM_E_INVALID;
  tfm_ctx *ctx = m->ctx;
  if (nx < m->ncols) return fail(ctx, TFM_E_INVALID, "x shorter than ncols");
  int rc = set_device(ctx);
  if (rc) return rc;
  int64_t ny = std::max(m->nrows, nx);
  double *dx = nullptr, *dy = nullptr;
  if ((rc = dalloc(ctx, &dx, nx)) || (rc = dalloc(ctx, &dy, ny))) return rc;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(dx, x, sizeof(double) * nx, hipMemcpyHostToDevice, ctx->stream));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(dy, 0, sizeof(double) * ny, ctx->stream));
  if (m->nrows > 0)
    hipLaunchKernelGGL(k_matvec_exact, dim3(grid1(m->nrows, 256)), dim3(256), 0, ctx->stream, m->nrows, m->d_rp,
                       m->d_ci, m->d_v, dx, dy);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(y, dy, sizeof(double) * ny, hipMemcpyDeviceToHost, ctx->stream));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dfree(dx);
  dfree(dy);
  return TFM_OK;
}

extern "C" void tfm_matrix_destroy(tfm_matrix *m) {
  if (!m) return;
  dfree(m->d_rp);
  dfree(m->d_ci);
  dfree(m->d_v);
  delete m;
}

// ---------------------------------------------------------------------------
// Reduced systems
// ---------------------------------------------------------------------------
static int collect_bcs(tfm_ctx *ctx, int64_t ndof, int64_t n_bc, const int64_t *bc_dof, const double *bc_val,
                       tfm_system *s, std::vector<unsigned char> &fixed, std::vector<double> &first) {
  fixed.assign((size_t)ndof, 0);
  first.assign((size_t)ndof, 0.0);
  s->bc_dof.assign(bc_dof, bc_dof + n_bc);
  s->bc_val.assign(bc_val, bc_val + n_bc);
  for (int64_t b = 0; b < n_bc; b++) {
    int64_t d = bc_dof[b];
    if (d < 0 || d >= ndof) return fail(ctx, TFM_E_INVALID, "Dirichlet DOF out of range");
    if (!fixed[d]) first[d] = bc_val[b];  // lifting uses the first match (assembly/src/lib.rs:525)
    fixed[d] = 1;
  }
  s->free_dofs.clear();
  for (int64_t i = 0; i < ndof; i++)
    if (!fixed[i]) s->free_dofs.push_back(i);
  s->n_free = (int64_t)s->free_dofs.size();
  return TFM_OK;
}

// fred = rhs - K_fc g over each free row, in ascending column order, exact.
__global__ void k_lift(int64_t ndof, const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
                       const unsigned char *__restrict__ fixed, const double *__restrict__ gval,
                       const double *__restrict__ rhs, double *__restrict__ f, int nrows) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndof) return;
  if (fixed[i]) { f[i] = 0.0; return; }
  double r = rhs[i];
  if (i < nrows)
    for (int j = rp[i]; j < rp[i + 1]; j++) {
      int c = ci[j];
      if (c < ndof && fixed[c]) r -= v[j] * gval[c];
    }
  f[i] = r;
}

// Row counts / fill of the reduced pattern of a full matrix.
__global__ void k_reduce_count(int64_t ndof, int nrows, const int *rp, const int *ci, const unsigned char *fixed,
                               int64_t ncol_lim, int *cnt, int *bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndof) return;
  int c = 0;
  if (!fixed[i] && i < nrows)
    for (int j = rp[i]; j < rp[i + 1]; j++) {
      int col = ci[j];
      if (col >= ncol_lim) { atomicExch(bad, 1); continue; }
      if (!fixed[col]) c++;
    }
  cnt[i] = c;
}

__global__ void k_reduce_fill(int64_t ndof, int nrows, const int *rp, const int *ci, const double *v,
                              const unsigned char *fixed, const int *rp2, int *ci2, double *v2) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndof || fixed[i] || i >= nrows) return;
  int o = rp2[i];
  for (int j = rp[i]; j < rp[i + 1]; j++) {
    int col = ci[j];
    if (col < ndof && !fixed[col]) {
      ci2[o] = col;
      v2[o] = v[j];
      o++;
    }
  }
}

extern "C" int tfm_system_from_matrix(tfm_ctx *ctx, const tfm_matrix *k, int64_t n, const double *rhs, int64_t n_bc,
                                      const int64_t *bc_dof, const double *bc_val, tfm_system **out) {
  if (!ctx || !k || !out || (n && !rhs) || (n_bc && (!bc_dof || !bc_val))) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_system *s = new tfm_system();
  s->ctx = ctx;
  s->dpn = 1;
  s->ndof = n;
  std::vector<unsigned char> fixed;
  std::vector<double> first;
  if ((rc = collect_bcs(ctx, n, n_bc, bc_dof, bc_val, s, fixed, first))) {
    delete s;
    return rc;
  }
  // reduce_system indexes row_ptrs[fdof + 1] for every free DOF: the matrix
  // must have a row for each (the reference panics otherwise).
  for (int64_t f : s->free_dofs)
    if (f >= k->nrows) {
      delete s;
      return fail(ctx, TFM_E_INVALID, "reduce_system: free DOF has no row in the assembled matrix");
    }
  double *d_g = nullptr, *d_rhs = nullptr;
  int *cnt = nullptr, *bad = nullptr;
  s->owns_pattern = true;
  if ((rc = dalloc(ctx, &s->d_fixed, n)) || (rc = dalloc(ctx, &d_g, n)) || (rc = dalloc(ctx, &d_rhs, n)) ||
      (rc = dalloc(ctx, &s->d_f, n)) || (rc = dalloc(ctx, &cnt, n + 1)) || (rc = dalloc(ctx, &bad, 1)) ||
      (rc = dalloc(ctx, &s->d_rp, n + 1))) {
    tfm_system_destroy(s);
    return rc;
  }
  hipStream_t st = ctx->stream;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(s->d_fixed, fixed.data(), n, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(d_g, first.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(d_rhs, rhs, sizeof(double) * n, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(bad, 0, sizeof(int), st));
  if (n > 0) {
    hipLaunchKernelGGL(k_reduce_count, dim3(grid1(n, 256)), dim3(256), 0, st, n, (int)k->nrows, k->d_rp, k->d_ci,
                       s->d_fixed, n, cnt, bad);
    hipLaunchKernelGGL(k_lift, dim3(grid1(n, 256)), dim3(256), 0, st, n, k->d_rp, k->d_ci, k->d_v, s->d_fixed, d_g,
                       d_rhs, s->d_f, (int)k->nrows);
  }
  int64_t nnz = 0;
  if ((rc = exclusive_scan_i32(ctx, cnt, s->d_rp, n, &nnz))) {
    tfm_system_destroy(s);
    return rc;
  }
  int hbad = 0;
  TFM_CHECK_HIP(ctx, hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost));
  if (hbad) {
    tfm_system_destroy(s);
    return fail(ctx, TFM_E_INVALID, "reduce_system: column beyond the rhs length");
  }
  s->nnz = nnz;
  s->nrows = n;
  if ((rc = dalloc(ctx, &s->d_ci, nnz)) || (rc = dalloc(ctx, &s->d_v64, nnz))) {
    tfm_system_destroy(s);
    return rc;
  }
  if (n > 0)
    hipLaunchKernelGGL(k_reduce_fill, dim3(grid1(n, 256)), dim3(256), 0, st, n, (int)k->nrows, k->d_rp, k->d_ci,
                       k->d_v, s->d_fixed, s->d_rp, s->d_ci, s->d_v64);
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  dfree(d_g);
  dfree(d_rhs);
  dfree(cnt);
  dfree(bad);
  *out = s;
  return TFM_OK;
}

extern "C" int tfm_system_create(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn, int64_t n_bc,
                                 const int64_t *bc_dof, const double *bc_val, const double *rhs, int flags,
                                 tfm_system **out) {
  if (!ctx || !mesh || !op || !out || (n_bc && (!bc_dof || !bc_val))) return TFM_E_INVALID;
  if (!(flags & (TFM_SYS_FP64 | TFM_SYS_FP32))) flags = TFM_SYS_FP64;
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_system *s = new tfm_system();
  s->ctx = ctx;
  s->dpn = dpn;
  s->ndof = mesh->n_nodes * dpn;
  std::vector<unsigned char> fixed;
  std::vector<double> first;
  if ((rc = collect_bcs(ctx, s->ndof, n_bc, bc_dof, bc_val, s, fixed, first))) {
    delete s;
    return rc;
  }
  bool homogeneous = true;
  for (int64_t b = 0; b < n_bc; b++) homogeneous &= (bc_val[b] == 0.0);
  if (!homogeneous) {
    // General lifting needs K[f, c] of fixed columns: assemble the full
    // matrix and reduce it (reduce_system's own order of operations).
    tfm_matrix *K = nullptr;
    delete s;
    if ((rc = tfm_assemble(ctx, mesh, op, dpn, &K))) return rc;
    std::vector<double> zero;
    if (!rhs) zero.assign((size_t)(mesh->n_nodes * dpn), 0.0);
    rc = tfm_system_from_matrix(ctx, K, mesh->n_nodes * dpn, rhs ? rhs : zero.data(), n_bc, bc_dof, bc_val, out);
    tfm_matrix_destroy(K);
    return rc;
  }
  s->has_plan = true;
  if ((rc = plan_create(ctx, mesh, op, dpn, n_bc ? fixed.data() : nullptr, flags & (TFM_SYS_FP64 | TFM_SYS_FP32),
                       &s->plan))) {
    tfm_system_destroy(s);
    return rc;
  }
  if (!s->plan.d_fixed) {
    if ((rc = dalloc(ctx, &s->plan.d_fixed, s->ndof))) { tfm_system_destroy(s); return rc; }
    TFM_CHECK_HIP(ctx, hipMemsetAsync(s->plan.d_fixed, 0, s->ndof, ctx->stream));
  }
  s->d_rp = s->plan.d_rp;
  s->d_ci = s->plan.d_ci;
  s->d_fixed = s->plan.d_fixed;
  s->nnz = s->plan.nnz;
  s->nrows = s->ndof;
  if ((flags & TFM_SYS_FP64) && (rc = dalloc(ctx, &s->d_v64, s->nnz))) { tfm_system_destroy(s); return rc; }
  if ((flags & TFM_SYS_FP32) && (rc = dalloc(ctx, &s->d_v32, s->nnz))) { tfm_system_destroy(s); return rc; }
  if ((rc = plan_numeric(ctx, &s->plan, s->d_v64, s->d_v32))) { tfm_system_destroy(s); return rc; }
  s->note_values(true, s->d_v64 != nullptr, s->d_v32 != nullptr);
  if ((rc = dalloc(ctx, &s->d_f, s->ndof))) { tfm_system_destroy(s); return rc; }
  // Homogeneous Dirichlet values: rhs[f] - K[f,c] * 0 leaves every free entry
  // unchanged, so fred is rhs restricted to the free DOFs.
  std::vector<double> f((size_t)s->ndof, 0.0);
  if (rhs)
    for (int64_t i = 0; i < s->ndof; i++) f[i] = fixed[i] ? 0.0 : rhs[i];
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(s->d_f, f.data(), sizeof(double) * s->ndof, hipMemcpyHostToDevice, ctx->stream));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = s;
  return TFM_OK;
}

extern "C" int tfm_system_assemble(tfm_system *s, const tfm_operator *op, int which) {
  if (!s) return TFM_E_INVALID;
  tfm_ctx *ctx = s->ctx;
  if (!s->has_plan) return fail(ctx, TFM_E_INVALID, "system was not built from a mesh");
  int rc = set_device(ctx);
  if (rc) return rc;
  if (op) {
    if (op->kind != s->plan.op.kind || op->model != s->plan.op.model || op->quad_order != s->plan.op.quad_order)
      return fail(ctx, TFM_E_INVALID, "re-assembly must keep the operator kind, model and rule");
    if ((rc = plan_set_coefficients(ctx, &s->plan, op))) return rc;
    if (op->kind == TFM_OP_SIMP)
      TFM_CHECK_HIP(ctx, hipMemcpyAsync(s->plan.d_tab + (size_t)s->plan.nq * kRuleStride, s->plan.ke0,
                                        sizeof(double) * 64, hipMemcpyHostToDevice, ctx->stream));
  }
  if (which == 0) which = TFM_SYS_FP64 | TFM_SYS_FP32;
  double *w64 = (which & TFM_SYS_FP64) ? s->d_v64 : nullptr;
  float *w32 = (which & TFM_SYS_FP32) ? s->d_v32 : nullptr;
  if ((rc = plan_numeric(ctx, &s->plan, w64, w32))) return rc;
  s->note_values(op != nullptr, w64 != nullptr, w32 != nullptr);
  return TFM_OK;
}

extern "C" int tfm_system_info(const tfm_system *s, int64_t *n_free, int64_t *nnz, int64_t *n_dof) {
  if (!s) return TFM_E_INVALID;
  if (n_free) *n_free = s->n_free;
  if (nnz) *nnz = s->nnz;
  if (n_dof) *n_dof = s->ndof;
  return TFM_OK;
}

extern "C" int tfm_debug_mf_apply(tfm_system *s, const double *x, double *y) {
  if (!s || !x || !y) return TFM_E_INVALID;
  tfm_ctx *ctx = s->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  MfBox box;
  if (!mf_box_setup(s, &box)) return fail(ctx, TFM_E_UNSUPPORTED, "system has no matrix-free operator");
  const int64_t n = s->ndof;
  hipStream_t st = ctx->stream;
  std::vector<float> hx((size_t)n), hq((size_t)n);
  for (int64_t i = 0; i < n; i++) hx[i] = (float)x[i];
  Scal *sc = nullptr;
  double *part = nullptr;
  float *z = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr, *mfc = nullptr;
  auto cleanup = [&]() { dfree(sc); dfree(part); dfree(z); dfree(p0); dfree(p1); dfree(q); dfree(mfc); };
  if ((rc = dalloc(ctx, &sc, 1)) || (rc = dalloc(ctx, &part, P_SLOTS * kMaxBlocks)) || (rc = dalloc(ctx, &z, n)) ||
      (rc = dalloc(ctx, &p0, n)) || (rc = dalloc(ctx, &p1, n)) || (rc = dalloc(ctx, &q, n)) ||
      (rc = mf_coef32(ctx, s, &box, &mfc))) {
    cleanup();
    return rc;
  }
  // iteration 0 of a fresh scalar block: beta = 0, so p = z = x and q = K x
  TFM_CHECK_HIP(ctx, hipMemsetAsync(sc, 0, sizeof(Scal), st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(p0, 0, sizeof(float) * n, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(z, hx.data(), sizeof(float) * n, hipMemcpyHostToDevice, st));
  mf_dir_launch(box, st, mf_grid(box), z, p0, p1, q, sc, part, 0);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hq.data(), q, sizeof(float) * n, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  cleanup();
  for (int64_t i = 0; i < n; i++) y[i] = (double)hq[i];
  return TFM_OK;
}

extern "C" int tfm_system_assembly_info(const tfm_system *s, int *path, int *tile_pairs, int *source, double *ms) {
  if (!s) return TFM_E_INVALID;
  const bool tet = s->has_plan && s->plan.use_tet;
  if (path) *path = tet ? 1 : !s->has_plan ? 0 : s->plan.use_coo ? 2 : s->plan.use_simp_grid ? 3 : 0;
  if (tile_pairs) *tile_pairs = tet ? s->plan.tet.tp : 0;
  if (source) *source = tet ? s->plan.tune.source : 0;
  if (ms) {
    ms[0] = tet ? s->plan.tune.ms[0] : 0.0;
    ms[1] = tet ? s->plan.tune.ms[1] : 0.0;
  }
  return TFM_OK;
}

// The reduced system in reference order; values from the FP32 array when
// fp32 is set (or the system holds no FP64 values).
static int system_download(const tfm_system *s, int64_t *row_ptrs, int64_t *col_ind, double *values, double *fred,
                           int64_t *free_dofs, bool fp32) {
  tfm_ctx *ctx = s->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  int64_t n = s->ndof;
  const bool use32 = fp32 || !s->d_v64;
  std::vector<int> rp(n + 1), ci(s->nnz);
  std::vector<double> v(!use32 ? s->nnz : 0), f(n);
  std::vector<float> v32(use32 && s->d_v32 ? s->nnz : 0);
  TFM_CHECK_HIP(ctx, hipMemcpy(rp.data(), s->d_rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  if (s->nnz) TFM_CHECK_HIP(ctx, hipMemcpy(ci.data(), s->d_ci, sizeof(int) * s->nnz, hipMemcpyDeviceToHost));
  if (s->nnz && v.size()) TFM_CHECK_HIP(ctx, hipMemcpy(v.data(), s->d_v64, sizeof(double) * s->nnz, hipMemcpyDeviceToHost));
  if (s->nnz && v32.size()) TFM_CHECK_HIP(ctx, hipMemcpy(v32.data(), s->d_v32, sizeof(float) * s->nnz, hipMemcpyDeviceToHost));
  TFM_CHECK_HIP(ctx, hipMemcpy(f.data(), s->d_f, sizeof(double) * n, hipMemcpyDeviceToHost));
  std::vector<int64_t> red((size_t)n, -1);
  for (int64_t i = 0; i < s->n_free; i++) red[s->free_dofs[i]] = i;
  int64_t o = 0;
  if (row_ptrs) row_ptrs[0] = 0;
  for (int64_t q = 0; q < s->n_free; q++) {
    int64_t g = s->free_dofs[q];
    for (int j = rp[g]; j < rp[g + 1]; j++) {
      if (col_ind) col_ind[o] = red[ci[j]];
      if (values) values[o] = use32 ? (double)v32[j] : v[j];
      o++;
    }
    if (row_ptrs) row_ptrs[q + 1] = o;
    if (fred) fred[q] = f[g];
    if (free_dofs) free_dofs[q] = g;
  }
  return TFM_OK;
}

extern "C" int tfm_system_download(const tfm_system *s, int64_t *row_ptrs, int64_t *col_ind, double *values,
                                   double *fred, int64_t *free_dofs) {
  if (!s) return TFM_E_INVALID;
  return system_download(s, row_ptrs, col_ind, values, fred, free_dofs, false);
}

extern "C" int tfm_debug_system_values32(const tfm_system *s, double *values) {
  if (!s) return TFM_E_INVALID;
  if (!s->d_v32) return fail(s->ctx, TFM_E_INVALID, "system holds no FP32 values");
  return system_download(s, nullptr, nullptr, values, nullptr, nullptr, true);
}

// Rows per CSR-stream chunk of the system's matrix (from its longest row).
static int system_chunk_rows(tfm_system *s, int *S) {
  if (s->csr_chunk > 0) {  // the pattern never changes
    *S = s->csr_chunk;
    return TFM_OK;
  }
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int n = (int)s->ndof;
  int mnz = 0, *d_m = nullptr;
  int rc = dalloc(ctx, &d_m, 1);
  if (rc) return rc;
  TFM_CHECK_HIP(ctx, hipMemsetAsync(d_m, 0, sizeof(int), st));
  if (n > 0)
    hipLaunchKernelGGL(k_max_row_nnz, dim3(std::min<unsigned>(grid1(n, 256), 2048)), dim3(256), 0, st, n, s->d_rp,
                       d_m);
  hipError_t e = hipMemcpy(&mnz, d_m, sizeof(int), hipMemcpyDeviceToHost);
  dfree(d_m);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  if (mnz > kWaveCap) return fail(ctx, TFM_E_UNSUPPORTED, "a matrix row is longer than the SpMV chunk buffer");
  *S = rows_per_chunk(mnz);
  s->csr_chunk = *S;
  return TFM_OK;
}

// The reference's direct solve (solve_multi, crates/tpt-fem-sparse/src/lib.rs:
// 244-278) of the reduced system for nrhs right-hand sides (d_F: device
// [nrhs][ndof], 0 at fixed DOFs; NULL: the system's own rhs), then the
// solve_with_dirichlet scatter into u [nrhs][ndof].  Each solution's FP64
// residual is certified: a matrix that is singular to working precision
// (factors without an exact zero pivot but leaves a residual above 1e-6 of
// ||f||) is reported as singular.
int tfm::system_solve_direct(tfm_system *s, int nrhs, const double *d_F, double *u, tfm_solve_stats *stats) {
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int n = (int)s->ndof;
  int S = 1, rc;
  if (s->n_free == 0)
    return fail(ctx, TFM_E_NUMERIC,
                "sparse numeric factorization/solve failed: rhs length 0 does not match matrix dimension 1");
  if (!s->d_v64) return fail(ctx, TFM_E_INVALID, "solve needs FP64 values (TFM_SYS_FP64)");
  if ((rc = system_chunk_rows(s, &S))) return rc;
  const double *F = d_F ? d_F : s->d_f;
  double *x = nullptr, *r = nullptr, *part = nullptr;
  Scal *sc = nullptr;
  auto cleanup = [&]() {
    dfree(x);
    dfree(r);
    dfree(part);
    dfree(sc);
  };
  if ((rc = dalloc(ctx, &x, (size_t)n * nrhs)) || (rc = dalloc(ctx, &r, n)) ||
      (rc = dalloc(ctx, &part, P_SLOTS * kMaxBlocks)) || (rc = dalloc(ctx, &sc, 1))) {
    cleanup();
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev0, st));
  if ((rc = direct_solve_dev(s, nrhs, F, x))) {
    cleanup();
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev1, st));
  const SpOp<double> K64 = csr_op<double>(n, s->d_rp, s->d_ci, s->d_v64, S);
  double worst = 0.0;
  std::vector<double> hx((size_t)n * nrhs);
  for (int k = 0; k < nrhs && !rc; k++) {
    residual_dispatch(S, st, K64.nb, n, s->d_rp, s->d_ci, s->d_v64, x + (size_t)k * n, F + (size_t)k * n, r, nullptr,
                      sc, part);
    Scal h{};
    hipError_t e = hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail(ctx, TFM_E_HIP, std::string("residual: ") + hipGetErrorString(e));
    const double rel = h.rr0 > 0 ? std::sqrt(h.rr / h.rr0) : std::sqrt(h.rr);
    worst = std::max(worst, rel);
  }
  if (!rc && hipMemcpy(hx.data(), x, sizeof(double) * hx.size(), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(ctx, TFM_E_HIP, "solution download failed");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  cleanup();
  if (rc) return rc;
  if (!(worst <= 1e-6))
    return fail(ctx, TFM_E_NUMERIC, "sparse numeric factorization/solve failed: singular matrix");
  for (int k = 0; k < nrhs; k++) {
    double *uk = u + (size_t)k * n;
    const double *xk = hx.data() + (size_t)k * n;
    for (int64_t i = 0; i < n; i++) uk[i] = 0.0;
    for (int64_t f : s->free_dofs) uk[f] = xk[f];
    for (size_t b = 0; b < s->bc_dof.size(); b++) uk[s->bc_dof[b]] = s->bc_val[b];
  }
  if (stats)
    for (int k = 0; k < nrhs; k++) stats[k] = tfm_solve_stats{0, 0, worst, ms * 1e-3};
  return TFM_OK;
}

static int system_solve_iter(tfm_system *s, const tfm_solve_opts &o, double *u, tfm_solve_stats *stats,
                             bool *not_spd);

extern "C" int tfm_system_solve(tfm_system *s, const tfm_solve_opts *opts, double *u, tfm_solve_stats *stats) {
  if (!s || !u) return TFM_E_INVALID;
  tfm_ctx *ctx = s->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-10, 20000, 1e-5};
  if (opts) o = *opts;
  if (o.max_iter <= 0) o.max_iter = 20000;
  if (o.mode != TFM_SOLVE_PCG64 && o.mode != TFM_SOLVE_MIXED && o.mode != TFM_SOLVE_PCG64_MG &&
      o.mode != TFM_SOLVE_DIRECT && o.mode != TFM_SOLVE_MIXED_CSR)
    return fail(ctx, TFM_E_INVALID, "unknown solve mode");
  if (o.mode == TFM_SOLVE_DIRECT) return system_solve_direct(s, 1, nullptr, u, stats);
  if (o.mode == TFM_SOLVE_PCG64_MG) {
    if (s->n_free == 0)
      return fail(ctx, TFM_E_NUMERIC,
                  "sparse numeric factorization/solve failed: rhs length 0 does not match matrix dimension 1");
    return system_solve_mg(s, o, u, stats);  // the batched solver with one problem, kept on the system
  }
  bool not_spd = false;
  rc = system_solve_iter(s, o, u, stats, &not_spd);
  // The reference factors every reduced system directly, SPD or not: a system
  // on which CG breaks down or stalls (not SPD, or too ill-conditioned for
  // the iteration budget) gets the dense LU when it is small enough.
  if ((not_spd || rc == TFM_E_NOT_CONVERGED) && s->n_free <= kDirectMax)
    return system_solve_direct(s, 1, nullptr, u, stats);
  return rc;
}

static int system_solve_iter(tfm_system *s, const tfm_solve_opts &o, double *u, tfm_solve_stats *stats,
                             bool *not_spd) {
  tfm_ctx *ctx = s->ctx;
  int rc = TFM_OK;
  const int n = (int)s->ndof;
  const bool mixed = o.mode == TFM_SOLVE_MIXED || o.mode == TFM_SOLVE_MIXED_CSR;
  // The precision form's FP32 operator: matrix-free on box meshes of P1 tets
  // (TFM_SOLVE_MIXED), else the assembled FP32 values.
  MfBox box;
  const bool mf = o.mode == TFM_SOLVE_MIXED && mf_box_setup(s, &box);
  // No free DOF: the reference's solve gets the 1 x 1 CSR of an empty
  // triplet list against a length-0 rhs and fails (sparse/src/lib.rs:244-258).
  if (s->n_free == 0)
    return fail(ctx, TFM_E_NUMERIC,
                "sparse numeric factorization/solve failed: rhs length 0 does not match matrix dimension 1");
  if (!s->d_v64) return fail(ctx, TFM_E_INVALID, "solve needs FP64 values (TFM_SYS_FP64)");
  if (mf && s->ver64 != s->version)
    return fail(ctx, TFM_E_INVALID, "mixed solve: the FP64 values predate the last coefficients");
  if (mixed && !mf && !s->d_v32) return fail(ctx, TFM_E_INVALID, "mixed solve needs FP32 values (TFM_SYS_FP32)");
  if (mixed && !mf && !s->fp32_current())
    return fail(ctx, TFM_E_INVALID, "mixed solve: the FP32 values predate the last FP64 re-assembly");
  hipStream_t st = ctx->stream;
  // rows per CSR-stream chunk, from the longest row
  int S = 1;
  {
    const int rc0 = system_chunk_rows(s, &S);
    if (rc0) return rc0;
  }
  // device work space
  Scal *sc = nullptr;
  double *part = nullptr, *x = nullptr;
  int *bad = nullptr;
  PcgBufs<double> B64;
  PcgBufs<float> B32;
  float *x32 = nullptr, *mfc = nullptr;
  auto cleanup = [&]() {
    dfree(sc); dfree(part); dfree(x); dfree(bad); dfree(x32); dfree(mfc);
    dfree(B64.r); dfree(B64.z); dfree(B64.p); dfree(B64.q); dfree(B64.dinv); dfree(B64.p2);
    dfree(B32.r); dfree(B32.z); dfree(B32.p); dfree(B32.q); dfree(B32.dinv); dfree(B32.p2);
  };
  if ((rc = dalloc(ctx, &sc, 1)) || (rc = dalloc(ctx, &part, P_SLOTS * kMaxBlocks)) || (rc = dalloc(ctx, &x, n)) ||
      (rc = dalloc(ctx, &bad, 1)) || (rc = dalloc(ctx, &B64.r, n)) || (rc = dalloc(ctx, &B64.dinv, n))) {
    cleanup();
    return rc;
  }
  if (!mixed && ((rc = dalloc(ctx, &B64.z, n)) || (rc = dalloc(ctx, &B64.p, n)) || (rc = dalloc(ctx, &B64.q, n)) ||
                 (rc = dalloc(ctx, &B64.p2, n)))) {
    cleanup();
    return rc;
  }
  if (mixed && ((rc = dalloc(ctx, &x32, n)) || (rc = dalloc(ctx, &B32.r, n)) || (rc = dalloc(ctx, &B32.z, n)) ||
                (rc = dalloc(ctx, &B32.p, n)) || (rc = dalloc(ctx, &B32.q, n)) || (rc = dalloc(ctx, &B32.dinv, n)) ||
                (rc = dalloc(ctx, &B32.p2, n)))) {
    cleanup();
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipMemsetAsync(sc, 0, sizeof(Scal), st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(x, 0, sizeof(double) * n, st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(bad, 0, sizeof(int), st));
  int nb1 = (int)grid1(n, 256);
  hipLaunchKernelGGL(k_diag_inv<double>, dim3(nb1), dim3(256), 0, st, n, s->d_rp, s->d_ci, s->d_v64, s->d_fixed,
                     B64.dinv, bad);
  if (mixed && !mf)
    hipLaunchKernelGGL(k_diag_inv<float>, dim3(nb1), dim3(256), 0, st, n, s->d_rp, s->d_ci, s->d_v32, s->d_fixed,
                       B32.dinv, bad);
  if (mf) hipLaunchKernelGGL(k_round_f32, dim3(nb1), dim3(256), 0, st, n, B64.dinv, B32.dinv);
  int hbad = 0;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  if (hbad) {
    cleanup();
    *not_spd = true;
    return fail(ctx, TFM_E_NUMERIC,
                "sparse numeric factorization/solve failed: singular matrix (a free DOF has no positive diagonal)");
  }
  // node-blocked sliced copy of the matrix for the iteration (structure once
  // per system, values on every solve)
  if ((rc = bsell_build(s))) {
    cleanup();
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev0, st));
  const SpOp<double> K64 = csr_op<double>(n, s->d_rp, s->d_ci, s->d_v64, S);
  SpOp<double> A64 = K64;
  SpOp<float> A32 = csr_op<float>(n, s->d_rp, s->d_ci, s->d_v32, S);
  if (mf && (rc = mf_coef32(ctx, s, &box, &mfc))) {
    cleanup();
    return rc;
  }
  if (mf) {
    A32.mf = true;
    A32.box = box;
    A32.nb = mf_grid(box);
  }
  SpOp<double> R64 = K64;  // the FP64 residuals' operator
  if (s->bsell.state > 0) {
    // FP64 values in the sliced copy: the plain iteration's operator, and
    // the precision form's residual replacements
    if (!(rc = bsell_values<double>(s, s->d_v64, &s->bsell.d_v64))) {
      R64 = bsell_op<double>(n, s->bsell, s->bsell.d_v64);
      R64.rp = s->d_rp;
      if (!mixed) A64 = R64;
    }
    if (mixed && !mf && !(rc = bsell_values<float>(s, s->d_v32, &s->bsell.d_v32)))
      A32 = bsell_op<float>(n, s->bsell, s->bsell.d_v32);
    if (rc) {
      cleanup();
      return rc;
    }
  }
  int total_it = 0, outer = 0, brk = 0;
  bool pcg_conv = false;
  double relres = 0.0;
  const int nbs = K64.nb;
  if (!mixed) {
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(B64.r, s->d_f, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    // Restarted PCG: when the recurrence residual says converged but the true
    // FP64 residual f - K x is still above rtol (recurrence drift), restart
    // from the true residual (at most 4 restarts).
    double target = o.rtol;
    for (outer = 0; outer < 5 && !rc; outer++) {
      double rr = 0;
      int it = 0;
      double rel_in = 1.0;
      if (outer > 0) {
        residual_dispatch(S, st, nbs, n, s->d_rp, s->d_ci, s->d_v64, x, s->d_f, B64.r, nullptr, sc, part);
        Scal h{};
        TFM_CHECK_HIP(ctx, hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
        TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
        rel_in = h.rr0 > 0 ? std::sqrt(h.rr / h.rr0) : 0.0;
        if (rel_in <= o.rtol) break;
        target = o.rtol / rel_in;  // relative to the restart residual
      } else {
        TFM_CHECK_HIP(ctx, hipMemcpyAsync(B64.r, s->d_f, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      }
      rc = pcg_run<double, double>(ctx, A64, B64, x, target, o.max_iter - total_it, sc, part, &it, &brk, &rr);
      total_it += it;
      pcg_conv = rr <= target;
      if (brk || !pcg_conv || total_it >= o.max_iter) break;
    }
    outer = std::max(outer, 1);
  } else {
    // FP32 PCG with FP64 reliable-update refinement; inner_rtol is the
    // residual reduction between FP64 replacements (0 picks 0.3).
    const double delta = (o.inner_rtol > 0.0 && o.inner_rtol < 1.0) ? o.inner_rtol : 0.3;
    rc = mixed_run(ctx, R64, A32, s->d_f, B32, B64.r, x32, x, o.rtol, delta,
                   o.max_iter, sc, part, &total_it, &outer, &brk, &relres);
  }
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev1, st));
  // final FP64 residual certificate
  residual_dispatch(S, st, nbs, n, s->d_rp, s->d_ci, s->d_v64, x, s->d_f, B64.r, nullptr, sc, part);
  Scal h{};
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&h, sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
  std::vector<double> hx((size_t)n);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hx.data(), x, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  relres = h.rr0 > 0 ? std::sqrt(h.rr / h.rr0) : 0.0;
  cleanup();
  if (rc) return rc;
  // solve_with_dirichlet scatter: u[free] = x, then u[bc] = value in order.
  for (int64_t i = 0; i < n; i++) u[i] = 0.0;
  for (int64_t f : s->free_dofs) u[f] = hx[f];
  for (size_t b = 0; b < s->bc_dof.size(); b++) u[s->bc_dof[b]] = s->bc_val[b];
  if (stats) {
    stats->iterations = total_it;
    stats->outer = outer;
    stats->rel_residual = relres;
    stats->seconds = ms * 1e-3;
  }
  if (brk) {
    *not_spd = true;
    return fail(ctx, TFM_E_BREAKDOWN, "PCG breakdown: p'Kp <= 0 (matrix not SPD)");
  }
  // PCG64 stops on its recurrence residual (the true FP64 residual can sit
  // slightly above it at tight tolerances: it is reported, not tested); the
  // refinement loop is driven by the true FP64 residual itself.
  bool ok = mixed ? (relres <= o.rtol) : (pcg_conv && relres <= o.rtol * 10.0);
  if (!ok) {
    char buf[160];
    snprintf(buf, sizeof buf, "PCG did not reach the tolerance: rel. residual %.3e after %d iterations", relres,
             total_it);
    return fail(ctx, TFM_E_NOT_CONVERGED, buf);
  }
  return TFM_OK;
}

// ---- allocation-free solve -------------------------------------------------
// The work space of tfm_system_solve_noalloc: PCG scalars, block partial
// sums, flags and the seven work vectors of an n-DOF system, in one
// file-scope device array (64 MiB, so n <= 1,197,000 DOFs).  Each GPU holds
// its own copy for the life of the process; its address on a device is
// looked up once, when the first context on that device is created.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
constexpr int kNoAllocDevices = 64;
namespace tfm_noalloc {
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];
}
static void *g_scratch_dev[kNoAllocDevices] = {};
static std::mutex g_scratch_mu;

int tfm::noalloc_init(tfm_ctx *ctx) {  // ctx->device is current
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (ctx->device < 0 || ctx->device >= kNoAllocDevices) return TFM_OK;  // the entry point refuses such devices
  if (g_scratch_dev[ctx->device]) return TFM_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(tfm_noalloc::scratch));
  if (e != hipSuccess || !p)
    return fail(ctx, TFM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_dev[ctx->device] = p;
  return TFM_OK;
}

struct NoAllocLayout {
  Scal *sc = nullptr;
  double *part = nullptr;
  int *flag = nullptr;  // [0] bad diagonal, [1] longest row
  double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr, *p2 = nullptr;
  size_t bytes = 0;
};

static NoAllocLayout noalloc_layout(char *base, int64_t n) {
  NoAllocLayout L;
  size_t off = 0;
  auto take = [&](size_t b) {
    char *q = base ? base + off : nullptr;
    off += (b + 255) / 256 * 256;
    return q;
  };
  L.sc = (Scal *)take(sizeof(Scal));
  L.part = (double *)take(sizeof(double) * P_SLOTS * kMaxBlocks);
  L.flag = (int *)take(2 * sizeof(int));
  double **v[7] = {&L.x, &L.r, &L.z, &L.p, &L.q, &L.dinv, &L.p2};
  for (double **d : v) *d = (double *)take(sizeof(double) * (size_t)std::max<int64_t>(n, 1));
  L.bytes = off;
  return L;
}

__global__ void k_zero_bytes(unsigned char *__restrict__ p, int64_t nbytes) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_zero_f64(double *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

extern "C" int tfm_system_solve_noalloc(tfm_system *s, const tfm_solve_opts *opts, double *u,
                                        tfm_solve_stats *stats) {
  if (!s || !u) return TFM_E_INVALID;
  tfm_ctx *ctx = s->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-10, 20000, 1e-5};
  if (opts) o = *opts;
  if (o.max_iter <= 0) o.max_iter = 20000;
  if (o.mode != TFM_SOLVE_PCG64)
    return fail(ctx, TFM_E_UNSUPPORTED, "the allocation-free solve runs the FP64 Jacobi PCG (TFM_SOLVE_PCG64) only");
  if (ctx->device < 0 || ctx->device >= kNoAllocDevices || !g_scratch_dev[ctx->device])
    return fail(ctx, TFM_E_INVALID, "allocation-free solve: no work space on this device");
  if (s->n_free == 0)
    return fail(ctx, TFM_E_NUMERIC,
                "sparse numeric factorization/solve failed: rhs length 0 does not match matrix dimension 1");
  if (!s->d_v64) return fail(ctx, TFM_E_INVALID, "solve needs FP64 values (TFM_SYS_FP64)");
  const int n = (int)s->ndof;
  const NoAllocLayout need = noalloc_layout(nullptr, n);
  if (need.bytes > kNoAllocBytes) {
    char buf[200];
    snprintf(buf, sizeof buf,
             "allocation-free solve: %d DOFs need %zu bytes of work space, the fixed array holds %zu "
             "(at most about 1,197,000 DOFs)",
             n, need.bytes, kNoAllocBytes);
    return fail(ctx, TFM_E_UNSUPPORTED, buf);
  }
  hipStream_t st = ctx->stream;
  const NoAllocLayout L = noalloc_layout((char *)g_scratch_dev[ctx->device], n);
  // the array keeps the previous call's contents: clear what must start at 0
  hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(256), 0, st, (unsigned char *)L.sc, (int64_t)sizeof(Scal));
  hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(64), 0, st, (unsigned char *)L.flag, (int64_t)(2 * sizeof(int)));
  hipLaunchKernelGGL(k_zero_f64, dim3(std::min<unsigned>(grid1(n, 256), 4096)), dim3(256), 0, st, L.x, (int64_t)n);
  if (n > 0) hipLaunchKernelGGL(k_max_row_nnz, dim3(grid1(n, 256)), dim3(256), 0, st, n, s->d_rp, L.flag + 1);
  hipLaunchKernelGGL(k_diag_inv<double>, dim3(grid1(n, 256)), dim3(256), 0, st, n, s->d_rp, s->d_ci, s->d_v64,
                     s->d_fixed, L.dinv, L.flag);
  int hflag[2] = {0, 0};
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hflag, L.flag, sizeof hflag, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  if (hflag[0])
    return fail(ctx, TFM_E_NUMERIC,
                "sparse numeric factorization/solve failed: singular matrix (a free DOF has no positive diagonal)");
  if (hflag[1] > kWaveCap) return fail(ctx, TFM_E_UNSUPPORTED, "a matrix row is longer than the SpMV chunk buffer");
  const int S = rows_per_chunk(hflag[1]);
  PcgBufs<double> B;
  B.r = L.r;
  B.z = L.z;
  B.p = L.p;
  B.q = L.q;
  B.dinv = L.dinv;
  B.p2 = L.p2;
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev0, st));
  const SpOp<double> K64 = csr_op<double>(n, s->d_rp, s->d_ci, s->d_v64, S);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.r, s->d_f, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  // restarted PCG on the true FP64 residual, as tfm_system_solve runs it
  int total_it = 0, outer = 0, brk = 0;
  bool pcg_conv = false;
  double target = o.rtol;
  for (outer = 0; outer < 5 && !rc; outer++) {
    double rr = 0, rel_in = 1.0;
    int it = 0;
    if (outer > 0) {
      residual_dispatch(S, st, K64.nb, n, s->d_rp, s->d_ci, s->d_v64, L.x, s->d_f, B.r, nullptr, L.sc, L.part);
      Scal h{};
      TFM_CHECK_HIP(ctx, hipMemcpyAsync(&h, L.sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
      TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
      rel_in = h.rr0 > 0 ? std::sqrt(h.rr / h.rr0) : 0.0;
      if (rel_in <= o.rtol) break;
      target = o.rtol / rel_in;
    }
    rc = pcg_run<double, double>(ctx, K64, B, L.x, target, o.max_iter - total_it, L.sc, L.part, &it, &brk, &rr,
                                 false);
    total_it += it;
    pcg_conv = rr <= target;
    if (brk || !pcg_conv || total_it >= o.max_iter) break;
  }
  outer = std::max(outer, 1);
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev1, st));
  residual_dispatch(S, st, K64.nb, n, s->d_rp, s->d_ci, s->d_v64, L.x, s->d_f, B.r, nullptr, L.sc, L.part);
  Scal h{};
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&h, L.sc, sizeof(Scal), hipMemcpyDeviceToHost, st));
  // x is 0 at the fixed DOFs (their Jacobi weight is 0): u = x, then the
  // Dirichlet values in bcs order (solve_with_dirichlet's scatter)
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(u, L.x, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  if (rc) return rc;
  for (size_t b = 0; b < s->bc_dof.size(); b++) u[s->bc_dof[b]] = s->bc_val[b];
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  const double relres = h.rr0 > 0 ? std::sqrt(h.rr / h.rr0) : 0.0;
  if (stats) *stats = tfm_solve_stats{total_it, outer, relres, ms * 1e-3};
  if (brk) return fail(ctx, TFM_E_BREAKDOWN, "PCG breakdown: p'Kp <= 0 (matrix not SPD)");
  if (!(pcg_conv && relres <= o.rtol * 10.0)) {
    char buf[160];
    snprintf(buf, sizeof buf, "PCG did not reach the tolerance: rel. residual %.3e after %d iterations", relres,
             total_it);
    return fail(ctx, TFM_E_NOT_CONVERGED, buf);
  }
  return TFM_OK;
}

extern "C" void tfm_system_destroy(tfm_system *s) {
  if (!s) return;
  if (s->has_plan) {
    plan_destroy(&s->plan);
    s->d_rp = nullptr;
    s->d_ci = nullptr;
    s->d_fixed = nullptr;
  } else {
    dfree(s->d_rp);
    dfree(s->d_ci);
    dfree(s->d_fixed);
  }
  dfree(s->d_v64);
  dfree(s->d_v32);
  bsell_destroy(&s->bsell);
  bsolver_destroy(s->mg_solver);
  bsolver_destroy(s->batch_solver);
  dfree(s->d_f);
  delete s;
}

// ---------------------------------------------------------------------------
// Drivers
// ---------------------------------------------------------------------------
extern "C" int tfm_solve_poisson(tfm_ctx *ctx, const tfm_mesh *mesh, double kappa, int quad_order, double f,
                                 int64_t n_bc, const int64_t *bc_dof, const double *bc_val, const tfm_solve_opts *opts,
                                 double *u) {
  if (!ctx || !mesh || !u) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  std::vector<double> rhs((size_t)mesh->n_nodes, 0.0);
  if ((rc = tfm_load_source(ctx, mesh, quad_order, f, rhs.data()))) return rc;
  tfm_operator op{};
  op.kind = TFM_OP_POISSON;
  op.quad_order = quad_order;
  op.coef = kappa;
  tfm_system *s = nullptr;
  if ((rc = tfm_system_create(ctx, mesh, &op, 1, n_bc, bc_dof, bc_val, rhs.data(), TFM_SYS_FP64, &s))) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-13, 100000, 1e-5};
  if (opts) o = *opts;
  rc = tfm_system_solve(s, &o, u, nullptr);
  tfm_system_destroy(s);
  return rc;
}

extern "C" int tfm_solve_poisson_nr(tfm_ctx *ctx, const tfm_mesh *mesh, double kappa, int quad_order, double f,
                                    const double *neumann_g, const double *robin_h, int64_t n_bc,
                                    const int64_t *bc_dof, const double *bc_val, const tfm_solve_opts *opts,
                                    double *u) {
  if (!ctx || !mesh || !u) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  std::vector<double> rhs((size_t)mesh->n_nodes, 0.0);
  if ((rc = tfm_load_source(ctx, mesh, quad_order, f, rhs.data()))) return rc;
  if (neumann_g && (rc = tfm_load_neumann(ctx, mesh, 1, *neumann_g, rhs.data()))) return rc;
  tfm_operator op{};
  op.kind = TFM_OP_POISSON;
  op.quad_order = quad_order;
  op.coef = kappa;
  tfm_matrix *K = nullptr;
  rc = robin_h ? tfm_assemble_robin(ctx, mesh, &op, 1, *robin_h, -1, 0.0, 0.0, &K) : tfm_assemble(ctx, mesh, &op, 1, &K);
  if (rc) return rc;
  tfm_system *s = nullptr;
  rc = tfm_system_from_matrix(ctx, K, mesh->n_nodes, rhs.data(), n_bc, bc_dof, bc_val, &s);
  tfm_matrix_destroy(K);
  if (rc) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-13, 100000, 1e-5};
  if (opts) o = *opts;
  rc = tfm_system_solve(s, &o, u, nullptr);
  tfm_system_destroy(s);
  return rc;
}

extern "C" int tfm_solve_poisson_qp(tfm_ctx *ctx, const tfm_mesh *mesh, double kappa, int quad_order,
                                    const double *sq, const double *neumann_gq, const double *robin_hq, int64_t n_bc,
                                    const int64_t *bc_dof, const double *bc_val, const tfm_solve_opts *opts,
                                    double *u) {
  if (!ctx || !mesh || !u || (!sq && mesh->n_elem > 0)) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_mesh *m = const_cast<tfm_mesh *>(mesh);
  const int64_t n = mesh->n_nodes;
  double *d_rhs = nullptr;
  if ((rc = dalloc(ctx, &d_rhs, n))) return rc;
  TFM_CHECK_HIP(ctx, hipMemsetAsync(d_rhs, 0, sizeof(double) * std::max<int64_t>(n, 1), ctx->stream));
  rc = load_source_qp_dev(ctx, m, quad_order, sq, d_rhs);
  // apply_neumann / apply_robin use FACE_QUAD_ORDER = 2 (assembly/src/lib.rs:596, :679)
  if (!rc && neumann_gq) rc = load_neumann_qp_dev(ctx, m, 1, 2, neumann_gq, d_rhs);
  std::vector<double> rhs((size_t)n, 0.0);
  if (!rc && n > 0 && hipMemcpy(rhs.data(), d_rhs, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(ctx, TFM_E_HIP, "rhs download failed");
  dfree(d_rhs);
  if (rc) return rc;
  tfm_operator op{};
  op.kind = TFM_OP_POISSON;
  op.quad_order = quad_order;
  op.coef = kappa;
  tfm_system *s = nullptr;
  if (robin_hq) {
    tfm_matrix *K = nullptr;
    if ((rc = assemble_robin_dev(ctx, mesh, &op, 1, 0.0, -1, 0.0, 0.0, 2, robin_hq, &K))) return rc;
    rc = tfm_system_from_matrix(ctx, K, n, rhs.data(), n_bc, bc_dof, bc_val, &s);
    tfm_matrix_destroy(K);
  } else {
    rc = tfm_system_create(ctx, mesh, &op, 1, n_bc, bc_dof, bc_val, rhs.data(), TFM_SYS_FP64, &s);
  }
  if (rc) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-13, 100000, 1e-5};
  if (opts) o = *opts;
  rc = tfm_system_solve(s, &o, u, nullptr);
  tfm_system_destroy(s);
  return rc;
}

extern "C" int tfm_solve_elasticity_qp(tfm_ctx *ctx, const tfm_mesh *mesh, int model, double young, double nu,
                                       int quad_order, const double *bq, int64_t n_bc, const int64_t *bc_dof,
                                       const double *bc_val, const tfm_solve_opts *opts, double *u) {
  if (!ctx || !mesh || !u || (!bq && mesh->n_elem > 0)) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  const int dim = mesh->dim;
  std::vector<double> rhs((size_t)(mesh->n_nodes * dim), 0.0);
  if ((rc = tfm_load_body_qp(ctx, mesh, quad_order, bq, rhs.data()))) return rc;
  tfm_operator op{};
  op.kind = TFM_OP_ELASTICITY;
  op.model = model;
  op.quad_order = quad_order;
  op.coef = young;
  op.nu = nu;
  tfm_system *s = nullptr;
  if ((rc = tfm_system_create(ctx, mesh, &op, dim, n_bc, bc_dof, bc_val, rhs.data(), TFM_SYS_FP64, &s))) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-13, 100000, 1e-5};
  if (opts) o = *opts;
  rc = tfm_system_solve(s, &o, u, nullptr);
  tfm_system_destroy(s);
  return rc;
}

extern "C" int tfm_solve_elasticity(tfm_ctx *ctx, const tfm_mesh *mesh, int model, double young, double nu,
                                    int quad_order, const double *body, int64_t n_bc, const int64_t *bc_dof,
                                    const double *bc_val, const tfm_solve_opts *opts, double *u) {
  if (!ctx || !mesh || !u) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  int dim = mesh->dim;
  std::vector<double> rhs((size_t)(mesh->n_nodes * dim), 0.0);
  double zero[3] = {0, 0, 0};
  if ((rc = tfm_load_body(ctx, mesh, quad_order, body ? body : zero, rhs.data()))) return rc;
  tfm_operator op{};
  op.kind = TFM_OP_ELASTICITY;
  op.model = model;
  op.quad_order = quad_order;
  op.coef = young;
  op.nu = nu;
  tfm_system *s = nullptr;
  if ((rc = tfm_system_create(ctx, mesh, &op, dim, n_bc, bc_dof, bc_val, rhs.data(), TFM_SYS_FP64, &s))) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-13, 100000, 1e-5};
  if (opts) o = *opts;
  rc = tfm_system_solve(s, &o, u, nullptr);
  tfm_system_destroy(s);
  return rc;
}
