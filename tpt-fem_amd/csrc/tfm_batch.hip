// tfm_batch.hip -- many independent problems in one launch.
//
// Batching form: P independent meshes (uneven box meshes, or SIMP designs on
// one grid) are concatenated into one mesh, so one assembly pass fills one
// block-diagonal reduced system.  The solver then runs P x NR independent
// Jacobi-PCG iterations at once (NR right-hand sides per problem, vectors
// interleaved [dof][rhs]): every block of a launch belongs to exactly one
// problem, dot products are reduced per problem (per-block partials, the
// problem's last block sums them in order), and each (problem, rhs) segment
// has its own alpha, beta and convergence flag.  Converged segments stop
// updating; a problem whose segments are all converged skips its blocks.
// Iterations are replayed from a captured hipGraph.  The precision form wraps
// the FP32 segments in per-segment FP64 refinement.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "tfm_elem.h"
#include "tfm_internal.h"

namespace tfm {

constexpr int kBB = 256;       // threads per block
constexpr int kRowsPerBlk = 512;

struct Seg {
  double rz, pq, rr, alpha, beta, tol2, rr0, ff, fu;
  int done, iters, breakdown, pad;
};

struct BatchDev {
  int P = 0, NR = 1, nblk = 0;
  int *blk_prob = nullptr, *blk_r0 = nullptr, *blk_r1 = nullptr;
  int *prob_blk0 = nullptr, *prob_nblk = nullptr;
  unsigned int *ticket = nullptr;  // [P * 4]
  double *part = nullptr;          // [nblk * NR * 3]
  Seg *seg = nullptr;              // [P * NR]
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Per-problem deterministic reduction of K quantities for each of NR rhs.
// vals[r*K + k] is this thread's contribution.  Returns true in thread 0 of
// the problem's last block, with totals in out[r*K + k].
template <int NR, int K>
__device__ bool prob_reduce(const double (&vals)[NR * K], const BatchDev &B, int b, int p, unsigned int *tk,
                            double (&out)[NR * K]) {
  __shared__ double sh[NR * K][kBB / 64];
  __shared__ bool last;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NR * K; i++) {
    double s = wsum(vals[i]);
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < NR * K) {
    double t = 0.0;
    for (int q = 0; q < kBB / 64; q++) t += sh[threadIdx.x][q];
    B.part[(size_t)b * NR * K + threadIdx.x] = t;
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = (atomicAdd(tk, 1u) == (unsigned)B.prob_nblk[p] - 1u);
  __syncthreads();
  if (!last) return false;
  __threadfence();
  if (threadIdx.x == 0) {
    const int b0 = B.prob_blk0[p], nb = B.prob_nblk[p];
    for (int i = 0; i < NR * K; i++) out[i] = 0.0;
    for (int q = 0; q < nb; q++)
      for (int i = 0; i < NR * K; i++) out[i] += B.part[(size_t)(b0 + q) * NR * K + i];
    *tk = 0u;
    return true;
  }
  return false;
}

template <int NR>
__device__ __forceinline__ bool all_done(const Seg *s) {
  bool d = true;
#pragma unroll
  for (int r = 0; r < NR; r++) d &= s[r].done != 0;
  return d;
}

// z = dinv r, p = z, per-segment rz, rr; tol2 = rtol^2 * rr0
template <typename VT, int NR>
__global__ __launch_bounds__(kBB) void kb_init(BatchDev B, const VT *__restrict__ r, VT *__restrict__ z,
                                               VT *__restrict__ p, const VT *__restrict__ dinv, double rtol2) {
  const int b = blockIdx.x, pr = B.blk_prob[b];
  Seg *S = B.seg + pr * NR;
  if (all_done<NR>(S)) return;
  double v[NR * 2];
#pragma unroll
  for (int i = 0; i < NR * 2; i++) v[i] = 0.0;
  for (int row = B.blk_r0[b] + threadIdx.x; row < B.blk_r1[b]; row += kBB) {
#pragma unroll
    for (int k = 0; k < NR; k++) {
      size_t i = (size_t)row * NR + k;
      VT zi = dinv[row] * r[i];
      z[i] = zi;
      p[i] = zi;
      v[k * 2] += (double)r[i] * (double)zi;
      v[k * 2 + 1] += (double)r[i] * (double)r[i];
    }
  }
  double out[NR * 2];
  if (prob_reduce<NR, 2>(v, B, b, pr, &B.ticket[pr * 4 + 0], out)) {
    for (int k = 0; k < NR; k++) {
      Seg &s = S[k];
      if (s.done) continue;
      s.rz = out[k * 2];
      s.rr = out[k * 2 + 1];
      s.rr0 = out[k * 2 + 1];
      s.tol2 = rtol2 * s.rr0;
      s.iters = 0;
      s.breakdown = 0;
      s.done = (s.rr0 == 0.0) ? 1 : 0;
    }
  }
}

// q = K p for NR vectors; partial p.q -> alpha.  Lanes: S per row (NR == 1)
// or one lane per (row, rhs) (NR > 1, the matrix entry is shared by the
// NR lanes of a row).
template <typename VT, int NR, int S>
__global__ __launch_bounds__(kBB) void kb_spmv(BatchDev B, const int *__restrict__ rp, const int *__restrict__ ci,
                                               const VT *__restrict__ v, const VT *__restrict__ p, VT *__restrict__ q) {
  const int b = blockIdx.x, pr = B.blk_prob[b];
  Seg *Sg = B.seg + pr * NR;
  if (all_done<NR>(Sg)) return;
  constexpr int L = NR > 1 ? NR : S;  // lanes per row
  const int lane = threadIdx.x % L, grp = threadIdx.x / L;
  double dv[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) dv[k] = 0.0;
  for (int row = B.blk_r0[b] + grp; row < B.blk_r1[b]; row += kBB / L) {
    const int rb = rp[row], re = rp[row + 1];
    if constexpr (NR == 1) {
      VT acc = 0;
      for (int j = rb + lane; j < re; j += S) acc += v[j] * p[ci[j]];
#pragma unroll
      for (int o = S / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, S);
      if (lane == 0) {
        q[row] = acc;
        dv[0] += (double)p[row] * (double)acc;
      }
    } else {
      VT acc = 0;
      for (int j = rb; j < re; j++) acc += v[j] * p[(size_t)ci[j] * NR + lane];
      q[(size_t)row * NR + lane] = acc;
      // lane k of the row owns rhs k: scatter into dv[k] via a select
#pragma unroll
      for (int k = 0; k < NR; k++)
        if (k == lane) dv[k] += (double)p[(size_t)row * NR + lane] * (double)acc;
    }
  }
  double out[NR];
  if (prob_reduce<NR, 1>(dv, B, b, pr, &B.ticket[pr * 4 + 1], out)) {
    for (int k = 0; k < NR; k++) {
      Seg &s = Sg[k];
      if (s.done) continue;
      s.pq = out[k];
      if (!(out[k] > 0.0)) {
        s.breakdown = 1;
        s.done = 1;
        s.alpha = 0.0;
      } else {
        s.alpha = s.rz / out[k];
      }
    }
  }
}

template <typename VT, typename XT, int NR>
__global__ __launch_bounds__(kBB) void kb_update(BatchDev B, XT *__restrict__ x, VT *__restrict__ r,
                                                 VT *__restrict__ z, const VT *__restrict__ p, const VT *__restrict__ q,
                                                 const VT *__restrict__ dinv) {
  const int b = blockIdx.x, pr = B.blk_prob[b];
  Seg *S = B.seg + pr * NR;
  if (all_done<NR>(S)) return;
  double a[NR];
  bool act[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    act[k] = !S[k].done;
    a[k] = S[k].alpha;
  }
  double v[NR * 2];
#pragma unroll
  for (int i = 0; i < NR * 2; i++) v[i] = 0.0;
  for (int row = B.blk_r0[b] + threadIdx.x; row < B.blk_r1[b]; row += kBB) {
#pragma unroll
    for (int k = 0; k < NR; k++) {
      if (!act[k]) continue;
      size_t i = (size_t)row * NR + k;
      x[i] += (XT)(a[k] * (double)p[i]);
      VT ri = r[i] - (VT)a[k] * q[i];
      r[i] = ri;
      VT zi = dinv[row] * ri;
      z[i] = zi;
      v[k * 2] += (double)ri * (double)zi;
      v[k * 2 + 1] += (double)ri * (double)ri;
    }
  }
  double out[NR * 2];
  if (prob_reduce<NR, 2>(v, B, b, pr, &B.ticket[pr * 4 + 2], out)) {
    for (int k = 0; k < NR; k++) {
      Seg &s = S[k];
      if (s.done) continue;
      s.beta = out[k * 2] / s.rz;
      s.rz = out[k * 2];
      s.rr = out[k * 2 + 1];
      s.iters += 1;
      if (out[k * 2 + 1] <= s.tol2) s.done = 1;
    }
  }
}

template <typename VT, int NR>
__global__ __launch_bounds__(kBB) void kb_pdir(BatchDev B, VT *__restrict__ p, const VT *__restrict__ z) {
  const int b = blockIdx.x, pr = B.blk_prob[b];
  const Seg *S = B.seg + pr * NR;
  if (all_done<NR>(S)) return;
  for (int row = B.blk_r0[b] + threadIdx.x; row < B.blk_r1[b]; row += kBB)
#pragma unroll
    for (int k = 0; k < NR; k++) {
      if (S[k].done) continue;
      size_t i = (size_t)row * NR + k;
      p[i] = z[i] + (VT)S[k].beta * p[i];
    }
}

// FP64 refinement residual r = f - K u per (row, rhs); per-segment ||r||^2,
// ||f||^2 and f.u.  Segments already converged in the outer loop keep theirs.
template <int NR>
__global__ __launch_bounds__(kBB) void kb_resid64(BatchDev B, const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const double *__restrict__ v, const double *__restrict__ u,
                                                  const double *__restrict__ f, double *__restrict__ r,
                                                  float *__restrict__ r32, const int *__restrict__ outer_done) {
  const int b = blockIdx.x, pr = B.blk_prob[b];
  double vals[NR * 3];
#pragma unroll
  for (int i = 0; i < NR * 3; i++) vals[i] = 0.0;
  for (int row = B.blk_r0[b] + threadIdx.x; row < B.blk_r1[b]; row += kBB) {
    const int rb = rp[row], re = rp[row + 1];
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) acc[k] = 0.0;
    for (int j = rb; j < re; j++) {
      const double vj = v[j];
      const size_t c = (size_t)ci[j] * NR;
#pragma unroll
      for (int k = 0; k < NR; k++) acc[k] += vj * u[c + k];
    }
#pragma unroll
    for (int k = 0; k < NR; k++) {
      size_t i = (size_t)row * NR + k;
      double ri = (rb == re) ? 0.0 : f[i] - acc[k];
      bool od = outer_done && outer_done[pr * NR + k];
      r[i] = od ? 0.0 : ri;
      if (r32) r32[i] = od ? 0.0f : (float)ri;
      vals[k * 3] += ri * ri;
      vals[k * 3 + 1] += f[i] * f[i];
      vals[k * 3 + 2] += f[i] * u[i];
    }
  }
  double out[NR * 3];
  if (prob_reduce<NR, 3>(vals, B, b, pr, &B.ticket[pr * 4 + 3], out)) {
    for (int k = 0; k < NR; k++) {
      Seg &s = B.seg[pr * NR + k];
      s.rr = out[k * 3];
      s.ff = out[k * 3 + 1];
      s.fu = out[k * 3 + 2];
    }
  }
}

__global__ void kb_add(int64_t n, double *u, const float *d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] += (double)d[i];
}

// Jacobi for the batched system (shared by all rhs of a row).
template <typename VT>
__global__ void kb_dinv(int n, const int *__restrict__ rp, const int *__restrict__ ci, const VT *__restrict__ v,
                        const unsigned char *__restrict__ fixed, VT *__restrict__ dinv, int *bad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int rb = rp[i], re = rp[i + 1];
  int lo = rb, hi = re;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ci[mid] < i) lo = mid + 1; else hi = mid;
  }
  VT d = (lo < re && ci[lo] == i) ? v[lo] : (VT)0;
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

__global__ void kb_set_done(Seg *seg, int nseg, const int *flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nseg) {
    seg[i].done = flags[i];
    seg[i].iters = 0;
  }
}

static int batch_setup(tfm_ctx *ctx, int P, const int64_t *dof_off, int NR, BatchDev &B) {
  std::vector<int> bp, b0, b1, pb0(P), pnb(P);
  for (int p = 0; p < P; p++) {
    int64_t r0 = dof_off[p], r1 = dof_off[p + 1];
    pb0[p] = (int)bp.size();
    int64_t nb = std::max<int64_t>(1, (r1 - r0 + kRowsPerBlk - 1) / kRowsPerBlk);
    for (int64_t q = 0; q < nb; q++) {
      bp.push_back(p);
      b0.push_back((int)std::min(r1, r0 + q * kRowsPerBlk));
      b1.push_back((int)std::min(r1, r0 + (q + 1) * kRowsPerBlk));
    }
    pnb[p] = (int)nb;
  }
  B.P = P;
  B.NR = NR;
  B.nblk = (int)bp.size();
  int rc;
  if ((rc = dalloc(ctx, &B.blk_prob, B.nblk)) || (rc = dalloc(ctx, &B.blk_r0, B.nblk)) ||
      (rc = dalloc(ctx, &B.blk_r1, B.nblk)) || (rc = dalloc(ctx, &B.prob_blk0, P)) ||
      (rc = dalloc(ctx, &B.prob_nblk, P)) || (rc = dalloc(ctx, &B.ticket, (size_t)P * 4)) ||
      (rc = dalloc(ctx, &B.part, (size_t)B.nblk * NR * 3)) || (rc = dalloc(ctx, &B.seg, (size_t)P * NR)))
    return rc;
  hipStream_t st = ctx->stream;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.blk_prob, bp.data(), sizeof(int) * B.nblk, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.blk_r0, b0.data(), sizeof(int) * B.nblk, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.blk_r1, b1.data(), sizeof(int) * B.nblk, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.prob_blk0, pb0.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.prob_nblk, pnb.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(B.ticket, 0, sizeof(unsigned) * P * 4, st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(B.seg, 0, sizeof(Seg) * P * NR, st));
  return TFM_OK;
}

static void batch_free(BatchDev &B) {
  dfree(B.blk_prob); dfree(B.blk_r0); dfree(B.blk_r1); dfree(B.prob_blk0); dfree(B.prob_nblk);
  dfree(B.ticket); dfree(B.part); dfree(B.seg);
}

template <typename VT, int NR>
static void launch_spmv(const BatchDev &B, hipStream_t st, int S, const int *rp, const int *ci, const VT *v,
                        const VT *p, VT *q) {
  dim3 g(B.nblk), t(kBB);
  if constexpr (NR > 1) {
    hipLaunchKernelGGL((kb_spmv<VT, NR, 1>), g, t, 0, st, B, rp, ci, v, p, q);
  } else {
    switch (S) {
    case 2: hipLaunchKernelGGL((kb_spmv<VT, 1, 2>), g, t, 0, st, B, rp, ci, v, p, q); break;
    case 4: hipLaunchKernelGGL((kb_spmv<VT, 1, 4>), g, t, 0, st, B, rp, ci, v, p, q); break;
    case 8: hipLaunchKernelGGL((kb_spmv<VT, 1, 8>), g, t, 0, st, B, rp, ci, v, p, q); break;
    case 16: hipLaunchKernelGGL((kb_spmv<VT, 1, 16>), g, t, 0, st, B, rp, ci, v, p, q); break;
    default: hipLaunchKernelGGL((kb_spmv<VT, 1, 32>), g, t, 0, st, B, rp, ci, v, p, q); break;
    }
  }
}

// Batched PCG from r (x = 0).  Stops per segment at ||r|| <= rtol ||r0||.
template <typename VT, typename XT, int NR>
static int bpcg(tfm_ctx *ctx, const BatchDev &B, int n, const int *rp, const int *ci, const VT *v, VT *r, VT *z,
                VT *p, VT *q, const VT *dinv, XT *x, double rtol, int max_iter, int S, std::vector<Seg> &hs) {
  hipStream_t st = ctx->stream;
  dim3 g(B.nblk), t(kBB);
  hipLaunchKernelGGL((kb_init<VT, NR>), g, t, 0, st, B, r, z, p, dinv, rtol * rtol);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  const int CH = 16;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  TFM_CHECK_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < CH; k++) {
    launch_spmv<VT, NR>(B, st, S, rp, ci, v, p, q);
    hipLaunchKernelGGL((kb_update<VT, XT, NR>), g, t, 0, st, B, x, r, z, p, q, dinv);
    hipLaunchKernelGGL((kb_pdir<VT, NR>), g, t, 0, st, B, p, z);
  }
  TFM_CHECK_HIP(ctx, hipStreamEndCapture(st, &graph));
  TFM_CHECK_HIP(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  const int nseg = B.P * B.NR;
  hs.resize(nseg);
  int rc = TFM_OK;
  for (int it = 0; it <= max_iter; it += CH) {
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(hs.data(), B.seg, sizeof(Seg) * nseg, hipMemcpyDeviceToHost, st));
    TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
    bool all = true;
    for (int s = 0; s < nseg; s++) all &= hs[s].done != 0;
    if (all || it >= max_iter) break;
    if (hipGraphLaunch(exec, st) != hipSuccess) {
      rc = fail(ctx, TFM_E_HIP, "hipGraphLaunch failed");
      break;
    }
  }
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hs.data(), B.seg, sizeof(Seg) * nseg, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  return rc;
}

template <int NR>
static int solve_batch_nr(tfm_system *s, int P, const int64_t *dof_off, const double *rhs,
                          const tfm_solve_opts &o, double *u, tfm_solve_stats *stats) {
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int n = (int)s->ndof;
  const size_t nv = (size_t)n * NR;
  const bool mixed = o.mode == TFM_SOLVE_MIXED;
  int S = 4;
  {
    double avg = s->n_free ? (double)s->nnz / (double)s->n_free : 1.0;
    S = avg <= 6 ? 2 : avg <= 12 ? 4 : avg <= 24 ? 8 : avg <= 48 ? 16 : 32;
  }
  BatchDev B;
  int rc = batch_setup(ctx, P, dof_off, NR, B);
  double *f = nullptr, *x = nullptr, *r64 = nullptr, *d64 = nullptr, *z64 = nullptr, *p64 = nullptr, *q64 = nullptr;
  float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr, *d32 = nullptr, *x32 = nullptr;
  int *bad = nullptr, *odone = nullptr;
  auto cleanup = [&]() {
    batch_free(B);
    dfree(f); dfree(x); dfree(r64); dfree(d64); dfree(z64); dfree(p64); dfree(q64);
    dfree(r32); dfree(z32); dfree(p32); dfree(q32); dfree(d32); dfree(x32); dfree(bad); dfree(odone);
  };
  if (rc || (rc = dalloc(ctx, &f, nv)) || (rc = dalloc(ctx, &x, nv)) || (rc = dalloc(ctx, &r64, nv)) ||
      (rc = dalloc(ctx, &d64, n)) || (rc = dalloc(ctx, &bad, 1)) || (rc = dalloc(ctx, &odone, P * NR))) {
    cleanup();
    return rc;
  }
  if (!mixed && ((rc = dalloc(ctx, &z64, nv)) || (rc = dalloc(ctx, &p64, nv)) || (rc = dalloc(ctx, &q64, nv)))) {
    cleanup();
    return rc;
  }
  if (mixed && ((rc = dalloc(ctx, &r32, nv)) || (rc = dalloc(ctx, &z32, nv)) || (rc = dalloc(ctx, &p32, nv)) ||
                (rc = dalloc(ctx, &q32, nv)) || (rc = dalloc(ctx, &d32, n)) || (rc = dalloc(ctx, &x32, nv)))) {
    cleanup();
    return rc;
  }
  // rhs: homogeneous-Dirichlet systems take f[fixed] = 0 (fred on the free DOFs)
  {
    std::vector<unsigned char> fx((size_t)n, 0);
    TFM_CHECK_HIP(ctx, hipMemcpy(fx.data(), s->d_fixed, n, hipMemcpyDeviceToHost));
    std::vector<double> hf(nv, 0.0);
    if (rhs) {
      for (int k = 0; k < NR; k++)
        for (int i = 0; i < n; i++) hf[(size_t)i * NR + k] = fx[i] ? 0.0 : rhs[(size_t)k * n + i];
    } else {
      std::vector<double> fr((size_t)n);
      TFM_CHECK_HIP(ctx, hipMemcpy(fr.data(), s->d_f, sizeof(double) * n, hipMemcpyDeviceToHost));
      for (int k = 0; k < NR; k++)
        for (int i = 0; i < n; i++) hf[(size_t)i * NR + k] = fr[i];
    }
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(f, hf.data(), sizeof(double) * nv, hipMemcpyHostToDevice, st));
  }
  TFM_CHECK_HIP(ctx, hipMemsetAsync(x, 0, sizeof(double) * nv, st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(bad, 0, sizeof(int), st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(odone, 0, sizeof(int) * P * NR, st));
  unsigned g1 = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(kb_dinv<double>, dim3(g1), dim3(256), 0, st, n, s->d_rp, s->d_ci, s->d_v64, s->d_fixed, d64, bad);
  if (mixed)
    hipLaunchKernelGGL(kb_dinv<float>, dim3(g1), dim3(256), 0, st, n, s->d_rp, s->d_ci, s->d_v32, s->d_fixed, d32, bad);
  int hbad = 0;
  TFM_CHECK_HIP(ctx, hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost));
  if (hbad) {
    cleanup();
    return fail(ctx, TFM_E_NUMERIC, "sparse numeric factorization/solve failed: singular matrix (a free DOF has no positive diagonal)");
  }
  const int nseg = P * NR;
  std::vector<Seg> hs(nseg);
  std::vector<int> iters(nseg, 0), cycles(nseg, 0), brk(nseg, 0), od(nseg, 0);
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev0, st));
  dim3 g(B.nblk), t(kBB);
  int total_it = 0;
  if (!mixed) {
    hipLaunchKernelGGL((kb_resid64<NR>), g, t, 0, st, B, s->d_rp, s->d_ci, s->d_v64, x, f, r64, (float *)nullptr,
                       (const int *)nullptr);
    rc = bpcg<double, double, NR>(ctx, B, n, s->d_rp, s->d_ci, s->d_v64, r64, z64, p64, q64, d64, x, o.rtol,
                                  o.max_iter, S, hs);
    for (int q = 0; q < nseg; q++) {
      iters[q] = hs[q].iters;
      brk[q] = hs[q].breakdown;
      cycles[q] = 1;
      od[q] = hs[q].done && !hs[q].breakdown;  // converged on its recurrence residual
    }
  } else {
    for (int cyc = 0; cyc < o.max_outer && !rc; cyc++) {
      TFM_CHECK_HIP(ctx, hipMemcpyAsync(odone, od.data(), sizeof(int) * nseg, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL((kb_resid64<NR>), g, t, 0, st, B, s->d_rp, s->d_ci, s->d_v64, x, f, r64, r32, odone);
      TFM_CHECK_HIP(ctx, hipMemcpyAsync(hs.data(), B.seg, sizeof(Seg) * nseg, hipMemcpyDeviceToHost, st));
      TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
      bool all = true;
      for (int q = 0; q < nseg; q++) {
        double rel = hs[q].ff > 0 ? std::sqrt(hs[q].rr / hs[q].ff) : 0.0;
        if (!od[q] && (rel <= o.rtol || iters[q] >= o.max_iter || brk[q])) od[q] = 1;
        all &= od[q] != 0;
      }
      if (all) break;
      TFM_CHECK_HIP(ctx, hipMemcpyAsync(odone, od.data(), sizeof(int) * nseg, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(kb_set_done, dim3((nseg + 255) / 256), dim3(256), 0, st, B.seg, nseg, odone);
      TFM_CHECK_HIP(ctx, hipMemsetAsync(x32, 0, sizeof(float) * nv, st));
      rc = bpcg<float, float, NR>(ctx, B, n, s->d_rp, s->d_ci, s->d_v32, r32, z32, p32, q32, d32, x32, o.inner_rtol,
                                  o.max_iter, S, hs);
      for (int q = 0; q < nseg; q++)
        if (!od[q]) {
          iters[q] += hs[q].iters;
          cycles[q] += 1;
          brk[q] |= hs[q].breakdown;
        }
      hipLaunchKernelGGL(kb_add, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, (int64_t)nv, x, x32);
    }
  }
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev1, st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(odone, 0, sizeof(int) * nseg, st));
  hipLaunchKernelGGL((kb_resid64<NR>), g, t, 0, st, B, s->d_rp, s->d_ci, s->d_v64, x, f, r64, (float *)nullptr,
                     (const int *)nullptr);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hs.data(), B.seg, sizeof(Seg) * nseg, hipMemcpyDeviceToHost, st));
  std::vector<double> hx(nv);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hx.data(), x, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  cleanup();
  if (rc) return rc;
  // u[k][*]: free entries from x, then Dirichlet values in bcs order
  bool worst_ok = true;
  for (int k = 0; k < NR; k++) {
    double *uk = u + (size_t)k * n;
    for (int i = 0; i < n; i++) uk[i] = 0.0;
    for (int64_t fdof : s->free_dofs) uk[fdof] = hx[(size_t)fdof * NR + k];
    for (size_t b = 0; b < s->bc_dof.size(); b++) uk[s->bc_dof[b]] = s->bc_val[b];
  }
  for (int q = 0; q < nseg; q++) {
    double rel = hs[q].ff > 0 ? std::sqrt(hs[q].rr / hs[q].ff) : 0.0;
    // plain: PCG converged on its recurrence residual (the true residual is
    // reported); mixed: the FP64 refinement residual itself met rtol
    worst_ok &= (mixed ? rel <= o.rtol : od[q] != 0) && !brk[q];
    total_it += iters[q];
    if (stats) {
      stats[q].iterations = iters[q];
      stats[q].outer = cycles[q];
      stats[q].rel_residual = rel;
      stats[q].seconds = ms * 1e-3;
    }
  }
  (void)total_it;
  if (!worst_ok) return fail(ctx, TFM_E_NOT_CONVERGED, "a batched segment missed the tolerance");
  return TFM_OK;
}

}  // namespace tfm

using namespace tfm;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
__global__ void kb_box_batch(double *coords, int *conn, const long long *desc, int P) {
  // desc per problem: nx, ny, nz, node_off, elem_off
  int p = blockIdx.y;
  if (p >= P) return;
  const long long *d = desc + 5 * p;
  long long nx = d[0], ny = d[1], nz = d[2], noff = d[3], eoff = d[4];
  long long nnod = (nx + 1) * (ny + 1) * (nz + 1), nbr = nx * ny * nz;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < nnod; id += (long long)gridDim.x * blockDim.x) {
    long long k = id % (nz + 1), j = (id / (nz + 1)) % (ny + 1), i = id / ((nz + 1) * (ny + 1));
    double t = (double)i / (double)nx, u = (double)j / (double)ny, w = (double)k / (double)nz;
    coords[3 * (noff + id) + 0] = 0.0 + t * (1.0 - 0.0);
    coords[3 * (noff + id) + 1] = 0.0 + u * (1.0 - 0.0);
    coords[3 * (noff + id) + 2] = 0.0 + w * (1.0 - 0.0);
  }
  const int OFF[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const int TETS[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
  for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b < nbr; b += (long long)gridDim.x * blockDim.x) {
    long long k = b % nz, j = (b / nz) % ny, i = b / (nz * ny);
    int ids[8];
    for (int c = 0; c < 8; c++)
      ids[c] = (int)(noff + ((i + OFF[c][0]) * (ny + 1) + (j + OFF[c][1])) * (nz + 1) + (k + OFF[c][2]));
    for (int tt = 0; tt < 6; tt++) {
      int *o = conn + 4 * (eoff + 6 * b + tt);
      for (int a = 0; a < 4; a++) o[a] = ids[TETS[tt][a]];
    }
  }
}

// P unit-cube box meshes [0,1]^3 with per-problem (nx, ny, nz), concatenated
// (nodes then elements in problem order).  The unit-cube brick orientation is
// positive for every (nx, ny, nz), so no tet needs the 1<->2 swap.
extern "C" int tfm_mesh_box_batch(tfm_ctx *ctx, int64_t P, const int64_t *n3, tfm_mesh **out, int64_t *node_off,
                                  int64_t *elem_off) {
  if (!ctx || !out || !n3 || P <= 0) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  std::vector<long long> desc(5 * P);
  int64_t nn = 0, ne = 0;
  for (int64_t p = 0; p < P; p++) {
    int64_t nx = n3[3 * p], ny = n3[3 * p + 1], nz = n3[3 * p + 2];
    if (nx < 1 || ny < 1 || nz < 1) return fail(ctx, TFM_E_INVALID, "batched box needs n >= 1 per axis");
    desc[5 * p] = nx; desc[5 * p + 1] = ny; desc[5 * p + 2] = nz; desc[5 * p + 3] = nn; desc[5 * p + 4] = ne;
    if (node_off) node_off[p] = nn;
    if (elem_off) elem_off[p] = ne;
    nn += (nx + 1) * (ny + 1) * (nz + 1);
    ne += 6 * nx * ny * nz;
  }
  if (node_off) node_off[P] = nn;
  if (elem_off) elem_off[P] = ne;
  std::vector<double> none;
  tfm_mesh *m = nullptr;
  // allocate through the generic constructor path (sizes only)
  m = new tfm_mesh();
  m->ctx = ctx; m->cell = TFM_TET; m->nn = 4; m->pdim = 3; m->dim = 3; m->n_nodes = nn; m->n_elem = ne;
  if (nn >= INT32_MAX || ne * 8 >= INT32_MAX) { delete m; return fail(ctx, TFM_E_UNSUPPORTED, "batch too large for int32 indexing"); }
  long long *d_desc = nullptr;
  if ((rc = dalloc(ctx, &m->d_coords, nn * 3)) || (rc = dalloc(ctx, &m->d_conn, ne * 4)) ||
      (rc = dalloc(ctx, &d_desc, 5 * P))) {
    tfm_mesh_destroy(m);
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(d_desc, desc.data(), sizeof(long long) * 5 * P, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(kb_box_batch, dim3(64, (unsigned)P), dim3(256), 0, ctx->stream, m->d_coords, m->d_conn, d_desc, (int)P);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  dfree(d_desc);
  if (e != hipSuccess) {
    tfm_mesh_destroy(m);
    return fail(ctx, TFM_E_HIP, "batched box kernel failed");
  }
  *out = m;
  return TFM_OK;
}

// P copies of the topopt Grid (one per SIMP design), concatenated.
__global__ void kb_grid_batch(double *coords, int *conn, int64_t nx, int64_t ny, double h, int P) {
  int64_t nn = nx * ny, ne = (nx - 1) * (ny - 1);
  int p = blockIdx.y;
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < nn || id < ne;
       id += (int64_t)gridDim.x * blockDim.x) {
    if (id < nn) {
      int64_t i = id % nx, j = id / nx;
      coords[2 * (p * nn + id)] = (double)i * h;
      coords[2 * (p * nn + id) + 1] = (double)j * h;
    }
    if (id < ne) {
      int64_t ei = id % (nx - 1), ej = id / (nx - 1);
      int *o = conn + 4 * (p * ne + id);
      int off = (int)(p * nn);
      o[0] = off + (int)(ej * nx + ei);
      o[1] = off + (int)(ej * nx + ei + 1);
      o[2] = off + (int)((ej + 1) * nx + ei + 1);
      o[3] = off + (int)((ej + 1) * nx + ei);
    }
  }
}

extern "C" int tfm_mesh_grid_batch(tfm_ctx *ctx, int64_t P, int64_t nx, int64_t ny, double h, tfm_mesh **out) {
  if (!ctx || !out || P <= 0 || nx < 2 || ny < 2) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  int64_t nn = P * nx * ny, ne = P * (nx - 1) * (ny - 1);
  if (nn >= INT32_MAX || ne * 8 >= INT32_MAX) return fail(ctx, TFM_E_UNSUPPORTED, "batch too large for int32 indexing");
  tfm_mesh *m = new tfm_mesh();
  m->ctx = ctx; m->cell = TFM_QUAD; m->nn = 4; m->pdim = 2; m->dim = 2; m->n_nodes = nn; m->n_elem = ne;
  if ((rc = dalloc(ctx, &m->d_coords, nn * 2)) || (rc = dalloc(ctx, &m->d_conn, ne * 4))) {
    tfm_mesh_destroy(m);
    return rc;
  }
  hipLaunchKernelGGL(kb_grid_batch, dim3(64, (unsigned)P), dim3(256), 0, ctx->stream, m->d_coords, m->d_conn, nx, ny, h,
                     (int)P);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    tfm_mesh_destroy(m);
    return fail(ctx, TFM_E_HIP, "batched grid kernel failed");
  }
  *out = m;
  return TFM_OK;
}

extern "C" int tfm_system_solve_batch(tfm_system *s, int64_t P, const int64_t *dof_off, int nrhs, const double *rhs,
                                      const tfm_solve_opts *opts, double *u, tfm_solve_stats *stats) {
  if (!s || !dof_off || !u || P <= 0) return TFM_E_INVALID;
  tfm_ctx *ctx = s->ctx;
  if (nrhs != 1 && nrhs != 8) return fail(ctx, TFM_E_UNSUPPORTED, "batched solve supports 1 or 8 right-hand sides");
  if (dof_off[0] != 0 || dof_off[P] != s->ndof) return fail(ctx, TFM_E_INVALID, "dof offsets must cover the system");
  for (int64_t p = 0; p < P; p++)
    if (dof_off[p + 1] < dof_off[p]) return fail(ctx, TFM_E_INVALID, "dof offsets must ascend");
  if (nrhs > 1 && !rhs) return fail(ctx, TFM_E_INVALID, "several right-hand sides need rhs");
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-10, 20000, 1e-4, 60};
  if (opts) o = *opts;
  if (o.mode == TFM_SOLVE_MIXED && !s->d_v32) return fail(ctx, TFM_E_INVALID, "mixed solve needs FP32 values");
  if (!s->d_v64) return fail(ctx, TFM_E_INVALID, "solve needs FP64 values");
  // A segment's rows must not couple to another problem's rows (block
  // diagonal): the kernels reduce per problem and would silently mix them.
  if (nrhs == 1) return solve_batch_nr<1>(s, (int)P, dof_off, rhs, o, u, stats);
  return solve_batch_nr<8>(s, (int)P, dof_off, rhs, o, u, stats);
}
