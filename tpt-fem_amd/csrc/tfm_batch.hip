// tfm_batch.hip -- many independent problems in one launch.
//
// Batching form: P independent meshes (uneven box meshes, or SIMP designs on
// one grid) are concatenated into one mesh, so one assembly pass fills one
// block-diagonal reduced system.  The solver then runs P x NR independent
// Jacobi-PCG iterations at once (NR right-hand sides per problem, vectors
// interleaved [dof][rhs]) on a node-blocked sliced copy of the matrix whose
// slices never straddle two problems; every block of a launch works on one
// problem.  Dot products are reduced per problem without fences: each block
// writes its partials, and every block of the next kernel re-sums its
// problem's partials in the same fixed order, so all of a problem's blocks
// agree on alpha, beta and convergence.  Each (problem, rhs) segment has its
// own scalars and stops on its own; a problem whose segments are all done
// skips its blocks.  Iterations replay from a captured hipGraph.  The
// precision form wraps FP32 segments in per-segment FP64 refinement.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <type_traits>

#include "tfm_bsell.h"
#include "tfm_elem.h"
#include "tfm_internal.h"
#include "tfm_mg.h"

namespace tfm {

constexpr int kBW = 4;         // waves (= slices) per block
constexpr double kFloor2 = 1e-12;  // precision form: FP32 residual floor (squared, relative to the last replacement)
constexpr int kBT = 64 * kBW;  // threads per block
constexpr int kScalarBlocks = 512;  // blocks of one problem beyond which its CG scalars are formed once
// TFM_SCALAR_BLOCKS=<n> moves that threshold (both modes give the same values).
static int scalar_blocks() {
  const char *e = getenv("TFM_SCALAR_BLOCKS");
  return e ? atoi(e) : kScalarBlocks;
}

// Partial-sum slots, each [nblk][NR]: p.q; r.z and r.r by iteration parity;
// refinement residual ||f - K u||^2, ||f||^2, f.u.
enum { Q_PQ = 0, Q_RZ0, Q_RZ1, Q_RR0, Q_RR1, Q_RES, Q_FF, Q_FU, Q_N };

struct Seg {
  double rr, rr0;          // latest recurrence ||r||^2 and the initial one
  double rt, ff, fu;       // refinement residual: ||f - K u||^2, ||f||^2, f.u
  double beta;             // scalar mode: this iteration's beta
  double alpha;            // scalar mode: this iteration's alpha
  int done, iters, breakdown, pad;
};

// Device view of the batch (passed by value).
struct BView {
  int nblk, bs;
  const int *blk_prob, *blk_s0, *blk_s1;  // block -> problem, slice range
  const int *prob_blk0, *prob_nblk;        // problem -> blocks
  const int *sb, *soff, *bcol;             // slices (block rows sb[s]..sb[s+1]) and stored blocks
  double *part;                            // [Q_N][nblk][NR]
  Seg *seg;                                // [P][NR]
  int *it;                                 // iteration counter
  int max_it;
  double rtol2;
};

// Sum over the block of NR per-thread values, returned to every thread
// (fixed order).
template <int NR>
__device__ __forceinline__ void bsum(double (&v)[NR]) {
  __shared__ double sh[NR][kBW];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NR; k++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  __syncthreads();
  if (l == 0)
#pragma unroll
    for (int k = 0; k < NR; k++) sh[k][w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < kBW; q++) t += sh[k][q];
    v[k] = t;
  }
}

// A problem's total of one slot, identical in every block of that problem:
// each thread takes a fixed strided subset of the problem's block partials,
// then the fixed block tree.  The partials come from an earlier kernel, so no
// fence is needed.
template <int NR>
__device__ __forceinline__ void prob_sum(const BView &V, int slot, int p, double (&out)[NR]) {
  const int b0 = V.prob_blk0[p], nb = V.prob_nblk[p];
  const double *q = V.part + ((size_t)slot * V.nblk + b0) * NR;
#pragma unroll
  for (int k = 0; k < NR; k++) out[k] = 0.0;
  // unrolled so the loads issue ahead of the adds, whose order is unchanged
#pragma unroll 8
  for (int i = threadIdx.x; i < nb; i += kBT)
#pragma unroll
    for (int k = 0; k < NR; k++) out[k] += q[(size_t)i * NR + k];
  bsum<NR>(out);
}

template <int NR>
__device__ __forceinline__ void put_part(const BView &V, int slot, double (&v)[NR]) {
  bsum<NR>(v);
  if (threadIdx.x < NR) {
    double t = v[0];
#pragma unroll
    for (int k = 1; k < NR; k++)
      if ((int)threadIdx.x == k) t = v[k];
    V.part[((size_t)slot * V.nblk + blockIdx.x) * NR + threadIdx.x] = t;
  }
}

// Thread-strided element range of a block: the interleaved entries
// [row0*NR, row1*NR) of the DOF rows of its slices.  With kBT a multiple of
// NR a thread always sees the same rhs (threadIdx.x % NR).
__device__ __forceinline__ void blk_rows(const BView &V, int b, int &r0, int &r1) {
  r0 = V.sb[V.blk_s0[b]] * V.bs;
  r1 = V.sb[V.blk_s1[b]] * V.bs;
}

// Sum over the block of one value per thread for rhs threadIdx.x % NR.
template <int NR>
__device__ __forceinline__ void rhs_sum(double x, double (&out)[NR]) {
#pragma unroll
  for (int k = 0; k < NR; k++) out[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NR; k++)
    if ((int)(threadIdx.x % NR) == k) out[k] = x;
}

// z = Dinv r (dinv null: z is given), both direction buffers zero, partial
// r.z and r.r into the parity-1 slots (they play "iteration -1").
template <typename VT, int NR>
__global__ __launch_bounds__(kBT) void kbx_init(BView V, const VT *__restrict__ r, VT *__restrict__ z,
                                                VT *__restrict__ p0, VT *__restrict__ p1, const VT *__restrict__ dinv) {
  int r0, r1;
  blk_rows(V, blockIdx.x, r0, r1);
  double rz = 0.0, rr = 0.0;
  for (int i = r0 * NR + threadIdx.x; i < r1 * NR; i += kBT) {
    const VT zi = dinv ? dinv[i / NR] * r[i] : z[i];
    if (dinv) z[i] = zi;
    p0[i] = 0;
    p1[i] = 0;
    rz += (double)r[i] * (double)zi;
    rr += (double)r[i] * (double)r[i];
  }
  double a[NR], c[NR];
  rhs_sum<NR>(rz, a);
  rhs_sum<NR>(rr, c);
  put_part<NR>(V, Q_RZ1, a);
  put_part<NR>(V, Q_RR1, c);
}

// Iteration i (parity e): convergence of each segment, beta, p_i = z_i +
// beta p_{i-1} on the fly, q = K p_i on the node-blocked slices (one lane per
// block row, NR right-hand sides at once), partial p.q.
template <typename VT, int BS, int NR, bool SC>
__global__ __launch_bounds__(kBT) void kbx_dir(BView V, const VT *__restrict__ bv, const VT *__restrict__ z,
                                               const VT *__restrict__ pold, VT *__restrict__ pnew,
                                               VT *__restrict__ q, int e) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  Seg *S = V.seg + (size_t)pr * NR;
  VT beta[NR];
  if constexpr (SC) {
    // scalar mode: kbx_beta has decided the stop test and beta
    __shared__ int s_any;
    __shared__ double s_beta[NR];
    if (threadIdx.x == 0) {
      int a = 0;
      for (int kk = 0; kk < NR; kk++) a |= !S[kk].done;
      s_any = a;
    }
    if (threadIdx.x < NR) s_beta[threadIdx.x] = S[threadIdx.x].done ? 0.0 : S[threadIdx.x].beta;
    __syncthreads();
    if (!s_any) return;  // uniform per block
#pragma unroll
    for (int k = 0; k < NR; k++) beta[k] = (VT)s_beta[k];
  } else {
  __shared__ int s_it, s_done[NR];
  if (threadIdx.x == 0) s_it = *V.it;
  if (threadIdx.x < NR) s_done[threadIdx.x] = S[threadIdx.x].done;
  __syncthreads();
  const int it = s_it;
  double rr[NR], rzn[NR], rzo[NR];
  prob_sum<NR>(V, e ? Q_RR0 : Q_RR1, pr, rr);
  prob_sum<NR>(V, e ? Q_RZ0 : Q_RZ1, pr, rzn);
  prob_sum<NR>(V, e ? Q_RZ1 : Q_RZ0, pr, rzo);
  const bool first = b == V.prob_blk0[pr] && threadIdx.x == 0;
  bool any = false;
#pragma unroll
  for (int k = 0; k < NR; k++) {
    // rr0: the reference of the stop test, set before the run (||f||^2) or,
    // when negative, taken from the first residual seen (the start, or the
    // one a residual replacement installed)
    const bool fresh = S[k].rr0 < 0.0;
    const double rr0 = fresh ? rr[k] : S[k].rr0;
    const bool d = s_done[k] || rr0 == 0.0 || rr[k] <= V.rtol2 * rr0 || it >= V.max_it;
    any |= !d;
    beta[k] = (VT)((it > 0 && !d && rzo[k] > 0.0) ? rzn[k] / rzo[k] : 0.0);
    if (first && !s_done[k]) {
      if (fresh) S[k].rr0 = rr[k];
      S[k].rr = rr[k];
      if (d) {
        S[k].done = 1;
        S[k].iters = it;
      }
    }
  }
  if (!any) return;  // uniform per block
  }
  constexpr int BB = BS * BS;
  const int lane = threadIdx.x & 63;
  const int s = V.blk_s0[b] + (threadIdx.x >> 6);
  double dot[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) dot[k] = 0.0;
  if (s < V.blk_s1[b]) {
    const int br = V.sb[s] + lane;
    const int o = V.soff[s], w = (V.soff[s + 1] - o) >> 6;
    const int *cp = V.bcol + o + lane;
    const VT *vp = bv + (int64_t)o * BB + lane;
    VT acc[BS][NR];
#pragma unroll
    for (int i = 0; i < BS; i++)
#pragma unroll
      for (int k = 0; k < NR; k++) acc[i][k] = 0;
    for (int kk = 0; kk < w; kk++) {
      const int64_t c = (int64_t)cp[kk * 64] * BS * NR;
      VT xv[BS][NR];
#pragma unroll
      for (int j = 0; j < BS; j++)
#pragma unroll
        for (int k = 0; k < NR; k++) xv[j][k] = z[c + j * NR + k] + beta[k] * pold[c + j * NR + k];
      const VT *vk = vp + (int64_t)kk * (64 * BB);
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int j = 0; j < BS; j++) {
          const VT a = vk[(i * BS + j) * 64];
#pragma unroll
          for (int k = 0; k < NR; k++) acc[i][k] += a * xv[j][k];
        }
    }
    if (br < V.sb[s + 1]) {
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int k = 0; k < NR; k++) {
          const int64_t idx = ((int64_t)br * BS + i) * NR + k;
          const VT pn = z[idx] + beta[k] * pold[idx];
          pnew[idx] = pn;
          q[idx] = acc[i][k];
          dot[k] += (double)pn * (double)acc[i][k];
        }
    }
  }
  put_part<NR>(V, Q_PQ, dot);
}

// Multigrid mode splits kbx_dir in three.  kbx_beta, one block per problem:
// the stop test and beta of each of its segments (into the Seg).
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_beta(BView V, int e) {
  const int pr = blockIdx.x;
  Seg *S = V.seg + (size_t)pr * NR;
  const int it = *V.it;
  double rr[NR], rzn[NR], rzo[NR];
  prob_sum<NR>(V, e ? Q_RR0 : Q_RR1, pr, rr);
  prob_sum<NR>(V, e ? Q_RZ0 : Q_RZ1, pr, rzn);
  prob_sum<NR>(V, e ? Q_RZ1 : Q_RZ0, pr, rzo);
  if (threadIdx.x != 0) return;
  for (int k = 0; k < NR; k++) {
    if (S[k].done) continue;
    const bool fresh = S[k].rr0 < 0.0;
    const double rr0 = fresh ? rr[k] : S[k].rr0;
    const bool d = rr0 == 0.0 || rr[k] <= V.rtol2 * rr0 || it >= V.max_it;
    S[k].beta = (it > 0 && !d && rzo[k] > 0.0) ? rzn[k] / rzo[k] : 0.0;
    if (fresh) S[k].rr0 = rr[k];
    S[k].rr = rr[k];
    if (d) {
      S[k].done = 1;
      S[k].iters = it;
    }
  }
}

// Scalar mode, one block per problem: alpha = r.z / p.q of each live
// segment (the same fixed-order sums the update's blocks would form), a
// breakdown (p.q <= 0) ends the segment.
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_alpha(BView V, int e) {
  const int pr = blockIdx.x;
  Seg *S = V.seg + (size_t)pr * NR;
  double pq[NR], rz[NR];
  prob_sum<NR>(V, Q_PQ, pr, pq);
  prob_sum<NR>(V, e ? Q_RZ0 : Q_RZ1, pr, rz);
  if (threadIdx.x != 0) return;
  for (int k = 0; k < NR; k++) {
    S[k].alpha = 0.0;
    if (S[k].done) continue;
    if (!(pq[k] > 0.0)) {
      S[k].breakdown = 1;
      S[k].done = 1;
      S[k].iters = *V.it + 1;  // as the update's counter reads once advanced
    } else {
      S[k].alpha = rz[k] / pq[k];
    }
  }
}

// kbx_dir_p: p = z + beta pold on the block's rows.
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_dir_p(BView V, const double *__restrict__ z,
                                                 const double *__restrict__ pold, double *__restrict__ pnew) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  const Seg *S = V.seg + (size_t)pr * NR;
  __shared__ int s_any;
  __shared__ double s_beta[NR];
  if (threadIdx.x == 0) {
    int a = 0;
    for (int kk = 0; kk < NR; kk++) a |= !S[kk].done;
    s_any = a;
  }
  if (threadIdx.x < NR) s_beta[threadIdx.x] = S[threadIdx.x].done ? 0.0 : S[threadIdx.x].beta;
  __syncthreads();
  if (!s_any) return;  // uniform per block
  const double bk = s_beta[threadIdx.x % NR];
  int r0, r1;
  blk_rows(V, b, r0, r1);
  for (int i = r0 * NR + threadIdx.x; i < r1 * NR; i += kBT) pnew[i] = z[i] + bk * pold[i];
}

// kbx_q_mg: q = K p matrix-free on SIMP grids (element moduli times the unit
// stiffness) and the partial p.q; rows where `mask` (the finest Jacobi
// diagonal inverse) is 0 -- the fixed DOFs -- get q = 0.  A block's rows are
// whole nodes of one design.  (Forming p here from z and pold at every
// neighbour instead of in kbx_dir_p doubles the loads and measured slower.)
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_q_mg(BView V, MgOp o, const double *__restrict__ mask,
                                                const double *__restrict__ pv, double *__restrict__ q) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  const Seg *S = V.seg + (size_t)pr * NR;
  __shared__ int s_any;
  if (threadIdx.x == 0) {
    int a = 0;
    for (int kk = 0; kk < NR; kk++) a |= !S[kk].done;
    s_any = a;
  }
  __syncthreads();
  if (!s_any) return;  // uniform per block, like kbx_dir_p's exit
  const int k = threadIdx.x % NR;
  int r0, r1;
  blk_rows(V, b, r0, r1);
  const int nn = o.nx * o.ny;
  double dot = 0.0;
  for (int idx = (r0 >> 1) * NR + threadIdx.x; idx < (r1 >> 1) * NR; idx += kBT) {
    const int node = idx / NR;
    const int p = node / nn, n = node - p * nn;
    double y[2];
    apply_node<true>(o, p, n % o.nx, n / o.nx, [&](int d) { return pv[d * NR + k]; }, y);
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const int g = 2 * node + a;
      const double qa = mask[g] == 0.0 ? 0.0 : y[a];
      q[g * NR + k] = qa;
      dot += pv[g * NR + k] * qa;
    }
  }
  double c[NR];
  rhs_sum<NR>(dot, c);
  put_part<NR>(V, Q_PQ, c);
}

// kbx_smooth_rz: the V-cycle's last step on the finest level, z = y + w Dinv
// (r - K y) (y the coarse-corrected start), fused with the partial r.z of
// the CG; the segments already done are skipped.
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_smooth_rz(BView V, MgOp o, const double *__restrict__ dinv,
                                                     const double *__restrict__ r, const double *__restrict__ yv,
                                                     double *__restrict__ z, int slot) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  const Seg *S = V.seg + (size_t)pr * NR;
  __shared__ int s_done[NR];
  if (threadIdx.x < NR) s_done[threadIdx.x] = S[threadIdx.x].done;
  __syncthreads();
  const int k = threadIdx.x % NR;
  const bool act = !s_done[k];
  int r0, r1;
  blk_rows(V, b, r0, r1);
  const int nn = o.nx * o.ny;
  double dot = 0.0;
  if (act)
    for (int idx = (r0 >> 1) * NR + threadIdx.x; idx < (r1 >> 1) * NR; idx += kBT) {
      const int node = idx / NR;
      const int p = node / nn, n = node - p * nn;
      double y[2];
      apply_node<true>(o, p, n % o.nx, n / o.nx, [&](int d) { return yv[d * NR + k]; }, y);
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int g = 2 * node + a;
        const double ra = r[g * NR + k];
        const double za = yv[g * NR + k] + kMgOmega * dinv[g] * (ra - y[a]);
        z[g * NR + k] = za;
        dot += ra * za;
      }
    }
  double c[NR];
  rhs_sum<NR>(dot, c);
  put_part<NR>(V, slot, c);
}

// alpha = r.z / p.q per segment; x += alpha p, r -= alpha q, z = Dinv r on
// the active segments; partial r.z, r.r.  Block 0 advances the counter.  With
// dinv null z is left to a separate preconditioner and kbx_rz.  SC (scalar
// mode): kbx_alpha has formed alpha and marked breakdowns.
template <typename VT, typename XT, int NR, bool SC = false>
__global__ __launch_bounds__(kBT) void kbx_update(BView V, XT *__restrict__ x, VT *__restrict__ r,
                                                  VT *__restrict__ z, const VT *__restrict__ p,
                                                  const VT *__restrict__ q, const VT *__restrict__ dinv, int e) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  Seg *S = V.seg + (size_t)pr * NR;
  if (b == 0 && threadIdx.x == 0) *V.it = *V.it + 1;
  __shared__ int s_done[NR];
  if (threadIdx.x < NR) s_done[threadIdx.x] = S[threadIdx.x].done;
  __syncthreads();
  bool any = false;
#pragma unroll
  for (int k = 0; k < NR; k++) any |= !s_done[k];
  if (!any) return;  // uniform per block
  double al[NR];
  if constexpr (SC) {
#pragma unroll
    for (int k = 0; k < NR; k++) al[k] = s_done[k] ? 0.0 : S[k].alpha;
  } else {
  double pq[NR], rz[NR];
  prob_sum<NR>(V, Q_PQ, pr, pq);
  prob_sum<NR>(V, e ? Q_RZ0 : Q_RZ1, pr, rz);
  const bool first = b == V.prob_blk0[pr] && threadIdx.x == 0;
#pragma unroll
  for (int k = 0; k < NR; k++) {
    al[k] = 0.0;
    if (s_done[k]) continue;
    if (!(pq[k] > 0.0)) {
      if (first) {
        S[k].breakdown = 1;
        S[k].done = 1;
        S[k].iters = *V.it;
      }
    } else {
      al[k] = rz[k] / pq[k];
    }
  }
  }
  const int k = threadIdx.x % NR;
  double a = al[0];
#pragma unroll
  for (int kk = 1; kk < NR; kk++)
    if (k == kk) a = al[kk];
  int r0, r1;
  blk_rows(V, b, r0, r1);
  double zr = 0.0, rr = 0.0;
  if (a != 0.0) {
    const VT af = (VT)a;
    for (int i = r0 * NR + threadIdx.x; i < r1 * NR; i += kBT) {
      x[i] += (XT)(a * (double)p[i]);
      const VT ri = r[i] - af * q[i];
      r[i] = ri;
      if (dinv) {
        const VT zi = dinv[i / NR] * ri;
        z[i] = zi;
        zr += (double)ri * (double)zi;
      }
      rr += (double)ri * (double)ri;
    }
  }
  double c1[NR], c2[NR];
  rhs_sum<NR>(zr, c1);
  rhs_sum<NR>(rr, c2);
  put_part<NR>(V, e ? Q_RZ1 : Q_RZ0, c1);
  put_part<NR>(V, e ? Q_RR1 : Q_RR0, c2);
}

// Partial r.z into `slot` after an external preconditioner wrote z.
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_rz(BView V, const double *__restrict__ r, const double *__restrict__ z,
                                              int slot) {
  int r0, r1;
  blk_rows(V, blockIdx.x, r0, r1);
  double s = 0.0;
  for (int i = r0 * NR + threadIdx.x; i < r1 * NR; i += kBT) s += r[i] * z[i];
  double a[NR];
  rhs_sum<NR>(s, a);
  put_part<NR>(V, slot, a);
}

// FP64 residual r = f - K u (CSR, fixed order) over the block's rows and
// partial ||r||^2, ||f||^2, f.u; segments flagged in `keep` get r = 0 (their
// refinement is over); the FP32 copy is refreshed for the segments in `sel`.
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_resid64(BView V, const int *__restrict__ rp, const int *__restrict__ ci,
                                                   const double *__restrict__ v, const double *__restrict__ u,
                                                   const double *__restrict__ f, double *__restrict__ r,
                                                   float *__restrict__ r32, const int *__restrict__ keep,
                                                   const int *__restrict__ sel) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  int r0, r1;
  blk_rows(V, b, r0, r1);
  const int k = threadIdx.x % NR;
  const bool kept = keep && keep[pr * NR + k];
  // the FP32 copy is written only for the selected segments (all when sel is null)
  const bool w32 = r32 && (!sel || sel[pr * NR + k]);
  double rr = 0.0, ff = 0.0, fu = 0.0;
  for (int i = r0 * NR + threadIdx.x; i < r1 * NR; i += kBT) {
    const int row = i / NR;
    const int rb = rp[row], re = rp[row + 1];
    double acc = 0.0;
    for (int j = rb; j < re; j++) acc += v[j] * u[(size_t)ci[j] * NR + k];
    const double ri = rb == re ? 0.0 : f[i] - acc;
    r[i] = kept ? 0.0 : ri;
    if (w32) r32[i] = kept ? 0.0f : (float)ri;
    rr += ri * ri;
    ff += f[i] * f[i];
    fu += f[i] * u[i];
  }
  double a[NR], c[NR], d[NR];
  rhs_sum<NR>(rr, a);
  rhs_sum<NR>(ff, c);
  rhs_sum<NR>(fu, d);
  put_part<NR>(V, Q_RES, a);
  put_part<NR>(V, Q_FF, c);
  put_part<NR>(V, Q_FU, d);
}

// kbx_resid64 on the node-blocked sliced copy (one lane per block row, NR
// right-hand sides at once), the same outputs; empty (fixed) rows have zero
// blocks and f = 0 there, so r = 0 as in the CSR form.
template <int BS, int NR>
__global__ __launch_bounds__(kBT) void kbx_resid_bs(BView V, const double *__restrict__ bv,
                                                    const double *__restrict__ u, const double *__restrict__ f,
                                                    double *__restrict__ r, float *__restrict__ r32,
                                                    const int *__restrict__ keep, const int *__restrict__ sel) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  if (sel) {
    // a problem with no selected segment keeps its partials (and so the
    // residual its segments were last replaced with)
    bool any = false;
#pragma unroll
    for (int k = 0; k < NR; k++) any |= sel[pr * NR + k] != 0;
    if (!any) return;
  }
  constexpr int BB = BS * BS;
  const int lane = threadIdx.x & 63;
  const int s = V.blk_s0[b] + (threadIdx.x >> 6);
  double rr[NR], ff[NR], fu[NR];
  bool kept[NR], w32[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    rr[k] = ff[k] = fu[k] = 0.0;
    kept[k] = keep && keep[pr * NR + k];
    w32[k] = r32 && (!sel || sel[pr * NR + k]);
  }
  if (s < V.blk_s1[b]) {
    const int br = V.sb[s] + lane;
    const int o = V.soff[s], w = (V.soff[s + 1] - o) >> 6;
    const int *cp = V.bcol + o + lane;
    const double *vp = bv + (int64_t)o * BB + lane;
    double acc[BS][NR];
#pragma unroll
    for (int i = 0; i < BS; i++)
#pragma unroll
      for (int k = 0; k < NR; k++) acc[i][k] = 0.0;
    for (int kk = 0; kk < w; kk++) {
      const int64_t c = (int64_t)cp[kk * 64] * BS * NR;
      double xv[BS][NR];
#pragma unroll
      for (int j = 0; j < BS; j++)
#pragma unroll
        for (int k = 0; k < NR; k++) xv[j][k] = u[c + j * NR + k];
      const double *vk = vp + (int64_t)kk * (64 * BB);
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int j = 0; j < BS; j++) {
          const double a = vk[(i * BS + j) * 64];
#pragma unroll
          for (int k = 0; k < NR; k++) acc[i][k] += a * xv[j][k];
        }
    }
    if (br < V.sb[s + 1]) {
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int k = 0; k < NR; k++) {
          const int64_t idx = ((int64_t)br * BS + i) * NR + k;
          const double fi = f[idx], ri = fi - acc[i][k];
          r[idx] = kept[k] ? 0.0 : ri;
          if (w32[k]) r32[idx] = kept[k] ? 0.0f : (float)ri;
          rr[k] += ri * ri;
          ff[k] += fi * fi;
          fu[k] += fi * u[idx];
        }
    }
  }
  put_part<NR>(V, Q_RES, rr);
  put_part<NR>(V, Q_FF, ff);
  put_part<NR>(V, Q_FU, fu);
}

// Multigrid on a box system (one problem, one rhs), level 0 on the sliced
// copy -- its FP32 values when the system has them (MT float: the V-cycle
// only preconditions, so its fine-level products may use the rounded
// matrix; both sweeps use the same one, which keeps the V-cycle symmetric;
// the CG's own product stays FP64).  kbx_mg3_resid: r0 = r - K z (z the
// V-cycle's first sweep).
template <int BS, typename MT>
__global__ __launch_bounds__(kBT) void kbx_mg3_resid(BView V, const MT *__restrict__ bv,
                                                     const double *__restrict__ r, const double *__restrict__ z,
                                                     double *__restrict__ r0) {
  const int b = blockIdx.x;
  if (V.seg[V.blk_prob[b]].done) return;
  constexpr int BB = BS * BS;
  const int lane = threadIdx.x & 63;
  const int s = V.blk_s0[b] + (threadIdx.x >> 6);
  if (s >= V.blk_s1[b]) return;
  const int br = V.sb[s] + lane;
  const int o = V.soff[s], w = (V.soff[s + 1] - o) >> 6;
  const int *cp = V.bcol + o + lane;
  const MT *vp = bv + (int64_t)o * BB + lane;
  double acc[BS];
#pragma unroll
  for (int i = 0; i < BS; i++) acc[i] = 0.0;
  for (int kk = 0; kk < w; kk++) {
    const int64_t c = (int64_t)cp[kk * 64] * BS;
    double xv[BS];
#pragma unroll
    for (int j = 0; j < BS; j++) xv[j] = z[c + j];
    const MT *vk = vp + (int64_t)kk * (64 * BB);
#pragma unroll
    for (int i = 0; i < BS; i++)
#pragma unroll
      for (int j = 0; j < BS; j++) acc[i] += (double)vk[(i * BS + j) * 64] * xv[j];
  }
  if (br < V.sb[s + 1])
#pragma unroll
    for (int i = 0; i < BS; i++) {
      const int64_t idx = (int64_t)br * BS + i;
      r0[idx] = r[idx] - acc[i];
    }
}

// kbx_mg3_smooth: the V-cycle's closing sweep on level 0, z = y + w Dinv
// (r - K y), fused with the CG's r.z partial sums into `slot`.
template <int BS, typename MT>
__global__ __launch_bounds__(kBT) void kbx_mg3_smooth(BView V, const MT *__restrict__ bv,
                                                      const double *__restrict__ r, const double *__restrict__ y,
                                                      const double *__restrict__ dinv, double *__restrict__ z,
                                                      int slot) {
  const int b = blockIdx.x;
  if (V.seg[V.blk_prob[b]].done) return;  // uniform per block
  constexpr int BB = BS * BS;
  const int lane = threadIdx.x & 63;
  const int s = V.blk_s0[b] + (threadIdx.x >> 6);
  double dot[1] = {0.0};
  if (s < V.blk_s1[b]) {
    const int br = V.sb[s] + lane;
    const int o = V.soff[s], w = (V.soff[s + 1] - o) >> 6;
    const int *cp = V.bcol + o + lane;
    const MT *vp = bv + (int64_t)o * BB + lane;
    double acc[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) acc[i] = 0.0;
    for (int kk = 0; kk < w; kk++) {
      const int64_t c = (int64_t)cp[kk * 64] * BS;
      double xv[BS];
#pragma unroll
      for (int j = 0; j < BS; j++) xv[j] = y[c + j];
      const MT *vk = vp + (int64_t)kk * (64 * BB);
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int j = 0; j < BS; j++) acc[i] += (double)vk[(i * BS + j) * 64] * xv[j];
    }
    if (br < V.sb[s + 1])
#pragma unroll
      for (int i = 0; i < BS; i++) {
        const int64_t idx = (int64_t)br * BS + i;
        const double zi = y[idx] + kMg3Omega * dinv[idx] * (r[idx] - acc[i]);
        z[idx] = zi;
        dot[0] += r[idx] * zi;
      }
  }
  put_part<1>(V, slot, dot);
}

// One block per problem: the refinement residual totals into its segments.
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_resid_finish(BView V) {
  const int pr = blockIdx.x;
  double a[NR], c[NR], d[NR];
  prob_sum<NR>(V, Q_RES, pr, a);
  prob_sum<NR>(V, Q_FF, pr, c);
  prob_sum<NR>(V, Q_FU, pr, d);
  if (threadIdx.x == 0)
    for (int k = 0; k < NR; k++) {
      Seg &g = V.seg[(size_t)pr * NR + k];
      g.rt = a[k];
      g.ff = c[k];
      g.fu = d[k];
    }
}

// Reliable-update helpers of the precision form, for the segments in sel:
// fold the FP32 correction into the FP64 solution ...
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_fold(BView V, double *__restrict__ x, float *__restrict__ x32,
                                                const int *__restrict__ sel) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  if (!sel[pr * NR + threadIdx.x % NR]) return;  // kBT % NR == 0: a thread keeps one rhs
  int r0, r1;
  blk_rows(V, b, r0, r1);
  for (int i = r0 * NR + threadIdx.x; i < r1 * NR; i += kBT) {
    x[i] += (double)x32[i];
    x32[i] = 0.0f;
  }
}

// ... and, after the FP64 residual replaced r, z = Dinv r with the r.z / r.r
// partials of the coming iteration (parity-1 slots); the direction is kept.
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_replace(BView V, const float *__restrict__ r, float *__restrict__ z,
                                                   const float *__restrict__ dinv, const int *__restrict__ sel) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  const bool on = sel[pr * NR + threadIdx.x % NR];
  int r0, r1;
  blk_rows(V, b, r0, r1);
  double rz = 0.0, rr = 0.0;
  if (on)
    for (int i = r0 * NR + threadIdx.x; i < r1 * NR; i += kBT) {
      const float zi = dinv[i / NR] * r[i];
      z[i] = zi;
      rz += (double)r[i] * (double)zi;
      rr += (double)r[i] * (double)r[i];
    }
  double a[NR], c[NR];
  rhs_sum<NR>(rz, a);
  rhs_sum<NR>(rr, c);
  bsum<NR>(a);
  bsum<NR>(c);
  if (threadIdx.x < NR && sel[pr * NR + threadIdx.x]) {
    double ta = a[0], tc = c[0];
#pragma unroll
    for (int k = 1; k < NR; k++)
      if ((int)threadIdx.x == k) {
        ta = a[k];
        tc = c[k];
      }
    V.part[((size_t)Q_RZ1 * V.nblk + b) * NR + threadIdx.x] = ta;
    V.part[((size_t)Q_RR1 * V.nblk + b) * NR + threadIdx.x] = tc;
  }
}

__global__ void kb_mark_done(Seg *seg, int nseg, const int *flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nseg && flags[i]) seg[i].done = 1;
}

// Resume of the flagged segments after a residual replacement: running again
// with the replaced residual as the new stop reference; segments flagged in
// zp (an FP32 breakdown) also restart their direction (both buffers zero).
template <int NR>
__global__ __launch_bounds__(kBT) void kbx_resume(BView V, float *__restrict__ p0, float *__restrict__ p1,
                                                  const int *__restrict__ flags, const int *__restrict__ zp) {
  const int b = blockIdx.x, pr = V.blk_prob[b];
  if (b == V.prob_blk0[pr] && threadIdx.x < NR && flags[pr * NR + threadIdx.x]) {
    Seg &g = V.seg[(size_t)pr * NR + threadIdx.x];
    g.done = 0;
    g.breakdown = 0;
    g.rr0 = -1.0;
  }
  if (!zp[pr * NR + threadIdx.x % NR]) return;
  int r0, r1;
  blk_rows(V, b, r0, r1);
  for (int i = r0 * NR + threadIdx.x; i < r1 * NR; i += kBT) {
    p0[i] = 0.0f;
    p1[i] = 0.0f;
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

// Segment state before an iteration run: done = flags (segments whose
// refinement is over), counters cleared, the stop reference ||f||^2 (ref_ff,
// when nonzero) or the first residual.
__global__ void kb_reset(Seg *seg, int nseg, const int *flags, int *it, int ref_ff) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *it = 0;
  if (i < nseg) {
    seg[i].done = flags ? flags[i] : 0;
    seg[i].iters = 0;
    seg[i].breakdown = 0;
    seg[i].rr0 = (ref_ff && seg[i].ff > 0.0) ? seg[i].ff : -1.0;
  }
}

// Host side: slices aligned with the problems, blocks of <= kBW slices of one
// problem, the sliced copy of the system matrix and the device view.
struct BBatch {
  int P = 0, NR = 1, bs = 1, nblk = 0, nslices = 0;
  int max_pnb = 0;  // most blocks of one problem
  int64_t nblkst = 0;
  int *blk_prob = nullptr, *blk_s0 = nullptr, *blk_s1 = nullptr, *prob_blk0 = nullptr, *prob_nblk = nullptr;
  int *sb = nullptr, *soff = nullptr, *bcol = nullptr, *map = nullptr;  // map: slot -> CSR position (-1: none)
  double *v64 = nullptr;
  float *v32 = nullptr;
  double *part = nullptr;
  Seg *seg = nullptr;
  int *it = nullptr;
  Seg *hseg = nullptr;         // pinned host copy of seg, read while the next chunk runs
  hipEvent_t ev_seg = nullptr;  // recorded after each copy into hseg
  BView view(int max_it, double rtol2) const {
    return BView{nblk, bs, blk_prob, blk_s0, blk_s1, prob_blk0, prob_nblk, sb, soff, bcol, part, seg, it,
                 max_it, rtol2};
  }
};

static void bb_free(BBatch &B) {
  dfree(B.blk_prob); dfree(B.blk_s0); dfree(B.blk_s1); dfree(B.prob_blk0); dfree(B.prob_nblk);
  dfree(B.sb); dfree(B.soff); dfree(B.bcol); dfree(B.map); dfree(B.v64); dfree(B.v32);
  dfree(B.part); dfree(B.seg); dfree(B.it);
  if (B.hseg) (void)hipHostFree(B.hseg);
  if (B.ev_seg) (void)hipEventDestroy(B.ev_seg);
  B.hseg = nullptr;
  B.ev_seg = nullptr;
}

template <int BS>
static void bb_launch_build(tfm_system *s, BBatch &B, int *cnt, int *w, int nbr, hipStream_t st, int phase) {
  const unsigned g = (unsigned)((nbr + 255) / 256), gs = (unsigned)((B.nslices + 255) / 256),
                 gf = (unsigned)(((int64_t)B.nslices * 64 + 255) / 256);
  if (phase == 0) {
    hipLaunchKernelGGL(k_bs_count<BS>, dim3(g), dim3(256), 0, st, nbr, s->d_rp, s->d_ci, cnt);
    hipLaunchKernelGGL(k_bs_width, dim3(gs), dim3(256), 0, st, nbr, B.nslices, (const int *)B.sb, cnt, w);
  } else {
    hipLaunchKernelGGL(k_bs_fill<BS>, dim3(gf), dim3(256), 0, st, nbr, B.nslices, (const int *)B.sb, s->d_rp,
                       s->d_ci, B.soff, B.bcol, B.map);
  }
}

static int bb_setup(tfm_system *s, int P, const int64_t *dof_off, int NR, BBatch &B) {
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  int bs = (s->dpn >= 1 && s->dpn <= 3) ? s->dpn : 1;
  for (int p = 0; p <= P && bs > 1; p++)
    if (dof_off[p] % bs) bs = 1;  // problems must start on a node
  const int64_t nbr = s->ndof / bs;
  // slices of <= 64 block rows, aligned with the problems; blocks of <= kBW
  // slices of one problem
  std::vector<int> sb, bp, b0, b1, pb0(P), pnb(P);
  for (int p = 0; p < P; p++) {
    const int64_t q0 = dof_off[p] / bs, q1 = dof_off[p + 1] / bs;
    pb0[p] = (int)bp.size();
    const int64_t ns = std::max<int64_t>(1, (q1 - q0 + 63) / 64);
    const int first = (int)sb.size();
    for (int64_t t = 0; t < ns; t++) sb.push_back((int)std::min(q1, q0 + 64 * t));
    for (int64_t t = 0; t < ns; t += kBW) {
      bp.push_back(p);
      b0.push_back(first + (int)t);
      b1.push_back(first + (int)std::min<int64_t>(ns, t + kBW));
    }
    pnb[p] = (int)bp.size() - pb0[p];
  }
  sb.push_back((int)nbr);
  B.P = P;
  B.NR = NR;
  B.bs = bs;
  B.nblk = (int)bp.size();
  B.nslices = (int)sb.size() - 1;
  B.max_pnb = pnb.empty() ? 0 : *std::max_element(pnb.begin(), pnb.end());
  int rc;
  int *cnt = nullptr, *w = nullptr;
  if ((rc = dalloc(ctx, &B.blk_prob, B.nblk)) || (rc = dalloc(ctx, &B.blk_s0, B.nblk)) ||
      (rc = dalloc(ctx, &B.blk_s1, B.nblk)) || (rc = dalloc(ctx, &B.prob_blk0, P)) ||
      (rc = dalloc(ctx, &B.prob_nblk, P)) || (rc = dalloc(ctx, &B.sb, sb.size())) ||
      (rc = dalloc(ctx, &B.soff, B.nslices + 1)) || (rc = dalloc(ctx, &B.part, (size_t)Q_N * B.nblk * NR)) ||
      (rc = dalloc(ctx, &B.seg, (size_t)P * NR)) || (rc = dalloc(ctx, &B.it, 1)) || (rc = dalloc(ctx, &cnt, nbr)) ||
      (rc = dalloc(ctx, &w, B.nslices))) {
    dfree(cnt);
    dfree(w);
    return rc;
  }
  if (hipHostMalloc((void **)&B.hseg, sizeof(Seg) * P * NR, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&B.ev_seg, hipEventDisableTiming) != hipSuccess) {
    dfree(cnt);
    dfree(w);
    return fail(ctx, TFM_E_HIP, "pinned segment buffer allocation failed");
  }
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.blk_prob, bp.data(), sizeof(int) * B.nblk, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.blk_s0, b0.data(), sizeof(int) * B.nblk, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.blk_s1, b1.data(), sizeof(int) * B.nblk, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.prob_blk0, pb0.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.prob_nblk, pnb.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(B.sb, sb.data(), sizeof(int) * sb.size(), hipMemcpyHostToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(B.seg, 0, sizeof(Seg) * P * NR, st));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(B.part, 0, sizeof(double) * Q_N * B.nblk * NR, st));
  switch (bs) {
  case 3: bb_launch_build<3>(s, B, cnt, w, (int)nbr, st, 0); break;
  case 2: bb_launch_build<2>(s, B, cnt, w, (int)nbr, st, 0); break;
  default: bb_launch_build<1>(s, B, cnt, w, (int)nbr, st, 0); break;
  }
  rc = exclusive_scan_i32(ctx, w, B.soff, B.nslices, &B.nblkst);
  dfree(cnt);
  dfree(w);
  if (rc) return rc;
  if (B.nblkst * bs * bs >= INT32_MAX) return fail(ctx, TFM_E_UNSUPPORTED, "batched system too large for int32 slots");
  if ((rc = dalloc(ctx, &B.bcol, B.nblkst)) || (rc = dalloc(ctx, &B.map, s->nnz))) return rc;
  switch (bs) {
  case 3: bb_launch_build<3>(s, B, nullptr, nullptr, (int)nbr, st, 1); break;
  case 2: bb_launch_build<2>(s, B, nullptr, nullptr, (int)nbr, st, 1); break;
  default: bb_launch_build<1>(s, B, nullptr, nullptr, (int)nbr, st, 1); break;
  }
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return bs_invert_map(ctx, &B.map, s->nnz, B.nblkst * bs * bs);  // B.map: slot -> CSR position
}

template <typename VT>
static int bb_values(tfm_system *s, BBatch &B, const VT *src, VT **dst) {
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t nv = B.nblkst * B.bs * B.bs;
  int rc = *dst ? TFM_OK : dalloc(ctx, dst, nv);
  if (rc) return rc;
  if (nv > 0)
    hipLaunchKernelGGL((k_bs_values<VT, VT>), dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, nv, B.map, src,
                       *dst);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

template <typename VT, int NR, bool SC = false>
static void bb_dir(const BBatch &B, const BView &V, hipStream_t st, const VT *bv, const VT *z, const VT *pold,
                   VT *pnew, VT *q, int e) {
  dim3 g(B.nblk), t(kBT);
  switch (B.bs) {
  case 3: hipLaunchKernelGGL((kbx_dir<VT, 3, NR, SC>), g, t, 0, st, V, bv, z, pold, pnew, q, e); break;
  case 2: hipLaunchKernelGGL((kbx_dir<VT, 2, NR, SC>), g, t, 0, st, V, bv, z, pold, pnew, q, e); break;
  default: hipLaunchKernelGGL((kbx_dir<VT, 1, NR, SC>), g, t, 0, st, V, bv, z, pold, pnew, q, e); break;
  }
}

// FP64 residual r = f - K u and its partials: on the sliced copy when it
// holds FP64 values, else on the CSR.
template <int NR>
static void bb_resid(const tfm_system *s, const BBatch &B, const BView &V, hipStream_t st, const double *u,
                     const double *f, double *r, float *r32, const int *keep, const int *sel) {
  dim3 g(B.nblk), t(kBT);
  if (!B.v64) {
    hipLaunchKernelGGL((kbx_resid64<NR>), g, t, 0, st, V, s->d_rp, s->d_ci, s->d_v64, u, f, r, r32, keep, sel);
    return;
  }
  switch (B.bs) {
  case 3: hipLaunchKernelGGL((kbx_resid_bs<3, NR>), g, t, 0, st, V, (const double *)B.v64, u, f, r, r32, keep, sel); break;
  case 2: hipLaunchKernelGGL((kbx_resid_bs<2, NR>), g, t, 0, st, V, (const double *)B.v64, u, f, r, r32, keep, sel); break;
  default: hipLaunchKernelGGL((kbx_resid_bs<1, NR>), g, t, 0, st, V, (const double *)B.v64, u, f, r, r32, keep, sel); break;
  }
}

// ---------------------------------------------------------------------------
// One batched solve as a sequence of captured graphs.
//
// A solve runs as four stages, each captured into a hipGraph on its first use
// and replayed afterwards (TFM_NO_GRAPH=1 launches the same kernels
// directly, for tracers):
//   A  the caller's pre-stage (the SIMP loop's numeric assembly pass) and the
//      refresh of the solver's copy: sliced values, Jacobi diagonals,
//      multigrid levels;
//   B  the initial FP64 residual and the CG state;
//   K  a chunk of iterations -- plain/multigrid: 16 (8) PCG iterations;
//      precision form: 16 FP32 iterations, then the refinement step, decided
//      per segment on the device (fold the FP32 correction, FP64 residual,
//      replace the FP32 residual, restart or stop);
//   C  the FP64 residual certificate of every segment and the caller's
//      post-stage (the SIMP loop's sensitivities, filter and OC update).
// The host launches A, B, then K until every segment is done -- it reads a
// small pinned snapshot per chunk, one chunk behind, so the check overlaps
// the next chunk -- then C, and reads the segments' state once at the end.
// ---------------------------------------------------------------------------
struct Stage {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t key = 0;
  void clear() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key = 0;
  }
};

static uint64_t key_of(std::initializer_list<uint64_t> parts) {
  uint64_t h = 1469598103934665603ull;
  for (uint64_t v : parts) h = (h ^ v) * 1099511628211ull + 0x9e3779b97f4a7c15ull;
  return h | 1ull;  // never 0, the empty stage's key
}

static uint64_t bits_of(double d) {
  uint64_t u;
  std::memcpy(&u, &d, sizeof(u));
  return u;
}

// Enqueue a stage: replay its graph, capturing it first when its key (the
// stop rule and the callers' tags, which fix every kernel argument) changed.
template <typename F>
static int stage_run(tfm_ctx *ctx, Stage &S, uint64_t key, bool use_graph, F &&enqueue) {
  hipStream_t st = ctx->stream;
  if (!use_graph) return enqueue();
  if (!S.exec || S.key != key) {
    S.clear();
    TFM_CHECK_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc || e != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return rc ? rc : fail(ctx, TFM_E_HIP, std::string("graph capture failed: ") + hipGetErrorString(e));
    }
    S.graph = g;
    if (hipGraphInstantiate(&S.exec, g, nullptr, nullptr, 0) != hipSuccess) {
      S.clear();
      return fail(ctx, TFM_E_HIP, "hipGraphInstantiate failed");
    }
    S.key = key;
  }
  if (hipGraphLaunch(S.exec, st) != hipSuccess) return fail(ctx, TFM_E_HIP, "hipGraphLaunch failed");
  return TFM_OK;
}

// Clears inside captured stages are fill kernels (no memset nodes).
template <typename T>
__global__ void kb_fill(int64_t n, T *__restrict__ p, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
template <typename T>
static void fill(hipStream_t st, T *p, int64_t n, T v) {
  if (n <= 0) return;
  const int64_t g = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(kb_fill<T>, dim3((unsigned)g), dim3(256), 0, st, n, p, v);
}

// Precision form, refinement state of one segment (device).
struct Mix {
  double rr_ref, best;  // FP32 residual^2 at the last replacement; best since
  int since;            // chunks without a 10 % improvement of best
  int od;               // refinement over (converged, max_iter or failed)
  int cycles, iters, nbrk, brk;
  int sel, rst;         // this chunk: replace the residual; restart the direction
};
constexpr int kMixCh = 16;  // FP32 iterations per chunk
constexpr int kStall = 64;  // chunks without a 10 % residual improvement

// After the first FP64 residual: segments already within rtol are over.
__global__ void kb_mix_init(const Seg *__restrict__ seg, int nseg, Mix *__restrict__ mx, int *__restrict__ odone,
                            int *__restrict__ mit, double rtol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *mit = 0;
  if (i >= nseg) return;
  const Seg g = seg[i];
  const double rel = g.ff > 0 ? sqrt(g.rt / g.ff) : 0.0;
  Mix m{};
  m.od = rel <= rtol;
  m.rr_ref = m.best = g.rt;
  m.cycles = 1;
  mx[i] = m;
  odone[i] = m.od;
}

// After a chunk: which running segments replace their residual -- those
// whose FP32 residual fell by delta since the last replacement or below the
// target, that froze (FP32 floor or breakdown), stalled, or spent max_iter.
// One block; advances the iteration count by a chunk.
__global__ void kb_mix_select(const Seg *__restrict__ seg, int nseg, Mix *__restrict__ mx, int *__restrict__ sel,
                              int *__restrict__ mit, double delta2, double rtol2, int max_iter) {
  const int done_it = *mit + kMixCh;
  for (int q = threadIdx.x; q < nseg; q += blockDim.x) {
    Mix m = mx[q];
    int s = 0;
    if (!m.od) {
      const Seg g = seg[q];
      if (g.rr < 0.81 * m.best) {
        m.best = g.rr;
        m.since = 0;
      } else {
        m.since++;
      }
      if (g.rr <= delta2 * m.rr_ref || g.rr <= rtol2 * g.ff || g.done || m.since >= kStall || done_it >= max_iter)
        s = 1;
    }
    m.sel = s;
    sel[q] = s;
    mx[q] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) *mit = done_it;
}

// After the FP64 residual of the selected segments: stop those within rtol
// (or out of iterations, or failing on repeated FP32 breakdowns), restart
// the direction of frozen ones; the others resume with the new residual.
__global__ void kb_mix_after(const Seg *__restrict__ seg, int nseg, Mix *__restrict__ mx, int *__restrict__ sel,
                             int *__restrict__ rst, int *__restrict__ odone, const int *__restrict__ mit, double rtol,
                             int max_iter) {
  const int done_it = *mit;
  for (int q = threadIdx.x; q < nseg; q += blockDim.x) {
    Mix m = mx[q];
    int r = 0;
    if (m.sel) {
      const Seg g = seg[q];
      const double rel = g.ff > 0 ? sqrt(g.rt / g.ff) : 0.0;
      m.cycles += 1;
      m.rr_ref = m.best = g.rt;
      m.since = 0;
      if (rel <= rtol || done_it >= max_iter) {
        m.od = 1;
        m.iters = done_it;
        m.sel = 0;
      } else if (g.breakdown && ++m.nbrk > 4) {
        // an FP32 breakdown (p.q <= 0) that keeps coming back fails the segment
        m.brk = 1;
        m.od = 1;
        m.iters = done_it;
        m.sel = 0;
      } else if (g.done) {
        r = 1;  // frozen (floor or breakdown): restart the direction
      }
    }
    rst[q] = r;
    sel[q] = m.sel;
    odone[q] = m.od;
    mx[q] = m;
  }
}

// Host side of a batched solve: the sliced matrix, work vectors and the
// device solution x ([n_dof][NR], kept between runs for warm starts).
struct BatchSolver {
  tfm_system *s = nullptr;
  int P = 0, NR = 1;
  bool mixed = false;
  BBatch B;
  double *f = nullptr, *x = nullptr, *r64 = nullptr, *d64 = nullptr, *z64 = nullptr, *p64 = nullptr,
         *p64b = nullptr, *q64 = nullptr;
  float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *p32b = nullptr, *q32 = nullptr, *d32 = nullptr,
        *x32 = nullptr;
  int *bad = nullptr, *odone = nullptr;
  Mix *mix = nullptr;            // precision form: per-segment refinement state
  int *mit = nullptr;            // precision form: iterations run
  int *sel = nullptr, *rst = nullptr;
  Mix *hmix = nullptr;           // pinned snapshot of mix (od), one chunk behind
  hipEvent_t ev_mix = nullptr;
  MgData *mg = nullptr;    // multigrid preconditioner (TFM_SOLVE_PCG64_MG), SIMP grids
  Mg3Data *mg3 = nullptr;  // the same for one system on a box mesh
  Stage sa, sb, sk, sc;    // the solve's captured stages
};

void bsolver_destroy(BatchSolver *b) {
  if (!b) return;
  b->sa.clear();
  b->sb.clear();
  b->sk.clear();
  b->sc.clear();
  bb_free(b->B);
  dfree(b->f); dfree(b->x); dfree(b->r64); dfree(b->d64); dfree(b->z64); dfree(b->p64); dfree(b->p64b);
  dfree(b->q64); dfree(b->r32); dfree(b->z32); dfree(b->p32); dfree(b->p32b); dfree(b->q32); dfree(b->d32);
  dfree(b->x32); dfree(b->bad); dfree(b->odone); dfree(b->mix); dfree(b->mit); dfree(b->sel); dfree(b->rst);
  if (b->hmix) (void)hipHostFree(b->hmix);
  if (b->ev_mix) (void)hipEventDestroy(b->ev_mix);
  mg_destroy(b->mg);
  mg3_destroy(b->mg3);
  delete b;
}

int bsolver_create(tfm_system *s, int P, const int64_t *dof_off, int NR, int mode, BatchSolver **out) {
  tfm_ctx *ctx = s->ctx;
  *out = nullptr;
  if (NR != 1 && NR != 8) return fail(ctx, TFM_E_UNSUPPORTED, "batched solve supports 1 or 8 right-hand sides");
  if (mode != TFM_SOLVE_PCG64 && mode != TFM_SOLVE_MIXED && mode != TFM_SOLVE_PCG64_MG)
    return fail(ctx, TFM_E_INVALID, "unknown solve mode");
  const bool mixed = mode == TFM_SOLVE_MIXED;
  bool mg2d = false, mg3d = false;
  if (mode == TFM_SOLVE_PCG64_MG) {
    // the V-cycles know two layouts: the SIMP operator on P concatenated
    // grids, problem p owning design p's 2 nx ny DOFs; and one system with
    // 1 or 3 DOFs per node on a box mesh, one right-hand side
    const tfm_mesh *m = s->has_plan ? s->plan.mesh : nullptr;
    mg2d = m && s->plan.op.kind == TFM_OP_SIMP && s->dpn == 2 && m->grid_P == P && s->plan.d_rho;
    for (int p = 0; mg2d && p <= P; p++) mg2d = dof_off[p] == (int64_t)p * 2 * m->grid_nx * m->grid_ny;
    mg3d = !mg2d && m && m->box_n[0] > 0 && P == 1 && NR == 1 && (s->dpn == 1 || s->dpn == 3) &&
           dof_off[0] == 0 && dof_off[1] == s->ndof &&
           s->ndof == (int64_t)s->dpn * (m->box_n[0] + 1) * (m->box_n[1] + 1) * (m->box_n[2] + 1);
    if (!mg2d && !mg3d)
      return fail(ctx, TFM_E_UNSUPPORTED,
                  "multigrid solve needs a SIMP system on a grid_batch mesh (one problem per design) or one "
                  "system on a box mesh with one right-hand side");
  }
  if (mixed && !s->d_v32) return fail(ctx, TFM_E_INVALID, "mixed solve needs FP32 values");
  if (mixed && !s->fp32_current())
    return fail(ctx, TFM_E_INVALID, "mixed solve: the FP32 values predate the last FP64 re-assembly");
  if (!s->d_v64) return fail(ctx, TFM_E_INVALID, "solve needs FP64 values");
  auto *b = new BatchSolver();
  b->s = s;
  b->P = P;
  b->NR = NR;
  b->mixed = mixed;
  const int64_t n = s->ndof;
  const size_t nv = (size_t)n * NR;
  const int nseg = P * NR;
  int rc = bb_setup(s, P, dof_off, NR, b->B);
  if (!rc && !((rc = dalloc(ctx, &b->f, nv)) || (rc = dalloc(ctx, &b->x, nv)) || (rc = dalloc(ctx, &b->r64, nv)) ||
               (rc = dalloc(ctx, &b->d64, n)) || (rc = dalloc(ctx, &b->bad, 1)) ||
               (rc = dalloc(ctx, &b->odone, nseg)))) {
    if (!mixed)
      (void)((rc = dalloc(ctx, &b->z64, nv)) || (rc = dalloc(ctx, &b->p64, nv)) ||
             (rc = dalloc(ctx, &b->p64b, nv)) || (rc = dalloc(ctx, &b->q64, nv)));
    else
      (void)((rc = dalloc(ctx, &b->r32, nv)) || (rc = dalloc(ctx, &b->z32, nv)) ||
             (rc = dalloc(ctx, &b->p32, nv)) || (rc = dalloc(ctx, &b->p32b, nv)) ||
             (rc = dalloc(ctx, &b->q32, nv)) || (rc = dalloc(ctx, &b->d32, n)) || (rc = dalloc(ctx, &b->x32, nv)) ||
             (rc = dalloc(ctx, &b->mix, nseg)) || (rc = dalloc(ctx, &b->mit, 1)) || (rc = dalloc(ctx, &b->sel, nseg)) ||
             (rc = dalloc(ctx, &b->rst, nseg)));
  }
  if (!rc && mixed &&
      (hipHostMalloc((void **)&b->hmix, sizeof(Mix) * nseg, hipHostMallocDefault) != hipSuccess ||
       hipEventCreateWithFlags(&b->ev_mix, hipEventDisableTiming) != hipSuccess))
    rc = fail(ctx, TFM_E_HIP, "pinned refinement-state buffer allocation failed");
  if (!rc && mg2d)
    rc = mg_create(ctx, P, (int)s->plan.mesh->grid_nx, (int)s->plan.mesh->grid_ny, NR, s->plan.ke0, &b->mg);
  if (!rc && mg3d) {
    const int64_t *bn = s->plan.mesh->box_n;
    const int n3[3] = {(int)bn[0] + 1, (int)bn[1] + 1, (int)bn[2] + 1};
    rc = mg3_create(ctx, n3, s->dpn, &b->mg3);
  }
  if (!rc && hipMemsetAsync(b->x, 0, sizeof(double) * nv, ctx->stream) != hipSuccess)
    rc = fail(ctx, TFM_E_HIP, "hipMemsetAsync failed");
  if (rc) {
    bsolver_destroy(b);
    return rc;
  }
  *out = b;
  return TFM_OK;
}

// Right-hand sides: host [NR][n_dof], or null for the system's own (NR = 1);
// homogeneous-Dirichlet systems take f[fixed] = 0 (fred on the free DOFs).
int bsolver_set_rhs(BatchSolver *b, const double *rhs) {
  tfm_system *s = b->s;
  tfm_ctx *ctx = s->ctx;
  const int64_t n = s->ndof;
  const int NR = b->NR;
  if (!rhs && NR == 1) {  // the system's own rhs is already 0 at fixed DOFs
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(b->f, s->d_f, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    return TFM_OK;
  }
  std::vector<unsigned char> fx((size_t)n, 0);
  TFM_CHECK_HIP(ctx, hipMemcpy(fx.data(), s->d_fixed, n, hipMemcpyDeviceToHost));
  std::vector<double> hf((size_t)n * NR, 0.0);
  if (rhs) {
    for (int k = 0; k < NR; k++)
      for (int64_t i = 0; i < n; i++) hf[(size_t)i * NR + k] = fx[i] ? 0.0 : rhs[(size_t)k * n + i];
  } else {
    std::vector<double> fr((size_t)n);
    TFM_CHECK_HIP(ctx, hipMemcpy(fr.data(), s->d_f, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < NR; k++)
      for (int64_t i = 0; i < n; i++) hf[(size_t)i * NR + k] = fr[i];
  }
  TFM_CHECK_HIP(ctx, hipMemcpy(b->f, hf.data(), sizeof(double) * hf.size(), hipMemcpyHostToDevice));
  return TFM_OK;
}

// Refresh of the solver's copy after a (re-)assembly, in two parts: the
// host-side checks and first-use allocations (run directly), then the
// kernels (captured into stage A): sliced values, Jacobi diagonals (a free
// DOF without a positive diagonal raises the `bad` flag, checked when the
// solve ends) and the multigrid levels.
static int refresh_prepare(BatchSolver *b) {
  tfm_system *s = b->s;
  tfm_ctx *ctx = s->ctx;
  const int64_t nv = b->B.nblkst * b->B.bs * b->B.bs;
  if (b->mixed && !s->fp32_current())
    return fail(ctx, TFM_E_INVALID, "mixed solve: the FP32 values predate the last FP64 re-assembly");
  int rc = b->B.v64 ? TFM_OK : dalloc(ctx, &b->B.v64, nv);
  // The box multigrid's two fine passes may read the FP32 copy, but only
  // while it holds the same matrix as the FP64 values (same version);
  // otherwise they run on the FP64 copy.
  const bool want32 = b->mixed || (b->mg3 && s->fp32_current());
  if (!rc && want32 && !b->B.v32) rc = dalloc(ctx, &b->B.v32, nv);
  if (!rc && !want32 && b->B.v32) dfree(b->B.v32);
  return rc;
}

static int refresh_enqueue(BatchSolver *b) {
  tfm_system *s = b->s;
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int n = (int)s->ndof;
  const int64_t nv = b->B.nblkst * b->B.bs * b->B.bs;
  if (nv > 0) {
    hipLaunchKernelGGL((k_bs_values<double, double>), dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, nv,
                       b->B.map, (const double *)s->d_v64, b->B.v64);
    if (b->B.v32)
      hipLaunchKernelGGL((k_bs_values<float, float>), dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, nv,
                         b->B.map, (const float *)s->d_v32, b->B.v32);
  }
  fill(st, b->bad, 1, 0);
  const unsigned g1 = (unsigned)((n + 255) / 256);
  if (n > 0) {
    hipLaunchKernelGGL(kb_dinv<double>, dim3(g1), dim3(256), 0, st, n, s->d_rp, s->d_ci, s->d_v64, s->d_fixed,
                       b->d64, b->bad);
    if (b->mixed)
      hipLaunchKernelGGL(kb_dinv<float>, dim3(g1), dim3(256), 0, st, n, s->d_rp, s->d_ci, s->d_v32, s->d_fixed,
                         b->d32, b->bad);
  }
  int rc;
  if (b->mg && (rc = mg_update(b->mg, s->plan.d_rho, s->plan.op.penal, s->plan.op.e0, s->plan.op.emin, s->d_fixed)))
    return rc;
  if (b->mg3 && (rc = mg3_update(b->mg3, s->d_rp, s->d_ci, s->d_v64, b->d64))) return rc;
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

static int bad_check(BatchSolver *b) {
  tfm_ctx *ctx = b->s->ctx;
  int hbad = 0;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&hbad, b->bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (hbad)
    return fail(ctx, TFM_E_NUMERIC,
                "sparse numeric factorization/solve failed: singular matrix (a free DOF has no positive diagonal)");
  return TFM_OK;
}

// The refresh on its own (run directly), for callers that solve many times
// against one matrix (time stepping).
int bsolver_refresh(BatchSolver *b) {
  int rc = refresh_prepare(b);
  if (!rc) rc = refresh_enqueue(b);
  if (!rc) rc = bad_check(b);
  return rc;
}

const double *bsolver_x(const BatchSolver *b) { return b->x; }
double *bsolver_x_mut(BatchSolver *b) { return b->x; }

// f[i] = fixed[i] ? 0 : d[i] (one right-hand side, already on the device).
__global__ void kb_rhs_dev(int64_t n, const double *__restrict__ d, const unsigned char *__restrict__ fixed,
                           double *__restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (fixed && fixed[i]) ? 0.0 : d[i];
}

int bsolver_set_rhs_dev(BatchSolver *b, const double *d_rhs) {
  tfm_system *s = b->s;
  if (b->NR != 1) return fail(s->ctx, TFM_E_INVALID, "device right-hand side needs one rhs per problem");
  const int64_t n = s->ndof;
  if (n > 0)
    hipLaunchKernelGGL(kb_rhs_dev, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->ctx->stream, n, d_rhs,
                       (const unsigned char *)s->d_fixed, b->f);
  TFM_CHECK_HIP(s->ctx, hipGetLastError());
  return TFM_OK;
}

// ---- the plain and multigrid CG (FP64) ------------------------------------
// Segments stop at ||r|| <= rtol ||f|| (recurrence residual against the
// segment's ||f||^2) or max_iter.  Two kernels per iteration (Jacobi);
// multigrid: z = M r by a V-cycle instead of Dinv r, skipping the segments
// already done; the update kernels write z = w Dinv r (the V-cycle's first
// sweep) and the closing fine sweep replaces their r.z partials; beta has
// its own kernel and K p is taken matrix-free (SIMP grids).
template <int NR>
struct PlainCg {
  BatchSolver *b;
  hipStream_t st;
  BView V;
  dim3 g, t;
  const double *jac;
  MgOp mop{};
  const double *mmask = nullptr;
  const int *sdone;
  int sstride;
  // Scalar mode for problems spread over many blocks: the stop test, beta
  // and alpha formed once per problem (kbx_beta, kbx_alpha) instead of by
  // every block re-summing all of its problem's partials; same values.
  bool sc;
  PlainCg(BatchSolver *b_, const tfm_solve_opts &o, bool scalar)
      : b(b_), st(b_->s->ctx->stream), V(b_->B.view(o.max_iter, o.rtol * o.rtol)), g(b_->B.nblk), t(kBT),
        jac(b_->d64), sc(scalar) {
    static_assert(sizeof(Seg) % sizeof(int) == 0, "Seg stride in ints");
    sdone = reinterpret_cast<const int *>(reinterpret_cast<const char *>(b->B.seg) + offsetof(Seg, done));
    sstride = (int)(sizeof(Seg) / sizeof(int));
    if (b->mg) {
      jac = mg_sweep_weights(b->mg);
      mop = mg_fine_op(b->mg);
      mmask = mg_fine_dinv(b->mg);
    }
    if (b->mg3) jac = mg3_sweep_weights(b->mg3);
  }
  int chunk_len() const { return (b->mg || b->mg3) ? 8 : 16; }
  // z = M r and the partial r.z into `slot`: the V-cycle up to the finest
  // level's last sweep, which runs fused with the dot product
  void precond(int slot) {
    double *r = b->r64, *z = b->z64;
    if (b->mg) {
      const double *yv = mg_apply(b->mg, r, z, st, sdone, sstride, false);
      if (yv)
        hipLaunchKernelGGL((kbx_smooth_rz<NR>), g, t, 0, st, V, mop, mmask, (const double *)r, yv, z, slot);
      else
        hipLaunchKernelGGL((kbx_rz<NR>), g, t, 0, st, V, (const double *)r, (const double *)z, slot);
      return;
    }
    if constexpr (NR == 1) {
      // the same on a box system: level 0 on the sliced copy around mg3_cycle
      Mg3Data *mg3 = b->mg3;
      const BBatch &B = b->B;
      if (mg3_depth(mg3) == 0) {
        mg3_direct(mg3, r, z, st, sdone);
        hipLaunchKernelGGL((kbx_rz<NR>), g, t, 0, st, V, (const double *)r, (const double *)z, slot);
        return;
      }
      double *r0 = mg3_fine_r(mg3);
      const double *d0 = mg3_fine_dinv(mg3);
      const float *b32 = B.v32;
      const double *bv = B.v64;
      if (b32) {
        if (B.bs == 3) hipLaunchKernelGGL((kbx_mg3_resid<3, float>), g, t, 0, st, V, b32, r, z, r0);
        else hipLaunchKernelGGL((kbx_mg3_resid<1, float>), g, t, 0, st, V, b32, r, z, r0);
      } else {
        if (B.bs == 3) hipLaunchKernelGGL((kbx_mg3_resid<3, double>), g, t, 0, st, V, bv, r, z, r0);
        else hipLaunchKernelGGL((kbx_mg3_resid<1, double>), g, t, 0, st, V, bv, r, z, r0);
      }
      const double *yv = mg3_cycle(mg3, z, st, sdone);
      if (b32) {
        if (B.bs == 3) hipLaunchKernelGGL((kbx_mg3_smooth<3, float>), g, t, 0, st, V, b32, r, yv, d0, z, slot);
        else hipLaunchKernelGGL((kbx_mg3_smooth<1, float>), g, t, 0, st, V, b32, r, yv, d0, z, slot);
      } else {
        if (B.bs == 3) hipLaunchKernelGGL((kbx_mg3_smooth<3, double>), g, t, 0, st, V, bv, r, yv, d0, z, slot);
        else hipLaunchKernelGGL((kbx_mg3_smooth<1, double>), g, t, 0, st, V, bv, r, yv, d0, z, slot);
      }
    }
  }
  // stage B: r = f - K x (x = 0 or the warm start) with ||f||^2 as the stop
  // reference, then the CG's first state
  int init(bool warm) {
    BatchSolver &S = *b;
    const BBatch &B = S.B;
    const int nseg = S.P * NR;
    const BView V0 = B.view(V.max_it, 0.0);
    if (!warm) fill(st, S.x, S.s->ndof * NR, 0.0);
    fill(st, S.odone, nseg, 0);
    bb_resid<NR>(S.s, B, V0, st, S.x, S.f, S.r64, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL((kbx_resid_finish<NR>), dim3(S.P), t, 0, st, V0);
    hipLaunchKernelGGL(kb_reset, dim3((nseg + 255) / 256), dim3(256), 0, st, B.seg, nseg, (const int *)nullptr, B.it,
                       1);
    hipLaunchKernelGGL((kbx_init<double, NR>), g, t, 0, st, V, S.r64, S.z64, S.p64, S.p64b, jac);
    if (S.mg || S.mg3) precond(Q_RZ1);
    TFM_CHECK_HIP(S.s->ctx, hipGetLastError());
    return TFM_OK;
  }
  // stage K: chunk_len() iterations (even: the buffer parity is static)
  int chunk() {
    BatchSolver &S = *b;
    const BBatch &B = S.B;
    double *pb[2] = {S.p64, S.p64b};
    for (int k = 0; k < chunk_len(); k++) {
      const int e = k & 1;
      if (S.mg) {
        hipLaunchKernelGGL((kbx_beta<NR>), dim3(S.P), t, 0, st, V, e);
        hipLaunchKernelGGL((kbx_dir_p<NR>), g, t, 0, st, V, (const double *)S.z64, (const double *)pb[1 - e], pb[e]);
        hipLaunchKernelGGL((kbx_q_mg<NR>), g, t, 0, st, V, mop, mmask, (const double *)pb[e], S.q64);
      } else if (sc) {
        hipLaunchKernelGGL((kbx_beta<NR>), dim3(S.P), t, 0, st, V, e);
        bb_dir<double, NR, true>(B, V, st, B.v64, S.z64, pb[1 - e], pb[e], S.q64, e);
      } else {
        bb_dir<double, NR>(B, V, st, B.v64, S.z64, pb[1 - e], pb[e], S.q64, e);
      }
      if (sc) {
        hipLaunchKernelGGL((kbx_alpha<NR>), dim3(S.P), t, 0, st, V, e);
        hipLaunchKernelGGL((kbx_update<double, double, NR, true>), g, t, 0, st, V, S.x, S.r64, S.z64, pb[e], S.q64,
                           jac, e);
      } else {
        hipLaunchKernelGGL((kbx_update<double, double, NR>), g, t, 0, st, V, S.x, S.r64, S.z64, pb[e], S.q64, jac, e);
      }
      if (S.mg || S.mg3) precond(e ? Q_RZ1 : Q_RZ0);
    }
    TFM_CHECK_HIP(S.s->ctx, hipGetLastError());
    return TFM_OK;
  }
};

// ---- the precision form ----------------------------------------------------
// FP32 PCG on the correction with reliable updates, per segment: whenever a
// segment's recurrence residual has dropped by delta since its last
// replacement (or stalls), its FP32 correction is folded into the FP64
// solution and its residual replaced by the true FP64 f - K u, keeping its
// search direction; a segment ends when that residual meets rtol.
// Replacements happen at chunk boundaries (the buffer and partial-slot
// parity is then consistent).  Inside a chunk a segment freezes only at the
// FP32 accuracy floor -- its residual 1e-6 below the last replacement -- or
// on a breakdown, so a small problem never iterates on rounding noise; a
// frozen segment restarts its direction.
template <int NR>
struct MixedCg {
  BatchSolver *b;
  hipStream_t st;
  BView V, V32;
  dim3 g, t;
  double rtol, delta;
  int max_iter;
  MixedCg(BatchSolver *b_, const tfm_solve_opts &o)
      : b(b_), st(b_->s->ctx->stream), V(b_->B.view(o.max_iter, 0.0)), V32(b_->B.view(o.max_iter, kFloor2)),
        g(b_->B.nblk), t(kBT), rtol(o.rtol),
        delta((o.inner_rtol > 0.0 && o.inner_rtol < 1.0) ? o.inner_rtol : 0.3), max_iter(o.max_iter) {}
  int chunk_len() const { return kMixCh; }
  int init(bool warm) {
    BatchSolver &S = *b;
    const BBatch &B = S.B;
    const int nseg = S.P * NR;
    const int nsb = (nseg + 255) / 256;
    if (!warm) fill(st, S.x, S.s->ndof * NR, 0.0);
    fill(st, S.odone, nseg, 0);
    // initial FP64 residual of every segment, rounded into the FP32 residual
    bb_resid<NR>(S.s, B, V, st, S.x, S.f, S.r64, S.r32, S.odone, nullptr);
    hipLaunchKernelGGL((kbx_resid_finish<NR>), dim3(S.P), t, 0, st, V);
    hipLaunchKernelGGL(kb_mix_init, dim3(nsb), dim3(256), 0, st, (const Seg *)B.seg, nseg, S.mix, S.odone, S.mit,
                       rtol);
    fill(st, S.x32, S.s->ndof * NR, 0.0f);
    hipLaunchKernelGGL(kb_reset, dim3(nsb), dim3(256), 0, st, B.seg, nseg, (const int *)S.odone, B.it, 0);
    hipLaunchKernelGGL((kbx_init<float, NR>), g, t, 0, st, V32, S.r32, S.z32, S.p32, S.p32b, (const float *)S.d32);
    TFM_CHECK_HIP(S.s->ctx, hipGetLastError());
    return TFM_OK;
  }
  int chunk() {
    BatchSolver &S = *b;
    const BBatch &B = S.B;
    const int nseg = S.P * NR;
    const int nsb = (nseg + 255) / 256;
    float *pb[2] = {S.p32, S.p32b};
    for (int k = 0; k < kMixCh; k++) {
      const int e = k & 1;
      bb_dir<float, NR>(B, V32, st, B.v32, S.z32, pb[1 - e], pb[e], S.q32, e);
      hipLaunchKernelGGL((kbx_update<float, float, NR>), g, t, 0, st, V32, S.x32, S.r32, S.z32, pb[e], S.q32,
                         (const float *)S.d32, e);
    }
    // the refinement step; a no-op for segments not selected
    hipLaunchKernelGGL(kb_mix_select, dim3(1), dim3(256), 0, st, (const Seg *)B.seg, nseg, S.mix, S.sel, S.mit,
                       delta * delta, rtol * rtol, max_iter);
    hipLaunchKernelGGL((kbx_fold<NR>), g, t, 0, st, V, S.x, S.x32, (const int *)S.sel);
    bb_resid<NR>(S.s, B, V, st, S.x, S.f, S.r64, S.r32, S.odone, S.sel);
    hipLaunchKernelGGL((kbx_resid_finish<NR>), dim3(S.P), t, 0, st, V);
    hipLaunchKernelGGL(kb_mix_after, dim3(1), dim3(256), 0, st, (const Seg *)B.seg, nseg, S.mix, S.sel, S.rst,
                       S.odone, (const int *)S.mit, rtol, max_iter);
    hipLaunchKernelGGL(kb_mark_done, dim3(nsb), dim3(256), 0, st, B.seg, nseg, (const int *)S.odone);
    hipLaunchKernelGGL((kbx_replace<NR>), g, t, 0, st, V32, (const float *)S.r32, S.z32, (const float *)S.d32,
                       (const int *)S.sel);
    hipLaunchKernelGGL((kbx_resume<NR>), g, t, 0, st, V32, S.p32, S.p32b, (const int *)S.sel, (const int *)S.rst);
    TFM_CHECK_HIP(S.s->ctx, hipGetLastError());
    return TFM_OK;
  }
};

template <int NR>
static int bsolver_solve_nr(BatchSolver *b, const tfm_solve_opts &o, bool warm, bool refresh, const SolveHooks *hk,
                            tfm_solve_stats *stats, double *fu) {
  tfm_system *s = b->s;
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  BBatch &B = b->B;
  const int P = b->P, nseg = P * NR;
  const bool use_graph = !getenv("TFM_NO_GRAPH");
  int rc = TFM_OK;
  if (refresh && (rc = refresh_prepare(b))) return rc;
  const uint64_t tag = hk ? hk->tag : 0;
  const bool scalar = !b->mixed && B.max_pnb > scalar_blocks();
  const uint64_t key = key_of({(uint64_t)o.mode, (uint64_t)o.max_iter, bits_of(o.rtol), bits_of(o.inner_rtol),
                               (uint64_t)warm, (uint64_t)refresh, tag, (uint64_t)(B.v32 != nullptr),
                               (uint64_t)scalar});
  // A: the caller's pre-stage and the refresh
  if ((refresh || (hk && hk->pre)) && (rc = stage_run(ctx, b->sa, key, use_graph, [&]() -> int {
        int r = TFM_OK;
        if (hk && hk->pre) r = hk->pre();
        if (!r && refresh) r = refresh_enqueue(b);
        return r;
      })))
    return rc;
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev0, st));
  std::vector<Seg> hs(nseg);
  std::vector<Mix> hm;
  int chunks = 0;
  auto drive = [&](auto &cg, void *snap, size_t snap_bytes, hipEvent_t ev, const void *src, auto all_done) -> int {
    int r = stage_run(ctx, b->sb, key, use_graph, [&]() { return cg.init(warm); });
    if (r) return r;
    const int CH = cg.chunk_len();
    const int cap = o.max_iter / CH + 4;
    auto launch = [&]() { return stage_run(ctx, b->sk, key, use_graph, [&]() { return cg.chunk(); }); };
    auto snapshot = [&]() -> bool {
      return hipMemcpyAsync(snap, src, snap_bytes, hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipEventRecord(ev, st) == hipSuccess;
    };
    // chunk k + 1 is queued before the host reads the state after chunk k;
    // a chunk past convergence only runs kernels that exit at once
    if ((r = launch())) return r;
    chunks = 1;
    if (!snapshot()) return fail(ctx, TFM_E_HIP, "batched solve: snapshot copy failed");
    for (;;) {
      if ((r = launch())) return r;
      chunks++;
      if (hipEventSynchronize(ev) != hipSuccess) return fail(ctx, TFM_E_HIP, "batched solve: event wait failed");
      if (all_done() || chunks > cap) break;
      if (!snapshot()) return fail(ctx, TFM_E_HIP, "batched solve: snapshot copy failed");
    }
    return TFM_OK;
  };
  if (!b->mixed) {
    PlainCg<NR> cg(b, o, scalar);
    rc = drive(cg, B.hseg, sizeof(Seg) * nseg, B.ev_seg, B.seg, [&]() {
      for (int q = 0; q < nseg; q++)
        if (!B.hseg[q].done) return false;
      return true;
    });
  } else {
    MixedCg<NR> cg(b, o);
    rc = drive(cg, b->hmix, sizeof(Mix) * nseg, b->ev_mix, b->mix, [&]() {
      for (int q = 0; q < nseg; q++)
        if (!b->hmix[q].od) return false;
      return true;
    });
  }
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev1, st));
  // C: the FP64 residual certificate of every segment, the caller's post-stage
  if (!rc)
    rc = stage_run(ctx, b->sc, key, use_graph, [&]() -> int {
      const BView V0 = B.view(o.max_iter, 0.0);
      bb_resid<NR>(s, B, V0, st, b->x, b->f, b->r64, nullptr, nullptr, nullptr);
      hipLaunchKernelGGL((kbx_resid_finish<NR>), dim3(P), dim3(kBT), 0, st, V0);
      TFM_CHECK_HIP(ctx, hipGetLastError());
      return (hk && hk->post) ? hk->post() : TFM_OK;
    });
  if (rc) {
    (void)hipStreamSynchronize(st);
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hs.data(), B.seg, sizeof(Seg) * nseg, hipMemcpyDeviceToHost, st));
  if (b->mixed) {
    hm.resize(nseg);
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(hm.data(), b->mix, sizeof(Mix) * nseg, hipMemcpyDeviceToHost, st));
  }
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  if (refresh && (rc = bad_check(b))) return rc;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  bool worst_ok = true;
  for (int q = 0; q < nseg; q++) {
    const double rel = hs[q].ff > 0 ? std::sqrt(hs[q].rt / hs[q].ff) : 0.0;
    int iters, cycles, brk, od;
    if (!b->mixed) {
      iters = hs[q].iters;
      brk = hs[q].breakdown;
      cycles = 1;
      od = hs[q].done && !hs[q].breakdown && hs[q].iters < o.max_iter;  // converged on its recurrence residual
    } else {
      iters = hm[q].iters ? hm[q].iters : hs[q].iters;  // segments still running count their own
      brk = hm[q].brk;
      cycles = hm[q].cycles;
      od = hm[q].od;
    }
    // plain: PCG converged on its recurrence residual (the true residual is
    // reported); mixed: the FP64 refinement residual itself met rtol
    worst_ok &= (b->mixed ? rel <= o.rtol : od != 0) && !brk;
    if (stats) {
      stats[q].iterations = iters;
      stats[q].outer = cycles;
      stats[q].rel_residual = rel;
      stats[q].seconds = ms * 1e-3;
    }
    if (fu) fu[q] = hs[q].fu;
  }
  if (!worst_ok) return fail(ctx, TFM_E_NOT_CONVERGED, "a batched segment missed the tolerance");
  return TFM_OK;
}

// One batched solve from x = 0 (or, warm, from the last solution): per
// segment stats and f.u (the compliance of a load case) when asked for.
// refresh: copy the system's current values into the solver first.
int bsolver_solve(BatchSolver *b, const tfm_solve_opts &o, bool warm, bool refresh, const SolveHooks *hooks,
                  tfm_solve_stats *stats, double *fu) {
  return b->NR == 1 ? bsolver_solve_nr<1>(b, o, warm, refresh, hooks, stats, fu)
                    : bsolver_solve_nr<8>(b, o, warm, refresh, hooks, stats, fu);
}

int bsolver_run(BatchSolver *b, const tfm_solve_opts &o, bool warm, tfm_solve_stats *stats, double *fu) {
  return bsolver_solve(b, o, warm, false, nullptr, stats, fu);
}

int system_solve_mg(tfm_system *s, const tfm_solve_opts &o, double *u, tfm_solve_stats *stats) {
  tfm_ctx *ctx = s->ctx;
  int rc;
  if (!s->mg_solver) {
    const int64_t off[2] = {0, s->ndof};
    if ((rc = bsolver_create(s, 1, off, 1, TFM_SOLVE_PCG64_MG, &s->mg_solver))) return rc;
  }
  BatchSolver *b = s->mg_solver;
  if ((rc = bsolver_set_rhs(b, nullptr))) return rc;
  rc = bsolver_solve(b, o, false, true, nullptr, stats, nullptr);
  if (rc && rc != TFM_E_NOT_CONVERGED) return rc;
  // x is 0 at the fixed DOFs: u = x, then the Dirichlet values in bcs order
  const hipError_t e = hipMemcpy(u, b->x, sizeof(double) * s->ndof, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  for (size_t q = 0; q < s->bc_dof.size(); q++) u[s->bc_dof[q]] = s->bc_val[q];
  return rc;
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
  m->grid_P = P; m->grid_nx = nx; m->grid_ny = ny;
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

// The system's batched solver for (P, DOF ranges, rhs count, mode), built on
// first use and kept with its captured graphs.
static int batch_solver_for(tfm_system *s, int64_t P, const int64_t *dof_off, int nrhs, int mode, BatchSolver **out) {
  std::vector<uint64_t> parts{(uint64_t)P, (uint64_t)nrhs, (uint64_t)mode};
  uint64_t key = key_of({(uint64_t)P, (uint64_t)nrhs, (uint64_t)mode});
  for (int64_t p = 0; p <= P; p++) key = key_of({key, (uint64_t)dof_off[p]});
  if (!s->batch_solver || s->batch_key != key) {
    bsolver_destroy(s->batch_solver);
    s->batch_solver = nullptr;
    s->batch_key = 0;
    const int rc = bsolver_create(s, (int)P, dof_off, nrhs, mode, &s->batch_solver);
    if (rc) return rc;
    s->batch_key = key;
  }
  *out = s->batch_solver;
  return TFM_OK;
}

static int solve_batch(tfm_system *s, int64_t P, const int64_t *dof_off, int nrhs, const double *rhs,
                       const tfm_solve_opts *opts, double *u, tfm_solve_stats *stats, const SolveHooks *hooks) {
  if (!s || !dof_off || !u || P <= 0 || nrhs < 1) return TFM_E_INVALID;
  tfm_ctx *ctx = s->ctx;
  if (dof_off[0] != 0 || dof_off[P] != s->ndof) return fail(ctx, TFM_E_INVALID, "dof offsets must cover the system");
  for (int64_t p = 0; p < P; p++)
    if (dof_off[p + 1] < dof_off[p]) return fail(ctx, TFM_E_INVALID, "dof offsets must ascend");
  if (nrhs > 1 && !rhs) return fail(ctx, TFM_E_INVALID, "several right-hand sides need rhs");
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-10, 20000, 0.3};
  if (opts) o = *opts;
  const int64_t n = s->ndof;
  if (o.mode == TFM_SOLVE_DIRECT) {
    // solve_multi: one factorisation of the whole (block-diagonal) system,
    // every right-hand side solved against it
    double *d_F = nullptr;
    if (rhs) {
      if ((rc = dalloc(ctx, &d_F, (size_t)n * nrhs))) return rc;
      if (hipMemcpy(d_F, rhs, sizeof(double) * n * nrhs, hipMemcpyHostToDevice) != hipSuccess) {
        dfree(d_F);
        return fail(ctx, TFM_E_HIP, "rhs upload failed");
      }
    }
    std::vector<tfm_solve_stats> st((size_t)nrhs);
    rc = system_solve_direct(s, nrhs, d_F, u, st.data());
    dfree(d_F);
    if (!rc && stats)
      for (int64_t p = 0; p < P; p++)
        for (int k = 0; k < nrhs; k++) stats[p * nrhs + k] = st[k];
    return rc;
  }
  if (nrhs != 1 && nrhs != 8) {
    // the kernels are built for 1 and 8 interleaved right-hand sides: chunks
    // of 8, the remainder one at a time
    std::vector<tfm_solve_stats> st((size_t)P * 8);
    int worst = TFM_OK;
    for (int k0 = 0; k0 < nrhs;) {
      const int c = nrhs - k0 >= 8 ? 8 : 1;
      rc = solve_batch(s, P, dof_off, c, rhs + (size_t)k0 * n, &o, u + (size_t)k0 * n, st.data(), nullptr);
      if (rc && rc != TFM_E_NOT_CONVERGED) return rc;
      if (rc) worst = rc;
      if (stats)
        for (int64_t p = 0; p < P; p++)
          for (int j = 0; j < c; j++) stats[p * nrhs + k0 + j] = st[(size_t)p * c + j];
      k0 += c;
    }
    return worst;
  }
  // A segment's rows must not couple to another problem's rows (block
  // diagonal): the kernels reduce per problem and would silently mix them.
  BatchSolver *b = nullptr;
  if ((rc = batch_solver_for(s, P, dof_off, nrhs, o.mode, &b)) || (rc = bsolver_set_rhs(b, rhs))) return rc;
  rc = bsolver_solve(b, o, false, true, hooks, stats, nullptr);
  if (rc && rc != TFM_E_NOT_CONVERGED) return rc;
  if (nrhs == 1) {
    // x is 0 at the fixed DOFs: u = x, then the Dirichlet values in bcs order
    hipError_t e = hipMemcpy(u, b->x, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    for (size_t q = 0; q < s->bc_dof.size(); q++) u[s->bc_dof[q]] = s->bc_val[q];
    return rc;
  }
  std::vector<double> hx((size_t)n * nrhs);
  hipError_t e = hipMemcpy(hx.data(), b->x, sizeof(double) * hx.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  // u[k][*]: free entries from x, then Dirichlet values in bcs order
  for (int k = 0; k < nrhs; k++) {
    double *uk = u + (size_t)k * n;
    for (int64_t i = 0; i < n; i++) uk[i] = 0.0;
    for (int64_t fdof : s->free_dofs) uk[fdof] = hx[(size_t)fdof * nrhs + k];
    for (size_t q = 0; q < s->bc_dof.size(); q++) uk[s->bc_dof[q]] = s->bc_val[q];
  }
  return rc;
}

extern "C" int tfm_system_solve_batch(tfm_system *s, int64_t P, const int64_t *dof_off, int nrhs, const double *rhs,
                                      const tfm_solve_opts *opts, double *u, tfm_solve_stats *stats) {
  return solve_batch(s, P, dof_off, nrhs, rhs, opts, u, stats, nullptr);
}

extern "C" int tfm_system_assemble_solve_batch(tfm_system *s, const tfm_operator *op, int64_t P,
                                               const int64_t *dof_off, int nrhs, const double *rhs,
                                               const tfm_solve_opts *opts, double *u, tfm_solve_stats *stats) {
  if (!s) return TFM_E_INVALID;
  tfm_ctx *ctx = s->ctx;
  if (!s->has_plan) return fail(ctx, TFM_E_INVALID, "system was not built from a mesh");
  int rc = set_device(ctx);
  if (rc) return rc;
  const int mode = opts ? opts->mode : TFM_SOLVE_PCG64;
  if (mode == TFM_SOLVE_DIRECT || (nrhs != 1 && nrhs != 8)) {
    // no graph sequence for these: the re-assembly, then the solve
    if ((rc = tfm_system_assemble(s, op, 0))) return rc;
    return solve_batch(s, P, dof_off, nrhs, rhs, opts, u, stats, nullptr);
  }
  if (op) {
    if (op->kind != s->plan.op.kind || op->model != s->plan.op.model || op->quad_order != s->plan.op.quad_order)
      return fail(ctx, TFM_E_INVALID, "re-assembly must keep the operator kind, model and rule");
    if ((rc = plan_set_coefficients(ctx, &s->plan, op))) return rc;
    if (op->kind == TFM_OP_SIMP)
      TFM_CHECK_HIP(ctx, hipMemcpyAsync(s->plan.d_tab + (size_t)s->plan.nq * kRuleStride, s->plan.ke0,
                                        sizeof(double) * 64, hipMemcpyHostToDevice, ctx->stream));
  }
  // every value array the system holds is re-assembled in the solve's first
  // stage (the versions are noted first: the solver checks them)
  double *w64 = s->d_v64;
  float *w32 = s->d_v32;
  s->note_values(op != nullptr, w64 != nullptr, w32 != nullptr);
  SolveHooks hk;
  hk.pre = [s, ctx, w64, w32]() { return plan_numeric(ctx, &s->plan, w64, w32); };
  hk.tag = key_of({0x61736d62ull, (uint64_t)(w64 != nullptr), (uint64_t)(w32 != nullptr)});
  rc = solve_batch(s, P, dof_off, nrhs, rhs, opts, u, stats, &hk);
  if (rc && rc != TFM_E_NOT_CONVERGED) (void)plan_numeric(ctx, &s->plan, w64, w32);  // values match the coefficients
  return rc;
}
