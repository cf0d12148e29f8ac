// tfm_dense.hip -- the reference's direct solve on the GPU, for reduced
// systems small enough to factor densely.
//
// What is computed.  solve / solve_multi (crates/tpt-fem-sparse/src/lib.rs:
// 225-278): the reduced CSR scattered into a dense n x n matrix (duplicates
// summed), factored by partial-pivot LU and solved for every right-hand side.
// The pivot of column k is the first row i >= k with the largest |A(i,k)|;
// rows k and p swap across all columns (and in b); l = A(i,k) / A(k,k) is
// stored in place, and rows with l != 0 are updated A(i,j) -= l * A(k,j),
// b[i] -= l * b[k]; back substitution sums s -= A(i,j) * x[j] for j
// ascending and divides by A(i,i).  Each of these operations is one IEEE
// operation on the same operands as the sequential loop (this translation
// unit is built without FMA contraction), so factors and solutions are
// bit-identical to a sequential execution of the same loop.
//
// How.  Per column k two launches: a one-block panel kernel (pivot search by
// a block reduction that keeps the smallest index among equal maxima, the
// row swap, the multipliers and the right-hand-side update) and a trailing
// update over (i > k, j > k).  Back substitution runs one wavefront per
// right-hand side: the lanes form the products A(i,j) * x[j] of a 64-wide
// chunk, lane 0 subtracts them in order, x stays in LDS.  Dense storage is
// row-major, n^2 doubles, so the path is limited to kDirectMax unknowns.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "tfm_internal.h"

#pragma clang fp contract(off)

namespace tfm {

constexpr int kPanelThreads = 1024;

// Reduced CSR (global DOF numbering, fixed rows empty, fixed columns absent)
// into the dense row-major matrix of the free DOFs; red maps a global DOF to
// its reduced index.
__global__ void kd_fill(int nf, const int *__restrict__ free_g, const int *__restrict__ red,
                        const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
                        double *__restrict__ A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nf) return;
  const int g = free_g[q];
  double *row = A + (size_t)q * nf;
  for (int j = rp[g]; j < rp[g + 1]; j++) {
    const int c = red[ci[j]];
    if (c >= 0) row[c] = row[c] + v[j];
  }
}

// b[r][q] = f[free_g[q]] for right-hand side r (f: [nrhs][ndof]).
__global__ void kd_rhs(int nf, int nrhs, int64_t ndof, const int *__restrict__ free_g, const double *__restrict__ f,
                       double *__restrict__ b) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (q >= nf || r >= nrhs) return;
  b[(size_t)r * nf + q] = f[(size_t)r * ndof + free_g[q]];
}

// Column k: pivot, swap, multipliers, right-hand sides.  One block.
__global__ __launch_bounds__(kPanelThreads) void kd_panel(int n, int k, int nrhs, double *__restrict__ A,
                                                          double *__restrict__ b, int *__restrict__ piv,
                                                          int *__restrict__ singular) {
  __shared__ double s_best[kPanelThreads];
  __shared__ int s_idx[kPanelThreads];
  const int t = threadIdx.x;
  // local scan over i = k + t, k + t + T, ... keeps the first maximum
  double best = -1.0;
  int idx = n;
  for (int i = k + t; i < n; i += kPanelThreads) {
    const double a = fabs(A[(size_t)i * n + k]);
    if (i == k || a > best) {
      best = a;
      idx = i;
    }
  }
  s_best[t] = best;
  s_idx[t] = idx;
  __syncthreads();
  for (int w = kPanelThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      const double b2 = s_best[t + w];
      const int i2 = s_idx[t + w];
      if (b2 > s_best[t] || (b2 == s_best[t] && i2 < s_idx[t])) {
        s_best[t] = b2;
        s_idx[t] = i2;
      }
    }
    __syncthreads();
  }
  const int p = s_idx[0];
  if (s_best[0] == 0.0) {
    if (t == 0) *singular = 1;
    return;
  }
  if (t == 0) piv[k] = p;
  if (p != k) {
    for (int j = t; j < n; j += kPanelThreads) {
      const double x = A[(size_t)k * n + j];
      A[(size_t)k * n + j] = A[(size_t)p * n + j];
      A[(size_t)p * n + j] = x;
    }
    for (int r = t; r < nrhs; r += kPanelThreads) {
      const double x = b[(size_t)r * n + k];
      b[(size_t)r * n + k] = b[(size_t)r * n + p];
      b[(size_t)r * n + p] = x;
    }
  }
  __syncthreads();
  const double pv = A[(size_t)k * n + k];
  for (int i = k + 1 + t; i < n; i += kPanelThreads) {
    const double l = A[(size_t)i * n + k] / pv;
    A[(size_t)i * n + k] = l;
    if (l != 0.0)
      for (int r = 0; r < nrhs; r++) b[(size_t)r * n + i] -= l * b[(size_t)r * n + k];
  }
}

// Trailing update A(i,j) -= l_i * A(k,j) for i > k, j > k (rows with l_i = 0
// unchanged).  blockIdx.y: a row, blockIdx.x * 256 + tid: a column offset.
__global__ void kd_update(int n, int k, double *__restrict__ A) {
  const int i = k + 1 + blockIdx.y;
  const int j = k + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || j >= n) return;
  const double l = A[(size_t)i * n + k];
  if (l != 0.0) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
}

// Back substitution, one 64-lane block per right-hand side; x in LDS.
__global__ __launch_bounds__(64) void kd_backsub(int n, const double *__restrict__ A, const double *__restrict__ b,
                                                 double *__restrict__ x) {
  extern __shared__ double s_x[];  // n + 64
  double *s_p = s_x + n;
  const int r = blockIdx.x, t = threadIdx.x;
  const double *br = b + (size_t)r * n;
  for (int i = n - 1; i >= 0; i--) {
    const double *Ai = A + (size_t)i * n;
    double s = br[i];
    for (int j0 = i + 1; j0 < n; j0 += 64) {
      const int j = j0 + t;
      if (j < n) s_p[t] = Ai[j] * s_x[j];
      __syncthreads();
      if (t == 0) {
        const int m = min(64, n - j0);
        for (int u = 0; u < m; u++) s = s - s_p[u];
      }
      __syncthreads();
    }
    if (t == 0) s_x[i] = s / Ai[i];
    __syncthreads();
  }
  for (int q = t; q < n; q += 64) x[(size_t)r * n + q] = s_x[q];
}

// x (reduced, [nrhs][nf]) into global numbering ([nrhs][ndof], fixed DOFs 0).
__global__ void kd_scatter(int nf, int nrhs, int64_t ndof, const int *__restrict__ free_g,
                           const double *__restrict__ x, double *__restrict__ xg) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (q >= nf || r >= nrhs) return;
  xg[(size_t)r * ndof + free_g[q]] = x[(size_t)r * nf + q];
}

int direct_solve_dev(tfm_system *s, int nrhs, const double *d_f, double *d_x) {
  tfm_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t nf64 = s->n_free, ndof = s->ndof;
  if (nf64 > kDirectMax)
    return fail(ctx, TFM_E_UNSUPPORTED, "direct solve: more free DOFs than the dense path holds");
  if (!s->d_v64) return fail(ctx, TFM_E_INVALID, "solve needs FP64 values (TFM_SYS_FP64)");
  const int nf = (int)nf64;
  if (nf == 0)
    return fail(ctx, TFM_E_NUMERIC,
                "sparse numeric factorization/solve failed: rhs length 0 does not match matrix dimension 1");
  std::vector<int> fg((size_t)nf), red((size_t)ndof, -1);
  for (int q = 0; q < nf; q++) {
    fg[q] = (int)s->free_dofs[q];
    red[fg[q]] = q;
  }
  int *d_fg = nullptr, *d_red = nullptr, *d_piv = nullptr, *d_sing = nullptr;
  double *A = nullptr, *b = nullptr, *x = nullptr;
  auto cleanup = [&]() {
    dfree(d_fg);
    dfree(d_red);
    dfree(d_piv);
    dfree(d_sing);
    dfree(A);
    dfree(b);
    dfree(x);
  };
  int rc;
  if ((rc = dalloc(ctx, &d_fg, nf)) || (rc = dalloc(ctx, &d_red, ndof)) || (rc = dalloc(ctx, &d_piv, nf)) ||
      (rc = dalloc(ctx, &d_sing, 1)) || (rc = dalloc(ctx, &A, (size_t)nf * nf)) ||
      (rc = dalloc(ctx, &b, (size_t)nf * nrhs)) || (rc = dalloc(ctx, &x, (size_t)nf * nrhs))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipMemcpyAsync(d_fg, fg.data(), sizeof(int) * nf, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_red, red.data(), sizeof(int) * ndof, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(A, 0, sizeof(double) * nf * nf, st);
  if (e == hipSuccess) e = hipMemsetAsync(d_sing, 0, sizeof(int), st);
  if (e != hipSuccess) {
    cleanup();
    return fail(ctx, TFM_E_HIP, std::string("direct solve setup: ") + hipGetErrorString(e));
  }
  const unsigned g1 = (unsigned)((nf + 255) / 256);
  hipLaunchKernelGGL(kd_fill, dim3(g1), dim3(256), 0, st, nf, (const int *)d_fg, (const int *)d_red, s->d_rp,
                     s->d_ci, (const double *)s->d_v64, A);
  hipLaunchKernelGGL(kd_rhs, dim3(g1, (unsigned)nrhs), dim3(256), 0, st, nf, nrhs, ndof, (const int *)d_fg, d_f, b);
  for (int k = 0; k < nf; k++) {
    hipLaunchKernelGGL(kd_panel, dim3(1), dim3(kPanelThreads), 0, st, nf, k, nrhs, A, b, d_piv, d_sing);
    const int m = nf - k - 1;
    if (m > 0) hipLaunchKernelGGL(kd_update, dim3((unsigned)((m + 255) / 256), (unsigned)m), dim3(256), 0, st, nf, k, A);
    // a zero pivot stops the factorisation: check every 256 columns
    if ((k & 255) == 255 || k == nf - 1) {
      int hs = 0;
      e = hipMemcpyAsync(&hs, d_sing, sizeof(int), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) {
        cleanup();
        return fail(ctx, TFM_E_HIP, std::string("direct solve: ") + hipGetErrorString(e));
      }
      if (hs) {
        cleanup();
        return fail(ctx, TFM_E_NUMERIC, "sparse numeric factorization/solve failed: singular matrix");
      }
    }
  }
  hipLaunchKernelGGL(kd_backsub, dim3((unsigned)nrhs), dim3(64), sizeof(double) * (nf + 64), st, nf,
                     (const double *)A, (const double *)b, x);
  e = hipMemsetAsync(d_x, 0, sizeof(double) * ndof * nrhs, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(kd_scatter, dim3(g1, (unsigned)nrhs), dim3(256), 0, st, nf, nrhs, ndof, (const int *)d_fg,
                       (const double *)x, d_x);
    e = hipStreamSynchronize(st);
  }
  cleanup();
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("direct solve: ") + hipGetErrorString(e));
  return TFM_OK;
}

}  // namespace tfm
