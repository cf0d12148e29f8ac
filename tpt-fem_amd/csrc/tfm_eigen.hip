// tfm_eigen.hip -- Lanczos eigenpairs of an assembled matrix on the GPU.
//
// lanczos_eigs (crates/tpt-fem-eigen/src/lib.rs:223-268): a Lanczos basis of
// dimension m from e_0 with full (modified Gram-Schmidt) re-orthogonalisation
// against every earlier vector (:135-167), the m x m tridiagonal T solved by
// the reference's cyclic Jacobi (:170-218) on the host, and the selected Ritz
// vectors formed back on the device.  The basis lives in HBM as m vectors of
// n doubles; the SpMV is the reference's CSR row sum (from 0.0, CSR order)
// and every dot product a fixed-order two-pass reduction, so a run is
// deterministic.  Only the Lanczos scalars and the Ritz vectors cross PCIe.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tfm_internal.h"

namespace tfm {

constexpr int kEB = 256;  // threads per block
constexpr int kEG = 512;  // blocks of a dot-product pass (fixed: the sum order never changes)

// matvec (:24-35): y[r] = sum over row r in CSR order, from 0.0.
__global__ __launch_bounds__(kEB) void ke_spmv(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                               const double *__restrict__ v, const double *__restrict__ x,
                                               double *__restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int c = rp[r]; c < rp[r + 1]; c++) s += v[c] * x[ci[c]];
  y[r] = s;
}

// Per-block partial of a . b (thread-strided, then a fixed block tree).
__global__ __launch_bounds__(kEB) void ke_dot_part(int n, const double *__restrict__ a, const double *__restrict__ b,
                                                   double *__restrict__ part) {
  __shared__ double sh[kEB / 64];
  double s = 0.0;
  for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += kEG * kEB) s += a[i] * b[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kEB / 64; q++) t += sh[q];
    part[blockIdx.x] = t;
  }
}

// out[0] = sum of the kEG partials (fixed order), or its square root.
__global__ __launch_bounds__(64) void ke_dot_fin(const double *__restrict__ part, double *__restrict__ out,
                                                 int root) {
  double s = 0.0;
  for (int i = threadIdx.x; i < kEG; i += 64) s += part[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x == 0) out[0] = root ? sqrt(s) : s;
}

// w -= p * v with p on the device (one Gram-Schmidt projection).
__global__ __launch_bounds__(kEB) void ke_axpy(int n, double *__restrict__ w, const double *__restrict__ v,
                                               const double *__restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] -= p[0] * v[i];
}

// out = w / b with b on the device (the reference divides, it does not
// multiply by the reciprocal).
__global__ __launch_bounds__(kEB) void ke_div(int n, const double *__restrict__ w, const double *__restrict__ b,
                                              double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = w[i] / b[0];
}

// Ritz vector: out[d] = sum_r c[r] * V[r][d] in basis order, from 0.0.
__global__ __launch_bounds__(kEB) void ke_ritz(int n, int m, const double *__restrict__ V,
                                               const double *__restrict__ c, double *__restrict__ out) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n) return;
  double s = 0.0;
  for (int r = 0; r < m; r++) s += c[r] * V[(size_t)r * n + d];
  out[d] = s;
}

// jacobi (:170-218), host: cyclic sweeps over p < q rotating every pair whose
// off-diagonal magnitude exceeds 1e-12, until a sweep changes nothing or
// max_sweeps; a (row-major m x m) is overwritten, q receives the
// eigenvectors as columns.
static void host_jacobi(int m, std::vector<double> &a, int max_sweeps, std::vector<double> &eig,
                        std::vector<double> &q) {
  q.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++) q[(size_t)i * m + i] = 1.0;
  for (int sw = 0; sw < max_sweeps; sw++) {
    bool changed = false;
    for (int p = 0; p < m; p++)
      for (int r = p + 1; r < m; r++) {
        const double apr = a[(size_t)p * m + r];
        if (!(std::fabs(apr) > 1e-12)) continue;
        changed = true;
        const double theta = (a[(size_t)r * m + r] - a[(size_t)p * m + p]) / (2.0 * apr);
        const double t = theta >= 0.0 ? 1.0 / (theta + std::sqrt(1.0 + theta * theta))
                                      : 1.0 / (theta - std::sqrt(1.0 + theta * theta));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = t * c;
        for (int i = 0; i < m; i++) {
          const double aip = a[(size_t)i * m + p], aiq = a[(size_t)i * m + r];
          a[(size_t)i * m + p] = c * aip - s * aiq;
          a[(size_t)i * m + r] = s * aip + c * aiq;
        }
        for (int i = 0; i < m; i++) {
          const double api = a[(size_t)p * m + i], aqi = a[(size_t)r * m + i];
          a[(size_t)p * m + i] = c * api - s * aqi;
          a[(size_t)r * m + i] = s * api + c * aqi;
        }
        for (int i = 0; i < m; i++) {
          const double vip = q[(size_t)i * m + p], viq = q[(size_t)i * m + r];
          q[(size_t)i * m + p] = c * vip - s * viq;
          q[(size_t)i * m + r] = s * vip + c * viq;
        }
      }
    if (!changed) break;
  }
  eig.resize(m);
  for (int i = 0; i < m; i++) eig[i] = a[(size_t)i * m + i];
}

}  // namespace tfm

using namespace tfm;

extern "C" int tfm_lanczos_eigs(const tfm_matrix *A, int num, int which, int lanczos_dim, double *evals,
                                double *evecs, int *count) {
  if (!A) return TFM_E_INVALID;
  tfm_ctx *ctx = A->ctx;
  if (lanczos_dim < 1 || num < 0 || (which != TFM_EIG_SMALLEST && which != TFM_EIG_LARGEST) ||
      (num > 0 && !evals))
    return fail(ctx, TFM_E_INVALID, "lanczos_eigs: lanczos_dim >= 1, num >= 0 and a valid end of the spectrum");
  if (A->nrows != A->ncols) return fail(ctx, TFM_E_INVALID, "lanczos_eigs: the matrix must be square");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const int n = (int)A->nrows, md = lanczos_dim;
  const unsigned gn = (unsigned)((n + kEB - 1) / kEB);
  double *V = nullptr, *w = nullptr, *part = nullptr, *sc = nullptr, *cq = nullptr, *vec = nullptr;
  auto cleanup = [&]() {
    dfree(V); dfree(w); dfree(part); dfree(sc); dfree(cq); dfree(vec);
  };
  // sc: [0] projection p, [1] norm, [2 ..] alpha (md), then beta (md)
  if ((rc = dalloc(ctx, &V, (size_t)md * n)) || (rc = dalloc(ctx, &w, n)) || (rc = dalloc(ctx, &part, kEG)) ||
      (rc = dalloc(ctx, &sc, 2 + 2 * (size_t)md)) || (rc = dalloc(ctx, &cq, md)) || (rc = dalloc(ctx, &vec, n))) {
    cleanup();
    return rc;
  }
  auto dot = [&](const double *a, const double *b, double *out, int root) {
    hipLaunchKernelGGL(ke_dot_part, dim3(kEG), dim3(kEB), 0, st, n, a, b, part);
    hipLaunchKernelGGL(ke_dot_fin, dim3(1), dim3(64), 0, st, (const double *)part, out, root);
  };
  const double one = 1.0;
  TFM_CHECK_HIP(ctx, hipMemsetAsync(V, 0, sizeof(double) * n, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  TFM_CHECK_HIP(ctx, hipMemcpy(V, &one, sizeof(double), hipMemcpyHostToDevice));  // e_0, norm 1
  double *alpha = sc + 2, *beta = sc + 2 + md;
  int m = 1;
  for (int j = 0; j < md; j++) {
    const double *vj = V + (size_t)j * n;
    hipLaunchKernelGGL(ke_spmv, dim3(gn), dim3(kEB), 0, st, n, (const int *)A->d_rp, (const int *)A->d_ci,
                       (const double *)A->d_v, vj, w);
    dot(w, vj, alpha + j, 0);
    for (int k = 0; k <= j; k++) {
      const double *vk = V + (size_t)k * n;
      dot(w, vk, sc, 0);
      hipLaunchKernelGGL(ke_axpy, dim3(gn), dim3(kEB), 0, st, n, w, vk, (const double *)sc);
    }
    if (j + 1 == md) break;
    dot(w, w, beta + j, 1);
    double b = 0.0;
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(&b, beta + j, sizeof(double), hipMemcpyDeviceToHost, st));
    TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
    if (b < 1e-12) break;  // an invariant subspace: the basis stops here
    hipLaunchKernelGGL(ke_div, dim3(gn), dim3(kEB), 0, st, n, (const double *)w, (const double *)(beta + j),
                       V + (size_t)(j + 1) * n);
    m++;
  }
  std::vector<double> h(2 * (size_t)md);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(h.data(), sc + 2, sizeof(double) * 2 * md, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  std::vector<double> t((size_t)m * m, 0.0), eig, q;
  for (int i = 0; i < m; i++) t[(size_t)i * m + i] = h[i];
  for (int i = 0; i + 1 < m; i++) t[(size_t)i * m + i + 1] = t[(size_t)(i + 1) * m + i] = h[md + i];
  host_jacobi(m, t, 200, eig, q);
  // stable ascending order of the Ritz values (sort_by with partial_cmp)
  std::vector<int> idx(m);
  for (int i = 0; i < m; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return eig[a] < eig[b]; });
  const int take = std::min(num, m);
  std::vector<double> col(m);
  for (int s = 0; s < take; s++) {
    const int i = which == TFM_EIG_SMALLEST ? idx[s] : idx[m - 1 - s];
    evals[s] = eig[i];
    if (!evecs) continue;
    for (int r = 0; r < m; r++) col[r] = q[(size_t)r * m + i];
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(cq, col.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(ke_ritz, dim3(gn), dim3(kEB), 0, st, n, m, (const double *)V, (const double *)cq, vec);
    dot(vec, vec, sc + 1, 1);
    hipLaunchKernelGGL(ke_div, dim3(gn), dim3(kEB), 0, st, n, (const double *)vec, (const double *)(sc + 1), w);
    TFM_CHECK_HIP(ctx, hipMemcpyAsync(evecs + (size_t)s * n, w, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  }
  if (count) *count = take;
  hipError_t e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("lanczos_eigs: ") + hipGetErrorString(e));
  return TFM_OK;
}
