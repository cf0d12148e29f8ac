// tfm_dynamic.hip -- time integration of M u'' + C u' + K u = f(t) on the GPU.
//
// CachedSystem (crates/tpt-fem-dynamic/src/lib.rs:133-166) is three device
// CSR matrices; the integrators keep u, v and a in HBM for the whole run and
// write every step's displacements into a device history that is copied out
// once.  The load is f(t_k) = fscale[k] * F.
//
//   central_difference (:289-377): lumped mass = M's row sums, the CFL guard
//     from max_i sum_j |K_ij| / m_i, then one kernel per step (two CSR row
//     sums, the acceleration, the half-step velocity and the displacement),
//     in the reference's operation order: bit-identical to it.
//   newmark (:194-263): the constant effective stiffness
//     K_hat = K + M/(b dt^2) + C g/(b dt) is formed once on the device (row
//     merge of the three patterns, each entry summed in that order); every
//     step solves K_hat u = rhs by the batched PCG, warm-started from the
//     previous displacements, where the reference factors a dense matrix.
#include <cmath>
#include <vector>

#include "tfm_internal.h"

namespace tfm {

struct DCsr {
  int nrows;
  const int *rp, *ci;
  const double *v;
};

static DCsr dcsr(const tfm_matrix *m) { return DCsr{(int)m->nrows, m->d_rp, m->d_ci, m->d_v}; }

// Csr::matvec row (crates/tpt-fem-sparse/src/lib.rs:195-205): from 0.0 in CSR
// order; rows past the matrix are 0.
__device__ __forceinline__ double row_dot(const DCsr &a, int r, const double *__restrict__ x) {
  double s = 0.0;
  if (r < a.nrows)
    for (int c = a.rp[r]; c < a.rp[r + 1]; c++) s += a.v[c] * x[a.ci[c]];
  return s;
}

// Lumped mass m_r (row sum of M from 0.0) and the CFL quantity
// sum_j |K_rj| / m_r for rows with m_r > 0 (-1 elsewhere).
__global__ void kd_lump(int n, DCsr M, DCsr K, double *__restrict__ md, double *__restrict__ rho) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int c = M.rp[r]; c < M.rp[r + 1]; c++) s += M.v[c];
  md[r] = s;
  double q = -1.0;
  if (s > 0.0) {
    double rs = 0.0;
    for (int c = K.rp[r]; c < K.rp[r + 1]; c++) rs += fabs(K.v[c]);
    q = rs / s;
  }
  rho[r] = q;
}

// a0 = (f0 - C v0 - K u0) / m, v_{1/2} = v0 - 0.5 dt a0, history row 0.
__global__ void kd_cd_start(int n, DCsr D, DCsr K, const double *__restrict__ F, double s0, double dt,
                            const double *__restrict__ md, const double *__restrict__ u0,
                            const double *__restrict__ v0, double *__restrict__ u, double *__restrict__ vh,
                            double *__restrict__ hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = (s0 * F[i] - row_dot(D, i, v0) - row_dot(K, i, u0)) / md[i];
  u[i] = u0[i];
  vh[i] = v0[i] - 0.5 * dt * a;
  hist[i] = u0[i];
}

// One step: a = (f_k - C v_{1/2} - K u) / m, v += dt a, u += dt v.
__global__ void kd_cd_step(int n, DCsr D, DCsr K, const double *__restrict__ F, double sk, double dt,
                           const double *__restrict__ md, const double *__restrict__ u,
                           const double *__restrict__ vh, double *__restrict__ un, double *__restrict__ vhn,
                           double *__restrict__ hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = (sk * F[i] - row_dot(D, i, vh) - row_dot(K, i, u)) / md[i];
  const double v = vh[i] + dt * a;
  const double x = u[i] + dt * v;
  vhn[i] = v;
  un[i] = x;
  hist[i] = x;
}

// ---- K_hat = K + s1 M + s2 C (rows of n, pattern union) ----------------------
__device__ __forceinline__ int next_col(const DCsr &a, int r, int j) {
  return (r < a.nrows && j < a.rp[r + 1]) ? a.ci[j] : INT_MAX;
}

__global__ void kd_union_count(int n, DCsr A, DCsr B, DCsr Cm, int *__restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int ja = r < A.nrows ? A.rp[r] : 0, jb = r < B.nrows ? B.rp[r] : 0, jc = r < Cm.nrows ? Cm.rp[r] : 0;
  int k = 0;
  for (;;) {
    const int ca = next_col(A, r, ja), cb = next_col(B, r, jb), cc = next_col(Cm, r, jc);
    const int m = min(ca, min(cb, cc));
    if (m == INT_MAX) break;
    ja += ca == m;
    jb += cb == m;
    jc += cc == m;
    k++;
  }
  cnt[r] = k;
}

// Each entry is Coo::to_csr's collapse of coo_add(K, coo_add(s1 M, s2 C))
// (dynamic/src/lib.rs:221-229): the first present value, then the others
// added in that order.
__global__ void kd_union_fill(int n, DCsr A, DCsr B, DCsr Cm, double s1, double s2, const int *__restrict__ rp,
                              int *__restrict__ ci, double *__restrict__ v) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int ja = r < A.nrows ? A.rp[r] : 0, jb = r < B.nrows ? B.rp[r] : 0, jc = r < Cm.nrows ? Cm.rp[r] : 0;
  int o = rp[r];
  for (;;) {
    const int ca = next_col(A, r, ja), cb = next_col(B, r, jb), cc = next_col(Cm, r, jc);
    const int m = min(ca, min(cb, cc));
    if (m == INT_MAX) break;
    bool have = false;
    double s = 0.0;
    if (ca == m) { s = A.v[ja++]; have = true; }
    if (cb == m) { const double t = B.v[jb++] * s1; s = have ? s + t : t; have = true; }
    if (cc == m) { const double t = Cm.v[jc++] * s2; s = have ? s + t : t; }
    ci[o] = m;
    v[o] = s;
    o++;
  }
}

// ---- Newmark per-step vectors -------------------------------------------------
// p_m = u/(b dt^2) + v/(b dt) + (0.5/b - 1) a and
// p_c = g/(b dt) u + (g/b - 1) v + dt (g/(2b) - 1) a (dynamic/src/lib.rs:242-247)
struct NmCoef {
  double d_u, d_v, d_a;       // g/(b dt), g/b - 1, dt (g/(2b) - 1)
  double bdt2, bdt, h, dt, g; // b dt^2, b dt, 0.5/b - 1, dt, g
};

__global__ void kd_nm_p(int n, NmCoef c, const double *__restrict__ u, const double *__restrict__ v,
                        const double *__restrict__ a, double *__restrict__ pm, double *__restrict__ pc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pm[i] = u[i] / c.bdt2 + v[i] / c.bdt + c.h * a[i];
  pc[i] = c.d_u * u[i] + c.d_v * v[i] + c.d_a * a[i];
}

__global__ void kd_nm_rhs(int n, DCsr M, DCsr D, const double *__restrict__ F, double sk,
                          const double *__restrict__ pm, const double *__restrict__ pc, double *__restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rhs[i] = sk * F[i] + row_dot(M, i, pm) + row_dot(D, i, pc);
}

__global__ void kd_nm_update(int n, NmCoef c, const double *__restrict__ un, double *__restrict__ u,
                             double *__restrict__ v, double *__restrict__ a, double *__restrict__ hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double an = (un[i] - u[i]) / c.bdt2 - v[i] / c.bdt - c.h * a[i];
  v[i] = v[i] + c.dt * ((1.0 - c.g) * a[i] + c.g * an);
  a[i] = an;
  u[i] = un[i];
  hist[i] = un[i];
}

__global__ void kd_sub3(int n, const double *__restrict__ F, double s0, DCsr D, DCsr K,
                        const double *__restrict__ u0, const double *__restrict__ v0, double *__restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = s0 * F[i] - row_dot(D, i, v0) - row_dot(K, i, u0);
}

static int check_ops(tfm_ctx *ctx, const tfm_matrix *M, const tfm_matrix *C, const tfm_matrix *K, int64_t n) {
  if (!M || !C || !K) return fail(ctx, TFM_E_INVALID, "dynamic: null operator");
  if (n < 1 || n >= INT32_MAX) return fail(ctx, TFM_E_INVALID, "dynamic: bad vector length");
  if (M->nrows < n || K->nrows < n)
    return fail(ctx, TFM_E_INVALID, "dynamic: mass and stiffness need a row for every DOF");
  // Csr::matvec indexes x by column: every column must be a DOF
  if (M->ncols > n || C->ncols > n || K->ncols > n)
    return fail(ctx, TFM_E_INVALID, "dynamic: an operator has a column beyond the vector length");
  if (M->ctx != ctx || C->ctx != ctx || K->ctx != ctx)
    return fail(ctx, TFM_E_INVALID, "dynamic: operators from different contexts");
  return TFM_OK;
}

}  // namespace tfm

using namespace tfm;

extern "C" int tfm_central_difference(const tfm_matrix *M, const tfm_matrix *C, const tfm_matrix *K, int64_t n,
                                      const double *u0, const double *v0, const double *f, const double *fscale,
                                      double dt, int nsteps, double *history, double *critical) {
  if (!M) return TFM_E_INVALID;
  tfm_ctx *ctx = M->ctx;
  int rc = check_ops(ctx, M, C, K, n);
  if (rc) return rc;
  if (!u0 || !v0 || !f || !history || nsteps < 0) return fail(ctx, TFM_E_INVALID, "dynamic: null vector");
  if ((rc = set_device(ctx))) return rc;
  hipStream_t st = ctx->stream;
  const int nn = (int)n;
  const unsigned g = (unsigned)((nn + 255) / 256);
  double *md = nullptr, *rho = nullptr, *F = nullptr, *buf = nullptr, *hist = nullptr;
  auto cleanup = [&]() { dfree(md); dfree(rho); dfree(F); dfree(buf); dfree(hist); };
  // buf: u0, v0, u, vh, u', vh'
  if ((rc = dalloc(ctx, &md, n)) || (rc = dalloc(ctx, &rho, n)) || (rc = dalloc(ctx, &F, n)) ||
      (rc = dalloc(ctx, &buf, 6 * n)) || (rc = dalloc(ctx, &hist, (size_t)(nsteps + 1) * n))) {
    cleanup();
    return rc;
  }
  double *du0 = buf, *dv0 = buf + n, *u = buf + 2 * n, *vh = buf + 3 * n, *u1 = buf + 4 * n, *vh1 = buf + 5 * n;
  TFM_CHECK_HIP(ctx, hipMemcpy(du0, u0, sizeof(double) * n, hipMemcpyHostToDevice));
  TFM_CHECK_HIP(ctx, hipMemcpy(dv0, v0, sizeof(double) * n, hipMemcpyHostToDevice));
  TFM_CHECK_HIP(ctx, hipMemcpy(F, f, sizeof(double) * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(kd_lump, dim3(g), dim3(256), 0, st, nn, dcsr(M), dcsr(K), md, rho);
  std::vector<double> hr(n);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hr.data(), rho, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  double rmax = 0.0;
  bool has_mass = false;
  for (int64_t i = 0; i < n; i++)
    if (hr[i] >= 0.0 || std::isnan(hr[i])) {  // rows with positive mass
      rmax = std::fmax(rmax, hr[i]);
      has_mass = true;
    }
  const double safe = has_mass && rmax > 0.0 ? 2.0 / std::sqrt(rmax) : INFINITY;
  if (critical) *critical = safe;
  if (dt > safe) {
    cleanup();
    char msg[160];
    snprintf(msg, sizeof msg, "central-difference timestep %e exceeds the critical (CFL) limit %e", dt, safe);
    return fail(ctx, TFM_E_CFL, msg);
  }
  auto sc = [&](int k) { return fscale ? fscale[k] : 1.0; };
  hipLaunchKernelGGL(kd_cd_start, dim3(g), dim3(256), 0, st, nn, dcsr(C), dcsr(K), (const double *)F, sc(0), dt,
                     (const double *)md, (const double *)du0, (const double *)dv0, u, vh, hist);
  for (int k = 1; k <= nsteps; k++) {
    hipLaunchKernelGGL(kd_cd_step, dim3(g), dim3(256), 0, st, nn, dcsr(C), dcsr(K), (const double *)F, sc(k), dt,
                       (const double *)md, (const double *)u, (const double *)vh, u1, vh1, hist + (size_t)k * n);
    std::swap(u, u1);
    std::swap(vh, vh1);
  }
  hipError_t e = hipMemcpy(history, hist, sizeof(double) * (size_t)(nsteps + 1) * n, hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("central difference: ") + hipGetErrorString(e));
  return TFM_OK;
}

extern "C" int tfm_newmark(const tfm_matrix *M, const tfm_matrix *C, const tfm_matrix *K, int64_t n,
                           const double *u0, const double *v0, const double *f, const double *fscale, double dt,
                           double beta, double gamma, int nsteps, const tfm_solve_opts *opts, double *history,
                           tfm_solve_stats *stats) {
  if (!M) return TFM_E_INVALID;
  tfm_ctx *ctx = M->ctx;
  int rc = check_ops(ctx, M, C, K, n);
  if (rc) return rc;
  if (!u0 || !v0 || !f || !history || nsteps < 0) return fail(ctx, TFM_E_INVALID, "dynamic: null vector");
  if (M->nrows != n || M->ncols != n) return fail(ctx, TFM_E_NUMERIC, "sparse numeric factorization/solve failed: mass matrix must be n x n");
  if ((rc = set_device(ctx))) return rc;
  hipStream_t st = ctx->stream;
  const int nn = (int)n;
  const unsigned g = (unsigned)((nn + 255) / 256);
  tfm_solve_opts o{TFM_SOLVE_PCG64, 1e-12, 100000, 0.3};
  if (opts) o = *opts;
  o.mode = TFM_SOLVE_PCG64;
  double *F = nullptr, *buf = nullptr, *hist = nullptr;
  int *cnt = nullptr;
  tfm_matrix *Kh = nullptr;
  tfm_system *sm = nullptr, *sk = nullptr;
  BatchSolver *bs = nullptr;
  auto cleanup = [&]() {
    bsolver_destroy(bs);
    if (sk) tfm_system_destroy(sk);
    if (sm) tfm_system_destroy(sm);
    if (Kh) tfm_matrix_destroy(Kh);
    dfree(F); dfree(buf); dfree(hist); dfree(cnt);
  };
  // buf: u0, v0, u, v, a, pm, pc, rhs
  if ((rc = dalloc(ctx, &F, n)) || (rc = dalloc(ctx, &buf, 8 * n)) ||
      (rc = dalloc(ctx, &hist, (size_t)(nsteps + 1) * n)) || (rc = dalloc(ctx, &cnt, n))) {
    cleanup();
    return rc;
  }
  double *du0 = buf, *dv0 = buf + n, *u = buf + 2 * n, *v = buf + 3 * n, *a = buf + 4 * n, *pm = buf + 5 * n,
         *pc = buf + 6 * n, *rhs = buf + 7 * n;
  TFM_CHECK_HIP(ctx, hipMemcpy(du0, u0, sizeof(double) * n, hipMemcpyHostToDevice));
  TFM_CHECK_HIP(ctx, hipMemcpy(dv0, v0, sizeof(double) * n, hipMemcpyHostToDevice));
  TFM_CHECK_HIP(ctx, hipMemcpy(F, f, sizeof(double) * n, hipMemcpyHostToDevice));
  auto sc = [&](int k) { return fscale ? fscale[k] : 1.0; };
  // a0 = M^-1 (f0 - C v0 - K u0), by PCG on M
  hipLaunchKernelGGL(kd_sub3, dim3(g), dim3(256), 0, st, nn, (const double *)F, sc(0), dcsr(C), dcsr(K),
                     (const double *)du0, (const double *)dv0, rhs);
  std::vector<double> h0(n), ha(n);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(h0.data(), rhs, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  if ((rc = tfm_system_from_matrix(ctx, M, n, h0.data(), 0, nullptr, nullptr, &sm)) ||
      (rc = tfm_system_solve(sm, &o, ha.data(), nullptr))) {
    cleanup();
    return rc;
  }
  TFM_CHECK_HIP(ctx, hipMemcpy(a, ha.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  // K_hat = K + M / (b dt^2) + C g / (b dt)
  const double b = beta, gm = gamma;
  const double s1 = 1.0 / (b * dt * dt), s2 = gm / (b * dt);
  hipLaunchKernelGGL(kd_union_count, dim3(g), dim3(256), 0, st, nn, dcsr(K), dcsr(M), dcsr(C), cnt);
  Kh = new tfm_matrix();
  Kh->ctx = ctx;
  Kh->nrows = Kh->ncols = n;
  if ((rc = dalloc(ctx, &Kh->d_rp, n + 1)) || (rc = exclusive_scan_i32(ctx, cnt, Kh->d_rp, n, &Kh->nnz)) ||
      (rc = dalloc(ctx, &Kh->d_ci, Kh->nnz)) || (rc = dalloc(ctx, &Kh->d_v, Kh->nnz))) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(kd_union_fill, dim3(g), dim3(256), 0, st, nn, dcsr(K), dcsr(M), dcsr(C), s1, s2,
                     (const int *)Kh->d_rp, Kh->d_ci, Kh->d_v);
  std::vector<double> zero(n, 0.0);
  const int64_t off[2] = {0, n};
  if ((rc = tfm_system_from_matrix(ctx, Kh, n, zero.data(), 0, nullptr, nullptr, &sk)) ||
      (rc = bsolver_create(sk, 1, off, 1, TFM_SOLVE_PCG64, &bs)) || (rc = bsolver_refresh(bs))) {
    cleanup();
    return rc;
  }
  NmCoef c{};
  c.bdt2 = b * dt * dt;
  c.bdt = b * dt;
  c.h = 0.5 / b - 1.0;
  c.d_u = gm / (b * dt);
  c.d_v = gm / b - 1.0;
  c.d_a = dt * (gm / (2.0 * b) - 1.0);
  c.dt = dt;
  c.g = gm;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(u, du0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(v, dv0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hist, du0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(bsolver_x_mut(bs), du0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  tfm_solve_stats tot{0, 0, 0.0, 0.0}, s{};
  for (int k = 1; k <= nsteps && !rc; k++) {
    hipLaunchKernelGGL(kd_nm_p, dim3(g), dim3(256), 0, st, nn, c, (const double *)u, (const double *)v,
                       (const double *)a, pm, pc);
    hipLaunchKernelGGL(kd_nm_rhs, dim3(g), dim3(256), 0, st, nn, dcsr(M), dcsr(C), (const double *)F, sc(k),
                       (const double *)pm, (const double *)pc, rhs);
    if ((rc = bsolver_set_rhs_dev(bs, rhs)) || (rc = bsolver_run(bs, o, true, &s, nullptr))) break;
    tot.iterations += s.iterations;
    tot.outer += 1;
    tot.rel_residual = std::max(tot.rel_residual, s.rel_residual);
    tot.seconds += s.seconds;
    hipLaunchKernelGGL(kd_nm_update, dim3(g), dim3(256), 0, st, nn, c, bsolver_x(bs), u, v, a,
                       hist + (size_t)k * n);
  }
  if (!rc) {
    hipError_t e = hipMemcpy(history, hist, sizeof(double) * (size_t)(nsteps + 1) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(ctx, TFM_E_HIP, std::string("newmark: ") + hipGetErrorString(e));
  }
  if (stats) *stats = tot;
  cleanup();
  return rc;
}
