// tfm_topopt.hip -- SIMP minimum-compliance loop on the GPU.
//
// topopt_simp (crates/tpt-fem-topopt/src/lib.rs:342-385) for P designs at
// once: the P grids are concatenated into one mesh, so every OC iteration is
// one SIMP assembly over all designs into one block-diagonal system, one
// batched PCG (warm-started from the previous iteration's displacements),
// then three element kernels -- raw sensitivity, sensitivity filter and the
// optimality-criteria update -- that never leave the device.  The densities
// live in the assembly plan's per-element array, so the update feeds the next
// assembly directly.  Only the per-design compliances come back each
// iteration.
//
// Element data layout (HBM): rho, raw, sens  double [P * ne], design-major,
// element e = ej * (nx - 1) + ei inside a design (Grid::new order,
// topopt/src/lib.rs:89-116); displacements from the batched solver,
// u[2 n + c] for node n of the concatenated mesh.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tfm_internal.h"

namespace tfm {

// crates/tpt-fem-topopt/src/lib.rs:53-61
constexpr double kEMinFactor = 1e-3;
constexpr double kRhoMin = 1e-3;
constexpr double kRhoMax = 1.0;
constexpr int kOcSteps = 60;  // oc_update bisection steps (:296)
constexpr int kOcThreads = 1024;

// element_quad_form (:241-254) of every element of every design and the raw
// sensitivity -p rho^(p-1) (E0 - Emin) u_e^T K0 u_e (:370-373), one thread per
// element, sums in the reference's i-then-j order.
__global__ __launch_bounds__(256) void k_topo_sens(int n, const int *__restrict__ conn, const double *__restrict__ u,
                                                   const double *__restrict__ ke0, const double *__restrict__ rho,
                                                   double penal, double de, double *__restrict__ raw) {
  __shared__ double k0[64];
  if (threadIdx.x < 64) k0[threadIdx.x] = ke0[threadIdx.x];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double ue[8];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int nd = conn[4 * e + a];
    ue[2 * a] = u[2 * nd];
    ue[2 * a + 1] = u[2 * nd + 1];
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) s += ue[i] * k0[i * 8 + j] * ue[j];
  raw[e] = -penal * pow(rho[e], penal - 1.0) * de * s;
}

// filter_sensitivity (:259-279): weights max(0, R - |c_e - c_f|) over the
// elements f with |c_e - c_f|^2 <= R^2.  The reference visits every element
// of the design; only those within W = floor(R / h) + 1 cells of e per axis can
// pass the test, so the window is visited instead, in the same ascending f
// order -- the skipped elements contribute nothing in the reference either.
__global__ __launch_bounds__(256) void k_topo_filter(int n, int ne, int nex, int ney, double h, double radius, int W,
                                                     const double *__restrict__ raw, double *__restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int p = idx / ne, e = idx - p * ne;
  const int ei = e % nex, ej = e / nex;
  const double *rw = raw + (size_t)p * ne;
  const double r2 = radius * radius;
  const double cex = ((double)ei + 0.5) * h, cey = ((double)ej + 0.5) * h;
  const int i0 = max(0, ei - W), i1 = min(nex - 1, ei + W);
  const int j0 = max(0, ej - W), j1 = min(ney - 1, ej + W);
  double num = 0.0, den = 0.0;
  for (int fj = j0; fj <= j1; fj++) {
    const double dy = cey - ((double)fj + 0.5) * h;
    for (int fi = i0; fi <= i1; fi++) {
      const double dx = cex - ((double)fi + 0.5) * h;
      const double d2 = dx * dx + dy * dy;
      if (d2 <= r2) {
        const double w = radius - sqrt(d2);
        num += w * rw[fj * nex + fi];
        den += w;
      }
    }
  }
  out[idx] = den > 0.0 ? num / den : rw[e];
}

// Candidate density of oc_update (:300-312): rho (B / lambda)^0.5 clamped to
// [max(rho - m, RHO_MIN), min(rho + m, RHO_MAX)].  The damping exponent is
// the constant 0.5, so the power is a square root (correctly rounded, as a
// correctly rounded pow(x, 0.5) is).
__device__ __forceinline__ double oc_move(double rho, double b, double lambda, double move) {
  const double ratio = b > 0.0 ? sqrt(b / lambda) : 0.0;
  double c = rho * ratio;
  const double lo = fmax(rho - move, kRhoMin), hi = fmin(rho + move, kRhoMax);
  if (c < lo) c = lo;  // f64::clamp
  if (c > hi) c = hi;
  return c;
}

// Sum over the block, the same value in every thread (fixed order).
__device__ __forceinline__ double oc_block_sum(double v) {
  __shared__ double sh[kOcThreads / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int q = 0; q < kOcThreads / 64; q++) t += sh[q];
  return t;
}

// oc_update (:284-331), one block per design: 60 bisection steps on lambda in
// [1e-9, 1e9] against the volume target vol_frac * ne, then the update at the
// last midpoint, in place.
__global__ __launch_bounds__(kOcThreads) void k_topo_oc(int ne, const double *__restrict__ vol_frac, double move,
                                                        double *__restrict__ rho, const double *__restrict__ sens) {
  const int p = blockIdx.x;
  double *r = rho + (size_t)p * ne;
  const double *s = sens + (size_t)p * ne;
  const double target = vol_frac[p] * (double)ne;
  double lo = 1e-9, hi = 1e9, lambda = 1.0;
  for (int it = 0; it < kOcSteps; it++) {
    lambda = 0.5 * (lo + hi);
    double sum = 0.0;
    for (int e = threadIdx.x; e < ne; e += kOcThreads) sum += oc_move(r[e], -s[e], lambda, move);
    sum = oc_block_sum(sum);
    if (sum > target) lo = lambda;
    else hi = lambda;
  }
  // every read of r above finished before the last block sum's barriers
  for (int e = threadIdx.x; e < ne; e += kOcThreads) r[e] = oc_move(r[e], -s[e], lambda, move);
}

}  // namespace tfm

using namespace tfm;

extern "C" int tfm_topopt_simp_batch(tfm_ctx *ctx, int64_t P, const tfm_topopt_params *prm, const double *f,
                                     const double *vol_frac, int64_t n_bc, const int64_t *bc_dof,
                                     const double *bc_val, const tfm_solve_opts *opts, double *densities,
                                     double *compliance, int *iterations, tfm_solve_stats *stats) {
  if (!ctx) return TFM_E_INVALID;
  if (!prm || !f || !densities || !compliance || P < 1 || (n_bc > 0 && (!bc_dof || !bc_val)) || n_bc < 0)
    return fail(ctx, TFM_E_INVALID, "topopt: null argument or empty batch");
  if (prm->nx < 2 || prm->ny < 2) return fail(ctx, TFM_E_INVALID, "topopt: the grid needs nx, ny >= 2");
  if (prm->max_iter < 0) return fail(ctx, TFM_E_INVALID, "topopt: max_iter must be >= 0");
  // f64::clamp panics unless min <= max, i.e. move_limit >= 0
  if (!(prm->move_limit >= 0.0)) return fail(ctx, TFM_E_INVALID, "topopt: move_limit must be >= 0");
  const int64_t nx = prm->nx, ny = prm->ny, nn = nx * ny, ndof = 2 * nn, nex = nx - 1, ney = ny - 1, ne = nex * ney;
  if (P * ne >= INT32_MAX / 8 || P * ndof >= INT32_MAX)
    return fail(ctx, TFM_E_UNSUPPORTED, "topopt: batch too large for int32 indexing");
  for (int64_t b = 0; b < n_bc; b++) {
    if (bc_dof[b] < 0 || bc_dof[b] >= ndof) return fail(ctx, TFM_E_INVALID, "topopt: Dirichlet DOF out of range");
    if (bc_val[b] != 0.0)
      return fail(ctx, TFM_E_UNSUPPORTED, "topopt: batched solve takes homogeneous Dirichlet values only");
  }
  tfm_solve_opts o{TFM_SOLVE_PCG64_MG, 1e-12, 100000, 0.3};
  if (opts) o = *opts;
  const bool mixed = o.mode == TFM_SOLVE_MIXED;
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const int64_t nc = (int64_t)prm->max_iter + 1;
  std::vector<double> vf(P), rho0((size_t)(P * ne));
  for (int64_t p = 0; p < P; p++) {
    vf[p] = vol_frac ? vol_frac[p] : prm->vol_frac;
    std::fill(rho0.begin() + p * ne, rho0.begin() + (p + 1) * ne, vf[p]);
  }
  // supports replicated per design; design p owns DOFs [p*ndof, (p+1)*ndof)
  std::vector<int64_t> dofs((size_t)(P * n_bc)), off(P + 1);
  std::vector<double> vals((size_t)(P * n_bc), 0.0);
  for (int64_t p = 0; p < P; p++)
    for (int64_t b = 0; b < n_bc; b++) dofs[p * n_bc + b] = p * ndof + bc_dof[b];
  for (int64_t p = 0; p <= P; p++) off[p] = p * ndof;
  tfm_operator op{};
  op.kind = TFM_OP_SIMP;
  op.model = TFM_PLANE_STRESS;
  op.quad_order = 2;
  op.nu = prm->nu;
  op.rho = rho0.data();
  op.penal = prm->penal;
  op.e0 = prm->e0;
  op.emin = kEMinFactor * prm->e0;  // :346
  op.h = prm->h;
  tfm_mesh *mesh = nullptr;
  tfm_system *sys = nullptr;
  BatchSolver *bs = nullptr;
  double *d_raw = nullptr, *d_sens = nullptr, *d_vf = nullptr, *d_ke0 = nullptr;
  auto cleanup = [&]() {
    bsolver_destroy(bs);
    if (sys) tfm_system_destroy(sys);
    if (mesh) tfm_mesh_destroy(mesh);
    dfree(d_raw);
    dfree(d_sens);
    dfree(d_vf);
    dfree(d_ke0);
  };
  if ((rc = tfm_mesh_grid_batch(ctx, P, nx, ny, prm->h, &mesh)) ||
      (rc = tfm_system_create(ctx, mesh, &op, 2, P * n_bc, dofs.data(), vals.data(), f,
                              TFM_SYS_FP64 | (mixed ? TFM_SYS_FP32 : 0), &sys)) ||
      (rc = bsolver_create(sys, (int)P, off.data(), 1, o.mode, &bs)) || (rc = bsolver_set_rhs(bs, nullptr)) ||
      (rc = dalloc(ctx, &d_raw, P * ne)) || (rc = dalloc(ctx, &d_sens, P * ne)) || (rc = dalloc(ctx, &d_vf, P)) ||
      (rc = dalloc(ctx, &d_ke0, 64))) {
    cleanup();
    return rc;
  }
  if (hipMemcpyAsync(d_vf, vf.data(), sizeof(double) * P, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(d_ke0, sys->plan.ke0, sizeof(double) * 64, hipMemcpyHostToDevice, st) != hipSuccess) {
    cleanup();
    return fail(ctx, TFM_E_HIP, "topopt: upload failed");
  }
  // filter window (cells per axis); R / h that is not finite or spans the
  // grid visits the whole design
  const int64_t span = std::max(nex, ney);
  const double q = std::fabs(prm->filter_radius) / std::fabs(prm->h);
  const int W = (int)(q < (double)span ? (int64_t)std::floor(q) + 1 : span);
  const int n = (int)(P * ne);
  const unsigned ge = (unsigned)((n + 255) / 256);
  std::vector<double> fu(P);
  std::vector<tfm_solve_stats> ss(P);
  if (stats)
    for (int64_t p = 0; p < P; p++) stats[p] = tfm_solve_stats{0, 0, 0.0, 0.0};
  // the OC step: sensitivities of the last solution, filter, density update
  auto oc_step = [&]() -> int {
    hipLaunchKernelGGL(k_topo_sens, dim3(ge), dim3(256), 0, st, n, (const int *)mesh->d_conn, bsolver_x(bs),
                       (const double *)d_ke0, (const double *)sys->plan.d_rho, prm->penal, prm->e0 - op.emin, d_raw);
    hipLaunchKernelGGL(k_topo_filter, dim3(ge), dim3(256), 0, st, n, (int)ne, (int)nex, (int)ney, prm->h,
                       prm->filter_radius, W, (const double *)d_raw, d_sens);
    hipLaunchKernelGGL(k_topo_oc, dim3((unsigned)P), dim3(kOcThreads), 0, st, (int)ne, (const double *)d_vf,
                       prm->move_limit, sys->plan.d_rho, (const double *)d_sens);
    if (hipGetLastError() != hipSuccess) return fail(ctx, TFM_E_HIP, "topopt: element kernel launch failed");
    return TFM_OK;
  };
  // Iterations k >= 2 run as one sequence of captured graphs: the numeric
  // assembly with the new densities and the solver's refresh, the warm
  // started solve, then the certificate and the OC step.
  SolveHooks loop;
  loop.pre = [&]() { return plan_numeric(ctx, &sys->plan, sys->d_v64, mixed ? sys->d_v32 : nullptr); };
  loop.post = oc_step;
  loop.tag = 0x70706f63ull;  // this loop's hooks
  for (int64_t k = 0; k < nc && !rc; k++) {
    if (k == 0) {
      rc = bsolver_solve(bs, o, false, true, nullptr, ss.data(), fu.data());
    } else if (k >= 2) {
      sys->note_values(true, true, mixed && sys->d_v32);
      rc = bsolver_solve(bs, o, true, true, &loop, ss.data(), fu.data());
    }
    if (rc) break;
    for (int64_t p = 0; p < P; p++) {
      compliance[p * nc + k] = k == 1 ? compliance[p * nc] : fu[p];
      if (stats && k != 1) {
        stats[p].iterations += ss[p].iterations;
        stats[p].outer += 1;
        stats[p].rel_residual = std::max(stats[p].rel_residual, ss[p].rel_residual);
        stats[p].seconds += ss[p].seconds;
      }
    }
    // iteration 1 re-solves the uniform start (:356-367): same system, same
    // u, then its OC step
    if (k == 1) rc = oc_step();
  }
  if (!rc) {
    if (hipMemcpyAsync(densities, sys->plan.d_rho, sizeof(double) * P * ne, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = fail(ctx, TFM_E_HIP, "topopt: download failed");
    if (iterations)
      for (int64_t p = 0; p < P; p++) iterations[p] = prm->max_iter;
  }
  cleanup();
  return rc;
}

extern "C" int tfm_topopt_simp(tfm_ctx *ctx, const tfm_topopt_params *prm, const double *f, int64_t n_bc,
                               const int64_t *bc_dof, const double *bc_val, const tfm_solve_opts *opts,
                               double *densities, double *compliance, int *iterations, tfm_solve_stats *stats) {
  return tfm_topopt_simp_batch(ctx, 1, prm, f, nullptr, n_bc, bc_dof, bc_val, opts, densities, compliance,
                               iterations, stats);
}
