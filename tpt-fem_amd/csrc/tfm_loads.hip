// tfm_loads.hip -- right-hand-side loads on the GPU.
//
// Each kernel is a per-node gather over the node's element list (element-id
// order), which is the order in which the reference's drivers scatter into
// rhs, so FP64 sums match bit for bit:
//   * source vector  -- poisson_source_vector + solve_poisson's scatter
//     (crates/tpt-fem-thermal/src/lib.rs:163-196, 223-230);
//   * body vector    -- elasticity_body_vector + solve_elasticity's scatter
//     (crates/tpt-fem-elasticity/src/lib.rs:312-348, 370-379);
//   * Neumann flux   -- boundary_faces + apply_neumann_order
//     (crates/tpt-fem-assembly/src/lib.rs:563-667): faces seen by exactly one
//     element, visited in (element, face, quadrature point) order.
#include <cmath>

#include "tfm_elem.h"
#include "tfm_internal.h"

#pragma clang fp contract(off)

namespace tfm {

int build_rule(int cell, int order, std::vector<double> &rule);

__global__ void k_load_source(const double *__restrict__ coords, int pdim, const int *__restrict__ conn,
                              int nn, int cell, const int *__restrict__ n2e_ptr, const int *__restrict__ n2e,
                              int64_t n_nodes, const double *__restrict__ rule, int nq, double s_const,
                              const double *__restrict__ bvec, int dim, double *__restrict__ rhs) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  const bool body = bvec != nullptr;
  const int dpn = body ? dim : 1;
  double acc[3];
  for (int a = 0; a < dpn; a++) acc[a] = rhs[n * dpn + a];
  for (int p = n2e_ptr[n]; p < n2e_ptr[n + 1]; p++) {
    int v = n2e[p], e = v >> 3, ln = v & 7;
    double X[8][3];
    for (int a = 0; a < nn; a++)
      for (int d = 0; d < 3; d++) X[a][d] = d < dim ? coords[(int64_t)conn[(int64_t)e * nn + a] * pdim + d] : 0.0;
    double f[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < nq; q++) {
      const double *R = rule + q * kRuleStride;
      double w = R[0];
      const double *ns = R + 4;
      const double *lg = R + 12;
      double g[8][3];
      double det;
      switch (cell) {
      case TFM_LINE: { double Xl[2][3] = {{X[0][0], X[0][1], X[0][2]}, {X[1][0], X[1][1], X[1][2]}};
        double gl[2][3]; det = map_point<2, 1>(Xl, lg, gl); break; }
      case TFM_TRI: { double Xt[3][3]; for (int a = 0; a < 3; a++) for (int d = 0; d < 3; d++) Xt[a][d] = X[a][d];
        double gt[3][3]; det = map_point<3, 2>(Xt, lg, gt); break; }
      case TFM_QUAD: { double Xq[4][3]; for (int a = 0; a < 4; a++) for (int d = 0; d < 3; d++) Xq[a][d] = X[a][d];
        double gq[4][3]; det = map_point<4, 2>(Xq, lg, gq); break; }
      case TFM_TET: { double Xt[4][3]; for (int a = 0; a < 4; a++) for (int d = 0; d < 3; d++) Xt[a][d] = X[a][d];
        double gt[4][3]; det = map_point<4, 3>(Xt, lg, gt); break; }
      default: { det = map_point<8, 3>(X, lg, g); break; }
      }
      det = fabs(det);
      if (body) {
        for (int a = 0; a < dpn; a++) f[a] += w * ns[ln] * bvec[a] * det;
      } else {
        f[0] += w * s_const * ns[ln] * det;
      }
    }
    for (int a = 0; a < dpn; a++) acc[a] += f[a];
  }
  for (int a = 0; a < dpn; a++) rhs[n * dpn + a] = acc[a];
}

// Boundary-face tables (crates/tpt-fem-assembly/src/lib.rs:73-117, P1 cells).
struct FaceTab {
  int nf, fn_;          // faces per cell, nodes per face
  int fcell;            // reference cell of the face
  int idx[6][4];
};
__constant__ FaceTab c_faces[5] = {
    {2, 1, TFM_LINE, {{0}, {1}}},
    {3, 2, TFM_LINE, {{1, 2}, {2, 0}, {0, 1}}},
    {4, 2, TFM_LINE, {{0, 1}, {1, 2}, {2, 3}, {3, 0}}},
    {4, 3, TFM_TRI, {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}}},
    {6, 4, TFM_QUAD, {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 2, 6, 7}, {0, 4, 7, 3}, {1, 5, 6, 2}}},
};

// A face of element e is on the boundary iff no other element of node n's
// adjacency (every element holding the face holds n) has all its nodes
// (boundary_faces' count-of-one, crates/tpt-fem-assembly/src/lib.rs:563-583).
__device__ __forceinline__ bool face_on_boundary(const FaceTab &FT, int fi, const int *en, int e, int nn,
                                                 const int *__restrict__ conn, const int *__restrict__ n2e,
                                                 int pb, int pe) {
  for (int p2 = pb; p2 < pe; p2++) {
    const int e2 = n2e[p2] >> 3;
    if (e2 == e) continue;
    bool all = true;
    for (int k = 0; k < FT.fn_ && all; k++) {
      const int node = en[FT.idx[fi][k]];
      bool found = false;
      for (int a = 0; a < nn; a++) found |= conn[(int64_t)e2 * nn + a] == node;
      all = found;
    }
    if (all) return false;
  }
  return true;
}

// Face quadrature point q (apply_neumann_order / apply_robin_order,
// assembly/src/lib.rs:620-650, 700-730): weight, shape values, physical
// point and surface measure |t| (Line faces) or |t0 x t1| (Tri/Quad faces).
__device__ __forceinline__ void face_point(const FaceTab &FT, const double (&P)[4][3], int pdim,
                                           const double *__restrict__ R, double &w, const double *&ns,
                                           double (&xp)[3], double &measure) {
  w = R[0];
  ns = R + 4;
  const double *ng = R + 12;
  const int fdim = FT.fcell == TFM_LINE ? 1 : 2;
  xp[0] = xp[1] = xp[2] = 0.0;
  for (int k = 0; k < FT.fn_; k++)
    for (int d = 0; d < pdim; d++) xp[d] += ns[k] * P[k][d];
  double t[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int m = 0; m < fdim; m++)
    for (int k = 0; k < FT.fn_; k++)
      for (int d = 0; d < pdim; d++) t[m][d] += ng[k * fdim + m] * P[k][d];
  if (fdim == 1) {
    double s = t[0][0] * t[0][0];
    for (int d = 1; d < pdim; d++) s = s + t[0][d] * t[0][d];
    measure = sqrt(s);
  } else {
    const double c0 = t[0][1] * t[1][2] - t[0][2] * t[1][1];
    const double c1 = t[0][2] * t[1][0] - t[0][0] * t[1][2];
    const double c2 = t[0][0] * t[1][1] - t[0][1] * t[1][0];
    measure = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  }
}

// One scalar Neumann pass: for every boundary face containing node n, in
// (element, face) order, rhs[n*dpn + c] += w*measure*g*N_k at each face
// quadrature point, with g = gconst, or -- when axis >= 0 -- gconst only at
// points with |x[axis] - coord| <= tol and 0 elsewhere.  comp < 0 tiles the
// load over all dpn components (apply_neumann's behaviour).
//
// skip_far (axis >= 0 and rhs holding +0.0 at every node, as in the traction
// composition): a face whose nodes all lie outside the band |x[axis] -
// coord| <= tol (with a rounding margin) cannot have a quadrature point in
// it -- the points are convex combinations of the nodes -- so it would only
// add +0.0 to sums that start at +0.0; such faces are skipped before the
// boundary test, which is the expensive part.
__global__ void k_load_neumann(const double *__restrict__ coords, int pdim, const int *__restrict__ conn, int nn,
                               int cell, const int *__restrict__ n2e_ptr, const int *__restrict__ n2e,
                               int64_t n_nodes, const double *__restrict__ frule, int fnq, int dpn, int comp,
                               double gconst, int axis, double coord, double tol, int skip_far,
                               double *__restrict__ rhs) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  const FaceTab &FT = c_faces[cell];
  const int pb = n2e_ptr[n], pe = n2e_ptr[n + 1];
  double acc[3];
  for (int c = 0; c < dpn; c++) acc[c] = rhs[n * dpn + c];
  bool touched = false;
  for (int p = pb; p < pe; p++) {
    int v = n2e[p], e = v >> 3, ln = v & 7;
    int en[8];
    for (int a = 0; a < nn; a++) en[a] = conn[(int64_t)e * nn + a];
    for (int fi = 0; fi < FT.nf; fi++) {
      int kself = -1;
      for (int k = 0; k < FT.fn_; k++)
        if (FT.idx[fi][k] == ln) kself = k;
      if (kself < 0) continue;
      if (skip_far) {
        double lo = coords[(int64_t)en[FT.idx[fi][0]] * pdim + axis], hi = lo;
        bool finite = lo == lo;
        for (int k = 1; k < FT.fn_; k++) {
          const double x = coords[(int64_t)en[FT.idx[fi][k]] * pdim + axis];
          finite &= x == x;
          lo = x < lo ? x : lo;
          hi = x > hi ? x : hi;
        }
        const double margin = 8.0 * 2.220446049250313e-16 * (fabs(lo) + fabs(hi)) + 1e-300;
        if (finite && (hi + margin < coord - tol || lo - margin > coord + tol)) continue;
      }
      if (!face_on_boundary(FT, fi, en, e, nn, conn, n2e, pb, pe)) continue;
      if (FT.fn_ == 1) continue;  // point faces of Line2 cells: no surface measure
      double P[4][3];
      for (int k = 0; k < FT.fn_; k++)
        for (int d = 0; d < 3; d++) P[k][d] = d < pdim ? coords[(int64_t)en[FT.idx[fi][k]] * pdim + d] : 0.0;
      for (int q = 0; q < fnq; q++) {
        double w, measure, xp[3];
        const double *ns;
        face_point(FT, P, pdim, frule + q * kRuleStride, w, ns, xp, measure);
        double g = gconst;
        if (axis >= 0 && !(fabs(xp[axis] - coord) <= tol)) g = 0.0;
        double val = w * measure * g * ns[kself];
        touched = true;
        if (comp < 0) {
          for (int c = 0; c < dpn; c++) acc[c] += val;
        } else {
          acc[comp] += val;
        }
      }
    }
  }
  if (touched)
    for (int c = 0; c < dpn; c++) rhs[n * dpn + c] = acc[c];
}

static int upload_rule(tfm_ctx *ctx, int cell, int order, double **d_rule, int *nq) {
  std::vector<double> rule;
  int q = build_rule(cell, order, rule);
  if (q <= 0) return fail(ctx, TFM_E_INVALID, "gauss_legendre: order must be in 1..=5");
  int rc = dalloc(ctx, d_rule, rule.size());
  if (rc) return rc;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(*d_rule, rule.data(), sizeof(double) * rule.size(), hipMemcpyHostToDevice,
                                    ctx->stream));
  *nq = q;
  return TFM_OK;
}

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int load_source_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, double f, double *d_rhs) {
  int rc;
  if ((rc = ensure_adjacency(mesh))) return rc;
  double *d_rule = nullptr;
  int nq = 0;
  if ((rc = upload_rule(ctx, mesh->cell, order, &d_rule, &nq))) return rc;
  if (mesh->n_nodes > 0)
    hipLaunchKernelGGL(k_load_source, dim3(grid1(mesh->n_nodes, 128)), dim3(128), 0, ctx->stream, mesh->d_coords,
                       mesh->pdim, mesh->d_conn, mesh->nn, mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, mesh->n_nodes,
                       d_rule, nq, f, (const double *)nullptr, mesh->dim, d_rhs);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  dfree(d_rule);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, "source-vector kernel failed");
  return TFM_OK;
}

int load_body_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, const double *b, double *d_rhs) {
  int rc;
  if ((rc = ensure_adjacency(mesh))) return rc;
  double *d_rule = nullptr, *d_b = nullptr;
  int nq = 0;
  if ((rc = upload_rule(ctx, mesh->cell, order, &d_rule, &nq))) return rc;
  if ((rc = dalloc(ctx, &d_b, 3))) return rc;
  double hb[3] = {0, 0, 0};
  for (int a = 0; a < mesh->dim; a++) hb[a] = b[a];
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(d_b, hb, sizeof hb, hipMemcpyHostToDevice, ctx->stream));
  if (mesh->n_nodes > 0)
    hipLaunchKernelGGL(k_load_source, dim3(grid1(mesh->n_nodes, 128)), dim3(128), 0, ctx->stream, mesh->d_coords,
                       mesh->pdim, mesh->d_conn, mesh->nn, mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, mesh->n_nodes,
                       d_rule, nq, 0.0, d_b, mesh->dim, d_rhs);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  dfree(d_rule);
  dfree(d_b);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, "body-vector kernel failed");
  return TFM_OK;
}

int load_neumann_dev(tfm_ctx *ctx, tfm_mesh *mesh, int dpn, int comp, double g, int axis, double coord,
                     double tol, double *d_rhs, bool zeroed) {
  int rc;
  if ((rc = ensure_adjacency(mesh))) return rc;
  // face_quad (assembly/src/lib.rs:336-358) at FACE_QUAD_ORDER = 2: Line
  // faces Gauss-Legendre 2, Tri faces the Degree2 rule, Quad faces 2x2.
  int fcell = mesh->cell == TFM_TET ? TFM_TRI : (mesh->cell == TFM_HEX ? TFM_QUAD : TFM_LINE);
  double *d_rule = nullptr;
  int nq = 0;
  if ((rc = upload_rule(ctx, fcell, 2, &d_rule, &nq))) return rc;
  if (mesh->n_nodes > 0)
    hipLaunchKernelGGL(k_load_neumann, dim3(grid1(mesh->n_nodes, 128)), dim3(128), 0, ctx->stream, mesh->d_coords,
                       mesh->pdim, mesh->d_conn, mesh->nn, mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, mesh->n_nodes,
                       d_rule, nq, dpn, comp, g, axis, coord, tol, (int)(zeroed && axis >= 0 && axis < mesh->pdim),
                       d_rhs);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  dfree(d_rule);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, "Neumann kernel failed");
  return TFM_OK;
}

// Robin rows (apply_robin_order, crates/tpt-fem-assembly/src/lib.rs:684-753,
// appended to assemble()'s triplets before to_csr): one thread per matrix row
// r = n*dpn + c.  The row starts as the element CSR row; then, for every
// boundary face containing node n in (element, face) order and every face
// quadrature point, each face node l adds w*measure*h*N_k*N_l (k = n's place
// in the face) to column l*dpn + c -- onto the element sum, or starting a new
// entry -- which is to_csr's insertion-order collapse.  h applies on the band
// |x[axis] - coord| <= tol when axis >= 0 and is 0 elsewhere (zero entries
// are still pushed: the reference does not skip them).  FILL = 0 counts the
// merged row (and its largest column), FILL = 1 writes it.
constexpr int kRobinCap = 160;  // merged row length bound (Hex8, 3 DOFs: 81 + boundary extras)

template <int FILL>
__global__ void k_robin(const double *__restrict__ coords, int pdim, const int *__restrict__ conn, int nn,
                        int cell, const int *__restrict__ n2e_ptr, const int *__restrict__ n2e, int64_t nrows,
                        int dpn, const double *__restrict__ frule, int fnq, double h, int axis, double coord,
                        double tol, const int *__restrict__ krp, const int *__restrict__ kci,
                        const double *__restrict__ kv, int64_t knrows, int *__restrict__ cnt,
                        const int *__restrict__ orp, int *__restrict__ oci, double *__restrict__ ov,
                        int *__restrict__ maxcol, int *__restrict__ err) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int64_t n = r / dpn;
  const int c = (int)(r % dpn);
  const FaceTab &FT = c_faces[cell];
  int cols[kRobinCap];
  double vals[kRobinCap];
  int m = 0;
  if (r < knrows)
    for (int j = krp[r]; j < krp[r + 1]; j++) {
      if (m == kRobinCap) { atomicExch(err, 1); return; }
      cols[m] = kci[j];
      vals[m] = kv[j];
      m++;
    }
  const int pb = n2e_ptr[n], pe = n2e_ptr[n + 1];
  for (int p = pb; p < pe; p++) {
    const int v = n2e[p], e = v >> 3, ln = v & 7;
    int en[8];
    for (int a = 0; a < nn; a++) en[a] = conn[(int64_t)e * nn + a];
    for (int fi = 0; fi < FT.nf; fi++) {
      int kself = -1;
      for (int k = 0; k < FT.fn_; k++)
        if (FT.idx[fi][k] == ln) kself = k;
      if (kself < 0 || FT.fn_ == 1) continue;
      if (!face_on_boundary(FT, fi, en, e, nn, conn, n2e, pb, pe)) continue;
      double P[4][3];
      for (int k = 0; k < FT.fn_; k++)
        for (int d = 0; d < 3; d++) P[k][d] = d < pdim ? coords[(int64_t)en[FT.idx[fi][k]] * pdim + d] : 0.0;
      for (int q = 0; q < fnq; q++) {
        double w, measure, xp[3];
        const double *ns;
        face_point(FT, P, pdim, frule + q * kRuleStride, w, ns, xp, measure);
        const double hq = (axis < 0 || fabs(xp[axis] - coord) <= tol) ? h : 0.0;
        const double base = w * measure * hq * ns[kself];
        for (int l = 0; l < FT.fn_; l++) {
          const int col = en[FT.idx[fi][l]] * dpn + c;
          const double val = base * ns[l];
          int i = 0;
          while (i < m && cols[i] < col) i++;
          if (i < m && cols[i] == col) {
            if (FILL) vals[i] += val;
          } else {
            if (m == kRobinCap) { atomicExch(err, 1); return; }
            for (int t = m; t > i; t--) {
              cols[t] = cols[t - 1];
              vals[t] = vals[t - 1];
            }
            cols[i] = col;
            vals[i] = val;
            m++;
          }
        }
      }
    }
  }
  if (FILL) {
    const int o = orp[r];
    for (int i = 0; i < m; i++) {
      oci[o + i] = cols[i];
      ov[o + i] = vals[i];
    }
  } else {
    cnt[r] = m;
    if (m > 0) {
      atomicMax(maxcol, cols[m - 1]);
      atomicMax(maxcol + 1, (int)r);
    }
  }
}

// K = to_csr(assemble(...) ++ apply_robin(...)) as a new matrix.
int assemble_robin_dev(tfm_ctx *ctx, const tfm_mesh *cmesh, const tfm_operator *op, int dpn, double h, int axis,
                       double coord, double tol, tfm_matrix **out) {
  tfm_mesh *mesh = const_cast<tfm_mesh *>(cmesh);
  int rc;
  tfm_matrix *K = nullptr;
  if ((rc = tfm_assemble(ctx, mesh, op, dpn, &K))) return rc;
  if ((rc = ensure_adjacency(mesh))) {
    tfm_matrix_destroy(K);
    return rc;
  }
  hipStream_t st = ctx->stream;
  const int fcell = mesh->cell == TFM_TET ? TFM_TRI : (mesh->cell == TFM_HEX ? TFM_QUAD : TFM_LINE);
  const int64_t nrows = mesh->n_nodes * dpn;
  double *d_rule = nullptr;
  int nq = 0, *cnt = nullptr, *aux = nullptr;
  tfm_matrix *R = new tfm_matrix();
  R->ctx = ctx;
  auto cleanup = [&]() {
    dfree(d_rule);
    dfree(cnt);
    dfree(aux);
    tfm_matrix_destroy(K);
  };
  if ((rc = upload_rule(ctx, fcell, 2, &d_rule, &nq)) || (rc = dalloc(ctx, &cnt, nrows)) ||
      (rc = dalloc(ctx, &aux, 3)) || (rc = dalloc(ctx, &R->d_rp, nrows + 1))) {
    cleanup();
    tfm_matrix_destroy(R);
    return rc;
  }
  const int init[3] = {-1, -1, 0};  // largest column, largest non-empty row, overflow flag
  TFM_CHECK_HIP(ctx, hipMemcpy(aux, init, sizeof init, hipMemcpyHostToDevice));
  const unsigned g = (unsigned)((nrows + 127) / 128);
  if (nrows > 0)
    hipLaunchKernelGGL(k_robin<0>, dim3(g), dim3(128), 0, st, mesh->d_coords, mesh->pdim, mesh->d_conn, mesh->nn,
                       mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, nrows, dpn, (const double *)d_rule, nq, h, axis,
                       coord, tol, (const int *)K->d_rp, (const int *)K->d_ci, (const double *)K->d_v, K->nnz ? K->nrows : 0,
                       cnt, (const int *)nullptr, (int *)nullptr, (double *)nullptr, aux, aux + 2);
  int64_t nnz = 0;
  if ((rc = exclusive_scan_i32(ctx, cnt, R->d_rp, nrows, &nnz))) {
    cleanup();
    tfm_matrix_destroy(R);
    return rc;
  }
  int haux[3];
  TFM_CHECK_HIP(ctx, hipMemcpy(haux, aux, sizeof haux, hipMemcpyDeviceToHost));
  if (haux[2]) {
    cleanup();
    tfm_matrix_destroy(R);
    return fail(ctx, TFM_E_UNSUPPORTED, "Robin rows longer than the kernel's row buffer");
  }
  R->nnz = nnz;
  R->nrows = haux[1] >= 0 ? (int64_t)haux[1] + 1 : 1;  // Coo::to_csr: max row + 1, empty -> 1 x 1
  R->ncols = haux[0] >= 0 ? (int64_t)haux[0] + 1 : 1;
  if ((rc = dalloc(ctx, &R->d_ci, nnz)) || (rc = dalloc(ctx, &R->d_v, nnz))) {
    cleanup();
    tfm_matrix_destroy(R);
    return rc;
  }
  if (nrows > 0 && nnz > 0)
    hipLaunchKernelGGL(k_robin<1>, dim3(g), dim3(128), 0, st, mesh->d_coords, mesh->pdim, mesh->d_conn, mesh->nn,
                       mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, nrows, dpn, (const double *)d_rule, nq, h, axis,
                       coord, tol, (const int *)K->d_rp, (const int *)K->d_ci, (const double *)K->d_v, K->nnz ? K->nrows : 0,
                       (int *)nullptr, (const int *)R->d_rp, R->d_ci, R->d_v, aux, aux + 2);
  hipError_t e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) {
    tfm_matrix_destroy(R);
    return fail(ctx, TFM_E_HIP, "Robin kernel failed");
  }
  *out = R;
  return TFM_OK;
}

}  // namespace tfm

using namespace tfm;

extern "C" int tfm_assemble_robin(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn, double h,
                                  int axis, double coord, double tol, tfm_matrix **out) {
  if (!ctx || !mesh || !op || !out || dpn < 1 || dpn > 3) return TFM_E_INVALID;
  if (axis >= mesh->pdim) return fail(ctx, TFM_E_INVALID, "Robin band axis beyond the coordinates");
  int rc = set_device(ctx);
  if (rc) return rc;
  return assemble_robin_dev(ctx, mesh, op, dpn, h, axis, coord, tol, out);
}

static int host_rhs_roundtrip(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, double *rhs,
                              int (*fn)(tfm_ctx *, tfm_mesh *, double *, const void *), const void *arg) {
  int rc = set_device(ctx);
  if (rc) return rc;
  int64_t n = mesh->n_nodes * dpn;
  double *d = nullptr;
  if ((rc = dalloc(ctx, &d, n))) return rc;
  hipError_t e = hipMemcpy(d, rhs, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) { dfree(d); return fail(ctx, TFM_E_HIP, "rhs upload failed"); }
  rc = fn(ctx, const_cast<tfm_mesh *>(mesh), d, arg);
  if (!rc && hipMemcpy(rhs, d, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(ctx, TFM_E_HIP, "rhs download failed");
  dfree(d);
  return rc;
}

struct SrcArg { int order; double f; };
struct BodyArg { int order; const double *b; };
struct NeuArg { int dpn, comp; double g; int axis; double coord, tol; };

extern "C" int tfm_load_source(tfm_ctx *ctx, const tfm_mesh *mesh, int quad_order, double f, double *rhs) {
  if (!ctx || !mesh || !rhs) return TFM_E_INVALID;
  SrcArg a{quad_order, f};
  return host_rhs_roundtrip(ctx, mesh, 1, rhs,
                            [](tfm_ctx *c, tfm_mesh *m, double *d, const void *p) {
                              const SrcArg *s = (const SrcArg *)p;
                              return load_source_dev(c, m, s->order, s->f, d);
                            },
                            &a);
}

extern "C" int tfm_load_body(tfm_ctx *ctx, const tfm_mesh *mesh, int quad_order, const double *b, double *rhs) {
  if (!ctx || !mesh || !rhs || !b) return TFM_E_INVALID;
  BodyArg a{quad_order, b};
  return host_rhs_roundtrip(ctx, mesh, mesh->dim, rhs,
                            [](tfm_ctx *c, tfm_mesh *m, double *d, const void *p) {
                              const BodyArg *s = (const BodyArg *)p;
                              return load_body_dev(c, m, s->order, s->b, d);
                            },
                            &a);
}

extern "C" int tfm_load_neumann(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, double g, double *rhs) {
  if (!ctx || !mesh || !rhs || dpn < 1 || dpn > 3) return TFM_E_INVALID;
  NeuArg a{dpn, -1, g, -1, 0.0, 0.0};
  return host_rhs_roundtrip(ctx, mesh, dpn, rhs,
                            [](tfm_ctx *c, tfm_mesh *m, double *d, const void *p) {
                              const NeuArg *s = (const NeuArg *)p;
                              return load_neumann_dev(c, m, s->dpn, s->comp, s->g, s->axis, s->coord, s->tol, d,
                                                      false);
                            },
                            &a);
}

// The reference's composition for a vector traction: one scalar pass per
// non-zero component into a zeroed buffer, then rhs[n*dpn + c] += buf[n].
__global__ void k_interleave_add(const double *buf, int64_t n, int dpn, int c, double *rhs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rhs[i * dpn + c] += buf[i];
}

extern "C" int tfm_load_plane_traction(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, int axis, double coord,
                                       double tol, const double *t, double *rhs) {
  if (!ctx || !mesh || !rhs || !t || dpn < 1 || dpn > 3 || axis < 0) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  int64_t nn = mesh->n_nodes;
  double *d_rhs = nullptr, *d_buf = nullptr;
  if ((rc = dalloc(ctx, &d_rhs, nn * dpn)) || (rc = dalloc(ctx, &d_buf, nn))) return rc;
  TFM_CHECK_HIP(ctx, hipMemcpy(d_rhs, rhs, sizeof(double) * nn * dpn, hipMemcpyHostToDevice));
  for (int c = 0; c < dpn && !rc; c++) {
    if (t[c] == 0.0) continue;
    if (hipMemsetAsync(d_buf, 0, sizeof(double) * nn, ctx->stream) != hipSuccess) {
      rc = fail(ctx, TFM_E_HIP, "memset failed");
      break;
    }
    rc = load_neumann_dev(ctx, const_cast<tfm_mesh *>(mesh), 1, 0, t[c], axis, coord, tol, d_buf, true);
    if (!rc && nn > 0)
      hipLaunchKernelGGL(k_interleave_add, dim3(grid1(nn, 256)), dim3(256), 0, ctx->stream, d_buf, nn, dpn, c, d_rhs);
  }
  if (!rc && hipMemcpy(rhs, d_rhs, sizeof(double) * nn * dpn, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(ctx, TFM_E_HIP, "rhs download failed");
  dfree(d_rhs);
  dfree(d_buf);
  return rc;
}
