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

// qv (optional): the closure's values at every element quadrature point,
// [n_elem][nq][dpn], in place of the constant s_const / bvec.
__global__ void k_load_source(const double *__restrict__ coords, int pdim, const int *__restrict__ conn,
                              int nn, int cell, const int *__restrict__ n2e_ptr, const int *__restrict__ n2e,
                              int64_t n_nodes, const double *__restrict__ rule, int nq, double s_const,
                              const double *__restrict__ bvec, int dim, const double *__restrict__ qv, int body,
                              double *__restrict__ rhs) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
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
      const double *qe = qv ? qv + ((int64_t)e * nq + q) * dpn : nullptr;
      if (body) {
        for (int a = 0; a < dpn; a++) f[a] += w * ns[ln] * (qe ? qe[a] : bvec[a]) * det;
      } else {
        f[0] += w * (qe ? qe[0] : s_const) * ns[ln] * det;
      }
    }
    for (int a = 0; a < dpn; a++) acc[a] += f[a];
  }
  for (int a = 0; a < dpn; a++) rhs[n * dpn + a] = acc[a];
}

// Physical positions of every element quadrature point, formed as the
// reference forms a closure's argument (crates/tpt-fem-thermal/src/lib.rs:180-185,
// crates/tpt-fem-elasticity/src/lib.rs:331-336): x[d] = sum_k ns[k] * X[k][d]
// from 0.0 in node order.  One thread per (element, point); out
// [n_elem][nq][pdim].
__global__ void k_qp_coords(const double *__restrict__ coords, int pdim, const int *__restrict__ conn, int nn,
                            int64_t n_elem, const double *__restrict__ rule, int nq, double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_elem * nq) return;
  const int64_t e = t / nq;
  const int q = (int)(t % nq);
  const double *ns = rule + q * kRuleStride + 4;
  double x[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nn; k++) {
    const int64_t nd = conn[e * nn + k];
    for (int d = 0; d < pdim; d++) x[d] += ns[k] * coords[nd * pdim + d];
  }
  for (int d = 0; d < pdim; d++) out[t * pdim + d] = x[d];
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

// Outward unit normal of a face quadrature point, as apply_neumann_order /
// apply_robin_order form the closure's second argument
// (crates/tpt-fem-assembly/src/lib.rs:394-420 surface, :634-651): the in-plane
// perpendicular (-t1, t0) of a Line face or the cross product of a surface
// face's tangents, divided by its norm, flipped when it points against face
// centroid minus element centroid (mean(): sum in node order, then divide).
__device__ __forceinline__ void face_normal(const FaceTab &FT, const double (&P)[4][3], const double *ecen,
                                            int pdim, const double *__restrict__ R, double (&nrm)[3]) {
  const double *ng = R + 12;
  const int fdim = FT.fcell == TFM_LINE ? 1 : 2;
  double t[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int m = 0; m < fdim; m++)
    for (int k = 0; k < FT.fn_; k++)
      for (int d = 0; d < pdim; d++) t[m][d] += ng[k * fdim + m] * P[k][d];
  nrm[0] = nrm[1] = nrm[2] = 0.0;
  if (fdim == 1) {
    nrm[0] = -t[0][1];
    nrm[1] = t[0][0];
    double s = nrm[0] * nrm[0];
    for (int d = 1; d < pdim; d++) s = s + nrm[d] * nrm[d];
    const double ln = sqrt(s);
    for (int d = 0; d < pdim; d++) nrm[d] = nrm[d] / ln;
  } else {
    const double c[3] = {t[0][1] * t[1][2] - t[0][2] * t[1][1], t[0][2] * t[1][0] - t[0][0] * t[1][2],
                         t[0][0] * t[1][1] - t[0][1] * t[1][0]};
    const double ln = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int d = 0; d < 3; d++) nrm[d] = c[d] / ln;
  }
  double fcen[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < FT.fn_; k++)
    for (int d = 0; d < pdim; d++) fcen[d] += P[k][d];
  for (int d = 0; d < pdim; d++) fcen[d] = fcen[d] / (double)FT.fn_;
  double dd = nrm[0] * (fcen[0] - ecen[0]);
  for (int d = 1; d < pdim; d++) dd = dd + nrm[d] * (fcen[d] - ecen[d]);
  if (dd < 0.0)
    for (int d = 0; d < pdim; d++) nrm[d] = -nrm[d];
}

// Which faces of each element lie on the boundary (bit fi of bmask[e]; point
// faces of Line2 cells excluded), and how many: boundary_faces' order is
// (element, face), so the index of face fi of element e among all boundary
// faces is boff[e] + popcount(bmask[e] & ((1 << fi) - 1)).
__global__ void k_bface_mask(const int *__restrict__ conn, int nn, int cell, const int *__restrict__ n2e_ptr,
                             const int *__restrict__ n2e, int64_t n_elem, unsigned char *__restrict__ bmask,
                             int *__restrict__ cnt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  const FaceTab &FT = c_faces[cell];
  int en[8];
  for (int a = 0; a < nn; a++) en[a] = conn[e * nn + a];
  unsigned m = 0;
  if (FT.fn_ > 1)
    for (int fi = 0; fi < FT.nf; fi++) {
      const int n0 = en[FT.idx[fi][0]];
      if (face_on_boundary(FT, fi, en, (int)e, nn, conn, n2e, n2e_ptr[n0], n2e_ptr[n0 + 1])) m |= 1u << fi;
    }
  bmask[e] = (unsigned char)m;
  cnt[e] = __popc(m);
}

__device__ __forceinline__ int64_t bface_point(const unsigned char *bmask, const int *boff, int64_t e, int fi,
                                               int fnq, int q) {
  return ((int64_t)boff[e] + __popc((unsigned)bmask[e] & ((1u << fi) - 1u))) * fnq + q;
}

// Position and outward normal of every boundary-face quadrature point, one
// thread per element.
__global__ void k_bface_points(const double *__restrict__ coords, int pdim, const int *__restrict__ conn, int nn,
                               int cell, int64_t n_elem, const unsigned char *__restrict__ bmask,
                               const int *__restrict__ boff, const double *__restrict__ frule, int fnq,
                               double *__restrict__ xout, double *__restrict__ nout) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem || bmask[e] == 0) return;
  const FaceTab &FT = c_faces[cell];
  int en[8];
  for (int a = 0; a < nn; a++) en[a] = conn[e * nn + a];
  double ecen[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < nn; a++)
    for (int d = 0; d < pdim; d++) ecen[d] += coords[(int64_t)en[a] * pdim + d];
  for (int d = 0; d < pdim; d++) ecen[d] = ecen[d] / (double)nn;
  for (int fi = 0; fi < FT.nf; fi++) {
    if (!((bmask[e] >> fi) & 1)) continue;
    double P[4][3];
    for (int k = 0; k < FT.fn_; k++)
      for (int d = 0; d < 3; d++) P[k][d] = d < pdim ? coords[(int64_t)en[FT.idx[fi][k]] * pdim + d] : 0.0;
    for (int q = 0; q < fnq; q++) {
      double w, measure, xp[3], nrm[3];
      const double *ns;
      face_point(FT, P, pdim, frule + q * kRuleStride, w, ns, xp, measure);
      face_normal(FT, P, ecen, pdim, frule + q * kRuleStride, nrm);
      const int64_t ip = bface_point(bmask, boff, e, fi, fnq, q);
      for (int d = 0; d < pdim; d++) {
        if (xout) xout[ip * pdim + d] = xp[d];
        if (nout) nout[ip * pdim + d] = nrm[d];
      }
    }
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
                               const double *__restrict__ gq, const unsigned char *__restrict__ bmask,
                               const int *__restrict__ boff, double *__restrict__ rhs) {
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
      if (FT.fn_ == 1) continue;  // point faces of Line2 cells: no surface measure
      if (bmask ? !((bmask[e] >> fi) & 1) : !face_on_boundary(FT, fi, en, e, nn, conn, n2e, pb, pe)) continue;
      double P[4][3];
      for (int k = 0; k < FT.fn_; k++)
        for (int d = 0; d < 3; d++) P[k][d] = d < pdim ? coords[(int64_t)en[FT.idx[fi][k]] * pdim + d] : 0.0;
      for (int q = 0; q < fnq; q++) {
        double w, measure, xp[3];
        const double *ns;
        face_point(FT, P, pdim, frule + q * kRuleStride, w, ns, xp, measure);
        double g = gq ? gq[bface_point(bmask, boff, e, fi, fnq, q)] : gconst;
        if (!gq && axis >= 0 && !(fabs(xp[axis] - coord) <= tol)) g = 0.0;
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

int element_rule_points(tfm_ctx *ctx, int cell, int order, int *nq) {
  std::vector<double> rule;
  int q = build_rule(cell, order, rule);
  if (q <= 0) return fail(ctx, TFM_E_INVALID, "gauss_legendre: order must be in 1..=5");
  *nq = q;
  return TFM_OK;
}

// Source (body = false) or body-force (body = true) load: constant values
// (f, or b[dim]) or, when h_qv is given, the closure's values at every
// element quadrature point (host, [n_elem][nq][1 or dim]).
static int load_elem_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, bool body, double f, const double *b,
                         const double *h_qv, double *d_rhs) {
  int rc;
  if ((rc = ensure_adjacency(mesh))) return rc;
  double *d_rule = nullptr, *d_b = nullptr, *d_qv = nullptr;
  int nq = 0;
  if ((rc = upload_rule(ctx, mesh->cell, order, &d_rule, &nq))) return rc;
  const int dpn = body ? mesh->dim : 1;
  auto cleanup = [&]() {
    dfree(d_rule);
    dfree(d_b);
    dfree(d_qv);
  };
  if ((rc = dalloc(ctx, &d_b, 3))) { cleanup(); return rc; }
  double hb[3] = {0, 0, 0};
  if (body && b)
    for (int a = 0; a < mesh->dim; a++) hb[a] = b[a];
  hipError_t e = hipMemcpyAsync(d_b, hb, sizeof hb, hipMemcpyHostToDevice, ctx->stream);
  const size_t nv = (size_t)mesh->n_elem * nq * dpn;
  if (e == hipSuccess && h_qv) {
    if ((rc = dalloc(ctx, &d_qv, nv))) { cleanup(); return rc; }
    e = hipMemcpyAsync(d_qv, h_qv, sizeof(double) * nv, hipMemcpyHostToDevice, ctx->stream);
  }
  if (e == hipSuccess && mesh->n_nodes > 0)
    hipLaunchKernelGGL(k_load_source, dim3(grid1(mesh->n_nodes, 128)), dim3(128), 0, ctx->stream, mesh->d_coords,
                       mesh->pdim, mesh->d_conn, mesh->nn, mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, mesh->n_nodes,
                       d_rule, nq, f, (const double *)d_b, mesh->dim, (const double *)d_qv, (int)body, d_rhs);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  cleanup();
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("element load kernel: ") + hipGetErrorString(e));
  return TFM_OK;
}

int load_source_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, double f, double *d_rhs) {
  return load_elem_dev(ctx, mesh, order, false, f, nullptr, nullptr, d_rhs);
}

int load_body_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, const double *b, double *d_rhs) {
  return load_elem_dev(ctx, mesh, order, true, 0.0, b, nullptr, d_rhs);
}

int load_source_qp_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, const double *h_sq, double *d_rhs) {
  return load_elem_dev(ctx, mesh, order, false, 0.0, nullptr, h_sq, d_rhs);
}

int load_body_qp_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, const double *h_bq, double *d_rhs) {
  return load_elem_dev(ctx, mesh, order, true, 0.0, nullptr, h_bq, d_rhs);
}

static int face_cell(int cell) { return cell == TFM_TET ? TFM_TRI : (cell == TFM_HEX ? TFM_QUAD : TFM_LINE); }

// The boundary-face index of a mesh (k_bface_mask + scan): per element the
// mask of its boundary faces and the index of its first one.
struct BFaceIndex {
  unsigned char *bmask = nullptr;
  int *boff = nullptr;
  int64_t nbf = 0;
  ~BFaceIndex() {
    dfree(bmask);
    dfree(boff);
  }
};

static int bface_index(tfm_ctx *ctx, tfm_mesh *mesh, BFaceIndex *bi) {
  int rc;
  if ((rc = ensure_adjacency(mesh))) return rc;
  int *cnt = nullptr;
  if ((rc = dalloc(ctx, &bi->bmask, mesh->n_elem)) || (rc = dalloc(ctx, &bi->boff, mesh->n_elem + 1)) ||
      (rc = dalloc(ctx, &cnt, mesh->n_elem)))
    return rc;
  if (mesh->n_elem > 0)
    hipLaunchKernelGGL(k_bface_mask, dim3(grid1(mesh->n_elem, 128)), dim3(128), 0, ctx->stream, mesh->d_conn,
                       mesh->nn, mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, mesh->n_elem, bi->bmask, cnt);
  rc = exclusive_scan_i32(ctx, cnt, bi->boff, mesh->n_elem, &bi->nbf);
  dfree(cnt);
  return rc;
}

// Neumann pass with the face rule of `order` (FACE_QUAD_ORDER = 2 in
// apply_neumann) and, when h_gq is given, the flux closure's values at every
// boundary point (host, in tfm_mesh_boundary_points order) in place of g.
static int load_neumann_dev_ex(tfm_ctx *ctx, tfm_mesh *mesh, int dpn, int comp, double g, int axis, double coord,
                               double tol, int order, const double *h_gq, double *d_rhs, bool zeroed) {
  int rc;
  if ((rc = ensure_adjacency(mesh))) return rc;
  // face_quad (assembly/src/lib.rs:336-358): Line faces Gauss-Legendre,
  // Tri faces the triangle rule, Quad faces the tensor rule.
  double *d_rule = nullptr, *d_gq = nullptr;
  int nq = 0;
  if ((rc = upload_rule(ctx, face_cell(mesh->cell), order, &d_rule, &nq))) return rc;
  BFaceIndex bi;
  if (h_gq) {
    if ((rc = bface_index(ctx, mesh, &bi)) || (rc = dalloc(ctx, &d_gq, (size_t)bi.nbf * nq))) {
      dfree(d_rule);
      return rc;
    }
    if (bi.nbf > 0 && hipMemcpyAsync(d_gq, h_gq, sizeof(double) * bi.nbf * nq, hipMemcpyHostToDevice,
                                     ctx->stream) != hipSuccess) {
      dfree(d_rule);
      dfree(d_gq);
      return fail(ctx, TFM_E_HIP, "flux upload failed");
    }
  }
  if (mesh->n_nodes > 0)
    hipLaunchKernelGGL(k_load_neumann, dim3(grid1(mesh->n_nodes, 128)), dim3(128), 0, ctx->stream, mesh->d_coords,
                       mesh->pdim, mesh->d_conn, mesh->nn, mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, mesh->n_nodes,
                       d_rule, nq, dpn, comp, g, axis, coord, tol,
                       (int)(!h_gq && zeroed && axis >= 0 && axis < mesh->pdim), (const double *)d_gq,
                       (const unsigned char *)(h_gq ? bi.bmask : nullptr), (const int *)(h_gq ? bi.boff : nullptr),
                       d_rhs);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  dfree(d_rule);
  dfree(d_gq);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, "Neumann kernel failed");
  return TFM_OK;
}

int load_neumann_dev(tfm_ctx *ctx, tfm_mesh *mesh, int dpn, int comp, double g, int axis, double coord,
                     double tol, double *d_rhs, bool zeroed) {
  return load_neumann_dev_ex(ctx, mesh, dpn, comp, g, axis, coord, tol, 2, nullptr, d_rhs, zeroed);
}

int load_neumann_qp_dev(tfm_ctx *ctx, tfm_mesh *mesh, int dpn, int order, const double *h_gq, double *d_rhs) {
  return load_neumann_dev_ex(ctx, mesh, dpn, -1, 0.0, -1, 0.0, 0.0, order, h_gq, d_rhs, false);
}

int boundary_points_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, double *h_x, double *h_n, int64_t *count) {
  int rc;
  double *d_rule = nullptr, *d_x = nullptr, *d_n = nullptr;
  int nq = 0;
  if ((rc = upload_rule(ctx, face_cell(mesh->cell), order, &d_rule, &nq))) return rc;
  BFaceIndex bi;
  if ((rc = bface_index(ctx, mesh, &bi))) { dfree(d_rule); return rc; }
  const int64_t np = bi.nbf * nq;
  *count = np;
  if ((h_x || h_n) && np > 0) {
    if ((rc = dalloc(ctx, &d_x, (size_t)np * mesh->pdim)) || (rc = dalloc(ctx, &d_n, (size_t)np * mesh->pdim))) {
      dfree(d_rule);
      dfree(d_x);
      return rc;
    }
    hipLaunchKernelGGL(k_bface_points, dim3(grid1(mesh->n_elem, 128)), dim3(128), 0, ctx->stream, mesh->d_coords,
                       mesh->pdim, mesh->d_conn, mesh->nn, mesh->cell, mesh->n_elem, (const unsigned char *)bi.bmask,
                       (const int *)bi.boff, (const double *)d_rule, nq, d_x, d_n);
    hipError_t e = hipSuccess;
    if (h_x) e = hipMemcpyAsync(h_x, d_x, sizeof(double) * np * mesh->pdim, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && h_n)
      e = hipMemcpyAsync(h_n, d_n, sizeof(double) * np * mesh->pdim, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(d_x);
    dfree(d_n);
    if (e != hipSuccess) { dfree(d_rule); return fail(ctx, TFM_E_HIP, "boundary points failed"); }
  }
  dfree(d_rule);
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
                        double tol, const double *__restrict__ hqv, const unsigned char *__restrict__ bmask,
                        const int *__restrict__ boff, const int *__restrict__ krp, const int *__restrict__ kci,
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
      if (bmask ? !((bmask[e] >> fi) & 1) : !face_on_boundary(FT, fi, en, e, nn, conn, n2e, pb, pe)) continue;
      double P[4][3];
      for (int k = 0; k < FT.fn_; k++)
        for (int d = 0; d < 3; d++) P[k][d] = d < pdim ? coords[(int64_t)en[FT.idx[fi][k]] * pdim + d] : 0.0;
      for (int q = 0; q < fnq; q++) {
        double w, measure, xp[3];
        const double *ns;
        face_point(FT, P, pdim, frule + q * kRuleStride, w, ns, xp, measure);
        const double hq = hqv ? hqv[bface_point(bmask, boff, e, fi, fnq, q)]
                              : ((axis < 0 || fabs(xp[axis] - coord) <= tol) ? h : 0.0);
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
                       double coord, double tol, int order, const double *h_hq, tfm_matrix **out) {
  tfm_mesh *mesh = const_cast<tfm_mesh *>(cmesh);
  int rc;
  tfm_matrix *K = nullptr;
  if ((rc = tfm_assemble(ctx, mesh, op, dpn, &K))) return rc;
  if ((rc = ensure_adjacency(mesh))) {
    tfm_matrix_destroy(K);
    return rc;
  }
  hipStream_t st = ctx->stream;
  const int64_t nrows = mesh->n_nodes * dpn;
  double *d_rule = nullptr, *d_hq = nullptr;
  int nq = 0, *cnt = nullptr, *aux = nullptr;
  tfm_matrix *R = new tfm_matrix();
  R->ctx = ctx;
  BFaceIndex bi;
  auto cleanup = [&]() {
    dfree(d_rule);
    dfree(d_hq);
    dfree(cnt);
    dfree(aux);
    tfm_matrix_destroy(K);
  };
  if ((rc = upload_rule(ctx, face_cell(mesh->cell), order, &d_rule, &nq)) || (rc = dalloc(ctx, &cnt, nrows)) ||
      (rc = dalloc(ctx, &aux, 3)) || (rc = dalloc(ctx, &R->d_rp, nrows + 1))) {
    cleanup();
    tfm_matrix_destroy(R);
    return rc;
  }
  if (h_hq) {
    if ((rc = bface_index(ctx, mesh, &bi)) || (rc = dalloc(ctx, &d_hq, (size_t)bi.nbf * nq))) {
      cleanup();
      tfm_matrix_destroy(R);
      return rc;
    }
    if (bi.nbf > 0 && hipMemcpy(d_hq, h_hq, sizeof(double) * bi.nbf * nq, hipMemcpyHostToDevice) != hipSuccess) {
      cleanup();
      tfm_matrix_destroy(R);
      return fail(ctx, TFM_E_HIP, "Robin coefficient upload failed");
    }
  }
  const unsigned char *bm = h_hq ? bi.bmask : nullptr;
  const int *bo = h_hq ? bi.boff : nullptr;
  const int init[3] = {-1, -1, 0};  // largest column, largest non-empty row, overflow flag
  TFM_CHECK_HIP(ctx, hipMemcpy(aux, init, sizeof init, hipMemcpyHostToDevice));
  const unsigned g = (unsigned)((nrows + 127) / 128);
  if (nrows > 0)
    hipLaunchKernelGGL(k_robin<0>, dim3(g), dim3(128), 0, st, mesh->d_coords, mesh->pdim, mesh->d_conn, mesh->nn,
                       mesh->cell, mesh->d_n2e_ptr, mesh->d_n2e, nrows, dpn, (const double *)d_rule, nq, h, axis,
                       coord, tol, (const double *)d_hq, bm, bo, (const int *)K->d_rp, (const int *)K->d_ci,
                       (const double *)K->d_v, K->nnz ? K->nrows : 0, cnt, (const int *)nullptr, (int *)nullptr,
                       (double *)nullptr, aux, aux + 2);
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
                       coord, tol, (const double *)d_hq, bm, bo, (const int *)K->d_rp, (const int *)K->d_ci,
                       (const double *)K->d_v, K->nnz ? K->nrows : 0, (int *)nullptr, (const int *)R->d_rp, R->d_ci,
                       R->d_v, aux, aux + 2);
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
  return assemble_robin_dev(ctx, mesh, op, dpn, h, axis, coord, tol, 2, nullptr, out);
}

extern "C" int tfm_assemble_robin_qp(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn,
                                     int order, const double *hq, tfm_matrix **out) {
  if (!ctx || !mesh || !op || !out || !hq || dpn < 1 || dpn > 3) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  return assemble_robin_dev(ctx, mesh, op, dpn, 0.0, -1, 0.0, 0.0, order, hq, out);
}

extern "C" int tfm_mesh_boundary_points(const tfm_mesh *mesh, int order, double *x, double *normal,
                                        int64_t *count) {
  if (!mesh || !count) return TFM_E_INVALID;
  tfm_ctx *ctx = mesh->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  return boundary_points_dev(ctx, const_cast<tfm_mesh *>(mesh), order, x, normal, count);
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

extern "C" int tfm_mesh_quad_points(const tfm_mesh *mesh, int quad_order, double *xq, int64_t *nq) {
  if (!mesh || !nq) return TFM_E_INVALID;
  tfm_ctx *ctx = mesh->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  int q = 0;
  if ((rc = element_rule_points(ctx, mesh->cell, quad_order, &q))) return rc;
  *nq = q;
  if (!xq || mesh->n_elem == 0) return TFM_OK;
  double *d_rule = nullptr, *d_x = nullptr;
  if ((rc = upload_rule(ctx, mesh->cell, quad_order, &d_rule, &q))) return rc;
  const int64_t nt = mesh->n_elem * q;
  if ((rc = dalloc(ctx, &d_x, (size_t)nt * mesh->pdim))) { dfree(d_rule); return rc; }
  hipLaunchKernelGGL(k_qp_coords, dim3(grid1(nt, 256)), dim3(256), 0, ctx->stream, mesh->d_coords, mesh->pdim,
                     mesh->d_conn, mesh->nn, mesh->n_elem, d_rule, q, d_x);
  hipError_t e = hipMemcpyAsync(xq, d_x, sizeof(double) * nt * mesh->pdim, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dfree(d_rule);
  dfree(d_x);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("quadrature points: ") + hipGetErrorString(e));
  return TFM_OK;
}

struct QpArg { int order; const double *qv; };

extern "C" int tfm_load_source_qp(tfm_ctx *ctx, const tfm_mesh *mesh, int quad_order, const double *sq,
                                  double *rhs) {
  if (!ctx || !mesh || !rhs || (!sq && mesh->n_elem > 0)) return TFM_E_INVALID;
  QpArg a{quad_order, sq};
  return host_rhs_roundtrip(ctx, mesh, 1, rhs,
                            [](tfm_ctx *c, tfm_mesh *m, double *d, const void *p) {
                              const QpArg *s = (const QpArg *)p;
                              return load_source_qp_dev(c, m, s->order, s->qv, d);
                            },
                            &a);
}

extern "C" int tfm_load_body_qp(tfm_ctx *ctx, const tfm_mesh *mesh, int quad_order, const double *bq,
                                double *rhs) {
  if (!ctx || !mesh || !rhs || (!bq && mesh->n_elem > 0)) return TFM_E_INVALID;
  QpArg a{quad_order, bq};
  return host_rhs_roundtrip(ctx, mesh, mesh->dim, rhs,
                            [](tfm_ctx *c, tfm_mesh *m, double *d, const void *p) {
                              const QpArg *s = (const QpArg *)p;
                              return load_body_qp_dev(c, m, s->order, s->qv, d);
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

struct NeuQpArg { int dpn, order; const double *gq; };

extern "C" int tfm_load_neumann_qp(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, int order, const double *gq,
                                   double *rhs) {
  if (!ctx || !mesh || !rhs || !gq || dpn < 1 || dpn > 3) return TFM_E_INVALID;
  NeuQpArg a{dpn, order, gq};
  return host_rhs_roundtrip(ctx, mesh, dpn, rhs,
                            [](tfm_ctx *c, tfm_mesh *m, double *d, const void *p) {
                              const NeuQpArg *s = (const NeuQpArg *)p;
                              return load_neumann_qp_dev(c, m, s->dpn, s->order, s->gq, d);
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
