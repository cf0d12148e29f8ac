// tfm_mf.hip -- matrix-free FP32 stiffness operator for P1-tet systems on box
// meshes, the inner operator of the precision form's PCG.
//
// The precision form integrates the elements in FP32 (with FP64 Jacobian
// determinants) and runs its CG in FP32; the outer residual r = f - K u uses
// the assembled FP64 matrix.  On a box_mesh (crates/tpt-fem-mesh-gen/src/
// lib.rs:422-470: node (i, j, k) is (i (ny + 1) + j) (nz + 1) + k, element
// 6 cube + t with cube (i ny + j) nz + k, tet t spanning cube corners
// {0, a_t, b_t, 6}) the FP32 product q = K p needs no stored matrix: every
// element's contribution is recomputed from the node coordinates, the
// element's coefficient and p, which moves about 150 bytes per node instead of
// the FP32 matrix's ~600.
//
// One block pass covers a brick of 4 x 8 x 8 nodes (i, j, k).  Its threads
// first evaluate every cube that touches the brick (5 x 9 x 9, the brick's
// cubes plus one layer below in each axis), each cube's six tets once, and
// leave the cube's eight corner forces in LDS; then each thread sums its
// node's forces from its eight cubes in a fixed order.  Cubes of the halo
// layer are evaluated by both neighbouring bricks with identical arithmetic,
// so the product is deterministic and independent of the launch shape.
//
// Per tet with edge vectors d_a, d_b, d_6 from corner 0 (FP64) and
// det = d_a . (d_b x d_6) (FP64): scaled gradients G1 = d_b x d_6,
// G2 = d_6 x d_a, G3 = d_a x d_b, G0 = -(G1 + G2 + G3) (FP32), so that the
// physical gradients are G_i / det.  With H = sum_i (p_i - p_0) (x) G_i the
// element force at node i is s sigma(H) G_i, s = coef / (6 |det|), where
// sigma(H) = lambda tr(H) I + mu (H + H^T) for elasticity (lambda, mu per
// unit modulus) and sigma(H) = H for Poisson: vol B^T D B p with
// vol = |det| / 6 and B = G / det.
#include <algorithm>
#include <climits>
#include <cmath>

#include "tfm_elem.h"
#include "tfm_internal.h"
#include "tfm_pcg.h"

#pragma clang fp contract(fast)

namespace tfm {

constexpr int kMfBI = 4, kMfBJ = 8, kMfBK = 8;                    // brick nodes per axis
constexpr int kMfCI = kMfBI + 1, kMfCJ = kMfBJ + 1, kMfCK = kMfBK + 1;  // cubes touching it
constexpr int kMfCubes = kMfCI * kMfCJ * kMfCK;
static_assert(kMfBI * kMfBJ * kMfBK == kBlock, "one brick node per thread");

// Cube corners (the box_mesh order) at offsets (di, dj, dk):
// 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0) 4 (0,0,1) 5 (1,0,1) 6 (1,1,1) 7 (0,1,1).
__device__ __forceinline__ int corner_di(int c) { return (c == 1 || c == 2 || c == 5 || c == 6) ? 1 : 0; }
__device__ __forceinline__ int corner_dj(int c) { return (c == 2 || c == 3 || c == 6 || c == 7) ? 1 : 0; }
__device__ __forceinline__ int corner_dk(int c) { return c >= 4 ? 1 : 0; }
// corner at offset (oi, oj, ok)
__device__ __forceinline__ int corner_at(int oi, int oj, int ok) {
  const int lo = oj ? (oi ? 2 : 3) : (oi ? 1 : 0);
  return ok ? (lo == 0 ? 4 : lo == 1 ? 5 : lo == 2 ? 6 : 7) : lo;
}

template <typename T>
__device__ __forceinline__ void cross3(const T (&u)[3], const T (&v)[3], T (&w)[3]) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

// The corner forces of one cube, summed over its six tets in order, stored
// into sf[(c * DPN + a) * kMfCubes] as soon as corner c is final (corners 0
// and 6 belong to every tet, corner 2 only to tets 0-1, 3 to 1-2, 7 to 2-3,
// 4 to 3-4, 1 and 5 to 0 and 5).
template <int DPN>
__device__ __forceinline__ void mf_cube(const MfBox &B, int ci, int cj, int ck, const float *__restrict__ z,
                                        const float *__restrict__ pold, float beta, float *sf) {
  const int n1 = B.n[1] + 1, n2 = B.n[2] + 1;
  int id[8];
#pragma unroll
  for (int c = 0; c < 8; c++) id[c] = ((ci + corner_di(c)) * n1 + (cj + corner_dj(c))) * n2 + ck + corner_dk(c);
  constexpr int TA[6] = {1, 2, 3, 7, 4, 5}, TB[6] = {2, 3, 7, 4, 5, 1};
  // FP64 edge vectors from corner 0, the determinants, then FP32 copies
  float df[8][3], cf[8][3], s[6];
  {
    double x0[3], d[8][3];
#pragma unroll
    for (int q = 0; q < 3; q++) x0[q] = B.coords[id[0] * 3 + q];
#pragma unroll
    for (int c = 1; c < 8; c++)
#pragma unroll
      for (int q = 0; q < 3; q++) d[c][q] = B.coords[id[c] * 3 + q] - x0[q];
    const int e0 = 6 * ((ci * B.n[1] + cj) * B.n[2] + ck);
#pragma unroll
    for (int t = 0; t < 6; t++) {
      double cb[3];
      cross3<double>(d[TB[t]], d[6], cb);
      const double det = d[TA[t]][0] * cb[0] + d[TA[t]][1] * cb[1] + d[TA[t]][2] * cb[2];
      // s = coef / (6 |det|) = coef vol / det^2
      s[t] = (B.coef32 ? B.coef32[e0 + t] : B.coef) * __builtin_amdgcn_rcpf(6.0f * fabsf((float)det));
#pragma unroll
      for (int q = 0; q < 3; q++) cf[TB[t]][q] = (float)cb[q];
    }
#pragma unroll
    for (int c = 1; c < 8; c++)
#pragma unroll
      for (int q = 0; q < 3; q++) df[c][q] = (float)d[c][q];
  }
  // p at the corners (p = z + beta p_old), as differences from corner 0
  float dp[8][DPN];
  {
    float p0[DPN];
#pragma unroll
    for (int a = 0; a < DPN; a++) p0[a] = z[id[0] * DPN + a] + beta * pold[id[0] * DPN + a];
#pragma unroll
    for (int c = 1; c < 8; c++)
#pragma unroll
      for (int a = 0; a < DPN; a++) dp[c][a] = (z[id[c] * DPN + a] + beta * pold[id[c] * DPN + a]) - p0[a];
  }
  float F[8][DPN];
#pragma unroll
  for (int c = 0; c < 8; c++)
#pragma unroll
    for (int a = 0; a < DPN; a++) F[c][a] = 0.f;
  auto flush = [&](int c) {
#pragma unroll
    for (int a = 0; a < DPN; a++) sf[(c * DPN + a) * kMfCubes] = F[c][a];
  };
#pragma unroll
  for (int t = 0; t < 6; t++) {
    const int a = TA[t], b = TB[t];
    float G[4][3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      G[1][q] = cf[b][q];
      G[2][q] = -cf[a][q];
    }
    cross3<float>(df[a], df[b], G[3]);
#pragma unroll
    for (int q = 0; q < 3; q++) G[0][q] = -(G[1][q] + G[2][q] + G[3][q]);
    const int v[4] = {0, a, b, 6};
    if constexpr (DPN == 1) {
      float h[3];
#pragma unroll
      for (int q = 0; q < 3; q++) h[q] = s[t] * (dp[a][0] * G[1][q] + dp[b][0] * G[2][q] + dp[6][0] * G[3][q]);
#pragma unroll
      for (int i = 0; i < 4; i++) F[v[i]][0] += h[0] * G[i][0] + h[1] * G[i][1] + h[2] * G[i][2];
    } else {
      float H[3][3];
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int q = 0; q < 3; q++) H[r][q] = dp[a][r] * G[1][q] + dp[b][r] * G[2][q] + dp[6][r] * G[3][q];
      const float lt = s[t] * B.lam * (H[0][0] + H[1][1] + H[2][2]), sm = s[t] * B.mu;
      float S[3][3];
#pragma unroll
      for (int r = 0; r < 3; r++) {
        S[r][r] = lt + 2.0f * sm * H[r][r];
#pragma unroll
        for (int q = r + 1; q < 3; q++) S[r][q] = S[q][r] = sm * (H[r][q] + H[q][r]);
      }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) F[v[i]][r] += S[r][0] * G[i][0] + S[r][1] * G[i][1] + S[r][2] * G[i][2];
    }
    if (t == 1) flush(2);
    if (t == 2) flush(3);
    if (t == 3) flush(7);
    if (t == 4) flush(4);
  }
  flush(0);
  flush(1);
  flush(5);
  flush(6);
}

// PCG direction step on the matrix-free operator: p_i = z_i + beta_i p_{i-1}
// on the fly, q = K p_i (rows of fixed DOFs zero), partial p_i . q -- the
// same contract as the assembled forms' direction kernels.
template <int DPN>
__global__ __launch_bounds__(kBlock) void k_pcg_dir_mf(MfBox B, const float *__restrict__ z,
                                                        const float *__restrict__ pold, float *__restrict__ pnew,
                                                        float *__restrict__ q, Scal *sc, double *part, int e) {
  __shared__ float s_f[8 * DPN * kMfCubes];
  double beta_d;
  if (!dir_prologue(sc, part, e, &beta_d)) return;  // uniform per block
  const float beta = (float)beta_d;
  const int t = threadIdx.x;
  const int nbi = (B.n[0] + kMfBI) / kMfBI, nbj = (B.n[1] + kMfBJ) / kMfBJ, nbk = (B.n[2] + kMfBK) / kMfBK;
  const int nbricks = nbi * nbj * nbk;
  const int per = (nbricks + gridDim.x - 1) / gridDim.x;
  const int w = xcd_item(blockIdx.x, gridDim.x);  // consecutive bricks on one XCD
  const int b1 = min(nbricks, (w + 1) * per);
  const int n1 = B.n[1] + 1, n2 = B.n[2] + 1;
  double dot = 0.0;
  for (int br = w * per; br < b1; br++) {
    const int bk = br % nbk, bj = (br / nbk) % nbj, bi = br / (nbk * nbj);
    const int i0 = bi * kMfBI, j0 = bj * kMfBJ, k0 = bk * kMfBK;
    // cubes (i0 - 1 .. i0 + BI - 1) x ... into LDS (zeros outside the mesh)
    for (int lc = t; lc < kMfCubes; lc += kBlock) {
      const int uk = lc % kMfCK, uj = (lc / kMfCK) % kMfCJ, ui = lc / (kMfCK * kMfCJ);
      const int ci = i0 - 1 + ui, cj = j0 - 1 + uj, ck = k0 - 1 + uk;
      if (ci >= 0 && ci < B.n[0] && cj >= 0 && cj < B.n[1] && ck >= 0 && ck < B.n[2]) {
        mf_cube<DPN>(B, ci, cj, ck, z, pold, beta, s_f + lc);
      } else {
#pragma unroll
        for (int c = 0; c < 8 * DPN; c++) s_f[c * kMfCubes + lc] = 0.f;
      }
    }
    __syncthreads();
    const int tk = t % kMfBK, tj = (t / kMfBK) % kMfBJ, ti = t / (kMfBK * kMfBJ);
    const int i = i0 + ti, j = j0 + tj, k = k0 + tk;
    if (i <= B.n[0] && j <= B.n[1] && k <= B.n[2]) {
      float acc[DPN];
#pragma unroll
      for (int a = 0; a < DPN; a++) acc[a] = 0.f;
      // the node's cubes (i - 1 + di, ...) hold it at corner offset (1 - di, ...)
#pragma unroll
      for (int di = 0; di < 2; di++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++)
#pragma unroll
          for (int dk = 0; dk < 2; dk++) {
            const int lc = ((ti + di) * kMfCJ + (tj + dj)) * kMfCK + (tk + dk);
            const int c = corner_at(1 - di, 1 - dj, 1 - dk);
#pragma unroll
            for (int a = 0; a < DPN; a++) acc[a] += s_f[(c * DPN + a) * kMfCubes + lc];
          }
      const int g0 = (i * n1 + j) * n2 + k;
#pragma unroll
      for (int a = 0; a < DPN; a++) {
        const int g = g0 * DPN + a;
        const float pn = z[g] + beta * pold[g];
        const float qv = B.fixed[g] ? 0.f : acc[a];
        pnew[g] = pn;
        q[g] = qv;
        dot += (double)pn * (double)qv;
      }
    }
    __syncthreads();
  }
  put_part(part, P_PQ, dot);
}

__global__ void k_mf_round(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

bool mf_box_setup(const tfm_system *s, MfBox *out) {
  if (!s->has_plan || !s->d_fixed) return false;
  const AsmPlan &pl = s->plan;
  const tfm_mesh *m = pl.mesh;
  if (!m || m->cell != TFM_TET || m->box_n[0] <= 0 || m->box_n[1] <= 0 || m->box_n[2] <= 0) return false;
  const int64_t nx = m->box_n[0], ny = m->box_n[1], nz = m->box_n[2];
  if ((nx + 1) * (ny + 1) * (nz + 1) != m->n_nodes || 6 * nx * ny * nz != m->n_elem) return false;
  if (m->n_nodes * pl.dpn >= INT32_MAX || m->n_elem >= INT32_MAX) return false;  // int32 node, DOF and element ids
  const bool poisson = pl.op.kind == TFM_OP_POISSON && pl.dpn == 1;
  const bool elastic = pl.op.kind == TFM_OP_ELASTICITY && pl.op.model == TFM_CONTINUUM3D && pl.dpn == 3;
  if (!poisson && !elastic) return false;
  MfBox B{};
  B.dpn = pl.dpn;
  B.n[0] = (int)nx;
  B.n[1] = (int)ny;
  B.n[2] = (int)nz;
  B.coords = m->d_coords;
  B.coef = (float)pl.op.coef;
  B.fixed = s->d_fixed;
  if (elastic) {
    const double nu = pl.op.nu;
    B.lam = (float)(nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
    B.mu = (float)(1.0 / (2.0 * (1.0 + nu)));
  }
  *out = B;
  return true;
}

int mf_coef32(tfm_ctx *ctx, const tfm_system *s, MfBox *B, float **owned) {
  *owned = nullptr;
  const double *ce = s->plan.d_coef_e;
  if (!ce) return TFM_OK;
  const int64_t ne = s->plan.mesh->n_elem;
  int rc = dalloc(ctx, owned, ne);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mf_round, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, ctx->stream, ne, ce, *owned);
  B->coef32 = *owned;
  return TFM_OK;
}

// One block per resident slot: a persistent grid larger than what the CUs
// hold at once would leave a tail of lone blocks.
int mf_grid(const MfBox &B) {
  const int64_t nbr = (int64_t)((B.n[0] + kMfBI) / kMfBI) * ((B.n[1] + kMfBJ) / kMfBJ) * ((B.n[2] + kMfBK) / kMfBK);
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  hipError_t e = B.dpn == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pcg_dir_mf<1>, kBlock, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pcg_dir_mf<3>, kBlock, 0);
  if (e != hipSuccess || per < 1) per = 1;
  int64_t nb = std::min<int64_t>((int64_t)cus * per, nbr);
  nb = std::min<int64_t>(kMaxBlocks / kXcds * kXcds, std::max<int64_t>(kXcds, (nb + kXcds - 1) / kXcds * kXcds));
  return (int)nb;
}

void mf_dir_launch(const MfBox &B, hipStream_t st, int nb, const float *z, const float *pold, float *pnew, float *q,
                   Scal *sc, double *part, int e) {
  if (B.dpn == 1)
    hipLaunchKernelGGL(k_pcg_dir_mf<1>, dim3(nb), dim3(kBlock), 0, st, B, z, pold, pnew, q, sc, part, e);
  else
    hipLaunchKernelGGL(k_pcg_dir_mf<3>, dim3(nb), dim3(kBlock), 0, st, B, z, pold, pnew, q, sc, part, e);
}

}  // namespace tfm
