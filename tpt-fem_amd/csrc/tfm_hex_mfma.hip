// tfm_hex_mfma.hip -- FP32 element matrices of 3-D isotropic elasticity on
// Hex8 cells with the matrix cores (precision form).
//
// The element matrix (crates/tpt-fem-elasticity/src/lib.rs:284-307: sum over
// the quadrature points of w det B^T D B) of an isotropic material is
//   K[(i,a),(j,b)] = E (lambda P[(i,a),(j,b)] + mu P[(i,b),(j,a)]
//                       + mu delta_ab sum_c P[(i,c),(j,c)])
// with P = H H^T, H[(i,a), q] = sqrt(w_q |det J_q|) dN_i/dx_a (x_q): a 24 x nq
// by nq x 24 product, which runs on v_mfma_f32_32x32x2f32 (rows and columns
// padded to 32, nq padded to even; four instructions for the 2x2x2 rule).
// Each element's matrix is formed once, where the node-gather kernel forms
// it once per element node; the values go to a per-element table
// (ke[e][24][24]) from which the tile kernel's numeric pass sums the CSR
// entries in ascending element order.
//
// One wavefront per element at a time: lanes (q, i) reduce the Jacobian of
// quadrature point q over the eight nodes with cross-lane shuffles (FP64,
// determinant in FP64), write their node's scaled gradients into the wave's
// slice of LDS, the MFMA accumulates P, and the lanes combine P's entries
// into the 576 values of K, written contiguously.
#include <cmath>

#include "tfm_elem.h"
#include "tfm_internal.h"

namespace tfm {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kHexWaves = 4;       // wavefronts per block
constexpr int kHexKMax = 32;       // quadrature points (padded): up to GL order 3 (27)
constexpr int kHexHS = kHexKMax + 1;
constexpr int kHexPS = 25;         // LDS row stride of P

__device__ __forceinline__ void hex_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
  return __shfl_xor(v, m, 64);
}

__global__ __launch_bounds__(64 * kHexWaves) void k_hex_ke_mfma(int64_t ne, const double *__restrict__ coords,
                                                                 const int *__restrict__ conn,
                                                                 const double *__restrict__ coef_e, double coef,
                                                                 float lam, float mu, const double *__restrict__ rule,
                                                                 int nq, float *__restrict__ ke) {
  __shared__ double s_x[kHexWaves][8][3];
  __shared__ float s_h[kHexWaves][24 * kHexHS];
  __shared__ float s_p[kHexWaves][24 * kHexPS];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * kHexWaves;
  const int kpad = (nq + 1) & ~1;
  double (*X)[3] = s_x[w];
  float *H = s_h[w];
  float *Pm = s_p[w];
  for (int64_t e = (int64_t)blockIdx.x * kHexWaves + w; e < ne; e += nwaves) {
    if (l < 8) {
      const int id = conn[e * 8 + l];
#pragma unroll
      for (int r = 0; r < 3; r++) X[l][r] = coords[(int64_t)id * 3 + r];
    }
    const float E = (float)(coef_e ? coef_e[e] : coef);
    hex_wave_sync();
    // scaled gradients: lane (q = q0 + l / 8, node i = l % 8)
    const int i = l & 7;
    for (int q0 = 0; q0 < kpad; q0 += 8) {
      const int q = q0 + (l >> 3);
      float g[3] = {0.f, 0.f, 0.f};
      if (q < nq) {
        const double *R = rule + q * kRuleStride;
        const double dn[3] = {R[12 + 3 * i], R[12 + 3 * i + 1], R[12 + 3 * i + 2]};
        // J[r][c] = sum_i X_i[r] dN_i/dxi_c over the 8 lanes of the group
        double J[9];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) J[r * 3 + c] = X[i][r] * dn[c];
#pragma unroll
        for (int m = 1; m < 8; m <<= 1)
#pragma unroll
          for (int k = 0; k < 9; k++) J[k] += shfl_xor_d(J[k], m);
        const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                           J[2] * (J[3] * J[7] - J[4] * J[6]);
        const double id = 1.0 / det;
        // inverse: iv[c][r] = d xi_c / d x_r
        float iv[9];
        iv[0] = (float)((J[4] * J[8] - J[5] * J[7]) * id);
        iv[1] = (float)((J[2] * J[7] - J[1] * J[8]) * id);
        iv[2] = (float)((J[1] * J[5] - J[2] * J[4]) * id);
        iv[3] = (float)((J[5] * J[6] - J[3] * J[8]) * id);
        iv[4] = (float)((J[0] * J[8] - J[2] * J[6]) * id);
        iv[5] = (float)((J[2] * J[3] - J[0] * J[5]) * id);
        iv[6] = (float)((J[3] * J[7] - J[4] * J[6]) * id);
        iv[7] = (float)((J[1] * J[6] - J[0] * J[7]) * id);
        iv[8] = (float)((J[0] * J[4] - J[1] * J[3]) * id);
        const float sw = sqrtf((float)(R[0] * fabs(det)));
        const float d0 = (float)dn[0], d1 = (float)dn[1], d2 = (float)dn[2];
#pragma unroll
        for (int r = 0; r < 3; r++) g[r] = sw * (d0 * iv[r] + d1 * iv[3 + r] + d2 * iv[6 + r]);
      }
      if (q < kpad) {
#pragma unroll
        for (int r = 0; r < 3; r++) H[(3 * i + r) * kHexHS + q] = g[r];
      }
    }
    hex_wave_sync();
    // P = H H^T: A = H (rows padded to 32), B = H^T (the same lane values)
    f32x16 acc;
#pragma unroll
    for (int k = 0; k < 16; k++) acc[k] = 0.f;
    const int row = l & 31, kh = l >> 5;
    for (int s = 0; s < kpad; s += 2) {
      const float v = row < 24 ? H[row * kHexHS + s + kh] : 0.f;
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v, v, acc, 0, 0, 0);
    }
    // output register k of lane l: row 8 (k / 4) + 4 (l / 32) + k % 4, column l % 32
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int pr = 8 * (k >> 2) + 4 * kh + (k & 3);
      if (pr < 24 && row < 24) Pm[pr * kHexPS + row] = acc[k];
    }
    hex_wave_sync();
    float *out = ke + e * 576;
    for (int x = l; x < 576; x += 64) {
      const int r = x / 24, c = x - r * 24;
      const int ii = r / 3, a = r - 3 * ii, jj = c / 3, b = c - 3 * jj;
      float v = lam * Pm[r * kHexPS + c] + mu * Pm[(3 * ii + b) * kHexPS + 3 * jj + a];
      if (a == b)
        v += mu * (Pm[(3 * ii) * kHexPS + 3 * jj] + Pm[(3 * ii + 1) * kHexPS + 3 * jj + 1] +
                   Pm[(3 * ii + 2) * kHexPS + 3 * jj + 2]);
      out[x] = E * v;
    }
    hex_wave_sync();  // the slices are rewritten by the next element
  }
}

bool hex_mfma_eligible(const AsmPlan *pl) {
  return pl->mesh->cell == TFM_HEX && pl->dpn == 3 && pl->op.kind == TFM_OP_ELASTICITY &&
         pl->op.model == TFM_CONTINUUM3D && pl->nq >= 1 && pl->nq <= kHexKMax && !getenv("TFM_NO_MFMA");
}

int hex_ke_mfma(tfm_ctx *ctx, const AsmPlan *pl, float *ke) {
  const int64_t ne = pl->mesh->n_elem;
  if (ne == 0) return TFM_OK;
  const double nu = pl->op.nu;
  const float lam = (float)(nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
  const float mu = (float)(1.0 / (2.0 * (1.0 + nu)));
  const int64_t blocks = std::min<int64_t>((ne + kHexWaves - 1) / kHexWaves, 8192);
  hipLaunchKernelGGL(k_hex_ke_mfma, dim3((unsigned)blocks), dim3(64 * kHexWaves), 0, ctx->stream, ne,
                     pl->mesh->d_coords, pl->mesh->d_conn, pl->d_coef_e, pl->op.coef, lam, mu, pl->d_tab, pl->nq,
                     ke);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  return TFM_OK;
}

}  // namespace tfm
