// tfm_pcg.h -- the PCG iteration's shared device pieces: launch shape,
// partial-sum slots, the scalar block and the fence-free reductions.
//
// Every kernel of one PCG iteration runs on the same grid (at most kMaxBlocks
// blocks of kBlock threads).  Each block writes its partial sums into a slot
// (put_part); every block of the next kernel re-sums the partials of that
// slot in one fixed order (sum_part), so all blocks agree on alpha, beta and
// convergence without fences or atomics, and the results are deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

struct tfm_system;
struct tfm_ctx;

namespace tfm {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 1024;

// Partial-sum slots (each kMaxBlocks doubles): p.q, r.z and r.r by iteration
// parity, and the refinement residual's ||r||^2, ||f||^2.
enum { P_PQ = 0, P_RZ0 = 1, P_RZ1 = 2, P_RR0 = 3, P_RR1 = 4, P_RES = 5, P_FF = 6, P_SLOTS = 8 };

struct Scal {
  double rr, rr0, tol2;
  int done, breakdown, it, max_it;
};

// Block sum of v, returned to every thread of the block (fixed order).
__device__ __forceinline__ double block_sum_all(double v) {
  __shared__ double sh[kBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; i++) t += sh[i];
  return t;
}

// Sum of the nb partials of one slot, identical in every block (each thread
// takes a fixed strided subset, then the fixed block tree).  The partials
// were written by an earlier kernel, so no fence is needed.
__device__ __forceinline__ double sum_part(const double *part, int slot, int nb) {
  const double *q = part + slot * kMaxBlocks;
  double t = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) t += q[i];
  return block_sum_all(t);
}

__device__ __forceinline__ void put_part(double *part, int slot, double v) {
  v = block_sum_all(v);
  if (threadIdx.x == 0) part[slot * kMaxBlocks + blockIdx.x] = v;
}

// Start of iteration i (parity e) in every block: converged / iteration
// limit (sc->done) decided identically by all blocks from the partials of the
// previous update; beta_i = rz_i / rz_{i-1}.  Returns false (uniformly per
// block) when the block has nothing to do.
__device__ __forceinline__ bool dir_prologue(Scal *sc, const double *part, int e, double *beta) {
  __shared__ int s_done, s_it;
  if (threadIdx.x == 0) {
    s_done = sc->done;
    s_it = sc->it;
  }
  __syncthreads();
  if (s_done) return false;
  const int nb = gridDim.x;
  const int it = s_it;
  if (it > 0) {
    double rr = sum_part(part, e ? P_RR0 : P_RR1, nb);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->rr = rr;  // latest recurrence residual, for the host
    if (rr <= sc->tol2 || it >= sc->max_it) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->done = 1;
        sc->rr = rr;
      }
      return false;
    }
  }
  *beta = 0.0;
  if (it > 0) {
    const double rzo = sum_part(part, e ? P_RZ1 : P_RZ0, nb);
    const double rzn = sum_part(part, e ? P_RZ0 : P_RZ1, nb);
    if (rzo > 0.0) *beta = rzn / rzo;  // 0 after a restart whose old slot is empty
  }
  return true;
}

// Matrix-free FP32 operator of a P1-tet system on a box mesh (tfm_mf.hip):
// cells per axis, the mesh coordinates, the element coefficients (or the
// scalar), the Lame factors per unit modulus (elasticity) and the fixed-DOF
// flags of the reduced system.
struct MfBox {
  int dpn = 1;
  int n[3] = {0, 0, 0};
  const double *coords = nullptr;
  const float *coef32 = nullptr;  // per element (FP32 copy), or nullptr: coef
  float coef = 0.f;
  float lam = 0.f, mu = 0.f;
  const unsigned char *fixed = nullptr;
};
// True (and *out filled) when the system qualifies: built from a box_mesh of
// P1 tets, Poisson (1 DOF per node) or 3-D elasticity (3), with fixed flags.
bool mf_box_setup(const tfm_system *s, MfBox *out);
// FP32 copy of the per-element coefficients into *owned (the caller frees it
// after the solve; nothing when the coefficient is a scalar).
int mf_coef32(tfm_ctx *ctx, const tfm_system *s, MfBox *B, float **owned);
// Blocks of the iteration on this operator (a multiple of 8, <= kMaxBlocks).
int mf_grid(const MfBox &B);
// The direction step of the PCG iteration (see k_pcg_dir_mf).
void mf_dir_launch(const MfBox &B, hipStream_t st, int nb, const float *z, const float *pold, float *pnew, float *q,
                   Scal *sc, double *part, int e);

}  // namespace tfm
