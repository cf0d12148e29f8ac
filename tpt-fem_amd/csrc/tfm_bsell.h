// tfm_bsell.h -- building the node-blocked sliced ELL copy of a CSR matrix.
//
// Block row b = the BS DOF rows of node b.  A slice is a run of <= 64
// consecutive block rows stored lane-interleaved: for k < width, the block
// column of (slice, k, lane) sits at soff[slice] + 64k + lane and entry (i, j)
// of that block at (soff[slice] + 64k) * BS^2 + (i*BS + j) * 64 + lane.
// Slices are either uniform (slice s = block rows [64s, 64s + 64)) or given by
// explicit bounds sb[s] .. sb[s+1] (the batched solver aligns them with its
// problems).  Padding blocks point at column 0 with zero values.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "tfm_internal.h"

namespace tfm {

// Smallest column block among the current heads of a block row's BS sorted rows.
template <int BS>
__device__ __forceinline__ int bs_merge_step(const int *__restrict__ ci, const int *j, const int *e) {
  int m = INT_MAX;
#pragma unroll
  for (int i = 0; i < BS; i++)
    if (j[i] < e[i]) m = min(m, ci[j[i]] / BS);
  return m;
}

// Distinct column blocks of each block row.
template <int BS>
__global__ void k_bs_count(int nbr, const int *__restrict__ rp, const int *__restrict__ ci, int *__restrict__ cnt) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbr) return;
  int j[BS], e[BS];
#pragma unroll
  for (int i = 0; i < BS; i++) {
    j[i] = rp[b * BS + i];
    e[i] = rp[b * BS + i + 1];
  }
  int k = 0;
  for (int m; (m = bs_merge_step<BS>(ci, j, e)) != INT_MAX; k++)
#pragma unroll
    for (int i = 0; i < BS; i++)
      while (j[i] < e[i] && ci[j[i]] / BS == m) j[i]++;
  cnt[b] = k;
}

__device__ __forceinline__ void bs_slice(const int *sb, int nbr, int s, int &r0, int &r1) {
  if (sb) {
    r0 = sb[s];
    r1 = sb[s + 1];
  } else {
    r0 = s * 64;
    r1 = min(nbr, r0 + 64);
  }
}

// 64 * (widest block row) per slice.
static __global__ void k_bs_width(int nbr, int nslices, const int *__restrict__ sb, const int *__restrict__ cnt,
                           int *__restrict__ w) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int r0, r1;
  bs_slice(sb, nbr, s, r0, r1);
  int m = 0;
  for (int b = r0; b < r1; b++) m = max(m, cnt[b]);
  w[s] = 64 * m;
}

// Block columns, padding and the CSR -> value-slot map; one thread per
// (slice, lane).
template <int BS>
__global__ void k_bs_fill(int nbr, int nslices, const int *__restrict__ sb, const int *__restrict__ rp,
                          const int *__restrict__ ci, const int *__restrict__ soff, int *__restrict__ bcol,
                          int *__restrict__ map) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslices * 64) return;
  constexpr int BB = BS * BS;
  const int s = t >> 6, lane = t & 63;
  int r0, r1;
  bs_slice(sb, nbr, s, r0, r1);
  const int b = r0 + lane;
  const int o = soff[s], w = (soff[s + 1] - o) >> 6;
  int k = 0;
  if (b < r1) {
    int j[BS], e[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) {
      j[i] = rp[b * BS + i];
      e[i] = rp[b * BS + i + 1];
    }
    for (int m; (m = bs_merge_step<BS>(ci, j, e)) != INT_MAX; k++) {
      bcol[o + k * 64 + lane] = m;
#pragma unroll
      for (int i = 0; i < BS; i++)
        for (; j[i] < e[i] && ci[j[i]] / BS == m; j[i]++)
          map[j[i]] = (o + k * 64) * BB + (i * BS + ci[j[i]] % BS) * 64 + lane;
    }
  }
  for (; k < w; k++) bcol[o + k * 64 + lane] = 0;
}

// Turns the fill's CSR -> slot map into the slot -> CSR map (slots of absent
// entries keep -1, the caller's clear).
static __global__ void k_bs_invert(int64_t nnz, const int *__restrict__ map, int *__restrict__ inv) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nnz) inv[map[j]] = (int)j;
}

// Values of the CSR into their slots, one thread per slot (coalesced stores;
// slots of absent entries get zero).
template <typename VT, typename ST>
__global__ void k_bs_values(int64_t nslots, const int *__restrict__ inv, const ST *__restrict__ src,
                            VT *__restrict__ dst) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nslots) {
    const int j = inv[q];
    dst[q] = j >= 0 ? (VT)src[j] : (VT)0;
  }
}

// Replaces *map (CSR -> slot, nnz entries) by its inverse (nslots entries).
static inline int bs_invert_map(tfm_ctx *ctx, int **map, int64_t nnz, int64_t nslots) {
  int *inv = nullptr;
  int rc = dalloc(ctx, &inv, nslots > 0 ? nslots : 1);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  TFM_CHECK_HIP(ctx, hipMemsetAsync(inv, 0xff, sizeof(int) * (nslots > 0 ? nslots : 1), st));
  if (nnz > 0)
    hipLaunchKernelGGL(k_bs_invert, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, nnz, (const int *)*map, inv);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  dfree(*map);  // waits for the inversion
  *map = inv;
  return TFM_OK;
}

}  // namespace tfm
