// tfm_core.hip -- context, meshes, node->element adjacency, generators.
//
// Mesh generation follows box_mesh (crates/tpt-fem-mesh-gen/src/lib.rs:422-495)
// and topopt Grid::new (crates/tpt-fem-topopt/src/lib.rs:89-116) node for
// node; both run as GPU kernels so a 128^3 box never round-trips through the
// host.  Built with -ffp-contract=off: coordinates must be the reference's
// min + (i/n)*(max-min), rounded exactly as Rust rounds it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "tfm_internal.h"

#pragma clang fp contract(off)

namespace tfm {

int fail(tfm_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int set_device(tfm_ctx *ctx) {
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return fail(ctx, TFM_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  return TFM_OK;
}

int scratch(tfm_ctx *ctx, size_t bytes, void **out) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->d_scratch) (void)hipFree(ctx->d_scratch);
    ctx->d_scratch = nullptr;
    ctx->scratch_bytes = 0;
    size_t want = std::max<size_t>(bytes, 1 << 20);
    TFM_CHECK_HIP(ctx, hipMalloc(&ctx->d_scratch, want));
    ctx->scratch_bytes = want;
  }
  *out = ctx->d_scratch;
  return TFM_OK;
}

// Sum of n non-negative int32 counts in int64 (integer addition, so the
// atomic order does not matter): the range check must not trust a sum that
// an int32 scan has already wrapped.
__global__ void k_sum_i64(const int *__restrict__ in, int64_t n, unsigned long long *__restrict__ out,
                          int *__restrict__ neg) {
  long long s = 0;
  int anyneg = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int v = in[i];
    anyneg |= v < 0;
    s += v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (anyneg) atomicOr(neg, 1);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)s);
}

int exclusive_scan_i32(tfm_ctx *ctx, const int *d_in, int *d_out, int64_t n, int64_t *total) {
  // out[0..n-1] = exclusive prefix, out[n] = total.  The total is formed in
  // int64 first; the int32 scan only runs when every prefix fits.
  size_t tmp = 0;
  TFM_CHECK_HIP(ctx, rocprim::exclusive_scan(nullptr, tmp, d_in, d_out, 0, (size_t)n,
                                             rocprim::plus<int>(), ctx->stream));
  void *d_tmp = nullptr;
  const size_t off = (tmp + 255) / 256 * 256;
  int rc = scratch(ctx, off + 16, &d_tmp);
  if (rc) return rc;
  unsigned long long *d_sum = (unsigned long long *)((char *)d_tmp + off);
  int *d_neg = (int *)(d_sum + 1);
  TFM_CHECK_HIP(ctx, hipMemsetAsync(d_sum, 0, 16, ctx->stream));
  if (n > 0) {
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_sum_i64, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d_in, n, d_sum, d_neg);
    TFM_CHECK_HIP(ctx, hipGetLastError());
  }
  unsigned long long hsum[2] = {0, 0};
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(hsum, d_sum, 16, hipMemcpyDeviceToHost, ctx->stream));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if ((int)(hsum[1] & 0xffffffffu)) return fail(ctx, TFM_E_INVALID, "scan: negative count");
  const int64_t t = (int64_t)hsum[0];
  if (t < 0 || t >= (int64_t)INT32_MAX) return fail(ctx, TFM_E_UNSUPPORTED, "scan total exceeds int32 range");
  if (n > 0)
    TFM_CHECK_HIP(ctx, rocprim::exclusive_scan(d_tmp, tmp, d_in, d_out, 0, (size_t)n,
                                               rocprim::plus<int>(), ctx->stream));
  int ti = (int)t;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(d_out + n, &ti, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *total = t;
  return TFM_OK;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__global__ void k_box_nodes(double *coords, int64_t nx, int64_t ny, int64_t nz, double mn0,
                            double mn1, double mn2, double mx0, double mx1, double mx2) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (nx + 1) * (ny + 1) * (nz + 1);
  if (id >= total) return;
  int64_t k = id % (nz + 1);
  int64_t j = (id / (nz + 1)) % (ny + 1);
  int64_t i = id / ((nz + 1) * (ny + 1));
  double t = nx > 0 ? (double)i / (double)nx : 0.0;
  double u = ny > 0 ? (double)j / (double)ny : 0.0;
  double w = nz > 0 ? (double)k / (double)nz : 0.0;
  coords[3 * id + 0] = mn0 + t * (mx0 - mn0);
  coords[3 * id + 1] = mn1 + u * (mx1 - mn1);
  coords[3 * id + 2] = mn2 + w * (mx2 - mn2);
}

// Six tets per brick sharing the 0-6 diagonal; tv[1]/tv[2] swapped where the
// brick-corner orientation (evaluated on the brick-local 0/1 corner offsets,
// as the reference's corner() closure does) is negative.
__global__ void k_box_tets(int *conn, int64_t nx, int64_t ny, int64_t nz, const int *swap6) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nx * ny * nz) return;
  int64_t k = b % nz, j = (b / nz) % ny, i = b / (nz * ny);
  const int OFF[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                         {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const int TETS[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                          {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
  int ids[8];
  for (int c = 0; c < 8; c++)
    ids[c] = (int)(((i + OFF[c][0]) * (ny + 1) + (j + OFF[c][1])) * (nz + 1) + (k + OFF[c][2]));
  for (int t = 0; t < 6; t++) {
    int v0 = ids[TETS[t][0]], v1 = ids[TETS[t][1]], v2 = ids[TETS[t][2]], v3 = ids[TETS[t][3]];
    if (swap6[t]) { int tmp = v1; v1 = v2; v2 = tmp; }
    int *o = conn + 4 * (6 * b + t);
    o[0] = v0; o[1] = v1; o[2] = v2; o[3] = v3;
  }
}

__global__ void k_grid(double *coords, int *conn, int64_t nx, int64_t ny, double h) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id < nx * ny) {
    int64_t i = id % nx, j = id / nx;
    coords[2 * id] = (double)i * h;
    coords[2 * id + 1] = (double)j * h;
  }
  if (id < (nx - 1) * (ny - 1)) {
    int64_t ei = id % (nx - 1), ej = id / (nx - 1);
    conn[4 * id + 0] = (int)(ej * nx + ei);
    conn[4 * id + 1] = (int)(ej * nx + ei + 1);
    conn[4 * id + 2] = (int)((ej + 1) * nx + ei + 1);
    conn[4 * id + 3] = (int)((ej + 1) * nx + ei);
  }
}

__device__ __host__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_jitter(double *coords, int64_t nx, int64_t ny, int64_t nz, double frac,
                         uint64_t seed) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (nx + 1) * (ny + 1) * (nz + 1)) return;
  int64_t k = id % (nz + 1), j = (id / (nz + 1)) % (ny + 1), i = id / ((nz + 1) * (ny + 1));
  if (i == 0 || i == nx || j == 0 || j == ny || k == 0 || k == nz) return;
  double h[3] = {1.0 / (double)nx, 1.0 / (double)ny, 1.0 / (double)nz};
  for (int a = 0; a < 3; a++) {
    uint64_t ctr = (uint64_t)(3 * id + a);
    uint64_t s = seed ^ (ctr * 0xD1B54A32D192ED03ull);
    double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
    coords[3 * id + a] += (2.0 * u - 1.0) * (frac * h[a]);
  }
}

__global__ void k_adj_count(const int *conn, int64_t n_elem, int nn, int *cnt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_elem * nn) return;
  atomicAdd(&cnt[conn[t]], 1);
}

__global__ void k_adj_fill(const int *conn, int64_t n_elem, int nn, const int *ptr, int *cursor,
                           int *n2e) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_elem * nn) return;
  int node = conn[t];
  int pos = atomicAdd(&cursor[node], 1);
  int e = (int)(t / nn), l = (int)(t % nn);
  n2e[ptr[node] + pos] = (e << 3) | l;
}

// Sort each node's list by element id (insertion sort; lists are short).
// Atomics make the fill order arbitrary; the sort makes it deterministic.
__global__ __launch_bounds__(256) void k_adj_sort(const int *ptr, int *n2e, int64_t n_nodes, int *maxadj) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int len = 0;
  if (n < n_nodes) {
    int b = ptr[n], e = ptr[n + 1];
    for (int i = b + 1; i < e; i++) {
      int v = n2e[i];
      int j = i - 1;
      while (j >= b && n2e[j] > v) {
        n2e[j + 1] = n2e[j];
        j--;
      }
      n2e[j + 1] = v;
    }
    len = e - b;
  }
  block_max(len, maxadj);
}

__global__ void k_nodes_on_plane(const double *coords, int64_t n, int pdim, int axis, double coord,
                                 double tol, unsigned char *flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double c = axis < pdim ? coords[i * pdim + axis] : 0.0;
  flag[i] = fabs(c - coord) <= tol ? 1 : 0;
}

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int ensure_adjacency(tfm_mesh *m) {
  if (m->d_n2e_ptr) return TFM_OK;
  tfm_ctx *ctx = m->ctx;
  int *cnt = nullptr, *cursor = nullptr, *dmax = nullptr;
  int rc;
  int64_t tot = m->n_elem * m->nn;
  if ((rc = dalloc(ctx, &cnt, m->n_nodes + 1))) return rc;
  if ((rc = dalloc(ctx, &m->d_n2e_ptr, m->n_nodes + 1))) return rc;
  if ((rc = dalloc(ctx, &m->d_n2e, tot))) return rc;
  if ((rc = dalloc(ctx, &cursor, m->n_nodes))) return rc;
  if ((rc = dalloc(ctx, &dmax, 1))) return rc;
  TFM_CHECK_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int) * (m->n_nodes + 1), ctx->stream));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(cursor, 0, sizeof(int) * m->n_nodes, ctx->stream));
  TFM_CHECK_HIP(ctx, hipMemsetAsync(dmax, 0, sizeof(int), ctx->stream));
  if (tot > 0) {
    hipLaunchKernelGGL(k_adj_count, dim3(grid1(tot, 256)), dim3(256), 0, ctx->stream, m->d_conn,
                       m->n_elem, m->nn, cnt);
    TFM_CHECK_HIP(ctx, hipGetLastError());
  }
  int64_t total = 0;
  if ((rc = exclusive_scan_i32(ctx, cnt, m->d_n2e_ptr, m->n_nodes, &total))) return rc;
  if (tot > 0) {
    hipLaunchKernelGGL(k_adj_fill, dim3(grid1(tot, 256)), dim3(256), 0, ctx->stream, m->d_conn,
                       m->n_elem, m->nn, m->d_n2e_ptr, cursor, m->d_n2e);
    TFM_CHECK_HIP(ctx, hipGetLastError());
  }
  if (m->n_nodes > 0) {
    hipLaunchKernelGGL(k_adj_sort, dim3(grid1(m->n_nodes, 256)), dim3(256), 0, ctx->stream,
                       m->d_n2e_ptr, m->d_n2e, m->n_nodes, dmax);
    TFM_CHECK_HIP(ctx, hipGetLastError());
  }
  int mx = 0;
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(&mx, dmax, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  m->h_n2e_ptr.resize(m->n_nodes + 1);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(m->h_n2e_ptr.data(), m->d_n2e_ptr, sizeof(int) * (m->n_nodes + 1),
                                    hipMemcpyDeviceToHost, ctx->stream));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  m->max_adj = mx;
  dfree(cnt);
  dfree(cursor);
  dfree(dmax);
  return TFM_OK;
}

}  // namespace tfm

using namespace tfm;

// ---------------------------------------------------------------------------
// C ABI: context
// ---------------------------------------------------------------------------
extern "C" int tfm_ctx_create(int device, tfm_ctx **out) {
  if (!out) return TFM_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return TFM_E_HIP;
  if (device < 0 || device >= ndev) return TFM_E_INVALID;
  tfm_ctx *c = new tfm_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return TFM_E_HIP;
  }
  // the allocation-free solve's work space on this device (looked up once)
  if (noalloc_init(c) != TFM_OK) {
    tfm_ctx_destroy(c);
    return TFM_E_HIP;
  }
  *out = c;
  return TFM_OK;
}

extern "C" void tfm_ctx_destroy(tfm_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_scratch) (void)hipFree(ctx->d_scratch);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char *tfm_last_error(const tfm_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

extern "C" int tfm_ctx_synchronize(tfm_ctx *ctx) {
  if (!ctx) return TFM_E_INVALID;
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return TFM_OK;
}

extern "C" int tfm_ctx_timer_begin(tfm_ctx *ctx) {
  if (!ctx) return TFM_E_INVALID;
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  return TFM_OK;
}

extern "C" int tfm_ctx_timer_end(tfm_ctx *ctx, double *ms) {
  if (!ctx || !ms) return TFM_E_INVALID;
  TFM_CHECK_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  TFM_CHECK_HIP(ctx, hipEventSynchronize(ctx->ev1));
  float f = 0.f;
  TFM_CHECK_HIP(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  *ms = f;
  return TFM_OK;
}

extern "C" int tfm_debug_scan_total(tfm_ctx *ctx, const int32_t *counts, int64_t n, int64_t *total) {
  if (!ctx || !total || n < 0 || (n > 0 && !counts)) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  int *d_in = nullptr, *d_out = nullptr;
  if ((rc = dalloc(ctx, &d_in, (size_t)n)) || (rc = dalloc(ctx, &d_out, (size_t)n + 1))) {
    dfree(d_in);
    return rc;
  }
  hipError_t e = n > 0 ? hipMemcpy(d_in, counts, sizeof(int) * (size_t)n, hipMemcpyHostToDevice) : hipSuccess;
  if (e != hipSuccess) rc = fail(ctx, TFM_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  if (!rc) rc = exclusive_scan_i32(ctx, d_in, d_out, n, total);
  dfree(d_in);
  dfree(d_out);
  return rc;
}

// ---------------------------------------------------------------------------
// C ABI: meshes
// ---------------------------------------------------------------------------
static int mesh_alloc(tfm_ctx *ctx, int cell, int64_t n_nodes, int pdim, int64_t n_elem,
                      tfm_mesh **out) {
  int nn = cell_nodes(cell);
  if (nn == 0) return fail(ctx, TFM_E_INVALID, "unknown cell type");
  if (pdim < cell_dim(cell) || pdim > 3) return fail(ctx, TFM_E_INVALID, "coordinate count below the cell dimension");
  if (n_nodes < 0 || n_elem < 0 || n_nodes >= INT32_MAX || n_elem * 8 >= INT32_MAX)
    return fail(ctx, TFM_E_UNSUPPORTED, "mesh too large for int32 indexing");
  tfm_mesh *m = new tfm_mesh();
  m->ctx = ctx;
  m->cell = cell;
  m->nn = nn;
  m->pdim = pdim;
  m->dim = cell_dim(cell);
  m->n_nodes = n_nodes;
  m->n_elem = n_elem;
  int rc;
  if ((rc = dalloc(ctx, &m->d_coords, n_nodes * pdim)) || (rc = dalloc(ctx, &m->d_conn, n_elem * nn))) {
    tfm_mesh_destroy(m);
    return rc;
  }
  *out = m;
  return TFM_OK;
}

extern "C" int tfm_mesh_create(tfm_ctx *ctx, int cell, int64_t n_nodes, int pdim, const double *coords,
                               int64_t n_elem, const int64_t *conn, tfm_mesh **out) {
  if (!ctx || !out || (n_nodes && !coords) || (n_elem && !conn)) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_mesh *m = nullptr;
  if ((rc = mesh_alloc(ctx, cell, n_nodes, pdim, n_elem, &m))) return rc;
  std::vector<int> c32((size_t)(n_elem * m->nn));
  for (size_t i = 0; i < c32.size(); i++) {
    if (conn[i] < 0 || conn[i] >= n_nodes) {
      tfm_mesh_destroy(m);
      return fail(ctx, TFM_E_INVALID, "element references a node index out of range");
    }
    c32[i] = (int)conn[i];
  }
  hipError_t e1 = hipMemcpy(m->d_coords, coords, sizeof(double) * n_nodes * pdim, hipMemcpyHostToDevice);
  hipError_t e2 = hipMemcpy(m->d_conn, c32.data(), sizeof(int) * c32.size(), hipMemcpyHostToDevice);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    tfm_mesh_destroy(m);
    return fail(ctx, TFM_E_HIP, "mesh upload failed");
  }
  *out = m;
  return TFM_OK;
}

static double host_orient3(const double *a, const double *b, const double *c, const double *d) {
  double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double ad[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  return ab[0] * (ac[1] * ad[2] - ac[2] * ad[1]) - ab[1] * (ac[0] * ad[2] - ac[2] * ad[0]) +
         ab[2] * (ac[0] * ad[1] - ac[1] * ad[0]);
}

extern "C" int tfm_mesh_box(tfm_ctx *ctx, const double mn[3], const double mx[3], int64_t nx, int64_t ny,
                            int64_t nz, tfm_mesh **out) {
  if (!ctx || !out || !mn || !mx || nx < 0 || ny < 0 || nz < 0) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  int64_t nnod = (nx + 1) * (ny + 1) * (nz + 1), nel = 6 * nx * ny * nz;
  tfm_mesh *m = nullptr;
  if ((rc = mesh_alloc(ctx, TFM_TET, nnod, 3, nel, &m))) return rc;
  // The orientation test depends only on the brick-local corner offsets
  // (the reference evaluates corner(0|1, 0|1, 0|1)), so it is the same for
  // every brick: evaluate it once on the host.
  int swap6[6];
  {
    const int OFF[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const int TETS[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                            {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
    double co[8][3];
    for (int c = 0; c < 8; c++) {
      double t = nx > 0 ? (double)OFF[c][0] / (double)nx : 0.0;
      double u = ny > 0 ? (double)OFF[c][1] / (double)ny : 0.0;
      double w = nz > 0 ? (double)OFF[c][2] / (double)nz : 0.0;
      co[c][0] = mn[0] + t * (mx[0] - mn[0]);
      co[c][1] = mn[1] + u * (mx[1] - mn[1]);
      co[c][2] = mn[2] + w * (mx[2] - mn[2]);
    }
    for (int t = 0; t < 6; t++)
      swap6[t] = host_orient3(co[TETS[t][0]], co[TETS[t][1]], co[TETS[t][2]], co[TETS[t][3]]) < 0.0;
  }
  int *d_swap = nullptr;
  if ((rc = dalloc(ctx, &d_swap, 6))) { tfm_mesh_destroy(m); return rc; }
  (void)hipMemcpyAsync(d_swap, swap6, sizeof swap6, hipMemcpyHostToDevice, ctx->stream);
  if (nnod > 0)
    hipLaunchKernelGGL(k_box_nodes, dim3(grid1(nnod, 256)), dim3(256), 0, ctx->stream, m->d_coords, nx,
                       ny, nz, mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]);
  if (nel > 0)
    hipLaunchKernelGGL(k_box_tets, dim3(grid1(nx * ny * nz, 256)), dim3(256), 0, ctx->stream, m->d_conn,
                       nx, ny, nz, d_swap);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  dfree(d_swap);
  if (e != hipSuccess || hipGetLastError() != hipSuccess) {
    tfm_mesh_destroy(m);
    return fail(ctx, TFM_E_HIP, "box_mesh kernel failed");
  }
  m->box_n[0] = nx;
  m->box_n[1] = ny;
  m->box_n[2] = nz;
  *out = m;
  return TFM_OK;
}

extern "C" int tfm_mesh_grid(tfm_ctx *ctx, int64_t nx, int64_t ny, double h, tfm_mesh **out) {
  if (!ctx || !out || nx < 2 || ny < 2) return TFM_E_INVALID;
  int rc = set_device(ctx);
  if (rc) return rc;
  tfm_mesh *m = nullptr;
  if ((rc = mesh_alloc(ctx, TFM_QUAD, nx * ny, 2, (nx - 1) * (ny - 1), &m))) return rc;
  hipLaunchKernelGGL(k_grid, dim3(grid1(nx * ny, 256)), dim3(256), 0, ctx->stream, m->d_coords, m->d_conn,
                     nx, ny, h);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess || hipGetLastError() != hipSuccess) {
    tfm_mesh_destroy(m);
    return fail(ctx, TFM_E_HIP, "grid kernel failed");
  }
  *out = m;
  return TFM_OK;
}

extern "C" int tfm_mesh_jitter(tfm_mesh *m, int64_t nx, int64_t ny, int64_t nz, double frac, uint64_t seed) {
  if (!m) return TFM_E_INVALID;
  tfm_ctx *ctx = m->ctx;
  if (m->pdim != 3 || (nx + 1) * (ny + 1) * (nz + 1) != m->n_nodes || nx < 1 || ny < 1 || nz < 1)
    return fail(ctx, TFM_E_INVALID, "jitter needs the (nx,ny,nz) box_mesh grid the mesh was built from");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_jitter, dim3(grid1(m->n_nodes, 256)), dim3(256), 0, ctx->stream, m->d_coords, nx, ny,
                     nz, frac, seed);
  TFM_CHECK_HIP(ctx, hipGetLastError());
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return TFM_OK;
}

extern "C" int tfm_mesh_info(const tfm_mesh *m, int *cell, int64_t *n_nodes, int *pdim, int64_t *n_elem) {
  if (!m) return TFM_E_INVALID;
  if (cell) *cell = m->cell;
  if (n_nodes) *n_nodes = m->n_nodes;
  if (pdim) *pdim = m->pdim;
  if (n_elem) *n_elem = m->n_elem;
  return TFM_OK;
}

extern "C" int tfm_mesh_download(const tfm_mesh *m, double *coords, int64_t *conn) {
  if (!m) return TFM_E_INVALID;
  tfm_ctx *ctx = m->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  if (coords)
    TFM_CHECK_HIP(ctx, hipMemcpy(coords, m->d_coords, sizeof(double) * m->n_nodes * m->pdim, hipMemcpyDeviceToHost));
  if (conn) {
    std::vector<int> c32((size_t)(m->n_elem * m->nn));
    TFM_CHECK_HIP(ctx, hipMemcpy(c32.data(), m->d_conn, sizeof(int) * c32.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < c32.size(); i++) conn[i] = c32[i];
  }
  return TFM_OK;
}

extern "C" int tfm_mesh_nodes_on_plane(const tfm_mesh *m, int axis, double coord, double tol, int64_t *out,
                                       int64_t *count) {
  if (!m || !out || !count || axis < 0) return TFM_E_INVALID;
  tfm_ctx *ctx = m->ctx;
  int rc = set_device(ctx);
  if (rc) return rc;
  unsigned char *flag = nullptr;
  if ((rc = dalloc(ctx, &flag, m->n_nodes))) return rc;
  if (m->n_nodes > 0)
    hipLaunchKernelGGL(k_nodes_on_plane, dim3(grid1(m->n_nodes, 256)), dim3(256), 0, ctx->stream,
                       m->d_coords, m->n_nodes, m->pdim, axis, coord, tol, flag);
  std::vector<unsigned char> h((size_t)m->n_nodes);
  hipError_t e = hipMemcpyAsync(h.data(), flag, m->n_nodes, hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  dfree(flag);
  if (e != hipSuccess || e2 != hipSuccess) return fail(ctx, TFM_E_HIP, "nodes_on_plane failed");
  int64_t c = 0;
  for (int64_t i = 0; i < m->n_nodes; i++)
    if (h[i]) out[c++] = i;
  *count = c;
  return TFM_OK;
}

namespace tfm {
void mesh_retain(const tfm_mesh *m) { m->refs.fetch_add(1, std::memory_order_relaxed); }
void mesh_release(const tfm_mesh *cm) {
  if (cm->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  tfm_mesh *m = const_cast<tfm_mesh *>(cm);
  dfree(m->d_coords);
  dfree(m->d_conn);
  dfree(m->d_n2e_ptr);
  dfree(m->d_n2e);
  delete m;
}
}  // namespace tfm

// The caller's handle; the device memory goes with the last holder (systems
// built on the mesh hold it too).
extern "C" void tfm_mesh_destroy(tfm_mesh *m) {
  if (m) tfm::mesh_release(m);
}
