// tfm_vtk.hip -- legacy binary VTK result files from device-resident meshes.
//
// write_vtk / write_vtk_with_data (crates/tpt-fem-io-vtk/src/lib.rs:117-183):
// mesh_to_vtk's unstructured grid exported as legacy BINARY, which is big
// endian.  The byte-swapped blocks -- points padded to three f64 components,
// cells as u32 (vertex count, then node ids), cell types as i32 and one f64
// block per scalar field -- are produced on the GPU next to the mesh; the
// host only writes the header lines between them.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tfm_internal.h"

namespace tfm {

__device__ __forceinline__ unsigned long long bswap64(unsigned long long x) {
  return __builtin_bswap64(x);
}
__device__ __forceinline__ unsigned bswap32(unsigned x) { return __builtin_bswap32(x); }

// points: n x 3 big-endian doubles, coordinates past pdim are 0.0
__global__ void kv_points(int64_t n, int pdim, const double *__restrict__ coords,
                          unsigned long long *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n) return;
  const int64_t node = i / 3;
  const int d = (int)(i % 3);
  const double x = d < pdim ? coords[node * pdim + d] : 0.0;
  out[i] = bswap64((unsigned long long)__double_as_longlong(x));
}

// cells: per element nn, then its node ids, big-endian u32
__global__ void kv_cells(int64_t ne, int nn, const int *__restrict__ conn, unsigned *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne * (nn + 1)) return;
  const int64_t e = i / (nn + 1);
  const int k = (int)(i % (nn + 1));
  const unsigned v = k == 0 ? (unsigned)nn : (unsigned)conn[e * nn + k - 1];
  out[i] = bswap32(v);
}

__global__ void kv_fill32(int64_t n, unsigned v, unsigned *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bswap32(v);
}

__global__ void kv_swap64(int64_t n, const double *__restrict__ in, unsigned long long *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = bswap64((unsigned long long)__double_as_longlong(in[i]));
}

}  // namespace tfm

using namespace tfm;

extern "C" int tfm_write_vtk(const tfm_mesh *mesh, int n_fields, const char *const *names,
                             const double *const *values, const char *path) {
  if (!mesh || !path || n_fields < 0 || (n_fields > 0 && (!names || !values))) return TFM_E_INVALID;
  tfm_ctx *ctx = mesh->ctx;
  // vtk_cell_type (io-vtk/src/lib.rs:101-114) for the P1 cells
  static const unsigned kType[5] = {3, 5, 9, 10, 12};
  if (mesh->cell < 0 || mesh->cell > 4) return fail(ctx, TFM_E_UNSUPPORTED, "write_vtk: unknown cell type");
  for (int f = 0; f < n_fields; f++)
    if (!names[f] || !values[f] || std::strpbrk(names[f], " \n"))
      return fail(ctx, TFM_E_INVALID, "write_vtk: a field needs a name without blanks and its values");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const int64_t n = mesh->n_nodes, ne = mesh->n_elem, nn = mesh->nn, nv = ne * (nn + 1);
  unsigned long long *d_pts = nullptr, *d_fld = nullptr;
  unsigned *d_cells = nullptr, *d_types = nullptr;
  double *d_in = nullptr;
  auto cleanup = [&]() { dfree(d_pts); dfree(d_fld); dfree(d_cells); dfree(d_types); dfree(d_in); };
  if ((rc = dalloc(ctx, &d_pts, 3 * n)) || (rc = dalloc(ctx, &d_cells, nv)) || (rc = dalloc(ctx, &d_types, ne)) ||
      (n_fields > 0 && ((rc = dalloc(ctx, &d_fld, n)) || (rc = dalloc(ctx, &d_in, n))))) {
    cleanup();
    return rc;
  }
  auto grid = [](int64_t k) { return dim3((unsigned)((k + 255) / 256)); };
  if (n > 0) hipLaunchKernelGGL(kv_points, grid(3 * n), dim3(256), 0, st, n, mesh->pdim, mesh->d_coords, d_pts);
  if (nv > 0) hipLaunchKernelGGL(kv_cells, grid(nv), dim3(256), 0, st, ne, (int)nn, mesh->d_conn, d_cells);
  if (ne > 0) hipLaunchKernelGGL(kv_fill32, grid(ne), dim3(256), 0, st, ne, kType[mesh->cell], d_types);
  std::vector<unsigned char> pts(24 * n), cells(4 * nv), types(4 * ne), fld(8 * n);
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(pts.data(), d_pts, pts.size(), hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(cells.data(), d_cells, cells.size(), hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipMemcpyAsync(types.data(), d_types, types.size(), hipMemcpyDeviceToHost, st));
  TFM_CHECK_HIP(ctx, hipStreamSynchronize(st));
  FILE *fp = std::fopen(path, "wb");
  if (!fp) {
    cleanup();
    return fail(ctx, TFM_E_INVALID, std::string("I/O error on VTK file: cannot open ") + path);
  }
  bool ok = true;
  auto put = [&](const void *p, size_t k) { ok = ok && (k == 0 || std::fwrite(p, 1, k, fp) == k); };
  auto text = [&](const std::string &s) { put(s.data(), s.size()); };
  text("# vtk DataFile Version 4.1\ntpt-fem mesh\nBINARY\n\nDATASET UNSTRUCTURED_GRID\nPOINTS " + std::to_string(n) +
       " double\n");
  put(pts.data(), pts.size());
  text("\n\nCELLS " + std::to_string(ne) + " " + std::to_string(nv) + "\n");
  put(cells.data(), cells.size());
  text("\n\nCELL_TYPES " + std::to_string(ne) + "\n");
  put(types.data(), types.size());
  text("\n\nPOINT_DATA " + std::to_string(n) + "\n");
  for (int f = 0; f < n_fields && ok; f++) {
    if (n > 0) {
      if (hipMemcpyAsync(d_in, values[f], sizeof(double) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        ok = false;
        break;
      }
      hipLaunchKernelGGL(kv_swap64, grid(n), dim3(256), 0, st, n, (const double *)d_in, d_fld);
      if (hipMemcpyAsync(fld.data(), d_fld, fld.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        ok = false;
        break;
      }
    }
    text(std::string("\nSCALARS ") + names[f] + " double 1\nLOOKUP_TABLE default\n");
    put(fld.data(), fld.size());
    text("\n");
  }
  text("\nCELL_DATA " + std::to_string(ne) + "\n\n");
  ok = (std::fclose(fp) == 0) && ok;
  cleanup();
  if (!ok) return fail(ctx, TFM_E_INVALID, std::string("I/O error on VTK file: write failed for ") + path);
  return TFM_OK;
}
