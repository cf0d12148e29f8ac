// tfm_internal.h -- internal types of the MI355X assembly/solve library.
//
// Device data layout (HBM):
//   coords   double [n_nodes * pdim]   node-major, components interleaved
//   conn     int32  [n_elem * nn]      element-major
//   n2e_ptr  int32  [n_nodes + 1]      node -> element adjacency (CSR)
//   n2e      int32  [n_elem * nn]      (eid << 3) | local index, eid ascending
//   CSR      int32 row_ptr [nrows+1], int32 col_ind [nnz], double/float values
// Reduced systems keep global DOF numbering on the device: fixed rows are
// empty and fixed columns absent, which is entry-for-entry the reference's
// kred (assembly/src/lib.rs:504-535) with the free map applied; the reduced
// numbering is produced only when a caller downloads it.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/tptfem.h"

#define TFM_CHECK_HIP(ctx, call)                                               \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      return tfm::fail((ctx), TFM_E_HIP, std::string(#call) + ": " +            \
                                             hipGetErrorString(e_));           \
    }                                                                          \
  } while (0)

struct tfm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // small device scratch for reductions / scalars (grow-only)
  void *d_scratch = nullptr;
  size_t scratch_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct tfm_mesh {
  tfm_ctx *ctx = nullptr;
  int cell = 0, nn = 0, pdim = 0, dim = 0;
  int64_t n_nodes = 0, n_elem = 0;
  double *d_coords = nullptr;
  int *d_conn = nullptr;
  // node -> element adjacency, built on first use
  int *d_n2e_ptr = nullptr;
  int *d_n2e = nullptr;
  int max_adj = 0;
  std::vector<int> h_n2e_ptr;  // kept for tile scheduling
  // tfm_mesh_grid_batch: P copies of an nx x ny node grid (0 otherwise)
  int64_t grid_P = 0, grid_nx = 0, grid_ny = 0;
  // tfm_mesh_box: cells per axis (0 otherwise; jitter keeps the node grid)
  int64_t box_n[3] = {0, 0, 0};
  // holders: the caller's handle and every assembly plan built on the mesh
  // (systems keep theirs); the last release frees it
  mutable std::atomic<int> refs{1};
};

struct tfm_matrix {
  tfm_ctx *ctx = nullptr;
  int64_t nrows = 0, ncols = 0, nnz = 0;
  int *d_rp = nullptr;
  int *d_ci = nullptr;
  double *d_v = nullptr;
};

namespace tfm {

// Maximum of v over the block (a whole number of wavefronts, at most 1024
// threads: wave shuffles, then LDS) into *out with one atomic per block:
// same-word atomics from every wave of a large grid serialise at L2.  Every
// thread of the block must call it.
__device__ __forceinline__ void block_max(int v, int *out) {
  __shared__ int s_max[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = s_max[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = max(m, s_max[w]);
    atomicMax(out, m);
  }
}

int fail(tfm_ctx *ctx, int code, const std::string &msg);
int set_device(tfm_ctx *ctx);
int ensure_adjacency(tfm_mesh *m);
void mesh_retain(const tfm_mesh *m);
void mesh_release(const tfm_mesh *m);
int scratch(tfm_ctx *ctx, size_t bytes, void **out);

template <typename T>
int dalloc(tfm_ctx *ctx, T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(ctx, TFM_E_HIP, std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                    " bytes): " + hipGetErrorString(e));
  }
  return TFM_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

inline int cell_nodes(int cell) {
  static const int nn[5] = {2, 3, 4, 4, 8};
  return (cell >= 0 && cell < 5) ? nn[cell] : 0;
}
inline int cell_dim(int cell) {
  static const int dd[5] = {1, 2, 2, 3, 3};
  return (cell >= 0 && cell < 5) ? dd[cell] : 0;
}

// Exclusive scan of n int32 counts into out[0..n] (out[n] = total) on the
// context stream; returns the total through *total (synchronises).
int exclusive_scan_i32(tfm_ctx *ctx, const int *d_in, int *d_out, int64_t n, int64_t *total);

// P1-tet assembly plan (tfm_asm_tet.hip), built once per pattern: the tile
// schedule and one item record per row block (n, m) of each tile's nodes
// (20 bytes: 3 x u32 contribution words, u32 row positions, u32 pmk), stored
// in the order the numeric pass takes them.
// How the tet tile size was chosen: source 1 = timed here, 2 = reused from
// an earlier system of the same shape class, 3 = forced by the environment;
// ms = measured pass time of each candidate (256, 192 pairs) when timed.
struct TetTune {
  int tile_pairs = 0, source = 0;
  float ms[2] = {0.f, 0.f};
};

struct TetPlan {
  int *d_tiles = nullptr;  // first node of each tile (ntiles + 1)
  int *d_tpa = nullptr;    // first (node, element) pair of each tile
  int *d_tib = nullptr;    // first item of each tile
  int *d_tnd = nullptr;    // diagonal blocks of each tile
  int *d_tw = nullptr;     // longest row (entries) of each tile's nodes: the staging stride
  int ntiles = 0;
  int tp = 0;              // pairs (and threads) per tile: 192 or 256
  // One item per row block (n, m) of a tile's nodes, in tile order: the
  // tile's diagonal blocks first, then the off-diagonal ones in node and
  // column order.  Contribution words [3][nitems]: a diagonal block's first
  // pair and pair count; an off-diagonal block's LDS element offsets of its
  // (<= 6) contributions in pair order, 16 bits each; or, for a block with
  // more contributions, the pair mask and the 2-bit LDS slots of its pairs.
  unsigned *d_off = nullptr;
  unsigned *d_pos = nullptr;  // per row a, bits 8a..8a+7: the block's first present entry relative to
                              // the row's start; bits 24..29: the node's index in its tile
  unsigned *d_pmk = nullptr;  // presence masks (4 bits per row) | count << 12 | first pair of n << 16
  int64_t nitems = 0;
};

// Assembly engine (tfm_assemble.hip).
struct AsmPlan {
  const tfm_mesh *mesh = nullptr;
  tfm_operator op{};
  int dpn = 1;
  int64_t ndof = 0;
  // device copies of per-element coefficient arrays (owned)
  double *d_coef_e = nullptr;
  double *d_rho = nullptr;
  double ke0[64] = {0};
  // quadrature rule table then SIMP unit stiffness, one device buffer (owned)
  double *d_tab = nullptr;
  int nq = 0;
  // fixed-DOF flags (owned), nullptr when no DOF is fixed
  unsigned char *d_fixed = nullptr;
  // tile schedule (owned)
  int *d_tiles = nullptr;
  int ntiles = 0;
  int tile_pairs = 0;
  // CSR pattern in global DOF numbering (owned)
  int *d_rp = nullptr;
  int *d_ci = nullptr;
  int64_t nrows = 0, ncols = 0, nnz = 0;
  // P1-tet fast path (tfm_asm_tet.hip), when its plan could be built, and
  // how its tile size was chosen
  bool use_tet = false;
  TetPlan tet;
  TetTune tune;
  // Elements whose coefficient was exactly 0 when the pattern was built (all
  // their entries were skipped, so they own no pattern entries); zero_all:
  // the scalar coefficient was 0.  A re-assembly may not make them non-zero.
  std::vector<int64_t> zero_elems;
  bool zero_all = false;
  // FP32 Hex8 elasticity element matrices ([n_elem][24][24], owned), made by
  // the first FP32 numeric pass that forms them on the matrix cores
  mutable float *d_ke32 = nullptr;
  // sorted-triplet assembly (tfm_asm_coo.hip), for node valences beyond the
  // tile kernels: per CSR entry its first triplet (d_coo_seg, nnz + 1), the
  // triplets' element-major positions in key order (d_coo_perm), and the
  // dense element-value tables of the numeric passes (made on first use)
  // SIMP grids: rows formed directly (tfm_assemble.hip), E(rho_e) per element
  bool use_simp_grid = false;
  mutable double *d_erho = nullptr;
  bool use_coo = false;
  int64_t coo_ntrip = 0;
  int64_t *d_coo_perm = nullptr;
  int *d_coo_perm32 = nullptr;  // the same positions in 32 bits, when the table has < 2^31 entries
  int *d_coo_seg = nullptr;
  mutable double *d_coo_v64 = nullptr;
  mutable float *d_coo_v32 = nullptr;
};
struct OpParams;
int tet_plan_create(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, const int *d_rp, const int *d_ci, int tile_pairs,
                    TetPlan *tp);
// Builds the tet plan at the tile size the autotuner picks for this shape
// class, timing each candidate into scratch value arrays of nnz entries for
// the arrays `vals` (TFM_SYS_FP64 | TFM_SYS_FP32) names.
int tet_plan_tuned(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, const int *d_rp, const int *d_ci, int64_t nnz,
                   const OpParams &P, int vals, TetPlan *tp, TetTune *info);
void tet_plan_destroy(TetPlan *tp);
int tet_numeric(tfm_ctx *ctx, const tfm_mesh *mesh, int dpn, const TetPlan &tp, const int *d_rp, const OpParams &P,
                double *v64, float *v32);
// `vals` names the value arrays later numeric passes write (TFM_SYS_FP64 |
// TFM_SYS_FP32): what the tile-size autotuner times.
int plan_create(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn,
                const unsigned char *h_fixed, int vals, AsmPlan *plan);
int plan_set_coefficients(tfm_ctx *ctx, AsmPlan *plan, const tfm_operator *op);
int plan_tables(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn, AsmPlan *plan, OpParams *P);
int plan_numeric(tfm_ctx *ctx, const AsmPlan *plan, double *d_v64, float *d_v32);
int coo_plan_create(tfm_ctx *ctx, AsmPlan *plan, const OpParams &P);
int coo_numeric(tfm_ctx *ctx, const AsmPlan *plan, const OpParams &P, double *v64, float *v32);
void coo_plan_destroy(AsmPlan *plan);
// 3-D elasticity on Hex8 cells with Gauss-Legendre order 1-3 is assembled by
// sorted triplets from per-element tables (FP64 exact, FP32 on the matrix
// cores); TFM_HEX_TILE in the environment keeps the tile kernels.
bool hex_table_route(const tfm_mesh *m, int dpn, const tfm_operator *op);
// FP32 element matrices of 3-D isotropic elasticity on Hex8 cells by MFMA
// (tfm_hex_mfma.hip); TFM_NO_MFMA in the environment switches them off.
bool hex_mfma_eligible(const AsmPlan *plan);
int hex_ke_mfma(tfm_ctx *ctx, const AsmPlan *plan, float *ke);
void plan_destroy(AsmPlan *plan);
int validate_operator(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn);

// Solver copy of a system matrix in node-blocked sliced ELL (tfm_solve.hip):
// block row b holds the bs DOF rows of node b; 64 consecutive block rows form
// a slice stored lane-interleaved, so the SpMV runs one lane per block row
// with coalesced loads and one column index per bs x bs block.
//   soff  int32 [nslices + 1]   first block of each slice (multiple of 64)
//   bcol  int32 [nblk]          block column of (slice, k, lane) at soff + 64k + lane;
//                               padding blocks point at column 0 with zero values
//   map   int32 [nnz]           CSR entry -> value index
//   value index of entry (i, j) of block (slice, k, lane):
//                               (soff + 64k) * bs^2 + (i * bs + j) * 64 + lane
struct BSell {
  int state = 0;  // 0 not built, 1 built, -1 not usable (index range)
  int bs = 0, nbr = 0, nslices = 0;
  int64_t nblk = 0, nnz = 0;
  int *d_soff = nullptr;
  int *d_bcol = nullptr;
  int *d_map = nullptr;  // slot -> CSR position (-1: a padding or absent entry)
  double *d_v64 = nullptr;
  float *d_v32 = nullptr;
};
void bsell_destroy(BSell *b);

// Batched solver (tfm_batch.hip) over a block-diagonal system of P problems
// (DOF ranges dof_off[p]..dof_off[p+1], homogeneous Dirichlet values), NR
// (1 or 8) right-hand sides each; the device solution stays in place between
// runs, so a run can warm-start from the previous one.
struct BatchSolver;
// mode: TFM_SOLVE_PCG64 (Jacobi), TFM_SOLVE_MIXED (FP32 + FP64 refinement) or
// TFM_SOLVE_PCG64_MG (multigrid V-cycle; SIMP systems on a tfm_mesh_grid_batch
// mesh with one problem per design).
int bsolver_create(tfm_system *s, int P, const int64_t *dof_off, int NR, int mode, BatchSolver **out);
int bsolver_set_rhs(BatchSolver *b, const double *rhs);
int bsolver_refresh(BatchSolver *b);
int bsolver_run(BatchSolver *b, const tfm_solve_opts &o, bool warm, tfm_solve_stats *stats, double *fu);
// One solve as a sequence of captured graphs (stage A: hooks->pre, e.g. the
// numeric assembly pass, and, with refresh, the solver's copy of the
// system's current values; B: initial residual and CG state; K: chunks of
// iterations until every segment is done; C: the FP64 residual certificate,
// then hooks->post).  The hooks only enqueue work on the context's stream,
// with arguments fixed for a given tag (graphs are re-captured when the tag
// or the stop rule changes).  bsolver_run is bsolver_solve without refresh
// or hooks.
struct SolveHooks {
  std::function<int()> pre, post;
  uint64_t tag = 0;
};
int bsolver_solve(BatchSolver *b, const tfm_solve_opts &o, bool warm, bool refresh, const SolveHooks *hooks,
                  tfm_solve_stats *stats, double *fu);
const double *bsolver_x(const BatchSolver *b);  // [n_dof][NR], fixed DOFs 0
double *bsolver_x_mut(BatchSolver *b);          // the same, to seed a warm start
int bsolver_set_rhs_dev(BatchSolver *b, const double *d_rhs);  // NR = 1, device [n_dof]
void bsolver_destroy(BatchSolver *b);
// tfm_system_solve in the multigrid mode: one problem, the system's own rhs;
// the solver (sliced copy, hierarchy) is kept on the system for later solves.
int system_solve_mg(tfm_system *s, const tfm_solve_opts &o, double *u, tfm_solve_stats *stats);

// Dense partial-pivot LU of the reduced system (tfm_dense.hip), the
// reference's direct solve, for at most kDirectMax free DOFs (x stays in one
// workgroup's LDS during back substitution): d_F [nrhs][ndof] right-hand
// sides (0 at fixed DOFs) -> d_x [nrhs][ndof] (0 at fixed DOFs).
constexpr int64_t kDirectMax = 8000;
int direct_solve_dev(tfm_system *s, int nrhs, const double *d_F, double *d_x);
// The same with the FP64 certificate and the Dirichlet scatter into host u
// [nrhs][ndof]; d_F NULL: the system's own rhs (nrhs = 1).
int system_solve_direct(tfm_system *s, int nrhs, const double *d_F, double *u, tfm_solve_stats *stats);

// Looks up this device's copy of the allocation-free solve's work space
// (tfm_solve.hip); tfm_ctx_create calls it with the context's device current.
int noalloc_init(tfm_ctx *ctx);

// Geometric multigrid for P batched topopt grids (tfm_mg.hip).
struct MgData;
int mg_levels(int nx, int ny, std::vector<int> *dims);
int mg_create(tfm_ctx *ctx, int P, int nx, int ny, int NR, const double *ke0, MgData **out);
int mg_update(MgData *mg, const double *d_rho, double penal, double e0, double emin, const unsigned char *d_fixed);
// z = M r ([n_dof][NR]); segment (p, k) is skipped when done[(p NR + k) stride] != 0
// on entry z holds mg_sweep_weights .* r (when those exist).  last false:
// the finest level's closing sweep z = y + w Dinv (r - K y) is left to the
// caller, who gets y back (null when the grid has one level: z is complete).
const double *mg_apply(const MgData *mg, const double *r, double *z, hipStream_t st, const int *done, int stride,
                       bool last = true);
const double *mg_sweep_weights(const MgData *mg);  // [n_dof] w Dinv of the finest level, or null
const double *mg_fine_dinv(const MgData *mg);      // [n_dof] finest Jacobi diagonal inverse (0: fixed)
struct MgOp;
MgOp mg_fine_op(const MgData *mg);

// Geometric multigrid for one system on a box mesh (tfm_mg3.hip), one rhs.
// The caller runs level 0 on its own matrix copy: z = mg3_sweep_weights .* r,
// mg3_fine_r = r - K z, y = mg3_cycle(z), then z = y + w Dinv (r - K y) with
// w = kMg3Omega and Dinv = mg3_fine_dinv; with depth 0, mg3_direct.
struct Mg3Data;
int mg3_levels(const int n[3], int dpn, std::vector<int> *dims);
int mg3_create(tfm_ctx *ctx, const int n[3], int dpn, Mg3Data **out);  // n: nodes per axis
int mg3_update(Mg3Data *mg, const int *rp, const int *ci, const double *v, const double *dinv0);
const double *mg3_cycle(const Mg3Data *mg, const double *z, hipStream_t st, const int *done);
void mg3_direct(const Mg3Data *mg, const double *r, double *z, hipStream_t st, const int *done);
int mg3_depth(const Mg3Data *mg);
double *mg3_fine_r(const Mg3Data *mg);
const double *mg3_fine_dinv(const Mg3Data *mg);
const double *mg3_sweep_weights(const Mg3Data *mg);
void mg3_destroy(Mg3Data *mg);
void mg_destroy(MgData *mg);

// Loads (tfm_loads.hip): device rhs of n_nodes*dpn doubles, accumulated.
int load_source_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, double f, double *d_rhs);
int load_body_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, const double *b, double *d_rhs);
// The same from the closure's values at every element quadrature point (host
// arrays [n_elem][nq] and [n_elem][nq][dim]).
int load_source_qp_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, const double *h_sq, double *d_rhs);
int load_body_qp_dev(tfm_ctx *ctx, tfm_mesh *mesh, int order, const double *h_bq, double *d_rhs);
int element_rule_points(tfm_ctx *ctx, int cell, int order, int *nq);
// Neumann pass from the flux closure's values at every boundary-face point
// (host array in tfm_mesh_boundary_points order), tiled over dpn components.
int load_neumann_qp_dev(tfm_ctx *ctx, tfm_mesh *mesh, int dpn, int order, const double *h_gq, double *d_rhs);
// to_csr(assemble ++ apply_robin): h constant (on an optional band) or, when
// h_hq is given, the coefficient closure's values at every boundary point.
int assemble_robin_dev(tfm_ctx *ctx, const tfm_mesh *mesh, const tfm_operator *op, int dpn, double h, int axis,
                       double coord, double tol, int order, const double *h_hq, tfm_matrix **out);
// zeroed: d_rhs holds +0.0 everywhere, so faces that can only add +0.0 may be skipped.
int load_neumann_dev(tfm_ctx *ctx, tfm_mesh *mesh, int dpn, int comp, double g, int axis,
                     double coord, double tol, double *d_rhs, bool zeroed);

}  // namespace tfm

struct tfm_system {
  tfm_ctx *ctx = nullptr;
  int dpn = 1;
  int64_t ndof = 0, n_free = 0;
  bool has_plan = false;
  tfm::AsmPlan plan;          // when built from a mesh
  // working matrix (global DOF numbering, fixed rows empty / cols absent)
  int *d_rp = nullptr;        // aliases plan.d_rp when has_plan
  int *d_ci = nullptr;
  int64_t nrows = 0, nnz = 0;
  double *d_v64 = nullptr;
  float *d_v32 = nullptr;
  double *d_f = nullptr;      // reduced rhs (fred) scattered to global DOFs, 0 at fixed
  unsigned char *d_fixed = nullptr;  // aliases plan.d_fixed when has_plan
  std::vector<int64_t> bc_dof;
  std::vector<double> bc_val;
  std::vector<int64_t> free_dofs;
  bool owns_pattern = false;
  tfm::BSell bsell;           // solver copy, built on the first solve
  tfm::BatchSolver *mg_solver = nullptr;  // multigrid solver of tfm_system_solve, built on its first use
  // batched solver of tfm_system_solve_batch / tfm_system_assemble_solve_batch,
  // kept (with its captured graphs) while P, the DOF ranges, the rhs count
  // and the mode stay the same
  tfm::BatchSolver *batch_solver = nullptr;
  uint64_t batch_key = 0;
  // Coefficient versions: `version` counts the coefficient sets the system
  // has been assembled with; each numeric pass stamps the value arrays it
  // wrote with the current one.  FP32 values describe the same matrix as
  // the FP64 values only while ver32 == ver64 (an FP64-only re-assembly with
  // new coefficients leaves them stale until an FP32 pass runs).
  uint64_t version = 0, ver64 = 0, ver32 = 0;
  void note_values(bool new_coefficients, bool w64, bool w32) {
    if (new_coefficients || version == 0) ++version;
    if (w64) ver64 = version;
    if (w32) ver32 = version;
  }
  bool fp32_current() const { return d_v32 != nullptr && ver32 == ver64; }
  int csr_chunk = 0;  // rows per CSR-stream chunk (from the longest row), once known
};
