#!/usr/bin/env python3
"""Benchmark: elements assembled/s + CG time to 1e-8 rel. error vs FP64,
mixed vs plain, on 128^3-cell jittered P1-tet linear elasticity.

Workload (one MI355X): box_mesh 128x128x128 cells x 6 tets (12,582,912 tets,
2,146,689 nodes, 6,440,067 DOF); interior nodes jittered by a seeded
uniform +-0.2h per axis; E per element log-uniform in [0.1, 10], nu = 0.3; all
DOFs clamped on x = 0; traction (0, 0, -1e-3) on x = 1.

A step is one numeric assembly pass of the reduced stiffness over the whole
mesh (the hot path), with mesh, coefficients and the CSR pattern resident in
HBM.  value = elements assembled per second by the FP64 (plain) pass.  After
the timed steps the reduced system is solved with FP64 PCG to a relative
residual of 1e-12 (the FP64 reference solution); then plain FP64 PCG and
FP32 PCG with FP64 residual replacement (on the matrix-free FP32 operator,
and on the assembled FP32 values) are each timed to 1e-8 relative error
against that solution (the loosest residual tolerance of a sweep that
reaches it), and the line reports the times and their ratios, with the
multigrid CG beside them.  Sub-records c2 (64^3 elasticity), c4 (256 uneven
Poisson boxes, batched) and c5 (64 SIMP designs x 8 loads, batched) carry
each workload's assembly passes (time, algorithmic bytes, HBM fraction) and
its solvers' times to 1e-8 error against an FP64 solve; hex8 (64^3 jittered
Hex8 elasticity) the same, with the FP32 pass on the matrix cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tpt-fem_amd"))

PEAK_HBM_GBS = 8000.0


def asm_kernel(tile_pairs, vt="double", dpn=3):
    """Name rocprofv3 gives the P1-tet assembly kernel at a tile size."""
    return f"tfm::k_asm_tet<{dpn}, {vt}, {tile_pairs}>"


def asm_bytes(n_nodes, n_elem, nnz, n_free, sv, per_elem_coef=True):
    """Algorithmic HBM bytes of one numeric assembly pass: FP64 coordinates,
    int32 connectivity, the per-element modulus, values written once plus
    col_ind read once, and the row pointers."""
    return 24 * n_nodes + 16 * n_elem + (8 * n_elem if per_elem_coef else 0) + (sv + 4) * nnz + 4 * (n_free + 1)


def host_cpu():
    """The host's CPU model and logical CPU count (for the CPU baseline record)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


class one_core:
    """Pins the calling thread to one core (the lowest it may run on) for the
    CPU baseline, as the single-threaded reference would run; restores the
    previous affinity on exit."""

    def __enter__(self):
        self.prev = os.sched_getaffinity(0)
        self.core = min(self.prev)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.prev)
        return False


def cpu_baseline(n, seed=7):
    """Single-core oracle (restated reference CPU path): assemble triplets with
    the exact-zero skip, Coo::to_csr and reduce_system, on an n^3 sample of
    the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (n, n, n))
    coords = O.jitter_interior(coords, (n, n, n), 0.2, seed)
    ye = 10.0 ** np.random.default_rng(seed).uniform(-1.0, 1.0, len(conn))
    bcs = [(int(i) * 3 + k, 0.0) for i in O.nodes_on_plane(coords, 0, 0.0, 1e-9) for k in range(3)]
    rhs = np.zeros(len(coords) * 3)
    op = O.make_op(O.TET, 3, 1, coef_e=ye, nu=0.3)
    with one_core():
        t0 = time.perf_counter()
        coo = O.assemble(coords, conn, op, 3)
        O.reduce_system(coo, rhs, bcs)
        dt = time.perf_counter() - t0
    return len(conn) / dt, len(conn), dt


def cpu_baseline_c1(reps=20):
    """C1's reference pipeline on one core: solve_poisson on the 8^3 box
    (assemble, to_csr, reduce_system, dense LU), mean wall time of `reps`
    calls, and the solution."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    coords, conn = O.box_mesh([0, 0, 0], [1, 1, 1], (8, 8, 8))
    bcs = [(int(v), 0.0) for v in sorted({int(v) for a in range(3) for c in (0.0, 1.0)
                                          for v in O.nodes_on_plane(coords, a, c, 1e-9)})]
    with one_core():
        u = O.solve_poisson(coords, O.TET, conn, 1.0, 2, 1.0, bcs)
        t0 = time.perf_counter()
        for _ in range(reps):
            O.solve_poisson(coords, O.TET, conn, 1.0, 2, 1.0, bcs)
        dt = (time.perf_counter() - t0) / reps
    return dt, u


def bench_c1(args, T, log, reps=20):
    """The CPU-runnable configuration: Poisson -lap u = 1 on the unit cube,
    box_mesh 8^3 x 6 P1 tets, k = 1, order 2, u = 0 on the six faces, FP64
    assembly and the direct (dense LU) solve, as one solve_poisson call on
    the GPU; checked against the discrete solution's closed form (DST-I:
    u_max = 0.05491766911624052 at node (4, 4, 4))."""
    ctx = T.Context(0)
    mesh = T.Mesh.box(ctx, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], (8, 8, 8))
    bcs = [(int(v), 0.0) for v in sorted({int(v) for a in range(3) for c in (0.0, 1.0)
                                          for v in mesh.nodes_on_plane(a, c, 1e-9)})]
    u = T.solve_poisson(mesh, 1.0, 2, 1.0, bcs, mode=T.SOLVE_DIRECT)  # warm-up (code objects)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        u = T.solve_poisson(mesh, 1.0, 2, 1.0, bcs, mode=T.SOLVE_DIRECT)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / reps
    umax = 0.05491766911624052
    rec = {"workload": "Poisson box_mesh 8^3 x6 P1 tets, k 1, order 2, u = 0 on all faces, dense LU",
           "n_elem": mesh.n_elem, "n_nodes": mesh.n_nodes,
           "gpu_solve_poisson_ms": dt * 1e3, "u_max": float(u.max()),
           "u_max_rel_err_vs_closed_form": abs(float(u.max()) - umax) / umax}
    log(f"c1 solve_poisson (GPU, dense LU): {dt * 1e3:.2f} ms per call, u_max {u.max():.17f}")
    if not args.skip_cpu:
        dtc, uc = cpu_baseline_c1(reps)
        rec["cpu_baseline"] = {"ms": dtc * 1e3, "cores": 1, "kind": "port",
                               "sample": "oracle solve_poisson (assemble, to_csr, reduce_system, dense LU), the "
                                         "whole configuration",
                               "rel_diff_gpu_vs_cpu": float(np.abs(u - uc).max() / np.abs(uc).max())}
        log(f"c1 oracle solve_poisson (1 core): {dtc * 1e3:.2f} ms per call")
    mesh.close()
    ctx.close()
    return rec


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` on the n^3 workload, from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a 1 GiB copy kernel), or None when no
    summary for this kernel and size is committed."""
    path = os.path.join(ROOT, "profiles", "asm_tet_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("n") != n:
        return None
    return rec.get("bytes_per_launch")


def copy_probe():
    """Five copies of 1 GiB (2^28 float32) between two device buffers: a
    kernel whose HBM traffic is known (1 GiB read + 1 GiB written per launch),
    profiled beside the assembly to scale the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_launch": 1 << 30, "launches": 5}), flush=True)


def box_face_nodes(n):
    """Node ids of box_mesh(n) on any of its six faces (node (i, j, k) is
    (i*(ny+1) + j)*(nz+1) + k)."""
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    on = (i == 0) | (i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz)
    return ((i * (ny + 1) + j) * (nz + 1) + k)[on]


def pcg_iter_bytes(nnz, n, sv):
    """Algorithmic bytes of one PCG iteration on the reduced CSR system:
    (s_v + 4) nnz + 4 (n + 1) + 2 s_v n for the SpMV, 16 s_v n for the vector
    updates (s_v: value bytes)."""
    return (sv + 4) * nnz + 4 * (n + 1) + 2 * sv * n + 16 * sv * n


def timed_assembly(ctx, s, steps, warmup):
    """Mean event-timed numeric passes (ms) of a system: FP64, FP32 and both
    arrays in one call."""
    out = {}
    for which, name in ((T_SYS[0], "fp64"), (T_SYS[1], "fp32"), (T_SYS[0] | T_SYS[1], "fp64_fp32")):
        for _ in range(warmup):
            s.assemble(None, which)
        ts = []
        for _ in range(steps):
            ctx.timer_begin()
            s.assemble(None, which)
            ts.append(ctx.timer_end())
        out[name] = float(np.mean(ts))
    s.assemble(None, 0)
    return out


def assembly_record(s, mesh, ms, per_elem_coef, cell_dpn):
    """Assembly times, algorithmic GB/s and HBM roofline fraction per pass."""
    info = s.assembly_info()
    rec = {"fp64_ms": ms["fp64"], "fp32_ms": ms["fp32"], "fp64_fp32_ms": ms["fp64_fp32"], "autotune": info}
    for name, sv in (("fp64", 8), ("fp32", 4)):
        b = asm_bytes(mesh.n_nodes, mesh.n_elem, s.nnz, s.n_free, sv, per_elem_coef)
        gbs = b / (ms[name] * 1e-3) / 1e9
        vt = "double" if sv == 8 else "float"
        if info["kernel"] == "p1_tet":
            kern = asm_kernel(info["tile_pairs"], vt, cell_dpn[1])
        elif info["kernel"] == "simp_grid":
            kern = f"tfm::k_simp_rows<{vt}, false>"
        elif info["kernel"] == "triplets":
            kern = f"tfm::k_coo_sum<{vt}, int>"
        else:
            kern = f"tfm::k_asm_tile<{cell_dpn[0]}, {cell_dpn[1]}, {2 if sv == 8 else 3}>"
        rec[name + "_roofline"] = {"kernel": kern, "algorithmic_bytes": b, "GBps": gbs, "frac": gbs / PEAK_HBM_GBS}
    return rec


def sweep(run, err_of, rtols, target, log, name):
    """The loosest tolerance of `rtols` whose solution is within `target`
    (relative error) of the FP64 reference: (that solve's record, every try)."""
    tried = []
    for rtol in rtols:
        rec = run(rtol)
        rec["rel_err_vs_fp64"] = err_of(rec.pop("u"))
        rec["rtol"] = rtol
        tried.append(rec)
        log(f"{name}: rtol {rtol:.0e}: {rec['iters']} it, {rec['s']:.3f} s, err {rec['rel_err_vs_fp64']:.2e}")
        if rec["rel_err_vs_fp64"] <= target:
            return rec, tried
    return None, tried


def single_run(T, ctx, s, mode, inner_rtol):
    def run(rtol):
        ctx.synchronize()
        w0 = time.perf_counter()
        u = s.solve(mode, rtol=rtol, max_iter=200000, inner_rtol=inner_rtol)
        wall = time.perf_counter() - w0
        st = s.stats
        return {"s": st.seconds, "iters": st.iterations, "cycles": st.outer,
                "ms_per_iter": 1e3 * st.seconds / max(1, st.iterations), "wall_s": wall, "u": u}
    return run


def batch_run(T, ctx, s, mode, off, rhs, nrhs, inner_rtol):
    def run(rtol):
        ctx.synchronize()
        w0 = time.perf_counter()
        u = s.solve_batch(off, rhs=rhs, nrhs=nrhs, mode=mode, rtol=rtol, max_iter=200000, inner_rtol=inner_rtol)
        wall = time.perf_counter() - w0
        st = s.batch_stats
        return {"s": st[0][3], "iters": max(x[0] for x in st), "max_rel_res": max(x[2] for x in st),
                "wall_s": wall, "u": u}
    return run


def segment_err(u, u_ref, off):
    """Largest relative error over the problems (and right-hand sides)."""
    u, u_ref = np.atleast_2d(u), np.atleast_2d(u_ref)
    worst = 0.0
    for k in range(u.shape[0]):
        for p in range(len(off) - 1):
            a, b = off[p], off[p + 1]
            nr = np.linalg.norm(u_ref[k, a:b])
            if nr > 0:
                worst = max(worst, float(np.linalg.norm(u[k, a:b] - u_ref[k, a:b]) / nr))
    return worst


BATCH_RTOLS = [1e-8, 3e-9, 1e-9, 3e-10, 1e-10, 3e-11, 1e-11]
# The multigrid CG's error sits orders of magnitude below its residual
# tolerance, so its sweep starts looser; every solver is timed at the
# loosest tolerance of its sweep that meets the error target.
MG_LOOSE = [1e-4, 3e-5, 1e-5, 3e-6, 1e-6, 3e-7, 1e-7, 3e-8]


def rtols_for(mode, rtols, T):
    """The sweep of a solver mode: multigrid's looser head, then rtols."""
    if mode != T.SOLVE_PCG64_MG:
        return rtols
    return [r for r in MG_LOOSE if r > rtols[0]] + list(rtols)


def bench_c2(args, T, log):
    """box_mesh 64^3 P1-tet elasticity, E = 1, nu = 0.3, x = 0 clamped,
    traction on x = 1: assembly and the solvers to 1e-8 error vs FP64."""
    ctx = T.Context(0)
    n = 64
    mesh = T.Mesh.box(ctx, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], (n, n, n))
    bcs = [(int(i) * 3 + k, 0.0) for i in mesh.nodes_on_plane(0, 0.0, 1e-9) for k in range(3)]
    rhs = T.plane_traction(mesh, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    s = T.System.create(mesh, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2), 3, bcs, rhs,
                        flags=T.SYS_FP64 | T.SYS_FP32)
    rec = {"workload": "P1-tet elasticity box_mesh 64^3 x6, E 1, nu 0.3, x=0 clamped, traction on x=1",
           "n_elem": mesh.n_elem, "n_free": s.n_free, "nnz": s.nnz,
           "assembly": assembly_record(s, mesh, timed_assembly(ctx, s, args.steps, args.warmup), False, (3, 3))}
    if not args.skip_solve:
        u_ref = s.solve(T.SOLVE_PCG64_MG, rtol=1e-12, max_iter=5000)
        nref = np.linalg.norm(u_ref)
        err = lambda u: float(np.linalg.norm(u - u_ref) / nref)
        solves = {}
        # the configuration's own solve: Jacobi PCG to relative residual 1e-12
        r12 = single_run(T, ctx, s, T.SOLVE_PCG64, args.inner_rtol)(1e-12)
        r12["rel_err_vs_multigrid_1e-12"] = err(r12.pop("u"))
        r12["rel_residual"] = s.stats.rel_residual
        r12["rtol"] = 1e-12
        solves["jacobi_fp64_rtol1e-12"] = r12
        log(f"c2 jacobi to 1e-12: {r12['iters']} it, {r12['s']:.3f} s, true residual {r12['rel_residual']:.2e}")
        for mode, name in ((T.SOLVE_PCG64, "plain_fp64"), (T.SOLVE_MIXED, "mixed_fp32_fp64refine"),
                           (T.SOLVE_PCG64_MG, "multigrid_fp64")):
            solves[name], _ = sweep(single_run(T, ctx, s, mode, args.inner_rtol), err,
                                    rtols_for(mode, args.rtol_sweep, T), args.err_target, log, "c2 " + name)
        rp, rm = solves["plain_fp64"], solves["mixed_fp32_fp64refine"]
        solves["mixed_speedup_vs_plain"] = rp["s"] / rm["s"] if rp and rm else None
        rec["solve"] = solves
    s.close()
    mesh.close()
    ctx.close()
    return rec


def bench_batch(args, T, log, kind, emit=True):
    """The batching forms: 256 uneven Poisson boxes (kappa_e = exp(N(0,1)),
    f = 1, u = 0 on the faces), or 64 SIMP Q4 cantilever designs x 8 point
    loads.  One assembly pass fills the block-diagonal system; one segmented
    PCG solves every problem (and right-hand side) together; each solver is
    timed to 1e-8 relative error (worst problem and load) against an FP64
    solve of the batch."""
    ctx = T.Context(0)
    rng = np.random.default_rng(2024)
    t0 = time.perf_counter()
    if kind == "c4":
        dims = [tuple(int(v) for v in d) for d in rng.integers(8, 49, size=(256, 3))]
        mesh, noff, eoff = T.Mesh.box_batch(ctx, dims)
        kappa = np.exp(rng.standard_normal(mesh.n_elem))
        ids = np.concatenate([noff[p] + box_face_nodes(d) for p, d in enumerate(dims)])
        bcs = np.stack([ids.astype(np.float64), np.zeros(len(ids))], axis=1)
        rhs = T.source_vector(mesh, 1.0)
        op = T.Operator.poisson(1.0, 2, kappa_e=kappa)
        dpn, nrhs, off, f, cell = 1, 1, noff, None, 3
        desc = (f"256 Poisson boxes, n per axis U{{8..48}}, {mesh.n_elem} tets, kappa_e = exp(N(0,1)), "
                "f = 1, u = 0 on faces")
        ref_mode = T.SOLVE_PCG64
        modes = [(T.SOLVE_PCG64, "plain_fp64"), (T.SOLVE_MIXED, "mixed_fp32_fp64refine")]
    else:
        P, nx, ny = 64, 257, 129
        mesh = T.Mesh.grid_batch(ctx, P, nx, ny, 1.0)
        rho = rng.uniform(0.01, 1.0, mesh.n_elem)
        nd1 = 2 * nx * ny
        left = (np.arange(P)[:, None] * nx * ny + np.arange(ny)[None, :] * nx).ravel()
        ids = np.concatenate([2 * left, 2 * left + 1])
        bcs = np.stack([ids.astype(np.float64), np.zeros(len(ids))], axis=1)
        f = np.zeros((8, P * nd1))
        rows = rng.integers(0, ny, size=(P, 8))
        for p in range(P):
            for k in range(8):
                f[k, 2 * (p * nx * ny + rows[p, k] * nx + nx - 1) + 1] = -1.0
        op = T.Operator.simp(rho, nu=0.3, penal=3.0, e0=1.0, emin=1e-9, h=1.0)
        dpn, nrhs, off, rhs, cell = 2, 8, np.arange(P + 1) * nd1, None, 2
        desc = f"64 SIMP Q4 designs 257x129 nodes x 8 point loads, rho U[0.01,1], p 3, {mesh.n_elem} quads"
        ref_mode = T.SOLVE_PCG64_MG
        modes = [(T.SOLVE_PCG64, "plain_fp64"), (T.SOLVE_MIXED, "mixed_fp32_fp64refine"),
                 (T.SOLVE_PCG64_MG, "multigrid_fp64")]
    s = T.System.create(mesh, op, dpn, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    ctx.synchronize()
    log(f"{kind} setup {time.perf_counter() - t0:.2f}s: {desc}; n_free {s.n_free}, nnz {s.nnz}")
    ms = timed_assembly(ctx, s, args.steps, args.warmup)
    log(f"{kind} assembly fp64 {ms['fp64']:.3f} ms, fp32 {ms['fp32']:.3f} ms")
    rec = {"workload": desc, "n_elem": mesh.n_elem, "n_free": s.n_free, "nnz": s.nnz,
           "assembly": assembly_record(s, mesh, ms, True, (cell, dpn)),
           "elements_per_s_fp64": mesh.n_elem / (ms["fp64"] * 1e-3)}
    if not args.skip_solve:
        u_ref = s.solve_batch(off, rhs=f, nrhs=nrhs, mode=ref_mode, rtol=1e-12, max_iter=200000)
        err = lambda u: segment_err(u, u_ref, off)
        solves = {"reference": {"mode": "plain_fp64" if ref_mode == T.SOLVE_PCG64 else "multigrid_fp64",
                                "rtol": 1e-12}}
        for mode, name in [m for m in modes if m[1].split("_")[0] in args.solvers]:
            solves[name], _ = sweep(batch_run(T, ctx, s, mode, off, f, nrhs, args.inner_rtol), err,
                                    rtols_for(mode, BATCH_RTOLS, T), args.err_target, log, f"{kind} {name}")
        rp, rm = solves.get("plain_fp64"), solves.get("mixed_fp32_fp64refine")
        solves["mixed_speedup_vs_plain"] = rp["s"] / rm["s"] if rp and rm else None
        rec["solve"] = solves
    s.close()
    mesh.close()
    ctx.close()
    if emit:
        line = {"metric": "batched assembly + solve", "workload": kind, "config": {"workload": desc}, **rec}
        print(json.dumps(line), flush=True)
    return rec


def hex_box(n, seed=1):
    """n^3 Hex8 box on the unit cube, interior nodes jittered by +-0.2h."""
    i, j, k = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    coords = np.stack([i, j, k], -1).reshape(-1, 3).astype(np.float64) / n
    rng = np.random.default_rng(seed)
    inner = ((i > 0) & (i < n) & (j > 0) & (j < n) & (k > 0) & (k < n)).ravel()
    coords[inner] += rng.uniform(-0.2 / n, 0.2 / n, (int(inner.sum()), 3))
    idx = lambda a, b, c: (a * (n + 1) + b) * (n + 1) + c
    I, J, K = [v.ravel() for v in np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")]
    conn = np.stack([idx(I, J, K), idx(I + 1, J, K), idx(I + 1, J + 1, K), idx(I, J + 1, K),
                     idx(I, J, K + 1), idx(I + 1, J, K + 1), idx(I + 1, J + 1, K + 1), idx(I, J + 1, K + 1)], -1)
    return coords, conn, rng


def bench_hex(args, T, log, n=64):
    """Hex8 3-D elasticity on a jittered n^3 box (E per element log-uniform
    [0.1, 10], nu 0.3, 2x2x2 Gauss points, x = 0 clamped, traction on x = 1):
    FP64 and FP32 assembly passes, the FP32 one with the element matrices on
    the matrix cores and, for comparison, with the per-pair FP32 kernel
    (TFM_NO_MFMA); then plain and precision-form Jacobi PCG to 1e-8 error."""
    ctx = T.Context(0)
    coords, conn, rng = hex_box(n)
    mesh = T.Mesh.from_arrays(ctx, T.HEX, coords, conn)
    ye = 10.0 ** rng.uniform(-1, 1, len(conn))
    bcs = [(int(q) * 3 + c, 0.0) for q in np.flatnonzero(coords[:, 0] == 0.0) for c in range(3)]
    rhs = T.plane_traction(mesh, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    s = T.System.create(mesh, T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye), 3, bcs, rhs,
                        flags=T.SYS_FP64 | T.SYS_FP32)
    ms = timed_assembly(ctx, s, args.steps, args.warmup)
    os.environ["TFM_NO_MFMA"] = "1"
    try:
        ms_pair = timed_assembly(ctx, s, args.steps, args.warmup)["fp32"]
    finally:
        del os.environ["TFM_NO_MFMA"]
    s.assemble(None, 0)
    log(f"hex8 assembly fp64 {ms['fp64']:.3f} ms, fp32 (MFMA) {ms['fp32']:.3f} ms, fp32 per pair {ms_pair:.3f} ms")
    kpad = 8
    rec = {"workload": f"Hex8 elasticity {n}^3 jittered box, E log-uniform [0.1,10], nu 0.3, 2x2x2 Gauss",
           "n_elem": mesh.n_elem, "n_free": s.n_free, "nnz": s.nnz,
           "assembly": assembly_record(s, mesh, ms, True, (4, 3)),
           "fp32_per_pair_kernel_ms": ms_pair, "fp32_mfma_speedup": ms_pair / ms["fp32"],
           "mfma_flops_per_pass": mesh.n_elem * 2 * 32 * 32 * kpad}
    if not args.skip_solve:
        u_ref = s.solve(T.SOLVE_PCG64, rtol=1e-12, max_iter=200000)
        nref = np.linalg.norm(u_ref)
        err = lambda u: float(np.linalg.norm(u - u_ref) / nref)
        solves = {}
        for mode, name in ((T.SOLVE_PCG64, "plain_fp64"), (T.SOLVE_MIXED, "mixed_fp32_fp64refine")):
            solves[name], _ = sweep(single_run(T, ctx, s, mode, args.inner_rtol), err, args.rtol_sweep,
                                    args.err_target, log, "hex8 " + name)
        rp, rm = solves["plain_fp64"], solves["mixed_fp32_fp64refine"]
        solves["mixed_speedup_vs_plain"] = rp["s"] / rm["s"] if rp and rm else None
        rec["solve"] = solves
    s.close()
    mesh.close()
    ctx.close()
    return rec


T_SYS = (1, 2)  # TFM_SYS_FP64, TFM_SYS_FP32


def bench_simp(args, T, log):
    """The SIMP outer loop, batched: 64 cantilever designs on Grid::new(257,
    129, 1) (256 x 128 quads), each with a unit point load (-y) at a seeded
    right-edge node, vol_frac 0.5, p 3, E0 1, nu 0.3, filter radius 2.4, move
    0.2 (the reference's own example parameters), args.simp_iters OC
    iterations.  Whole-loop wall time per OC iteration of all designs with
    three state solvers: Jacobi FP64 PCG, the precision form and FP64 CG with
    the multigrid V-cycle (the default; `value` is its rate)."""
    ctx = T.Context(0)
    rng = np.random.default_rng(2025)
    P, nx, ny, iters = 64, 257, 129, args.simp_iters
    nd1 = 2 * nx * ny
    F = np.zeros((P, nd1))
    rows = rng.integers(0, ny, size=P)
    for p in range(P):
        F[p, 2 * (rows[p] * nx + nx - 1) + 1] = -1.0
    bcs = [(2 * j * nx + c, 0.0) for j in range(ny) for c in (0, 1)]
    desc = (f"{P} SIMP cantilever designs, Grid {nx}x{ny} nodes ({(nx - 1) * (ny - 1)} quads), point load at a "
            f"seeded right-edge node, vol 0.5, p 3, R 2.4, move 0.2, {iters} OC iterations")
    out = {}
    modes = ((T.SOLVE_PCG64, "plain_fp64", 1e-10), (T.SOLVE_MIXED, "mixed_fp32_fp64refine", 1e-10),
             (T.SOLVE_PCG64_MG, "multigrid_fp64", 1e-10))
    for mode, name, rtol in [m for m in modes if m[1].split("_")[0] in args.solvers]:
        T.topopt_simp(ctx, nx, ny, 1.0, 1.0, 0.3, 0.5, 3.0, 2.4, 1, 0.2, F[:4], bcs, mode=mode, rtol=rtol)  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        r = T.topopt_simp(ctx, nx, ny, 1.0, 1.0, 0.3, 0.5, 3.0, 2.4, iters, 0.2, F, bcs, mode=mode, rtol=rtol)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        st = r.stats
        rec = {"rtol": rtol, "wall_s": wall, "ms_per_oc_iteration": 1e3 * wall / iters,
               "solve_device_s_max": max(x[3] for x in st),
               "pcg_iterations_mean_per_design": float(np.mean([x[0] for x in st])),
               "worst_rel_residual": max(x[2] for x in st),
               "compliance_first_last_mean": [float(r.compliance[:, 0].mean()), float(r.compliance[:, -1].mean())],
               "volume_error_max": float(np.abs(r.densities.sum(axis=1) - 0.5 * r.densities.shape[1]).max())}
        out[name] = rec
        log(f"simp {name}: {rec}")
    vname = "multigrid_fp64" if "multigrid_fp64" in out else next(iter(out))
    line = {"metric": "SIMP OC iterations (all designs) per second", "workload": "simp", "config": {"workload": desc},
            "value": iters / out[vname]["wall_s"], "unit": "OC iterations/s",
            "value_solver": vname + (" (the default state solver)" if vname == "multigrid_fp64" else ""),
            "loops": out}
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        cn, cm, ci = 49, 25, 5
        f1, b1 = O.cantilever_load(cn, cm, 1.0)
        with one_core():
            t0 = time.perf_counter()
            O.topopt_simp(cn, cm, 1.0, 1.0, 0.3, 0.5, 3.0, 2.4, ci, 0.2, f1, b1)
            dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": ci / dt, "unit": "OC iterations/s of ONE design", "cores": 1,
                                "kind": "port", "host": host_cpu(),
                                "sample": f"oracle topopt_simp (dense LU, as the reference), 1 design Grid {cn}x{cm}, "
                                          f"{ci} OC iterations; the 257x129 grid needs a 34.9 GB dense LU per solve"}
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_c3(args, T, log):
    """The headline workload: timed assembly passes, then the solvers to the
    error target against an FP64 reference solve."""
    n = args.n
    ctx = T.Context(0)
    t_setup = time.perf_counter()
    mesh = T.Mesh.box(ctx, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], (n, n, n))
    mesh.jitter((n, n, n), 0.2, 2024)
    ye = 10.0 ** np.random.default_rng(2024).uniform(-1.0, 1.0, mesh.n_elem)
    clamp = mesh.nodes_on_plane(0, 0.0, 1e-9)
    bcs = [(int(i) * 3 + k, 0.0) for i in clamp for k in range(3)]
    rhs = T.plane_traction(mesh, 3, 0, 1.0, 1e-9, [0.0, 0.0, -1e-3])
    op = T.Operator.elasticity(T.CONTINUUM3D, 1.0, 0.3, 2, young_e=ye)
    ctx.synchronize()
    t_create = time.perf_counter()
    sysm = T.System.create(mesh, op, 3, bcs, rhs, flags=T.SYS_FP64 | T.SYS_FP32)
    ctx.synchronize()
    create_s = time.perf_counter() - t_create
    tune = sysm.assembly_info()
    kernel = asm_kernel(tune["tile_pairs"])
    log(f"setup {time.perf_counter() - t_setup:.2f}s (system create {create_s:.3f}s): {mesh.n_elem} tets, "
        f"{mesh.n_nodes} nodes, n_free {sysm.n_free}, nnz {sysm.nnz}; assembly kernel {kernel} ({tune['source']})")

    # ---- timed assembly --------------------------------------------------
    for _ in range(args.warmup):
        sysm.assemble(None, T.SYS_FP64)
        sysm.assemble(None, T.SYS_FP32)
    ctx.synchronize()
    t64, t32 = [], []
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.timer_begin()
        sysm.assemble(None, T.SYS_FP64)
        t64.append(ctx.timer_end())
    ctx.synchronize()
    wall = time.perf_counter() - wall0
    for _ in range(args.steps):
        ctx.timer_begin()
        sysm.assemble(None, T.SYS_FP32)
        t32.append(ctx.timer_end())
    # both arrays of the precision form in one call: one pass writing the
    # FP64 values and their FP32 roundings
    tb = []
    for _ in range(args.steps):
        ctx.timer_begin()
        sysm.assemble(None, T.SYS_FP64 | T.SYS_FP32)
        tb.append(ctx.timer_end())
    ms64 = float(np.mean(t64))
    ms32 = float(np.mean(t32))
    msb = float(np.mean(tb))
    b64 = asm_bytes(mesh.n_nodes, mesh.n_elem, sysm.nnz, sysm.n_free, 8)
    b32 = asm_bytes(mesh.n_nodes, mesh.n_elem, sysm.nnz, sysm.n_free, 4)
    gbs64 = b64 / (ms64 * 1e-3) / 1e9
    gbs32 = b32 / (ms32 * 1e-3) / 1e9
    bb = b64 + 4 * sysm.nnz
    gbsb = bb / (msb * 1e-3) / 1e9
    log(f"assembly fp64 {ms64:.3f} ms ({gbs64:.0f} GB/s), fp32 {ms32:.3f} ms ({gbs32:.0f} GB/s), "
        f"both {msb:.3f} ms ({gbsb:.0f} GB/s)")
    sysm.assemble(None, 0)

    # ---- solves ------------------------------------------------------------
    solve = {}
    if not args.skip_solve:
        u_ref = sysm.solve(T.SOLVE_PCG64, rtol=1e-12, max_iter=200000)
        st = sysm.stats
        solve["ref_fp64_rtol1e-12"] = {"s": st.seconds, "iters": st.iterations, "rel_res": st.rel_residual}
        log(f"fp64 ref: {st.iterations} it, {st.seconds:.3f}s, res {st.rel_residual:.2e}")
        nref = np.linalg.norm(u_ref)
        err = lambda u: float(np.linalg.norm(u - u_ref) / nref)
        run = lambda mode: single_run(T, ctx, sysm, mode, args.inner_rtol)
        legs = [(T.SOLVE_PCG64, "plain_fp64", "plain"), (T.SOLVE_MIXED, "mixed_fp32_fp64refine", "mixed"),
                (T.SOLVE_MIXED_CSR, "mixed_csr_fp32_fp64refine", "mixed")]
        sw = {}
        for mode, name, key in legs:
            if key in args.solvers:
                solve[name], sw[name] = sweep(run(mode), err, args.rtol_sweep, args.err_target, log, name)
        rp, rm, rc = solve.get("plain_fp64"), solve.get("mixed_fp32_fp64refine"), solve.get("mixed_csr_fp32_fp64refine")
        # per-iteration HBM rate of the assembled-matrix legs: the algorithmic
        # bytes of one PCG iteration (CSR SpMV + 16 s n of vector traffic)
        # over the measured time per iteration
        for name, sv in (("plain_fp64", 8), ("mixed_csr_fp32_fp64refine", 4)):
            r = solve.get(name)
            if r and r["ms_per_iter"] > 0:
                b = pcg_iter_bytes(sysm.nnz, sysm.n_free, sv)
                gbs = b / (r["ms_per_iter"] * 1e-3) / 1e9
                r["iteration_roofline"] = {"algorithmic_bytes": b, "GBps": gbs, "frac": gbs / PEAK_HBM_GBS,
                                           "kernels": "k_pcg_dir_bsell + k_pcg_update"}
        solve["err_target"] = args.err_target
        solve["mixed_operator"] = "matrix-free FP32 (box mesh of P1 tets)"
        solve["mixed_speedup_vs_plain"] = (rp["s"] / rm["s"]) if (rp and rm and rm["s"] > 0) else None
        solve["mixed_csr_speedup_vs_plain"] = (rp["s"] / rc["s"]) if (rp and rc and rc["s"] > 0) else None
        solve["sweep"] = sw
        if "multigrid" in args.solvers:
            # FP64 CG with the box-mesh multigrid V-cycle.  Its first call
            # builds the solver kept on the system (sliced copy, level
            # storage), as the reference solve above built the other forms'
            # sliced copy; every timed call still refreshes the values and the
            # Galerkin hierarchy (in wall_s, not in s).
            ctx.synchronize()
            w0 = time.perf_counter()
            sysm.solve(T.SOLVE_PCG64_MG, rtol=1e-2, max_iter=1000)
            first_wall = time.perf_counter() - w0
            rg, tg = sweep(run(T.SOLVE_PCG64_MG), err, rtols_for(T.SOLVE_PCG64_MG, args.rtol_sweep, T),
                           args.err_target, log, "multigrid")
            if rg:
                rg["first_call_wall_s"] = first_wall
            solve["multigrid_fp64"] = rg
            sw["multigrid_fp64"] = tg
            if rp and rg:
                solve["multigrid_speedup_vs_plain_wall"] = rp["wall_s"] / rg["wall_s"]

    cpu = None
    if not args.skip_cpu:
        eps, ne, dt = cpu_baseline(args.cpu_n)
        cpu = {"value": eps, "unit": "elements/s", "cores": 1, "kind": "port", "host": host_cpu(),
               "scope": "symbolic + numeric (assemble, to_csr, reduce_system) of a fresh system; `value` times "
                        "the numeric pass on a resident pattern, assembly.end_to_end_elements_per_s the GPU's "
                        "figure for the CPU's scope",
               "sample": f"oracle assemble+to_csr+reduce_system, {args.cpu_n}^3 jittered elasticity "
                         f"({ne} tets, {dt:.1f}s); a port of the reference's loop in C at -O3 without its "
                         "per-element heap allocations, so faster than the reference itself"}
        log(f"cpu baseline {eps:.0f} elements/s")

    line = {
        "metric": "elements assembled/s + CG time to 1e-8 rel err vs FP64, mixed vs plain",
        "value": mesh.n_elem / (ms64 * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jitter and moduli)",
        "config": {"workload": f"P1-tet elasticity box_mesh {n}^3 x6, jitter 0.2h, E log-uniform [0.1,10], "
                               "nu 0.3, x=0 clamped, traction on x=1",
                   "n_elem": mesh.n_elem, "n_nodes": mesh.n_nodes, "n_free": sysm.n_free, "nnz_reduced": sysm.nnz},
        "assembly": {"fp64_ms": ms64, "fp32_ms": ms32, "fp32_elements_per_s": mesh.n_elem / (ms32 * 1e-3),
                     "fp64_GBps": gbs64, "fp32_GBps": gbs32, "autotune": tune,
                     "fp32_roofline": {"kernel": asm_kernel(tune["tile_pairs"], "float"),
                                       "algorithmic_bytes": b32, "frac": gbs32 / PEAK_HBM_GBS},
                     "fp64_fp32_ms": msb,
                     "fp64_fp32_roofline": {"kernel": f"tfm::k_asm_tet_dual<3, {tune['tile_pairs']}>",
                                            "algorithmic_bytes": bb, "GBps": gbsb, "frac": gbsb / PEAK_HBM_GBS},
                     "system_create_s": create_s,
                     "end_to_end_elements_per_s": mesh.n_elem / create_s},
        "solve": solve,
        "roofline": {"bound": "hbm", "achieved": gbs64, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": gbs64 / PEAK_HBM_GBS, "traffic": measured_traffic(kernel, n),
                     "kernel": kernel, "algorithmic_bytes_per_launch": b64},
        "cpu_baseline": cpu,
    }
    sysm.close()
    mesh.close()
    ctx.close()
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="cells per axis")
    ap.add_argument("--err-target", type=float, default=1e-8,
                    help="relative error vs the FP64 solution (rtol 1e-12) that the timed solves must reach")
    ap.add_argument("--rtol-sweep", type=float, nargs="+", default=[1e-6, 7e-7, 5e-7, 4e-7, 3e-7, 2e-7, 1.5e-7, 1e-7, 7e-8, 5e-8, 3e-8, 2e-8, 1e-8, 5e-9, 2e-9, 1e-9],
                    help="residual tolerances tried, loosest first, until the error target is met")
    ap.add_argument("--inner-rtol", type=float, default=0.3,
                    help="FP32 residual reduction between FP64 residual replacements")
    ap.add_argument("--cpu-n", type=int, default=48, help="cells per axis of the CPU-baseline sample (about 10 s on one core)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-solve", action="store_true")
    ap.add_argument("--no-sub", action="store_true",
                    help="c3 only: leave out the c1 / c2 / c4 / c5 / hex8 sub-records of the line")
    ap.add_argument("--workload", choices=["c3", "c1", "c2", "c4", "c5", "hex8", "simp"], default="c3",
                    help="c3: the headline single problem (with c1 / c2 / c4 / c5 / hex8 sub-records); c1 / c2 / c4 / c5 / hex8 alone; "
                         "simp: the batched SIMP loop (own JSON lines)")
    ap.add_argument("--simp-iters", type=int, default=30,
                    help="OC iterations of the simp workload (the reference example's max_iter)")
    ap.add_argument("--solvers", nargs="+", choices=["plain", "mixed", "multigrid"],
                    default=["plain", "mixed", "multigrid"],
                    help="solvers timed by every workload (c3 times plain and mixed to the error target, and the "
                         "multigrid CG beside them)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run five 1 GiB device-to-device copies (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding a problem over several GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    import tptfem as T

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.workload == "simp":
        bench_simp(args, T, log)
        return
    if args.workload == "c1":
        print(json.dumps({"metric": "assembly + direct solve", "workload": "c1", **bench_c1(args, T, log)}), flush=True)
        return
    if args.workload == "c2":
        print(json.dumps({"metric": "assembly + solve", "workload": "c2", **bench_c2(args, T, log)}), flush=True)
        return
    if args.workload == "hex8":
        print(json.dumps({"metric": "assembly + solve", "workload": "hex8", **bench_hex(args, T, log)}), flush=True)
        return
    if args.workload in ("c4", "c5"):
        bench_batch(args, T, log, args.workload)
        return
    line = bench_c3(args, T, log)
    if not args.no_sub:
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = min(args.steps, 10), min(args.warmup, 2)
        line["c1"] = bench_c1(sub, T, log)
        line["c2"] = bench_c2(sub, T, log)
        line["c4"] = bench_batch(sub, T, log, "c4", emit=False)
        line["c5"] = bench_batch(sub, T, log, "c5", emit=False)
        line["hex8"] = bench_hex(sub, T, log)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
