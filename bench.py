#!/usr/bin/env python3
"""Benchmark: batched mixed-precision field solves on one MI355X.

Workload (default): 3D ellipsoid 80x65x90 mm, linear tetrahedra (250,047
nodes, 1,429,968 Tet4), 100 MHz complex system with seed-drawn mua/mus' blob
maps, 64 Gaussian boundary sources, diagonal preconditioning, relative
residual tolerance 1e-10 (the reference criterion ||r|| <= tol ||q||).

A "step" is one CalcFields over all 64 sources: complex-FP32 block COCG
(conjugate orthogonal CG for the complex symmetric system; 64 columns, one
block) inside FP64 reliable updates, right-hand sides and fields resident in
HBM.  value = field solves per second (columns / s).  --solver BICGSTAB
times the reference's method instead; either way the other method is timed
beside it (other_solver).

Also measured in the same run:
  - the plain FP64 single-right-hand-side GPU path (the reference's method,
    BiCGSTAB, same tolerance) on a subset of the columns -> fp64 solves / s,
    speedup;
  - the FP64 block path (the headline's method, all columns, same
    tolerance) -> the gain of precision alone;
  - field error of every mixed-precision column against an FP64 GPU block
    solve at relative residual 1e-13 (max relative L2);
  - roofline of the dominant kernel (block SpMM) timed with HIP events on
    the library stream;
  - cpu_baseline: the FP64 CPU oracle, one source per host thread (the
    reference's THREAD_LEVEL 2 solves sources concurrently, fwdsolver.cc:619-643).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the inner SpMM, from two rocprofv3 --pmc passes
# (FETCH_SIZE, WRITE_SIZE) scaled by a 1 GiB copy, one file per inner
# solver (the SpMM's fused dot reads a different W); the files give the
# commands
TRAFFIC_FILES = {
    "COCG": os.path.join(ROOT, "profiles", "spmm_pmc_traffic_ellipsoid250k_64src_cocg.md"),
    "BICGSTAB": os.path.join(ROOT, "profiles", "spmm_pmc_traffic_ellipsoid250k_64src.md"),
}


# the 64-column SpMM variants: kernel name and what it does
SPMM_KERNELS = {
    "direct": "k_spmm64<float2,0> (inner block SpMM + <W,Y> dots, complex FP32, 64 columns, gathers per nonzero)",
    "rowblock": "k_spmm_rb<float2,0> (inner block SpMM + <W,Y> dots, complex FP32, 64 columns, blocks of 8 rows "
                "sharing their neighbour-row gathers)",
    "lds": "k_spmm_lds<float2,0,...> (inner block SpMM, neighbour rows staged in LDS through registers)",
    "dma": "k_spmm_dma<...> (inner block SpMM, neighbour rows staged in LDS by the LDS DMA, double-buffered)",
    "rowblock_pipelined": "k_spmm_rbp<0,%s> (inner block SpMM + %s dots, complex FP32, 64 columns, row blocks of 8 "
                          "with packed masks, double-buffered gather batches)",
    "rowblock_mfma": "k_spmm_rbm<4,%s> (inner block SpMM + %s dots, complex FP32, 64 columns, blocks of 8 rows as "
                     "16x16x4 FP32 matrix-core products over their shared neighbour rows)",
}


def spmm_kernel_name(variant, solver):
    """Name of the inner SpMM kernel the solve launches: BiCGSTAB dots with
    the packed sign-pattern shadow residual, COCG with the direction
    (unconjugated)."""
    name = SPMM_KERNELS[variant]
    if variant in ("rowblock_pipelined", "rowblock_mfma"):
        name = name % (("true", "<rd,Y>") if solver == "BICGSTAB" else ("false", "p^T Y"))
    return name


def measured_traffic(workload, tile, kernel, solver):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (None unless it was measured for this workload, tile, kernel and inner
    solver)."""
    path = TRAFFIC_FILES.get(solver)
    if workload != "ellipsoid3d" or tile != 64 or not path or not os.path.exists(path):
        return None
    found, traffic = None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_kernel"):
                found = line.split("=")[1].strip()
            if line.startswith("traffic_bytes_per_launch"):
                traffic = int(line.split("=")[1])
    return traffic if found == kernel.split(" ")[0] else None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def dense_node_major(csc_nq):
    """(n x nQ) sparse -> dense complex128 node-major array."""
    return np.ascontiguousarray(csc_nq.toarray().astype(np.complex128))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ellipsoid3d")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--solver", default="COCG", choices=["BICGSTAB", "COCG"],
                    help="inner Krylov method of the timed solves; the other one is timed beside it")
    ap.add_argument("--tile", type=int, default=64)
    ap.add_argument("--fp64-cols", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-jacobian", action="store_true")
    ap.add_argument("--jac-workload", default="cyl3d_jac")
    ap.add_argument("--jac-reps", type=int, default=3)
    ap.add_argument("--jacobian-only", action="store_true",
                    help="run only the Jacobian leg (fields, projections, nodal and raster J) and print its "
                         "record; for kernel traces and counter passes of the Jacobian kernels")
    ap.add_argument("--raster", type=int, default=64, help="raster basis grid size per axis")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-multifreq", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--spmm-variant", default=None, choices=list(SPMM_KERNELS),
                    help="64-column SpMM variant of the solvers (default: the library's; the autotuner may change it)")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="arrival window of the serving leg")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--fp64-only", action="store_true",
                    help="run only the plain FP64 single-column leg (for a kernel trace of that path) and print "
                         "its record as JSON")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="short run for hardware-counter passes: one mixed solve with the default settings "
                         "(no autotuner: its timings are distorted under the profiler), 5 timed launches of the "
                         "inner SpMM, then three 1 GiB device-to-device copies that calibrate FETCH_SIZE / "
                         "WRITE_SIZE; prints no JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(sources would shard by column blocks, one block of sources per GPU). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import toastpp_amd as T
    from toastpp_amd import _lib as L
    from toastpp_amd import meshgen as G

    dev, hbm = L.device_info()
    log("device", dev, "%.0f GiB" % hbm)
    if args.jacobian_only:
        print(json.dumps({"jacobian": jacobian_leg(args)}), flush=True)
        return
    nodes, el, Q, M, mua, mus, ref, freq, desc = G.workload(args.workload)
    n = len(nodes)
    log("workload", desc, "nodes", n, "elements", len(el))
    mesh = T.Mesh(data=(nodes, el))
    mesh.SetQM(Q, M)
    sol = T.Solution(mua, mus, ref)
    omega = freq * 2 * np.pi * 1e-6
    nq = Q.shape[0]
    qv = mesh.Qvec(width=2.0)                    # (n x nQ)
    B = dense_node_major(qv)
    dB = L.DeviceArray.from_host(B)
    dX = L.DeviceArray(B.shape, np.complex128)

    def make_solver(mode, tol, tile=None, solver=None):
        f = T.FwdSolver(mesh, solver or args.solver, tol, complex=True)
        f.SetMode(mode)
        f.SetPrecon("DIAG")
        if tile:
            f.SetTile(tile)
        if args.spmm_variant:
            f.SetSpmmVariant(args.spmm_variant)
        f.Allocate()
        t0 = time.perf_counter()
        f.Reset(sol, omega)
        return f, (time.perf_counter() - t0) * 1e3

    fm, asm_ms = make_solver("mixed", args.tol, args.tile)
    log("assembly + preconditioner (host call incl. upload) %.1f ms" % asm_ms)
    tuned = None
    if not args.no_autotune and not args.pmc_pass and not args.fp64_only:
        t0 = time.perf_counter()
        tuned = fm.Autotune(nq, dB.ptr, dX.ptr, nq)
        log("autotune (%.1f s): %s" % (time.perf_counter() - t0, tuned))
    if args.fp64_only:
        print(json.dumps({"fp64_single": fp64_single_leg(args, make_solver, L, B, n, nq, None)}), flush=True)
        return
    if args.pmc_pass:
        fm.CalcFieldsDevice(nq, dB.ptr, dX.ptr)
        fm.TimeKernel(0, reps=5)
        size = 1 << 30
        a, b = L.DeviceArray((size,), np.uint8), L.DeviceArray((size,), np.uint8)
        for _ in range(3):
            L.call("tmi_memcpy_d2d", b._p, a._p, size)
        log("pmc pass done: 5 SpMM launches, 3 copies of %d bytes" % size)
        return

    # reference fields: FP64 block solve at tight tolerance (untimed)
    fr, _ = make_solver("fp64_block", 1e-13, args.tile)
    dR = L.DeviceArray(B.shape, np.complex128)
    t0 = time.perf_counter()
    st_ref = fr.CalcFieldsDevice(nq, dB.ptr, dR.ptr)
    log("FP64 block reference solve %.2f s, iters/col %.0f, max relres %.2e" %
        (time.perf_counter() - t0, np.mean([s[0] for s in st_ref]), max(s[2] for s in st_ref)))
    Xref = dR.to_host()
    del fr

    # warmup + timed mixed-precision steps
    for w in range(args.warmup):
        fm.CalcFieldsDevice(nq, dB.ptr, dX.ptr)
        log("warmup %d: %.1f ms, iters/col %.0f" % (w, *fm.LastTiming()))
    L.synchronize()
    t0 = time.perf_counter()
    gpu_ms = []
    for k in range(args.steps):
        stats = fm.CalcFieldsDevice(nq, dB.ptr, dX.ptr)
        gpu_ms.append(fm.LastTiming()[0])
    L.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    launches, graph_launches, host_syncs = fm.LastCounts()   # of the last timed solve
    value = nq * args.steps / elapsed
    iters = [s[0] for s in stats]
    sweeps = [s[1] for s in stats]
    conv = all(s[3] for s in stats)
    log("mixed: %.1f ms/step (GPU events %.1f), %.1f solves/s, iters/col %.0f, sweeps %d-%d, converged %s" %
        (ms_per_step, np.mean(gpu_ms), value, np.mean(iters), min(sweeps), max(sweeps), conv))
    X = dX.to_host()
    err = np.linalg.norm(X - Xref, axis=0) / np.linalg.norm(Xref, axis=0)
    log("field error vs FP64 reference: max %.2e median %.2e" % (err.max(), np.median(err)))

    # dominant kernel: the block SpMM of the inner solve
    spmm_kernel = spmm_kernel_name(tuned["spmm_variant"] if tuned else (args.spmm_variant or "rowblock_mfma"), args.solver)
    spmm_ms, spmm_bytes = fm.TimeKernel(0, reps=20)
    res_ms, res_bytes = fm.TimeKernel(1, reps=10)
    achieved = spmm_bytes / (spmm_ms * 1e-3) / 1e9
    log("SpMM: %.3f ms/launch, %.1f MB algorithmic, %.0f GB/s (%.1f%% of %.0f)" %
        (spmm_ms, spmm_bytes / 1e6, achieved, 100 * achieved / HBM_PEAK_GBS, HBM_PEAK_GBS))
    log("FP64 residual SpMM: %.3f ms/launch, %.0f GB/s" % (res_ms, res_bytes / (res_ms * 1e-3) / 1e9))

    # plain FP64 single-RHS path on a subset of columns
    fp64 = None
    if not args.no_fp64 and args.fp64_cols > 0:
        fp64 = fp64_single_leg(args, make_solver, L, B, n, nq, Xref)

    # FP64 block Krylov over all columns at the same tolerance: the gain of
    # batching alone, beside which the mixed mode shows the gain of precision
    fp64b = None
    if not args.no_fp64:
        fb, _ = make_solver("fp64_block", args.tol, args.tile)
        dXb = L.DeviceArray(B.shape, np.complex128)
        fb.CalcFieldsDevice(nq, dB.ptr, dXb.ptr)    # warm
        L.synchronize()
        t0 = time.perf_counter()
        stb = fb.CalcFieldsDevice(nq, dB.ptr, dXb.ptr)
        L.synchronize()
        tb = time.perf_counter() - t0
        eb = np.linalg.norm(dXb.to_host() - Xref, axis=0) / np.linalg.norm(Xref, axis=0)
        fp64b = dict(solver=args.solver.lower(), solves_per_s=nq / tb, columns=nq, iters_per_col=float(np.mean([s[0] for s in stb])),
                     max_rel_err=float(eb.max()), converged=bool(all(s[3] for s in stb)))
        log("fp64 block: %.1f solves/s over %d columns, iters/col %.0f, err %.2e" %
            (nq / tb, nq, fp64b["iters_per_col"], eb.max()))
        del fb, dXb

    # the other complex Krylov method (BiCGSTAB / COCG), mixed mode, same
    # tolerance and SpMM variant: solves/s, SpMMs per column, field error
    other = None
    if not args.no_fp64:
        oname = "COCG" if args.solver == "BICGSTAB" else "BICGSTAB"
        fo, _ = make_solver("mixed", args.tol, tuned["tile"] if tuned else args.tile, solver=oname)
        if tuned:
            fo.SetSpmmVariant(tuned["spmm_variant"])
        dXo = L.DeviceArray(B.shape, np.complex128)
        fo.CalcFieldsDevice(nq, dB.ptr, dXo.ptr)    # warm (captures the graphs)
        L.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            sto = fo.CalcFieldsDevice(nq, dB.ptr, dXo.ptr)
        L.synchronize()
        to = (time.perf_counter() - t0) / args.steps
        eo = np.linalg.norm(dXo.to_host() - Xref, axis=0) / np.linalg.norm(Xref, axis=0)
        ito = float(np.mean([s[0] for s in sto]))
        other = dict(solver=oname.lower(), solves_per_s=nq / to, ms_per_step=to * 1e3, iters_per_col=ito,
                     spmm_per_col=ito * (2 if oname == "BICGSTAB" else 1), max_rel_err=float(eo.max()),
                     converged=bool(all(s[3] for s in sto)))
        log("mixed %s: %.1f solves/s, iters/col %.0f (%.0f SpMMs), err %.2e" %
            (oname, nq / to, ito, other["spmm_per_col"], eo.max()))
        del fo, dXo

    del fm
    mfreq = None
    if not args.no_multifreq:
        mfreq = multifreq_leg(args, T, L, mesh, sol, B)
    jac = None
    if not args.no_jacobian:
        jac = jacobian_leg(args)

    batch = serve = None
    if not args.no_serve:
        batch = multi_mesh_leg(args)
        serve = serving_leg(args)

    # CPU legs on the FP64 oracle: the error of the mixed fields against the
    # oracle's own solve, and the CPU baseline (one source per host thread)
    cpu, oerr = None, None
    if not args.no_cpu:
        from oracle import oracle as O
        t0 = time.perf_counter()
        om = O.Mesh(nodes=nodes, elems=el)
        cm, ck, c2 = O.coeffs(mua, mus, ref)
        system = (om,) + tuple(om.sysmat(cm, ck, c2, omega))
        log("cpu oracle setup (mesh, FP64 assembly) %.1f s" % (time.perf_counter() - t0))
        order = np.argsort(iters, kind="stable")
        cols = sorted({int(order[0]), int(order[1]), int(order[-2]), int(order[-1])})
        t0 = time.perf_counter()
        errs, orr = oracle_errors(system, B, X, cols)
        oerr = {"columns": cols, "iters": [int(iters[c]) for c in cols], "rel_l2": errs, "oracle_tol": 1e-12,
                "oracle_max_relres": float(orr)}
        log("field error vs FP64 oracle (columns %s: fastest and slowest, oracle tol 1e-12, %.1f s): %s"
            % (cols, time.perf_counter() - t0, " ".join("%.2e" % e for e in errs)))
        cpu = cpu_baseline(system, Q, args.tol, args.cpu_seconds,
                           int(fp64["iters_per_col"]) if fp64 else int(iters[0]))

    out = {
        "metric": "field solves/s & Jacobian s, mixed vs FP64 GPU at 1e-8 rel err; p50/p99 latency",
        "value": round(value, 3),
        "unit": "field solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex-fp32 Krylov + fp64 refinement (c64/c128)",
        "data": "synthetic (generated mesh, seed-drawn parameter maps)",
        "config": {"workload": "%s: %s" % (args.workload, desc), "nodes": n, "elements": int(len(el)),
                   "sources": nq, "tile": tuned["tile"] if tuned else args.tile, "tol": args.tol, "precon": "diag",
                   "solver": args.solver.lower(), "autotuned": tuned},
        "max_rel_err_vs_fp64": float(err.max()),
        "max_rel_err_vs_oracle": (max(oerr["rel_l2"]) if oerr else None),
        "oracle_check": oerr,
        "iters_per_col": float(np.mean(iters)),
        "spmm_per_col": float(np.mean(iters)) * (2 if args.solver == "BICGSTAB" else 1),
        "per_solve": {"kernel_launches": int(launches), "graph_launches": int(graph_launches),
                      "host_syncs": int(host_syncs)},
        "other_solver": other,
        "converged": bool(conv),
        "fp64_single": fp64,
        "speedup_vs_fp64_single": (round(value / fp64["solves_per_s"], 3) if fp64 else None),
        "fp64_block": fp64b,
        "speedup_vs_fp64_block": (round(value / fp64b["solves_per_s"], 3) if fp64b else None),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.workload, args.tile, spmm_kernel, args.solver),
                     "kernel": spmm_kernel,
                     "avg_launch_ms": round(spmm_ms, 4), "algorithmic_bytes": int(spmm_bytes)},
        "jacobian": jac,
        "multi_mesh": batch,
        "multi_frequency": mfreq,
        "serving": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def fp64_single_leg(args, make_solver, L, B, n, nq, Xref):
    """The plain FP64 single-right-hand-side GPU path (the reference's
    method, BiCGSTAB, one column at a time) on the first columns: solves/s,
    iterations, time per iteration, launches / graph launches / host
    synchronisations per iteration; the field error against Xref when given."""
    ff, _ = make_solver("fp64", args.tol, solver="BICGSTAB")   # the reference's method, plainly
    k = min(args.fp64_cols, nq)
    Bs = L.DeviceArray.from_host(np.ascontiguousarray(B[:, :k]))
    Xs = L.DeviceArray((n, k), np.complex128)
    B1 = L.DeviceArray.from_host(np.ascontiguousarray(B[:, :1]))
    X1 = L.DeviceArray((n, 1), np.complex128)
    ff.CalcFieldsDevice(1, B1.ptr, X1.ptr)      # warm the single-column kernels
    L.synchronize()
    t0 = time.perf_counter()
    st64 = ff.CalcFieldsDevice(k, Bs.ptr, Xs.ptr)
    L.synchronize()
    t64 = time.perf_counter() - t0
    c64 = ff.LastCounts()
    e64 = (np.linalg.norm(Xs.to_host() - Xref[:, :k], axis=0) / np.linalg.norm(Xref[:, :k], axis=0)
           if Xref is not None else np.full(k, np.nan))
    fp64 = dict(solver="bicgstab", solves_per_s=k / t64, columns=k, iters_per_col=float(np.mean([s[0] for s in st64])),
                max_rel_err=(float(e64.max()) if Xref is not None else None))
    ms_it = t64 * 1e3 / max(1.0, sum(s[0] for s in st64))
    fp64["ms_per_iteration"] = ms_it
    its64 = max(1.0, sum(s[0] for s in st64))
    fp64["launches_per_iteration"] = round(c64[0] / its64, 2)
    fp64["graph_launches_per_iteration"] = round(c64[1] / its64, 3)
    fp64["host_syncs_per_iteration"] = round(c64[2] / its64, 3)
    log("fp64 single-RHS: %.2f solves/s over %d columns, iters/col %.0f, %.3f ms per iteration, err %.2e" %
        (k / t64, k, fp64["iters_per_col"], ms_it, e64.max()))
    return fp64


def jacobian_leg(args):
    """Nodal Jacobian of log amplitude and phase w.r.t. cmua and ckappa on the
    ~103k-node cylinder, 32 x 32 source-detector pairs, 100 MHz: direct and
    adjoint fields from one 64-column mixed-precision block solve, linear
    projections, then GenerateJacobian on the GPU (J = 2048 x 2n doubles,
    resident in HBM), and the same Jacobian on a 64^3 raster basis.  Rows of
    two sources (nodal) and one source (raster) are checked against the FP64
    oracle's Jacobians computed from the same fields."""
    import ctypes as C

    from scipy import sparse

    import toastpp_amd as T
    from toastpp_amd import _lib as L
    from toastpp_amd import meshgen as G
    from oracle import oracle as O

    nodes, el, Q, M, mua, mus, ref, freq, desc = G.workload(args.jac_workload)
    n = len(nodes)
    log("jacobian workload", desc, "nodes", n, "elements", len(el))
    mesh = T.Mesh(data=(nodes, el))
    mesh.SetQM(Q, M)
    nq, nm, nqm = mesh.qm_dims()
    B = np.ascontiguousarray(np.hstack([mesh.Qvec(width=2.0).toarray(),
                                        mesh.Mvec(width=2.0, ref=ref).toarray()]).astype(np.complex128))
    dB = L.DeviceArray.from_host(B)
    dX = L.DeviceArray(B.shape, np.complex128)
    f = T.FwdSolver(mesh, args.solver, args.tol, complex=True)
    f.SetMode("mixed")
    f.SetPrecon("DIAG")
    f.SetTile(64)
    f.Allocate()
    f.Reset(T.Solution(mua, mus, ref), freq * 2 * np.pi * 1e-6)
    f.CalcFieldsDevice(nq + nm, dB.ptr, dX.ptr)            # warm
    L.synchronize()
    t0 = time.perf_counter()
    st = f.CalcFieldsDevice(nq + nm, dB.ptr, dX.ptr)
    L.synchronize()
    t_fields = time.perf_counter() - t0
    X = dX.to_host()
    mrows = sparse.csr_matrix(B[:, nq:].T)
    f.ProjectAllDevice(mrows, dX.ptr, nq + nm, scale="lin")
    t0 = time.perf_counter()
    proj = f.ProjectAllDevice(mrows, dX.ptr, nq + nm, scale="lin")
    t_proj = time.perf_counter() - t0
    dP = L.DeviceArray.from_host(proj)
    dJ = L.DeviceArray((2 * nqm, 2 * n), np.float64)
    h = mesh.handle_ptr()
    base = dX.ptr

    def run():
        L.call("tmi_jacobian_nodal_device", h, nq, nm, C.c_void_p(base), nq + nm, C.c_void_p(base + 16 * nq),
               nq + nm, C.c_void_p(dP.ptr), L.DATA_LOG, C.c_void_p(dJ.ptr))

    run()
    L.synchronize()
    ts = []
    for _ in range(max(1, args.jac_reps)):
        t0 = time.perf_counter()
        run()
        L.synchronize()
        ts.append(time.perf_counter() - t0)
    t_jac = min(ts)
    jbytes = 4.0 * nqm * n * 8 + n * (nq + nm) * 16
    log("jacobian: fields %.1f ms (64 columns, iters/col %.0f), projections %.1f ms, nodal J %.1f ms "
        "(%.0f GB/s of J + fields)" % (t_fields * 1e3, np.mean([s[0] for s in st]), t_proj * 1e3, t_jac * 1e3,
                                       jbytes / t_jac / 1e9))
    # parity: rows of sources 0 and nq-1 against the oracle on the same fields
    lptr = np.arange(nq + 1, dtype=np.int32) * nm
    rows = []
    for q in (0, nq - 1):
        rows += list(range(lptr[q], lptr[q + 1]))
    rows = np.array(rows)
    Jg = np.empty((2 * len(rows), 2 * n))
    for t, k in enumerate(rows):
        for blk in (0, 1):
            r = k + blk * nqm
            dst = Jg[t + blk * len(rows)]
            L.call("tmi_memcpy_d2h", dst.ctypes.data_as(C.c_void_p), C.c_void_p(dJ.ptr + int(r) * 2 * n * 8), 2 * n * 8)
    t0 = time.perf_counter()
    om = O.Mesh(nodes=nodes, elems=el)
    dsub = np.ascontiguousarray(X[:, [0, nq - 1]].T)
    asub = np.ascontiguousarray(X[:, nq:].T)
    lp2 = np.array([0, nm, 2 * nm], np.int32)
    li2 = np.tile(np.arange(nm, dtype=np.int32), 2)
    Jo = O.jacobian_nodal(om, dsub, asub, lp2, li2, proj[rows], logscale=True)
    k2 = len(rows)
    worst = 0.0
    for r in (slice(0, k2), slice(k2, 2 * k2)):
        for c in (slice(0, n), slice(n, 2 * n)):
            worst = max(worst, np.max(np.abs(Jg[r, c] - Jo[r, c])) / np.max(np.abs(Jo[r, c])))
    log("jacobian parity (%d rows vs FP64 oracle, %.1f s): worst block-relative diff %.2e" %
        (2 * k2, time.perf_counter() - t0, worst))
    del dJ
    out = {"workload": "%s: %s" % (args.jac_workload, desc), "nodes": n, "measurements": nqm,
           "fields_s": round(t_fields, 4), "projections_s": round(t_proj, 4), "nodal_jacobian_s": round(t_jac, 4),
           "total_s": round(t_fields + t_proj + t_jac, 4), "jacobian_GBps": round(jbytes / t_jac / 1e9, 1),
           "parity_rows": int(2 * k2), "max_block_rel_diff_vs_oracle": float(worst)}

    # raster basis: 64^3 Raster_Pixel over the mesh bounding box
    g = args.raster
    t0 = time.perf_counter()
    R = T.Raster(mesh, np.array([g, g, g]))
    t_rb = time.perf_counter() - t0
    slen = R.SLen()
    dJr = L.DeviceArray((2 * nqm, 2 * slen), np.float64)
    hr = R.handle_ptr()

    def run_r():
        L.call("tmi_jacobian_raster_device", hr, nq, nm, C.c_void_p(base), nq + nm, C.c_void_p(base + 16 * nq),
               nq + nm, C.c_void_p(dP.ptr), L.DATA_LOG, C.c_void_p(dJr.ptr))

    run_r()
    L.synchronize()
    ts = []
    for _ in range(max(1, args.jac_reps)):
        t0 = time.perf_counter()
        run_r()
        L.synchronize()
        ts.append(time.perf_counter() - t0)
    t_jr = min(ts)
    rbytes = 4.0 * nqm * slen * 8
    log("raster %d^3: build %.1f ms, %d solution voxels; grid J %.1f ms (%.0f GB/s of J)" %
        (g, t_rb * 1e3, slen, t_jr * 1e3, rbytes / t_jr / 1e9))
    # parity: the rows of source 0 against the oracle's grid Jacobian
    rows0 = np.arange(nm)
    Jgr = np.empty((2 * nm, 2 * slen))
    for t, k in enumerate(rows0):
        for blk in (0, 1):
            L.call("tmi_memcpy_d2h", Jgr[t + blk * nm].ctypes.data_as(C.c_void_p),
                   C.c_void_p(dJr.ptr + int(k + blk * nqm) * 2 * slen * 8), 2 * slen * 8)
    t0 = time.perf_counter()
    orr = O.Raster(om, [g, g, g])
    Jor = O.jacobian_grid(orr, np.ascontiguousarray(X[:, :1].T), asub, np.array([0, nm], np.int32),
                          np.arange(nm, dtype=np.int32), proj[rows0], logscale=True)
    wr = np.max(np.abs(Jgr - Jor)) / np.max(np.abs(Jor))
    log("raster Jacobian parity (%d rows vs FP64 oracle, %.1f s): max diff / max |J| %.2e" %
        (2 * nm, time.perf_counter() - t0, wr))
    del dJr, f
    out.update({"raster": "%d^3" % g, "raster_solution_voxels": int(slen), "raster_build_s": round(t_rb, 4),
                "raster_jacobian_s": round(t_jr, 4), "raster_jacobian_GBps": round(rbytes / t_jr / 1e9, 1),
                "raster_parity_rows": int(2 * nm), "raster_max_rel_diff_vs_oracle": float(wr)})
    return out


def _prepare(T, L, nodes, el, Q, mua, mus, ref, freqs, tol, solver="COCG"):
    """one device mesh, its source block (complex and real) and a prepared
    solver per frequency (`solver` for complex systems, PCG for real CW ones)"""
    mesh = T.Mesh(data=(nodes, el))
    mesh.SetQM(Q, Q[:1])
    B = mesh.Qvec(width=2.0).toarray()
    dBc = L.DeviceArray.from_host(np.ascontiguousarray(B))
    dBr = L.DeviceArray.from_host(np.ascontiguousarray(B.real))
    solvers = {}
    for freq in freqs:
        cplx = freq != 0
        f = T.FwdSolver(mesh, solver if cplx else "CG", tol, complex=cplx)
        f.SetPrecon("DIAG")
        f.Allocate()
        f.Reset(T.Solution(mua, mus, ref), freq * 2 * np.pi * 1e-6)
        solvers[freq] = (f, dBc if cplx else dBr)
    return mesh, solvers


def multi_mesh_leg(args):
    """Three ellipsoids (20k, 80k, 250k nodes, 64 sources each, 100 MHz):
    one after another, then as three concurrent requests of the serving
    loop (one stream each)."""
    import ctypes as C

    import toastpp_amd as T
    from toastpp_amd import _lib as L
    from toastpp_amd import meshgen as G
    from toastpp_amd.serve import Server

    meshes = []
    for name, nodes, el, Q, mua, mus, ref in G.ellipsoid_batch():
        mesh, solvers = _prepare(T, L, nodes, el, Q, mua, mus, ref, [100.0], args.tol, args.solver)
        f, dB = solvers[100.0]
        meshes.append((mesh, f, dB, name))
    outs = [L.DeviceArray((dB.shape[0], 64), np.complex128) for _, _, dB, _ in meshes]
    srv = Server(slots=3, hbm_budget_gib=200.0)

    def run(concurrent):
        ids = []
        for dX in outs:
            L.call("tmi_memcpy_h2d", dX._p, np.zeros(dX.shape, np.complex128).ctypes.data_as(C.c_void_p), dX.nbytes)
        t0 = time.perf_counter()
        for (mesh, f, dB, name), dX in zip(meshes, outs):
            ids.append(srv.submit(f, 64, dB.ptr, dX.ptr, 64))
            if not concurrent:
                srv.wait(ids[-1])
        if concurrent:
            for r in ids:
                srv.wait(r)
        return time.perf_counter() - t0

    run(True)   # warm
    t_seq = run(False)
    t_con = run(True)
    srv.close()
    cols = 3 * 64
    log("multi-mesh 20k+80k+250k x 64 sources: one after another %.1f ms, concurrent %.1f ms (%.2fx), %.0f solves/s"
        % (t_seq * 1e3, t_con * 1e3, t_seq / t_con, cols / t_con))
    return {"meshes": [m[3] for m in meshes], "sources_each": 64, "sequential_s": round(t_seq, 4),
            "concurrent_s": round(t_con, 4), "speedup": round(t_seq / t_con, 3),
            "field_solves_per_s": round(cols / t_con, 2)}


def multifreq_leg(args, T, L, mesh, sol, B):
    """Several frequencies of one mesh in one block solve: 16 of the
    workload's sources at each of 0, 50, 100 and 200 MHz as the 64 columns
    of one mixed COCG solve, against four solves of 16 columns, each after
    a Reset at its frequency.  Reports both times and the largest field
    difference between them."""
    freqs = np.array([0.0, 50.0, 100.0, 200.0])
    w = freqs * 2 * np.pi * 1e-6
    n = B.shape[0]
    cols = np.arange(64) % 16 * (B.shape[1] // 16)
    om = w[np.arange(64) // 16]
    Bm = np.ascontiguousarray(B[:, cols])
    dB = L.DeviceArray.from_host(Bm)
    dX = L.DeviceArray(Bm.shape, np.complex128)

    def solver():
        f = T.FwdSolver(mesh, "COCG", args.tol, complex=True)
        f.SetMode("mixed")
        f.SetPrecon("DIAG")
        f.Allocate()
        f.Reset(sol, w[2])
        return f

    f = solver()
    f.CalcFieldsFreqsDevice(64, dB.ptr, dX.ptr, om)   # warm: builds the omega = 1 operator
    L.synchronize()
    t0 = time.perf_counter()
    st = f.CalcFieldsFreqsDevice(64, dB.ptr, dX.ptr, om)
    L.synchronize()
    t_one = time.perf_counter() - t0
    X1 = dX.to_host()
    conv = all(s[3] for s in st)
    del f
    g = solver()
    parts = [L.DeviceArray.from_host(np.ascontiguousarray(Bm[:, 16 * i:16 * i + 16])) for i in range(4)]
    outs = [L.DeviceArray((n, 16), np.complex128) for _ in range(4)]
    for i in range(4):   # warm
        g.Reset(sol, w[i])
        g.CalcFieldsDevice(16, parts[i].ptr, outs[i].ptr)
    L.synchronize()
    t0 = time.perf_counter()
    for i in range(4):
        g.Reset(sol, w[i])
        g.CalcFieldsDevice(16, parts[i].ptr, outs[i].ptr)
    L.synchronize()
    t_four = time.perf_counter() - t0
    X4 = np.hstack([o.to_host() for o in outs])
    diff = float((np.linalg.norm(X1 - X4, axis=0) / np.linalg.norm(X4, axis=0)).max())
    log("multi-frequency (16 sources x 0/50/100/200 MHz): one 64-column solve %.1f ms, four 16-column solves "
        "(with their Resets) %.1f ms (%.2fx); max field difference %.1e" % (t_one * 1e3, t_four * 1e3,
                                                                            t_four / t_one, diff))
    return {"frequencies_MHz": freqs.tolist(), "sources_each": 16, "one_solve_s": round(t_one, 4),
            "separate_solves_s": round(t_four, 4), "speedup": round(t_four / t_one, 3),
            "max_rel_field_difference": diff, "converged": bool(conv)}


def serving_leg(args):
    """Poisson arrivals (seeded) of CalcFields requests: meshes from 3k-node
    disks to 250k-node ellipsoids, frequency in {0, 50, 100, 200} MHz, 8-64
    sources, tol as the main leg; matrices resident, continuous batching in
    the serving loop.  Latency = finish - arrival."""
    import ctypes as C

    import toastpp_amd as T
    from toastpp_amd import _lib as L
    from toastpp_amd import meshgen as G
    from toastpp_amd.serve import Server, latency_stats

    pool = G.serving_pool()
    freqs = [0.0, 50.0, 100.0, 200.0]
    prepared, keep = {}, []
    t0 = time.perf_counter()
    for mi, (name, nodes, el, Q, mua, mus, ref) in enumerate(pool):
        mesh, solvers = _prepare(T, L, nodes, el, Q, mua, mus, ref, freqs, args.tol, args.solver)
        keep.append(mesh)
        for f in freqs:
            prepared[(mi, f)] = solvers[f]
    log("serving: %d meshes x %d frequencies prepared in %.1f s" % (len(pool), len(freqs), time.perf_counter() - t0))
    rng = np.random.default_rng(4)
    t, arrivals = 0.0, []
    while True:
        t += rng.exponential(1.0 / args.serve_rate)
        if t > args.serve_seconds:
            break
        arrivals.append((t, int(rng.integers(len(pool))), freqs[int(rng.integers(len(freqs)))],
                         int(rng.integers(8, 65))))
    srv = Server(slots=args.serve_slots, hbm_budget_gib=200.0)
    reqs = []
    start = time.perf_counter()
    for (ta, mi, f, nq) in arrivals:
        solver, dB = prepared[(mi, f)]
        n = dB.shape[0]
        dX = L.DeviceArray((n, 64), np.complex128 if solver.complex else np.float64)
        L.call("tmi_memcpy_h2d", dX._p, np.zeros(dX.shape, dX.dtype).ctypes.data_as(C.c_void_p), dX.nbytes)
        delay = ta - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        reqs.append((srv.submit(solver, nq, dB.ptr, dX.ptr, 64), dX, nq, mi, f))
    times, worst, cols, notconv = [], 0.0, 0, 0
    for rid, dX, nq, mi, f in reqs:
        st, tm, status = srv.wait(rid)
        times.append(tm)
        worst = max(worst, max(x[2] for x in st))
        notconv += sum(1 for x in st if not x[3])
        cols += nq
    makespan = max(tm[2] for tm in times) - min(tm[0] for tm in times)
    srv.close()
    lat = latency_stats(times)
    svc = np.array([tm[2] - tm[1] for tm in times])
    out = {"requests": len(reqs), "arrival_rate_per_s": args.serve_rate, "slots": args.serve_slots,
           "latency_p50_ms": round(lat["p50_ms"], 2), "latency_p99_ms": round(lat["p99_ms"], 2),
           "latency_mean_ms": round(lat["mean_ms"], 2), "service_p50_ms": round(float(np.median(svc)) * 1e3, 2),
           "throughput_requests_per_s": round(len(reqs) / makespan, 2),
           "field_solves_per_s": round(cols / makespan, 1), "max_relres": worst, "not_converged": notconv,
           "meshes": [p[0] for p in pool], "freqs_MHz": freqs}
    log("serving: %d requests at %.0f/s: latency p50 %.1f ms p99 %.1f ms, %.1f req/s, %.0f solves/s, max relres %.1e"
        % (len(reqs), args.serve_rate, lat["p50_ms"], lat["p99_ms"], out["throughput_requests_per_s"],
           out["field_solves_per_s"], worst))
    return out


def host_info():
    """CPU model and core counts of the host the bench runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "affinity_cores": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}


# host threads of the CPU legs: the cores this process may use, at most the
# 16-thread CPU share one GPU box gives a process
CPU_THREAD_CAP = 16


def cpu_baseline(system, Q, tol, budget_s, gpu_iters):
    """FP64 oracle (plain C restatement of the reference path): BiCGSTAB +
    diagonal preconditioner, one source per host thread, as the reference's
    THREAD_LEVEL 2 runs CalcFields (fwdsolver.cc:619-643); the oracle's C
    solver runs outside the GIL.  If the solves would exceed the time
    budget, they are cut at the budget and the wall time per iteration is
    scaled to the iteration count the FP64 solve needs."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    om, rp, ci, val = system
    info = host_info()
    nthr = max(1, min(CPU_THREAD_CAP, info["affinity_cores"], len(Q)))
    qs = om.qvec_gaussian(Q[:nthr], 2.0).astype(np.complex128)
    t1 = time.perf_counter()
    _, it0, _ = O.bicgstab(rp, ci, val, qs[0], tol, 5, True)
    per_it = (time.perf_counter() - t1) / max(1, it0 + 1)
    maxit = max(10, int(budget_s / per_it))
    t1 = time.perf_counter()
    with ThreadPoolExecutor(nthr) as ex:
        res = list(ex.map(lambda q: O.bicgstab(rp, ci, val, q, tol, maxit, True), qs))
    wall = time.perf_counter() - t1
    done = all(r[2] <= tol for r in res)
    its = [r[1] for r in res]
    if done:
        t_col = wall / nthr
        sample = ("%d of 64 columns, one per thread, solved to tol %.0e in %d-%d iterations (%.1f s wall)"
                  % (nthr, tol, min(its), max(its), wall))
    else:
        t_col = wall / max(1, max(its) + 1) * gpu_iters / nthr
        sample = ("%d columns, one per thread, %d BiCGSTAB iterations each timed (%.1f s wall), scaled to "
                  "the %d iterations the solve needs" % (nthr, max(its) + 1, wall, gpu_iters))
    log("cpu oracle: %s -> %.3f s per field solve (%d threads; host %s, %d cores usable, %d online)"
        % (sample, t_col, nthr, info["cpu_model"], info["affinity_cores"], info["os_cpu_count"]))
    return {"value": round(1.0 / t_col, 5), "unit": "field solves/s", "cores": nthr, "kind": "port",
            "sample": sample, "threads": nthr, "host_cores": info["affinity_cores"],
            "host_cpu_count": info["os_cpu_count"], "cpu_model": info["cpu_model"]}


def oracle_errors(system, B, X, cols, tol=1e-12):
    """Relative L2 error of GPU field columns against the FP64 oracle's
    BiCGSTAB (diagonal preconditioner) at tol, one column per host thread."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    om, rp, ci, val = system
    nthr = max(1, min(CPU_THREAD_CAP, len(os.sched_getaffinity(0)), len(cols)))
    with ThreadPoolExecutor(nthr) as ex:
        res = list(ex.map(lambda c: O.bicgstab(rp, ci, val, np.ascontiguousarray(B[:, c]), tol, 50000, True), cols))
    errs = [float(np.linalg.norm(X[:, c] - r[0]) / np.linalg.norm(r[0])) for c, r in zip(cols, res)]
    return errs, max(r[2] for r in res)


if __name__ == "__main__":
    main()
