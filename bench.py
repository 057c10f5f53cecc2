"""GeMSLR-FGMRES solve benchmark on one MI355X.

Workload (default): 3D 7-point Laplacian 128^3 (n = 2,097,152), GeMSLR with
3 levels, 64 subdomains per level, ILUT drop 1e-3 / fill 20, rank-16
thick-restart Arnoldi low-rank corrections, inner Schur iteration off;
FGMRES(50), tolerance 1e-8 on ||b - A x|| / ||b||; x* ~ U[-1, 1] (seed c per
column c), b = A x*.

A step is one batched solve of --nrhs right-hand sides with the mixed form
(preconditioner apply in FP32 with FP32 accumulation; residual SpMV, delayed
CGS2 orthogonalization, Givens and the solution update in FP64).  The plain FP64 form (single RHS,
FP64 apply) is timed beside it.  Setup runs once, outside the timed region;
inputs are resident in HBM when timing starts.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CONFIGS = {
    # 3D Laplacian 128^3, GeMSLR 3 levels / 64 subdomains / ILUT 1e-3,20 / rank 16: the headline workload
    "c2": dict(grid=(128, 128, 128), nlev=3, ncomp=64, droptol=1e-3, rownnz=20, rank=16, kdim=50, tol=1e-8),
    # small case for quick runs
    "c1": dict(grid=(64, 64, 1), nlev=2, ncomp=4, droptol=1e-2, rownnz=10, rank=0, kdim=30, tol=1e-8),
    "m64": dict(grid=(64, 64, 64), nlev=3, ncomp=64, droptol=1e-3, rownnz=20, rank=16, kdim=50, tol=1e-8),
    # shifted indefinite 128^3 Laplacian (sigma = 0.3), b ~ N(0, 1), rank 32, tol 1e-10
    "c3": dict(grid=(128, 128, 128), nlev=3, ncomp=64, droptol=1e-3, rownnz=20, rank=32, kdim=50, tol=1e-10,
               shift=0.3, rhs="normal"),
    # upwind convection-diffusion 96^3 (c = 0.3), b ~ N(0, 1)
    "c4": dict(grid=(96, 96, 96), nlev=3, ncomp=64, droptol=1e-3, rownnz=20, rank=16, kdim=50, tol=1e-8,
               convdiff=0.3, rhs="normal"),
}
HBM_PEAK_GBS = 8000.0
# HBM bytes read + written per level-0 ring-solve launch of the 128^3 operator
# (FP32 apply), by right-hand sides per launch, measured with rocprofv3
# FETCH_SIZE and WRITE_SIZE passes over `bench.py --probe --nrhs N`, each
# scaled by what the counter reads for a 1 GiB device copy (8-column groups: of
# 128 threads on whole steps at 32 columns, one wave on half steps at 64)
TRAFFIC_LEVEL0 = {32: 1667.6e6, 64: 2742.2e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--nrhs", type=int, default=64, help="right-hand sides per batched solve")
    ap.add_argument("--single-solves", type=int, default=5, help="timed single-RHS solves per form")
    ap.add_argument("--fp64-batch-steps", type=int, default=3, help="timed batched solves in the plain FP64 form")
    ap.add_argument("--host-steps", type=int, default=3,
                    help="timed batched solves handed host arrays (PCIe-inclusive rate, not the value)")
    ap.add_argument("--profile-steps", type=int, default=2,
                    help="batched solves with the level-0 solve timed per launch (roofline), after the timed steps")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU time budget of each oracle baseline")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the OpenMP-over-blocks baseline (0: the host's CPUs, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the short serving leg of the default run")
    ap.add_argument("--serve-requests", type=int, default=12, help="requests of the default run's serving leg")
    ap.add_argument("--serve-rate", type=float, default=3.0, help="arrivals per second of that leg")
    ap.add_argument("--serve", action="store_true",
                    help="run the serving stream (resident operators, Poisson arrivals) instead")
    ap.add_argument("--requests", type=int, default=32, help="requests of the serving stream")
    ap.add_argument("--rate", type=float, default=2.0, help="mean request arrivals per second")
    ap.add_argument("--graphs", type=int, default=1, help="serving: replay captured FGMRES steps")
    ap.add_argument("--orth", default="dcgs2", choices=["mgs", "cgs2", "dcgs2"],
                    help="orthogonalization of the GPU solves (the CPU baseline uses MGS, as the reference)")
    ap.add_argument("--schur", default="",
                    help="inner Schur iteration 'maxits,tol' (e.g. 5,1e-2, the reference's default); off if empty")
    ap.add_argument("--probe", action="store_true",
                    help="counter probe: --nrhs FP32 applies of the 128^3 operator and a 1 GiB copy")
    ap.add_argument("--probe-fgmres", action="store_true",
                    help="counter workload: the calibration copy, then one batched mixed FGMRES solve")
    ap.add_argument("--trsv-sweep", default="",
                    help="comma-separated column counts: level-set solve times per launch (FP32 and FP64 applies"
                         " of the 128^3 operator, the launch's own geometry or the one PGS_* variables force)")
    ap.add_argument("--spmv-probe", action="store_true",
                    help="one-column SpMV of the 128^3 Laplacian with 8, 4, 2 lanes per row and one thread per row,"
                         " and 32-column products with 4 and 8 columns per thread (for a kernel trace; each variant"
                         " is its own kernel instantiation)")
    ap.add_argument("--restart", default="",
                    help="restart path: save the --config operator to this file, load it, upload it and compare"
                         " the time with the setup")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is not built",
              file=sys.stderr)
        sys.exit(2)
    if args.serve:
        serve(args)
        return
    if args.probe or args.probe_fgmres:
        probe(args)
        return
    if args.spmv_probe:
        spmv_probe(args)
        return
    if args.trsv_sweep:
        trsv_sweep(args)
        return
    if args.restart:
        restart(args)
        return

    import pargemslr_amd as pg

    cfg = CONFIGS[args.config]
    dev = 0  # HIP device 0 through the library's own runtime
    nr = args.nrhs
    t0 = time.time()
    if "convdiff" in cfg:
        A = pg.convdiff(*cfg["grid"], cfg["convdiff"])
    else:
        A = pg.laplacian(*cfg["grid"], shift=cfg.get("shift", 0.0))
    log(f"[bench] matrix n={A.n} nnz={A.nnz} ({time.time() - t0:.1f}s)")
    schur = None
    if args.schur:
        k, t = args.schur.split(",")
        schur = (int(k), float(t))
    params = pg.config_params(cfg["nlev"], cfg["ncomp"], cfg["droptol"], cfg["rownnz"], cfg["rank"], kdim=cfg["kdim"],
                              schur=schur)
    G = pg.Gemslr(A, params, seed=0)
    inf = G.info()
    log(f"[bench] setup {inf.setup_seconds[3]:.1f}s (partition {inf.setup_seconds[0]:.1f}, ilut {inf.setup_seconds[1]:.1f},"
        f" lowrank {inf.setup_seconds[2]:.1f}); nlev={inf.nlev} nnz_ilu={inf.nnz_ilu} nnz_lr={inf.nnz_lr}")
    levels = G.levels()
    # the same setup with the low-rank corrections built on the device
    # (pgs_gemslr_setup_device), timed beside the host build; its ranks and
    # bases are compared with the host build's (the solves below use the
    # host-built operator)
    Gd = pg.Gemslr(A, params, seed=0, lowrank_device=0)
    infd = Gd.info()
    lr_dist, ranks_equal = 0.0, True
    for lh, ld in zip(levels, Gd.levels()):
        ranks_equal = ranks_equal and lh["rank"] == ld["rank"]
        if lh["rank"] and ld["rank"]:
            Wh, Wd = lh["W"], ld["W"]
            lr_dist = max(lr_dist, float(np.linalg.norm(Wd - Wh @ (Wh.conj().T @ Wd))))
    del Gd
    setup_lr = {"host_lowrank_s": round(inf.setup_seconds[2], 3), "device_lowrank_s": round(infd.setup_seconds[2], 3),
                "device_setup_s": round(infd.setup_seconds[3], 2), "ranks_equal": ranks_equal,
                "max_subspace_dist": lr_dist}
    log(f"[bench] low-rank build on the device {infd.setup_seconds[2]:.2f}s (host {inf.setup_seconds[2]:.2f}s), "
        f"ranks equal {ranks_equal}, subspace distance {lr_dist:.1e}")
    D = pg.DevOp(G, prec_mask=pg.PGS_PREC_FP64 | pg.PGS_PREC_FP32)
    W1 = pg.Workspace(D, max_nrhs=1, kdim=cfg["kdim"])
    WB = pg.Workspace(D, max_nrhs=nr, kdim=cfg["kdim"])
    log(f"[bench] device operator {D.bytes() / 1e9:.2f} GB")

    n = A.n
    if cfg.get("rhs") == "normal":
        Xs = None
        Bh = np.stack([pg.rng_normal(c, n) for c in range(nr)], axis=1)
    else:
        Xs = np.stack([pg.rng_uniform(c, n) for c in range(nr)], axis=1)
        Bh = A.matvec(Xs)
    B = pg.DeviceArray.from_numpy(Bh, device=dev)
    b1 = pg.DeviceArray.from_numpy(np.ascontiguousarray(Bh[:, 0]), device=dev)
    x1 = pg.DeviceArray(n, device=dev)
    XB = pg.DeviceArray((n, nr), device=dev)
    kd, tol = cfg["kdim"], cfg["tol"]
    orth = {"mgs": pg.PGS_ORTH_MGS, "cgs2": pg.PGS_ORTH_CGS2, "dcgs2": pg.PGS_ORTH_DCGS2}[args.orth]

    def solve1(prec):
        x1.zero()
        return W1.fgmres(b1, x1, kdim=kd, maxits=1000, tol=tol, prec=prec, orth=orth)[0]

    def solveB(prec=pg.PGS_PREC_FP32, graph=False):
        XB.zero()
        return WB.fgmres(B, XB, nrhs=nr, kdim=kd, maxits=1000, tol=tol, prec=prec, orth=orth, use_graph=graph)

    # plain FP64 form, single RHS
    r64 = solve1(pg.PGS_PREC_FP64)
    pg.synchronize(dev)
    lat64 = []
    for _ in range(args.single_solves):
        pg.synchronize(dev)
        ts = time.perf_counter()
        r64 = solve1(pg.PGS_PREC_FP64)
        pg.synchronize(dev)
        lat64.append(time.perf_counter() - ts)
    x64 = x1.numpy()
    res64 = np.linalg.norm(Bh[:, 0] - A.matvec(x64)) / np.linalg.norm(Bh[:, 0])
    log(f"[bench] fp64 single: status {r64[0]} iters {r64[1]} true rel res {res64:.2e} "
        f"median {np.median(lat64) * 1e3:.1f} ms")

    # mixed form, single RHS
    r32 = solve1(pg.PGS_PREC_FP32)
    pg.synchronize(dev)
    lat32 = []
    for _ in range(args.single_solves):
        pg.synchronize(dev)
        ts = time.perf_counter()
        r32 = solve1(pg.PGS_PREC_FP32)
        pg.synchronize(dev)
        lat32.append(time.perf_counter() - ts)
    x32 = x1.numpy()
    res32 = np.linalg.norm(Bh[:, 0] - A.matvec(x32)) / np.linalg.norm(Bh[:, 0])
    log(f"[bench] mixed single: status {r32[0]} iters {r32[1]} true rel res {res32:.2e} "
        f"median {np.median(lat32) * 1e3:.1f} ms")

    # headline: mixed, batched, direct launches, no per-launch timing
    for _ in range(args.warmup):
        solveB()
    pg.synchronize(dev)
    step_t = []
    pg.synchronize(dev)
    t_start = time.perf_counter()
    resB = None
    for _ in range(args.steps):
        ts = time.perf_counter()
        resB = solveB()
        pg.synchronize(dev)
        step_t.append(time.perf_counter() - ts)
    pg.synchronize(dev)
    T = time.perf_counter() - t_start
    XBh = XB.numpy()

    # the same steps replaying HIP graphs of the inner FGMRES steps (captured
    # in the first, untimed solve)
    solveB(graph=True)
    pg.synchronize(dev)
    graph_t = []
    for _ in range(args.steps):
        ts = time.perf_counter()
        resG = solveB(graph=True)
        pg.synchronize(dev)
        graph_t.append(time.perf_counter() - ts)
    graph_same = [r[:2] for r in resG] == [r[:2] for r in resB] and bool(np.array_equal(XB.numpy(), XBh))
    log(f"[bench] mixed batched with graph replay: {nr * len(graph_t) / sum(graph_t):.2f} solves/s, "
        f"same iterations and bits as direct launches: {graph_same}")

    # plain FP64 form, batched (the same columns, FP64 apply): splits the
    # speedup into batching and precision
    solveB(prec=pg.PGS_PREC_FP64)
    pg.synchronize(dev)
    f64b_t = []
    for _ in range(args.fp64_batch_steps):
        ts = time.perf_counter()
        res64B = solveB(prec=pg.PGS_PREC_FP64)
        pg.synchronize(dev)
        f64b_t.append(time.perf_counter() - ts)
    X64B = XB.numpy()
    true64B = max(float(np.linalg.norm(Bh[:, c] - A.matvec(X64B[:, c])) / np.linalg.norm(Bh[:, c])) for c in range(nr))
    ok64B = all(r[0] == 0 for r in res64B) and true64B <= tol
    fp64_batched = nr * len(f64b_t) / sum(f64b_t)
    log(f"[bench] fp64 batched nrhs={nr}: {fp64_batched:.2f} solves/s, iters {min(r[1] for r in res64B)}-"
        f"{max(r[1] for r in res64B)}, max true rel res {true64B:.2e}, ok={ok64B}")

    # the mixed batched solve handed host arrays, as a caller with host
    # memory sees it: right-hand sides copied in and solutions copied out
    # (pageable numpy memory) inside the timed region.  Reported beside the
    # value, which keeps its inputs resident in HBM.
    host_t, h2d_t, d2h_t = [], [], []
    for _ in range(args.host_steps):
        pg.synchronize(dev)
        ts = time.perf_counter()
        B.upload(Bh)
        pg.synchronize(dev)
        t1 = time.perf_counter()
        solveB()
        pg.synchronize(dev)
        t2 = time.perf_counter()
        XB.numpy()
        pg.synchronize(dev)
        t3 = time.perf_counter()
        host_t.append(t3 - ts)
        h2d_t.append(t1 - ts)
        d2h_t.append(t3 - t2)
    pcie = None
    if host_t:
        pcie = {"solves_per_s": round(nr * len(host_t) / sum(host_t), 3), "steps": len(host_t),
                "h2d_ms": round(1e3 * float(np.median(h2d_t)), 1), "d2h_ms": round(1e3 * float(np.median(d2h_t)), 1),
                "bytes_each_way": int(Bh.nbytes), "host_memory": "pageable"}
        log(f"[bench] mixed batched with host arrays: {pcie['solves_per_s']:.2f} solves/s "
            f"(copy in {pcie['h2d_ms']:.1f} ms, out {pcie['d2h_ms']:.1f} ms for {Bh.nbytes / 1e9:.2f} GB each way)")

    # per-launch timing (HIP events on the launch stream) of every kernel
    # family over separate mixed batched solves: the roofline figures
    WB.profile(True)
    for _ in range(args.profile_steps):
        solveB()
    pg.synchronize(dev)
    prof = WB.profile_read()
    fams = WB.profile_families()
    WB.profile(False)
    tot_ms = sum(f["total_ms"] for f in fams.values())
    kernel_table = {}
    log(f"[bench] kernel families over {args.profile_steps} batched solves (direct launches, events per launch):")
    for name, f in fams.items():
        if f["launches"] == 0:
            continue
        gbs = f["total_bytes"] / (f["total_ms"] * 1e-3) / 1e9 if f["total_ms"] > 0 and f["total_bytes"] > 0 else None
        kernel_table[name] = {"launches": f["launches"], "ms": round(f["total_ms"], 3),
                              "share": round(f["total_ms"] / tot_ms, 4) if tot_ms > 0 else None,
                              "gb": round(f["total_bytes"] / 1e9, 3),
                              "gbs": round(gbs, 1) if gbs else None,
                              "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
        log(f"[bench]   {name:16s} {f['launches']:7d} launches {f['total_ms']:9.2f} ms "
            f"({100 * f['total_ms'] / tot_ms:5.1f} %)  {f['total_bytes'] / 1e9:8.2f} GB  "
            f"{(gbs or 0):7.0f} GB/s  {(gbs or 0) / HBM_PEAK_GBS:5.3f} of peak")
    true_res = [float(np.linalg.norm(Bh[:, c] - A.matvec(XBh[:, c])) / np.linalg.norm(Bh[:, c])) for c in range(nr)]
    err_vs_fp64 = float(np.linalg.norm(XBh[:, 0] - x64) / np.linalg.norm(x64))
    err_vs_xstar = float(np.linalg.norm(XBh - Xs) / np.linalg.norm(Xs)) if Xs is not None else None
    iters = [r[1] for r in resB]
    ok = all(r[0] == 0 for r in resB) and max(true_res) <= tol
    solves_per_s = nr * args.steps / T
    fp64_solves_per_s = 1.0 / float(np.median(lat64))
    log(f"[bench] mixed batched nrhs={nr}: {solves_per_s:.2f} solves/s, iters {min(iters)}-{max(iters)}, "
        f"max true rel res {max(true_res):.2e}, ok={ok}")

    ach = prof["total_bytes"] / (prof["total_ms"] * 1e-3) / 1e9 if prof["total_ms"] > 0 else 0.0
    # HBM bytes per level-0 launch of this workload (summary under profiles/)
    traffic = TRAFFIC_LEVEL0.get(nr) if args.config == "c2" else None
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_trsv_ring (level-0 block ILU solve, L+U sweeps)",
                "launches": prof["launches"],
                "avg_launch_ms": round(prof["total_ms"] / max(1, prof["launches"]), 4),
                "bytes_per_launch": round(prof["total_bytes"] / max(1, prof["launches"]))}

    cpu = None
    if not args.no_cpu:
        # the oracle (FP64, modified Gram-Schmidt as the reference's
        # FlexGmresClass) solving the same columns to the same tolerance, whole
        # solves until about --cpu-seconds have been timed: once on one thread,
        # once with the diagonal blocks' ILU solves spread over OpenMP threads
        # (as the reference's SolveB, gemslr.cpp:3653-3655; the results are
        # the same bits)
        from oracle import oracle as orc
        O = orc.GemslrOracle(n, G.perm(), levels)
        host_cpus = os.cpu_count() or 1
        nthreads = args.cpu_threads if args.cpu_threads > 0 else min(16, host_cpus)

        def cpu_leg(threads):
            orc.set_threads(threads)
            t_cpu, its_cpu, solves = 0.0, [], 0
            while solves < nr and (solves == 0 or t_cpu < args.cpu_seconds):
                ts = time.perf_counter()
                _, st_c, it_c, _, _ = orc.fgmres(A, Bh[:, solves], prec=O, kdim=kd, maxits=1000, tol=tol, orth=0)
                t_cpu += time.perf_counter() - ts
                its_cpu.append(it_c)
                solves += 1
            orc.set_threads(1)
            log(f"[bench] cpu baseline, {threads} thread(s): {solves} solves in {t_cpu:.1f} s -> "
                f"{solves / t_cpu:.5f} solves/s (iterations {its_cpu})")
            return solves, t_cpu, its_cpu

        s1, t1, i1 = cpu_leg(1)
        sN, tN, iN = cpu_leg(nthreads)
        cpu = {"value": round(sN / tN, 5), "unit": "solves/s", "cores": nthreads, "kind": "port",
               "sample": f"oracle FP64 GeMSLR-FGMRES({kd}), {sN} whole solves of columns 0-{sN - 1} to {tol} on the "
                         f"same operator with the diagonal blocks' ILU solves over {nthreads} OpenMP threads "
                         f"({tN:.1f} s, iterations {iN}); host has {host_cpus} CPUs",
               "value_1thread": round(s1 / t1, 5), "sample_1thread": f"{s1} solves in {t1:.1f} s on 1 thread",
               "host_cpus": host_cpus}

    out = {
        "metric": "FGMRES solves/s at 1e-8 true relative residual (mixed FP32-apply/FP64-Krylov, batched)",
        "value": round(solves_per_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(T / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-apply/f64-krylov",
        "data": "synthetic",
        "config": {"workload": f"{'convdiff' if 'convdiff' in cfg else 'laplacian'}3d_{cfg['grid'][0]}^3"
                               f"{' shift=' + str(cfg['shift']) if 'shift' in cfg else ''} gemslr nlev={cfg['nlev']} "
                               f"ncomp={cfg['ncomp']} ilut={cfg['droptol']}/{cfg['rownnz']} rank={cfg['rank']} "
                               f"fgmres({kd}) tol={tol}"
                               f"{' inner=fgmres(%d,%g)' % schur if schur else ''} orth={args.orth}",
                   "nrhs_per_step": nr, "n": n, "nnz": A.nnz},
        "latency_ms": {"batched_step_p50": round(float(np.percentile(step_t, 50)) * 1e3, 2),
                       "batched_step_p99": round(float(np.percentile(step_t, 99)) * 1e3, 2),
                       "fp64_single_p50": round(float(np.percentile(lat64, 50)) * 1e3, 2),
                       "fp64_single_p99": round(float(np.percentile(lat64, 99)) * 1e3, 2),
                       "mixed_single_p50": round(float(np.percentile(lat32, 50)) * 1e3, 2),
                       "mixed_single_p99": round(float(np.percentile(lat32, 99)) * 1e3, 2),
                       "fp64_batched_step_p50": round(float(np.percentile(f64b_t, 50)) * 1e3, 2)},
        "fp64_single_solves_per_s": round(fp64_solves_per_s, 3),
        "mixed_single_solves_per_s": round(1.0 / float(np.median(lat32)), 3),
        "fp64_batched_solves_per_s": round(fp64_batched, 3),
        "pcie_inclusive": pcie,
        "mixed_batched_graph_solves_per_s": round(nr * len(graph_t) / sum(graph_t), 3),
        "graph_replay_same_bits": graph_same,
        "speedup_vs_fp64": round(solves_per_s / fp64_solves_per_s, 2),
        "speedup_split": {"batching_fp64": round(fp64_batched / fp64_solves_per_s, 2),
                          "precision_batched": round(solves_per_s / fp64_batched, 2)},
        "timed_solves": {"fp64_single": len(lat64), "mixed_single": len(lat32), "mixed_batched_steps": args.steps,
                         "fp64_batched_steps": len(f64b_t)},
        "iters": {"fp64": r64[1], "mixed_single": r32[1], "batched_min": min(iters), "batched_max": max(iters),
                  "fp64_batched_min": min(r[1] for r in res64B), "fp64_batched_max": max(r[1] for r in res64B)},
        "true_rel_res": {"fp64": float(res64), "mixed_single": float(res32), "batched_max": max(true_res),
                         "fp64_batched_max": true64B},
        "err_mixed_vs_fp64_solution": err_vs_fp64,
        "err_vs_xstar": err_vs_xstar,
        "converged": ok and ok64B,
        "setup_s": round(inf.setup_seconds[3], 2),
        "setup_lowrank": setup_lr,
        "roofline": roofline,
        "kernel_table": kernel_table,
        "cpu_baseline": cpu,
    }
    if not args.no_serve and args.config == "c2":
        # a short stream of the serving workload (four resident operators,
        # Poisson arrivals, 1-16 columns per request): its solves/s and request
        # latency percentiles beside the batched figures
        for D_ in (W1, WB, D):
            D_.close()
        sv = serving_stream(args.serve_requests, args.serve_rate, bool(args.graphs), orth, None)
        out["serving"] = {k: sv[k] for k in ("value", "unit", "latency_ms", "batch_nrhs_mean", "max_true_rel_res",
                                             "converged", "wall_s")}
        out["serving"]["requests"] = sv["config"]["requests"]
        out["serving"]["rhs_total"] = sv["config"]["rhs_total"]
        out["serving"]["workload"] = sv["config"]["workload"]
    print(json.dumps(out), flush=True)


def restart(args):
    """Restart without setup: the --config operator is set up once, written to
    an operator file, read back and uploaded; the line reports the setup time
    beside the save, load and upload times, and checks that the reloaded
    operator's batched FP32 apply equals the original's bit for bit."""
    import pargemslr_amd as pg

    cfg = CONFIGS[args.config]
    if "convdiff" in cfg:
        A = pg.convdiff(*cfg["grid"], cfg["convdiff"])
    else:
        A = pg.laplacian(*cfg["grid"], shift=cfg.get("shift", 0.0))
    params = pg.config_params(cfg["nlev"], cfg["ncomp"], cfg["droptol"], cfg["rownnz"], cfg["rank"], kdim=cfg["kdim"])
    t0 = time.perf_counter()
    G = pg.Gemslr(A, params, seed=0)
    t_setup = time.perf_counter() - t0
    log(f"[restart] setup {t_setup:.2f}s")
    t0 = time.perf_counter()
    G.save(args.restart)
    t_save = time.perf_counter() - t0
    size = os.path.getsize(args.restart)
    log(f"[restart] saved {size / 1e6:.1f} MB in {t_save:.2f}s")
    t0 = time.perf_counter()
    H = pg.Gemslr.load(args.restart)
    t_load = time.perf_counter() - t0
    t0 = time.perf_counter()
    DH = pg.DevOp(H, prec_mask=pg.PGS_PREC_FP64 | pg.PGS_PREC_FP32)
    pg.synchronize()
    t_upload = time.perf_counter() - t0
    log(f"[restart] loaded in {t_load:.2f}s, uploaded in {t_upload:.2f}s")
    nr = args.nrhs
    V = pg.DeviceArray.from_numpy(np.stack([pg.rng_normal(c, A.n) for c in range(nr)], axis=1))
    outs = []
    for D in (pg.DevOp(G, prec_mask=pg.PGS_PREC_FP32), DH):
        W = pg.Workspace(D, max_nrhs=nr, kdim=2)
        Z = pg.DeviceArray(V.shape)
        W.apply(V, Z, nrhs=nr, prec=pg.PGS_PREC_FP32)
        pg.synchronize()
        outs.append(Z.numpy())
    same = bool(np.array_equal(outs[0], outs[1]))
    os.remove(args.restart)
    print(json.dumps({"restart": f"{args.config}: grid {cfg['grid']}, nlev={cfg['nlev']} ncomp={cfg['ncomp']}", "setup_s": round(t_setup, 3), "save_s": round(t_save, 3),
                      "file_mb": round(size / 1e6, 1), "load_s": round(t_load, 3),
                      "upload_s": round(t_upload, 3), "load_plus_upload_s": round(t_load + t_upload, 3),
                      "apply_bit_identical": same}), flush=True)
    if not same:
        sys.exit(1)


def trsv_sweep(args):
    """Level-set solve times of the 128^3 operator (C2 setup) per column
    count: one warm-up apply, then five applies with the workspace profiler on;
    prints ms per launch of level 0 and of levels 1+ (the profiler's trsv
    families) and checks every width's columns against the one-column apply."""
    import pargemslr_amd as pg

    cfg = CONFIGS["c2"]
    A = pg.laplacian(*cfg["grid"])
    G = pg.Gemslr(A, pg.config_params(cfg["nlev"], cfg["ncomp"], cfg["droptol"], cfg["rownnz"], cfg["rank"]), seed=0)
    nrs = [int(x) for x in args.trsv_sweep.split(",")]
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=max(nrs), kdim=2)
    V = np.stack([pg.rng_normal(c, A.n) for c in range(max(nrs))], axis=1)
    for name, prec in (("fp32", pg.PGS_PREC_FP32), ("fp64", pg.PGS_PREC_FP64)):
        z1 = pg.DeviceArray(A.n)
        W.apply(pg.DeviceArray.from_numpy(np.ascontiguousarray(V[:, 0])), z1, prec=prec)
        pg.synchronize()
        ref = z1.numpy()
        for nr in nrs:
            Vd = pg.DeviceArray.from_numpy(np.ascontiguousarray(V[:, :nr]))
            Z = pg.DeviceArray((A.n, nr))
            W.apply(Vd, Z, nrhs=nr, prec=prec)
            pg.synchronize()
            W.profile(True)
            for _ in range(5):
                W.apply(Vd, Z, nrhs=nr, prec=prec)
            pg.synchronize()
            f = W.profile_families()
            W.profile(False)
            col0 = Z.numpy().reshape(A.n, nr)[:, 0]
            diff = np.linalg.norm(col0 - ref) / np.linalg.norm(ref)
            l0, l1 = f["trsv level 0"], f["trsv levels 1+"]
            log(f"[trsv-sweep] {name} nrhs {nr:3d}: level 0 {l0['total_ms'] / l0['launches']:.3f} ms/launch,"
                f" levels 1+ {l1['total_ms'] / max(l1['launches'], 1):.3f} ms/launch,"
                f" column 0 vs the one-column apply {diff:.1e}")
    W.close()
    D.close()


def spmv_probe(args):
    """y = A x with the 128^3 Laplacian held by itself on the device
    (pgs_dev_csr), 30 products per lanes-per-row setting; the products of
    each setting are checked against the first setting's."""
    import ctypes as C

    import pargemslr_amd as pg

    A = pg.laplacian(128, 128, 128)
    L = pg.lib()
    x = pg.DeviceArray.from_numpy(pg.rng_normal(1, A.n))
    y = pg.DeviceArray(A.n)
    one = (C.c_double * 2)(1.0, 0.0)
    zero = (C.c_double * 2)(0.0, 0.0)
    ref = None
    for lpr in ("8", "4", "2", "1"):
        os.environ["PGS_SPMV_LPR"] = lpr
        h = C.c_void_p()
        pg.capi._chk(L.pgs_dev_csr_create(0, A.n, pg.capi._iptr(A.ia), pg.capi._iptr(A.ja),
                                          A.a.ctypes.data_as(C.POINTER(C.c_double)), 0, C.byref(h)), "create")
        for _ in range(30):
            pg.capi._chk(L.pgs_dev_csr_spmv(h, one, x.ptr, zero, y.ptr, None), "spmv")
        pg.synchronize()
        out = y.numpy()
        ref = out if ref is None else ref
        log(f"[spmv-probe] lanes per row {lpr}: max |y - y(8 lanes)| {np.abs(out - ref).max():.2e}")
        L.pgs_dev_csr_destroy(h)
    del os.environ["PGS_SPMV_LPR"]
    # 32-column products with 4 and 8 columns per thread (FP64 x and y)
    G = pg.Gemslr(A, pg.config_params(1, 1, 1e-1, 2, 0), seed=0)
    nr = 32
    X = pg.DeviceArray.from_numpy(np.stack([pg.rng_normal(c, A.n) for c in range(nr)], axis=1))
    Y = pg.DeviceArray((A.n, nr))
    ref = None
    for cols in ("4", "8"):
        os.environ["PGS_SPMM_COLS"] = cols
        D = pg.DevOp(G, prec_mask=pg.PGS_PREC_FP64)
        W = pg.Workspace(D, max_nrhs=nr, kdim=1)
        for _ in range(20):
            W.spmv(X, Y, nrhs=nr)
        pg.synchronize()
        out = Y.numpy()
        ref = out if ref is None else ref
        log(f"[spmv-probe] {nr} columns, {cols} per thread: max |Y - Y(4)| {np.abs(out - ref).max():.2e}")
        W.close()
        D.close()
    del os.environ["PGS_SPMM_COLS"]


def probe(args):
    """Workload for hardware-counter passes (rocprofv3 --pmc): one warm-up
    apply of the 128^3 operator (FP32, --nrhs right-hand sides), a 1 GiB
    device-to-device copy whose known size calibrates FETCH_SIZE and
    WRITE_SIZE, then three applies; with --probe-fgmres the copy, then one
    batched mixed FGMRES solve of the benchmark (delayed CGS2) instead."""
    import pargemslr_amd as pg

    cfg = CONFIGS["c2"]
    A = pg.laplacian(*cfg["grid"])
    G = pg.Gemslr(A, pg.config_params(cfg["nlev"], cfg["ncomp"], cfg["droptol"], cfg["rownnz"], cfg["rank"]), seed=0)
    nr = args.nrhs
    if args.probe_fgmres:
        D = pg.DevOp(G, prec_mask=pg.PGS_PREC_FP64 | pg.PGS_PREC_FP32)
        W = pg.Workspace(D, max_nrhs=nr, kdim=cfg["kdim"])
        Xs = np.stack([pg.rng_uniform(c, A.n) for c in range(nr)], axis=1)
        B = pg.DeviceArray.from_numpy(A.matvec(Xs))
        X = pg.DeviceArray((A.n, nr))
        nb = 1 << 30
        src, dst = pg.DeviceArray(nb // 8), pg.DeviceArray(nb // 8)
        pg.capi._chk(pg.lib().pgs_memcpy_dtod(0, dst.ptr, src.ptr, nb, None), "pgs_memcpy_dtod")
        pg.synchronize()
        log(f"[probe] copied {nb} bytes")
        res = W.fgmres(B, X, nrhs=nr, kdim=cfg["kdim"], maxits=1000, tol=cfg["tol"], prec=pg.PGS_PREC_FP32,
                       orth=pg.PGS_ORTH_DCGS2)
        pg.synchronize()
        log(f"[probe] fgmres done, iterations {min(r[1] for r in res)}-{max(r[1] for r in res)}")
        return
    D = pg.DevOp(G, prec_mask=pg.PGS_PREC_FP32)
    W = pg.Workspace(D, max_nrhs=nr, kdim=2)
    V = pg.DeviceArray.from_numpy(np.stack([pg.rng_normal(c, A.n) for c in range(nr)], axis=1))
    Z = pg.DeviceArray(V.shape)
    W.apply(V, Z, nrhs=nr, prec=pg.PGS_PREC_FP32)
    pg.synchronize()
    nb = 1 << 30
    src, dst = pg.DeviceArray(nb // 8), pg.DeviceArray(nb // 8)
    pg.capi._chk(pg.lib().pgs_memcpy_dtod(0, dst.ptr, src.ptr, nb, None), "pgs_memcpy_dtod")
    log(f"[probe] copied {nb} bytes")
    for _ in range(3):
        W.apply(V, Z, nrhs=nr, prec=pg.PGS_PREC_FP32)
    pg.synchronize()
    log("[probe] applies done")


def serve(args):
    """The serving stream alone (--serve): one JSON line."""
    orth = {"mgs": 0, "cgs2": 1, "dcgs2": 2}[args.orth]
    schur = None
    if args.schur:
        k, t = args.schur.split(",")
        schur = (int(k), float(t))
    print(json.dumps(serving_stream(args.requests, args.rate, bool(args.graphs), orth, schur)), flush=True)


def serving_stream(requests, rate, graphs, orth, schur):
    """Serving stream: four resident operators (2D Laplacian 1024^2, 3D
    Laplacian 64^3, upwind convection-diffusion 96^3 with c = 0.3, complex-
    shifted 3D Laplacian 80^3 with shift 0.2+0.02i; GeMSLR 3 levels, 64
    subdomains, ILUT 1e-3/20, rank 16), requests with Poisson arrivals, each
    on a uniformly chosen operator with 1-16 N(0,1) right-hand sides (complex
    N(0,1) for the complex operator), solved to 1e-8 by the request server
    (FP32 / complex64 apply inside FP64 / complex128 FGMRES, continuous
    batching)."""
    import threading

    import pargemslr_amd as pg

    rng = np.random.default_rng(0)
    mats = [("laplacian2d_1024^2", pg.laplacian(1024, 1024, 1)), ("laplacian3d_64^3", pg.laplacian(64, 64, 64)),
            ("convdiff3d_96^3_c0.3", pg.convdiff(96, 96, 96, 0.3)),
            ("complex_laplacian3d_80^3_shift0.2+0.02i", pg.laplacian(80, 80, 80, shift=0.2 + 0.02j, complex_=True))]

    def rhs(seed, A):
        if A.is_complex:
            return pg.rng_normal(seed, 2 * A.n).view(np.complex128)
        return pg.rng_normal(seed, A.n)

    params = pg.config_params(3, 64, 1e-3, 20, 16, kdim=50, schur=schur)
    ops = []
    for name, A in mats:
        t0 = time.time()
        G = pg.Gemslr(A, params, seed=0)
        D = pg.DevOp(G, prec_mask=pg.PGS_PREC_FP32)
        ops.append((name, A, G, D))
        log(f"[serve] {name}: n={A.n} setup {time.time() - t0:.1f}s, device {D.bytes() / 1e9:.2f} GB")
    S = pg.Server(max_batch=16, slots=3, mem_reserve=8 << 30, kdim=50, tol=1e-8, orth=orth, use_graph=graphs)
    ids = [S.add_operator(D) for _, _, _, D in ops]
    # warm-up: one request per operator (workspaces, graphs)
    for i, (_, A, _, _) in enumerate(ops):
        S.wait(S.submit(ids[i], rhs(7, A)))
    plan = []
    for k in range(requests):
        op = int(rng.integers(0, len(ops)))
        nr = int(rng.integers(1, 17))
        B = np.stack([rhs(100000 + 64 * k + c, ops[op][1]) for c in range(nr)], axis=1)
        plan.append((op, B, float(rng.exponential(1.0 / rate))))
    reqs = [None] * len(plan)

    def feed():
        for k, (op, B, gap) in enumerate(plan):
            time.sleep(gap)
            reqs[k] = S.submit(ids[op], B)

    t_start = time.perf_counter()
    th = threading.Thread(target=feed)
    th.start()
    th.join()
    out = [S.wait(r) for r in reqs]
    T = time.perf_counter() - t_start
    S.close()
    lat = np.array([info["total_ms"] for _, _, info in out])
    cols = sum(B.shape[1] for _, B, _ in plan)
    worst = 0.0
    for (op, B, _), (X, its, info) in zip(plan, out):
        A = ops[op][1]
        for c in range(B.shape[1]):
            worst = max(worst, float(np.linalg.norm(B[:, c] - A.matvec(X[:, c])) / np.linalg.norm(B[:, c])))
    ok = all(info["status"] == 0 for _, _, info in out) and worst <= 1e-8
    per_op = {}
    for i, (name, _, _, _) in enumerate(ops):
        sel = [(B, r) for (op, B, _), r in zip(plan, out) if op == i]
        if not sel:
            continue
        l_i = np.array([r[2]["total_ms"] for _, r in sel])
        per_op[name] = {"requests": len(sel), "rhs": int(sum(B.shape[1] for B, _ in sel)),
                        "iters_mean": round(float(np.mean(np.concatenate([r[1] for _, r in sel]))), 1),
                        "p50_ms": round(float(np.percentile(l_i, 50)), 2),
                        "p99_ms": round(float(np.percentile(l_i, 99)), 2)}
    line = {
        "metric": "serving: FGMRES solves/s and request latency at 1e-8 true relative residual (mixed, batched)",
        "value": round(cols / T, 3), "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32-apply/f64-krylov (complex64/complex128 for the complex operator)", "data": "synthetic",
        "config": {"workload": "serving stream: " + ", ".join(n for n, _, _, _ in ops),
                   "requests": len(plan), "rhs_total": cols, "arrival_rate_per_s": rate,
                   "max_batch": 16, "slots": 3, "graphs": graphs,
                   "inner_schur": list(schur) if schur else None},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2),
                       "mean_queue": round(float(np.mean([i["queue_ms"] for _, _, i in out])), 2)},
        "batch_nrhs_mean": round(float(np.mean([i["batch_nrhs"] for _, _, i in out])), 2),
        "per_operator": per_op, "max_true_rel_res": worst, "converged": ok, "wall_s": round(T, 2),
    }
    for _, _, _, D in ops:
        D.close()
    return line


if __name__ == "__main__":
    main()
