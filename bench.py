#!/usr/bin/env python3
"""Time-to-solution benchmark of the MI355X SaP solver.

Workload (default): synthetic banded system N=1,000,000, half-bandwidth
K=200, off-diagonals U[-1,1), diagonal 0.8 x row off-diagonal abs-sum (not
diagonally dominant), x_true ~ U[-1,1), b = A x_true.  SaP-C (truncated
SPIKE, LU_UL) with P=64 partitions, FP32 factors and sweeps, FP64
BiCGStab(2) with FP64 refinement, stopping at relative residual 1e-10.

One step = setup() + solve() from scratch with A and b resident in HBM
(the reference's Stats.timeSetup + Stats.timeSolve).  The plain path (FP64
SaP-D, same P, no refinement) is timed beside it for the speedup.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-solution @ rel. residual 1e-10, N=1M K=200 banded; speedup vs plain"

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, GB/s
FP32_MFMA_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix
FP32_VALU_PEAK_TFLOPS = 78.6    # MI355X FP32 vector without packed math

WORKLOADS = {
    # name: (n, k, dominance, P, seed_A, seed_x)
    "c1": (1_000_000, 200, 0.8, 64, 10, 11),
    "c0": (200_000, 50, 1.2, 16, 0, 1),
}


def band_nnz(n, k):
    return n * (2 * k + 1) - k * (k + 1)


def part_sizes(n, P):
    q, r = divmod(n, P)
    return [q + 1 if p < r else q for p in range(P)]


def factor_nnz(n, k, P):
    # stored entries of the partition-diagonal blocks of the band
    return sum(ln * (2 * k + 1) - k * (k + 1) for ln in part_sizes(n, P))


def algorithmic(name, n, k, P, sp, spike):
    """(bytes or flops, unit, bound) per launch of a kernel family."""
    if name == "spmv":
        return 8 * band_nnz(n, k) + 16 * n, "GB/s", "hbm"
    if name == "sweep_full":
        extra = 2 * k * (P - 1) * sp if spike else 0
        return sp * factor_nnz(n, k, P) + 16 * n + extra, "GB/s", "hbm"
    if name == "sweep_interface":
        # L of LU (partitions 0..P-2) and U of UL (1..P-1), plus D and v
        sizes = part_sizes(n, P)
        half = sum(ln * (k + 1) - k * (k + 1) // 2 for ln in sizes[:-1]) + \
            sum(ln * (k + 1) - k * (k + 1) // 2 for ln in sizes[1:])
        return sp * half + 8 * (2 * n), "GB/s", "hbm"
    if name == "band_factor":
        # 2 n_p k^2 flops per LU or UL of a partition
        sizes = part_sizes(n, P)
        jobs = (sizes[:-1] + sizes[1:]) if spike else sizes
        return sum(2.0 * ln * k * k for ln in jobs), "TFLOP/s", "mfma"
    return None, None, None


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c1_sapc_fp32_pmc_traffic.md")
FAMILY_KERNEL = {"band_factor": "band_factor_resident_kernel", "sweep_full": "sweep_full_kernel",
                 "sweep_interface": "sweep_interface_kernel", "spmv": "band_spmv_kernel",
                 "band_to_colmajor": "band_to_colmajor_pipe_kernel"}


def measured_traffic(family):
    """HBM bytes per launch (read + write) of a kernel family, from the
    committed PMC summary (FETCH_SIZE/WRITE_SIZE runs, copy-calibrated)."""
    if family not in FAMILY_KERNEL or not os.path.exists(TRAFFIC_FILE):
        return None
    in_table = False
    for line in open(TRAFFIC_FILE):
        if line.startswith("| kernel | HBM read"):
            in_table = True
            continue
        if in_table and line.startswith("|") and FAMILY_KERNEL[family] in line:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                return float(cells[1]) + float(cells[2])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS) + ["c2", "c3", "c4"])
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device-to-device (a known-size kernel for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build runs on one GPU "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    import saplibrary_amd as sap

    if args.workload == "c3":
        return bench_c3(args, torch, sap)
    if args.workload == "c2":
        return bench_c2(args, torch, sap)
    if args.workload == "c4":
        return bench_c4(args, torch, sap)

    if args.copy_calibration:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] calibration copy: 1073741824 bytes read, 1073741824 bytes written", flush=True)
        del src, dst

    n, k, dom, P, sa, sx = WORKLOADS[args.workload]
    spike = args.workload == "c1"
    print(f"[bench] workload {args.workload}: n={n} k={k} d={dom} P={P}", flush=True)

    A = sap.BandedMatrix.generate(n, k, dom, sa)
    xt = sap.gen_vector(n, sx)
    b = torch.empty_like(xt)
    sap.band_spmv(A, xt, b)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()

    def options(plain, bcr=0):
        # autotune: the warm-up step times the conversion tile widths and
        # sweep staging depths for this k (the results do not change)
        if plain:
            return sap.Options(precondType=sap.Block, factMethod=sap.LU_UL, precValueType=sap.PREC_DOUBLE,
                               relTol=1e-10, profile=1, autotune=1)
        return sap.Options(precondType=sap.Spike if spike else sap.Block, factMethod=sap.LU_UL,
                           precValueType=sap.PREC_FLOAT, relTol=1e-10, refine=8, profile=1, autotune=1,
                           useBCR=bcr)

    def run(plain, steps, warmup, label, bcr=0):
        # one solver object: buffers allocated by the warm-up setup are reused
        s = sap.Solver(P, options(plain, bcr))
        times, last = [], None
        for it in range(warmup + steps):
            if it == warmup:
                s.reset_profile()
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s.setup(A)
            ok = s.solve(None, b, x)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            st = s.getStats()
            if it >= warmup:
                times.append((t1 - t0) * 1e3)
                last = (s, st, ok)
            print(f"[bench] {label} step {it}: {(t1 - t0) * 1e3:.2f} ms ok={ok} "
                  f"its={st.numIterations} rel={st.relResidualNorm:.2e} setup={st.timeSetup:.2f} "
                  f"solve={st.timeSolve:.2f}", flush=True)
        return times, last

    times, (solver, st, ok) = run(False, args.steps, args.warmup, "SaP-C fp32" if spike else "SaP-D fp32")
    prof = solver.profile()
    ms = sum(times) / len(times)

    # independent checks: true residual and forward error against x_true
    r = torch.empty_like(b)
    sap.band_spmv(A, x, r)
    rel = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
    ferr = float(torch.linalg.norm(x - xt) / torch.linalg.norm(xt))

    # dominant kernel family (per step) and its roofline
    steps_profiled = args.steps
    fam = max(prof.items(), key=lambda kv: kv[1][1])
    name, (launches, total_ms) = fam
    algo, unit, bound = algorithmic(name, n, k, P, 4, spike)
    roof = None
    if algo is not None and launches > 0:
        avg_s = total_ms / launches / 1e3
        if unit == "GB/s":
            achieved = algo / avg_s / 1e9
            peak = HBM_PEAK_GBS
        else:
            achieved = algo / avg_s / 1e12
            peak = FP32_MFMA_PEAK_TFLOPS
        traffic = measured_traffic(name)
        roof = {"kernel": name, "bound": bound, "achieved": round(achieved, 2), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 4), "traffic": traffic,
                "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if traffic else None,
                "avg_launch_ms": round(total_ms / launches, 4), "launches_per_step": launches / steps_profiled,
                "algorithmic_per_launch": algo}
    kernels = {kname: {"launches_per_step": round(l / steps_profiled, 2),
                       "ms_per_step": round(t / steps_profiled, 3)} for kname, (l, t) in prof.items()}

    if roof is not None and name == "band_factor":
        # one workgroup per (partition, LU or UL) job holds a CU for the whole
        # launch: the peak of the CUs the launch occupies
        jobs = 2 * (P - 1) if spike and P > 1 else P
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        roof["workgroups"] = jobs
        roof["frac_of_occupied_cus_peak"] = round(achieved / (peak * min(jobs, cus) / cus), 4)

    # the practical streaming-read bandwidth (the library's own read kernel)
    read_gbs = sap.stream_read_gbs()
    if roof is not None and roof["unit"] == "GB/s":
        roof["frac_of_stream_read"] = round(roof["achieved"] / read_gbs, 4)
    stream_fracs = {}
    for fam in ("spmv", "sweep_full", "sweep_interface"):
        if fam in prof and prof[fam][0] > 0:
            algo_b, _, _ = algorithmic(fam, n, k, P, 4, spike)
            gbs = algo_b / (prof[fam][1] / prof[fam][0] / 1e3) / 1e9
            stream_fracs[fam] = {"GB/s": round(gbs, 1), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4),
                                 "frac_of_stream_read": round(gbs / read_gbs, 4)}

    # the SaP-D workload with block cyclic reduction of the partitions
    # (useBCR) instead of their sequential banded LU
    bcr = None
    if not spike:
        btimes, (bs, bst, bok) = run(False, args.steps, args.warmup, "SaP-D fp32 BCR", bcr=1)
        bprof = bs.profile()
        bcr = {"ms": round(sum(btimes) / len(btimes), 3), "converged": bool(bok),
               "setup_ms": round(bst.timeSetup, 3), "solve_ms": round(bst.timeSolve, 3),
               "iterations": bst.numIterations,
               "kernels": {kn: {"launches_per_step": round(l / args.steps, 2), "ms_per_step": round(t / args.steps, 3)}
                           for kn, (l, t) in bprof.items()}}

    plain_ms = None
    if not args.no_plain:
        ptimes, (_, pst, pok) = run(True, max(1, args.steps), args.warmup, "plain fp64 SaP-D")
        plain_ms = sum(ptimes) / len(ptimes)

    # BiCGStab(2) per-iteration time on a solve that iterates: the c1 matrix
    # with SaP-D (FP32 block-diagonal preconditioner, no refinement, the
    # reference's default tolerance 1e-6), one right-hand side through the
    # host-driven loop and two through the batched, graph-replayed loop
    krylov = None
    if not args.no_plain:
        ko = dict(precondType=sap.Block, factMethod=sap.LU_UL, precValueType=sap.PREC_FLOAT, relTol=1e-6)
        ks = sap.Solver(P, sap.Options(**ko))
        ks.setup(A)
        kx = torch.zeros_like(b)
        ks.solve(None, b, kx)                     # warm-up
        kx.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ks.solve(None, b, kx)
        torch.cuda.synchronize()
        t_single = (time.perf_counter() - t0) * 1e3
        its_single = ks.getStats().numIterations
        B2 = torch.stack([b, 0.5 * b], dim=1)
        X2 = torch.zeros_like(B2)
        ks.solve(None, B2, X2)                    # warm-up (graph capture)
        X2.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ks.solve(None, B2, X2)
        torch.cuda.synchronize()
        t_batch = (time.perf_counter() - t0) * 1e3
        its_batch = ks.getStats().numIterations
        krylov = {"workload": "c1 matrix, SaP-D P=64 FP32, BiCGStab(2), rel tol 1e-6, no refinement",
                  "iterations": its_single, "solve_ms": round(t_single, 3),
                  "ms_per_iteration": round(t_single / max(its_single, 0.25), 3),
                  "batched_2rhs_iterations": its_batch, "batched_2rhs_solve_ms": round(t_batch, 3),
                  "batched_ms_per_iteration_per_rhs": round(t_batch / max(its_batch, 0.25) / 2, 3)}
        print(f"[bench] krylov: {krylov}", flush=True)
        del ks

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.workload, args.cpu_threads)

    out = {
        "metric": METRIC,
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 preconditioner + f64 Krylov",
        "data": "synthetic (xoshiro256** generator, seeds %d/%d)" % (sa, sx),
        "config": {"workload": f"banded N={n} K={k} diag={dom}x off-diag abs-sum; "
                               f"{'SaP-C LU_UL' if spike else 'SaP-D'} P={P}; BiCGStab(2) + FP64 refinement; "
                               f"rel tol 1e-10",
                   "n": n, "k": k, "partitions": st.numPartitions},
        "converged": bool(ok),
        "rel_residual": rel,
        "forward_error": ferr,
        "iterations": st.numIterations,
        "setup_ms": round(st.timeSetup, 3),
        "solve_ms": round(st.timeSolve, 3),
        "speedup_vs_plain": round(plain_ms / ms, 3) if plain_ms else None,
        "plain_ms": round(plain_ms, 3) if plain_ms else None,
        "kernels": kernels,
        "roofline": roof,
        "krylov": krylov,
        "bcr": bcr,
        "stream_read_gbs": round(read_gbs, 1),
        "hbm_kernels": stream_fracs,
        "cpu_baseline": cpu,
        "autotune": [{"family": ["conversion_tile", "sweep_nbuf"][e["family"]], "k": e["k"],
                      "elem_bytes": e["elem_bytes"], "value": e["value"], "default": e["default_value"],
                      "best_ms": round(e["best_ms"], 4), "default_ms": round(e["default_ms"], 4)}
                     for e in sap.autotune_entries()],
    }
    print(json.dumps(out), flush=True)


def _line(metric, value, unit, args, ms, dtype, data, config, hib=False, **extra):
    out = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": hib, "scaling": "strong",
           "vs_baseline": None, "dtype": dtype, "data": data, "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


C3_SYSTEMS = 32
C3_RHS = 16


def c3_systems():
    """32 banded systems, N ~ U[10k, 300k], K ~ U[10, 150] (numpy seed 30),
    the banded generator with diagonal 1.2x the off-diagonal abs-sum (seeds
    30 + i), 16 right-hand sides each (seeds 1000 + 16 i + r)."""
    import numpy as np
    rng = np.random.default_rng(30)
    ns = rng.integers(10_000, 300_001, C3_SYSTEMS)
    ks = rng.integers(10, 151, C3_SYSTEMS)
    return [(int(n), int(k), 30 + i) for i, (n, k) in enumerate(zip(ns, ks))]


def bench_c3(args, torch, sap):
    """Batching workload: (a) 32 uneven banded systems x 16 right-hand sides
    in one batched solve (every kernel family of the Krylov loop one launch
    for all 512 columns, each system's preconditioner launches covering its
    16 columns, one graph replay per iteration) against solving the same
    columns one after another; (b) 64 update() steps on one N=400k, K=120
    banded pattern, values x (1 + U[-0.05, 0.05]) per step, each followed by
    a solve."""
    P = 16
    opts = dict(precondType=sap.Block, factMethod=sap.LU_UL, precValueType=sap.PREC_FLOAT, relTol=1e-10, refine=8,
                autotune=1)
    specs = c3_systems()
    mats, B = [], []
    for n, k, seed in specs:
        mats.append(sap.BandedMatrix.generate(n, k, 1.2, seed))
        B.append(torch.stack([sap.gen_vector(n, 1000 + 16 * seed + r) for r in range(C3_RHS)], dim=1))
    torch.cuda.synchronize()
    print(f"[bench] c3: {C3_SYSTEMS} systems, N {min(s[0] for s in specs)}..{max(s[0] for s in specs)}, "
          f"K {min(s[1] for s in specs)}..{max(s[1] for s in specs)}, {C3_RHS} rhs each", flush=True)

    sb = sap.SystemBatch([P] * C3_SYSTEMS, sap.Options(**opts))
    t_batch, ok_all = [], True
    for it in range(args.warmup + args.steps):
        X = [torch.zeros_like(b) for b in B]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sb.setup(mats)
        ok = sb.solve(B, X)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if it >= args.warmup:
            t_batch.append((t1 - t0) * 1e3)
            ok_all &= ok
        print(f"[bench] c3 batched step {it}: {(t1 - t0) * 1e3:.2f} ms ok={ok}", flush=True)
    res = sb.results()
    worst = 0.0
    for (n, k, seed), A, b, x in zip(specs, mats, B, X):
        for r in range(C3_RHS):
            y = torch.empty(n, dtype=torch.float64, device="cuda")
            sap.band_spmv(A, x[:, r].contiguous(), y)
            worst = max(worst, float(torch.linalg.norm(b[:, r] - y) / torch.linalg.norm(b[:, r])))
    setup_ms, solve_ms = sb.times()

    # the same columns one system and one column at a time
    t_seq = []
    solvers = [sap.Solver(P, sap.Options(batch=0, **opts)) for _ in specs]
    for it in range(1 + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s, A, b in zip(solvers, mats, B):
            s.setup(A)
            x = torch.zeros_like(b)
            s.solve(None, b, x)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if it >= 1:
            t_seq.append((t1 - t0) * 1e3)
        print(f"[bench] c3 one-by-one step {it}: {(t1 - t0) * 1e3:.2f} ms", flush=True)
    del solvers

    # update() sequence
    n, k = 400_000, 120
    base = sap.BandedMatrix.generate(n, k, 1.2, 40)
    xt = sap.gen_vector(n, 41)
    s = sap.Solver(P, sap.Options(trackReordering=1, **opts))
    s.setup(base)
    g = torch.Generator(device="cuda")
    g.manual_seed(42)
    vals = base.band.clone()
    seq_ms, seq_ok, seq_worst = [], True, 0.0
    b = torch.empty_like(xt)
    for step in range(64):
        vals = vals * (1.0 + (torch.rand(vals.shape, generator=g, device="cuda", dtype=torch.float64) - 0.5) * 0.1)
        sap.band_spmv(sap.BandedMatrix(vals), xt, b)
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assert s.update(vals)
        ok = s.solve(None, b, x)
        torch.cuda.synchronize()
        seq_ms.append((time.perf_counter() - t0) * 1e3)
        seq_ok &= ok
        seq_worst = max(seq_worst, s.getStats().relResidualNorm)
    print(f"[bench] c3 update sequence: 64 steps, mean {sum(seq_ms) / 64:.2f} ms per update+solve, ok={seq_ok}",
          flush=True)

    ms = sum(t_batch) / len(t_batch)
    cols = C3_SYSTEMS * C3_RHS
    _line("time-to-solution, 32 banded systems x 16 right-hand sides (batched) @ rel. residual 1e-10", ms, "ms",
          args, ms, "f32 preconditioner + f64 Krylov", "synthetic (xoshiro256** banded generator, seeds 30..61)",
          {"workload": "32 banded systems N~U[10k,300k] K~U[10,150] diag 1.2x, 16 rhs each; SaP-D P=16; "
                       "BiCGStab(2) + FP64 refinement batched over 512 columns, graph-replayed iterations; "
                       "64-step update() sequence N=400k K=120",
           "systems": C3_SYSTEMS, "rhs_per_system": C3_RHS},
          converged=bool(ok_all), worst_rel_residual=worst, setup_ms=round(setup_ms, 3),
          solve_ms=round(solve_ms, 3), one_by_one_ms=round(sum(t_seq) / len(t_seq), 3),
          batching_speedup=round(sum(t_seq) / len(t_seq) / ms, 3), columns=cols,
          max_iterations=max(r[1] for r in res),
          update_sequence={"steps": 64, "n": n, "k": k, "ms_per_step": round(sum(seq_ms) / 64, 3),
                           "converged": bool(seq_ok), "worst_rel_residual": seq_worst})


def bench_c2(args, torch, sap):
    """General sparse workload (CSR N=500k, ~12 entries per row, random band
    of width 300 under a symmetric permutation, 20% weak diagonals): host
    DB + RCM + drop-off, then SaP.  The specified recipe (1% drop-off, SaP-C
    with 32 partitions, FP32 factors) is timed as specified and reported
    with its outcome; the fastest configuration found that converges on this
    matrix (no drop-off, 4 partitions, FP64 factors) is timed beside it."""
    from saplibrary_amd import workloads as W
    n = 500_000
    rp, ci, val = W.general_sparse(n)
    A = sap.CsrMatrix.from_host(rp, ci, val)
    xt = torch.from_numpy(W.uniform_vector(n, 22)).cuda()
    b = torch.empty_like(xt)
    sap.csr_spmv(A, xt, b)
    out = {}
    for name, P, o in [("specified", 32, dict(dropOffFraction=0.01, maxNumIterations=30, refine=3)),
                       ("converging", 4, dict(dropOffFraction=0.0, precValueType=sap.PREC_DOUBLE, refine=4))]:
        times = []
        for it in range(args.warmup + max(1, min(args.steps, 2))):
            s = sap.Solver(P, sap.Options(relTol=1e-10, **o))
            x = torch.zeros_like(b)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s.setup(A)
            ok = s.solve(None, b, x)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if it >= args.warmup:
                times.append((t1 - t0) * 1e3)
            st = s.getStats()
            r = torch.empty_like(b)
            sap.csr_spmv(A, x, r)
            rel = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
            print(f"[bench] c2 {name} step {it}: {(t1 - t0) * 1e3:.1f} ms ok={ok} rel={rel:.3e} "
                  f"setup={st.timeSetup:.1f} (DB {st.time_DB:.1f}) solve={st.timeSolve:.1f}", flush=True)
        out[name] = {"ms": round(sum(times) / len(times), 1), "converged": bool(ok), "rel_residual": rel,
                     "partitions": st.numPartitions, "k": st.bandwidth, "k_db": st.bandwidthDB,
                     "k_reorder": st.bandwidthReorder, "drop_off": st.actualDropOff, "setup_ms": round(st.timeSetup, 1),
                     "time_DB_ms": round(st.time_DB, 1), "solve_ms": round(st.timeSolve, 1),
                     "iterations": st.numIterations, "options": {k2: v for k2, v in o.items()}}
    conv = out["converging"]
    _line("time-to-solution, general sparse CSR N=500k @ rel. residual 1e-10", conv["ms"], "ms", args, conv["ms"],
          "f64 preconditioner + f64 Krylov", "synthetic (numpy generator, seeds 20/21/22)",
          {"workload": "CSR N=500000 ~12 nnz/row, band width 300 under a symmetric permutation, 20% weak "
                       "diagonals; DB + scaling + RCM on the host, then SaP (LU_only, variable-bandwidth path)",
           "n": n, "nnz": int(rp[-1])}, specified=out["specified"], converging=conv)


def bench_c4(args, torch, sap):
    """Learned option selector: the network is trained on the GPU (the
    library's own kernels) on the committed corpus of solver-measured
    timings (tests/golden/selector_corpus.npz), then fresh matrices of the
    corpus generators (never seen in training) are solved with the default
    options and with the chosen ones (falling back to the defaults when the
    choice does not reach the tolerance), both timed live."""
    import numpy as np
    from saplibrary_amd import selector as S
    z = np.load(os.path.join(ROOT, "tests", "golden", "selector_corpus.npz"))
    F, T, OK = z["features"], z["times"], z["ok"]
    keep = OK.any(axis=1)
    t0 = time.perf_counter()
    sel = S.Selector(sap, torch, seed=1)
    losses = sel.fit(F[keep], T[keep], OK[keep], epochs=3000)
    train_ms = (time.perf_counter() - t0) * 1e3
    print(f"[bench] c4: trained on {int(keep.sum())} matrices in {train_ms:.0f} ms, loss {losses[0]:.3g} -> "
          f"{losses[-1]:.3g}", flush=True)
    count = 40
    rows = []
    for q in range(count):
        i = 5000 + q
        A, b, f, is_csr, desc = S.corpus_matrix(sap, torch, i, np.random.default_rng((4, i)))
        g = int(sel.choose(f[None, :])[0])
        S.time_config(sap, torch, A, b, S.DEFAULT, is_csr)      # warm-up of this size
        td, okd = S.time_config(sap, torch, A, b, S.DEFAULT, is_csr, reps=2)
        tl, okl = S.time_config(sap, torch, A, b, S.GRID[g], is_csr, reps=2)
        if not okl:
            tl += td
        rows.append((td, okd, tl, okl or okd, g, desc))
        print(f"[bench] c4 matrix {q}: {desc} default {td:.1f} ms ok={okd}, learned {S.GRID[g]} {tl:.1f} ms",
              flush=True)
    both = [(td, tl) for td, okd, tl, ok, g, d in rows if okd]
    ratio = np.array([tl / td for td, tl in both])
    frac = float(np.mean(ratio <= 1.0)) if len(ratio) else None
    gm = float(np.exp(np.mean(np.log(ratio)))) if len(ratio) else None
    solved_default = sum(1 for r in rows if r[1])
    solved_learned = sum(1 for r in rows if r[3])
    total_d = sum(r[0] for r in rows)
    total_l = sum(r[2] for r in rows)
    _line("learned options vs defaults, time-to-solution ratio (geometric mean) @ rel. residual 1e-10", gm,
          "ratio", args, total_l / count, "f32 preconditioner + f64 Krylov; f32 network",
          "synthetic (corpus generators, matrices 5000..5039 not in the training corpus)",
          {"workload": "40 fresh matrices (60% banded N 20k-100k K 10-150 dominance 0.5-2, 40% CSR N 10k-30k "
                       "width 40-300 with weak diagonals), default options (Spike LU_only DB P=16) vs the "
                       "choice of an ensemble of 5 networks from 24 features over a 60-point option grid",
           "grid": len(S.GRID), "features": S.N_FEATURES},
          no_slower_fraction=frac, matrices_default_converged=solved_default,
          matrices_learned_converged=solved_learned, total_default_ms=round(total_d, 1),
          total_learned_ms=round(total_l, 1), train_ms=round(train_ms, 1), train_matrices=int(keep.sum()))


def cpu_baseline(workload, threads):
    """The CPU oracle (a restatement of the reference algorithm, OpenMP over
    partitions) timed on the host cores for the same workload."""
    if threads > 0:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    import numpy as np
    from oracle import pyoracle as O

    n, k, dom, P, sa, sx = WORKLOADS[workload]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in out.splitlines() if ln.startswith("Model name")), None)
    except Exception:
        pass
    nproc = os.cpu_count()
    band = O.gen_banded(n, k, dom, sa)
    b = O.band_spmv(band, O.gen_vector(n, sx))
    # one warm-up run, then the median of five (setup + solve each)
    times = []
    for it in range(6):
        S = O.Solver(P, rel_tol=1e-10, fact="LU_UL", precond="Spike" if workload == "c1" else "Block", refine=8)
        t0 = time.perf_counter()
        rc = S.setup(band)
        ok, _ = S.solve(b)
        t1 = time.perf_counter()
        st = S.stats()
        if it > 0:
            times.append((t1 - t0) * 1e3)
        print(f"[bench] cpu oracle run {it}: {(t1 - t0):.2f} s ok={ok} rc={rc} its={st['num_iterations']}",
              flush=True)
    del band, b
    times.sort()
    return {"value": round(times[len(times) // 2], 1), "unit": "ms", "cores": cores, "kind": "port",
            "sample": f"the full {workload} workload, setup+solve, median of 5 after one warm-up, "
                      f"OMP_NUM_THREADS={cores}",
            "runs_ms": [round(t, 1) for t in times], "nproc": nproc, "cpu_model": model,
            "converged": bool(ok), "rel_residual": st["rel_residual_norm"]}


if __name__ == "__main__":
    main()
