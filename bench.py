#!/usr/bin/env python3
"""HGEMM evaluation benchmark on one MI355X.

Workload (H2-b, the headline configuration): N = 65,536 points in d = 28 from
a 16-component Gaussian mixture (per-component diagonal sigma in [0.5, 2]),
standardised; Gaussian kernel h = 1.5, leaf 256, rank cap 256, ID tolerance
1e-5, near budget 0.03; W is 65,536 x 1,024, N(0,1), FP64.  Synthetic data
from the package's splitmix64/Box-Muller generators (points seed 1002,
W seed 2002, sampled rows seed 3002).

A step is one evaluation Y = K~ W in the mixed form (split-BF16x3 near and
far products with FP32 accumulation, FP32 tree transfers), with W and Y in
FP64 on the device (the FP64->FP32 input and FP32->FP64 output conversions
are inside the timed region).  The plain FP64 evaluation is timed the same
way for the speedup.  The compressed matrix is built on the host before the
timed region, as in the reference driver (matroxTest-V1/main.cpp:84-108).

Prints one JSON line (rank 0, N=1).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_PEAK = {"bf16": 2500.0, "fp32": 157.3, "fp64": 78.6, "hbm_gbs": 8000.0}

WORKLOAD = dict(name="h2b_gmm16_n65536_d28", n=65536, d=28, ncomp=16, sig_lo=0.5, sig_hi=2.0, diag=True,
                h=1.5, leaf=256, rank=256, tol=1e-5, budget=0.03, nrhs=1024,
                seed_points=1002, seed_w=2002, seed_rows=3002, check_rows=1000)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nrhs", type=int, default=WORKLOAD["nrhs"])
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-nrhs", type=int, default=128)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a plain 1-buffer copy of this size (HBM-counter calibration)")
    ap.add_argument("--skip-checks", action="store_true", help="skip the dense-row error checks")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a hipGraph")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this evaluator runs on one GPU; multi-GPU runs (--gpus > 1) are not part of this "
              "build (the sharding design is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from matrox_ru_amd import _native, hgemm, host
    for p in (_native.HOST_LIB, _native.HGEMM_LIB):
        if not os.path.exists(p):
            from matrox_ru_amd import build as b
            b.build_all()
            break
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    wl = dict(WORKLOAD)
    nrhs = args.nrhs

    t0 = time.time()
    X = host.gaussian_mixture(wl["n"], wl["d"], wl["ncomp"], sig_lo=wl["sig_lo"], sig_hi=wl["sig_hi"],
                              diag=wl["diag"], seed=wl["seed_points"])
    hm = host.build_hmatrix(X, leaf_size=wl["leaf"], max_rank=wl["rank"], tol=wl["tol"], budget=wl["budget"],
                            h=wl["h"])
    t_build = time.time() - t0
    log(f"[bench] host build {t_build:.1f}s: nleaf={hm.nleaf} ncount={hm.ncount} fcount={hm.fcount} "
        f"ranks {hm.slen[1:].min()}..{hm.slen[1:].max()}")
    W = host.normal((hm.N, nrhs), seed=wl["seed_w"])
    TW = hm.transform(W)
    del W
    prec = args.precision
    plan = hgemm.Plan(hm, precisions=("fp64", prec))
    perm = hm.lids[: hm.N]
    plan.set_points(X[perm], wl["h"])
    Wd = torch.from_numpy(TW).to(dev)
    del TW
    Y64 = torch.empty_like(Wd)
    Ymx = torch.empty_like(Wd)
    flops64, _ = plan.flops("fp64", nrhs)
    flopsmx, parts = plan.flops(prec, nrhs)
    log(f"[bench] plan ready; algorithmic flops/eval = {flops64 / 1e12:.4f} TFLOP")

    stream = torch.cuda.current_stream()

    def timed(precision, Y, steps, warmup):
        for _ in range(warmup):
            plan.eval(Wd, Y, nrhs, precision=precision)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            plan.eval(Wd, Y, nrhs, precision=precision)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    plan.set_graph(not args.no_graph)
    ms64 = timed("fp64", Y64, max(3, args.steps // 4), 1)
    log(f"[bench] fp64 {ms64:.3f} ms/eval")
    msmx = timed(prec, Ymx, args.steps, args.warmup)
    log(f"[bench] {prec} {msmx:.3f} ms/eval")
    # FP32 I/O variant of the mixed path (no conversions)
    W32 = Wd.float()
    Y32 = torch.empty_like(W32)
    for _ in range(2):
        plan.eval(W32, Y32, nrhs, precision=prec)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        plan.eval(W32, Y32, nrhs, precision=prec)
    e1.record(stream)
    torch.cuda.synchronize()
    ms32io = e0.elapsed_time(e1) / args.steps
    del W32, Y32

    # per-launch timing on the evaluation stream (separate pass)
    plan.set_profiling(True)
    per_kind = {}
    nprof = 3
    per_launch = None
    for _ in range(nprof):
        plan.eval(Wd, Ymx, nrhs, precision=prec)
        lt = plan.launch_times()
        per_launch = [ms for _, ms in lt] if per_launch is None else [a + ms for a, (_, ms) in zip(per_launch, lt)]
        for k, ms in lt:
            per_kind.setdefault(k, []).append(ms)
    log("[bench] per-launch ms: " + " ".join(f"{x / nprof:.3f}" for x in per_launch))
    log("[bench] per-launch GFLOP: " + " ".join(f"{f / 1e9:.1f}" for f in plan.launch_flops(prec, nrhs)))
    plan.set_profiling(False)
    kind_total = {k: sum(v) / nprof for k, v in per_kind.items()}
    kind_launches = {k: len(v) // nprof for k, v in per_kind.items()}
    dom = max((k for k in kind_total if k != "convert"), key=lambda k: kind_total[k])
    dom_flops = parts.get(dom, 0.0)
    dom_avg_ms = kind_total[dom] / max(1, kind_launches[dom])
    dom_flops_per_launch = dom_flops / max(1, kind_launches[dom])
    achieved = dom_flops_per_launch / (dom_avg_ms * 1e-3) / 1e12
    peak = {"bf16x3": MI355X_PEAK["bf16"], "seg_gemm_f32": MI355X_PEAK["fp32"],
            "seg_gemm_f64": MI355X_PEAK["fp64"]}[dom]
    roofline = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 5), "traffic": None,
                "launches_per_eval": kind_launches[dom], "avg_launch_ms": round(dom_avg_ms, 5),
                "flops_per_launch": dom_flops_per_launch}
    if dom == "bf16x3":
        roofline["frac_effective_x3"] = round(achieved / (peak / 3), 5)
    tfile = os.path.join(ROOT, "profiles", "traffic_h2b_n65536.txt")
    if nrhs == WORKLOAD["nrhs"] and os.path.exists(tfile):
        kv = dict(l.strip().split("=", 1) for l in open(tfile) if "=" in l and not l.startswith("#"))
        if kv.get("kernel") == dom:
            roofline["traffic"] = float(kv["fetch_bytes_per_launch"]) + float(kv["write_bytes_per_launch"])
            roofline["traffic_source"] = "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/x3_kernel_h2b_n65536.md"
            roofline["algorithmic_bytes_per_launch"] = round(4.67e9 / kind_launches[dom], -6)
    kinds_report = {k: {"ms_per_eval": round(kind_total[k], 4), "launches": kind_launches[k],
                        "tflops": round(parts.get(k, 0.0) / (kind_total[k] * 1e-3) / 1e12, 3)
                        if parts.get(k) else None} for k in kind_total}
    log(f"[bench] kernels: {kinds_report}")

    # errors
    d = Ymx - Y64
    err_vs64 = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(Y64))
    del d
    checks = {}
    if not args.skip_checks:
        rng = np.random.default_rng(wl["seed_rows"])
        rows = np.sort(rng.choice(hm.N, wl["check_rows"], replace=False)).astype(np.int32)
        dense = plan.dense_rows(rows, Wd, nrhs)
        r64 = plan.rows_of(Y64, rows, nrhs)
        rmx = plan.rows_of(Ymx, rows, nrhs)
        nd = torch.linalg.vector_norm(dense)
        checks["rel_err_fp64_vs_dense"] = float(torch.linalg.vector_norm(r64 - dense) / nd)
        checks["rel_err_mixed_vs_dense"] = float(torch.linalg.vector_norm(rmx - dense) / nd)
        # `acc` (computeError): first 10 points of leaf 0, mean relative row error
        arows = np.arange(min(10, int(hm.Ddim[0])), dtype=np.int32)
        da = plan.dense_rows(arows, Wd, nrhs)
        for name, Y in (("acc_fp64", Y64), ("acc_mixed", Ymx)):
            ya = plan.rows_of(Y, arows, nrhs)
            checks[name] = float(((torch.linalg.vector_norm(ya - da, dim=1) /
                                   torch.linalg.vector_norm(da, dim=1)).mean()))
        log(f"[bench] checks: {checks}")

    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib * (1 << 20) // 8
        a = torch.ones(nb, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calib_copy_mib} MiB done")
        del a, b

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        ns = args.cpu_sample_nrhs
        Ws = host.normal((hm.N, ns), seed=wl["seed_w"] + 1)
        TWs = hm.transform(Ws)
        oracle.hgemm(hm, TWs, ns)
        ts = []
        for _ in range(3):
            t1 = time.time()
            oracle.hgemm(hm, TWs, ns)
            ts.append(time.time() - t1)
        tcpu = statistics.median(ts)
        fl = oracle.flops(hm, ns)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(fl / tcpu / 1e12, 5), "unit": "TFLOP/s", "cores": cores, "kind": "port",
               "sample": f"oracle FP64 evaluation (restated HGEMM_gen.h sweeps, OpenMP, own column-major "
                         f"GEMM) of the same compressed matrix with nrhs={ns} instead of {nrhs}; median of 3 "
                         f"runs = {tcpu:.3f} s",
               "ms_per_eval_scaled_to_nrhs": round(tcpu * 1e3 * nrhs / ns, 1)}
        log(f"[bench] cpu baseline: {cpu}")

    out = {
        "metric": "HGEMM eval TFLOP/s+ms, mixed vs FP64 plain, N=64K nrhs=1024, rel err<=1e-5",
        "value": round(flopsmx / (msmx * 1e-3) / 1e12, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msmx, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-f32acc" if prec == "bf16x3" else "f32",
        "data": "synthetic (splitmix64/Box-Muller Gaussian mixture points, N(0,1) W)",
        "graph": not args.no_graph,
        "config": {"workload": wl["name"], "N": hm.N, "d": wl["d"], "nrhs": nrhs, "leaf": wl["leaf"],
                   "max_rank": wl["rank"], "tol": wl["tol"], "budget": wl["budget"], "h": wl["h"],
                   "nleaf": int(hm.nleaf), "near_blocks": int(hm.ncount), "far_blocks": int(hm.fcount),
                   "mean_rank": round(float(hm.slen[1:].mean()), 2), "io": "f64 (conversions timed)"},
        "algorithmic_tflop_per_eval": round(flops64 / 1e12, 5),
        "fp64_ms": round(ms64, 4),
        "fp64_tflops": round(flops64 / (ms64 * 1e-3) / 1e12, 3),
        "fp64_frac_of_peak": round(flops64 / (ms64 * 1e-3) / 1e12 / MI355X_PEAK["fp64"], 4),
        "speedup_vs_fp64": round(ms64 / msmx, 3),
        "mixed_ms_f32_io": round(ms32io, 4),
        "rel_err_mixed_vs_fp64": err_vs64,
        **checks,
        "kernels": kinds_report,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_build_s": round(t_build, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
