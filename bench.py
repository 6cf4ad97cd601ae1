#!/usr/bin/env python3
"""HGEMM evaluation benchmark on one MI355X.

Default workload (H2-b, the headline configuration): N = 65,536 points in
d = 28 from a 16-component Gaussian mixture (per-component diagonal sigma in
[0.5, 2]), standardised; Gaussian kernel h = 1.5, leaf 256, rank cap 256, ID
tolerance 1e-5, near budget 0.03; W is 65,536 x 1,024, N(0,1), FP64.
Synthetic data from the package's splitmix64/Box-Muller generators (points
seed 1002, W seed 2002, sampled rows seed 3002).

A step is one evaluation Y = K~ W in the mixed form (split-BF16x3 products
with FP32 accumulation on MFMA), with W and Y in FP64 on the device (the
input split and the FP64 output are inside the timed region).  The plain
FP64 evaluation is timed the same way for the speedup, and the adjoint
dW = K~^T dY and the backward pass with block gradients are timed beside it.
The compressed matrix is built on the host before the timed region, as in
the reference driver (matroxTest-V1/main.cpp:84-108).

Other workloads (--workload): hss_n16384 (plain HSS FP64, leaf 128, nrhs 256),
adjoint_n32768 (FP32 forward and backward with block gradients, nrhs 128,
checked against FP64 central differences on 64 directions) and batched64
(64 independent H2-b matrices of N log-uniform in [2048, 32768] evaluated as
one batch, FP32 storage) and scale_n524288 (N = 524,288, 2048 leaves, nrhs 512;
near and coupling blocks generated on the device; BF16x3, FP32 and FP64
plans one after another, each checked against dense FP64 rows).

Prints one JSON line (N=1).
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

HEADLINE_METRIC = "HGEMM eval TFLOP/s+ms, mixed vs FP64 plain, N=64K nrhs=1024, rel err<=1e-5"

WORKLOADS = {
    "h2b_n65536": dict(name="h2b_gmm16_n65536_d28", n=65536, d=28, ncomp=16, sig_lo=0.5, sig_hi=2.0, diag=True,
                       h=1.5, leaf=256, rank=256, tol=1e-5, budget=0.03, nrhs=1024, precision="bf16x3",
                       seed_points=1002, seed_w=2002, seed_rows=3002, check_rows=1000),
    "hss_n16384": dict(name="hss_gmm8_n16384_d18", n=16384, d=18, ncomp=8, sig_lo=1.0, sig_hi=1.0, diag=False,
                       h=1.0, leaf=128, rank=128, tol=1e-5, budget=0.0, nrhs=256, precision="bf16x3",
                       seed_points=1001, seed_w=2001, seed_rows=3001, check_rows=1000),
    "adjoint_n32768": dict(name="adjoint_gmm32_n32768_d54", n=32768, d=54, ncomp=32, sig_lo=1.0, sig_hi=1.0,
                           diag=False, h=2.0, leaf=256, rank=192, tol=1e-5, budget=0.03, nrhs=128,
                           precision="fp32", seed_points=1004, seed_w=2004, seed_rows=3004, check_rows=1000,
                           seed_target=4004, seed_dirs=5004),
}


BATCHED = dict(name="batched64_h2b_gmm16", count=64, n_lo=2048, n_hi=32768, dims=(8, 18, 28, 54), h_lo=0.8,
               h_hi=2.0, leaves=(128, 256), rank=128, tol=1e-5, budget=0.03, nrhs=(32, 128, 512), ncomp=16,
               sig_lo=0.5, sig_hi=2.0, seed=3, seed_points=1003, seed_w=2003, seed_rows=3003, check_rows=100)


SCALE = dict(name="h2b_gmm64_n524288_d27", n=524288, d=27, ncomp=64, sig_lo=0.5, sig_hi=1.5, diag=False, h=1.2,
             leaf=256, rank=256, tol=1e-5, budget=0.03, nrhs=512, seed_points=1005, seed_w=2005, seed_rows=3005,
             check_rows=2000)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="h2b_n65536", choices=sorted(WORKLOADS) + ["batched64", "scale_n524288"])
    ap.add_argument("--nrhs", type=int, default=0, help="override the workload's nrhs")
    ap.add_argument("--precision", default="", choices=["", "bf16x3", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-nrhs", type=int, default=0,
                    help="columns of the CPU baseline's evaluation (0: the workload's nrhs)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a plain 1-buffer copy of this size (HBM-counter calibration)")
    ap.add_argument("--skip-checks", action="store_true", help="skip the dense-row error checks")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint and backward timings")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a hipGraph")
    ap.add_argument("--launch-variants", default="",
                    help="comma-separated BF16x3 kernel variant per forward launch under --schedule, "
                         "replacing the autotuner's choice (to profile a fixed configuration)")
    ap.add_argument("--fp64-seg-variants", default="",
                    help="SCHEDULE:v1,v2,... FP64 forward schedule and kernel variant per launch, replacing the "
                         "autotuner's choice (to profile a fixed configuration)")
    ap.add_argument("--schedule", default="fused", choices=["fused", "split", "fused_balanced", "split_balanced", "coarse",
                                                            "coarse_split"],
                    help="forward schedule used with --launch-variants")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip timing the drop-in HGEMM entry point on host arrays (PCIe-inclusive)")
    return ap.parse_args()


class Timer:
    """Mean time of a callable over `steps` calls after `warmup` calls, with
    events on the stream the evaluations are enqueued on."""

    def __init__(self, torch, stream):
        self.torch = torch
        self.stream = stream

    def __call__(self, fn, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(self.stream)
        for _ in range(steps):
            fn()
        e1.record(self.stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps


def build_matrix(wl, host):
    t0 = time.time()
    X = host.gaussian_mixture(wl["n"], wl["d"], wl["ncomp"], sig_lo=wl["sig_lo"], sig_hi=wl["sig_hi"],
                              diag=wl["diag"], seed=wl["seed_points"])
    hm = host.build_hmatrix(X, leaf_size=wl["leaf"], max_rank=wl["rank"], tol=wl["tol"], budget=wl["budget"],
                            h=wl["h"])
    t_build = time.time() - t0
    log(f"[bench] host build {t_build:.1f}s: nleaf={hm.nleaf} ncount={hm.ncount} fcount={hm.fcount} "
        f"ranks {hm.slen[1:].min()}..{hm.slen[1:].max()} hss={hm.is_hss}")
    return X, hm, t_build


def launch_profile(plan, run, prec, nrhs, parts, nprof=3):
    """Per-launch durations (HIP events on the evaluation stream) -> per-kind
    totals and the roofline entry of the dominant kernel."""
    plan.set_profiling(True)
    per_kind = {}
    per_launch = None
    for _ in range(nprof):
        run()
        lt = plan.launch_times()
        per_launch = [ms for _, ms in lt] if per_launch is None else [a + ms for a, (_, ms) in zip(per_launch, lt)]
        for k, ms in lt:
            per_kind.setdefault(k, []).append(ms)
    plan.set_profiling(False)
    log("[bench] per-launch ms: " + " ".join(f"{x / nprof:.3f}" for x in per_launch))
    log("[bench] per-launch GFLOP: " + " ".join(f"{f / 1e9:.1f}" for f in plan.launch_flops(prec, nrhs)))
    kind_total = {k: sum(v) / nprof for k, v in per_kind.items()}
    kind_launches = {k: len(v) // nprof for k, v in per_kind.items()}
    dom = max((k for k in kind_total if k != "convert"), key=lambda k: kind_total[k])
    dom_flops = parts.get(dom, 0.0)
    dom_avg_ms = kind_total[dom] / max(1, kind_launches[dom])
    dom_flops_per_launch = dom_flops / max(1, kind_launches[dom])
    achieved = dom_flops_per_launch / (dom_avg_ms * 1e-3) / 1e12
    peak = {"bf16x3": MI355X_PEAK["bf16"], "gemm_f32": MI355X_PEAK["fp32"],
            "gemm_f64": MI355X_PEAK["fp64"]}[dom]
    roofline = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 5), "traffic": None,
                "launches_per_eval": kind_launches[dom], "avg_launch_ms": round(dom_avg_ms, 5),
                "flops_per_launch": dom_flops_per_launch}
    if dom == "bf16x3":
        roofline["frac_effective_x3"] = round(achieved / (peak / 3), 5)
    kinds = {k: {"ms_per_eval": round(kind_total[k], 4), "launches": kind_launches[k],
                 "tflops": round(parts.get(k, 0.0) / (kind_total[k] * 1e-3) / 1e12, 3) if parts.get(k) else None}
             for k in kind_total}
    log(f"[bench] kernels: {kinds}")
    return roofline, kinds


def dense_checks(plan, hm, wl, Wd, nrhs, results, np, torch):
    rng = np.random.default_rng(wl["seed_rows"])
    rows = np.sort(rng.choice(hm.N, min(wl["check_rows"], hm.N), replace=False)).astype(np.int32)
    dense = plan.dense_rows(rows, Wd, nrhs)
    nd = torch.linalg.vector_norm(dense)
    checks = {}
    arows = np.arange(min(10, int(hm.Ddim[0])), dtype=np.int32)
    da = plan.dense_rows(arows, Wd, nrhs)
    for name, Y in results.items():
        r = plan.rows_of(Y, rows, nrhs).double()
        checks[f"rel_err_{name}_vs_dense"] = float(torch.linalg.vector_norm(r - dense) / nd)
        # `acc` (computeError, matroxTest-V1/Util.h:276-332): first 10 points of leaf 0
        ya = plan.rows_of(Y, arows, nrhs).double()
        checks[f"acc_{name}"] = float(((torch.linalg.vector_norm(ya - da, dim=1) /
                                        torch.linalg.vector_norm(da, dim=1)).mean()))
    log(f"[bench] checks: {checks}")
    return checks


def host_cores():
    """(threads the oracle's OpenMP loops use, physical cores among the CPUs
    this process may run on)."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id") as f:
                core = f.read().strip()
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id") as f:
                pkg = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    threads = int(os.environ.get("OMP_NUM_THREADS", len(cpus)))
    return threads, len(cores)


def cpu_baseline(hm, wl, nrhs, ns, host, own_ns=128):
    """The oracle (restated reference evaluation, OpenMP loops) on the host
    cores, on the same compressed matrix with ns columns (the workload's
    nrhs unless a smaller sample is asked for).  Its block products go
    through MKL cblas_dgemm (sequential, inside the OpenMP loops, as the
    reference runs) when the MKL runtime is present, else through the
    oracle's own GEMM; the own GEMM is also timed, on own_ns columns."""
    from oracle import oracle

    def timed(cols):
        Ws = host.normal((hm.N, cols), seed=wl["seed_w"] + 1)
        TWs = hm.transform(Ws)
        oracle.hgemm(hm, TWs, cols)
        ts = []
        for _ in range(3):
            t1 = time.time()
            oracle.hgemm(hm, TWs, cols)
            ts.append(time.time() - t1)
        return statistics.median(ts), oracle.flops(hm, cols)

    own_ns = min(own_ns, ns)
    t_own, fl_own = timed(own_ns)
    mkl = oracle.use_mkl()
    t_mkl = fl = None
    if mkl:
        try:
            t_mkl, fl = timed(ns)
        finally:
            oracle.use_blas(None)
    if t_mkl is None:
        t_mkl, fl = timed(ns)   # the own GEMM at the full sample
    blas = f"MKL cblas_dgemm ({mkl}, sequential)" if mkl else "own column-major GEMM"
    threads, phys = host_cores()
    cpu = {"value": round(fl / t_mkl / 1e12, 5), "unit": "TFLOP/s", "cores": threads, "kind": "port",
           "physical_cores": phys,
           "sample": f"oracle FP64 evaluation (restated HGEMM_gen.h sweeps, OpenMP on {threads} threads, {blas}) "
                     f"of the same compressed matrix with nrhs={ns}; median of 3 runs after one = {t_mkl:.3f} s",
           "blas": blas,
           "ms_per_eval": round(t_mkl * 1e3 * nrhs / ns, 1),
           "own_gemm_tflops": round(fl_own / t_own / 1e12, 5), "own_gemm_nrhs": own_ns}
    log(f"[bench] cpu baseline: {cpu}")
    return cpu


def adjoint_timings(plan, timer, Wd, nrhs, prec, ms_fwd_mixed, ms_fwd_64, steps, torch):
    """Adjoint dW = K~^T dY (mixed and FP64) and the backward pass with block
    gradients (FP32 and FP64), on the same matrix and sizes as the forward."""
    dY = Wd   # any N x nrhs input; the timing does not depend on its values
    G = torch.empty_like(Wd)
    Y = torch.empty_like(Wd)
    out = {}
    ms_adj = timer(lambda: plan.eval_adjoint(dY, G, nrhs, prec), steps, 2)
    G64 = torch.empty_like(Wd)
    ms_adj64 = timer(lambda: plan.eval_adjoint(dY, G64, nrhs, "fp64"), max(3, steps // 4), 1)
    d = G - G64
    out["rel_err_adjoint_mixed_vs_fp64"] = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(G64))
    del d
    out.update({"adjoint_mixed_ms": round(ms_adj, 4), "adjoint_mixed_ratio_vs_forward": round(ms_adj / ms_fwd_mixed, 3),
                "adjoint_fp64_ms": round(ms_adj64, 4), "adjoint_fp64_ratio_vs_forward": round(ms_adj64 / ms_fwd_64, 3)})
    # backward with block gradients: forward skeletons from a forward evaluation
    grad_flops = plan.grad_flops(nrhs)
    for gp in ("bf16x3", "fp32", "fp64"):
        if gp not in plan.precisions:
            continue
        plan.eval(Wd, Y, nrhs, gp)
        ws, us = plan.skeletons(nrhs, torch.float64 if gp == "fp64" else torch.float32)
        ms_b = timer(lambda: plan.backward(Wd, dY, nrhs, gp, skeletons=(ws, us)), max(3, steps // 4), 1)
        ms_fwd = timer(lambda: plan.eval(Wd, Y, nrhs, gp), max(3, steps // 4), 1)
        plan.set_profiling(True)
        plan.backward(Wd, dY, nrhs, gp, skeletons=(ws, us))
        lt = plan.launch_times()
        plan.set_profiling(False)
        ms_g = sum(ms for k, ms in lt if k.startswith("grad"))
        out[f"backward_{gp}_ms"] = round(ms_b, 4)
        out[f"forward_{gp}_ms"] = round(ms_fwd, 4)
        out[f"backward_{gp}_ratio_vs_forward"] = round(ms_b / ms_fwd, 3)
        out[f"block_grad_{gp}_ms"] = round(ms_g, 4)
        out[f"block_grad_{gp}_tflops"] = round(grad_flops / (ms_g * 1e-3) / 1e12, 3) if ms_g > 0 else None
        del ws, us
    out["block_grad_tflop"] = round(grad_flops / 1e12, 5)
    log(f"[bench] adjoint: {out}")
    return out


def dropin_timings(hm, Wd, nrhs, flops64, np, hgemm, reps=3):
    """The reference's own call (matroxTest-V1/main.cpp:94-108): HGEMM on
    host arrays, FP64, timed around the call alone as the reference driver
    times it.  Every call copies W in and apres, wskel and uskel out; the
    first call also builds the device plan.  Host buffers are allocated and
    touched before the timing; uskel and apres are zeroed between calls (the
    reference's precondition, main.cpp:47-48)."""
    call = hgemm.Dropin(hm, nrhs)
    TWh = Wd.cpu().numpy()
    bufs = [np.full(n, 0.0) for n in (hm.N * nrhs, call.sk, call.sk)]   # apres, wskel, uskel (touched)

    def once():
        bufs[0][:] = 0.0
        bufs[2][:] = 0.0
        t0 = time.perf_counter()
        call(TWh, *bufs)
        return (time.perf_counter() - t0) * 1e3

    out = {}
    for mode in ("verify", "trust"):
        hgemm.set_hgemm_cache(mode)   # drops the kept plan: the next call builds it
        first = once()
        ms = statistics.median(once() for _ in range(reps))
        out[mode] = {"first_call_ms": round(first, 1), "ms": round(ms, 1), "tflops": round(flops64 / ms / 1e9, 3)}
    hgemm.set_hgemm_cache("verify")
    moved = (hm.N * nrhs + (hm.N * nrhs + 2 * call.sk)) * 8
    out["bytes_host_device"] = moved
    out["note"] = ("HGEMM (drop-in) on host arrays, FP64; per call: W in, apres/wskel/uskel out "
                   f"({moved / 1e9:.2f} GB over PCIe); verify = blocks hashed on the host each call, "
                   "trust = blocks assumed unchanged (mx_hgemm_cache)")
    log(f"[bench] drop-in HGEMM on host arrays: {out}")
    return out


def fd_check(plan, hm, wl, nrhs, torch, np, host, ndir=64, eps=1e-6):
    """Gradients of L = 1/2|K~W - T|^2 (W, T FP32; FP32, FP64 and, when the
    plan has it, split-BF16x3 backward) against FP64 central differences on
    ndir random directions (a quarter each in W, D, B, VT), and dW against the
    oracle on the transposed structure."""
    from matrox_ru_amd.autograd import hgemm_apply
    from oracle import oracle
    dev = torch.device("cuda:0")
    W32 = torch.from_numpy(host.normal((hm.N * nrhs,), seed=wl["seed_w"], dtype=np.float32)).to(dev)
    T32 = torch.from_numpy(host.normal((hm.N * nrhs,), seed=wl["seed_target"], dtype=np.float32)).to(dev)
    D0 = torch.from_numpy(np.ascontiguousarray(hm.D)).to(dev)
    B0 = torch.from_numpy(np.ascontiguousarray(hm.B)).to(dev)
    VT0 = torch.from_numpy(np.ascontiguousarray(hm.VT)).to(dev)
    res = {}
    grads = {}
    gps = [gp for gp in ("fp32", "fp64", "bf16x3") if gp in plan.precisions]
    for gp in gps:
        Wg = (W32.double() if gp == "fp64" else W32).clone().requires_grad_()
        D, B, VT = (t.clone().requires_grad_() for t in (D0, B0, VT0))
        Y = hgemm_apply(plan, Wg, D, B, VT, nrhs=nrhs, precision=gp)
        L = 0.5 * ((Y - T32.to(Y.dtype)) ** 2).sum()
        L.backward()
        grads[gp] = [Wg.grad.double(), D.grad, B.grad, VT.grad]
        if gp == "fp32":
            dY = (Y.detach().double() - T32.double()).cpu().numpy()
    Wo, _, _ = oracle.hgemm(hm.transposed(), dY, nrhs)
    res["rel_err_dW_fp32_vs_oracle_transposed"] = float(np.linalg.norm(grads["fp32"][0].cpu().numpy() - Wo)
                                                        / np.linalg.norm(Wo))
    W64, T64 = W32.double(), T32.double()
    lens = [hm.N * nrhs, len(hm.D), len(hm.B), len(hm.VT)]
    an = {gp: [] for gp in gps}
    fd = []
    group = []
    for i in range(ndir):
        g = i % 4
        if lens[g] == 0:
            continue
        r = torch.from_numpy(host.normal((lens[g],), seed=wl["seed_dirs"] + i)).to(dev)
        r /= r.norm()
        v = [torch.zeros(n, dtype=torch.float64, device=dev) if j != g else r for j, n in enumerate(lens)]
        Ys = []
        for s in (1.0, -1.0):
            plan.set_blocks(D0 + s * eps * v[1], B0 + s * eps * v[2], VT0 + s * eps * v[3])
            Y = torch.empty_like(W64)
            plan.eval(W64 + s * eps * v[0], Y, nrhs, "fp64")
            Ys.append(Y)
        fd.append(0.5 * float(((Ys[0] - Ys[1]) * (Ys[0] + Ys[1] - 2 * T64)).sum()) / (2 * eps))
        for gp in an:
            an[gp].append(float(torch.dot(grads[gp][g], r)))
        group.append(g)
    plan.set_blocks(D0, B0, VT0)
    fd = np.array(fd)
    group = np.array(group)
    for gp in an:
        a = np.array(an[gp])
        res[f"fd_rel_err_{gp}"] = float(np.linalg.norm(a - fd) / np.linalg.norm(fd))
        for g, nm in enumerate(["W", "D", "B", "VT"]):
            m = group == g
            if m.any():
                res[f"fd_rel_err_{gp}_{nm}"] = float(np.linalg.norm(a[m] - fd[m]) / np.linalg.norm(fd[m]))
    res["fd_directions"] = int(len(fd))
    res["fd_eps"] = eps
    log(f"[bench] fd check: {res}")
    return res


def batched_instances(wl, host, np):
    """The 64 instance descriptions (seeded; instance i offsets seeds by 100 i)."""
    rng = np.random.default_rng(wl["seed"])
    inst = []
    for i in range(wl["count"]):
        n = int(round(np.exp(rng.uniform(np.log(wl["n_lo"]), np.log(wl["n_hi"])))))
        inst.append(dict(n=n, d=int(rng.choice(wl["dims"])), h=float(rng.uniform(wl["h_lo"], wl["h_hi"])),
                         leaf=int(rng.choice(wl["leaves"])), nrhs=int(rng.choice(wl["nrhs"])),
                         seed_points=wl["seed_points"] + 100 * i, seed_w=wl["seed_w"] + 100 * i,
                         seed_rows=wl["seed_rows"] + 100 * i))
    return inst


def run_batched(args):
    """Many-instance evaluation: every instance of one nrhs in one tree per
    group, groups concurrent, the whole batch replayed from one hipGraph."""
    import numpy as np
    import torch
    from matrox_ru_amd import hgemm, host
    wl = dict(BATCHED)
    inst = batched_instances(wl, host, np)
    t0 = time.time()
    hms, Xs = [], []
    for it in inst:
        X = host.gaussian_mixture(it["n"], it["d"], wl["ncomp"], sig_lo=wl["sig_lo"], sig_hi=wl["sig_hi"], diag=True,
                                  seed=it["seed_points"])
        hms.append(host.build_hmatrix(X, leaf_size=it["leaf"], max_rank=wl["rank"], tol=wl["tol"],
                                      budget=wl["budget"], h=it["h"]))
        Xs.append(X)
    t_build = time.time() - t0
    log(f"[bench] host build of {len(hms)} instances {t_build:.1f}s; total N {sum(hm.N for hm in hms)}; "
        f"HSS-degenerate {sum(int(hm.is_hss) for hm in hms)}")
    prec = args.precision or "bf16x3"
    b = hgemm.Batch(hms, [it["nrhs"] for it in inst], precisions=("fp64", "fp32", "bf16x3"))
    log(f"[bench] groups (nrhs, rows): {b.groups}")
    W64 = b.buffers(torch.float64, fill=0.0)
    for i, (hm, it) in enumerate(zip(hms, inst)):
        b.instance_view(W64, i).copy_(torch.from_numpy(hm.transform(host.normal((hm.N, it["nrhs"]),
                                                                                seed=it["seed_w"]))))
    W32 = [w.float() for w in W64]
    Y64 = b.buffers(torch.float64)
    Y32 = b.buffers(torch.float32)
    Ymx = b.buffers(torch.float32)
    for p in ("fp32", "bf16x3"):
        b.autotune(p, "f32", reps=2)
    b.autotune("fp64", "f64", reps=1)
    b.set_graph(not args.no_graph)
    timer = Timer(torch, torch.cuda.current_stream())
    ms64 = timer(lambda: b.eval(W64, Y64, "fp64"), max(3, args.steps // 4), 1)
    ms32 = timer(lambda: b.eval(W32, Y32, "fp32"), args.steps, args.warmup)
    msmx = timer(lambda: b.eval(W32, Ymx, "bf16x3"), args.steps, args.warmup)
    # the batched sweep (every BF16x3 group in one launch) against the
    # groups' own concurrent launches, whichever the autotuner kept
    sweep_pick = b.get_sweep("forward")
    launch_forms = {}
    Yt = b.buffers(torch.float32)
    for v in (0, 1, 5, 6):
        b.set_sweep("forward", v)
        launch_forms["one_launch_w" + str(v) if v else "group_launches"] = round(
            timer(lambda: b.eval(W32, Yt, "bf16x3"), args.steps, args.warmup), 4)
    b.set_sweep("forward", sweep_pick)
    del Yt
    log(f"[bench] batch bf16x3 launch forms (ms): {launch_forms}; autotuned sweep width {sweep_pick}")
    fl = b.flops("fp64")
    log(f"[bench] batch fp64 {ms64:.3f} ms, fp32 {ms32:.3f} ms, bf16x3 {msmx:.3f} ms; {fl / 1e12:.4f} TFLOP")
    # kernel roofline per group, each group's plan run alone (HIP events per
    # launch); small nrhs is HBM-bound on the packed blocks: every 4-byte
    # split element serves 2 nrhs flops, so the ceiling is min(BF16 peak / 3,
    # nrhs / 2 flop/B x 8 TB/s)
    roof_groups = []
    x3_fl, x3_ms = 0.0, 0.0
    for g, (nr, rows) in enumerate(b.groups):
        gp = b.plan(g)
        try:
            _, parts_g = gp.flops("bf16x3", nr)
        except hgemm.MxError:   # a group the accuracy guard routed to FP32 has no BF16x3 lists
            roof_groups.append({"group_nrhs": nr, "group_rows": rows, "precision": "fp32 (accuracy guard)"})
            continue
        Yg = torch.empty_like(W32[g])
        roof, kinds = launch_profile(gp, lambda: gp.eval(W32[g], Yg, nr, "bf16x3"), "bf16x3", nr, parts_g, nprof=2)
        ceiling = min(MI355X_PEAK["bf16"] / 3, nr / 2 * MI355X_PEAK["hbm_gbs"] / 1e3)
        roof.update({"group_nrhs": nr, "group_rows": rows, "ceiling_x3": round(ceiling, 1),
                     "bound": "hbm" if ceiling < MI355X_PEAK["bf16"] / 3 else "mfma",
                     "frac_of_ceiling": round(roof["achieved"] / ceiling, 4)})
        roof_groups.append(roof)
        x3_fl += roof["flops_per_launch"] * roof["launches_per_eval"]
        x3_ms += roof["avg_launch_ms"] * roof["launches_per_eval"]
        del Yg
    errs32, errsmx, dense = [], [], []
    for i, (hm, it) in enumerate(zip(hms, inst)):
        y64 = b.instance_view(Y64, i)
        n64 = torch.linalg.vector_norm(y64)
        errs32.append(float(torch.linalg.vector_norm(b.instance_view(Y32, i).double() - y64) / n64))
        errsmx.append(float(torch.linalg.vector_norm(b.instance_view(Ymx, i).double() - y64) / n64))
        if not args.skip_checks:
            b.set_points(i, Xs[i][hm.lids[:hm.N]], it["h"])
            rows = np.sort(np.random.default_rng(it["seed_rows"]).choice(hm.N, min(wl["check_rows"], hm.N),
                                                                          replace=False)).astype(np.int32)
            dr = b.dense_rows(i, rows, b.instance_view(W64, i))
            r = hgemm.leaf_rows(hm.wrow, hm.Ddim, y64, rows, it["nrhs"])
            dense.append(float(torch.linalg.vector_norm(r - dr) / torch.linalg.vector_norm(dr)))
    per = [{"N": hm.N, "d": it["d"], "h": round(it["h"], 3), "leaf": it["leaf"], "nrhs": it["nrhs"],
            "hss": bool(hm.is_hss), "rel_err_fp32_vs_fp64": errs32[i], "rel_err_bf16x3_vs_fp64": errsmx[i],
            "rel_err_fp64_vs_dense": dense[i] if dense else None} for i, (hm, it) in enumerate(zip(hms, inst))]
    ms = msmx if prec == "bf16x3" else ms32
    out = {
        "metric": f"batched HGEMM eval TFLOP/s+ms, {len(hms)} instances, {prec} vs FP64 plain",
        "value": round(fl / (ms * 1e-3) / 1e12, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-f32acc" if prec == "bf16x3" else "f32",
        "data": "synthetic (splitmix64/Box-Muller Gaussian mixture points, N(0,1) W), FP32 storage",
        "graph": not args.no_graph,
        "config": {"workload": wl["name"], "instances": len(hms), "total_N": sum(hm.N for hm in hms),
                   "groups": [{"nrhs": g[0], "rows": g[1]} for g in b.groups]},
        "algorithmic_tflop_per_eval": round(fl / 1e12, 5),
        "fp64_ms": round(ms64, 4), "fp32_ms": round(ms32, 4), "bf16x3_ms": round(msmx, 4),
        "bf16x3_launch_forms_ms": launch_forms, "batched_sweep_width": sweep_pick,
        "fp32_tflops": round(fl / (ms32 * 1e-3) / 1e12, 3), "bf16x3_tflops": round(fl / (msmx * 1e-3) / 1e12, 3),
        "fp64_tflops": round(fl / (ms64 * 1e-3) / 1e12, 3),
        "max_rel_err_fp32_vs_fp64": max(errs32), "max_rel_err_bf16x3_vs_fp64": max(errsmx),
        "median_rel_err_bf16x3_vs_fp64": float(np.median(errsmx)),
        "roofline": {"bound": "mfma", "kernel": "bf16x3", "unit": "TFLOP/s", "peak": MI355X_PEAK["bf16"],
                     "achieved": round(x3_fl / (x3_ms * 1e-3) / 1e12, 3),
                     "frac": round(x3_fl / (x3_ms * 1e-3) / 1e12 / MI355X_PEAK["bf16"], 5), "traffic": None,
                     "note": "groups run one at a time; per-group ceilings in roofline_groups"},
        "roofline_groups": roof_groups,
        "instances": per,
        "host_build_s": round(t_build, 1),
    }
    print(json.dumps(out), flush=True)


def run_scale(args):
    """Precision at scale: the blocks (FP64 ~160 GB) are never stored on the
    host; each precision's plan generates them on the device in its own
    representation, one plan at a time."""
    import numpy as np
    import torch
    from matrox_ru_amd import hgemm, host
    wl = dict(SCALE)
    nrhs = args.nrhs or wl["nrhs"]
    t0 = time.time()
    X = host.gaussian_mixture(wl["n"], wl["d"], wl["ncomp"], sig_lo=wl["sig_lo"], sig_hi=wl["sig_hi"],
                              diag=wl["diag"], seed=wl["seed_points"])
    hm = host.build_hmatrix(X, leaf_size=wl["leaf"], max_rank=wl["rank"], tol=wl["tol"], budget=wl["budget"],
                            h=wl["h"], skip_blocks=1)
    t_build = time.time() - t0
    log(f"[bench] host build {t_build:.1f}s: nleaf={hm.nleaf} ncount={hm.ncount} fcount={hm.fcount} "
        f"D {hm.scalars['D_len'] * 8 / 1e9:.1f} GB B {hm.scalars['B_len'] * 8 / 1e9:.1f} GB (FP64, not built)")
    W = torch.from_numpy(hm.transform(host.normal((hm.N, nrhs), seed=wl["seed_w"]))).cuda()
    rows = np.sort(np.random.default_rng(wl["seed_rows"]).choice(hm.N, wl["check_rows"], replace=False))
    rows = rows.astype(np.int32)
    timer = Timer(torch, torch.cuda.current_stream())
    res, Y64, dense = {}, None, None
    for prec in ("bf16x3", "fp32", "fp64"):
        t1 = time.time()
        plan = hgemm.Plan(hm, precisions=(prec,), generate=(X, wl["h"]))
        t_plan = time.time() - t1
        Wp = W if prec != "fp32" else W.float()
        Y = torch.empty_like(Wp)
        t1 = time.time()
        plan.autotune(prec, nrhs, "f32" if prec == "fp32" else "f64", reps=1)
        t_tune = time.time() - t1
        plan.set_graph(not args.no_graph)
        steps = max(2, args.steps // (8 if prec == "fp64" else 2))
        ms = timer(lambda: plan.eval(Wp, Y, nrhs, prec), steps, 1)
        fl, parts = plan.flops(prec, nrhs)
        roof, _ = launch_profile(plan, lambda: plan.eval(Wp, Y, nrhs, prec), prec, nrhs, parts, nprof=1)
        r = {"ms": round(ms, 3), "tflops": round(fl / (ms * 1e-3) / 1e12, 3), "plan_s": round(t_plan, 1),
             "autotune_s": round(t_tune, 1), "schedule": plan.get_schedule("forward", prec), "roofline": roof}
        if prec == "fp64":
            Y64 = Y
            if not args.skip_checks:
                plan.set_points(X[hm.lids[:hm.N]], wl["h"])
                dense = plan.dense_rows(rows, W, nrhs)
        res[prec] = (r, Y.double() if prec != "fp64" else Y)
        log(f"[bench] {prec}: {r}")
        plan.close()
        del plan, Wp
        torch.cuda.empty_cache()
    fl64 = fl
    out_prec = {}
    for prec, (r, Y) in res.items():
        if prec != "fp64":
            r["rel_err_vs_fp64"] = float(torch.linalg.vector_norm(Y - Y64) / torch.linalg.vector_norm(Y64))
        if dense is not None:
            yr = hgemm.leaf_rows(hm.wrow, hm.Ddim, Y, rows, nrhs)
            r["rel_err_vs_dense"] = float(torch.linalg.vector_norm(yr - dense) / torch.linalg.vector_norm(dense))
        out_prec[prec] = r
    log(f"[bench] scale: {out_prec}")
    ms = out_prec["bf16x3"]["ms"]
    out = {
        "metric": f"HGEMM eval TFLOP/s+ms at N=524288 nrhs={nrhs}, bf16x3 vs FP32 vs FP64, generated blocks",
        "value": round(fl64 / (ms * 1e-3) / 1e12, 3),
        "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3-f32acc",
        "data": "synthetic (splitmix64/Box-Muller Gaussian mixture points, N(0,1) W)",
        "config": {"workload": wl["name"], "N": hm.N, "d": wl["d"], "nrhs": nrhs, "nleaf": int(hm.nleaf),
                   "near_blocks": int(hm.ncount), "far_blocks": int(hm.fcount),
                   "D_GB_fp64": round(hm.scalars["D_len"] * 8 / 1e9, 2),
                   "B_GB_fp64": round(hm.scalars["B_len"] * 8 / 1e9, 2),
                   "mean_rank": round(float(hm.slen[1:].mean()), 2), "check_rows": wl["check_rows"]},
        "algorithmic_tflop_per_eval": round(fl64 / 1e12, 4),
        "precisions": out_prec,
        "roofline": out_prec["bf16x3"]["roofline"],
        "host_build_s": round(t_build, 1),
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this evaluator runs on one GPU; runs on more than one GPU (--gpus > 1) are not "
              "supported (the tree would shard by subtrees, with one exchange of skeleton blocks before the "
              "coupling phase; not built).", file=sys.stderr)
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
    if args.workload == "batched64":
        return run_batched(args)
    if args.workload == "scale_n524288":
        return run_scale(args)
    dev = torch.device("cuda:0")
    wl = dict(WORKLOADS[args.workload])
    nrhs = args.nrhs or wl["nrhs"]
    prec = args.precision or wl["precision"]
    adjoint_wl = args.workload == "adjoint_n32768"

    X, hm, t_build = build_matrix(wl, host)
    W = host.normal((hm.N, nrhs), seed=wl["seed_w"])
    TW = hm.transform(W)
    del W
    precs = ["fp64", prec] + (["fp32"] if prec != "fp32" and not args.no_adjoint else [])
    if adjoint_wl and "bf16x3" not in precs:
        precs.append("bf16x3")   # mixed-precision backward in the finite-difference check
    plan = hgemm.Plan(hm, precisions=tuple(precs), adjoint=not args.no_adjoint or adjoint_wl)
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
    timer = Timer(torch, stream)
    # autotuner: schedule and pipeline depth per direction, on this nrhs
    tuning = {}
    for p in plan.precisions:
        ms = plan.autotune(p, nrhs, "f64", reps=2)
        tuning[p] = {d: plan.get_schedule(d, p) for d in (["forward", "adjoint"] if plan.adjoint else ["forward"])}
        tuning[p]["ms"] = [round(m, 4) for m in ms]
        if p == "bf16x3":
            tuning[p]["launch_variants"] = {d: plan.launch_variants(d)
                                            for d in (["forward", "adjoint"] if plan.adjoint else ["forward"])}
        else:
            tuning[p]["seg_variants"] = {d: plan.seg_variants(p, d)
                                         for d in (["forward", "adjoint"] if plan.adjoint else ["forward"])}
    if args.launch_variants and "bf16x3" in plan.precisions:
        plan.set_schedule("forward", "bf16x3", args.schedule)
        plan.set_launch_variants([int(v) for v in args.launch_variants.split(",")], "forward")
        tuning["bf16x3"]["forward"] = plan.get_schedule("forward", "bf16x3")
        tuning["bf16x3"]["launch_variants"]["forward"] = plan.launch_variants("forward")
        tuning["bf16x3"]["pinned"] = True
    if args.fp64_seg_variants and "fp64" in plan.precisions:
        sched, vs = args.fp64_seg_variants.split(":")
        plan.set_schedule("forward", "fp64", sched)
        plan.set_seg_variants([int(v) for v in vs.split(",")], "fp64", "forward")
        tuning["fp64"]["forward"] = plan.get_schedule("forward", "fp64")
        tuning["fp64"]["seg_variants"]["forward"] = plan.seg_variants("fp64", "forward")
        tuning["fp64"]["pinned"] = True
    log(f"[bench] autotuned: {tuning}")
    plan.set_graph(not args.no_graph)
    ms64 = timer(lambda: plan.eval(Wd, Y64, nrhs, precision="fp64"), max(3, args.steps // 4), 1)
    log(f"[bench] fp64 {ms64:.3f} ms/eval")
    msmx = timer(lambda: plan.eval(Wd, Ymx, nrhs, precision=prec), args.steps, args.warmup)
    log(f"[bench] {prec} {msmx:.3f} ms/eval")
    # FP32 I/O variant of the mixed path (no FP64 conversions)
    W32 = Wd.float()
    Y32 = torch.empty_like(W32)
    ms32io = timer(lambda: plan.eval(W32, Y32, nrhs, precision=prec), args.steps, 2)
    log(f"[bench] {prec} with FP32 I/O {ms32io:.3f} ms/eval")
    del W32

    roofline, kinds_report = launch_profile(plan, lambda: plan.eval(Wd, Ymx, nrhs, precision=prec), prec, nrhs,
                                            parts)
    tfile = os.path.join(ROOT, "profiles", "traffic_h2b_n65536.txt")
    if args.workload == "h2b_n65536" and nrhs == wl["nrhs"] and os.path.exists(tfile):
        kv = dict(l.strip().split("=", 1) for l in open(tfile) if "=" in l and not l.startswith("#"))
        nl = roofline["launches_per_eval"]
        if kv.get("kernel") == roofline["kernel"] and f"fetch_bytes_per_eval_{nl}" in kv:
            roofline["traffic"] = round((float(kv[f"fetch_bytes_per_eval_{nl}"]) +
                                         float(kv[f"write_bytes_per_eval_{nl}"])) / nl, -6)
            roofline["traffic_source"] = "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/x3_kernel_h2b_n65536.md"
            roofline["algorithmic_bytes_per_launch"] = round(4.67e9 / nl, -6)

    d = Ymx - Y64
    err_vs64 = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(Y64))
    del d
    log(f"[bench] rel err mixed vs fp64 {err_vs64:.3e}")
    checks = {}
    if not args.skip_checks:
        checks = dense_checks(plan, hm, wl, Wd, nrhs, {"fp64": Y64, "mixed": Ymx}, np, torch)
    del Y32

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_timings(plan, timer, Wd, nrhs, prec, msmx, ms64, args.steps, torch)
        if adjoint_wl:
            adjoint.update(fd_check(plan, hm, wl, nrhs, torch, np, host))

    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib * (1 << 20) // 8
        a = torch.ones(nb, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calib_copy_mib} MiB done")
        del a, b

    dropin = None
    if not args.no_dropin:
        dropin = dropin_timings(hm, Wd, nrhs, flops64, np, hgemm)

    cpu = None if args.no_cpu_baseline else cpu_baseline(hm, wl, nrhs, min(args.cpu_sample_nrhs or nrhs, nrhs), host)

    metric = HEADLINE_METRIC if args.workload == "h2b_n65536" else \
        f"HGEMM eval TFLOP/s+ms, {prec} vs FP64 plain, {wl['name']} nrhs={nrhs}"
    out = {
        "metric": metric,
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
                   "hss": bool(hm.is_hss), "mean_rank": round(float(hm.slen[1:].mean()), 2),
                   "io": "f64 (conversions timed)"},
        "algorithmic_tflop_per_eval": round(flops64 / 1e12, 5),
        "fp64_ms": round(ms64, 4),
        "fp64_tflops": round(flops64 / (ms64 * 1e-3) / 1e12, 3),
        "fp64_frac_of_peak": round(flops64 / (ms64 * 1e-3) / 1e12 / MI355X_PEAK["fp64"], 4),
        "speedup_vs_fp64": round(ms64 / msmx, 3),
        "mixed_ms_f32_io": round(ms32io, 4),
        "rel_err_mixed_vs_fp64": err_vs64,
        **checks,
        "adjoint": adjoint,
        "tuning": tuning,
        "dropin_host_arrays": dropin,
        "kernels": kinds_report,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_build_s": round(t_build, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
