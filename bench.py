"""Benchmark: converged pixels/s of the MI355X Milne–Eddington inversion.

A step is one inversion of the whole synthetic map with the profiles already
resident in HBM (me_invert_device); `value` is converged pixels (chi2/normp <=
chisq_criteria) per second of step time.  The map is generated on the device
by the FP64 forward model (Philox truths + noise).  The roofline object uses
the algorithmic op model (fixed per-event FLOP constants x device-counted
events) over the kernel time measured with HIP events on the context stream;
the timed steps run without the FP32 path's per-call Faddeeva region
counting, and one extra untimed step with it supplies the event counts.
The cpu_baseline leg times the CPU FP64 oracle on a bounded sample of the
same map on the host's cores.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hmi4k]
                       [--precision fp32|fp64] [--pixels N] [--cpu-seconds S]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "milne-eddington_amd"))

# per-event algorithmic FLOP constants (one FMA = 2, a transcendental = 1)
C_SYN, C_JAC, C_LOSS, C_HESS = 88, 1140, 16, 684
C_F = [274, 97, 12, 20, 35, 45, 60, 4]   # Faddeeva by region: 916, Hui p6, asymptotic 1-4 terms, y = 0, nu = 0
C_SVD, C_BK = 15309, 171
P32, P64 = 157.3e12, 78.6e12             # MI355X dense FP32 / FP64 peaks


def op_model(st, nl: int, fp64: bool):
    f_core = nl * (st.n_jac * (C_SYN + C_JAC + C_LOSS) + st.n_trial * (C_SYN + C_LOSS) + st.n_hess * C_HESS)
    f_fadd = sum(n * c for n, c in zip(st.n_fadd, C_F))
    f_svd = st.n_svd * C_SVD + st.n_bk * C_BK
    if fp64:
        return 0.0, float(f_core + f_fadd + f_svd)
    return float(f_core + f_fadd), float(f_svd)


def roofline(st, nl, fp64, kernel_s, traffic=None):
    f32, f64 = op_model(st, nl, fp64)
    t_min = f32 / P32 + f64 / P64
    achieved = (f32 + f64) / kernel_s / 1e12
    peak = (f32 + f64) / t_min / 1e12 if t_min > 0 else 0.0
    return {"bound": "compute", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4) if peak else 0.0, "traffic": traffic,
            "flops_fp32": f32, "flops_fp64": f64,
            "note": "vector-ALU bound: FP32 synthesis/Faddeeva/Jacobian/normal matrix + FP64 9x9 SVD; "
                    "peak = (F32+F64)/(F32/157.3T + F64/78.6T)"}


def hbm_traffic(workload: str, mode: str, precision: str):
    """HBM bytes per inversion launch, read from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (already scaled by the plain-copy calibration)."""
    p = REPO / "profiles" / "invert_kernel_hbm_traffic.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        f = line.split()
        if len(f) >= 5 and f[:3] == [workload, mode, precision]:
            return float(f[3]) + float(f[4])
    return None


def copy_probe(mib: int):
    """A plain device copy of a known size: what FETCH_SIZE / WRITE_SIZE read
    for it calibrates the counters (a copy reads and writes `mib` MiB)."""
    import torch
    a = torch.ones(mib * 2**18, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy probe: 3 copies of {mib} MiB", flush=True)


def cpu_baseline(w, cfg, target_s: float, threads: int):
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O
    orc = O.Oracle(cfg, w.wavelengths())
    # size the sample: grow a calibration run until it takes ~1/10 of the
    # target, then scale to the target
    n0 = 16 * threads
    while True:
        prof, _ = orc.gen_map(w.seed, w.ranges(), n0, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma)
        t = time.perf_counter()
        orc.invert(prof.astype(np.float64), nthreads=threads)
        dt = time.perf_counter() - t
        if dt >= target_s / 10 or n0 >= w.n:
            break
        n0 = min(w.n, n0 * 4)
    n = int(min(w.n, max(n0, n0 * target_s / max(dt, 1e-3))))
    prof, _ = orc.gen_map(w.seed, w.ranges(), n, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma)
    t = time.perf_counter()
    out = orc.invert(prof.astype(np.float64), nthreads=threads)
    dt = time.perf_counter() - t
    conv = int(np.sum((out["res"][:, 9] > 0) & (out["res"][:, 9] <= cfg.chisq_criteria)))
    return {"value": round(conv / dt, 2), "unit": "converged pixels/s", "cores": threads, "kind": "port",
            "sample": f"first {n} pixels of the same map, FP64 oracle, {threads} threads, {dt:.1f} s",
            "converged_fraction": round(conv / n, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hmi4k")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--pixels", type=int, default=0, help="invert only the first N pixels of the map")
    ap.add_argument("--num-run", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--tune-pixels", type=int, default=1 << 20,
                    help="autotune the launch shape on this many pixels first (0: keep --block/--waves-per-cu)")
    ap.add_argument("--mode", default="both", choices=["plain", "learned", "both"])
    ap.add_argument("--train-samples", type=int, default=4 * 1024 * 1024)
    ap.add_argument("--train-epochs", type=int, default=24)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(pixels are independent: a multi-GPU run would give each GPU a row slab)", file=sys.stderr)
        sys.exit(2)

    import torch
    from meinv import ME_FP32, ME_FP64, ME_INIT_LEARNED, Inversion
    from meinv.workloads import get

    w = get(args.workload)
    n = w.n if args.pixels <= 0 else min(args.pixels, w.n)
    fp64 = args.precision == "fp64"
    over = dict(precision=ME_FP64 if fp64 else ME_FP32, block_size=args.block, waves_per_cu=args.waves_per_cu)
    if args.num_run:
        over["num_run"] = args.num_run
    cfg = w.config(**over)
    inv = Inversion(cfg, w.wavelengths())
    d_prof = torch.empty((n, 4, w.nl), dtype=torch.float32, device="cuda:0")
    d_res = torch.empty((n, 10), dtype=torch.float64, device="cuda:0")
    inv.gen_map_device(w.seed, w.ranges(), n, d_prof.data_ptr(), None, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                       sigma=w.sigma)
    torch.cuda.synchronize()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    log(f"bench: {w.name} n={n} nl={w.nl} {args.precision} num_run={cfg.num_run} mode={args.mode}")

    def timed(engine, label):
        # the launch shape is autotuned first (untimed) unless given; timed
        # steps run without the FP32 path's per-call region counting; one extra
        # untimed step with it gives the event counts of the op model (the
        # inversion is deterministic: every step does the same work)
        engine.set_region_counts(fp64)
        tune = None
        if args.tune_pixels > 0 and args.block == 0 and args.waves_per_cu == 0:
            tune = engine.autotune(d_prof.data_ptr(), min(n, args.tune_pixels))
            log(f"{label} autotuned launch: block {tune['block_size']}, {tune['waves_per_cu']} waves per CU "
                f"({tune['kernel_ms']:.1f} ms on {min(n, args.tune_pixels)} px)")
        for i in range(args.warmup):
            st = engine.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
            log(f"{label} warmup {i}: {st.total_ms:.1f} ms (kernel {st.kernel_ms:.1f}), converged "
                f"{st.n_converged}/{st.n_valid}")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sts = []
        for i in range(args.steps):
            st = engine.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
            sts.append(st)
            log(f"{label} step {i}: {st.total_ms:.1f} ms (kernel {st.kernel_ms:.1f}), converged "
                f"{st.n_converged}/{st.n_valid}")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        engine.set_region_counts(True)
        counted = engine.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
        return wall / args.steps * 1e3, sts, counted, tune

    def summary(ms, sts):
        last = sts[-1]
        v = max(1, last.n_valid)
        return {"converged_pixels_per_s": round(last.n_converged / (ms / 1e3), 1), "ms_per_step": round(ms, 2),
                "converged": int(last.n_converged), "valid": int(last.n_valid),
                "converged_fraction": round(last.n_converged / v, 4),
                "runs_per_pixel": round(last.n_runs / v, 3), "iters_per_pixel": round(last.n_iters / v, 3),
                "trials_per_pixel": round(last.n_trial / v, 3), "fallback_fraction": round(last.n_fallback / v, 4),
                "kernel_ms": round(float(np.mean([s.kernel_ms for s in sts])), 2), "svd_fail": int(last.n_svd_fail)}

    results = {}
    plain_ms = plain_sts = plain_counted = None
    if args.mode in ("plain", "both"):
        plain_ms, plain_sts, plain_counted, tune = timed(inv, "plain")
        results["plain"] = summary(plain_ms, plain_sts)
        if tune:
            results["plain"]["launch"] = {"block": tune["block_size"], "waves_per_cu": tune["waves_per_cu"]}
    learned_ms = learned_sts = learned_counted = None
    train_s = None
    if args.mode in ("learned", "both"):
        linv = Inversion(w.config(init_mode=ME_INIT_LEARNED, **over), w.wavelengths())
        t = time.perf_counter()
        loss = linv.mlp_train(w.seed + 100, w.ranges(), args.train_samples, batch=4096, epochs=args.train_epochs,
                              lr=2e-3, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma)
        train_s = time.perf_counter() - t
        log(f"MLP trained on {args.train_samples} device-synthesised profiles x {args.train_epochs} epochs in "
            f"{train_s:.1f} s, final loss {loss:.4g}")
        learned_ms, learned_sts, learned_counted, tune = timed(linv, "learned")
        results["learned"] = summary(learned_ms, learned_sts)
        if tune:
            results["learned"]["launch"] = {"block": tune["block_size"], "waves_per_cu": tune["waves_per_cu"]}
        results["learned"]["train_s"] = round(train_s, 2)
        results["learned"]["train_loss"] = loss
        results["learned"]["train_profiles"] = args.train_samples
        results["learned"]["train_epochs"] = args.train_epochs
    headline = "learned" if learned_sts is not None else "plain"
    ms_per_step = learned_ms if headline == "learned" else plain_ms
    sts = learned_sts if headline == "learned" else plain_sts
    counted = learned_counted if headline == "learned" else plain_counted
    last = sts[-1]
    conv = last.n_converged
    value = conv / (ms_per_step / 1e3)
    kernel_s = np.mean([s.kernel_ms for s in sts]) / 1e3
    res = d_res.cpu().numpy()
    threads = min(os.cpu_count() or 1, 16)
    cpu = cpu_baseline(w, cfg, args.cpu_seconds, threads) if args.cpu_seconds > 0 else None

    line = {
        "metric": "converged pixels/s at fixed chi2",
        "value": round(value, 1),
        "unit": "converged pixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64" if fp64 else "fp32+fp64",
        "data": "synthetic (Philox truths, FP64 forward model, noise 1e-3 S)",
        "config": {"workload": f"{w.name}: Fe I {w.lambda0} A, nl={w.nl}, {w.nx}x{w.ny} px"
                               + (f", {int(w.ff_frac * 100)}% field-free" if w.ff_frac else ""),
                   "pixels": n, "init": "learned MLP + LM, fallback plain" if headline == "learned" else "heuristic",
                   "num_run": cfg.num_run, "chisq_criteria": cfg.chisq_criteria, "precision": args.precision},
        "roofline": roofline(counted, w.nl, fp64, kernel_s, hbm_traffic(w.name, headline, args.precision)),
        "cpu_baseline": cpu,
        "details": dict(results, median_chi2=float(np.median(res[:, 9])),
                        speedup_learned_vs_plain=(round(results["learned"]["converged_pixels_per_s"]
                                                        / results["plain"]["converged_pixels_per_s"], 3)
                                                  if "learned" in results and "plain" in results else None)),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
