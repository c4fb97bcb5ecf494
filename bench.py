"""Benchmark: converged pixels/s of the MI355X Milne–Eddington inversion.

A step is one inversion of the whole synthetic map with the profiles already
resident in HBM (me_invert_device); `value` is converged pixels (chi2/normp <=
chisq_criteria) per second of step time.  The map is generated on the device
by the FP64 forward model (Philox truths + noise).  The roofline object uses
the algorithmic op model (fixed per-event FLOP constants x device-counted
events) over the kernel time measured with HIP events on the context stream;
the timed steps run without the FP32 path's per-call Faddeeva region
counting, and one extra untimed step with it supplies the event counts.

Beside the headline (details):
  * plain random-restart LM on the same map and GPU (the learned-vs-plain ratio),
    also at the other restart budgets of num_run 1, 5 and 10 (the headline runs
    the workload's: 5 for hmi4k and niris1k, where configurations 1 and 3 put
    Philox restarts);
  * the reference's own precision (FP64 synthesis, --no-fp64 skips it);
  * the opt-in FP32 SVD (svd_fp32 = 1, the two-waves-per-SIMD kernel build)
    on the plain map: time and how many records match the default FP64-SVD
    records within 0.1 G;
  * the learned run under the reference's LM settings (nine trials per LM
    iteration instead of the learned run's one, S bounded to 1.4 Imax instead
    of 2 Imax; --learned-lambda-limit 0 --learned-s-limit 0 make them the
    headline): time, trials, converged fraction, identical records, chi2
    quantiles;
  * the drop-in path from host buffers (me_invert_batch: staging + kernel +
    records back; --no-e2e skips it);
  * learned records against the FP64 oracle's learned policy fed the same
    network predictions, and against the oracle's plain records of the
    cpu_baseline sample (fraction within 0.1 G / 0.01 deg / 1 m/s, medians);
  * throughput with the network's training charged to this one map, and
    (--mlp-in) the saved-network path that trains nothing.
The cpu_baseline leg times the CPU FP64 oracle on a bounded sample of the
same map on the host's CPU share.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hmi4k]
                       [--precision fp32|fp64] [--pixels N] [--cpu-seconds S]
                       [--mlp-in NET | --save-mlp NET] [--num-run R]
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
DEG = 180.0 / np.pi
TRAFFIC_FILE = "profiles/invert_kernel_hbm_traffic.txt"


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
    out = {"bound": "compute", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TFLOP/s",
           "frac": round(achieved / peak, 4) if peak else 0.0, "traffic": traffic[0] if traffic else None,
           "flops_fp32": f32, "flops_fp64": f64,
           "note": "vector-ALU bound: FP32 synthesis/Faddeeva/Jacobian/normal matrix + FP64 9x9 SVD; "
                   "peak = (F32+F64)/(F32/157.3T + F64/78.6T)"}
    if traffic:
        out["traffic_source"] = traffic[1]
    return out


def hbm_traffic(workload: str, mode: str, precision: str, num_run: int):
    """HBM bytes per inversion launch from the committed rocprofv3 FETCH_SIZE
    / WRITE_SIZE summary (scaled by the plain-copy calibration), with the
    file and the kernel build it was measured on; None when not measured."""
    p = REPO / TRAFFIC_FILE
    if not p.exists():
        return None
    build = ""
    for line in p.read_text().splitlines():
        if line.startswith("# build:"):
            build = line[len("# build:"):].strip()
        f = line.split()
        if len(f) >= 6 and f[:4] == [workload, mode, precision, str(num_run)]:
            return float(f[4]) + float(f[5]), f"{TRAFFIC_FILE} ({build})" if build else TRAFFIC_FILE
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


def host_cpus():
    """Threads for the CPU baseline: the CPUs this process may run on, capped
    by OMP_NUM_THREADS when set (a GPU box gives one GPU's job a share of its
    cores and sets it to that share); plus the machine's CPU count and model."""
    avail = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(omp)) if omp.isdigit() and int(omp) > 0 else avail
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, os.cpu_count() or avail, model


def cpu_baseline(w, cfg, target_s: float):
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O
    threads, ncpu, model = host_cpus()
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
    line = {"value": round(conv / dt, 2), "unit": "converged pixels/s", "cores": threads, "kind": "port",
            "sample": f"first {n} pixels of the same map, FP64 oracle (plain LM, num_run {cfg.num_run}), "
                      f"{threads} threads, {dt:.1f} s",
            "host_cpus": ncpu, "cpu_model": model, "converged_fraction": round(conv / n, 4)}
    return line, out["res"]


def agreement(a, b, crit):
    """Records a against b: fraction of all pixels and of pixels converged in
    both within 0.1 G / 0.01 deg / 1 m/s, and the median differences."""
    def phys(x, y):
        ok = np.abs(x[:, 0] - y[:, 0]) <= 0.1
        ok &= np.abs(x[:, 1] - y[:, 1]) * DEG <= 0.01
        dphi = np.abs(x[:, 2] - y[:, 2])
        ok &= np.minimum(dphi, np.pi - dphi) * DEG <= 0.01
        ok &= np.abs(x[:, 3] - y[:, 3]) <= 1e-3
        return ok
    both = (a[:, 9] > 0) & (a[:, 9] <= crit) & (b[:, 9] > 0) & (b[:, 9] <= crit)
    d = np.abs(a[both] - b[both])
    dphi = np.minimum(d[:, 2], np.pi - d[:, 2]) if both.any() else d[:, 2]
    names = ["B_G", "theta_deg", "phi_deg", "vlos_m_s", "dopp_mA", "damp", "eta0", "S", "beta"]
    scale = [1.0, DEG, DEG, 1e3, 1, 1, 1, 1, 1]
    med = {}
    for k, nm in enumerate(names):
        v = dphi if k == 2 else d[:, k]
        med[nm] = float(np.median(v) * scale[k]) if v.size else None
    return {"pixels": int(len(a)), "within_tol_all": round(float(phys(a, b).mean()), 4),
            "converged_both": int(both.sum()),
            "within_tol_converged_both": round(float(phys(a[both], b[both]).mean()), 4) if both.any() else None,
            "converged_fraction": [round(float(np.mean((a[:, 9] > 0) & (a[:, 9] <= crit))), 4),
                                   round(float(np.mean((b[:, 9] > 0) & (b[:, 9] <= crit))), 4)],
            "median_abs_diff_converged_both": med}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hmi4k")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--pixels", type=int, default=0, help="invert only the first N pixels of the map")
    ap.add_argument("--num-run", type=int, default=0, help="LM restart budget (default: the workload's, 5 for hmi4k)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--svd-fp32", type=int, default=0, choices=[0, 1],
                    help="FP32 path: SVD solve in FP64 (0, default) or FP32 (1, the two-wave build)")
    ap.add_argument("--tune-pixels", type=int, default=1 << 22,
                    help="autotune the launch shape on this many pixels first, about 64 per resident lane "
                         "(0: keep --block/--waves-per-cu)")
    ap.add_argument("--mode", default="both", choices=["plain", "learned", "both"])
    ap.add_argument("--train-samples", type=int, default=8 * 1024 * 1024,
                    help="device-synthesised training profiles (8M x 24 epochs: 6.2 s on hmi4k)")
    ap.add_argument("--train-epochs", type=int, default=24)
    ap.add_argument("--mlp-in", default="", help="use this saved network instead of training one")
    ap.add_argument("--save-mlp", default="", help="write the trained network to this file")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 (reference precision) legs")
    ap.add_argument("--no-e2e", action="store_true", help="skip the host-buffer me_invert_batch leg")
    ap.add_argument("--restarts", type=int, default=1,
                    help="also time plain and learned at the other num_run of 1, 5, 10 (0 skips)")
    ap.add_argument("--oracle-pixels", type=int, default=16384,
                    help="pixels of the learned-vs-oracle check (0 skips it)")
    ap.add_argument("--learned-lambda-limit", type=float, default=4e-3,
                    help="damping bound of the learned run (4e-3: one trial per iteration, 6e-3: two; 0: lambda_limits[1], "
                         "the reference's nine); the fallback's heuristic runs keep the reference bound")
    ap.add_argument("--learned-s-limit", type=float, default=2.0,
                    help="upper bound of S in the learned run in units of the profile's I maximum (0: "
                         "Init_Guess's 1.4, which every heuristic run keeps)")
    ap.add_argument("--no-ref-bound", action="store_true",
                    help="skip the learned run under the reference's LM settings")
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
    from meinv.network import load_network, save_network
    from meinv.workloads import get

    w = get(args.workload)
    n = w.n if args.pixels <= 0 else min(args.pixels, w.n)
    fp64 = args.precision == "fp64"
    over = dict(precision=ME_FP64 if fp64 else ME_FP32, block_size=args.block, waves_per_cu=args.waves_per_cu,
                svd_fp32=args.svd_fp32)
    over["num_run"] = args.num_run if args.num_run else w.bench_num_run
    cfg = w.config(**over)
    over_l = dict(over, init_mode=ME_INIT_LEARNED, learned_lambda_limit=args.learned_lambda_limit,
                  learned_s_limit=args.learned_s_limit)
    inv = Inversion(cfg, w.wavelengths())
    d_prof = torch.empty((n, 4, w.nl), dtype=torch.float32, device="cuda:0")
    d_res = torch.empty((n, 10), dtype=torch.float64, device="cuda:0")
    d_truth = torch.empty((n, 9), dtype=torch.float64, device="cuda:0")
    inv.gen_map_device(w.seed, w.ranges(), n, d_prof.data_ptr(), d_truth.data_ptr(), ff_frac=w.ff_frac,
                       ff_bmax=w.ff_bmax, sigma=w.sigma)
    truth = d_truth.cpu().numpy()
    del d_truth
    torch.cuda.synchronize()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    log(f"bench: {w.name} n={n} nl={w.nl} {args.precision} num_run={cfg.num_run} mode={args.mode}")

    def timed(engine, label, steps, warmup, fp64_path, tune_ok=True, count=True):
        # the launch shape is autotuned first (untimed) unless given; timed
        # steps run without the FP32 path's per-call region counting; one extra
        # untimed step with it gives the event counts of the op model (the
        # inversion is deterministic: every step does the same work)
        engine.set_region_counts(fp64_path)
        tune = None
        if tune_ok and args.tune_pixels > 0 and args.block == 0 and args.waves_per_cu == 0:
            tune = engine.autotune(d_prof.data_ptr(), min(n, args.tune_pixels))
            log(f"{label} autotuned launch: block {tune['block_size']}, {tune['waves_per_cu']} waves per CU "
                f"({tune['kernel_ms']:.1f} ms on {min(n, args.tune_pixels)} px)")
        for i in range(warmup):
            st = engine.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
            log(f"{label} warmup {i}: {st.total_ms:.1f} ms (kernel {st.kernel_ms:.1f}), converged "
                f"{st.n_converged}/{st.n_valid}")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sts = []
        for i in range(steps):
            st = engine.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
            sts.append(st)
            log(f"{label} step {i}: {st.total_ms:.1f} ms (kernel {st.kernel_ms:.1f}), converged "
                f"{st.n_converged}/{st.n_valid}")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        counted = None
        if count:
            engine.set_region_counts(True)
            counted = engine.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
        return wall / steps * 1e3, sts, counted, tune

    def vs_truth(res):
        """Median error against the map's truths over the converged pixels
        (the inverse problem's own degeneracy at chi2 <= chisq_criteria, not
        an engine error)."""
        ok = (res[:, 9] > 0) & (res[:, 9] <= cfg.chisq_criteria)
        d = np.abs(res[ok, :9] - truth[ok])
        dphi = np.minimum(d[:, 2], np.pi - d[:, 2])
        return {"B_G": round(float(np.median(d[:, 0])), 3), "theta_deg": round(float(np.median(d[:, 1]) * DEG), 4),
                "phi_deg": round(float(np.median(dphi) * DEG), 4), "vlos_m_s": round(float(np.median(d[:, 3]) * 1e3), 3)}

    def svd_precision(learned):
        return "fp32" if not fp64 and args.svd_fp32 == 1 else "fp64"

    def summary(ms, sts, tune=None):
        last = sts[-1]
        v = max(1, last.n_valid)
        out = {"converged_pixels_per_s": round(last.n_converged / (ms / 1e3), 1), "ms_per_step": round(ms, 2),
               "converged": int(last.n_converged), "valid": int(last.n_valid),
               "converged_fraction": round(last.n_converged / v, 4),
               "runs_per_pixel": round(last.n_runs / v, 3), "iters_per_pixel": round(last.n_iters / v, 3),
               "trials_per_pixel": round(last.n_trial / v, 3), "fallback_fraction": round(last.n_fallback / v, 4),
               "kernel_ms": round(float(np.mean([s.kernel_ms for s in sts])), 2), "svd_fail": int(last.n_svd_fail)}
        if tune:
            out["launch"] = {"block": tune["block_size"], "waves_per_cu": tune["waves_per_cu"]}
        return out

    def get_network(engine):
        """Train the initialiser on device-synthesised profiles, or import a
        saved one; returns (training seconds or None, final loss or None)."""
        if args.mlp_in:
            params, ranges = load_network(args.mlp_in)
            engine.mlp_import(params, ranges)
            log(f"MLP read from {args.mlp_in}")
            return None, None
        t = time.perf_counter()
        loss = engine.mlp_train(w.seed + 100, w.ranges(), args.train_samples, batch=4096, epochs=args.train_epochs,
                                lr=2e-3, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma)
        train_s = time.perf_counter() - t
        log(f"MLP trained on {args.train_samples} device-synthesised profiles x {args.train_epochs} epochs in "
            f"{train_s:.1f} s, final loss {loss:.4g}")
        if args.save_mlp:
            save_network(args.save_mlp, engine.mlp_params(), w.ranges())
        return train_s, loss

    results = {}
    plain_ms = plain_sts = plain_counted = plain_res = None
    if args.mode in ("plain", "both"):
        plain_ms, plain_sts, plain_counted, tune = timed(inv, "plain", args.steps, args.warmup, fp64)
        plain_res = d_res.cpu().numpy()
        results["plain"] = summary(plain_ms, plain_sts, tune)
        results["plain"]["median_error_vs_truth_converged"] = vs_truth(plain_res)
    learned_ms = learned_sts = learned_counted = None
    linv = None
    learned_res = None
    if args.mode in ("learned", "both"):
        linv = Inversion(w.config(**over_l), w.wavelengths())
        train_s, loss = get_network(linv)
        learned_ms, learned_sts, learned_counted, tune = timed(linv, "learned", args.steps, args.warmup, fp64)
        learned_res = d_res.cpu().numpy()
        results["learned"] = summary(learned_ms, learned_sts, tune)
        results["learned"]["median_error_vs_truth_converged"] = vs_truth(learned_res)
        if train_s is not None:
            conv = learned_sts[-1].n_converged
            results["learned"].update(
                train_s=round(train_s, 2), train_loss=loss, train_profiles=args.train_samples,
                train_epochs=args.train_epochs,
                converged_pixels_per_s_training_charged=round(conv / (learned_ms / 1e3 + train_s), 1))
        else:
            results["learned"]["network"] = f"saved network {args.mlp_in} (no training in this run)"
    headline = "learned" if learned_sts is not None else "plain"
    ms_per_step = learned_ms if headline == "learned" else plain_ms
    sts = learned_sts if headline == "learned" else plain_sts
    counted = learned_counted if headline == "learned" else plain_counted
    last = sts[-1]
    conv = last.n_converged
    value = conv / (ms_per_step / 1e3)
    kernel_s = np.mean([s.kernel_ms for s in sts]) / 1e3
    head_res = learned_res if headline == "learned" else d_res.cpu().numpy()
    line_roof = roofline(counted, w.nl, fp64, kernel_s, hbm_traffic(w.name, headline, args.precision, cfg.num_run))

    # ---- the reference's own precision: FP64 synthesis, the same network
    if not fp64 and not args.no_fp64:
        o64 = dict(over, precision=ME_FP64)
        if args.mode in ("plain", "both"):
            inv64 = Inversion(w.config(**o64), w.wavelengths())
            ms, s64, c64, _ = timed(inv64, "plain fp64", 1, 0, True, tune_ok=False)
            results["plain_fp64"] = summary(ms, s64)
            results["plain_fp64"]["roofline_frac"] = roofline(c64, w.nl, True, s64[-1].kernel_ms / 1e3)["frac"]
            inv64.close()
        if linv is not None:
            l64 = Inversion(w.config(**dict(over_l, precision=ME_FP64)), w.wavelengths())
            l64.mlp_import(linv.mlp_params(), w.ranges())
            ms, s64, c64, _ = timed(l64, "learned fp64", 1, 0, True, tune_ok=False)
            results["learned_fp64"] = summary(ms, s64)
            results["learned_fp64"]["roofline_frac"] = roofline(c64, w.nl, True, s64[-1].kernel_ms / 1e3)["frac"]
            l64.close()

    # ---- the learned run under the reference's LM settings (nine trials per
    # iteration, S bounded to 1.4 Imax as Init_Guess does), same network and
    # launch shape: what the learned run's own bounds change in time, trials,
    # converged fraction and records
    if linv is not None and (args.learned_lambda_limit != 0 or args.learned_s_limit != 0) and not args.no_ref_bound:
        shp = results["learned"].get("launch", {})
        e = Inversion(w.config(**dict(over_l, learned_lambda_limit=0.0, learned_s_limit=0.0,
                                      block_size=shp.get("block", 0), waves_per_cu=shp.get("waves_per_cu", 0))),
                      w.wavelengths())
        e.mlp_import(linv.mlp_params(), w.ranges())
        ms, sr, _, _ = timed(e, "learned, reference LM settings", 1, 1, fp64, tune_ok=False, count=False)
        ref_res = d_res.cpu().numpy()
        key = "learned_reference_settings"
        results[key] = summary(ms, sr)
        results[key]["identical_records"] = round(float(np.mean(np.all(ref_res == learned_res, axis=1))), 5)
        chi_a, chi_b = learned_res[learned_res[:, 9] > 0, 9], ref_res[ref_res[:, 9] > 0, 9]
        q = [0.1, 0.5, 0.9, 0.99]
        results[key]["chi2_quantiles"] = {
            "quantiles": q, "learned_bounds": [round(float(x), 5) for x in np.quantile(chi_a, q)],
            "reference_settings": [round(float(x), 5) for x in np.quantile(chi_b, q)]}
        results[key]["median_error_vs_truth_converged"] = vs_truth(ref_res)
        results["speedup_learned_bounds"] = round(
            results["learned"]["converged_pixels_per_s"] / results[key]["converged_pixels_per_s"], 3)
        e.close()

    # ---- the opt-in FP32 SVD (svd_fp32 = 1: the two-wave kernel build) on the
    # plain map, with its records against the default (FP64 SVD) plain
    # records of the same map
    if "plain" in results and not fp64 and args.svd_fp32 == 0 and args.restarts:
        e = Inversion(w.config(**dict(over, svd_fp32=1)), w.wavelengths())
        ms, sr, _, tune32 = timed(e, "plain, FP32 SVD", 1, 1, fp64, count=False)
        s32 = d_res.cpu().numpy()
        key = "plain_svd_fp32"
        results[key] = summary(ms, sr, tune32)
        both = (s32[:, 9] > 0) & (plain_res[:, 9] > 0)
        dB = np.abs(s32[both, 0] - plain_res[both, 0])
        results[key]["same_as_fp64_svd_within_0.1G"] = round(float(np.mean(dB <= 0.1)), 4)
        results[key]["median_error_vs_truth_converged"] = vs_truth(s32)
        results[key]["speedup_vs_fp64_svd"] = round(results["plain"]["ms_per_step"] / ms, 3)
        e.close()

    # ---- other restart budgets: num_run 1 (no Gaussian jumps), 5 and 10
    # (configuration 3's bound), plain and learned (the fallback runs num_run),
    # on the launch shape tuned above
    if args.restarts:
        shape = {}
        for key in ("plain", "learned"):
            if key in results and "launch" in results[key]:
                shape[key] = dict(block_size=results[key]["launch"]["block"],
                                  waves_per_cu=results[key]["launch"]["waves_per_cu"])
        for nr in [r for r in (1, 5, 10) if r != cfg.num_run]:
            if "plain" in results:
                e = Inversion(w.config(**dict(over, num_run=nr, **shape.get("plain", {}))), w.wavelengths())
                ms, sr, _, _ = timed(e, f"plain num_run={nr}", 1, 0, fp64, tune_ok=False, count=False)
                results[f"plain_num_run_{nr}"] = summary(ms, sr)
                e.close()
            if linv is not None:
                e = Inversion(w.config(**dict(over_l, num_run=nr, **shape.get("learned", {}))), w.wavelengths())
                e.mlp_import(linv.mlp_params(), w.ranges())
                ms, sr, _, _ = timed(e, f"learned num_run={nr}", 1, 0, fp64, tune_ok=False, count=False)
                results[f"learned_num_run_{nr}"] = summary(ms, sr)
                e.close()
            if f"plain_num_run_{nr}" in results and f"learned_num_run_{nr}" in results:
                results[f"speedup_learned_vs_plain_num_run_{nr}"] = round(
                    results[f"learned_num_run_{nr}"]["converged_pixels_per_s"]
                    / results[f"plain_num_run_{nr}"]["converged_pixels_per_s"], 3)

    # ---- the drop-in path: host buffers in and out (me_invert_batch)
    if not args.no_e2e:
        eng = linv if linv is not None else inv
        prof_h = d_prof.cpu().numpy()
        eng.set_region_counts(fp64)
        t = time.perf_counter()
        out = eng.invert(prof_h)
        wall = time.perf_counter() - t
        st = out["stats"]
        results["host_buffers"] = {
            "entry_point": "me_invert_batch (float32 host profiles in, float64 host records out)",
            "mode": headline, "total_ms": round(st.total_ms, 2), "kernel_ms": round(st.kernel_ms, 2),
            "wall_ms": round(wall * 1e3, 2), "converged": int(st.n_converged),
            "converged_pixels_per_s": round(st.n_converged / (st.total_ms / 1e3), 1),
            "bytes_in": int(prof_h.nbytes), "bytes_out": int(n * 80)}
        del prof_h, out

    cpu = None
    cpu_res = None
    if args.cpu_seconds > 0:
        cpu, cpu_res = cpu_baseline(w, cfg, args.cpu_seconds)
    # ---- learned records judged by the FP64 oracle
    if linv is not None and args.oracle_pixels > 0:
        sys.path.insert(0, str(REPO / "oracle"))
        import oracle as O
        m = min(args.oracle_pixels, n)
        d_init = torch.empty((m, 9), dtype=torch.float64, device="cuda:0")
        linv.mlp_predict_device(d_prof.data_ptr(), m, d_init.data_ptr())
        torch.cuda.synchronize()
        pred = d_init.cpu().numpy()
        prof_m = d_prof[:m].cpu().numpy().astype(np.float64)
        threads = host_cpus()[0]
        orc = O.Oracle(w.config(**over_l), w.wavelengths())
        ref = orc.invert(prof_m, init=pred, learned=True, nthreads=threads)["res"]
        results["learned_vs_oracle_learned_policy"] = agreement(learned_res[:m], ref, cfg.chisq_criteria)
        results["learned_vs_oracle_learned_policy"]["oracle"] = (
            "FP64 oracle, learned policy, seeded with the device's own network predictions")
        if cpu_res is not None:
            k = min(len(cpu_res), n)
            results["learned_vs_oracle_plain"] = agreement(learned_res[:k], cpu_res[:k], cfg.chisq_criteria)
            results["learned_vs_oracle_plain"]["oracle"] = "FP64 oracle, plain LM (the cpu_baseline sample)"

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
                   "num_run": cfg.num_run, "chisq_criteria": cfg.chisq_criteria, "precision": args.precision,
                   "learned_lambda_limit": args.learned_lambda_limit if headline == "learned" else None,
                   "learned_s_limit": args.learned_s_limit if headline == "learned" else None,
                   "svd": svd_precision(headline == "learned")},
        "roofline": line_roof,
        "cpu_baseline": cpu,
        "details": dict(results, median_chi2=float(np.median(head_res[:, 9])),
                        speedup_learned_vs_plain=(round(results["learned"]["converged_pixels_per_s"]
                                                        / results["plain"]["converged_pixels_per_s"], 3)
                                                  if "learned" in results and "plain" in results else None)),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
