#!/usr/bin/env python3
"""MCMC sweeps/s of vcfpop's Bayesian clustering on one MI355X.

Workload (default): plain drop-in at N=2,000 individuals, L=100,000 biallelic
diploid SNPs packed at 2 bits, K=5, admixture model with alpha Metropolis,
FP64 (synthetic data: Fst 0.05 drift, 2% missing).  A step is one full
sweep (UpdateP, UpdateQ [Metropolis every 10th], UpdateZ, UpdateAlpha,
Record every 10th) with inputs resident in HBM.

Prints one JSON line.  `roofline` is for the dominant kernel (the Z pass),
timed with HIP events on the chain's stream; `cpu_baseline` times the CPU
oracle (C, OpenMP, this build's restatement of the reference path) on a
bounded number of sweeps of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_VALU_PEAK_TOPS = 39.3     # 78.6 TFLOP/s counts an FMA as 2; these kernels issue no FMA


def zpass_algorithmic(N, L, K, bits_per_geno, maxploidy, KT):
    """algorithmic bytes and FP64 ops of one Z pass (per-slot counts: K mul,
    K +eps, K-1 sum, (U-1)*s, <= K-1 subtract and compare)"""
    slots = N * L * maxploidy
    bytes_ = N * L * bits_per_geno / 8 + K * KT * 8 + N * K * 8 + K * KT * 4 + N * K * 4
    fp64 = slots * (5 * K + 2)
    return bytes_, fp64


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="admix_n2000_l100k_k5")
    ap.add_argument("--topology", default="parallel", choices=["parallel", "reference"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB megabytes (scales traffic counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import vcfpop_amd as V
        nbytes = args.copy_calibration * 1000 * 1000
        ms = V.debug_copy(nbytes, iters=5)
        print(json.dumps({"copy_bytes": nbytes, "launches": 6, "ms_per_copy": ms,
                          "GB/s (read+write)": 2 * nbytes / ms / 1e6}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (chains shard across GPUs "
              "with no communication, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import vcfpop_amd as V
    spec = V.WORKLOADS[args.workload]
    K = spec["K"] if isinstance(spec["K"], int) else spec["K"][0]
    print(f"[bench] generating {args.workload}", file=sys.stderr, flush=True)
    s = V.make_synth(args.workload)
    ds = V.Dataset(s, 0)
    topo = V.RNG_PARALLEL if args.topology == "parallel" else V.RNG_REFERENCE
    prec = V.PREC_MIXED if args.precision == "mixed" else V.PREC_FP64
    p = V.params_default(admix_val=1, locpriori_val=1 if spec.get("locpriori") else 2,
                         nburnin=spec["nburnin"], nreps=spec["nreps"], g_seed=1,
                         rng_topology=topo, precision=prec)
    ch = V.Chain(ds, p, K, 1)
    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    ch.step(args.warmup)
    ch.sync()
    ch.enable_timing(True)
    t0 = time.perf_counter()
    ch.step(args.steps)
    ch.sync()
    dt = time.perf_counter() - t0
    z_ms, z_n = ch.kernel_time(0)
    sweeps_s = args.steps / dt

    bits = 2 if args.workload == "admix_n2000_l100k_k5" else 2
    b_alg, f_alg = zpass_algorithmic(ds.N, ds.L, K, bits, ds.maxploidy, ds.KT)
    z_avg_s = (z_ms / 1e3) / max(z_n, 1)
    gbs = b_alg / z_avg_s / 1e9
    tops = f_alg / z_avg_s / 1e12
    traffic = None
    tp = os.path.join(ROOT, "profiles", "zpass_admix_n2000_l100k_k5_traffic.json")
    if os.path.exists(tp) and args.workload == "admix_n2000_l100k_k5":
        traffic = json.load(open(tp)).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(s, p, K, args.cpu_seconds)

    out = {
        "metric": "MCMC sweeps/s at N=2000,L=100k,K=5",
        "value": round(sweeps_s, 3),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == V.PREC_FP64 else "f32-weights/f64",
        "data": "synthetic (seeded generator: Fst 0.05 drift, 2% missing)",
        "config": {"workload": args.workload, "N": ds.N, "L": ds.L, "K": K, "model": "admixture+alpha Metropolis",
                   "rng_topology": args.topology, "precision": args.precision,
                   "record_every": 10, "metro_every": 10},
        "roofline": {
            "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": traffic,
            "kernel": "z_admix", "kernel_avg_ms": round(z_avg_s * 1e3, 4), "launches": z_n,
            "alg_bytes_per_launch": int(b_alg), "alg_fp64_ops_per_launch": int(f_alg),
            "valu": {"bound": "fp64-valu (binding)", "achieved": round(tops, 3), "peak": FP64_VALU_PEAK_TOPS,
                     "unit": "TFLOP/s", "frac": round(tops / FP64_VALU_PEAK_TOPS, 4)},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(s, p, K, budget_s):
    """CPU oracle sweeps of the same workload on the host cores (bounded sample)"""
    import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    O.set_threads(cores)
    od = O.Data(s.desc)
    oc = O.Chain(od, p, K, 1)
    t0 = time.perf_counter()
    n = 0
    while True:
        oc.sweep(1)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 200:
            break
    return {"value": round(n / el, 4), "unit": "sweeps/s", "cores": cores, "kind": "port",
            "sample": f"{n} sweeps from m=0 of the same workload, same RNG topology"}


if __name__ == "__main__":
    main()
