#!/usr/bin/env python3
"""MCMC sweeps/s of vcfpop's Bayesian clustering on one MI355X, plus the
batched, served and mixed-precision forms.

Headline (`value`): the drop-in at N=2,000 individuals, L=100,000 biallelic
diploid SNPs packed at 2 bits, K=5, admixture model with alpha Metropolis,
FP64 (synthetic data: Fst 0.05 drift, 2% missing), drawing the reference's
own random streams (rng_topology "reference": every draw, and so every
result, equals vcfpop's).  A step is one full sweep (UpdateP, UpdateQ
[Metropolis every 10th], UpdateZ, UpdateAlpha, Record every 10th) with
inputs resident in HBM.

Extra objects on the same JSON line:
  roofline      the Z pass, timed with HIP events on the chain's stream
                (every 10th pass of the timed region)
  cpu_baseline  the CPU oracle (C, OpenMP) on a bounded number of sweeps of
                the same workload on the host cores
  parallel_topology  the same workload with one stream per (k, l) and per
                individual for P and Q (same distributions, other draws)
  real_float    the same workload as BAYESIAN<float> (-g_float=single):
                REAL32 (-g_fastsingle=no) and FAST32 (vcfpop's default)
  batched       64 chains (K=2..9 x 8 seeds) on N=1,000 x L=50,000: chains/hr
                one chain at a time vs one graph launch per sweep of all 64
  served        a Poisson job stream (N ~ U[100,5000], L ~ logU[1k,200k],
                ploidy 2 or 4, K ~ U[1,10], admix / no-admix / F model) through
                the job server vs the same jobs one after another; p50/p99
                latency and chains/hr
  precision     N=3,000, L=200,000 mixed ploidy 2/4/6, 20% multiallelic, K=6,
                locprior: sweeps/s FP64 vs FP32 weights, and the FP32 run's
                deviation from the FP64 run
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector, published (an FMA counts as 2 flops)


def zpass_algorithmic(N, L, K, bits_per_geno, maxploidy, KT):
    """algorithmic bytes and FP64 ops of one Z pass (per-slot counts: K mul,
    K +eps, K-1 sum, (U-1)*s, <= K-1 subtract and compare)"""
    slots = N * L * maxploidy
    bytes_ = N * L * bits_per_geno / 8 + K * KT * 8 + N * K * 8 + K * KT * 4 + N * K * 4
    fp64 = slots * (5 * K + 2)
    return bytes_, fp64


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="admix_n2000_l100k_k5")
    ap.add_argument("--topology", default="reference", choices=["parallel", "reference"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "real32", "fast32"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline workload (no other-topology, float, batched, served or precision legs)")
    ap.add_argument("--batch-sweeps", type=int, default=40)
    ap.add_argument("--batch-form", default="all",
                    choices=["all", "plain", "batched", "batched_mixed", "plain_drop_in", "batched_drop_in",
                             "batched_mixed_drop_in"],
                    help="with --leg batched: time only this form of the 64 chains (for a profiler run)")
    ap.add_argument("--serve-jobs", type=int, default=24)
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-sweep-scale", type=int, default=20,
                    help="job sweep counts U[2000,10000] are divided by this")
    ap.add_argument("--leg", choices=["small", "batched", "served", "precision"], default=None,
                    help="only this extra leg (its JSON object on one line), for profiling")
    ap.add_argument("--served-window", type=float, default=0.0, metavar="SECONDS",
                    help="only the served job stream at full size (sweeps U[2000,10000]), arrivals over "
                         "the first SECONDS of the Poisson stream, against the same jobs one at a time; "
                         "prints one JSON line and exits")
    ap.add_argument("--serve-max-active", type=int, nargs="+", default=[32],
                    help="jobs the server keeps on the device at once (--served-window: one served run each)")
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
    if args.leg:
        import vcfpop_amd as V
        fn = {"small": lambda: bench_small(V, V.RNG_REFERENCE if args.topology == "reference" else V.RNG_PARALLEL),
              "batched": lambda: bench_batched(V, args.batch_sweeps, args.batch_form),
              "served": lambda: bench_served(V, args.serve_jobs, args.serve_rate, args.serve_sweep_scale),
              "precision": lambda: bench_precision(V)}[args.leg]
        print(json.dumps({args.leg: fn()}), flush=True)
        return
    if args.served_window > 0:
        import vcfpop_amd as V
        print(json.dumps(bench_served_window(V, args.served_window, args.serve_rate, args.serve_max_active)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (chains shard across GPUs "
              "with no communication, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import vcfpop_amd as V
    spec = V.WORKLOADS[args.workload]
    K = spec["K"] if isinstance(spec["K"], int) else spec["K"][0]
    log("generating", args.workload)
    s = V.make_synth(args.workload)
    ds = V.Dataset(s, 0)
    topo = V.RNG_PARALLEL if args.topology == "parallel" else V.RNG_REFERENCE
    prec = {"fp64": V.PREC_FP64, "mixed": V.PREC_MIXED, "real32": V.PREC_REAL32, "fast32": V.PREC_FAST32}[args.precision]
    p = V.params_default(admix_val=1, locpriori_val=1 if spec.get("locpriori") else 2,
                         nburnin=spec["nburnin"], nreps=spec["nreps"], g_seed=1,
                         rng_topology=topo, precision=prec)
    ch = V.Chain(ds, p, K, 1)
    log("warmup", args.warmup)
    ch.step(args.warmup)
    ch.sync()
    ch.enable_timing(True, every=10)
    t0 = time.perf_counter()
    ch.step(args.steps)
    ch.sync()
    dt = time.perf_counter() - t0
    z_ms, z_n = ch.kernel_time(0)
    sweeps_s = args.steps / dt
    G, Nc = ch.chunking()

    b_alg, f_alg = zpass_algorithmic(ds.N, ds.L, K, 2, ds.maxploidy, ds.KT)
    z_avg_s = (z_ms / 1e3) / max(z_n, 1)
    gbs = b_alg / z_avg_s / 1e9
    tflops = f_alg / z_avg_s / 1e12
    traffic = None
    tp = os.path.join(ROOT, "profiles", "zpass_admix_n2000_l100k_k5_traffic.json")
    if os.path.exists(tp) and args.workload == "admix_n2000_l100k_k5":
        tj = json.load(open(tp))   # rocprof FETCH_SIZE/WRITE_SIZE per launch, by chunk count
        traffic = tj.get("bytes_per_launch_by_G", {}).get(str(G), tj.get("bytes_per_launch") if G == tj.get("chunks_G") else None)
    ch.close()

    # the same workload with the other RNG topology, same steps
    other = "reference" if args.topology == "parallel" else "parallel"
    other_leg = None
    if not args.no_extra:
        log("topology", other)
        po = V.params_default(**{f: getattr(p, f) for f, _ in p._fields_})
        po.rng_topology = V.RNG_REFERENCE if other == "reference" else V.RNG_PARALLEL
        co = V.Chain(ds, po, K, 1)
        co.step(args.warmup)
        co.sync()
        t0 = time.perf_counter()
        co.step(args.steps)
        co.sync()
        dto = time.perf_counter() - t0
        co.close()
        other_leg = {"rng_topology": other, "sweeps_per_s": round(args.steps / dto, 3),
                     "ms_per_step": round(dto * 1e3 / args.steps, 4), "steps": args.steps, "warmup": args.warmup}

    # the same workload as BAYESIAN<float>, same topology and steps
    real_float = {}
    for name, pr in (("real32", V.PREC_REAL32), ("fast32", V.PREC_FAST32)):
        if pr == prec or args.no_extra:
            continue
        log("precision", name)
        pf = V.params_default(**{f: getattr(p, f) for f, _ in p._fields_})
        pf.precision = pr
        cf = V.Chain(ds, pf, K, 1)
        cf.step(args.warmup)
        cf.sync()
        t0 = time.perf_counter()
        cf.step(args.steps)
        cf.sync()
        dtf = time.perf_counter() - t0
        cf.close()
        real_float[name] = {"sweeps_per_s": round(args.steps / dtf, 3), "ms_per_step": round(dtf * 1e3 / args.steps, 4),
                            "rng_topology": args.topology}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(V, s, p, K, args.cpu_seconds)

    out = {
        "metric": ("MCMC sweeps/s at N=2000,L=100k,K=5" if args.workload == "admix_n2000_l100k_k5"
                   else f"MCMC sweeps/s ({args.workload})"),
        "value": round(sweeps_s, 3),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {V.PREC_FP64: "f64", V.PREC_MIXED: "f32-weights/f64", V.PREC_REAL32: "f32-state/f64",
                  V.PREC_FAST32: "f32"}[prec],
        "data": "synthetic (seeded generator: Fst 0.05 drift, 2% missing)",
        "config": {"workload": args.workload, "N": ds.N, "L": ds.L, "K": K, "model": "admixture+alpha Metropolis" + ("+locprior" if spec.get("locpriori") else ""),
                   "rng_topology": args.topology, "precision": args.precision,
                   "record_every": 10, "metro_every": 10,
                   "z_chunks": {"G": G, "individuals_per_chunk": Nc, "chosen_by": "timed candidates"}},
        # the Z pass is bound by VALU issue (FP64 arithmetic plus the integer
        # RNG, count and decode instructions), not by HBM: the binding roof is
        # the published FP64 peak; the HBM figures are secondary
        "roofline": {
            "bound": "fp64", "achieved": round(tflops, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
            "kernel": "z_admix (sweeps without Record)", "kernel_avg_ms": round(z_avg_s * 1e3, 4), "launches_timed": z_n,
            "alg_bytes_per_launch": int(b_alg), "alg_fp64_flops_per_launch": int(f_alg),
            "hbm": {"achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
                    "measured_bytes_per_launch": traffic},
        },
        "cpu_baseline": cpu,
        f"{other}_topology": other_leg,
        "real_float": real_float,
    }
    ds.close()
    if not args.no_extra:
        for name, fn in (("admix_n200_l2k_k3", lambda: bench_small(V, p.rng_topology)),
                         ("batched", lambda: bench_batched(V, args.batch_sweeps)),
                         ("served", lambda: bench_served(V, args.serve_jobs, args.serve_rate, args.serve_sweep_scale)),
                         ("precision", lambda: bench_precision(V))):
            log(name)
            try:
                out[name] = fn()
            except Exception as e:   # an extra leg must not hide the headline line
                out[name] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out), flush=True)


def host_cpu():
    """host threads the CPU baseline may use, and what the host reports.  The
    affinity mask bounds the threads; OMP_NUM_THREADS, when set, is the CPU
    share the host grants this process and bounds them too."""
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = os.cpu_count() or 1
    threads = aff
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": max(1, threads), "affinity_cpus": aff, "nproc": os.cpu_count(),
            "omp_num_threads": omp or None, "model": model}


def cpu_baseline(V, s, p, K, budget_s):
    """The CPU oracle (C, OpenMP over individuals and 64-locus tiles) on the
    host cores: a bounded sample of the headline workload, plus the plain
    CPU-runnable workload (N=200, L=2,000, K=3) and the 64-chain workload
    (N=1,000, L=50,000, K=2..9) as chains/hr of one chain at a time."""
    import oracle as O
    hc = host_cpu()
    O.set_threads(hc["threads"])
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
    out = {"value": round(n / el, 4), "unit": "sweeps/s", "cores": hc["threads"], "kind": "port",
           "sample": f"{n} sweeps from m=0 of the same workload, same RNG topology",
           "host": hc}
    # the CPU-runnable workload: 300 sweeps of a full chain (1,000 burn-in + 2,000)
    s0 = V.make_synth("admix_n200_l2k_k3")
    p0 = V.params_default(admix_val=1, nburnin=1000, nreps=2000, g_seed=1, rng_topology=p.rng_topology)
    c0 = O.Chain(O.Data(s0.desc), p0, 3, 1)
    t0 = time.perf_counter()
    c0.sweep(300)
    out["admix_n200_l2k_k3_sweeps_per_s"] = round(300 / (time.perf_counter() - t0), 2)
    # the 64-chain workload: 3 sweeps per K after one untimed sweep
    s2 = V.make_synth("batch_n1000_l50k_k2to9x8")
    d2 = O.Data(s2.desc)
    per_k = {}
    for k in range(2, 10):
        c2 = O.Chain(d2, V.params_default(admix_val=1, nburnin=0, nreps=5000, g_seed=1,
                                          rng_topology=p.rng_topology), k, 1)
        c2.sweep(1)
        t0 = time.perf_counter()
        c2.sweep(3)
        per_k[k] = (time.perf_counter() - t0) / 3
    mean_chain_s = 5000 * sum(per_k.values()) / len(per_k)
    out["batch_n1000_l50k_k2to9x8_chains_per_hr"] = round(3600.0 / mean_chain_s, 3)
    out["batch_n1000_l50k_k2to9x8_ms_per_sweep_by_K"] = {k: round(v * 1e3, 2) for k, v in per_k.items()}
    return out


def bench_small(V, topology):
    """the plain CPU-runnable workload on the device: one full chain of
    N=200, L=2,000 biallelic diploid, K=3, 1,000 burn-in + 2,000 sampling
    sweeps, FP64, from chain creation to the arranged result"""
    s = V.make_synth("admix_n200_l2k_k3")
    ds = V.Dataset(s, 0)
    p = V.params_default(admix_val=1, nburnin=1000, nreps=2000, g_seed=1, rng_topology=topology)
    t0 = time.perf_counter()
    c = V.Chain(ds, p, 3, 1)
    t1 = time.perf_counter()
    c.step(3000)
    c.finish()
    t2 = time.perf_counter()
    c.close()
    ds.close()
    return {"sweeps": 3000, "sweeps_per_s": round(3000 / (t2 - t1), 2), "chain_wall_s": round(t2 - t0, 4),
            "chain_create_s": round(t1 - t0, 4)}


def bench_batched(V, S, form="all"):
    """64 chains (K=2..9 x 8 replicate seeds) over one N=1,000 x L=50,000
    dataset, 5,000 sweeps each: chains/hr when chains run one at a time
    (the reference's one-chain-per-device leasing) and as one batch, in the
    parallel draw topology and in the reference's (the drop-in as wired into
    vcfpop, whose batched FP64 chains equal the plain ones bit for bit)"""
    s = V.make_synth("batch_n1000_l50k_k2to9x8")
    ds = V.Dataset(s, 0)

    def chains(precision=V.PREC_FP64, topology=V.RNG_PARALLEL):
        return [V.Chain(ds, V.params_default(admix_val=1, nburnin=0, nreps=5000, g_seed=1,
                                             rng_topology=topology, precision=precision), K, r)
                for K in range(2, 10) for r in range(1, 9)]

    def run_batch(cs):
        b = V.Batch(cs)
        b.step(2)
        b.sync()
        t0 = time.perf_counter()
        b.step(S)
        b.sync()
        t = (time.perf_counter() - t0) / S
        st = b.stats()
        b.close()
        for c in cs:
            c.close()
        return t, st

    def run_plain(cs):
        for c in cs:
            c.step(2)
            c.sync()
        t0 = time.perf_counter()
        for c in cs:
            c.step(S)
            c.sync()
        t = (time.perf_counter() - t0) / S
        for c in cs:
            c.close()
        return t

    per_hr = lambda t: 64 * 3600.0 / (t * 5000)
    if form != "all":
        # one form alone (for a profiler run)
        prec = V.PREC_MIXED if "mixed" in form else V.PREC_FP64
        topo = V.RNG_REFERENCE if form.endswith("drop_in") else V.RNG_PARALLEL
        cs = chains(prec, topo)
        t = run_plain(cs) if form.startswith("plain") else run_batch(cs)[0]
        ds.close()
        return {"workload": "batch_n1000_l50k_k2to9x8", "form": form, "measured_sweeps": S,
                "ms_per_sweep_of_64": round(t * 1e3, 3), "chains_per_hr": round(per_hr(t), 1)}

    plain = run_plain(chains())
    batched, st = run_batch(chains())
    # the same batch with FP32 weights (FP64 lnL, draws and Metropolis steps)
    mixed, _ = run_batch(chains(V.PREC_MIXED))
    # the drop-in's own draws (the reference's serial P and Q streams)
    log("batched: reference topology")
    plain_ref = run_plain(chains(topology=V.RNG_REFERENCE))
    batched_ref, _ = run_batch(chains(topology=V.RNG_REFERENCE))
    mixed_ref, _ = run_batch(chains(V.PREC_MIXED, V.RNG_REFERENCE))
    ds.close()
    return {"workload": "batch_n1000_l50k_k2to9x8", "chains": 64, "sweeps_per_chain": 5000,
            "measured_sweeps": S, "rng_topology": "parallel", "plain_chains_per_hr": round(per_hr(plain), 1),
            "batched_chains_per_hr": round(per_hr(batched), 1), "speedup": round(plain / batched, 3),
            "batched_mixed_chains_per_hr": round(per_hr(mixed), 1),
            "speedup_mixed_vs_plain_fp64": round(plain / mixed, 3),
            "ms_per_sweep_of_64": {"plain": round(plain * 1e3, 3), "batched": round(batched * 1e3, 3),
                                   "batched_mixed": round(mixed * 1e3, 3)},
            "graphs": st,
            "speedup_mixed_vs_plain_drop_in": round(plain_ref / mixed, 3),
            "reference_topology": {
                "plain_drop_in_chains_per_hr": round(per_hr(plain_ref), 1),
                "batched_chains_per_hr": round(per_hr(batched_ref), 1),
                "speedup": round(plain_ref / batched_ref, 3),
                "batched_mixed_chains_per_hr": round(per_hr(mixed_ref), 1),
                "speedup_mixed_vs_plain_drop_in": round(plain_ref / mixed_ref, 3),
                "ms_per_sweep_of_64": {"plain": round(plain_ref * 1e3, 3), "batched": round(batched_ref * 1e3, 3),
                                       "batched_mixed": round(mixed_ref * 1e3, 3)}}}


def make_jobs(V, n, rate, scale, seed=7, window=None):
    """n jobs of the served stream (or every job arriving before `window`
    seconds): Poisson arrivals at `rate`, N ~ U[100,5000], L ~ logU[1k,200k],
    ploidy 2 or 4, K ~ U[1,10], admixture / no-admixture / F model,
    U[2000,10000] sweeps divided by `scale`"""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs, t = [], 0.0
    for j in range(n if window is None else 1 << 30):
        t += rng.exponential(1.0 / rate)
        if window is not None and t > window:
            break
        N = int(rng.integers(100, 5001))
        L = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
        pl = [(2, 1.0)] if rng.random() < 0.5 else [(4, 1.0)]
        K = int(rng.integers(1, 11))
        model = ("admix", "noadmix", "fmodel")[int(rng.integers(0, 3))]
        sweeps = int(rng.integers(2000, 10001)) // scale
        s = V.Synth(n_ind=N, n_loc=L, k_true=3, alpha_q=0.5, missing=0.01, ploidy=pl, seed=1000 + j)
        kw = (dict(admix_val=1) if model == "admix" else
              dict(admix_val=2, nadmburnin=sweeps // 10) if model == "noadmix" else dict(admix_val=1, f_val=1))
        burn = sweeps // 4
        # z_chunk=-1: the chunk model, as the server uses it, for the plain runs too (no timing trials)
        p = V.params_default(nburnin=burn, nreps=sweeps - burn - kw.get("nadmburnin", 0),
                             rng_topology=V.RNG_PARALLEL, g_seed=j, z_chunk=-1, **kw)
        jobs.append(dict(arrival=t, synth=s, params=p, K=K, N=N, L=L, sweeps=sweeps, model=model))
    return jobs


def bench_served_window(V, window, rate, max_active=(32,)):
    """The served stream as specified (full job sizes and sweep counts) for
    the jobs arriving in the first `window` seconds: open-loop latency
    through the server (queueing included, the stream drained) against the
    same jobs run one at a time in arrival order on plain chains"""
    import numpy as np
    jobs = make_jobs(V, 0, rate, 1, window=window)
    log(f"served window {window} s: {len(jobs)} jobs")
    total = lambda j: j["params"].nburnin + j["params"].nreps + (j["params"].nadmburnin if j["params"].admix_val != 1 else 0)
    dur = []
    for x, j in enumerate(jobs):
        a = time.perf_counter()
        d = V.Dataset(j["synth"], 0)
        c = V.Chain(d, j["params"], j["K"], 1)
        c.step(total(j))
        c.finish()
        c.close()
        d.close()
        dur.append(time.perf_counter() - a)
        log(f"plain job {x + 1}/{len(jobs)}: N={j['N']} L={j['L']} K={j['K']} {j['model']} {j['sweeps']} sweeps {dur[-1]:.2f} s")
    end, lat_plain = 0.0, []
    for j, d_ in zip(jobs, dur):
        end = max(end, j["arrival"]) + d_
        lat_plain.append(end - j["arrival"])
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)
    out = {"arrival_window_s": window, "arrival_rate_per_s": rate, "jobs": len(jobs),
           "sweeps_per_job": "U[2000,10000]", "model_mix": "admix/no-admix/F",
           "plain": {"device_seconds": round(sum(dur), 2), "p50_s": pct(lat_plain, 50), "p99_s": pct(lat_plain, 99),
                     "makespan_s": round(end, 2), "chains_per_hr": round(len(jobs) * 3600.0 / sum(dur), 1)},
           "served": {},
           "policy": "FIFO admission under a 128 GB budget, at most max_active jobs on the device; "
                     "latency counts queueing; the stream is drained"}
    for ma in max_active:
        srv = V.Server(hbm_budget_bytes=128 << 30, max_active=ma, slice_sweeps=16, loaders=2)
        ids = []
        t0 = time.perf_counter()
        for j in jobs:
            now = time.perf_counter() - t0
            if now < j["arrival"]:
                time.sleep(j["arrival"] - now)
            ids.append(srv.submit(j["synth"], j["params"], j["K"], 1))
        for x, i in enumerate(ids):
            srv.wait(i, 3600.0)
            log(f"max_active {ma}: served job {x + 1}/{len(ids)} done at {time.perf_counter() - t0:.1f} s")
        span = time.perf_counter() - t0
        lat = [x["t_end"] - x["t_submit"] for x in (srv.info(i) for i in ids)]
        st = srv.stats()
        srv.close()
        out["served"][f"max_active_{ma}"] = {"p50_s": pct(lat, 50), "p99_s": pct(lat, 99), "makespan_s": round(span, 2),
                                             "chains_per_hr": round(len(jobs) * 3600.0 / span, 1),
                                             "peak_reserved_GB": round(st["peak_reserved_bytes"] / 1e9, 3)}
    return {"served_window": out}


def bench_served(V, n, rate, scale):
    """The job stream through the server against the same jobs run one after
    another on plain chains.

    open loop: jobs submitted at their Poisson arrival times; latency =
      completion - arrival (queueing included).  The plain latencies replay
      the measured per-job durations through a one-job-at-a-time queue.
    saturated: every job submitted at once; chains/hr = jobs / makespan."""
    import numpy as np
    log(f"served: generating {n} jobs")
    jobs = make_jobs(V, n, rate, scale)
    total = lambda j: j["params"].nburnin + j["params"].nreps + (j["params"].nadmburnin if j["params"].admix_val != 1 else 0)
    def with_(p, **kw):
        q = V.params_default(**{f: getattr(p, f) for f, _ in p._fields_})
        for k, v in kw.items():
            setattr(q, k, v)
        return q

    # plain, back to back: per-job durations and the saturated makespan
    def run_plain(topology):
        dur = []
        t0 = time.perf_counter()
        for j in jobs:
            a = time.perf_counter()
            d = V.Dataset(j["synth"], 0)
            c = V.Chain(d, with_(j["params"], rng_topology=topology), j["K"], 1)
            c.step(total(j))
            c.finish()
            c.close()
            d.close()
            dur.append(time.perf_counter() - a)
        return dur, time.perf_counter() - t0

    dur, plain_sat = run_plain(V.RNG_PARALLEL)
    # the drop-in as wired into vcfpop: the reference's own P and Q streams
    log("served: plain drop-in (reference topology)")
    _, plain_ref_sat = run_plain(V.RNG_REFERENCE)
    end, lat_plain = 0.0, []
    for j, d_ in zip(jobs, dur):
        end = max(end, j["arrival"]) + d_
        lat_plain.append(end - j["arrival"])

    def serve(at_arrival, mixed=False):
        srv = V.Server(hbm_budget_bytes=128 << 30, max_active=32, slice_sweeps=16, loaders=2)
        ids = []
        t0 = time.perf_counter()
        for j in jobs:
            now = time.perf_counter() - t0
            if at_arrival and now < j["arrival"]:
                time.sleep(j["arrival"] - now)
            p = with_(j["params"], precision=V.PREC_MIXED) if mixed else j["params"]
            ids.append(srv.submit(j["synth"], p, j["K"], 1))
        for i in ids:
            srv.wait(i, 900.0)
        span = time.perf_counter() - t0
        lat = [x["t_end"] - x["t_submit"] for x in (srv.info(i) for i in ids)]
        st = srv.stats()
        srv.close()
        return span, lat, st

    open_span, lat, st = serve(True)
    sat_span, _, st_sat = serve(False)
    mix_span, _, _ = serve(False, mixed=True)
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)
    return {"jobs": n, "arrival_rate_per_s": rate,
            "sweeps_per_job": f"U[2000,10000]/{scale}", "model_mix": "admix/no-admix/F",
            "open_loop": {"served_p50_s": pct(lat, 50), "served_p99_s": pct(lat, 99),
                          "plain_p50_s": pct(lat_plain, 50), "plain_p99_s": pct(lat_plain, 99),
                          "served_makespan_s": round(open_span, 3)},
            "saturated": {"served_makespan_s": round(sat_span, 3), "plain_makespan_s": round(plain_sat, 3),
                          "served_chains_per_hr": round(n * 3600.0 / sat_span, 1),
                          "plain_chains_per_hr": round(n * 3600.0 / plain_sat, 1),
                          "speedup": round(plain_sat / sat_span, 3),
                          "served_mixed_makespan_s": round(mix_span, 3),
                          "served_mixed_chains_per_hr": round(n * 3600.0 / mix_span, 1),
                          "speedup_mixed_vs_plain_fp64": round(plain_sat / mix_span, 3),
                          "plain_drop_in_makespan_s": round(plain_ref_sat, 3),
                          "plain_drop_in_chains_per_hr": round(n * 3600.0 / plain_ref_sat, 1),
                          "speedup_mixed_vs_plain_drop_in": round(plain_ref_sat / mix_span, 3),
                          "peak_reserved_GB": round(st_sat["peak_reserved_bytes"] / 1e9, 3)},
            "policy": "FIFO admission under a 128 GB budget, at most 32 jobs on the device; "
                      "latency counts queueing; the stream is drained"}


def bench_precision(V, S=20):
    """FP64 vs FP32 weights on the mixed-ploidy locprior workload; the FP32
    chain's deviation from the FP64 chain (same seed, same streams)"""
    import numpy as np
    spec = V.WORKLOADS["locprior_n3000_l200k_k6_mixed"]
    s = V.make_synth("locprior_n3000_l200k_k6_mixed")
    ds = V.Dataset(s, 0)
    res = {"workload": "locprior_n3000_l200k_k6_mixed", "N": ds.N, "L": ds.L, "K": spec["K"],
           "maxploidy": ds.maxploidy}
    states = {}
    for name, prec in (("fp64", V.PREC_FP64), ("mixed", V.PREC_MIXED)):
        p = V.params_default(admix_val=1, locpriori_val=1, nburnin=0, nreps=10000, g_seed=1,
                             rng_topology=V.RNG_PARALLEL, precision=prec)
        c = V.Chain(ds, p, spec["K"], 1)
        c.step(2)
        c.sync()
        t0 = time.perf_counter()
        c.step(S)
        c.sync()
        res[f"{name}_sweeps_per_s"] = round(S / (time.perf_counter() - t0), 3)
        c.record()
        states[name] = {n: c.get(n) for n in ("Q", "Freq", "lnL", "Mi")}
        c.close()
    # the per-sweep FP32 error: one sweep from identical FP64 state
    from vcfpop_amd.precision import one_sweep_error
    p1 = V.params_default(admix_val=1, locpriori_val=1, nburnin=0, nreps=10000, g_seed=1, rng_topology=V.RNG_REFERENCE)
    res["one_sweep"] = one_sweep_error(ds, p1, spec["K"])
    a, b = states["fp64"], states["mixed"]
    res["after_sweeps"] = S + 2
    res["max_abs_dQ"] = float(np.abs(a["Q"] - b["Q"]).max())
    res["max_abs_dP"] = float(np.abs(a["Freq"] - b["Freq"]).max())
    res["rel_dlnL"] = float(abs(a["lnL"][0] - b["lnL"][0]) / abs(a["lnL"][0]))
    res["frac_Mi_cells_differing"] = float((a["Mi"] != b["Mi"]).mean())
    ds.close()
    res["posterior"] = posterior_precision(V)
    return res


def posterior_ancestry(V, ds, K, seed, prec):
    """posterior-mean ancestry (normalised MiSum) and mean lnL of one full chain
    of the plain admixture workload (1,000 burn-in + 2,000 sampling sweeps)"""
    p = V.params_default(admix_val=1, nburnin=1000, nreps=2000, g_seed=seed, rng_topology=V.RNG_PARALLEL,
                         precision=prec)
    c = V.Chain(ds, p, K, 1)
    c.step(3000)
    c.finish()
    ms = c.get("MiSum")
    out = ms / ms.sum(axis=1, keepdims=True), float(c.get("rout")[0])
    c.close()
    return out


def posterior_precision(V, seeds=(1, 2, 3, 4, 5, 6, 7, 8, 9, 10)):
    """FP32 weights against FP64 over full chains of the plain admixture
    workload: posterior-mean ancestry and mean lnL, beside the seed-to-seed
    spread of the FP64 chains (the Monte Carlo error of the same estimate)"""
    import itertools

    import numpy as np
    s = V.make_synth("admix_n200_l2k_k3")
    ds = V.Dataset(s, 0)
    d_mix, d_seed, l_mix, l_seed = [], [], [], []
    for seed in seeds:
        q64, l64 = posterior_ancestry(V, ds, 3, seed, V.PREC_FP64)
        q32, l32 = posterior_ancestry(V, ds, 3, seed, V.PREC_MIXED)
        q64b, l64b = posterior_ancestry(V, ds, 3, seed + 1000, V.PREC_FP64)
        d_mix.append(np.abs(q32 - q64))
        # the other seed's chain may settle on other cluster labels
        perm = min(itertools.permutations(range(3)), key=lambda pm: np.abs(q64b[:, list(pm)] - q64).mean())
        d_seed.append(np.abs(q64b[:, list(perm)] - q64))
        l_mix.append(abs(l32 - l64) / abs(l64))
        l_seed.append(abs(l64b - l64) / abs(l64))
    ds.close()
    d_mix, d_seed = np.stack(d_mix), np.stack(d_seed)
    return {"workload": "admix_n200_l2k_k3", "chains": len(seeds), "sweeps": 3000,
            "mean_abs_dQ_mixed_vs_fp64": float(d_mix.mean()), "max_abs_dQ_mixed_vs_fp64": float(d_mix.max()),
            "mean_abs_dQ_fp64_other_seed": float(d_seed.mean()), "max_abs_dQ_fp64_other_seed": float(d_seed.max()),
            "max_rel_dmeanlnL_mixed_vs_fp64": float(max(l_mix)),
            "max_rel_dmeanlnL_fp64_other_seed": float(max(l_seed))}


if __name__ == "__main__":
    main()
