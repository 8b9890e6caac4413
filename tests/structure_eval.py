#!/usr/bin/env python3
"""Per-pass times and state digests of one clustering sweep for the oracle,
the FP64 device path and the mixed-precision device path: the analogue of
vcfpop's -structure_eval mode (src/function.cpp:4743-4766), which runs K = 5,
one sweep, no burn-in, for the admixture and no-admixture models at each
precision.

Integer state (Z, Ni, Mi) is digested exactly; floating state (P, Q, alpha)
is reported as the largest relative difference from the oracle's.  Both
chains start from the same seed, so InitAdmix agrees bit for bit and every
pass then runs on each side's own state.

    python tests/structure_eval.py                         # a small mixed-ploidy dataset
    python tests/structure_eval.py --workload admix_n2000_l100k_k5
    python tests/structure_eval.py --n-ind 500 --n-loc 20000 --ploidy 2:0.6,4:0.4 --json

Test infrastructure (it runs the oracle); needs a GPU for the device columns.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import oracle as O          # noqa: E402
import vcfpop_amd as V      # noqa: E402

MODELS = {
    # the passes of a first sweep (m = 0), in MCMC order (src/structure.cpp:2112-2166)
    "admix": (dict(admix_val=1), ["update_p", "update_q_admix", "update_z_admix", "update_alpha", "record"]),
    "noadmix": (dict(admix_val=2), ["update_p", "update_q_noadmix", "update_z_noadmix", "record"]),
}
INT_STATE = ("Z", "Ni", "Mi")
FP_STATE = ("Freq", "Q", "Alpha")


def digest(a):
    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()[:12]


def max_rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def run_eval(synth, K=5, seed=1, models=("admix", "noadmix"), backends=("oracle", "fp64", "mixed"), threads=8):
    """One sweep, pass by pass, on every backend; returns a list of rows
    {model, pass, backend, ms, Z, Ni, Mi, d_Freq, d_Q, d_Alpha}."""
    O.set_threads(threads)
    od = O.Data(synth.desc)
    ds = V.Dataset(synth, 0) if any(b != "oracle" for b in backends) else None
    rows = []
    for model in models:
        kw, passes = MODELS[model]
        params = {b: V.params_default(nburnin=0, nreps=1, nthinning=1, nadmburnin=0, g_seed=seed, z_chunk=-1,
                                      precision=V.PREC_MIXED if b == "mixed" else V.PREC_FP64, **kw)
                  for b in backends}
        for b in backends:      # a throwaway sweep, so that module loading is not timed
            if b != "oracle":
                w = V.Chain(ds, params[b], K, 2)
                for name in passes:
                    getattr(w, name)()
                w.sync()
                w.close()
        chains = {b: O.Chain(od, params[b], K, 1) if b == "oracle" else V.Chain(ds, params[b], K, 1)
                  for b in backends}
        for b, c in chains.items():
            if b != "oracle":
                c.sync()
        for name in ["init"] + passes:
            ref = None
            for b in backends:
                c = chains[b]
                t0 = time.perf_counter()
                if name != "init":
                    if b == "oracle":
                        c.call(name)
                    else:
                        getattr(c, name)()
                        c.sync()
                ms = (time.perf_counter() - t0) * 1e3
                row = {"model": model, "pass": name, "backend": b, "ms": round(ms, 3)}
                state = {n: c.get(n) for n in INT_STATE + FP_STATE}
                for n in INT_STATE:
                    row[n] = digest(state[n].astype(np.int64))
                if b == "oracle":
                    ref = state
                for n in FP_STATE:
                    row["d_" + n] = max_rel(state[n], ref[n]) if ref is not None else None
                rows.append(row)
        for c in chains.values():
            c.close() if hasattr(c, "close") else None
    if ds is not None:
        ds.close()
    return rows


def print_table(rows):
    print(f"{'model':8s} {'pass':18s} {'backend':7s} {'ms':>9s}  {'Z':12s} {'Ni':12s} {'Mi':12s} "
          f"{'dP':>9s} {'dQ':>9s} {'dAlpha':>9s}")
    for r in rows:
        f = lambda x: "-" if x is None else f"{x:9.2e}"
        print(f"{r['model']:8s} {r['pass']:18s} {r['backend']:7s} {r['ms']:9.3f}  {r['Z']} {r['Ni']} {r['Mi']} "
              f"{f(r['d_Freq'])} {f(r['d_Q'])} {f(r['d_Alpha'])}")


def parse_ploidy(text):
    out = []
    for part in text.split(","):
        v, w = part.split(":")
        out.append((int(v), float(w)))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", help="a named synthetic workload (vcfpop_amd.WORKLOADS)")
    ap.add_argument("--n-ind", type=int, default=300)
    ap.add_argument("--n-loc", type=int, default=3000)
    ap.add_argument("--ploidy", default="2:0.6,4:0.3,6:0.1", help="ploidy:weight,...")
    ap.add_argument("--frac-multi", type=float, default=0.2)
    ap.add_argument("--missing", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--K", type=int, default=5)
    ap.add_argument("--backends", default="oracle,fp64,mixed")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--json", action="store_true", help="one JSON line instead of the table")
    args = ap.parse_args()
    if args.workload:
        s = V.make_synth(args.workload)
    else:
        s = V.Synth(n_ind=args.n_ind, n_loc=args.n_loc, k_true=3, alpha_q=0.5, missing=args.missing,
                    frac_multi=args.frac_multi, min_alleles=3, max_alleles=8, multi_dir=0.3,
                    ploidy=parse_ploidy(args.ploidy), seed=args.seed)
    rows = run_eval(s, K=args.K, seed=args.seed, backends=tuple(args.backends.split(",")), threads=args.threads)
    if args.json:
        print(json.dumps(rows))
    else:
        print_table(rows)


if __name__ == "__main__":
    main()
