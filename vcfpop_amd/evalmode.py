"""Pass-by-pass evaluation of one clustering sweep: vcfpop's -structure_eval
mode.  The reference's mode (src/function.cpp:4743-4766) fixes K = 5, one
run, one sweep with no burn-in, and runs the admixture and no-admixture
models at each precision; each pass then times every implementation and
prints its time with hashes of the integer state it wrote (for example
src/structure.cpp:1484 for UpdateZAdmix).

Here the implementations are the device's FP64 and mixed-precision paths,
plus any caller-supplied backends placed first (a test harness adds a CPU
restatement that way).  Every backend runs its own chain from the same seed,
so initialisation agrees bit for bit and each pass then runs on that
backend's own state.  Integer state (Z, Ni, Mi) is reported as a digest;
floating state (P, Q, alpha) as the largest relative difference from the
first backend's.

    python -m vcfpop_amd.evalmode --workload admix_n2000_l100k_k5
    python -m vcfpop_amd.evalmode --n-ind 500 --n-loc 20000 --ploidy 2:0.6,4:0.4 --json
"""
import argparse
import hashlib
import json
import sys
import time

import numpy as np

from ._lib import PREC_FP64, PREC_MIXED, Chain, Dataset, Synth, params_default

# the passes of a first sweep (m = 0), in MCMC order (src/structure.cpp:2112-2166)
MODELS = {
    "admix": (dict(admix_val=1), ["update_p", "update_q_admix", "update_z_admix", "update_alpha", "record"]),
    "noadmix": (dict(admix_val=2), ["update_p", "update_q_noadmix", "update_z_noadmix", "record"]),
}
INT_STATE = ("Z", "Ni", "Mi")
FP_STATE = ("Freq", "Q", "Alpha")
DEVICE_BACKENDS = {"fp64": PREC_FP64, "mixed": PREC_MIXED}


def digest(a):
    """12 hex digits of the SHA-1 of an array's int64 values"""
    return hashlib.sha1(np.ascontiguousarray(np.asarray(a).astype(np.int64)).tobytes()).hexdigest()[:12]


def max_rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


class DeviceBackend:
    """a device chain driven pass by pass"""

    def __init__(self, ds, params, K, run_id=1):
        self.chain = Chain(ds, params, K, run_id)
        self.chain.sync()

    def run(self, name):
        getattr(self.chain, name)()
        self.chain.sync()

    def get(self, name):
        return self.chain.get(name)

    def close(self):
        self.chain.close()


def eval_params(model, precision, seed, base=None):
    """the evaluation's options: `base` (other -structure_* options) with the
    run length, model and precision of the evaluation"""
    kw = dict(base or {})
    kw.update(MODELS[model][0])
    kw.update(nburnin=0, nreps=1, nthinning=1, nadmburnin=0, g_seed=seed, z_chunk=-1, precision=precision)
    return params_default(**kw)


def run_eval(ds, K=5, seed=1, models=("admix", "noadmix"), backends=("fp64", "mixed"), extra=None, base=None):
    """One sweep, pass by pass, on every backend.

    ds: a Dataset; backends: names in DEVICE_BACKENDS; extra: optional
    {name: factory(params, K) -> object with run(pass), get(field), close()},
    run before the device backends (the first backend is the reference of the
    d_* columns); base: other options (params_default keywords).  Returns rows
    {model, pass, backend, ms, Z, Ni, Mi, d_Freq, d_Q, d_Alpha}."""
    extra = dict(extra or {})
    order = list(extra) + list(backends)
    rows = []
    for model in models:
        _, passes = MODELS[model]
        params = {b: eval_params(model, DEVICE_BACKENDS.get(b, PREC_FP64), seed, base) for b in order}
        for b in backends:      # a throwaway sweep, so that module loading is not timed
            w = DeviceBackend(ds, params[b], K, 2)
            for name in passes:
                w.run(name)
            w.close()
        chains = {b: extra[b](params[b], K) if b in extra else DeviceBackend(ds, params[b], K) for b in order}
        for name in ["init"] + passes:
            ref = None
            for b in order:
                c = chains[b]
                t0 = time.perf_counter()
                if name != "init":
                    c.run(name)
                ms = (time.perf_counter() - t0) * 1e3
                row = {"model": model, "pass": name, "backend": b, "ms": round(ms, 3)}
                state = {n: c.get(n) for n in INT_STATE + FP_STATE}
                for n in INT_STATE:
                    row[n] = digest(state[n])
                if ref is None:
                    ref = state
                for n in FP_STATE:
                    row["d_" + n] = max_rel(state[n], ref[n])
                rows.append(row)
        for c in chains.values():
            c.close()
    return rows


def print_table(rows, file=None):
    file = file or sys.stdout
    print(f"{'model':8s} {'pass':18s} {'backend':7s} {'ms':>9s}  {'Z':12s} {'Ni':12s} {'Mi':12s} "
          f"{'dP':>9s} {'dQ':>9s} {'dAlpha':>9s}", file=file)
    for r in rows:
        print(f"{r['model']:8s} {r['pass']:18s} {r['backend']:7s} {r['ms']:9.3f}  {r['Z']} {r['Ni']} {r['Mi']} "
              f"{r['d_Freq']:9.2e} {r['d_Q']:9.2e} {r['d_Alpha']:9.2e}", file=file)


def parse_ploidy(text):
    out = []
    for part in text.split(","):
        v, w = part.split(":")
        out.append((int(v), float(w)))
    return out


def synth_from_args(args):
    from .workloads import make_synth
    if args.workload:
        return make_synth(args.workload)
    return Synth(n_ind=args.n_ind, n_loc=args.n_loc, k_true=3, alpha_q=0.5, missing=args.missing,
                 frac_multi=args.frac_multi, min_alleles=3, max_alleles=8, multi_dir=0.3,
                 ploidy=parse_ploidy(args.ploidy), seed=args.seed)


def add_args(ap):
    ap.add_argument("--workload", help="a named synthetic workload (vcfpop_amd.WORKLOADS)")
    ap.add_argument("--n-ind", type=int, default=300)
    ap.add_argument("--n-loc", type=int, default=3000)
    ap.add_argument("--ploidy", default="2:0.6,4:0.3,6:0.1", help="ploidy:weight,...")
    ap.add_argument("--frac-multi", type=float, default=0.2)
    ap.add_argument("--missing", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--K", type=int, default=5)
    ap.add_argument("--json", action="store_true", help="one JSON line instead of the table")


def main(argv=None):
    ap = argparse.ArgumentParser(description="pass-by-pass times and state digests of one clustering sweep")
    add_args(ap)
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    s = synth_from_args(args)
    ds = Dataset(s, args.device)
    rows = run_eval(ds, K=args.K, seed=args.seed)
    ds.close()
    if args.json:
        print(json.dumps(rows))
    else:
        print_table(rows)


if __name__ == "__main__":
    sys.exit(main())
