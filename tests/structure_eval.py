#!/usr/bin/env python3
"""vcfpop's -structure_eval mode with the oracle as the first column: one
clustering sweep, pass by pass, on the oracle (the CPU restatement), the FP64
device path and the mixed-precision device path (vcfpop_amd.evalmode runs
the device columns; this harness adds the oracle in front, so the d_*
columns are differences from the oracle).

    python tests/structure_eval.py                         # a small mixed-ploidy dataset
    python tests/structure_eval.py --workload admix_n2000_l100k_k5
    python tests/structure_eval.py --n-ind 500 --n-loc 20000 --ploidy 2:0.6,4:0.4 --json

Test infrastructure (it runs the oracle); needs a GPU for the device columns.
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import oracle as O                  # noqa: E402
import vcfpop_amd as V              # noqa: E402
from vcfpop_amd import evalmode     # noqa: E402


class OracleBackend:
    """the oracle's chain driven pass by pass"""

    def __init__(self, od, params, K):
        self.chain = O.Chain(od, params, K, 1)

    def run(self, name):
        self.chain.call(name)

    def get(self, name):
        return self.chain.get(name)

    def close(self):
        if hasattr(self.chain, "close"):
            self.chain.close()


def run_eval(synth, K=5, seed=1, models=("admix", "noadmix"), backends=("oracle", "fp64", "mixed"), threads=8):
    """rows of evalmode.run_eval with the oracle first (when named)"""
    extra = {}
    if "oracle" in backends:
        O.set_threads(threads)
        od = O.Data(synth.desc)
        extra["oracle"] = lambda params, k: OracleBackend(od, params, k)
    ds = V.Dataset(synth, 0)
    try:
        return evalmode.run_eval(ds, K=K, seed=seed, models=models,
                                 backends=tuple(b for b in backends if b != "oracle"), extra=extra)
    finally:
        ds.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    evalmode.add_args(ap)
    ap.add_argument("--backends", default="oracle,fp64,mixed")
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    s = evalmode.synth_from_args(args)
    rows = run_eval(s, K=args.K, seed=args.seed, backends=tuple(args.backends.split(",")), threads=args.threads)
    if args.json:
        print(json.dumps(rows))
    else:
        evalmode.print_table(rows)


if __name__ == "__main__":
    main()
