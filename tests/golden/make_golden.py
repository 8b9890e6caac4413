#!/usr/bin/env python3
"""Writes tests/golden/*.npz: fixed input/output vectors of the oracle.

The reference ships no golden vectors, known-answer tests or fixtures for
this path, so these are self-consistency fixtures produced by the oracle
(itself pinned by tests/test_oracle_*.py): they freeze its behaviour and are
the fixed-input vectors the device path is checked against.

  rng.npz     xorshift128+ outputs, Uniform, Normal, Gamma and Next draws for
              fixed (seed, salt) pairs; HashULong / MurmurHash64 answers
  chain.npz   a small admixture chain (N=100, L=500, K=3): state after init and
              after sweeps 1, 2 and 10 (P, Q, Z, Ni, Mi, alpha, lnL)
  zpass.npz   injected-state Z passes on a mixed-ploidy (2/4/6), 25%
              multiallelic dataset with locations: given P, Q and m, the
              resulting Ni and Mi, for K = 2, 5, 9

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O          # noqa: E402
import vcfpop_amd as V      # noqa: E402

SALTS = [0xDA310D821043E1D3, 0xE7853EB09C92935A, 0x9B83516C63E249C3, 0xD4EA1DB746DF14E3]
SEEDS = [0, 1, 12345, 0xFFFFFFFFFFFF]

CHAIN_SYNTH = dict(n_ind=100, n_loc=500, k_true=3, alpha_q=0.5, fst=0.05, missing=0.02, seed=77)
CHAIN_PARAMS = dict(admix_val=1, nburnin=4, nreps=20, nthinning=2, g_seed=9)
ZPASS_SYNTH = dict(n_ind=60, n_loc=400, k_true=4, alpha_q=0.5, fst=0.05, missing=0.03, frac_multi=0.25,
                   min_alleles=3, max_alleles=8, multi_dir=0.3, ploidy=[(2, 0.6), (4, 0.3), (6, 0.1)], n_pop=4,
                   seed=78)


def rng_vectors():
    out = {}
    for a, seed in enumerate(SEEDS):
        for b, salt in enumerate(SALTS):
            r = O.Rng(seed, salt)
            out[f"xs_{a}_{b}"] = np.array([r.xorshift() for _ in range(16)], dtype=np.uint64)
            r = O.Rng(seed, salt)
            out[f"unif_{a}_{b}"] = np.array([r.uniform() for _ in range(8)])
            r = O.Rng(seed, salt)
            out[f"normal_{a}_{b}"] = np.array([r.normal() for _ in range(8)])
            r = O.Rng(seed, salt)
            out[f"gamma_{a}_{b}"] = np.array([r.gamma(al) for al in (0.3, 1.0, 2.5, 11.0)])
            r = O.Rng(seed, salt)
            out[f"next_{a}_{b}"] = np.array([r.next(7) for _ in range(16)], dtype=np.uint64)
    lib = O.load()
    vals = np.array([0, 1, 2, 0xDEADBEEF, 0x0123456789ABCDEF, 2 ** 64 - 1], dtype=np.uint64)
    out["hash_in"] = vals
    out["hash_ulong"] = np.array([lib.or_hash_ulong(int(v)) for v in vals], dtype=np.uint32)
    out["murmur64"] = np.array([lib.or_murmur64(int(v), SALTS[2]) for v in vals], dtype=np.uint64)
    return out


def chain_vectors():
    s = V.Synth(**CHAIN_SYNTH)
    d = O.Data(s.desc)
    p = V.params_default(**CHAIN_PARAMS)
    c = O.Chain(d, p, 3, 1)
    out = {}
    done = 0
    for tag, upto in (("init", 0), ("m1", 1), ("m2", 2), ("m10", 10)):
        c.sweep(upto - done)
        done = upto
        for n in ("Freq", "Q", "Z", "Ni", "Mi", "Alpha"):
            out[f"{tag}_{n}"] = c.get(n)
        out[f"{tag}_lnL"] = np.array([c.record_pass()])
        # record_pass adds Mi into MiSum; nothing later reads MiSum here
    return out


def zpass_vectors():
    s = V.Synth(**ZPASS_SYNTH)
    d = O.Data(s.desc)
    orig, afo, vt = d.order()
    k2 = s.arrays()["n_allele"][orig]
    rng = np.random.default_rng(5)
    out = {}
    for K in (2, 5, 9):
        p = V.params_default(admix_val=1, g_seed=11)
        c = O.Chain(d, p, K, 1)
        Freq = np.zeros((K, d.KT))
        for l in range(len(afo)):
            Freq[:, afo[l]: afo[l] + k2[l]] = rng.dirichlet(np.ones(k2[l]), size=K)
        Q = rng.dirichlet(np.ones(K) * 0.7, size=d.N)
        m = int(rng.integers(0, 10000))
        c.set("Freq", Freq)
        c.set("Q", Q)
        c.set("m", [m])
        c.call("update_z_admix")
        out[f"K{K}_Freq"], out[f"K{K}_Q"], out[f"K{K}_m"] = Freq, Q, np.array([m])
        out[f"K{K}_Ni"], out[f"K{K}_Mi"] = c.get("Ni"), c.get("Mi")
    return out


def main():
    O.set_threads(4)
    np.savez_compressed(os.path.join(HERE, "rng.npz"), **rng_vectors())
    np.savez_compressed(os.path.join(HERE, "chain.npz"), **chain_vectors())
    np.savez_compressed(os.path.join(HERE, "zpass.npz"), **zpass_vectors())
    for f in ("rng.npz", "chain.npz", "zpass.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
