"""Shared test helpers: dense genotype views of small datasets, comparisons."""
import numpy as np


def dense_genotypes(arrays, order):
    """typed allele copies of (sorted locus l, individual i) as a list of arrays"""
    orig = order[0]
    N, L = arrays["N"], arrays["L"]
    gb = arrays["geno_bits"]
    out = []
    for l in range(L):
        o = orig[l]
        b = int(arrays["bitsize"][o])
        off = int(arrays["geno_byte_offset"][o])
        g0 = int(arrays["gtab_offset"][o])
        row = []
        for i in range(N):
            bit = i * b
            w = int.from_bytes(bytes(gb[off + (bit >> 3): off + (bit >> 3) + 4].tolist() + [0, 0, 0, 0])[:4], "little")
            gid = (w >> (bit & 7)) & ((1 << b) - 1)
            v = int(arrays["gt_ploidy"][g0 + gid])
            ao = int(arrays["gt_allele_offset"][g0 + gid])
            row.append(arrays["allele_pool"][ao: ao + v].astype(np.int64))
        out.append(row)
    return out


def lnl_direct(geno, afo, Freq, Q):
    """sum over typed copies of log(sum_k Q[i,k] P[k, afo[l] + x])"""
    s = 0.0
    for l, row in enumerate(geno):
        for i, als in enumerate(row):
            for x in als:
                s += np.log(np.dot(Q[i], Freq[:, afo[l] + x]))
    return s


def metro_direct(geno, afo, Freq, Q, Qp):
    out = np.zeros(len(Q))
    for l, row in enumerate(geno):
        for i, als in enumerate(row):
            for x in als:
                p = Freq[:, afo[l] + x]
                out[i] += np.log(np.dot(Qp[i], p) / np.dot(Q[i], p))
    return out


def noadmix_direct(geno, afo, Freq, N, K):
    out = np.zeros((N, K))
    for l, row in enumerate(geno):
        for i, als in enumerate(row):
            for x in als:
                out[i] += np.log(Freq[:, afo[l] + x])
    return out


def random_state(rng, N, K, KT, afo, k2):
    """a Dirichlet P per (k, locus) and a Dirichlet Q per individual"""
    Freq = np.zeros((K, KT))
    for l in range(len(afo)):
        Freq[:, afo[l]: afo[l] + k2[l]] = rng.dirichlet(np.ones(k2[l]), size=K)
    Q = rng.dirichlet(np.ones(K) * 0.7, size=N)
    return Freq, Q
