"""The device genotype image (tiles, transposed id words, decode tables) built
on the host: every (locus, individual) decodes through the tile layout to the
genotype the packed description holds (LSB-first ids of CeilLog2(ngeno) bits,
src/math2.h:503-512; genotype tables as in src/structure.cpp:2202-2273's
sorted locus order).  Covers biallelic loci of every id width the biallelic
kernels take (up to 64 genotypes), multiallelic loci and missing genotypes."""
import numpy as np
import pytest


def raw_genotype(a, o, i):
    b = int(a["bitsize"][o])
    bit = int(a["geno_byte_offset"][o]) * 8 + i * b
    gid = 0
    for x in range(b):
        byte = a["geno_bits"][(bit + x) >> 3]
        gid |= ((int(byte) >> ((bit + x) & 7)) & 1) << x
    g = int(a["gtab_offset"][o]) + gid
    v = int(a["gt_ploidy"][g])
    s = int(a["gt_allele_offset"][g])
    return list(a["allele_pool"][s: s + v])


@pytest.mark.parametrize("ploidy,frac_multi,missing", [
    ([(2, 1.0)], 0.0, 0.02),
    ([(2, 0.6), (4, 0.3), (6, 0.1)], 0.2, 0.05),
    ([(4, 0.5), (8, 0.5)], 0.1, 0.3),
])
def test_every_genotype_decodes(vps, ploidy, frac_multi, missing):
    s = vps.Synth(n_ind=37, n_loc=300, k_true=3, alpha_q=0.5, missing=missing, frac_multi=frac_multi,
                  min_alleles=3, max_alleles=6, ploidy=ploidy, seed=31)
    a = s.arrays()
    ds = vps.Dataset(s, -1)
    orig, afo, vt = ds.order()
    for l in range(ds.L):
        o = int(orig[l])
        for i in range(ds.N):
            assert list(ds.decode(l, i)) == raw_genotype(a, o, i), (l, o, i)


def test_wide_biallelic_tables_present(vps):
    """Mixed ploidy with missing data gives biallelic loci of 4-5 id bits."""
    s = vps.Synth(n_ind=200, n_loc=200, k_true=3, alpha_q=0.5, missing=0.05, frac_multi=0.0,
                  ploidy=[(2, 0.6), (4, 0.3), (6, 0.1)], seed=32)
    a = s.arrays()
    assert (a["n_allele"] <= 2).all() and (a["bitsize"] > 3).any()
