"""GPU parity against the oracle, through the C ABI (vcfpop_amd.Chain).

Bar: integer state (Z, Ni, Mi) bit-exact; FP64 state (P, Q, alpha, lnL,
log ratios) within 1e-12 relative (numbers that are sums of many logs of
mixed sign: 1e-13 of the sum of absolute terms).  The oracle's state is fed
to the device before each pass, as the device and host math libraries may
differ in the last ulp of log/pow (Gamma draws) and long chains may drift.
"""
import sys

import numpy as np
import pytest

from helpers import random_state

pytestmark = pytest.mark.gpu

RTOL = 1e-12

DATASETS = {
    # biallelic diploid (BI tiles only)
    "bi_diploid": dict(n_ind=96, n_loc=700, k_true=3, alpha_q=0.5, fst=0.05, missing=0.02, seed=21),
    # N a multiple of 64: chunks that end exactly on a 64-individual row group
    "bi_diploid_128": dict(n_ind=128, n_loc=300, k_true=3, alpha_q=0.5, fst=0.05, missing=0.02, seed=25),
    # mixed ploidy 2/4/6 with multiallelic loci and locations (BI and GEN tiles)
    "mixed": dict(n_ind=75, n_loc=500, k_true=4, alpha_q=0.5, fst=0.05, missing=0.03, frac_multi=0.25,
                  min_alleles=3, max_alleles=8, multi_dir=0.3, ploidy=[(2, 0.6), (4, 0.3), (6, 0.1)], n_pop=4,
                  seed=22),
    # ragged: N not a multiple of anything, a single short tile
    # enough individuals that UpdateAlpha's LogProd columns span several blocks
    "many_ind": dict(n_ind=5000, n_loc=24, n_pop=2, k_true=3, alpha_q=0.5, fst=0.05, missing=0.02, seed=27),
    "ragged": dict(n_ind=13, n_loc=37, k_true=2, alpha_q=1.0, missing=0.2, ploidy=[(1, 0.3), (3, 0.7)], seed=23),
    # the maximum ploidy (8 and 10 copies; BI tiles on the runtime slot loop, GEN descriptors of 10 alleles)
    "high_ploidy": dict(n_ind=40, n_loc=200, k_true=3, alpha_q=0.5, missing=0.05, frac_multi=0.4, min_alleles=3,
                        max_alleles=12, multi_dir=0.5, ploidy=[(8, 0.5), (10, 0.5)], n_pop=2, seed=26),
    # many alleles: K x alleles tables beyond one block's LDS at large K
    "many_alleles": dict(n_ind=40, n_loc=160, k_true=3, alpha_q=0.5, missing=0.05, frac_multi=0.6, min_alleles=12,
                         max_alleles=40, multi_dir=0.5, ploidy=[(2, 0.5), (4, 0.5)], n_pop=2, seed=24),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pairs(vps, orc):
    orc.set_threads(8)
    out = {}
    for name, spec in DATASETS.items():
        s = vps.Synth(**spec)
        out[name] = (s, vps.Dataset(s, 0), orc.Data(s.desc))
    return out


def make(vps, orc, pairs, name, K, run_id=1, **kw):
    s, ds, od = pairs[name]
    p = vps.params_default(**kw)
    return ds, od, vps.Chain(ds, p, K, run_id), orc.Chain(od, p, K, run_id)


def feed(g, o, names):
    for n in names:
        g.set(n, o.get(n))


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.allclose(a, b, rtol=rtol, atol=0)


@pytest.mark.parametrize("name", list(DATASETS))
@pytest.mark.parametrize("K", [1, 2, 3, 5, 8])
def test_init_admix(vps, orc, pairs, name, K):
    say("init", name, K)
    ds, od, g, o = make(vps, orc, pairs, name, K, admix_val=1, g_seed=5)
    for n in ("Z", "Mi", "Ni"):
        assert (g.get(n) == o.get(n)).all(), n
    # single-Z start (no admixture, no burn-in)
    ds, od, g, o = make(vps, orc, pairs, name, K, admix_val=2, nadmburnin=0, g_seed=6)
    for n in ("Z", "Mi", "Ni"):
        assert (g.get(n) == o.get(n)).all(), n


@pytest.mark.parametrize("name", list(DATASETS))
@pytest.mark.parametrize("K", [1, 2, 3, 4, 5, 6, 8, 11, 16])
def test_z_admix_bit_exact(vps, orc, pairs, name, K):
    say("z_admix", name, K)
    ds, od, g, o = make(vps, orc, pairs, name, K, admix_val=1, g_seed=7)
    orig, afo, vt = ds.order()
    k2 = pairs[name][0].arrays()["n_allele"][orig]
    rng = np.random.default_rng(K)
    for m in (0, 3, 1000):
        Freq, Q = random_state(rng, ds.N, K, ds.KT, afo, k2)
        for c in (g, o):
            c.set("Freq", Freq)
            c.set("Q", Q)
            c.set("m", [m])
        g.update_z_admix()
        o.call("update_z_admix")
        assert (g.get("Ni") == o.get("Ni")).all(), m
        assert (g.get("Mi") == o.get("Mi")).all(), m


@pytest.mark.parametrize("K", [12, 16])
def test_count_tables_beyond_lds(vps, orc, pairs, K):
    """initial counts and the no-admixture count pass when the K x alleles
    table does not fit in LDS (counts by global atomics)"""
    say("counts beyond LDS", K)
    ds, od, g, o = make(vps, orc, pairs, "many_alleles", K, admix_val=1, g_seed=15)
    assert ds.maxK * K * 64 * 4 > 64 * 1024
    for n in ("Z", "Mi", "Ni"):
        assert (g.get(n) == o.get(n)).all(), n
    Z = np.random.default_rng(K).integers(0, K, ds.N).astype(np.uint16)
    for c in (g, o):
        c.set("Z", Z)
    g.set("binaryq", [1])
    g.update_z_noadmix()
    o.call("update_z_noadmix")
    assert (g.get("Ni") == o.get("Ni")).all()
    assert (g.get("Mi") == o.get("Mi")).all()


@pytest.mark.parametrize("name", ["bi_diploid", "bi_diploid_128", "mixed"])
@pytest.mark.parametrize("z_chunk", [1, 7, 32, 64, 512])
def test_z_chunking_invariance(vps, orc, pairs, name, z_chunk):
    """every individual chunking of the genotype passes gives the oracle's
    Z pass (admixture and no-admixture) bit for bit"""
    say("z chunking", name, z_chunk)
    ds, od, g, o = make(vps, orc, pairs, name, 3, admix_val=1, g_seed=9, z_chunk=z_chunk)
    G, Nc = g.chunking()
    # z_chunk fixes the chunk count; chunks are then balanced
    assert G == -(-ds.N // z_chunk) and Nc == -(-ds.N // G)
    orig, afo, vt = ds.order()
    k2 = pairs[name][0].arrays()["n_allele"][orig]
    rng = np.random.default_rng(z_chunk)
    Freq, Q = random_state(rng, ds.N, 3, ds.KT, afo, k2)
    for c in (g, o):
        c.set("Freq", Freq)
        c.set("Q", Q)
        c.set("m", [11])
    g.update_z_admix()
    o.call("update_z_admix")
    assert (g.get("Ni") == o.get("Ni")).all()
    assert (g.get("Mi") == o.get("Mi")).all()
    Z = rng.integers(0, 3, ds.N).astype(np.uint16)
    for c in (g, o):
        c.set("Z", Z)
    g.update_z_noadmix()
    o.call("update_z_noadmix")
    assert (g.get("Ni") == o.get("Ni")).all()


def test_z_admix_full_size_admix_n2000_l100k_k5(vps, orc):
    """one Z pass at the benchmark size (N=2,000, L=100,000, K=5), bit-exact"""
    say("z_admix full size")
    orc.set_threads(16)
    s = vps.make_synth("admix_n2000_l100k_k5")
    ds, od = vps.Dataset(s, 0), orc.Data(s.desc)
    p = vps.params_default(admix_val=1, g_seed=1)
    g, o = vps.Chain(ds, p, 5, 1), orc.Chain(od, p, 5, 1)
    rng = np.random.default_rng(0)
    orig, afo, vt = ds.order()
    k2 = s.arrays()["n_allele"][orig]
    Freq, Q = random_state(rng, ds.N, 5, ds.KT, afo, k2)
    for c in (g, o):
        c.set("Freq", Freq)
        c.set("Q", Q)
        c.set("m", [42])
    g.update_z_admix()
    o.call("update_z_admix")
    assert (g.get("Ni") == o.get("Ni")).all()
    assert (g.get("Mi") == o.get("Mi")).all()
    # invariants at full size
    assert (g.get("Mi").sum(1) == vt).all()
    orc.set_threads(8)


@pytest.mark.parametrize("name", list(DATASETS))
@pytest.mark.parametrize("K", [2, 5])
def test_likelihood_passes(vps, orc, pairs, name, K):
    say("lik", name, K)
    ds, od, g, o = make(vps, orc, pairs, name, K, admix_val=1, g_seed=8)
    orig, afo, vt = ds.order()
    k2 = pairs[name][0].arrays()["n_allele"][orig]
    rng = np.random.default_rng(100 + K)
    Freq, Q = random_state(rng, ds.N, K, ds.KT, afo, k2)
    for c in (g, o):
        c.set("Freq", Freq)
        c.set("Q", Q)
    # Record lnL (admixture)
    g.record()
    ref = o.record_pass()
    assert abs(g.get("lnL")[0] - ref) <= RTOL * abs(ref)
    # Metropolis: proposals, per-individual log ratio, acceptance.  Each
    # copy's ratio is formed with the reference's rounding (products, sums,
    # then the division), so only the order of the product differs: 1e-12
    # relative per individual (absolute below |r| = 1).
    for c in (g, o):
        c.set("Q", Q)
        c.set("m", [10])
    g.update_q_metro()
    o.call("update_q_metro")
    r_g, r_o = g.get("bufN1"), o.get("bufN1")
    qp = o.get("bufNK1")
    assert close(g.get("bufNK1")[vt > 0], qp[vt > 0])
    tol = RTOL * np.maximum(np.abs(r_o), 1.0)
    assert (np.abs(r_g - r_o) <= tol).all()
    assert close(g.get("Q"), o.get("Q"))


@pytest.mark.parametrize("name", list(DATASETS))
@pytest.mark.parametrize("topology", [0, 1])
def test_q_noadmix(vps, orc, pairs, name, topology):
    say("q_noadmix", name, topology)
    K = 3
    ds, od, g, o = make(vps, orc, pairs, name, K, admix_val=2, nadmburnin=0, g_seed=9, rng_topology=topology)
    orig, afo, vt = ds.order()
    k2 = pairs[name][0].arrays()["n_allele"][orig]
    Freq, Q = random_state(np.random.default_rng(7), ds.N, K, ds.KT, afo, k2)
    for c in (g, o):
        c.set("Freq", Freq)
        c.set("m", [4])
    g.update_q_noadmix()
    o.call("update_q_noadmix")
    assert (g.get("Z")[vt > 0] == o.get("Z")[vt > 0]).all()
    assert (g.get("Q") == o.get("Q")).all()
    # the count pass for binary Q
    g.set("binaryq", [1])
    g.update_z_noadmix()
    o.call("update_z_noadmix")
    assert (g.get("Ni") == o.get("Ni")).all()
    assert (g.get("Mi") == o.get("Mi")).all()


@pytest.mark.parametrize("name", list(DATASETS))
@pytest.mark.parametrize("topology", [0, 1])
def test_dirichlet_updates(vps, orc, pairs, name, topology):
    say("dirichlet", name, topology)
    K = 4
    ds, od, g, o = make(vps, orc, pairs, name, K, admix_val=1, g_seed=10, rng_topology=topology)
    o.set("m", [3])
    o.call("update_z_admix")        # a realistic Ni / Mi
    feed(g, o, ["Ni", "Mi", "m"])
    g.update_p()
    o.call("update_p")
    assert close(g.get("Freq"), o.get("Freq"))
    g.update_q_admix()
    o.call("update_q_admix")
    assert close(g.get("Q"), o.get("Q"))


@pytest.mark.parametrize("name", ["bi_diploid", "many_ind"])
@pytest.mark.parametrize("kw", [dict(), dict(diffalpha_val=1), dict(uniformalpha_val=2), dict(locpriori_val=1)])
def test_update_alpha(vps, orc, pairs, name, kw):
    """global, per-cluster and per-location alpha; on many_ind every LogProd
    column spans several blocks (2,048 values a block)"""
    say("alpha", name, kw)
    ds, od, g, o = make(vps, orc, pairs, name, 3, admix_val=1, g_seed=11, **kw)
    Q = np.random.default_rng(3).dirichlet(np.ones(3), size=ds.N)
    for m in range(5):
        for c in (g, o):
            c.set("Q", Q)
            c.set("m", [m])
        feed(g, o, ["Alpha", "loc"] if kw.get("locpriori_val") else ["Alpha"])
        g.update_alpha()
        o.call("update_alpha")
        assert close(g.get("Alpha"), o.get("Alpha"))
        if kw.get("locpriori_val"):
            assert close(g.get("loc"), o.get("loc"))


def test_update_alpha_many_locations(vps, orc):
    """locprior alpha with more location x cluster columns (80 x 16) than the
    LogProd partial buffer's default capacity (1,024): one block per column"""
    say("alpha, 80 locations, K = 16")
    s = vps.Synth(n_ind=400, n_loc=24, n_pop=80, k_true=3, alpha_q=0.5, fst=0.05, missing=0.02, seed=29)
    ds, od = vps.Dataset(s, 0), orc.Data(s.desc)
    p = vps.params_default(admix_val=1, locpriori_val=1, g_seed=13)
    g, o = vps.Chain(ds, p, 16, 1), orc.Chain(od, p, 16, 1)
    Q = np.random.default_rng(5).dirichlet(np.ones(16), size=ds.N)
    for m in range(3):
        for c in (g, o):
            c.set("Q", Q)
            c.set("m", [m])
        feed(g, o, ["Alpha", "loc"])
        g.update_alpha()
        o.call("update_alpha")
        assert close(g.get("Alpha"), o.get("Alpha"))
        assert close(g.get("loc"), o.get("loc"))
    g.close()
    o.close()
    ds.close()
    od.close()


SWEEP_MODELS = {
    "admix": dict(admix_val=1),
    "admix_parallel": dict(admix_val=1, rng_topology=1),
    "noadmix": dict(admix_val=2, nadmburnin=2),
    "noadmix_parallel": dict(admix_val=2, nadmburnin=2, rng_topology=1),
    "locprior_admix": dict(admix_val=1, locpriori_val=1),
    "locprior_noadmix": dict(admix_val=2, locpriori_val=1),
    "fmodel": dict(admix_val=1, f_val=1),
    "fmodel_parallel": dict(admix_val=1, f_val=1, rng_topology=1),
    "fmodel_single": dict(admix_val=1, f_val=1, singlef_val=1),
    "lambda": dict(admix_val=1, inferlambda_val=0, difflambda_val=1),
    "lambda_single": dict(admix_val=1, inferlambda_val=0, difflambda_val=2),
    "lambda_parallel": dict(admix_val=1, inferlambda_val=0, difflambda_val=1, rng_topology=1),
    "locprior_admix_parallel": dict(admix_val=1, locpriori_val=1, rng_topology=1),
}

FED = ["Freq", "Q", "Z", "Ni", "Mi", "Alpha", "Lambda", "loc", "Fst", "FreqA", "binaryq", "MiSum", "FreqM", "rout"]


@pytest.mark.parametrize("model", list(SWEEP_MODELS))
def test_sweep_tracks_oracle(vps, orc, pairs, model):
    """full sweeps (all updates, Record every 2nd sweep), oracle state fed in first"""
    say("sweep", model)
    kw = dict(nburnin=3, nreps=8, nthinning=2, writelnl_val=1, g_seed=12)
    kw.update(SWEEP_MODELS[model])
    ds, od, g, o = make(vps, orc, pairs, "mixed", 3, run_id=2, **kw)
    for sweep in range(9):
        feed(g, o, FED + ["m"])
        g.step(1)
        o.sweep(1)
        for n in ("Z", "Ni", "Mi", "MiSum"):
            assert (g.get(n) == o.get(n)).all(), (sweep, n)
        for n in ("Freq", "Q", "Alpha", "Lambda", "loc", "Fst", "FreqA", "FreqM", "rout"):
            assert close(g.get(n), o.get(n), rtol=1e-11), (sweep, n)
    g.finish()
    o.finish()
    assert close(g.get("MiSum"), o.get("MiSum"))
    assert close(g.get("rout"), o.get("rout"), rtol=1e-11)


@pytest.mark.parametrize("topology", [0, 1])
def test_fmodel_sweeps_many_alleles(vps, orc, pairs, topology):
    """the F model's UpdateA at loci of 12-40 alleles and K = 4: on the
    reference's streams, A1's (cluster, allele) terms span several rounds of
    64 lanes; A1 and A2 both occur over the sweeps (each sweep picks one)"""
    say("F model, many alleles, topology", topology)
    kw = dict(admix_val=1, f_val=1, nburnin=3, nreps=8, nthinning=2, g_seed=17, rng_topology=topology)
    ds, od, g, o = make(vps, orc, pairs, "many_alleles", 4, run_id=1, **kw)
    for sweep in range(8):
        feed(g, o, FED + ["m"])
        g.step(1)
        o.sweep(1)
        for n in ("Z", "Ni", "Mi"):
            assert (g.get(n) == o.get(n)).all(), (sweep, n)
        for n in ("Freq", "Q", "Fst", "FreqA"):
            assert close(g.get(n), o.get(n), rtol=1e-11), (sweep, n)
    g.close()
    o.close()


def test_mixed_precision_z_agreement(vps, orc, pairs):
    """FP32 weights: same streams, so assignments differ from FP64 only where
    the float rounding crosses a cluster boundary; counts stay consistent"""
    say("mixed precision")
    K = 4
    ds, od, g, o = make(vps, orc, pairs, "mixed", K, admix_val=1, g_seed=13, precision=1)
    orig, afo, vt = ds.order()
    k2 = pairs["mixed"][0].arrays()["n_allele"][orig]
    Freq, Q = random_state(np.random.default_rng(9), ds.N, K, ds.KT, afo, k2)
    for c in (g, o):
        c.set("Freq", Freq)
        c.set("Q", Q)
        c.set("m", [2])
    g.update_z_admix()
    o.call("update_z_admix")
    Mg, Mo = g.get("Mi"), o.get("Mi")
    assert (Mg.sum(1) == vt).all()
    assert np.abs(Mg - Mo).sum() <= 0.001 * Mo.sum()
