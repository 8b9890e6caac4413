"""Pins the oracle's genotype passes and dataset handling on small inputs:
closed forms for the likelihood passes (direct sums of logs), exact count
invariants and a brute-force distribution check for the Z pass, and the
locus order of CalcBayesian's prologue."""
import numpy as np
import pytest
from scipy import stats

from helpers import dense_genotypes, lnl_direct, metro_direct, noadmix_direct, random_state


def _small(vps, **kw):
    spec = dict(n_ind=40, n_loc=150, k_true=3, alpha_q=0.5, missing=0.05, frac_multi=0.25,
                ploidy=[(2, 0.5), (3, 0.2), (4, 0.3)], n_pop=3, seed=11)
    spec.update(kw)
    return vps.Synth(**spec)


def _setup(vps, orc, K=3, **kw):
    s = _small(vps, **kw)
    a = s.arrays()
    d = orc.Data(s.desc)
    order = d.order()
    p = vps.params_default(admix_val=1, g_seed=3)
    ch = orc.Chain(d, p, K, 1)
    geno = dense_genotypes(a, order)
    k2 = [int(a["n_allele"][o]) for o in order[0]]
    return s, a, d, order, ch, geno, k2


def test_locus_order_is_stable_bitsize_sort(vps, orc):
    s = _small(vps)
    a = s.arrays()
    orig = orc.Data(s.desc).order()[0]
    expect = np.argsort(a["bitsize"], kind="stable")
    assert (orig == expect).all()


def test_vt_counts_typed_copies(vps, orc):
    s, a, d, order, ch, geno, k2 = _setup(vps, orc)
    vt = order[2]
    for i in range(a["N"]):
        assert vt[i] == sum(len(geno[l][i]) for l in range(a["L"]))


def test_record_lnl_closed_form(vps, orc):
    s, a, d, order, ch, geno, k2 = _setup(vps, orc)
    rng = np.random.default_rng(1)
    Freq, Q = random_state(rng, a["N"], 3, d.KT, order[1], k2)
    ch.set("Freq", Freq)
    ch.set("Q", Q)
    got = ch.record_pass()
    exact = lnl_direct(geno, order[1], Freq, Q)
    assert abs(got - exact) <= 1e-12 * abs(exact)


def test_metro_logratio_closed_form(vps, orc):
    s, a, d, order, ch, geno, k2 = _setup(vps, orc)
    rng = np.random.default_rng(2)
    Freq, Q = random_state(rng, a["N"], 3, d.KT, order[1], k2)
    Qp = rng.dirichlet(np.ones(3), size=a["N"])
    ch.set("Freq", Freq)
    ch.set("Q", Q)
    got = ch.metro_logratio(Qp)
    exact = metro_direct(geno, order[1], Freq, Q, Qp)
    scale = sum(abs(np.log(np.dot(Qp[i], Freq[:, order[1][l] + x]) / np.dot(Q[i], Freq[:, order[1][l] + x])))
                for l, row in enumerate(geno) for i, als in enumerate(row) for x in als)
    assert np.max(np.abs(got - exact)) <= 1e-13 * scale


def test_noadmix_loglik_closed_form(vps, orc):
    s, a, d, order, ch, geno, k2 = _setup(vps, orc)
    rng = np.random.default_rng(3)
    Freq, Q = random_state(rng, a["N"], 3, d.KT, order[1], k2)
    ch.set("Freq", Freq)
    got = ch.noadmix_loglik()
    exact = noadmix_direct(geno, order[1], Freq, a["N"], 3)
    assert np.allclose(got, exact, rtol=1e-12, atol=0)


def test_z_admix_count_invariants(vps, orc):
    s, a, d, order, ch, geno, k2 = _setup(vps, orc, K=4)
    rng = np.random.default_rng(4)
    Freq, Q = random_state(rng, a["N"], 4, d.KT, order[1], k2)
    ch.set("Freq", Freq)
    ch.set("Q", Q)
    for m in (0, 1, 17):
        ch.set("m", [m])
        ch.call("update_z_admix")
        Ni, Mi = ch.get("Ni"), ch.get("Mi")
        afo = order[1]
        # sum_k Ni = typed copies per allele; sum_k Mi = vt; column sums agree
        for l in range(a["L"]):
            cnt = np.zeros(k2[l], dtype=np.int64)
            for als in geno[l]:
                for x in als:
                    cnt[x] += 1
            assert (Ni[:, afo[l]: afo[l] + k2[l]].sum(0) == cnt).all()
        assert (Mi.sum(1) == order[2]).all()
        assert (Ni.sum(1) == Mi.sum(0)).all()


def test_z_admix_brute_force_distribution(vps, orc):
    """one typed copy, K clusters: over many sweeps (independent streams) the
    cluster frequencies match q_k p_k + eps normalised"""
    import ctypes as C
    N, L = 1, 1
    arrays = dict(N=1, L=1, S=1, ind_pop=np.zeros(1, np.int32), geno_bits=np.zeros(8, np.uint8),
                  geno_byte_offset=np.zeros(1, np.uint64), bitsize=np.ones(1, np.uint8),
                  n_allele=np.array([2], np.int32), gtab_offset=np.array([0, 1], np.int64),
                  gt_ploidy=np.array([1], np.uint8), gt_allele_offset=np.zeros(1, np.int64),
                  allele_pool=np.array([1], np.uint16))
    from vcfpop_amd._lib import desc_from_arrays
    desc = desc_from_arrays(arrays)
    d = orc.Data(desc)
    ch = orc.Chain(d, vps.params_default(admix_val=1), 3, 1)
    Freq = np.array([[0.7, 0.3], [0.2, 0.8], [0.5, 0.5]])
    Q = np.array([[0.2, 0.5, 0.3]])
    ch.set("Freq", Freq)
    ch.set("Q", Q)
    counts = np.zeros(3)
    for m in range(6000):
        ch.set("m", [m])
        ch.call("update_z_admix")
        counts += ch.get("Mi")[0]
    w = Q[0] * Freq[:, 1] + 1e-10
    assert stats.chisquare(counts, w / w.sum() * counts.sum()).pvalue > 1e-3


def test_z_admix_thread_invariance(vps, orc):
    s, a, d, order, ch, geno, k2 = _setup(vps, orc, K=3)
    rng = np.random.default_rng(5)
    Freq, Q = random_state(rng, a["N"], 3, d.KT, order[1], k2)
    res = []
    for nt in (1, 3, 8):
        orc.set_threads(nt)
        ch.set("Freq", Freq)
        ch.set("Q", Q)
        ch.set("m", [5])
        ch.call("update_z_admix")
        res.append((ch.get("Ni"), ch.get("Mi")))
    orc.set_threads(1)
    for r in res[1:]:
        assert (r[0] == res[0][0]).all() and (r[1] == res[0][1]).all()


def test_full_sweeps_run_all_models(vps, orc):
    """every model runs and keeps Q rows on the simplex"""
    s = _small(vps)
    d = orc.Data(s.desc)
    for kw in (dict(admix_val=1), dict(admix_val=2, nadmburnin=3), dict(admix_val=1, locpriori_val=1),
               dict(admix_val=2, locpriori_val=1), dict(admix_val=1, f_val=1), dict(admix_val=1, diffalpha_val=1),
               dict(admix_val=1, inferlambda_val=0, difflambda_val=1)):
        p = vps.params_default(nburnin=4, nreps=6, nthinning=2, g_seed=9, **kw)
        ch = orc.Chain(d, p, 3, 2)
        ch.sweep(10)
        Q = ch.get("Q")
        vt = d.order()[2]
        assert np.allclose(Q[vt > 0].sum(1), 1.0, atol=1e-9), kw
        ch.finish()
        assert np.isfinite(ch.get("rout")).all(), kw


def test_z_noadmix_counts_closed_form(vps, orc):
    """UpdateZNoAdmix (src/structure.cpp:1259-1301): Mi[i][Z_i] = typed
    copies of i, Ni[Z_i][l][x] counts every typed copy x, from a dense decode"""
    s, a, d, order, ch, geno, k2 = _setup(vps, orc)
    K = 3
    rng = np.random.default_rng(4)
    Z = rng.integers(0, K, a["N"]).astype(np.uint16)
    ch.set("Z", Z)
    ch.call("update_z_noadmix")
    afo, vt = order[1], order[2]
    Ni = np.zeros((K, d.KT), dtype=np.int64)
    Mi = np.zeros((a["N"], K), dtype=np.int64)
    for l, row in enumerate(geno):
        for i, als in enumerate(row):
            for x in als:
                Ni[Z[i], afo[l] + x] += 1
    Mi[np.arange(a["N"]), Z] = vt
    assert (ch.get("Ni") == Ni).all()
    assert (ch.get("Mi") == Mi).all()


# UPDATEQ stream salt (src/global.h)
_SALT_UPDATE_Q = 0xE7853EB09C92935A


@pytest.mark.parametrize("topology", [0, 1])
def test_q_noadmix_draw_restated(vps, orc, topology):
    """UpdateQNoAdmix's draw (src/structureAVX.cpp:279-287): per typed
    individual, in order, PolyLog over its K log likelihoods (pinned above)
    picks Z_i and Q_i becomes that one-hot row; the reference topology draws
    from one UPDATEQ stream, the parallel one from one stream per individual"""
    s, a, d, order, ch, geno, k2 = _setup(vps, orc)
    K, N = 3, a["N"]
    p = vps.params_default(admix_val=2, g_seed=5, rng_topology=topology, nadmburnin=0)
    ch = orc.Chain(d, p, K, 1)
    seed = ch.seed()
    rng = np.random.default_rng(6)
    vt = order[2]
    for t in range(5):
        m = 300 + t
        Freq, Q = random_state(rng, N, K, d.KT, order[1], k2)
        ch.set("Freq", Freq)
        ch.set("m", [m])
        ll = ch.noadmix_loglik()
        ch.call("update_q_noadmix")
        Zg, Qg = ch.get("Z"), ch.get("Q")
        g = orc.Rng(seed + m, _SALT_UPDATE_Q) if topology == 0 else None
        for i in range(N):
            if not vt[i]:
                assert (Qg[i] == 0).all()
                continue
            if topology == 1:
                g = orc.Rng(seed + m * N + i, _SALT_UPDATE_Q)
            k = g.polylog(ll[i])
            assert Zg[i] == k, (t, i)
            assert Qg[i, k] == 1 and Qg[i].sum() == 1
