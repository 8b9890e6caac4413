"""Pins the oracle's scalar updates by what their targets fix:

* UpdateP and UpdateQAdmix draw Dirichlet posteriors, so with fixed counts
  the sample mean over many sweeps must be (prior + counts) / total
  (src/structure.cpp:812-862, 986-1004);
* UpdateAlpha is a Metropolis step whose stationary law, with Q held fixed,
  is prior(alpha) * prod_i Dir(Q_i | alpha 1_K) (src/structure.cpp:1501-1594);
  the chain's samples are compared with that density integrated on a grid.
  A dropped or mis-signed term in the acceptance ratio moves the law.
"""
import numpy as np
import pytest
from scipy import special, stats


def _data(vps, orc, **kw):
    spec = dict(n_ind=30, n_loc=40, k_true=3, alpha_q=0.8, missing=0.0, seed=31)
    spec.update(kw)
    s = vps.Synth(**spec)
    return s, orc.Data(s.desc)


@pytest.mark.parametrize("topology", [0, 1])
def test_update_p_posterior_mean(vps, orc, topology):
    s, d = _data(vps, orc, frac_multi=0.3, min_alleles=3, max_alleles=5)
    K = 2
    p = vps.params_default(admix_val=1, g_seed=4, rng_topology=topology)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(0)
    Ni = rng.integers(0, 12, size=(K, d.KT)).astype(np.int32)
    ch.set("Ni", Ni)
    M = 3000
    acc = np.zeros((K, d.KT))
    acc2 = np.zeros((K, d.KT))
    for m in range(M):
        ch.set("m", [m])
        ch.call("update_p")
        f = ch.get("Freq")
        acc += f
        acc2 += f * f
    mean = acc / M
    sd = np.sqrt(np.maximum(acc2 / M - mean ** 2, 1e-30) / M)
    _, afo, _ = d.order()
    k2 = np.diff(np.append(afo, d.KT))
    lam = 1.0   # default lambda
    expect = np.zeros_like(mean)
    for l in range(len(afo)):
        a = Ni[:, afo[l]: afo[l] + k2[l]] + lam
        expect[:, afo[l]: afo[l] + k2[l]] = a / a.sum(1, keepdims=True)
    z = np.abs(mean - expect) / sd
    assert z.max() < 5.5, z.max()


@pytest.mark.parametrize("topology", [0, 1])
def test_update_q_admix_posterior_mean(vps, orc, topology):
    s, d = _data(vps, orc)
    K = 3
    p = vps.params_default(admix_val=1, g_seed=5, rng_topology=topology)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(1)
    Mi = rng.integers(0, 30, size=(d.N, K)).astype(np.int64)
    alpha = np.array([0.7, 1.3, 0.4])
    ch.set("Mi", Mi)
    ch.set("Alpha", alpha)
    M = 3000
    acc = np.zeros((d.N, K))
    acc2 = np.zeros((d.N, K))
    for m in range(M):
        ch.set("m", [m])
        ch.call("update_q_admix")
        q = ch.get("Q")
        acc += q
        acc2 += q * q
    mean = acc / M
    sd = np.sqrt(np.maximum(acc2 / M - mean ** 2, 1e-30) / M)
    a = Mi + alpha
    expect = a / a.sum(1, keepdims=True)
    z = np.abs(mean - expect) / sd
    assert z.max() < 5.5, z.max()


def _alpha_posterior_grid(Q, prior, amax):
    """normalised density of alpha on a grid: prior * prod_i Dir(Q_i | alpha)"""
    N, K = Q.shape
    sl = np.log(Q).sum()
    g = np.linspace(1e-4, amax, 20001)
    lp = N * (special.gammaln(K * g) - K * special.gammaln(g)) + (g - 1) * sl + prior(g)
    w = np.exp(lp - lp.max())
    return g, w / np.trapezoid(w, g)


@pytest.mark.parametrize("uniform", [True, False])
def test_update_alpha_stationary_law(vps, orc, uniform):
    s, d = _data(vps, orc)
    K = 3
    kw = dict(admix_val=1, g_seed=6, stdalpha=0.05)
    if not uniform:
        kw.update(uniformalpha_val=2, alphapriora=2.0, alphapriorb=0.5)
    p = vps.params_default(**kw)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(2)
    Q = rng.dirichlet(np.ones(K) * 0.6, size=d.N)
    ch.set("Q", Q)
    ch.set("Alpha", np.full(K, 0.6))
    burn, M, thin = 2000, 60000, 30
    out = []
    for m in range(burn + M):
        ch.set("m", [m])
        ch.call("update_alpha")
        if m >= burn and (m - burn) % thin == 0:
            out.append(ch.get("Alpha")[0])
    out = np.array(out)
    if uniform:
        prior = lambda g: np.zeros_like(g)
        amax = p.maxalpha
    else:
        prior = lambda g: (p.alphapriora - 1) * np.log(g) - g / p.alphapriorb
        amax = 10.0
    g, dens = _alpha_posterior_grid(Q, prior, amax)
    cdf = np.cumsum(dens) * (g[1] - g[0])
    mean = np.trapezoid(g * dens, g)
    sd = np.sqrt(np.trapezoid((g - mean) ** 2 * dens, g))
    # thinned Metropolis samples: compare moments and the KS distance loosely
    assert abs(out.mean() - mean) < 0.1 * sd + 4 * sd / np.sqrt(len(out) / 5), (out.mean(), mean, sd)
    assert abs(out.std() - sd) < 0.15 * sd, (out.std(), sd)
    ks = stats.kstest(out, lambda x: np.interp(x, g, cdf)).statistic
    assert ks < 0.06, ks


# UPDATEAlpha stream salt (src/global.h: salts of the per-purpose streams)
_SALT_UPDATE_ALPHA = 0x8A26E2C4D1BC5E54


@pytest.mark.parametrize("uniform", [True, False])
def test_update_alpha_per_cluster_restated(vps, orc, uniform):
    """Per-cluster alpha (diffalpha, src/structure.cpp:1561-1580) against a
    restatement with scipy's log-gamma and plain log sums: one Metropolis
    pass over k with normal proposals from the UPDATEAlpha stream, a uniform
    drawn only when the log ratio is below -1e-10, and the Dirichlet
    normaliser taken at the sum of the alphas before the pass (the
    reference does not update that sum as the alphas change, so the
    chain's stationary law is not the posterior and no law test applies)."""
    s, d = _data(vps, orc)
    K = 3
    kw = dict(admix_val=1, g_seed=8, stdalpha=0.3, diffalpha_val=1)
    if not uniform:
        kw.update(uniformalpha_val=2, alphapriora=2.0, alphapriorb=0.5)
    p = vps.params_default(**kw)
    ch = orc.Chain(d, p, K, 1)
    seed = ch.seed()
    rng = np.random.default_rng(3)
    moved = 0
    for t in range(200):
        Q = rng.dirichlet(np.ones(K) * 0.7, size=d.N)
        alpha = rng.uniform(0.2, 2.0, K)
        m = 50 + t
        ch.set("Q", Q)
        ch.set("Alpha", alpha)
        ch.set("m", [m])
        ch.call("update_alpha")
        got = ch.get("Alpha")
        r = orc.Rng(seed + m, _SALT_UPDATE_ALPHA)
        a = alpha.copy()
        S = alpha.sum()
        for k in range(K):
            am = r.normal(a[k], p.stdalpha)
            if am <= 0 or (am >= p.maxalpha and uniform):
                continue
            ao = a[k]
            dd = am - ao
            dl = 0.0 if uniform else (p.alphapriora - 1) * np.log(am / ao) + (ao - am) / p.alphapriorb
            dl += dd * np.log(Q[:, k]).sum() + (special.gammaln(S + dd) - special.gammaln(S)
                                               - special.gammaln(am) + special.gammaln(ao)) * d.N
            if dl >= -1e-10 or r.uniform() < np.exp(dl):
                a[k] = am
        assert np.allclose(got, a, rtol=1e-13, atol=0), (t, got, a)
        moved += int(np.any(got != alpha))
    assert 10 < moved < 200, moved


def _loc_pack(K, S, r, eta, gamma, alocal, di, sumalpha):
    """the "loc" state: r, Eta[K], Gamma[S][K], AlphaLocal[S][K], Di[S][K], SumAlpha[S]"""
    return np.concatenate([[r], eta, gamma.ravel(), alocal.ravel(), di.ravel(), sumalpha])


def _loc_unpack(v, K, S):
    o = 1 + K
    return (v[0], v[1:o], v[o:o + S * K].reshape(S, K), v[o + S * K:o + 2 * S * K].reshape(S, K),
            v[o + 2 * S * K:o + 3 * S * K].reshape(S, K), v[o + 3 * S * K:])


@pytest.mark.parametrize("m0", [40, 41])
def test_update_alpha_locprior_restated(vps, orc, m0):
    """The location-prior branch of UpdateAlpha (src/structure.cpp:1503-1556)
    against a restatement with scipy's log-gamma and plain log sums: global
    alpha_k given the local alphas, SumAlpha refreshed on sweeps divisible
    by 10, then each local alpha_sk given the Q rows of location s, the
    running SumAlpha updated on acceptance.  m0 = 40 refreshes SumAlpha
    first, 41 does not."""
    s, d = _data(vps, orc, n_pop=3)
    ind_pop = np.asarray(s.arrays()["ind_pop"])
    assert np.all(np.diff(ind_pop) >= 0)          # individuals grouped by location
    K, S = 3, 3
    p = vps.params_default(admix_val=1, locpriori_val=1, g_seed=9, stdalpha=0.3)
    ch = orc.Chain(d, p, K, 1)
    seed = ch.seed()
    rng = np.random.default_rng(5)
    moved = 0
    for t in range(60):
        m = m0 + 10 * t
        Q = rng.dirichlet(np.ones(K) * 0.7, size=d.N)
        alpha = rng.uniform(0.3, 2.0, K)
        r = rng.uniform(0.5, 3.0)
        al = rng.uniform(0.2, 2.0, (S, K))
        sa = al.sum(1) + rng.uniform(-0.1, 0.1, S)    # stale sums: only every 10th sweep refreshes them
        eta, gam, di = rng.uniform(0.1, 1, K), rng.uniform(0.1, 1, (S, K)), rng.uniform(0.1, 1, (S, K))
        ch.set("Q", Q)
        ch.set("Alpha", alpha)
        ch.set("loc", _loc_pack(K, S, r, eta, gam, al, di, sa))
        ch.set("m", [m])
        ch.call("update_alpha")
        got_a = ch.get("Alpha")
        _, _, _, got_al, _, got_sa = _loc_unpack(ch.get("loc"), K, S)
        g = orc.Rng(seed + m, _SALT_UPDATE_ALPHA)
        a, AL, SA = alpha.copy(), al.copy(), sa.copy()
        for k in range(K):
            am = g.normal(a[k], p.stdalpha)
            if 0 < am < p.maxalpha:
                dd = am - a[k]
                dl = S * (dd * r * np.log(r) + special.gammaln(r * a[k]) - special.gammaln(r * am))
                dl += r * dd * np.log(AL[:, k]).sum()
                if dl >= -1e-10 or g.uniform() < np.exp(dl):
                    a[k] = am
        if m % 10 == 0:
            SA = AL.sum(1)
        for sidx in range(S):
            rows = ind_pop == sidx
            n_s = int(rows.sum())
            for k in range(K):
                ao = AL[sidx, k]
                am = g.normal(ao, p.stdalpha)
                if 0 < am < p.maxalpha:
                    dd = am - ao
                    dl = (r * a[k] - 1) * np.log(am / ao) - dd * r + n_s * (
                        special.gammaln(SA[sidx] + dd) - special.gammaln(am) - special.gammaln(SA[sidx]) + special.gammaln(ao))
                    dl += dd * np.log(Q[rows, k]).sum()
                    if dl >= -1e-10 or g.uniform() < np.exp(dl):
                        AL[sidx, k] = am
                        SA[sidx] += dd
        assert np.allclose(got_a, a, rtol=1e-13, atol=0), (t, got_a, a)
        assert np.allclose(got_al, AL, rtol=1e-13, atol=0), (t, got_al, AL)
        assert np.allclose(got_sa, SA, rtol=1e-12, atol=1e-12), (t, got_sa, SA)
        moved += int(np.any(got_al != al))
    assert moved > 10, moved


# UPDATELoc stream salt (src/global.h)
_SALT_UPDATE_LOC = 0x95EE7991226E2D14


def _restate_locpriori(orc, seed, m, admix, p, K, S, alpha, r, eta, gam, al, di):
    """UpdateLocPriori, src/structure.cpp:1170-1255, with scipy's log-gamma and plain logs"""
    g = orc.Rng(seed + m, _SALT_UPDATE_LOC)
    lg = special.gammaln
    eta, gam = eta.copy(), gam.copy()
    rm = g.uniform(r - p.stdr, r + p.stdr)
    if admix:
        if 0 < rm < p.maxr:
            d = rm - r
            dt1 = rm * np.log(rm) - r * np.log(r)
            dl = 0.0
            for k in range(K):
                dl += S * (alpha[k] * dt1 - lg(rm * alpha[k]) + lg(r * alpha[k]))
                dl += alpha[k] * d * np.log(al[:, k]).sum() - al[:, k].sum() * d
            if dl >= -1e-10 or g.uniform() < np.exp(dl):
                r = rm
        return r, eta, gam
    if 0 < rm < p.maxr:
        dl = S * (lg(rm) - lg(r) + sum(lg(r * eta[k]) - lg(rm * eta[k]) for k in range(K)))
        dl += sum((rm - r) * eta[k] * np.log(gam[:, k]).sum() for k in range(K))
        if dl >= -1e-10 or g.uniform() < np.exp(dl):
            r = rm
    if K >= 2:
        i1 = g.next(K)
        j1 = g.next_avoid(K, i1)
        delta = g.uniform(0, p.epseta)
        e1, e2 = eta[i1] + delta, eta[j1] - delta
        if e1 < 1 and e2 > 0:
            dl = S * (lg(r * eta[i1]) + lg(r * eta[j1]) - lg(r * e1) - lg(r * e2))
            dl += r * delta * np.log(gam[:, i1] / gam[:, j1]).sum()
            if dl >= -1e-10 or g.uniform() < np.exp(dl):
                eta[i1] += delta
                eta[j1] -= delta
        for s in range(S):
            i1 = g.next(K)
            j1 = g.next_avoid(K, i1)
            delta = g.uniform(0, p.epsgamma)
            if gam[s, i1] + delta < 1 and gam[s, j1] - delta > 0:
                dl = ((r * eta[i1] - 1.0 + di[s, i1]) * np.log(1 + delta / gam[s, i1]) +
                      (r * eta[j1] - 1.0 + di[s, j1]) * np.log(1 - delta / gam[s, j1]))
                if dl >= -1e-10 or g.uniform() < np.exp(dl):
                    gam[s, i1] += delta
                    gam[s, j1] -= delta
    if m % 10 == 0:   # Unify, src/math2.cpp:2268-2285: (x + MIN_FREQ) / (sum + n MIN_FREQ), MIN_FREQ = 1e-10
        eta = (eta + 1e-10) / (eta.sum() + K * 1e-10)
        gam = (gam + 1e-10) / (gam.sum(1, keepdims=True) + K * 1e-10)
    return r, eta, gam


@pytest.mark.parametrize("admix", [True, False])
@pytest.mark.parametrize("m0", [30, 31])
def test_update_locpriori_restated(vps, orc, admix, m0):
    """r (and, without admixture, eta and gamma) Metropolis of the location
    prior against a restatement with scipy's log-gamma and plain logs; the
    uniform for acceptance is drawn only when the log ratio is below -1e-10,
    and eta and gamma are renormalised (with the 1e-10 floor) on sweeps
    divisible by 10 (m0 = 30)"""
    s, d = _data(vps, orc, n_pop=3)
    K, S = 3, 3
    p = vps.params_default(admix_val=1 if admix else 2, locpriori_val=1, g_seed=10, stdr=0.4, epseta=0.1,
                           epsgamma=0.1, nadmburnin=0)
    ch = orc.Chain(d, p, K, 1)
    seed = ch.seed()
    rng = np.random.default_rng(6)
    moved = 0
    for t in range(80):
        m = m0 + 10 * t
        alpha = rng.uniform(0.3, 2.0, K)
        r = rng.uniform(0.5, 3.0)
        eta = rng.dirichlet(np.ones(K) * 2)
        gam = rng.dirichlet(np.ones(K) * 2, size=S)
        al, di = rng.uniform(0.2, 2.0, (S, K)), rng.uniform(0.0, 20.0, (S, K))
        sa = al.sum(1)
        ch.set("Alpha", alpha)
        ch.set("loc", _loc_pack(K, S, r, eta, gam, al, di, sa))
        ch.set("m", [m])
        ch.call("update_locpriori")
        got_r, got_eta, got_gam, _, _, _ = _loc_unpack(ch.get("loc"), K, S)
        er, eeta, egam = _restate_locpriori(orc, seed, m, admix, p, K, S, alpha, r, eta, gam, al, di)
        assert np.isclose(got_r, er, rtol=1e-13, atol=0), (t, got_r, er)
        assert np.allclose(got_eta, eeta, rtol=1e-12, atol=1e-15), (t, got_eta, eeta)
        assert np.allclose(got_gam, egam, rtol=1e-12, atol=1e-15), (t, got_gam, egam)
        moved += int(got_r != r) + int(np.any(got_gam != gam))
    assert moved > 10, moved


def test_arrange_closed_form(vps, orc):
    """Arrange (src/structure.cpp:1862-1869): the run accumulators divided by
    the number of records, Var lnL = E[lnL^2] - E[lnL]^2, lnP(D) = E[lnL] -
    Var/2, FreqM divided by the number of records, and each MiSum row
    unified with the 1e-10 floor, (x + 1e-10) / (sum + K 1e-10)
    (src/math2.cpp:2268-2285), checked on a chain run to completion"""
    s, d = _data(vps, orc)
    K = 3
    nburnin, nreps, nthin = 20, 47, 5
    p = vps.params_default(admix_val=1, g_seed=12, nburnin=nburnin, nreps=nreps, nthinning=nthin)
    ch = orc.Chain(d, p, K, 1)
    ch.sweep(nburnin + nreps)
    nr = len(range(0, nreps, nthin))               # Record at m >= nburnin, (m - nburnin) % nthin == 0
    rout, misum, freqm = ch.get("rout").copy(), ch.get("MiSum").copy(), ch.get("FreqM").copy()
    assert misum.sum() > 0 and rout[0] < 0
    ch.finish()
    r2 = rout / nr
    var = r2[1] - r2[0] ** 2
    assert np.allclose(ch.get("rout")[4:], r2[4:], rtol=1e-15, atol=0)
    assert np.isclose(ch.get("rout")[0], r2[0], rtol=1e-15) and np.isclose(ch.get("rout")[1], r2[1], rtol=1e-15)
    assert np.isclose(ch.get("rout")[2], var, rtol=1e-9, atol=1e-9 * abs(r2[1]))
    assert np.isclose(ch.get("rout")[3], r2[0] - var / 2, rtol=1e-12)
    assert np.allclose(ch.get("FreqM"), freqm / nr, rtol=1e-15, atol=0)
    ms = misum.astype(np.float64)
    expect = (ms + 1e-10) / (ms.sum(1, keepdims=True) + K * 1e-10)
    assert np.allclose(ch.get("MiSum"), expect, rtol=1e-14, atol=0)


# UPDATEA / UPDATEF stream salts (src/global.h)
_SALT_UPDATE_A = 0x49913AD9EDBD2EB8
_SALT_UPDATE_F = 0xF82104A5ADDAAF9F


def _fmodel_state(rng, K, KT, k2s, afos):
    """random cluster and ancestral frequencies (per locus on the simplex) and F"""
    P = np.empty((K, KT))
    A = np.zeros((3, KT))
    for k2, o in zip(k2s, afos):
        P[:, o:o + k2] = rng.dirichlet(np.ones(k2), size=K)
        A[0, o:o + k2] = rng.dirichlet(np.ones(k2) * 2)
    F = rng.uniform(0.05, 0.5, K)
    return P, A, F


@pytest.mark.parametrize("topology", [0, 1])
def test_update_a_restated(vps, orc, topology):
    """UpdateA of the F model (src/structure.cpp:1639-1804) against a
    restatement with scipy's log-gamma and plain logs: HashULong(seed + m) & 1
    picks the Dirichlet proposal (UpdateA1) or the pairwise shift (UpdateA2);
    the reference topology draws from 32 streams over consecutive locus
    blocks, the parallel one from one stream per locus"""
    s, d = _data(vps, orc, frac_multi=0.4, min_alleles=3, max_alleles=5)
    K = 3
    p = vps.params_default(admix_val=1, f_val=1, g_seed=13, rng_topology=topology)
    ch = orc.Chain(d, p, K, 1)
    seed, L, N, KT = ch.seed(), d.L, d.N, d.KT
    orig, afo, _ = d.order()
    k2 = np.asarray(s.arrays()["n_allele"])[orig]
    lib = orc.load()
    rng = np.random.default_rng(7)
    kinds = set()
    for t in range(24):
        m = 100 + t
        P, A, F = _fmodel_state(rng, K, KT, k2, afo)
        f = (1 - F) / F
        lam = rng.uniform(0.5, 2.0)
        ch.set("Freq", P)
        ch.set("FreqA", A)
        ch.set("Fst", np.concatenate([F, f]))
        ch.set("Lambda", np.full(K, lam))
        ch.set("m", [m])
        ch.call("update_a")
        got = ch.get("FreqA")[0]
        a1 = lib.or_hash_ulong(seed + m) & 1
        kinds.add(a1)
        pa = A[0].copy()

        def locus(g, l):
            n, o = int(k2[l]), int(afo[l])
            if a1:
                pb = lam + (P[:, o:o + n] * f[:, None]).sum(0)
                pc = g.dirichlet_a(pb)
                dl = ((pb - lam) * np.log(pa[o:o + n] / pc)).sum()
                for k in range(K):
                    dl += (special.gammaln(f[k] * pa[o:o + n]) - special.gammaln(f[k] * pc)
                           + f[k] * (pc - pa[o:o + n]) * np.log(P[k, o:o + n])).sum()
                if dl >= -1e-10 or g.uniform() < np.exp(dl):
                    pa[o:o + n] = pc
            else:
                m1 = g.next(n)
                n1 = g.next_avoid(n, m1)
                delta = g.uniform(0, N ** -0.5)
                pm0, pn0 = pa[o + m1], pa[o + n1]
                pm1, pn1 = pm0 + delta, pn0 - delta
                if pm1 >= 1 or pn1 <= 0:
                    return
                dl = (lam - 1) * np.log(pm1 * pn1 / pm0 / pn0)
                for k in range(K):
                    dl += (special.gammaln(f[k] * pm0) + special.gammaln(f[k] * pn0) - special.gammaln(f[k] * pm1)
                           - special.gammaln(f[k] * pn1) + f[k] * delta * np.log(P[k, o + m1] / P[k, o + n1]))
                if dl >= -1e-10 or g.uniform() < np.exp(dl):
                    pa[o + m1], pa[o + n1] = pm1, pn1

        if topology == 0:
            for b in range(32):
                g = orc.Rng(seed + m * 32 + b, _SALT_UPDATE_A)
                for l in range(b * L // 32, (b + 1) * L // 32):
                    locus(g, l)
        else:
            for l in range(L):
                locus(orc.Rng(seed + m * L + l, _SALT_UPDATE_A), l)
        assert np.allclose(got, pa, rtol=1e-12, atol=1e-15), (t, a1, np.abs(got - pa).max())
        assert np.any(got != A[0])
    assert kinds == {0, 1}


@pytest.mark.parametrize("fsame", [False, True])
def test_update_f_restated(vps, orc, fsame):
    """UpdateF (src/structure.cpp:1808-1858) against a restatement with
    scipy's log-gamma and plain logs: F_k by a normal proposal, the Gamma-like
    prior terms, the ancestral-frequency likelihood over every allele; with a
    single F the first proposal moves all clusters at once"""
    s, d = _data(vps, orc, frac_multi=0.4, min_alleles=3, max_alleles=5)
    K = 3
    p = vps.params_default(admix_val=1, f_val=1, g_seed=14, singlef_val=1 if fsame else 2, stdf=0.05)
    ch = orc.Chain(d, p, K, 1)
    seed, L, KT = ch.seed(), d.L, d.KT
    orig, afo, _ = d.order()
    k2 = np.asarray(s.arrays()["n_allele"])[orig]
    rng = np.random.default_rng(8)
    moved = 0
    for t in range(60):
        m = 200 + t
        P, A, F = _fmodel_state(rng, K, KT, k2, afo)
        if fsame:
            F[:] = F[0]
        f = (1 - F) / F
        ch.set("Freq", P)
        ch.set("FreqA", A)
        ch.set("Fst", np.concatenate([F, f]))
        ch.set("m", [m])
        ch.call("update_f")
        got = ch.get("Fst").ravel()
        g = orc.Rng(seed + m, _SALT_UPDATE_F)
        F2, f2 = F.copy(), f.copy()
        pa = A[0]
        for k in range(K):
            Fo = F2[k]
            Fm = g.normal(Fo, p.stdf)
            if Fm < 0 or Fm > 1:
                if fsame:
                    break
                continue
            fo, fm = f2[k], (1 - Fm) / Fm
            dl = (p.pmeanf * (Fo - Fm) / p.pstdf ** 2 + (p.pmeanf ** 2 / p.pstdf ** 2 - 1) * np.log(Fm / Fo)
                  + (K if fsame else 1) * L * (special.gammaln(fm) - special.gammaln(fo)))
            for kk in range(k, K):
                dl += (pa * (fm - fo) * np.log(P[kk]) + special.gammaln(fo * pa) - special.gammaln(fm * pa)).sum()
                if not fsame:
                    break
            if dl >= -1e-10 or g.uniform() < np.exp(dl):
                if fsame:
                    F2[:], f2[:] = Fm, fm
                else:
                    F2[k], f2[k] = Fm, fm
            if fsame:
                break
        assert np.allclose(got, np.concatenate([F2, f2]), rtol=1e-13, atol=0), (t, got, F2, f2)
        moved += int(np.any(got[:K] != F))
    assert moved > 5, moved


# UPDATELambda stream salt (src/global.h)
_SALT_UPDATE_LAMBDA = 0x67099E2FE2510092


@pytest.mark.parametrize("diff, fmodel", [(False, False), (True, False), (False, True)])
def test_update_lambda_restated(vps, orc, diff, fmodel):
    """UpdateLambda (src/structure.cpp:1598-1635; never reached in the
    reference, whose option parsing cannot make inferlambda true, but kept
    reachable here with inferlambda_val = 0) against a restatement with
    scipy's log-gamma: the Dirichlet normalisers weighted by the number of
    loci with i alleles, and the log product of the cluster (or, in the F
    model, ancestral) frequencies"""
    s, d = _data(vps, orc, frac_multi=0.4, min_alleles=3, max_alleles=5)
    K = 3
    p = vps.params_default(admix_val=1, f_val=1 if fmodel else 2, inferlambda_val=0,
                           difflambda_val=1 if diff else 2, stdlambda=0.2, g_seed=15)
    ch = orc.Chain(d, p, K, 1)
    seed, KT = ch.seed(), d.KT
    orig, afo, _ = d.order()
    k2 = np.asarray(s.arrays()["n_allele"])[orig]
    kdis = np.bincount(k2)
    rng = np.random.default_rng(9)
    lg = special.gammaln
    moved = 0
    for t in range(60):
        m = 400 + t
        P, A, F = _fmodel_state(rng, K, KT, k2, afo)
        lam = rng.uniform(0.3, 2.0, K) if diff else np.full(K, rng.uniform(0.3, 2.0))
        ch.set("Freq", P)
        ch.set("FreqA", A)
        ch.set("Lambda", lam)
        ch.set("m", [m])
        ch.call("update_lambda")
        got = ch.get("Lambda")
        g = orc.Rng(seed + m, _SALT_UPDATE_LAMBDA)
        e = lam.copy()

        def dl_of(lo, lm, np_, logprod):
            gd = lg(lo) - lg(lm)
            s_ = sum(kdis[i] * np_ * (lg(i * lm) - lg(i * lo) + i * gd) for i in range(2, len(kdis)))
            return s_ + (lm - lo) * logprod

        if diff:
            for k in range(K):
                lo = e[k]
                lm = g.normal(lo, p.stdlambda)
                if lm <= 0 or lm >= p.maxlambda:
                    continue
                dl = dl_of(lo, lm, 1, np.log(P[k]).sum())
                if dl >= -1e-10 or g.uniform() < np.exp(dl):
                    e[k] = lm
        else:
            lo = e[0]
            lm = g.normal(lo, p.stdlambda)
            if 0 < lm < p.maxlambda:
                lp = np.log(A[0]).sum() if fmodel else np.log(P).sum()
                dl = dl_of(lo, lm, 1 if fmodel else K, lp)
                if dl >= -1e-10 or g.uniform() < np.exp(dl):
                    e[:] = lm
        assert np.allclose(got, e, rtol=1e-13, atol=0), (t, got, e)
        moved += int(np.any(got != lam))
    assert moved > 5, moved
