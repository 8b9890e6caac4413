"""Pins the oracle's Metropolis steps by the laws they target.

Each step of the location prior (src/structure.cpp:1170-1255), of UpdateAlpha
(:1501-1594), UpdateLambda (:1598-1635), UpdateA1/A2 (:1650-1804) and UpdateF
(:1808-1858) is a Metropolis-Hastings move whose target, with the rest of the
state held, is a known conditional posterior.  The oracle step is iterated
with that state held, and the moved parameter's samples are compared with the
target density integrated on a grid (Kolmogorov-Smirnov, scaled by the
chain's effective sample size).

Steps not under test are held by the reference's own mechanics: a proposal
width of 0 proposes the current value (accepted with log ratio 0), a
parameter above its upper bound never moves, or the held state is reset
before each call.  Sweep indices avoid multiples of 10, where Eta and Gamma
are renormalised.

Each test also checks the same samples against the target with one term
dropped or one sign flipped and requires those laws to be rejected: the
checks are sensitive to exactly the mistakes a transcription can make.
"""
import numpy as np
import pytest
from scipy import integrate, special, stats

lg = special.gammaln


def _data(vps, orc, **kw):
    spec = dict(n_ind=30, n_loc=40, k_true=3, alpha_q=0.8, missing=0.0, seed=31)
    spec.update(kw)
    s = vps.Synth(**spec)
    return s, orc.Data(s.desc)


def _ks(x, g, logd):
    w = np.exp(logd - logd.max())
    cdf = integrate.cumulative_trapezoid(w, g, initial=0.0)
    cdf /= cdf[-1]
    return stats.kstest(x, lambda v: np.interp(v, g, cdf)).statistic


def check_law(x, g, logd, wrong=(), label=""):
    """KS distance of chain samples x from the law exp(logd) on grid g,
    within 2/sqrt(n_eff) (n_eff from the lag-1 autocorrelation); each
    mutated law in `wrong` must sit at least twice that far away"""
    x = np.asarray(x, float)
    n = len(x)
    r1 = float(np.corrcoef(x[:-1], x[1:])[0, 1])
    neff = n * (1 - r1) / (1 + r1) if r1 > 0 else n
    thr = 2.0 / np.sqrt(neff)
    k0 = _ks(x, g, logd)
    assert k0 < thr, (label, "law", k0, thr, neff)
    for name, lw in wrong:
        kw = _ks(x, g, lw)
        assert kw > 2 * thr, (label, "mutation not rejected", name, kw, thr)
    return k0, thr


def _loc_pack(r, eta, gamma, alocal, di, sumalpha):
    """the "loc" state: r, Eta[K], Gamma[S][K], AlphaLocal[S][K], Di[S][K], SumAlpha[S]"""
    return np.concatenate([[r], eta, gamma.ravel(), alocal.ravel(), di.ravel(), sumalpha])


def _loc_unpack(v, K, S):
    o = 1 + K
    return (v[0], v[1:o], v[o:o + S * K].reshape(S, K), v[o + S * K:o + 2 * S * K].reshape(S, K),
            v[o + 2 * S * K:o + 3 * S * K].reshape(S, K), v[o + 3 * S * K:])


def _sweep_indices(n, start=1):
    """n sweep indices, none divisible by 10"""
    m = np.arange(start, start + n + n // 9 + 2)
    return m[m % 10 != 0][:n]


def _run(ch, call, n, read, reset=None, burn=1000, thin=10, start=1):
    out = []
    for t, m in enumerate(_sweep_indices(burn + n * thin, start)):
        if reset:
            reset()
        ch.set("m", [int(m)])
        ch.call(call)
        if t >= burn and (t - burn) % thin == 0:
            out.append(read())
    return np.array(out)


# ---------------------------------------------------------------------------
# UpdateLocPriori, src/structure.cpp:1170-1255
# ---------------------------------------------------------------------------
def test_locprior_r_admix_law(vps, orc):
    """admixture: r targets prod_{s,k} Gamma(AlphaLocal_sk | shape r alpha_k,
    rate r) on (0, maxr) (uniform prior, symmetric uniform proposal)"""
    K, S = 3, 4
    s, d = _data(vps, orc, n_pop=S)
    p = vps.params_default(admix_val=1, locpriori_val=1, g_seed=21, stdr=1.5)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(1)
    alpha = np.array([0.8, 1.5, 2.5])
    r0 = 3.0
    x = rng.gamma(r0 * alpha, 1 / r0, size=(S, K))
    ch.set("Alpha", alpha)
    ch.set("loc", _loc_pack(1.0, np.full(K, 1 / K), np.full((S, K), 1 / K), x, np.zeros((S, K)), x.sum(1)))
    r = _run(ch, "update_locpriori", 4000, lambda: ch.get("loc")[0])
    g = np.linspace(1e-4, p.maxr, 40001)[1:]
    sl, sx = np.log(x).sum(0), x.sum(0)

    def law(drop=None):
        v = np.zeros_like(g)
        for k in range(K):
            a = alpha[k]
            v += S * (g * a * np.log(g) - lg(g * a))
            if drop != "log x":
                v += g * a * sl[k]
            if drop != "sum x":
                v -= g * sx[k]
        return v

    check_law(r, g, law(), [("without sum x", law("sum x")), ("without log x", law("log x"))], "r admix")


def test_locprior_r_noadmix_law(vps, orc):
    """no admixture: r targets prod_s Dir(Gamma_s | r Eta) on (0, maxr); Eta
    and Gamma held by zero proposal widths"""
    K, S = 3, 4
    s, d = _data(vps, orc, n_pop=S)
    p = vps.params_default(admix_val=2, locpriori_val=1, nadmburnin=0, g_seed=22, stdr=2.0, epseta=0.0,
                           epsgamma=0.0)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(2)
    eta = np.array([0.2, 0.3, 0.5])
    gam = rng.dirichlet(6.0 * eta, size=S)
    ch.set("loc", _loc_pack(1.0, eta, gam, np.ones((S, K)), np.zeros((S, K)), np.zeros(S)))
    r = _run(ch, "update_locpriori", 4000, lambda: ch.get("loc")[0])
    _, e2, g2, _, _, _ = _loc_unpack(ch.get("loc"), K, S)
    assert np.array_equal(e2, eta) and np.array_equal(g2, gam)   # held
    g = np.linspace(1e-4, p.maxr, 40001)[1:]
    sl = np.log(gam).sum(0)

    def law(drop=None, flip=False):
        v = S * lg(g) * (-1 if flip else 1)
        for k in range(K):
            v -= S * lg(g * eta[k])
            if drop != "log gamma":
                v += g * eta[k] * sl[k]
        return v

    check_law(r, g, law(), [("without log Gamma", law("log gamma")), ("lgamma(r) sign", law(flip=True))],
              "r no-admix")


def test_locprior_eta_law(vps, orc):
    """Eta (K = 2) targets prod_s Dir(Gamma_s | r Eta) on the simplex: pair
    moves of a uniform step; r held (stdr = 0), Gamma held (epsgamma = 0)"""
    K, S = 2, 5
    s, d = _data(vps, orc, n_pop=S)
    p = vps.params_default(admix_val=2, locpriori_val=1, nadmburnin=0, g_seed=23, stdr=0.0, epseta=0.15,
                           epsgamma=0.0)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(3)
    r0 = 4.0
    gam = rng.dirichlet(r0 * np.array([0.3, 0.7]), size=S)
    ch.set("loc", _loc_pack(r0, np.array([0.5, 0.5]), gam, np.ones((S, K)), np.zeros((S, K)), np.zeros(S)))
    e = _run(ch, "update_locpriori", 4000, lambda: ch.get("loc")[1])
    assert ch.get("loc")[0] == r0
    g = np.linspace(0, 1, 20001)[1:-1]
    l1, l2 = np.log(gam[:, 0]).sum(), np.log(gam[:, 1]).sum()

    def law(swap=False, drop=None):
        v = -S * (lg(r0 * g) + lg(r0 * (1 - g)))
        if drop != "log gamma":
            v += (r0 * g - 1) * (l2 if swap else l1) + (r0 * (1 - g) - 1) * (l1 if swap else l2)
        return v

    check_law(e, g, law(), [("Gamma columns swapped", law(swap=True)), ("without log Gamma", law(drop="log gamma"))],
              "eta")


def test_locprior_gamma_law(vps, orc):
    """Gamma_s (K = 2) targets Dir(r Eta + D_s), independently per location:
    each location's first component is Beta-distributed"""
    K, S = 2, 4
    s, d = _data(vps, orc, n_pop=S)
    p = vps.params_default(admix_val=2, locpriori_val=1, nadmburnin=0, g_seed=24, stdr=0.0, epseta=0.0,
                           epsgamma=0.25)
    ch = orc.Chain(d, p, K, 1)
    r0, eta = 3.0, np.array([0.4, 0.6])
    D = np.array([[7.0, 3.0], [2.0, 9.0], [5.0, 5.0], [0.0, 4.0]])
    ch.set("loc", _loc_pack(r0, eta, np.full((S, K), 0.5), np.ones((S, K)), D, np.zeros(S)))
    gs = _run(ch, "update_locpriori", 4000, lambda: _loc_unpack(ch.get("loc"), K, S)[2][:, 0].copy())
    g = np.linspace(0, 1, 20001)[1:-1]
    for sidx in range(S):
        a, b = r0 * eta[0] + D[sidx, 0], r0 * eta[1] + D[sidx, 1]
        law = stats.beta.logpdf(g, a, b)
        wrong = [("without D", stats.beta.logpdf(g, r0 * eta[0], r0 * eta[1])),
                 ("D transposed", stats.beta.logpdf(g, r0 * eta[0] + D[sidx, 1], r0 * eta[1] + D[sidx, 0]))]
        if D[sidx, 0] == D[sidx, 1]:
            wrong = wrong[:1]
        check_law(gs[:, sidx], g, law, wrong, f"gamma s={sidx}")


# ---------------------------------------------------------------------------
# UpdateAlpha, src/structure.cpp:1501-1594
# ---------------------------------------------------------------------------
def test_locprior_global_alpha_law(vps, orc):
    """location prior, global alpha_k: targets prod_s Gamma(AlphaLocal_sk |
    shape r alpha_k, rate r) on (0, maxalpha); the local alphas (moved in the
    same call) are reset before each call"""
    K, S = 2, 6
    s, d = _data(vps, orc, n_pop=S)
    p = vps.params_default(admix_val=1, locpriori_val=1, g_seed=25, stdalpha=0.6)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(5)
    r0, a_true = 2.5, np.array([1.2, 3.0])
    x = rng.gamma(r0 * a_true, 1 / r0, size=(S, K))
    ch.set("Q", rng.dirichlet(np.ones(K), size=d.N))
    ch.set("Alpha", a_true.copy())
    loc = _loc_pack(r0, np.full(K, 0.5), np.full((S, K), 0.5), x, np.zeros((S, K)), x.sum(1))
    a = _run(ch, "update_alpha", 4000, lambda: ch.get("Alpha").copy(), reset=lambda: ch.set("loc", loc))
    g = np.linspace(1e-4, p.maxalpha, 40001)[1:]
    for k in range(K):
        sl = np.log(x[:, k]).sum()
        law = S * (r0 * g * np.log(r0) - lg(r0 * g)) + r0 * g * sl
        wrong = [("without r log r", -S * lg(r0 * g) + r0 * g * sl), ("without log x", S * (r0 * g * np.log(r0) - lg(r0 * g)))]
        check_law(a[:, k], g, law, wrong, f"global alpha k={k}")


def test_locprior_local_alpha_law(vps, orc):
    """location prior, global and local alphas together (three locations,
    K = 2, Q and r held): the pair of steps targets
        p(alpha, a) ~ prod_{s,k} Gamma(a_sk | shape r alpha_k, rate r)
                      prod_s prod_{i in s} Dir(Q_i | a_s)
    on (0, maxalpha).  The marginals of a global and a local alpha are
    integrated on grids: with G[alpha, a] the Gamma density and L_s[a1, a2]
    location s's Dirichlet likelihood, h_s = G L_s G^T and p(alpha) ~ prod_s
    h_s.  The shared alpha pools the locations, so a local step without its
    Gamma term, or a Dirichlet likelihood without lgamma(sum a), moves the
    local marginal."""
    K, S = 2, 3
    s, d = _data(vps, orc, n_pop=S, n_ind=30)
    ind_pop = np.asarray(s.arrays()["ind_pop"])
    p = vps.params_default(admix_val=1, locpriori_val=1, g_seed=26, stdalpha=0.5, maxalpha=8.0)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(6)
    r0 = 3.0
    a_true = np.array([[1.0, 3.0], [2.0, 2.0], [3.0, 1.2]])
    Q = np.concatenate([rng.dirichlet(a_true[x], size=int((ind_pop == x).sum())) for x in range(S)])
    ch.set("Q", Q)
    ch.set("Alpha", np.array([2.0, 2.0]))
    ch.set("loc", _loc_pack(r0, np.full(K, 0.5), np.full((S, K), 0.5), a_true.copy(), np.zeros((S, K)),
                            a_true.sum(1)))
    draws = _run(ch, "update_alpha", 4000, lambda: np.concatenate([ch.get("Alpha"), ch.get("loc")[1 + K + S * K:1 + K + S * K + 2]]),
                 thin=20)
    ga = np.linspace(0.01, p.maxalpha, 321)      # local alphas
    gA = np.linspace(0.01, p.maxalpha, 321)      # global alphas
    logG = r0 * gA[:, None] * np.log(r0) - lg(r0 * gA[:, None]) + (r0 * gA[:, None] - 1) * np.log(ga[None, :]) - r0 * ga[None, :]
    G = np.exp(logG)
    A1, A2 = np.meshgrid(ga, ga, indexing="ij")

    def lik(x, with_sum=True):
        rows = ind_pop == x
        sl = np.log(Q[rows]).sum(0)
        ll = rows.sum() * ((lg(A1 + A2) if with_sum else 0) - lg(A1) - lg(A2)) + (A1 - 1) * sl[0] + (A2 - 1) * sl[1]
        return np.exp(ll - ll.max())

    def marginals(with_sum=True, local_prior=True):
        Ls = [lik(x, with_sum) for x in range(S)]
        hs = [G @ L @ G.T for L in Ls]
        pA = np.prod([h / h.max() for h in hs], axis=0)
        # location 0's first local alpha
        if local_prior:
            w = pA / hs[0]
            B = Ls[0] @ G.T                          # [a1, alpha2]
            C = w @ B.T                              # [alpha1, a1]
            ma = (G * C).sum(0)
        else:
            ma = Ls[0].sum(1)
        return np.log(np.maximum(pA.sum(1), 1e-300)), np.log(np.maximum(ma, 1e-300))

    law_A, law_a = marginals()
    _, wrong_sum = marginals(with_sum=False)
    _, wrong_prior = marginals(local_prior=False)
    check_law(draws[:, 0], gA, law_A, [], "global alpha_0 (joint)")
    check_law(draws[:, 2], ga, law_a, [("without lgamma(sum a)", wrong_sum), ("local step without its Gamma term", wrong_prior)],
              "local alpha s=0 k=0 (joint)")


def _dir_alpha_loglik(Q, a_full, k, g):
    """sum_i log Dir(Q_i | alpha) as a function of alpha_k = g, the others fixed"""
    N = Q.shape[0]
    rest = a_full.sum() - a_full[k]
    return N * (lg(rest + g) - lg(g)) + g * np.log(Q[:, k]).sum()


@pytest.mark.parametrize("uniform", [True, False])
def test_per_cluster_alpha_first_law(vps, orc, uniform):
    """per-cluster alpha, first cluster of the pass with the others reset
    before each call: alpha_0 targets prior(alpha_0) prod_i Dir(Q_i | alpha)
    (the pass's sum of alphas is exact for the first cluster)"""
    K = 3
    s, d = _data(vps, orc, n_ind=40)
    kw = dict(admix_val=1, g_seed=27, stdalpha=0.4, diffalpha_val=1)
    if not uniform:
        kw.update(uniformalpha_val=2, alphapriora=6.0, alphapriorb=0.15)   # prior mean 0.9, sd 0.37
    p = vps.params_default(**kw)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(7)
    a_true = np.array([1.4, 0.6, 2.2])
    Q = rng.dirichlet(a_true, size=d.N)
    ch.set("Q", Q)
    ch.set("Alpha", a_true.copy())
    held = a_true.copy()

    def reset():
        a = ch.get("Alpha")
        a[1:] = held[1:]
        ch.set("Alpha", a)

    a0 = _run(ch, "update_alpha", 4000, lambda: ch.get("Alpha")[0], reset=reset)
    top = p.maxalpha if uniform else 15.0
    g = np.linspace(1e-4, top, 40001)[1:]
    prior = np.zeros_like(g) if uniform else (p.alphapriora - 1) * np.log(g) - g / p.alphapriorb
    law = _dir_alpha_loglik(Q, held, 0, g) + prior
    N = d.N
    wrong = [("without lgamma(sum alpha)", -N * lg(g) + g * np.log(Q[:, 0]).sum() + prior),
             ("another cluster's log Q", _dir_alpha_loglik(Q[:, [1, 0, 2]], held, 0, g) + prior)]
    if not uniform:
        wrong.append(("without the prior", _dir_alpha_loglik(Q, held, 0, g)))
    check_law(a0, g, law, wrong, f"alpha_0 uniform={uniform}")


def test_per_cluster_alpha_later_law(vps, orc):
    """per-cluster alpha, second cluster: the first cluster's alpha is above
    maxalpha and never moves (uniform prior), the third is reset before each
    call, so alpha_1 targets prod_i Dir(Q_i | alpha) as a function of alpha_1"""
    K = 3
    s, d = _data(vps, orc, n_ind=60)
    p = vps.params_default(admix_val=1, g_seed=28, stdalpha=0.4, diffalpha_val=1)
    ch = orc.Chain(d, p, K, 1)
    rng = np.random.default_rng(8)
    held = np.array([p.maxalpha + 3.0, 1.1, 2.0])   # a proposal below maxalpha is 7.5 sd away
    Q = rng.dirichlet(held, size=d.N)
    ch.set("Q", Q)
    ch.set("Alpha", held.copy())

    def reset():
        a = ch.get("Alpha")
        assert a[0] == held[0]
        a[2] = held[2]
        ch.set("Alpha", a)

    a1 = _run(ch, "update_alpha", 4000, lambda: ch.get("Alpha")[1], reset=reset)
    g = np.linspace(1e-4, p.maxalpha, 40001)[1:]
    law = _dir_alpha_loglik(Q, held, 1, g)
    wrong = [("sum without alpha_0", _dir_alpha_loglik(Q, held * np.array([0, 1, 1]), 1, g))]
    check_law(a1, g, law, wrong, "alpha_1")


# ---------------------------------------------------------------------------
# UpdateLambda, src/structure.cpp:1598-1635 (enabled with inferlambda_val = 0)
# ---------------------------------------------------------------------------
def _lambda_loglik(P_rows, g):
    """sum over rows of log Dir(row | lambda 1), rows of varying length"""
    v = np.zeros_like(g)
    for row in P_rows:
        n = len(row)
        v += lg(n * g) - n * lg(g) + (g - 1) * np.log(row).sum()
    return v


def _lambda_loglik_biallelic_norm(P_rows, g):
    """the same with every locus's Dirichlet normaliser taken as biallelic
    (a misreading of the per-allele-count table kdis)"""
    v = np.zeros_like(g)
    for row in P_rows:
        v += lg(2 * g) - 2 * lg(g) + (g - 1) * np.log(row).sum()
    return v


def _freq_rows(P, afo, k2):
    return [P[afo[l]: afo[l] + k2[l]] for l in range(len(afo))]


@pytest.mark.parametrize("diff", [True, False])
def test_lambda_law(vps, orc, diff):
    """lambda targets prod Dir(P_kl | lambda 1) over its clusters' rows
    (per cluster with difflambda, over all clusters otherwise) on
    (0, maxlambda)"""
    K = 2
    s, d = _data(vps, orc, n_loc=30, frac_multi=0.3, min_alleles=3, max_alleles=5)
    p = vps.params_default(admix_val=1, g_seed=29, inferlambda_val=0, difflambda_val=1 if diff else 2,
                           stdlambda=0.4)
    ch = orc.Chain(d, p, K, 1)
    _, afo, _ = d.order()
    k2 = np.diff(np.append(afo, d.KT))
    rng = np.random.default_rng(9)
    lam_true = [0.9, 3.0] if diff else [1.8, 1.8]
    P = np.zeros((K, d.KT))
    for k in range(K):
        for l in range(len(afo)):
            P[k, afo[l]: afo[l] + k2[l]] = rng.dirichlet(np.full(k2[l], lam_true[k]))
    ch.set("Freq", P)
    ch.set("Lambda", np.array(lam_true))
    lam = _run(ch, "update_lambda", 4000, lambda: ch.get("Lambda").copy())
    g = np.linspace(1e-4, p.maxlambda, 40001)[1:]
    if diff:
        for k in range(K):
            rows = _freq_rows(P[k], afo, k2)
            wrong = [("biallelic normalisers", _lambda_loglik_biallelic_norm(rows, g)),
                     ("the other cluster", _lambda_loglik(_freq_rows(P[1 - k], afo, k2), g))]
            check_law(lam[:, k], g, _lambda_loglik(rows, g), wrong, f"lambda k={k}")
    else:
        assert np.all(lam[:, 0] == lam[:, 1])
        rows = _freq_rows(P[0], afo, k2) + _freq_rows(P[1], afo, k2)
        check_law(lam[:, 0], g, _lambda_loglik(rows, g), [("one cluster only", _lambda_loglik(rows[:len(afo)], g))],
                  "lambda shared")


def test_lambda_fmodel_law(vps, orc):
    """F model, shared lambda: targets prod_l Dir(A_l | lambda 1) over the
    ancestral frequencies"""
    K = 2
    s, d = _data(vps, orc, n_loc=30, frac_multi=0.3, min_alleles=3, max_alleles=5)
    p = vps.params_default(admix_val=1, f_val=1, g_seed=30, inferlambda_val=0, stdlambda=0.4)
    ch = orc.Chain(d, p, K, 1)
    _, afo, _ = d.order()
    k2 = np.diff(np.append(afo, d.KT))
    rng = np.random.default_rng(10)
    A = np.zeros(d.KT)
    for l in range(len(afo)):
        A[afo[l]: afo[l] + k2[l]] = rng.dirichlet(np.full(k2[l], 2.5))
    FA = np.zeros((3, d.KT))
    FA[0] = A
    ch.set("FreqA", FA)
    lam = _run(ch, "update_lambda", 4000, lambda: ch.get("Lambda")[0])
    g = np.linspace(1e-4, p.maxlambda, 40001)[1:]
    rows = _freq_rows(A, afo, k2)
    P = ch.get("Freq")
    check_law(lam, g, _lambda_loglik(rows, g), [("cluster frequencies instead", _lambda_loglik(_freq_rows(P[0], afo, k2), g))],
              "lambda F model")


# ---------------------------------------------------------------------------
# UpdateA (A1 Dirichlet proposal / A2 pair shift), UpdateF
# ---------------------------------------------------------------------------
def _fmodel_chain(vps, orc, seed, Ftrue, n_loc=12, singlef=2):
    K = len(Ftrue)
    s, d = _data(vps, orc, n_loc=n_loc, n_ind=40)
    p = vps.params_default(admix_val=1, f_val=1, g_seed=seed, singlef_val=singlef, stdf=0.05)
    ch = orc.Chain(d, p, K, 1)
    _, afo, _ = d.order()
    k2 = np.diff(np.append(afo, d.KT))
    rng = np.random.default_rng(seed)
    A = np.zeros(d.KT)
    P = np.zeros((K, d.KT))
    f = (1 - np.asarray(Ftrue)) / np.asarray(Ftrue)
    for l in range(len(afo)):
        a = rng.dirichlet(np.full(k2[l], 3.0))
        A[afo[l]: afo[l] + k2[l]] = a
        for k in range(K):
            P[k, afo[l]: afo[l] + k2[l]] = rng.dirichlet(f[k] * a)
    ch.set("Freq", P)
    FA = np.zeros((3, d.KT))
    FA[0] = A
    ch.set("FreqA", FA)
    ch.set("Fst", np.concatenate([Ftrue, f]))
    return ch, p, d, afo, k2, A, P, f


def test_update_a_law(vps, orc):
    """ancestral frequencies of a biallelic locus: A targets
    Dir(A | lambda) prod_k Dir(P_kl | f_k A) (A1's independence proposal and
    A2's pair shift, chosen per sweep by HashULong(seed + m) & 1)"""
    ch, p, d, afo, k2, A, P, f = _fmodel_chain(vps, orc, 31, [0.4, 0.3])
    lam = 3.5
    ch.set("Lambda", np.full(2, lam))
    bi = [l for l in range(len(afo)) if k2[l] == 2][:4]
    assert len(bi) == 4
    a = _run(ch, "update_a", 3000, lambda: ch.get("FreqA")[0][[afo[l] for l in bi]].copy(), thin=5)
    g = np.linspace(0, 1, 20001)[1:-1]
    for j, l in enumerate(bi):
        p1, p2 = P[:, afo[l]], P[:, afo[l] + 1]

        def law(drop=None, swap=False):
            v = (lam - 1) * (np.log(g) + np.log(1 - g)) if drop != "prior" else np.zeros_like(g)
            for k in range(len(f)):
                v += -lg(f[k] * g) - lg(f[k] * (1 - g))
                if drop != "log p":
                    q1, q2 = (p2[k], p1[k]) if swap else (p1[k], p2[k])
                    v += f[k] * g * np.log(q1) + f[k] * (1 - g) * np.log(q2)
            return v

        check_law(a[:, j], g, law(), [("without log P", law(drop="log p")), ("alleles swapped", law(swap=True)),
                                      ("without the Dirichlet(lambda) prior", law(drop="prior"))], f"A locus {l}")


@pytest.mark.parametrize("singlef", [2, 1])
def test_update_f_law(vps, orc, singlef):
    """F_k targets Gamma(F | pmeanf^2/pstdf^2, pmeanf/pstdf^2) prod_l
    Dir(P_kl | f A_l), f = (1 - F)/F, on (0, 1); one F over all clusters with
    a single F"""
    Ft = [0.12, 0.12] if singlef == 1 else [0.2, 0.06]
    ch, p, d, afo, k2, A, P, f = _fmodel_chain(vps, orc, 32 + singlef, Ft, n_loc=20, singlef=singlef)
    F = _run(ch, "update_f", 4000, lambda: ch.get("Fst")[0].copy())
    g = np.linspace(0, 1, 20001)[1:-1]
    fg = (1 - g) / g
    shape, rate = p.pmeanf ** 2 / p.pstdf ** 2, p.pmeanf / p.pstdf ** 2
    rows = range(len(afo))

    def cluster_ll(k, drop=None):
        v = np.zeros_like(g)
        for l in rows:
            a = A[afo[l]: afo[l] + k2[l]]
            pk = P[k, afo[l]: afo[l] + k2[l]]
            v += lg(fg) - sum(lg(fg * a[x]) for x in range(k2[l]))
            if drop != "log p":
                v += fg * (a * np.log(pk)).sum()
        return v

    prior = (shape - 1) * np.log(g) - rate * g
    if singlef == 1:
        assert np.all(F[:, 0] == F[:, 1])
        law = prior + cluster_ll(0) + cluster_ll(1)
        check_law(F[:, 0], g, law, [("one cluster only", prior + cluster_ll(0)),
                                    ("without log P", prior + cluster_ll(0, "log p") + cluster_ll(1, "log p"))], "single F")
    else:
        for k in range(2):
            check_law(F[:, k], g, prior + cluster_ll(k), [("the other cluster", prior + cluster_ll(1 - k)),
                                                          ("without log P", prior + cluster_ll(k, "log p"))], f"F k={k}")


# ---------------------------------------------------------------------------
# RNG<float> (src/statistics.cpp:325-393): xorshift96 and its uniform
# ---------------------------------------------------------------------------
def _gf2_matpow_is_identity(T, e):
    """T^e == I over GF(2) for a 0/1 integer matrix"""
    n = T.shape[0]
    R = np.eye(n, dtype=np.int64)
    B = T.copy()
    while e:
        if e & 1:
            R = (R @ B) & 1
        B = (B @ B) & 1
        e >>= 1
    return np.array_equal(R, np.eye(n, dtype=np.int64))


def _xs96_matrix(orc):
    """the 96x96 GF(2) matrix of one RngF step, column j = image of basis bit j"""
    T = np.zeros((96, 96), dtype=np.int64)
    for j in range(96):
        r = orc.RngF(0, 0)
        r.x, r.y, r.z = [(1 << (j - 32 * w)) if j // 32 == w else 0 for w in range(3)]
        r.xorshift()
        state = r.x | (r.y << 32) | (r.z << 64)
        T[:, j] = [(state >> i) & 1 for i in range(96)]
    return T


def test_rngf_full_period(orc):
    """xorshift96 with shifts (16, 5, 1) is linear over GF(2) and has period
    2^96 - 1: T^(2^96-1) = I and T^((2^96-1)/q) != I for each prime q"""
    T = _xs96_matrix(orc)
    # linearity: the step of an XOR of states is the XOR of the steps
    rng = np.random.default_rng(11)
    for _ in range(20):
        wa, wb = rng.integers(0, 2 ** 32, 3), rng.integers(0, 2 ** 32, 3)
        ra, rb, rab = orc.RngF(0, 0), orc.RngF(0, 0), orc.RngF(0, 0)
        ra.x, ra.y, ra.z = (int(v) for v in wa)
        rb.x, rb.y, rb.z = (int(v) for v in wb)
        rab.x, rab.y, rab.z = (int(v) for v in wa ^ wb)
        for r in (ra, rb, rab):
            r.xorshift()
        assert (rab.x, rab.y, rab.z) == (ra.x ^ rb.x, ra.y ^ rb.y, ra.z ^ rb.z)
    n = 2 ** 96 - 1
    primes = [3, 5, 7, 13, 17, 97, 193, 241, 257, 673, 65537, 22253377]
    prod = 1
    for q in primes:
        prod *= q
    assert prod * 3 == n                      # 2^96 - 1 = 3^2 5 7 13 17 97 193 241 257 673 65537 22253377
    assert _gf2_matpow_is_identity(T, n)
    for q in primes:
        assert not _gf2_matpow_is_identity(T, n // q), q


def test_rngf_seeding_and_uniform(orc):
    """seeding: s = MurmurHash32(seed, salt), x = 0x075BCD15 ^ s,
    y = 0x159A55E5 ^ (s << 3), z = 0x1F123BB5 ^ (s << 6) (32-bit); Uniform
    puts the output's low 23 bits in a float's mantissa in [1, 2) and
    subtracts 1; the outputs are uniform"""
    import ctypes
    lib = orc.load()
    for seed, salt in ((0, 0), (12345, 0x9B83516C63E249C3), (2 ** 40 + 7, 0xDA310D821043E1D3)):
        r = orc.RngF(seed, salt)
        s = lib.or_murmur32(seed, salt)
        assert (r.x, r.y, r.z) == (0x075BCD15 ^ s, (0x159A55E5 ^ (s << 3)) & 0xFFFFFFFF,
                                   (0x1F123BB5 ^ (s << 6)) & 0xFFFFFFFF)
        twin = orc.RngF(seed, salt)
        for _ in range(200):
            u = lib.or_rngf_uniform(ctypes.byref(r), 2.0, 5.0)
            w = twin.xorshift()
            bits = np.array([0x3F800000 | (w & 0x7FFFFF)], dtype=np.uint32).view(np.float32)[0]
            want = np.float32((np.float32(bits) - np.float32(1.0)) * np.float32(3.0) + np.float32(2.0))
            assert np.float32(u) == want
    r = orc.RngF(99, 0x9B83516C63E249C3)
    x = np.array([lib.or_rngf_uniform(ctypes.byref(r), 0.0, 1.0) for _ in range(60000)])
    counts, _ = np.histogram(x, bins=50, range=(0, 1))
    assert stats.chisquare(counts).pvalue > 1e-4
    assert x.min() >= 0 and x.max() < 1
