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
