"""Whole chains against the CPU oracle, free-running (no state fed between
sweeps): the plain admixture workload (N=200, L=2,000 biallelic diploid,
K=3, 1,000 burn-in + 2,000 sampling sweeps, FP64) across 10 seeds.

Bar: the cluster assignments and count tables stay bit-identical through
every sweep (checked every 250 sweeps), the reported ancestry (normalised
MiSum) is identical, and Q, mean lnL and the run accumulators agree to
1e-12 relative."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def config0(vps, orc):
    orc.set_threads(16)
    s = vps.make_synth("admix_n200_l2k_k3")
    return s, vps.Dataset(s, 0), orc.Data(s.desc)


@pytest.mark.parametrize("seed,topology", [(s, 0) for s in range(1, 11)] + [(1, 1), (2, 1)])
def test_full_chain_matches_oracle(vps, orc, config0, seed, topology):
    s, ds, od = config0
    say("full chain seed", seed, "topology", topology)
    p = vps.params_default(admix_val=1, nburnin=1000, nreps=2000, g_seed=seed, rng_topology=topology)
    g, o = vps.Chain(ds, p, 3, 1), orc.Chain(od, p, 3, 1)
    for blk in range(12):
        g.step(250)
        o.sweep(250)
        for n in ("Z", "Mi", "Ni"):
            assert (g.get(n) == o.get(n)).all(), (n, (blk + 1) * 250)
    g.finish()
    o.finish()
    assert (g.get("MiSum") == o.get("MiSum")).all()
    qg, qo = g.get("Q"), o.get("Q")
    assert np.abs(qg - qo).max() <= 1e-12
    rg, ro = g.get("rout"), o.get("rout")
    # rout = [mean lnL, mean lnL^2, var = mean lnL^2 - mean^2, lnP(D) = mean - var/2, ...]
    # (Arrange, src/structure.cpp:1862-1869).  The variance cancels two ~2e11
    # terms, so its error bound is relative to mean lnL^2, not to itself.
    tol = 1e-12 * np.abs(ro)
    tol[2] = tol[3] = 1e-12 * abs(ro[1])
    assert (np.abs(rg - ro) <= tol).all(), (rg, ro)
    assert (g.get("Alpha") == o.get("Alpha")).all() or np.allclose(g.get("Alpha"), o.get("Alpha"), rtol=1e-12)


def test_tables_match_oracle_run(vps, orc, config0, tmp_path):
    """the result tables of a device run equal those written from the oracle
    run's state (same seed, same options)"""
    s, ds, od = config0
    p = vps.params_default(admix_val=1, nburnin=200, nreps=400, writelnl_val=1, g_seed=3)
    g, o = vps.Chain(ds, p, 3, 1), orc.Chain(od, p, 3, 1)
    g.step(600)
    o.sweep(600)
    g.finish()
    o.finish()
    g.write_tables(str(tmp_path / "gpu"), diversity=1)
    st = dict(Q=o.get("Q"), ancestry=o.get("MiSum"), FreqM=o.get("FreqM"), rout=o.get("rout"),
              lnltrace=o.get("lnltrace"), seed=o.seed())
    vps.write_tables_host(ds, p, 3, 1, st, str(tmp_path / "cpu"), diversity=1)
    # the variance of lnL (and ln P(D) through it) cancels two ~2e11 terms:
    # those two lines agree to 1e-12 of mean lnL^2; every other line is identical
    scale = 1e-12 * abs(st["rout"][1])
    for suffix in (".structure.k=3_rep=1_id=1.txt", ".structure.k=3_rep=1_id=1.lnl.txt"):
        a = open(str(tmp_path / "gpu") + suffix, newline="").read().split("\r\n")
        b = open(str(tmp_path / "cpu") + suffix, newline="").read().split("\r\n")
        assert len(a) == len(b), suffix
        for x, y in zip(a, b):
            if x.startswith(("Variance of ln likelihood=", "Estimated Ln Prob of Data=")):
                assert abs(float(x.split("=")[1]) - float(y.split("=")[1])) <= max(scale, 1e-5), (x, y)
            else:
                assert x == y, (suffix, x, y)


def _ancestry(vps, ds, seed, prec):
    p = vps.params_default(admix_val=1, nburnin=1000, nreps=2000, g_seed=seed, rng_topology=1, precision=prec)
    c = vps.Chain(ds, p, 3, 1)
    c.step(3000)
    c.finish()
    ms = c.get("MiSum")
    out = ms / ms.sum(axis=1, keepdims=True), float(c.get("rout")[0])
    c.close()
    return out


def _aligned(q, ref):
    """q with its cluster columns permuted to match ref best (chains of
    different seeds may settle on different cluster labels)"""
    import itertools
    perms = list(itertools.permutations(range(q.shape[1])))
    best = min(perms, key=lambda pm: np.abs(q[:, list(pm)] - ref).mean())
    return q[:, list(best)]


def test_mixed_posterior_within_monte_carlo_error(vps, config0):
    """FP32 weights: the chain leaves the FP64 trajectory at its first
    differing draw, so its posterior-mean ancestry is another Monte Carlo
    estimate of the same quantity.  Its distance from the FP64 estimate of the
    same seed stays within the spread between FP64 estimates of two seeds,
    and mean lnL agrees to that spread as well."""
    s, ds, od = config0
    d_mix, d_seed, l_mix, l_seed = [], [], [], []
    for seed in (1, 2, 3, 4):
        say("posterior ancestry, seed", seed)
        q64, l64 = _ancestry(vps, ds, seed, vps.PREC_FP64)
        q32, l32 = _ancestry(vps, ds, seed, vps.PREC_MIXED)
        q64b, l64b = _ancestry(vps, ds, seed + 1000, vps.PREC_FP64)
        # once off the FP64 trajectory the FP32 chain may settle on other labels too
        d_mix.append(np.abs(_aligned(q32, q64) - q64).mean())
        d_seed.append(np.abs(_aligned(q64b, q64) - q64).mean())
        l_mix.append(abs(l32 - l64))
        l_seed.append(abs(l64b - l64))
    say("mean |dQ| mixed", d_mix, "other seed", d_seed, "|d mean lnL| mixed", l_mix, "other seed", l_seed)
    assert np.mean(d_mix) <= np.mean(d_seed)
    assert max(d_mix) < 0.01
    assert np.mean(l_mix) <= 3 * np.mean(l_seed) + 1.0


def test_finish_before_sampling_is_refused(vps):
    """Arrange divides by the recorded sampling sweeps: finishing a chain
    still in burn-in is a state error, and the chain can go on and finish"""
    s = vps.Synth(n_ind=40, n_loc=200, k_true=2, seed=3)
    ds = vps.Dataset(s, 0)
    c = vps.Chain(ds, vps.params_default(admix_val=1, nburnin=5, nreps=4, g_seed=1), 2, 1)
    c.step(5)
    with pytest.raises(vps.VpsError):
        c.finish()
    c.step(4)
    c.finish()
    r = c.get("rout")
    assert np.isfinite(r).all()
    c.close()
    ds.close()
