"""REAL = float on the device (BAYESIAN<float>: precision REAL32 for
-g_float=single -g_fastsingle=no, FAST32 for vcfpop's default
-g_fastsingle=yes) against the oracle's single-precision path.

Bar as for FP64: Z, Ni, Mi and MiSum bit for bit; P, Q, FreqM, the
parameters and the run summary to 1e-11 relative, with the oracle's state fed
before each sweep.  The device's P, Q and FreqM are single-precision values.
FAST32's Z pass draws from RNG<float> (xorshift96) streams split over chunks
by GF(2) jumps: its counts equal the oracle's at every chunking.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DATA = {
    "bi_diploid": dict(n_ind=96, n_loc=700, k_true=3, alpha_q=0.5, fst=0.05, missing=0.02, seed=21),
    "mixed": dict(n_ind=75, n_loc=500, k_true=4, alpha_q=0.5, fst=0.05, missing=0.03, frac_multi=0.25,
                  min_alleles=3, max_alleles=8, multi_dir=0.3, ploidy=[(2, 0.6), (4, 0.3), (6, 0.1)], n_pop=4,
                  seed=22),
}

MODELS = {
    "admix": dict(admix_val=1, metrofreq=3),
    "admix_parallel": dict(admix_val=1, metrofreq=3, rng_topology=1),
    "noadmix": dict(admix_val=2, nadmburnin=2),
    "locprior_admix": dict(admix_val=1, locpriori_val=1, metrofreq=4),
    "locprior_noadmix": dict(admix_val=2, locpriori_val=1),
}

FED = ["Freq", "Q", "Z", "Ni", "Mi", "Alpha", "Lambda", "loc", "Fst", "FreqA", "binaryq", "MiSum", "FreqM", "rout", "m"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def close(a, b, rtol):
    return np.allclose(np.asarray(a, np.float64), np.asarray(b, np.float64), rtol=rtol, atol=0)


def single(a):
    a = np.asarray(a, np.float64)
    return bool((a.astype(np.float32).astype(np.float64) == a).all())


@pytest.fixture(scope="module")
def pairs(vps, orc):
    orc.set_threads(8)
    out = {}
    for name, spec in DATA.items():
        s = vps.Synth(**spec)
        out[name] = (s, vps.Dataset(s, 0), orc.Data(s.desc))
    return out


@pytest.mark.parametrize("prec", ["REAL32", "FAST32"])
@pytest.mark.parametrize("model", list(MODELS))
@pytest.mark.parametrize("name", list(DATA))
def test_sweeps_track_oracle(vps, orc, pairs, prec, model, name):
    say("REAL = float sweeps", prec, model, name)
    s, ds, od = pairs[name]
    kw = dict(nburnin=3, nreps=8, nthinning=2, writelnl_val=1, g_seed=12, precision=getattr(vps, "PREC_" + prec))
    kw.update(MODELS[model])
    p = vps.params_default(**kw)
    g, o = vps.Chain(ds, p, 3, 2), orc.Chain(od, p, 3, 2)
    assert single(g.get("Freq")) and single(g.get("Q"))
    for sweep in range(11):
        for n in FED:
            g.set(n, o.get(n))
        g.step(1)
        o.sweep(1)
        for n in ("Z", "Ni", "Mi", "MiSum"):
            assert (g.get(n) == o.get(n)).all(), (sweep, n)
        for n in ("Freq", "Q", "Alpha", "Lambda", "loc", "FreqM", "rout"):
            assert close(g.get(n), o.get(n), 1e-11), (sweep, n)
        assert single(g.get("Freq")) and single(g.get("Q")) and single(g.get("FreqM")), sweep
    g.finish()
    o.finish()
    assert close(g.get("MiSum"), o.get("MiSum"), 1e-12)
    assert close(g.get("rout"), o.get("rout"), 1e-11)
    assert close(g.get("FreqM"), o.get("FreqM"), 1e-11) and single(g.get("FreqM"))
    g.close()
    o.close()


@pytest.mark.parametrize("z_chunk", [1, 7, 64, 512])
def test_fast32_z_pass_chunking(vps, orc, pairs, z_chunk):
    """the xorshift96 streams jumped to each chunk's first individual"""
    s, ds, od = pairs["mixed"]
    p = vps.params_default(admix_val=1, g_seed=5, precision=vps.PREC_FAST32, z_chunk=z_chunk)
    g, o = vps.Chain(ds, p, 4, 3), orc.Chain(od, p, 4, 3)
    rng = np.random.default_rng(z_chunk)
    for m in (0, 1, 9):
        Freq = o.get("Freq")
        Q = rng.dirichlet(np.ones(4), size=od.N).astype(np.float32).astype(np.float64)
        for c in (g, o):
            c.set("Q", Q)
            c.set("Freq", Freq)
            c.set("m", [m])
        g.update_z_admix()
        o.call("update_z_admix")
        for n in ("Ni", "Mi"):
            assert (g.get(n) == o.get(n)).all(), (z_chunk, m, n)
    g.close()
    o.close()


@pytest.mark.parametrize("prec", ["REAL32", "FAST32"])
def test_whole_run_equals_oracle(vps, orc, pairs, prec):
    """a short run from the initial state, nothing fed"""
    s, ds, od = pairs["bi_diploid"]
    p = vps.params_default(admix_val=1, nburnin=10, nreps=20, nthinning=2, metrofreq=5, g_seed=31,
                           precision=getattr(vps, "PREC_" + prec))
    g, o = vps.Chain(ds, p, 3, 1), orc.Chain(od, p, 3, 1)
    g.step(30)
    o.sweep(30)
    for n in ("Z", "Ni", "Mi"):
        assert (g.get(n) == o.get(n)).all(), n
    assert close(g.get("Q"), o.get("Q"), 1e-11) and close(g.get("Freq"), o.get("Freq"), 1e-11)
    g.finish()
    o.finish()
    assert (g.get("MiSum") == o.get("MiSum")).all()
    # variance and lnP(D) are differences of the lnL moments: their error
    # scales with E[lnL^2]
    ro, rg = o.get("rout"), g.get("rout")
    tol = 1e-11 * np.abs(ro)
    tol[2] = tol[3] = 1e-11 * abs(ro[1])
    assert (np.abs(rg - ro) <= tol).all(), (rg, ro)
    g.close()
    o.close()


def test_f_model_needs_double(vps, pairs):
    s, ds, od = pairs["bi_diploid"]
    with pytest.raises(vps.VpsError, match="F model"):
        vps.Chain(ds, vps.params_default(admix_val=1, f_val=1, precision=vps.PREC_REAL32), 2, 1)
