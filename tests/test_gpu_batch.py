"""Batched chains (vps_batch): every chain of a batch ends bit-identical to
the same chain stepped alone, across graph replays with eager work between
them, for mixed models, K values and schedules."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("Freq", "Q", "Z", "Ni", "Mi", "Alpha", "MiSum", "FreqM", "rout", "lnltrace", "m")

MODELS = {
    "admix_metro": dict(admix_val=1),
    "admix_diffalpha": dict(admix_val=1, diffalpha_val=1),
    "noadmix_burnin": dict(admix_val=2, nadmburnin=5),
    "locprior_admix": dict(admix_val=1, locpriori_val=1),
    "locprior_noadmix": dict(admix_val=2, locpriori_val=1),
    "fmodel": dict(admix_val=1, f_val=1),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def data(vps):
    s = vps.Synth(n_ind=75, n_loc=500, k_true=3, alpha_q=0.5, fst=0.05, missing=0.03, frac_multi=0.25,
                  min_alleles=3, max_alleles=6, multi_dir=0.3, ploidy=[(2, 0.7), (4, 0.3)], n_pop=4, seed=41)
    return s, vps.Dataset(s, 0)


def make_chains(vps, ds, specs):
    out = []
    for run_id, (model, K, extra) in enumerate(specs, start=1):
        kw = dict(MODELS[model])
        # z_chunk=-1: the modelled chunking, the same for every copy of a
        # chain (the Record lnL sums in chunk order; timed trials may pick
        # different chunkings for two copies)
        kw.update(nburnin=6, nreps=14, nthinning=3, metrofreq=4, writelnl_val=1, g_seed=17, z_chunk=-1)
        kw.update(extra)
        out.append(vps.Chain(ds, vps.params_default(**kw), K, run_id))
    return out


def assert_same(a, b, tag):
    for n in STATE:
        x, y = a.get(n), b.get(n)
        assert x.shape == y.shape and (x == y).all(), (tag, n)


SPECS = [("admix_metro", 2, {}), ("admix_metro", 3, {}), ("admix_diffalpha", 4, {}),
         ("noadmix_burnin", 3, {}), ("locprior_admix", 3, {}), ("locprior_noadmix", 2, {}),
         ("fmodel", 3, {}), ("admix_metro", 5, dict(rng_topology=1)), ("admix_metro", 3, dict(precision=1))]


def test_batch_equals_chains_alone(vps, data):
    s, ds = data
    say("batch vs alone:", [x[0] for x in SPECS])
    alone = make_chains(vps, ds, SPECS)
    batched = make_chains(vps, ds, SPECS)
    b = vps.Batch(batched)
    done = 0
    for n in (5, 1, 7, 7):   # graph replays with host reads between them
        b.step(n)
        b.sync()
        for c in alone:
            c.step(n)
            c.sync()
        done += n
        for i, (a, c) in enumerate(zip(alone, batched)):
            assert_same(a, c, (SPECS[i], done))
    st = b.stats()
    assert st["graph_launches"] + st["eager_sweeps"] == done
    # one capture per distinct vector of sweep plans; these schedules repeat
    assert 0 < st["captures"] < st["graph_launches"], st
    b.close()
    for a, c in zip(alone, batched):
        a.finish()
        c.finish()
        assert (a.get("MiSum") == c.get("MiSum")).all()
        assert (a.get("rout") == c.get("rout")).all()


def test_batch_mixed_schedules(vps, data):
    """chains whose schedules differ (thinning, Metropolis period) still
    match; each distinct plan vector is captured once"""
    s, ds = data
    specs = [("admix_metro", 2, dict(nthinning=2)), ("admix_metro", 3, dict(metrofreq=3)),
             ("noadmix_burnin", 2, dict(nadmburnin=2))]
    alone = make_chains(vps, ds, specs)
    batched = make_chains(vps, ds, specs)
    b = vps.Batch(batched)
    b.step(20)
    b.sync()
    for c in alone:
        c.step(20)
    for a, c in zip(alone, batched):
        a.sync()
        assert_same(a, c, "mixed")
    b.close()


def test_batch_then_chain_calls(vps, data):
    """plain chain steps after batch steps see the batch's work"""
    s, ds = data
    specs = [("admix_metro", 3, {}), ("admix_metro", 2, {})]
    alone = make_chains(vps, ds, specs)
    batched = make_chains(vps, ds, specs)
    b = vps.Batch(batched)
    b.step(4)
    for c in batched:
        c.step(3)    # queued behind the batch on the chain's stream
    b.step(2)
    for c in alone:
        c.step(9)
    for a, c in zip(alone, batched):
        a.sync()
        c.sync()
        assert_same(a, c, "interleaved")
    b.close()


@pytest.mark.parametrize("model", list(MODELS))
def test_sweep_graphs_equal_eager(vps, data, model):
    """vps_chain_step with cached per-plan sweep graphs gives the same chain
    as kernel-by-kernel launches"""
    s, ds = data
    g = make_chains(vps, ds, [(model, 3, dict(sweep_graphs=1))])[0]
    e = make_chains(vps, ds, [(model, 3, dict(sweep_graphs=0))])[0]
    for n in (4, 9, 7):
        g.step(n)
        e.step(n)
        g.sync()
        e.sync()
        assert_same(g, e, (model, n))
