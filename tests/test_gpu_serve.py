"""Serving (vps_server): jobs of uneven N, L, K, ploidy and model run
concurrently under an HBM budget; each job's results are bit-identical to
the same job run alone; admission respects the budget."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


MODELS = {
    "admix": dict(admix_val=1),
    "noadmix": dict(admix_val=2, nadmburnin=6),
    "fmodel": dict(admix_val=1, f_val=1),
    "locprior": dict(admix_val=1, locpriori_val=1),
}


def job_specs():
    rng = np.random.default_rng(5)
    out = []
    for j in range(9):
        model = list(MODELS)[j % 4]
        ploidy = [(2, 1.0)] if j % 3 else [(2, 0.5), (4, 0.5)]
        synth = dict(n_ind=int(rng.integers(40, 300)), n_loc=int(rng.integers(200, 2500)), k_true=3,
                     alpha_q=0.5, fst=0.05, missing=0.02, ploidy=ploidy, seed=100 + j,
                     n_pop=4 if model == "locprior" else 0)
        par = dict(MODELS[model], nburnin=int(rng.integers(8, 20)), nreps=int(rng.integers(15, 40)),
                   nthinning=3, writelnl_val=1, g_seed=3, z_chunk=-1)   # the server's chunking
        out.append((synth, par, int(rng.integers(1, 7)), j + 1))
    return out


def run_alone(vps, s, p, K, run_id, sweeps):
    ds = vps.Dataset(s, 0)
    ch = vps.Chain(ds, p, K, run_id)
    ch.step(sweeps)
    ch.finish()
    out = {n: ch.get(n) for n in ("Q", "MiSum", "rout", "lnltrace", "lnL")}
    ch.close()
    ds.close()
    return out


def test_served_results_equal_alone(vps):
    specs = job_specs()
    srv = vps.Server(hbm_budget_bytes=8 << 30, max_active=3, slice_sweeps=5)
    jobs = []
    for synth, par, K, run_id in specs:
        s = vps.Synth(**synth)
        p = vps.params_default(**par)
        jid = srv.submit(s, p, K, run_id)
        jobs.append((jid, s, p, K, run_id))
    say("submitted", len(jobs))
    for jid, s, p, K, run_id in jobs:
        srv.wait(jid, 120.0)
        inf = srv.info(jid)
        assert inf["state"] == "done"
        alone = run_alone(vps, s, p, K, run_id, inf["sweeps_total"])
        for n in ("Q", "MiSum", "rout", "lnltrace", "lnL"):
            got = srv.result(jid, n)
            assert got.shape == alone[n].reshape(got.shape).shape and (got == alone[n].reshape(got.shape)).all(), (jid, n)
        assert inf["t_submit"] <= inf["t_start"] <= inf["t_end"]
        say("job", jid, "N", s.desc.n_ind, "K", K, "ok")
    st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0 and st["reserved_bytes"] == 0
    srv.close()


def test_budget_admission(vps):
    s = vps.Synth(n_ind=200, n_loc=3000, k_true=3, seed=9)
    p = vps.params_default(admix_val=1, nburnin=5, nreps=30)
    probe = vps.Server(hbm_budget_bytes=8 << 30, max_active=1)
    j0 = probe.submit(s, p, 4, 1)
    probe.wait(j0, 60.0)
    need = probe.info(j0)["device_bytes"]
    probe.close()
    assert need > 0
    # room for two such jobs at a time
    budget = int(2.5 * need)
    srv = vps.Server(hbm_budget_bytes=budget, max_active=0, slice_sweeps=4)
    ids = [srv.submit(s, p, 4, r) for r in range(1, 7)]
    for j in ids:
        srv.wait(j, 120.0)
    st = srv.stats()
    assert st["completed"] == 6
    assert st["peak_reserved_bytes"] <= budget, st
    # a job that can never fit is refused at submission
    with pytest.raises(vps.VpsError):
        srv.submit(vps.Synth(n_ind=4000, n_loc=20000, k_true=3, seed=2), p, 10, 1)
    srv.close()


def test_destroy_with_running_jobs(vps):
    srv = vps.Server(hbm_budget_bytes=8 << 30, max_active=2, slice_sweeps=2)
    s = vps.Synth(n_ind=300, n_loc=4000, k_true=3, seed=10)
    p = vps.params_default(admix_val=1, nburnin=100, nreps=5000)
    for r in range(1, 5):
        srv.submit(s, p, 3, r)
    srv.close()   # abandons the running and queued jobs without hanging


def test_submit_checks_run_length(vps):
    """a job's sweep count is its run length: nburnin + nreps (+ nadmburnin
    for no-admixture without a location prior); another count is refused"""
    s = vps.Synth(n_ind=50, n_loc=300, k_true=2, seed=11)
    p = vps.params_default(admix_val=1, nburnin=4, nreps=6)
    srv = vps.Server(hbm_budget_bytes=2 << 30, max_active=1)
    for bad in (4, 9, 11):
        with pytest.raises(vps.VpsError):
            srv.submit(s, p, 2, 1, sweeps=bad)
    j = srv.submit(s, p, 2, 1, sweeps=0)        # 0 takes the run length
    srv.wait(j, 60.0)
    assert srv.info(j)["sweeps_total"] == 10
    q = vps.params_default(admix_val=2, nburnin=3, nreps=5, nadmburnin=2)
    with pytest.raises(vps.VpsError):
        srv.submit(s, q, 2, 2, sweeps=8)
    j = srv.submit(s, q, 2, 2, sweeps=10)
    srv.wait(j, 60.0)
    assert srv.result(j, "rout").size > 0
    srv.close()
