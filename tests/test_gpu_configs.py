"""The benchmark configurations at their own sizes, against the CPU oracle.

* 64 chains (K = 2..9 x 8 replicates, N = 1,000, L = 50,000) as one batch in
  the reference's draw topology: every chain equals itself stepped alone, and
  the K = 2, 5 and 9 chains equal the oracle's (Z, Ni, Mi bit for bit; P and
  Q to 1e-12).
* One full-size Z pass of the mixed-ploidy multiallelic location-prior
  workload (N = 3,000, L = 200,000, ploidy 2/4/6, K = 6), with the chunking
  that takes the 8-wavefront multiallelic kernel, against the oracle's.
* A served mix with a job at the job-stream distribution's extremes
  (N = 5,000, L = 200,000, ploidy 4, K = 10) and jobs of each model
  (admixture, no admixture, F model): every job equals itself run alone, and
  one job per model equals the oracle's run (ancestry bit for bit; Q and the
  run summary to 1e-12).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.allclose(a, b, rtol=rtol, atol=0)


def test_batch_64_chains_equal_alone_and_oracle(vps, orc):
    say("64-chain batch: N=1000 L=50000 K=2..9 x 8, reference topology")
    s = vps.make_synth("batch_n1000_l50k_k2to9x8")
    ds = vps.Dataset(s, 0)
    od = orc.Data(s.desc)
    specs = [(K, rep, (K - 2) * 8 + rep) for K in range(2, 10) for rep in range(1, 9)]   # run id: K outer, replicate inner

    def params():
        return vps.params_default(admix_val=1, nburnin=0, nreps=5000, g_seed=1, rng_topology=vps.RNG_REFERENCE,
                                  z_chunk=-1)

    batched = [vps.Chain(ds, params(), K, rid) for K, _, rid in specs]
    b = vps.Batch(batched)
    S = 5
    b.step(S)
    b.sync()
    st = b.stats()
    assert st["graph_launches"] + st["eager_sweeps"] == S
    for (K, rep, rid), c in zip(specs, batched):
        a = vps.Chain(ds, params(), K, rid)
        a.step(S)
        a.sync()
        for n in ("Z", "Ni", "Mi", "Q", "Freq"):
            assert (a.get(n) == c.get(n)).all(), (K, rep, n)
        a.close()
    say("batch equals alone for all 64 chains")
    for K in (2, 5, 9):
        rid = (K - 2) * 8 + 1
        c = batched[rid - 1]
        o = orc.Chain(od, params(), K, rid)
        o.sweep(S)
        for n in ("Z", "Ni", "Mi"):
            assert (o.get(n) == c.get(n)).all(), (K, n)
        assert close(c.get("Q"), o.get("Q")) and close(c.get("Freq"), o.get("Freq")), K
        say("K", K, "equals the oracle after", S, "sweeps")
        o.close()
    b.close()
    for c in batched:
        c.close()
    ds.close()


def test_mixed_ploidy_full_size_zpass_gen8(vps, orc):
    say("Z pass: N=3000 L=200000 ploidy 2/4/6 20% multiallelic K=6 locprior, FP64")
    s = vps.make_synth("locprior_n3000_l200k_k6_mixed")
    ds = vps.Dataset(s, 0)
    od = orc.Data(s.desc)
    K = 6
    # 12 chunks of 250 individuals: the 8-wavefront GEN block is taken when
    # G >= 8 and LDS bounds the resident blocks, as at this shape
    p = vps.params_default(admix_val=1, locpriori_val=1, g_seed=3, z_chunk=250)
    g = vps.Chain(ds, p, K, 1)
    assert g.chunking() == (12, 250)
    assert g.gen_waves() == 8, g.gen_waves()
    o = orc.Chain(od, p, K, 1)
    o.set("m", [0])
    o.sweep(1)                        # a realistic P and Q
    for n in ("Freq", "Q", "m"):
        g.set(n, o.get(n))
    g.update_z_admix()
    o.call("update_z_admix")
    for n in ("Ni", "Mi"):
        assert (g.get(n) == o.get(n)).all(), n
    say("Ni and Mi equal the oracle's")
    g.close()
    o.close()
    ds.close()


# (synth spec, params, K, compare with the oracle).  The largest job (the
# stream's extremes: N = 5,000, L = 200,000, ploidy 4, K = 10) is checked
# against its run alone; one medium job per model against the oracle, whose
# CPU passes at the largest size take minutes.
SERVED = [
    (dict(n_ind=5000, n_loc=200000, k_true=3, alpha_q=0.5, missing=0.01, ploidy=[(4, 1.0)], seed=41),
     dict(admix_val=1, nburnin=0, nreps=2, nthinning=2), 10, False),
    (dict(n_ind=1200, n_loc=40000, k_true=3, alpha_q=0.5, missing=0.01, ploidy=[(4, 1.0)], seed=47),
     dict(admix_val=1, nburnin=1, nreps=2, nthinning=1), 7, True),
    (dict(n_ind=2000, n_loc=50000, k_true=3, alpha_q=0.5, missing=0.01, ploidy=[(2, 1.0)], seed=42),
     dict(admix_val=2, nadmburnin=1, nburnin=1, nreps=2, nthinning=1), 6, True),
    (dict(n_ind=1500, n_loc=30000, k_true=3, alpha_q=0.5, missing=0.01, ploidy=[(4, 1.0)], seed=43),
     dict(admix_val=1, f_val=1, nburnin=1, nreps=2, nthinning=1), 4, True),
    (dict(n_ind=140, n_loc=2000, k_true=3, alpha_q=0.5, missing=0.01, ploidy=[(2, 1.0)], seed=44),
     dict(admix_val=1, nburnin=3, nreps=9, nthinning=2), 1, False),
    (dict(n_ind=900, n_loc=7000, k_true=3, alpha_q=0.5, missing=0.01, ploidy=[(4, 1.0)], seed=45),
     dict(admix_val=2, nadmburnin=2, nburnin=2, nreps=6, nthinning=3), 8, False),
]


def test_served_mix_equals_alone_and_oracle(vps, orc):
    srv = vps.Server(hbm_budget_bytes=64 << 30, max_active=6, slice_sweeps=1, loaders=2)
    jobs = []
    for j, (spec, par, K, _) in enumerate(SERVED):
        s = vps.Synth(**spec)
        p = vps.params_default(g_seed=7, z_chunk=-1, **par)
        jobs.append((srv.submit(s, p, K, j + 1), s, p, K))
        say("submitted job", j + 1, "N", spec["n_ind"], "L", spec["n_loc"], "K", K, par)
    names = ("Q", "MiSum", "rout", "lnltrace", "lnL")
    for j, (jid, s, p, K) in enumerate(jobs):
        srv.wait(jid, 600.0)
        got = {n: srv.result(jid, n) for n in names}
        ds = vps.Dataset(s, 0)
        c = vps.Chain(ds, p, K, j + 1)
        c.step(srv.info(jid)["sweeps_total"])
        c.finish()
        for n in names:
            want = c.get(n)
            assert (got[n] == want.reshape(got[n].shape)).all(), (j + 1, n)
        c.close()
        ds.close()
        say("job", j + 1, "equals its run alone")
        if SERVED[j][3]:
            o = orc.Chain(orc.Data(s.desc), p, K, j + 1)
            for x in range(srv.info(jid)["sweeps_total"]):
                o.sweep(1)
                say("  oracle sweep", x + 1)
            o.finish()
            assert (o.get("MiSum") == got["MiSum"].reshape(o.get("MiSum").shape)).all(), j + 1
            assert close(o.get("Q"), got["Q"].reshape(o.get("Q").shape)), j + 1
            # rout = [mean lnL, mean lnL^2, var, lnP(D), ...] (Arrange,
            # src/structure.cpp:1862-1869): the variance cancels two ~lnL^2
            # terms, so its bound is relative to mean lnL^2
            ro = o.get("rout")
            tol = RTOL * np.abs(ro)
            tol[2] = tol[3] = RTOL * abs(ro[1])
            assert (np.abs(got["rout"] - ro) <= tol).all(), (j + 1, got["rout"], ro)
            o.close()
            say("job", j + 1, "equals the oracle")
    st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    srv.close()
