"""The mixed-precision mode's error per sweep, at the mixed-ploidy multiallelic
location-prior workload's full size (N = 3,000, L = 200,000, ploidy 2/4/6,
20% multiallelic, K = 6): one sweep from identical FP64 state, FP32 weights
against FP64 (the FP64 path equals the oracle, tests/test_gpu_configs.py and
tests/test_gpu_parity.py).

Measured on MI355X: 450 of 1.74e9 assignments move (2.6e-7); after the
next UpdateP/UpdateQ, max|dP| 6.9e-4, max|dQ| 5.1e-6, relative dlnL 5.7e-11.
Bounds: at most 1 in 10^5 assignments moves; Q moves by at most 1e-4, P by
at most 0.05 (more than Q at loci with few copies of an allele in a cluster,
where one copy is a large share), the Record lnL by at most 1e-8 relative."""
import sys

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("topology", ["RNG_PARALLEL", "RNG_REFERENCE"])
@pytest.mark.parametrize("prec", ["PREC_MIXED", "PREC_FAST32"])
def test_single_precision_q_copy_follows_q(vps, topology, prec):
    """the FP32-weight Z passes read Q rounded to single precision: written
    beside Q by the Dirichlet UpdateQ, rounded before the pass otherwise
    (Metropolis sweeps); after every sweep it equals Q rounded, for every
    typed individual (an untyped one's slots are never counted)"""
    import numpy as np
    print("Q copy:", topology, prec, flush=True)
    s = vps.Synth(n_ind=150, n_loc=400, k_true=3, alpha_q=0.5, fst=0.05, missing=0.02, seed=41)
    ds = vps.Dataset(s, 0)
    vt = ds.order()[2]
    p = vps.params_default(admix_val=1, nburnin=4, nreps=8, metrofreq=3, g_seed=9,
                           rng_topology=getattr(vps, topology), precision=getattr(vps, prec))
    c = vps.Chain(ds, p, 4, 1)
    for sweep in range(8):
        c.step(1)
        q, qf = c.get("Q"), c.get("Qf")
        assert (qf[vt > 0] == q[vt > 0].astype(np.float32)).all(), sweep
    c.close()
    ds.close()


def test_mixed_one_sweep_error_full_size(vps):
    from vcfpop_amd.precision import one_sweep_error
    print("mixed-precision error: N=3000 L=200000 ploidy 2/4/6 K=6 locprior", flush=True)
    s = vps.make_synth("locprior_n3000_l200k_k6_mixed")
    ds = vps.Dataset(s, 0)
    p = vps.params_default(admix_val=1, locpriori_val=1, nburnin=0, nreps=10000, g_seed=1)
    e = one_sweep_error(ds, p, 6)
    print(e, flush=True)
    sys.stdout.flush()
    assert e["typed_allele_copies"] > 10 ** 9
    assert e["frac_moved_assignments"] <= 1e-5
    assert e["moved_by_allele_table_min"] <= 2 * max(e["moved_assignments_min"], 1) + 100
    assert e["max_abs_dQ_next_sweep"] <= 1e-4
    assert e["max_abs_dP_next_sweep"] <= 0.05
    assert e["rel_dlnL_next_sweep"] <= 1e-8
    ds.close()
