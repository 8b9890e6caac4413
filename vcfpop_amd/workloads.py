"""Benchmark workloads: the five configurations the engine is measured on.

Each entry gives the synthetic-data recipe (vps_synth_config fields) and the
run shape.  Sizes are exactly those of the benchmark definitions; tests use
scaled-down copies.
"""
from ._lib import Synth

WORKLOADS = {
    # plain admixture model, CPU-runnable: diploid, N=200, L=2,000 biallelic, K=3,
    # 1,000 burn-in + 2,000 sampling sweeps, P ~ Dir(1,1), Q ~ Dir(0.5 x 3), 1% missing
    "admix_n200_l2k_k3": dict(
        synth=dict(n_ind=200, n_loc=2000, k_true=3, alpha_q=0.5, fst=0.0, missing=0.01, seed=1),
        K=3, nburnin=1000, nreps=2000, admix=True),
    # plain drop-in on GPU: diploid, N=2,000, L=100,000 biallelic (2 bits/genotype), K=5,
    # admixture with alpha Metropolis, 10,000 sweeps, Fst 0.05 drift, 2% missing
    "admix_n2000_l100k_k5": dict(
        synth=dict(n_ind=2000, n_loc=100000, k_true=5, alpha_q=0.5, fst=0.05, missing=0.02, seed=2),
        K=5, nburnin=0, nreps=10000, admix=True),
    # batched chains: N=1,000, L=50,000, true K=4; 64 chains (K=2..9 x 8 seeds), 5,000 sweeps each
    "batch_n1000_l50k_k2to9x8": dict(
        synth=dict(n_ind=1000, n_loc=50000, k_true=4, alpha_q=0.5, fst=0.05, missing=0.02, seed=3),
        K=list(range(2, 10)), replicates=8, nburnin=0, nreps=5000, admix=True),
    # mixed precision, mixed ploidy 2/4/6 (60/30/10%), L=200,000, 20% multiallelic (3-8 alleles,
    # Dirichlet(0.3)), K=6, locprior with 8 sampling locations
    "locprior_n3000_l200k_k6_mixed": dict(
        synth=dict(n_ind=3000, n_loc=200000, k_true=6, alpha_q=0.5, fst=0.05, missing=0.02,
                   frac_multi=0.2, min_alleles=3, max_alleles=8, multi_dir=0.3, n_pop=8,
                   ploidy=[(2, 0.6), (4, 0.3), (6, 0.1)], seed=4),
        K=6, nburnin=0, nreps=10000, admix=True, locpriori=True, precision="mixed"),
}


def make_synth(name, scale=None, **override):
    """Synthetic dataset of a workload; `scale` shrinks N and L (tests)."""
    spec = dict(WORKLOADS[name]["synth"])
    if scale:
        spec["n_ind"] = max(8, int(spec["n_ind"] * scale))
        spec["n_loc"] = max(8, int(spec["n_loc"] * scale))
    spec.update(override)
    return Synth(**spec)
