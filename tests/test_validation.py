"""Input validation at the C ABI, on the host (datasets built with device
-1; chain arguments are checked before any device call): empty and
malformed genotype images and out-of-range chain options are refused with
a status and a message, never a crash."""
import ctypes as C

import numpy as np
import pytest


@pytest.fixture(scope="module")
def arrays(vps):
    s = vps.Synth(n_ind=30, n_loc=80, k_true=2, alpha_q=0.5, missing=0.05, frac_multi=0.3, min_alleles=3,
                  max_alleles=6, ploidy=[(2, 0.5), (4, 0.5)], seed=61)
    return {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for k, v in s.arrays().items()}


def edited(arrays, **kw):
    a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
    a.update(kw)
    return a


def test_valid_image_builds(vps, arrays):
    ds = vps.Dataset(arrays, -1)
    assert (ds.N, ds.L) == (30, 80)
    ds.close()


@pytest.mark.parametrize("case, match", [
    ("no_individuals", "empty dataset"),
    ("no_loci", "empty dataset"),
    ("bitsize_zero", "bitsize 0 out of range"),
    ("too_many_genotypes", "genotypes for"),
    ("no_alleles", "0 alleles"),
    ("too_many_alleles", "supports at most"),
])
def test_malformed_image_is_refused(vps, arrays, case, match):
    if case == "no_individuals":
        a = edited(arrays, N=0)
    elif case == "no_loci":
        a = edited(arrays, L=0)
    elif case == "bitsize_zero":
        b = arrays["bitsize"].copy()
        b[5] = 0
        a = edited(arrays, bitsize=b)
    elif case == "too_many_genotypes":
        b = arrays["bitsize"].copy()
        ng = np.diff(arrays["gtab_offset"])
        l = int(np.argmax(ng))
        b[l] = max(1, int(np.ceil(np.log2(ng[l]))) - 1)   # fewer bits than the genotype table needs
        a = edited(arrays, bitsize=b)
    elif case == "no_alleles":
        k = arrays["n_allele"].copy()
        k[7] = 0
        a = edited(arrays, n_allele=k)
    else:
        k = arrays["n_allele"].copy()
        k[7] = 1000
        a = edited(arrays, n_allele=k)
    with pytest.raises(vps.VpsError, match=match):
        vps.Dataset(a, -1)


def test_chain_options_are_checked(vps, arrays):
    ds = vps.Dataset(arrays, -1)
    lib = vps.load()
    h = C.c_void_p()
    for K, kw, match in [(0, {}, "K=0"), (17, {}, "K=17"), (3, {"nthinning": 0}, "nthinning"),
                         (3, {"rng_topology": 7}, "rng_topology"), (3, {"precision": 9}, "precision")]:
        p = vps.params_default(**kw)
        rc = lib.vps_chain_create(ds.h, C.byref(p), K, 1, C.byref(h))
        assert rc != 0
        assert match in lib.vps_last_error().decode(), (K, kw)
    ds.close()
