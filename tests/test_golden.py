"""The committed fixtures under tests/golden (made by tests/golden/make_golden.py
from the oracle) are reproduced exactly: RNG streams and hashes, a small
chain's states after init and sweeps 1, 2, 10, and injected-state Z passes on
a mixed-ploidy multiallelic dataset."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden as MG  # noqa: E402


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def test_rng_and_hash_vectors(orc):
    g = load("rng.npz")
    now = MG.rng_vectors()
    assert set(g.files) == set(now)
    for k in g.files:
        assert (g[k] == now[k]).all(), k


def test_chain_states(orc):
    g = load("chain.npz")
    now = MG.chain_vectors()
    for k in g.files:
        assert (g[k] == now[k]).all(), k


def test_zpass_vectors(orc):
    g = load("zpass.npz")
    now = MG.zpass_vectors()
    for k in g.files:
        assert (g[k] == now[k]).all(), k


@pytest.mark.gpu
def test_device_reproduces_chain_fixture(vps):
    g = load("chain.npz")
    s = vps.Synth(**MG.CHAIN_SYNTH)
    ds = vps.Dataset(s, 0)
    c = vps.Chain(ds, vps.params_default(**MG.CHAIN_PARAMS), 3, 1)
    done = 0
    for tag, upto in (("init", 0), ("m1", 1), ("m2", 2), ("m10", 10)):
        c.step(upto - done)
        done = upto
        for n in ("Z", "Ni", "Mi"):
            assert (c.get(n) == g[f"{tag}_{n}"]).all(), (tag, n)
        for n in ("Freq", "Q", "Alpha"):
            assert np.allclose(c.get(n), g[f"{tag}_{n}"], rtol=1e-12, atol=0), (tag, n)
        c.record()
        assert abs(c.get("lnL")[0] - g[f"{tag}_lnL"][0]) <= 1e-12 * abs(g[f"{tag}_lnL"][0]), tag


@pytest.mark.gpu
@pytest.mark.parametrize("K", [2, 5, 9])
def test_device_reproduces_zpass_fixture(vps, K):
    g = load("zpass.npz")
    s = vps.Synth(**MG.ZPASS_SYNTH)
    ds = vps.Dataset(s, 0)
    c = vps.Chain(ds, vps.params_default(admix_val=1, g_seed=11), K, 1)
    c.set("Freq", g[f"K{K}_Freq"])
    c.set("Q", g[f"K{K}_Q"])
    c.set("m", g[f"K{K}_m"].astype(np.int32))
    c.update_z_admix()
    assert (c.get("Ni") == g[f"K{K}_Ni"]).all()
    assert (c.get("Mi") == g[f"K{K}_Mi"]).all()
