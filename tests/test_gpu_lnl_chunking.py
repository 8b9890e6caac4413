"""Record's lnL summed inside the FP64 Z pass does not depend on how the
individuals are chunked: partials are kept per (tile, group of 32
individuals), and every automatic chunking cuts at group boundaries.  The
same seed then reports the same lnL bits whatever chunking the tuner picks,
and the value agrees with the oracle (src/structure.cpp:1943-2045) to 1e-12.
A chunking that splits a group (an explicit z_chunk) takes Record's own
likelihood pass instead."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def data(vps, orc):
    orc.set_threads(8)
    out = {}
    for name, spec in {
        "bi": dict(n_ind=768, n_loc=2500, k_true=3, alpha_q=0.5, fst=0.05, missing=0.02, seed=31),
        "mixed": dict(n_ind=320, n_loc=800, k_true=3, alpha_q=0.5, fst=0.05, missing=0.03, frac_multi=0.25,
                      min_alleles=3, max_alleles=6, multi_dir=0.3, ploidy=[(2, 0.6), (4, 0.4)], seed=32),
    }.items():
        s = vps.Synth(**spec)
        out[name] = (s, vps.Dataset(s, 0), orc.Data(s.desc))
    return out


# N is a multiple of 32, so each of these z_chunk values but the last gives
# balanced chunks of whole groups (the last splits groups)
CHUNKS = {"bi": (32, 128, 256, 50), "mixed": (32, 64, 160, 50)}


@pytest.mark.parametrize("name", ["bi", "mixed"])
def test_record_lnl_bits_equal_across_chunkings(vps, orc, data, name):
    s, ds, od = data[name]
    # nthinning 1: every sweep records; Metropolis every 10th, so 9 of 10
    # Record sweeps sum their lnL inside the Z pass
    kw = dict(admix_val=1, nburnin=0, nreps=40, nthinning=1, g_seed=5)
    traces = {}
    a, b, cc, split = CHUNKS[name]
    for zc in CHUNKS[name]:
        say(name, "z_chunk", zc)
        c = vps.Chain(ds, vps.params_default(z_chunk=zc, **kw), 3, 1)
        G, Nc = c.chunking()
        assert (Nc % 32 == 0) == (zc != split), (zc, G, Nc)
        tr = []
        for _ in range(12):
            c.step(1)
            tr.append(float(c.get("lnL")[0]))
        traces[zc] = np.array(tr)
        c.close()
    o = orc.Chain(od, vps.params_default(**kw), 3, 1)
    ot = []
    for _ in range(12):
        o.sweep(1)
        ot.append(float(o.get("lnL")[0]))
    ot = np.array(ot)
    # whole groups: the same bits
    assert (traces[a] == traces[b]).all() and (traces[a] == traces[cc]).all()
    # the oracle and a group-splitting chunking: to 1e-12
    assert np.allclose(traces[a], ot, rtol=1e-12, atol=0)
    assert np.allclose(traces[split], ot, rtol=1e-12, atol=0)


def test_automatic_chunks_keep_groups_only_when_needed(vps, data):
    """a chain whose Record sweeps can sum in the Z pass (thinning not a
    multiple of metrofreq) gets chunks of whole groups from the tuner; with
    the default schedule every Record sweep is a Metropolis sweep and the
    chunks stay balanced"""
    s, ds, od = data["bi"]
    c = vps.Chain(ds, vps.params_default(admix_val=1, nburnin=0, nreps=40, nthinning=3, g_seed=5), 3, 1)
    G, Nc = c.chunking()
    c.close()
    assert G == 1 or Nc % 32 == 0, (G, Nc)
    c = vps.Chain(ds, vps.params_default(admix_val=1, nburnin=20, nreps=40, g_seed=5), 3, 1)
    G, Nc = c.chunking()
    c.close()
    assert Nc == -(-ds.N // G), (G, Nc)
