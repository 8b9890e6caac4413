"""vps_record_lnl_noalloc: Record's likelihood pass on caller-owned device
buffers, its partials in the library's static scratch array, against the
same pass of vps_pass_record on a chain holding the same P and Q (or Z).

Bar: bit-identical lnL (the same kernels, the same partition into locus
ranges, the same fixed-order sum).  Two calls in a row, the second on a
smaller dataset, so the second call reads only partials it wrote itself."""
import ctypes
import sys

import numpy as np
import pytest

from helpers import random_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _Hip:
    """device buffers through the HIP runtime the library itself links
    (libamdhip64.so.7, already loaded with it), not through another copy"""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def _ok(self, e, what):
        assert e == 0, (what, e)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        self._ok(self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))), "hipMalloc")
        self.bufs.append(p)
        self._ok(self.rt.hipMemcpy(p, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy H2D")
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self._ok(self.rt.hipMemcpy(ctypes.c_void_p(out.ctypes.data), p, ctypes.c_size_t(out.nbytes), 2), "hipMemcpy D2H")
        return out

    def sync(self):
        self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _case(vps, synth, K, seed, binary):
    ds = vps.Dataset(synth, 0)
    orig, afo, vt = ds.order()
    k2 = synth.arrays()["n_allele"][orig]
    rng = np.random.default_rng(seed)
    Freq, Q = random_state(rng, ds.N, K, ds.KT, afo, k2)
    Z = rng.integers(0, K, size=ds.N).astype(np.uint16)
    ch = vps.Chain(ds, vps.params_default(admix_val=2 if binary else 1, g_seed=seed, z_chunk=-1), K, 1)
    ch.set("Freq", Freq)
    if binary:
        ch.set("Z", Z)
        ch.set("binaryq", [1])
    else:
        ch.set("Q", Q)
    ch.record()
    want = ch.get("lnL")[0]
    ch.close()

    hip = _Hip()
    f = hip.upload(np.concatenate([Freq.ravel(), np.zeros(2)]))
    q = hip.upload(Q.ravel())
    z = hip.upload(Z)
    nan = np.full(1, np.nan)
    out = hip.upload(nan)
    lib = vps.load()
    rc = lib.vps_record_lnl_noalloc(ds.h, K, f, None if binary else q, z if binary else None, out, None)
    assert rc == 0, lib.vps_last_error()
    hip.sync()
    got = hip.download(out, nan)[0]
    hip.close()
    ds.close()
    return got, want


@pytest.mark.parametrize("binary", [False, True])
def test_noalloc_record_equals_chain_record(vps, binary):
    big = vps.Synth(n_ind=700, n_loc=5000, k_true=3, alpha_q=0.5, missing=0.02, frac_multi=0.2,
                    min_alleles=3, max_alleles=6, ploidy=[(2, 0.6), (4, 0.4)], seed=21)
    small = vps.Synth(n_ind=90, n_loc=600, k_true=2, alpha_q=0.5, missing=0.05, seed=22)
    for name, synth, K in (("N=700 L=5000 mixed ploidy", big, 5), ("N=90 L=600 diploid", small, 3)):
        say("noalloc record:", name, "K", K, "binary", binary)
        got, want = _case(vps, synth, K, 7, binary)
        say("  lnL", got, "chain", want)
        assert np.isfinite(want)
        assert got == want, (name, got, want)


def test_noalloc_rejects_bad_arguments(vps):
    lib = vps.load()
    s = vps.Synth(n_ind=40, n_loc=100, k_true=2, seed=23)
    ds = vps.Dataset(s, 0)
    dummy = ctypes.c_void_p(16)
    assert lib.vps_record_lnl_noalloc(ds.h, 0, dummy, dummy, None, dummy, None) != 0     # K out of range
    assert lib.vps_record_lnl_noalloc(ds.h, 3, None, dummy, None, dummy, None) != 0      # no P
    assert lib.vps_record_lnl_noalloc(ds.h, 3, dummy, None, None, dummy, None) != 0      # neither Q nor Z
    ds.close()
