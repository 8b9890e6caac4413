"""The vcfpop-side adapter (vcfpop_amd/adapter/structureVPS.cpp), compiled
against the vcfpop.h fixture in tests/adapter_mock.

* It exports the reference's runtime seam: the extern "C" set of
  src/structureCUDA.h:6-52, link-compatible with structureCUDA.o, and the
  BAYESIAN<REAL> members behind `usecuda` for REAL = double and float
  (src/structure.h:186-306).
* Driven through those symbols with host-owned state, as vcfpop's pass
  dispatchers drive them, each member equals the oracle's pass: the Z passes'
  counts bit for bit, P, Q and the likelihoods to 1e-12, and the whole-run
  hand-off equals the oracle's run.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOCK = os.path.join(ROOT, "tests", "adapter_mock", "libvps_adapter_mock.so")
RTOL = 1e-12

# src/structureCUDA.h:6-52
SEAM = ["WaitCUDA", "Malloc", "MallocHostCUDA", "FreeHostCUDA", "MallocDeviceCUDA", "FreeCUDA", "MemsetCUDA",
        "MemcpyCUDA", "FreeStructureMemory", "CopyStructureMemory", "GetDeviceCountCUDA", "AllocMemoryCUDA",
        "FreeMemoryCUDA", "CreateStreamCUDA", "DestroyStreamCUDA", "ShowDevicesCUDA", "ResetDeviceCUDA",
        "Svd64CUDA", "Svd32CUDA", "Eig64CUDA", "Eig32CUDA", "MatrixMul64CUDA", "MatrixMul32CUDA"]
MEMBERS = ["InitCUDA", "UninitCUDA", "UpdatePCUDA", "UpdateQNoAdmixCUDA", "UpdateQMetroCUDA", "UpdateZAdmixCUDA",
           "UpdateZNoAdmixCUDA", "RecordCUDA"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_adapter_exports_the_reference_seam(vps):
    assert os.path.exists(MOCK), "run __graft_entry__.build()"
    out = subprocess.run(["nm", "-DC", "--defined-only", MOCK], capture_output=True, text=True, check=True).stdout
    plain = {line.split()[-1] for line in out.splitlines() if line.split()[1] in "TW"}
    missing = [s for s in SEAM if s not in plain]
    assert not missing, missing
    for real in ("double", "float"):
        for m in MEMBERS:
            assert f"BAYESIAN<{real}>::{m}()" in out, (real, m)
        assert f"VpsMCMC<{real}>(BAYESIAN<{real}>&)" in out


class Mock:
    def __init__(self):
        self.lib = C.CDLL(MOCK)
        L = self.lib
        L.mock_set_image.argtypes = [C.c_void_p]
        L.mock_set_options.argtypes = [C.c_void_p]
        L.mock_run_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.mock_run_new.restype = C.c_void_p
        L.mock_run_free.argtypes = [C.c_void_p]
        L.mock_run_field.argtypes = [C.c_void_p, C.c_char_p]
        L.mock_run_field.restype = C.c_void_p
        L.mock_run_call.argtypes = [C.c_void_p, C.c_char_p]
        L.mock_seam.argtypes = [C.c_char_p, C.c_int]
        L.mock_last_exit.restype = C.c_char_p

    def seam(self, fn, dev=0):
        assert self.lib.mock_seam(fn.encode(), dev) == 0, self.lib.mock_last_exit()


class Run:
    """a BAYESIAN<REAL> of the fixture (REAL = float with real32) with host
    buffers as numpy views"""
    DT = {"Ni": (np.int32, "KKT"), "Mi": (np.int64, "NK"), "MiSum": (np.int64, "NK"), "Z": (np.uint16, "N"),
          "Freq": (np.float64, "KKT"), "FreqM": (np.float64, "KKT"), "FreqA": (np.float64, "3KT"),
          "Q": (np.float64, "NK"), "Alpha": (np.float64, "K"), "Lambda": (np.float64, "K"), "f": (np.float64, "K"),
          "F": (np.float64, "K"), "Eta": (np.float64, "K"), "Gamma": (np.float64, "SK"), "Di": (np.float64, "SK"),
          "AlphaLocal": (np.float64, "SK"), "SumAlpha": (np.float64, "S"), "bufNK1": (np.float64, "NK"),
          "bufN1": (np.float64, "N"), "rout": (np.float64, "R"), "m": (np.int32, "1"), "nr": (np.int32, "1"),
          "r": (np.float64, "1")}

    REAL = ("Freq", "FreqM", "FreqA", "Q", "Di")   # REAL arrays of BAYESIAN<REAL>

    def __init__(self, mock, K, run_id, N, KT, S, rlen, real32=False):
        self.m, self.h = mock, mock.lib.mock_run_new(K, run_id, rlen, 0, 1 if real32 else 0)
        self.real32 = real32
        self.n = {"KKT": K * KT, "NK": N * K, "N": N, "3KT": 3 * KT, "K": K, "SK": S * K, "S": S, "R": rlen, "1": 1}
        self.K, self.S = K, S

    def __getitem__(self, name):
        dt, n = self.DT[name]
        if self.real32 and name in self.REAL:
            dt = np.float32
        p = self.m.lib.mock_run_field(self.h, name.encode())
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(self.n[n],))

    def call(self, member):
        assert self.m.lib.mock_run_call(self.h, member.encode()) == 0, self.m.lib.mock_last_exit()

    def load_from(self, o):
        """the oracle chain's state into the host fields"""
        K, S = self.K, self.S
        for n in ("Freq", "Q", "Z", "Ni", "Mi", "Alpha", "Lambda"):
            self[n][:] = o.get(n).ravel()
        loc = o.get("loc")
        self["r"][0] = loc[0]
        self["Eta"][:] = loc[1:1 + K]
        self["Gamma"][:] = loc[1 + K:1 + K + S * K]
        self["AlphaLocal"][:] = loc[1 + K + S * K:1 + K + 2 * S * K]
        self["Di"][:] = loc[1 + K + 2 * S * K:1 + K + 3 * S * K]
        self["SumAlpha"][:] = loc[1 + K + 3 * S * K:]
        fst = o.get("Fst")
        self["F"][:], self["f"][:] = fst[0], fst[1]
        self["FreqA"][:] = o.get("FreqA").ravel()
        self["m"][0] = o.get("m")[0]

    def close(self):
        self.m.lib.mock_run_free(self.h)


def close(a, b, rtol=RTOL):
    return np.allclose(np.asarray(a, np.float64), np.asarray(b, np.float64), rtol=rtol, atol=0)


@pytest.mark.gpu
def test_members_through_the_seam_equal_the_oracle(vps, orc):
    mock = Mock()
    s = vps.Synth(n_ind=70, n_loc=400, k_true=3, alpha_q=0.5, missing=0.04, frac_multi=0.2, min_alleles=3,
                  max_alleles=5, ploidy=[(2, 0.6), (4, 0.4)], n_pop=3, seed=61)
    od = orc.Data(s.desc)
    N, KT, S = od.N, od.KT, 3
    assert mock.lib.mock_set_image(C.byref(s.desc)) == 0
    mock.seam("CopyStructureMemory", 0)
    mock.seam("CreateStreamCUDA", 0)
    K = 3

    # admixture with the location prior
    p = vps.params_default(admix_val=1, locpriori_val=1, g_seed=5, nburnin=2, nreps=6)
    mock.lib.mock_set_options(C.byref(p))
    o = orc.Chain(od, p, K, 2)
    o.sweep(3)
    run = Run(mock, K, 2, N, KT, S, o.get("rout").size)
    run.load_from(o)
    run.call("InitCUDA")
    run.call("UpdateZAdmixCUDA")
    o.call("update_z_admix")
    for n in ("Ni", "Mi"):
        assert (run[n] == o.get(n).ravel()).all(), n
    say("UpdateZAdmixCUDA equals the oracle")
    run.call("UpdatePCUDA")
    o.call("update_p")
    assert close(run["Freq"], o.get("Freq").ravel())
    say("UpdatePCUDA equals the oracle")
    o.set("m", [10])
    run["m"][0] = 10
    run.call("UpdateQMetroCUDA")
    o.call("update_q_metro")
    assert close(run["Q"], o.get("Q").ravel())
    r_o = o.get("bufN1")
    assert (np.abs(run["bufN1"] - r_o) <= RTOL * np.maximum(np.abs(r_o), 1.0)).all()
    say("UpdateQMetroCUDA equals the oracle")
    run.call("RecordCUDA")
    ref = o.record_pass()
    assert abs(run["bufNK1"][0] - ref) <= RTOL * abs(ref)
    say("RecordCUDA equals the oracle")
    run.call("UninitCUDA")
    run.close()
    o.close()

    # no admixture: the PolyLog Q draw and the binary-Q counts
    p = vps.params_default(admix_val=2, nadmburnin=0, g_seed=6, nburnin=2, nreps=6)
    mock.lib.mock_set_options(C.byref(p))
    o = orc.Chain(od, p, K, 3)
    o.sweep(2)
    run = Run(mock, K, 3, N, KT, S, o.get("rout").size)
    run.load_from(o)
    run.call("InitCUDA")
    run.call("UpdateQNoAdmixCUDA")
    o.call("update_q_noadmix")
    assert (run["Z"] == o.get("Z")).all() and (run["Q"] == o.get("Q").ravel()).all()
    run.call("UpdateZNoAdmixCUDA")
    o.call("update_z_noadmix")
    for n in ("Ni", "Mi"):
        assert (run[n] == o.get(n).ravel()).all(), n
    say("UpdateQNoAdmixCUDA and UpdateZNoAdmixCUDA equal the oracle")
    run.call("UninitCUDA")
    run.close()
    o.close()

    # the whole-run hand-off from the host's initial state
    p = vps.params_default(admix_val=1, g_seed=7, nburnin=20, nreps=40, nthinning=4)
    mock.lib.mock_set_options(C.byref(p))
    o = orc.Chain(od, p, K, 4)
    run = Run(mock, K, 4, N, KT, S, o.get("rout").size)
    run.load_from(o)
    run.call("InitCUDA")
    run.call("MCMC")
    o.sweep(60)
    o.finish()
    assert run["m"][0] == 60 and run["nr"][0] == 10
    assert (run["MiSum"].view(np.float64) == o.get("MiSum").ravel()).all()
    assert close(run["Q"], o.get("Q").ravel())
    ro = o.get("rout")
    tol = RTOL * np.abs(ro)
    tol[2] = tol[3] = RTOL * abs(ro[1])
    assert (np.abs(run["rout"] - ro) <= tol).all()
    say("the whole-run hand-off equals the oracle's run")
    run.call("UninitCUDA")
    run.close()
    mock.seam("FreeStructureMemory", 0)


@pytest.mark.gpu
@pytest.mark.parametrize("fastsingle", [1, 2])
def test_float_members_equal_the_oracle(vps, orc, fastsingle):
    """BAYESIAN<float> (-g_float=single): -g_fastsingle=yes runs the engine's
    FAST32 mode, =no its REAL32 mode; members and the whole run equal the
    oracle's single-precision path"""
    mock = Mock()
    C.c_int.in_dll(mock.lib, "g_fastsingle_val").value = fastsingle
    prec = vps.PREC_FAST32 if fastsingle == 1 else vps.PREC_REAL32
    s = vps.Synth(n_ind=70, n_loc=400, k_true=3, alpha_q=0.5, missing=0.04, frac_multi=0.2, min_alleles=3,
                  max_alleles=5, ploidy=[(2, 0.6), (4, 0.4)], n_pop=3, seed=61)
    od = orc.Data(s.desc)
    N, KT, S = od.N, od.KT, 3
    assert mock.lib.mock_set_image(C.byref(s.desc)) == 0
    mock.seam("CopyStructureMemory", 0)
    mock.seam("CreateStreamCUDA", 0)
    K = 3
    p = vps.params_default(admix_val=1, g_seed=5, nburnin=2, nreps=6, precision=prec)
    mock.lib.mock_set_options(C.byref(p))
    o = orc.Chain(od, p, K, 2)
    o.sweep(3)
    run = Run(mock, K, 2, N, KT, S, o.get("rout").size, real32=True)
    run.load_from(o)
    assert (run["Freq"].astype(np.float64) == o.get("Freq").ravel()).all()   # single-precision state
    run.call("InitCUDA")
    run.call("UpdateZAdmixCUDA")
    o.call("update_z_admix")
    for n in ("Ni", "Mi"):
        assert (run[n] == o.get(n).ravel()).all(), n
    run.call("UpdatePCUDA")
    o.call("update_p")
    assert (run["Freq"].astype(np.float64) == o.get("Freq").ravel()).all()
    o.set("m", [9])
    run["m"][0] = 9
    run.call("UpdateQMetroCUDA")
    o.call("update_q_metro")
    assert (run["Q"].astype(np.float64) == o.get("Q").ravel()).all()
    run.call("RecordCUDA")
    ref = o.record_pass()
    assert abs(run["bufNK1"][0] - ref) <= RTOL * abs(ref)
    say("BAYESIAN<float> members equal the oracle, fastsingle", fastsingle)
    run.call("UninitCUDA")
    run.close()
    o.close()

    p = vps.params_default(admix_val=1, g_seed=7, nburnin=20, nreps=40, nthinning=4, precision=prec)
    mock.lib.mock_set_options(C.byref(p))
    o = orc.Chain(od, p, K, 4)
    run = Run(mock, K, 4, N, KT, S, o.get("rout").size, real32=True)
    run.load_from(o)
    run.call("InitCUDA")
    run.call("MCMC")
    o.sweep(60)
    o.finish()
    assert run["m"][0] == 60 and run["nr"][0] == 10
    assert (run["MiSum"].view(np.float64) == o.get("MiSum").ravel()).all()
    assert close(run["Q"], o.get("Q").ravel())
    assert close(run["FreqM"], o.get("FreqM").ravel())
    ro = o.get("rout")
    tol = RTOL * np.abs(ro)
    tol[2] = tol[3] = RTOL * abs(ro[1])
    assert (np.abs(run["rout"] - ro) <= tol).all()
    say("BAYESIAN<float> whole run equals the oracle's")
    run.call("UninitCUDA")
    run.close()
    mock.seam("FreeStructureMemory", 0)
