"""The C ABI: every function declared in include/*.h is exported by the
built library and bound by the Python mirror.  No device calls."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vps_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_function_is_exported(vps):
    lib = ctypes.CDLL(vps.LIB_PATH)
    names = declared("vps.h") + declared("vps_synth.h") + declared("vps_ingest.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_function(vps):
    from vcfpop_amd import _lib
    names = declared("vps.h") + declared("vps_synth.h") + declared("vps_ingest.h")
    unbound = [n for n in names if n not in _lib._SIGS]
    assert not unbound, unbound


def test_params_struct_matches_header(vps):
    p = vps.params_default()
    assert p.admix_val == 2 and p.nburnin == 1000 and p.nreps == 10000 and p.nthinning == 10
    assert p.alpha == 1.0 and p.metrofreq == 10 and p.z_chunk == 0


def test_abi_version(vps):
    assert vps.load().vps_abi_version() == 1
