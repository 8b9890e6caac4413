"""ctypes wrapper of the oracle (oracle/liboracle.so).

The oracle is a CPU restatement of vcfpop's Bayesian clustering path and is
TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker.  The engine (vcfpop_amd/) never
loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    u64, u32, i64, dbl, vp = C.c_uint64, C.c_uint32, C.c_int64, C.c_double, C.c_void_p
    sig = {
        "or_crc32c_u64": ([u32, u64], u32),
        "or_hash_ulong": ([u64], u32),
        "or_hash_string": ([vp, C.c_int], u32),
        "or_murmur2": ([u32, u32], u32),
        "or_murmur64": ([u64, u64], u64),
        "or_murmur32": ([u64, u64], u32),
        "or_rng_init": ([vp, u64, u64], None),
        "or_rng_xorshift": ([vp], u64),
        "or_rng_uniform": ([vp, dbl, dbl], dbl),
        "or_rng_normal": ([vp, dbl, dbl], dbl),
        "or_rng_gamma": ([vp, dbl, dbl], dbl),
        "or_rng_next": ([vp, u64], u64),
        "or_rng_next_avoid": ([vp, u64, u64], u64),
        "or_rng_poly_scalar": ([vp, vp, C.c_int], C.c_int),
        "or_rng_poly_avx": ([vp, vp, C.c_int], C.c_int),
        "or_rng_polylog": ([vp, vp, C.c_int], C.c_int),
        "or_rng_dirichlet_ab": ([vp, vp, vp, vp, C.c_int, C.c_int], None),
        "or_rng_dirichlet_af": ([vp, vp, vp, C.c_int, dbl], None),
        "or_rngf_init": ([vp, u64, u64], None),
        "or_rngf_xorshift": ([vp], u32),
        "or_rngf_uniform": ([vp, C.c_float, C.c_float], C.c_float),
        "or_rngf_poly_avx": ([vp, vp, C.c_int], C.c_int),
        "or_rng_dirichlet_ab32": ([vp, vp, vp, vp, C.c_int, C.c_int], None),
        "or_rng_dirichlet_af32": ([vp, vp, vp, C.c_int, dbl], None),
        "or_mylog": ([dbl], dbl),
        "or_mysin": ([dbl], dbl),
        "or_mycos": ([dbl], dbl),
        "or_mysqrt": ([dbl], dbl),
        "or_charge_log": ([vp, vp, dbl], None),
        "or_close_log": ([i64, dbl], dbl),
        "or_logprod_avx": ([vp, i64], dbl),
        "or_logprod_sep": ([vp, i64, i64], dbl),
        "or_sum_avx": ([vp, i64], dbl),
        "or_unify_avx": ([vp, i64], None),
        "or_data_create": ([vp], vp),
        "or_data_destroy": ([vp], None),
        "or_data_kt": ([vp], i64),
        "or_data_maxploidy": ([vp], C.c_int),
        "or_data_order": ([vp, vp, vp, vp], None),
        "or_data_gid": ([vp, i64, C.c_int], C.c_int),
        "or_set_threads": ([C.c_int], None),
        "or_chain_create": ([vp, vp, C.c_int, C.c_int], vp),
        "or_chain_destroy": ([vp], None),
        "or_chain_get_state": ([vp, C.c_int, vp, C.c_size_t], C.c_int),
        "or_chain_set_state": ([vp, C.c_int, vp, C.c_size_t], C.c_int),
        "or_chain_sizes": ([vp, C.POINTER(i64), C.POINTER(i64)], None),
        "or_chain_seed": ([vp], u64),
        "or_record_pass": ([vp], dbl),
        "or_q_metro_logratio": ([vp, vp, vp], None),
        "or_q_noadmix_loglik": ([vp, vp], None),
        "or_sweep": ([vp, C.c_int], None),
    }
    for name in ("or_update_p", "or_update_q", "or_update_q_admix", "or_update_q_metro", "or_update_q_noadmix",
                 "or_update_locpriori", "or_update_z", "or_update_z_admix", "or_update_z_noadmix",
                 "or_update_alpha", "or_update_lambda", "or_update_a", "or_update_f", "or_record",
                 "or_arrange"):
        sig[name] = ([vp], None)
    for name, (a, r) in sig.items():
        f = getattr(lib, name)
        f.argtypes = a
        f.restype = r
    _lib = lib
    return lib


class Rng(C.Structure):
    _fields_ = [("x", C.c_uint64), ("y", C.c_uint64), ("U1", C.c_double), ("U2", C.c_double)]

    def __init__(self, seed, salt):
        super().__init__()
        load().or_rng_init(C.byref(self), seed, salt)

    def xorshift(self):
        return load().or_rng_xorshift(C.byref(self))

    def uniform(self, a=0.0, b=1.0):
        return load().or_rng_uniform(C.byref(self), a, b)

    def normal(self, mean=0.0, sd=1.0):
        return load().or_rng_normal(C.byref(self), mean, sd)

    def gamma(self, alpha, beta=1.0):
        return load().or_rng_gamma(C.byref(self), alpha, beta)

    def next(self, mx):
        return load().or_rng_next(C.byref(self), mx)

    def next_avoid(self, mx, av):
        return load().or_rng_next_avoid(C.byref(self), mx, av)

    def poly_avx(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        return load().or_rng_poly_avx(C.byref(self), w.ctypes.data, len(w))

    def poly_scalar(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        return load().or_rng_poly_scalar(C.byref(self), a.ctypes.data, len(a))

    def polylog(self, a):
        a = np.array(a, dtype=np.float64)
        return load().or_rng_polylog(C.byref(self), a.ctypes.data, len(a))

    def dirichlet_ab(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.int64)
        res = np.empty(len(a))
        load().or_rng_dirichlet_ab(C.byref(self), res.ctypes.data, a.ctypes.data, b.ctypes.data, 1, len(a))
        return res

    def dirichlet_a(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        res = np.empty(len(a))
        load().or_rng_dirichlet_af(C.byref(self), res.ctypes.data, a.ctypes.data, len(a), 1.0)
        return res

    def dirichlet_ab32(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.int64)
        res = np.empty(len(a))
        load().or_rng_dirichlet_ab32(C.byref(self), res.ctypes.data, a.ctypes.data, b.ctypes.data, 1, len(a))
        return res

    def dirichlet_a32(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        res = np.empty(len(a))
        load().or_rng_dirichlet_af32(C.byref(self), res.ctypes.data, a.ctypes.data, len(a), 1.0)
        return res


class RngF(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]

    def __init__(self, seed, salt):
        super().__init__()
        load().or_rngf_init(C.byref(self), seed, salt)

    def xorshift(self):
        return load().or_rngf_xorshift(C.byref(self))

    def poly_avx(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        return load().or_rngf_poly_avx(C.byref(self), w.ctypes.data, len(w))


def charge_log(values):
    """sequential OpenLog/ChargeLog/CloseLog over values"""
    lib = load()
    s, p = C.c_int64(0), C.c_double(1.0)
    for v in values:
        lib.or_charge_log(C.byref(s), C.byref(p), float(v))
    return lib.or_close_log(s.value, p.value)


def set_threads(n):
    load().or_set_threads(int(n))


class Data:
    """Oracle view of a dataset description (VpsDatasetDesc)."""

    def __init__(self, desc):
        self._desc = desc
        self.h = load().or_data_create(C.addressof(desc))
        self.N, self.L = desc.n_ind, desc.n_loc
        self.KT = load().or_data_kt(self.h)
        self.maxploidy = load().or_data_maxploidy(self.h)

    def order(self):
        orig = np.empty(self.L, dtype=np.int64)
        afo = np.empty(self.L, dtype=np.int64)
        vt = np.empty(self.N, dtype=np.int64)
        load().or_data_order(self.h, orig.ctypes.data, afo.ctypes.data, vt.ctypes.data)
        return orig, afo, vt

    def gid(self, l, i):
        return load().or_data_gid(self.h, int(l), int(i))

    def close(self):
        if getattr(self, "h", None):
            load().or_data_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# state ids as include/vps.h
_STATE = {"Freq": 1, "Q": 2, "Z": 3, "Ni": 4, "Mi": 5, "Alpha": 6, "Lambda": 7, "m": 8, "bufN1": 9,
          "bufNK1": 10, "lnL": 11, "MiSum": 12, "FreqM": 13, "rout": 14, "FreqA": 15, "Fst": 16,
          "loc": 17, "binaryq": 18, "lnltrace": 19}
_DT = {"Freq": np.float64, "Q": np.float64, "Z": np.uint16, "Ni": np.int32, "Mi": np.int64,
       "Alpha": np.float64, "Lambda": np.float64, "m": np.int32, "bufN1": np.float64, "bufNK1": np.float64,
       "lnL": np.float64, "MiSum": np.int64, "FreqM": np.float64, "rout": np.float64, "FreqA": np.float64,
       "Fst": np.float64, "loc": np.float64, "binaryq": np.int32, "lnltrace": np.float64}


class Chain:
    """Oracle chain with the same state names as vcfpop_amd.Chain."""

    def __init__(self, data, params, K, run_id=1):
        self.data = data
        self.K, self.N, self.KT = int(K), data.N, data.KT
        self.S = data._desc.n_pop if data._desc.n_pop > 0 else 1
        self.params = params
        self.h = load().or_chain_create(data.h, C.addressof(params), int(K), int(run_id))
        if not self.h:
            raise ValueError("the oracle keeps REAL = double for the F model")
        self._finished = False

    def _shape(self, name):
        K, N, KT, S = self.K, self.N, self.KT, self.S
        r, n = C.c_int64(), C.c_int64()
        load().or_chain_sizes(self.h, C.byref(r), C.byref(n))
        return {"Freq": (K, KT), "Q": (N, K), "Z": (N,), "Ni": (K, KT), "Mi": (N, K), "Alpha": (K,),
                "Lambda": (K,), "m": (1,), "bufN1": (N,), "bufNK1": (N, K), "lnL": (1,), "MiSum": (N, K),
                "FreqM": (K, KT), "rout": (r.value,), "FreqA": (3, KT), "Fst": (2, K),
                "loc": (1 + K + 3 * S * K + S,), "binaryq": (1,), "lnltrace": (n.value,)}[name]

    def get(self, name):
        a = np.empty(self._shape(name), dtype=_DT[name])
        rc = load().or_chain_get_state(self.h, _STATE[name], a.ctypes.data, a.nbytes)
        assert rc == 0, name
        if name == "MiSum" and self._finished:
            a = a.view(np.float64)
        return a

    def set(self, name, value):
        a = np.ascontiguousarray(np.asarray(value, dtype=_DT[name]).reshape(self._shape(name)))
        rc = load().or_chain_set_state(self.h, _STATE[name], a.ctypes.data, a.nbytes)
        assert rc == 0, name

    def seed(self):
        return load().or_chain_seed(self.h)

    def call(self, name):
        getattr(load(), "or_" + name)(self.h)

    def sweep(self, n=1):
        load().or_sweep(self.h, int(n))

    def record_pass(self):
        return load().or_record_pass(self.h)

    def metro_logratio(self, qprop):
        q = np.ascontiguousarray(qprop, dtype=np.float64)
        out = np.empty(self.N)
        load().or_q_metro_logratio(self.h, q.ctypes.data, out.ctypes.data)
        return out

    def noadmix_loglik(self):
        out = np.empty((self.N, self.K))
        load().or_q_noadmix_loglik(self.h, out.ctypes.data)
        return out

    def finish(self):
        load().or_arrange(self.h)
        self._finished = True

    def close(self):
        if getattr(self, "h", None):
            load().or_chain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
