"""ctypes bindings of libvps.so (include/vps.h, include/vps_synth.h)."""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libvps.so")

RNG_REFERENCE = 0
RNG_PARALLEL = 1
PREC_FP64 = 0
PREC_MIXED = 1
PREC_REAL32 = 2
PREC_FAST32 = 3

_lib = None


class VpsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vps error {code}: {msg}")
        self.code = code


class VpsDatasetDesc(C.Structure):
    _fields_ = [
        ("n_ind", C.c_int32),
        ("n_loc", C.c_int64),
        ("n_pop", C.c_int32),
        ("ind_pop", C.POINTER(C.c_int32)),
        ("geno_bits", C.POINTER(C.c_uint8)),
        ("geno_byte_offset", C.POINTER(C.c_uint64)),
        ("bitsize", C.POINTER(C.c_uint8)),
        ("n_allele", C.POINTER(C.c_int32)),
        ("gtab_offset", C.POINTER(C.c_int64)),
        ("gt_ploidy", C.POINTER(C.c_uint8)),
        ("gt_allele_offset", C.POINTER(C.c_int64)),
        ("allele_pool", C.POINTER(C.c_uint16)),
        ("ind_name", C.POINTER(C.c_char_p)),
        ("pop_name", C.POINTER(C.c_char_p)),
        ("loc_name", C.POINTER(C.c_char_p)),
        ("pop_region", C.POINTER(C.c_char_p)),
    ]


class VpsParams(C.Structure):
    _fields_ = [
        ("admix_val", C.c_int32), ("locpriori_val", C.c_int32), ("f_val", C.c_int32),
        ("nburnin", C.c_int32), ("nreps", C.c_int32), ("nthinning", C.c_int32),
        ("nruns", C.c_int32), ("nadmburnin", C.c_int32),
        ("lambda_", C.c_double), ("stdlambda", C.c_double), ("maxlambda", C.c_double),
        ("inferlambda_val", C.c_int32), ("difflambda_val", C.c_int32), ("diversity_val", C.c_int32),
        ("alpha", C.c_double), ("stdalpha", C.c_double), ("maxalpha", C.c_double),
        ("alphapriora", C.c_double), ("alphapriorb", C.c_double),
        ("inferalpha_val", C.c_int32), ("diffalpha_val", C.c_int32), ("uniformalpha_val", C.c_int32),
        ("metrofreq", C.c_int32),
        ("r", C.c_double), ("maxr", C.c_double), ("stdr", C.c_double), ("epseta", C.c_double),
        ("epsgamma", C.c_double),
        ("pmeanf", C.c_double), ("pstdf", C.c_double), ("stdf", C.c_double),
        ("singlef_val", C.c_int32),
        ("writelnl_val", C.c_int32),
        ("krange_min", C.c_int32), ("krange_max", C.c_int32),
        ("g_seed", C.c_uint64),
        ("rng_topology", C.c_int32),
        ("precision", C.c_int32),
        ("z_chunk", C.c_int32),
        ("sweep_graphs", C.c_int32),
    ]


class VpsResultHeader(C.Structure):
    _fields_ = [("input", C.c_char_p), ("output", C.c_char_p), ("argv", C.POINTER(C.c_char_p)), ("argc", C.c_int32),
                ("filters", C.c_int32), ("time", C.c_int64)]


class VpsOutputOptions(C.Structure):
    _fields_ = [("decimal", C.c_int32), ("scientific", C.c_int32), ("delimiter", C.c_char), ("pad", C.c_char * 7),
                ("linebreak", C.c_char_p), ("diversity", C.c_int32), ("rep", C.c_int32)]


class VpsRunSummary(C.Structure):
    _fields_ = [("id", C.c_int32), ("k", C.c_int32), ("rep", C.c_int32), ("pad", C.c_int32),
                ("mean_lnl", C.c_double), ("var_lnl", C.c_double), ("lnpd", C.c_double)]


class VpsRunState(C.Structure):
    _fields_ = [("Q", C.c_void_p), ("ancestry", C.c_void_p), ("FreqM", C.c_void_p), ("rout", C.c_void_p),
                ("rout_len", C.c_int64), ("lnltrace", C.c_void_p), ("lnltrace_len", C.c_int64), ("seed", C.c_uint64)]


class VpsServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_active", C.c_int32), ("hbm_budget_bytes", C.c_int64),
                ("slice_sweeps", C.c_int32), ("loaders", C.c_int32)]


class VpsJobInfo(C.Structure):
    _fields_ = [("job_id", C.c_int64), ("state", C.c_int32), ("status", C.c_int32), ("n_ind", C.c_int32),
                ("K", C.c_int32), ("n_loc", C.c_int64), ("sweeps_done", C.c_int64), ("sweeps_total", C.c_int64),
                ("device_bytes", C.c_int64), ("estimated_bytes", C.c_int64), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_end", C.c_double), ("rout_len", C.c_int64), ("lnltrace_len", C.c_int64)]


class VpsServerStats(C.Structure):
    _fields_ = [("pending", C.c_int64), ("loading", C.c_int64), ("running", C.c_int64), ("completed", C.c_int64),
                ("failed", C.c_int64), ("reserved_bytes", C.c_int64), ("peak_reserved_bytes", C.c_int64),
                ("uptime_s", C.c_double)]


class VpsSynthConfig(C.Structure):
    _fields_ = [
        ("n_ind", C.c_int32),
        ("n_loc", C.c_int64),
        ("k_true", C.c_int32),
        ("alpha_q", C.c_double),
        ("fst", C.c_double),
        ("missing", C.c_double),
        ("frac_multi", C.c_double),
        ("min_alleles", C.c_int32),
        ("max_alleles", C.c_int32),
        ("multi_dir", C.c_double),
        ("ploidy_levels", C.POINTER(C.c_int32)),
        ("ploidy_frac", C.POINTER(C.c_double)),
        ("n_ploidy", C.c_int32),
        ("n_pop", C.c_int32),
        ("seed", C.c_uint64),
    ]


class VpsDatasetInfo(C.Structure):
    _fields_ = [
        ("n_ind", C.c_int32), ("n_loc", C.c_int64), ("kt", C.c_int64),
        ("max_ploidy", C.c_int32), ("min_ploidy", C.c_int32), ("max_alleles", C.c_int32),
        ("n_pop", C.c_int32), ("n_tiles", C.c_int64), ("device_bytes", C.c_int64),
    ]


# exported symbols and their signatures (argtypes, restype)
_P = C.c_void_p
_SIGS = {
    "vps_abi_version": ([], C.c_int),
    "vps_last_error": ([], C.c_char_p),
    "vps_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "vps_params_default": ([C.POINTER(VpsParams)], None),
    "vps_dataset_create": ([C.POINTER(VpsDatasetDesc), C.c_int, C.POINTER(_P)], C.c_int),
    "vps_dataset_destroy": ([_P], C.c_int),
    "vps_dataset_get_info": ([_P, C.POINTER(VpsDatasetInfo)], C.c_int),
    "vps_dataset_get_order": ([_P, _P, _P, _P], C.c_int),
    "vps_chain_create": ([_P, C.POINTER(VpsParams), C.c_int, C.c_int, C.POINTER(_P)], C.c_int),
    "vps_chain_destroy": ([_P], C.c_int),
    "vps_chain_step": ([_P, C.c_int], C.c_int),
    "vps_chain_finish": ([_P], C.c_int),
    "vps_chain_sync": ([_P], C.c_int),
    "vps_chain_get_state": ([_P, C.c_int, _P, C.c_size_t], C.c_int),
    "vps_chain_set_state": ([_P, C.c_int, _P, C.c_size_t], C.c_int),
    "vps_chain_get_sizes": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "vps_record_lnl_noalloc": ([_P, C.c_int, _P, _P, _P, _P, _P], C.c_int),
    "vps_device_alloc": ([C.c_int, C.c_size_t, C.POINTER(_P)], C.c_int),
    "vps_device_free": ([_P], C.c_int),
    "vps_device_memset": ([_P, C.c_int, C.c_size_t], C.c_int),
    "vps_device_memcpy": ([_P, _P, C.c_size_t, C.c_int], C.c_int),
    "vps_device_synchronize": ([C.c_int], C.c_int),
    "vps_device_reset": ([C.c_int], C.c_int),
    "vps_device_name": ([C.c_int, C.c_char_p, C.c_size_t], C.c_int),
    "vps_pass_update_z_admix": ([_P], C.c_int),
    "vps_pass_update_z_noadmix": ([_P], C.c_int),
    "vps_pass_update_q_noadmix": ([_P], C.c_int),
    "vps_pass_update_q_metro": ([_P], C.c_int),
    "vps_pass_record": ([_P], C.c_int),
    "vps_pass_update_p": ([_P], C.c_int),
    "vps_pass_update_q_admix": ([_P], C.c_int),
    "vps_pass_update_alpha": ([_P], C.c_int),
    "vps_chain_enable_timing": ([_P, C.c_int], C.c_int),
    "vps_chain_kernel_time": ([_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)], C.c_int),
    "vps_server_config_default": ([C.POINTER(VpsServerConfig)], None),
    "vps_server_create": ([C.POINTER(VpsServerConfig), C.POINTER(_P)], C.c_int),
    "vps_server_submit": ([_P, C.POINTER(VpsDatasetDesc), C.POINTER(VpsParams), C.c_int, C.c_int, C.c_int64,
                           C.POINTER(C.c_int64)], C.c_int),
    "vps_server_wait": ([_P, C.c_int64, C.c_double], C.c_int),
    "vps_server_job_info": ([_P, C.c_int64, C.POINTER(VpsJobInfo)], C.c_int),
    "vps_server_job_result": ([_P, C.c_int64, C.c_int, _P, C.c_size_t], C.c_int),
    "vps_server_release": ([_P, C.c_int64], C.c_int),
    "vps_server_stats": ([_P, C.POINTER(VpsServerStats)], C.c_int),
    "vps_server_destroy": ([_P], C.c_int),
    "vps_output_options_default": ([C.POINTER(VpsOutputOptions)], None),
    "vps_chain_write_tables": ([_P, C.c_char_p, C.POINTER(VpsOutputOptions), C.POINTER(VpsRunSummary)], C.c_int),
    "vps_write_tables_host": ([_P, C.POINTER(VpsParams), C.c_int, C.c_int, C.POINTER(VpsRunState), C.c_char_p,
                               C.POINTER(VpsOutputOptions), C.POINTER(VpsRunSummary)], C.c_int),
    "vps_write_summary": ([C.c_char_p, C.POINTER(VpsRunSummary), C.c_int, C.POINTER(VpsOutputOptions),
                           C.POINTER(VpsResultHeader)], C.c_int),
    "vps_batch_create": ([C.POINTER(_P), C.c_int, C.POINTER(_P)], C.c_int),
    "vps_batch_step": ([_P, C.c_int], C.c_int),
    "vps_batch_sync": ([_P], C.c_int),
    "vps_batch_destroy": ([_P], C.c_int),
    "vps_batch_stats": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "vps_debug_decode": ([_P, C.c_int64, C.c_int32, C.POINTER(C.c_int32), _P], C.c_int),
    "vps_debug_jump": ([C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
    "vps_debug_jump96": ([C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)], C.c_int),
    "vps_debug_mt_accept": ([C.POINTER(C.c_double)] * 4 + [C.c_int64, C.POINTER(C.c_uint8), C.POINTER(C.c_int64)], C.c_int),
    "vps_debug_chunking": ([_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "vps_debug_gen_waves": ([_P, C.POINTER(C.c_int32)], C.c_int),
    "vps_debug_copy": ([C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "vps_synth_default": ([C.POINTER(VpsSynthConfig)], None),
    "vps_synth_create": ([C.POINTER(VpsSynthConfig), C.POINTER(_P)], C.c_int),
    "vps_synth_desc": ([_P, C.POINTER(VpsDatasetDesc)], C.c_int),
    "vps_synth_truth": ([_P, _P], C.c_int),
    "vps_synth_sizes": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "vps_synth_destroy": ([_P], C.c_int),
    "vps_image_read_polygene": ([C.c_char_p, C.POINTER(C.c_void_p)], C.c_int),
    "vps_image_read_vcf": ([C.c_char_p, C.POINTER(C.c_void_p)], C.c_int),
    "vps_image_desc": ([_P, C.POINTER(VpsDatasetDesc)], C.c_int),
    "vps_image_assign_indtext": ([_P, C.c_char_p], C.c_int),
    "vps_image_sizes": ([_P] + [C.POINTER(C.c_int64)] * 5, C.c_int),
    "vps_image_destroy": ([_P], C.c_int),
}


def load():
    """Load libvps.so (built by __graft_entry__.build / `make -C vcfpop_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VpsError(-1, f"{LIB_PATH} is not built; run `make -C vcfpop_amd` or __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise VpsError(rc, load().vps_last_error().decode(errors="replace"))


def device_count():
    n = C.c_int(0)
    rc = load().vps_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def params_default(**kw):
    p = VpsParams()
    load().vps_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, "lambda_" if k == "lambda" else k, v)
    return p


def debug_jump(D, x, y):
    xx, yy = C.c_uint64(x), C.c_uint64(y)
    _check(load().vps_debug_jump(C.c_uint64(D), C.byref(xx), C.byref(yy)))
    return xx.value, yy.value


def debug_jump96(D, x, y, z):
    xx, yy, zz = C.c_uint32(x), C.c_uint32(y), C.c_uint32(z)
    _check(load().vps_debug_jump96(C.c_uint64(D), C.byref(xx), C.byref(yy), C.byref(zz)))
    return xx.value, yy.value, zz.value


def debug_mt_accept(u, t, v, d):
    """the device Gamma draws' acceptance decision (host build of the same
    code): (accept flags, number decided by the exact double form)"""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, t, v, d)]
    n = len(arrs[0])
    out = np.zeros(n, dtype=np.uint8)
    ex = C.c_int64()
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    _check(load().vps_debug_mt_accept(*[ptr(a) for a in arrs], C.c_int64(n),
                                       out.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(ex)))
    return out.astype(bool), ex.value


def debug_copy(nbytes, iters=10, device=0):
    ms = C.c_double()
    _check(load().vps_debug_copy(int(device), int(nbytes), int(iters), C.byref(ms)))
    return ms.value


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Synth:
    """Synthetic dataset in vcfpop's packed layout (include/vps_synth.h)."""

    def __init__(self, **kw):
        lib = load()
        cfg = VpsSynthConfig()
        lib.vps_synth_default(C.byref(cfg))
        ploidy = kw.pop("ploidy", None)
        self._keep = []
        if ploidy:
            levels = np.array([p for p, _ in ploidy], dtype=np.int32)
            frac = np.array([f for _, f in ploidy], dtype=np.float64)
            self._keep += [levels, frac]
            cfg.ploidy_levels = _ptr(levels, C.c_int32)
            cfg.ploidy_frac = _ptr(frac, C.c_double)
            cfg.n_ploidy = len(levels)
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.cfg = cfg
        h = C.c_void_p()
        _check(lib.vps_synth_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.desc = VpsDatasetDesc()
        _check(lib.vps_synth_desc(h, C.byref(self.desc)))

    def truth(self):
        q = np.empty((self.cfg.n_ind, self.cfg.k_true), dtype=np.float64)
        _check(load().vps_synth_truth(self.h, q.ctypes.data))
        return q

    def arrays(self):
        """numpy copies of the packed description (for tests)."""
        d, L, N = self.desc, self.desc.n_loc, self.desc.n_ind
        pb, ng, na = C.c_int64(), C.c_int64(), C.c_int64()
        _check(load().vps_synth_sizes(self.h, C.byref(pb), C.byref(ng), C.byref(na)))
        out = dict(
            N=N, L=L, S=d.n_pop,
            ind_pop=np.ctypeslib.as_array(d.ind_pop, (N,)).copy(),
            geno_bits=np.ctypeslib.as_array(d.geno_bits, (pb.value,)).copy(),
            geno_byte_offset=np.ctypeslib.as_array(d.geno_byte_offset, (L,)).copy(),
            bitsize=np.ctypeslib.as_array(d.bitsize, (L,)).copy(),
            n_allele=np.ctypeslib.as_array(d.n_allele, (L,)).copy(),
            gtab_offset=np.ctypeslib.as_array(d.gtab_offset, (L + 1,)).copy(),
            gt_ploidy=np.ctypeslib.as_array(d.gt_ploidy, (ng.value,)).copy(),
            gt_allele_offset=np.ctypeslib.as_array(d.gt_allele_offset, (ng.value,)).copy(),
            allele_pool=np.ctypeslib.as_array(d.allele_pool, (max(na.value, 1),)).copy(),
        )
        return out

    def close(self):
        if getattr(self, "h", None):
            load().vps_synth_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _PackedArrays:
    """numpy copies of a packed description (for tests)."""

    def _arrays(self, pb, ng, na):
        d, L, N = self.desc, self.desc.n_loc, self.desc.n_ind
        return dict(
            N=N, L=L, S=d.n_pop,
            ind_pop=np.ctypeslib.as_array(d.ind_pop, (N,)).copy(),
            geno_bits=np.ctypeslib.as_array(d.geno_bits, (max(pb, 1),)).copy(),
            geno_byte_offset=np.ctypeslib.as_array(d.geno_byte_offset, (L,)).copy(),
            bitsize=np.ctypeslib.as_array(d.bitsize, (L,)).copy(),
            n_allele=np.ctypeslib.as_array(d.n_allele, (L,)).copy(),
            gtab_offset=np.ctypeslib.as_array(d.gtab_offset, (L + 1,)).copy(),
            gt_ploidy=np.ctypeslib.as_array(d.gt_ploidy, (ng,)).copy(),
            gt_allele_offset=np.ctypeslib.as_array(d.gt_allele_offset, (ng,)).copy(),
            allele_pool=np.ctypeslib.as_array(d.allele_pool, (max(na, 1),)).copy(),
        )


class Image(_PackedArrays):
    """A genotype file read into vcfpop's packed layout (include/vps_ingest.h)."""

    def __init__(self, path, fmt="polygene"):
        """fmt: "polygene", or "vcf" (plain or gzip/BGZF-compressed)"""
        readers = {"polygene": "vps_image_read_polygene", "vcf": "vps_image_read_vcf"}
        if fmt not in readers:
            raise ValueError(f"unsupported genotype format {fmt!r}")
        lib = load()
        h = C.c_void_p()
        _check(getattr(lib, readers[fmt])(os.fsencode(str(path)), C.byref(h)))
        self.h = h
        self.desc = VpsDatasetDesc()
        _check(lib.vps_image_desc(h, C.byref(self.desc)))
        v = [C.c_int64() for _ in range(5)]
        _check(lib.vps_image_sizes(h, *[C.byref(x) for x in v]))
        self.loci_read, self.loci_kept, self._pb, self._ng, self._na = (x.value for x in v)

    @property
    def N(self):
        return self.desc.n_ind

    @property
    def L(self):
        return self.desc.n_loc

    def assign_indtext(self, indtext):
        """populations and regions from vcfpop's -g_indtext (None: one
        population, "Total"); individuals move to population order"""
        lib = load()
        _check(lib.vps_image_assign_indtext(self.h, indtext.encode() if indtext else None))
        _check(lib.vps_image_desc(self.h, C.byref(self.desc)))

    def names(self, kind):
        """'ind', 'pop', 'loc' or 'reg' (each population's level-1 region) names"""
        d = self.desc
        if kind == "reg" and not d.pop_region:
            return ["Total"] * d.n_pop
        arr, n = {"ind": (d.ind_name, d.n_ind), "pop": (d.pop_name, d.n_pop), "loc": (d.loc_name, d.n_loc),
                  "reg": (d.pop_region, d.n_pop)}[kind]
        return [arr[i].decode() for i in range(n)]

    def arrays(self):
        return self._arrays(self._pb, self._ng, self._na)

    def close(self):
        if getattr(self, "h", None):
            load().vps_image_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def desc_from_arrays(a):
    """Build a VpsDatasetDesc over numpy arrays (the arrays must stay alive)."""
    d = VpsDatasetDesc()
    d.n_ind, d.n_loc, d.n_pop = int(a["N"]), int(a["L"]), int(a["S"])
    d.ind_pop = _ptr(a["ind_pop"], C.c_int32)
    d.geno_bits = _ptr(a["geno_bits"], C.c_uint8)
    d.geno_byte_offset = _ptr(a["geno_byte_offset"], C.c_uint64)
    d.bitsize = _ptr(a["bitsize"], C.c_uint8)
    d.n_allele = _ptr(a["n_allele"], C.c_int32)
    d.gtab_offset = _ptr(a["gtab_offset"], C.c_int64)
    d.gt_ploidy = _ptr(a["gt_ploidy"], C.c_uint8)
    d.gt_allele_offset = _ptr(a["gt_allele_offset"], C.c_int64)
    d.allele_pool = _ptr(a["allele_pool"], C.c_uint16)
    return d


class Dataset:
    """Device genotype image (vps_dataset); device=-1 builds it on the host only."""

    def __init__(self, desc, device=0):
        self._src = desc
        if isinstance(desc, (Synth, Image)):
            d = desc.desc
        elif isinstance(desc, dict):
            d = desc_from_arrays(desc)
        else:
            d = desc
        h = C.c_void_p()
        _check(load().vps_dataset_create(C.byref(d), int(device), C.byref(h)))
        self.h = h
        inf = VpsDatasetInfo()
        _check(load().vps_dataset_get_info(h, C.byref(inf)))
        self.N, self.L, self.KT = inf.n_ind, inf.n_loc, inf.kt
        self.maxploidy, self.minploidy, self.maxK = inf.max_ploidy, inf.min_ploidy, inf.max_alleles
        self.S, self.ntiles, self.device_bytes = inf.n_pop, inf.n_tiles, inf.device_bytes

    def order(self):
        orig = np.empty(self.L, dtype=np.int64)
        afo = np.empty(self.L, dtype=np.int64)
        vt = np.empty(self.N, dtype=np.int64)
        _check(load().vps_dataset_get_order(self.h, orig.ctypes.data, afo.ctypes.data, vt.ctypes.data))
        return orig, afo, vt

    def decode(self, l, i):
        v = C.c_int32()
        als = np.zeros(16, dtype=np.uint16)
        _check(load().vps_debug_decode(self.h, int(l), int(i), C.byref(v), als.ctypes.data))
        return als[: v.value].copy()

    def close(self):
        if getattr(self, "h", None):
            load().vps_dataset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# state id, dtype, shape(chain)
STATE = {
    "Freq": (1, np.float64, lambda c: (c.K, c.KT)),
    "Q": (2, np.float64, lambda c: (c.N, c.K)),
    "Z": (3, np.uint16, lambda c: (c.N,)),
    "Ni": (4, np.int32, lambda c: (c.K, c.KT)),
    "Mi": (5, np.int64, lambda c: (c.N, c.K)),
    "Alpha": (6, np.float64, lambda c: (c.K,)),
    "Lambda": (7, np.float64, lambda c: (c.K,)),
    "m": (8, np.int32, lambda c: (1,)),
    "bufN1": (9, np.float64, lambda c: (c.N,)),
    "bufNK1": (10, np.float64, lambda c: (c.N, c.K)),
    "lnL": (11, np.float64, lambda c: (1,)),
    "MiSum": (12, np.int64, lambda c: (c.N, c.K)),
    "FreqM": (13, np.float64, lambda c: (c.K, c.KT)),
    "rout": (14, np.float64, lambda c: (c.rlen(),)),
    "FreqA": (15, np.float64, lambda c: (3, c.KT)),
    "Fst": (16, np.float64, lambda c: (2, c.K)),
    "loc": (17, np.float64, lambda c: (1 + c.K + 3 * c.S * c.K + c.S,)),
    "binaryq": (18, np.int32, lambda c: (1,)),
    "lnltrace": (19, np.float64, lambda c: (c.nlnl(),)),
    "Qf": (20, np.float32, lambda c: (c.N, c.K)),
}


class Chain:
    """One MCMC run (BAYESIAN<double>) on the device."""

    def __init__(self, ds, params, K, run_id=1):
        self.ds = ds
        self.K, self.N, self.KT, self.S = int(K), ds.N, ds.KT, ds.S
        self.params = params
        h = C.c_void_p()
        _check(load().vps_chain_create(ds.h, C.byref(params), int(K), int(run_id), C.byref(h)))
        self.h = h

    def rlen(self):
        r, n = C.c_int64(), C.c_int64()
        _check(load().vps_chain_get_sizes(self.h, C.byref(r), C.byref(n)))
        return r.value

    def nlnl(self):
        r, n = C.c_int64(), C.c_int64()
        _check(load().vps_chain_get_sizes(self.h, C.byref(r), C.byref(n)))
        return n.value

    def get(self, name):
        sid, dt, shp = STATE[name]
        a = np.empty(shp(self), dtype=dt)
        _check(load().vps_chain_get_state(self.h, sid, a.ctypes.data, a.nbytes))
        if name == "MiSum" and getattr(self, "_finished", False):
            a = a.view(np.float64)
        return a

    def set(self, name, value):
        sid, dt, shp = STATE[name]
        a = np.ascontiguousarray(np.asarray(value, dtype=dt).reshape(shp(self)))
        _check(load().vps_chain_set_state(self.h, sid, a.ctypes.data, a.nbytes))

    def step(self, n=1):
        _check(load().vps_chain_step(self.h, int(n)))

    def sync(self):
        _check(load().vps_chain_sync(self.h))

    def finish(self):
        _check(load().vps_chain_finish(self.h))
        self._finished = True

    def update_z_admix(self):
        _check(load().vps_pass_update_z_admix(self.h))

    def update_z_noadmix(self):
        _check(load().vps_pass_update_z_noadmix(self.h))

    def update_q_noadmix(self):
        _check(load().vps_pass_update_q_noadmix(self.h))

    def update_q_metro(self):
        _check(load().vps_pass_update_q_metro(self.h))

    def update_q_admix(self):
        _check(load().vps_pass_update_q_admix(self.h))

    def update_p(self):
        _check(load().vps_pass_update_p(self.h))

    def update_alpha(self):
        _check(load().vps_pass_update_alpha(self.h))

    def record(self):
        _check(load().vps_pass_record(self.h))

    def enable_timing(self, on=True, every=1):
        """time the Z pass with HIP events: every n-th pass (0 / False = off)"""
        _check(load().vps_chain_enable_timing(self.h, int(every) if on else 0))

    def kernel_time(self, kernel=0):
        ms, n = C.c_double(), C.c_int64()
        _check(load().vps_chain_kernel_time(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def write_tables(self, prefix, **opts):
        """WriteStructure after finish(); returns the run's summary row"""
        o = output_options(**opts)
        sm = VpsRunSummary()
        _check(load().vps_chain_write_tables(self.h, str(prefix).encode(), C.byref(o), C.byref(sm)))
        return _summary_dict(sm)

    def chunking(self):
        """(G, Nc): individual chunks of the genotype passes"""
        g, nc = C.c_int32(), C.c_int32()
        _check(load().vps_debug_chunking(self.h, C.byref(g), C.byref(nc)))
        return g.value, nc.value

    def gen_waves(self):
        """wavefronts per block of the multiallelic Z-pass kernel (4 or 8)"""
        w = C.c_int32()
        _check(load().vps_debug_gen_waves(self.h, C.byref(w)))
        return w.value

    def close(self):
        if getattr(self, "h", None):
            load().vps_chain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Chains on one device advanced together, one graph launch per sweep
    (vps_batch_*).  Keeps references to its chains."""

    def __init__(self, chains):
        self.chains = list(chains)
        arr = (_P * len(self.chains))(*[c.h for c in self.chains])
        h = _P()
        _check(load().vps_batch_create(arr, len(self.chains), C.byref(h)))
        self.h = h

    def step(self, n=1):
        _check(load().vps_batch_step(self.h, int(n)))

    def sync(self):
        _check(load().vps_batch_sync(self.h))

    def stats(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(load().vps_batch_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"graph_launches": a.value, "captures": b.value, "eager_sweeps": c.value}

    def close(self):
        if getattr(self, "h", None):
            load().vps_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


JOB_STATES = {0: "pending", 1: "loading", 2: "running", 3: "done", 4: "failed"}


class Server:
    """Job server on one device (vps_server_*): HBM-budget admission and
    concurrent slices of every admitted job."""

    def __init__(self, device=0, hbm_budget_bytes=64 << 30, max_active=64, slice_sweeps=16, loaders=2):
        cfg = VpsServerConfig()
        load().vps_server_config_default(C.byref(cfg))
        cfg.device, cfg.hbm_budget_bytes, cfg.max_active, cfg.slice_sweeps = device, hbm_budget_bytes, max_active, slice_sweeps
        cfg.loaders = loaders
        h = _P()
        _check(load().vps_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._shapes = {}

    def submit(self, source, params, K, run_id=1, sweeps=None):
        """source: a Synth (or anything with .desc / n_ind); sweeps defaults to
        the params' burn-in + sampling length"""
        if sweeps is None:
            sweeps = params.nburnin + params.nreps + (params.nadmburnin if params.admix_val != 1 and params.locpriori_val != 1 else 0)
        jid = C.c_int64()
        _check(load().vps_server_submit(self.h, C.byref(source.desc), C.byref(params), int(K), int(run_id),
                                        int(sweeps), C.byref(jid)))
        self._shapes[jid.value] = (source.desc.n_ind, int(K))
        return jid.value

    def wait(self, job, timeout=-1.0):
        _check(load().vps_server_wait(self.h, int(job), float(timeout)))

    def info(self, job):
        inf = VpsJobInfo()
        _check(load().vps_server_job_info(self.h, int(job), C.byref(inf)))
        d = {f: getattr(inf, f) for f, _ in VpsJobInfo._fields_}
        d["state"] = JOB_STATES[d["state"]]
        return d

    def result(self, job, name):
        """final Q, MiSum (normalised ancestry), rout, lnltrace or lnL of a done job"""
        N, K = self._shapes[int(job)]
        if name == "lnL":
            a = np.empty(1)
        elif name in ("Q", "MiSum"):
            a = np.empty((N, K))
        elif name in ("rout", "lnltrace"):
            a = np.empty(self.info(job)[name + "_len" if name == "rout" else "lnltrace_len"])
        else:
            raise VpsError(f"the server keeps no {name}")
        _check(load().vps_server_job_result(self.h, int(job), STATE[name][0], a.ctypes.data, a.nbytes))
        return a

    def release(self, job):
        _check(load().vps_server_release(self.h, int(job)))
        self._shapes.pop(int(job), None)

    def stats(self):
        st = VpsServerStats()
        _check(load().vps_server_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in VpsServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            load().vps_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def output_options(**kw):
    """vps_output_options with the reference defaults, fields overridden by kw
    (linebreak as str, delimiter as a one-character str)"""
    o = VpsOutputOptions()
    load().vps_output_options_default(C.byref(o))
    for k, v in kw.items():
        if k == "linebreak":
            o._lb = v.encode()
            o.linebreak = o._lb
        elif k == "delimiter":
            o.delimiter = v.encode()
        else:
            setattr(o, k, v)
    return o


def _summary_dict(sm):
    return {"id": sm.id, "k": sm.k, "rep": sm.rep, "mean_lnl": sm.mean_lnl, "var_lnl": sm.var_lnl, "lnpd": sm.lnpd}


def write_tables_host(ds, params, K, run_id, state, prefix, **opts):
    """WriteStructure from host arrays: state = dict(Q, ancestry, FreqM, rout,
    lnltrace (optional), seed)"""
    keep = {n: np.ascontiguousarray(state[n], dtype=np.float64) for n in ("Q", "ancestry", "FreqM", "rout")}
    tr = state.get("lnltrace")
    if tr is not None:
        keep["lnltrace"] = np.ascontiguousarray(tr, dtype=np.float64)
    st = VpsRunState(keep["Q"].ctypes.data, keep["ancestry"].ctypes.data, keep["FreqM"].ctypes.data,
                     keep["rout"].ctypes.data, keep["rout"].size,
                     keep["lnltrace"].ctypes.data if tr is not None else None,
                     keep["lnltrace"].size if tr is not None else 0, int(state["seed"]))
    o = output_options(**opts)
    sm = VpsRunSummary()
    _check(load().vps_write_tables_host(ds.h, C.byref(params), int(K), int(run_id), C.byref(st),
                                        str(prefix).encode(), C.byref(o), C.byref(sm)))
    return _summary_dict(sm)


def write_summary(path, summaries, header=None, **opts):
    """WriteStructureSummary over the runs' summary dicts; header = dict(input,
    output, argv (list of str), filters (bool), time (epoch seconds, 0 = now))
    for the result-file head, or None for the title line only"""
    arr = (VpsRunSummary * max(1, len(summaries)))()
    for i, d in enumerate(summaries):
        arr[i] = VpsRunSummary(d["id"], d["k"], d["rep"], 0, d["mean_lnl"], d["var_lnl"], d["lnpd"])
    o = output_options(**opts)
    h = None
    if header is not None:
        argv = [a.encode() for a in header.get("argv", [])]
        av = (C.c_char_p * max(1, len(argv)))(*argv)
        h = VpsResultHeader(header.get("input", "").encode(), header.get("output", "").encode(), av, len(argv),
                            1 if header.get("filters") else 0, int(header.get("time", 0)))
    _check(load().vps_write_summary(str(path).encode(), arr, len(summaries), C.byref(o),
                                    C.byref(h) if h is not None else None))
