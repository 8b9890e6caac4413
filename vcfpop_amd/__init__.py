"""vcfpop_amd — MI355X (gfx950) engine for vcfpop's Bayesian clustering
(`-structure`).

The engine is the C-ABI library ``vcfpop_amd/libvps.so`` (declared in
``include/vps.h`` and ``include/vps_synth.h``); this package is a thin ctypes
mirror of it for tests and benchmarks.  There is no Python or CPU fallback:
importing a compute entry point without the built library raises.

Names follow the reference's BAYESIAN<REAL> (src/structure.h:50-307):
``Chain.update_z_admix`` is UpdateZAdmixCUDA, ``Chain.record`` is RecordCUDA,
and state arrays carry the reference's names (Freq, Q, Z, Ni, Mi, ...).
"""
from ._lib import (  # noqa: F401
    LIB_PATH,
    VpsError,
    load,
    params_default,
    VpsParams,
    VpsDatasetDesc,
    VpsSynthConfig,
    Synth,
    Image,
    Dataset,
    Chain,
    Batch,
    Server,
    output_options,
    write_tables_host,
    write_summary,
    STATE,
    RNG_REFERENCE,
    RNG_PARALLEL,
    PREC_FP64,
    PREC_MIXED,
    PREC_REAL32,
    PREC_FAST32,
    device_count,
    debug_jump,
    debug_jump96,
    debug_mt_accept,
    debug_copy,
)
from .workloads import WORKLOADS, make_synth  # noqa: F401
