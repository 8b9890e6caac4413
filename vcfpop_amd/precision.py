"""The error of the mixed-precision mode (FP32 per-genotype weights; FP64
lnL, Dirichlet draws and Metropolis steps) from one sweep of identical state.

Two chains, FP64 and mixed, take the same state of an FP64 run (P, Q, Z,
counts, alpha, lambda, location prior; same sweep index) and run one sweep:
the two Z passes see the same P, Q and random streams, so their assignments
differ only where a weight rounded to FP32 moves a cluster boundary past the
draw.  The next sweep's UpdateP and UpdateQ draw from the two count tables
with the same streams, so the FP32 error then reaches P, Q and the Record lnL.

The FP64 chain is the engine's FP64 path, which equals the CPU oracle
(Z, Ni, Mi bit for bit; P, Q, lnL to 1e-12).
"""
import numpy as np

from ._lib import PREC_FP64, PREC_MIXED, Chain, params_default

# chain state an FP64 run hands to both chains
FED = ("Freq", "Q", "Z", "Ni", "Mi", "Alpha", "Lambda", "loc", "Fst", "FreqA", "m")


def one_sweep_error(ds, params, K, warm=4, run_id=1):
    """Per-sweep FP32 error at one state after `warm` FP64 sweeps.

    Returns a dict: assignments that moved (at least sum|dMi|/2, a lower bound
    exact unless two moves in one individual cancel), their fraction of the
    typed allele copies, the same count from the allele table Ni, and after
    the next UpdateP/UpdateQ: max|dP|, max|dQ| and the relative difference of
    the Record lnL."""
    fields = {f: getattr(params, f) for f, _ in params._fields_}
    base = Chain(ds, params_default(**dict(fields, precision=PREC_FP64)), K, run_id)
    base.step(warm)
    base.sync()
    state = {n: base.get(n) for n in FED}
    base.close()
    res = {}
    for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
        c = Chain(ds, params_default(**dict(fields, precision=prec)), K, run_id)
        for n in FED:
            c.set(n, state[n])
        c.step(1)
        c.sync()
        r = {"Mi": c.get("Mi"), "Ni": c.get("Ni")}
        c.update_p()
        c.update_q_admix()
        c.record()
        r.update(Freq=c.get("Freq"), Q=c.get("Q"), lnL=float(c.get("lnL")[0]))
        c.close()
        res[name] = r
    a, b = res["fp64"], res["mixed"]
    copies = int(a["Mi"].sum())
    moved_mi = int(np.abs(a["Mi"] - b["Mi"]).sum()) // 2
    moved_ni = int(np.abs(a["Ni"].astype(np.int64) - b["Ni"]).sum()) // 2
    return {"after_fp64_sweeps": warm, "typed_allele_copies": copies,
            "moved_assignments_min": moved_mi, "frac_moved_assignments": moved_mi / max(copies, 1),
            "moved_by_allele_table_min": moved_ni,
            "max_abs_dP_next_sweep": float(np.abs(a["Freq"] - b["Freq"]).max()),
            "max_abs_dQ_next_sweep": float(np.abs(a["Q"] - b["Q"]).max()),
            "rel_dlnL_next_sweep": abs(a["lnL"] - b["lnL"]) / abs(a["lnL"])}
