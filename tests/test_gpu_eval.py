"""The -structure_eval analogue (tests/structure_eval.py): one sweep, pass
by pass, on the oracle, the FP64 device path and the mixed-precision path,
each chain on its own state from the same seed."""
import pytest

pytestmark = pytest.mark.gpu


def test_structure_eval_fp64_tracks_oracle(vps):
    import structure_eval as E
    s = vps.Synth(n_ind=120, n_loc=900, k_true=3, alpha_q=0.5, missing=0.03, frac_multi=0.2, min_alleles=3,
                  max_alleles=6, ploidy=[(2, 0.6), (4, 0.4)], seed=51)
    rows = E.run_eval(s, K=5)
    by = {(r["model"], r["pass"], r["backend"]): r for r in rows}
    assert {r["backend"] for r in rows} == {"oracle", "fp64", "mixed"}
    assert {r["pass"] for r in rows} >= {"init", "update_p", "update_z_admix", "update_z_noadmix", "record"}
    for (model, name, b), r in by.items():
        if b != "fp64":
            continue
        o = by[(model, name, "oracle")]
        for n in ("Z", "Ni", "Mi"):
            assert r[n] == o[n], (model, name, n)
        for n in ("d_Freq", "d_Q", "d_Alpha"):
            assert r[n] <= 1e-11, (model, name, n, r[n])
