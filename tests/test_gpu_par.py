"""The example parameter file end to end on the device: the polygene file
read into the packed layout with the -g_indtext populations and regions,
every (K, rep) run of the -structure options as one batch, WriteStructure
tables and the summary (with vcfpop's result-file head) written; and the
first runs equal the oracle's runs on the same ingested image."""
import os
import shutil
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PAR = os.path.join(HERE, "golden", "par.polygene.txt")
RTOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_example_par_first_runs_equal_oracle(vps, orc):
    """runs 1-5 (K = 2 replicates 1-4, K = 3 replicate 1) of the example's
    model (admixture, location prior, F model, metrofreq 10) for 40 sweeps,
    batched on the device as the runner batches them, against the oracle:
    Z, Ni, Mi and the ancestry (MiSum) bit for bit; Q, P and the run
    summary to 1e-12"""
    from vcfpop_amd import par
    job = par.Job(PAR, sweeps=40)
    ds = vps.Dataset(job.image, 0)
    od = orc.Data(job.image.desc)
    sel = list(enumerate(job.runs))[:5]
    say("example par: runs", [i + 1 for i, _ in sel], "x", job.total, "sweeps, reference topology")
    chains = [vps.Chain(ds, job.params, K, i + 1) for i, (K, rep) in sel]
    b = vps.Batch(chains)
    b.step(job.total)
    b.sync()
    for c, (i, (K, rep)) in zip(chains, sel):
        o = orc.Chain(od, job.params, K, i + 1)
        o.sweep(job.total)
        for n in ("Z", "Ni", "Mi"):
            assert (c.get(n) == o.get(n)).all(), (i + 1, n)
        assert np.allclose(c.get("Q"), o.get("Q"), rtol=RTOL, atol=0), i + 1
        assert np.allclose(c.get("Freq"), o.get("Freq"), rtol=RTOL, atol=0), i + 1
        c.finish()
        o.finish()
        assert (c.get("MiSum") == o.get("MiSum")).all(), i + 1
        ro, rc = o.get("rout"), c.get("rout")
        tol = RTOL * np.abs(ro)
        tol[2] = tol[3] = RTOL * abs(ro[1])
        assert (np.abs(rc - ro) <= tol).all(), (i + 1, rc, ro)
        say("run", i + 1, f"(K={K}, rep={rep}) equals the oracle")
        o.close()
    b.close()
    for c in chains:
        c.close()
    ds.close()
    job.close()


def test_example_par_runs(vps, tmp_path):
    from vcfpop_amd import par
    shutil.copy(os.path.join(HERE, "golden", "data.polygene.txt"), tmp_path / "data.polygene.txt")
    text = open(os.path.join(HERE, "golden", "par.polygene.txt")).read().replace("polygene\\out", "res\\out")
    (tmp_path / "par.txt").write_text(text)
    print("example par: K 2..6 x 4 runs, 60 sweeps", flush=True)
    summ = par.run(str(tmp_path / "par.txt"), device=0, sweeps=60, log=lambda *a: None)
    assert [(s["id"], s["k"], s["rep"]) for s in summ] == [(i + 1, 2 + i // 4, 1 + i % 4) for i in range(20)]
    for s in summ:
        f = tmp_path / "res" / f"out.structure.k={s['k']}_rep={s['rep']}_id={s['id']}.txt"
        assert f.exists(), f
        text = f.read_bytes().decode()
        assert "Model=ADM,LOC,F" in text and f"Number of cluster={s['k']}" in text
        anc = text[text.index("Inferred ancestry of individuals"):].split("\r\n")
        assert anc[2].split("\t")[:3] == ["Ind", "Pop", "RegL1"]
        assert anc[3].split("\t")[:3] == ["Ind1", "pop1", "reg1"] and anc[122].split("\t")[:3] == ["Ind120", "pop4", "reg2"]
    raw = (tmp_path / "res" / "out.structure.txt").read_bytes().decode()
    assert not raw.endswith("\r\n")
    rows = raw.split("\r\n")
    assert rows[0] == "##Bayesian clustering calculated by vcfpop v1.08b"
    assert rows[1:3] == ["#  Input: data.polygene.txt", "#  Output: res/out"] and rows[3].startswith("#  Time: ")
    assert rows[4] == "#  Parameter: -g_output=res\\out" and rows[5] == "#    -g_decimal=6"
    assert "#    -structure_krange=[2,6]" in rows and not any("indtext" in r for r in rows)
    assert sum(1 for r in rows if r.split("\t")[0].isdigit()) == 20


def test_example_par_eval_mode(vps, tmp_path):
    """-structure_eval=yes: K = 5, one sweep, both models, every pass on the
    FP64 and mixed device paths (vcfpop's evaluation mode)"""
    from vcfpop_amd import par
    shutil.copy(os.path.join(HERE, "golden", "data.polygene.txt"), tmp_path / "data.polygene.txt")
    text = open(os.path.join(HERE, "golden", "par.polygene.txt")).read().replace("polygene\\out", "res\\out")
    (tmp_path / "par.txt").write_text(text + "\n-structure_eval=yes\n")
    print("example par, eval mode", flush=True)
    lines = []
    rows = par.run(str(tmp_path / "par.txt"), device=0, log=lines.append)
    assert {r["backend"] for r in rows} == {"fp64", "mixed"}
    assert {(r["model"], r["pass"]) for r in rows} >= {("admix", "update_z_admix"), ("noadmix", "update_z_noadmix"),
                                                      ("admix", "record"), ("noadmix", "update_q_noadmix")}
    by = {(r["model"], r["pass"], r["backend"]): r for r in rows}
    for model in ("admix", "noadmix"):    # the same seed: initialisation agrees on both paths
        assert by[(model, "init", "fp64")]["Mi"] == by[(model, "init", "mixed")]["Mi"]
    assert any("update_z_admix" in ln for ln in lines)
