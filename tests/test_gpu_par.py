"""The example parameter file end to end on the device: the polygene file
read into the packed layout, every (K, rep) run of the -structure options
as one batch, WriteStructure tables and the summary written."""
import os
import shutil

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


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
        text = f.read_text()
        assert "Model=ADM,LOC,F" in text and f"Number of cluster={s['k']}" in text
    rows = (tmp_path / "res" / "out.structure.txt").read_text().splitlines()
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
