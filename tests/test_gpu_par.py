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


def _write_vcf(path, n_ind=90, n_loc=400, seed=3):
    """a seeded VCF: diploid and tetraploid samples, 2-4 alleles per variant
    with cluster-dependent frequencies, some missing genotypes, some phased"""
    rng = np.random.default_rng(seed)
    ploidy = np.where(rng.random(n_ind) < 0.8, 2, 4)
    q = rng.dirichlet([0.5, 0.5, 0.5], n_ind)
    rows = ["##fileformat=VCFv4.2",
            "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(f"S{i + 1}" for i in range(n_ind))]
    for l in range(n_loc):
        k = int(rng.choice([2, 2, 2, 3, 4])) if l % 50 else 2
        p = rng.dirichlet(np.full(k, 0.7), 3)
        alt = ",".join("ACGT"[x] for x in range(1, k)) if l % 50 else "."   # an ALT of '.' is one allele
        f = []
        for i in range(n_ind):
            v = int(ploidy[i])
            if rng.random() < 0.03:
                f.append("/".join(["."] * v))
                continue
            mix = q[i] @ p
            al = rng.choice(k, size=v, p=mix / mix.sum())
            sep = "|" if rng.random() < 0.2 else "/"
            f.append(sep.join(str(int(a)) for a in al) + ":7")
        rows.append(f"{1 + l // 200}\t{100 + l}\t{'.' if l % 7 == 0 else f'v{l}'}\tA\t{alt}\t.\tPASS\t.\tGT:DP\t" + "\t".join(f))
    path.write_text("\n".join(rows) + "\n")


def test_vcf_par_runs_equal_oracle(vps, orc, tmp_path):
    """a VCF through a vcfpop parameter file (-g_format=vcf, populations from
    -g_indtext): the image read by the VCF reader, runs 1-3 (K = 2 and 3)
    batched on the device for 30 sweeps, against the oracle on the same
    image: Z, Ni, Mi and the ancestry bit for bit; Q and P to 1e-12"""
    from vcfpop_amd import par
    _write_vcf(tmp_path / "data.vcf")
    (tmp_path / "par.txt").write_text(
        "-g_input=data.vcf\n-g_format=vcf\n-g_output=res/out\n-g_seed=7\n"
        "-g_indtext=\"popA:#1-#30 popB:#31-#90\"\n"
        "-structure\n-structure_admix=yes\n-structure_krange=[2,3]\n-structure_nruns=2\n")
    job = par.Job(str(tmp_path / "par.txt"), sweeps=30)
    assert job.image.loci_read == 400 and job.image.names("pop") == ["popA", "popB"]
    ds = vps.Dataset(job.image, 0)
    od = orc.Data(job.image.desc)
    sel = list(enumerate(job.runs))[:3]
    say("vcf par: runs", [i + 1 for i, _ in sel], "x", job.total, "sweeps; loci kept", job.image.loci_kept)
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
        say("vcf run", i + 1, f"(K={K}, rep={rep}) equals the oracle")
        o.close()
    b.close()
    for c in chains:
        c.close()
    ds.close()
    job.close()
