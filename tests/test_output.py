"""Result tables (WriteStructure, src/structure.cpp:88-377) written from a
finished run's host state: sections, layout and number format, and that each
table carries the right quantity (final-Q membership per location, MiSum
ancestry, net nucleotide distance from FreqM).  Runs on the CPU: the run is
the oracle's, the genotype image a host-only dataset."""
import numpy as np

FMT = "%0.5f"


def run(vps, orc, tmp_path, **par):
    s = vps.Synth(n_ind=30, n_loc=60, k_true=3, alpha_q=0.5, missing=0.05, frac_multi=0.3,
                  ploidy=[(2, 0.6), (4, 0.4)], n_pop=3, seed=12)
    ds = vps.Dataset(s, -1)
    kw = dict(admix_val=1, nburnin=4, nreps=8, nthinning=2, writelnl_val=1, g_seed=2)
    kw.update(par)
    p = vps.params_default(**kw)
    o = orc.Chain(orc.Data(s.desc), p, 3, 7)
    o.sweep(kw["nburnin"] + kw["nreps"])
    o.finish()
    st = dict(Q=o.get("Q"), ancestry=o.get("MiSum"), FreqM=o.get("FreqM"), rout=o.get("rout"),
              lnltrace=o.get("lnltrace"), seed=o.seed())
    prefix = str(tmp_path / "out")
    summ = vps.write_tables_host(ds, p, 3, 7, st, prefix, rep=2, diversity=1)
    text = open(prefix + ".structure.k=3_rep=2_id=7.txt", newline="").read()
    return s, ds, p, st, text, summ, prefix


def section(lines, title):
    i = lines.index(title)
    out = []
    for ln in lines[i + 1:]:
        if ln == "":
            break
        out.append(ln)
    return out


def test_run_table_sections_and_values(vps, orc, tmp_path):
    s, ds, p, st, text, summ, prefix = run(vps, orc, tmp_path)
    assert "\r\n" in text and "\n\n" not in text.replace("\r\n", "\x00")
    lines = text.split("\r\n")
    assert f"Seed={st['seed']}" in lines
    assert "Model=ADM,NOLOC,NOF" in lines
    assert "Number of cluster=3" in lines and "Replicate=2" in lines and "Run id=7" in lines
    r = st["rout"]
    assert f"Mean value of ln likelihood={FMT % r[0]}" in lines
    assert f"Variance of ln likelihood={FMT % r[2]}" in lines
    assert f"Estimated Ln Prob of Data={FMT % r[3]}" in lines
    assert summ == {"id": 7, "k": 3, "rep": 2, "mean_lnl": r[0], "var_lnl": r[2], "lnpd": r[3]}
    # membership of each location from the final Q (sum, then Unify with MIN_FREQ)
    ind_pop = s.arrays()["ind_pop"]
    mem = section(lines, "Proportion of membership of each pre-defined population in each of the 3 clusters")
    assert mem[0] == "\tCluster" and mem[1] == "Pop\t1\t2\t3"
    for sidx, row in enumerate(mem[2:]):
        o = st["Q"][ind_pop == sidx].sum(0)
        o = (o + 1e-10) / (o.sum() + 3e-10)
        assert row == "\t".join([f"Pop{sidx + 1}"] + [FMT % v for v in o])
    # ancestry = normalised MiSum
    anc = section(lines, "Inferred ancestry of individuals")
    assert anc[0] == "\t\tCluster" and anc[1] == "Ind\tPop\tRegL1\t1\t2\t3"
    for i, row in enumerate(anc[2:]):
        assert row == "\t".join([f"Ind{i + 1}", f"Pop{ind_pop[i] + 1}", "Total"] + [FMT % v for v in st["ancestry"][i]])
    # net nucleotide distance from FreqM
    orig, afo, vt = ds.order()
    k2 = np.diff(np.append(afo, ds.KT))
    F = st["FreqM"]
    H = np.array([np.mean([1 - (F[k, afo[l]:afo[l] + k2[l]] ** 2).sum() for l in range(ds.L)]) for k in range(3)])
    D = np.array([[np.mean([1 - (F[i, afo[l]:afo[l] + k2[l]] * F[j, afo[l]:afo[l] + k2[l]]).sum() for l in range(ds.L)])
                   - (H[i] + H[j]) / 2 for j in range(3)] for i in range(3)])
    dist = section(lines, "Allele-frequency divergence among pops (Net nucleotide distance)")
    assert dist[0] == "\tCluster" and dist[1] == "Cluster\t1\t2\t3"
    for i, row in enumerate(dist[2:]):
        got = [float(x) for x in row.split("\t")[1:]]
        assert np.allclose(got, D[i], atol=6e-6)
    # diversity table: one Average row per cluster, then K rows per locus
    div = section(lines, "Estimated genetic diversity in each cluster")
    assert div[0].startswith("Locus\tCluster\tHeterozygosity\tAllele0")
    assert len(div) == 1 + 3 + 3 * ds.L
    assert div[1].startswith("Average\t1")


def test_lnl_trace_file(vps, orc, tmp_path):
    s, ds, p, st, text, summ, prefix = run(vps, orc, tmp_path)
    rows = open(prefix + ".structure.k=3_rep=2_id=7.lnl.txt", newline="").read().split("\r\n")
    assert rows[0] == "4" and rows[1] == "Iteration\tlnL"
    body = [r for r in rows[2:] if r]
    assert [int(r.split("\t")[0]) for r in body] == [0, 2, 4, 6, 8, 10]
    assert [r.split("\t")[1] for r in body] == [FMT % v for v in st["lnltrace"]]


def test_options_and_summary(vps, orc, tmp_path):
    s, ds, p, st, text, summ, prefix = run(vps, orc, tmp_path, locpriori_val=1)
    assert "Model=ADM,LOC,NOF" in text and "Mean value of r=" in text
    out = vps.write_tables_host(ds, p, 3, 7, st, str(tmp_path / "csv"), delimiter=",", linebreak="\n",
                                decimal=3, rep=1)
    t2 = open(str(tmp_path / "csv") + ".structure.k=3_rep=1_id=7.txt", newline="").read()
    assert "\r\n" not in t2 and "Ind,Pop,RegL1,1,2,3" in t2
    assert f"Mean value of ln likelihood={'%0.3f' % st['rout'][0]}" in t2
    vps.write_summary(tmp_path / "sum.txt", [summ, out])
    rows = open(tmp_path / "sum.txt", newline="").read().split("\r\n")
    assert rows[:3] == ["##Bayesian clustering calculated by vcfpop v1.08b", "", ""]
    assert rows[3] == "ID\tk\trep\tMean lnL\tVar lnL\tlnP(D)"
    assert rows[4].split("\t")[:3] == ["7", "3", "2"]
    assert len(rows) == 6 and rows[5].split("\t")[0] == "7"    # no line break after the last row


def test_summary_head(vps, tmp_path):
    """OpenResFile (src/file.cpp:269-287) and WriteParameters
    (src/parameters.cpp:1408-1418): title, Input, Output, local Time,
    Parameter = argv[0], then the later -g_* (not -g_indtext / -g_indtab),
    -structure* and, with -f, -f_* arguments"""
    import time as _time
    t = 1767225600 + 3 * 3600 + 25 * 60 + 7
    argv = ["-p=run.txt", "-g_input=a.txt", "-G_seed=3", "-g_indtext=p:#1-#2", "-g_indtab=x", "-structure",
            "-STRUCTURE_k=2", "-structures", "-f_qual=[1,2]", "-fst", "-gx", "-g_"]
    runs = [dict(id=1, k=2, rep=1, mean_lnl=-1.5, var_lnl=2.0, lnpd=-3.25)]
    lt = _time.localtime(t)
    stamp = "%04d-%02d-%02d %02d:%02d:%02d" % lt[:6]
    common = ["##Bayesian clustering calculated by vcfpop v1.08b", "#  Input: a.txt", "#  Output: out/x",
              "#  Time: " + stamp, "#  Parameter: -p=run.txt", "#    -g_input=a.txt", "#    -G_seed=3",
              "#    -structure", "#    -STRUCTURE_k=2", "#    -structures"]
    tail = ["#    -g_", "", "", "ID\tk\trep\tMean lnL\tVar lnL\tlnP(D)", "1\t2\t1\t-1.500\t2.000\t-3.250"]
    for filters, extra in ((False, []), (True, ["#    -f_qual=[1,2]"])):
        hdr = dict(input="a.txt", output="out/x", argv=argv, filters=filters, time=t)
        vps.write_summary(tmp_path / "s.txt", runs, header=hdr, decimal=3, linebreak="\n")
        rows = open(tmp_path / "s.txt", newline="").read().split("\n")
        assert rows == common + extra + tail


def test_ancestry_regions_from_indtext(vps, orc, tmp_path):
    """the RegL1 column (src/structure.cpp:283) carries each population's
    level-1 region from -g_indtext"""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data.polygene.txt")
    im = vps.Image(path)
    im.assign_indtext("w:#1-#40 x:#41-#60 y:#61-#100 z:#101-#120 #REG north:x,z south:w,y")
    assert im.names("pop") == ["x", "z", "w", "y"] and im.names("reg") == ["north", "north", "south", "south"]
    ds = vps.Dataset(im, -1)
    p = vps.params_default(admix_val=1, nburnin=2, nreps=4, nthinning=2, g_seed=4)
    o = orc.Chain(orc.Data(im.desc), p, 2, 1)
    o.sweep(6)
    o.finish()
    st = dict(Q=o.get("Q"), ancestry=o.get("MiSum"), FreqM=o.get("FreqM"), rout=o.get("rout"), seed=o.seed())
    vps.write_tables_host(ds, p, 2, 1, st, str(tmp_path / "r"), rep=1)
    lines = open(str(tmp_path / "r") + ".structure.k=2_rep=1_id=1.txt", newline="").read().split("\r\n")
    anc = section(lines, "Inferred ancestry of individuals")
    assert anc[1] == "Ind\tPop\tRegL1\t1\t2"
    want = ([(f"Ind{i}", "x", "north") for i in range(41, 61)] + [(f"Ind{i}", "z", "north") for i in range(101, 121)]
            + [(f"Ind{i}", "w", "south") for i in range(1, 41)] + [(f"Ind{i}", "y", "south") for i in range(61, 101)])
    assert [tuple(r.split("\t")[:3]) for r in anc[2:]] == want
    mem = section(lines, "Proportion of membership of each pre-defined population in each of the 2 clusters")
    assert [r.split("\t")[0] for r in mem[2:]] == ["x", "z", "w", "y"]
    ds.close()
    im.close()
