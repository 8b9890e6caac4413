"""Genotype text read into the packed layout (include/vps_ingest.h): the
polygene reader (IND::polygene, src/load.cpp:3047-3120), genotype ids in
order of first appearance with each ploidy's missing genotype registered
first, allele indices in order of first appearance over the genotype table
(IndexAlleleLength, src/load.cpp:1535-1590), loci with fewer than two
alleles dropped, CeilLog2 packing.  Runs on the CPU (host-only datasets)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLE = os.path.join(HERE, "golden", "data.polygene.txt")

SMALL = ("ID\tPop\tPloidy\tA\tB\tC\n"
         "i1\tp1\t2\t3,1\t7,7\t5,5\n"
         "i2\tp1\t2\t\t7,9\t5,5\n"
         "i3\tp2\t4\t1,1,3,3\t9,9,9,7\t5,5,5,5\n"
         "i4\tp2\t2\t1,3\t \t5,5\r\n")


def unpack_ids(a, l, N):
    b = int(a["bitsize"][l])
    base = int(a["geno_byte_offset"][l]) * 8
    out = []
    for i in range(N):
        v = 0
        for x in range(b):
            bit = base + i * b + x
            v |= ((int(a["geno_bits"][bit >> 3]) >> (bit & 7)) & 1) << x
        out.append(v)
    return out


def table(a, l):
    g0, g1 = int(a["gtab_offset"][l]), int(a["gtab_offset"][l + 1])
    out = []
    for g in range(g0, g1):
        v, s = int(a["gt_ploidy"][g]), int(a["gt_allele_offset"][g])
        out.append(tuple(int(x) for x in a["allele_pool"][s:s + v]) if v else "missing")
    return out


def test_known_answer(vps, tmp_path):
    p = tmp_path / "small.txt"
    p.write_text(SMALL)
    im = vps.Image(p)
    assert (im.N, im.L, im.loci_read, im.loci_kept) == (4, 2, 3, 2)
    assert im.names("loc") == ["A", "B"] and im.names("pop") == ["p1", "p2"]
    assert im.names("ind") == ["i1", "i2", "i3", "i4"]
    a = im.arrays()
    assert list(a["ind_pop"]) == [0, 0, 1, 1]
    assert list(a["n_allele"]) == [2, 2]
    assert list(a["bitsize"]) == [2, 3]
    # locus A: missing(2), (1,3), missing(4), (1,1,3,3); alleles 1 -> 0, 3 -> 1
    assert table(a, 0) == ["missing", (0, 1), "missing", (0, 0, 1, 1)]
    assert list(a["gt_ploidy"][:4]) == [0, 2, 0, 4]
    assert unpack_ids(a, 0, 4) == [1, 0, 3, 1]
    # locus B: missing(2), (7,7), (7,9), missing(4), (7,9,9,9); 7 -> 0, 9 -> 1
    assert table(a, 1) == ["missing", (0, 0), (0, 1), "missing", (0, 1, 1, 1)]
    assert unpack_ids(a, 1, 4) == [1, 2, 4, 0]
    # the device image decodes the same genotypes
    ds = vps.Dataset(im, -1)
    orig, afo, vt = ds.order()
    assert list(vt) == [4, 2, 8, 2]
    got = {int(orig[l]): [list(ds.decode(l, i)) for i in range(4)] for l in range(ds.L)}
    assert got[0] == [[0, 1], [], [0, 0, 1, 1], [0, 1]]
    assert got[1] == [[0, 0], [0, 1], [0, 1, 1, 1], []]


def restate(text):
    """the reader's rules, restated independently: per locus, tables and ids"""
    rows = [r.split("\t") for r in text.replace("\r", "").split("\n") if r.strip()]
    L = len(rows[0]) - 3
    out = []
    for l in range(L):
        keys, tab, ids = {}, [], []
        for r in rows[1:]:
            v, f = int(r[2]), r[3 + l].strip()
            g = ("m", v) if not f else ("t", v, tuple(sorted(int(x) for x in f.split(","))))
            for e in ([("m", v)] if g[0] == "m" else [("m", v), g]):
                if e not in keys:
                    keys[e] = len(tab)
                    tab.append(e)
            ids.append(keys[g])
        amap = {}
        for e in tab:
            for x in (e[2] if e[0] == "t" else ()):
                amap.setdefault(x, len(amap))
        if len(amap) < 2:
            continue
        out.append((len(amap), ["missing" if e[0] == "m" else tuple(sorted(amap[x] for x in e[2])) for e in tab], ids))
    return out


def test_example_file_matches_restatement(vps):
    im = vps.Image(EXAMPLE)
    assert (im.N, im.L, im.loci_read) == (120, 10, 10)
    a = im.arrays()
    assert list(np.bincount(a["ind_pop"])) == [30, 30, 30, 30]
    exp = restate(open(EXAMPLE).read())
    assert len(exp) == im.L
    for l, (k, tab, ids) in enumerate(exp):
        assert a["n_allele"][l] == k
        assert table(a, l) == tab
        assert a["bitsize"][l] == max(1, int(np.ceil(np.log2(len(tab)))))
        assert unpack_ids(a, l, im.N) == ids


@pytest.mark.parametrize("body,msg", [
    ("x\tp\t2\t1,2,3\n", "alleles for ploidy"),
    ("x\tp\t11\t1,2\n", "ploidy"),
    ("x\tp\t2\t1,a\n", "not an integer"),
    ("x\tp\t2\n", "fields"),
])
def test_reader_errors(vps, tmp_path, body, msg):
    p = tmp_path / "bad.txt"
    p.write_text("ID\tPop\tPloidy\tA\n" + body)
    with pytest.raises(vps.VpsError, match=msg):
        vps.Image(p)


# --- populations and regions from -g_indtext (vcfpop_amd/csrc/vps_ingest.cpp,
# vcfpop's Initialize src/function.cpp:4352-4500, AssignInds
# src/load.cpp:88-160, AssignVInds/SetVReg src/filter.cpp:166-398)

SIX = ("ID\tPop\tPloidy\tA\tB\n"
       "a1\tx\t2\t1,2\t1,1\n"
       "a2\tx\t2\t2,2\t1,3\n"
       "a3\tx\t2\t1,1\t3,3\n"
       "a4\ty\t2\t3,3\t1,3\n"
       "a5\ty\t2\t1,3\t\n"
       "a6\ty\t2\t2,3\t3,3\n")


def six(vps, tmp_path):
    p = tmp_path / "six.txt"
    p.write_text(SIX)
    return vps.Image(str(p))


def test_indtext_default_is_one_population(vps, tmp_path):
    im = six(vps, tmp_path)
    assert im.names("pop") == ["x", "y"]      # the population column, until -g_indtext is applied
    im.assign_indtext(None)
    assert im.desc.n_pop == 1 and im.names("pop") == ["Total"] and im.names("reg") == ["Total"]
    assert list(im.arrays()["ind_pop"]) == [0] * 6
    im.assign_indtext("")
    assert im.names("pop") == ["Total"]


def test_indtext_orders_individuals_by_region_walk(vps, tmp_path):
    im = six(vps, tmp_path)
    before = im.arrays()
    ids0 = [unpack_ids(before, l, 6) for l in range(2)]
    # pB holds a2 (by name) and a5..a6 (a range), pA holds #1 and #4;
    # a3 is in no population (DefPop).  Regions in parse order: r2 = {pB}
    # then r1 = {pA}; DefPop falls in DefRegL1, the first region walked.
    im.assign_indtext("pA:#1,#4\npB:a2,#5-#6 #REG r2:#2 r1:pA")
    assert im.names("pop") == ["DefPop", "pB", "pA"]
    assert im.names("reg") == ["DefRegL1", "r2", "r1"]
    assert im.names("ind") == ["a3", "a2", "a5", "a6", "a1", "a4"]
    a = im.arrays()
    assert list(a["ind_pop"]) == [0, 1, 1, 1, 2, 2]
    order = [2, 1, 4, 5, 0, 3]
    for l in range(2):    # genotype tables unchanged, ids moved with the individuals
        assert unpack_ids(a, l, 6) == [ids0[l][i] for i in order]
        assert table(a, l) == table(before, l)


def test_indtext_single_region_reads_total(vps, tmp_path):
    im = six(vps, tmp_path)
    im.assign_indtext("p1:#1-#3 p2:#4-#6 #REG only:#1-#2")
    assert im.names("pop") == ["p1", "p2"] and im.names("reg") == ["Total", "Total"]
    im.assign_indtext("p1:#1-#3 p2:#4-#6")
    assert im.names("reg") == ["Total", "Total"]
    assert im.names("ind") == ["a1", "a2", "a3", "a4", "a5", "a6"]


def test_indtext_two_region_levels(vps, tmp_path):
    im = six(vps, tmp_path)
    # level 1: ra = {p3}, rb = {p1, p2}; level 2: big = {rb, ra} in that
    # order by index range, so the walk visits ra (index 1) before rb
    im.assign_indtext("p1:#1-#2 p2:#3-#4 p3:#5-#6 #REG ra:p3 rb:#1-#2 #REG big:#1-#2 other:")
    assert im.names("pop") == ["p3", "p1", "p2"]
    assert im.names("reg") == ["ra", "rb", "rb"]
    assert im.names("ind") == ["a5", "a6", "a1", "a2", "a3", "a4"]


@pytest.mark.parametrize("text,msg", [
    ("p:#1-#3 p:#4", "same name"),
    ("p:#1-#3 q:#3-#4", "same individual"),
    ("p:#1-#7", "exceeds range"),
    ("p:zz", "not in the genotype data"),
    ("p:#9", "cannot find individual"),
    ("p:#1 q:#2 #REG r:#1-#3", "exceeds range"),
    ("p:#1 q:#2 #REG r:#1 s:#1", "the same member"),
    ("p:#1 q:#2 #REG r:#1,p", "twice"),
    ("p:#1 q:#2 #REG p:#2", "repeat population/region name"),
    ("p:#1 q:#2 #REG r:nope", "cannot find nope"),
    ("p:#1 q", "has no ':'"),
])
def test_indtext_errors(vps, tmp_path, text, msg):
    im = six(vps, tmp_path)
    with pytest.raises(vps.VpsError, match=msg):
        im.assign_indtext(text)
