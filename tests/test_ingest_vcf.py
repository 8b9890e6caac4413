"""The VCF reader (include/vps_ingest.h): LoadVCF and IND::AddVCFGenotype
(src/load.cpp:2111-2330, 3495-3560), the VCF LOCUS (src/function.cpp:
3166-3240), ReadVCFGenoString and CountPloidy (src/string2.cpp:41-60,
src/load.h:86-93), and the monomorphic-locus filter every input passes
(src/filter.cpp:100-133).  Known answers counted by hand; runs on the CPU."""
import gzip

import pytest

from test_ingest import table, unpack_ids

# fields are tab-separated; the header's ## lines are skipped
VCF = "\n".join([
    "##fileformat=VCFv4.2",
    "##contig=<ID=1>",
    "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\ts1\ts2\ts3\ts4\ts5",
    # biallelic; a phased heterozygote is the same genotype as 0/1
    "1\t100\trs1\tA\tG\t.\tPASS\t.\tGT\t0/0\t0/1\t./.\t1/1\t0|1",
    # three alleles, no ID (named POS_REF_ALT), a haploid missing '.'
    "1\t200\t.\tC\tT,G\t50\tPASS\tDP=9\tGT:DP\t0/2:5\t./.:3\t0/0\t2/2:1\t.:7",
    # only the reference allele is typed: dropped
    "1\t300\trs3\tG\tA\t.\tPASS\t.\tGT\t0/0\t0/0\t./.\t0/0\t0/0",
    # ALT '.' still counts as one allele (k = 2); tetraploid and haploid samples
    "2\t50\trs4\tT\t.\t.\tPASS\t.\tGT\t0/1\t0/0\t0/0/1/1\t./././.\t1",
    # four alleles; a '.' allele makes the genotype missing; copies sorted
    "2\t60\trs5\tA\tC,T,G\t.\tPASS\t.\tGT:AD\t0/.:1,0\t3/1\t1/3\t./.:2,2\t2/2",
]) + "\n"


def check(im):
    assert (im.N, im.L, im.loci_read, im.loci_kept) == (5, 4, 5, 4)
    assert im.names("ind") == ["s1", "s2", "s3", "s4", "s5"]
    assert im.names("loc") == ["rs1", "200_C_T,G", "rs4", "rs5"]
    assert im.names("pop") == ["Total"]
    a = im.arrays()
    assert list(a["ind_pop"]) == [0] * 5
    assert list(a["n_allele"]) == [2, 3, 2, 4]
    assert list(a["bitsize"]) == [2, 3, 3, 2]
    assert table(a, 0) == ["missing", (0, 0), (0, 1), (1, 1)]
    assert unpack_ids(a, 0, 5) == [1, 2, 0, 3, 2]
    assert table(a, 1) == ["missing", (0, 2), (0, 0), (2, 2), "missing"]
    assert unpack_ids(a, 1, 5) == [1, 0, 2, 3, 4]
    assert table(a, 2) == ["missing", (0, 1), (0, 0), "missing", (0, 0, 1, 1), "missing", (1,)]
    assert unpack_ids(a, 2, 5) == [1, 2, 4, 3, 6]
    assert table(a, 3) == ["missing", (1, 3), (2, 2)]
    assert unpack_ids(a, 3, 5) == [0, 1, 1, 0, 2]
    return a


def test_known_answer(vps, tmp_path):
    p = tmp_path / "small.vcf"
    p.write_text(VCF)
    im = vps.Image(p, "vcf")
    check(im)
    # the device image decodes the same genotypes (host-only dataset)
    ds = vps.Dataset(im, -1)
    orig, afo, vt = ds.order()
    assert list(vt) == [6, 6, 8, 4, 5]
    got = {int(orig[l]): [list(ds.decode(l, i)) for i in range(5)] for l in range(ds.L)}
    assert got[0] == [[0, 0], [0, 1], [], [1, 1], [0, 1]]
    assert got[2] == [[0, 1], [0, 0], [0, 0, 1, 1], [], [1]]
    assert got[3] == [[], [1, 3], [1, 3], [], [2, 2]]


def test_gzip_and_crlf(vps, tmp_path):
    p = tmp_path / "small.vcf.gz"
    with gzip.open(p, "wt", newline="") as f:
        f.write(VCF.replace("\n", "\r\n"))
    check(vps.Image(p, "vcf"))


def test_populations_from_indtext(vps, tmp_path):
    p = tmp_path / "small.vcf"
    p.write_text(VCF)
    im = vps.Image(p, "vcf")
    im.assign_indtext("A:s4,s1 B:#2-#3")
    # the default population first, individuals in population order (their
    # file order kept within one)
    assert im.names("pop") == ["DefPop", "A", "B"]
    assert im.names("ind") == ["s5", "s1", "s4", "s2", "s3"]


@pytest.mark.parametrize("bad, msg", [
    ("1\t100\trs1\tA\tG\t.\tPASS\t.\tDP\t3\t3\t3\t3\t3", "GT"),            # no GT in FORMAT
    ("1\t100\trs1\tA\tG\t.\tPASS\t.\tGT\t0/0\t0/2\t0/0\t0/0\t0/0", "0/2"),  # allele index >= k
    ("1\t100\trs1\tA\tG\t.\tPASS\t.\tGT\t0/0\t0/1\t0/0\t0/0", "fields"),    # a sample left out
    ("1\t100\trs1\tA\tG\t.\tPASS\t.\tGT\t0/0\t0/1\t" + "/".join(["0"] * 11) + "\t0/0\t0/0", "ploidy"),
    ("1\t100\trs1\tA\tG\t.\tPASS\t.\tGT\t0/0\tx/1\t0/0\t0/0\t0/0", "x/1"),  # not a number
])
def test_errors(vps, tmp_path, bad, msg):
    lines = VCF.split("\n")
    p = tmp_path / "bad.vcf"
    p.write_text("\n".join(lines[:3] + [bad]) + "\n")
    with pytest.raises(vps.VpsError, match=msg):
        vps.Image(p, "vcf")
