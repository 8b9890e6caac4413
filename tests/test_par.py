"""The clustering options of a vcfpop parameter file (vcfpop_amd.par):
-g_* inputs and the -structure_* option surface (src/parameters.cpp:
1230-1315) mapped onto vps_params.  CPU only; the device run is in
tests/test_gpu_par.py."""
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PAR = os.path.join(HERE, "golden", "par.polygene.txt")


def test_example_parameter_file(vps):
    from vcfpop_amd import par
    g, st, argv = par.read_par(PAR)
    assert g["input"] == "data.polygene.txt" and g["format"] == "polygene" and g["seed"] == "1"
    kw, krange = par.structure_params(st, int(g["seed"]))
    assert krange == (2, 6)
    assert kw["admix_val"] == 1 and kw["locpriori_val"] == 1 and kw["f_val"] == 1
    assert kw["uniformalpha_val"] == 1 and kw["inferalpha_val"] == 2 and kw["diffalpha_val"] == 2
    assert (kw["nburnin"], kw["nreps"], kw["nthinning"], kw["nruns"], kw["metrofreq"]) == (1000, 10000, 10, 4, 10)
    p = vps.params_default(**kw)
    assert (p.krange_min, p.krange_max, p.g_seed) == (2, 6, 1)


def test_unknown_structure_option_is_refused(vps):
    from vcfpop_amd import par
    with pytest.raises(ValueError, match="unrecognized"):
        par.structure_params({"bogus": "1"}, 0)
    with pytest.raises(ValueError, match="yes or no"):
        par.structure_params({"admix": "maybe"}, 0)


def test_eval_option(vps):
    from vcfpop_amd import par
    par.structure_params({"eval": "yes"}, 0)
    par.structure_params({"eval": "no"}, 0)
    with pytest.raises(ValueError, match="yes or no"):
        par.structure_params({"eval": "1"}, 0)


def test_parameter_file_tokens(vps):
    """SetParameters (src/parameters.cpp:1345-1362) splits a parameter file
    at blanks outside double quotes and trims the quotes (TrimQuote,
    src/string2.cpp:148-163)"""
    from vcfpop_amd import par
    assert par.par_tokens('-a=1  -b="x y"\r\n\t"-c=z"\n-d\n') == ["-a=1", "-b=x y", "-c=z", "-d"]
    assert par.trim_quote('"-e="q"') == "-e=q"
    g, st, argv = par.read_par(PAR)
    assert argv[0] == "-g_output=polygene\\out" and argv[1] == "-g_decimal=6"
    assert "-g_indtext=pop1:#1-#30 pop2:#31-#60 pop3:#61-#90 pop4:#91-#120 #REG reg1:#1-#2 reg2:#3-#4" in argv
    assert g["indtext"].endswith("#REG reg1:#1-#2 reg2:#3-#4")
    assert "Huang2014,Huang2015,Ritland1996_modified,Loiselle1995_modified,Ritland1996,Loiselle1995,Weir1996" in argv
    assert par.path_value(g["output"]) == "polygene/out"


def test_example_job_populations(vps):
    """the example's -g_indtext: four populations of 30 in file order, two
    level-1 regions"""
    from vcfpop_amd import par
    job = par.Job(PAR, sweeps=10)
    im = job.image
    assert im.names("pop") == ["pop1", "pop2", "pop3", "pop4"]
    assert im.names("reg") == ["reg1", "reg1", "reg2", "reg2"]
    assert im.names("ind") == [f"Ind{i}" for i in range(1, 121)]
    assert list(im.arrays()["ind_pop"]) == [i // 30 for i in range(120)]
    assert job.runs[:5] == [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1)] and job.total == 10
    assert job.header["input"] == "data.polygene.txt" and job.header["output"] == "polygene/out"
    job.close()


def test_float_options(vps):
    """-g_float=single runs BAYESIAN<float>: with vcfpop's default
    -g_fastsingle=yes the fast single-precision passes"""
    from vcfpop_amd import par
    assert par.real_precision({}) == vps.PREC_FP64
    assert par.real_precision({"float": "single"}) == vps.PREC_FAST32
    assert par.real_precision({"float": "single", "fastsingle": "no"}) == vps.PREC_REAL32
    assert par.real_precision({"float": "double", "fastsingle": "no"}) == vps.PREC_FP64
    with pytest.raises(ValueError, match="single or double"):
        par.real_precision({"float": "half"})
