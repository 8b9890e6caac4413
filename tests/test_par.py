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
    g, st = par.read_par(PAR)
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
