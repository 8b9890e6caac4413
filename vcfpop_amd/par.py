"""Run the Bayesian clustering of a vcfpop parameter file on the device.

    python -m vcfpop_amd.par par.polygene.txt [--device 0] [--sweeps N] [--topology reference|parallel]

The file is split as vcfpop splits it (SetParameters,
src/parameters.cpp:1345-1362): blank-separated tokens, blanks inside double
quotes kept, quotes trimmed.  Reads the options the clustering uses and
ignores the rest (other analyses are not part of this engine):
  -g_input, -g_format (polygene or vcf; a .vcf.gz input is read through
    zlib), -g_output, -g_seed, -g_decimal
  -g_float, -g_fastsingle: REAL = float runs (PREC_FAST32 / PREC_REAL32)
  -g_indtext: populations and regions (vcfpop's default without it: one
    population; the file's population column is not read, as in vcfpop)
  -structure_* (src/parameters.cpp:1230-1315): models, K range, runs,
  burn-in / sampling / thinning lengths, lambda/alpha/r/F settings.
The runs are the reference's task list (src/structure.cpp:2402-2413): for
K = krange_min..krange_max and rep = 1..nruns, run id = 1, 2, ... in that
order and run seed = HashULong(g_seed + id).  All runs are one batch on
one device; each finished run writes its WriteStructure table and the
runs' WriteStructureSummary table, with the result-file head of OpenResFile
(src/file.cpp:269-287), goes to <g_output>.structure.txt.

-structure_eval=yes switches to vcfpop's evaluation mode instead
(src/function.cpp:4743-4766): K = 5, one run, one sweep, both models, each
pass timed on the FP64 and the mixed-precision device paths with digests of
its integer state (vcfpop_amd.evalmode).
"""
import argparse
import os
import re
import sys
import time

from ._lib import (PREC_FAST32, PREC_FP64, PREC_REAL32, RNG_PARALLEL, RNG_REFERENCE, Batch, Chain, Dataset, Image,
                   params_default, write_summary)

_YESNO = {"yes": 1, "no": 2}
_BOOL_OPTS = {"admix", "locpriori", "f", "inferlambda", "difflambda", "diversity", "inferalpha", "diffalpha",
              "uniformalpha", "singlef", "writelnl"}
_INT_OPTS = {"nburnin", "nreps", "nthinning", "nruns", "nadmburnin", "metrofreq"}
_DBL_OPTS = {"lambda", "stdlambda", "maxlambda", "alpha", "stdalpha", "maxalpha", "alphapriora", "alphapriorb",
             "r", "maxr", "stdr", "epseta", "epsgamma", "pmeanf", "pstdf", "stdf"}
_IGNORED = {"plot", "nstream"}


def trim_quote(a):
    """TrimQuote (src/string2.cpp:148-163): a leading and a trailing double
    quote dropped, and one right after the first '='"""
    if not a:
        return a
    if a[0] == '"':
        a = a[1:]
    if a and a[-1] == '"':
        a = a[:-1]
    i = a.find("=")
    if i != -1 and i + 1 < len(a) and a[i + 1] == '"':
        a = a[:i + 1] + a[i + 2:]
    return a


def par_tokens(text):
    """a parameter file's arguments as vcfpop reads them: split at line
    breaks, tabs and spaces outside double quotes, quotes trimmed"""
    text = text.replace("\r\n", "\n")
    toks, cur, quoted = [], [], False
    for ch in text:
        if ch == '"':
            quoted = not quoted
        if ch in "\n\t " and not quoted:
            if cur:
                toks.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    if cur:
        toks.append("".join(cur))
    return [trim_quote(t) for t in toks]


def real_precision(g):
    """-g_float (single|double, default double) and -g_fastsingle (yes|no,
    default yes; src/parameters.cpp:443-444, 795-798): BAYESIAN<float> runs
    as PREC_FAST32 (fast single) or PREC_REAL32, BAYESIAN<double> as FP64"""
    fl = g.get("float", "double").lower()
    fs = g.get("fastsingle", "yes").lower()
    if fl not in ("single", "double"):
        raise ValueError(f"-g_float={fl}: expected single or double")
    if fs not in _YESNO:
        raise ValueError(f"-g_fastsingle={fs}: expected yes or no")
    if fl == "double":
        return PREC_FP64
    return PREC_FAST32 if fs == "yes" else PREC_REAL32


def path_value(v):
    """ReplacePathDelim (src/parameters.cpp:364-375): backslashes as '/'"""
    return v.replace("\\", "/")


def read_par(path):
    """(global options, structure options, arguments) of a parameter file"""
    g, st = {}, {}
    with open(path, encoding="utf-8", errors="replace", newline="") as f:
        argv = par_tokens(f.read())
    for a in argv:
        m = re.match(r"^-(g|structure)_([A-Za-z0-9]+)=(.*)$", a, re.S)
        if not m:
            continue
        (g if m.group(1).lower() == "g" else st)[m.group(2).lower()] = m.group(3)
    return g, st, argv


def structure_params(st, g_seed):
    """vps_params from -structure_* options (the reference's defaults otherwise)"""
    kw = {}
    krange = (1, 5)
    for k, v in st.items():
        if k in _BOOL_OPTS:
            if v.lower() not in _YESNO:
                raise ValueError(f"-structure_{k}={v}: expected yes or no")
            kw[k + "_val"] = _YESNO[v.lower()]
        elif k in _INT_OPTS:
            kw[k] = int(v)
        elif k in _DBL_OPTS:
            kw[k] = float(v)
        elif k == "eval":
            if v.lower() not in _YESNO:
                raise ValueError(f"-structure_{k}={v}: expected yes or no")
        elif k == "krange":
            m = re.match(r"^\[?\s*(\d+)\s*,\s*(\d+)\s*\]?$", v)
            if not m:
                raise ValueError(f"-structure_krange={v}: expected [min,max]")
            krange = (int(m.group(1)), int(m.group(2)))
        elif k not in _IGNORED:
            raise ValueError(f"unrecognized parameter -structure_{k}")
    kw["g_seed"] = g_seed
    kw["krange_min"], kw["krange_max"] = krange
    return kw, krange


class _LogFile:
    """a file-like adapter that sends whole lines to a log function"""

    def __init__(self, log):
        self.log, self.buf = log, ""

    def write(self, text):
        self.buf += text
        while "\n" in self.buf:
            line, self.buf = self.buf.split("\n", 1)
            self.log(line)

    def flush(self):
        pass


class Job:
    """A parameter file's clustering job: the genotype image with vcfpop's
    populations and regions, the run parameters, the run list and the
    result-file head."""

    def __init__(self, par_path, sweeps=None, topology="reference"):
        g, st, argv = read_par(par_path)
        base = os.path.dirname(os.path.abspath(par_path))
        self.format = g.get("format", "vcf").lower()
        if self.format not in ("polygene", "vcf"):
            raise ValueError(f"-g_format={g['format']}: this engine reads the polygene and VCF formats")
        if "indtab" in g:
            raise ValueError("-g_indtab is not supported: give the populations with -g_indtext")
        inp = g.get("input")
        if not inp:
            raise ValueError("-g_input is missing")
        self.header = {"input": path_value(inp), "output": path_value(g.get("output", "vcfpop")), "argv": argv,
                       "filters": any(a.lower() == "-f" for a in argv), "time": 0}
        self.input = os.path.join(base, path_value(inp).replace("/", os.sep))
        self.out = os.path.join(base, path_value(g.get("output", "vcfpop")).replace("/", os.sep))
        self.decimal = int(g.get("decimal", 5))
        self.seed = int(g.get("seed", 0))
        kw, (kmin, kmax) = structure_params(st, self.seed)
        if sweeps is not None:   # a shorter run: burn-in a fifth, sampling the rest
            kw["nburnin"], kw["nreps"] = sweeps // 5, sweeps - sweeps // 5
        kw["rng_topology"] = RNG_PARALLEL if topology == "parallel" else RNG_REFERENCE
        kw["precision"] = real_precision(g)
        self.kw = kw
        self.params = p = params_default(**kw)
        nadm = p.nadmburnin if (p.admix_val == 2 and p.locpriori_val == 2) else 0
        self.total = nadm + p.nburnin + p.nreps
        # run id i + 1 for (K, rep) in the reference's task order
        self.runs = [(k, rep) for k in range(kmin, kmax + 1) for rep in range(1, p.nruns + 1)]
        self.eval = st.get("eval", "no").lower() == "yes"
        self.image = Image(self.input, self.format)
        indtext = g.get("indtext")
        self.image.assign_indtext(indtext.replace("#n", "\n") if indtext else None)

    def close(self):
        self.image.close()


def run(par_path, device=0, sweeps=None, topology="reference", log=print):
    """Run a parameter file's clustering; returns the runs' summaries (or,
    with -structure_eval=yes, the evaluation rows)"""
    job = Job(par_path, sweeps, topology)
    im, p, out = job.image, job.params, job.out
    log(f"{job.input}: {im.N} individuals in {im.desc.n_pop} populations, {im.L} of {im.loci_read} loci "
        "(fewer than 2 alleles dropped)")
    ds = Dataset(im, device)
    if job.eval:
        from . import evalmode
        base = {k: v for k, v in job.kw.items() if k not in ("krange_min", "krange_max", "nruns", "rng_topology", "precision")}
        rows = evalmode.run_eval(ds, K=5, seed=job.seed, base=base)
        evalmode.print_table(rows, file=_LogFile(log))
        ds.close()
        job.close()
        return rows
    os.makedirs(os.path.dirname(out) or ".", exist_ok=True)
    chains = [Chain(ds, p, k, i + 1) for i, (k, rep) in enumerate(job.runs)]
    t0 = time.perf_counter()
    b = Batch(chains)
    b.step(job.total)
    b.sync()
    dt = time.perf_counter() - t0
    log(f"{len(job.runs)} runs x {job.total} sweeps in {dt:.2f} s")
    summaries = []
    for c, (k, rep) in zip(chains, job.runs):
        c.finish()
        summaries.append(c.write_tables(out, rep=rep, decimal=job.decimal, diversity=p.diversity_val))
    write_summary(out + ".structure.txt", summaries, header=job.header, decimal=job.decimal)
    b.close()
    for c in chains:
        c.close()
    ds.close()
    job.close()
    return summaries


def main(argv=None):
    ap = argparse.ArgumentParser(description="Bayesian clustering of a vcfpop parameter file on the device")
    ap.add_argument("par")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--sweeps", type=int, default=None, help="override burn-in + sampling length (a quick run)")
    ap.add_argument("--topology", choices=["reference", "parallel"], default="reference")
    args = ap.parse_args(argv)
    for s in run(args.par, args.device, args.sweeps, args.topology):
        if "id" in s:
            print(f"id {s['id']} K={s['k']} rep {s['rep']}: mean lnL {s['mean_lnl']:.3f}, lnP(D) {s['lnpd']:.3f}")


if __name__ == "__main__":
    sys.exit(main())
