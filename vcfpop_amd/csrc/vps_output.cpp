// vps_output.cpp — the run's result tables (WriteStructure and
// WriteStructureSummary, src/structure.cpp:88-401), written on the host from
// a finished chain's state: run header, mean lnL / variance / ln P(D), the
// mean hyper-parameters, the membership of each sampling location (from the
// final Q), the inferred ancestry of each individual (normalised MiSum), the
// net nucleotide distance among clusters and, optionally, the per-locus
// diversity of each cluster (from the mean allele frequencies FreqM).
// Number format, delimiter and line break follow the reference's -decimal,
// -scientific, -delimiter and -linebreak options (WriteReal,
// src/string2.cpp:1174-1178).
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <time.h>
#include <string>
#include <vector>
#include "vps_runtime.h"

using namespace vps;

namespace {

struct Writer {
    FILE* f = nullptr;
    char fmt[16];
    char del = '\t';
    std::string lb = "\r\n";
    explicit Writer(const vps_output_options& o)
    {
        snprintf(fmt, sizeof(fmt), o.scientific == 1 ? "%%0.%de" : "%%0.%df", o.decimal);
        del = o.delimiter;
        lb = o.linebreak ? o.linebreak : "\r\n";
    }
    void real(double v)
    {
        if (isnan(v) || isinf(v)) fputs("nan", f);
        else fprintf(f, fmt, v);
    }
    void s(const char* x) { fputs(x, f); }
    void br() { fputs(lb.c_str(), f); }
    void d() { fputc(del, f); }
};

std::string name_or(const std::vector<std::string>& v, size_t i, const char* prefix)
{
    if (i < v.size() && !v[i].empty()) return v[i];
    return std::string(prefix) + std::to_string(i + 1);
}

}  // namespace

extern "C" void vps_output_options_default(vps_output_options* o)
{
    memset(o, 0, sizeof(*o));
    o->decimal = 5;         // -decimal (src/parameters.cpp:432-433)
    o->scientific = 2;
    o->delimiter = '\t';    // -delimiter (src/parameters.cpp:458)
    o->linebreak = "\r\n";  // -linebreak (src/parameters.cpp:459)
    o->diversity = 2;       // -structure_diversity
    o->rep = 1;
}

extern "C" int vps_write_tables_host(const vps_dataset* ds, const vps_params* par, int K, int run_id,
                                     const vps_run_state* st, const char* prefix, const vps_output_options* opt,
                                     vps_run_summary* summary)
{
    if (!ds || !par || !st || !prefix || !opt) return fail(VPS_ERR_INVALID, "null argument");
    if (!st->Q || !st->ancestry || !st->FreqM || !st->rout || st->rout_len < 4)
        return fail(VPS_ERR_INVALID, "incomplete run state");
    if (opt->decimal < 1 || opt->decimal > 15) return fail(VPS_ERR_INVALID, "decimal %d", opt->decimal);
    const int N = ds->d.N, S = std::max(1, ds->d.S);
    const int64_t L = ds->d.L, KT = ds->d.KT;
    const bool admix = par->admix_val == 1, locpriori = par->locpriori_val == 1, fmodel = par->f_val == 1;
    const bool inferalpha = par->inferalpha_val == 1, diffalpha = par->diffalpha_val == 1;
    const bool inferlambda = par->inferlambda_val == 0, difflambda = par->difflambda_val == 1;
    const bool fsame = par->singlef_val == 1;
    const int nburnin = par->nburnin + ((!locpriori && !admix) ? par->nadmburnin : 0);
    const int nthin = std::max(1, par->nthinning);
    const double* rout = st->rout;
    const int64_t rlen = st->rout_len;
    Writer w(*opt);
    char path[4096];

    // lnL trace (structure_writelnl)
    if (par->writelnl_val == 1 && st->lnltrace) {
        snprintf(path, sizeof(path), "%s.structure.k=%d_rep=%d_id=%d.lnl.txt", prefix, K, opt->rep, run_id);
        w.f = fopen(path, "wb");
        if (!w.f) return fail(VPS_ERR_INVALID, "cannot write %s", path);
        fprintf(w.f, "%d", nburnin); w.br();
        w.s("Iteration"); w.d(); w.s("lnL"); w.br();
        int64_t x = 0;
        for (int i = 0; i < nburnin; i += nthin) {
            fprintf(w.f, "%d", i); w.d();
            w.real(x < st->lnltrace_len ? st->lnltrace[x] : NAN); ++x; w.br();
        }
        for (int i = 0; i < par->nreps; i += nthin) {
            fprintf(w.f, "%d", i + nburnin); w.d();
            w.real(x < st->lnltrace_len ? st->lnltrace[x] : NAN); ++x; w.br();
        }
        fclose(w.f);
    }

    snprintf(path, sizeof(path), "%s.structure.k=%d_rep=%d_id=%d.txt", prefix, K, opt->rep, run_id);
    w.f = fopen(path, "wb");
    if (!w.f) return fail(VPS_ERR_INVALID, "cannot write %s", path);
    w.br(); w.br(); w.s("Parameters:"); w.br();
    fprintf(w.f, "Seed=%lld", (long long)st->seed); w.br();
    fprintf(w.f, "Model=%s,%s,%s", admix ? "ADM" : "NOADM", locpriori ? "LOC" : "NOLOC", fmodel ? "F" : "NOF"); w.br();
    fprintf(w.f, "Number of cluster=%d", K); w.br();
    fprintf(w.f, "Replicate=%d", opt->rep); w.br();
    fprintf(w.f, "Run id=%d", run_id); w.br();
    w.s("Mean value of ln likelihood="); w.real(rout[0]); w.br();
    w.s("Variance of ln likelihood="); w.real(rout[2]); w.br();
    w.s("Estimated Ln Prob of Data="); w.real(rout[3]); w.br(); w.br();
    if (summary) {
        summary->id = run_id; summary->k = K; summary->rep = opt->rep;
        summary->mean_lnl = rout[0]; summary->var_lnl = rout[2]; summary->lnpd = rout[3];
    }

    const int nf = fmodel ? (fsame ? 1 : K) : 0;
    const int nl = difflambda ? K : 1;
    if (inferlambda) {
        if (!difflambda) { w.br(); w.s("Mean value of lambda="); w.real(rout[4]); }
        else for (int j = 0; j < K; ++j) { w.br(); fprintf(w.f, "Mean value of lambda_%d=", j + 1); w.real(rout[4 + j]); }
    }
    if (locpriori) {
        w.br(); w.s("Mean value of r="); w.real(rout[4 + nl]);
        w.br(); fprintf(w.f, "Mean value of global %s", admix ? "alpha" : "eta");
        w.br(); w.d(); w.s("Cluster"); w.br();
        for (int j = 0; j < K; ++j) { w.d(); fprintf(w.f, "%d", j + 1); }
        w.br();
        for (int j = 0; j < K; ++j) { w.d(); w.real(rout[4 + nl + 1 + j]); }
        w.br(); fprintf(w.f, "Mean value of local %s for each location", admix ? "alpha" : "gamma");
        w.br(); w.d(); w.s("Cluster"); w.br(); w.s("Pop");
        for (int j = 0; j < K; ++j) { w.d(); fprintf(w.f, "%d", j + 1); }
        for (int s = 0; s < ds->d.S; ++s) {
            w.br(); w.s(name_or(ds->pop_names, s, "Pop").c_str());
            for (int j = 0; j < K; ++j) { w.d(); w.real(rout[4 + nl + 1 + K + s * K + j]); }
        }
    } else if (admix && inferalpha) {
        if (!diffalpha) { w.br(); w.s("Mean value of alpha="); w.real(rout[4 + nl]); }
        else for (int j = 0; j < K; ++j) { w.br(); fprintf(w.f, "Mean value of alpha_%d=", j + 1); w.real(rout[4 + nl + j]); }
    }
    if (fmodel) {
        if (fsame) { w.br(); w.s("Mean value of Fst="); w.real(rout[rlen - 1]); }
        else {
            w.br(); w.s("Mean value of Fst");
            w.br(); w.d(); w.s("Cluster"); w.br();
            for (int j = 0; j < K; ++j) { w.d(); fprintf(w.f, "%d", j + 1); }
            w.br();
            for (int i = 0; i < nf; ++i) { w.d(); w.real(rout[rlen - nf + i]); }
        }
    }

    // membership of each location: sum of the final Q rows, rows normalised
    // (Add + Unify with MIN_FREQ, src/structure.cpp:249-257)
    std::vector<double> O((size_t)S * K, 0.0);
    for (int i = 0; i < N; ++i) {
        const int s = ds->d.S > 0 ? ds->popid[i] : 0;
        for (int k = 0; k < K; ++k) O[(size_t)s * K + k] += st->Q[(size_t)i * K + k];
    }
    for (int s = 0; s < S; ++s) unify_row(O.data() + (size_t)s * K, K);
    w.br(); w.br();
    fprintf(w.f, "Proportion of membership of each pre-defined population in each of the %d clusters", K);
    w.br(); w.d(); w.s("Cluster"); w.br(); w.s("Pop");
    for (int k = 0; k < K; ++k) { w.d(); fprintf(w.f, "%d", k + 1); }
    for (int s = 0; s < S; ++s) {
        w.br(); w.s(name_or(ds->pop_names, s, "Pop").c_str());
        for (int k = 0; k < K; ++k) { w.d(); w.real(O[(size_t)s * K + k]); }
    }

    // inferred ancestry: normalised allele-copy assignments (MiSum)
    w.br(); w.br(); w.s("Inferred ancestry of individuals");
    w.br(); w.d(); w.d(); w.s("Cluster");
    w.br(); w.s("Ind"); w.d(); w.s("Pop"); w.d(); w.s("RegL1");
    for (int k = 0; k < K; ++k) { w.d(); fprintf(w.f, "%d", k + 1); }
    for (int i = 0; i < N; ++i) {
        const int s = ds->d.S > 0 ? ds->popid[i] : 0;
        w.br(); w.s(name_or(ds->ind_names, i, "Ind").c_str()); w.d();
        w.s(name_or(ds->pop_names, s, "Pop").c_str()); w.d();
        w.s((size_t)s < ds->pop_regions.size() && !ds->pop_regions[s].empty() ? ds->pop_regions[s].c_str() : "Total");
        for (int k = 0; k < K; ++k) { w.d(); w.real(st->ancestry[(size_t)i * K + k]); }
    }

    // heterozygosity of each cluster and net nucleotide distance
    // (src/structure.cpp:292-339)
    std::vector<double> H(K, 0.0), D((size_t)K * K, 0.0), H2((size_t)K * L, 0.0);
    for (int k = 0; k < K; ++k) {
        const double* p = st->FreqM + (size_t)k * KT;
        for (int64_t l = 0; l < L; ++l) {
            const int k2 = ds->k2[l];
            double ss = 0;
            for (int a = 0; a < k2; ++a) ss += p[a] * p[a];
            H2[(size_t)k * L + l] = 1 - ss;
            H[k] += 1 - ss;
            p += k2;
        }
        H[k] /= (double)L;
        for (int j = 0; j <= k; ++j) {
            const double* p1 = st->FreqM + (size_t)k * KT;
            const double* p2 = st->FreqM + (size_t)j * KT;
            double dt = 0;
            for (int64_t l = 0; l < L; ++l) {
                const int k2 = ds->k2[l];
                double sp = 0;
                for (int a = 0; a < k2; ++a) sp += p1[a] * p2[a];
                dt += 1 - sp;
                p1 += k2;
                p2 += k2;
            }
            D[(size_t)j * K + k] = D[(size_t)k * K + j] = dt / (double)L - (H[k] + H[j]) / 2;
        }
    }
    w.br(); w.br(); w.s("Allele-frequency divergence among pops (Net nucleotide distance)");
    w.br(); w.d(); w.s("Cluster"); w.br(); w.s("Cluster");
    for (int j = 0; j < K; ++j) { w.d(); fprintf(w.f, "%d", j + 1); }
    for (int i = 0; i < K; ++i) {
        w.br(); fprintf(w.f, "%d", i + 1);
        for (int j = 0; j < K; ++j) { w.d(); w.real(D[(size_t)i * K + j]); }
    }

    if (opt->diversity == 1) {
        w.br(); w.br(); w.s("Estimated genetic diversity in each cluster");
        w.br(); w.s("Locus"); w.d(); w.s("Cluster"); w.d(); w.s("Heterozygosity");
        for (int a = 0; a < ds->d.maxK; ++a) { w.d(); fprintf(w.f, "Allele%d", a); }
        for (int k = 0; k < K; ++k) {
            w.br(); w.s(k ? "" : "Average"); w.d(); fprintf(w.f, "%d", k + 1);
            w.real(H[k]);
        }
        for (int64_t l = 0; l < L; ++l)
            for (int k = 0; k < K; ++k) {
                w.br(); w.s(k ? "" : name_or(ds->loc_names, (size_t)l, "Locus").c_str());
                w.d(); fprintf(w.f, "%d", k + 1); w.d();
                w.real(H2[(size_t)k * L + l]);
                const double* p = st->FreqM + (size_t)k * KT + ds->afo[l];
                for (int a = 0; a < ds->k2[l]; ++a) { w.d(); w.real(p[a]); }
            }
    }
    fclose(w.f);
    return VPS_OK;
}

extern "C" int vps_chain_write_tables(vps_chain* ch, const char* prefix, const vps_output_options* opt,
                                      vps_run_summary* summary)
{
    if (!ch || !prefix || !opt) return fail(VPS_ERR_INVALID, "null argument");
    if (!ch->finished) return fail(VPS_ERR_STATE, "write tables after vps_chain_finish");
    const int64_t N = ch->c.N, K = ch->K, KT = ch->c.KT;
    int64_t rlen = 0, nl = 0;
    int rc = vps_chain_get_sizes(ch, &rlen, &nl);
    if (rc) return rc;
    std::vector<double> Q(N * K), anc(N * K), fm(K * KT), rout(std::max<int64_t>(rlen, 1)), trace(std::max<int64_t>(nl, 1));
    if (!rc) rc = vps_chain_get_state(ch, VPS_STATE_Q, Q.data(), Q.size() * 8);
    if (!rc) rc = vps_chain_get_state(ch, VPS_STATE_MISUM, anc.data(), anc.size() * 8);
    if (!rc) rc = vps_chain_get_state(ch, VPS_STATE_FREQM, fm.data(), fm.size() * 8);
    if (!rc) rc = vps_chain_get_state(ch, VPS_STATE_ROUT, rout.data(), rlen * 8);
    if (!rc && nl) rc = vps_chain_get_state(ch, VPS_STATE_LNLTRACE, trace.data(), nl * 8);
    if (rc) return rc;
    vps_run_state st{};
    st.Q = Q.data(); st.ancestry = anc.data(); st.FreqM = fm.data();
    st.rout = rout.data(); st.rout_len = rlen;
    st.lnltrace = nl ? trace.data() : nullptr; st.lnltrace_len = nl;
    st.seed = ch->c.seed;
    return vps_write_tables_host(ch->ds, &ch->par, (int)K, ch->id, &st, prefix, opt, summary);
}

// LwrLineCmp(prefix, arg) == 0 (src/string2.cpp:247-258): arg starts with
// prefix, letters compared in lower case
static bool starts_lower(const char* arg, const char* prefix)
{
    for (; *prefix; ++prefix, ++arg) {
        const int a = (*prefix >= 'A' && *prefix <= 'Z') ? *prefix + 32 : *prefix;
        const int b = (*arg >= 'A' && *arg <= 'Z') ? *arg + 32 : *arg;
        if (a != b) return false;
    }
    return true;
}

extern "C" int vps_write_summary(const char* path, const vps_run_summary* runs, int n, const vps_output_options* opt,
                                 const vps_result_header* hdr)
{
    if (!path || (!runs && n) || !opt) return fail(VPS_ERR_INVALID, "null argument");
    if (hdr && hdr->argc && !hdr->argv) return fail(VPS_ERR_INVALID, "null argv");
    Writer w(*opt);
    w.f = fopen(path, "wb");
    if (!w.f) return fail(VPS_ERR_INVALID, "cannot write %s", path);
    // OpenResFile("-structure", "#Bayesian clustering"): a title starting with
    // '#' prefixes every head line with '#'
    w.s("##Bayesian clustering calculated by vcfpop v" VPS_VCFPOP_VERSION);
    w.br();
    if (hdr) {
        fprintf(w.f, "#  Input: %s", hdr->input ? hdr->input : ""); w.br();
        fprintf(w.f, "#  Output: %s", hdr->output ? hdr->output : ""); w.br();
        const time_t t = hdr->time ? (time_t)hdr->time : time(nullptr);
        struct tm tmv;
        localtime_r(&t, &tmv);
        fprintf(w.f, "#  Time: %04d-%02d-%02d %02d:%02d:%02d", tmv.tm_year + 1900, tmv.tm_mon + 1, tmv.tm_mday,
                tmv.tm_hour, tmv.tm_min, tmv.tm_sec);
        w.br();
        fprintf(w.f, "#  Parameter: %s", hdr->argc > 0 && hdr->argv[0] ? hdr->argv[0] : ""); w.br();
        for (int i = 1; i < hdr->argc; ++i) {
            const char* a = hdr->argv[i] ? hdr->argv[i] : "";
            const bool g = starts_lower(a, "-g_") && !starts_lower(a, "-g_indtext") && !starts_lower(a, "-g_indtab");
            if ((hdr->filters && starts_lower(a, "-f_")) || g || starts_lower(a, "-structure")) {
                fprintf(w.f, "#    %s", a); w.br();
            }
        }
    }
    w.br(); w.br();
    w.s("ID"); w.d(); w.s("k"); w.d(); w.s("rep"); w.d(); w.s("Mean lnL"); w.d(); w.s("Var lnL"); w.d(); w.s("lnP(D)");
    for (int i = 0; i < n; ++i) {
        w.br(); fprintf(w.f, "%d", runs[i].id); w.d(); fprintf(w.f, "%d", runs[i].k); w.d();
        fprintf(w.f, "%d", runs[i].rep); w.d();
        w.real(runs[i].mean_lnl); w.d(); w.real(runs[i].var_lnl); w.d(); w.real(runs[i].lnpd);
    }
    fclose(w.f);
    return VPS_OK;
}
