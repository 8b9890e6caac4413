// vps_synth.cpp — deterministic synthetic genotype data in vcfpop's packed
// layout (used by the benchmark workloads and the serving job stream).
//
// Recipe: Q_i ~ Dirichlet(alpha_q * 1_K); per locus either
//   fst == 0:  P[k][l] ~ Dirichlet(1, ..., 1)                  (biallelic)
//   fst  > 0:  p0 ~ U(0.05, 0.95), P[k][l] ~ Beta(p0 (1-F)/F, (1-p0)(1-F)/F)
// a fraction of loci is multiallelic with U{min..max} alleles and
// P[k][l] ~ Dirichlet(multi_dir); each allele copy of individual i draws its
// cluster from Q_i and its allele from P[cluster][l]; a genotype is missing
// with probability `missing`.  Individual ploidy follows the configured mix.
// Genotype ids are assigned per locus in order of first appearance, the
// missing genotype of each ploidy being its own entry (vcfpop's
// missing-with-ploidy patterns, src/global.h:29), and packed with
// CeilLog2(ngeno) bits (src/math2.h:503-512) LSB-first per locus.
#include <algorithm>
#include <string.h>
#include <thread>
#include <unordered_map>
#include "vps_internal.h"
#include "../../include/vps_synth.h"

using namespace vps;

struct vps_synth {
    int32_t N = 0;
    int64_t L = 0;
    int32_t S = 1;
    int32_t Kt = 0;
    std::vector<uint8_t> bits;
    std::vector<uint64_t> byteoff;
    std::vector<uint8_t> bitsize;
    std::vector<int32_t> nallele;
    std::vector<int64_t> gtoff;
    std::vector<uint8_t> gtploidy;
    std::vector<int64_t> gtaoff;
    std::vector<uint16_t> pool;
    std::vector<int32_t> pop;
    std::vector<double> qtrue;
    std::vector<int32_t> ploidy;
};

namespace {

int ceil_log2(int64_t v)
{
    if (v <= 1) return 1;
    int b = 0;
    uint64_t x = (uint64_t)(v - 1);
    while (x) { ++b; x >>= 1; }
    return b;
}

double beta_draw(Rng& r, double a, double b)
{
    const double x = r.gamma(a), y = r.gamma(b);
    return x / (x + y);
}

struct LocusOut {
    int k2 = 2;
    int bits = 1;
    std::vector<uint32_t> ids;           // [N]
    std::vector<uint8_t> gploidy;        // [ngeno]
    std::vector<uint16_t> galleles;      // concatenated
    std::vector<int64_t> gaoff;          // [ngeno] relative offsets
};

}  // namespace

extern "C" void vps_synth_default(vps_synth_config* c)
{
    memset(c, 0, sizeof(*c));
    c->n_ind = 200; c->n_loc = 2000; c->k_true = 3; c->alpha_q = 0.5; c->fst = 0.0; c->missing = 0.01;
    c->frac_multi = 0.0; c->min_alleles = 3; c->max_alleles = 8; c->multi_dir = 0.3;
    c->n_ploidy = 0; c->n_pop = 1; c->seed = 1;
}

extern "C" int vps_synth_create(const vps_synth_config* cfg, vps_synth** out)
{
    if (!cfg || !out) return fail(VPS_ERR_INVALID, "null argument");
    *out = nullptr;
    const int N = cfg->n_ind, Kt = cfg->k_true;
    const int64_t L = cfg->n_loc;
    if (N < 1 || L < 1 || Kt < 1) return fail(VPS_ERR_INVALID, "synth sizes N=%d L=%lld K=%d", N, (long long)L, Kt);
    if (cfg->max_alleles > VPS_MAX_ALLELES || cfg->min_alleles < 2 || cfg->min_alleles > cfg->max_alleles)
        return fail(VPS_ERR_INVALID, "allele range %d..%d", cfg->min_alleles, cfg->max_alleles);
    vps_synth* s = new vps_synth();
    s->N = N; s->L = L; s->Kt = Kt; s->S = std::max(1, cfg->n_pop);

    // individuals: ploidy, population, admixture proportions
    Rng ri;
    ri.init(cfg->seed, 0x5EED0001ull);
    s->ploidy.resize(N);
    for (int i = 0; i < N; ++i) {
        int v = 2;
        if (cfg->n_ploidy > 0) {
            double u = ri.uniform(), acc = 0;
            v = cfg->ploidy_levels[cfg->n_ploidy - 1];
            for (int j = 0; j < cfg->n_ploidy; ++j) {
                acc += cfg->ploidy_frac[j];
                if (u < acc) { v = cfg->ploidy_levels[j]; break; }
            }
        }
        if (v < 1 || v > VPS_MAX_PLOIDY) { delete s; return fail(VPS_ERR_INVALID, "ploidy %d", v); }
        s->ploidy[i] = v;
    }
    s->pop.resize(N);
    for (int i = 0; i < N; ++i) s->pop[i] = (int32_t)((int64_t)i * s->S / N);
    s->qtrue.resize((size_t)N * Kt);
    {
        std::vector<double> a(Kt, cfg->alpha_q);
        for (int i = 0; i < N; ++i) dirichlet_a(ri, s->qtrue.data() + (size_t)i * Kt, a.data(), Kt, false);
    }

    // loci in parallel, each from its own stream
    std::vector<LocusOut> lo(L);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            std::vector<double> P, a, cum;
            std::unordered_map<uint64_t, uint32_t> gmap;
            std::vector<uint16_t> cop;
            for (int64_t l = t; l < L; l += nt) {
                Rng r;
                r.init(cfg->seed * 0x9E3779B97F4A7C15ull + (uint64_t)l, 0x5EED0002ull);
                LocusOut& o = lo[l];
                const bool multi = cfg->frac_multi > 0 && r.uniform() < cfg->frac_multi;
                int k2 = 2;
                if (multi) k2 = cfg->min_alleles + (int)r.next((uint64_t)(cfg->max_alleles - cfg->min_alleles + 1));
                o.k2 = k2;
                P.assign((size_t)Kt * k2, 0);
                if (multi) {
                    a.assign(k2, cfg->multi_dir);
                    for (int k = 0; k < Kt; ++k) dirichlet_a(r, P.data() + (size_t)k * k2, a.data(), k2, false);
                } else if (cfg->fst > 0) {
                    const double F = cfg->fst, p0 = r.uniform(0.05, 0.95);
                    for (int k = 0; k < Kt; ++k) {
                        double p = beta_draw(r, p0 * (1 - F) / F, (1 - p0) * (1 - F) / F);
                        P[(size_t)k * 2] = p; P[(size_t)k * 2 + 1] = 1 - p;
                    }
                } else {
                    a.assign(2, 1.0);
                    for (int k = 0; k < Kt; ++k) dirichlet_a(r, P.data() + (size_t)k * 2, a.data(), 2, false);
                }
                // cumulative allele distributions per cluster
                cum.assign((size_t)Kt * k2, 0);
                for (int k = 0; k < Kt; ++k) {
                    double acc = 0;
                    for (int x = 0; x < k2; ++x) { acc += P[(size_t)k * k2 + x]; cum[(size_t)k * k2 + x] = acc; }
                }
                gmap.clear();
                o.ids.resize(N);
                o.gploidy.clear(); o.galleles.clear(); o.gaoff.clear();
                for (int i = 0; i < N; ++i) {
                    const int v = s->ploidy[i];
                    uint64_t key;
                    const bool miss = r.uniform() < cfg->missing;
                    cop.assign(v, 0);
                    if (miss) {
                        key = ~0ull - (uint64_t)v;   // above every typed key (top nibble = ploidy <= 10)
                    } else {
                        const double* q = s->qtrue.data() + (size_t)i * Kt;
                        for (int c = 0; c < v; ++c) {
                            double u = r.uniform(), acc = 0;
                            int z = Kt - 1;
                            for (int k = 0; k < Kt; ++k) { acc += q[k]; if (u < acc) { z = k; break; } }
                            const double u2 = r.uniform() * cum[(size_t)z * k2 + k2 - 1];
                            int x = k2 - 1;
                            for (int y = 0; y < k2; ++y) if (u2 < cum[(size_t)z * k2 + y]) { x = y; break; }
                            cop[c] = (uint16_t)x;
                        }
                        std::sort(cop.begin(), cop.end());
                        key = (uint64_t)v << 60;
                        for (int c = 0; c < v; ++c) key |= (uint64_t)cop[c] << (6 * c);
                    }
                    auto it = gmap.find(key);
                    uint32_t id;
                    if (it == gmap.end()) {
                        id = (uint32_t)o.gploidy.size();
                        gmap.emplace(key, id);
                        o.gploidy.push_back(miss ? 0 : (uint8_t)v);
                        o.gaoff.push_back((int64_t)o.galleles.size());
                        if (!miss) o.galleles.insert(o.galleles.end(), cop.begin(), cop.end());
                    } else id = it->second;
                    o.ids[i] = id;
                }
                o.bits = ceil_log2((int64_t)o.gploidy.size());
            }
        });
    for (auto& x : th) x.join();

    // concatenate and pack (GENO_WRITER layout, src/function.cpp:3580-3631)
    s->byteoff.resize(L); s->bitsize.resize(L); s->nallele.resize(L); s->gtoff.resize(L + 1);
    uint64_t nbytes = 0;
    int64_t ng = 0, np = 0;
    for (int64_t l = 0; l < L; ++l) {
        s->byteoff[l] = nbytes;
        s->bitsize[l] = (uint8_t)lo[l].bits;
        s->nallele[l] = lo[l].k2;
        s->gtoff[l] = ng;
        nbytes += ((uint64_t)lo[l].bits * N + 7) / 8;
        ng += (int64_t)lo[l].gploidy.size();
        np += (int64_t)lo[l].galleles.size();
    }
    s->gtoff[L] = ng;
    s->bits.assign(nbytes + 8, 0);
    s->gtploidy.resize(ng); s->gtaoff.resize(ng); s->pool.resize(np);
    {
        int64_t g = 0, p = 0;
        for (int64_t l = 0; l < L; ++l) {
            for (size_t j = 0; j < lo[l].gploidy.size(); ++j) {
                s->gtploidy[g + j] = lo[l].gploidy[j];
                s->gtaoff[g + j] = p + lo[l].gaoff[j];
            }
            std::copy(lo[l].galleles.begin(), lo[l].galleles.end(), s->pool.begin() + p);
            g += (int64_t)lo[l].gploidy.size();
            p += (int64_t)lo[l].galleles.size();
        }
    }
    th.clear();
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int64_t l = t; l < L; l += nt) {
                const int b = lo[l].bits;
                uint8_t* base = s->bits.data() + s->byteoff[l];
                for (int i = 0; i < N; ++i) {
                    const uint64_t bit = (uint64_t)i * b;
                    const uint32_t v = lo[l].ids[i];
                    for (int x = 0; x < b; ++x)
                        if ((v >> x) & 1) base[(bit + x) >> 3] |= (uint8_t)(1u << ((bit + x) & 7));
                }
                std::vector<uint32_t>().swap(lo[l].ids);
            }
        });
    for (auto& x : th) x.join();
    *out = s;
    return VPS_OK;
}

extern "C" int vps_synth_desc(const vps_synth* s, vps_dataset_desc* d)
{
    if (!s || !d) return fail(VPS_ERR_INVALID, "null argument");
    memset(d, 0, sizeof(*d));
    d->n_ind = s->N;
    d->n_loc = s->L;
    d->n_pop = s->S;
    d->ind_pop = s->pop.data();
    d->geno_bits = s->bits.data();
    d->geno_byte_offset = s->byteoff.data();
    d->bitsize = s->bitsize.data();
    d->n_allele = s->nallele.data();
    d->gtab_offset = s->gtoff.data();
    d->gt_ploidy = s->gtploidy.data();
    d->gt_allele_offset = s->gtaoff.data();
    d->allele_pool = s->pool.data();
    return VPS_OK;
}

extern "C" int vps_synth_truth(const vps_synth* s, double* q)
{
    if (!s || !q) return fail(VPS_ERR_INVALID, "null argument");
    memcpy(q, s->qtrue.data(), s->qtrue.size() * sizeof(double));
    return VPS_OK;
}

extern "C" int vps_synth_sizes(const vps_synth* s, int64_t* packed_bytes, int64_t* n_genotypes, int64_t* n_alleles)
{
    if (!s) return fail(VPS_ERR_INVALID, "null argument");
    if (packed_bytes) *packed_bytes = (int64_t)s->bits.size();
    if (n_genotypes) *n_genotypes = (int64_t)s->gtploidy.size();
    if (n_alleles) *n_alleles = (int64_t)s->pool.size();
    return VPS_OK;
}

extern "C" int vps_synth_destroy(vps_synth* s)
{
    delete s;
    return VPS_OK;
}
