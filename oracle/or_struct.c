/*
 * or_struct.c — oracle restatement of BAYESIAN<double> (src/structure.cpp)
 * with the genotype-pass semantics of the AVX path (src/structureAVX.cpp).
 * Test infrastructure only (see vps_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "vps_oracle.h"

static int g_threads = 1;
void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s ? s : 1); if (!p) abort(); return p; }

/* ------------------------------------------------------------------ */
/* dataset                                                              */
/* ------------------------------------------------------------------ */
struct or_data {
    int N, S, maxploidy, minploidy, maxK;
    int64_t L, KT;
    uint8_t*  bits;       /* packed ids, copied with 16 bytes of tail padding */
    uint64_t* byteoff;    /* [L] sorted */
    uint8_t*  bitsize;    /* [L] sorted */
    int32_t*  k2;         /* [L] sorted */
    int64_t*  afo;        /* [L] allele_freq_offset in sorted order */
    int64_t*  orig;       /* [L] original index of sorted locus */
    int64_t*  gtoff;      /* [L] first genotype of sorted locus */
    uint8_t*  gtploidy;   /* [G] */
    int64_t*  gtaoff;     /* [G] */
    uint16_t* alleles;    /* pool */
    int64_t*  vt;         /* [N] */
    int32_t*  popid;      /* [N] */
    int32_t*  pop_ind0;   /* [S] */
    int32_t*  pop_nind;   /* [S] */
    int       size_min, size_max;
    int64_t   lend[40];
    int64_t   coprime, coprime64[40];
    double*   totfreq;    /* [KT] total-population allele frequency (F model) */
    int32_t*  nhaplo;     /* [L] */
};

int64_t or_data_kt(const or_data* d) { return d->KT; }
int or_data_maxploidy(const or_data* d) { return d->maxploidy; }

static inline int gid_at(const or_data* d, int64_t l, int i)
{
    uint64_t bit = (uint64_t)i * d->bitsize[l];
    const uint8_t* p = d->bits + d->byteoff[l] + (bit >> 3);
    uint64_t w;
    memcpy(&w, p, 8);
    return (int)((w >> (bit & 7)) & ((1ull << d->bitsize[l]) - 1ull));
}
int or_data_gid(const or_data* d, int64_t l, int i) { return gid_at(d, l, i); }

/* IsCoprime, src/structure.cpp:2186-2198 */
static int is_coprime(int64_t x, int64_t y)
{
    int64_t z = 1;
    while (x % y != 0) { z = x % y; x = y; y = z; }
    return z == 1;
}

static const uint8_t* g_sort_bits;
static int cmp_locus(const void* a, const void* b)
{
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    if (g_sort_bits[x] != g_sort_bits[y]) return g_sort_bits[x] < g_sort_bits[y] ? -1 : 1;
    return x < y ? -1 : (x > y);
}

or_data* or_data_create(const vps_dataset_desc* in)
{
    or_data* d = (or_data*)xcalloc(1, sizeof(or_data));
    d->N = in->n_ind; d->L = in->n_loc; d->S = in->n_pop;
    int64_t L = d->L, N = d->N;

    /* total packed bytes */
    uint64_t nbytes = 0;
    for (int64_t l = 0; l < L; ++l) {
        uint64_t e = in->geno_byte_offset[l] + ((uint64_t)in->bitsize[l] * N + 7) / 8;
        if (e > nbytes) nbytes = e;
    }
    d->bits = (uint8_t*)xcalloc(nbytes + 16, 1);
    memcpy(d->bits, in->geno_bits, nbytes);

    /* locus order: ascending bitsize, ties by original index (QSLocus2,
       src/structure.cpp:2202-2273, a total order, hence a stable sort) */
    int64_t* idx = (int64_t*)xcalloc(L, sizeof(int64_t));
    for (int64_t l = 0; l < L; ++l) idx[l] = l;
    g_sort_bits = in->bitsize;
    qsort(idx, L, sizeof(int64_t), cmp_locus);

    d->byteoff = (uint64_t*)xcalloc(L, 8); d->bitsize = (uint8_t*)xcalloc(L, 1);
    d->k2 = (int32_t*)xcalloc(L, 4); d->afo = (int64_t*)xcalloc(L, 8);
    d->orig = idx; d->gtoff = (int64_t*)xcalloc(L, 8);
    int64_t G = in->gtab_offset[L];
    d->gtploidy = (uint8_t*)xcalloc(G, 1); memcpy(d->gtploidy, in->gt_ploidy, G);
    d->gtaoff = (int64_t*)xcalloc(G, 8); memcpy(d->gtaoff, in->gt_allele_offset, G * 8);
    int64_t na = 0;
    for (int64_t g = 0; g < G; ++g) if (in->gt_allele_offset[g] + in->gt_ploidy[g] > na) na = in->gt_allele_offset[g] + in->gt_ploidy[g];
    d->alleles = (uint16_t*)xcalloc(na + 1, 2); memcpy(d->alleles, in->allele_pool, na * 2);

    d->KT = 0; d->maxK = 0;
    for (int64_t l = 0; l < L; ++l) {
        int64_t o = idx[l];
        d->byteoff[l] = in->geno_byte_offset[o];
        d->bitsize[l] = in->bitsize[o];
        d->k2[l] = in->n_allele[o];
        d->gtoff[l] = in->gtab_offset[o];
        d->afo[l] = d->KT;                 /* src/structure.cpp:2374-2381 */
        d->KT += d->k2[l];
        if (d->k2[l] > d->maxK) d->maxK = d->k2[l];
    }

    /* vt, vmin/vmax -> maxploidy/minploidy (src/function.cpp:4904-4915) */
    d->vt = (int64_t*)xcalloc(N, 8);
    int* vmin = (int*)xcalloc(N, 4); int* vmax = (int*)xcalloc(N, 4);
    for (int i = 0; i < N; ++i) vmin[i] = 100;
    d->totfreq = (double*)xcalloc(d->KT, 8); d->nhaplo = (int32_t*)xcalloc(L, 4);
    for (int64_t l = 0; l < L; ++l) {
        double* p = d->totfreq + d->afo[l];
        int nh = 0;
        for (int i = 0; i < N; ++i) {
            int64_t g = d->gtoff[l] + gid_at(d, l, i);
            int v = d->gtploidy[g];
            if (!v) continue;
            d->vt[i] += v;
            if (v < vmin[i]) vmin[i] = v;
            if (v > vmax[i]) vmax[i] = v;
            nh += v;
            for (int a = 0; a < v; ++a) p[d->alleles[d->gtaoff[g] + a]] += 1;
        }
        d->nhaplo[l] = nh;
        or_unify_avx(p, d->k2[l]);        /* CalcFreqGcount, src/function.cpp:3994-4017 */
    }
    d->maxploidy = 0; d->minploidy = 100;
    for (int i = 0; i < N; ++i) {
        int mn = vmin[i] == 100 ? 0 : vmin[i];
        if (mn < d->minploidy) d->minploidy = mn;
        if (vmax[i] > d->maxploidy) d->maxploidy = vmax[i];
    }
    free(vmin); free(vmax);

    d->popid = (int32_t*)xcalloc(N, 4);
    d->pop_ind0 = (int32_t*)xcalloc(d->S, 4); d->pop_nind = (int32_t*)xcalloc(d->S, 4);
    for (int s = 0; s < d->S; ++s) d->pop_ind0[s] = -1;
    for (int i = 0; i < N; ++i) {
        int s = in->ind_pop ? in->ind_pop[i] : 0;
        d->popid[i] = s;
        if (d->pop_ind0[s] < 0) d->pop_ind0[s] = i;
        d->pop_nind[s]++;
    }
    for (int s = 0; s < d->S; ++s) if (d->pop_ind0[s] < 0) d->pop_ind0[s] = 0;

    /* bitsize groups and coprime strides, src/structure.cpp:2358-2395 */
    d->size_min = d->bitsize[0]; d->size_max = d->bitsize[L - 1];
    memset(d->lend, 0, sizeof(d->lend));
    int64_t csize = d->size_min;
    for (int64_t l = 0; l < L; ++l)
        while ((int64_t)d->bitsize[l] != csize) { d->lend[csize] = l; csize++; }
    d->lend[d->size_max] = L;
    d->coprime = L <= 67 ? 1 : 67;
    while (!is_coprime(L, d->coprime)) d->coprime++;
    if (d->coprime >= L) d->coprime = 1;
    for (int s = d->size_min; s <= d->size_max; ++s) {
        int64_t lstart = d->lend[s - 1], lend = d->lend[s];
        int64_t lend0 = (lend - lstart + 63) / 64;
        d->coprime64[s] = lend0 <= 11 ? 1 : 11;
        while (!is_coprime(lend0, d->coprime64[s])) d->coprime64[s]++;
        if (d->coprime >= L) d->coprime64[s] = 1;
    }
    return d;
}

void or_data_destroy(or_data* d)
{
    if (!d) return;
    free(d->bits); free(d->byteoff); free(d->bitsize); free(d->k2); free(d->afo); free(d->orig);
    free(d->gtoff); free(d->gtploidy); free(d->gtaoff); free(d->alleles); free(d->vt);
    free(d->popid); free(d->pop_ind0); free(d->pop_nind); free(d->totfreq); free(d->nhaplo);
    free(d);
}

void or_data_order(const or_data* d, int64_t* orig, int64_t* afo, int64_t* vt)
{
    if (orig) memcpy(orig, d->orig, d->L * 8);
    if (afo) memcpy(afo, d->afo, d->L * 8);
    if (vt) memcpy(vt, d->vt, d->N * 8);
}

/* ------------------------------------------------------------------ */
/* chain                                                                */
/* ------------------------------------------------------------------ */
struct or_chain {
    or_data* d;
    vps_params p;
    int K, id, N, S;
    int64_t L, KT;
    uint64_t seed;
    int m, nr, nlnL, rlen;
    int admix, locpriori, fmodel, inferlambda, difflambda, inferalpha, diffalpha, uniformalpha, fsame;
    int nburnin, nreps, nthinning, nadmburnin, metrofreq, writelnl;
    double lambda, stdlambda, maxlambda, alpha, stdalpha, maxalpha, alphapriora, alphapriorb;
    double r, maxr, epsr, epseta, epsgamma, pmeanf, pstdf, stdf;
    int binaryq, singlez, topology;
    double *Freq, *FreqM, *FreqA, *Lambda, *Alpha, *Q, *f, *F, *Eta, *Gamma, *Di, *AlphaLocal, *SumAlpha;
    int64_t *MiSum, *Mi, *kdis;
    int32_t* Ni;
    uint16_t* Z;
    double *bufNK1, *bufN1, *lnL, *rout, lastlnl;
    int misum_is_double;
    /* REAL = float (-g_float=single): every REAL array of BAYESIAN<float>
       holds single-precision values; fast32 adds -g_fastsingle=yes (the
       single-precision genotype passes and RNG<float> Z streams) */
    int real32, fast32;
};

/* a value stored into a REAL array */
static inline double rstore(const or_chain* c, double v) { return c->real32 ? (double)(float)v : v; }

uint64_t or_chain_seed(or_chain* c) { return c->seed; }

static void init_fix(or_chain* c)      /* InitFix, src/structure.cpp:468-553 */
{
    or_data* d = c->d;
    int K = c->K; int64_t KT = c->KT;
    c->nr = 0; c->r = 1;
    c->Freq = (double*)xcalloc(K * KT, 8);
    c->FreqM = (double*)xcalloc(K * KT, 8);
    c->FreqA = (double*)xcalloc(3 * KT, 8);
    c->bufNK1 = (double*)xcalloc((size_t)c->N * K + 64, 8);
    c->bufN1 = (double*)xcalloc(c->N, 8);
    c->MiSum = (int64_t*)xcalloc((size_t)c->N * K, 8);
    for (int64_t j = 0; j < K * KT; ++j) c->Freq[j] = 0.5;
    c->Ni = (int32_t*)xcalloc(K * KT, 4);
    c->Mi = (int64_t*)xcalloc((size_t)c->N * K, 8);
    c->Q = (double*)xcalloc((size_t)c->N * K, 8);
    for (int64_t j = 0; j < (int64_t)c->N * K; ++j) c->Q[j] = rstore(c, 1.0 / K);
    c->Alpha = (double*)xcalloc(K, 8); c->Lambda = (double*)xcalloc(K, 8);
    for (int k = 0; k < K; ++k) { c->Alpha[k] = c->alpha; c->Lambda[k] = c->lambda; }
    int nf = c->fmodel ? (c->fsame ? 1 : K) : 0;
    int na = c->admix ? (c->diffalpha ? K : 1) : 0;
    int nl = c->difflambda ? K : 1;
    c->rlen = c->locpriori ? 4 + nl + 1 + K + K * c->S + nf : 4 + nl + na + nf;
    c->rout = (double*)xcalloc(c->rlen, 8);
    /* the reference sizes this nburnin/nthinning + nreps/nthinning + 1, which is
       one short when both divisions truncate; the oracle sizes by ceiling */
    c->lnL = (double*)xcalloc((c->nburnin + c->nthinning - 1) / c->nthinning + (c->nreps + c->nthinning - 1) / c->nthinning + 1, 8);
    c->nlnL = 0;
    for (int64_t l = 0; l < c->L; ++l) {
        int k2 = d->k2[l];
        if (k2 != 2)
            for (int k = 0; k < K; ++k)
                for (int a = 0; a < k2; ++a) c->Freq[k * KT + d->afo[l] + a] = rstore(c, 1.0 / k2);
    }
}

static void init_admix(or_chain* c)    /* InitAdmix, src/structure.cpp:557-649 */
{
    or_data* d = c->d;
    int K = c->K, N = c->N; int64_t KT = c->KT;
    c->Z = (uint16_t*)xcalloc(N + 32, 2);
    c->singlez = c->nadmburnin == 0 && !c->admix;
    if (c->singlez) {
        or_rng rng; or_rng_init(&rng, c->seed - 1, OR_SALT_INITADM);
        for (int i = 0; i < N; ++i) {
            c->Z[i] = (uint16_t)or_rng_next(&rng, K);
            c->Mi[i * K + c->Z[i]] = d->vt[i];
        }
        for (int64_t l = 0; l < c->L; ++l) {
            for (int i = 0; i < N; ++i) {
                int64_t g = d->gtoff[l] + gid_at(d, l, i);
                int v = d->gtploidy[g];
                if (!v) continue;
                int32_t* ni = c->Ni + (int64_t)c->Z[i] * KT + d->afo[l];
                for (int a = 0; a < v; ++a) ni[d->alleles[d->gtaoff[g] + a]]++;
            }
        }
    } else {
        for (int64_t l = 0; l < c->L; ++l) {
            or_rng rng; or_rng_init(&rng, c->seed + l, OR_SALT_INITADM);
            for (int i = 0; i < N; ++i) {
                int64_t g = d->gtoff[l] + gid_at(d, l, i);
                int v = d->gtploidy[g];
                if (!v) continue;
                for (int a = 0; a < v; ++a) {
                    int k2 = (int)or_rng_next(&rng, K);
                    c->Mi[(int64_t)i * K + k2]++;
                    c->Ni[k2 * KT + d->afo[l] + d->alleles[d->gtaoff[g] + a]]++;
                }
            }
        }
    }
}

static void init_locpriori(or_chain* c)  /* src/structure.cpp:653-675 */
{
    int K = c->K, S = c->S;
    c->SumAlpha = (double*)xcalloc(S, 8);
    c->AlphaLocal = (double*)xcalloc(S * K, 8);
    c->Eta = (double*)xcalloc(K, 8);
    c->Gamma = (double*)xcalloc(S * K, 8);
    c->Di = (double*)xcalloc(S * K, 8);
    for (int j = 0; j < S * K; ++j) c->AlphaLocal[j] = c->alpha;
    for (int k = 0; k < K; ++k) c->Eta[k] = 1.0 / K;
    for (int j = 0; j < S * K; ++j) c->Gamma[j] = 1.0 / K;
}

static void init_fmodel(or_chain* c)     /* src/structure.cpp:679-708 */
{
    or_data* d = c->d;
    int K = c->K; int64_t KT = c->KT;
    c->f = (double*)xcalloc(K, 8); c->F = (double*)xcalloc(K, 8);
    if (!c->fmodel) return;
    for (int k = 0; k < K; ++k) { c->F[k] = c->pmeanf; c->f[k] = (1 - c->F[k]) / c->F[k]; }
    for (int k = 0; k < K; ++k) for (int64_t a = 0; a < KT; ++a) c->Freq[k * KT + a] = c->Lambda[k];
    double* p = c->FreqA;
    for (int64_t a = 0; a < KT; ++a) p[a] = c->lambda;
    for (int64_t l = 0; l < c->L; ++l) {
        int k2 = d->k2[l];
        const double* p2 = d->totfreq + d->afo[l];
        int nh = d->nhaplo[l];
        for (int a = 0; a < k2; ++a) p[a] = (c->lambda + p2[a] * nh) / (k2 * c->lambda + nh);
        p += k2;
    }
}

or_chain* or_chain_create(or_data* d, const vps_params* p, int K, int run_id)
{
    or_chain* c = (or_chain*)xcalloc(1, sizeof(or_chain));
    c->d = d; c->p = *p; c->K = K; c->id = run_id;
    c->N = d->N; c->L = d->L; c->KT = d->KT; c->S = d->S;
    /* ReadPar, src/structure.cpp:405-464 */
    c->admix = p->admix_val == 1; c->locpriori = p->locpriori_val == 1; c->fmodel = p->f_val == 1;
    c->nburnin = p->nburnin; c->nreps = p->nreps; c->nthinning = p->nthinning;
    c->lambda = p->lambda; c->stdlambda = p->stdlambda; c->maxlambda = p->maxlambda;
    c->inferlambda = p->inferlambda_val == 0; c->difflambda = p->difflambda_val == 1;
    c->alpha = p->alpha; c->inferalpha = p->inferalpha_val == 1; c->diffalpha = p->diffalpha_val == 1;
    c->uniformalpha = p->uniformalpha_val == 1; c->stdalpha = p->stdalpha; c->maxalpha = p->maxalpha;
    c->alphapriora = p->alphapriora; c->alphapriorb = p->alphapriorb; c->metrofreq = p->metrofreq;
    c->r = p->r; c->maxr = p->maxr; c->epsr = p->stdr; c->epseta = p->epseta; c->epsgamma = p->epsgamma;
    c->pmeanf = p->pmeanf; c->pstdf = p->pstdf; c->stdf = p->stdf; c->fsame = p->singlef_val == 1;
    c->writelnl = p->writelnl_val == 1;
    c->topology = p->rng_topology;
    c->real32 = p->precision == VPS_PREC_REAL32 || p->precision == VPS_PREC_FAST32;
    c->fast32 = p->precision == VPS_PREC_FAST32;
    if (c->real32 && c->fmodel) { free(c); return NULL; }   /* the F model keeps REAL = double here */
    c->kdis = (int64_t*)xcalloc(d->maxK + 1, 8);
    for (int64_t l = 0; l < c->L; ++l) c->kdis[d->k2[l]]++;
    c->nadmburnin = (!c->locpriori && !c->admix) ? p->nadmburnin : 0;
    c->nburnin += c->nadmburnin;
    c->seed = or_hash_ulong(p->g_seed + (uint64_t)run_id);
    c->m = 0;
    init_fix(c);
    init_admix(c);
    init_locpriori(c);
    init_fmodel(c);
    return c;
}

void or_chain_destroy(or_chain* c)
{
    if (!c) return;
    free(c->Freq); free(c->FreqM); free(c->FreqA); free(c->Lambda); free(c->Alpha); free(c->Q);
    free(c->f); free(c->F); free(c->Eta); free(c->Gamma); free(c->Di); free(c->AlphaLocal);
    free(c->SumAlpha); free(c->MiSum); free(c->Mi); free(c->kdis); free(c->Ni); free(c->Z);
    free(c->bufNK1); free(c->bufN1); free(c->lnL); free(c->rout);
    free(c);
}

void or_chain_sizes(or_chain* c, int64_t* rlen, int64_t* nlnl)
{
    if (rlen) *rlen = c->rlen;
    if (nlnl) *nlnl = c->nlnL;
}

static int state_ptr(or_chain* c, int which, void** ptr, size_t* bytes)
{
    int K = c->K, N = c->N, S = c->S; int64_t KT = c->KT;
    switch (which) {
    case VPS_STATE_FREQ:   *ptr = c->Freq;   *bytes = (size_t)K * KT * 8; return 0;
    case VPS_STATE_Q:      *ptr = c->Q;      *bytes = (size_t)N * K * 8; return 0;
    case VPS_STATE_Z:      *ptr = c->Z;      *bytes = (size_t)N * 2; return 0;
    case VPS_STATE_NI:     *ptr = c->Ni;     *bytes = (size_t)K * KT * 4; return 0;
    case VPS_STATE_MI:     *ptr = c->Mi;     *bytes = (size_t)N * K * 8; return 0;
    case VPS_STATE_ALPHA:  *ptr = c->Alpha;  *bytes = (size_t)K * 8; return 0;
    case VPS_STATE_LAMBDA: *ptr = c->Lambda; *bytes = (size_t)K * 8; return 0;
    case VPS_STATE_ITER:   *ptr = &c->m;     *bytes = 4; return 0;
    case VPS_STATE_BUFN1:  *ptr = c->bufN1;  *bytes = (size_t)N * 8; return 0;
    case VPS_STATE_BUFNK1: *ptr = c->bufNK1; *bytes = (size_t)N * K * 8; return 0;
    case VPS_STATE_LNL:    *ptr = &c->lastlnl; *bytes = 8; return 0;
    case VPS_STATE_MISUM:  *ptr = c->MiSum;  *bytes = (size_t)N * K * 8; return 0;
    case VPS_STATE_FREQM:  *ptr = c->FreqM;  *bytes = (size_t)K * KT * 8; return 0;
    case VPS_STATE_ROUT:   *ptr = c->rout;   *bytes = (size_t)c->rlen * 8; return 0;
    case VPS_STATE_FREQA:  *ptr = c->FreqA;  *bytes = (size_t)3 * KT * 8; return 0;
    case VPS_STATE_BINARYQ:*ptr = &c->binaryq; *bytes = 4; return 0;
    case VPS_STATE_LNLTRACE: *ptr = c->lnL;  *bytes = (size_t)c->nlnL * 8; return 0;
    default: (void)S; return -1;
    }
}

/* composite states are copied piecewise */
int or_chain_get_state(or_chain* c, int which, void* host, size_t bytes)
{
    int K = c->K, S = c->S;
    if (which == VPS_STATE_FST) {
        if (bytes != (size_t)2 * K * 8) return -1;
        memcpy(host, c->F, K * 8); memcpy((double*)host + K, c->f, K * 8); return 0;
    }
    if (which == VPS_STATE_LOC) {
        size_t need = (size_t)(1 + K + 3 * S * K + S) * 8;
        if (bytes != need) return -1;
        double* h = (double*)host;
        h[0] = c->r; memcpy(h + 1, c->Eta, K * 8);
        memcpy(h + 1 + K, c->Gamma, S * K * 8);
        memcpy(h + 1 + K + S * K, c->AlphaLocal, S * K * 8);
        memcpy(h + 1 + K + 2 * S * K, c->Di, S * K * 8);
        memcpy(h + 1 + K + 3 * S * K, c->SumAlpha, S * 8);
        return 0;
    }
    void* p; size_t b;
    if (state_ptr(c, which, &p, &b) || b != bytes) return -1;
    memcpy(host, p, b);
    return 0;
}

int or_chain_set_state(or_chain* c, int which, const void* host, size_t bytes)
{
    int K = c->K, S = c->S;
    if (which == VPS_STATE_FST) {
        if (bytes != (size_t)2 * K * 8) return -1;
        memcpy(c->F, host, K * 8); memcpy(c->f, (const double*)host + K, K * 8); return 0;
    }
    if (which == VPS_STATE_LOC) {
        size_t need = (size_t)(1 + K + 3 * S * K + S) * 8;
        if (bytes != need) return -1;
        const double* h = (const double*)host;
        c->r = h[0]; memcpy(c->Eta, h + 1, K * 8);
        memcpy(c->Gamma, h + 1 + K, S * K * 8);
        memcpy(c->AlphaLocal, h + 1 + K + S * K, S * K * 8);
        memcpy(c->Di, h + 1 + K + 2 * S * K, S * K * 8);
        memcpy(c->SumAlpha, h + 1 + K + 3 * S * K, S * 8);
        return 0;
    }
    void* p; size_t b;
    if (state_ptr(c, which, &p, &b) || b != bytes) return -1;
    memcpy(p, host, b);
    return 0;
}

/* ------------------------------------------------------------------ */
/* UpdateP, src/structure.cpp:812-862                                   */
/* ------------------------------------------------------------------ */
static void set_prior_freq(or_chain* c)
{
    int K = c->K; int64_t KT = c->KT;
    if (c->fmodel) {
        if (c->fsame) {
            for (int64_t a = 0; a < KT; ++a) c->Freq[a] = c->FreqA[a] * c->f[0];
            for (int k = 1; k < K; ++k) memcpy(c->Freq + k * KT, c->Freq, KT * 8);
        } else
            for (int k = 0; k < K; ++k)
                for (int64_t a = 0; a < KT; ++a) c->Freq[k * KT + a] = c->FreqA[a] * c->f[k];
    } else
        for (int k = 0; k < K; ++k)
            for (int64_t a = 0; a < KT; ++a) c->Freq[k * KT + a] = rstore(c, c->Lambda[k]);
}

void or_update_p(or_chain* c)
{
    or_data* d = c->d;
    int K = c->K; int64_t L = c->L, KT = c->KT;
    set_prior_freq(c);
    if (c->topology == VPS_RNG_REFERENCE) {
        /* 128 fixed RNG blocks over kl = k*L + l */
        #pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
        for (int b = 0; b < 128; ++b) {
            int64_t kl_start = (int64_t)b * K * L / 128, kl_end = (int64_t)(b + 1) * K * L / 128;
            or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m * 128 + b, OR_SALT_UPDATEP);
            for (int64_t kl = kl_start; kl < kl_end; ++kl) {
                int64_t k = kl / L, l = kl % L;
                int32_t* ni = c->Ni + k * KT + d->afo[l];
                double* p = c->Freq + k * KT + d->afo[l];
                if (c->real32) or_rng_dirichlet_ab32(&rng, p, p, ni, 0, d->k2[l]);
                else or_rng_dirichlet_ab(&rng, p, p, ni, 0, d->k2[l]);
            }
        }
    } else {
        /* parallel topology: one stream per (k, l) */
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int64_t kl = 0; kl < (int64_t)K * L; ++kl) {
            int64_t k = kl / L, l = kl % L;
            or_rng rng; or_rng_init(&rng, c->seed + ((uint64_t)c->m * K + k) * L + l, OR_SALT_UPDATEP);
            int32_t* ni = c->Ni + k * KT + d->afo[l];
            double* p = c->Freq + k * KT + d->afo[l];
            if (c->real32) or_rng_dirichlet_ab32(&rng, p, p, ni, 0, d->k2[l]);
            else or_rng_dirichlet_ab(&rng, p, p, ni, 0, d->k2[l]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* 64-locus tile helpers for the AVX-order reductions                   */
/* ------------------------------------------------------------------ */
/* product of 64 lanes in the order of the AVX tree: 16 vectors of 4 lanes
   (vector kk = lanes 4kk..4kk+3) folded kk+8, kk+4, kk+2, kk+1, then
   (v0*v2)*(v1*v3) (_mm256_reduce_mul_pd, src/mathAVX.h:88-92) */
static double tree64(const double* f)
{
    double v[16][4];
    for (int kk = 0; kk < 16; ++kk) for (int e = 0; e < 4; ++e) v[kk][e] = f[4 * kk + e];
    for (int KK = 8; KK >= 1; KK >>= 1)
        for (int kk = 0; kk < KK; ++kk) for (int e = 0; e < 4; ++e) v[kk][e] = v[kk][e] * v[kk + KK][e];
    return (v[0][0] * v[0][2]) * (v[0][1] * v[0][3]);
}

typedef struct tile { int64_t l0; int n; } tile;

/* tiles in the single-subthread processing order: bitsize groups ascending,
   within a group ia = 0.. mapped by the coprime stride (src/structureAVX.cpp:
   305-312) */
static tile* make_tiles(const or_data* d, int64_t* ntiles)
{
    int64_t cnt = 0;
    for (int s = d->size_min; s <= d->size_max; ++s) cnt += (d->lend[s] - d->lend[s - 1] + 63) / 64;
    tile* t = (tile*)xcalloc(cnt, sizeof(tile));
    int64_t j = 0;
    for (int s = d->size_min; s <= d->size_max; ++s) {
        int64_t lstart = d->lend[s - 1], lend = d->lend[s];
        int64_t lend0 = (lend - lstart + 63) / 64;
        for (int64_t ia = 0; ia < lend0; ++ia) {
            int64_t l = lstart + ((ia * d->coprime64[s] % lend0) << 6);
            t[j].l0 = l;
            t[j].n = (int)(lend - l >= 64 ? 64 : lend - l);
            j++;
        }
    }
    *ntiles = cnt;
    return t;
}

/* per-individual maxv of a tile (src/structureAVX.cpp:364-370) */
static int tile_maxv(const or_data* d, const tile* t, const int* vv)
{
    if (d->maxploidy == d->minploidy) return d->maxploidy;
    int mv = 0;
    for (int j = 0; j < t->n; ++j) if (vv[j] > mv) mv = vv[j];
    return mv;
}

/* ------------------------------------------------------------------ */
/* UpdateQAdmix, src/structure.cpp:986-1004                             */
/* ------------------------------------------------------------------ */
static const double* alpha_row(or_chain* c, int i)
{
    return c->locpriori ? c->AlphaLocal + (int64_t)c->d->popid[i] * c->K : c->Alpha;
}

void or_update_q_admix(or_chain* c)
{
    int K = c->K, N = c->N;
    c->binaryq = 0;
    if (c->topology == VPS_RNG_REFERENCE) {
        or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m, OR_SALT_UPDATEQ);
        for (int i = 0; i < N; ++i)
            if (c->d->vt[i]) (c->real32 ? or_rng_dirichlet_ab32 : or_rng_dirichlet_ab)(&rng, c->Q + (int64_t)i * K, alpha_row(c, i), c->Mi + (int64_t)i * K, 1, K);
    } else {
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int i = 0; i < N; ++i) {
            if (!c->d->vt[i]) continue;
            or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m * N + i, OR_SALT_UPDATEQ);
            (c->real32 ? or_rng_dirichlet_ab32 : or_rng_dirichlet_ab)(&rng, c->Q + (int64_t)i * K, alpha_row(c, i), c->Mi + (int64_t)i * K, 1, K);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Metropolis log ratio pass, UpdateQMetroAVX body, src/structureAVX.cpp: */
/* 600-780: per (i, tile, slot) the 64-lane product of Σq'p / Σqp is     */
/* charged into individual i's accumulator, tiles in processing order.  */
/* ------------------------------------------------------------------ */
void or_q_metro_logratio(or_chain* c, const double* qprop, double* out)
{
    or_data* d = c->d;
    int K = c->K, N = c->N; int64_t KT = c->KT;
    int64_t nt; tile* tl = make_tiles(d, &nt);
    #pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
    for (int i = 0; i < N; ++i) {
        int64_t slog = 0; double prod = 1;
        const double* q = c->Q + (int64_t)i * K;
        const double* qi = qprop + (int64_t)i * K;
        for (int64_t t = 0; t < nt; ++t) {
            int vv[64]; const uint16_t* als[64];
            for (int j = 0; j < 64; ++j) { vv[j] = 0; als[j] = NULL; }
            for (int j = 0; j < tl[t].n; ++j) {
                int64_t l = tl[t].l0 + j;
                int64_t g = d->gtoff[l] + gid_at(d, l, i);
                vv[j] = d->gtploidy[g]; als[j] = d->alleles + d->gtaoff[g];
            }
            int maxv = tile_maxv(d, &tl[t], vv);
            for (int a = 0; a < maxv; ++a) {
                double f[64];
                for (int j = 0; j < 64; ++j) {
                    if (j < tl[t].n && a < vv[j]) {
                        int64_t l = tl[t].l0 + j;
                        const double* p = c->Freq + d->afo[l] + als[j][a];
                        if (c->fast32) {
                            /* the fast single-precision branch, src/structureAVX.cpp:725-760 */
                            float f1 = 0, f2 = 0;
                            for (int k = 0; k < K; ++k) {
                                f1 = f1 + (float)p[k * KT] * (float)qi[k];
                                f2 = f2 + (float)p[k * KT] * (float)q[k];
                            }
                            f[j] = (double)(f1 / f2);
                        } else {
                            double f1 = 0, f2 = 0;
                            for (int k = 0; k < K; ++k) {
                                f1 = f1 + p[k * KT] * qi[k];
                                f2 = f2 + p[k * KT] * q[k];
                            }
                            f[j] = f1 / f2;
                        }
                    } else f[j] = 1.0;
                }
                or_charge_log(&slog, &prod, tree64(f));
            }
        }
        out[i] = or_close_log(slog, prod);
    }
    free(tl);
}

/* UpdateQMetro, src/structure.cpp:1090-1136 with the AVX body */
void or_update_q_metro(or_chain* c)
{
    int K = c->K, N = c->N;
    c->binaryq = 0;
    double* bufi = c->bufNK1;
    if (c->topology == VPS_RNG_REFERENCE) {
        or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m, OR_SALT_UPDATEQ);
        for (int i = 0; i < N; ++i)
            if (c->d->vt[i]) (c->real32 ? or_rng_dirichlet_af32 : or_rng_dirichlet_af)(&rng, bufi + (int64_t)i * K, alpha_row(c, i), K, 1.0);
        or_q_metro_logratio(c, bufi, c->bufN1);
        for (int i = 0; i < N; ++i) {
            if (!c->d->vt[i]) continue;
            if (c->bufN1[i] >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(c->bufN1[i]))
                memcpy(c->Q + (int64_t)i * K, bufi + (int64_t)i * K, K * 8);
        }
    } else {
        or_rng* rs = (or_rng*)xcalloc(N, sizeof(or_rng));
        for (int i = 0; i < N; ++i) {
            or_rng_init(&rs[i], c->seed + (uint64_t)c->m * N + i, OR_SALT_UPDATEQ);
            if (c->d->vt[i]) (c->real32 ? or_rng_dirichlet_af32 : or_rng_dirichlet_af)(&rs[i], bufi + (int64_t)i * K, alpha_row(c, i), K, 1.0);
        }
        or_q_metro_logratio(c, bufi, c->bufN1);
        for (int i = 0; i < N; ++i) {
            if (!c->d->vt[i]) continue;
            if (c->bufN1[i] >= OR_NZERO || or_rng_uniform(&rs[i], 0, 1) < exp(c->bufN1[i]))
                memcpy(c->Q + (int64_t)i * K, bufi + (int64_t)i * K, K * 8);
        }
        free(rs);
    }
}

/* ------------------------------------------------------------------ */
/* UpdateQNoAdmix, src/structureAVX.cpp:254-428                         */
/* ------------------------------------------------------------------ */
void or_q_noadmix_loglik(or_chain* c, double* out)
{
    or_data* d = c->d;
    int K = c->K, N = c->N; int64_t KT = c->KT;
    int64_t nt; tile* tl = make_tiles(d, &nt);
    #pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
    for (int i = 0; i < N; ++i) {
        int64_t slog[VPS_MAX_K * 4]; double prod[VPS_MAX_K * 4];
        for (int k = 0; k < K; ++k) { slog[k] = 0; prod[k] = 1; }
        if (c->locpriori && d->vt[i])
            for (int k = 0; k < K; ++k) or_charge_log(&slog[k], &prod[k], c->Gamma[(int64_t)d->popid[i] * K + k]);
        for (int64_t t = 0; t < nt; ++t) {
            int vv[64]; const uint16_t* als[64];
            for (int j = 0; j < 64; ++j) { vv[j] = 0; als[j] = NULL; }
            for (int j = 0; j < tl[t].n; ++j) {
                int64_t l = tl[t].l0 + j;
                int64_t g = d->gtoff[l] + gid_at(d, l, i);
                vv[j] = d->gtploidy[g]; als[j] = d->alleles + d->gtaoff[g];
            }
            int maxv = tile_maxv(d, &tl[t], vv);
            for (int a = 0; a < maxv; ++a)
                for (int k = 0; k < K; ++k) {
                    double f[64];
                    for (int j = 0; j < 64; ++j)
                        f[j] = (j < tl[t].n && a < vv[j]) ? c->Freq[k * KT + d->afo[tl[t].l0 + j] + als[j][a]] : 1.0;
                    or_charge_log(&slog[k], &prod[k], tree64(f));
                }
        }
        for (int k = 0; k < K; ++k) out[(int64_t)i * K + k] = or_close_log(slog[k], prod[k]);
    }
    free(tl);
}

void or_update_q_noadmix(or_chain* c)
{
    int K = c->K, N = c->N;
    memset(c->Q, 0, (size_t)N * K * 8);
    or_q_noadmix_loglik(c, c->bufNK1);
    or_rng rng;
    if (c->topology == VPS_RNG_REFERENCE) or_rng_init(&rng, c->seed + (uint64_t)c->m, OR_SALT_UPDATEQ);
    for (int i = 0; i < N; ++i) {
        if (!c->d->vt[i]) continue;
        if (c->topology != VPS_RNG_REFERENCE) or_rng_init(&rng, c->seed + (uint64_t)c->m * N + i, OR_SALT_UPDATEQ);
        uint16_t k2 = (uint16_t)or_rng_polylog(&rng, c->bufNK1 + (int64_t)i * K, K);
        c->Q[(int64_t)i * K + k2] = 1;
        c->Z[i] = k2;
    }
    c->binaryq = 1;
}

/* UpdateQ, src/structure.cpp:1140-1166 */
void or_update_q(or_chain* c)
{
    int K = c->K, N = c->N;
    if (!c->admix) {
        if (c->m >= c->nadmburnin) or_update_q_noadmix(c);
        else or_update_q_admix(c);
        if (c->locpriori) {
            memset(c->Di, 0, (size_t)c->S * K * 8);
            for (int i = 0; i < N; ++i)
                if (c->d->vt[i])
                    for (int k = 0; k < K; ++k) c->Di[(int64_t)c->d->popid[i] * K + k] += c->Q[(int64_t)i * K + k];
        }
    } else {
        if (c->metrofreq > 0 && c->m % c->metrofreq == 0) or_update_q_metro(c);
        else or_update_q_admix(c);
    }
}

/* ------------------------------------------------------------------ */
/* UpdateLocPriori, src/structure.cpp:1170-1255                          */
/* ------------------------------------------------------------------ */
void or_update_locpriori(or_chain* c)
{
    if (!c->locpriori) return;
    int K = c->K, S = c->S;
    or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m, OR_SALT_UPDATELOC);
    double r = c->r;
    if (c->admix) {
        double rm = or_rng_uniform(&rng, r - c->epsr, r + c->epsr);
        if (rm > 0 && rm < c->maxr) {
            double dlnL = 0, dd = rm - r, dt1 = rm * or_mylog(rm) - r * or_mylog(r);
            for (int k = 0; k < K; ++k) {
                dlnL += S * (c->Alpha[k] * dt1 - lgamma(rm * c->Alpha[k]) + lgamma(r * c->Alpha[k]));
                double sa = 0;
                for (int s = 0; s < S; ++s) sa += c->AlphaLocal[s * K + k];
                dlnL += c->Alpha[k] * dd * or_logprod_sep(c->AlphaLocal + k, S, K) - sa * dd;
            }
            if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) c->r = rm;
        }
    } else {
        double rm = or_rng_uniform(&rng, r - c->epsr, r + c->epsr);
        if (rm > 0 && rm < c->maxr) {
            double dlnL = lgamma(rm) - lgamma(r);
            for (int k = 0; k < K; ++k) dlnL += lgamma(r * c->Eta[k]) - lgamma(rm * c->Eta[k]);
            dlnL *= S;
            for (int k = 0; k < K; ++k) dlnL += (rm - r) * c->Eta[k] * or_logprod_sep(c->Gamma + k, S, K);
            if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) c->r = rm;
        }
        r = c->r;
        if (K >= 2) {
            int64_t i1 = (int64_t)or_rng_next(&rng, K), j1 = (int64_t)or_rng_next_avoid(&rng, K, i1);
            double delta = or_rng_uniform(&rng, 0, c->epseta);
            double e1 = c->Eta[i1] + delta, e2 = c->Eta[j1] - delta;
            if (e1 < 1 && e2 > 0) {
                double dlnL = S * (lgamma(r * c->Eta[i1]) + lgamma(r * c->Eta[j1]) - lgamma(r * e1) - lgamma(r * e2));
                /* LogProdDiv(A, B, S, K), src/mathAVX.cpp (scalar loop) */
                int64_t s1 = 0, s2 = 0; double p1 = 1, p2 = 1;
                for (int s = 0; s < S; ++s) {
                    or_charge_log(&s1, &p1, c->Gamma[s * K + i1]);
                    or_charge_log(&s2, &p2, c->Gamma[s * K + j1]);
                }
                dlnL += r * delta * (or_close_log(s1, p1) - or_close_log(s2, p2));
                if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) {
                    c->Eta[i1] += delta;
                    c->Eta[j1] -= delta;
                }
            }
        }
        if (K >= 2) {
            for (int s = 0; s < S; ++s) {
                int64_t i1 = (int64_t)or_rng_next(&rng, K), j1 = (int64_t)or_rng_next_avoid(&rng, K, i1);
                double delta = or_rng_uniform(&rng, 0, c->epsgamma);
                double* G = c->Gamma + (int64_t)s * K;
                if (G[i1] + delta < 1 && G[j1] - delta > 0) {
                    double dlnL = (r * c->Eta[i1] - 1.0 + c->Di[s * K + i1]) * or_mylog(1 + delta / G[i1]) +
                                  (r * c->Eta[j1] - 1.0 + c->Di[s * K + j1]) * or_mylog(1 - delta / G[j1]);
                    if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) {
                        G[i1] += delta;
                        G[j1] -= delta;
                    }
                }
            }
        }
        if (c->m % 10 == 0) {
            or_unify_avx(c->Eta, K);
            for (int s = 0; s < S; ++s) or_unify_avx(c->Gamma + (int64_t)s * K, K);
        }
    }
}

/* ------------------------------------------------------------------ */
/* UpdateZNoAdmix, src/structureAVX.cpp:432-552                          */
/* ------------------------------------------------------------------ */
void or_update_z_noadmix(or_chain* c)
{
    or_data* d = c->d;
    int K = c->K, N = c->N; int64_t KT = c->KT;
    memset(c->Mi, 0, (size_t)N * K * 8);
    memset(c->Ni, 0, (size_t)K * KT * 4);
    for (int i = 0; i < N; ++i) c->Mi[(int64_t)i * K + c->Z[i]] = d->vt[i];
    #pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads)
    for (int64_t l = 0; l < c->L; ++l) {
        for (int i = 0; i < N; ++i) {
            int64_t g = d->gtoff[l] + gid_at(d, l, i);
            int v = d->gtploidy[g];
            int32_t* ni = c->Ni + (int64_t)c->Z[i] * KT + d->afo[l];
            for (int a = 0; a < v; ++a) ni[d->alleles[d->gtaoff[g] + a]]++;
        }
    }
}

/* ------------------------------------------------------------------ */
/* UpdateZAdmix, src/structureAVX.cpp:788-981: per locus one stream      */
/* RNG(seed + m*L + l, UPDATEZ) (lane j of RNGAVX is seeded seed + j,     */
/* src/mathAVX.cpp:24-71); exactly one draw per (i, slot < maxploidy).    */
/* ------------------------------------------------------------------ */
void or_update_z_admix(or_chain* c)
{
    or_data* d = c->d;
    int K = c->K, N = c->N; int64_t KT = c->KT, L = c->L;
    int mp = d->maxploidy;
    memset(c->Ni, 0, (size_t)K * KT * 4);
    int nth = g_threads;
    int64_t* mi_all = (int64_t*)xcalloc((size_t)nth * N * K, 8);
    #pragma omp parallel num_threads(nth)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int64_t* mi = mi_all + (size_t)tid * N * K;
        double w[VPS_MAX_K * 4];
        float w32[VPS_MAX_K * 4];
        #pragma omp for schedule(dynamic, 64)
        for (int64_t l = 0; l < L; ++l) {
            or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m * L + l, OR_SALT_UPDATEZ);
            or_rngf rngf; or_rngf_init(&rngf, c->seed + (uint64_t)c->m * L + l, OR_SALT_UPDATEZ);
            const double* p = c->Freq + d->afo[l];
            int32_t* ni = c->Ni + d->afo[l];
            for (int i = 0; i < N; ++i) {
                int64_t g = d->gtoff[l] + gid_at(d, l, i);
                int v = d->gtploidy[g];
                const uint16_t* als = d->alleles + d->gtaoff[g];
                const double* q = c->Q + (int64_t)i * K;
                for (int a = 0; a < mp; ++a) {
                    if (a >= v) {
                        if (c->fast32) or_rngf_xorshift(&rngf);
                        else or_rng_xorshift(&rng);
                        continue;
                    }
                    int k2;
                    if (c->fast32) {
                        /* REAL = float with -g_fastsingle: weights and draw in single
                           precision from RNG<float> (src/structureAVX.cpp:942-947) */
                        for (int k = 0; k < K; ++k) w32[k] = (float)p[k * KT + als[a]] * (float)q[k] + (float)OR_MIN_FREQ;
                        k2 = or_rngf_poly_avx(&rngf, w32, K);
                    } else {
                        for (int k = 0; k < K; ++k) w[k] = p[k * KT + als[a]] * q[k] + OR_MIN_FREQ;
                        k2 = or_rng_poly_avx(&rng, w, K);
                    }
                    mi[(int64_t)i * K + k2]++;
                    ni[k2 * KT + als[a]]++;
                }
            }
        }
    }
    memset(c->Mi, 0, (size_t)N * K * 8);
    for (int t = 0; t < nth; ++t)
        for (int64_t j = 0; j < (int64_t)N * K; ++j) c->Mi[j] += mi_all[(size_t)t * N * K + j];
    free(mi_all);
}

void or_update_z(or_chain* c)
{
    if (c->binaryq) or_update_z_noadmix(c);
    else or_update_z_admix(c);
}

/* ------------------------------------------------------------------ */
/* UpdateAlpha, src/structure.cpp:1501-1594                              */
/* ------------------------------------------------------------------ */
void or_update_alpha(or_chain* c)
{
    or_data* d = c->d;
    int K = c->K, N = c->N, S = c->S;
    or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m, OR_SALT_UPDATEALPHA);
    if (c->locpriori && c->admix) {
        double r = c->r;
        for (int k = 0; k < K; ++k) {
            double am = or_rng_normal(&rng, c->Alpha[k], c->stdalpha);
            if (am > 0 && am < c->maxalpha) {
                double dd = am - c->Alpha[k];
                double dlnL = S * (dd * r * or_mylog(r) + lgamma(r * c->Alpha[k]) - lgamma(r * am));
                dlnL += r * dd * or_logprod_sep(c->AlphaLocal + k, S, K);
                if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) c->Alpha[k] = am;
            }
        }
        if (c->m % 10 == 0)
            for (int s = 0; s < S; ++s) {
                double sm = 0;
                for (int k = 0; k < K; ++k) sm += c->AlphaLocal[s * K + k];
                c->SumAlpha[s] = sm;
            }
        for (int s = 0; s < S; ++s) {
            int snind = d->pop_nind[s];
            const double* sq = c->Q + (int64_t)d->pop_ind0[s] * K;
            for (int k = 0; k < K; ++k) {
                double ao = c->AlphaLocal[s * K + k], am = or_rng_normal(&rng, ao, c->stdalpha);
                if (am > 0 && am < c->maxalpha) {
                    double dd = am - ao;
                    double dlnL = (r * c->Alpha[k] - 1) * or_mylog(am / ao) - dd * r +
                        d->pop_nind[s] * (lgamma(c->SumAlpha[s] + dd) - lgamma(am) - lgamma(c->SumAlpha[s]) + lgamma(ao));
                    dlnL += dd * or_logprod_sep(sq + k, snind, K);
                    if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) {
                        c->AlphaLocal[s * K + k] = am;
                        c->SumAlpha[s] += dd;
                    }
                }
            }
        }
    } else if (c->inferalpha) {
        if (!c->admix && c->m >= c->nadmburnin) return;
        if (c->diffalpha) {
            double sumalpha = 0;
            for (int k = 0; k < K; ++k) sumalpha += c->Alpha[k];
            for (int k = 0; k < K; ++k) {
                double am = or_rng_normal(&rng, c->Alpha[k], c->stdalpha);
                if (am <= 0 || (am >= c->maxalpha && c->uniformalpha)) continue;
                double ao = c->Alpha[k];
                double dd = am - ao;
                double dlnL = c->uniformalpha ? 0 : ((c->alphapriora - 1) * or_mylog(am / ao) + (ao - am) / c->alphapriorb);
                dlnL += (am - ao) * or_logprod_sep(c->Q + k, N, K) +
                        (lgamma(sumalpha + dd) - lgamma(sumalpha) - lgamma(am) + lgamma(ao)) * N;
                if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) c->Alpha[k] = am;
            }
        } else {
            double am = or_rng_normal(&rng, c->Alpha[0], c->stdalpha);
            if (am <= 0 || (am >= c->maxalpha && c->uniformalpha)) return;
            double ao = c->Alpha[0];
            double dlnL = c->uniformalpha ? 0 : ((c->alphapriora - 1) * or_mylog(am / ao) + (ao - am) / c->alphapriorb);
            dlnL += (am - ao) * or_logprod_avx(c->Q, (int64_t)N * K) +
                    (lgamma(K * am) - lgamma(K * ao) - K * lgamma(am) + K * lgamma(ao)) * N;
            if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL))
                for (int k = 0; k < K; ++k) c->Alpha[k] = am;
        }
    }
}

/* ------------------------------------------------------------------ */
/* UpdateLambda, src/structure.cpp:1598-1635 (inferlambda is false for    */
/* every value the option parser can produce, src/structure.cpp:430)     */
/* ------------------------------------------------------------------ */
void or_update_lambda(or_chain* c)
{
    if (!c->inferlambda) return;
    or_data* d = c->d;
    int K = c->K; int64_t KT = c->KT;
    or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m, OR_SALT_UPDATELAMBDA);
    if (c->difflambda) {
        for (int k = 0; k < K; ++k) {
            double lo = c->Lambda[k], lm = or_rng_normal(&rng, lo, c->stdlambda);
            if (lm <= 0 || lm >= c->maxlambda) continue;
            double dlnL = 0, gd = lgamma(lo) - lgamma(lm);
            for (int i = 2; i <= d->maxK; ++i) dlnL += c->kdis[i] * (lgamma(i * lm) - lgamma(i * lo) + i * gd);
            dlnL += (lm - lo) * or_logprod_avx(c->Freq + k * KT, KT);
            if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) c->Lambda[k] = lm;
        }
    } else {
        int np = c->fmodel ? 1 : K;
        double lo = c->Lambda[0], lm = or_rng_normal(&rng, lo, c->stdlambda);
        if (lm <= 0 || lm >= c->maxlambda) return;
        double dlnL = 0, gd = lgamma(lo) - lgamma(lm);
        for (int i = 2; i <= d->maxK; ++i) dlnL += c->kdis[i] * np * (lgamma(i * lm) - lgamma(i * lo) + i * gd);
        dlnL += (lm - lo) * (c->fmodel ? or_logprod_avx(c->FreqA, KT) : or_logprod_avx(c->Freq, K * KT));
        if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL))
            for (int k = 0; k < K; ++k) c->Lambda[k] = lm;
    }
}

/* ------------------------------------------------------------------ */
/* UpdateA / A1 / A2, src/structure.cpp:1639-1804                        */
/* ------------------------------------------------------------------ */
static void update_a1_locus(or_chain* c, or_rng* rng, int64_t l, double* pb, double* pc)
{
    or_data* d = c->d;
    int K = c->K; int64_t KT = c->KT;
    double lam = c->Lambda[0];
    int k2 = d->k2[l];
    double* pa = c->FreqA + d->afo[l];
    for (int a = 0; a < k2; ++a) pb[a] = lam;
    for (int k = 0; k < K; ++k)
        for (int a = 0; a < k2; ++a) pb[a] += c->Freq[k * KT + d->afo[l] + a] * c->f[k];
    or_rng_dirichlet_af(rng, pc, pb, k2, 1.0);
    double dlnL = 0;
    for (int a = 0; a < k2; ++a) dlnL += (pb[a] - lam) * (or_mylog(pa[a] / pc[a]));
    for (int k = 0; k < K; ++k) {
        const double* pk = c->Freq + k * KT + d->afo[l];
        double fr = c->f[k];
        for (int a = 0; a < k2; ++a)
            dlnL += lgamma(fr * pa[a]) - lgamma(fr * pc[a]) + fr * (pc[a] - pa[a]) * or_mylog(pk[a]);
    }
    if (dlnL >= OR_NZERO || or_rng_uniform(rng, 0, 1) < exp(dlnL))
        memcpy(pa, pc, k2 * 8);
}

static void update_a2_locus(or_chain* c, or_rng* rng, int64_t l)
{
    or_data* d = c->d;
    int K = c->K; int64_t KT = c->KT;
    double lam = c->Lambda[0];
    int k2 = d->k2[l];
    double* pa = c->FreqA + d->afo[l];
    const double* p = c->Freq + d->afo[l];
    int m1 = (int)or_rng_next(rng, k2), n1 = (int)or_rng_next_avoid(rng, k2, m1);
    double delta = or_rng_uniform(rng, 0, pow(c->N, -0.5));
    double pm0 = pa[m1], pm1 = pm0 + delta, pn0 = pa[n1], pn1 = pn0 - delta;
    if (pm1 >= 1 || pn1 <= 0) return;
    double dlnL = (lam - 1) * or_mylog(pm1 * pn1 / pm0 / pn0);
    for (int k = 0; k < K; ++k) {
        double fr = c->f[k];
        dlnL += lgamma(fr * pm0) + lgamma(fr * pn0) - lgamma(fr * pm1) - lgamma(fr * pn1) +
                (fr * delta) * or_mylog(p[k * KT + m1] / p[k * KT + n1]);
    }
    if (dlnL >= OR_NZERO || or_rng_uniform(rng, 0, 1) < exp(dlnL)) { pa[m1] = pm1; pa[n1] = pn1; }
}

void or_update_a(or_chain* c)
{
    if (!c->fmodel) return;
    or_data* d = c->d;
    int use_a1 = or_hash_ulong(c->seed + (uint64_t)c->m) & 1;
    int64_t L = c->L;
    if (c->topology == VPS_RNG_REFERENCE) {
        #pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
        for (int b = 0; b < 32; ++b) {
            double pb[VPS_MAX_ALLELES * 4 + 16], pc[VPS_MAX_ALLELES * 4 + 16];
            int64_t l_start = (int64_t)b * L / 32, l_end = (int64_t)(b + 1) * L / 32;
            or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m * 32 + b, OR_SALT_UPDATEA);
            for (int64_t l = l_start; l < l_end; ++l) {
                if (use_a1) update_a1_locus(c, &rng, l, pb, pc);
                else update_a2_locus(c, &rng, l);
            }
        }
    } else {
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int64_t l = 0; l < L; ++l) {
            double pb[VPS_MAX_ALLELES * 4 + 16], pc[VPS_MAX_ALLELES * 4 + 16];
            or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m * L + l, OR_SALT_UPDATEA);
            if (use_a1) update_a1_locus(c, &rng, l, pb, pc);
            else update_a2_locus(c, &rng, l);
        }
    }
    (void)d;
}

/* UpdateF, src/structure.cpp:1808-1858 */
void or_update_f(or_chain* c)
{
    if (!c->fmodel) return;
    int K = c->K; int64_t KT = c->KT;
    or_rng rng; or_rng_init(&rng, c->seed + (uint64_t)c->m, OR_SALT_UPDATEF);
    for (int k = 0; k < K; ++k) {
        double Fo = c->F[k];
        double Fm = or_rng_normal(&rng, Fo, c->stdf);
        if (Fm < 0 || Fm > 1) { if (c->fsame) break; else continue; }
        double fo = c->f[k], fm = (1 - Fm) / Fm;
        double dlnL = c->pmeanf * (Fo - Fm) / (c->pstdf * c->pstdf) +
            (c->pmeanf * c->pmeanf / (c->pstdf * c->pstdf) - 1) * or_mylog(Fm / Fo) +
            (double)((int64_t)(c->fsame ? K : 1) * c->L) * (lgamma(fm) - lgamma(fo));
        for (int k2 = k; k2 < K; ++k2) {
            const double* p = c->Freq + k2 * KT;
            const double* pa = c->FreqA;
            for (int64_t a = 0; a < KT; ++a) {
                double pa2 = pa[a];
                dlnL += pa2 * (fm - fo) * or_mylog(p[a]) + lgamma(fo * pa2) - lgamma(fm * pa2);
            }
            if (!c->fsame) break;
        }
        if (dlnL >= OR_NZERO || or_rng_uniform(&rng, 0, 1) < exp(dlnL)) {
            if (c->fsame) { for (int j = 0; j < K; ++j) { c->F[j] = Fm; c->f[j] = fm; } }
            else { c->F[k] = Fm; c->f[k] = fm; }
        }
        if (c->fsame) break;
    }
}

/* ------------------------------------------------------------------ */
/* Record, RecordAVX body src/structureAVX.cpp:1006-1187: per tile, per  */
/* individual, per slot the 64-lane product is charged into one          */
/* accumulator.  Products are formed in parallel and charged in order.  */
/* ------------------------------------------------------------------ */
double or_record_pass(or_chain* c)
{
    or_data* d = c->d;
    int K = c->K, N = c->N; int64_t KT = c->KT;
    for (int64_t j = 0; j < (int64_t)N * K; ++j) c->MiSum[j] += c->Mi[j];
    int64_t nt; tile* tl = make_tiles(d, &nt);
    int mp = d->maxploidy;
    double* prods = (double*)xcalloc((size_t)nt * N * mp, 8);
    int* nslot = (int*)xcalloc((size_t)nt * N, 4);
    int isadmix = !c->binaryq;
    #pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int64_t t = 0; t < nt; ++t) {
        for (int i = 0; i < N; ++i) {
            int vv[64]; const uint16_t* als[64];
            for (int j = 0; j < 64; ++j) { vv[j] = 0; als[j] = NULL; }
            for (int j = 0; j < tl[t].n; ++j) {
                int64_t l = tl[t].l0 + j;
                int64_t g = d->gtoff[l] + gid_at(d, l, i);
                vv[j] = d->gtploidy[g]; als[j] = d->alleles + d->gtaoff[g];
            }
            int maxv = tile_maxv(d, &tl[t], vv);
            nslot[t * N + i] = maxv;
            const double* q = c->Q + (int64_t)i * K;
            for (int a = 0; a < maxv; ++a) {
                double f[64];
                for (int j = 0; j < 64; ++j) {
                    if (j < tl[t].n && a < vv[j]) {
                        const double* p = c->Freq + d->afo[tl[t].l0 + j] + als[j][a];
                        if (isadmix && c->fast32) {
                            /* src/structureAVX.cpp:1148-1168: the mixture in float */
                            float s = 0;
                            for (int k = 0; k < K; ++k) s = s + (float)p[k * KT] * (float)q[k];
                            f[j] = (double)s;
                        } else if (isadmix) {
                            double s = 0;
                            for (int k = 0; k < K; ++k) s = s + p[k * KT] * q[k];
                            f[j] = s;
                        } else f[j] = p[(int64_t)c->Z[i] * KT];
                    } else f[j] = 1.0;
                }
                prods[((size_t)t * N + i) * mp + a] = tree64(f);
            }
        }
    }
    int64_t slog = 0; double prod = 1;
    for (int64_t t = 0; t < nt; ++t)
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < nslot[t * N + i]; ++a)
                or_charge_log(&slog, &prod, prods[((size_t)t * N + i) * mp + a]);
    free(prods); free(nslot); free(tl);
    c->lastlnl = or_close_log(slog, prod);
    return c->lastlnl;
}

/* Record, src/structure.cpp:1943-2045 */
void or_record(or_chain* c)
{
    int K = c->K, S = c->S; int64_t KT = c->KT;
    int m = c->m;
    if ((c->writelnl && m < c->nburnin && m % c->nthinning == 0) ||
        (m >= c->nburnin && (m - c->nburnin) % c->nthinning == 0)) {
        double lik = or_record_pass(c);
        if (c->writelnl) c->lnL[c->nlnL++] = lik;
        if (m >= c->nburnin) {
            c->nr++;
            c->rout[0] += lik;
            c->rout[1] += lik * lik;
            int nf = c->fmodel ? (c->fsame ? 1 : K) : 0;
            int na = c->admix ? (c->diffalpha ? K : 1) : 0;
            int nl = c->difflambda ? K : 1;
            for (int j = 0; j < nl; ++j) c->rout[4 + j] += c->Lambda[j];
            if (c->locpriori) {
                c->rout[4 + nl] += c->r;
                for (int k = 0; k < K; ++k) c->rout[4 + nl + 1 + k] += c->admix ? c->Alpha[k] : c->Eta[k];
                for (int j = 0; j < S * K; ++j) c->rout[4 + nl + 1 + K + j] += c->admix ? c->AlphaLocal[j] : c->Gamma[j];
            } else if (c->admix)
                for (int j = 0; j < na; ++j) c->rout[4 + nl + j] += c->Alpha[j];
            for (int j = 0; j < nf; ++j) c->rout[c->rlen - nf + j] += c->F[j];
            if (c->real32)   /* Add(FreqM, Freq) on float arrays */
                for (int64_t j = 0; j < K * KT; ++j) c->FreqM[j] = (double)((float)c->FreqM[j] + (float)c->Freq[j]);
            else
                for (int64_t j = 0; j < K * KT; ++j) c->FreqM[j] += c->Freq[j];
        }
    }
}

/* Arrange, src/structure.cpp:1862-1869 */
void or_arrange(or_chain* c)
{
    int K = c->K, N = c->N; int64_t KT = c->KT;
    double* ms = (double*)c->MiSum;
    for (int64_t j = 0; j < (int64_t)N * K; ++j) ms[j] = (double)c->MiSum[j];
    for (int i = 0; i < N; ++i) or_unify_avx(ms + (int64_t)i * K, K);
    c->misum_is_double = 1;
    double inv = 1.0 / c->nr;
    for (int j = 0; j < c->rlen; ++j) c->rout[j] *= inv;
    c->rout[2] = c->rout[1] - c->rout[0] * c->rout[0];
    c->rout[3] = c->rout[0] - c->rout[2] / 2;
    if (c->real32) {   /* Mul(FreqM, (REAL)(1.0 / nr)) on a float array */
        float invf = (float)(1.0 / c->nr);
        for (int64_t j = 0; j < K * KT; ++j) c->FreqM[j] = (double)((float)c->FreqM[j] * invf);
    } else {
        double invf = 1.0 / c->nr;
        for (int64_t j = 0; j < K * KT; ++j) c->FreqM[j] *= invf;
    }
}

/* MCMC body, src/structure.cpp:2122-2158 */
void or_sweep(or_chain* c, int n)
{
    for (int s = 0; s < n; ++s) {
        or_update_p(c);
        or_update_q(c);
        or_update_locpriori(c);
        or_update_z(c);
        or_update_alpha(c);
        or_update_lambda(c);
        or_update_a(c);
        or_update_f(c);
        or_record(c);
        c->m++;
    }
}
